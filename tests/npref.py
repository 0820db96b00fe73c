"""Independent FP64 numpy formulation of the three update rules (test helper).

Written in tensor form (population arrays of shape (19, nz, ny, nx),
np.roll for streaming) from the algorithm's description, not from the
oracle's loops, so a dropped term, a wrong sign or a swapped index in either
shows up as a mismatch.  The velocity ordering is the one the reference lists
in its comments (lbm3d_legacy.cu:16-21); the bounce-back partner is the
opposite direction.
"""
import numpy as np

E = np.array([
    (0, 0, 0),
    (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0),
    (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
    (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)], dtype=np.int64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = np.array([int(np.where((E == -e).all(axis=1))[0][0]) for e in E])
CS2 = 1.0 / 3.0
SOLID, BAND, INLET, OUTLET = 1, 2, 4, 8


def grid(a, nx, ny, nz):
    return np.asarray(a, np.float64).reshape(nz, ny, nx)


def feq(rho, u):
    """rho (...), u (3, ...) -> (19, ...)"""
    eu = np.tensordot(E.astype(np.float64), u, axes=(1, 0))
    uu = (u * u).sum(0)
    return W.reshape(-1, *([1] * rho.ndim)) * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uu)


def pull(post, solid):
    """incoming[q](c) = post[q](c - e_q), or post[opp q](c) when c - e_q is solid"""
    out = np.empty_like(post)
    for q in range(19):
        sh = (int(E[q, 2]), int(E[q, 1]), int(E[q, 0]))
        src = np.roll(post[q], sh, axis=(0, 1, 2))
        src_solid = np.roll(solid, sh, axis=(0, 1, 2))
        out[q] = np.where(src_solid, post[OPP[q]], src)
    return out


def moments(f):
    rho = f.sum(0)
    j = np.tensordot(E.T.astype(np.float64), f, axes=(1, 0))
    M = np.einsum("qa,qb,q...->ab...", E.astype(np.float64), E.astype(np.float64), f)
    return rho, j, M


class PopRef:
    """BGK (and Hybrid: with equal rates every cell is BGK) on populations."""

    def __init__(self, nx, ny, nz, tau, force=(0, 0, 0), flags=None, inlet=None):
        self.shape = (nz, ny, nx)
        self.omega = 1.0 / tau
        self.F = np.array(force, np.float64).reshape(3, 1, 1, 1)
        fl = np.zeros(self.shape, np.uint8) if flags is None else np.asarray(flags).reshape(
            self.shape)
        self.solid = (fl & SOLID) != 0
        self.inlet = (fl & INLET) != 0
        self.outlet = (fl & OUTLET) != 0
        self.uin = np.zeros((nz, ny)) if inlet is None else np.asarray(inlet, np.float64).reshape(
            nz, ny)
        self.f = None

    def init_feq(self, rho, u):
        self.f = feq(rho, u)
        self.rho, self.u = rho.copy(), u.copy()

    def step(self):
        f = self.f
        if self.outlet.any():
            f[:, self.outlet] = np.roll(f, 1, axis=3)[:, self.outlet]
        rho, j, _ = moments(f)
        u = j / rho + self.F
        post = f + self.omega * (feq(rho, u) - f)
        if self.inlet.any():
            uin = np.zeros((3,) + self.shape)
            uin[0] = self.uin[:, :, None]
            post[:, self.inlet] = feq(np.ones(self.shape), uin)[:, self.inlet]
            rho = np.where(self.inlet, 1.0, rho)
            u = np.where(self.inlet[None], uin, u)
        new = pull(post, self.solid)
        new[:, self.solid] = f[:, self.solid]
        self.rho = np.where(self.solid, 1.0, rho)
        self.u = np.where(self.solid[None], 0.0, u)
        self.f = new


class HomeRef:
    """HOME: 10 moments per cell; regularised post-collision populations
    f*_q = feq(rho,u) + (1-omega) * w_q/(2 cs^4) * (e_q e_q - cs2 I) : S."""

    def __init__(self, nx, ny, nz, tau, force=(0, 0, 0), flags=None, inlet=None):
        self.pop = PopRef(nx, ny, nz, tau, force, flags, inlet)
        self.omega = 1.0 / tau

    def init(self, rho, u):
        self.rho, self.u = rho.copy(), u.copy()
        self.S = np.zeros((3, 3) + rho.shape)

    def step(self):
        p = self.pop
        Q = np.einsum("qa,qb->qab", E.astype(np.float64), E.astype(np.float64)) - CS2 * np.eye(3)
        fneq = np.einsum("qab,ab...->q...", Q, self.S) * (W / (2 * CS2 * CS2)).reshape(
            -1, 1, 1, 1)
        post = feq(self.rho, self.u) + (1 - self.omega) * fneq
        inc = pull(post, p.solid)
        rho, j, M = moments(inc)
        u = j / rho + p.F
        S = M - rho * (CS2 * np.eye(3)[:, :, None, None, None] + u[:, None] * u[None, :])
        if p.outlet.any():
            rho = np.where(p.outlet, np.roll(rho, 1, axis=2), rho)
            u = np.where(p.outlet[None], np.roll(u, 1, axis=3), u)
            S = np.where(p.outlet[None, None], np.roll(S, 1, axis=4), S)
        if p.inlet.any():
            rho = np.where(p.inlet, 1.0, rho)
            u[0] = np.where(p.inlet, p.uin[:, :, None], u[0])
            u[1:] = np.where(p.inlet[None], 0.0, u[1:])
            S = np.where(p.inlet[None, None], 0.0, S)
        self.rho = np.where(p.solid, 1.0, rho)
        self.u = np.where(p.solid[None], 0.0, u)
        self.S = np.where(p.solid[None, None], 0.0, S)

    def moment_planes(self):
        """(rho, ux, uy, uz, Sxx, Sxy, Sxz, Syy, Syz, Szz) flattened"""
        S = self.S
        planes = [self.rho, self.u[0], self.u[1], self.u[2], S[0, 0], S[0, 1], S[0, 2], S[1, 1],
                  S[1, 2], S[2, 2]]
        return np.concatenate([a.reshape(-1) for a in planes])
