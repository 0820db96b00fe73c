"""Differentiable FP64 torch formulation of the three update rules, with the
momentum-exchange drag (test helper: the reference for the adjoint kernels).

Tensor form -- populations (19, nz, ny, nx), torch.roll for streaming and
torch.where for walls, inlets and outlets -- written from the update rules,
not from the HIP kernels or the oracle's loops.  Every rule matches the
solver's step: BGK f* = f + omega (feq(rho, u) - f) with u = j/rho + F (the
Hybrid band uses omega_L, the bulk omega_S); HOME rebuilds the regularised
post-collision populations from 10 moments, streams them and re-projects;
halfway bounce-back; inlet cells emit feq(1, (u_in, 0, 0)); an outlet cell
emits what its upstream neighbour x-1 emits (populations) or takes its new
moments (HOME).  The end fields are the ones the solver reports: the
collision moments of the last step (populations) or the end moments (HOME).

The drag of step n is D^n = sum over links (x non-solid, x + e_q a body
cell, e_qx != 0) of 2 e_qx f*_q(x), f* the post-collision populations.
"""
import numpy as np
import torch

E = np.array([
    (0, 0, 0),
    (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0),
    (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
    (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)], dtype=np.int64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = [int(np.where((E == -e).all(axis=1))[0][0]) for e in E]
# the reference's pairing table (mirror images for q >= 7), HOME only
OPP_REF = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]
CS2 = 1.0 / 3.0
SOLID, BAND, INLET, OUTLET = 1, 2, 4, 8
BGK, HOME, HYBRID = 0, 1, 2


def _shift(a, q, sign=1):
    """a(x - sign*e_q): roll by +sign*e_q over the trailing (z, y, x) axes."""
    ex, ey, ez = (int(v) * sign for v in E[q])
    return torch.roll(a, shifts=(ez, ey, ex), dims=(-3, -2, -1))


class TorchLBM:
    def __init__(self, method, nx, ny, nz, tau, force=(0.0, 0.0, 0.0), flags=None,
                 omegas=None, bounce=0, body=None, dtype=torch.float64, device="cpu"):
        self.method = method
        self.shape = (nz, ny, nx)
        self.dt = dict(dtype=dtype, device=device)
        fl = np.zeros(self.shape, np.uint8) if flags is None else np.asarray(
            flags, np.uint8).reshape(self.shape)
        b = lambda m: torch.as_tensor((fl & m) != 0, device=device)  # noqa: E731
        self.solid, self.band, self.inlet, self.outlet = b(SOLID), b(BAND), b(INLET), b(OUTLET)
        body = self.solid if body is None else torch.as_tensor(
            np.asarray(body).reshape(self.shape) != 0, device=device) & self.solid
        om = 1.0 / tau
        oL, oH, oS = (om, om, om) if omegas is None else omegas
        if method == HYBRID:
            self.omega = torch.where(self.band, torch.tensor(oL, **self.dt),
                                     torch.tensor(oS, **self.dt))
        else:
            self.omega = torch.full(self.shape, oL, **self.dt)
        self.omega_l = oL
        self.F = torch.tensor(force, **self.dt).reshape(3, 1, 1, 1)
        self.table = OPP if bounce == 0 else OPP_REF
        self.E = torch.tensor(E, **self.dt)
        self.W = torch.tensor(W, **self.dt)
        # link masks: x non-solid and x + e_q a body cell, for e_qx != 0
        self.links = {q: (~self.solid) & _shift(body, q, -1) for q in range(1, 19) if E[q, 0]}

    # ---- pieces
    def feq(self, rho, u):
        out = []
        uu = (u * u).sum(0)
        for q in range(19):
            eu = E[q, 0] * u[0] + E[q, 1] * u[1] + E[q, 2] * u[2]
            out.append(W[q] * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uu))
        return torch.stack(out)

    def moments(self, f):
        rho = f.sum(0)
        j = torch.stack([sum(E[q, a] * f[q] for q in range(19) if E[q, a]) for a in range(3)])
        M = {}
        for a in range(3):
            for b in range(a, 3):
                M[a, b] = sum(E[q, a] * E[q, b] * f[q] for q in range(19) if E[q, a] * E[q, b])
        return rho, j, M

    def drag(self, post):
        return sum(2.0 * E[q, 0] * (post[q] * m).sum() for q, m in self.links.items())

    def inlet_u(self, uin):
        zero = torch.zeros(self.shape, **self.dt)
        return torch.stack([uin[:, :, None].expand(self.shape), zero, zero])

    # ---- runs
    def run(self, rho0, u0, uin, T):
        """rho0 (nz,ny,nx), u0 (3,nz,ny,nx), uin (nz,ny) -> (rho, u, drag[T])"""
        if self.method == HOME:
            return self._run_home(rho0, u0, uin, T)
        return self._run_pop(rho0, u0, uin, T)

    def _run_pop(self, rho0, u0, uin, T):
        f = self.feq(rho0, u0)
        one = torch.ones(self.shape, **self.dt)
        feq_in = self.feq(one, self.inlet_u(uin))
        drags = []
        rho = u = None
        for _ in range(T):
            rho, j, _ = self.moments(f)
            u = j / rho + self.F
            post = f + self.omega * (self.feq(rho, u) - f)
            post = torch.where(self.inlet, feq_in, post)
            rho = torch.where(self.inlet, one, rho)
            u = torch.where(self.inlet, self.inlet_u(uin), u)
            post = torch.where(self.outlet, torch.roll(post, 1, dims=-1), post)
            rho = torch.where(self.outlet, torch.roll(rho, 1, dims=-1), rho)
            u = torch.where(self.outlet, torch.roll(u, 1, dims=-1), u)
            drags.append(self.drag(post))
            new = []
            for q in range(19):
                if q == 0:
                    new.append(post[0])
                    continue
                new.append(torch.where(_shift(self.solid, q), post[OPP[q]], _shift(post[q], q)))
            f = torch.where(self.solid, f, torch.stack(new))
            rho = torch.where(self.solid, one, rho)
            u = torch.where(self.solid, torch.zeros_like(u), u)
        return rho, u, torch.stack(drags)

    def rec(self, rho, u, S):
        """regularised post-collision populations from the moments"""
        feq = self.feq(rho, u)
        out = []
        for q in range(19):
            c = [float(v) for v in E[q]]
            contr = 0.0
            for a in range(3):
                contr = contr + (c[a] * c[a] - CS2) * S[a, a]
            for a, b in ((0, 1), (0, 2), (1, 2)):
                if c[a] * c[b]:
                    contr = contr + 2.0 * c[a] * c[b] * S[a, b]
            out.append(feq[q] + (1.0 - self.omega_l) * 4.5 * W[q] * contr)
        return torch.stack(out)

    def _run_home(self, rho0, u0, uin, T):
        zero = torch.zeros(self.shape, **self.dt)
        one = torch.ones(self.shape, **self.dt)
        rho, u = rho0, u0
        S = {(a, b): zero for a in range(3) for b in range(a, 3)}
        drags = []
        for _ in range(T):
            post = self.rec(rho, u, S)
            drags.append(self.drag(post))
            inc = [post[0]]
            for q in range(1, 19):
                inc.append(torch.where(_shift(self.solid, q), post[self.table[q]],
                                       _shift(post[q], q)))
            r, j, M = self.moments(torch.stack(inc))
            un = j / r + self.F
            Sn = {}
            for a in range(3):
                for b in range(a, 3):
                    Sn[a, b] = M[a, b] - (r * (CS2 if a == b else 0.0) + r * un[a] * un[b])

            def fix(v, at_inlet, at_solid):
                v = torch.where(self.outlet, torch.roll(v, 1, dims=-1), v)
                v = torch.where(self.inlet, at_inlet, v)
                return torch.where(self.solid, at_solid, v)

            rho = fix(r, one, one)
            u = torch.stack([fix(un[0], uin[:, :, None].expand(self.shape), zero),
                             fix(un[1], zero, zero), fix(un[2], zero, zero)])
            S = {k: fix(v, zero, zero) for k, v in Sn.items()}
        return rho, u, torch.stack(drags)
