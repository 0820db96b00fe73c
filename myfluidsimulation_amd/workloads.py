"""Synthetic workloads: 3D Taylor-Green vortex fields, seeded jittered
icosphere obstacles written as binary STL, and obstacle-channel flag grids.

Random numbers come from splitmix64 (Steele, Lea & Flood, OOPSLA 2014) so
every consumer (tests, bench, the C++ executables' inputs) gets the same
geometry from the same seed.
"""
from __future__ import annotations

import math
import struct

import numpy as np

from .capi import FLAG_INLET, FLAG_OUTLET, FLAG_SOLID

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in [0, 1)."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))


# ---------------------------------------------------------------- Taylor-Green
def taylor_green_3d(nx: int, ny: int, nz: int, u0: float = 0.05, rho0: float = 1.0):
    """3D TGV at cell centres x = i + 0.5 with k = 2*pi/N per axis:
    u = u0 sin(kx)cos(ky)cos(kz), v = -u0 cos(kx)sin(ky)cos(kz), w = 0,
    rho = rho0 + (3 rho0 u0^2/16)(cos2kx + cos2ky)(cos2kz + 2).
    Returns float32 arrays (rho, ux, uy, uz) of nx*ny*nz, x fastest.
    The trigonometric factors are evaluated per axis and broadcast: every
    element goes through the same float64 operations in the same order as a
    full-grid evaluation, 20x faster at 512^3."""
    kx, ky, kz = 2 * np.pi / nx, 2 * np.pi / ny, 2 * np.pi / nz
    X = (kx * (np.arange(nx) + 0.5))[None, None, :]
    Y = (ky * (np.arange(ny) + 0.5))[None, :, None]
    Z = (kz * (np.arange(nz) + 0.5))[:, None, None]
    f32 = lambda a: np.ascontiguousarray(a.reshape(-1), dtype=np.float32)  # noqa: E731
    cz = np.cos(Z)
    u = f32(u0 * np.sin(X) * np.cos(Y) * cz)
    v = f32(-u0 * np.cos(X) * np.sin(Y) * cz)
    rho = f32(rho0 + (3.0 * rho0 * u0 * u0 / 16.0) * (np.cos(2 * X) + np.cos(2 * Y)) * (
        np.cos(2 * Z) + 2.0))
    return rho, u, v, np.zeros_like(u)


def tgv_energy_analytic(t, u0, tau, n):
    """Mean kinetic energy 0.5<|u|^2> of the linear-decay 3D TGV on an n^3
    periodic box: (u0^2/8) exp(-6 nu k^2 t), nu = (tau - 1/2)/3, k = 2pi/n."""
    nu = (tau - 0.5) / 3.0
    k = 2 * np.pi / n
    return (u0 * u0 / 8.0) * np.exp(-6.0 * nu * k * k * np.asarray(t, dtype=np.float64))


def kinetic_energy(rho, ux, uy, uz) -> float:
    """0.5 <|u|^2> over all cells (float64 accumulation)."""
    u = np.asarray(ux, np.float64)
    v = np.asarray(uy, np.float64)
    w = np.asarray(uz, np.float64)
    return float(0.5 * np.mean(u * u + v * v + w * w))


def fourier_modes(nx, ny, nz, n_modes=8, amp=0.01, seed=1):
    """Seeded velocity perturbation: a sum of n_modes periodic Fourier modes
    amp * d_m * sin(k_m . x + phi_m) with integer wave numbers k_m in
    [-2, 2]^3 (not all zero), phase phi_m ~ U[0, 2pi) and a unit direction
    d_m perpendicular to k_m (so each mode is divergence-free), at cell
    centres.  Returns float32 (ux, uy, uz) of nx*ny*nz, x fastest."""
    rng = SplitMix64(seed)
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5,
                          indexing="ij")
    u = [np.zeros_like(x) for _ in range(3)]
    for _ in range(n_modes):
        k = [0, 0, 0]
        while k == [0, 0, 0]:
            k = [min(4, int(rng.uniform() * 5)) - 2 for _ in range(3)]
        phi = 2 * np.pi * rng.uniform()
        r = np.array([2 * rng.uniform() - 1 for _ in range(3)])
        kv = np.array(k, float)
        d = r - kv * (r @ kv) / (kv @ kv)
        if np.linalg.norm(d) < 1e-6:
            d = np.cross(kv, [1.0, 0.0, 0.0] if abs(kv[0]) < abs(kv).max() else [0.0, 1.0, 0.0])
        d = d / np.linalg.norm(d)
        ph = 2 * np.pi * (k[0] * x / nx + k[1] * y / ny + k[2] * z / nz) + phi
        s = amp * np.sin(ph)
        for a in range(3):
            u[a] = u[a] + d[a] * s
    return tuple(np.ascontiguousarray(c.reshape(-1), dtype=np.float32) for c in u)


# ---------------------------------------------------------------- icosphere
def icosphere(subdiv: int = 3):
    """Unit icosphere: (vertices (V,3) float64, faces (F,3) int), outward
    counter-clockwise faces; 12 -> 42 -> 162 -> 642 vertices, 20*4^s faces."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, float) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts), np.array(faces, dtype=np.int64)


def jittered_icosphere(center, radius, seed, jitter=0.10, subdiv=3):
    """Icosphere with each unique vertex at radius R*(1 + jitter*U[-1,1])
    (splitmix64 from `seed`, one draw per vertex in vertex order), so the
    mesh stays closed.  Returns triangle soup (F, 9) float64."""
    v, f = icosphere(subdiv)
    rng = SplitMix64(seed)
    r = np.array([radius * (1.0 + jitter * (2.0 * rng.uniform() - 1.0)) for _ in range(len(v))])
    p = np.asarray(center, float)[None, :] + v * r[:, None]
    return p[f].reshape(-1, 9)


def write_binary_stl(path: str, tri_xyz) -> None:
    """Binary STL; the 80-byte header does not start with 'solid' (a file
    starting with 'solid' is parsed as ASCII first, stl_loader.cpp:108)."""
    t = np.asarray(tri_xyz, np.float32).reshape(-1, 9)
    header = b"fluidsim binary stl".ljust(80, b"\0")
    with open(path, "wb") as fh:
        fh.write(header)
        fh.write(struct.pack("<I", t.shape[0]))
        for tri in t:
            a, b, c = tri[0:3], tri[3:6], tri[6:9]
            n = np.cross(b - a, c - a)
            ln = np.linalg.norm(n)
            n = n / ln if ln > 0 else n
            fh.write(struct.pack("<3f", *n.astype(np.float32)))
            fh.write(tri.astype("<f4").tobytes())
            fh.write(b"\0\0")


def obstacle_stl(path, nx, ny, nz, center, radius, seed):
    """Write the jittered icosphere in unit-cube coordinates (x/nx, y/ny,
    z/nz): the voxelizer maps voxel (i,j,k) to ((i+.5)/nx, ...), so the
    sphere is pre-distorted per axis.  Returns (scale, translate) that make
    the fit transform the identity: scale = the bbox's largest side (so
    s = scale/maxDim = 1) and translate = the bbox centre."""
    soup = jittered_icosphere(center, radius, seed)
    unit = soup.reshape(-1, 3) / np.array([nx, ny, nz], float)
    unit = unit.astype(np.float32).reshape(-1, 9)
    write_binary_stl(path, unit)
    pts = unit.reshape(-1, 3)
    bmin, bmax = pts.min(axis=0), pts.max(axis=0)
    size = bmax - bmin
    scale = float(np.float32(max(size[0], size[1], size[2])))
    center32 = np.float32(0.5) * (bmin + bmax)
    return scale, tuple(float(c) for c in center32)


# ---------------------------------------------------------------- channel
def channel_flags(nx, ny, nz, solid=None):
    """Flag grid of an obstacle channel: solid y/z wall planes OR-ed with the
    obstacle, equilibrium inlet cells at x=0, zero-gradient outlet cells at
    x=nx-1 (fluid cells only; an outlet needs a fluid cell at x-1)."""
    fl = np.zeros((nz, ny, nx), np.uint8)
    if solid is not None:
        fl |= (np.asarray(solid, np.uint8).reshape(nz, ny, nx) != 0).astype(np.uint8) * FLAG_SOLID
    fl[:, 0, :] |= FLAG_SOLID
    fl[:, ny - 1, :] |= FLAG_SOLID
    fl[0, :, :] |= FLAG_SOLID
    fl[nz - 1, :, :] |= FLAG_SOLID
    fluid = (fl & FLAG_SOLID) == 0
    fl[:, :, 0] |= np.where(fluid[:, :, 0], FLAG_INLET, 0).astype(np.uint8)
    ok_out = fluid[:, :, nx - 1] & fluid[:, :, nx - 2]
    fl[:, :, nx - 1] |= np.where(ok_out, FLAG_OUTLET, 0).astype(np.uint8)
    return fl.reshape(-1)


def uniform_flow(n, u_in):
    rho = np.ones(n, np.float32)
    ux = np.full(n, u_in, np.float32)
    z = np.zeros(n, np.float32)
    return rho, ux, z, z.copy()


# ---------------------------------------------------------------- batched sweep
SWEEP_EDGES = tuple(range(32, 129, 16))  # 32, 48, ..., 128


def batch_sweep_specs(n_domains=64, seed=7, channel_u=0.04):
    """The batched B0 / tau sweep: each domain draws its three edges from
    {32, 48, ..., 128}, tau ~ U[0.52, 0.9], a method from {BGK, HOME, Hybrid}
    and a Hybrid band from {1, 2, 3}; the first half are Taylor-Green
    vortices (u0 ~ U[0.02, 0.08]), the second half channels with a seeded
    jittered icosphere of radius 0.15 * (smallest edge)."""
    rng = SplitMix64(seed)
    pick = lambda seq: seq[min(len(seq) - 1, int(rng.uniform() * len(seq)))]  # noqa: E731
    specs = []
    for i in range(n_domains):
        nx, ny, nz = pick(SWEEP_EDGES), pick(SWEEP_EDGES), pick(SWEEP_EDGES)
        tau = 0.52 + 0.38 * rng.uniform()
        method = pick((0, 1, 2))
        band = pick((1, 2, 3))
        spec = dict(index=i, method=method, nx=nx, ny=ny, nz=nz, tau=tau, band=band)
        if i < n_domains // 2:
            spec.update(kind="tgv", u0=0.02 + 0.06 * rng.uniform())
        else:
            spec.update(kind="channel", u_in=channel_u,
                        radius=0.15 * min(nx, ny, nz), center=(nx / 4.0, ny / 2.0, nz / 2.0),
                        stl_seed=seed * 1000 + i)
        specs.append(spec)
    return specs


def build_sweep_domain(spec, tmpdir):
    """Flags (or None), initial fields and inlet profile (or None) of one
    sweep domain; the channel obstacle is written as STL and voxelized on the
    GPU.  Returns (flags, fields, inlet)."""
    import os

    from . import solver as S

    nx, ny, nz = spec["nx"], spec["ny"], spec["nz"]
    if spec["kind"] == "tgv":
        return None, taylor_green_3d(nx, ny, nz, spec["u0"]), None
    path = os.path.join(tmpdir, f"sweep_{spec['index']}.stl")
    scale, trans = obstacle_stl(path, nx, ny, nz, spec["center"], spec["radius"],
                                spec["stl_seed"])
    tris, bmin, bmax = S.load_stl(path)
    solid = S.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
    flags = channel_flags(nx, ny, nz, solid)
    if spec["method"] == 2:
        band = S.band((flags & FLAG_SOLID).astype(np.uint8), nx, ny, nz, spec["band"])
        flags = flags | (band * 2).astype(np.uint8)
    n = nx * ny * nz
    return flags, uniform_flow(n, spec["u_in"]), np.full(ny * nz, spec["u_in"], np.float32)
