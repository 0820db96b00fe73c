"""Writes the regression fixtures in this directory (run from the repository
root: python tests/golden/make_golden.py).

The reference ships no golden vectors, known-answer tests or fixtures for this
path, and none of its hot path is Python, so nothing of it is imported or run
here.  These fixtures are the FP32 oracle's own outputs on small seeded inputs,
committed once the oracle was pinned (tests/test_oracle_lbm.py,
tests/test_oracle_geom.py, tests/test_torch_ref.py): tests/test_golden.py then
fails if a later change moves any of them.  Contents:

  lbm_channel.npz  a 20x12x10 obstacle channel (walls, inlet 0.04, outlet,
                   box obstacle, Hybrid band d0=1), 25 steps from uniform flow
                   with small seeded perturbations: inputs (flags, rho, ux,
                   uy, uz, inlet) and, per method (bgk, home, hybrid), the
                   four macro fields and the full state
  voxel.npz        a seeded jittered icosphere (1280 faces) in a 40x32x24 grid
                   at refine 1 and 2: the triangle soup, its bounding box and
                   the flag grids, plus the band (d0 = 1, 2) of the refine-1
                   grid
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from myfluidsimulation_amd import workloads as wl  # noqa: E402

NX, NY, NZ, STEPS, TAU = 20, 12, 10, 25, 0.62


def lbm_inputs():
    rng = np.random.default_rng(20261203)
    n = NX * NY * NZ
    obst = np.zeros((NZ, NY, NX), np.uint8)
    obst[3:6, 4:7, 6:9] = 1
    flags = wl.channel_flags(NX, NY, NZ, obst.reshape(-1))
    band = pyoracle.band((flags & 1).astype(np.uint8), NX, NY, NZ, 1)
    flags = flags | (band * 2).astype(np.uint8)
    rho = (1.0 + 1e-3 * rng.standard_normal(n)).astype(np.float32)
    ux = (0.04 + 2e-3 * rng.standard_normal(n)).astype(np.float32)
    uy = (2e-3 * rng.standard_normal(n)).astype(np.float32)
    uz = (2e-3 * rng.standard_normal(n)).astype(np.float32)
    inlet = (0.04 + 2e-3 * rng.standard_normal(NY * NZ)).astype(np.float32)
    return flags, rho, ux, uy, uz, inlet


def lbm_outputs(flags, rho, ux, uy, uz, inlet):
    out = {}
    for name, m in (("bgk", 0), ("home", 1), ("hybrid", 2)):
        o = pyoracle.Oracle(m, NX, NY, NZ, TAU, (1e-5, 0.0, 0.0))
        o.set_flags(flags)
        o.set_inlet(inlet)
        if m == 2:
            o.set_omegas(1.0 / TAU, 1.3, 1.45)
        o.reinit(rho, ux, uy, uz)
        o.step(STEPS)
        for k, a in enumerate(o.macros()):
            out[f"{name}_macro{k}"] = a
        out[f"{name}_state"] = o.state()
    return out


def voxel_case(tmp):
    nx, ny, nz = 40, 32, 24
    path = os.path.join(tmp, "ico.stl")
    scale, trans = wl.obstacle_stl(path, nx, ny, nz, (18.0, 15.0, 12.0), 7.5, seed=777)
    tris, bmin, bmax = pyoracle.load_stl(path)
    out = dict(tris=tris, bmin=bmin, bmax=bmax, scale=np.float32(scale),
               trans=np.array(trans, np.float32), dims=np.array([nx, ny, nz]))
    for r in (1, 2):
        out[f"solid_r{r}"] = pyoracle.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans, r, 0.5)
    for d0 in (1, 2):
        out[f"band_d{d0}"] = pyoracle.band(out["solid_r1"], nx, ny, nz, d0)
    return out


def main():
    import tempfile

    pyoracle.build()
    inp = lbm_inputs()
    outs = lbm_outputs(*inp)
    np.savez_compressed(os.path.join(HERE, "lbm_channel.npz"),
                        **dict(zip(("flags", "rho", "ux", "uy", "uz", "inlet"), inp)), **outs)
    with tempfile.TemporaryDirectory() as td:
        np.savez_compressed(os.path.join(HERE, "voxel.npz"), **voxel_case(td))
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
