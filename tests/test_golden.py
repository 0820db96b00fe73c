"""The committed fixtures of tests/golden (FP32 oracle outputs on small
seeded inputs; tests/golden/make_golden.py says how they were made):

* CPU: the oracle still reproduces them bit for bit (a change that moves the
  oracle's arithmetic, its streaming or its voxelizer fails here);
* GPU: libfluidsim reproduces them bit for bit (the solver, all three methods
  with walls, obstacle, inlet, outlet, forcing and Hybrid rates, the
  voxelizer at refine 1 and 2, the band).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as mg  # noqa: E402


def load(name):
    with np.load(os.path.join(HERE, "golden", name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_oracle_reproduces_lbm_fixture(oracle):
    g = load("lbm_channel.npz")
    inp = mg.lbm_inputs()
    for k, a in zip(("flags", "rho", "ux", "uy", "uz", "inlet"), inp):
        assert np.array_equal(g[k], a), k
    out = mg.lbm_outputs(*inp)
    for k, a in out.items():
        assert np.array_equal(g[k], a), k


def test_oracle_reproduces_voxel_fixture(oracle, tmp_path):
    g = load("voxel.npz")
    out = mg.voxel_case(str(tmp_path))
    for k, a in out.items():
        assert np.array_equal(g[k], np.asarray(a)), k


@pytest.mark.gpu
@pytest.mark.parametrize("name,method", [("bgk", 0), ("home", 1), ("hybrid", 2)])
def test_gpu_reproduces_lbm_fixture(fs, name, method):
    g = load("lbm_channel.npz")
    s = fs.Solver(method, mg.NX, mg.NY, mg.NZ, mg.TAU, (1e-5, 0.0, 0.0))
    s.set_flags(g["flags"])
    s.set_inlet_profile(g["inlet"])
    if method == 2:
        s.set_omegas(1.0 / mg.TAU, 1.3, 1.45)
    s.reinit_from_macro(g["rho"], g["ux"], g["uy"], g["uz"])
    s.step(mg.STEPS)
    for k in range(4):
        assert np.array_equal(s.macro(k), g[f"{name}_macro{k}"]), k
    assert np.array_equal(s.state(), g[f"{name}_state"])
    s.close()


@pytest.mark.gpu
def test_gpu_reproduces_voxel_fixture(fs):
    g = load("voxel.npz")
    nx, ny, nz = (int(v) for v in g["dims"])
    for r in (1, 2):
        got = fs.voxelize(g["tris"], g["bmin"], g["bmax"], nx, ny, nz, float(g["scale"]),
                          tuple(float(t) for t in g["trans"]), refine=r, threshold=0.5)
        assert np.array_equal(got, g[f"solid_r{r}"]), r
    for d0 in (1, 2):
        assert np.array_equal(fs.band(g["solid_r1"], nx, ny, nz, d0), g[f"band_d{d0}"])
