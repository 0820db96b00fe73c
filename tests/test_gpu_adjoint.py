"""GPU: gradients of T-step runs from the adjoint kernels (fs_adjoint_*,
through the C ABI with host staging; no torch in this process) against the
FP64 reference gradients of tests/golden/adjoint32.npz, which
tests/golden/make_adjoint_golden.py computed on the CPU with torch autograd
of the FP64 formulation tests/torch_lbm.py (pinned against the FP64 oracle by
tests/test_torch_ref.py), and against the FP64 central differences on 32
seeded random directions stored with them.

Objectives (32^3, T = 40, FP32 kernels):
  (a) Taylor-Green inverse problem: J = 1/2 |u(T) - target|^2, the target
      the end field of a run from the vortex plus 8 seeded Fourier modes
      (amplitude 0.01); gradient with respect to the initial velocity.
  (b) channel with a seeded icosphere: J = mean over the steps of the drag
      on the obstacle; gradient with respect to the inlet profile.
Tolerances: relative L2 error of the gradient against FP64 autograd
<= 1e-4; every directional derivative within 1e-4 of the central
difference, relative to max(|central difference|, |grad| / sqrt(dim)) (a
direction nearly orthogonal to the gradient makes the difference itself
close to zero, and a plain relative error then measures only rounding).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_adjoint_golden as mg  # noqa: E402  (direction_set and the case constants)

pytestmark = pytest.mark.gpu

G = np.load(os.path.join(HERE, "golden", "adjoint32.npz"))
N, T, TAU = int(G["N"]), int(G["T"]), float(G["tau"])
METHODS = [name for name, _ in mg.METHODS]
CODE = dict(mg.METHODS)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def check_directions(grad, fd, shape):
    """<grad, v> against the stored central differences along the same
    seeded directions"""
    g = np.asarray(grad, np.float64).reshape(shape)
    rms = float(np.linalg.norm(g)) / np.sqrt(g.size)
    worst = 0.0
    for v, d in zip(mg.direction_set(shape, int(G["ndir"]), int(G["dir_seed"])), fd):
        ad = float((g * v).sum())
        rel = abs(ad - d) / max(abs(d), rms)
        worst = max(worst, rel)
        assert rel <= 1e-4, f"directional derivative {ad} vs central difference {d}"
    return worst


def make_solver(fs, name, flags):
    s = fs.Solver(CODE[name], N, N, N, TAU)
    s.set_flags(flags)
    if name == "hybrid":
        s.set_omegas(*[float(w) for w in G["hyb_omegas"]])
    return s


def channel_flags(name):
    fl = G["channel_flags"]
    return fl | (G["channel_band"] * 2).astype(np.uint8) if name == "hybrid" else fl


@pytest.mark.parametrize("name", METHODS)
def test_tgv_inverse_gradient(fs, name):
    from myfluidsimulation_amd.solver import Adjoint

    rho, u, _ = mg.tgv_inputs()
    print(f"adjoint tgv {name}: N {N} T {T}", flush=True)
    s = make_solver(fs, name, np.zeros(N ** 3, np.uint8))
    adj = Adjoint(s, T)
    try:
        _, ux, uy, uz, _ = adj.forward(rho, *u)
        tgt = G[f"tgv_{name}_target"]
        du = [a - t for a, t in zip((ux, uy, uz), tgt)]
        J = 0.5 * sum(float((d.astype(np.float64) ** 2).sum()) for d in du)
        assert J == pytest.approx(float(G[f"tgv_{name}_J"]), rel=1e-4)
        _, gx, gy, gz, _ = adj.backward(None, *du, None, want=(False, True, True, True, False))
        grad = np.stack([gx, gy, gz])
        err = rel_l2(grad, G[f"tgv_{name}_grad"])
        print(f"{name}: J {J:.6e}  rel L2 err vs FP64 autograd {err:.2e}")
        assert err <= 1e-4
        worst = check_directions(grad, G[f"tgv_{name}_fd"], (3, N, N, N))
        print(f"{name}: worst directional rel err vs central differences {worst:.2e}")
    finally:
        adj.close()
        s.close()


@pytest.mark.parametrize("name", METHODS)
def test_channel_drag_gradient(fs, name):
    from myfluidsimulation_amd.solver import Adjoint

    print(f"adjoint channel {name}: N {N} T {T}", flush=True)
    s = make_solver(fs, name, channel_flags(name))
    adj = Adjoint(s, T, body=G["channel_body"])
    try:
        *_, drag = adj.forward(G["channel_rho"], G["channel_ux"], G["channel_uy"],
                               G["channel_uz"], G["channel_inlet"])
        J = float(drag.astype(np.float64).mean())
        assert abs(J) > 1e-3
        assert J == pytest.approx(float(G[f"channel_{name}_J"]), rel=1e-4)
        gd = np.full(T, 1.0 / T, np.float32)
        *_, gin = adj.backward(None, None, None, None, gd, want=(False, False, False, False, True))
        err = rel_l2(gin, G[f"channel_{name}_grad"])
        print(f"{name}: drag {J:.6f}  rel L2 err vs FP64 autograd {err:.2e}")
        assert err <= 1e-4
        worst = check_directions(gin, G[f"channel_{name}_fd"], (N, N))
        print(f"{name}: worst directional rel err vs central differences {worst:.2e}")
    finally:
        adj.close()
        s.close()


@pytest.mark.parametrize("name", ["bgk", "home"])
def test_segments_give_identical_gradients(fs, name):
    """recomputation is bit-exact, so every checkpoint spacing gives the same
    gradient bits, and a second backward pass repeats the first"""
    from myfluidsimulation_amd.solver import Adjoint

    steps = 12
    flags = channel_flags(name)
    rng = np.random.default_rng(3)
    gux = (1e-3 * rng.standard_normal(N ** 3)).astype(np.float32)
    gd = np.linspace(0.5, 1.5, steps).astype(np.float32)
    out = []
    for seg in (1, 5, steps):
        print(f"adjoint segments {name}: segment {seg}", flush=True)
        s = make_solver(fs, name, flags)
        adj = Adjoint(s, steps, segment=seg, body=G["channel_body"])
        try:
            assert adj.info()["checkpoints"] == -(-steps // seg)
            adj.forward(G["channel_rho"], G["channel_ux"], G["channel_uy"], G["channel_uz"],
                        G["channel_inlet"])
            first = adj.backward(None, gux, None, None, gd)
            again = adj.backward(None, gux, None, None, gd)
            for a, b in zip(first, again):
                assert np.array_equal(a, b)
            assert np.abs(first[1]).max() > 0 and np.abs(first[4]).max() > 0
            out.append(first)
        finally:
            adj.close()
            s.close()
    for o in out[1:]:
        for a, b in zip(o, out[0]):
            assert np.array_equal(a, b)


def test_adjoint_rejects_bad_use(fs):
    from myfluidsimulation_amd.capi import FS_ERR_STATE, FluidSimError
    from myfluidsimulation_amd.solver import Adjoint

    s = fs.Solver(0, 16, 8, 8, 0.6)
    with pytest.raises(FluidSimError):
        Adjoint(s, 0)
    a = Adjoint(s, 4)
    with pytest.raises(FluidSimError) as e:
        a.backward_ptrs(*([None] * 10))
    assert e.value.code == FS_ERR_STATE
    a.close()
    fl = np.zeros((8, 8, 16), np.uint8)
    fl[:, :, 3] = 4  # inlet cells away from x = 0
    s.set_flags(fl.reshape(-1))
    with pytest.raises(FluidSimError):
        Adjoint(s, 4)
    s.close()
