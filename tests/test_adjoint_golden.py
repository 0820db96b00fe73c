"""CPU: the adjoint fixtures (tests/golden/adjoint32.npz, written by
tests/golden/make_adjoint_golden.py from FP64 torch autograd of
tests/torch_lbm.py) agree with the FP32 oracle's forward runs: the stored
Taylor-Green objective J is reproduced by the oracle's end field within
float32 accuracy, and the stored channel inputs are what the oracle
voxelizer and band give for the seeded obstacle today."""
import os
import sys
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_adjoint_golden as mg  # noqa: E402

G = np.load(os.path.join(HERE, "golden", "adjoint32.npz"))
N, T, TAU = int(G["N"]), int(G["T"]), float(G["tau"])


@pytest.mark.parametrize("name,method", mg.METHODS)
def test_tgv_objective_matches_oracle(oracle, name, method):
    rho, u, _ = mg.tgv_inputs()
    o = oracle.Oracle(method, N, N, N, TAU)
    if method == 2:
        o.set_omegas(*[float(w) for w in G["hyb_omegas"]])
    o.reinit(rho, *u)
    o.step(T)
    _, ux, uy, uz = o.macros()
    tgt = G[f"tgv_{name}_target"]
    J = 0.5 * sum(float(((a.astype(np.float64) - t) ** 2).sum()) for a, t in zip((ux, uy, uz), tgt))
    assert J == pytest.approx(float(G[f"tgv_{name}_J"]), rel=1e-4)
    o.close()


def test_channel_inputs_regenerate():
    with tempfile.TemporaryDirectory() as td:
        ch = mg.channel_inputs(td)
    for k, v in ch.items():
        assert np.array_equal(v, G[f"channel_{k}"]), k
    assert ch["body"].sum() > 100 and ch["band"].sum() > ch["body"].sum()
