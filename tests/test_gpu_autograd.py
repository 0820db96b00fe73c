"""GPU: the torch.autograd.Function front end (DifferentiableRun) on the
Taylor-Green inverse problem of tests/golden/adjoint32.npz, in a child
process that imports torch first (the GPU test process itself stays free of
torch).

torch and libfluidsim can share a process only when torch's bundled HIP
runtime is the major.minor version the library was built against: imported
first, torch's runtime serves libfluidsim too (same soname); imported after,
a second runtime is mapped, and the library refuses both mismatches
(fs_device_count / fs_solver_create / fs_solver_set_stream return
FS_ERR_DEVICE).  The test runs only where the versions agree.
"""
import os
import re
import subprocess
import sys

import pytest

from myfluidsimulation_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _torch_hip():
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return None
    txt = open(os.path.join(os.path.dirname(spec.origin), "version.py")).read()
    m = re.search(r"^hip\b.*?=\s*'(\d+)\.(\d+)", txt, re.M)
    return (int(m.group(1)), int(m.group(2))) if m else None


def _built():
    b, _ = capi.hip_versions()
    return b // 10000000, (b // 100000) % 100


pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(_torch_hip() != _built(), reason=(
        f"torch bundles HIP {_torch_hip()} but libfluidsim is built against HIP {_built()}: "
        "one process cannot hold both, and the library refuses the mismatch")),
]

CHILD = r"""
import os, sys
import torch  # first: libfluidsim then binds to torch's HIP runtime
import numpy as np
sys.path[:0] = [sys.argv[1], os.path.join(sys.argv[1], "tests", "golden")]
import make_adjoint_golden as mg
from myfluidsimulation_amd import Solver
from myfluidsimulation_amd.adjoint import DifferentiableRun

G = np.load(os.path.join(sys.argv[1], "tests", "golden", "adjoint32.npz"))
N, T, tau = int(G["N"]), int(G["T"]), float(G["tau"])
rho, u, _ = mg.tgv_inputs()
dev = "cuda"
s = Solver(0, N, N, N, tau)
run = DifferentiableRun(s, T)
u0 = [torch.tensor(a, device=dev).requires_grad_() for a in u]
_, ux, uy, uz, _ = run(*u0, rho0=torch.tensor(rho, device=dev))
tg = torch.tensor(G["tgv_bgk_target"], device=dev)
loss = 0.5 * (((ux - tg[0]) ** 2).sum() + ((uy - tg[1]) ** 2).sum() + ((uz - tg[2]) ** 2).sum())
loss.backward()
g = torch.stack([x.grad for x in u0]).double().cpu().numpy()
ref = G["tgv_bgk_grad"].astype(np.float64)
err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
print("rel", err)
assert err <= 1e-4
run.close()
s.close()
print("AUTOGRAD_OK")
"""


def test_autograd_function_tgv(fs):
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True,
                       timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "AUTOGRAD_OK" in r.stdout
