"""GPU: the 64^3 Taylor-Green case (tau 0.6, u0 0.05, 1000 steps) against the
FP64 oracle and the analytic kinetic-energy decay (u0^2/8) exp(-6 nu k^2 t).

test_gpu_parity.py shows the FP32 fields equal the FP32 oracle bit for bit
after 1000 steps; here the same GPU run, sampled every 10 steps, is set
against the FP64 oracle (the FP32 rounding error: end fields within 1e-4 of
rho0 and u0, the energy history within 5e-5 relative) and against the
analytic decay.  The analytic form is the linear (low-Reynolds) decay; at
u0 = 0.05 (Re about 15) the vortex's nonlinear energy transfer makes the
lattice energy fall below it, by 23% after 1000 steps, in the FP64 oracle as
on the GPU.  A low-amplitude control (u0 = 0.005, Re about 1.5) follows the
analytic decay to 1.5% over the same 1000 steps, which separates that
physics from any error of the step.
"""
import numpy as np
import pytest

from myfluidsimulation_amd import workloads as wl

pytestmark = pytest.mark.gpu

N, TAU, STEPS, EVERY = 64, 0.6, 1000, 10


def gpu_history(fs, method, u0):
    fields = wl.taylor_green_3d(N, N, N, u0)
    with fs.Solver(method, N, N, N, TAU) as s:
        s.reinit_from_macro(*fields)
        e = []
        for _ in range(STEPS // EVERY):
            s.step(EVERY)
            e.append(wl.kinetic_energy(*s.macros()))
        return np.array(e), s.macros()


def times(method):
    # BGK / Hybrid report the moments entering their last step
    lag = 0 if method == 1 else 1
    return np.arange(EVERY, STEPS + 1, EVERY) - lag


@pytest.mark.parametrize("method", [0, 1, 2], ids=["bgk", "home", "hybrid"])
def test_tgv64_fp64_and_energy(fs, oracle, method):
    u0 = 0.05
    e_gpu, end = gpu_history(fs, method, u0)
    o = oracle.Oracle(method, N, N, N, TAU, precision=oracle.F64)
    o.reinit(*wl.taylor_green_3d(N, N, N, u0))
    e64 = []
    for _ in range(STEPS // EVERY):
        o.step(EVERY)
        e64.append(wl.kinetic_energy(*o.macros(dtype=np.float64)))
    ref = o.macros(dtype=np.float64)
    o.close()
    e64 = np.array(e64)
    drho = float(np.abs(end[0].astype(np.float64) - ref[0]).max())
    du = max(float(np.abs(end[k].astype(np.float64) - ref[k]).max()) for k in (1, 2, 3)) / u0
    de = float(np.abs(e_gpu / e64 - 1).max())
    ea = wl.tgv_energy_analytic(times(method), u0, TAU, N)
    dev = e_gpu / ea - 1
    print(f"\nmethod {method}: FP32 GPU vs FP64 oracle after {STEPS} steps: |drho|max {drho:.2e}, "
          f"|du|max/u0 {du:.2e}; energy history max rel diff {de:.2e}")
    print(f"method {method}: E/E_analytic - 1 every 100 steps: "
          + " ".join(f"{x:+.4f}" for x in dev[9::10]))
    assert drho <= 1e-4 and du <= 1e-4
    assert de <= 5e-5
    # early times are linear: within 1% up to t = 100
    assert np.all(np.abs(dev[:10]) <= 1e-2)
    # the FP64 oracle shows the same departure from the linear decay
    assert np.all(np.abs(e64 / ea - 1 - dev) <= 1e-4)


@pytest.mark.parametrize("method", [0, 1], ids=["bgk", "home"])
def test_tgv64_low_amplitude_follows_analytic_decay(fs, method):
    u0 = 0.005
    e_gpu, _ = gpu_history(fs, method, u0)
    dev = e_gpu / wl.tgv_energy_analytic(times(method), u0, TAU, N) - 1
    print(f"\nmethod {method}, u0 {u0}: E/E_analytic - 1 every 100 steps: "
          + " ".join(f"{x:+.4f}" for x in dev[9::10]))
    assert np.all(np.abs(dev) <= 1.5e-2)
