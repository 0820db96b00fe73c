"""GPU: fluidsim_tgv (the Taylor-Green harness of main_taylor_green.cpp)
end to end.  Its CSV rows (energy, L2 and Linf error against the analytic
vortex for Legacy and HOME) must equal the same run replayed on the oracle
and sampled here in float64; the analytic column must be the closed form;
the [cfg ] / [step] lines must be the reference's; and the simulated energy
must follow the analytic decay (2D extruded and the --3d extension)."""
import csv
import math
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def exact(nx, ny, nz, t, u0, tau, k, three_d):
    """the harness's analytic fields (float64, libm sin/cos via math)"""
    cs2 = 1.0 / 3.0
    kx, ky = 2 * math.pi * k / nx, 2 * math.pi * k / ny
    kz = 2 * math.pi * k / nz if three_d else 0.0
    k2 = kx * kx + ky * ky + kz * kz
    nu = (tau - 0.5) * cs2
    dv, dr = math.exp(-nu * k2 * t), math.exp(-2 * nu * k2 * t)
    n = nx * ny * nz
    u, v, r = np.empty(n), np.empty(n), np.empty(n)
    i = 0
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                X, Y, Z = (x + 0.5) * kx, (y + 0.5) * ky, (z + 0.5) * kz
                if three_d:
                    u[i] = u0 * math.sin(X) * math.cos(Y) * math.cos(Z) * dv
                    v[i] = -u0 * math.cos(X) * math.sin(Y) * math.cos(Z) * dv
                    r[i] = 1.0 + (3.0 * u0 * u0 * dr / 16.0) * (math.cos(2 * X) + math.cos(2 * Y)) * (
                        math.cos(2 * Z) + 2.0)
                else:
                    u[i] = u0 * math.sin(X) * math.cos(Y) * dv
                    v[i] = -u0 * math.cos(X) * math.sin(Y) * dv
                    r[i] = 1.0 + (u0 * u0 * dr / (4.0 * cs2)) * (math.cos(2 * X) + math.cos(2 * Y))
                i += 1
    e = (0.125 if three_d else 0.25) * u0 * u0 * math.exp(-2 * nu * k2 * t)
    return r, u, v, e, nu, kx, ky, k2


def sample(o, ex_u, ex_v):
    _, u, v, w = (a.astype(np.float64) for a in o.macros())
    d2 = (u - ex_u) ** 2 + (v - ex_v) ** 2 + w ** 2
    n = u.size
    return (math.fsum(0.5 * (u * u + v * v + w * w)) / n, math.sqrt(math.fsum(d2) / n),
            math.sqrt(d2.max()))


@pytest.mark.parametrize("three_d", [False, True])
def test_tgv_harness_matches_oracle_replay(oracle, tmp_path, three_d):
    nx, ny, nz, steps, every, u0, tau = 32, 32, (32 if three_d else 4), 40, 10, 0.05, 0.6
    path = str(tmp_path / "tgv.csv")
    exe = os.path.join(ROOT, "myfluidsimulation_amd", "bin", "fluidsim_tgv")
    args = [exe, "--nx", str(nx), "--ny", str(ny), "--nz", str(nz), "--steps", str(steps),
            "--sample-every", str(every), "--u0", repr(u0), "--tau", repr(tau), "--csv", path]
    if three_d:
        args.append("--3d")
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rho0, ux0, uy0, _, nu, kx, ky, k2 = exact(nx, ny, nz, 0.0, u0, tau, 1.0, three_d)
    lines = r.stdout.splitlines()
    assert lines[0] == f"[cfg ] Nx={nx} Ny={ny} Nz={nz} steps={steps} sample={every}"
    assert lines[1] == (f"[cfg ] tau={tau:g} nu={nu:g} u0={u0:g} kx_wave={kx:g} "
                        f"ky_wave={ky:g} k^2={k2:g}")
    assert lines[2] == f"[cfg ] csv={path}"
    assert lines[-1] == f"[write] saved {steps // every + 1} samples to {path}"
    with open(path) as fh:
        rows = list(csv.DictReader(fh))
    assert [int(rw["step"]) for rw in rows] == list(range(0, steps + 1, every))

    f32 = lambda a: np.asarray(a, np.float32)  # noqa: E731
    solvers = {}
    for name, m in (("legacy", 0), ("home", 1)):
        o = oracle.Oracle(m, nx, ny, nz, np.float32(tau))
        o.set_flags(np.zeros(nx * ny * nz, np.uint8))
        o.reinit(f32(rho0), f32(ux0), f32(uy0), np.zeros(nx * ny * nz, np.float32))
        solvers[name] = o
    done = 0
    for rw in rows:
        step = int(rw["step"])
        for o in solvers.values():
            o.step(step - done)
        done = step
        _, eu, ev, e_exact, *_ = exact(nx, ny, nz, float(step), u0, tau, 1.0, three_d)
        assert float(rw["energy_exact"]) == pytest.approx(e_exact, rel=1e-8)
        for name, o in solvers.items():
            e, l2, linf = sample(o, eu, ev)
            assert float(rw[f"energy_{name}"]) == pytest.approx(e, rel=1e-7), (step, name)
            assert float(rw[f"l2_{name}"]) == pytest.approx(l2, rel=1e-6, abs=1e-15), (step, name)
            assert float(rw[f"linf_{name}"]) == pytest.approx(linf, rel=1e-6, abs=1e-15)
            # physics: the decay follows the analytic rate
            assert float(rw[f"energy_{name}"]) == pytest.approx(e_exact, rel=0.03)
        step_line = [ln for ln in lines if ln.startswith(f"[step {step:5d}]")]
        assert step_line and f"E_exact={e_exact:12g}" in step_line[0]
