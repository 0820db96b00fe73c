"""GPU: fluidsim_sweep, the comparer's headless Hybrid(B0) band sweep
(main_compare.cpp:505-699) on the batched engine.  Every CSV column but the
timings must equal the same configurations replayed one by one on the
oracle (voxelizer, band, slab density, Legacy / HOME / Hybrid solvers) and
reduced here in float64; the header and the log lines are the reference's."""
import csv
import io
import math
import os
import subprocess

import numpy as np
import pytest

from myfluidsimulation_amd import workloads as wl

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = ("d0,legacy_cells,legacy_ratio,mean_u_legacy,mean_u_home,mean_u_hybrid,"
          "mean_absdiff_home,mean_absdiff_hybrid,mean_rho_legacy,mean_rho_home,mean_rho_hybrid,"
          "absdiff_rho_home,absdiff_rho_hybrid,ms_per_step_legacy,ms_per_step_home,"
          "ms_per_step_hybrid,amr_factor,amr_threshold,fluid_cells,near_wall_cells")


def near_wall(solid, nx, ny, nz):
    s = solid.reshape(nz, ny, nx).astype(bool)
    nb = np.zeros_like(s)
    nb[:, :, 1:] |= s[:, :, :-1]
    nb[:, :, :-1] |= s[:, :, 1:]
    nb[:, 1:, :] |= s[:, :-1, :]
    nb[:, :-1, :] |= s[:, 1:, :]
    nb[1:, :, :] |= s[:-1, :, :]
    nb[:-1, :, :] |= s[1:, :, :]
    return np.flatnonzero((nb & ~s).reshape(-1))


def test_sweep_matches_oracle_replay(oracle, tmp_path):
    nx, ny, nz, dmax, steps, tau, fx = 40, 24, 20, 3, 30, 0.58, 1e-5
    stl = str(tmp_path / "body.stl")
    wl.obstacle_stl(stl, nx, ny, nz, (14.0, 12.0, 10.0), 4.0, seed=99)
    exe = os.path.join(ROOT, "myfluidsimulation_amd", "bin", "fluidsim_sweep")
    r = subprocess.run([exe, "--nx", str(nx), "--ny", str(ny), "--nz", str(nz), "--stl", stl,
                        "--hybrid-sweep-max", str(dmax), "--hybrid-sweep-steps", str(steps),
                        "--init", "slab", "--axis", "y", "--amp", "0.002", "--width", "0.25",
                        "--force", repr(fx)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    at = out.index("d0,legacy_cells")
    log, table = out[:at].splitlines(), out[at:]
    assert table.splitlines()[0] == HEADER
    rows = list(csv.DictReader(io.StringIO(table)))
    assert [int(rw["d0"]) for rw in rows] == list(range(dmax + 1))

    tris, bmin, bmax = oracle.load_stl(stl)
    solid = oracle.voxelize(tris, bmin, bmax, nx, ny, nz, 0.5, (0.5, 0.5, 0.5))
    near = near_wall(solid, nx, ny, nz)
    fluid = solid == 0
    n_fluid = int(fluid.sum())
    assert f"[wall] near-wall fluid cells: {near.size}" in log
    rho0 = oracle.rho_slab(nx, ny, nz, 1, np.float32(0.002), np.float32(0.5), np.float32(0.25))

    def run(method, flags):
        o = oracle.Oracle(method, nx, ny, nz, np.float32(tau), (np.float32(fx), 0.0, 0.0))
        if method == 1:
            o.set_bounce(1)  # the reference's pairing table, as LBM3D_Home
        o.set_flags(flags)
        o.reinit(rho0)
        o.step(steps)
        rr, u, v, w = o.macros()
        sp = np.sqrt(u * u + v * v + w * w)[near]  # float32, as launch_speed
        return rr, sp

    rl, sl = run(0, solid)
    rh, sh = run(1, solid)
    cnt = max(near.size, 1)
    mean = lambda a: math.fsum(a.astype(np.float64)) / cnt  # noqa: E731
    mrho = lambda rr: math.fsum(rr[fluid].astype(np.float64)) / n_fluid  # noqa: E731
    for rw in rows:
        d0 = int(rw["d0"])
        band = (oracle.band(solid, nx, ny, nz, d0) if d0 > 0 else np.zeros_like(solid))
        ry, sy = run(2, solid | (band * 2).astype(np.uint8))
        assert int(rw["legacy_cells"]) == int(band.sum())
        assert float(rw["legacy_ratio"]) == pytest.approx(band.sum() / n_fluid, abs=1e-6)
        expect = {
            "mean_u_legacy": mean(sl), "mean_u_home": mean(sh), "mean_u_hybrid": mean(sy),
            "mean_absdiff_home": mean(np.abs(sh - sl)), "mean_absdiff_hybrid": mean(np.abs(sy - sl)),
            "mean_rho_legacy": mrho(rl), "mean_rho_home": mrho(rh), "mean_rho_hybrid": mrho(ry),
        }
        expect["absdiff_rho_home"] = abs(expect["mean_rho_home"] - expect["mean_rho_legacy"])
        expect["absdiff_rho_hybrid"] = abs(expect["mean_rho_hybrid"] - expect["mean_rho_legacy"])
        for k, val in expect.items():
            assert float(rw[k]) == pytest.approx(val, rel=2e-9, abs=1e-15), (d0, k)
        assert int(rw["fluid_cells"]) == n_fluid and int(rw["near_wall_cells"]) == near.size
        assert int(rw["amr_factor"]) == 1 and float(rw["amr_threshold"]) == 0.5
        for k in ("ms_per_step_legacy", "ms_per_step_home", "ms_per_step_hybrid"):
            assert float(rw[k]) > 0
    assert float(rows[0]["mean_u_legacy"]) > 0  # the forcing drives a flow
