"""GPU: the benchmark configurations at their full sizes.

* 384x128x128 obstacle channel (seeded 1280-face icosphere R=20, inlet,
  outlet, walls; Hybrid with band 2): 500 steps of every method equal the
  FP32 oracle bit for bit (fields and full state).
* 512^3 periodic Taylor-Green vortex (BGK and HOME, 200 steps), properties
  that do not need a CPU run at this size: total mass is conserved, the
  kinetic energy follows the analytic decay (u0^2/8) exp(-6 nu k^2 t), and
  the mirror symmetries of the vortex (u odd and v even under x -> -x, the
  reverse under y -> -y, both even under z -> -z; the w the nonlinear term
  generates odd under z -> -z, even under x and y) hold; after 4 steps both
  equal the FP32 oracle bit for bit (fields and full state).
* The 768x256x256 channel's icosphere (R=40): the voxelized flag grid equals
  the oracle voxelizer's and its band (d0 = 1, 2, 4) the oracle band; BGK,
  HOME and Hybrid with bands 1, 2, 4 on that channel equal the FP32 oracle bit
  for bit after 11 steps (fields and full state).
"""
import os

import numpy as np
import pytest

from myfluidsimulation_amd import workloads as wl

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def channel(fs, tmp, nx, ny, nz, center, radius, seed=2024):
    path = os.path.join(tmp, "obstacle.stl")
    scale, trans = wl.obstacle_stl(path, nx, ny, nz, center, radius, seed)
    tris, bmin, bmax = fs.load_stl(path)
    return tris, bmin, bmax, scale, trans


@pytest.mark.parametrize("method", [0, 1, 2], ids=["bgk", "home", "hybrid"])
def test_channel384_500_steps_exact(fs, oracle, tmp_path, method):
    nx, ny, nz, tau, u_in, steps = 384, 128, 128, 0.56, 0.05, 500
    tris, bmin, bmax, scale, trans = channel(fs, str(tmp_path), nx, ny, nz, (96, 64, 64), 20.0)
    solid = fs.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
    flags = wl.channel_flags(nx, ny, nz, solid)
    if method == 2:
        flags = flags | (fs.band((flags & 1).astype(np.uint8), nx, ny, nz, 2) * 2).astype(np.uint8)
    n = nx * ny * nz
    fields = wl.uniform_flow(n, u_in)
    inlet = np.full(ny * nz, u_in, np.float32)
    s = fs.Solver(method, nx, ny, nz, tau)
    s.set_flags(flags)
    s.set_inlet_profile(inlet)
    s.reinit_from_macro(*fields)
    s.step(steps)
    o = oracle.Oracle(method, nx, ny, nz, tau)
    o.set_flags(flags)
    o.set_inlet(inlet)
    o.reinit(*fields)
    o.step(steps)
    for k in range(4):
        assert np.array_equal(s.macro(k), o.macro(k)), f"macro {k}"
    assert np.array_equal(s.state(), o.state())
    ux = s.macro(1).reshape(nz, ny, nx)
    assert ux[nz // 2, ny // 2, nx - 10] > 0.02  # the flow passes the obstacle
    s.close()
    o.close()


@pytest.mark.parametrize("method", [0, 1], ids=["bgk", "home"])
def test_tgv512_properties(fs, method):
    n, tau, u0, steps = 512, 0.6, 0.05, 200
    rho, ux, uy, uz = wl.taylor_green_3d(n, n, n, u0)
    mass0 = float(np.sum(rho, dtype=np.float64))
    s = fs.Solver(method, n, n, n, tau)
    s.reinit_from_macro(rho, ux, uy, uz)
    del rho, ux, uy, uz
    s.step(steps)
    r, u, v, w = (s.macro(k).reshape(n, n, n) for k in range(4))
    s.close()
    # BGK reports the moments entering its last step
    t = steps if method == 1 else steps - 1
    # float32 rounding of the collision (the reference rule, equal bit for bit
    # to the FP32 oracle on smaller grids) drifts the total by ~5e-6 relative
    # over 200 steps here; a lost or doubled population would move it by
    # orders of magnitude more
    assert abs(float(np.sum(r, dtype=np.float64)) - mass0) / mass0 < 5e-5
    e = 0.5 * float(np.mean(u.astype(np.float64) ** 2 + v.astype(np.float64) ** 2 +
                            w.astype(np.float64) ** 2))
    e_exact = float(wl.tgv_energy_analytic(t, u0, tau, n))
    assert e == pytest.approx(e_exact, rel=2e-3)
    # cell-centred mirrors: index i -> n-1-i
    fx = lambda a: a[:, :, ::-1]  # noqa: E731
    fy = lambda a: a[:, ::-1, :]  # noqa: E731
    fz = lambda a: a[::-1, :, :]  # noqa: E731
    # (float32 rounding differs between mirrored directions, so the mirror
    # images agree to rounding, not bit for bit)
    tol = 2e-5 * u0
    assert np.abs(u + fx(u)).max() < tol and np.abs(v - fx(v)).max() < tol
    assert np.abs(u - fy(u)).max() < tol and np.abs(v + fy(v)).max() < tol
    assert np.abs(u - fz(u)).max() < tol and np.abs(v - fz(v)).max() < tol
    # w, which the nonlinear term generates, is odd under z -> -z, even under x, y
    assert np.abs(w + fz(w)).max() < tol and np.abs(w - fx(w)).max() < tol
    assert np.abs(w - fy(w)).max() < tol and float(np.abs(w).max()) > 1e-5
    assert np.abs(r - fx(r)).max() < 5e-6 and np.abs(r - fz(r)).max() < 5e-6


@pytest.mark.parametrize("method", [0, 1], ids=["bgk", "home"])
def test_tgv512_exact_short(fs, oracle, method):
    """The 512^3 vortex's default kernels (two steps per launch: BGK 64x12x64
    tiles, HOME the 64x12x64 tile with its folded ring rows) equal the FP32
    oracle bit for bit after 4 steps (fields and full state)."""
    n, tau, u0, steps = 512, 0.6, 0.05, 4
    print(f"tgv512 exact method {method}", flush=True)
    fields = wl.taylor_green_3d(n, n, n, u0)
    s = fs.Solver(method, n, n, n, tau)
    assert s.variant == (8 if method == 0 else 15)
    s.reinit_from_macro(*fields)
    s.step(steps)
    o = oracle.Oracle(method, n, n, n, tau)
    o.reinit(*fields)
    del fields
    o.step(steps)
    for k in range(4):
        assert np.array_equal(s.macro(k), o.macro(k)), f"macro {k}"
    assert np.array_equal(s.state(), o.state())
    s.close()
    o.close()


def test_channel768_geometry_exact(fs, oracle, tmp_path):
    nx, ny, nz = 768, 256, 256
    tris, bmin, bmax, scale, trans = channel(fs, str(tmp_path), nx, ny, nz, (192, 128, 128), 40.0)
    g = fs.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
    o = oracle.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
    assert np.array_equal(g, o)
    vol = 4.0 / 3.0 * np.pi * 40.0 ** 3
    assert 0.7 * vol < g.sum() < 1.4 * vol
    flags = wl.channel_flags(nx, ny, nz, g)
    solid = (flags & 1).astype(np.uint8)
    for d0 in (1, 2, 4):
        assert np.array_equal(fs.band(solid, nx, ny, nz, d0), oracle.band(solid, nx, ny, nz, d0))


@pytest.fixture(scope="module")
def channel768_flags(fs, tmp_path_factory):
    nx, ny, nz = 768, 256, 256
    tris, bmin, bmax, scale, trans = channel(fs, str(tmp_path_factory.mktemp("c768")), nx, ny, nz,
                                             (192, 128, 128), 40.0)
    return wl.channel_flags(nx, ny, nz, fs.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans))


@pytest.mark.parametrize("method,band", [(0, 0), (1, 0), (2, 1), (2, 2), (2, 4)],
                         ids=["bgk", "home", "hybrid_band1", "hybrid_band2", "hybrid_band4"])
def test_channel768_solvers_exact(fs, oracle, channel768_flags, method, band):
    """The 768x256x256 roofline channel (tau 0.56, inlet u 0.05, outlet,
    walls, icosphere R=40): 11 steps (five two-step launches and one
    one-step launch for BGK / Hybrid) equal the FP32 oracle bit for bit,
    fields and full state."""
    nx, ny, nz, tau, u_in, steps = 768, 256, 256, 0.56, 0.05, 11
    print(f"channel768 method {method} band {band}", flush=True)
    flags = channel768_flags
    if band:
        flags = flags | (fs.band((flags & 1).astype(np.uint8), nx, ny, nz, band) * 2).astype(
            np.uint8)
    fields = wl.uniform_flow(nx * ny * nz, u_in)
    inlet = np.full(ny * nz, u_in, np.float32)
    s = fs.Solver(method, nx, ny, nz, tau)
    s.set_flags(flags)
    s.set_inlet_profile(inlet)
    s.reinit_from_macro(*fields)
    s.step(steps)
    o = oracle.Oracle(method, nx, ny, nz, tau)
    o.set_flags(flags)
    o.set_inlet(inlet)
    o.reinit(*fields)
    o.step(steps)
    for k in range(4):
        assert np.array_equal(s.macro(k), o.macro(k)), f"macro {k}"
    assert np.array_equal(s.state(), o.state())
    s.close()
    o.close()
