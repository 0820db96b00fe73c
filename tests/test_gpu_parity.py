"""GPU parity: libfluidsim.so (HIP, through the C ABI) against the FP32 CPU
oracle on identical float32 inputs.

The per-cell arithmetic of the kernels is the reference rule's, in the same
float order and without contraction, so fields are expected to be equal bit
for bit; the acceptance bar stated for this path is a relative max-norm of
1e-5 on rho, u (and HOME's moments) after 1000 steps (TOL below).
"""
import numpy as np
import pytest

from myfluidsimulation_amd import workloads as wl

TOL = 1e-5  # relative max-norm, rho/u/moments vs the FP32 oracle

pytestmark = pytest.mark.gpu


def rel_maxnorm(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = max(np.max(np.abs(ref)), 1e-30)
    return float(np.max(np.abs(got - ref)) / scale)


def run_pair(fs, oracle, method, dims, tau, steps, flags=None, fields=None, inlet=None,
             force=(0.0, 0.0, 0.0), substeps_split=None, variant=None):
    nx, ny, nz = dims
    n = nx * ny * nz
    s = fs.Solver(method, nx, ny, nz, tau, force)
    if variant is not None:
        s.set_variant(variant)
    o = oracle.Oracle(method, nx, ny, nz, tau, force)
    if flags is not None:
        s.set_flags(flags)
        o.set_flags(flags)
    if inlet is not None:
        s.set_inlet_profile(inlet)
        o.set_inlet(inlet)
    if fields is None:
        fields = (np.ones(n, np.float32), None, None, None)
    s.reinit_from_macro(*fields)
    o.reinit(*fields)
    chunks = substeps_split or [steps]
    for k in chunks:
        s.step(k)
    o.step(steps)
    out = dict(g_macro=s.macros(), o_macro=o.macros(), g_state=s.state(), o_state=o.state())
    s.close()
    o.close()
    return out


def assert_exact(res, what=""):
    for k, (g, r) in enumerate(zip(res["g_macro"], res["o_macro"])):
        assert np.array_equal(g, r), f"{what} macro {k}: rel {rel_maxnorm(g, r):.3e}"
    assert np.array_equal(res["g_state"], res["o_state"]), (
        f"{what} state: rel {rel_maxnorm(res['g_state'], res['o_state']):.3e}")


@pytest.mark.parametrize("method", [0, 1, 2], ids=["bgk", "home", "hybrid"])
def test_tgv64_1000_steps(fs, oracle, method):
    """Taylor-Green 64^3, tau 0.6, u0 0.05, 1000 steps (the CPU-runnable case)."""
    n = 64
    fields = wl.taylor_green_3d(n, n, n, 0.05)
    res = run_pair(fs, oracle, method, (n, n, n), 0.6, 1000, fields=fields)
    errs = [rel_maxnorm(g, r) for g, r in zip(res["g_macro"][:3], res["o_macro"][:3])]
    print(f"method {method}: rel max-norm rho/ux/uy {errs}")
    assert max(errs) <= TOL
    if method == 1:
        assert rel_maxnorm(res["g_state"], res["o_state"]) <= TOL
    assert_exact(res, f"tgv method {method}")


@pytest.mark.parametrize("method", [0, 1, 2], ids=["bgk", "home", "hybrid"])
@pytest.mark.parametrize("steps", [1, 2, 7, 40])
def test_obstacle_box_exact(fs, oracle, method, steps):
    """Periodic box with a solid block and a solid column touching the
    periodic seam: bounce-back, solid macros, both AA parities."""
    nx, ny, nz = 32, 24, 16
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[5:11, 8:15, 9:17] = 1
    solid[:, 0, 0] = 1
    solid[3, :, 31] = 1
    rng = np.random.default_rng(7)
    n = nx * ny * nz
    rho = (1.0 + 0.01 * rng.standard_normal(n)).astype(np.float32)
    ux = (0.02 * rng.standard_normal(n)).astype(np.float32)
    uy = (0.02 * rng.standard_normal(n)).astype(np.float32)
    uz = (0.02 * rng.standard_normal(n)).astype(np.float32)
    res = run_pair(fs, oracle, method, (nx, ny, nz), 0.7, steps, flags=solid.reshape(-1),
                   fields=(rho, ux, uy, uz), force=(1e-5, -2e-6, 0.0))
    assert_exact(res, f"box method {method} steps {steps}")


def test_step_split_invariance(fs, oracle):
    """step(3)+step(4)+step(5) equals step(12): macros are written on the
    last step of each call and the AA parity carries across calls."""
    nx, ny, nz = 24, 16, 8
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[2:5, 4:9, 6:12] = 1
    fields = wl.taylor_green_3d(nx, ny, nz, 0.04)
    for method in (0, 2):
        res = run_pair(fs, oracle, method, (nx, ny, nz), 0.8, 12, flags=solid.reshape(-1),
                       fields=fields, substeps_split=[3, 4, 5])
        assert_exact(res, f"split method {method}")


@pytest.mark.parametrize("method", [0, 1, 2], ids=["bgk", "home", "hybrid"])
def test_channel_inlet_outlet_band(fs, oracle, method):
    """Obstacle channel: y/z walls, equilibrium inlet, zero-gradient outlet,
    a box obstacle, the Hybrid band (d0 = 2) from the CPU band rule."""
    nx, ny, nz = 48, 20, 18
    obst = np.zeros((nz, ny, nx), np.uint8)
    obst[6:12, 7:13, 12:18] = 1
    flags = wl.channel_flags(nx, ny, nz, obst.reshape(-1))
    solid = (flags & 1).astype(np.uint8)
    band = oracle.band(solid, nx, ny, nz, 2)
    flags = flags | (band * 2).astype(np.uint8)
    inlet = np.full(ny * nz, 0.05, np.float32)
    fields = wl.uniform_flow(nx * ny * nz, 0.05)
    for steps in (1, 2, 301):
        res = run_pair(fs, oracle, method, (nx, ny, nz), 0.56, steps, flags=flags,
                       fields=fields, inlet=inlet)
        assert_exact(res, f"channel method {method} steps {steps}")


def test_hybrid_equals_legacy_when_rates_equal(fs):
    """omega_s = omega_h (lbm3d_hybrid.cu:264-267) makes every Hybrid cell
    a BGK cell: with u = 0 inits the two solvers agree bit for bit."""
    nx, ny, nz = 32, 16, 16
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[4:10, 5:11, 8:16] = 1
    out = []
    for method in (0, 2):
        with fs.Solver(method, nx, ny, nz, 0.6, (1e-5, 0, 0)) as s:
            s.set_solid_mask(solid.reshape(-1))
            if method == 2:
                s.set_legacy_mapping((solid.reshape(-1) == 0).astype(np.uint8))
            s.reinit_from_macro(np.ones(nx * ny * nz, np.float32))
            s.step(50)
            out.append(s.macros())
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_hybrid_two_rate_form(fs, oracle):
    """A shear rate different from the hydro rate exercises the two-rate
    bulk form; the band cells keep BGK."""
    nx, ny, nz = 24, 16, 12
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[3:7, 5:10, 6:12] = 1
    band = oracle.band(solid.reshape(-1), nx, ny, nz, 1)
    flags = solid.reshape(-1) | (band * 2)
    fields = wl.taylor_green_3d(nx, ny, nz, 0.03)
    s = fs.Solver(2, nx, ny, nz, 0.6)
    o = oracle.Oracle(2, nx, ny, nz, 0.6)
    for x in (s, o):
        x.set_flags(flags)
    s.set_omegas(1 / 0.6, 1 / 0.6, 1.4)
    o.set_omegas(float(np.float32(1 / 0.6)), float(np.float32(1 / 0.6)), float(np.float32(1.4)))
    s.reinit_from_macro(*fields)
    o.reinit(*fields)
    s.step(25)
    o.step(25)
    for g, r in zip(s.macros(), o.macros()):
        assert np.array_equal(g, r)


def test_reset_and_null_reinit(fs, oracle):
    nx, ny, nz = 16, 12, 10
    for method in (0, 1, 2):
        with fs.Solver(method, nx, ny, nz, 0.7) as s:
            s.reinit_from_macro(None)  # Legacy: reset(); HOME/Hybrid: rho = 1
            s.step(3)
            o = oracle.Oracle(method, nx, ny, nz, 0.7)
            o.reinit(None)
            o.step(3)
            for g, r in zip(s.macros(), o.macros()):
                assert np.array_equal(g, r)
            s.reset()
            np.testing.assert_array_equal(s.macro(0), np.ones(nx * ny * nz, np.float32))


def test_load_state_roundtrip(fs, oracle):
    nx, ny, nz = 20, 12, 8
    rng = np.random.default_rng(3)
    for method in (0, 1):
        with fs.Solver(method, nx, ny, nz, 0.65) as s:
            fields = wl.taylor_green_3d(nx, ny, nz, 0.05)
            s.reinit_from_macro(*fields)
            s.step(5)
            st = s.state()
            s.load_state(st)
            np.testing.assert_array_equal(s.state(), st)
            # an arbitrary state steps like the oracle from the same state
            pert = (st * (1 + 1e-3 * rng.standard_normal(st.size))).astype(np.float32)
            s.load_state(pert)
            o = oracle.Oracle(method, nx, ny, nz, 0.65)
            o.set_state(pert)
            s.step(3)
            o.step(3)
            np.testing.assert_array_equal(s.state(), o.state())


@pytest.mark.parametrize("dims", [(65, 12, 10), (130, 9, 17), (64, 8, 16)])
def test_home_tile_shapes(fs, oracle, dims):
    """HOME's tiled kernels on partial tiles and z-chunks, and the plain
    gather the solver keeps when nx-1 starts a 64-wide tile (nx = 65)."""
    nx, ny, nz = dims
    obst = np.zeros((nz, ny, nx), np.uint8)
    obst[2:5, 3:6, 10:14] = 1
    flags = wl.channel_flags(nx, ny, nz, obst.reshape(-1))
    inlet = np.full(ny * nz, 0.04, np.float32)
    fields = wl.uniform_flow(nx * ny * nz, 0.04)
    tiled = (nx - 1) % 64 != 0
    for variant in (tuple(range(1, 17)) if tiled else (0,)):
        res = run_pair(fs, oracle, 1, (nx, ny, nz), 0.62, 9, flags=flags, fields=fields,
                       inlet=inlet, force=(1e-6, 0.0, 0.0), variant=variant)
        assert_exact(res, f"home tile {dims} variant {variant}")
    if not tiled:
        with fs.Solver(1, nx, ny, nz, 0.62) as s, pytest.raises(fs.capi.FluidSimError):
            s.set_variant(1)


def test_home_variants_and_autotune(fs, oracle):
    """Every HOME kernel variant gives the oracle's bits; autotune leaves the
    state untouched and picks one of them."""
    nx, ny, nz = 130, 20, 40
    obst = np.zeros((nz, ny, nx), np.uint8)
    obst[10:20, 6:12, 30:40] = 1
    flags = wl.channel_flags(nx, ny, nz, obst.reshape(-1))
    inlet = np.full(ny * nz, 0.05, np.float32)
    fields = wl.uniform_flow(nx * ny * nz, 0.05)
    o = oracle.Oracle(1, nx, ny, nz, 0.58)
    o.set_flags(flags)
    o.set_inlet(inlet)
    o.reinit(*fields)
    o.step(6)
    ref = o.state()
    for variant in tuple(range(17)) + (-1,):
        with fs.Solver(1, nx, ny, nz, 0.58) as s:
            s.set_flags(flags)
            s.set_inlet_profile(inlet)
            s.set_variant(variant)
            s.reinit_from_macro(*fields)
            if variant == -1:
                before = s.state()
                chosen = s.autotune(3)
                assert chosen in range(17) and s.variant == chosen
                np.testing.assert_array_equal(s.state(), before)
            s.step(6)
            np.testing.assert_array_equal(s.state(), ref)


@pytest.mark.parametrize("variant", [12, 15], ids=["64x11", "64x12"])
def test_home_pair_outlet_rules(fs, oracle, variant):
    """The two-step HOME tile (variants 12 and 15) where an outlet sits on a tile's
    left ring column: a channel whose outlet column x = nx-1 (nx % 64 == 0)
    wraps onto the inlet keeps the two-step kernel; an outlet at x = 63 whose
    downstream cells at x = 64 are fluid sends the solver to a one-step
    kernel.  Odd and even step counts, both bounce rules: the oracle's bits."""
    nx, ny, nz = 128, 14, 10
    obst = np.zeros((nz, ny, nx), np.uint8)
    obst[3:6, 4:8, 30:36] = 1
    flags = wl.channel_flags(nx, ny, nz, obst.reshape(-1))
    assert (flags.reshape(nz, ny, nx)[1:-1, 1:-1, nx - 1] & 8).all()
    inlet = np.linspace(0.03, 0.05, ny * nz).astype(np.float32)
    fields = wl.uniform_flow(nx * ny * nz, 0.04)
    mid = flags.copy().reshape(nz, ny, nx)
    mid[4:7, 5:9, 63] = 8  # outlets inside the channel, fluid at x = 62 and 64
    for fl, pair in ((flags, True), (mid.reshape(-1), False)):
        for rule in (0, 1):
            with fs.Solver(1, nx, ny, nz, 0.6) as s:
                s.set_flags(fl)
                s.set_inlet_profile(inlet)
                s.set_bounce_rule(rule)
                s.set_variant(variant)
                assert (s.variant == variant) == pair
                s.reinit_from_macro(*fields)
                o = oracle.Oracle(1, nx, ny, nz, 0.6)
                o.set_flags(fl)
                o.set_inlet(inlet)
                o.set_bounce(rule)
                o.reinit(*fields)
                for k in (1, 4, 7):
                    s.step(k)
                    o.step(k)
                    what = f"pair {pair} rule {rule} chunk {k}"
                    np.testing.assert_array_equal(s.state(), o.state(), err_msg=what)
                o.close()


def pop_case(nx, ny, nz, seed=5):
    """walls, inlet, outlet, an obstacle block and a Hybrid band next to it"""
    obst = np.zeros((nz, ny, nx), np.uint8)
    obst[nz // 3:nz // 3 + 3, ny // 3:ny // 3 + 3, nx // 4:nx // 4 + 4] = 1
    flags = wl.channel_flags(nx, ny, nz, obst.reshape(-1))
    band = np.zeros((nz, ny, nx), np.uint8)
    band[:, :, nx // 4 - 2:nx // 4 + 6] = 2
    flags = flags | (band.reshape(-1) & ~((flags & 1) * 2).astype(np.uint8))
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    fields = (np.float32(1) + np.float32(0.01) * rng.standard_normal(n, dtype=np.float32),
              np.float32(0.04) + np.float32(0.01) * rng.standard_normal(n, dtype=np.float32),
              np.float32(0.01) * rng.standard_normal(n, dtype=np.float32),
              np.float32(0.01) * rng.standard_normal(n, dtype=np.float32))
    inlet = np.linspace(0.02, 0.05, ny * nz).astype(np.float32)
    return flags, fields, inlet


@pytest.mark.parametrize("method", [0, 2], ids=["bgk", "hybrid"])
def test_pop_variants_and_autotune(fs, oracle, method):
    """Every BGK / Hybrid kernel variant (the one-step AA kernel and the
    two-step tiles) gives the oracle's bits through step chunks of odd and even
    length, so runs switch between the two kernels and both representations;
    autotune leaves the state untouched and picks one of them."""
    nx, ny, nz = 130, 20, 40
    flags, fields, inlet = pop_case(nx, ny, nz)
    chunks = (1, 4, 3, 2, 5)
    o = oracle.Oracle(method, nx, ny, nz, 0.58, (2e-6, 0.0, 0.0))
    o.set_flags(flags)
    o.set_inlet(inlet)
    if method == 2:
        o.set_omegas(1.0 / 0.58, 1.3, 1.45)
    o.reinit(*fields)
    refs = []
    for k in chunks:
        o.step(k)
        refs.append((o.state(), o.macros()))
    for variant in (0, 1, 2, 3, 4, 5, 6, 7, 8, -1):
        with fs.Solver(method, nx, ny, nz, 0.58, (2e-6, 0.0, 0.0)) as s:
            s.set_flags(flags)
            s.set_inlet_profile(inlet)
            if method == 2:
                s.set_omegas(1.0 / 0.58, 1.3, 1.45)
            s.set_variant(variant)
            s.reinit_from_macro(*fields)
            if variant == -1:
                before, mac_before = s.state(), s.macros()
                chosen = s.autotune(3)
                assert chosen in range(9) and s.variant == chosen
                np.testing.assert_array_equal(s.state(), before)
                for g, r in zip(s.macros(), mac_before):  # the macro planes too
                    np.testing.assert_array_equal(g, r)
            else:
                assert s.variant == variant
            for k, (st, mac) in zip(chunks, refs):
                s.step(k)
                np.testing.assert_array_equal(s.state(), st, err_msg=f"variant {variant}")
                for g, r in zip(s.macros(), mac):
                    np.testing.assert_array_equal(g, r, err_msg=f"variant {variant}")


@pytest.mark.parametrize("dims", [(130, 9, 17), (64, 8, 16), (70, 5, 1), (3, 2, 2), (200, 3, 33)])
def test_pop_pair_tile_shapes(fs, oracle, dims):
    """The two-step kernel on partial tiles and z-chunks, one-plane and tiny
    grids (the ring cells wrap onto the tile itself)."""
    nx, ny, nz = dims
    n = nx * ny * nz
    rng = np.random.default_rng(11)
    fields = (np.float32(1) + np.float32(0.01) * rng.standard_normal(n, dtype=np.float32),
              np.float32(0.02) * rng.standard_normal(n, dtype=np.float32),
              np.float32(0.02) * rng.standard_normal(n, dtype=np.float32),
              np.float32(0.02) * rng.standard_normal(n, dtype=np.float32))
    flags = None
    if nx >= 8 and ny >= 3:
        f = np.zeros((nz, ny, nx), np.uint8)
        f[:, 0, :] = 1
        f[:, ny - 1, nx // 2:nx // 2 + 3] = 1
        flags = f.reshape(-1)
    for variant in (1, 2, 3, 4, 5, 6, 7, 8):
        nxyz = (nx, ny, nz)
        o = oracle.Oracle(0, *nxyz, 0.6, (1e-5, 0.0, 0.0))
        s = fs.Solver(0, *nxyz, 0.6, (1e-5, 0.0, 0.0))
        if flags is not None:
            o.set_flags(flags)
            s.set_flags(flags)
        o.reinit(*fields)
        s.reinit_from_macro(*fields)
        s.set_variant(variant)
        s.step(6)
        o.step(6)
        np.testing.assert_array_equal(s.state(), o.state(), err_msg=f"{dims} variant {variant}")
        for g, r in zip(s.macros(), o.macros()):
            np.testing.assert_array_equal(g, r)
        s.close()
        o.close()


@pytest.mark.parametrize("method", [0, 1, 2], ids=["bgk", "home", "hybrid"])
def test_outlet_at_tile_row_start(fs, oracle, method):
    """An outlet column at x = 64 (upstream x = 63 in the previous 64-wide
    row): the tiled kernels take an outlet's result from the lane at x-1, so
    the solver runs its one-step kernels and still gives the oracle's bits."""
    nx, ny, nz = 100, 6, 5
    f = np.zeros((nz, ny, nx), np.uint8)
    f[:, :, 0] = 4
    f[:, :, 64] = 8
    f[:, 0, :] = 1
    f[:, 0, 0] = 1
    f[:, 0, 64] = 1
    inlet = np.full(ny * nz, 0.03, np.float32)
    fields = wl.uniform_flow(nx * ny * nz, 0.03)
    with fs.Solver(method, nx, ny, nz, 0.6) as s:
        s.set_flags(f.reshape(-1))
        assert s.variant == 0
        s.set_variant(-1)
        assert s.variant == 0
    res = run_pair(fs, oracle, method, (nx, ny, nz), 0.6, 8, flags=f.reshape(-1), fields=fields,
                   inlet=inlet)
    assert_exact(res, "outlet at x=64")


@pytest.mark.parametrize("method", [0, 2], ids=["bgk", "hybrid"])
def test_outlet_at_narrow_tile_row_start(fs, oracle, method):
    """An outlet column at x = 96 starts a 32-wide row but not a 64-wide one:
    the 32-wide two-step variant falls back to the one-step kernel, the
    64-wide ones run; every choice gives the oracle's bits."""
    nx, ny, nz = 130, 6, 5
    f = np.zeros((nz, ny, nx), np.uint8)
    f[:, :, 0] = 4
    f[:, :, 96] = 8
    f[:, 0, :] = 1
    flags = f.reshape(-1)
    inlet = np.full(ny * nz, 0.03, np.float32)
    fields = wl.uniform_flow(nx * ny * nz, 0.03)
    for variant, effective in ((5, 0), (1, 1), (2, 2)):
        with fs.Solver(method, nx, ny, nz, 0.6) as s:
            s.set_flags(flags)
            s.set_variant(variant)
            assert s.variant == effective
        res = run_pair(fs, oracle, method, (nx, ny, nz), 0.6, 8, flags=flags, fields=fields,
                       inlet=inlet, variant=variant)
        assert_exact(res, f"outlet at x=96, variant {variant}")


@pytest.mark.parametrize("where", ["feeder", "outlet"])
def test_hybrid_band_next_to_outlet(fs, oracle, where):
    """An outlet collides as its upstream cell, band bit included: the band
    marked on the feeder column only, or on the outlet column only, gives the
    oracle's bits with the one-step kernel and both two-step tile widths."""
    nx, ny, nz = 100, 8, 6
    flags = wl.channel_flags(nx, ny, nz).reshape(nz, ny, nx)
    col = nx - 2 if where == "feeder" else nx - 1
    flags[:, :, col] |= np.where((flags[:, :, col] & 1) == 0, 2, 0).astype(np.uint8)
    flags = flags.reshape(-1)
    inlet = np.full(ny * nz, 0.04, np.float32)
    fields = wl.uniform_flow(nx * ny * nz, 0.04)
    o = oracle.Oracle(2, nx, ny, nz, 0.6, (1e-6, 0.0, 0.0))
    o.set_flags(flags)
    o.set_inlet(inlet)
    o.set_omegas(1.0 / 0.6, 1.25, 1.5)
    o.reinit(*fields)
    o.step(6)
    for variant in (0, 1, 5):
        with fs.Solver(2, nx, ny, nz, 0.6, (1e-6, 0.0, 0.0)) as s:
            s.set_flags(flags)
            s.set_inlet_profile(inlet)
            s.set_omegas(1.0 / 0.6, 1.25, 1.5)
            s.set_variant(variant)
            s.reinit_from_macro(*fields)
            s.step(6)
            np.testing.assert_array_equal(s.state(), o.state(), err_msg=f"variant {variant}")
            for g, r in zip(s.macros(), o.macros()):
                np.testing.assert_array_equal(g, r, err_msg=f"variant {variant}")
    o.close()


def random_pop_case(seed):
    """A small random domain: solid blobs, an optional inlet/outlet channel,
    Hybrid band cells, forcing and perturbed fields (seeded)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = int(rng.integers(6, 90)), int(rng.integers(3, 20)), int(rng.integers(1, 18))
    solid = (rng.random((nz, ny, nx)) < 0.08).astype(np.uint8)
    flags = solid.copy()
    if rng.random() < 0.6 and nx >= 4:
        flags[:, :, 0] = np.where(solid[:, :, 0] == 0, 4, 1)
        ok = (solid[:, :, nx - 1] == 0) & (flags[:, :, nx - 2] == 0)
        flags[:, :, nx - 1] = np.where(ok, 8, flags[:, :, nx - 1])
    band = (rng.random((nz, ny, nx)) < 0.3) & ((flags & 1) == 0)
    flags = (flags | (band * 2)).astype(np.uint8).reshape(-1)
    n = nx * ny * nz
    fields = (np.float32(1) + np.float32(0.01) * rng.standard_normal(n, dtype=np.float32),
              np.float32(0.03) * rng.standard_normal(n, dtype=np.float32),
              np.float32(0.03) * rng.standard_normal(n, dtype=np.float32),
              np.float32(0.03) * rng.standard_normal(n, dtype=np.float32))
    inlet = (np.float32(0.02) + np.float32(0.02) * rng.random(ny * nz, dtype=np.float32))
    method = int(rng.choice([0, 2]))
    variant = seed % 6
    chunks = [int(k) for k in rng.integers(1, 6, size=3)]
    force = (float(rng.normal(0, 1e-5)), float(rng.normal(0, 1e-5)), 0.0)
    return (nx, ny, nz), flags, fields, inlet, method, variant, chunks, force


@pytest.mark.parametrize("seed", range(16))
def test_pop_random_geometries(fs, oracle, seed):
    """Seeded random small domains through every BGK / Hybrid kernel variant
    and step chunking: the oracle's bits."""
    dims, flags, fields, inlet, method, variant, chunks, force = random_pop_case(seed)
    o = oracle.Oracle(method, *dims, 0.57, force)
    o.set_flags(flags)
    o.set_inlet(inlet)
    if method == 2:
        o.set_omegas(1.0 / 0.57, 1.3, 1.6)
    o.reinit(*fields)
    with fs.Solver(method, *dims, 0.57, force) as s:
        s.set_flags(flags)
        s.set_inlet_profile(inlet)
        if method == 2:
            s.set_omegas(1.0 / 0.57, 1.3, 1.6)
        s.set_variant(variant)
        s.reinit_from_macro(*fields)
        for k in chunks:
            s.step(k)
            o.step(k)
            what = f"seed {seed} dims {dims} method {method} variant {variant} chunk {k}"
            np.testing.assert_array_equal(s.state(), o.state(), err_msg=what)
            for g, r in zip(s.macros(), o.macros()):
                np.testing.assert_array_equal(g, r, err_msg=what)
    o.close()


@pytest.mark.parametrize("seed", range(30))
def test_home_random_geometries(fs, oracle, seed):
    """Seeded random small domains through every HOME kernel variant, each
    with both bounce-back rules: the oracle's bits."""
    dims, flags, fields, inlet, _, _, chunks, force = random_pop_case(100 + seed)
    flags = flags & np.uint8(0xFD)  # HOME has no band
    variant, rule = seed % 15, (seed // 15) % 2
    o = oracle.Oracle(1, *dims, 0.61, force)
    o.set_flags(flags)
    o.set_inlet(inlet)
    o.set_bounce(rule)
    o.reinit(*fields)
    with fs.Solver(1, *dims, 0.61, force) as s:
        s.set_flags(flags)
        s.set_inlet_profile(inlet)
        s.set_bounce_rule(rule)
        nx = dims[0]
        if variant == 0 or (nx - 1) % 64 != 0:
            s.set_variant(variant)
        s.reinit_from_macro(*fields)
        for k in chunks:
            s.step(k)
            o.step(k)
            what = f"seed {seed} dims {dims} variant {variant} rule {rule} chunk {k}"
            np.testing.assert_array_equal(s.state(), o.state(), err_msg=what)
            for g, r in zip(s.macros(), o.macros()):
                np.testing.assert_array_equal(g, r, err_msg=what)
    o.close()
