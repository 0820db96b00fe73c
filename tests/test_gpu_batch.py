"""GPU: the batched engine against single-domain solvers (which are checked
bit for bit against the oracle elsewhere) and directly against the oracle."""
import numpy as np
import pytest

from myfluidsimulation_amd import workloads as wl

pytestmark = pytest.mark.gpu


def small_domains():
    """Mixed methods, uneven sizes, periodic and channel domains."""
    out = []
    rng = np.random.default_rng(11)
    shapes = [(32, 48, 16), (48, 32, 32), (80, 16, 18), (64, 20, 24), (34, 18, 40), (96, 32, 8),
              (33, 17, 9)]
    for i, (nx, ny, nz) in enumerate(shapes):
        method = i % 3
        tau = 0.55 + 0.3 * rng.random()
        n = nx * ny * nz
        if i % 2 == 0:
            fields = wl.taylor_green_3d(nx, ny, nz, 0.03 + 0.01 * i)
            solid = np.zeros((nz, ny, nx), np.uint8)
            solid[nz // 3: nz // 2, ny // 3: ny // 2, nx // 4: nx // 2] = 1
            flags = solid.reshape(-1)
            inlet = None
        else:
            obst = np.zeros((nz, ny, nx), np.uint8)
            obst[nz // 3: nz // 2 + 1, ny // 3: ny // 2 + 1, nx // 4: nx // 3 + 1] = 1
            flags = wl.channel_flags(nx, ny, nz, obst.reshape(-1))
            fields = wl.uniform_flow(n, 0.04)
            inlet = np.full(ny * nz, 0.04, np.float32)
        if method == 2:
            import pyoracle

            band = pyoracle.band((flags & 1).astype(np.uint8), nx, ny, nz, 1 + i % 3)
            flags = flags | (band * 2).astype(np.uint8)
        out.append(dict(method=method, dims=(nx, ny, nz), tau=tau, flags=flags, fields=fields,
                        inlet=inlet, force=(1e-6 * i, 0.0, 0.0)))
    return out


def make_batch(fs, doms):
    b = fs.Batch([(d["method"], *d["dims"], d["tau"], d["force"]) for d in doms])
    for i, d in enumerate(doms):
        b.set_flags(i, d["flags"])
        if d["inlet"] is not None:
            b.set_inlet_profile(i, d["inlet"])
        b.reinit_from_macro(i, *d["fields"])
    return b


def single(fs, d, steps):
    with fs.Solver(d["method"], *d["dims"], d["tau"], d["force"]) as s:
        s.set_flags(d["flags"])
        if d["inlet"] is not None:
            s.set_inlet_profile(d["inlet"])
        s.reinit_from_macro(*d["fields"])
        s.step(steps)
        return s.macros(), s.state()


@pytest.mark.parametrize("steps", [1, 2, 35])
@pytest.mark.parametrize("launch", [(0, 0, True), (512, 256, True), (256, 64, False)])
def test_batch_matches_single_domains(fs, steps, launch):
    doms = small_domains()
    with make_batch(fs, doms) as b:
        b.set_launch(*launch)
        b.step(steps)
        for i, d in enumerate(doms):
            mac, st = single(fs, d, steps)
            for k in range(4):
                np.testing.assert_array_equal(b.macro(i, k), mac[k], err_msg=f"dom {i} macro {k}")
            np.testing.assert_array_equal(b.state(i), st, err_msg=f"dom {i} state")


def test_batch_against_oracle(fs, oracle):
    doms = small_domains()[:3]
    with make_batch(fs, doms) as b:
        b.step(20)
        for i, d in enumerate(doms):
            o = oracle.Oracle(d["method"], *d["dims"], d["tau"], d["force"])
            o.set_flags(d["flags"])
            if d["inlet"] is not None:
                o.set_inlet(d["inlet"])
            o.reinit(*d["fields"])
            o.step(20)
            np.testing.assert_array_equal(b.state(i), o.state())


def test_batch_split_steps_and_autotune(fs):
    doms = small_domains()
    with make_batch(fs, doms) as b:
        before = [b.state(i) for i in range(len(doms))]
        p, h = b.autotune(4)
        assert p >= 0 and h >= 0
        for i in range(len(doms)):  # autotune restores the state
            np.testing.assert_array_equal(b.state(i), before[i])
        b.step(17)
        b.step(20)
        for i, d in enumerate(doms):
            mac, st = single(fs, d, 37)
            np.testing.assert_array_equal(b.state(i), st)


def test_batch_rejects_bad_input(fs):
    from myfluidsimulation_amd.capi import FluidSimError

    with pytest.raises(FluidSimError):
        fs.Batch([(1, 1, 8, 8, 0.6)])
    nx, ny, nz = 33, 10, 10  # outlet at x = 32 starts a 32-wide HOME tile
    with fs.Batch([(1, nx, ny, nz, 0.6)]) as b:
        with pytest.raises(FluidSimError):
            b.set_flags(0, wl.channel_flags(nx, ny, nz))
        b.step(1)
        with pytest.raises(FluidSimError):
            b.reinit_from_macro(0, np.ones(nx * ny * nz, np.float32))


def test_batch_outlet_at_pair_tile_start(fs):
    """A population outlet at x = 32 starts a 32-wide tile row of the two-step
    batch kernel, so every population domain of the batch takes one step per
    launch; results still equal the single-domain solvers'."""
    doms = small_domains()[2:4]
    nx, ny, nz = 33, 10, 10
    doms.append(dict(method=0, dims=(nx, ny, nz), tau=0.6, flags=wl.channel_flags(nx, ny, nz),
                     fields=wl.uniform_flow(nx * ny * nz, 0.03),
                     inlet=np.full(ny * nz, 0.03, np.float32), force=(0.0, 0.0, 0.0)))
    with make_batch(fs, doms) as b:
        b.step(6)
        b.step(5)
        for i, d in enumerate(doms):
            mac, st = single(fs, d, 11)
            np.testing.assert_array_equal(b.state(i), st, err_msg=f"dom {i} state")
            for k in range(4):
                np.testing.assert_array_equal(b.macro(i, k), mac[k], err_msg=f"dom {i} macro {k}")


@pytest.mark.parametrize("seed", range(4))
def test_batch_random_domains(fs, seed):
    """Seeded random small domains of every method in one batch, stepped in
    random chunks: every domain equals its single-domain solver."""
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_gpu_parity import random_pop_case

    rng = np.random.default_rng(500 + seed)
    doms = []
    for i in range(4):
        dims, flags, fields, inlet, method, _, _, force = random_pop_case(1000 + 10 * seed + i)
        method = int(rng.integers(0, 3))
        nx = dims[0]
        if method == 1:
            flags = flags & np.uint8(0xFD)
            if (nx - 1) % 32 == 0:  # a HOME outlet may not start a 32-wide tile
                flags = np.where((flags & 12) != 0, 0, flags).astype(np.uint8)
        doms.append(dict(method=method, dims=dims, tau=0.6 + 0.1 * i, flags=flags,
                         fields=fields, inlet=inlet, force=force))
    chunks = [int(k) for k in rng.integers(1, 8, size=3)]
    with make_batch(fs, doms) as b:
        for k in chunks:
            b.step(k)
        for i, d in enumerate(doms):
            mac, st = single(fs, d, sum(chunks))
            np.testing.assert_array_equal(b.state(i), st, err_msg=f"seed {seed} dom {i}")
            for q in range(4):
                np.testing.assert_array_equal(b.macro(i, q), mac[q], err_msg=f"seed {seed} dom {i}")


def test_seeded_sweep_64_domains(fs, oracle, tmp_path):
    """The 64-domain batched sweep (seed 7: edges 32..128, tau 0.52..0.9,
    mixed BGK / HOME / Hybrid with bands 1..3, half Taylor-Green, half
    icosphere channels): after 101 batch steps (two-step launches and one
    one-step launch) every domain equals its single-domain solver bit for
    bit, and three domains (one per method) equal the FP32 oracle."""
    steps = 101
    specs = wl.batch_sweep_specs(64, 7)
    built = [wl.build_sweep_domain(sp, str(tmp_path)) for sp in specs]
    doms = []
    for sp, (flags, fields, inlet) in zip(specs, built):
        n = sp["nx"] * sp["ny"] * sp["nz"]
        doms.append(dict(method=sp["method"], dims=(sp["nx"], sp["ny"], sp["nz"]), tau=sp["tau"],
                         flags=np.zeros(n, np.uint8) if flags is None else flags, fields=fields,
                         inlet=inlet, force=(0.0, 0.0, 0.0)))
    with make_batch(fs, doms) as b:
        b.step(steps)
        for i, d in enumerate(doms):
            print(f"sweep domain {i}: method {d['method']} dims {d['dims']}", flush=True)
            mac, st = single(fs, d, steps)
            np.testing.assert_array_equal(b.state(i), st, err_msg=f"dom {i} state")
            for k in range(4):
                np.testing.assert_array_equal(b.macro(i, k), mac[k], err_msg=f"dom {i} macro {k}")
        picked = {}
        for i, d in enumerate(doms):  # the smallest domain of each method
            n = int(np.prod(d["dims"]))
            if d["method"] not in picked or n < picked[d["method"]][1]:
                picked[d["method"]] = (i, n)
        assert len(picked) == 3
        for i, _ in picked.values():
            d = doms[i]
            o = oracle.Oracle(d["method"], *d["dims"], d["tau"])
            o.set_flags(d["flags"])
            if d["inlet"] is not None:
                o.set_inlet(d["inlet"])
            o.reinit(*d["fields"])
            o.step(steps)
            np.testing.assert_array_equal(b.state(i), o.state(), err_msg=f"dom {i} vs oracle")
            for k in range(4):
                np.testing.assert_array_equal(b.macro(i, k), o.macro(k), err_msg=f"dom {i} macro")
