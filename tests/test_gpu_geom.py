"""GPU geometry and field helpers against the CPU oracle: voxelizer (bit for
bit), band BFS (bit for bit), |u|, density profiles, the stand-alone feq
re-initialisation."""
import os

import numpy as np
import pytest

from myfluidsimulation_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _stl(tmp_path, nx, ny, nz, center, radius, seed):
    path = os.path.join(tmp_path, f"ico_{seed}.stl")
    scale, trans = wl.obstacle_stl(path, nx, ny, nz, center, radius, seed)
    return path, scale, trans


@pytest.mark.parametrize("refine", [1, 2, 3])
def test_voxelize_icosphere_exact(fs, oracle, tmp_path, refine):
    nx, ny, nz = 96, 32, 32
    path, scale, trans = _stl(str(tmp_path), nx, ny, nz, (24, 16, 16), 6.0, 11 + refine)
    tris, bmin, bmax = fs.load_stl(path)
    otris, obmin, obmax = oracle.load_stl(path)
    np.testing.assert_array_equal(tris, otris)
    np.testing.assert_array_equal(bmin, obmin)
    np.testing.assert_array_equal(bmax, obmax)
    g = fs.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans, refine, 0.5)
    o = oracle.voxelize(otris, obmin, obmax, nx, ny, nz, scale, trans, refine, 0.5)
    assert g.sum() > 0
    assert np.array_equal(g, o), f"{int((g != o).sum())} voxels differ"


@pytest.mark.parametrize("shape", [(1, 12, 9), (2, 5, 30), (63, 17, 4), (64, 9, 7), (65, 11, 6),
                                   (129, 8, 5), (257, 4, 4), (520, 5, 4)])
def test_voxelize_line_lengths_exact(fs, oracle, tmp_path, shape):
    """x-line lengths around multiples of 64 cells and of 256 (the x sweep
    handles a line 64 cells at a time, four words per batch)."""
    nx, ny, nz = shape
    for k in range(2):
        center = (nx * (0.35 + 0.3 * k), ny * 0.5, nz * 0.45)
        radius = max(1.5, 0.35 * min(nx, ny, nz)) * (0.8 + 0.4 * k)
        path, scale, trans = _stl(str(tmp_path), nx, ny, nz, center, radius, 40 + k)
        tris, bmin, bmax = fs.load_stl(path)
        g = fs.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
        o = oracle.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
        assert np.array_equal(g, o), f"{int((g != o).sum())} voxels differ"


def _box(lo, hi):
    """12 triangles of an axis-aligned box."""
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    faces = [(0, 1, 2), (0, 2, 3), (4, 6, 5), (4, 7, 6), (0, 4, 5), (0, 5, 1),
             (3, 2, 6), (3, 6, 7), (0, 3, 7), (0, 7, 4), (1, 5, 6), (1, 6, 2)]
    return np.stack([v[list(f)].reshape(9) for f in faces])


def _shell(c0, c1, t, hole_side):
    """Hollow cube [c0, c1]^3 with walls of thickness t; hole_side +1 / -1
    cuts a square opening (y, z in [0.4, 0.6]) into the +x / -x wall, 0 none."""
    parts = []
    for axis in range(3):
        for side in (0, 1):
            lo, hi = [c0] * 3, [c1] * 3
            if side == 0:
                hi[axis] = c0 + t
            else:
                lo[axis] = c1 - t
            if axis == 0 and hole_side == (1 if side else -1):
                a, b = 0.4, 0.6  # the wall minus the opening: four slabs
                for ylo, yhi, zlo, zhi in ((c0, a, c0, c1), (b, c1, c0, c1), (a, b, c0, a),
                                           (a, b, b, c1)):
                    parts.append(_box((lo[0], ylo, zlo), (hi[0], yhi, zhi)))
            else:
                parts.append(_box(lo, hi))
    return np.concatenate(parts)


@pytest.mark.parametrize("shape", [(64, 64, 64), (130, 48, 40)])
def test_voxelize_nested_shells_exact(fs, oracle, tmp_path, shape):
    """Nested hollow cubes with openings on opposite sides around a closed
    one: the exterior reaches the space between the outer shells only by
    turning corners (several rounds of line sweeps), and the closed shell's
    inside is an enclosed cavity that ends up solid."""
    nx, ny, nz = shape
    soup = np.concatenate([_shell(0.1, 0.9, 0.05, +1), _shell(0.25, 0.75, 0.05, -1),
                           _shell(0.4, 0.6, 0.04, 0)])
    path = os.path.join(str(tmp_path), "shells.stl")
    wl.write_binary_stl(path, soup)
    tris, bmin, bmax = fs.load_stl(path)
    scale = float((bmax - bmin).max())
    trans = tuple(float(v) for v in 0.5 * (bmin + bmax))
    g = fs.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
    o = oracle.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
    assert np.array_equal(g, o), f"{int((g != o).sum())} voxels differ"
    m = g.reshape(nz, ny, nx)
    # between the outer two shells (reached through both openings) is fluid,
    # the innermost cube's centre is solid
    assert m[nz // 2, ny // 2, int(0.18 * nx)] == 0
    assert m[nz // 2, int(0.18 * ny), nx // 2] == 0
    assert m[nz // 2, ny // 2, int(0.32 * nx)] == 0
    assert m[nz // 2, ny // 2, nx // 2] == 1


def test_voxelize_reference_defaults_exact(fs, oracle, tmp_path):
    """Default VoxelParams path (scale 0.5 about the centre) on a non-cubic
    grid, plus thresholds 0.25 / 1.0 for the AMR downsample."""
    path = os.path.join(str(tmp_path), "s.stl")
    wl.write_binary_stl(path, wl.jittered_icosphere((0, 0, 0), 1.0, 5))
    tris, bmin, bmax = fs.load_stl(path)
    for thr in (0.25, 1.0):
        g = fs.voxelize(tris, bmin, bmax, 40, 30, 20, 0.5, (0.5, 0.5, 0.5), 2, thr)
        o = oracle.voxelize(tris, bmin, bmax, 40, 30, 20, 0.5, (0.5, 0.5, 0.5), 2, thr)
        assert np.array_equal(g, o)
    g = fs.voxelize(tris, bmin, bmax, 40, 30, 20)
    o = oracle.voxelize(tris, bmin, bmax, 40, 30, 20)
    assert np.array_equal(g, o)


def test_voxelize_channel_config(fs, oracle, tmp_path):
    """The obstacle of the 384x128x128 channel: jittered icosphere, radius
    20 cells, centre (96,64,64), 1280 faces."""
    nx, ny, nz = 384, 128, 128
    path, scale, trans = _stl(str(tmp_path), nx, ny, nz, (96, 64, 64), 20.0, 2024)
    tris, bmin, bmax = fs.load_stl(path)
    assert tris.shape == (1280, 9)
    g = fs.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
    o = oracle.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
    assert np.array_equal(g, o)
    vol = 4.0 / 3.0 * np.pi * 20.0 ** 3
    assert 0.7 * vol < g.sum() < 1.4 * vol


@pytest.mark.parametrize("nx", [40, 48], ids=["cellwise", "16-cell blocks"])
@pytest.mark.parametrize("d0", [0, 1, 2, 3, 5])
def test_band_exact(fs, oracle, d0, nx):
    """Rows of 40 cells take the one-cell kernels, rows of 48 (a multiple of
    16) the 16-cell block kernels."""
    ny, nz = 24, 20
    rng = np.random.default_rng(d0)
    solid = (rng.random(nx * ny * nz) < 0.02).astype(np.uint8)
    solid.reshape(nz, ny, nx)[:, 0, :] = 1
    g = fs.band(solid, nx, ny, nz, d0)
    o = oracle.band(solid, nx, ny, nz, d0)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("nx", [280, 288])
@pytest.mark.parametrize("d0", [200, 254, 255, 300])
def test_band_deep_exact(fs, oracle, d0, nx):
    """Bands of hundreds of rings on a long thin grid (Manhattan diameter
    299-307): distances are kept in one byte up to 254 rings (one-cell
    kernels for rows of 280 cells, 16-cell blocks for 288) and in an int
    beyond."""
    ny, nz = 12, 10
    solid = np.zeros(nx * ny * nz, np.uint8)
    solid.reshape(nz, ny, nx)[nz // 2, ny // 2, 3] = 1
    g = fs.band(solid, nx, ny, nz, d0)
    o = oracle.band(solid, nx, ny, nz, d0)
    assert np.array_equal(g, o)
    m = g.reshape(nz, ny, nx)
    assert m[nz // 2, ny // 2, 4] == 1
    assert m[nz // 2, ny // 2, nx - 1] == (1 if d0 >= (nx - 1 - 3) else 0)


def test_band_no_solid(fs):
    assert fs.band(np.zeros(8 * 8 * 8, np.uint8), 8, 8, 8, 3).sum() == 0


def test_fields_speed_and_profiles(fs, oracle):
    import ctypes as C

    from myfluidsimulation_amd import DeviceBuffer
    from myfluidsimulation_amd.capi import call

    nx, ny, nz = 30, 20, 10
    n = nx * ny * nz
    rng = np.random.default_rng(1)
    u, v, w = (rng.standard_normal(n).astype(np.float32) for _ in range(3))
    bu, bv, bw = (DeviceBuffer.from_host(a) for a in (u, v, w))
    out = DeviceBuffer(4 * n)
    call("fs_speed", C.c_void_p(bu.ptr), C.c_void_p(bv.ptr), C.c_void_p(bw.ptr),
         C.c_void_p(out.ptr), n)
    np.testing.assert_array_equal(out.to_host(np.float32, n), oracle.speed(u, v, w))
    call("fs_rho_slab", C.c_void_p(out.ptr), nx, ny, nz, 1, 2e-3, 0.4, 0.3)
    np.testing.assert_array_equal(out.to_host(np.float32, n),
                                  oracle.rho_slab(nx, ny, nz, 1, 2e-3, 0.4, 0.3))
    call("fs_rho_gauss", C.c_void_p(out.ptr), nx, ny, nz, 0.6, 0.5, 0.5, 0.08, 1e-3)
    # expf on the GPU vs std::exp on the host: tolerance, not bits
    np.testing.assert_allclose(out.to_host(np.float32, n),
                               oracle.rho_gauss(nx, ny, nz, 0.6, 0.5, 0.5, 0.08, 1e-3),
                               rtol=0, atol=2e-7)


def test_reinit_equilibrium_reference_entry(fs, oracle):
    """reinit_equilibrium_from_macro (lbm_reinit.cu:52-61) writes f[q*N+i] with
    the cs2 form: equal to the Legacy solver's reinit state."""
    import ctypes as C

    from myfluidsimulation_amd import DeviceBuffer
    from myfluidsimulation_amd.capi import lib

    nx, ny, nz = 16, 12, 8
    n = nx * ny * nz
    rho, ux, uy, uz = wl.taylor_green_3d(nx, ny, nz, 0.05)
    bufs = [DeviceBuffer.from_host(a) for a in (rho, ux, uy, uz)]
    f = DeviceBuffer(19 * 4 * n)
    lib().reinit_equilibrium_from_macro(*[C.c_void_p(b.ptr) for b in bufs], C.c_void_p(f.ptr), n)
    lib().fs_device_synchronize()
    o = oracle.Oracle(0, nx, ny, nz, 0.6)
    o.reinit(rho, ux, uy, uz)
    np.testing.assert_array_equal(f.to_host(np.float32, 19 * n), o.state())


def test_field_diff_and_shift(fs):
    """fs_diff / fs_shift and the reference-named launch_diff / launch_shift
    (field_ops.cu:20-41): a - b and a + bias, elementwise in float32."""
    import ctypes as C

    from myfluidsimulation_amd import DeviceBuffer
    from myfluidsimulation_amd.capi import call, lib

    n = 70001  # not a multiple of the block size
    rng = np.random.default_rng(3)
    a, b = (rng.standard_normal(n).astype(np.float32) for _ in range(2))
    ba, bb = DeviceBuffer.from_host(a), DeviceBuffer.from_host(b)
    out = DeviceBuffer(4 * n)
    call("fs_diff", C.c_void_p(ba.ptr), C.c_void_p(bb.ptr), C.c_void_p(out.ptr), n)
    np.testing.assert_array_equal(out.to_host(np.float32, n), a - b)
    call("fs_shift", C.c_void_p(ba.ptr), C.c_float(-1.0), C.c_void_p(out.ptr), n)
    np.testing.assert_array_equal(out.to_host(np.float32, n), a + np.float32(-1.0))
    lib().launch_diff(C.c_void_p(bb.ptr), C.c_void_p(ba.ptr), C.c_void_p(out.ptr), n)
    lib().fs_device_synchronize()
    np.testing.assert_array_equal(out.to_host(np.float32, n), b - a)
    lib().launch_shift(C.c_void_p(ba.ptr), C.c_float(0.25), C.c_void_p(out.ptr), n)
    lib().fs_device_synchronize()
    np.testing.assert_array_equal(out.to_host(np.float32, n), a + np.float32(0.25))


def read_binary_vtk(path, ncomp):
    data = open(path, "rb").read()
    lines, pos = [], 0
    n_header = 10 if ncomp == 1 else 9
    for _ in range(n_header):
        end = data.index(b"\n", pos)
        lines.append(data[pos:end].decode())
        pos = end + 1
    n = int(lines[7].split()[1])
    vals = np.frombuffer(data[pos:pos + 4 * n * ncomp], dtype=">f4")
    assert data[pos + 4 * n * ncomp:] == b"\n"
    return lines, vals.astype(np.float32)


def test_binary_vtk_writer(fs, oracle, tmp_path):
    """BINARY legacy VTK: the ASCII writer's header with BINARY, then the exact
    float32 values big-endian (vectors interleaved per point)."""
    import ctypes as C

    from myfluidsimulation_amd import DeviceBuffer
    from myfluidsimulation_amd.capi import call

    nx, ny, nz = 7, 5, 3
    n = nx * ny * nz
    rng = np.random.default_rng(8)
    u, v, w = (rng.standard_normal(n).astype(np.float32) for _ in range(3))
    bufs = [DeviceBuffer.from_host(a) for a in (u, v, w)]
    ps, pv = str(tmp_path / "s.vtk"), str(tmp_path / "v.vtk")
    call("fs_write_scalar_vtk_binary", ps.encode(), C.c_void_p(bufs[0].ptr), nx, ny, nz)
    call("fs_write_vector_vtk_binary", pv.encode(), *[C.c_void_p(b.ptr) for b in bufs], nx, ny, nz)
    lines, vals = read_binary_vtk(ps, 1)
    np.testing.assert_array_equal(vals, u)
    ascii_path = str(tmp_path / "a.vtk")
    oracle.write_scalar_vtk(ascii_path, u, nx, ny, nz)
    ascii_head = open(ascii_path).read().split("\n")[:10]
    assert lines == [ln.replace("ASCII", "BINARY") for ln in ascii_head]
    lines, vals = read_binary_vtk(pv, 3)
    np.testing.assert_array_equal(vals.reshape(n, 3), np.stack([u, v, w], axis=1))
    assert lines[2] == "BINARY" and lines[7] == f"POINT_DATA {n}"


def test_geometry_scratch_reuse_and_release(fs, oracle, tmp_path):
    """fs_voxelize / fs_band keep their device scratch and pinned staging
    between calls: a large call, a smaller one (odd cell counts, so packed
    words end mid-word), fs_geometry_release, and the small call again all
    equal the oracle; masks whose bytes are not 0/1 count as nonzero."""
    from myfluidsimulation_amd import capi

    path = str(tmp_path / "ico.stl")
    for nx, ny, nz in ((96, 64, 48), (37, 23, 19)):
        scale, trans = wl.obstacle_stl(path, nx, ny, nz, (nx / 3, ny / 2, nz / 2), min(ny, nz) / 4,
                                       seed=nx)
        tris, bmin, bmax = fs.load_stl(path)
        g = fs.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
        assert np.array_equal(g, oracle.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans))
        solid = (g * 7).astype(np.uint8)  # nonzero, not 1
        assert np.array_equal(fs.band(solid, nx, ny, nz, 2), oracle.band(g, nx, ny, nz, 2))
    capi.call("fs_geometry_release")
    g2 = fs.voxelize(tris, bmin, bmax, nx, ny, nz, scale, trans)
    assert np.array_equal(g2, g)
    assert np.array_equal(fs.band(g2, nx, ny, nz, 3), oracle.band(g2, nx, ny, nz, 3))
