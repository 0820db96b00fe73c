"""CPU: pins the oracle's geometry (voxelizer, STL loader, band BFS).

The reference voxelizer needs GLM, which this image lacks, so it cannot be
built here; the restatement is pinned by known answers instead: closed-form
voxel counts of axis-aligned boxes, an independent double-precision
separating-axis test away from ties, an exact rational one on dyadic inputs
(ties included), sphere volume, STL byte fixtures, and a brute-force taxicab
distance for the band.
"""
import os
import struct

import numpy as np
import pytest
from scipy import ndimage

from myfluidsimulation_amd import workloads as wl


def cube_tris(lo, hi):
    """Closed axis-aligned box as 12 triangles (soup, (12, 9))."""
    c = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]],
                  [lo[0], hi[1], lo[2]], [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]],
                  [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]], np.float64)
    faces = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (2, 3, 7),
             (2, 7, 6), (1, 2, 6), (1, 6, 5), (0, 4, 7), (0, 7, 3)]
    return np.array([np.concatenate([c[a], c[b], c[d]]) for a, b, d in faces], np.float32)


def sat64(c, h, v0, v1, v2):
    """Triangle / axis-aligned box overlap by the separating axis theorem in
    float64 (13 axes: 3 box normals, the face normal, 9 edge crosses);
    returns (overlap, margin) with margin the smallest |gap| seen."""
    v = np.array([v0, v1, v2], np.float64) - np.asarray(c, np.float64)
    h = np.asarray(h, np.float64)
    axes = [np.eye(3)[i] for i in range(3)]
    e = [v[1] - v[0], v[2] - v[1], v[0] - v[2]]
    axes.append(np.cross(e[0], e[1]))
    for ed in e:
        for i in range(3):
            axes.append(np.cross(ed, np.eye(3)[i]))
    margin = np.inf
    for a in axes:
        if np.dot(a, a) < 1e-24:
            continue
        p = v @ a
        r = np.abs(a) @ h
        gap = max(p.min() - r, -r - p.max())
        margin = min(margin, abs(gap) / np.linalg.norm(a))
        if gap > 0:
            return False, margin
    return True, margin


def test_sat_against_float64_sat(oracle):
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(4000):
        c = rng.random(3)
        h = 0.05 + 0.1 * rng.random(3)
        tri = c + 0.4 * rng.standard_normal((3, 3))
        ok64, margin = sat64(c, h, *tri)
        if margin < 1e-4:
            continue
        assert oracle.tri_box_overlap(c, h, *tri) == ok64
        checked += 1
    assert checked > 3000


def sat_exact(c, h, v0, v1, v2):
    """Triangle / axis-aligned box overlap by the separating axis theorem in
    exact rational arithmetic, with touching counted as overlap (separation
    only where a projection gap is strictly positive, as voxelize.cpp:23-81
    compares); returns (overlap, tie) where tie says some axis touched
    exactly (gap 0).  Axes with a zero edge or a zero cross product are
    skipped, as the reference's 1e-12 / 1e-20 thresholds skip them."""
    from fractions import Fraction as Fr

    c = [Fr(x) for x in c]
    h = [Fr(x) for x in h]
    v = [[Fr(a) - b for a, b in zip(p, c)] for p in (v0, v1, v2)]

    def cross(a, b):
        return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]

    def dot(a, b):
        return sum(x * y for x, y in zip(a, b))

    e = [[v[1][i] - v[0][i] for i in range(3)], [v[2][i] - v[1][i] for i in range(3)],
         [v[0][i] - v[2][i] for i in range(3)]]
    axes = [[Fr(int(i == j)) for j in range(3)] for i in range(3)]
    axes.append(cross(e[0], e[1]))
    for ed in e:
        if all(x == 0 for x in ed):
            continue
        for i in range(3):
            axes.append(cross(ed, [Fr(int(i == j)) for j in range(3)]))
    tie = False
    for k, a in enumerate(axes):
        if k > 3 and dot(a, a) == 0:
            continue
        p = [dot(a, vi) for vi in v]
        r = sum(abs(x) * y for x, y in zip(a, h))
        if k == 3:  # the face normal: |n . v0| against r
            gap = abs(p[0]) - r
        else:
            gap = max(min(p) - r, -r - max(p))
        tie = tie or gap == 0
        if gap > 0:
            return False, tie
    return True, tie


def test_sat_exact_ties_dyadic(oracle):
    """Dyadic vertices (multiples of 1/4 in [-2, 2]) and boxes (centres on the
    half grid, half-sizes 1/4 or 1/2) keep every float32 product and sum of
    the test exact, so the oracle must give the exact answer, ties included:
    a vertex, edge or face exactly on a box face counts as overlap.  A strict
    / non-strict comparison swapped on any axis fails here."""
    rng = np.random.default_rng(11)
    ties = 0
    for _ in range(8000):
        c = rng.integers(-2, 3, 3) / 2.0
        h = rng.choice([0.25, 0.5], 3)
        tri = rng.integers(-8, 9, (3, 3)) / 4.0
        ok, tie = sat_exact(c, h, *tri)
        assert oracle.tri_box_overlap(c, h, *tri) == ok, (c, h, tri)
        ties += tie and ok
    assert ties > 150


@pytest.mark.parametrize("n,lo,hi,expect", [
    (16, (0.25,) * 3, (0.75,) * 3, 10 ** 3),  # faces on voxel boundaries touch
    (16, (0.3,) * 3, (0.7,) * 3, 8 ** 3),
    (20, (0.11, 0.22, 0.33), (0.87, 0.61, 0.52), None),  # faces inside cells
])
def test_axis_aligned_box_voxel_count(oracle, n, lo, hi, expect):
    """The box mesh fit with scale = its largest side and translate = its
    centre is the identity map; a voxel is solid iff its closed cell meets
    the closed box."""
    tris = cube_tris(lo, hi)
    bmin, bmax = np.array(lo, np.float32), np.array(hi, np.float32)
    size = float(np.max(bmax - bmin))
    center = tuple(float(x) for x in np.float32(0.5) * (bmin + bmax))
    got = oracle.voxelize(tris, bmin, bmax, n, n, n, size, center).reshape(n, n, n)
    cnt = []
    for k in range(3):
        i = np.arange(n)
        cnt.append(int(np.sum((i / n <= hi[k]) & ((i + 1) / n >= lo[k]))))
    assert got.sum() == int(np.prod(cnt))
    if expect is not None:
        assert got.sum() == expect


def test_amr_downsample_closed_form(oracle):
    """refine 2 on the [0.25,0.75]^3 box: fine cells 7..24 per axis; coarse
    cells with coverage >= 0.5 are the 8^3 interior plus the 6 faces."""
    tris = cube_tris((0.25,) * 3, (0.75,) * 3)
    bmin, bmax = np.full(3, 0.25, np.float32), np.full(3, 0.75, np.float32)
    got = oracle.voxelize(tris, bmin, bmax, 16, 16, 16, 0.5, (0.5, 0.5, 0.5), 2, 0.5)
    assert got.sum() == 8 ** 3 + 6 * 8 ** 2
    got = oracle.voxelize(tris, bmin, bmax, 16, 16, 16, 0.5, (0.5, 0.5, 0.5), 2, 1.0)
    assert got.sum() == 8 ** 3
    got = oracle.voxelize(tris, bmin, bmax, 16, 16, 16, 0.5, (0.5, 0.5, 0.5), 2, 0.0)
    assert got.sum() == 10 ** 3  # cov > 0 still required


def test_sphere_pointwise_bounds(oracle):
    """Icosphere (inscribed in the ball of radius r, faces at distance >= r_in
    from the centre) voxelized about the cube centre: every voxel whose centre
    lies within r_in - sqrt(3)/2 is solid (interior fill), every voxel whose
    centre lies beyond r + sqrt(3)/2 is empty (no leak of the flood fill)."""
    soup = wl.jittered_icosphere((0, 0, 0), 1.0, 1, jitter=0.0)
    tris = soup.astype(np.float32)
    pts = tris.reshape(-1, 3).astype(np.float64)
    bmin, bmax = tris.reshape(-1, 3).min(0), tris.reshape(-1, 3).max(0)
    n = 64
    got = oracle.voxelize(tris, bmin, bmax, n, n, n, 0.5, (0.5, 0.5, 0.5)).reshape(n, n, n)
    t = pts.reshape(-1, 3, 3)
    nrm = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    r_in = np.min(np.abs(np.einsum("ij,ij->i", nrm, t[:, 0])) / np.linalg.norm(nrm, axis=1))
    scale = 0.5 / float(np.max(bmax - bmin)) * n  # cells per mesh unit
    c = (np.arange(n) + 0.5) - n / 2
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    d = np.sqrt(x * x + y * y + z * z)
    h = np.sqrt(3) / 2
    assert got[d < r_in * scale - h].all()
    assert not got[d > 1.0 * scale + h].any()


def _binary_stl(path, tris, header=b"fixture"):
    with open(path, "wb") as fh:
        fh.write(header.ljust(80, b"\0"))
        fh.write(struct.pack("<I", len(tris)))
        for t in tris:
            fh.write(struct.pack("<3f", 0, 0, 1))
            fh.write(np.asarray(t, "<f4").tobytes())
            fh.write(b"\0\0")


def _ascii_stl(path, tris):
    with open(path, "w") as fh:
        fh.write("solid fixture\n")
        for t in tris:
            fh.write("  facet normal 0 0 1\n    outer loop\n")
            for k in range(3):
                fh.write("      vertex {} {} {}\n".format(*t[3 * k:3 * k + 3]))
            fh.write("    endloop\n  endfacet\n")
        fh.write("endsolid fixture\n")


TRIS = [[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.5, 0.0],
        [0.1, -2.0, 3.0, 0.25, 0.5, 0.75, 1e-3, 2e5, -7.125]]


def test_stl_ascii_binary_and_failures(oracle, tmp_path):
    a = os.path.join(tmp_path, "a.stl")
    b = os.path.join(tmp_path, "b.stl")
    _ascii_stl(a, TRIS)
    _binary_stl(b, TRIS)
    for p in (a, b):
        tris, bmin, bmax = oracle.load_stl(p)
        np.testing.assert_array_equal(tris, np.array(TRIS, np.float32))
        np.testing.assert_array_equal(bmin, np.array(TRIS, np.float32).reshape(-1, 3).min(0))
        np.testing.assert_array_equal(bmax, np.array(TRIS, np.float32).reshape(-1, 3).max(0))
    # binary whose header starts with "solid": ASCII parse finds no vertex,
    # falls back to binary (stl_loader.cpp:108-118)
    c = os.path.join(tmp_path, "c.stl")
    _binary_stl(c, TRIS, header=b"solid but binary")
    np.testing.assert_array_equal(oracle.load_stl(c)[0], np.array(TRIS, np.float32))
    # truncated binary record -> failure
    d = os.path.join(tmp_path, "d.stl")
    with open(b, "rb") as fh:
        raw = fh.read()
    with open(d, "wb") as fh:
        fh.write(raw[:-10])
    assert oracle.load_stl(d) is None
    assert oracle.load_stl(os.path.join(tmp_path, "missing.stl")) is None


@pytest.mark.parametrize("d0", [0, 1, 2, 3])
def test_band_matches_taxicab_distance(oracle, d0):
    """6-neighbour BFS over the whole (non-periodic) grid = taxicab distance
    to the nearest solid cell; band = fluid cells within d0."""
    nx, ny, nz = 20, 14, 11
    rng = np.random.default_rng(d0)
    solid = rng.random((nz, ny, nx)) < 0.01
    solid[3:6, 4:8, 5:9] = True
    got = oracle.band(solid.reshape(-1).astype(np.uint8), nx, ny, nz, d0).reshape(nz, ny, nx)
    dist = ndimage.distance_transform_cdt(~solid, metric="taxicab")
    expect = (~solid) & (dist <= d0)
    np.testing.assert_array_equal(got.astype(bool), expect)
    dist_o = oracle.distance(solid.reshape(-1).astype(np.uint8), nx, ny, nz, 4).reshape(nz, ny, nx)
    ref = np.where(dist <= 4, dist, -1)
    np.testing.assert_array_equal(dist_o, ref)


def test_band_max_dist_at_least_one(oracle):
    """maxDist = max(1, d0) (sim_runner.hpp:222): d0 = 0 gives no band."""
    solid = np.zeros(6 * 6 * 6, np.uint8)
    solid[100] = 1
    assert oracle.band(solid, 6, 6, 6, 0).sum() == 0


def test_vtk_writer_format(oracle, tmp_path):
    p = os.path.join(tmp_path, "s.vtk")
    field = np.array([1.0, 0.5, 1.23456789, -2.5e-7, 1e20, 0.0], np.float32)
    oracle.write_scalar_vtk(p, field, 3, 2, 1)
    txt = open(p).read()
    head = ("# vtk DataFile Version 3.0\nscalar\nASCII\nDATASET STRUCTURED_POINTS\n"
            "DIMENSIONS 3 2 1\nORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA 6\n"
            "SCALARS s float 1\nLOOKUP_TABLE default\n")
    assert txt == head + "1\n0.5\n1.23457\n-2.5e-07\n1e+20\n0\n"
    q = os.path.join(tmp_path, "v.vtk")
    oracle.write_vector_vtk(q, field[:2], field[2:4], field[4:6], 2, 1, 1)
    assert open(q).read().endswith("VECTORS vec float\n1 1.23457 1e+20\n0.5 -2.5e-07 0\n")


# ---------------------------------------------------------------- field helpers
# kern_speed (field_ops.cu:12-18), kern_rho_slab / kern_rho_gauss
# (init_profiles.cu:7-29): pinned by closed forms and exact special cases
# chosen so that a dropped component, a corner-instead-of-centre coordinate
# or a sigma / 2 sigma^2 slip changes the result.

def test_speed_exact_quadruples_and_float64(oracle):
    # integer Pythagorean quadruples: every product, sum and the root are
    # exact in float32, whatever the component order or signs
    quads = [(2, 3, 6, 7), (1, 4, 8, 9), (2, 6, 9, 11), (1, 2, 2, 3), (4, 4, 7, 9),
             (2, 10, 11, 15), (0, 0, 5, 5), (0, 3, 0, 3)]
    u = []
    want = []
    for a, b, c, d in quads:
        for perm in ((a, b, c), (c, a, b), (b, c, a)):
            for sgn in ((1, 1, 1), (-1, 1, -1), (1, -1, -1)):
                for scale in (1.0, 2.0 ** -7, 2.0 ** 5):
                    u.append([s * p * scale for s, p in zip(sgn, perm)])
                    want.append(d * scale)
    u = np.array(u, np.float32)
    got = oracle.speed(u[:, 0], u[:, 1], u[:, 2])
    np.testing.assert_array_equal(got, np.array(want, np.float32))
    # random vectors: within 2 ulp of the float64 magnitude of the float32
    # inputs (products, two sums and the root each round once)
    rng = np.random.default_rng(17)
    v = (rng.standard_normal((20000, 3)) * 10.0 ** rng.uniform(-4, 1, (20000, 1))).astype(np.float32)
    got = oracle.speed(v[:, 0], v[:, 1], v[:, 2]).astype(np.float64)
    ref = np.sqrt((v.astype(np.float64) ** 2).sum(axis=1))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(got - ref) <= 2.0 * ulp)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_rho_slab_cell_centres(oracle, axis):
    """The slab holds the cells whose centre s = (i + 0.5)/N satisfies
    |s - centre| < width/2: with N = 20, centre 0.5, width 0.5 that is
    i = 5..14 (a corner coordinate i/N would give i = 6..14)."""
    dims = [24, 18, 14]
    dims[axis] = 20
    nx, ny, nz = dims
    amp = np.float32(3e-3)
    rho = oracle.rho_slab(nx, ny, nz, axis, amp, 0.5, 0.5).reshape(nz, ny, nx)
    idx = np.indices((nz, ny, nx))[2 - axis]
    inside = (idx >= 5) & (idx <= 14)
    assert np.array_equal(rho[inside], np.full(inside.sum(), np.float32(1) + amp))
    assert np.all(rho[~inside] == np.float32(1.0))
    # an off-centre slab, counted from the cell centres in float64
    rho = oracle.rho_slab(nx, ny, nz, axis, amp, 0.3125, 0.2).reshape(nz, ny, nx)
    line = np.moveaxis(rho, 2 - axis, 0)[:, 3, 2]
    n = dims[axis]
    s = (np.arange(n) + 0.5) / n
    cells = np.flatnonzero(np.abs(s - 0.3125) < 0.1)
    assert np.array_equal(np.flatnonzero(line != np.float32(1.0)), cells)
    assert np.all(np.moveaxis(rho, 2 - axis, 0) == line[:, None, None])


def test_rho_gauss_peak_sigma_and_closed_form(oracle):
    """Cell centres of a 16-cell axis are exact binary fractions: centred on
    cell 8 (0.53125), the peak is 1 + amp exactly, and with sigma = 1/16 the
    next cell along x sits at exactly one sigma, where the profile is
    1 + amp e^{-1/2} (not e^{-1}, as a missing factor 2 would give)."""
    n = 16
    amp, sig = np.float32(0.01), np.float32(1.0 / 16.0)
    c = 8.5 / 16
    rho = oracle.rho_gauss(n, n, n, c, c, c, sig, amp).reshape(n, n, n)
    assert rho[8, 8, 8] == np.float32(1.0) + amp
    one_sigma = 1.0 + float(amp) * np.exp(-0.5)
    for cell in ((8, 8, 9), (8, 9, 8), (9, 8, 8), (8, 8, 7)):
        assert abs(float(rho[cell]) - one_sigma) <= 2 * np.spacing(np.float32(1.0))
    assert abs(float(rho[8, 9, 9]) - (1.0 + float(amp) * np.exp(-1.0))) <= 2 * np.spacing(
        np.float32(1.0))
    # everywhere: the float64 closed form at the float32 inputs, within a few ulp of 1
    nx, ny, nz = 30, 22, 12
    cx, cy, cz = np.float32(0.41), np.float32(0.6), np.float32(0.37)
    sig, amp = np.float32(0.15), np.float32(0.02)
    rho = oracle.rho_gauss(nx, ny, nz, cx, cy, cz, sig, amp).reshape(nz, ny, nx).astype(np.float64)
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz, (np.arange(ny) + 0.5) / ny,
                          (np.arange(nx) + 0.5) / nx, indexing="ij")
    r2 = (x - float(cx)) ** 2 + (y - float(cy)) ** 2 + (z - float(cz)) ** 2
    ref = 1.0 + float(amp) * np.exp(-r2 / (2.0 * float(sig) ** 2))
    assert np.max(np.abs(rho - ref)) <= 4 * np.spacing(np.float32(1.0))
