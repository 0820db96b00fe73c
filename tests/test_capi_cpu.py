"""CPU: the C-ABI library loads, exports every function include/fluidsim.h
declares, validates arguments before touching the device, reports errors by
status code and message, and its host-only STL parser matches the oracle.
No compute call runs here (there is no GPU in this container)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from myfluidsimulation_amd import capi, workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fluidsim.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|const char\*)\s+\**(\w+)\s*\(", txt, re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) > 40
    lib = C.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    # and the Python binding declares a signature for each of them
    assert set(names) <= set(capi.SIGNATURES)


def test_version_and_device_count():
    L = capi.lib()
    assert L.fs_version() == capi.ABI_VERSION
    assert capi.device_count() >= 0


def _create(method, nx, ny, nz, tau):
    h = C.c_void_p()
    rc = capi.lib().fs_solver_create(method, nx, ny, nz, tau, 0.0, 0.0, 0.0, C.byref(h))
    return rc, h


@pytest.mark.parametrize("args", [(7, 8, 8, 8, 0.6), (0, 1, 8, 8, 0.6), (0, 8, 8, 0, 0.6),
                                  (1, 8, 8, 8, 0.5), (2, 8, 8, 8, float("nan")),
                                  (0, 2048, 2048, 1024, 0.6)])
def test_create_rejects_bad_arguments(args):
    rc, h = _create(*args)
    assert rc == capi.FS_ERR_INVALID
    assert not h.value
    assert capi.lib().fs_last_error()


def test_create_without_device_or_with_injected_fault(monkeypatch):
    if capi.device_count() == 0:
        rc, h = _create(0, 16, 16, 16, 0.6)
        assert rc == capi.FS_ERR_DEVICE
        assert b"device" in capi.lib().fs_last_error()
    monkeypatch.setenv("FS_FAULT_INJECT", "alloc")
    rc, h = _create(0, 16, 16, 16, 0.6)
    assert rc == capi.FS_ERR_ALLOC
    p = C.c_void_p()
    assert capi.lib().fs_device_malloc(C.byref(p), 64) == capi.FS_ERR_ALLOC


def test_null_handles_are_rejected():
    L = capi.lib()
    assert L.fs_solver_step(None, 1) == capi.FS_ERR_INVALID
    assert L.fs_solver_reset(None) == capi.FS_ERR_INVALID
    assert L.fs_solver_destroy(None) == capi.FS_OK
    assert L.fs_solver_set_flags(None, None) == capi.FS_ERR_INVALID
    assert L.fs_voxelize(None, 0, None, None, 0, 0, 0, 0.5, 0, 0, 0, 1, 0.5,
                         None) == capi.FS_ERR_INVALID
    assert L.fs_band(None, 4, 4, 4, 1, None) == capi.FS_ERR_INVALID


def test_stl_loader_matches_oracle(oracle, tmp_path):
    """fs_load_stl is host code: binary and ASCII files parse to the oracle's
    bits, a missing file is FS_ERR_IO (the reference's `false`)."""
    from myfluidsimulation_amd import load_stl

    p = os.path.join(tmp_path, "ico.stl")
    wl.write_binary_stl(p, wl.jittered_icosphere((0.3, -1.0, 2.0), 0.7, 9))
    got = load_stl(p)
    ref = oracle.load_stl(p)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    q = os.path.join(tmp_path, "a.stl")
    tris = got[0][:5]
    with open(q, "w") as fh:
        fh.write("solid x\n")
        for t in tris:
            fh.write("facet normal 0 0 0\nouter loop\n")
            for k in range(3):
                fh.write("vertex %.9g %.9g %.9g\n" % tuple(t[3 * k:3 * k + 3]))
            fh.write("endloop\nendfacet\n")
    for a, b in zip(load_stl(q), oracle.load_stl(q)):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(load_stl(q)[0], tris)
    with pytest.raises(capi.FluidSimError) as e:
        load_stl(os.path.join(tmp_path, "missing.stl"))
    assert e.value.code == capi.FS_ERR_IO


def test_flag_validation_before_device():
    """Inlet/outlet rules are checked on the host (no handle needed to see
    them fail: a null solver is rejected first, so use the oracle's rule)."""
    flags = wl.channel_flags(8, 6, 6)
    fl = flags.reshape(6, 6, 8)
    assert ((fl[:, :, 0] & capi.FLAG_INLET) != 0).sum() == 16
    assert ((fl[:, :, 7] & capi.FLAG_OUTLET) != 0).sum() == 16
    assert not np.any((fl & capi.FLAG_SOLID) & (fl & (capi.FLAG_INLET | capi.FLAG_OUTLET)))


def _bin(name):
    return os.path.join(ROOT, "myfluidsimulation_amd", "bin", name)


@pytest.mark.skipif(capi.device_count() > 0, reason="checks the no-device error path")
def test_cli_reports_missing_device_and_stl(tmp_path):
    r = subprocess.run([_bin("fluidsim_legacy"), "--nx", "16", "--ny", "16", "--nz", "8",
                        "--stl", os.path.join(tmp_path, "none.stl"), "--steps", "1"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "STL の読み込みに失敗しました" in r.stderr
    assert "[error]" in r.stderr


def _torch_hip_version():
    """major.minor of the HIP runtime the installed torch bundles, read from
    its version.py without importing torch (None when torch is absent or a
    CPU build)"""
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return None
    txt = open(os.path.join(os.path.dirname(spec.origin), "version.py")).read()
    m = re.search(r"^hip\b.*?=\s*'(\d+)\.(\d+)", txt, re.M)
    return (int(m.group(1)), int(m.group(2))) if m else None


_RUNTIME_PROBE = r"""
import ctypes as C, sys
order = sys.argv[1]
def load():
    L = C.CDLL(sys.argv[2])
    L.fs_last_error.restype = C.c_char_p
    return L
if order == "torch_first":
    import torch  # noqa: F401
    L = load()
else:
    L = load()
    import torch  # noqa: F401
n = C.c_int(-1)
rc = L.fs_device_count(C.byref(n))
print(rc)
print(L.fs_last_error().decode())
"""


@pytest.mark.parametrize("order", ["torch_first", "library_first"])
def test_foreign_hip_runtime_is_refused(order):
    """A process that imports torch maps torch's bundled HIP runtime: imported
    first, libfluidsim binds to it (same soname); imported after, a second
    runtime is mapped beside the one libfluidsim loaded.  The library refuses
    both unless the bundled runtime is the major.minor it was built against
    (and even then refuses two mapped runtimes)."""
    tv = _torch_hip_version()
    if tv is None:
        pytest.skip("no ROCm torch installed")
    built, _ = capi.hip_versions()
    same = tv == (built // 10000000, (built // 100000) % 100)
    r = subprocess.run(["python", "-c", _RUNTIME_PROBE, order, capi.LIB_PATH],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rc, msg = r.stdout.strip().split("\n", 1) if "\n" in r.stdout.strip() else (r.stdout.strip(), "")
    if order == "library_first":
        assert int(rc) == capi.FS_ERR_DEVICE and "HIP runtimes are mapped" in msg
    elif same:
        assert int(rc) == capi.FS_OK
    else:
        assert int(rc) == capi.FS_ERR_DEVICE and "built against HIP" in msg


def test_hip_versions_reported():
    built, runtime = capi.hip_versions()
    assert built // 100000 == runtime // 100000  # this process maps only the build's runtime
