"""The C-ABI library loads and exports every entry point include/*.h declares (CPU only: no device
calls), and the Python mirror binds each of them."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^[A-Za-z_][\w \*]*?\b(gfe_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


@pytest.fixture(scope="module")
def libgfe():
    import genefer22_amd
    if not os.path.exists(genefer22_amd.LIB_PATH):
        from genefer22_amd import build
        build.build()
    return ctypes.CDLL(genefer22_amd.LIB_PATH)


def test_headers_declare_entry_points():
    names = declared()
    for must in ("gfe_create", "gfe_square_dup", "gfe_mul", "gfe_init_multiplicand", "gfe_copy", "gfe_set",
                 "gfe_square_mul", "gfe_get_digits", "gfe_set_digits", "gfe_save", "gfe_load",
                 "gfe_batch_create", "gfe_batch_submit", "gfe_run_quick"):
        assert must in names


def test_library_exports_every_declared_symbol(libgfe):
    missing = [n for n in declared() if not hasattr(libgfe, n)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol():
    import genefer22_amd
    bound = {name for name, _, _ in genefer22_amd.ABI}
    assert set(declared()) <= bound, sorted(set(declared()) - bound)


def test_library_is_gfx950_code(libgfe):
    import genefer22_amd
    out = subprocess.run(["strings", genefer22_amd.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_no_device_no_fallback(libgfe):
    """Without a HIP device, creating an engine fails loudly (no CPU fallback)."""
    import genefer22_amd
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(genefer22_amd.GfeError):
        genefer22_amd.Transform(1_000_000_000, 12)


def test_host_residue_function(libgfe):
    """gfe_residue (the product's gint, host code) matches the oracle on random digit vectors."""
    import numpy as np
    import genefer22_amd
    import oracle
    rs = np.random.RandomState(5)
    for n, b in ((12, 1026), (12, 999_999_998), (13, 2_000_000_000)):
        d = rs.randint(-(b // 2), b // 2, 1 << n).astype(np.int32)
        mine = genefer22_amd.residue(d, n, b, canonical=True)
        want = oracle.residue(d, n, b)
        assert (mine["res64"], mine["old64"], mine["hash64"]) == (want.res64, want.old64, want.hash64)
        assert np.array_equal(mine["digits"], oracle.unbalance(d, n, b))


def test_argument_errors_need_no_device(libgfe):
    """Out-of-range candidates and malformed arguments are GFE_EINVAL (-1) with a message, before any
    device work, on every entry point that takes them."""
    import genefer22_amd as G
    L = G.lib()
    h = ctypes.c_void_p()
    r = G.Result()
    cases = [
        lambda: L.gfe_create(1026, 11, 3, 0, ctypes.byref(h)),
        lambda: L.gfe_create(1026, 24, 3, 0, ctypes.byref(h)),
        lambda: L.gfe_create(1022, 12, 3, 0, ctypes.byref(h)),
        lambda: L.gfe_create(2000000002, 12, 3, 0, ctypes.byref(h)),
        lambda: L.gfe_batch_create(11, 4, 0, ctypes.byref(h)),
        lambda: L.gfe_batch_create(15, 0, 0, ctypes.byref(h)),
        lambda: L.gfe_run_quick(1026, 24, 0, ctypes.byref(r)),
        lambda: L.gfe_run_prp(1000, 12, 0, 10, 0, None, ctypes.byref(r)),
        lambda: L.gfe_run_prp_trace(1026, 12, 0, 10, 0, None, None, ctypes.byref(r)),
        lambda: L.gfe_run_prp_resumable(1026, 12, 0, None, 0, 0, None, ctypes.byref(r)),
        lambda: L.gfe_run_prp_resumable(1026, 25, 0, b"/tmp/x.ctx", 0, 0, None, ctypes.byref(r)),
        lambda: L.gfe_run_proof(1026, 12, 0, 0, None, ctypes.byref(r)),
        lambda: L.gfe_write_proof(1026, 12, 0, 3, b"/tmp/x", 4, None, ctypes.byref(r)),
        lambda: L.gfe_write_proof(900, 12, 0, 3, b"/tmp/x", 0, None, ctypes.byref(r)),
    ]
    for i, call in enumerate(cases):
        assert call() == -1, i
        assert L.gfe_last_error(), i
