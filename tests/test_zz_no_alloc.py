"""The allocation-free dense Cholesky (gfnff_dense_cholesky_noalloc) against
the allocating entry point it varies (gfnff_dense_cholesky): same kernels on
the same input, so the factors must be bitwise equal.  Two calls in a row,
the second smaller, exercise the reuse of the shared scratch array; a
matrix that is not positive definite must report GFNFF_STATUS_EEQ in the
device status word; sizes beyond the maximum are refused before anything is
queued."""
import ctypes as C

import numpy as np
import pytest

from conftest import HipRuntime, gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

STATUS_ARGS, STATUS_EEQ = 1, 3
MAX_N = 1024


def spd(n, seed):
    """Symmetric, diagonally dominant (SPD) test matrix, seeded."""
    rng = np.random.default_rng(seed)
    b = rng.normal(size=(n, n))
    return np.ascontiguousarray(b @ b.T / n + np.eye(n))


def reference_factor(lib, a):
    n = a.shape[0]
    out = np.zeros_like(a)
    st = lib.gfnff_dense_cholesky(n, a.ctypes.data_as(C.POINTER(C.c_double)),
                                  out.ctypes.data_as(C.POINTER(C.c_double)))
    return st, out


def noalloc_factor(lib, hip, a):
    n = a.shape[0]
    a_dev = hip.upload(a)
    l_dev = hip.alloc(8 * n * n)
    s_dev = hip.upload(np.array([-1], dtype=np.int32))
    try:
        st = lib.gfnff_dense_cholesky_noalloc(n, a_dev, l_dev, s_dev, None)
        assert st == 0
        L = hip.download(l_dev, (n, n))
        status = int(hip.download(s_dev, (1,), np.int32)[0])
    finally:
        for p in (a_dev, l_dev, s_dev):
            hip.free(p)
    return status, L


def test_noalloc_cholesky_matches_allocating_entry():
    from curcuma_amd._lib import load_library
    lib = load_library()
    hip = HipRuntime()
    for n, seed in ((300, 1), (130, 2)):  # second call smaller: the scratch holds the first call's data
        a = spd(n, seed)
        print(f"noalloc cholesky n={n}", flush=True)
        st_ref, L_ref = reference_factor(lib, a)
        assert st_ref == 0
        status, L = noalloc_factor(lib, hip, a)
        assert status == 0
        assert np.array_equal(L, L_ref)
        assert np.abs(L @ L.T - a).max() <= 1e-12 * np.abs(a).max()


def test_noalloc_cholesky_not_spd_and_limits():
    from curcuma_amd._lib import load_library
    lib = load_library()
    hip = HipRuntime()
    a = spd(96, 3)
    a[50, 50] = -5.0
    print("noalloc cholesky: indefinite matrix", flush=True)
    st_ref, _ = reference_factor(lib, a)
    assert st_ref == STATUS_EEQ
    status, _ = noalloc_factor(lib, hip, a)
    assert status == STATUS_EEQ
    # an SPD call afterwards is clean again (the status word is cleared per call)
    b = spd(64, 4)
    status, L = noalloc_factor(lib, hip, b)
    st_ref, L_ref = reference_factor(lib, b)
    assert status == 0 and st_ref == 0 and np.array_equal(L, L_ref)
    # beyond the maximum: refused, nothing queued (no device pointer is touched)
    dummy = C.c_void_p(16)
    assert lib.gfnff_dense_cholesky_noalloc(MAX_N + 1, dummy, dummy, dummy, None) == STATUS_ARGS
    assert lib.gfnff_dense_cholesky_noalloc(0, dummy, dummy, dummy, None) == STATUS_ARGS
    assert lib.gfnff_dense_cholesky_noalloc(10, None, dummy, dummy, None) == STATUS_ARGS
