"""Kernel-level GPU tests: the blocked FP64 Cholesky (MFMA trailing update,
single-wavefront diagonal blocks) against numpy on random SPD matrices."""
import numpy as np
import pytest

from conftest import gpu_available
from curcuma_amd._lib import fptr, load_library

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)


@pytest.mark.parametrize("n", [5, 64, 65, 128, 200, 333])
def test_dense_cholesky(n):
    rng = np.random.default_rng(n)
    M = rng.normal(size=(n, n))
    A = M @ M.T + n * np.eye(n)
    L = np.zeros((n, n))
    st = load_library().gfnff_dense_cholesky(n, fptr(np.ascontiguousarray(A)), fptr(L))
    assert st == 0
    Lref = np.linalg.cholesky(A)
    err = np.abs(L - Lref).max() / np.abs(Lref).max()
    print(f"n={n} rel err {err:.2e}")
    assert err < 1e-12


def test_not_spd_is_flagged():
    A = np.eye(70)
    A[40, 40] = -1.0
    L = np.zeros((70, 70))
    assert load_library().gfnff_dense_cholesky(70, fptr(A), fptr(L)) == 3


@pytest.mark.parametrize("n", [64, 200, 333, 700])
def test_dense_cholesky_f32(n):
    """FP32 MFMA factorisation of the mixed-precision path: error at the level
    FP32 arithmetic allows for a well-conditioned SPD matrix."""
    rng = np.random.default_rng(100 + n)
    M = rng.normal(size=(n, n))
    A = M @ M.T + n * np.eye(n)
    A32 = np.ascontiguousarray(A, dtype=np.float32)
    L = np.zeros((n, n), dtype=np.float32)
    st = load_library().gfnff_dense_cholesky_f32(n, A32.ctypes.data, L.ctypes.data)
    assert st == 0
    Lref = np.linalg.cholesky(A32.astype(np.float64))
    err = np.abs(L - Lref).max() / np.abs(Lref).max()
    print(f"f32 n={n} rel err {err:.2e}")
    assert err < 2e-5
