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
