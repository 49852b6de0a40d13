import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np
import pytest

try:  # loaded before the engine library, so that tests handing torch device memory to it share one HIP runtime
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

from curcuma_amd._lib import NTERMS, ParamsView  # noqa: E402

FIXTURES = ROOT / "tests" / "golden" / "fixtures"
ORACLE_LIB = ROOT / "oracle" / "_build" / "liboracle.so"
BOHR = 0.529177210903

# per-component gradient slots written by the oracle (oracle/gfnff_oracle.c)
GCOMP = ["Bond", "Angle", "Torsion", "Repulsion", "NBRep", "ES", "Disp", "HB", "XB", "BATM", "ATM"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def load_fixture(name: str) -> dict:
    return json.loads((FIXTURES / f"{name}.json").read_text())


class Oracle:
    """ctypes wrapper of the CPU oracle (test infrastructure only)."""

    def __init__(self):
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"oracle not built: {ORACLE_LIB}")
        self.lib = C.CDLL(str(ORACLE_LIB))
        self.lib.gfnff_oracle_eval.argtypes = [C.POINTER(ParamsView), C.c_void_p, C.c_int, C.c_void_p,
                                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]

    def eval(self, view: ParamsView, xyz_bohr: np.ndarray, grad: bool = True, components: bool = False):
        n = view.natoms
        x = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        E = C.c_double()
        terms = np.zeros(NTERMS)
        g = np.zeros(3 * n)
        q = np.zeros(n)
        cn = np.zeros(n)
        gc = np.zeros(len(GCOMP) * 3 * n) if components else None
        st = self.lib.gfnff_oracle_eval(C.byref(view), x.ctypes.data, 1 if grad else 0, C.byref(E),
                                        terms.ctypes.data, g.ctypes.data, q.ctypes.data, cn.ctypes.data,
                                        gc.ctypes.data if components else None)
        if st != 0:
            raise RuntimeError(f"oracle status {st}")
        out = {"E": E.value, "terms": terms, "grad": g.reshape(n, 3), "q": q, "cn": cn}
        if components:
            out["comp"] = {k: gc.reshape(len(GCOMP), n, 3)[i] for i, k in enumerate(GCOMP)}
        return out

    def eval_mt(self, view: ParamsView, xyz_bohr: np.ndarray, nthreads: int, grad: bool = True, q_in=None):
        """Threaded oracle (FFWorkspace scheme); q_in replaces the EEQ solve."""
        lib = self.lib
        lib.gfnff_oracle_eval_mt.argtypes = [C.POINTER(ParamsView), C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        n = view.natoms
        x = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        qi = None if q_in is None else np.ascontiguousarray(q_in, dtype=np.float64)
        E = C.c_double()
        terms = np.zeros(NTERMS)
        g = np.zeros(3 * n)
        q = np.zeros(n)
        cn = np.zeros(n)
        st = lib.gfnff_oracle_eval_mt(C.byref(view), x.ctypes.data, 1 if grad else 0, int(nthreads),
                                      qi.ctypes.data if qi is not None else None, C.byref(E), terms.ctypes.data,
                                      g.ctypes.data, q.ctypes.data, cn.ctypes.data)
        if st != 0:
            raise RuntimeError(f"oracle status {st}")
        return {"E": E.value, "terms": terms, "grad": g.reshape(n, 3), "q": q, "cn": cn}


    def eval_alpb(self, view: ParamsView, sview, xyz_bohr: np.ndarray, grad: bool = True, nthreads: int = 1):
        """GFN-FF with ALPB solvation (sview: the engine's AlpbView)."""
        from curcuma_amd._lib import AlpbView
        lib = self.lib
        lib.gfnff_oracle_eval_alpb.argtypes = [C.POINTER(ParamsView), C.POINTER(AlpbView), C.c_void_p, C.c_int,
                                               C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        n = view.natoms
        x = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        E = C.c_double()
        terms = np.zeros(NTERMS)
        g = np.zeros(3 * n)
        q = np.zeros(n)
        parts = np.zeros(4)
        st = lib.gfnff_oracle_eval_alpb(C.byref(view), C.byref(sview), x.ctypes.data, 1 if grad else 0,
                                        int(nthreads), C.byref(E), terms.ctypes.data, g.ctypes.data, q.ctypes.data,
                                        parts.ctypes.data)
        if st != 0:
            raise RuntimeError(f"oracle status {st}")
        return {"E": E.value, "terms": terms, "grad": g.reshape(n, 3), "q": q, "parts": parts}

    def eeq_residual(self, view: ParamsView, xyz_bohr: np.ndarray, cn: np.ndarray, q: np.ndarray, nthreads: int = 1):
        """(|r|/|b| of the constrained EEQ system at charges q, max fragment
        charge-sum error); cn: the evaluation's log-capped CN."""
        lib = self.lib
        lib.gfnff_oracle_eeq_residual.argtypes = [C.POINTER(ParamsView), C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        x = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        cn = np.ascontiguousarray(cn, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        rel, qe = C.c_double(), C.c_double()
        st = lib.gfnff_oracle_eeq_residual(C.byref(view), x.ctypes.data, cn.ctypes.data, q.ctypes.data,
                                           int(nthreads), C.byref(rel), C.byref(qe))
        if st != 0:
            raise RuntimeError(f"oracle status {st}")
        return rel.value, qe.value


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


def gpu_available() -> bool:
    try:
        from curcuma_amd._lib import load_library
        return load_library().gfnff_device_count() > 0
    except Exception:
        return False


class HipRuntime:
    """Minimal ctypes access to the HIP runtime the engine library links
    (device buffers for the device-resident entry points in GPU tests)."""

    def __init__(self):
        from curcuma_amd._lib import load_library
        load_library()  # the engine's runtime is the one resolved below
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipStreamDestroy.argtypes = [C.c_void_p]
        self.rt.hipStreamSynchronize.argtypes = [C.c_void_p]

    def stream(self) -> C.c_void_p:
        s = C.c_void_p()
        assert self.rt.hipStreamCreate(C.byref(s)) == 0
        return s

    def upload(self, a: np.ndarray) -> C.c_void_p:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p

    def alloc(self, nbytes: int) -> C.c_void_p:
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(nbytes, 8)) == 0
        return p

    def download(self, p: C.c_void_p, shape, dtype=np.float64) -> np.ndarray:
        out = np.zeros(shape, dtype=dtype)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self, p: C.c_void_p) -> None:
        self.rt.hipFree(p)
