"""HIP-graph capture and replay of the device-resident evaluation
(gfnff_graph_capture / gfnff_graph_launch) against eager evaluations.

Each replay is checked against an eager evaluation of the same coordinates,
with eager work of the program and synchronisations between the replays
(replays after the first are where recorded memory nodes have gone wrong on
this hardware; the evaluation records none).  The kernels sum in a fixed
order, so replay and eager results are bitwise equal.  Covered: a batch of
molecules on the mixed path and a single molecule on the PCG path (the
persistent solver under graph launch), coordinates changed between
replays, and a capture dropped when the engine is rebuilt."""
import ctypes as C

import numpy as np
import pytest

from conftest import BOHR, HipRuntime, gpu_available
from curcuma_amd import GFNFFBatch, GFNFFMethod, Mol
from curcuma_amd._lib import NTERMS
from curcuma_amd.workloads import mixed_batch, water_cluster

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)


def run_case(m, hip, x0):
    n = len(x0)
    rng = np.random.default_rng(12)
    xs = [x0, x0 + rng.normal(0.0, 0.01, x0.shape), x0 + rng.normal(0.0, 0.01, x0.shape)]
    s = hip.stream()
    xg = hip.upload(np.ascontiguousarray(x0))
    eg, gg = hip.alloc(8 * (NTERMS + 1)), hip.alloc(8 * 3 * n)
    xe = hip.upload(np.ascontiguousarray(x0))
    ee, ge = hip.alloc(8 * (NTERMS + 1)), hip.alloc(8 * 3 * n)
    try:
        m.energyGradientDevice(xe.value, ee.value, ge.value, s.value)  # engine set-up before the capture
        m.captureGraph(xg.value, eg.value, gg.value, s.value)
        for rep in range(6):
            x = np.ascontiguousarray(xs[rep % 3])
            print(f"replay {rep}", flush=True)
            assert hip.rt.hipMemcpy(xg, x.ctypes.data, x.nbytes, 1) == 0
            assert hip.rt.hipMemcpy(xe, x.ctypes.data, x.nbytes, 1) == 0
            m.launchGraph(s.value)
            assert hip.rt.hipStreamSynchronize(s) == 0
            e_graph, g_graph = hip.download(eg, (NTERMS + 1,)), hip.download(gg, (n, 3))
            m.energyGradientDevice(xe.value, ee.value, ge.value, s.value)  # eager work between replays
            assert hip.rt.hipStreamSynchronize(s) == 0
            e_eager, g_eager = hip.download(ee, (NTERMS + 1,)), hip.download(ge, (n, 3))
            assert np.isfinite(e_graph[0])
            assert np.array_equal(e_graph, e_eager), (rep, e_graph[0], e_eager[0])
            assert np.array_equal(g_graph, g_eager), rep
        assert m.deviceStatus() == 0
    finally:
        for p in (xg, eg, gg, xe, ee, ge):
            hip.free(p)
        hip.rt.hipStreamDestroy(s)


def test_graph_replay_batch_mixed():
    off, Z, xyz = mixed_batch(24, 20, 200, sigma=0.03, seed=21)
    b = GFNFFBatch(precision="mixed")
    assert b.setBatch(off, Z, xyz), b.getErrorMessage()
    run_case(b, HipRuntime(), xyz * (1.0 / BOHR))


def test_graph_replay_single_pcg_and_rebuild():
    Z, xyz = water_cluster(64, seed=11)
    m = GFNFFMethod(precision="fp64")
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    hip = HipRuntime()
    run_case(m, hip, xyz * (1.0 / BOHR))
    # an option change rebuilds the engine: the capture is gone
    m.setParameters({"gather_lanes": 64})
    with pytest.raises(RuntimeError):
        m.launchGraph(None)
