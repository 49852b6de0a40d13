"""The launch-shape autotuner's search and cache (curcuma_amd/autotune.py),
without a GPU: coordinate descent on a synthetic cost, the JSON cache round
trip, and the workload signature."""
import numpy as np

from curcuma_amd import autotune


def test_descent_finds_the_separable_minimum():
    cost = {"a": {0: 3.0, 1: 1.0, 2: 2.0}, "b": {16: 0.5, 64: 0.2}}
    seen = []

    def measure(p):
        seen.append(dict(p))
        return 10.0 + sum(cost[k][p.get(k, 0 if k == "a" else 16)] for k in cost)

    best, timings = autotune.descend({"a": [0, 1, 2], "b": [16, 64]}, measure)
    assert best == {"a": 1, "b": 64}
    assert min(timings.values()) == 10.0 + 1.0 + 0.2
    assert len(seen) == len(timings)  # every configuration timed once


def test_cache_round_trip(tmp_path):
    path = tmp_path / "tune.json"
    sig = autotune.signature(np.array([6, 1, 1, 1, 1]), None, "mixed", 0)
    assert autotune.lookup("dev", sig, path) is None
    autotune.save_cache({f"dev|{sig}": {"params": {"batch_solver": 0}, "timings_ms": {"defaults": 1.0}}}, path)
    hit = autotune.lookup("dev", sig, path)
    assert hit["params"] == {"batch_solver": 0}
    # another workload or device misses
    assert autotune.lookup("other", sig, path) is None
    sig2 = autotune.signature(np.array([6, 1, 1, 1, 1]), np.array([0, 2, 5]), "mixed", 0)
    assert sig2 != sig and autotune.lookup("dev", sig2, path) is None


def test_signature_depends_on_composition_not_order():
    z = np.array([8, 1, 1, 6, 1, 1, 1, 1])
    assert autotune.signature(z, None, "fp64", 0) == autotune.signature(z[::-1], None, "fp64", 0)
    assert autotune.signature(z, None, "fp64", 0) != autotune.signature(z, None, "mixed", 0)


def test_tuned_keys_accepted_by_the_library():
    """Every value the tuner may set, and every engine default it starts
    from, is a legal gfnff_set_param value (the C-ABI validates without a
    device); an out-of-range value is refused."""
    from curcuma_amd import GFNFFMethod
    m = GFNFFMethod()
    lib, h = m._lib, m._h
    for space in (autotune.SPACE_BATCH, autotune.SPACE_SINGLE):
        for k, vals in space.items():
            for v in list(vals) + [autotune.ENGINE_DEFAULTS[k]]:
                assert lib.gfnff_set_param(h, k.encode(), float(v)) == 0, (k, v)
    assert lib.gfnff_set_param(h, b"batch_solver", 7.0) != 0
    assert lib.gfnff_set_param(h, b"batch_fp32", 2.0) != 0
