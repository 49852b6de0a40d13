"""Re-detection of the hydrogen/halogen-bond lists on geometry updates
(updateHBXBIfNeeded + shouldUpdateHBXB, gfnff_method.cpp:1597-1690,
2228-2259): lists are rebuilt from the set-up topology when
sqrt(sum dx^2)/N > 0.3 Bohr since the last detection.  Host-side only."""
import numpy as np

from conftest import BOHR
from curcuma_amd import GFNFFMethod, Mol
from curcuma_amd.workloads import water_cluster


def hb_lists(m):
    v = m.params_view()
    if v.nhb == 0:
        return [], []
    ahb = np.ctypeslib.as_array(v.hb_ahb, (3 * v.nhb,)).reshape(-1, 3).tolist()
    cs = np.ctypeslib.as_array(v.hb_case, (v.nhb,)).tolist()
    return ahb, cs


def test_lists_follow_large_moves_and_match_fresh_detection():
    Z, x0 = water_cluster(2, spacing=2.9, seed=11)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, x0))
    base = hb_lists(m)
    assert len(base[0]) > 0
    # pull the second water 30 A away: sqrt(3 * (30/BOHR)^2) / 6 = 16 Bohr > 0.3
    x1 = x0.copy()
    x1[3:] += np.array([30.0, 0.0, 0.0])
    assert m.updateGeometry(x1)
    far = hb_lists(m)
    fresh = GFNFFMethod()
    assert fresh.setMolecule(Mol(Z, x1))
    assert far == hb_lists(fresh)
    assert len(far[0]) < len(base[0])
    # back again: the original lists return
    assert m.updateGeometry(x0)
    assert hb_lists(m) == base


def test_small_moves_keep_the_lists():
    Z, x0 = water_cluster(2, spacing=2.9, seed=11)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, x0))
    base = hb_lists(m)
    # move the second water far but below the threshold of the reference's
    # sqrt(sum dx^2)/N measure: |dx| per moved atom d, sqrt(3) d / 6 <= 0.3 Bohr
    d = 0.9 * 0.3 * 6.0 / np.sqrt(3.0) * BOHR
    x1 = x0.copy()
    x1[3:] += np.array([d, 0.0, 0.0])
    assert m.updateGeometry(x1)
    assert hb_lists(m) == base
    # the threshold is a parameter: 0 forces a re-detection
    m.setParameters({"hb_update_rmsd_bohr": 0.0})
    x2 = x1.copy()
    x2[3:] += np.array([1e-6, 0.0, 0.0])
    assert m.updateGeometry(x2)
    fresh = GFNFFMethod()
    assert fresh.setMolecule(Mol(Z, x2))
    assert hb_lists(m) == hb_lists(fresh)
