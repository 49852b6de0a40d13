"""The oracle's EEQ residual check (gfnff_oracle_eeq_residual), used by the GPU
tests to state how far an engine solve is from the exact constrained EEQ
solution.  Pinned here on the CPU:
  * the oracle's own direct solve (Cholesky + Schur complement) has a
    residual at rounding level, on one-fragment and many-fragment systems;
  * the reference fixture charges (printed to ~1e-7 e) give a residual of
    that order, not zero and not large;
  * a planted error that keeps every fragment's charge sum (a charge moved
    between two atoms of one fragment) is seen in proportion to its size;
  * a charge added to one fragment is seen by the charge-sum check.
"""
import numpy as np
import pytest

from conftest import BOHR, load_fixture
from curcuma_amd import GFNFFMethod, Mol
from curcuma_amd.workloads import water_cluster


def setup_fixture(name):
    fx = load_fixture(name)
    Z = np.array(fx["Z"])
    xb = np.array(fx["xyz_bohr"])
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xb * BOHR)), m.getErrorMessage()
    return fx, m, (xb * BOHR) * (1.0 / BOHR)


@pytest.mark.parametrize("name", ["caffeine", "complex", "triose"])
def test_residual_of_oracle_solve_is_rounding(name, oracle):
    fx, m, x = setup_fixture(name)
    r = oracle.eval(m.params_view(), x, grad=False)
    rel, qe = oracle.eeq_residual(m.params_view(), x, r["cn"], r["q"])
    print(f"{name}: |r|/|b| = {rel:.2e}, charge-sum error {qe:.2e}")
    assert rel <= 1e-13
    assert qe <= 1e-12
    # the fixture's printed charges (Fortran solve, ~1e-7 e) are close, not exact
    if name != "triose":
        relf, _ = oracle.eeq_residual(m.params_view(), x, r["cn"], np.array(fx["energy_charges"]))
        print(f"{name}: fixture charges |r|/|b| = {relf:.2e}")
        assert 1e-10 < relf < 1e-4


def test_residual_many_fragments_and_planted_errors(oracle):
    Z, xyz = water_cluster(27, seed=5)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    v = m.params_view()
    assert v.nfrag == 27
    x = xyz * (1.0 / BOHR)
    r = oracle.eval(v, x, grad=False)
    rel, qe = oracle.eeq_residual(v, x, r["cn"], r["q"], nthreads=2)
    assert rel <= 1e-13 and qe <= 1e-12
    # charge moved inside fragment 0 (atoms 0 and 1): sums unchanged, residual grows linearly
    rels = []
    for eps in (1e-9, 1e-6):
        q = r["q"].copy()
        q[0] += eps
        q[1] -= eps
        rel_e, qe_e = oracle.eeq_residual(v, x, r["cn"], q)
        assert qe_e <= 1e-12
        rels.append(rel_e)
    assert rels[0] > 1e-11
    assert 300 < rels[1] / rels[0] < 3000
    # charge added to one fragment: the charge-sum check sees it
    q = r["q"].copy()
    q[4] += 1e-7
    _, qe_b = oracle.eeq_residual(v, x, r["cn"], q)
    assert abs(qe_b - 1e-7) < 1e-12
