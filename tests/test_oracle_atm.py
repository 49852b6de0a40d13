"""The oracle's D4 ATM term (calcATM / calcATMGradient,
ff_workspace_gfnff.cpp:1526-1658) pinned by mathematics: no reference
fixture reports it (the Fortran decomposition has no bonded-triplet ATM).

  * gradient: central finite differences of the ATM energy alone with the D4
    weights (so C6) held fixed, as the reference's analytic ATM gradient has
    no CN dependence of C6;
  * energy: an independent evaluation of every triplet through explicit
    angle cosines, E = sum c9/3 * scale * fdmp * (3 cos a cos b cos c + 1) /
    (r_ij r_ik r_jk)^3 with C6 = gw_i^T C6ref gw_j and c9 = sqrt|C6ij C6ik C6jk|,
    instead of the oracle's law-of-cosines product;
  * the full evaluation's ATM energy and ATM gradient component equal the
    term alone.
The gradient check settles the sign convention: the reference's -c9 in the
gradient together with its -(dgij + dgik) orientation is +dE/dx.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import BOHR, load_fixture
from curcuma_amd import GFNFFMethod, Mol
from curcuma_amd.workloads import alkane, glycine_peptide


def setup_lib(oracle):
    lib = oracle.lib
    lib.gfnff_oracle_atm.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.gfnff_oracle_cn_gw.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


def atm(lib, v, x, gw=None, grad=False):
    n = v.natoms
    E = C.c_double()
    g = np.zeros(3 * n)
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    st = lib.gfnff_oracle_atm(C.byref(v), xx.ctypes.data, gw.ctypes.data if gw is not None else None, C.byref(E),
                              g.ctypes.data if grad else None)
    assert st == 0
    return E.value, g.reshape(n, 3)


def weights(lib, v, x):
    n = v.natoms
    gw = np.zeros(7 * n)
    cn = np.zeros(n)
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    assert lib.gfnff_oracle_cn_gw(C.byref(v), xx.ctypes.data, cn.ctypes.data, gw.ctypes.data) == 0
    return gw, cn


def molecules():
    fx = load_fixture("polymer")
    xb = np.array(fx["xyz_bohr"])
    yield "polymer", np.array(fx["Z"]), xb * BOHR
    yield "peptide", *glycine_peptide(4, sigma=0.05, seed=2)
    yield "alkane", *alkane(12, sigma=0.05, seed=3)


def independent_atm(v, x, gw):
    """E_ATM from explicit cosines and the D4 C6 contraction (numpy)."""
    n, ne = v.natoms, v.nelem
    elem = np.ctypeslib.as_array(v.elem, (n,))
    nref = np.ctypeslib.as_array(v.nref, (ne,))
    c6ref = np.ctypeslib.as_array(v.c6ref, (ne * ne * 49,)).reshape(ne, ne, 7, 7)
    rcov = np.ctypeslib.as_array(v.rcov_d3, (n,))
    ijk = np.ctypeslib.as_array(v.atm_ijk, (3 * v.natm,)).reshape(-1, 3)
    sc = np.ctypeslib.as_array(v.atm_scale, (v.natm,))
    W = gw.reshape(n, 7).copy()
    for i in range(n):
        W[i, nref[elem[i]]:] = 0.0

    def c6(a, b):
        return np.einsum("ta,tab,tb->t", W[a], c6ref[elem[a], elem[b]], W[b])

    i, j, k = ijk[:, 0], ijk[:, 1], ijk[:, 2]
    c9 = np.sqrt(np.abs(c6(i, j) * c6(i, k) * c6(j, k)))
    rij, rik, rjk = x[j] - x[i], x[k] - x[i], x[k] - x[j]
    dij, dik, djk = (np.linalg.norm(r, axis=1) for r in (rij, rik, rjk))
    cos_i = np.einsum("tc,tc->t", rij, rik) / (dij * dik)
    cos_j = np.einsum("tc,tc->t", -rij, rjk) / (dij * djk)
    cos_k = np.einsum("tc,tc->t", -rik, -rjk) / (dik * djk)
    r0 = lambda a, b: 0.58 * np.sqrt(3.0 * rcov[a] * rcov[b]) + 4.80  # noqa: E731
    fd = 1.0 / (1.0 + 6.0 * (r0(i, j) * r0(i, k) * r0(j, k) / (dij * dik * djk)) ** (14.0 / 3.0))
    ang = (3.0 * cos_i * cos_j * cos_k + 1.0) / (dij * dik * djk) ** 3
    return float(np.sum(c9 / 3.0 * sc * fd * ang))


@pytest.mark.parametrize("case", [0, 1, 2])
def test_atm_gradient_is_fd_of_atm_energy(oracle, case):
    name, Z, xyz = list(molecules())[case]
    lib = setup_lib(oracle)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz))
    v = m.params_view()
    assert v.natm > 0
    x = xyz * (1.0 / BOHR)
    gw, _ = weights(lib, v, x)
    E0, g = atm(lib, v, x, gw, grad=True)
    assert E0 > 0.0  # the bonded-triplet ATM term is repulsive for these geometries
    h = 1e-5
    rng = np.random.default_rng(7)
    atoms = rng.choice(len(Z), min(len(Z), 12), replace=False)
    scale = np.abs(g).max()
    for a in atoms:
        for c in range(3):
            xp = x.copy()
            xp[a, c] += h
            xm = x.copy()
            xm[a, c] -= h
            fd = (atm(lib, v, xp, gw)[0] - atm(lib, v, xm, gw)[0]) / (2 * h)
            # a flipped sign would miss by 2|g|; FD truncation is ~h^2 |E'''|
            assert abs(fd - g[a, c]) <= 1e-6 * scale + 1e-15, (name, a, c, fd, g[a, c])
    # translation invariance of the term itself
    assert np.abs(g.sum(axis=0)).max() <= 1e-12 * max(1.0, scale * len(Z))


@pytest.mark.parametrize("case", [0, 1, 2])
def test_atm_energy_by_explicit_cosines(oracle, case):
    name, Z, xyz = list(molecules())[case]
    lib = setup_lib(oracle)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz))
    v = m.params_view()
    x = xyz * (1.0 / BOHR)
    gw, _ = weights(lib, v, x)
    E = atm(lib, v, x)[0]
    ref = independent_atm(v, x, gw)
    assert abs(E - ref) <= 1e-12 * abs(ref) + 1e-18, (name, E, ref)


def test_full_evaluation_carries_the_atm_term(oracle):
    name, Z, xyz = next(molecules())
    lib = setup_lib(oracle)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz))
    v = m.params_view()
    x = xyz * (1.0 / BOHR)
    r = oracle.eval(v, x, components=True)
    E, g = atm(lib, v, x, grad=True)
    assert r["terms"][10] == pytest.approx(E, rel=1e-14, abs=0)
    assert np.abs(r["comp"]["ATM"] - g).max() <= 1e-18 + 1e-13 * np.abs(g).max()
    # the polymer's ATM energy, reported beside the fixture (which has none)
    assert 7.0e-6 < E < 8.0e-6
