"""HIP engine against the CPU oracle on identical inputs (same parameter set,
same Bohr coordinates).  Every call goes through the C-ABI.

Tolerances: energies are FP64 on both sides and agree to rounding; the bar
asserted is |dE| <= 1e-9 Eh absolute or 1e-12 relative per term, charges
1e-9 e, gradient components 1e-9 Eh/Bohr (the north-star bar for the FP64
path is |dE| <= 1e-6 Eh/atom and gradient RMS <= 1e-5 Eh/Bohr).
"""
import numpy as np
import pytest

from conftest import BOHR, gpu_available, load_fixture
from curcuma_amd import GFNFFMethod, Mol
from curcuma_amd.workloads import alkane, glycine_peptide, water_cluster
from test_oracle_pi import bromomethane_ammonia, hexadienyne

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)


def run_both(Z, xyz_ang, oracle):
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz_ang)), m.getErrorMessage()
    e = m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    x = np.asarray(xyz_ang) * (1.0 / BOHR)
    ref = oracle.eval(m.params_view(), x)
    return m, e, ref


def check(m, e, ref, n, etol=1e-9, gtol=1e-9):
    print(f"N={n} E_gpu={e:.12f} E_cpu={ref['E']:.12f} dE={e - ref['E']:.3e}")
    assert abs(e - ref["E"]) <= max(etol, 1e-12 * abs(ref["E"]))
    t = m._terms
    for k in range(len(t)):
        assert abs(t[k] - ref["terms"][k]) <= max(etol, 1e-12 * abs(ref["terms"][k])), k
    g = m.getGradient()
    dg = np.abs(g - ref["grad"]).max()
    print(f"  max|dg|={dg:.3e}  max|dq|={np.abs(m.getCharges() - ref['q']).max():.3e}")
    assert dg <= gtol
    assert np.abs(m.getCharges() - ref["q"]).max() <= 1e-9
    assert np.abs(m.getCN() - ref["cn"]).max() <= 1e-12


@pytest.mark.parametrize("name", ["HH", "OH", "H2O", "water", "CH4", "methane", "CH3OH", "CH3OCH3", "HCl", "triose",
                                  "HCN", "O3", "C6H6", "acetic_acid_dimer", "caffeine", "complex", "polymer"])
def test_fixture_molecules(name, oracle):
    """Includes the pi fixtures: inversions (planar and saturated N) and
    hydrogen bonds of cases 3 (carbonyl acceptor) and 4 (sp2 N lone pair)."""
    fx = load_fixture(name)
    Z = np.array(fx["Z"])
    xyz = np.array(fx["xyz_bohr"]) * BOHR
    m, e, ref = run_both(Z, xyz, oracle)
    check(m, e, ref, len(Z))


def test_alkane_c40h82(oracle):
    Z, xyz = alkane(40, sigma=0.02, seed=1)
    m, e, ref = run_both(Z, xyz, oracle)
    check(m, e, ref, len(Z))


def test_glycine_peptide(oracle):
    """Amide pi systems, inversions and intramolecular hydrogen bonds."""
    Z, xyz = glycine_peptide(12, sigma=0.03, seed=2)
    m, e, ref = run_both(Z, xyz, oracle)
    assert m.params_view().ninv == 24
    check(m, e, ref, len(Z))


def test_triple_bond_torsion(oracle):
    Z, xyz = hexadienyne(37.0)
    m, e, ref = run_both(Z, xyz, oracle)
    assert m.params_view().nstors == 1 and abs(m._terms[12]) > 1e-5
    check(m, e, ref, len(Z))


def test_halogen_bond(oracle):
    Z, xyz = bromomethane_ammonia(3.0)
    xyz = xyz + np.random.default_rng(5).normal(0.0, 0.05, xyz.shape)
    m, e, ref = run_both(Z, xyz, oracle)
    assert m.params_view().nxb >= 1 and m._terms[9] < 0.0
    check(m, e, ref, len(Z))


def test_small_water_cluster(oracle):
    """Multi-fragment EEQ (27 waters) with hydrogen bonds of cases 1 and 2."""
    Z, xyz = water_cluster(27, seed=5)
    m, e, ref = run_both(Z, xyz, oracle)
    check(m, e, ref, len(Z))


def test_repeat_is_bitwise_identical():
    """Deterministic reductions: two evaluations give identical bits."""
    Z, xyz = water_cluster(64, seed=7)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz))
    e1 = m.calculateEnergy(True)
    g1 = m.getGradient()
    e2 = m.calculateEnergy(True)
    g2 = m.getGradient()
    assert e1 == e2 and np.array_equal(g1, g2)


def test_fd_gradient_gpu():
    """Central finite differences of the GPU energy (h = 1e-5 Bohr)."""
    Z, xyz = water_cluster(8, seed=9)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz))
    m.calculateEnergy(True)
    g = m.getGradient()
    h = 1e-5 * BOHR
    rng = np.random.default_rng(0)
    for a in rng.choice(len(Z), 6, replace=False):
        for c in range(3):
            xp = xyz.copy()
            xp[a, c] += h
            m.updateGeometry(xp)
            ep = m.calculateEnergy(False)
            xm = xyz.copy()
            xm[a, c] -= h
            m.updateGeometry(xm)
            em = m.calculateEnergy(False)
            fd = (ep - em) / (2 * 1e-5)
            assert abs(fd - g[a, c]) < 1e-7, (a, c, fd, g[a, c])


def brute_cn_lists(xb):
    """Pair builder #1 of the reference (cn_calculator.cpp:392-406): i < j scan,
    d2 = dx*dx + dy*dy + dz*dz (unfused, left to right), strict d2 < 36, j
    appended to i and i to j, so every list is ascending."""
    n = len(xb)
    lists = [[] for _ in range(n)]
    for i in range(n):
        d = xb[i] - xb[i + 1:]
        d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        for j in (np.nonzero(d2 < 36.0)[0] + i + 1):
            lists[i].append(int(j))
            lists[int(j)].append(i)
    return lists


@pytest.mark.parametrize("case", ["water64", "alkane40", "water3000", "bundle16"])
def test_cn_neighbour_lists_bit_exact(case):
    """water3000 (9000 atoms, 36 partner tiles) and a 16-chain alkane bundle
    (3872 atoms) run the tile-box pruning of the CN kernel; the two small
    cases fit one partner tile."""
    from curcuma_amd._lib import as_i32, iptr, load_library
    from curcuma_amd.workloads import alkane_bundle
    if case == "water64":
        Z, xyz = water_cluster(64, seed=11)
    elif case == "water3000":
        Z, xyz = water_cluster(3000, seed=2)
    elif case == "bundle16":
        Z, xyz = alkane_bundle(16, 80, seed=4)
    else:
        Z, xyz = alkane(40, sigma=0.02, seed=1)
    print(f"{case}: {len(Z)} atoms", flush=True)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
    m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    n = len(Z)
    maxnb = 128
    counts = as_i32(np.zeros(n))
    idx = as_i32(np.zeros(n * maxnb))
    assert load_library().gfnff_cn_neighbours(m.handle, iptr(counts), iptr(idx), maxnb) == 0
    ref = brute_cn_lists(np.asarray(xyz) * (1.0 / BOHR))
    total = 0
    for i in range(n):
        got = list(idx[i * maxnb:i * maxnb + counts[i]])
        assert got == ref[i], (i, got, ref[i])
        total += len(got)
    print(f"{case}: {n} atoms, {total // 2} pairs, lists identical")


def test_cholesky_lookahead_is_bitwise_identical():
    """The lookahead option only reorders independent tile updates across two
    streams; every tile sees the same operands, so results are bitwise equal."""
    Z, xyz = water_cluster(512, seed=3)
    out = []
    for la in (0, 1):
        m = GFNFFMethod(parameters={"cholesky_lookahead": la, "eeq_solver": 1})
        assert m.setMolecule(Mol(Z, xyz)), m.getErrorMessage()
        e = m.calculateEnergy(True)
        assert not m.hasError(), m.getErrorMessage()
        out.append((e, m.getGradient().copy(), m.getCharges().copy()))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def test_hb_lists_refreshed_on_large_moves(oracle):
    """After a geometry update beyond the re-detection threshold the device
    engine evaluates the re-detected hydrogen-bond lists: same energy and
    gradient as a fresh handle at that geometry, and as the oracle on the
    refreshed parameters."""
    Z, x0 = water_cluster(8, seed=12)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, x0))
    m.calculateEnergy(True)
    x1 = x0.copy()
    x1[12:] += np.array([12.0, 0.0, 0.0])
    assert m.updateGeometry(x1)
    e = m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    f = GFNFFMethod()
    assert f.setMolecule(Mol(Z, x1))
    ef = f.calculateEnergy(True)
    assert abs(e - ef) <= 1e-9 * max(1.0, abs(ef))
    assert np.abs(m.getGradient() - f.getGradient()).max() <= 1e-9
    ref = oracle.eval(m.params_view(), x1 * (1.0 / BOHR))
    assert abs(e - ref["E"]) <= 1e-9 * max(1.0, abs(ref["E"]))


@pytest.mark.parametrize("solver", [0, 1])
def test_eeq_solvers_agree(solver, oracle):
    """The iterative EEQ solve (default) and the blocked Cholesky give the
    oracle's charges and energy; PCG converges in few iterations on a water
    cluster (reduced matrix kappa ~ 6)."""
    Z, xyz = water_cluster(216, seed=9)
    m = GFNFFMethod()
    m.setParameters({"eeq_solver": solver})
    assert m.setMolecule(Mol(Z, xyz))
    e = m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    ref = oracle.eval(m.params_view(), np.asarray(xyz) * (1.0 / BOHR))
    check(m, e, ref, len(Z))
    its = m.solverIterations()
    print(f"solver {solver}: iterations {its}")
    if solver == 0:
        assert 0 < its < 60
    else:
        assert its == 0


def test_pcg_on_chain_molecule(oracle):
    """Single-fragment chains are less well conditioned (kappa of the reduced
    matrix ~1e3): PCG still converges to the oracle's charges."""
    Z, xyz = alkane(120, sigma=0.02, seed=4)
    m, e, ref = run_both(Z, xyz, oracle)
    check(m, e, ref, len(Z))
    assert 0 < m.solverIterations() < 400


def test_batch_cn_neighbour_lists_bit_exact():
    """The per-molecule CN kernel of batches (k_cn_mol): every molecule's lists
    equal the reference's builder run on that molecule alone (global atom
    indices)."""
    from curcuma_amd import GFNFFBatch
    from curcuma_amd._lib import as_i32, iptr, load_library
    from curcuma_amd.workloads import mixed_batch
    off, Z, xyz = mixed_batch(48, 20, 400, sigma=0.03, seed=3)
    m = GFNFFBatch()
    assert m.setBatch(off, Z, xyz), m.getErrorMessage()
    m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    n = len(Z)
    maxnb = 128
    counts = as_i32(np.zeros(n))
    idx = as_i32(np.zeros(n * maxnb))
    assert load_library().gfnff_cn_neighbours(m.handle, iptr(counts), iptr(idx), maxnb) == 0
    xb = np.asarray(xyz) * (1.0 / BOHR)
    total = 0
    for k in range(len(off) - 1):
        a0, a1 = int(off[k]), int(off[k + 1])
        ref = brute_cn_lists(xb[a0:a1])
        for i in range(a0, a1):
            got = list(idx[i * maxnb:i * maxnb + counts[i]])
            assert got == [a0 + j for j in ref[i - a0]], (k, i)
            total += len(got)
    print(f"batch of {len(off) - 1} molecules, {n} atoms, {total // 2} pairs, lists identical")


@pytest.mark.parametrize("ion,charge", [("H3O+", 1), ("OH-", -1)])
def test_charged_two_fragments(ion, charge, oracle):
    """A charged two-fragment system (charge placed by the lower EEQ energy
    of both placements) on the GPU against the oracle on the same parameters."""
    from test_abi import _ion_water
    Z, xyz, _ = _ion_water(ion, True)
    m = GFNFFMethod()
    assert m.setMolecule(Mol(Z, xyz, charge=charge)), m.getErrorMessage()
    e = m.calculateEnergy(True)
    assert not m.hasError(), m.getErrorMessage()
    ref = oracle.eval(m.params_view(), np.asarray(xyz) * (1.0 / BOHR))
    check(m, e, ref, len(Z))
    assert abs(m.getCharges().sum() - charge) < 1e-10
