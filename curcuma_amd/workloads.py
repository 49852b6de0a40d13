"""Seeded synthetic molecules for the benchmark and the parity tests.

Recipes (all coordinates in Angstrom):
  * n-alkane CnH(2n+2), all-trans, ideal geometry (C-C 1.53, C-H 1.09,
    tetrahedral angles), Gaussian coordinate noise.
  * water cluster: the first ``nmol`` sites of a cubic lattice filled in
    lexicographic (i, j, k) order, O-H 0.96, H-O-H 104.5 deg, uniform random
    rigid orientations, uniform jitter of the O position.
  * parallel alkane bundle: chains on a hexagonal lattice.
  * glycine oligopeptide H-(Gly)n-OH in the fully extended conformation
    (all backbone dihedrals 180 deg, planar zig-zag backbone): N-CA 1.46,
    CA-C 1.52, C-N 1.33, C=O 1.23, C-OH 1.34, N-H 1.01, O-H 0.97, CA-H 1.09,
    angles N-CA-C 111, CA-C-N 116, C-N-CA 122 deg; carbonyl O and amide H
    in the backbone plane on the outer bisector, CA and terminal NH2
    hydrogens tetrahedral.  7n + 3 atoms.
Random numbers come from numpy's PCG64 generator seeded per workload.
"""
from __future__ import annotations

import math

import numpy as np

TETRA = math.acos(-1.0 / 3.0)  # 109.47 deg


def alkane(nc: int, sigma: float = 0.0, seed: int = 0, cc: float = 1.53, ch: float = 1.09):
    """All-trans n-alkane C_nH_(2n+2); returns (Z, xyz[A])."""
    half = TETRA / 2.0
    dx, dy = cc * math.sin(half), cc * math.cos(half)
    # backbone with one virtual atom on each end for the terminal methyl hydrogens
    back = np.array([[(k) * dx, (k % 2) * dy, 0.0] for k in range(-1, nc + 1)])
    carb = back[1:-1]
    Z = [6] * nc
    H = []
    zax = np.array([0.0, 0.0, 1.0])
    ca, sa = math.cos(TETRA / 2.0), math.sin(TETRA / 2.0)
    for k in range(nc):
        c = back[k + 1]
        prv, nxt = back[k], back[k + 2]
        b = 2.0 * c - prv - nxt
        b /= np.linalg.norm(b)
        H.append(c + ch * (ca * b + sa * zax))
        H.append(c + ch * (ca * b - sa * zax))
        if k == 0:
            u = prv - c
            H.append(c + ch * u / np.linalg.norm(u))
        if k == nc - 1:
            u = nxt - c
            H.append(c + ch * u / np.linalg.norm(u))
    xyz = np.vstack([carb, np.array(H)])
    Z = np.array(Z + [1] * len(H), dtype=np.int32)
    if sigma > 0:
        rng = np.random.default_rng(seed)
        xyz = xyz + rng.normal(0.0, sigma, xyz.shape)
    return Z, xyz


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_cluster(nmol: int = 3000, spacing: float = 3.1, jitter: float = 0.05, seed: int = 2, oh: float = 0.96,
                  hoh_deg: float = 104.5):
    """Non-periodic water cluster on a cubic lattice; returns (Z, xyz[A])."""
    side = int(math.ceil(nmol ** (1.0 / 3.0) - 1e-9))
    rng = np.random.default_rng(seed)
    half = math.radians(hoh_deg) / 2.0
    local = np.array([[0.0, 0.0, 0.0], [oh * math.sin(half), oh * math.cos(half), 0.0],
                      [-oh * math.sin(half), oh * math.cos(half), 0.0]])
    Z, xyz = [], []
    count = 0
    for i in range(side):
        for j in range(side):
            for k in range(side):
                if count == nmol:
                    break
                site = np.array([i, j, k], dtype=float) * spacing + rng.uniform(-jitter, jitter, 3)
                R = _random_rotation(rng)
                for a, zz in zip(local, (8, 1, 1)):
                    xyz.append(site + R @ a)
                    Z.append(zz)
                count += 1
    return np.array(Z, dtype=np.int32), np.array(xyz)


def alkane_bundle(nchains: int = 80, nc: int = 80, a: float = 4.5, sigma: float = 0.02, seed: int = 4):
    """Parallel all-trans alkanes on the first nchains sites of a hexagonal lattice."""
    Z0, x0 = alkane(nc)
    rng = np.random.default_rng(seed)
    side = int(math.ceil(math.sqrt(nchains)))
    Z, xyz = [], []
    count = 0
    for r in range(side):
        for c in range(side):
            if count == nchains:
                break
            off = np.array([0.0, (c + 0.5 * (r % 2)) * a, r * a * math.sqrt(3.0) / 2.0])
            xyz.append(x0 + off)
            Z.append(Z0)
            count += 1
    xyz = np.vstack(xyz)
    xyz = xyz + rng.normal(0.0, sigma, xyz.shape)
    return np.concatenate(Z).astype(np.int32), xyz


def alkane_batch(nmol: int = 4096, nmin: int = 20, nmax: int = 400, sigma: float = 0.03, seed: int = 3):
    """Batch of independent all-trans n-alkanes: atom counts drawn uniformly in
    [nmin, nmax] and rounded down to the nearest CnH(2n+2) (3n + 2 atoms),
    ideal geometry plus Gaussian noise.  Returns (atom_off, Z, xyz[A])."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=nmol)
    Zs, xs, off = [], [], [0]
    for k, na in enumerate(sizes):
        nc = max(1, (int(na) - 2) // 3)
        Z, xyz = alkane(nc, sigma=sigma, seed=int(seed * 1_000_003 + k))
        Zs.append(Z)
        xs.append(xyz)
        off.append(off[-1] + len(Z))
    return np.asarray(off, dtype=np.int32), np.concatenate(Zs).astype(np.int32), np.vstack(xs)


def _tetra_pair(c, a, b, length):
    """Two tetrahedral substituents on atom c whose other neighbours are a and b."""
    u, v = a - c, b - c
    u, v = u / np.linalg.norm(u), v / np.linalg.norm(v)
    bis = -(u + v)
    bis /= np.linalg.norm(bis)
    nrm = np.cross(u, v)
    nrm /= np.linalg.norm(nrm)
    h = TETRA / 2.0
    return [c + length * (math.cos(h) * bis + math.sin(h) * nrm),
            c + length * (math.cos(h) * bis - math.sin(h) * nrm)]


def _trigonal(c, a, b, length):
    """In-plane substituent of the sp2 atom c on the outer bisector of a-c-b."""
    u, v = a - c, b - c
    d = -(u / np.linalg.norm(u) + v / np.linalg.norm(v))
    return c + length * d / np.linalg.norm(d)


def glycine_peptide(nres: int, sigma: float = 0.0, seed: int = 0):
    """H-(Gly)n-OH, fully extended; returns (Z, xyz[A]) with 7n + 3 atoms."""
    bonds = [1.46, 1.52, 1.33]                   # N-CA, CA-C, C-N(next)
    angles = [math.radians(a) for a in (122.0, 111.0, 116.0)]  # at N, CA, C
    nb = 3 * nres
    pos = [np.zeros(3)]
    heading = 0.0
    for k in range(1, nb + 1):  # the extra position is the C-terminal hydroxyl O
        L = 1.34 if k == nb else bonds[(k - 1) % 3]
        pos.append(pos[-1] + L * np.array([math.cos(heading), math.sin(heading), 0.0]))
        if k < nb:
            turn = math.pi - angles[k % 3]
            heading += turn if k % 2 else -turn
    pos = np.array(pos)
    Z, xyz = [], []
    for r in range(nres):
        N, CA, C = pos[3 * r], pos[3 * r + 1], pos[3 * r + 2]
        nxt = pos[3 * r + 3]
        Z += [7, 6, 6, 8]
        O = _trigonal(C, CA, nxt, 1.23)
        xyz += [N, CA, C, O]
        if r == 0:
            # terminal NH2: two tetrahedral hydrogens around the N-CA axis
            a = (CA - N) / np.linalg.norm(CA - N)
            e1 = np.cross(a, [0.0, 0.0, 1.0])
            e1 /= np.linalg.norm(e1)
            e2 = np.cross(a, e1)
            th = math.pi - TETRA
            for ph in (math.radians(60.0), math.radians(-60.0)):
                d = -a * math.cos(th) + (math.cos(ph) * e1 + math.sin(ph) * e2) * math.sin(th)
                Z.append(1)
                xyz.append(N + 1.01 * d)
        else:
            Z.append(1)
            xyz.append(_trigonal(N, pos[3 * r - 1], CA, 1.01))
        Z += [1, 1]
        xyz += _tetra_pair(CA, N, C, 1.09)
    OH = pos[nb]
    C = pos[nb - 1]
    Z += [8, 1]
    xyz += [OH, OH + 0.97 * _oh_dir(C, pos[nb - 2], OH)]
    xyz = np.array(xyz)
    if sigma > 0:
        rng = np.random.default_rng(seed)
        xyz = xyz + rng.normal(0.0, sigma, xyz.shape)
    return np.array(Z, dtype=np.int32), xyz


def _oh_dir(C, CA, O):
    """Hydroxyl H direction: in the carboxyl plane, anti to CA about the C-O bond
    rotated to a C-O-H angle of 107 deg."""
    a = (C - O) / np.linalg.norm(C - O)
    w = CA - C
    w = w - np.dot(w, a) * a
    w = -w / np.linalg.norm(w)
    th = math.radians(107.0)
    d = math.cos(th) * a + math.sin(th) * w
    return d / np.linalg.norm(d)


def mixed_batch(nmol: int = 4096, nmin: int = 20, nmax: int = 400, sigma: float = 0.03, seed: int = 3):
    """Batch of independent n-alkanes and glycine oligopeptides: per molecule a
    seeded coin picks the family and an atom count drawn uniformly in
    [nmin, nmax] is rounded down to that family's nearest size (alkane 3n + 2,
    peptide 7n + 3), ideal extended geometry plus Gaussian noise.  Returns
    (atom_off, Z, xyz[A])."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=nmol)
    fam = rng.integers(0, 2, size=nmol)
    Zs, xs, off = [], [], [0]
    for k, (na, f) in enumerate(zip(sizes, fam)):
        sk = int(seed * 1_000_003 + k)
        if f == 0:
            Z, xyz = alkane(max(1, (int(na) - 2) // 3), sigma=sigma, seed=sk)
        else:
            Z, xyz = glycine_peptide(max(1, (int(na) - 3) // 7), sigma=sigma, seed=sk)
        Zs.append(Z)
        xs.append(xyz)
        off.append(off[-1] + len(Z))
    return np.asarray(off, dtype=np.int32), np.concatenate(Zs).astype(np.int32), np.vstack(xs)


WORKLOADS = {
    "alkane_c40h82": lambda: alkane(40, sigma=0.02, seed=1),
    "water_3000": lambda: water_cluster(3000, seed=2),
    "alkane_bundle_80xc80": lambda: alkane_bundle(),
}
