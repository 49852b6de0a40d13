/*
 * roofline_counts.h -- algorithmic work per unit of the GFN-FF evaluation,
 * the fixed convention behind every roofline figure the engine reports.
 *
 * Counting rule: each add, sub, mul, div, sqrt, rsqrt and each exp, erf,
 * log, pow, sin, cos, acos or atan2 call is 1 flop; an FMA is 2; compares
 * and selects are 0; the internals of transcendental functions are not
 * counted.  Bytes are the compulsory HBM bytes of the unit.  Hardware
 * counters (rocprofv3 FETCH_SIZE / WRITE_SIZE) are reported beside these
 * figures, never in their place.
 *
 * Units count unordered pairs (i < j) inside the term's cut-off: Coulomb
 * r <= 100 Bohr (ff_workspace_gfnff.cpp:850-898), D4 dispersion r <= 50 Bohr
 * (:715-767), non-bonded repulsion r <= 20 Bohr (:814-842).
 */
#ifndef GFNFF_ROOFLINE_COUNTS_H
#define GFNFF_ROOFLINE_COUNTS_H

/* pair terms (energy + gradient) */
#define GFNFF_RF_COULOMB_PAIR_FLOPS 36
#define GFNFF_RF_COULOMB_ATOM_BYTES_IN 40  /* xyz, q, alpha */
#define GFNFF_RF_COULOMB_ATOM_BYTES_OUT 24 /* gradient */
#define GFNFF_RF_EEQ_ELEM_FLOPS 14         /* off-diagonal A_ij */
#define GFNFF_RF_EEQ_DIAG_FLOPS 3
#define GFNFF_RF_EEQ_ELEM_BYTES 8 /* per stored FP64 element */
#define GFNFF_RF_DISP_PAIR_FLOPS 53
#define GFNFF_RF_REP_PAIR_FLOPS 31
#define GFNFF_RF_CN_PAIR_FLOPS 33 /* CN + dCN, fused */
#define GFNFF_RF_CN_ATOM_FLOPS 8
#define GFNFF_RF_CN_ENTRY_BYTES 8
#define GFNFF_RF_CHAIN_PAIR_FLOPS 12
#define GFNFF_RF_CHAIN_ATOM_FLOPS 6

/* bonded lists: flops, parameter bytes; each atom gathered costs 24 B of
 * coordinates in and 24 B of gradient out */
#define GFNFF_RF_BOND_FLOPS 46
#define GFNFF_RF_BOND_PAR_BYTES 72
#define GFNFF_RF_ANGLE_FLOPS 130
#define GFNFF_RF_ANGLE_PAR_BYTES 28
#define GFNFF_RF_TORSION_FLOPS 240
#define GFNFF_RF_TORSION_PAR_BYTES 44
#define GFNFF_RF_INVERSION_FLOPS 230
#define GFNFF_RF_INVERSION_PAR_BYTES 44
#define GFNFF_RF_HB_FLOPS 300
#define GFNFF_RF_HB_NEIGHBOUR_FLOPS 40
#define GFNFF_RF_HB_PAR_BYTES 96
#define GFNFF_RF_BATM_FLOPS 110
#define GFNFF_RF_BATM_BYTES 36
#define GFNFF_RF_ATM_FLOPS 160
#define GFNFF_RF_ATM_BYTES 60
#define GFNFF_RF_ATOM_GATHER_BYTES 24

/* EEQ linear algebra on an n x n system with nfrag constraints:
 *   Cholesky n^3/3; triangular solves 2 n^2 (1 + nfrag); Schur (2/3) nfrag^3;
 *   one FP64 refinement step 2 n^2 (residual) + 2 n^2 (solve), 8 n^2 bytes;
 *   one conjugate-gradient iteration reads the packed symmetric matrix once:
 *   8 (n/64)(n/64 + 1)/2 64^2 bytes (n padded to 64).
 * A rank-NB trailing update of the lower triangle (diagonal included) of an
 * m x m block counts m (m + 1) NB flops (one FMA per element and step). */
#define GFNFF_RF_CHOL_DIV 3
#define GFNFF_RF_TRSV_FACTOR 2
#define GFNFF_RF_SCHUR_NUM 2
#define GFNFF_RF_SCHUR_DEN 3
#define GFNFF_RF_REFINE_FACTOR 4

#endif
