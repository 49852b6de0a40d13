// Host-side GFN-FF topology and parameter generation (built once per molecule).
//
// This is the engine's counterpart of the reference's initialisation path
// (src/core/energy_calculators/ff_methods/gfnff_method.cpp:709-850,
// calculateTopologyInfo :8813-9556, generateGFNFFParameterSet :3557-3747).
// It produces the flat structure-of-arrays parameter set that the device
// kernels and the CPU oracle consume; nothing here runs per evaluation.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "alpb.hpp"
#include "gfnff_params.h"

namespace gfnff {

// Unit constants.  The reference uses three different Angstrom/Bohr factors and
// each one is used where the reference uses it:
//   geometry input     1/0.529177210903   (src/core/units.h:81-82)
//   CN covalent radii  1.8897259886       (ff_methods/cn_calculator.cpp:375)
//   GFN-FF tables      1/0.52917726       (ff_methods/gfnff_par.h:276)
constexpr double kAngToBohrGeom = 1.0 / 0.529177210903;
constexpr double kAngToBohrCN = 1.8897259886;
constexpr double kAngToBohrTab = 1.0 / 0.52917726;

// Atom count from which the reference builds its spatial cell list for the
// hydrogen/halogen-bond candidates (nb_cell_list_min_atoms, gfnff_method.cpp:8078);
// it changes the XB cut-off compare from strict to inclusive.
constexpr int kCellListMinAtoms = 800;

struct Options {
    double cn_cutoff_bohr = 6.0;    // per-evaluation CN neighbour cutoff
    double hb_accuracy = 0.1;       // HB list thresholds hbthr1/hbthr2
    int threads = 1;                // host threads for the Phase-1 solve
    bool cholesky_lookahead = false;  // overlap trailing updates with the next super-panel (side stream)
    double hb_update_rmsd_bohr = 0.3;  // HB/XB re-detection threshold on sqrt(sum dx^2)/N
    int eeq_solver = 0;               // 0: preconditioned CG (single molecules), 1: blocked Cholesky
    double pcg_tol = 0.0;             // PCG relative residual |r| / |b|; 0: 1e-13 (FP64), 1e-10 (mixed)
    int batch_solver = 5;             // batches: right-looking per-molecule Cholesky with tile phases per
                                      // wavefront in two occupancy / prefetch variants 0, 1; 2 one FP64 tile
                                      // step at a time per workgroup; 3 one column block at a time over the
                                      // whole batch (three launches per block); 4 left-looking per molecule
                                      // with the matrix formed in place (FP64); 5 (default) the same by column
                                      // blocks over the whole batch
    int batch_fp32 = 0;               // mixed batches: 1 factors in FP32 (MFMA) with FP64 refinement
    int gather_lanes = 0;             // lanes per atom of the slot and HB gathers: 0 by list length, 16 or 64
    int batch_pairs = 0;              // batches: 0 each pair once in 64x64 tiles per molecule, 1 row kernel
    int pair_split = 0;               // 1: dispersion + repulsion before the EEQ solve, Coulomb after (tiles)
    int batch_streams = 1;            // batch_solver 5: 2 = the batch's factorisation in two halves on two streams
    int pcg_groups = 0;               // workgroups of the persistent PCG: 0 all that fit (at most 256), else at most this
};

// Error codes shared with the C-ABI.
enum Status : int {
    kOk = 0,
    kErrArgs = 1,
    kErrUnsupported = 2,
    kErrEEQ = 3,
    kErrGeometry = 4,
    kErrDevice = 5,
    kErrNaN = 6,
};

// Owns every array the gfnff_params_view points into.
struct ParamSet {
    int n = 0, nfrag = 0, nelem = 0, charge = 0;
    std::vector<int32_t> Z, elem, frag, present_Z, nref, metal;
    std::vector<double> chi_base, gam_eff, alpeeq, cnf, rcov_cn, rcov_d3, hb_rad, qa, zeta, sqrt_zr4r2,
        rep_anb, repz, hb_lpd, hb_bas, hb_aci, qfrag, refcn, c6ref;
    std::vector<int32_t> topo_off, topo_idx, topo_dist;

    std::vector<int32_t> bond_ij, bond_hbH, hbcn_pair;
    std::vector<double> bond_par;
    std::vector<int32_t> angle_ijk;
    std::vector<double> angle_par;
    std::vector<int32_t> tors_ijkl;
    std::vector<double> tors_par;
    std::vector<int32_t> inv_ijkl;
    std::vector<double> inv_par;
    std::vector<int32_t> stors_ijkl;
    std::vector<double> stors_par;
    std::vector<int32_t> brep_ij;
    std::vector<double> brep_par;
    std::vector<int32_t> hb_ahb, hb_case, hb_nbB_off, hb_nbB_idx, hb_parentC, hb_nbC_off, hb_nbC_idx;
    std::vector<double> hb_par;
    std::vector<int32_t> xb_axb;
    std::vector<double> xb_par;
    std::vector<int32_t> atm_ijk;
    std::vector<double> atm_scale;
    std::vector<int32_t> batm_ijk;
    std::vector<double> batm_par;

    // Diagnostics kept from the topology build (not used per evaluation).
    std::vector<int32_t> hyb, pi_frag, topo_bonds;   // topo_bonds: [2*nb] detected bonds
    std::vector<double> cn_topo;

    // batches of independent molecules (nmol > 1): atom offsets and, per term
    // list (bond, angle, torsion, bonded repulsion, HB, ATM, BATM, inversion,
    // triple-bond torsion, XB), the entry offsets of every molecule ([10][nmol + 1])
    // ALPB solvation of a single molecule (n == 0: gas phase)
    AlpbSet solv;

    int nmol = 1;
    std::vector<int32_t> mol_off, mol_seg;
    int mol_of_atom(int i) const;

    gfnff_params_view view() const;
};

// D4 element tables (reference counts, reference CNs, C6 grid) for P.present_Z.
void set_element_tables(ParamSet& P);

// Concatenates independently generated molecules into one parameter set with
// molecule offsets (the batched evaluation treats molecules as independent).
int merge_batch(const std::vector<ParamSet>& mols, ParamSet& out, std::string& msg);

// Builds the parameter set for one molecule.  xyz_bohr is N x 3 row-major.
// forced_bonds (pairs, 0-based) replaces geometric bond detection when nonempty
// (reference: gfnff_method.cpp:954-963).  Returns a Status; msg gets a reason.
int build_params(int n, const int32_t* Z, const double* xyz_bohr, int charge,
                 const int32_t* forced_bonds, int nforced, const Options& opt, ParamSet& out,
                 std::string& msg);

// Re-detects the hydrogen- and halogen-bond lists of a single molecule at the
// geometry xyz_bohr with the set-up topology and topology charges
// (updateHBXBIfNeeded, gfnff_method.cpp:1597-1690); the bond/HB cross
// reference of the set-up is kept.  changed tells whether any list differs.
int refresh_hbxb(ParamSet& P, const double* xyz_bohr, const Options& opt, bool& changed, std::string& msg);

// Dense symmetric solve of the bordered EEQ system used at set-up
// (A q + C^T lambda = x, C q = qfrag).  A is n x n row-major (overwritten).
bool eeq_solve_host(int n, int nfrag, std::vector<double>& A, const std::vector<double>& x,
                    const std::vector<int32_t>& frag, const std::vector<double>& qfrag,
                    std::vector<double>& q, int threads);

}  // namespace gfnff
