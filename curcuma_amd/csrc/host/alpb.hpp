// ALPB implicit solvation for GFN-FF: per-atom parameters and the angular
// grid, built on the host once per molecule and solvent.  The per-evaluation
// arithmetic (Born radii, surface areas, the Born matrix's entries in the EEQ
// solve, energy and gradient) runs on the device (device/engine.hip, K13).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gfnff_params.h"

namespace gfnff {

// Units: Bohr, Hartree.  n == 0: gas phase (no solvent).
struct AlpbSet {
    int n = 0;
    std::string solvent;
    double keps = 0.0;        // (1/eps - 1) / (1 + alpbet)
    double alpbet = 0.0;      // 0.571412 / eps
    double born_scale = 0.0;  // c1
    double gshift = 0.0;      // free-energy shift (gsolv state)
    double lrcut = 0.0;       // Born-radius pair cutoff
    double srcut = 0.0;       // surface-area neighbour cutoff
    int hbond = 0;            // any atom with a non-zero H-bond strength
    // per atom
    std::vector<double> vdwr, rho, svdw, vdwsa, wrp, trj_lo, trj_hi, gamsasa, hbmag;
    // angular grid: x, y, z, weight per point
    std::vector<double> grid;
    int nang = 0;

    gfnff_alpb_view view() const;
};

// Canonical solvent name for `name` (case-insensitive, aliases resolved), or
// an empty string when the solvent is unknown.
std::string alpb_solvent_name(const char* name);

// Parameters of `solvent` for the atoms Z[0..n-1]; returns kOk or kErrArgs
// with msg set (unknown solvent).
int alpb_setup(const char* solvent, int n, const int32_t* Z, AlpbSet& S, std::string& msg);

// The 230-point Lebedev-Laikov grid (unit sphere, weights summing to 1) as
// x, y, z, w per point.
std::vector<double> lebedev230();

}  // namespace gfnff
