/*
 * gfnff_params.h -- flat, read-only view of the GFN-FF parameter set that the
 * engine builds once per molecule (topology, EEQ corrections, term lists).
 *
 * This is the engine's equivalent of the reference's GFNFFParameterSet
 * (src/core/energy_calculators/ff_methods/gfnff_parameters.h:247-303) plus the
 * per-atom topology data the per-evaluation path reads (TopologyInfo,
 * ff_methods/gfnff.h:295-1003).  Everything is structure-of-arrays, 0-based,
 * lengths in Bohr, energies in Hartree.
 *
 * The view points into memory owned by a gfnff_handle; it stays valid until the
 * next gfnff_set_molecule() or gfnff_destroy() on that handle.
 */
#ifndef GFNFF_PARAMS_H
#define GFNFF_PARAMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFNFF_MAXREF 7

/* Packed per-term parameter record widths (doubles per entry). */
#define GFNFF_BOND_NPAR 8   /* r0b_i r0b_j cnfak_i cnfak_j rabshift ff alpha fc */
#define GFNFF_ANGLE_NPAR 2  /* fc theta0 */
#define GFNFF_TORS_NPAR 4   /* V n phi0 atcut */
#define GFNFF_INV_NPAR 3    /* fc omega0 type */
#define GFNFF_REP_NPAR 2    /* alpha repab */
#define GFNFF_HB_NPAR 7     /* bas_A bas_B aci_A aci_B q_H q_A q_B */
#define GFNFF_BATM_NPAR 3   /* zb3atm_i zb3atm_j zb3atm_k */
#define GFNFF_XB_NPAR 3     /* acidity_X q_X q_B */

typedef struct gfnff_params_view {
    int32_t natoms;
    int32_t nfrag;
    int32_t nelem;              /* number of distinct elements present */
    int32_t total_charge;

    /* per atom [natoms] */
    const int32_t* Z;
    const int32_t* elem;        /* index into present_Z */
    const int32_t* frag;        /* fragment id, 0-based */
    const double* chi_base;     /* -chi + dxi (-0.02 for amide H) */
    const double* gam_eff;      /* gam + dgam */
    const double* alpeeq;       /* (alpha + ff*qa)^2 */
    const double* cnf;          /* EEQ CN factor */
    const double* rcov_cn;      /* 4/3 * r_cov(Pyykko) * 1.8897259886 (CN radius) */
    const double* rcov_d3;      /* D3 covalent radius in Bohr, no 4/3 (damping radii) */
    const double* hb_rad;       /* Pyykko covalent radius as tabulated (Angstrom), HB/XB damping */
    const int32_t* metal;       /* metal class 0/1/2 of the element */
    const double* qa;           /* Phase-1 topology charges */
    const double* zeta;         /* D4 zeta(Z, qa) */
    const double* sqrt_zr4r2;   /* D4 sqrt(0.5*r4r2*sqrt(Z)) */
    const double* rep_anb;      /* non-bonded repulsion atomic exponent factor */
    const double* repz;         /* repulsion effective charge */
    const double* hb_lpd;       /* HB case-4 lone-pair distance 0.50 - 0.018 repz(Z), from the
                                   element table (independent of fitted repz scales) */
    const double* hb_bas;       /* HB/XB basicity per atom (element value with the carbonyl,
                                   nitro and carbene overrides) */
    const double* hb_aci;       /* HB acidity per atom (amide N scaled by 0.80) */
    const double* qfrag;        /* [nfrag] fragment charges */

    /* present elements [nelem] and D4 reference data */
    const int32_t* present_Z;
    const int32_t* nref;        /* [nelem] */
    const double* refcn;        /* [nelem*7] */
    const double* c6ref;        /* [nelem*nelem*7*7]  (ei,ej,ri,rj) */

    /* topological neighbourhood up to 3 bonds: CSR rows per atom (sorted j),
       dist 1 = bonded, 2 = 1,3, 3 = 1,4 */
    const int32_t* topo_off;    /* [natoms+1] */
    const int32_t* topo_idx;
    const int32_t* topo_dist;

    /* term lists */
    int32_t nbond;
    const int32_t* bond_ij;     /* [2*nbond] */
    const double* bond_par;     /* [8*nbond] */
    const int32_t* bond_hbH;    /* [nbond] H atom whose HB-CN scales alpha, or -1 */
    int32_t nhbcn;              /* H...B pairs entering the X-H bond HB coordination */
    const int32_t* hbcn_pair;   /* [2*nhbcn] (H, B) */

    int32_t nangle;
    const int32_t* angle_ijk;   /* [3*nangle], j = centre */
    const double* angle_par;

    int32_t ntors;
    const int32_t* tors_ijkl;   /* [4*ntors] */
    const double* tors_par;

    int32_t ninv;
    const int32_t* inv_ijkl;    /* [4*ninv] i = centre */
    const double* inv_par;

    int32_t nstors;             /* triple-bond torsions (undamped) */
    const int32_t* stors_ijkl;  /* [4*nstors] */
    const double* stors_par;    /* [nstors] erefhalf: E = erefhalf*(1 - cos 2 phi) */

    int32_t nbrep;
    const int32_t* brep_ij;
    const double* brep_par;

    int32_t nhb;
    const int32_t* hb_ahb;      /* [3*nhb] A, H, B */
    const int32_t* hb_case;     /* [nhb] 1..4 */
    const double* hb_par;       /* [7*nhb] */
    const int32_t* hb_nbB_off;  /* [nhb+1] CSR of neighbours of B */
    const int32_t* hb_nbB_idx;
    const int32_t* hb_parentC;  /* [nhb] acceptor parent C (case 3) or -1 */
    const int32_t* hb_nbC_off;  /* [nhb+1] neighbours of C other than B (case 3) */
    const int32_t* hb_nbC_idx;

    int32_t nxb;                /* halogen/chalcogen/pnictogen bonds A-X...B */
    const int32_t* xb_axb;      /* [3*nxb] A, X, B */
    const double* xb_par;       /* [3*nxb] */

    int32_t natm;
    const int32_t* atm_ijk;     /* [3*natm] */
    const double* atm_scale;    /* [natm] triple scale */

    int32_t nbatm;
    const int32_t* batm_ijk;    /* [3*nbatm] */
    const double* batm_par;     /* [3*nbatm] */
} gfnff_params_view;

/* ALPB implicit solvation of one molecule (units Bohr, Hartree): the
   per-atom parameters of alpb_solvation.cpp:362-396 and the angular grid. */
typedef struct gfnff_alpb_view {
    int32_t natoms;             /* 0: gas phase */
    int32_t hbond;              /* 1 when any atom has a H-bond strength */
    int32_t nang;               /* angular grid points */
    double keps;                /* (1/eps - 1) / (1 + alpbet) */
    double alpbet;              /* 0.571412 / eps */
    double born_scale;          /* c1 */
    double gshift;              /* free-energy shift */
    double lrcut;               /* Born-radius pair cutoff */
    double srcut;               /* surface neighbour cutoff */
    const double* vdwr;         /* D3 van der Waals radius */
    const double* rho;          /* descreening radius vdwr * sx */
    const double* svdw;         /* vdwr - Born offset */
    const double* vdwsa;        /* vdwr + probe radius */
    const double* wrp;          /* radial weight of the surface shell */
    const double* trj_lo;       /* (vdwsa - w)^2 */
    const double* trj_hi;       /* (vdwsa + w)^2 */
    const double* gamsasa;      /* surface tension */
    const double* hbmag;        /* H-bond strength (<= 0) */
    const double* grid;         /* [4*nang] x, y, z, weight */
} gfnff_alpb_view;

/* Per-term energy slots, in the order of the reference's FFEnergyComponents
   (ff_methods/ff_workspace.h:49-95) restricted to GFN-FF terms. */
enum gfnff_term {
    GFNFF_E_BOND = 0,
    GFNFF_E_ANGLE,
    GFNFF_E_TORSION,
    GFNFF_E_INVERSION,
    GFNFF_E_DISPERSION,
    GFNFF_E_BONDED_REP,
    GFNFF_E_NONBONDED_REP,
    GFNFF_E_COULOMB,
    GFNFF_E_HBOND,
    GFNFF_E_XBOND,
    GFNFF_E_ATM,
    GFNFF_E_BATM,
    GFNFF_E_STORS,
    GFNFF_E_SOLVATION,          /* ALPB solvation free energy (0 in the gas phase) */
    GFNFF_E_RESERVED14,
    GFNFF_NTERMS
};

#ifdef __cplusplus
}
#endif
#endif
