/*
 * gfnff_mi355x.h -- C-ABI of the MI355X GFN-FF energy/gradient engine.
 *
 * The entry points replace the reference's ComputationalMethod plug-in for
 * "-method gfnff" (src/core/energy_calculators/computational_method.h:43-360,
 * created by MethodFactory::create, method_factory.cpp:460-516, CPU adapter
 * qm_methods/gfnff_method.{h,cpp}).  Plain pointers and sizes only; no C++ or
 * torch types cross this boundary and nothing throws across it.  Every
 * function returns a status (0 = ok, see GFNFF_STATUS_*); the message of the
 * last failure is available from gfnff_last_error().
 *
 * Units follow the reference: coordinates in Angstrom on input (converted with
 * 1/0.529177210903, gfnff_method.cpp:757), energy in Hartree, gradient in
 * Hartree/Bohr (no conversion, gfnff_method.cpp:1831-1844).
 *
 * Threading: one handle is used by one host thread at a time (the reference's
 * GFNFFComputationalMethod::isThreadSafe() returns false,
 * qm_methods/gfnff_method.h:53).
 */
#ifndef GFNFF_MI355X_H
#define GFNFF_MI355X_H

#include <stdint.h>

#include "gfnff_params.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GFNFF_STATUS_OK 0
#define GFNFF_STATUS_ARGS 1
#define GFNFF_STATUS_UNSUPPORTED 2
#define GFNFF_STATUS_EEQ 3
#define GFNFF_STATUS_GEOMETRY 4
#define GFNFF_STATUS_DEVICE 5
#define GFNFF_STATUS_NAN 6

/* arithmetic modes for gfnff_set_precision */
#define GFNFF_PREC_FP64 0
#define GFNFF_PREC_MIXED 1

typedef struct gfnff_handle gfnff_handle;

/* Replaces MethodFactory::create("gfnff", cfg) (method_factory.cpp:460-516). */
gfnff_handle* gfnff_create(void);
void gfnff_destroy(gfnff_handle* h);

/* ComputationalMethod::setMolecule(const Mol&) (computational_method.h:56;
 * Mol layout global.h:69-99): Z[natoms], xyz_ang[natoms*3] row-major,
 * total charge, optional forced bonds [2*nbonds] (0-based) that replace
 * geometric bond detection when nbonds > 0 (gfnff_method.cpp:954-963).
 * Builds topology and parameters on the host; no device work. */
int gfnff_set_molecule(gfnff_handle* h, int32_t natoms, const int32_t* Z, const double* xyz_ang, int32_t charge,
                       const int32_t* forced_bonds, int32_t nbonds);

/* Batch of independent molecules (the batched workload; no counterpart in the
 * reference's single-molecule interface, which a caller would loop over):
 * molecule m is atoms [atom_off[m], atom_off[m+1]) of Z / xyz_ang (Angstrom),
 * total charge charges[m] (NULL: all neutral).  Each molecule's topology and
 * parameters are generated as by gfnff_set_molecule; every later evaluation
 * (gfnff_energy_gradient, gfnff_energy_gradient_device) treats the molecules
 * as independent systems in one launch per kernel: *energy is the sum over
 * the batch, the gradient covers all atoms.  Molecules must have at most 512
 * reduced EEQ dimensions (atoms minus fragments, rounded up to 64). */
int gfnff_set_batch(gfnff_handle* h, int32_t nmol, const int32_t* atom_off, const int32_t* Z, const double* xyz_ang,
                    const int32_t* charges);

/* Per-molecule energies (nmol) of the last evaluation of a batch: host copy
 * (synchronises) or device copy queued on `stream`. */
int gfnff_batch_energies(gfnff_handle* h, double* e_mol);
int gfnff_batch_energies_device(gfnff_handle* h, double* e_mol_dev, void* stream);

/* ComputationalMethod::updateGeometry(const Matrix&) (computational_method.h:63), Angstrom.
 * Like the reference's per-evaluation check (gfnff_method.cpp:1597-1690,
 * 2228-2259), the hydrogen/halogen-bond lists of a single molecule are
 * re-detected when sqrt(sum dx^2)/N exceeds hb_update_rmsd_bohr (0.3) since
 * the last detection; the same check runs before every host-buffer
 * evaluation (the first evaluation always re-detects).  The device-resident
 * entry point does not read coordinates back: it uses the lists of the last
 * host-side check. */
int gfnff_update_geometry(gfnff_handle* h, const double* xyz_ang);

/* ComputationalMethod::calculateEnergy(bool gradient) + copyGradientTo(Matrix&)
 * (computational_method.h:70,121-123).  energy: Hartree; grad: natoms*3
 * Hartree/Bohr (may be NULL when want_grad == 0); terms: GFNFF_NTERMS per-term
 * energies (may be NULL).  Runs on the GPU. */
int gfnff_energy_gradient(gfnff_handle* h, int32_t want_grad, double* energy, double* grad, double* terms);

/* Device-resident variant: xyz_bohr_dev (natoms*3, Bohr), energy_dev
 * (GFNFF_NTERMS+1 doubles, always written: the total followed by the
 * terms), grad_dev (natoms*3, may be NULL) are device pointers; work is queued
 * on `stream` (a hipStream_t, NULL = the default stream) and the call returns
 * without synchronising.  Failures detected on the device (see
 * GFNFF_FLAG_*) make energy_dev[0] NaN and are accumulated in the handle's
 * status word, read with gfnff_device_status. */
int gfnff_energy_gradient_device(gfnff_handle* h, const double* xyz_bohr_dev, double* energy_dev, double* grad_dev,
                                 void* stream);

/* Engine extension (no counterpart in the reference, whose CPU evaluation
 * has no launches to replay): records one device-resident evaluation with
 * these fixed buffers (as gfnff_energy_gradient_device) into a HIP graph,
 * both streams of the engine included, and instantiates it.  Nothing runs
 * during the capture.  `stream` must be a created hipStream_t (capture is not
 * possible on the default stream).  gfnff_graph_launch replays it on any
 * stream: every replay evaluates the coordinates then in xyz_bohr_dev and
 * writes energy_dev and grad_dev.  The graph is dropped when the engine is
 * rebuilt (new molecule, parameters, options, or a hydrogen-bond list change
 * found by gfnff_update_geometry); gfnff_graph_launch then returns
 * GFNFF_STATUS_ARGS and a new capture is needed. */
int gfnff_graph_capture(gfnff_handle* h, const double* xyz_bohr_dev, double* energy_dev, double* grad_dev,
                        void* stream);
int gfnff_graph_launch(gfnff_handle* h, void* stream);

/* Device-resident molecular dynamics (replaces the per-step energy calls of
 * SimpleMD: simplemd.cpp:3866-3890 FastEnergy inside SimpleMD::Verlet
 * :2580-2673).  Queues nsteps velocity-Verlet steps on `stream`, each one
 * device-resident evaluation of the handle's molecule plus a few update
 * kernels, with no host transfer per step.  xyz_bohr_dev, vel_dev (Bohr per
 * atomic time unit) and grad_dev (Eh/Bohr) are natoms x 3 device arrays
 * updated in place; grad_dev must hold the gradient at xyz_bohr_dev on entry
 * (one gfnff_energy_gradient_device call) and holds it on return.
 * mass_amu: natoms host masses in amu; dt_fs: time step in fs.  T0 > 0
 * couples to a Berendsen bath of time constant tau_fs (SimpleMD::Berendson,
 * simplemd.cpp:3967-3975; the bath scales after the second half-kick, as
 * there), T0 <= 0 runs NVE.  record_dev (device, nsteps x 4 doubles)
 * receives Epot, Ekin, T (K, 3N - 6 degrees of freedom) and Epot + Ekin per
 * step.  When the displacement since the last hydrogen/halogen-bond
 * detection, sqrt(sum dx^2) / N, exceeds hb_update_rmsd_bohr
 * (gfnff_method.cpp:2243-2251), GFNFF_FLAG_HB_STALE is raised in the status
 * word: the caller then reads the coordinates back, calls
 * gfnff_update_geometry (which re-detects the lists) and continues. */
int gfnff_md_steps(gfnff_handle* h, double* xyz_bohr_dev, double* vel_dev, double* grad_dev, const double* mass_amu,
                   double dt_fs, double T0, double tau_fs, int32_t nsteps, double* record_dev, void* stream);

/* Device-side failure flags of an evaluation (bit mask). */
#define GFNFF_FLAG_EEQ_NOT_SPD 1   /* EEQ matrix not positive definite (Cholesky pivot) */
#define GFNFF_FLAG_CN_OVERFLOW 2   /* more than 128 atoms within the 6 Bohr CN cut-off of one atom */
#define GFNFF_FLAG_SOLVE 4         /* EEQ solve did not converge or a bounded wait timed out */
#define GFNFF_FLAG_HB_STALE 8      /* gfnff_md_steps: moved past the HB/XB re-detection threshold (not a failure) */
#define GFNFF_FLAG_SOLV_NEIGHBOURS 16 /* ALPB: more than 512 atoms within the surface cut-off of one atom */

/* Engine extension (no counterpart in the reference, whose CPU evaluation
 * reports through hasError()): the GFNFF_FLAG_* bits raised by every
 * evaluation of the handle since the last call with clear != 0.
 * Synchronises the device.  0 in *flags means all evaluations were clean. */
int gfnff_device_status(gfnff_handle* h, int32_t* flags, int32_t clear);

/* ComputationalMethod::getCharges() / getCN() (computational_method.h:295-337):
 * EEQ charges and coordination numbers of the last evaluation. */
int gfnff_charges(gfnff_handle* h, double* q);
int gfnff_cn(gfnff_handle* h, double* cn);

/* ComputationalMethod::setParameters(json) for the parity-relevant keys
 * (gfnff.h:219-293): "cn_cutoff_bohr", "hb_accuracy", "threads",
 * "hb_update_rmsd_bohr" (gfnff_method.cpp:2250).  Engine keys:
 * "eeq_solver" 0 = preconditioned CG on the reduced EEQ system (default for
 * single molecules), 1 = blocked Cholesky, the direct solve of the reference
 * (batches always factor per molecule); "pcg_tol" the relative residual
 * |r| / |b| at which the CG solve stops (0 = default: 1e-13 for FP64, 1e-10
 * for the mixed path, at most 1e-6); "cholesky_lookahead" 0/1 overlaps the
 * trailing update with the next super-panel on a second stream (off by
 * default).  Launch-shape keys (results unchanged, chosen by the autotuner,
 * curcuma_amd/autotune.py): "batch_solver" 0 / 1 = per-molecule Cholesky
 * with the tile phases shared out over wavefronts in two occupancy /
 * operand-prefetch variants, 2 = one FP64 tile step at a time per
 * workgroup (factor and substitutions in one launch), 3 = the
 * factorisations advanced one column block at a time over the whole batch
 * (three launches per block: diagonal blocks, panel tiles, trailing tiles),
 * 4 = left-looking per molecule with the matrix entries formed inside the
 * factorisation (no separate matrix build; FP64 on both paths), 5 = the
 * same left-looking form advanced one column block at a time over the whole
 * batch (three launches per block: column tiles, diagonal blocks, panels;
 * the default), 6 = as 5 with the panel product done inside the tiles below
 * the diagonal, formed after the diagonal block (three launches per block:
 * diagonal tiles, diagonal blocks, tiles below);
 * "batch_streams" 2 = with batch_solver 5, the batch's factorisation in
 * two halves (molecules dealt alternately by size) on two streams (1, one
 * stream, is the default; 2 runs eagerly only: gfnff_graph_capture refuses
 * it, and it is not faster on the mixed path);
 * "batch_fp32" 1 = on the mixed path, batches factor in FP32 (FP32 MFMA)
 * and refine in FP64 (charges agree with the FP64 factor to about 1e-12 e;
 * batch_solver 0, 1 and 3); "pair_split" 1 = the charge-independent pair
 * terms (dispersion, repulsion) evaluated before the EEQ solve, beside it,
 * and the Coulomb term after it (pair tiles only; 0 = one pass after the
 * solve, the default); "batch_pairs" 0 = pair terms of batches with each
 * pair once in 64 x 64 tiles per molecule (default), 1 = one thread per row
 * atom over its molecule; "gather_lanes" 0 (by mean list length), 16 or 64
 * lanes per atom in the slot and hydrogen-bond gathers; "pcg_groups" 0 (all
 * workgroups that fit, at most 256) or an upper bound on the persistent
 * PCG's workgroups. */
int gfnff_set_param(gfnff_handle* h, const char* key, double value);

/* GFNFF_PREC_FP64 (default, the reference's arithmetic) or GFNFF_PREC_MIXED:
 * pair terms in FP32 with FP64 coordinate differences and FP64 row sums; the
 * EEQ solve in FP64 to |r| <= 1e-10 |b| (CG), or with "eeq_solver" 1 the
 * reduced EEQ matrix factored in FP32 (FP32 MFMA) with FP64 iterative
 * refinement against the FP64 matrix; everything else FP64. */
int gfnff_set_precision(gfnff_handle* h, int32_t mode);

/* GFNFF's "solvent" / "solvent_model" parameters (gfnff_method.cpp:811-835,
 * -gfnff.solvent NAME): ALPB implicit solvation, which GFN-FF always uses
 * (P16 Born kernel; a request for GBSA means ALPB there too).  NAME is one of
 * the reference's 25 GFN-FF solvents or their aliases (alpb_parameters.h:
 * 1856-1891; case-insensitive); NULL, "" or "none" select the gas phase.
 * Applies to the current molecule and to every molecule set afterwards; the
 * energy then includes the solvation free energy (gsolv state) in
 * GFNFF_E_SOLVATION, and the EEQ charges feel the reaction field
 * (gfnff_method.cpp:1129-1140).  An unknown name returns GFNFF_STATUS_ARGS
 * and leaves the handle in the gas phase, as the reference continues
 * gas-phase after a failed initialisation (gfnff_method.cpp:826-829).
 * Single molecules only: GFNFF_STATUS_UNSUPPORTED for batches. */
int gfnff_set_solvent(gfnff_handle* h, const char* solvent);

/* ALPBSolvation::getEnergyParts (alpb_solvation.cpp:1016-1041) of the last
 * evaluation: parts[4] = {gborn, ghb, gsasa, gshift} in Hartree. */
int gfnff_solvation_energy(gfnff_handle* h, double* parts);

/* Engine extension: read-only view of the solvation parameters of the
 * current molecule (natoms 0 in the gas phase). */
int gfnff_get_solvation(const gfnff_handle* h, gfnff_alpb_view* out);

/* ComputationalMethod::hasError() / getErrorMessage() (computational_method.h:190,201). */
int gfnff_has_error(const gfnff_handle* h);
const char* gfnff_last_error(const gfnff_handle* h);

/* Engine extension: read-only view of the parameter set built by
 * gfnff_set_molecule (no counterpart in the reference's interface; the
 * reference exposes the same data through GFNFFParameterSet). */
int gfnff_get_params(const gfnff_handle* h, gfnff_params_view* out);

/* Engine extension: topology diagnostics of the last set-up.  hyb and pi
 * (natoms each), bonds (2*nbonds_max); returns the number of detected bonds
 * in *nbonds. */
int gfnff_get_topology(const gfnff_handle* h, int32_t* hyb, int32_t* pi, int32_t* bonds, int32_t nbonds_max,
                       int32_t* nbonds);

/* Number of visible GPUs (0 when none). */
int gfnff_device_count(void);

/* Iterations of the last iterative EEQ solve of the handle (0 when the
 * Cholesky path ran, -1 without a device engine); synchronises. No
 * counterpart in the reference (solver diagnostics). */
int gfnff_solver_iterations(gfnff_handle* h);

/* Parameter adjoint for force-matching fits (no counterpart in the
 * reference's interface; used by a fitting harness).  The fitted parameters
 * are per-element scale factors theta (3 * nelem, elements in the order of
 * gfnff_params_view.present_Z):
 *   theta[e]             repulsion charge repz (bonded and non-bonded,
 *                        ff_workspace_gfnff.cpp:774-842)
 *   theta[nelem + e]     repulsion exponent (non-bonded rep_anb, bonded repa)
 *   theta[2 nelem + e]   D4 sqrt(Z r4r2) (ff_workspace_gfnff.cpp:715-767)
 * gfnff_param_count returns 3 * nelem (-1 without a molecule);
 * gfnff_param_scales sets theta (all 1 after gfnff_set_molecule);
 * gfnff_param_adjoint returns dtheta[k] = sum_i w_i . d g_i / d theta_k at the
 * last evaluation, for a weight field w (natoms * 3), e.g. w = 2 (g - g_ref)
 * for the loss sum |g - g_ref|^2 = sum |F - F_ref|^2.  Built for up to 4
 * elements (GFNFF_STATUS_UNSUPPORTED beyond). */
int gfnff_param_count(const gfnff_handle* h);
int gfnff_param_scales(gfnff_handle* h, const double* theta, int32_t ntheta);
int gfnff_param_adjoint(gfnff_handle* h, const double* w, double* dtheta, int32_t ntheta);

/* Engine extension: the CN neighbour lists of the last evaluation, as the
 * reference's pair builder #1 defines them (CNCalculator::
 * calculateGFNFFCNWithNeighbors, cn_calculator.cpp:367-428): every pair with
 * d^2 < 36 Bohr^2 (strict), each list ascending in j.  counts[natoms],
 * idx[natoms * maxnb], entries past a list's end are -1.  Lists are capped at
 * 128 entries (the evaluation fails with GFNFF_STATUS_GEOMETRY beyond that). */
int gfnff_cn_neighbours(gfnff_handle* h, int32_t* counts, int32_t* idx, int32_t maxnb);

/* Engine extension (no counterpart in the reference): HIP-event timing of one
 * kernel class on the evaluation stream, for roofline measurement.  Classes:
 * cn eeq_build chol_diag chol_panel chol_update trsv charges pairs pairs_aux nbrep
 * refine bonds angles torsions brep hb atm batm chain reduce; NULL or "" switches timing
 * off.  Every launch of the class in later evaluations is bracketed by a
 * start/stop event pair. */
int gfnff_profile_kernel(gfnff_handle* h, const char* name);

/* Waits for the recorded events and returns the summed launch time (ms) and
 * the number of launches since the last read or selection; restarts the count. */
int gfnff_profile_read(gfnff_handle* h, double* total_ms, int32_t* launches);

/* Engine extension (no counterpart in the reference): dense Cholesky of an
 * n x n SPD matrix (row-major, host buffers) with the engine's blocked FP64
 * factorisation kernels; l_out receives the lower factor.  Used to test the
 * EEQ factorisation in isolation. */
int gfnff_dense_cholesky(int32_t n, const double* a, double* l_out);

/* The same with the FP32 kernels of the mixed-precision path (FP32 MFMA). */
int gfnff_dense_cholesky_f32(int32_t n, const float* a, float* l_out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gfnff_dense_cholesky on device buffers, allocating
 * nothing.  a_dev: n x n SPD matrix (row-major, read only); l_dev: n x n,
 * receives L (zeros above the diagonal); status_dev: one int32, receives 0 or
 * GFNFF_STATUS_EEQ (a pivot was not positive) when the stream reaches it.
 * Work is queued on `stream` (hipStream_t, NULL = default) without
 * synchronising.  The padded working matrix, the inverses of its diagonal
 * blocks and the status word live in one static device array of
 * GFNFF_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), which
 * every call on that GPU shares: calls that overlap on one GPU race.
 * n > GFNFF_NOALLOC_MAX_N, n <= 0 or a NULL pointer return GFNFF_STATUS_ARGS
 * and queue nothing. */
#define GFNFF_NOALLOC_MAX_N 1024
#define GFNFF_NOALLOC_SCRATCH_BYTES (8 * (1024 * 1024 + 1024 * 64 + 32))
int gfnff_dense_cholesky_noalloc(int32_t n, const double* a_dev, double* l_dev, int32_t* status_dev, void* stream);

#ifdef __cplusplus
}
#endif
#endif
