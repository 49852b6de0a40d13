// Device engine: owns the HBM-resident parameter set and runs one GFN-FF
// energy+gradient evaluation as a sequence of HIP kernels on one stream.
#pragma once

#include <memory>
#include <string>

#include "../host/gfnff_topology.hpp"

namespace gfnff {

struct DeviceEngine;

// Uploads the parameter set; returns nullptr and sets status/msg on failure.
// opt.eeq_solver: 0 = preconditioned CG on the reduced EEQ system (single
// molecules; batches always use the per-molecule Cholesky), 1 = blocked
// Cholesky.  opt.pcg_tol: relative residual of the PCG solve (0: 1e-13 for
// FP64, 1e-10 for the mixed path, whose accuracy bar is 1e-6 Eh/atom).
// opt.cholesky_lookahead, opt.batch_solver: see Options.
DeviceEngine* device_create(const ParamSet& P, int precision, const Options& opt, int& status, std::string& msg);
void device_destroy(DeviceEngine* e);

// Host-buffer evaluation: uploads xyz (Bohr), evaluates, downloads
// e_terms[0] = total, e_terms[1..] = per-term energies, gradient, q and CN.
int device_eval_host(DeviceEngine* e, const double* xyz_bohr, bool want_grad, double* e_terms, double* grad,
                     double* q, double* cn, std::string& msg);

// Device-resident evaluation, asynchronous on `stream`.
int device_eval(DeviceEngine* e, const double* xyz_bohr_dev, double* energy_dev, double* grad_dev, void* stream,
                std::string& msg);

// Graph capture of one device-resident evaluation (fixed buffers, non-default
// stream) and its replay; the graph lives as long as the engine.
int device_graph_capture(DeviceEngine* e, const double* xyz_bohr_dev, double* energy_dev, double* grad_dev,
                         void* stream, std::string& msg);
int device_graph_launch(DeviceEngine* e, void* stream, std::string& msg);

// Device-resident velocity-Verlet MD: nsteps steps queued on `stream`, x/v/grad
// device arrays updated in place (grad must hold the gradient at x on entry),
// rec_dev receives (Epot, Ekin, T, Etot) per step; flag 8 in the status word
// when sqrt(sum |x - xref|^2) / N exceeds disp_thresh.  Units: see engine.hip K12.
int device_md_steps(DeviceEngine* e, double* x_dev, double* v_dev, double* grad_dev, const double* mass_amu,
                    const double* xref_bohr, double dt_fs, double T0, double tau_fs, double disp_thresh, int nsteps,
                    double* rec_dev, void* stream, std::string& msg);

int device_count();

// ALPB solvation energy parts of the last evaluation {gborn, ghb, gsasa,
// gshift} (synchronises); kErrArgs when the engine has no solvent.
int device_solvation_parts(DeviceEngine* e, double* parts, std::string& msg);

// Status flags (FLAG_* bits) raised by any evaluation since the last cleared
// read; synchronises the device.  clear resets the word.
int device_status(DeviceEngine* e, int* flags, bool clear, std::string& msg);

// Iterations of the last iterative EEQ solve (0 when the Cholesky path runs,
// -1 on a device error); synchronises.
int device_solver_iterations(DeviceEngine* e);

// CN neighbour lists of the last evaluation (ascending j, strict d^2 < 36 Bohr^2):
// counts[n], idx[n * maxnb] (-1 past the end of a list).
int device_cn_neighbours(DeviceEngine* e, int* counts, int* idx, int maxnb, std::string& msg);

// Per-molecule energies of the last evaluation of a batch (host or device destination).
int device_batch_energies(DeviceEngine* e, double* e_mol, bool device_dst, void* stream, std::string& msg);

// Re-uploads the scaled repulsion / dispersion parameters of P.
int device_update_params(DeviceEngine* e, const ParamSet& P, std::string& msg);
// Parameter adjoint at the last evaluation: out[p * nelem + e] = sum_i w_i .
// dg_i / dln(theta_pe) for p = repz, repulsion exponent, sqrt(Z r4r2).
int device_param_adjoint(DeviceEngine* e, const double* w, double* out, std::string& msg);

// HIP-event timing of one kernel class on the evaluation stream.
int device_profile_select(DeviceEngine* e, const char* name, std::string& msg);
int device_profile_read(DeviceEngine* e, double* total_ms, int* launches, std::string& msg);

// Dense SPD factorisation with the EEQ Cholesky kernels (host buffers).
int device_dense_cholesky(int n, const double* a, double* l_out, std::string& msg);
int device_dense_cholesky_f32(int n, const float* a, float* l_out, std::string& msg);
// The same on device buffers without any allocation (scratch in a static device
// array, n <= 1024); asynchronous on `stream`, status_dev receives 0 or kErrEEQ.
int device_dense_cholesky_noalloc(int n, const double* a_dev, double* l_dev, int* status_dev, void* stream,
                                  std::string& msg);

}  // namespace gfnff

// unique_ptr support
namespace std {
template <>
struct default_delete<gfnff::DeviceEngine> {
    void operator()(gfnff::DeviceEngine* e) const { gfnff::device_destroy(e); }
};
}  // namespace std
