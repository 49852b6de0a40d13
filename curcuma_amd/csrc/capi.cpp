// C-ABI of the engine (include/gfnff_mi355x.h): host handle, set-up on the
// host, evaluation on the GPU through the device engine (device/engine.hip).
//
// Error semantics follow the reference's adapter: a failed EEQ set-up or a
// non-finite energy marks the handle as failed (qm_methods/gfnff_method.cpp:
// 36-48, 83-91; energycalculator.cpp:455-480); nothing throws across the ABI.

#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "gfnff_mi355x.h"
#include "host/gfnff_topology.hpp"
#include "device/engine.hpp"

struct gfnff_handle {
    gfnff::Options opt;
    gfnff::ParamSet params;
    bool have_molecule = false;
    std::vector<double> xyz_bohr;     // current geometry, Bohr
    std::vector<double> last_q, last_cn;
    std::string err;
    bool failed = false;
    int precision = GFNFF_PREC_FP64;
    std::unique_ptr<gfnff::DeviceEngine> dev;
    bool dev_geom_dirty = true;
    // per-element parameter scales (repz, repulsion exponent, sqrt(Z r4r2)) and
    // the unscaled parameters they multiply
    std::vector<double> theta, base_repz, base_anb, base_szr, base_brep;
    // geometry (Bohr) at which the hydrogen/halogen-bond lists were last
    // detected during evaluation; empty until the first evaluation
    std::vector<double> hb_ref;
    // ALPB solvent (canonical name; empty: gas phase), applied to every
    // molecule set afterwards
    std::string solvent;
};

namespace {
int fail(gfnff_handle* h, int code, const std::string& msg) {
    if (h) {
        h->err = msg;
        h->failed = true;
    }
    return code;
}
}  // namespace

extern "C" {

gfnff_handle* gfnff_create(void) { return new (std::nothrow) gfnff_handle(); }

void gfnff_destroy(gfnff_handle* h) {
    try {
        delete h;
    } catch (...) {
    }
}

int gfnff_set_molecule(gfnff_handle* h, int32_t natoms, const int32_t* Z, const double* xyz_ang, int32_t charge,
                       const int32_t* forced_bonds, int32_t nbonds) {
    if (!h) return GFNFF_STATUS_ARGS;
    if (natoms <= 0 || !Z || !xyz_ang || nbonds < 0 || (nbonds > 0 && !forced_bonds))
        return fail(h, GFNFF_STATUS_ARGS, "gfnff_set_molecule: invalid arguments");
    try {
        h->failed = false;
        h->err.clear();
        h->have_molecule = false;
        h->dev.reset();
        h->xyz_bohr.resize((size_t)natoms * 3);
        for (size_t i = 0; i < h->xyz_bohr.size(); ++i) h->xyz_bohr[i] = xyz_ang[i] * gfnff::kAngToBohrGeom;
        std::string msg;
        int st = gfnff::build_params(natoms, Z, h->xyz_bohr.data(), charge, forced_bonds, nbonds, h->opt, h->params,
                                     msg);
        if (st != gfnff::kOk) return fail(h, st, "gfnff_set_molecule: " + msg);
        if (!h->solvent.empty()) {
            st = gfnff::alpb_setup(h->solvent.c_str(), natoms, Z, h->params.solv, msg);
            if (st != gfnff::kOk) return fail(h, st, "gfnff_set_molecule: " + msg);
        }
        h->have_molecule = true;
        h->dev_geom_dirty = true;
        h->hb_ref.clear();
        h->theta.assign(3 * h->params.nelem, 1.0);
        h->base_repz = h->params.repz;
        h->base_anb = h->params.rep_anb;
        h->base_szr = h->params.sqrt_zr4r2;
        h->base_brep = h->params.brep_par;
        h->last_q.assign(natoms, 0.0);
        h->last_cn.assign(natoms, 0.0);
        return GFNFF_STATUS_OK;
    } catch (const std::bad_alloc&) {
        return fail(h, GFNFF_STATUS_ARGS, "gfnff_set_molecule: out of host memory");
    } catch (const std::exception& e) {
        return fail(h, GFNFF_STATUS_ARGS, std::string("gfnff_set_molecule: ") + e.what());
    }
}

int gfnff_set_batch(gfnff_handle* h, int32_t nmol, const int32_t* atom_off, const int32_t* Z, const double* xyz_ang,
                    const int32_t* charges) {
    if (!h) return GFNFF_STATUS_ARGS;
    if (nmol < 1 || !atom_off || !Z || !xyz_ang || atom_off[0] != 0)
        return fail(h, GFNFF_STATUS_ARGS, "gfnff_set_batch: invalid arguments");
    for (int m = 0; m < nmol; ++m)
        if (atom_off[m + 1] <= atom_off[m]) return fail(h, GFNFF_STATUS_ARGS, "gfnff_set_batch: empty molecule");
    if (!h->solvent.empty())
        return fail(h, GFNFF_STATUS_UNSUPPORTED, "gfnff_set_batch: ALPB solvation is built for single molecules");
    try {
        h->failed = false;
        h->err.clear();
        h->have_molecule = false;
        h->dev.reset();
        const int n = atom_off[nmol];
        h->xyz_bohr.resize((size_t)n * 3);
        for (size_t i = 0; i < h->xyz_bohr.size(); ++i) h->xyz_bohr[i] = xyz_ang[i] * gfnff::kAngToBohrGeom;
        std::vector<gfnff::ParamSet> mols(nmol);
        std::vector<int> status(nmol, gfnff::kOk);
        std::vector<std::string> msgs(nmol);
        const int nthr = std::max(1, std::min<int>(nmol, (int)std::max(1u, std::thread::hardware_concurrency())));
        gfnff::Options one = h->opt;  // molecules in parallel, each built single-threaded
        one.threads = 1;
        std::vector<std::thread> pool;
        for (int w = 0; w < nthr; ++w)
            pool.emplace_back([&, w]() {
                for (int m = w; m < nmol; m += nthr) {
                    const int a0 = atom_off[m], na = atom_off[m + 1] - a0;
                    status[m] = gfnff::build_params(na, Z + a0, h->xyz_bohr.data() + 3 * (size_t)a0,
                                                    charges ? charges[m] : 0, nullptr, 0, one, mols[m], msgs[m]);
                }
            });
        for (auto& t : pool) t.join();
        for (int m = 0; m < nmol; ++m)
            if (status[m] != gfnff::kOk)
                return fail(h, status[m], "gfnff_set_batch: molecule " + std::to_string(m) + ": " + msgs[m]);
        std::string msg;
        int st = gfnff::merge_batch(mols, h->params, msg);
        if (st != gfnff::kOk) return fail(h, st, "gfnff_set_batch: " + msg);
        h->have_molecule = true;
        h->dev_geom_dirty = true;
        h->last_q.assign(n, 0.0);
        h->last_cn.assign(n, 0.0);
        h->theta.assign(3 * h->params.nelem, 1.0);
        h->base_repz = h->params.repz;
        h->base_anb = h->params.rep_anb;
        h->base_szr = h->params.sqrt_zr4r2;
        h->base_brep = h->params.brep_par;
        return GFNFF_STATUS_OK;
    } catch (const std::bad_alloc&) {
        return fail(h, GFNFF_STATUS_ARGS, "gfnff_set_batch: out of host memory");
    } catch (const std::exception& e) {
        return fail(h, GFNFF_STATUS_DEVICE, std::string("gfnff_set_batch: ") + e.what());
    }
}

int gfnff_batch_energies(gfnff_handle* h, double* e_mol) {
    if (!h || !e_mol || !h->have_molecule || !h->dev) return GFNFF_STATUS_ARGS;
    std::string msg;
    int st = gfnff::device_batch_energies(h->dev.get(), e_mol, false, nullptr, msg);
    return st ? fail(h, st, "gfnff_batch_energies: " + msg) : GFNFF_STATUS_OK;
}

int gfnff_batch_energies_device(gfnff_handle* h, double* e_mol_dev, void* stream) {
    if (!h || !e_mol_dev || !h->have_molecule || !h->dev) return GFNFF_STATUS_ARGS;
    std::string msg;
    int st = gfnff::device_batch_energies(h->dev.get(), e_mol_dev, true, stream, msg);
    return st ? fail(h, st, "gfnff_batch_energies_device: " + msg) : GFNFF_STATUS_OK;
}

static int refresh_pairs_if_needed(gfnff_handle* h);

int gfnff_update_geometry(gfnff_handle* h, const double* xyz_ang) {
    if (!h) return GFNFF_STATUS_ARGS;
    if (!h->have_molecule || !xyz_ang) return fail(h, GFNFF_STATUS_ARGS, "gfnff_update_geometry: no molecule");
    for (size_t i = 0; i < h->xyz_bohr.size(); ++i) {
        if (!std::isfinite(xyz_ang[i])) return fail(h, GFNFF_STATUS_GEOMETRY, "gfnff_update_geometry: non-finite");
        h->xyz_bohr[i] = xyz_ang[i] * gfnff::kAngToBohrGeom;
    }
    h->dev_geom_dirty = true;
    try {
        return refresh_pairs_if_needed(h);
    } catch (const std::exception& e) {
        return fail(h, GFNFF_STATUS_ARGS, std::string("gfnff_update_geometry: ") + e.what());
    }
}

// Re-detects the hydrogen/halogen-bond lists when the geometry has moved by
// sqrt(sum dx^2)/N > hb_update_rmsd_bohr since the last detection, and on the
// first evaluation (shouldUpdateHBXB + updateHBXBIfNeeded,
// gfnff_method.cpp:2228-2259, 1597-1690).  A changed list rebuilds the device
// engine's term lists.  Batches keep their set-up lists.
static int refresh_pairs_if_needed(gfnff_handle* h) {
    if (h->params.nmol != 1) return GFNFF_STATUS_OK;
    const size_t n3 = h->xyz_bohr.size();
    if (h->hb_ref.size() == n3) {
        double s = 0.0;
        for (size_t i = 0; i < n3; ++i) s += (h->xyz_bohr[i] - h->hb_ref[i]) * (h->xyz_bohr[i] - h->hb_ref[i]);
        if (!(std::sqrt(s) / (double)h->params.n > h->opt.hb_update_rmsd_bohr)) return GFNFF_STATUS_OK;
    }
    bool changed = false;
    std::string msg;
    int st = gfnff::refresh_hbxb(h->params, h->xyz_bohr.data(), h->opt, changed, msg);
    if (st) return fail(h, st, "hydrogen-bond list update: " + msg);
    h->hb_ref = h->xyz_bohr;
    if (changed) h->dev.reset();
    return GFNFF_STATUS_OK;
}

static int ensure_device(gfnff_handle* h) {
    if (h->dev) return GFNFF_STATUS_OK;
    std::string msg;
    int st = GFNFF_STATUS_OK;
    h->dev.reset(gfnff::device_create(h->params, h->precision, h->opt, st, msg));
    if (!h->dev) return fail(h, st ? st : GFNFF_STATUS_DEVICE, "device set-up: " + msg);
    return GFNFF_STATUS_OK;
}

int gfnff_energy_gradient(gfnff_handle* h, int32_t want_grad, double* energy, double* grad, double* terms) {
    if (!h) return GFNFF_STATUS_ARGS;
    if (!h->have_molecule) return fail(h, GFNFF_STATUS_ARGS, "gfnff_energy_gradient: no molecule");
    if (!energy) return fail(h, GFNFF_STATUS_ARGS, "gfnff_energy_gradient: energy pointer is NULL");
    try {
        int st = refresh_pairs_if_needed(h);
        if (st) return st;
        st = ensure_device(h);
        if (st) return st;
        std::string msg;
        double e_terms[GFNFF_NTERMS + 1];
        std::vector<double> g(want_grad ? h->xyz_bohr.size() : 0);
        st = gfnff::device_eval_host(h->dev.get(), h->xyz_bohr.data(), want_grad != 0, e_terms,
                                     want_grad ? g.data() : nullptr, h->last_q.data(), h->last_cn.data(), msg);
        if (st) return fail(h, st, "gfnff_energy_gradient: " + msg);
        h->failed = false;
        if (!std::isfinite(e_terms[0])) {
            *energy = 0.0;
            return fail(h, GFNFF_STATUS_NAN, "gfnff_energy_gradient: energy is not finite");
        }
        *energy = e_terms[0];
        if (terms) std::memcpy(terms, e_terms + 1, sizeof(double) * GFNFF_NTERMS);
        if (want_grad && grad) {
            bool finite = true;
            for (double v : g) finite = finite && std::isfinite(v);
            std::memcpy(grad, g.data(), sizeof(double) * g.size());
            if (!finite) return fail(h, GFNFF_STATUS_NAN, "gfnff_energy_gradient: gradient is not finite");
        }
        return GFNFF_STATUS_OK;
    } catch (const std::exception& e) {
        return fail(h, GFNFF_STATUS_DEVICE, std::string("gfnff_energy_gradient: ") + e.what());
    }
}

int gfnff_energy_gradient_device(gfnff_handle* h, const double* xyz_bohr_dev, double* energy_dev, double* grad_dev,
                                 void* stream) {
    if (!h) return GFNFF_STATUS_ARGS;
    if (!h->have_molecule || !xyz_bohr_dev || !energy_dev)
        return fail(h, GFNFF_STATUS_ARGS, "gfnff_energy_gradient_device: invalid arguments");
    try {
        int st = ensure_device(h);
        if (st) return st;
        std::string msg;
        st = gfnff::device_eval(h->dev.get(), xyz_bohr_dev, energy_dev, grad_dev, stream, msg);
        if (st) return fail(h, st, "gfnff_energy_gradient_device: " + msg);
        return GFNFF_STATUS_OK;
    } catch (const std::exception& e) {
        return fail(h, GFNFF_STATUS_DEVICE, std::string("gfnff_energy_gradient_device: ") + e.what());
    }
}

int gfnff_graph_capture(gfnff_handle* h, const double* xyz_bohr_dev, double* energy_dev, double* grad_dev,
                        void* stream) {
    if (!h) return GFNFF_STATUS_ARGS;
    if (!h->have_molecule || !xyz_bohr_dev || !energy_dev || !stream)
        return fail(h, GFNFF_STATUS_ARGS, "gfnff_graph_capture: invalid arguments");
    try {
        int st = ensure_device(h);
        if (st) return st;
        std::string msg;
        st = gfnff::device_graph_capture(h->dev.get(), xyz_bohr_dev, energy_dev, grad_dev, stream, msg);
        if (st) return fail(h, st, "gfnff_graph_capture: " + msg);
        return GFNFF_STATUS_OK;
    } catch (const std::exception& e) {
        return fail(h, GFNFF_STATUS_DEVICE, std::string("gfnff_graph_capture: ") + e.what());
    }
}

int gfnff_md_steps(gfnff_handle* h, double* xyz_bohr_dev, double* vel_dev, double* grad_dev, const double* mass_amu,
                   double dt_fs, double T0, double tau_fs, int32_t nsteps, double* record_dev, void* stream) {
    if (!h) return GFNFF_STATUS_ARGS;
    if (!h->have_molecule || !xyz_bohr_dev || !vel_dev || !grad_dev || !mass_amu || !record_dev || nsteps < 0 ||
        !(dt_fs > 0.0) || (T0 > 0.0 && !(tau_fs > 0.0)))
        return fail(h, GFNFF_STATUS_ARGS, "gfnff_md_steps: invalid arguments");
    try {
        int st = ensure_device(h);
        if (st) return st;
        // the coordinates of the last hydrogen/halogen-bond detection (set-up or update_geometry)
        const std::vector<double>& ref = h->hb_ref.size() == h->xyz_bohr.size() ? h->hb_ref : h->xyz_bohr;
        std::string msg;
        st = gfnff::device_md_steps(h->dev.get(), xyz_bohr_dev, vel_dev, grad_dev, mass_amu, ref.data(), dt_fs, T0,
                                    tau_fs, h->opt.hb_update_rmsd_bohr, nsteps, record_dev, stream, msg);
        if (st) return fail(h, st, "gfnff_md_steps: " + msg);
        return GFNFF_STATUS_OK;
    } catch (const std::exception& e) {
        return fail(h, GFNFF_STATUS_DEVICE, std::string("gfnff_md_steps: ") + e.what());
    }
}

int gfnff_graph_launch(gfnff_handle* h, void* stream) {
    if (!h) return GFNFF_STATUS_ARGS;
    if (!h->dev) return fail(h, GFNFF_STATUS_ARGS, "gfnff_graph_launch: no captured evaluation");
    std::string msg;
    const int st = gfnff::device_graph_launch(h->dev.get(), stream, msg);
    return st ? fail(h, st, "gfnff_graph_launch: " + msg) : GFNFF_STATUS_OK;
}

int gfnff_device_status(gfnff_handle* h, int32_t* flags, int32_t clear) {
    if (!h || !flags) return GFNFF_STATUS_ARGS;
    *flags = 0;
    if (!h->dev) return GFNFF_STATUS_OK;
    try {
        std::string msg;
        int f = 0;
        int st = gfnff::device_status(h->dev.get(), &f, clear != 0, msg);
        if (st) return fail(h, st, "gfnff_device_status: " + msg);
        *flags = f;
        return GFNFF_STATUS_OK;
    } catch (const std::exception& e) {
        return fail(h, GFNFF_STATUS_DEVICE, std::string("gfnff_device_status: ") + e.what());
    }
}

int gfnff_charges(gfnff_handle* h, double* q) {
    if (!h || !q || !h->have_molecule) return GFNFF_STATUS_ARGS;
    std::memcpy(q, h->last_q.data(), sizeof(double) * h->last_q.size());
    return GFNFF_STATUS_OK;
}

int gfnff_cn(gfnff_handle* h, double* cn) {
    if (!h || !cn || !h->have_molecule) return GFNFF_STATUS_ARGS;
    std::memcpy(cn, h->last_cn.data(), sizeof(double) * h->last_cn.size());
    return GFNFF_STATUS_OK;
}

int gfnff_set_param(gfnff_handle* h, const char* key, double value) {
    if (!h || !key) return GFNFF_STATUS_ARGS;
    std::string k(key);
    if (k == "cn_cutoff_bohr") {
        if (value != 6.0) return fail(h, GFNFF_STATUS_UNSUPPORTED, "cn_cutoff_bohr: only the default 6.0 is built");
        h->opt.cn_cutoff_bohr = value;
    } else if (k == "hb_accuracy") {
        if (!(value > 0.0)) return fail(h, GFNFF_STATUS_ARGS, "hb_accuracy must be > 0");
        h->opt.hb_accuracy = value;
    } else if (k == "threads") {
        h->opt.threads = value < 1 ? 1 : (int)value;
    } else if (k == "hb_update_rmsd_bohr") {
        if (!(value >= 0.0)) return fail(h, GFNFF_STATUS_ARGS, "hb_update_rmsd_bohr must be >= 0");
        h->opt.hb_update_rmsd_bohr = value;
    } else if (k == "eeq_solver") {
        if (value != 0.0 && value != 1.0) return fail(h, GFNFF_STATUS_ARGS, "eeq_solver: 0 (PCG) or 1 (Cholesky)");
        h->opt.eeq_solver = (int)value;
        h->dev.reset();
    } else if (k == "pcg_tol") {
        if (!(value >= 0.0 && value <= 1e-6)) return fail(h, GFNFF_STATUS_ARGS, "pcg_tol: 0 (default) or in (0, 1e-6]");
        h->opt.pcg_tol = value;
        h->dev.reset();
    } else if (k == "cholesky_lookahead") {
        h->opt.cholesky_lookahead = value != 0.0;
        h->dev.reset();
    } else if (k == "batch_solver") {
        if (!(value >= 0.0 && value <= 6.0) || value != std::floor(value))
            return fail(h, GFNFF_STATUS_ARGS, "batch_solver: 0 .. 6");
        h->opt.batch_solver = (int)value;
        h->dev.reset();
    } else if (k == "batch_fp32") {
        if (value != 0.0 && value != 1.0) return fail(h, GFNFF_STATUS_ARGS, "batch_fp32: 0 or 1");
        h->opt.batch_fp32 = (int)value;
        h->dev.reset();
    } else if (k == "pair_split") {
        if (value != 0.0 && value != 1.0) return fail(h, GFNFF_STATUS_ARGS, "pair_split: 0 or 1");
        h->opt.pair_split = (int)value;
        h->dev.reset();
    } else if (k == "batch_streams") {
        if (value != 1.0 && value != 2.0) return fail(h, GFNFF_STATUS_ARGS, "batch_streams: 1 or 2");
        h->opt.batch_streams = (int)value;
        h->dev.reset();
    } else if (k == "batch_pairs") {
        if (value != 0.0 && value != 1.0) return fail(h, GFNFF_STATUS_ARGS, "batch_pairs: 0 or 1");
        h->opt.batch_pairs = (int)value;
        h->dev.reset();
    } else if (k == "gather_lanes") {
        if (value != 0.0 && value != 16.0 && value != 64.0) return fail(h, GFNFF_STATUS_ARGS, "gather_lanes: 0, 16 or 64");
        h->opt.gather_lanes = (int)value;
        h->dev.reset();
    } else if (k == "pcg_groups") {
        if (!(value >= 0.0 && value <= 256.0) || value != std::floor(value))
            return fail(h, GFNFF_STATUS_ARGS, "pcg_groups: 0 (all that fit) or 1..256");
        h->opt.pcg_groups = (int)value;
        h->dev.reset();
    } else {
        return fail(h, GFNFF_STATUS_UNSUPPORTED, "unknown parameter: " + k);
    }
    return GFNFF_STATUS_OK;
}

int gfnff_set_precision(gfnff_handle* h, int32_t mode) {
    if (!h) return GFNFF_STATUS_ARGS;
    if (mode != GFNFF_PREC_FP64 && mode != GFNFF_PREC_MIXED) return fail(h, GFNFF_STATUS_ARGS, "unknown precision");
    if (mode != h->precision) {
        h->precision = mode;
        h->dev.reset();
    }
    return GFNFF_STATUS_OK;
}

int gfnff_set_solvent(gfnff_handle* h, const char* solvent) {
    if (!h) return GFNFF_STATUS_ARGS;
    try {
        const std::string req = solvent ? solvent : "";
        std::string name;
        if (!req.empty() && req != "none") {
            name = gfnff::alpb_solvent_name(solvent);
            if (name.empty()) {
                // the reference continues in the gas phase after an unknown
                // solvent (gfnff_method.cpp:826-829); so does the handle
                h->solvent.clear();
                h->params.solv = gfnff::AlpbSet();
                h->dev.reset();
                return fail(h, GFNFF_STATUS_ARGS, "gfnff_set_solvent: unknown solvent '" + req + "'");
            }
        }
        if (!name.empty() && h->have_molecule && h->params.nmol != 1)
            return fail(h, GFNFF_STATUS_UNSUPPORTED, "gfnff_set_solvent: ALPB solvation is built for single molecules");
        h->solvent = name;
        h->params.solv = gfnff::AlpbSet();
        if (!name.empty() && h->have_molecule) {
            std::string msg;
            int st = gfnff::alpb_setup(name.c_str(), h->params.n, h->params.Z.data(), h->params.solv, msg);
            if (st != gfnff::kOk) return fail(h, st, "gfnff_set_solvent: " + msg);
        }
        h->dev.reset();
        return GFNFF_STATUS_OK;
    } catch (const std::exception& e) {
        return fail(h, GFNFF_STATUS_ARGS, std::string("gfnff_set_solvent: ") + e.what());
    }
}

int gfnff_get_solvation(const gfnff_handle* h, gfnff_alpb_view* out) {
    if (!h || !out || !h->have_molecule) return GFNFF_STATUS_ARGS;
    *out = h->params.solv.view();
    return GFNFF_STATUS_OK;
}

int gfnff_solvation_energy(gfnff_handle* h, double* parts) {
    if (!h || !parts) return GFNFF_STATUS_ARGS;
    if (!h->have_molecule || !h->dev || h->params.solv.n == 0)
        return fail(h, GFNFF_STATUS_ARGS, "gfnff_solvation_energy: no solvated evaluation");
    std::string msg;
    int st = gfnff::device_solvation_parts(h->dev.get(), parts, msg);
    return st ? fail(h, st, "gfnff_solvation_energy: " + msg) : GFNFF_STATUS_OK;
}

int gfnff_has_error(const gfnff_handle* h) { return h ? (h->failed ? 1 : 0) : 1; }

const char* gfnff_last_error(const gfnff_handle* h) { return h ? h->err.c_str() : "null handle"; }

int gfnff_get_params(const gfnff_handle* h, gfnff_params_view* out) {
    if (!h || !out || !h->have_molecule) return GFNFF_STATUS_ARGS;
    *out = h->params.view();
    return GFNFF_STATUS_OK;
}

int gfnff_get_topology(const gfnff_handle* h, int32_t* hyb, int32_t* pi, int32_t* bonds, int32_t nbonds_max,
                       int32_t* nbonds) {
    if (!h || !h->have_molecule) return GFNFF_STATUS_ARGS;
    const auto& P = h->params;
    if (hyb) std::memcpy(hyb, P.hyb.data(), sizeof(int32_t) * P.n);
    if (pi) std::memcpy(pi, P.pi_frag.data(), sizeof(int32_t) * P.n);
    int nb = (int)P.topo_bonds.size() / 2;
    if (nbonds) *nbonds = nb;
    if (bonds) std::memcpy(bonds, P.topo_bonds.data(), sizeof(int32_t) * 2 * std::min(nb, (int)nbonds_max));
    return GFNFF_STATUS_OK;
}

int gfnff_device_count(void) { return gfnff::device_count(); }

int gfnff_solver_iterations(gfnff_handle* h) {
    if (!h || !h->dev) return -1;
    return gfnff::device_solver_iterations(h->dev.get());
}

int gfnff_cn_neighbours(gfnff_handle* h, int32_t* counts, int32_t* idx, int32_t maxnb) {
    if (!h || !counts || !idx || maxnb <= 0 || !h->dev) return GFNFF_STATUS_ARGS;
    std::string msg;
    int st = gfnff::device_cn_neighbours(h->dev.get(), counts, idx, maxnb, msg);
    return st ? fail(h, st, "gfnff_cn_neighbours: " + msg) : GFNFF_STATUS_OK;
}

int gfnff_profile_kernel(gfnff_handle* h, const char* name) {
    if (!h || !h->have_molecule) return GFNFF_STATUS_ARGS;
    int st = ensure_device(h);
    if (st) return st;
    std::string msg;
    st = gfnff::device_profile_select(h->dev.get(), name, msg);
    return st ? fail(h, st, "gfnff_profile_kernel: " + msg) : GFNFF_STATUS_OK;
}

int gfnff_profile_read(gfnff_handle* h, double* total_ms, int32_t* launches) {
    if (!h || !total_ms || !launches || !h->dev) return GFNFF_STATUS_ARGS;
    std::string msg;
    int nl = 0;
    int st = gfnff::device_profile_read(h->dev.get(), total_ms, &nl, msg);
    *launches = nl;
    return st ? fail(h, st, "gfnff_profile_read: " + msg) : GFNFF_STATUS_OK;
}

int gfnff_param_count(const gfnff_handle* h) {
    if (!h || !h->have_molecule) return -1;
    return (int)h->theta.size();
}

int gfnff_param_scales(gfnff_handle* h, const double* theta, int32_t ntheta) {
    if (!h || !theta || !h->have_molecule) return GFNFF_STATUS_ARGS;
    if (ntheta != (int32_t)h->theta.size())
        return fail(h, GFNFF_STATUS_ARGS, "gfnff_param_scales: expected 3 * nelem scales");
    for (int k = 0; k < ntheta; ++k)
        if (!(theta[k] > 0.0) || !std::isfinite(theta[k]))
            return fail(h, GFNFF_STATUS_ARGS, "gfnff_param_scales: scales must be positive and finite");
    auto& P = h->params;
    const int ne = P.nelem;
    h->theta.assign(theta, theta + ntheta);
    for (int i = 0; i < P.n; ++i) {
        const int e = P.elem[i];
        P.repz[i] = h->base_repz[i] * theta[e];
        P.rep_anb[i] = h->base_anb[i] * theta[ne + e];
        P.sqrt_zr4r2[i] = h->base_szr[i] * theta[2 * ne + e];
    }
    const int nb = (int)P.brep_par.size() / GFNFF_REP_NPAR;
    for (int b = 0; b < nb; ++b) {
        const int ei = P.elem[P.brep_ij[2 * b]], ej = P.elem[P.brep_ij[2 * b + 1]];
        P.brep_par[2 * b] = h->base_brep[2 * b] * std::sqrt(theta[ne + ei] * theta[ne + ej]);
        P.brep_par[2 * b + 1] = h->base_brep[2 * b + 1] * theta[ei] * theta[ej];
    }
    if (h->dev) {
        std::string msg;
        int st = gfnff::device_update_params(h->dev.get(), P, msg);
        if (st) return fail(h, st, "gfnff_param_scales: " + msg);
    }
    return GFNFF_STATUS_OK;
}

int gfnff_param_adjoint(gfnff_handle* h, const double* w, double* dtheta, int32_t ntheta) {
    if (!h || !w || !dtheta || !h->have_molecule) return GFNFF_STATUS_ARGS;
    if (!h->dev) return fail(h, GFNFF_STATUS_ARGS, "gfnff_param_adjoint: no evaluation yet");
    if (ntheta != (int32_t)h->theta.size())
        return fail(h, GFNFF_STATUS_ARGS, "gfnff_param_adjoint: expected 3 * nelem outputs");
    try {
        std::string msg;
        std::vector<double> dln(ntheta);
        int st = gfnff::device_param_adjoint(h->dev.get(), w, dln.data(), msg);
        if (st) return fail(h, st, "gfnff_param_adjoint: " + msg);
        for (int k = 0; k < ntheta; ++k) dtheta[k] = dln[k] / h->theta[k];
        return GFNFF_STATUS_OK;
    } catch (const std::exception& e) {
        return fail(h, GFNFF_STATUS_DEVICE, std::string("gfnff_param_adjoint: ") + e.what());
    }
}

int gfnff_dense_cholesky(int32_t n, const double* a, double* l_out) {
    if (n <= 0 || !a || !l_out) return GFNFF_STATUS_ARGS;
    std::string msg;
    try {
        return gfnff::device_dense_cholesky(n, a, l_out, msg);
    } catch (...) {
        return GFNFF_STATUS_DEVICE;
    }
}

int gfnff_dense_cholesky_noalloc(int32_t n, const double* a_dev, double* l_dev, int32_t* status_dev, void* stream) {
    if (n <= 0 || n > GFNFF_NOALLOC_MAX_N || !a_dev || !l_dev || !status_dev) return GFNFF_STATUS_ARGS;
    std::string msg;
    try {
        return gfnff::device_dense_cholesky_noalloc(n, a_dev, l_dev, status_dev, stream, msg);
    } catch (...) {
        return GFNFF_STATUS_DEVICE;
    }
}

int gfnff_dense_cholesky_f32(int32_t n, const float* a, float* l_out) {
    if (n <= 0 || !a || !l_out) return GFNFF_STATUS_ARGS;
    std::string msg;
    try {
        return gfnff::device_dense_cholesky_f32(n, a, l_out, msg);
    } catch (...) {
        return GFNFF_STATUS_DEVICE;
    }
}

}  // extern "C"
