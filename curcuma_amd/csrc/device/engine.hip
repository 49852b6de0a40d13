// GFN-FF energy + gradient on one MI355X (gfx950): FP64 path and mixed path,
// single molecules and batches of independent molecules.
//
// One evaluation is a fixed sequence of kernels on two streams (launch_eval):
//   second stream: the reduced EEQ matrix (k_eeq_reduced / k_eeq_reduced_batch)
//     and, for batches, the per-molecule factorisation (k_lstep_tiles,
//     k_bstep_diag, k_bstep_tiles or the per-molecule forms), then the
//     hydrogen-bond kernels of batches;
//   evaluation stream:
//   k_cn / k_cn_mol     CN (6 Bohr), log cap, dlogdcn, D4 weights, EEQ rhs, CN neighbour lists
//   k_d4_half           per-atom D4 half contraction of C6 and dC6/dCN
//   k_bonds .. k_batm   bonded term lists writing per-atom gradient slots
//   k_slot_gather       per-atom sum of the slots
//   k_hb_*              hydrogen bonds (single molecules on this stream)
//   -- join --
//   EEQ solve           k_eeq_pcg / k_chol_* + k_trsv_* (single), k_bsolve_subst (batches)
//   k_charges           charges from the reduced solution, self energy, qtmp
//   k_pairs_sym         Coulomb + D4 dispersion + repulsion, each pair once (64 x 64 tiles)
//   k_pairs_sym_reduce  fixed-order sum of the tile partials (no atomics)
//   k_atom_tail         bonded / 1,3 / 1,4 H-H corrections of the repulsion pair sum, dE/dCN totals
//   k_finish            CN chain rule + gradient assembly
//   k_reduce1/_final    fixed-order energy reductions
// Reference semantics: see oracle/gfnff_oracle.c (same formulas, same cut-offs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "engine.hpp"

namespace gfnff {

namespace {

#define HIP_OK(x)                                                   \
    do {                                                            \
        hipError_t e_ = (x);                                        \
        if (e_ != hipSuccess) {                                     \
            msg = std::string(#x) + ": " + hipGetErrorString(e_);   \
            return kErrDevice;                                      \
        }                                                           \
    } while (0)

constexpr int NB = 64;          // Cholesky block size
#define MFMA_F32_ROW_FROM_REG false  // FP32 16x16x4 result layout (see MfmaT)
constexpr int MAXNB_CN = 128;   // CN neighbour list capacity per atom
constexpr int TILE = 256;       // pair kernel tile / workgroup size

constexpr double KN = -7.5;
// CN count 0.5 (1 + erf(v)); for v <= -6, erf(v) rounds to -1 and the count
// is 0 (the branch skips erf for the many list members beyond the covalent
// distance; also the hydrogen-bond coordination count of X-H bonds)
__device__ inline double cn_count(double v) {
    double c = 0.0;
    if (v > -6.0) c = 0.5 * (1.0 + erf(v));
    return c;
}
constexpr double CNMAX = 4.4;
constexpr double SQRTPI_TRUNC = 1.77245385091;   // constant used by the reference's dCN
constexpr double TSQRT2PI = 0.797884560802866;
constexpr double TSQRT2PI_SELF = 0.797884560802865;

enum Flag { FLAG_CHOL = 1, FLAG_CNOVF = 2, FLAG_SOLVE = 4 };

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ inline double3 ld3(const double* x, int i) { return make_double3(x[3 * i], x[3 * i + 1], x[3 * i + 2]); }
__device__ inline double sq(double v) { return v * v; }
__device__ inline double cube(double v) { return v * v * v; }
__device__ inline double3 operator-(double3 a, double3 b) { return make_double3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline double3 operator+(double3 a, double3 b) { return make_double3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline double3 operator*(double s, double3 a) { return make_double3(s * a.x, s * a.y, s * a.z); }
__device__ inline double dot(double3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline double3 cross(double3 a, double3 b) {
    return make_double3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline double nrm(double3 a) { return sqrt(dot(a, a)); }
__device__ inline void st3(double* s, int slot, double3 v) {
    s[3 * slot] = v.x;
    s[3 * slot + 1] = v.y;
    s[3 * slot + 2] = v.z;
}

// ---------------------------------------------------------------------------
// K1: CN, log cap, dlogdcn, D4 weights, EEQ right-hand side, neighbour list
// cn_calculator.cpp:367-428; gfnff_method.cpp:5880-5930; d4param_generator.cpp:1079-1213,1692-1768
// ---------------------------------------------------------------------------
// One wavefront per atom i (four atoms per workgroup share the LDS j tiles).
// Lanes test 64 partners j at a time; __ballot and a prefix popcount give the
// position of every hit, so each neighbour list is ascending in j with the
// strict d^2 < 36 compare of the reference's builder (cn_calculator.cpp:
// 367-428).  The raw CN is summed per lane and combined by a fixed butterfly.
__device__ void cn_atom_finish(int i, double raw, int cnt, const int* __restrict__ elem, const int* __restrict__ nref,
                               const double* __restrict__ refcn, const double* __restrict__ chi_base,
                               const double* __restrict__ cnf, double* __restrict__ cn, double* __restrict__ dlog,
                               double* __restrict__ gw, double* __restrict__ dgw, double* __restrict__ rhs,
                               int* __restrict__ nbn, int* __restrict__ flags);

// exact bounding boxes of the 256-atom tiles of the atom order (k_cn's
// partner tiles): bb[6 t] = (min x, min y, min z, max x, max y, max z)
__global__ __launch_bounds__(256) void k_tile_box(int n, const double* __restrict__ x, double* __restrict__ bb) {
    __shared__ double red[6][256];
    const int t = blockIdx.x, tid = threadIdx.x, j = min(t * 256 + tid, n - 1);
    const double3 p = ld3(x, j);
    red[0][tid] = p.x;
    red[1][tid] = p.y;
    red[2][tid] = p.z;
    red[3][tid] = p.x;
    red[4][tid] = p.y;
    red[5][tid] = p.z;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                red[d][tid] = fmin(red[d][tid], red[d][tid + w]);
                red[3 + d][tid] = fmax(red[3 + d][tid], red[3 + d][tid + w]);
            }
        }
        __syncthreads();
    }
    if (tid < 6) bb[6 * t + tid] = red[tid][0];
}

__global__ __launch_bounds__(256) void k_cn(int n, const double* __restrict__ x, const double* __restrict__ rcov,
                                            const int* __restrict__ elem, const int* __restrict__ nref,
                                            const double* __restrict__ refcn, const double* __restrict__ chi_base,
                                            const double* __restrict__ cnf, double* __restrict__ cn,
                                            double* __restrict__ dlog, double* __restrict__ gw,
                                            double* __restrict__ dgw, double* __restrict__ rhs,
                                            int* __restrict__ nbn, int* __restrict__ nbl, int* __restrict__ flags,
                                            const int* __restrict__ mol_of, const int* __restrict__ mol_off,
                                            const double* __restrict__ tbox) {
    __shared__ double sx[256], sy[256], sz[256], sr[256];
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const bool act = i < n;
    double3 xi = act ? ld3(x, i) : make_double3(0, 0, 0);
    double ri = act ? rcov[i] : 0.0;
    double raw = 0.0;
    int cnt = 0;
    // partner range: all atoms, or in a batch the molecules of this block's atoms
    int jlo = 0, jhi = n, mlo = 0, mhi = n;
    if (mol_of) {
        const int first = blockIdx.x * 4, last = min(n, first + 4) - 1;
        jlo = mol_off[mol_of[first]];
        jhi = mol_off[mol_of[last] + 1];
        if (act) {
            mlo = mol_off[mol_of[i]];
            mhi = mol_off[mol_of[i] + 1];
        }
    }
    for (int t0 = jlo; t0 < jhi; t0 += 256) {
        // single molecule: skip a partner tile whose bounding box (tbox, the
        // 256-atom tiles of the atom order) is beyond the cut-off for all four
        // atoms; the margin covers the rounding of both distance tests
        bool near = true;
        if (!mol_of) {
            const double* bx = tbox + 6 * (t0 >> 8);
            const double ex = fmax(0.0, fmax(bx[0] - xi.x, xi.x - bx[3]));
            const double ey = fmax(0.0, fmax(bx[1] - xi.y, xi.y - bx[4]));
            const double ez = fmax(0.0, fmax(bx[2] - xi.z, xi.z - bx[5]));
            near = act && ex * ex + ey * ey + ez * ez <= 36.0 * (1.0 + 1e-9);
        }
        if (!__syncthreads_or(near)) continue;
        int jl = t0 + threadIdx.x;
        if (jl < jhi) {
            sx[threadIdx.x] = x[3 * jl];
            sy[threadIdx.x] = x[3 * jl + 1];
            sz[threadIdx.x] = x[3 * jl + 2];
            sr[threadIdx.x] = rcov[jl];
        }
        __syncthreads();
        const int m = min(256, jhi - t0);
        for (int s0 = 0; s0 < m; s0 += 64) {
            const int jj = s0 + lane, j = t0 + jj;
            bool hit = false;
            double contrib = 0.0;
            if (act && jj < m && j != i && j >= mlo && j < mhi) {
                double dx = xi.x - sx[jj], dy = xi.y - sy[jj], dz = xi.z - sz[jj];
                // unfused products and sums: the list membership test matches
                // the reference's ((dx*dx + dy*dy) + dz*dz) < 36 bit for bit
                double d2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
                if (d2 < 36.0) {
                    hit = true;
                    double r = sqrt(d2);
                    double rij = ri + sr[jj];
                    contrib = cn_count(KN * (r - rij) / rij);
                }
            }
            const unsigned long long mask = __ballot(hit);
            if (hit) {
                int pos = cnt + __popcll(mask & ((1ull << lane) - 1ull));
                if (pos < MAXNB_CN) nbl[(size_t)i * MAXNB_CN + pos] = j;
                raw += contrib;
            }
            cnt += __popcll(mask);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) raw += __shfl_xor(raw, off);
    if (!act || lane != 0) return;
    cn_atom_finish(i, raw, cnt, elem, nref, refcn, chi_base, cnf, cn, dlog, gw, dgw, rhs, nbn, flags);
}

// Per-atom part of the CN kernels once the neighbour scan is done: list
// length, log-capped CN and its derivative, D4 weights, EEQ right-hand side.
__device__ void cn_atom_finish(int i, double raw, int cnt, const int* __restrict__ elem, const int* __restrict__ nref,
                               const double* __restrict__ refcn, const double* __restrict__ chi_base,
                               const double* __restrict__ cnf, double* __restrict__ cn, double* __restrict__ dlog,
                               double* __restrict__ gw, double* __restrict__ dgw, double* __restrict__ rhs,
                               int* __restrict__ nbn, int* __restrict__ flags) {
    if (cnt > MAXNB_CN) atomicOr(flags, FLAG_CNOVF);
    nbn[i] = min(cnt, MAXNB_CN);
    double c = log(1.0 + exp(CNMAX)) - log(1.0 + exp(CNMAX - raw));
    cn[i] = c;
    dlog[i] = exp(CNMAX) / (exp(CNMAX) + exp(raw));
    int e = elem[i], nr = nref[e];
    double ex[7], dex[7], norm = 0.0, dnorm = 0.0;
#pragma unroll
    for (int r = 0; r < 7; ++r) {
        ex[r] = 0.0;
        dex[r] = 0.0;
        if (r < nr) {
            double df = c - refcn[e * 7 + r];
            ex[r] = exp(-4.0 * df * df);
            dex[r] = -8.0 * df * ex[r];
            norm += ex[r];
            dnorm += dex[r];
        }
    }
#pragma unroll
    for (int r = 0; r < 7; ++r) {
        double g = 0.0, dg = 0.0;
        if (norm > 1e-10) {
            if (r < nr) {
                g = ex[r] / norm;
                dg = (dex[r] * norm - ex[r] * dnorm) / (norm * norm);
            }
        } else if (r == 0 && nr > 0) {
            g = 1.0;
        }
        gw[(size_t)i * 7 + r] = g;
        dgw[(size_t)i * 7 + r] = dg;
    }
    rhs[i] = chi_base[i] + cnf[i] * sqrt(c);
}

// CN for a batch of molecules of at most CNMOL_MAX atoms each: one
// workgroup per molecule stages the molecule's coordinates and radii once;
// wave w scans atoms w, w + 4, ... against every atom of the molecule with
// the ballot-ordered list construction of k_cn (ascending partners) and
// leaves the raw CN and list length in LDS; the per-atom epilogue (log cap,
// D4 weights, EEQ right-hand side) then runs one thread per atom.  The LDS
// arrays are sized at launch for the batch's largest molecule (cap atoms,
// 44 bytes each), so that workgroups of small-molecule batches share a CU
// many at a time and hide the latency of each wavefront's scan.
constexpr int CNMOL_MAX = 1024;
constexpr int CNMOL_LDS_BYTES = 5 * 8 + 4;  // per atom of the cap
__global__ __launch_bounds__(256, 6) void k_cn_mol(int cap, const int* __restrict__ mol_off, const double* __restrict__ x,
                                                const double* __restrict__ rcov, const int* __restrict__ elem,
                                                const int* __restrict__ nref, const double* __restrict__ refcn,
                                                const double* __restrict__ chi_base, const double* __restrict__ cnf,
                                                double* __restrict__ cn, double* __restrict__ dlog,
                                                double* __restrict__ gw, double* __restrict__ dgw,
                                                double* __restrict__ rhs, int* __restrict__ nbn,
                                                int* __restrict__ nbl, int* __restrict__ flags) {
    extern __shared__ double cnm_lds[];
    double *sx = cnm_lds, *sy = sx + cap, *sz = sy + cap, *sr = sz + cap, *sraw = sr + cap;
    int* scnt = (int*)(sraw + cap);
    const int m = blockIdx.x, a0 = mol_off[m], nm = mol_off[m + 1] - a0;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int k = tid; k < nm; k += 256) {
        sx[k] = x[3 * (a0 + k)];
        sy[k] = x[3 * (a0 + k) + 1];
        sz[k] = x[3 * (a0 + k) + 2];
        sr[k] = rcov[a0 + k];
    }
    __syncthreads();
    for (int ii = w; ii < nm; ii += 4) {
        const int i = a0 + ii;
        const double xix = sx[ii], xiy = sy[ii], xiz = sz[ii], ri = sr[ii];
        double raw = 0.0;
        int cnt = 0;
        for (int s0 = 0; s0 < nm; s0 += 64) {
            const int jj = s0 + lane;
            bool hit = false;
            double contrib = 0.0;
            if (jj < nm && jj != ii) {
                const double dx = xix - sx[jj], dy = xiy - sy[jj], dz = xiz - sz[jj];
                // unfused products and sums, as in k_cn
                const double d2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
                if (d2 < 36.0) {
                    hit = true;
                    const double r = sqrt(d2);
                    const double rij = ri + sr[jj];
                    contrib = cn_count(KN * (r - rij) / rij);
                }
            }
            const unsigned long long mask = __ballot(hit);
            if (hit) {
                const int pos = cnt + __popcll(mask & ((1ull << lane) - 1ull));
                if (pos < MAXNB_CN) nbl[(size_t)i * MAXNB_CN + pos] = a0 + jj;
                raw += contrib;
            }
            cnt += __popcll(mask);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) raw += __shfl_xor(raw, off);
        if (lane == 0) {  // scan results; the per-atom epilogue runs below, one thread per atom
            sraw[ii] = raw;
            scnt[ii] = cnt;
        }
    }
    __syncthreads();
    for (int k = tid; k < nm; k += 256)
        cn_atom_finish(a0 + k, sraw[k], scnt[k], elem, nref, refcn, chi_base, cnf, cn, dlog, gw, dgw, rhs, nbn, flags);
}

// ---------------------------------------------------------------------------
// K2: reduced EEQ matrix.  Per fragment f with atoms a_1 < ... < a_m, the
// charges are q = qp + N p with qp = Q_f/m and columns e_a - e_last, so
//   Ared(a,b) = A(ia,ib) - A(ia,lb) - A(la,ib) + A(la,lb)
// with A the Coulomb matrix of eeq_solver.cpp:3571-3650.  Lower triangle of
// the padded (np x np) matrix; the padding is the identity.
// ---------------------------------------------------------------------------
// 1/sqrt(v) in FP64: the hardware estimate refined by two Newton steps
// (within about one ulp; each step doubles the correct bits)
__device__ inline double rsqrt_nr2(double v) {
    double y = __builtin_amdgcn_rsq(v);
    const double h = 0.5 * v;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// A(i, j) = erf(g r) / r with g = 1 / sqrt(alp_i + alp_j), evaluated as
// erf(g r2 / r) / r from one reciprocal square root of r2 (no divide or
// square root sequence on the dependency chain).  For g r >= 6, erfc(g r) <
// 2.2e-17 lies below half an ulp of 1 and erf(g r) is 1: the branch is
// uniform over a wavefront of far-apart atoms and skips the erf evaluation.
__device__ inline double eeq_a_r2(double r2, double ai, double aj) {
    const double rinv = rsqrt_nr2(r2);
    const double g = rsqrt_nr2(ai + aj);
    const double v = g * (r2 * rinv);
    double ef = 1.0;
    if (v < 6.0) ef = erf(v);
    return ef * rinv;
}

// erf(v) and exp(-v^2) for v >= 0 (screened Coulomb pair term and its
// derivative).  From v = 9 (FP64) or v = 5 (FP32) on, erf(v) rounds to 1 and
// the derivative's (2 / sqrt(pi)) v exp(-v^2) lies below the type's
// resolution relative to erf(v), so both are skipped.
__device__ inline void erf_exp(double v, double& ef, double& eg) {
    ef = 1.0;
    eg = 0.0;
    if (v < 9.0) {
        ef = erf(v);
        eg = exp(-v * v);
    }
}
__device__ inline void erf_exp(float v, float& ef, float& eg) {
    ef = 1.0f;
    eg = 0.0f;
    if (v < 5.0f) {
        ef = erff(v);
        eg = __expf(-v * v);
    }
}

__device__ inline double eeq_a(int i, int j, const double* __restrict__ x, const double* __restrict__ alp,
                               const double* __restrict__ diag) {
    if (i == j) return diag[i];
    double3 d = ld3(x, i) - ld3(x, j);
    return eeq_a_r2(dot(d, d), alp[i], alp[j]);
}

__global__ __launch_bounds__(256) void k_eeq_diag(int n, const double* __restrict__ gam,
                                                  const double* __restrict__ alp, double* __restrict__ diag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) diag[i] = gam[i] + TSQRT2PI / sqrt(alp[i]);
}

// Reduced variables of one fragment are consecutive and share the last atom
// la, so a 64-block of variables touches few distinct last atoms (32 for
// water, 1 for one large molecule).  Per tile, the A entries that involve a
// last atom, A(ia, lb), A(la, ib) and A(la, lb), are evaluated once per
// distinct last atom into LDS (ul: the block's distinct last atoms, lidx: a
// variable's index among them), and every element then needs one fresh
// A(ia, ib).  The sum is formed in the same order as the direct form, so R is
// bit-identical; tiles with more than RED_UMAX distinct last atoms use the
// direct four-term form.
constexpr int RED_UMAX = 32;
// eeq_a for two distinct atoms from staged positions and alphas (the same
// operations as eeq_a, so the same bits)
__device__ inline double eeq_a_v(double3 pi, double ai, double3 pj, double aj) {
    double3 d = pi - pj;
    return eeq_a_r2(dot(d, d), ai, aj);
}
struct RedSm {
    union {
        struct {
            double Aiu[NB][RED_UMAX + 1], Aui[RED_UMAX][NB + 1], Auu[RED_UMAX][RED_UMAX + 1];
        } c;
        double Tl[NB][NB + 1];  // the transposed tile, after the caches are no longer read
    };
    double3 rp[NB], cp[NB], urp[RED_UMAX], ucp[RED_UMAX];
    double ra_[NB], ca_[NB], ura[RED_UMAX], uca[RED_UMAX];
    int sia[NB], sib[NB], slr[NB], slc[NB], sur[RED_UMAX], suc[RED_UMAX];
};
constexpr int RED_T = 512;  // threads per tile of k_eeq_reduced
__global__ __launch_bounds__(RED_T) void k_eeq_reduced(int nr, int np, const int* __restrict__ ra,
                                                     const int* __restrict__ rl, const int* __restrict__ lidx,
                                                     const int* __restrict__ ul, const int* __restrict__ nul,
                                                     const double* __restrict__ x, const double* __restrict__ alp,
                                                     const double* __restrict__ diag, double* __restrict__ A,
                                                     int full) {
    // one 64x64 lower tile per workgroup of RED_T threads; grid enumerates
    // tiles (bi >= bj).  full == 0: lower triangle of the row-major np x np
    // matrix (Cholesky).  full == 1: also the mirrored upper tile.  full == 2:
    // packed symmetric storage for the iterative solver, tile t at A + t 64^2
    // (row-major), diagonal tiles completed from their lower triangle
    __shared__ RedSm sm;
    int t = blockIdx.x;
    int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    while (bi * (bi + 1) / 2 > t) --bi;
    int bj = t - bi * (bi + 1) / 2;
    const int tid = threadIdx.x;
    const int nur = nul[bi], nuc = nul[bj];
    const bool ded = nur <= RED_UMAX && nuc <= RED_UMAX;
    if (tid < NB) {
        const int r = bi * NB + tid, c = bj * NB + tid;
        const int ia = (r < nr) ? ra[r] : -1, ib = (c < nr) ? ra[c] : -1;
        sm.sia[tid] = ia;
        sm.slr[tid] = (r < nr) ? lidx[r] : 0;
        sm.sib[tid] = ib;
        sm.slc[tid] = (c < nr) ? lidx[c] : 0;
        sm.rp[tid] = ld3(x, max(ia, 0));
        sm.ra_[tid] = alp[max(ia, 0)];
        sm.cp[tid] = ld3(x, max(ib, 0));
        sm.ca_[tid] = alp[max(ib, 0)];
    } else if (ded && tid < NB + RED_UMAX) {
        const int u = tid - NB;
        if (u < nur) {
            const int a = ul[bi * NB + u];
            sm.sur[u] = a;
            sm.urp[u] = ld3(x, a);
            sm.ura[u] = alp[a];
        }
        if (u < nuc) {
            const int a = ul[bj * NB + u];
            sm.suc[u] = a;
            sm.ucp[u] = ld3(x, a);
            sm.uca[u] = alp[a];
        }
    }
    __syncthreads();
    if (ded) {
        for (int e = tid; e < NB * nuc; e += RED_T) {
            const int r = e / nuc, u = e - r * nuc;
            sm.c.Aiu[r][u] = (sm.sia[r] >= 0) ? eeq_a_v(sm.rp[r], sm.ra_[r], sm.ucp[u], sm.uca[u]) : 0.0;
        }
        for (int e = tid; e < nur * NB; e += RED_T) {
            const int u = e / NB, c = e - u * NB;
            sm.c.Aui[u][c] = (sm.sib[c] >= 0) ? eeq_a_v(sm.urp[u], sm.ura[u], sm.cp[c], sm.ca_[c]) : 0.0;
        }
        for (int e = tid; e < nur * nuc; e += RED_T) {
            const int u = e / nuc, w = e - u * nuc;
            sm.c.Auu[u][w] = (sm.sur[u] == sm.suc[w]) ? diag[sm.sur[u]]
                                                      : eeq_a_v(sm.urp[u], sm.ura[u], sm.ucp[w], sm.uca[w]);
        }
        __syncthreads();
    }
    // thread tid owns column cc = tid % 64 of rows rr = tid / 64 + (RED_T / 64) u
    constexpr int PER = NB * NB / RED_T, RS = RED_T / 64;
    const int cc = tid & 63, c = bj * NB + cc;
    double vals[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int rr = (tid >> 6) + RS * u, r = bi * NB + rr;
        double v = 0.0;
        if (c <= r) {
            if (r >= nr || c >= nr) {
                v = (r == c) ? 1.0 : 0.0;
            } else if (ded) {
                const int ia = sm.sia[rr], ib = sm.sib[cc];
                const double a1 = (ia == ib) ? diag[ia] : eeq_a_v(sm.rp[rr], sm.ra_[rr], sm.cp[cc], sm.ca_[cc]);
                v = a1 - sm.c.Aiu[rr][sm.slc[cc]] - sm.c.Aui[sm.slr[rr]][cc] + sm.c.Auu[sm.slr[rr]][sm.slc[cc]];
            } else {
                int ia = ra[r], la = rl[r], ib = ra[c], lb = rl[c];
                v = eeq_a(ia, ib, x, alp, diag) - eeq_a(ia, lb, x, alp, diag) - eeq_a(la, ib, x, alp, diag) +
                    eeq_a(la, lb, x, alp, diag);
            }
            if (full == 2)
                A[(size_t)t * NB * NB + rr * NB + cc] = v;
            else
                A[(size_t)r * np + c] = v;
        }
        vals[u] = v;
    }
    if (!full || (full == 2 && bi != bj)) return;
    __syncthreads();  // the caches are no longer read: Tl reuses their memory
#pragma unroll
    for (int u = 0; u < PER; ++u) sm.Tl[(tid >> 6) + RS * u][cc] = vals[u];
    __syncthreads();
    // upper counterpart: A(bj*64 + i, bi*64 + j) = T(j, i), row-contiguous stores
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int i = (tid >> 6) + RS * u, j = cc;
        if (bi == bj && j <= i) continue;
        if (full == 2)
            A[(size_t)t * NB * NB + i * NB + j] = sm.Tl[j][i];
        else
            A[(size_t)(bj * NB + i) * np + bi * NB + j] = sm.Tl[j][i];
    }
}

// reduced right-hand side: b(a) = (x - A qp)(ia) - (x - A qp)(la)
__global__ __launch_bounds__(256) void k_eeq_rhs_qp(int n, const double* __restrict__ x,
                                                    const double* __restrict__ alp, const double* __restrict__ diag,
                                                    const double* __restrict__ qp, const double* __restrict__ rhs,
                                                    double* __restrict__ r, const int* __restrict__ mol_of,
                                                    const int* __restrict__ mol_off) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    const int j0 = mol_of ? mol_off[mol_of[i]] : 0, j1 = mol_of ? mol_off[mol_of[i] + 1] : n;
    for (int j = j0; j < j1; ++j)
        if (qp[j] != 0.0) s += eeq_a(i, j, x, alp, diag) * qp[j];
    r[i] = rhs[i] - s;
}

__global__ __launch_bounds__(256) void k_eeq_rhs_reduce(int nr, int np, const int* __restrict__ ra,
                                                        const int* __restrict__ rl, const double* __restrict__ r,
                                                        double* __restrict__ b) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= np) return;
    b[a] = (a < nr && ra[a] >= 0) ? r[ra[a]] - r[rl[a]] : 0.0;  // ra < 0: padding of a batched block
}

// ---------------------------------------------------------------------------
// K3: blocked Cholesky, right-looking, lower, row-major with leading dim np.
// ---------------------------------------------------------------------------
// 1 / sqrt(d) from the hardware estimate refined by Newton steps
// (v_rsq_f64 ~23 bits: two steps reach double precision; f32: one step)
__device__ __forceinline__ double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
    y = y * (1.5 - 0.5 * d * y * y);
    return y * (1.5 - 0.5 * d * y * y);
}
__device__ __forceinline__ float rsqrt_nr(float d) {
    const float y = __builtin_amdgcn_rsqf(d);
    return y * (1.5f - 0.5f * d * y * y);
}

// Factorisation L L^T of one NB x NB diagonal block and the inverse of L by
// one workgroup of 256 threads, in steps of four columns, with the working
// matrix resident in LDS.  Thread (t, q) = (tid / 4, tid % 4) owns the
// entries (t, 4u + q).  Step b (columns k = 4b .. 4b + 3):
//  1. every thread reads the 4 x 4 pivot block and factors it redundantly
//     (four reciprocal square roots), reads its row's four entries and solves
//     them against the pivot factor: L(t, k .. k + 3);
//  2. the rows below the block store their L(t, k + q) (only their own quad
//     read those entries); barrier; the pivot rows store theirs;
//  3. every thread updates its entries A(t, j) -= sum_m L(t, k + m) L(j, k + m),
//     k + 3 < j <= t, reading the published columns; barrier.
// Two barriers per four columns (32 for the block), a rolled step loop (no
// register arrays, compact code).  The inverse X = L^{-1} is formed four
// rows at a time: thread (c, q) keeps X(4u + q, c) in registers; for rows
// 4b .. 4b + 3 the four partial sums L(r, :4b) X(:4b, c) are taken together
// (independent chains), combined over the quad by shuffles, and the 4 x 4
// lower-triangular system with the pivot factor is solved (no barrier).
// The block routine operates on caller-provided LDS so that one workgroup can
// also factor a whole small matrix (batched path); `col` is not used.
template <typename T>
__device__ __forceinline__ void chol_diag_block(int np, int k0, T* __restrict__ A, T* __restrict__ Linv, int* __restrict__ flags,
                                T (*sL)[NB + 1], T (*col)[NB], T* rdiag, bool write_L = true) {
    const int tid = threadIdx.x;
    const int t = tid >> 2, q = tid & 3, qb = tid & ~3 & 63;  // qb: lane of the quad's q = 0
    _Pragma("unroll") for (int u_ = 0, e = tid; u_ < NB * NB / 256; ++u_, e += 256) sL[e / NB][e % NB] = A[(size_t)(k0 + e / NB) * np + k0 + e % NB];
    __syncthreads();
    // the working matrix stays in LDS: thread (t, q) owns the entries (t, 4u + q)
    int bad = 0;
#pragma unroll 1
    for (int b = 0; b < NB / 4; ++b) {
        const int k = 4 * b;
        // pivot factor (lower 4 x 4), redundantly in every thread
        T p[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int jj = 0; jj <= i; ++jj) p[i][jj] = sL[k + i][k + jj];
        T l[4][4] = {}, il[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int jj = 0; jj < i; ++jj) {
                T s = p[i][jj];
#pragma unroll
                for (int m = 0; m < jj; ++m) s -= l[i][m] * l[jj][m];
                l[i][jj] = s * il[jj];
            }
            T d = p[i][i];
#pragma unroll
            for (int m = 0; m < i; ++m) d -= l[i][m] * l[i][m];
            if (!(d > T(0))) {
                bad = 1;
                d = T(1);
            }
            il[i] = rsqrt_nr(d);
            l[i][i] = d * il[i];
        }
        // this row's L(t, k .. k + 3): solved against the pivot factor below the block,
        // the pivot factor's row inside it
        T y[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            T sv = sL[t][k + m];
#pragma unroll
            for (int jj = 0; jj < m; ++jj) sv -= l[m][jj] * y[jj];
            y[m] = sv * il[m];
        }
        const bool piv_row = t >= k && t < k + 4;
        if (piv_row) {
            const int i = t - k;  // static register indices: select among the pivot factor's rows
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const T lm = (i == 0) ? l[0][m] : (i == 1) ? l[1][m] : (i == 2) ? l[2][m] : l[3][m];
                y[m] = (m <= i) ? lm : T(0);
            }
        }
        // rows below the pivot block publish now (only their own quad read these entries);
        // the pivot rows after the barrier, once every thread has read the pivot block
        if (t >= k + 4) sL[t][k + q] = y[q];
        __syncthreads();
        if (piv_row) sL[t][k + q] = y[q];
        // A(t, j) -= sum_m L(t, k + m) L(j, k + m) for this thread's columns j > k + 3
        if (t >= k + 4) {
            for (int j = k + 4 + q; j <= t; j += 4)
                sL[t][j] -= y[0] * sL[j][k] + y[1] * sL[j][k + 1] + y[2] * sL[j][k + 2] + y[3] * sL[j][k + 3];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < NB / 4; ++u)
        if (4 * u + q > t) sL[t][4 * u + q] = T(0);  // the factor's upper triangle
    __syncthreads();
    if (q == 0) rdiag[t] = T(1) / sL[t][t];
    if (write_L) {
        _Pragma("unroll") for (int u_ = 0, e = tid; u_ < NB * NB / 256; ++u_, e += 256) {
            int r = e / NB, c = e % NB;
            if (c <= r) A[(size_t)(k0 + r) * np + k0 + c] = sL[r][c];
        }
    }
    __syncthreads();
    const int c = t;
    T xr[NB / 4];
#pragma unroll
    for (int u = 0; u < NB / 4; ++u) xr[u] = T(0);
#pragma unroll
    for (int b = 0; b < NB / 4; ++b) {
        const int r = 4 * b;
        T part[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int u = 0; u < b; ++u)
#pragma unroll
            for (int m = 0; m < 4; ++m) part[m] += sL[r + m][4 * u + q] * xr[u];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            part[m] += __shfl_xor(part[m], 1);
            part[m] += __shfl_xor(part[m], 2);
        }
        T x[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            T s = ((r + m == c) ? T(1) : T(0)) - part[m];
#pragma unroll
            for (int jj = 0; jj < m; ++jj) s -= sL[r + m][r + jj] * x[jj];
            x[m] = s * rdiag[r + m];
        }
        xr[b] = x[q];
        __builtin_amdgcn_sched_barrier(0);
    }
    T* Li = Linv + (size_t)(k0 / NB) * NB * NB;
#pragma unroll
    for (int u = 0; u < NB / 4; ++u) Li[(4 * u + q) * NB + c] = xr[u];
    if (bad && tid == 0) atomicOr(flags, FLAG_CHOL);
}

template <typename T>
__global__ __launch_bounds__(256, 3) void k_chol_diag(int np, int k0, T* __restrict__ A, T* __restrict__ Linv,
                                                   int* __restrict__ flags) {
    __shared__ T sL[NB][NB + 1];
    __shared__ T col[2][NB];
    __shared__ T rdiag[NB];
    chol_diag_block<T>(np, k0, A, Linv, flags, sL, col, rdiag);
}

// panel: rows below the diagonal block, L(i, k-block) = A(i, k-block) * Linv^T
// panel tile: rows r0 .. r0 + 63 of the column block k0, L = A Linv_k^T
template <typename T>
__device__ __forceinline__ void chol_panel_tile(int np, int k0, int r0, T* __restrict__ A, const T* __restrict__ Linv,
                                T (*S)[NB + 1], T (*W)[NB + 1]) {
    int tid = threadIdx.x;
    const T* Li = Linv + (size_t)(k0 / NB) * NB * NB;
    _Pragma("unroll") for (int u_ = 0, e = tid; u_ < NB * NB / 256; ++u_, e += 256) {
        int r = e / NB, c = e % NB;
        S[r][c] = A[(size_t)(r0 + r) * np + k0 + c];
        W[r][c] = Li[e];
    }
    __syncthreads();
    // out(r, c) = sum_p S(r, p) * W(c, p), p <= c
    int tr = tid / 16, tc = tid % 16;
    T acc[4][4] = {};
    for (int p = 0; p < NB; ++p) {
        T a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a[u] = S[tr * 4 + u][p];
            b[u] = W[tc * 4 + u][p];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) A[(size_t)(r0 + tr * 4 + u) * np + k0 + tc * 4 + v] = acc[u][v];
}

template <typename T>
__global__ __launch_bounds__(256) void k_chol_panel(int np, int k0, T* __restrict__ A, const T* __restrict__ Linv) {
    __shared__ T S[NB][NB + 1];
    __shared__ T W[NB][NB + 1];
    chol_panel_tile<T>(np, k0, k0 + NB + blockIdx.x * NB, A, Linv, S, W);
}

// Update of a set of lower 64 x 64 tiles by a block of factored columns:
//   A(bi, bj) -= sum_{c in [kc0, kc0 + kw)} L(bi, c) L(bj, c)^T
// over the trapezoid of tiles with column blocks bj in [bj0, bj0 + w) and rows
// bi in [bj, nbk).  Tiles are numbered row by row (the first w rows form a
// triangle, every later row holds w tiles) and handed out XCD-contiguously:
// workgroups are dispatched round-robin over the 8 XCDs, so consecutive tile
// numbers (same block row, same L panel rows) are given to workgroups of one
// XCD and share its L2.  The rank kw is streamed through LDS in 64-column
// chunks.  FP64 MFMA (v_mfma_f64_16x16x4f64): each of the 4 wavefronts owns a
// 32x32 quarter of the tile as 2x2 blocks of 16x16.  Operand layout per lane:
// A(i = lane%16, k = lane/16), B(k = lane/16, j = lane%16); result element v
// of a block sits at (row 4*v + lane/16, column lane%16) (measured on gfx950).
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// Per-type 16x16x4 MFMA and its result layout: element v of a lane is at row
// 4*v + lane/16 (kRowFromReg) or 4*(lane/16) + v, column lane%16.
template <typename T>
struct MfmaT;
template <>
struct MfmaT<double> {
    typedef dbl4 V;
    static constexpr bool kRowFromReg = true;
    static __device__ __forceinline__ V op(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct MfmaT<float> {
    typedef flt4 V;
    static constexpr bool kRowFromReg = MFMA_F32_ROW_FROM_REG;
    static __device__ __forceinline__ V op(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

__device__ inline int xcd_tile(int b, int nb) {
    const int x = b & 7, l = b >> 3, per = nb >> 3, rem = nb & 7;
    return x * per + min(x, rem) + l;
}

// A(ri.., rj..) -= sum_{c in [kc0, kc0 + kw)} L(ri.., c) L(rj.., c)^T for one
// 64 x 64 tile (lower triangle only when ri == rj), FP64/FP32 MFMA
template <typename T>
__device__ __forceinline__ void chol_update_tile(int np, int ri, int rj, int kc0, int kw, T* __restrict__ A, T (*P)[NB + 1],
                                 T (*Q)[NB + 1]) {
    using M = MfmaT<T>;
    const int bi = ri, bj = rj;  // only compared for equality (diagonal tile)
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int li = lane % 16, lk = lane / 16;
    typename M::V acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = typename M::V{0, 0, 0, 0};
    // chunk kc of both panels: element u of this thread is row tid/64 + 4u,
    // column tid%64; the next chunk is fetched into registers while the
    // current one feeds the MFMAs
    T pr[NB * NB / 256], qr[NB * NB / 256];
    auto fetch = [&](int kc) {
#pragma unroll
        for (int u = 0; u < NB * NB / 256; ++u) {
            const int r = (tid >> 6) + 4 * u, c = tid & 63;
            pr[u] = A[(size_t)(ri + r) * np + kc + c];
            qr[u] = A[(size_t)(rj + r) * np + kc + c];
        }
    };
    fetch(kc0);
    for (int kc = kc0; kc < kc0 + kw; kc += NB) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NB * NB / 256; ++u) {
            const int r = (tid >> 6) + 4 * u, c = tid & 63;
            P[r][c] = pr[u];
            Q[r][c] = qr[u];
        }
        __syncthreads();
        if (kc + NB < kc0 + kw) fetch(kc + NB);
#pragma unroll 4
        for (int k = 0; k < NB; k += 4) {
            const T fa0 = P[wr + li][k + lk], fa1 = P[wr + 16 + li][k + lk];
            const T fb0 = Q[wc + li][k + lk], fb1 = Q[wc + 16 + li][k + lk];
            acc[0][0] = M::op(fa0, fb0, acc[0][0]);
            acc[0][1] = M::op(fa0, fb1, acc[0][1]);
            acc[1][0] = M::op(fa1, fb0, acc[1][0]);
            acc[1][1] = M::op(fa1, fb1, acc[1][1]);
        }
    }
    // off-diagonal tiles: all 16 read-modify-writes issued unconditionally;
    // diagonal tiles keep the lower triangle only
    T old[2][2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = ri + wr + 16 * x + (M::kRowFromReg ? 4 * v + lk : 4 * lk + v);
                const int c = rj + wc + 16 * y + li;
                old[x][y][v] = A[(size_t)r * np + c];
            }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = ri + wr + 16 * x + (M::kRowFromReg ? 4 * v + lk : 4 * lk + v);
                const int c = rj + wc + 16 * y + li;
                if (bi != bj || c <= r) A[(size_t)r * np + c] = old[x][y][v] - acc[x][y][v];
            }
}

template <typename T>
__global__ __launch_bounds__(256) void k_chol_update(int np, int kc0, int kw, int bj0, int w, int nbk,
                                                     T* __restrict__ A) {
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int head = w * (w + 1) / 2;
    int bi, bj;
    if (t < head) {
        int r = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((r + 1) * (r + 2) / 2 <= t) ++r;
        while (r * (r + 1) / 2 > t) --r;
        bi = bj0 + r;
        bj = bj0 + (t - r * (r + 1) / 2);
    } else {
        const int t2 = t - head;
        bi = bj0 + w + t2 / w;
        bj = bj0 + t2 % w;
    }
    if (bi >= nbk) return;
    __shared__ T P[NB][NB + 1];
    __shared__ T Q[NB][NB + 1];
    chol_update_tile<T>(np, bi * NB, bj * NB, kc0, kw, A, P, Q);
}

// Blocked right-looking Cholesky of the padded np x np matrix in two levels:
// column blocks of NB = 64 are factored (diag + panel) and applied at once
// only to the remaining columns of their super-panel of SP blocks; the rest of
// the matrix receives one rank-(SP NB) update per super-panel.  Every element
// of the trailing matrix is then read and written once per 256 columns
// instead of once per 64, which lifts the update above the HBM roofline.
constexpr int SP = 4;
constexpr int SYMV_SPLIT = 8;   // tile split of the refinement matvec
constexpr int REFINE_STEPS = 3; // FP32 solves per mixed-precision EEQ solve

// ---------------------------------------------------------------------------
// K4: triangular solves with the blocked factor
// ---------------------------------------------------------------------------
// Blocked substitution with the stored inverses of the diagonal blocks.
// Forward, block k: y_k = Linv_k b_k (every workgroup recomputes the 64x64
// product from LDS; workgroup 0 stores it), then rows below the block get
// b_r -= L(r, k) y_k.  Backward, block k (descending): x_k = Linv_k^T c_k,
// then c_j -= L(k, j)^T x_k for j < k0.
// Forward, block k: y_k = Linv_k b_k, then b_r -= L(r, k-block) y_k for the
// 64-row tile of this workgroup (rows r0 = k0 + NB + 64 blockIdx.x ...).  The
// tile is staged in LDS with coalesced row reads; thread (row, q) sums the
// products p = 4u + q and quad shuffles combine the four shares.  Every
// workgroup forms y_k itself (64 x 64 product from LDS); workgroup 0 stores it.
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_fwd(int np, int k0, const T* __restrict__ A,
                                                  const T* __restrict__ Linv, T* __restrict__ b,
                                                  T* __restrict__ y) {
    __shared__ T W[NB][NB + 1];
    __shared__ T Tt[NB][NB + 1];
    __shared__ T bk[NB], yk[NB];
    const int tid = threadIdx.x, t = tid >> 2, q = tid & 3;
    const int r0 = k0 + NB + blockIdx.x * NB;
    const T* Li = Linv + (size_t)(k0 / NB) * NB * NB;
    _Pragma("unroll") for (int u_ = 0, e = tid; u_ < NB * NB / 256; ++u_, e += 256) {
        const int rr = e / NB, p = e % NB;
        W[rr][p] = Li[e];
        const T v = A[(size_t)min(r0 + rr, np - 1) * np + k0 + p];  // clamped: unconditional load
        Tt[rr][p] = (r0 + rr < np) ? v : T(0);
    }
    if (tid < NB) bk[tid] = b[k0 + tid];
    __syncthreads();
    T s = T(0);
#pragma unroll
    for (int u = 0; u < NB / 4; ++u) s += W[t][4 * u + q] * bk[4 * u + q];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (q == 0) {
        yk[t] = s;
        if (blockIdx.x == 0) y[k0 + t] = s;
    }
    __syncthreads();
    s = T(0);
#pragma unroll
    for (int u = 0; u < NB / 4; ++u) s += Tt[t][4 * u + q] * yk[4 * u + q];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (q == 0 && r0 + t < np) b[r0 + t] -= s;
}
// Backward, block k (descending): x_k = Linv_k^T c_k, then c_j -= L(k-block, j)^T x_k
// for the 64 columns j of this workgroup (j < k0).  Wave g sums rows p in
// [16 g, 16 g + 16) with coalesced row reads; the four partial sums are
// combined in LDS in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_bwd(int np, int k0, const T* __restrict__ A,
                                                  const T* __restrict__ Linv, T* __restrict__ c,
                                                  T* __restrict__ xout) {
    __shared__ T W[NB][NB + 1];
    __shared__ T ck[NB], xk[NB];
    __shared__ T part[4][NB];
    const int tid = threadIdx.x, t = tid >> 2, q = tid & 3;
    const T* Li = Linv + (size_t)(k0 / NB) * NB * NB;
    _Pragma("unroll") for (int u_ = 0, e = tid; u_ < NB * NB / 256; ++u_, e += 256) W[e / NB][e % NB] = Li[e];
    if (tid < NB) ck[tid] = c[k0 + tid];
    __syncthreads();
    T s = T(0);
#pragma unroll
    for (int u = 0; u < NB / 4; ++u) s += W[4 * u + q][t] * ck[4 * u + q];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (q == 0) {
        xk[t] = s;
        if (blockIdx.x == 0) xout[k0 + t] = s;
    }
    __syncthreads();
    const int jj = tid & 63, g = tid >> 6;
    const int col = blockIdx.x * NB + jj;
    T acc = T(0);
    if (col < k0) {
#pragma unroll
        for (int p = 16 * g; p < 16 * g + 16; ++p) acc += A[(size_t)(k0 + p) * np + col] * xk[p];
    }
    part[g][jj] = acc;
    __syncthreads();
    if (g == 0 && col < k0) c[col] -= ((part[0][jj] + part[1][jj]) + part[2][jj]) + part[3][jj];
}

// Single-launch sweeps.  One workgroup per 64-row block; block w needs the
// solution blocks k < w (forward) or k > w (backward).  Each workgroup stages
// the L tile of the next dependency in LDS before waiting for its ready flag,
// folds L(w,k) y_k into its right-hand side as soon as y_k is published, then
// applies the stored inverse of its diagonal block and publishes y_w with an
// agent-scope release.  Workgroups wait only for blocks published by
// lower-numbered workgroups (the backward sweep numbers blocks in descending
// order), so progress never depends on a workgroup that has not started.
// Every wait is bounded: after SPIN_MAX polls the workgroup sets FLAG_SOLVE
// and continues, so the launch always terminates.
constexpr int SPIN_MAX = 1 << 18;

// Waits for ready flag f; gives up at once when any workgroup has already
// flagged a failed wait, so a broken sweep ends after one bounded wait.
__device__ inline bool wait_ready(const int* f, int* flags) {
    for (int it = 0; it < SPIN_MAX; ++it) {
        if (__hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != 0) return true;
        if (__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_SOLVE) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    atomicOr(flags, FLAG_SOLVE);
    return false;
}

template <typename T>
__global__ __launch_bounds__(256) void k_trsv_fwd_sweep(int np, const T* __restrict__ A, const T* __restrict__ Linv,
                                                        const T* __restrict__ b, T* __restrict__ y,
                                                        int* __restrict__ ready, int* __restrict__ flags) {
    __shared__ T Tt[NB][NB + 1];
    __shared__ T Wi[NB][NB + 1];  // inverse of the diagonal block, fetched first
    __shared__ T yk[NB];
    __shared__ T bw[NB];
    const int w = blockIdx.x, tid = threadIdx.x, t = tid >> 2, q = tid & 3;
    T acc = (q == 0) ? b[w * NB + t] : T(0);  // thread (t, q): share of row t
    {
        const T* Li = Linv + (size_t)w * NB * NB;
        _Pragma("unroll") for (int u_ = 0, e = tid; u_ < NB * NB / 256; ++u_, e += 256) Wi[e / NB][e % NB] = Li[e];
    }
    bool ok = true;
    for (int k = 0; k < w; ++k) {
        __syncthreads();
        _Pragma("unroll") for (int u_ = 0, e = tid; u_ < NB * NB / 256; ++u_, e += 256)
            Tt[e / NB][e % NB] = A[(size_t)(w * NB + e / NB) * np + k * NB + e % NB];
        if (tid < NB) {
            ok = wait_ready(ready + k, flags) && ok;
            yk[tid] = __hip_atomic_load(y + k * NB + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NB / 4; ++u) acc -= Tt[t][4 * u + q] * yk[4 * u + q];
    }
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    __syncthreads();
    if (q == 0) bw[t] = acc;
    __syncthreads();
    T s = T(0);
#pragma unroll
    for (int u = 0; u < NB / 4; ++u) s += Wi[t][4 * u + q] * bw[4 * u + q];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (q == 0) {
        __hip_atomic_store(y + w * NB + t, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();  // the block is visible device-wide before the flag
    }
    __syncthreads();
    if (tid == 0) __hip_atomic_store(ready + w, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (!ok && tid < NB) atomicOr(flags, FLAG_SOLVE);
}

// Backward: x_w = Linv_w^T (c_w - sum_{k > w} L(k, w)^T x_k), workgroup g
// handles block w = nbk - 1 - g.
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_bwd_sweep(int np, const T* __restrict__ A, const T* __restrict__ Linv,
                                                        const T* __restrict__ c, T* __restrict__ x,
                                                        int* __restrict__ ready, int* __restrict__ flags) {
    __shared__ T Tt[NB][NB + 1];
    __shared__ T Wi[NB][NB + 1];  // inverse of the diagonal block, fetched first
    __shared__ T xk[NB];
    __shared__ T cw[NB];
    const int nbk = np / NB;
    const int w = nbk - 1 - (int)blockIdx.x, tid = threadIdx.x, t = tid >> 2, q = tid & 3;
    T acc = (q == 0) ? c[w * NB + t] : T(0);  // thread (t, q): share of column t
    {
        const T* Li = Linv + (size_t)w * NB * NB;
        _Pragma("unroll") for (int u_ = 0, e = tid; u_ < NB * NB / 256; ++u_, e += 256) Wi[e / NB][e % NB] = Li[e];
    }
    bool ok = true;
    for (int k = nbk - 1; k > w; --k) {
        __syncthreads();
        _Pragma("unroll") for (int u_ = 0, e = tid; u_ < NB * NB / 256; ++u_, e += 256)
            Tt[e / NB][e % NB] = A[(size_t)(k * NB + e / NB) * np + w * NB + e % NB];
        if (tid < NB) {
            ok = wait_ready(ready + k, flags) && ok;
            xk[tid] = __hip_atomic_load(x + k * NB + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NB / 4; ++u) acc -= Tt[4 * u + q][t] * xk[4 * u + q];
    }
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    __syncthreads();
    if (q == 0) cw[t] = acc;
    __syncthreads();
    T s = T(0);
#pragma unroll
    for (int u = 0; u < NB / 4; ++u) s += Wi[4 * u + q][t] * cw[4 * u + q];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (q == 0) {
        __hip_atomic_store(x + w * NB + t, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();  // the block is visible device-wide before the flag
    }
    __syncthreads();
    if (tid == 0) __hip_atomic_store(ready + w, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (!ok && tid < NB) atomicOr(flags, FLAG_SOLVE);
}

// ---------------------------------------------------------------------------
// K4c: iterative EEQ solve.  The reduced EEQ matrices of non-bonded clusters
// are well conditioned (water clusters: kappa ~ 6), so preconditioned CG
// converges in ~15 matrix-vector products.  Each product streams the matrix
// once from HBM (row-major, both triangles stored), which replaces the
// latency-bound chain of the blocked factorisation.
// ---------------------------------------------------------------------------
// Block-Jacobi preconditioner: inverse factors of the 64 x 64 diagonal blocks
// of R (one workgroup per block; R is left intact).
// block-Jacobi factors: L^-1 of every diagonal block of the packed matrix
// (block I at Rp + (I (I + 1) / 2 + I) 64^2)
__global__ __launch_bounds__(256, 3) void k_bjacobi_setup(double* __restrict__ Rp, double* __restrict__ Lj,
                                                       int* __restrict__ flags) {
    __shared__ double sL[NB][NB + 1];
    __shared__ double col[2][NB];
    __shared__ double rdiag[NB];
    const size_t I = blockIdx.x;
    chol_diag_block<double>(NB, 0, Rp + (I * (I + 1) / 2 + I) * NB * NB, Lj + I * NB * NB, flags, sL, col, rdiag,
                            false);
}

constexpr int PCG_THREADS = 256;

// Sum over the 64 lanes of a wavefront (fixed butterfly; every lane gets it).
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Fixed-order sum over a workgroup of 256 threads: wavefront butterflies,
// then the four wavefront totals in order (every thread returns the total).
__device__ double wg_sum256(double v, double* red) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// Device-coherent access for data shared between the workgroups of a
// persistent kernel: relaxed agent-scope atomics, i.e. loads and stores that
// bypass the non-coherent caches (no cache write-back or invalidation, whose
// cost grows with the number of workgroups on a multi-die GPU).
__device__ inline double ld_c(const double* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void st_c(double* p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline unsigned long long ld_cu(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_cu(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// s_waitcnt vmcnt(0): this wave's loads and stores have completed (device-
// coherent stores are then visible to every device-coherent load).
__device__ inline void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Grid-wide all-reduce of two values per workgroup for a persistent kernel
// whose G <= 256 workgroups of 256 threads are all resident.  Mailboxes are
// 128-byte lines of 64-bit words holding MB_EMPTY (a NaN pattern no
// arithmetic produces) until written: at barrier j workgroup g writes its
// values to line (g, j mod 3), and thread t of every workgroup polls line
// (t, j mod 3) and reads producer t's values; every workgroup then adds the
// G values in the same fixed order, so all hold identical sums.  A writer
// clears its line (g, (j + 1) mod 3), last written at barrier j - 2, before
// writing barrier j: having passed barrier j - 1 it has seen every
// workgroup's barrier j - 1 value, each written after its writer finished
// reading barrier j - 2, so no reader still needs the cleared line, and the
// clear completes before the next value is written.  Every wait is bounded
// (FLAG_SOLVE on a timeout).  Ordering without cache maintenance: the data a
// workgroup shares is written with device-coherent stores that complete
// (vmcnt) before its mailbox value is written, and read with device-coherent
// loads issued after all values were seen, so every shared write before the
// barrier is visible to every workgroup after it.
constexpr unsigned long long MB_EMPTY = 0x7ff0deadbeef0001ull;
constexpr int MB_LINE = 16;                  // 64-bit words per 128-byte line
constexpr int MB_WORDS = 3 * 256 * MB_LINE;  // lines (g, k), k < 3, at (3 g + k) MB_LINE

__device__ bool grid_allreduce2(double v0, double v1, double* sums, unsigned long long* mb, int j, int* flags,
                                double* red) {
    __shared__ int bad_s;
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, k = j % 3;
    if (tid == 0) bad_s = 0;  // every thread read the previous value before the last call's final barrier
    wait_vm();
    __syncthreads();
    if (tid == 0) {
        unsigned long long* const own = mb + (3 * g + k) * MB_LINE;
        unsigned long long* const old = mb + (3 * g + (k + 1) % 3) * MB_LINE;
        st_cu(old, MB_EMPTY);
        st_cu(old + 1, MB_EMPTY);
        wait_vm();
        st_cu(own, __double_as_longlong(v0));
        st_cu(own + 1, __double_as_longlong(v1));
    }
    double u0 = 0.0, u1 = 0.0;
    if (tid < G) {
        const unsigned long long* const in = mb + (3 * tid + k) * MB_LINE;
        unsigned long long a0 = MB_EMPTY, a1 = MB_EMPTY;
        for (int it = 0;; ++it) {
            a0 = ld_cu(in);
            a1 = ld_cu(in + 1);
            if (a0 != MB_EMPTY && a1 != MB_EMPTY) break;
            if (it > SPIN_MAX || (__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_SOLVE)) {
                bad_s = 1;
                atomicOr(flags, FLAG_SOLVE);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        u0 = __longlong_as_double(a0);
        u1 = __longlong_as_double(a1);
    }
    u0 = wave_sum(u0);
    u1 = wave_sum(u1);
    __syncthreads();
    if ((tid & 63) == 0) {
        red[tid >> 6] = u0;
        red[4 + (tid >> 6)] = u1;
    }
    __syncthreads();
    sums[0] = (red[0] + red[1]) + (red[2] + red[3]);
    sums[1] = (red[4] + red[5]) + (red[6] + red[7]);
    const bool ok = bad_s == 0;
    __syncthreads();
    return ok;
}

// z_b = Linv_b^T (Linv_b r_b) for the 64-block b (r_b and the result in LDS;
// threads 0..255 compute, every thread synchronises).
__device__ void bjacobi_apply(const double* __restrict__ Li, const double* rb, double* wb, double* zb) {
    const int tid = threadIdx.x, t = (tid >> 2) & 63, q = tid & 3;
    const bool act = tid < 256;
    double s = 0.0;
    if (act) {
#pragma unroll
        for (int u = 0; u < NB / 4; ++u) s += Li[t * NB + 4 * u + q] * rb[4 * u + q];
    }
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    __syncthreads();
    if (act && q == 0) wb[t] = s;
    __syncthreads();
    s = 0.0;
    if (act) {
#pragma unroll
        for (int u = 0; u < NB / 4; ++u) s += Li[(4 * u + q) * NB + t] * wb[4 * u + q];
    }
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (act && q == 0) zb[t] = s;
    __syncthreads();
}

// Persistent PCG on R x = b with R in packed symmetric storage: the 64x64
// blocks (I, J), I >= J, at Rp + (I (I + 1) / 2 + J) 64^2, row-major, the
// diagonal blocks full.  Workgroup g owns the 64-row blocks g, g + G, ... (at
// most PCG_MAXB) for the vector updates.  The matrix-vector product runs over
// the stored blocks, each block read once by one wavefront (contiguous,
// balanced ranges of the packed order): lane (rg, cg) holds 8 rows x 8
// columns of the block, and the row sums R_IJ p_J and column sums
// R_IJ^T p_I come from 8-lane transpose-reduces.  Every (row block X, column block Y)
// product goes to its own partial part[X][Y][64], which the owner of X adds
// in Y order after the barrier; p.q is formed from the partials (it is
// linear in them).  Per iteration: (1) p = z + beta p_prev (formed on the fly
// by every reader; owners store their rows), block products; barrier;
// (2) alpha = rz / pq; owners: q, x += alpha p, r -= alpha q, z = M^-1 r,
// partials r.z and r.r; barrier; beta = rz_new / rz.  Stops when
// r.r <= tol2 * b.b (uniform decision from identical sums) or after maxit
// iterations (FLAG_SOLVE).  Cross-workgroup vectors (z, p, part) are written
// and read with device-coherent accesses (ld_c, st_c); an owner keeps its
// rows of x and r, and the Linv of a single owned block, in LDS for the
// whole solve and writes x at the end.
constexpr int PCG_MAXB = 8;
constexpr double PCG_TOL_FP64 = 1e-13;   // default |r| / |b| of the FP64 path
constexpr double PCG_TOL_MIXED = 1e-10;  // default of the mixed path (EEQ energy is variational in q)
constexpr int PCG_MAXIT = 1000;

// Sum of 8 per-lane values v[k] over the 8 lanes that differ in lane bits
// s, s + 1, s + 2 (recursive halving): lane l ends with the total of
// k(l) = 4 b(s+2) + 2 b(s+1) + b(s), b the bits of l; the eight lanes of a
// group hold the same bits.
template <int S>
__device__ inline double tr_reduce8(const double (&v)[8], int l) {
    double a4[4], a2[2];
    const bool h2 = (l >> (S + 2)) & 1, h1 = (l >> (S + 1)) & 1, h0 = (l >> S) & 1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double keep = h2 ? v[i + 4] : v[i], send = h2 ? v[i] : v[i + 4];
        a4[i] = keep + __shfl_xor(send, 4 << S);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double keep = h1 ? a4[i + 2] : a4[i], send = h1 ? a4[i] : a4[i + 2];
        a2[i] = keep + __shfl_xor(send, 2 << S);
    }
    return (h0 ? a2[1] : a2[0]) + __shfl_xor(h0 ? a2[0] : a2[1], 1 << S);
}

__global__ __launch_bounds__(256) void k_eeq_pcg(int np, const double* __restrict__ Rp,
                                                 const double* __restrict__ Lj, const double* __restrict__ b,
                                                 double* __restrict__ x, double* __restrict__ z,
                                                 double* __restrict__ pb,
                                                 double* __restrict__ part, unsigned long long* __restrict__ mb,
                                                 int* __restrict__ iters, double tol2, int maxit,
                                                 int* __restrict__ flags) {
    __shared__ double red[256];
    __shared__ double pown[PCG_MAXB][NB];
    __shared__ double qs[4][NB];
    __shared__ double rb[NB], wb[NB], zb[NB];
    __shared__ double xo[PCG_MAXB][NB], ro[PCG_MAXB][NB];  // owned rows of x and r, resident for the solve
    __shared__ double Lc[NB * NB];                          // Linv of the owned block when there is one
    __shared__ __align__(16) double pw[4][2][NB];            // per wavefront: p of the block's rows and columns
    // a diagonal block that is not positive definite (FLAG_CHOL, set before
    // this launch, so every workgroup reads the same value) leaves the solve to
    // the LU fallback that follows
    if (__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_CHOL) return;
    const int G = gridDim.x, g = blockIdx.x, nbk = np / NB;
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63;
    const int nown = (nbk - g + G - 1) / G;
    // this wavefront's blocks of the packed order: [t0, t1)
    const long T = (long)nbk * (nbk + 1) / 2, W = 4L * G, gw = 4L * g + wv;
    const int t0 = (int)(T * gw / W), t1 = (int)(T * (gw + 1) / W);
    int I0 = (int)((sqrt(8.0 * t0 + 1.0) - 1.0) * 0.5);
    while ((long)(I0 + 1) * (I0 + 2) / 2 <= t0) ++I0;
    while ((long)I0 * (I0 + 1) / 2 > t0) --I0;
    const int J0 = t0 - I0 * (I0 + 1) / 2;
    // lane (rg, cg) holds rows 8 rg + i (i < 8) and columns 2 cg + (k & 1) + 16 (k >> 1) (k < 8) of a
    // block: every 16-byte load instruction reads whole 128-byte row segments
    const int rg = l >> 3, cg = l & 7;
    const bool lcache = nown == 1;
    if (lcache)
        for (int e = tid; e < NB * NB; e += 256) Lc[e] = Lj[(size_t)g * NB * NB + e];
    // init: x = 0, r = b, z = M^-1 b; partials r.z and b.b
    double lrz = 0.0, lbb = 0.0;
    for (int o = 0; o < nown; ++o) {
        const int kb = g + o * G;
        if (tid < NB) {
            const double bv = b[kb * NB + tid];
            rb[tid] = bv;
            xo[o][tid] = 0.0;
            ro[o][tid] = bv;
        }
        __syncthreads();
        bjacobi_apply(lcache ? Lc : Lj + (size_t)kb * NB * NB, rb, wb, zb);
        if (tid < NB) {
            st_c(z + kb * NB + tid, zb[tid]);
            lrz += rb[tid] * zb[tid];
            lbb += rb[tid] * rb[tid];
        }
    }
    int nbar = 0;  // barriers passed
    double sums[2];
    const double s_rz = wg_sum256(lrz, red), s_bb = wg_sum256(lbb, red);
    if (!grid_allreduce2(s_rz, s_bb, sums, mb, nbar++, flags, red)) return;
    double rz = sums[0];
    const double bb = sums[1];
    double beta = 0.0;
    int it = 0;
    bool done = !(bb > 0.0);
    for (; !done && it < maxit; ++it) {
        const double* pprev = pb + (size_t)((it + 1) & 1) * np;
        double* pcur = pb + (size_t)(it & 1) * np;
        // (1) owners: p rows; every wavefront: its blocks' products
        for (int o = 0; o < nown; ++o) {
            if (tid < NB) {
                const int row = (g + o * G) * NB + tid;
                const double pc = (it == 0) ? 0.0 : ld_c(pprev + row);
                const double pv = ld_c(z + row) + beta * pc;
                pown[o][tid] = pv;
                st_c(pcur + row, pv);
            }
        }
        double lpq = 0.0;
        int I = I0, J = J0;
        for (int t = t0; t < t1; ++t) {
            const double2* Rb = reinterpret_cast<const double2*>(Rp + (size_t)t * NB * NB);
            double rv[8][8];
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const double2 v2 = Rb[((8 * rg + i) * NB + 16 * j + 2 * cg) / 2];
                    rv[i][2 * j] = v2.x;
                    rv[i][2 * j + 1] = v2.y;
                }
            const double pcj = (it == 0) ? 0.0 : ld_c(pprev + J * NB + l);
            const double pjl = ld_c(z + J * NB + l) + beta * pcj;  // p of column J * 64 + l
            const double pci = (it == 0) ? 0.0 : ld_c(pprev + I * NB + l);
            const double pil = ld_c(z + I * NB + l) + beta * pci;  // p of row I * 64 + l
            pw[wv][0][l] = pil;
            pw[wv][1][l] = pjl;
            __builtin_amdgcn_wave_barrier();
            double pi[8], pj[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                pi[k] = pw[wv][0][8 * rg + k];
                pj[k] = pw[wv][1][2 * cg + (k & 1) + 16 * (k >> 1)];
            }
            const int cend = 2 * cg + (rg & 1) + 16 * (rg >> 1);  // the column this lane's column sum ends on
            const double pjc = pw[wv][1][cend];
            __builtin_amdgcn_wave_barrier();
            double rs8[8], cs8[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                double a = 0.0, c = 0.0;
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    a += rv[k][m] * pj[m];
                    c += rv[m][k] * pi[m];
                }
                rs8[k] = a;
                cs8[k] = c;
            }
            const double rs = tr_reduce8<0>(rs8, l);  // row 8 rg + cg = l
            st_c(part + ((size_t)I * nbk + J) * NB + l, rs);
            lpq += pil * rs;
            if (I != J) {
                const double cs = tr_reduce8<3>(cs8, l);  // column cend
                st_c(part + ((size_t)J * nbk + I) * NB + cend, cs);
                lpq += pjc * cs;
            }
            if (++J > I) {
                ++I;
                J = 0;
            }
        }
        const double s_pq = wg_sum256(lpq, red);
        if (!grid_allreduce2(s_pq, 0.0, sums, mb, nbar++, flags, red)) return;
        const double pq = sums[0];
        const double alpha = rz / pq;
        // (2) owners: q rows from the partials (Y order), x, r, z and the
        // partials r.z, r.r
        double lrz2 = 0.0, lrr = 0.0;
        for (int o = 0; o < nown; ++o) {
            const int kb = g + o * G;
            {
                // thread (row l, part wv) adds Y = wv, wv + 4, ... in batches of 32 loads
                const double* pr = part + (size_t)kb * nbk * NB + l;
                double acc = 0.0;
                for (int y0 = wv; y0 < nbk; y0 += 128) {
                    double t32[32];
#pragma unroll
                    for (int u = 0; u < 32; ++u) t32[u] = ld_c(pr + (size_t)min(y0 + 4 * u, nbk - 1) * NB);
#pragma unroll
                    for (int u = 0; u < 32; ++u) acc += (y0 + 4 * u < nbk) ? t32[u] : 0.0;
                }
                qs[wv][l] = acc;
            }
            __syncthreads();
            if (tid < NB) {
                const double qv = (qs[0][tid] + qs[1][tid]) + (qs[2][tid] + qs[3][tid]);
                xo[o][tid] += alpha * pown[o][tid];
                const double rv = ro[o][tid] - alpha * qv;
                ro[o][tid] = rv;
                rb[tid] = rv;
            }
            __syncthreads();
            bjacobi_apply(lcache ? Lc : Lj + (size_t)kb * NB * NB, rb, wb, zb);
            if (tid < NB) {
                st_c(z + kb * NB + tid, zb[tid]);
                lrz2 += rb[tid] * zb[tid];
                lrr += rb[tid] * rb[tid];
            }
        }
        const double s_rz2 = wg_sum256(lrz2, red), s_rr = wg_sum256(lrr, red);
        if (!grid_allreduce2(s_rz2, s_rr, sums, mb, nbar++, flags, red)) return;
        const double rz2 = sums[0], rr = sums[1];
        beta = rz2 / rz;
        rz = rz2;
        done = rr <= tol2 * bb;
    }
    for (int o = 0; o < nown; ++o)
        if (tid < NB) x[(g + o * G) * NB + tid] = xo[o][tid];
    if (g == 0 && tid == 0) iters[0] = it;
    if (!done && g == 0 && tid == 0) atomicOr(flags, FLAG_SOLVE);
}

// ---------------------------------------------------------------------------
// K4b: mixed-precision EEQ.  The reduced matrix R is factored in FP32 and the
// FP64 solution is won back by iterative refinement against the FP64 R:
//   x = 0, r = b;  repeat: x += L32^-T L32^-1 r;  r = b - R x.
// ---------------------------------------------------------------------------
// FP32 copy of the lower tiles (bi >= bj) of R; tile t numbered row by row.
__global__ __launch_bounds__(256) void k_lower_to_f32(int np, const double* __restrict__ A, float* __restrict__ A32) {
    const int t = blockIdx.x;
    int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    while (bi * (bi + 1) / 2 > t) --bi;
    const int bj = t - bi * (bi + 1) / 2;
    _Pragma("unroll") for (int u_ = 0, e = (int)threadIdx.x; u_ < NB * NB / 256; ++u_, e += 256) {
        const size_t o = (size_t)(bi * NB + e / NB) * np + bj * NB + e % NB;
        A32[o] = (float)A[o];
    }
}

// Partial products of the symmetric matvec y = R x from the lower tiles of R:
// workgroup (I, sp) sums, for the 64 rows of block row I, the tiles J = sp,
// sp + nsplit, ...: (I, J) by rows when J < I, the diagonal tile completed
// symmetrically from its lower triangle, and (J, I)^T by columns when J > I.
// part[sp][row]; summed in a fixed order by k_eeq_resid.
__global__ __launch_bounds__(256) void k_symv_part(int np, int nsplit, const double* __restrict__ R,
                                                   const double* __restrict__ x, double* __restrict__ part) {
    __shared__ double Tt[NB][NB + 1];
    __shared__ double xs[NB];
    const int I = blockIdx.x, sp = blockIdx.y, nbk = np / NB;
    const int tid = threadIdx.x, t = tid >> 2, q = tid & 3;
    double acc = 0.0;
    for (int J = sp; J < nbk; J += nsplit) {
        __syncthreads();
        const bool below = J > I;
        const int r0 = below ? J * NB : I * NB, c0 = below ? I * NB : J * NB;
        _Pragma("unroll") for (int u_ = 0, e = tid; u_ < NB * NB / 256; ++u_, e += 256) {
            const int r = e / NB, c = e % NB;
            // diagonal tile: symmetric completion from its lower triangle
            Tt[r][c] = (J == I && c > r) ? R[(size_t)(r0 + c) * np + c0 + r] : R[(size_t)(r0 + r) * np + c0 + c];
        }
        if (tid < NB) xs[tid] = x[J * NB + tid];
        __syncthreads();
        if (!below) {
#pragma unroll
            for (int u = 0; u < NB / 4; ++u) acc += Tt[t][4 * u + q] * xs[4 * u + q];
        } else {
#pragma unroll
            for (int u = 0; u < NB / 4; ++u) acc += Tt[4 * u + q][t] * xs[4 * u + q];
        }
    }
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    if (q == 0) part[(size_t)sp * np + I * NB + t] = acc;
}

// r = b - sum_sp part[sp] (fixed order); rf = (float) r for the next FP32 solve.
__global__ __launch_bounds__(256) void k_eeq_resid(int np, int nsplit, const double* __restrict__ part,
                                                   const double* __restrict__ b, double* __restrict__ r,
                                                   float* __restrict__ rf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double y = 0.0;
    for (int sp = 0; sp < nsplit; ++sp) y += part[(size_t)sp * np + i];
    const double v = b[i] - y;
    r[i] = v;
    rf[i] = (float)v;
}

// x (+)= d for the FP32 correction d; first == 1 starts from x = d.
__global__ __launch_bounds__(256) void k_refine_add(int np, int first, const float* __restrict__ d,
                                                    double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    x[i] = first ? (double)d[i] : x[i] + (double)d[i];
}

__global__ __launch_bounds__(256) void k_to_f32(int n, const double* __restrict__ a, float* __restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------
// K5: charges, self energy (ff_workspace.cpp:426-449) and qtmp (:459-507)
// ---------------------------------------------------------------------------
// Fallback solve of a reduced EEQ system that is not positive definite
// (FLAG_CHOL raised by the Cholesky of a diagonal block or of the matrix): the
// reference falls back to an LU with partial pivoting of its bordered system
// when the Cholesky of A fails (eeq_solver.cpp:2143-2165, the augmented-system
// PartialPivLU).  The reduced system R p = b is the bordered system with the
// constraints eliminated, so Gaussian elimination with partial pivoting on it
// gives the same charges.  One workgroup; does nothing unless the flag is set;
// clears FLAG_CHOL (and FLAG_SOLVE of the failed solve) when it succeeds (a
// zero pivot keeps it: singular system).
// R: layout 0 = packed lower 64x64 blocks (block I >= J at (I(I+1)/2 + J) 64^2,
// row-major inside), layout 1 = row-major np x np, lower triangle used.
// W: nr x (nr + 1) work (the system augmented with b).  nr <= LU_MAX.
constexpr int LU_MAX = 1024;
constexpr int LU_T = 1024;
__global__ __launch_bounds__(LU_T) void k_eeq_lu_fallback(int np, int nr, int layout, const double* __restrict__ R,
                                                          const double* __restrict__ b, double* __restrict__ W,
                                                          double* __restrict__ x, int* __restrict__ flags) {
    __shared__ int s_flag, s_piv;
    __shared__ double s_best[LU_T / 64];
    __shared__ int s_idx[LU_T / 64];
    __shared__ double xs[LU_MAX];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) s_flag = __hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (!(s_flag & FLAG_CHOL) || nr > LU_MAX) return;
    const int m = nr, ld = nr + 1;
    for (long e = tid; e < (long)m * m; e += LU_T) {
        const int i = (int)(e / m), j = (int)(e % m);
        const int r = max(i, j), c = min(i, j);
        double v;
        if (layout == 0) {
            const long I = r / NB, J = c / NB;
            v = R[(I * (I + 1) / 2 + J) * NB * NB + (r % NB) * NB + (c % NB)];
        } else {
            v = R[(long)r * np + c];
        }
        W[(long)i * ld + j] = v;
    }
    for (int i = tid; i < m; i += LU_T) W[(long)i * ld + m] = b[i];
    __syncthreads();
    int ok = 1;
    for (int k = 0; k < m; ++k) {
        // pivot row: largest |W(i, k)| for i >= k, the lowest index on ties
        double best = -1.0;
        int bi = m;
        for (int i = k + tid; i < m; i += LU_T) {
            const double a = fabs(W[(long)i * ld + k]);
            if (a > best) {
                best = a;
                bi = i;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o);
            const int oi = __shfl_xor(bi, o);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        if (lane == 0) {
            s_best[wv] = best;
            s_idx[wv] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double bb = s_best[0];
            int ii = s_idx[0];
            for (int w = 1; w < LU_T / 64; ++w)
                if (s_best[w] > bb || (s_best[w] == bb && s_idx[w] < ii)) {
                    bb = s_best[w];
                    ii = s_idx[w];
                }
            s_piv = (bb > 0.0 && ii < m) ? ii : -1;
        }
        __syncthreads();
        const int p = s_piv;
        if (p < 0) {
            ok = 0;
            break;
        }
        if (p != k)
            for (int j = k + tid; j <= m; j += LU_T) {
                const double t = W[(long)k * ld + j];
                W[(long)k * ld + j] = W[(long)p * ld + j];
                W[(long)p * ld + j] = t;
            }
        __syncthreads();
        const double inv = 1.0 / W[(long)k * ld + k];
        const int cols = m - k;  // columns k+1 .. m (the last is b)
        for (long e = tid; e < (long)(m - k - 1) * cols; e += LU_T) {
            const int i = k + 1 + (int)(e / cols), j = k + 1 + (int)(e % cols);
            W[(long)i * ld + j] -= (W[(long)i * ld + k] * inv) * W[(long)k * ld + j];
        }
        __syncthreads();
    }
    if (!ok) return;
    for (int i = m - 1; i >= 0; --i) {
        double sacc = 0.0;
        for (int j = i + 1 + tid; j < m; j += LU_T) sacc += W[(long)i * ld + j] * xs[j];
        for (int o = 32; o > 0; o >>= 1) sacc += __shfl_xor(sacc, o);
        if (lane == 0) s_best[wv] = sacc;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < LU_T / 64; ++w) t += s_best[w];
            xs[i] = (W[(long)i * ld + m] - t) / W[(long)i * ld + i];
        }
        __syncthreads();
    }
    for (int i = tid; i < np; i += LU_T) x[i] = (i < m) ? xs[i] : 0.0;
    // the LU result replaces whatever the failed solve left (its bounded waits included)
    if (tid == 0) atomicAnd(flags, ~(FLAG_CHOL | FLAG_SOLVE));
}

// Per fragment f, the sum of the reduced solution over its atoms (the last
// atom's charge is qp minus it): one wavefront per fragment, lanes striding
// the fragment's atom list, fixed butterfly.
__global__ __launch_bounds__(256) void k_frag_sum(int nfrag, const int* __restrict__ frag_off,
                                                  const int* __restrict__ frag_atoms, const int* __restrict__ red_of,
                                                  const double* __restrict__ p, double* __restrict__ fsum) {
    const int f = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (f >= nfrag) return;
    double s = 0.0;
    for (int k = frag_off[f] + lane; k < frag_off[f + 1]; k += 64) {
        const int ra = red_of[frag_atoms[k]];
        if (ra >= 0) s += p[ra];
    }
    s = wave_sum(s);
    if (lane == 0) fsum[f] = s;
}

__global__ __launch_bounds__(256) void k_charges(int n, const int* __restrict__ red_of,
                                                 const double* __restrict__ fsum, const int* __restrict__ frag,
                                                 const double* __restrict__ qp, const double* __restrict__ p,
                                                 const double* __restrict__ chi_base, const double* __restrict__ cnf,
                                                 const double* __restrict__ gam, const double* __restrict__ alp,
                                                 const double* __restrict__ cn, double* __restrict__ q,
                                                 double* __restrict__ e_self, double* __restrict__ qtmp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ri = red_of[i];
    const double qi = (ri >= 0) ? qp[i] + p[ri] : qp[i] - fsum[frag[i]];  // ri < 0: last atom of its fragment
    q[i] = qi;
    double c = cn[i];
    double chi = chi_base[i] + cnf[i] * sqrt(fmax(c, 0.0));
    e_self[i] = -qi * chi + 0.5 * qi * qi * (gam[i] + TSQRT2PI_SELF / sqrt(alp[i]));
    qtmp[i] = qi * cnf[i] / (2.0 * sqrt(fmax(c, 0.0)) + 1e-16);
}

// ---------------------------------------------------------------------------
// K6: pair terms.  Row atom i per thread, partners j staged in LDS tiles.
//   Coulomb  (r <= 100)  ff_workspace_gfnff.cpp:850-898
//   D4       (r <= 50)   :715-767 with C6 from the current Gaussian weights
//   repulsion (r <= 20)  :814-842 with the element-pair factor; bonded and
//                        1,3/1,4 H-H pairs are corrected by k_atom_tail
// Each pair is evaluated from both sides; a row keeps half of each pair
// energy and its own full gradient and dE/dCN, so no atomics are needed.
// ---------------------------------------------------------------------------
struct PairArgs {
    int n, nelem;
    const double* x;
    const double* q;
    const double* alp;
    const double* gw;
    const double* dgw;
    const double* zeta;
    const double* szr;
    const double* anb;
    const double* repz;
    const int* Z;
    const int* elem;
    const int* nref;
    const int* metal;
    const double* c6ref;
    const double* h;  // [n][nelem][7] D4 half contraction (k_d4_half)
    const double* dh;
    const int* mol_of;   // batches: molecule of each atom (nullptr: one molecule)
    const int* mol_off;  // batches: atom offsets of the molecules
    double* g;        // [3n]
    double* dEdcn;    // [n]
    double* e_coul;   // [n]
    double* e_disp;   // [n]
    double* e_rep;    // [n]
};

__device__ inline double rep_ff(int zi, int zj, int mi, int mj) {
    if (zi == 1 && zj == 1) return 0.6290;
    if ((zi == 1 && mj > 0) || (zj == 1 && mi > 0)) return 0.85;
    if ((zi == 1 && zj == 6) || (zj == 1 && zi == 6)) return 0.91;
    if ((zi == 1 && zj == 8) || (zj == 1 && zi == 8)) return 1.04;
    return 1.0;
}

// Partner range of a pair-kernel workgroup: the j chunk of a single molecule,
// or, in a batch, the union of the molecules of the workgroup's rows (each
// row then keeps only partners of its own molecule).
__device__ inline void pair_range(const PairArgs& a, int cj, int& j0, int& j1) {
    if (!a.mol_of) {
        j0 = blockIdx.y * cj;
        j1 = min(a.n, j0 + cj);
    } else {
        const int first = blockIdx.x * TILE, last = min(a.n, first + TILE) - 1;
        j0 = a.mol_off[a.mol_of[first]];
        j1 = a.mol_off[a.mol_of[last] + 1];
    }
}

// D4 half contraction per atom i and partner element e:
//   h[i][e][b]  = sum_a gw_i[a]  c6ref[ei, e, a, b]
//   dh[i][e][b] = sum_a dgw_i[a] c6ref[ei, e, a, b]
// so that C6_ij = h[i][ej] . gw_j and dC6_ij/dCN_i = dh[i][ej] . gw_j
// (the reference's half-contraction, d4param_generator.cpp:1235).
__global__ __launch_bounds__(256) void k_d4_half(int n, int ne, const int* __restrict__ elem,
                                                 const int* __restrict__ nref, const double* __restrict__ c6ref,
                                                 const double* __restrict__ gw, const double* __restrict__ dgw,
                                                 double* __restrict__ h, double* __restrict__ dh) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * ne) return;
    int i = t / ne, e = t % ne;
    int ei = elem[i], nri = nref[ei];
    const double* c = c6ref + (size_t)(ei * ne + e) * 49;
    double hb[7] = {}, dhb[7] = {};
    for (int aa = 0; aa < nri; ++aa) {
        double g = gw[(size_t)i * 7 + aa], dg = dgw[(size_t)i * 7 + aa];
#pragma unroll
        for (int bb = 0; bb < 7; ++bb) {
            hb[bb] += g * c[aa * 7 + bb];
            dhb[bb] += dg * c[aa * 7 + bb];
        }
    }
#pragma unroll
    for (int bb = 0; bb < 7; ++bb) {
        h[(size_t)t * 7 + bb] = hb[bb];
        dh[(size_t)t * 7 + bb] = dhb[bb];
    }
}

// Pair terms over full rows, split into j chunks so that the launch fills the
// chip: block (bx, by) owns rows i in [256 bx, 256 bx + 256) and partners j in
// [by cj, by cj + cj).  Each thread accumulates its row's gradient, dE/dCN and
// three energies over the chunk and writes them to part[by][k][i]; the chunks
// are summed in a fixed order by k_pairs_reduce (no atomics, deterministic).
// NE > 0: the D4 half contraction of row i for every partner element sits in
// registers and is selected by the partner's element, which is uniform across
// the wavefront because all lanes visit the same j; NE == 0: read from memory.
constexpr int PART_K = 7;
template <int NE>
__global__ __launch_bounds__(TILE) void k_pairs_part(PairArgs a, int cj, double* __restrict__ part) {
    __shared__ double sx[TILE], sy[TILE], sz[TILE], sq[TILE], sal[TILE], szt[TILE], ssz[TILE], san[TILE],
        srz[TILE];
    __shared__ double sgw[TILE * 7];
    __shared__ int sZ[TILE], sel[TILE], smt[TILE];
    const int n = a.n, ne = a.nelem;
    const int i = blockIdx.x * TILE + threadIdx.x;
    const bool act = i < n;
    int j0, j1;
    pair_range(a, cj, j0, j1);
    const int mlo = (a.mol_of && act) ? a.mol_off[a.mol_of[i]] : 0;
    const int mhi = (a.mol_of && act) ? a.mol_off[a.mol_of[i] + 1] : n;
    double3 xi = act ? ld3(a.x, i) : make_double3(0, 0, 0);
    double qi = act ? a.q[i] : 0.0, ali = act ? a.alp[i] : 1.0, zti = act ? a.zeta[i] : 0.0;
    double szi = act ? a.szr[i] : 0.0, ani = act ? a.anb[i] : 0.0, rzi = act ? a.repz[i] : 0.0;
    int zi = act ? a.Z[i] : 0, mi = act ? a.metal[i] : 0;
    constexpr int NH = NE > 0 ? NE * 7 : 1;
    double hr[NH], dhr[NH];
    if (NE > 0) {
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            hr[k] = act ? a.h[(size_t)i * NH + k] : 0.0;
            dhr[k] = act ? a.dh[(size_t)i * NH + k] : 0.0;
        }
    }
    double gx = 0, gy = 0, gz = 0, dcn = 0, ec = 0, ed = 0, er = 0;
    for (int t0 = j0; t0 < j1; t0 += TILE) {
        __syncthreads();
        int j = t0 + threadIdx.x;
        if (j < j1) {
            sx[threadIdx.x] = a.x[3 * j];
            sy[threadIdx.x] = a.x[3 * j + 1];
            sz[threadIdx.x] = a.x[3 * j + 2];
            sq[threadIdx.x] = a.q[j];
            sal[threadIdx.x] = a.alp[j];
            szt[threadIdx.x] = a.zeta[j];
            ssz[threadIdx.x] = a.szr[j];
            san[threadIdx.x] = a.anb[j];
            srz[threadIdx.x] = a.repz[j];
            sZ[threadIdx.x] = a.Z[j];
            sel[threadIdx.x] = a.elem[j];
            smt[threadIdx.x] = a.metal[j];
#pragma unroll
            for (int r = 0; r < 7; ++r) sgw[threadIdx.x * 7 + r] = a.gw[(size_t)j * 7 + r];
        }
        __syncthreads();
        if (!act) continue;
        const int m = min(TILE, j1 - t0);
        for (int jj = 0; jj < m; ++jj) {
            if (t0 + jj == i || t0 + jj < mlo || t0 + jj >= mhi) continue;
            double dx = xi.x - sx[jj], dy = xi.y - sy[jj], dz = xi.z - sz[jj];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 10000.0) continue;  // beyond every pair cut-off (Coulomb 100 Bohr)
            double r = sqrt(r2);
            double dEdr = 0.0;
            // Coulomb
            if (r <= 100.0 && r >= 1e-10) {
                double gij = 1.0 / sqrt(ali + sal[jj]);
                double gr = gij * r;
                double ef, eg;
                erf_exp(gr, ef, eg);
                double qq = qi * sq[jj];
                ec += 0.5 * qq * ef / r;
                dEdr += qq * (gij * eg * (2.0 / 1.772453850905516) / r - ef / r2);
            }
            // D4 dispersion
            if (r <= 50.0 && r >= 1e-10) {
                const int ej = sel[jj];
                double c6 = 0.0, dci = 0.0;
                if (NE > 0) {
#pragma unroll
                    for (int e = 0; e < (NE > 0 ? NE : 1); ++e)
                        if (e == ej) {
#pragma unroll
                            for (int bb = 0; bb < 7; ++bb) {
                                double gj = sgw[jj * 7 + bb];
                                c6 += hr[e * 7 + bb] * gj;
                                dci += dhr[e * 7 + bb] * gj;
                            }
                        }
                } else {
                    const double* hp = a.h + ((size_t)i * ne + ej) * 7;
                    const double* dp = a.dh + ((size_t)i * ne + ej) * 7;
#pragma unroll
                    for (int bb = 0; bb < 7; ++bb) {
                        double gj = sgw[jj * 7 + bb];
                        c6 += hp[bb] * gj;
                        dci += dp[bb] * gj;
                    }
                }
                double r4r2 = 3.0 * szi * ssz[jj];
                double r0 = 0.58 * sqrt(r4r2) + 4.80;
                double r0s = r0 * r0;
                double zc = zti * szt[jj];
                double r6 = r2 * r2 * r2, r06 = r0s * r0s * r0s;
                double t6 = 1.0 / (r6 + r06), t8 = 1.0 / (r6 * r2 + r06 * r0s);
                double sum = t6 + 2.0 * r4r2 * t8;
                ed += 0.5 * (-c6 * sum * zc);
                double d6 = -6.0 * r2 * r2 * t6 * t6, d8 = -8.0 * r2 * r2 * r2 * t8 * t8;
                dEdr += -c6 * zc * (d6 + 2.0 * r4r2 * d8) * r;
                dcn -= dci * sum * zc;
            }
            // non-bonded repulsion (generic pair factor; bonded / 1,3 / 1,4
            // corrections follow in k_atom_tail)
            if (r <= 20.0 && r >= 1e-8) {
                double ff = rep_ff(zi, sZ[jj], mi, smt[jj]);
                double al = sqrt(ani * san[jj]) * ff;
                double ab = rzi * srz[jj] * 0.4270;
                double e = ab * exp(-al * r * sqrt(r)) / r;
                er += 0.5 * e;
                dEdr += -e / r - 1.5 * al * sqrt(r) * e;
            }
            double s = dEdr / r;
            gx += s * dx;
            gy += s * dy;
            gz += s * dz;
        }
    }
    if (!act) return;
    double* p = part + (size_t)blockIdx.y * PART_K * n + i;
    p[0] = gx;
    p[(size_t)n] = gy;
    p[(size_t)2 * n] = gz;
    p[(size_t)3 * n] = dcn;
    p[(size_t)4 * n] = ec;
    p[(size_t)5 * n] = ed;
    p[(size_t)6 * n] = er;
}

// Mixed-precision pair terms: the same rows, chunks and partial-sum layout as
// k_pairs_part.  Coordinate differences are formed in FP64 (no loss of the
// absolute positions), every term is evaluated in FP32 (hardware rsqrt, rcp,
// exp, erf paths) and the row sums are accumulated in FP64.
template <int NE>
__global__ __launch_bounds__(TILE) void k_pairs_part32(PairArgs a, int cj, double* __restrict__ part) {
    __shared__ double sx[TILE], sy[TILE], sz[TILE];
    __shared__ float sq[TILE], sal[TILE], szt[TILE], ssz[TILE], san[TILE], srz[TILE];
    __shared__ float sgw[TILE * 7];
    __shared__ int sZ[TILE], sel[TILE], smt[TILE];
    const int n = a.n, ne = a.nelem;
    const int i = blockIdx.x * TILE + threadIdx.x;
    const bool act = i < n;
    int j0, j1;
    pair_range(a, cj, j0, j1);
    const int mlo = (a.mol_of && act) ? a.mol_off[a.mol_of[i]] : 0;
    const int mhi = (a.mol_of && act) ? a.mol_off[a.mol_of[i] + 1] : n;
    const double3 xi = act ? ld3(a.x, i) : make_double3(0, 0, 0);
    const float qi = act ? (float)a.q[i] : 0.f, ali = act ? (float)a.alp[i] : 1.f;
    const float zti = act ? (float)a.zeta[i] : 0.f, szi = act ? (float)a.szr[i] : 0.f;
    const float ani = act ? (float)a.anb[i] : 0.f, rzi = act ? (float)a.repz[i] : 0.f;
    const int zi = act ? a.Z[i] : 0, mi = act ? a.metal[i] : 0;
    constexpr int NH = NE > 0 ? NE * 7 : 1;
    float hr[NH], dhr[NH];
    if (NE > 0) {
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            hr[k] = act ? (float)a.h[(size_t)i * NH + k] : 0.f;
            dhr[k] = act ? (float)a.dh[(size_t)i * NH + k] : 0.f;
        }
    }
    double gx = 0, gy = 0, gz = 0, dcn = 0, ec = 0, ed = 0, er = 0;
    for (int t0 = j0; t0 < j1; t0 += TILE) {
        __syncthreads();
        const int j = t0 + threadIdx.x;
        if (j < j1) {
            sx[threadIdx.x] = a.x[3 * j];
            sy[threadIdx.x] = a.x[3 * j + 1];
            sz[threadIdx.x] = a.x[3 * j + 2];
            sq[threadIdx.x] = (float)a.q[j];
            sal[threadIdx.x] = (float)a.alp[j];
            szt[threadIdx.x] = (float)a.zeta[j];
            ssz[threadIdx.x] = (float)a.szr[j];
            san[threadIdx.x] = (float)a.anb[j];
            srz[threadIdx.x] = (float)a.repz[j];
            sZ[threadIdx.x] = a.Z[j];
            sel[threadIdx.x] = a.elem[j];
            smt[threadIdx.x] = a.metal[j];
#pragma unroll
            for (int r = 0; r < 7; ++r) sgw[threadIdx.x * 7 + r] = (float)a.gw[(size_t)j * 7 + r];
        }
        __syncthreads();
        if (!act) continue;
        const int m = min(TILE, j1 - t0);
        for (int jj = 0; jj < m; ++jj) {
            if (t0 + jj == i || t0 + jj < mlo || t0 + jj >= mhi) continue;
            const double dxd = xi.x - sx[jj], dyd = xi.y - sy[jj], dzd = xi.z - sz[jj];
            const double r2d = dxd * dxd + dyd * dyd + dzd * dzd;
            if (r2d > 10000.0) continue;  // beyond every pair cut-off (Coulomb 100 Bohr)
            const float r2 = (float)r2d;
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r = r2 * rinv;
            float dEdr = 0.f;
            if (r2d >= 1e-20) {
                // Coulomb (r <= 100)
                const float gij = __builtin_amdgcn_rsqf(ali + sal[jj]);
                const float gr = gij * r;
                float ef, eg;
                erf_exp(gr, ef, eg);
                const float qq = qi * sq[jj];
                ec += (double)(0.5f * qq * ef * rinv);
                dEdr += qq * (gij * eg * 1.1283791671f * rinv - ef * rinv * rinv);
            }
            if (r2d <= 2500.0 && r2d >= 1e-20) {
                // D4 dispersion (r <= 50)
                const int ej = sel[jj];
                float c6 = 0.f, dci = 0.f;
                if (NE > 0) {
#pragma unroll
                    for (int e = 0; e < (NE > 0 ? NE : 1); ++e)
                        if (e == ej) {
#pragma unroll
                            for (int bb = 0; bb < 7; ++bb) {
                                const float gj = sgw[jj * 7 + bb];
                                c6 += hr[e * 7 + bb] * gj;
                                dci += dhr[e * 7 + bb] * gj;
                            }
                        }
                } else {
                    const double* hp = a.h + ((size_t)i * ne + ej) * 7;
                    const double* dp = a.dh + ((size_t)i * ne + ej) * 7;
#pragma unroll
                    for (int bb = 0; bb < 7; ++bb) {
                        const float gj = sgw[jj * 7 + bb];
                        c6 += (float)hp[bb] * gj;
                        dci += (float)dp[bb] * gj;
                    }
                }
                const float r4r2 = 3.f * szi * ssz[jj];
                const float r0 = 0.58f * __builtin_sqrtf(r4r2) + 4.80f;
                const float r0s = r0 * r0;
                const float zc = zti * szt[jj];
                const float r6 = r2 * r2 * r2, r06 = r0s * r0s * r0s;
                const float t6 = __builtin_amdgcn_rcpf(r6 + r06), t8 = __builtin_amdgcn_rcpf(r6 * r2 + r06 * r0s);
                const float sum = t6 + 2.f * r4r2 * t8;
                ed += (double)(0.5f * (-c6 * sum * zc));
                const float d6 = -6.f * r2 * r2 * t6 * t6, d8 = -8.f * r2 * r2 * r2 * t8 * t8;
                dEdr += -c6 * zc * (d6 + 2.f * r4r2 * d8) * r;
                dcn -= (double)(dci * sum * zc);
            }
            if (r2d <= 400.0 && r2d >= 1e-16) {
                // non-bonded repulsion (r <= 20), generic pair factor
                const float ff = (float)rep_ff(zi, sZ[jj], mi, smt[jj]);
                const float al = __builtin_sqrtf(ani * san[jj]) * ff;
                const float ab = rzi * srz[jj] * 0.4270f;
                const float sr = __builtin_sqrtf(r);
                const float e = ab * __expf(-al * r * sr) * rinv;
                er += (double)(0.5f * e);
                dEdr += -e * rinv - 1.5f * al * sr * e;
            }
            const double sgrad = (double)(dEdr * rinv);
            gx += sgrad * dxd;
            gy += sgrad * dyd;
            gz += sgrad * dzd;
        }
    }
    if (!act) return;
    double* p = part + (size_t)blockIdx.y * PART_K * n + i;
    p[0] = gx;
    p[(size_t)n] = gy;
    p[(size_t)2 * n] = gz;
    p[(size_t)3 * n] = dcn;
    p[(size_t)4 * n] = ec;
    p[(size_t)5 * n] = ed;
    p[(size_t)6 * n] = er;
}

// ---------------------------------------------------------------------------
// Pair terms with each unordered pair evaluated once (single molecules).
// Tiles of 64 rows x 64 columns of the lower triangle (row block R >= column
// block C, tile t = R (R + 1) / 2 + C), one wavefront each.  The column
// atoms' data sit in LDS; at step s lane l pairs its row atom with column
// (l + s) mod 64, so every lane meets every column once and the 64 lanes of
// one step touch 64 distinct columns: the column-side sums accumulate in LDS
// without atomics, in a fixed order.  Diagonal tiles take s = 1 .. 31 and,
// for l < 32, s = 32 (each unordered pair once).  Every tile writes its
// row-side and column-side sums (PART_K quantities x 64 atoms) to partials of
// its own, which k_pairs_sym_reduce adds per atom in a fixed order.  The pair
// arithmetic is that of k_pairs_part (T = double) or k_pairs_part32
// (T = float: FP64 coordinate differences, FP32 terms, every sum in FP64,
// row side and column side); the partner's dC6/dCN
// uses sum_b h_i[e_j][b] dgw_j[b], the transposed half contraction.
constexpr int PS_B = 64;
__device__ inline float ps_rsqrt(float v) { return __builtin_amdgcn_rsqf(v); }
__device__ inline double ps_rsqrt(double v) { return 1.0 / sqrt(v); }
__device__ inline float ps_sqrt(float v) { return __builtin_sqrtf(v); }
__device__ inline double ps_sqrt(double v) { return sqrt(v); }
__device__ inline float ps_rcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ inline double ps_rcp(double v) { return 1.0 / v; }
__device__ inline float ps_exp(float v) { return __expf(v); }
__device__ inline double ps_exp(double v) { return exp(v); }

// terms: PT_COUL | PT_DISP | PT_REP select the pair terms evaluated (all three
// in one pass, or the charge-independent dispersion and repulsion before the
// EEQ solve and the Coulomb term after it, into separate partials).
constexpr int PT_COUL = 1, PT_DISP = 2, PT_REP = 4, PT_ALL = 7;
template <typename T, int NE>
__global__ __launch_bounds__(PS_B) void k_pairs_sym(PairArgs a, double* __restrict__ part,
                                                    const int4* __restrict__ tiles, int terms) {
    __shared__ double cx[PS_B], cy[PS_B], cz[PS_B];
    __shared__ T cq[PS_B], cal[PS_B], czt[PS_B], csz[PS_B], can[PS_B], crz[PS_B];
    __shared__ T cgw[7][PS_B], cdgw[7][PS_B];
    __shared__ int czem[PS_B];        // Z | element << 8 | metal << 16
    __shared__ double acc[PART_K][PS_B];  // column-side sums of this tile, in FP64 on both paths
    const int t = blockIdx.x, l = threadIdx.x;
    // the tile's molecule (atoms a0 .. a0 + n - 1) and its row and column blocks
    int a0 = 0, n = a.n, R, C;
    if (tiles) {
        const int4 tt = tiles[t];
        a0 = tt.x;
        n = tt.y;
        R = tt.z;
        C = tt.w;
    } else {
        R = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((R + 1) * (R + 2) / 2 <= t) ++R;
        while (R * (R + 1) / 2 > t) --R;
        C = t - R * (R + 1) / 2;
    }
    const int ne = a.nelem;
    const int i = a0 + R * PS_B + l, jl = a0 + C * PS_B + l;
    const bool acti = R * PS_B + l < n;
    {
        const bool actj = C * PS_B + l < n;
        const int js = actj ? jl : a0;
        cx[l] = a.x[3 * js];
        cy[l] = a.x[3 * js + 1];
        cz[l] = a.x[3 * js + 2];
        cq[l] = (T)a.q[js];
        cal[l] = (T)a.alp[js];
        czt[l] = (T)a.zeta[js];
        csz[l] = (T)a.szr[js];
        can[l] = (T)a.anb[js];
        crz[l] = (T)a.repz[js];
        czem[l] = a.Z[js] | (a.elem[js] << 8) | (a.metal[js] << 16);
#pragma unroll
        for (int r = 0; r < 7; ++r) {
            cgw[r][l] = (T)a.gw[(size_t)js * 7 + r];
            cdgw[r][l] = (T)a.dgw[(size_t)js * 7 + r];
        }
#pragma unroll
        for (int k = 0; k < PART_K; ++k) acc[k][l] = 0.0;
    }
    __syncthreads();
    const int is = acti ? i : a0;
    const double3 xi = ld3(a.x, is);
    const T qi = (T)a.q[is], ali = (T)a.alp[is], zti = (T)a.zeta[is], szi = (T)a.szr[is];
    const T ani = (T)a.anb[is], rzi = (T)a.repz[is];
    const int zi = a.Z[is], mi = a.metal[is];
    constexpr int NH = NE > 0 ? NE * 7 : 1;
    T hr[NH], dhr[NH];
    if (NE > 0) {
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            hr[k] = (T)a.h[(size_t)is * NH + k];
            dhr[k] = (T)a.dh[(size_t)is * NH + k];
        }
    }
    const int jmax = min(PS_B, n - C * PS_B);  // columns present in this tile
    const bool diag = R == C;
    const int s0 = diag ? 1 : 0, s1 = diag ? (l < 32 ? 32 : 31) : PS_B - 1;
    double gx = 0, gy = 0, gz = 0, dcn = 0, ec = 0, ed = 0, er = 0;
    for (int s = s0; s <= s1; ++s) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);  // keep the LDS updates of consecutive steps in order
        const int jj = (l + s) & (PS_B - 1);
        if (!acti || jj >= jmax) continue;
        const double dxd = xi.x - cx[jj], dyd = xi.y - cy[jj], dzd = xi.z - cz[jj];
        const double r2d = dxd * dxd + dyd * dyd + dzd * dzd;
        if (r2d > ((terms & PT_COUL) ? 10000.0 : 2500.0)) continue;  // beyond every pair cut-off (Coulomb 100 Bohr)
        T dEdr = 0, hec = 0, hed = 0, her = 0, dci = 0, dcj = 0;
        const T r2 = (T)r2d;
        T r, rinv;
        if (sizeof(T) == 4) {
            rinv = ps_rsqrt(r2);
            r = r2 * rinv;
        } else {
            r = ps_sqrt(r2);
            rinv = T(1) / r;
        }
        const bool coul = (terms & PT_COUL) && ((sizeof(T) == 4) ? (r2d >= 1e-20) : (r <= T(100) && r >= T(1e-10)));
        const bool disp = (terms & PT_DISP) &&
                          ((sizeof(T) == 4) ? (r2d <= 2500.0 && r2d >= 1e-20) : (r <= T(50) && r >= T(1e-10)));
        const bool rep = (terms & PT_REP) && ((sizeof(T) == 4) ? (r2d <= 400.0 && r2d >= 1e-16) : (r <= T(20) && r >= T(1e-8)));
        if (coul) {
            const T gij = ps_rsqrt(ali + cal[jj]);
            const T gr = gij * r;
            T ef, eg;
            erf_exp(gr, ef, eg);
            const T qq = qi * cq[jj];
            if (sizeof(T) == 4) {
                hec = T(0.5) * qq * ef * rinv;
                dEdr += qq * (gij * eg * T(1.1283791671) * rinv - ef * rinv * rinv);
            } else {
                hec = T(0.5) * qq * ef / r;
                dEdr += qq * (gij * eg * T(2.0 / 1.772453850905516) / r - ef / r2);
            }
        }
        if (disp) {
            const int ej = (czem[jj] >> 8) & 0xff;
            T c6 = 0;
            if (NE > 0) {
#pragma unroll
                for (int e = 0; e < (NE > 0 ? NE : 1); ++e)
                    if (e == ej) {
#pragma unroll
                        for (int bb = 0; bb < 7; ++bb) {
                            const T gj = cgw[bb][jj], dgj = cdgw[bb][jj];
                            c6 += hr[e * 7 + bb] * gj;
                            dci += dhr[e * 7 + bb] * gj;
                            dcj += hr[e * 7 + bb] * dgj;
                        }
                    }
            } else {
                const double* hp = a.h + ((size_t)is * ne + ej) * 7;
                const double* dp = a.dh + ((size_t)is * ne + ej) * 7;
#pragma unroll
                for (int bb = 0; bb < 7; ++bb) {
                    const T gj = cgw[bb][jj], dgj = cdgw[bb][jj];
                    c6 += (T)hp[bb] * gj;
                    dci += (T)dp[bb] * gj;
                    dcj += (T)hp[bb] * dgj;
                }
            }
            const T r4r2 = T(3) * szi * csz[jj];
            const T r0 = T(0.58) * ps_sqrt(r4r2) + T(4.80);
            const T r0s = r0 * r0;
            const T zc = zti * czt[jj];
            const T r6 = r2 * r2 * r2, r06 = r0s * r0s * r0s;
            T t6, t8;
            if (sizeof(T) == 4) {
                t6 = ps_rcp(r6 + r06);
                t8 = ps_rcp(r6 * r2 + r06 * r0s);
            } else {
                t6 = T(1) / (r6 + r06);
                t8 = T(1) / (r6 * r2 + r06 * r0s);
            }
            const T sum = t6 + T(2) * r4r2 * t8;
            hed = T(0.5) * (-c6 * sum * zc);
            const T d6 = T(-6) * r2 * r2 * t6 * t6, d8 = T(-8) * r2 * r2 * r2 * t8 * t8;
            dEdr += -c6 * zc * (d6 + T(2) * r4r2 * d8) * r;
            dci = dci * sum * zc;
            dcj = dcj * sum * zc;
        }
        if (rep) {
            const T ff = (T)rep_ff(zi, czem[jj] & 0xff, mi, czem[jj] >> 16);
            const T al = ps_sqrt(ani * can[jj]) * ff;
            const T ab = rzi * crz[jj] * T(0.4270);
            const T sr = ps_sqrt(r);
            T e;
            if (sizeof(T) == 4) {
                e = ab * ps_exp(-al * r * sr) * rinv;
                dEdr += -e * rinv - T(1.5) * al * sr * e;
            } else {
                e = ab * ps_exp(-al * r * sr) / r;
                dEdr += -e / r - T(1.5) * al * sr * e;
            }
            her = T(0.5) * e;
        }
        const double sg = (sizeof(T) == 4) ? (double)(dEdr * rinv) : (double)dEdr / (double)r;
        const double fx = sg * dxd, fy = sg * dyd, fz = sg * dzd;
        gx += fx;
        gy += fy;
        gz += fz;
        ec += (double)hec;
        ed += (double)hed;
        er += (double)her;
        dcn -= (double)dci;
        acc[0][jj] -= fx;
        acc[1][jj] -= fy;
        acc[2][jj] -= fz;
        if (disp) acc[3][jj] -= (double)dcj;
        acc[4][jj] += (double)hec;
        if (disp) acc[5][jj] += (double)hed;
        if (rep) acc[6][jj] += (double)her;
    }
    __syncthreads();
    double* pr = part + (size_t)t * 2 * PART_K * PS_B;
    const double iv[PART_K] = {gx, gy, gz, dcn, ec, ed, er};
#pragma unroll
    for (int k = 0; k < PART_K; ++k) {
        pr[k * PS_B + l] = iv[k];
        pr[(PART_K + k) * PS_B + l] = acc[k][l];
    }
}

// Per atom a = A 64 + l: the row-side partials of tiles (A, C), C = 0 .. A,
// then the column-side partials of tiles (R, A), R = A .. nb - 1, in that
// order; one wavefront per block per quarter of the tiles, the four quarters
// added in order.
// Batches (blocks != nullptr): one workgroup per 64-atom block {a0, n, A, base}
// of a molecule (atoms a0 .. a0 + n - 1, its tiles numbered from base).
// part2 (nullable): a second set of tile partials (the split pair terms),
// added tile by tile after the first.
__global__ __launch_bounds__(256) void k_pairs_sym_reduce(int n, int nb, const double* __restrict__ part,
                                                          const double* __restrict__ part2,
                                                          double* __restrict__ g, double* __restrict__ dEdcn,
                                                          double* __restrict__ e_coul, double* __restrict__ e_disp,
                                                          double* __restrict__ e_rep, const int4* __restrict__ blocks) {
    __shared__ double qs[4][PART_K][PS_B];
    int A = blockIdx.x, a0 = 0, base = 0;
    if (blocks) {
        const int4 bb = blocks[blockIdx.x];
        a0 = bb.x;
        n = bb.y;
        A = bb.z;
        base = bb.w;
        nb = (n + PS_B - 1) / PS_B;
    }
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntile = nb + 1;  // A + 1 row-side tiles and nb - A column-side tiles
    double s[PART_K] = {0, 0, 0, 0, 0, 0, 0};
    for (int u = w; u < ntile; u += 4) {
        const bool rowside = u <= A;
        const int Rr = u - 1;  // column side: tile (R, A), R = u - 1 = A .. nb - 1
        const int t = base + (rowside ? A * (A + 1) / 2 + u : Rr * (Rr + 1) / 2 + A);
        const size_t o = ((size_t)t * 2 + (rowside ? 0 : 1)) * PART_K * PS_B + l;
        const double* p = part + o;
#pragma unroll
        for (int k = 0; k < PART_K; ++k) s[k] += p[k * PS_B];
        if (part2) {
#pragma unroll
            for (int k = 0; k < PART_K; ++k) s[k] += part2[o + k * PS_B];
        }
    }
#pragma unroll
    for (int k = 0; k < PART_K; ++k) qs[w][k][l] = s[k];
    __syncthreads();
    if (threadIdx.x >= PS_B || A * PS_B + (int)threadIdx.x >= n) return;
    const int i = a0 + A * PS_B + threadIdx.x;
    double v[PART_K];
#pragma unroll
    for (int k = 0; k < PART_K; ++k) v[k] = (qs[0][k][l] + qs[1][k][l]) + (qs[2][k][l] + qs[3][k][l]);
    g[3 * i] = v[0];
    g[3 * i + 1] = v[1];
    g[3 * i + 2] = v[2];
    dEdcn[i] = v[3];
    e_coul[i] = v[4];
    e_disp[i] = v[5];
    e_rep[i] = v[6];
}

// Sums the j-chunk partials of k_pairs_part in chunk order.
__global__ __launch_bounds__(256) void k_pairs_reduce(int n, int nchunk, const double* __restrict__ part,
                                                      double* __restrict__ g, double* __restrict__ dEdcn,
                                                      double* __restrict__ e_coul, double* __restrict__ e_disp,
                                                      double* __restrict__ e_rep) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc[PART_K] = {};
    for (int c = 0; c < nchunk; ++c) {
        const double* p = part + (size_t)c * PART_K * n + i;
#pragma unroll
        for (int k = 0; k < PART_K; ++k) acc[k] += p[(size_t)k * n];
    }
    g[3 * i] = acc[0];
    g[3 * i + 1] = acc[1];
    g[3 * i + 2] = acc[2];
    dEdcn[i] = acc[3];
    e_coul[i] = acc[4];
    e_disp[i] = acc[5];
    e_rep[i] = acc[6];
}

// ---------------------------------------------------------------------------
// K7: repulsion corrections over the topological neighbourhood
// (gfnff_method.cpp:9781-9872: bonded pairs excluded, H-H 1,3 / 1,4 factors)
// ---------------------------------------------------------------------------
__device__ inline void nbrep_fix_atom(int i, const double* __restrict__ x, const int* __restrict__ Z,
                                      const int* __restrict__ metal, const double* __restrict__ anb,
                                      const double* __restrict__ repz, const int* __restrict__ toff,
                                      const int* __restrict__ tidx, const int* __restrict__ tdist,
                                      double* __restrict__ g, double* __restrict__ e_rep) {
    double3 xi = ld3(x, i);
    double gx = 0, gy = 0, gz = 0, er = 0;
    for (int k = toff[i]; k < toff[i + 1]; ++k) {
        int j = tidx[k], td = tdist[k];
        int zi = Z[i], zj = Z[j];
        double fgen = rep_ff(zi, zj, metal[i], metal[j]);
        double fnew;
        if (td == 1) fnew = 0.0;
        else if (zi == 1 && zj == 1 && td == 2) fnew = 0.6290 * 1.4580;
        else if (zi == 1 && zj == 1 && td == 3) fnew = 0.6290 * 0.7080;
        else continue;
        double3 d = xi - ld3(x, j);
        double r = nrm(d);
        if (r > 20.0 || r < 1e-8) continue;
        double ab = repz[i] * repz[j] * 0.4270;
        double base = sqrt(anb[i] * anb[j]);
        double de = 0.0, dEdr = 0.0;
        {
            double al = base * fgen;
            double e = ab * exp(-al * r * sqrt(r)) / r;
            de -= e;
            dEdr -= -e / r - 1.5 * al * sqrt(r) * e;
        }
        if (fnew != 0.0) {
            double al = base * fnew;
            double e = ab * exp(-al * r * sqrt(r)) / r;
            de += e;
            dEdr += -e / r - 1.5 * al * sqrt(r) * e;
        }
        er += 0.5 * de;
        double s = dEdr / r;
        gx += s * d.x;
        gy += s * d.y;
        gz += s * d.z;
    }
    g[3 * i] += gx;
    g[3 * i + 1] += gy;
    g[3 * i + 2] += gz;
    e_rep[i] += er;
}

// ---------------------------------------------------------------------------
// K8: bonded term lists.  Each entry writes its atom gradients to fixed slots
// (gathered per atom in k_finish in a fixed order) and its energy to e_out.
// ---------------------------------------------------------------------------
// HB coordination of X-H hydrogens (ff_workspace_gfnff.cpp:46-97) and bonds (:104-180)
// bonds with an X-H hydrogen and a non-empty hydrogen-bond partner list
// (bsel: their indices): one wavefront per bond,
// lanes stride the bond's hydrogen-bond partner list (HB coordination and
// its gradient), fixed butterflies; lane 0 writes the bond's own slots

__global__ __launch_bounds__(256) void k_bonds(int nsel, const int* __restrict__ bsel, const int* __restrict__ ij,
                                               const double* __restrict__ par,
                                               const int* __restrict__ hb_off, const int* __restrict__ hb_B,
                                               const int* __restrict__ slot0, const double* __restrict__ x,
                                               const double* __restrict__ cn, const double* __restrict__ rcov_d3,
                                               const int* __restrict__ bond_H, double* __restrict__ slots,
                                               double* __restrict__ dcn_slot, double* __restrict__ e_out) {
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= nsel) return;
    const int b = bsel[w];
    int i = ij[2 * b], j = ij[2 * b + 1];
    const double* p = par + 8 * b;
    double3 d = ld3(x, i) - ld3(x, j);
    double r = nrm(d);
    double r0 = (p[0] + p[2] * cn[i] + p[1] + p[3] * cn[j] + p[4]) * p[5];
    double dr = r - r0;
    double a0 = p[6], al = a0;
    const int H = bond_H[b];
    const int k0 = (H >= 0) ? hb_off[b] : 0, k1 = (H >= 0) ? hb_off[b + 1] : 0;
    if (H >= 0) {
        const double3 xh = ld3(x, H);
        double hbcn = 0.0;
        for (int k = k0 + lane; k < k1; k += 64) {
            int B = hb_B[k];
            double3 v = ld3(x, B) - xh;
            double rr2 = dot(v, v);
            if (rr2 > 900.0) continue;
            double rr = sqrt(rr2);
            double rc = 1.78 * (4.0 / 3.0) * (rcov_d3[H] + rcov_d3[B]);
            hbcn += cn_count(-27.5 * (rr - rc) / rc);
        }
        hbcn = wave_sum(hbcn);
        al = (-(1.0 - 0.9) * hbcn + 1.0) * a0;
    }
    double e = p[7] * exp(-al * dr * dr);
    double dEdr = -2.0 * al * dr * e;
    int s = slot0[b];
    if (lane == 0) {
        e_out[b] = e;
        st3(slots, s, (dEdr / r) * d);
        st3(slots, s + 1, (-dEdr / r) * d);
        double yy = -dEdr;
        dcn_slot[2 * b] = yy * p[5] * p[2];
        dcn_slot[2 * b + 1] = yy * p[5] * p[3];
    }
    if (H >= 0) {
        double zz = 0.1 * a0 * dr * dr * e;
        const double3 xh = ld3(x, H);
        double3 gH = make_double3(0, 0, 0);
        for (int k = k0 + lane; k < k1; k += 64) {
            int B = hb_B[k];
            double3 v = ld3(x, B) - xh;
            double rr2 = dot(v, v);
            double3 gB = make_double3(0, 0, 0);
            if (rr2 <= 900.0) {
                double rr = sqrt(rr2);
                double rc = 1.78 * (4.0 / 3.0) * (rcov_d3[H] + rcov_d3[B]);
                double arg = -27.5 * (rr - rc) / rc;
                // exp(-arg^2) underflows to 0 for |arg| > 27.5: skipped there
                double dc = 0.0;
                if (arg > -27.5) dc = 0.5641895835477563 * (-27.5 / rc) * exp(-arg * arg) / rr;
                gH = gH + (-zz * dc) * v;
                gB = (zz * dc) * v;
            }
            st3(slots, s + 3 + (k - k0), gB);
        }
        gH.x = wave_sum(gH.x);
        gH.y = wave_sum(gH.y);
        gH.z = wave_sum(gH.z);
        if (lane == 0) st3(slots, s + 2, gH);
    }
}

// bonds without HB coordination (no X-H hydrogen, or an empty partner list,
// where hbcn = 0 leaves alpha = a0 exactly; bsel: their indices), one thread
// each: the arithmetic of k_bonds' lane 0; an X-H bond's hydrogen slot gets
// its zero gradient
__global__ __launch_bounds__(256) void k_bonds_plain(int nsel, const int* __restrict__ bsel,
                                                     const int* __restrict__ ij, const double* __restrict__ par,
                                                     const int* __restrict__ slot0, const int* __restrict__ bond_H,
                                                     const double* __restrict__ x, const double* __restrict__ cn,
                                                     double* __restrict__ slots, double* __restrict__ dcn_slot,
                                                     double* __restrict__ e_out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsel) return;
    const int b = bsel[t];
    const int i = ij[2 * b], j = ij[2 * b + 1];
    const double* p = par + 8 * b;
    const double3 d = ld3(x, i) - ld3(x, j);
    const double r = nrm(d);
    const double r0 = (p[0] + p[2] * cn[i] + p[1] + p[3] * cn[j] + p[4]) * p[5];
    const double dr = r - r0;
    const double al = p[6];
    const double e = p[7] * exp(-al * dr * dr);
    const double dEdr = -2.0 * al * dr * e;
    const int s = slot0[b];
    e_out[b] = e;
    st3(slots, s, (dEdr / r) * d);
    st3(slots, s + 1, (-dEdr / r) * d);
    const double yy = -dEdr;
    dcn_slot[2 * b] = yy * p[5] * p[2];
    dcn_slot[2 * b + 1] = yy * p[5] * p[3];
    if (bond_H[b] >= 0) st3(slots, s + 2, make_double3(0, 0, 0));
}

// angles: ff_workspace_gfnff.cpp:187-275
__global__ __launch_bounds__(256) void k_angles(int na, const int* __restrict__ ijk, const double* __restrict__ par,
                                                const int* __restrict__ slot0, const double* __restrict__ x,
                                                const double* __restrict__ rcov_d3, double* __restrict__ slots,
                                                double* __restrict__ e_out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na) return;
    int i = ijk[3 * t], j = ijk[3 * t + 1], k = ijk[3 * t + 2];
    double fc = par[2 * t], th0 = par[2 * t + 1];
    double3 u = ld3(x, i) - ld3(x, j), v = ld3(x, k) - ld3(x, j);
    double ru = nrm(u), rv = nrm(v);
    double cth = (ru < 1e-10 || rv < 1e-10) ? 1.0 : dot(u, v) / (ru * rv);
    cth = fmin(1.0, fmax(-1.0, cth));
    double th = acos(cth), e, dEdcos;
    if (fabs(M_PI - th0) < 1e-6) {
        double dt = th - th0;
        e = fc * dt * dt;
        double sn = sin(th);
        dEdcos = (sn > 1e-14) ? -2.0 * fc * dt / sn : 0.0;
    } else {
        double c0 = cos(th0);
        e = fc * (cth - c0) * (cth - c0);
        dEdcos = 2.0 * fc * (cth - c0);
    }
    double rci = rcov_d3[i] * (4.0 / 3.0), rcj = rcov_d3[j] * (4.0 / 3.0), rck = rcov_d3[k] * (4.0 / 3.0);
    double rc1 = 0.595 * (rci + rcj) * (rci + rcj), rc2 = 0.595 * (rcj + rck) * (rcj + rck);
    double r12 = ru * ru, r22 = rv * rv;
    double rr1 = r12 / rc1, rr2 = r22 / rc2;
    rr1 *= rr1;
    rr2 *= rr2;
    double d1 = 1.0 / (1.0 + rr1), d2 = 1.0 / (1.0 + rr2), damp = d1 * d2;
    e_out[t] = e * damp;
    double dd1 = (r12 > 1e-8) ? -4.0 * rr1 / (r12 * (1.0 + rr1) * (1.0 + rr1)) : 0.0;
    double dd2 = (r22 > 1e-8) ? -4.0 * rr2 / (r22 * (1.0 + rr2) * (1.0 + rr2)) : 0.0;
    double sinth = sqrt(fmax(0.0, 1.0 - cth * cth));
    double pre = (sinth < 1e-14 || ru < 1e-10 || rv < 1e-10) ? 0.0 : dEdcos * damp;
    double3 gi = (pre / ru) * ((1.0 / rv) * v - (cth / ru) * u) + (e * dd1 * d2) * u;
    double3 gk = (pre / rv) * ((1.0 / ru) * u - (cth / rv) * v) + (e * dd2 * d1) * v;
    int s = slot0[t];
    st3(slots, s, gi);
    st3(slots, s + 1, (-1.0) * (gi + gk));
    st3(slots, s + 2, gk);
}

// signed dihedral and its gradient (phi as in gfnff_geometry.h:60-110)
__device__ __forceinline__ double dihed(double3 ri, double3 rj, double3 rk, double3 rl, double3 g[4], bool& ok) {
    double3 v1 = rj - ri, v2 = rk - rj, v3 = rl - rk;
    double3 n1 = cross(v1, v2), n2 = cross(v2, v3);
    double nn1 = nrm(n1), nn2 = nrm(n2), nv2 = nrm(v2);
    ok = !(nn1 < 1e-10 || nn2 < 1e-10 || nv2 < 1e-10);
    if (!ok) {
        for (int a = 0; a < 4; ++a) g[a] = make_double3(0, 0, 0);
        return 0.0;
    }
    double c = dot(n1, n2) / (nn1 * nn2);
    double s = nv2 * dot(n1, v3) / nn1 / nn2;
    s = fmin(1.0, fmax(-1.0, s));
    double phi = atan2(s, c);
    double n1s = nn1 * nn1, n2s = nn2 * nn2, b2 = nv2 * nv2;
    double3 gi = (-nv2 / n1s) * n1, gl = (nv2 / n2s) * n2;
    double aa = dot(v1, v2) / b2, bb = dot(v3, v2) / b2;
    g[0] = gi;
    g[3] = gl;
    g[1] = (-1.0 - aa) * gi + bb * gl;
    g[2] = aa * gi + (-1.0 - bb) * gl;
    return phi;
}

// torsions (primary and extra): ff_workspace_gfnff.cpp:282-482
__global__ __launch_bounds__(256) void k_tors(int nt, const int* __restrict__ ijkl, const double* __restrict__ par,
                                              const int* __restrict__ slot0, const double* __restrict__ x,
                                              const double* __restrict__ rcov_d3, double* __restrict__ slots,
                                              double* __restrict__ e_out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int id[4] = {ijkl[4 * t], ijkl[4 * t + 1], ijkl[4 * t + 2], ijkl[4 * t + 3]};
    const double* p = par + 4 * t;
    double3 r[4];
    for (int a = 0; a < 4; ++a) r[a] = ld3(x, id[a]);
    double3 gp[4];
    bool ok;
    double phi = dihed(r[0], r[1], r[2], r[3], gp, ok);
    double V = p[0], nn = p[1], phi0 = p[2], atc = p[3];
    double c1 = nn * (phi - phi0) + M_PI;
    double e = V * (1.0 + cos(c1));
    double3 v[3] = {r[0] - r[1], r[1] - r[2], r[2] - r[3]};
    double dmp[3], rr[3], r2[3];
    for (int s = 0; s < 3; ++s) {
        r2[s] = dot(v[s], v[s]);
        double rs = rcov_d3[id[s]] * (4.0 / 3.0) + rcov_d3[id[s + 1]] * (4.0 / 3.0);
        double rc = atc * rs * rs;
        rr[s] = r2[s] / rc;
        rr[s] *= rr[s];
        dmp[s] = 1.0 / (1.0 + rr[s]);
    }
    double damp = dmp[0] * dmp[1] * dmp[2];
    e_out[t] = e * damp;
    double dEdphi = ok ? -V * nn * sin(c1) * damp : 0.0;
    double dd[3];
    for (int s = 0; s < 3; ++s) dd[s] = (r2[s] > 1e-8) ? -4.0 * rr[s] / (r2[s] * (1.0 + rr[s]) * (1.0 + rr[s])) : 0.0;
    double f0 = e * dd[0] * dmp[1] * dmp[2], f1 = e * dmp[0] * dd[1] * dmp[2], f2 = e * dmp[0] * dmp[1] * dd[2];
    int s = slot0[t];
    st3(slots, s, dEdphi * gp[0] + f0 * v[0]);
    st3(slots, s + 1, dEdphi * gp[1] + (-f0) * v[0] + f1 * v[1]);
    st3(slots, s + 2, dEdphi * gp[2] + (-f1) * v[1] + f2 * v[2]);
    st3(slots, s + 3, dEdphi * gp[3] + (-f2) * v[2]);
}

// out-of-plane angle of centre ri against the plane through its first
// neighbour rj (gfnff_geometry.h:207-300): n = (ri - rj) x (rk - rj),
// v = rl - ri, omega = asin(n.v / (|n||v|)), with its gradient from
// gn = v/(|n||v|) - s n/|n|^2, gv = n/(|n||v|) - s v/|v|^2.
__device__ __forceinline__ double oop(double3 ri, double3 rj, double3 rk, double3 rl, double3 g[4]) {
    double3 re = ri - rj, rd = rk - rj, rv = rl - ri;
    double3 nn = cross(re, rd);
    double N = nrm(nn), V = nrm(rv);
    for (int a = 0; a < 4; ++a) g[a] = make_double3(0, 0, 0);
    if (N < 1e-10 || V < 1e-10) return 0.0;
    double s = fmin(1.0, fmax(-1.0, dot(nn, rv) / (N * V)));
    double om = asin(s), co = cos(om);
    if (fabs(N * V * co) < 1e-14) return om;
    double3 gn = (1.0 / (N * V)) * rv + (-s / (N * N)) * nn;
    double3 gv = (1.0 / (N * V)) * nn + (-s / (V * V)) * rv;
    double3 dre = cross(rd, gn), drd = cross(gn, re);
    double ic = 1.0 / co;
    g[0] = ic * (dre + (-1.0) * gv);
    g[1] = (-ic) * (dre + drd);
    g[2] = ic * drd;
    g[3] = ic * gv;
    return om;
}

// inversions: ff_workspace_gfnff.cpp:489-604.  par: fc, omega0, type
// (0: V (1 - cos w), -1: V (cos w - cos w0)^2); damping over the distances of
// the first neighbour j to the centre and the other two neighbours.
__global__ __launch_bounds__(256) void k_inv(int nv, const int* __restrict__ ijkl, const double* __restrict__ par,
                                             const int* __restrict__ slot0, const double* __restrict__ x,
                                             const double* __restrict__ rcov_d3, double* __restrict__ slots,
                                             double* __restrict__ e_out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    int id[4] = {ijkl[4 * t], ijkl[4 * t + 1], ijkl[4 * t + 2], ijkl[4 * t + 3]};
    const double* p = par + 3 * t;
    double3 r[4];
    double rc[4];
    for (int a = 0; a < 4; ++a) {
        r[a] = ld3(x, id[a]);
        rc[a] = rcov_d3[id[a]] * (4.0 / 3.0);
    }
    double3 go[4];
    double om = oop(r[0], r[1], r[2], r[3], go);
    const int other[3] = {0, 2, 3};
    double3 dv[3];
    double r2[3], rr[3], dmp[3];
    for (int s = 0; s < 3; ++s) {
        int o = other[s];
        dv[s] = r[1] - r[o];
        r2[s] = dot(dv[s], dv[s]);
        double rcut = 0.505 * (rc[1] + rc[o]) * (rc[1] + rc[o]);
        rr[s] = (r2[s] / rcut) * (r2[s] / rcut);
        dmp[s] = 1.0 / (1.0 + rr[s]);
    }
    double damp = dmp[0] * dmp[1] * dmp[2];
    double V = p[0], et, dEdom;
    if (p[2] == 0.0) {
        et = V * (1.0 - cos(om));
        dEdom = V * sin(om) * damp;
    } else {
        double df = cos(om) - cos(p[1]);
        et = V * df * df;
        dEdom = -2.0 * V * sin(om) * df * damp;
    }
    e_out[t] = et * damp;
    double f[3];
    for (int s = 0; s < 3; ++s) {
        double dd = (r2[s] < 1e-8) ? 0.0 : -4.0 * rr[s] / (r2[s] * (1.0 + rr[s]) * (1.0 + rr[s]));
        f[s] = et * dd * dmp[(s + 1) % 3] * dmp[(s + 2) % 3];
    }
    int s = slot0[t];
    st3(slots, s, dEdom * go[0] + (-f[0]) * dv[0]);
    st3(slots, s + 1, dEdom * go[1] + f[0] * dv[0] + f[1] * dv[1] + f[2] * dv[2]);
    st3(slots, s + 2, dEdom * go[2] + (-f[1]) * dv[1]);
    st3(slots, s + 3, dEdom * go[3] + (-f[2]) * dv[2]);
}

// triple-bond torsions: ff_workspace_gfnff.cpp:611-648 (undamped),
// E = erefhalf (1 - cos 2 phi)
__global__ __launch_bounds__(256) void k_stors(int nt, const int* __restrict__ ijkl, const double* __restrict__ eh,
                                               const int* __restrict__ slot0, const double* __restrict__ x,
                                               double* __restrict__ slots, double* __restrict__ e_out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double3 gp[4];
    bool ok;
    double phi = dihed(ld3(x, ijkl[4 * t]), ld3(x, ijkl[4 * t + 1]), ld3(x, ijkl[4 * t + 2]), ld3(x, ijkl[4 * t + 3]),
                       gp, ok);
    double e = eh[t];
    e_out[t] = e - e * cos(2.0 * phi);
    double d = ok ? 2.0 * e * sin(2.0 * phi) : 0.0;
    int s = slot0[t];
    for (int a = 0; a < 4; ++a) st3(slots, s + a, d * gp[a]);
}

// bonded repulsion: ff_workspace_gfnff.cpp:774-807
__global__ __launch_bounds__(256) void k_brep(int nb, const int* __restrict__ ij, const double* __restrict__ par,
                                              const int* __restrict__ slot0, const double* __restrict__ x,
                                              double* __restrict__ slots, double* __restrict__ e_out) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int i = ij[2 * b], j = ij[2 * b + 1];
    double3 d = ld3(x, i) - ld3(x, j);
    double r = nrm(d);
    double e = 0.0, dEdr = 0.0;
    if (!(r > 20.0 || r < 1e-8)) {
        double al = par[2 * b], ab = par[2 * b + 1];
        e = ab * exp(-al * r * sqrt(r)) / r;
        dEdr = -e / r - 1.5 * al * sqrt(r) * e;
    }
    e_out[b] = e;
    int s = slot0[b];
    st3(slots, s, (dEdr / r) * d);
    st3(slots, s + 1, (-dEdr / r) * d);
}

// hydrogen bonds, cases 1-4: ff_workspace_gfnff.cpp:938-1404
// hb_core evaluates one triplet A-H...B: returns its energy and the gradient
// on A, H and B; with XTRA it also writes the gradient of the acceptor's
// neighbours to xs[k] (k-th neighbour of B) and, for case 3, of the
// neighbours D of the carbonyl carbon other than B to xs[nnb + d].  Case 3: B
// is a carbonyl O whose only neighbour is the carbon C; case 4: sp2 N
// acceptor with a virtual lone pair opposite its neighbours.  The per-triplet
// parameters of the reference (basicity, acidity, topology charges) are
// per-atom values, and the neighbour lists depend on B only.
struct HBArgs {
    const double* cb;      // basicity per atom
    const double* ca;      // acidity per atom
    const double* Qh;      // per atom e^(15 q) / (e^(15 q) + 1), q the topology charge (role H)
    const double* Qab;     // per atom e^(-15 q) / (e^(-15 q) + 1) (roles A, B)
    const int* nb_off;     // per atom: neighbours of an acceptor of cases 2-4 (CSR)
    const int* nb_idx;
    const int* parentC;    // per atom: carbonyl carbon of a case-3 acceptor, else -1
    const int* nbc_off;    // per atom: neighbours of parentC other than the acceptor (CSR)
    const int* nbc_idx;
    const double* hb_rad;
    const double* lpd;     // per atom: case-4 lone-pair distance (element table)
    const int* Z;
    const float4* pk;      // FP32 per atom {basicity, acidity, Qab, hb_rad} (roles A, B of the FP32 cores)
    const float* qh32;     // FP32 Qh (role H of the FP32 cores)
};
__device__ inline double pow6(double t) {
    const double t2 = t * t;
    return t2 * t2 * t2;
}

template <bool XTRA>
__device__ double hb_core(int A, int H, int B, int cs, const HBArgs& a, const double* __restrict__ x,
                          double3& ga_out, double3& gh_out, double3& gb_out, double* __restrict__ xs) {
    const double basA = a.cb[A], basB = a.cb[B], aciA = a.ca[A], aciB = a.ca[B];

    double3 pA = ld3(x, A), pH = ld3(x, H), pB = ld3(x, B);
    double rAH = nrm(pH - pA), rHB = nrm(pB - pH), rAB = nrm(pB - pA);
    const int nb0 = (cs >= 2) ? a.nb_off[B] : 0, nb1 = (cs >= 2) ? a.nb_off[B + 1] : 0;
    const int C3 = (cs == 3) ? a.parentC[B] : -1;
    const int c0 = (C3 >= 0) ? a.nbc_off[B] : 0, c1 = (C3 >= 0) ? a.nbc_off[B + 1] : 0;
    const int sD = nb1 - nb0;  // first extra slot of the D atoms
    double3 z3 = make_double3(0, 0, 0);
    double3 ga = z3, gb = z3, gh = z3;
    double E = 0.0;
    bool nb_written = false;
    if (rAB <= 50.0) {
        double rAH4 = rAH * rAH * rAH * rAH, rHB4 = rHB * rHB * rHB * rHB;
        double den = 1.0 / (rAH4 + rHB4);
        const double QH = a.Qh[H], QA = a.Qab[A], QB = a.Qab[B];
        double bas = (QA * basA * rAH4 + QB * basB * rHB4) * den;
        double aci = (aciB * rAH4 + aciA * rHB4) * den;
        double radab = a.hb_rad[A] + a.hb_rad[B];
        const double ratio3 = pow6(22.0 * radab / (rAB * rAB)), ratio1 = pow6(rAB * rAB / 85.0);
        double dsh = 1.0 / (1.0 + ratio3);
        double dlg = 1.0 / (1.0 + ratio1);
        double denv = dsh * dlg;
        double expo0 = (49.0 / radab) * ((rAH + rHB) / rAB - 1.0);
        double doutl = (expo0 > 15.0) ? 0.0 : 2.0 / (1.0 + exp(expo0));
        double outl_nb = 1.0;
        double nbcut = (a.Z[B] == 7 && nb1 - nb0 == 1) ? 2.0 : 11.20;
        if (cs >= 2)
            for (int k = nb0; k < nb1; ++k) {
                double3 pn = ld3(x, a.nb_idx[k]);
                double ex = (nbcut / radab) * ((nrm(pn - pA) + nrm(pn - pB)) / rAB - 1.0);
                outl_nb *= (2.0 / (1.0 + exp(-ex)) - 1.0);
            }
        double rdamp = (cs >= 2) ? denv * (1.8 / (rHB * rHB * rHB) - 0.8 / (rAB * rAB * rAB)) : denv / (rAB * rAB * rAB);
        // case 4: virtual lone pair lp = B - lpd v/|v|, v = sum_nb (x_nb - x_B)
        double outl_lp = 1.0, lpd = 0.0, xlp = 0.0;
        double3 lpv = z3, lp = pB;
        int nlp = 0;
        if (cs == 4) {
            lpd = a.lpd[B];
            for (int k = nb0; k < nb1; ++k) lpv = lpv + (ld3(x, a.nb_idx[k]) - pB);
            double vn = nrm(lpv);
            if (vn > 1e-10 && nb1 > nb0) {
                nlp = nb1 - nb0;
                lp = pB + (-lpd / vn) * lpv;
                xlp = (56.0 / radab) * ((nrm(pA - lp) + lpd + 1e-12) / rAB - 1.0);
                outl_lp = 2.0 / (1.0 + exp(xlp));
            }
        }
        // case 3: H...O=C bend and D-C=O...H torsion factors
        double eangl = 1.0, etors = 1.0;
        double3 dB3 = z3, dC3 = z3, dH3 = z3;  // d eangl / d(B, C, H)
        const double tfc = (1.0 - 0.94) / 2.0;
        if (C3 >= 0) {
            const double cc0 = cos(120.0 * M_PI / 180.0);
            const double kb = (1.0 - 0.20) / ((1.0 - cc0) * (1.0 - cc0));
            double3 u = ld3(x, C3) - pB, w = pH - pB;
            double ru = nrm(u), rw = nrm(w);
            double ca = fmin(1.0, fmax(-1.0, dot(u, w) / (ru * rw + 1e-14)));
            eangl = 1.0 - kb * (ca - cc0) * (ca - cc0);
            double dea = 2.0 * kb * (ca - cc0);
            dC3 = (-dea / ru) * ((1.0 / rw) * w + (-ca / ru) * u);
            dH3 = (-dea / rw) * ((1.0 / ru) * u + (-ca / rw) * w);
            dB3 = (-1.0) * (dC3 + dH3);
            double3 pC = ld3(x, C3);
            for (int k = c0; k < c1; ++k) {
                double3 gp[4];
                bool ok;
                double phi = dihed(ld3(x, a.nbc_idx[k]), pB, pC, pH, gp, ok);
                etors *= (1.0 + cos(2.0 * (phi - M_PI / 2.0) + M_PI)) * tfc + 0.94;
            }
        }
        double qhoutl = QH * doutl * outl_nb * outl_lp;
        double gscale = (cs == 2 || cs == 4) ? 0.268 : (cs == 3 ? 0.350 : 1.0);
        double cval = aciA * basB * QA * QB * gscale;
        E = (cs >= 2) ? -rdamp * qhoutl * cval * eangl * etors : -bas * aci * rdamp * qhoutl;
        double3 drab = pA - pB, drah = pA - pH, drbh = pB - pH;
        double rab2 = rAB * rAB, rbh2 = rHB * rHB, rah2 = rAH * rAH;
        double rahprbh = rAH + rHB + 1e-12;
        double ddamp = (-2.0 * 6.0 * ratio1 / (1.0 + ratio1)) + (2.0 * 6.0 * ratio3 / (1.0 + ratio3));
        double expo = (49.0 / radab) * (rahprbh / rAB - 1.0);
        if (!(expo > 15.0)) {
            double ratio2 = exp(expo);
            if (cs >= 2) {
                double ct = cval * eangl * etors;
                double rbhdamp = denv * 1.8 / (rbh2 * rHB), rabdamp = denv * -0.8 / (rab2 * rAB);
                double dterm = -qhoutl * ct, aterm = -rdamp * QH * outl_nb * outl_lp * ct;
                double nbterm = -rdamp * QH * doutl * outl_lp * ct;
                double gi = ((rabdamp + rbhdamp) * ddamp - 3.0 * rabdamp) / rab2 * dterm;
                ga = gi * drab;
                gb = (-gi) * drab;
                gi = -3.0 * rbhdamp / rbh2 * dterm;
                gb = gb + gi * drbh;
                gh = (-gi) * drbh;
                double tmp1 = -2.0 * aterm * ratio2 * expo / ((1.0 + ratio2) * (1.0 + ratio2)) / (rahprbh - rAB);
                gi = -tmp1 * rahprbh / rab2;
                ga = ga + gi * drab;
                gb = gb + (-gi) * drab;
                double3 dga = (tmp1 / rAH) * drah, dgb = (tmp1 / rHB) * drbh;
                ga = ga + dga;
                gb = gb + dgb;
                gh = gh + (-1.0) * (dga + dgb);
                // case 4: d outl_lp through s = |A - lp| + lpd and rAB
                double3 pg = z3;
                if (nlp > 0) {
                    double lpterm = -rdamp * QH * doutl * outl_nb * cval;
                    double3 dal = pA - lp;
                    double ral = nrm(dal), sl = ral + lpd + 1e-12;
                    double el = exp(xlp), fe = 2.0 * lpterm * el / ((1.0 + el) * (1.0 + el)) * (56.0 / radab);
                    double dEds = -fe / rAB, dEdrab = fe * sl / rab2;
                    double3 gA = (dEds / (ral + 1e-12)) * dal;
                    ga = ga + (dEdrab / rAB) * drab + gA;
                    gb = gb + (-dEdrab / rAB) * drab;
                    double3 glp = (-1.0) * gA;
                    double vn = nrm(lpv), vg = dot(lpv, glp);
                    pg = (lpd / vn) * (glp + (-vg / (vn * vn)) * lpv);
                    gb = gb + glp + (double)nlp * pg;
                }
                // case 3: bend and torsion factors
                double3 gC = z3;
                if (C3 >= 0) {
                    double bterm = -rdamp * qhoutl * etors * cval, tterm = -rdamp * qhoutl * eangl * cval;
                    gb = gb + bterm * dB3;
                    gC = bterm * dC3;
                    gh = gh + bterm * dH3;
                    double3 pC = ld3(x, C3);
                    for (int k = c0; k < c1; ++k) {
                        double3 gp[4];
                        bool ok;
                        double phi = dihed(ld3(x, a.nbc_idx[k]), pB, pC, pH, gp, ok);
                        double cc = 2.0 * (phi - M_PI / 2.0) + M_PI;
                        double etk = (1.0 + cos(cc)) * tfc + 0.94;
                        double f = ok ? (etors / etk) * tterm * (-2.0 * sin(cc) * tfc) : 0.0;
                        if (XTRA) st3(xs, sD + (k - c0), f * gp[0]);
                        gb = gb + f * gp[1];
                        gC = gC + f * gp[2];
                        gh = gh + f * gp[3];
                    }
                }
                nb_written = true;
                int kk = 3;
                for (int k = nb0; k < nb1; ++k, ++kk) {
                    double3 pn = ld3(x, a.nb_idx[k]);
                    double3 dranb = pA - pn, drbnb = pB - pn;
                    double ranb = nrm(dranb), rbnb = nrm(drbnb), sm = ranb + rbnb + 1e-12;
                    double ex = (nbcut / radab) * (sm / rAB - 1.0);
                    double r2n = exp(-ex);
                    double oi = 2.0 / (1.0 + r2n) - 1.0;
                    double others = (fabs(oi) > 1e-12) ? outl_nb / oi : 1.0;
                    double tmp2 = 2.0 * nbterm * others * r2n * ex / ((1.0 + r2n) * (1.0 + r2n)) / (sm - rAB);
                    double gi2 = -tmp2 * sm / rab2;
                    ga = ga + gi2 * drab;
                    gb = gb + (-gi2) * drab;
                    double3 da = (tmp2 / ranb) * dranb, db = (tmp2 / rbnb) * drbnb;
                    ga = ga + da;
                    gb = gb + db;
                    double3 gn = (-1.0) * (da + db + pg);
                    if (kk == 3 && C3 >= 0) gn = gn + gC;
                    if (XTRA) st3(xs, kk - 3, gn);
                }
            } else {
                double caa = QA * basA, cbb = QB * basB;
                double rterm = -aci * rdamp * qhoutl, dterm = -aci * bas * qhoutl, sterm = -rdamp * bas * qhoutl;
                double aterm = -aci * bas * rdamp * QH;
                double tmp = den * den * 4.0;
                double dd24a = rah2 * rHB4 * tmp, dd24b = rbh2 * rAH4 * tmp;
                ga = ((caa - cbb) * dd24a * rterm) * drah;
                gb = ((cbb - caa) * dd24b * rterm) * drbh;
                gh = (-1.0) * (ga + gb);
                double3 dga = ((aciB - aciA) * dd24a * sterm) * drah, dgb = ((aciA - aciB) * dd24b * sterm) * drbh;
                ga = ga + dga;
                gb = gb + dgb;
                gh = gh + (-1.0) * (dga + dgb);
                double gi = rdamp * (ddamp - 3.0) / rab2;
                ga = ga + (gi * dterm) * drab;
                gb = gb + (-gi * dterm) * drab;
                gi = aterm * 2.0 * ratio2 * expo * rahprbh / ((1.0 + ratio2) * (1.0 + ratio2)) / (rahprbh - rAB) / rab2;
                ga = ga + gi * drab;
                gb = gb + (-gi) * drab;
                double to = -2.0 * aterm * ratio2 * expo / ((1.0 + ratio2) * (1.0 + ratio2)) / (rahprbh - rAB);
                double3 da = (to / rAH) * drah, db = (to / rHB) * drbh;
                ga = ga + da;
                gb = gb + db;
                gh = gh + (-1.0) * (da + db);
            }
        }
    }
    ga_out = ga;
    gh_out = gh;
    gb_out = gb;
    if (XTRA && !nb_written) {
        for (int kk = 0; kk < nb1 - nb0; ++kk) st3(xs, kk, z3);
        for (int k = c0; k < c1; ++k) st3(xs, sD + (k - c0), z3);
    }
    return E;
}

// halogen bonds A-X...B: ff_workspace_gfnff.cpp:1411-1519.  The gradient is
// the reference's: its short- and long-range damping derivatives
// (-12 f_s t_s / (r (1 + t_s)) and -12 r f_l t_l / (70 (1 + t_l))) are kept as
// they are, so along r_XB it is not the derivative of the energy.
__global__ __launch_bounds__(256) void k_xb(int nx, const int* __restrict__ axb, const double* __restrict__ par,
                                            const int* __restrict__ slot0, const double* __restrict__ x,
                                            const double* __restrict__ hb_rad, double* __restrict__ slots,
                                            double* __restrict__ e_out) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nx) return;
    int A = axb[3 * b], X = axb[3 * b + 1], B = axb[3 * b + 2];
    const double* p = par + 3 * b;
    double3 pA = ld3(x, A), pX = ld3(x, X), pB = ld3(x, B);
    double3 vax = pX - pA, vxb = pB - pX, vab = pB - pA;
    double rax = nrm(vax), rxb = nrm(vxb), rab = nrm(vab);
    double3 z3 = make_double3(0, 0, 0), gA = z3, gX = z3, gB = z3;
    double E = 0.0;
    double radab = hb_rad[A] + hb_rad[B];
    double ex = 70.0 * ((rax + rxb) / rab - 1.0);
    if (!(rxb > 20.0) && !(ex > 15.0)) {
        double ts = pow(5.0 * radab / (rxb * rxb), 6.0), tl = pow(rxb * rxb / 70.0, 6.0);
        double fs = 1.0 / (1.0 + ts), fl = 1.0 / (1.0 + tl);
        double fo = 2.0 / (1.0 + exp(ex));
        double eX = exp(15.0 * p[1]), eB = exp(15.0 * -p[2]);
        double QX = eX / (eX + 0.03), QB = eB / (eB + 0.03);
        double r3 = rxb * rxb * rxb;
        double Rd = fs * fl * fo / r3;
        double pre = -QB * p[0] * QX;
        E = pre * Rd;
        double dfs = -2.0 * 6.0 * fs * ts / (rxb * (1.0 + ts));
        double dfl = -2.0 * 6.0 * rxb * fl * tl / (70.0 * (1.0 + tl));
        double et = exp(ex), de = (1.0 + et) * (1.0 + et);
        double dfo_ax = -2.0 * et * 70.0 / (rab * de);
        double dfo_ab = 2.0 * et * 70.0 * (rax + rxb) / (rab * rab * de);
        double dR_xb = (dfs * fl * fo + fs * dfl * fo + fs * fl * dfo_ax) / r3 - 3.0 * Rd / rxb;
        double gxb = pre * dR_xb / (rxb + 1e-14), gax = pre * (fs * fl * dfo_ax / r3) / (rax + 1e-14);
        double gab = pre * (fs * fl * dfo_ab / r3) / (rab + 1e-14);
        gA = (-gax) * vax + (-gab) * vab;
        gX = gax * vax + (-gxb) * vxb;
        gB = gxb * vxb + gab * vab;
    }
    e_out[b] = E;
    int s0 = slot0[b];
    st3(slots, s0, gA);
    st3(slots, s0 + 1, gX);
    st3(slots, s0 + 2, gB);
}

// Case-1 triplet (H bonded to neither A nor B): the cs = 1 path of hb_core
// without the neighbour, lone-pair and carbonyl factors (all 1).  T = double
// is the reference arithmetic; T = float (mixed path) forms the coordinate
// differences in FP64 and evaluates the term in FP32 with hardware rsqrt,
// rcp and exp.
__device__ inline double hb_sqrt(double v) { return sqrt(v); }
__device__ inline float hb_sqrt(float v) { return __builtin_sqrtf(v); }
__device__ inline double hb_rcp(double v) { return 1.0 / v; }
__device__ inline float hb_rcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ inline double hb_exp(double v) { return exp(v); }
__device__ inline float hb_exp(float v) { return __expf(v); }
template <typename T>
struct V3 {
    T x, y, z;
};
template <typename T>
__device__ inline V3<T> v3(double3 d) {
    return {(T)d.x, (T)d.y, (T)d.z};
}
template <typename T>
__device__ inline double3 d3(T s, V3<T> v) {
    return make_double3((double)(s * v.x), (double)(s * v.y), (double)(s * v.z));
}
// per-atom HB parameters of a triplet: FP64 arrays for the FP64 cores, the
// packed FP32 copies (two 16-byte loads and one 4-byte load) for T = float
template <typename T>
__device__ inline void hb_params(int A, int H, int B, const HBArgs& a, T& basA, T& basB, T& aciA, T& aciB, T& QH,
                                 T& QA, T& QB, T& radab) {
    if constexpr (sizeof(T) == sizeof(float)) {
        const float4 pa = a.pk[A], pb = a.pk[B];
        basA = pa.x;
        aciA = pa.y;
        QA = pa.z;
        basB = pb.x;
        aciB = pb.y;
        QB = pb.z;
        radab = pa.w + pb.w;
        QH = a.qh32[H];
    } else {
        basA = a.cb[A];
        basB = a.cb[B];
        aciA = a.ca[A];
        aciB = a.ca[B];
        QH = a.Qh[H];
        QA = a.Qab[A];
        QB = a.Qab[B];
        radab = a.hb_rad[A] + a.hb_rad[B];
    }
}
template <typename T>
__device__ double hb_core1(int A, int H, int B, const HBArgs& a, const double* __restrict__ x, double3& ga,
                           double3& gh, double3& gb) {
    const double3 z3 = make_double3(0, 0, 0);
    ga = z3;
    gh = z3;
    gb = z3;
    const double3 pA = ld3(x, A), pH = ld3(x, H), pB = ld3(x, B);
    // difference vectors drah = A - H, drbh = B - H, drab = A - B
    const V3<T> drah = v3<T>(pA - pH), drbh = v3<T>(pB - pH), drab = v3<T>(pA - pB);
    const T rah2 = drah.x * drah.x + drah.y * drah.y + drah.z * drah.z;
    const T rbh2 = drbh.x * drbh.x + drbh.y * drbh.y + drbh.z * drbh.z;
    const T rab2 = drab.x * drab.x + drab.y * drab.y + drab.z * drab.z;
    const T rAH = hb_sqrt(rah2), rHB = hb_sqrt(rbh2), rAB = hb_sqrt(rab2);
    if (!(rAB <= T(50))) return 0.0;
    T basA, basB, aciA, aciB, QH, QA, QB, radab;
    hb_params<T>(A, H, B, a, basA, basB, aciA, aciB, QH, QA, QB, radab);
    const T rAH4 = rah2 * rah2, rHB4 = rbh2 * rbh2;
    const T den = hb_rcp(rAH4 + rHB4);
    const T bas = (QA * basA * rAH4 + QB * basB * rHB4) * den;
    const T aci = (aciB * rAH4 + aciA * rHB4) * den;
    const T irab2 = hb_rcp(rab2), irAB = hb_rcp(rAB);
    const T t3 = T(22) * radab * irab2, t1 = rab2 * T(1.0 / 85.0);
    const T ratio3 = (t3 * t3) * (t3 * t3) * (t3 * t3), ratio1 = (t1 * t1) * (t1 * t1) * (t1 * t1);
    const T denv = hb_rcp(T(1) + ratio3) * hb_rcp(T(1) + ratio1);
    const T c49 = T(49) * hb_rcp(radab);
    const T expo0 = c49 * ((rAH + rHB) * irAB - T(1));
    const T doutl = (expo0 > T(15)) ? T(0) : T(2) * hb_rcp(T(1) + hb_exp(expo0));
    const T rdamp = denv * irab2 * irAB;
    const T qhoutl = QH * doutl;
    const double E = (double)(-bas * aci * rdamp * qhoutl);
    const T rahprbh = rAH + rHB + T(1e-12);
    const T ddamp = (T(-12) * ratio1 * hb_rcp(T(1) + ratio1)) + (T(12) * ratio3 * hb_rcp(T(1) + ratio3));
    const T expo = c49 * (rahprbh * irAB - T(1));
    if (expo > T(15)) return E;
    const T ratio2 = hb_exp(expo);
    const T caa = QA * basA, cbb = QB * basB;
    const T rterm = -aci * rdamp * qhoutl, dterm = -aci * bas * qhoutl, sterm = -rdamp * bas * qhoutl;
    const T aterm = -aci * bas * rdamp * QH;
    const T tmp = den * den * T(4);
    const T dd24a = rah2 * rHB4 * tmp, dd24b = rbh2 * rAH4 * tmp;
    // coefficients of drah, drbh, drab in the gradient on A and B (H: minus the sum)
    // expo / (rAH + rHB - rAB) = 49 / (radab rAB): the reference divides the two
    // (with a 1e-12 guard); the closed form stays finite for collinear A-H-B
    const T ie2 = hb_rcp((T(1) + ratio2) * (T(1) + ratio2)), eir = c49 * irAB;
    const T gi1 = rdamp * (ddamp - T(3)) * irab2 * dterm;
    const T gi2 = aterm * T(2) * ratio2 * eir * rahprbh * ie2 * irab2;
    const T to = T(-2) * aterm * ratio2 * ie2 * eir;
    const T cah = (caa - cbb) * dd24a * rterm + (aciB - aciA) * dd24a * sterm + to * hb_rcp(rAH);
    const T cbh = (cbb - caa) * dd24b * rterm + (aciA - aciB) * dd24b * sterm + to * hb_rcp(rHB);
    const T cab = gi1 + gi2;
    ga = make_double3((double)(cah * drah.x + cab * drab.x), (double)(cah * drah.y + cab * drab.y),
                      (double)(cah * drah.z + cab * drab.z));
    gb = make_double3((double)(cbh * drbh.x - cab * drab.x), (double)(cbh * drbh.y - cab * drab.y),
                      (double)(cbh * drbh.z - cab * drab.z));
    gh = make_double3(-(double)(cah * drah.x + cbh * drbh.x), -(double)(cah * drah.y + cbh * drbh.y),
                      -(double)(cah * drah.z + cbh * drbh.z));
    return E;
}

// Case-2 triplet (acceptor B with covalent neighbours, no lone-pair or
// carbonyl factor): the cs = 2 path of hb_core in the arithmetic of
// hb_core1 (T = float on the mixed path: FP64 coordinate differences, FP32
// terms).  The quotients expo / (rAH + rHB - rAB) and ex_k / (s_k - rAB) are
// taken in closed form, 49 / (radab rAB) and cnb / rAB.  xs[k]: gradient on
// the k-th neighbour of B.
template <typename T>
__device__ double hb_core2(int A, int H, int B, const HBArgs& a, const double* __restrict__ x, double3& ga,
                           double3& gh, double3& gb, double* __restrict__ xs) {
    const double3 z3 = make_double3(0, 0, 0);
    ga = z3;
    gh = z3;
    gb = z3;
    const int nb0 = a.nb_off[B], nb1 = a.nb_off[B + 1];
    const double3 pA = ld3(x, A), pH = ld3(x, H), pB = ld3(x, B);
    const V3<T> drah = v3<T>(pA - pH), drbh = v3<T>(pB - pH), drab = v3<T>(pA - pB);
    const T rah2 = drah.x * drah.x + drah.y * drah.y + drah.z * drah.z;
    const T rbh2 = drbh.x * drbh.x + drbh.y * drbh.y + drbh.z * drbh.z;
    const T rab2 = drab.x * drab.x + drab.y * drab.y + drab.z * drab.z;
    const T rAH = hb_sqrt(rah2), rHB = hb_sqrt(rbh2), rAB = hb_sqrt(rab2);
    double E = 0.0;
    bool written = false;
    if (rAB <= T(50)) {
        T basA, basB, aciA, aciB, QH, QA, QB, radab;
        hb_params<T>(A, H, B, a, basA, basB, aciA, aciB, QH, QA, QB, radab);
        const T irab2 = hb_rcp(rab2), irAB = hb_rcp(rAB), irHB = hb_rcp(rHB), irbh2 = irHB * irHB;
        const T t3 = T(22) * radab * irab2, t1 = rab2 * T(1.0 / 85.0);
        const T ratio3 = (t3 * t3) * (t3 * t3) * (t3 * t3), ratio1 = (t1 * t1) * (t1 * t1) * (t1 * t1);
        const T denv = hb_rcp(T(1) + ratio3) * hb_rcp(T(1) + ratio1);
        const T irad = hb_rcp(radab), c49 = T(49) * irad;
        const T expo0 = c49 * ((rAH + rHB) * irAB - T(1));
        const T doutl = (expo0 > T(15)) ? T(0) : T(2) * hb_rcp(T(1) + hb_exp(expo0));
        const T cnb = ((a.Z[B] == 7 && nb1 - nb0 == 1) ? T(2.0) : T(11.20)) * irad;
        T outl_nb = T(1);
        for (int k = nb0; k < nb1; ++k) {
            const double3 pn = ld3(x, a.nb_idx[k]);
            const V3<T> dan = v3<T>(pA - pn), dbn = v3<T>(pB - pn);
            const T sm = hb_sqrt(dan.x * dan.x + dan.y * dan.y + dan.z * dan.z) +
                         hb_sqrt(dbn.x * dbn.x + dbn.y * dbn.y + dbn.z * dbn.z);
            const T ex = cnb * (sm * irAB - T(1));
            outl_nb *= T(2) * hb_rcp(T(1) + hb_exp(-ex)) - T(1);
        }
        const T rbhdamp = denv * T(1.8) * irbh2 * irHB, rabdamp = denv * T(-0.8) * irab2 * irAB;
        const T rdamp = rbhdamp + rabdamp;
        const T cval = aciA * basB * QA * QB * T(0.268);
        const T qhoutl = QH * doutl * outl_nb;
        E = (double)(-rdamp * qhoutl * cval);
        const T rahprbh = rAH + rHB + T(1e-12);
        const T ddamp = (T(-12) * ratio1 * hb_rcp(T(1) + ratio1)) + (T(12) * ratio3 * hb_rcp(T(1) + ratio3));
        const T expo = c49 * (rahprbh * irAB - T(1));
        if (!(expo > T(15))) {
            written = true;
            const T ratio2 = hb_exp(expo);
            const T dterm = -qhoutl * cval, aterm = -rdamp * QH * outl_nb * cval;
            const T nbterm = -rdamp * QH * doutl * cval;
            const T ie2 = hb_rcp((T(1) + ratio2) * (T(1) + ratio2));
            const T tmp1 = T(-2) * aterm * ratio2 * ie2 * c49 * irAB;
            // coefficients of drab (A: +, B: -), drah (A: +, H: -), drbh (B: +, H: -)
            T cab = ((rabdamp + rbhdamp) * ddamp - T(3) * rabdamp) * irab2 * dterm - tmp1 * rahprbh * irab2;
            const T cah = tmp1 * hb_rcp(rAH);
            const T cbh = T(-3) * rbhdamp * irbh2 * dterm + tmp1 * irHB;
            T gax = T(0), gay = T(0), gaz = T(0), gbx = T(0), gby = T(0), gbz = T(0);
            int kk = 0;
            for (int k = nb0; k < nb1; ++k, ++kk) {
                const double3 pn = ld3(x, a.nb_idx[k]);
                const V3<T> dan = v3<T>(pA - pn), dbn = v3<T>(pB - pn);
                const T ranb = hb_sqrt(dan.x * dan.x + dan.y * dan.y + dan.z * dan.z);
                const T rbnb = hb_sqrt(dbn.x * dbn.x + dbn.y * dbn.y + dbn.z * dbn.z);
                const T sm = ranb + rbnb + T(1e-12);
                const T ex = cnb * (sm * irAB - T(1));
                const T r2n = hb_exp(-ex), ir = hb_rcp(T(1) + r2n);
                const T oi = T(2) * ir - T(1);
                const T others = (fabs(oi) > T(1e-12)) ? outl_nb * hb_rcp(oi) : T(1);
                const T tmp2 = T(2) * nbterm * others * r2n * ir * ir * cnb * irAB;
                cab -= tmp2 * sm * irab2;
                const T fa = tmp2 * hb_rcp(ranb), fb = tmp2 * hb_rcp(rbnb);
                gax += fa * dan.x;
                gay += fa * dan.y;
                gaz += fa * dan.z;
                gbx += fb * dbn.x;
                gby += fb * dbn.y;
                gbz += fb * dbn.z;
                st3(xs, kk, make_double3(-(double)(fa * dan.x + fb * dbn.x), -(double)(fa * dan.y + fb * dbn.y),
                                         -(double)(fa * dan.z + fb * dbn.z)));
            }
            ga = make_double3((double)(cah * drah.x + cab * drab.x + gax), (double)(cah * drah.y + cab * drab.y + gay),
                              (double)(cah * drah.z + cab * drab.z + gaz));
            gb = make_double3((double)(cbh * drbh.x - cab * drab.x + gbx), (double)(cbh * drbh.y - cab * drab.y + gby),
                              (double)(cbh * drbh.z - cab * drab.z + gbz));
            gh = make_double3(-(double)(cah * drah.x + cbh * drbh.x), -(double)(cah * drah.y + cbh * drbh.y),
                              -(double)(cah * drah.z + cbh * drbh.z));
        }
    }
    if (!written)
        for (int kk = 0; kk < nb1 - nb0; ++kk) st3(xs, kk, z3);
    return E;
}

// Case-1 hydrogen bonds by role: the case-1 triplets are kept three times,
// grouped by the atom in role A, H or B (entries: the other two atoms), each
// group cut into chunks of at most HB_CHUNK triplets.  One wavefront per
// chunk, one kernel per role, evaluates every triplet of its chunk (lanes
// stride the chunk; a triplet is evaluated once per role, which costs less
// than storing and gathering per-triplet gradients), adds the gradient on its
// own atom (and, for role A, the energy) with a fixed butterfly and writes
// one result per chunk.
constexpr int HB_CHUNK = 256;
struct HBRoles {
    const int2* ent;
    const int* goff;    // [ngt + 1]
    const int* gatom;   // [ngt]
    int ng0, ng1, ngt;  // role A chunks [0, ng0), H [ng0, ng1), B [ng1, ngt)
};
template <int ROLE, typename T>
__global__ __launch_bounds__(256) void k_hb_role(HBRoles r, HBArgs a, const double* __restrict__ x,
                                                 double* __restrict__ gout, double* __restrict__ egrp) {
    const int lane = threadIdx.x & 63;
    const int g0 = (ROLE == 0) ? 0 : (ROLE == 1 ? r.ng0 : r.ng1), g1 = (ROLE == 0) ? r.ng0 : (ROLE == 1 ? r.ng1 : r.ngt);
    const int w = g0 + blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= g1) return;
    const int at = r.gatom[w], e0 = r.goff[w], e1 = r.goff[w + 1];
    double3 g = make_double3(0, 0, 0);
    double E = 0.0;
    for (int e = e0 + lane; e < e1; e += 64) {
        const int2 en = r.ent[e];
        double3 ga, gh, gb;
        if (ROLE == 0) {
            E += hb_core1<T>(at, en.x, en.y, a, x, ga, gh, gb);
            g = g + ga;
        } else if (ROLE == 1) {
            hb_core1<T>(en.x, at, en.y, a, x, ga, gh, gb);
            g = g + gh;
        } else {
            hb_core1<T>(en.x, en.y, at, a, x, ga, gh, gb);
            g = g + gb;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        g.x += __shfl_xor(g.x, o);
        g.y += __shfl_xor(g.y, o);
        g.z += __shfl_xor(g.z, o);
        if (ROLE == 0) E += __shfl_xor(E, o);
    }
    if (lane == 0) {
        st3(gout, w, g);
        if (ROLE == 0) egrp[w] = E;
    }
}

// Hydrogen bonds of cases 2-4, one thread each: energy, and the gradient in
// the triplet's slots s0 (A), s0 + 1 (H), s0 + 2 (B), s0 + 3 + k (k-th
// neighbour of B), then the carbonyl's D atoms.  MODE 0: every entry by the
// FP64 core; MODE 1: case-2 entries only, by hb_core2<T> (the mixed path's
// FP32 terms); MODE 2: entries of cases 3 and 4 only, FP64.  ord (MODE 1
// and 2): the entries of the mode, case 2 first, so that no lane idles on
// an entry of the other mode; nh counts the mode's entries.
template <typename T, int MODE>
__global__ __launch_bounds__(256, 2) void k_hb_full(int nh, const int* __restrict__ ord, const int4* __restrict__ ahbc,
                                                 const int* __restrict__ slot0, HBArgs a, const double* __restrict__ x,
                                                 double* __restrict__ slots, double* __restrict__ e_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nh) return;
    const int h = MODE == 0 ? k : ord[k];
    const int4 t = ahbc[h];
    const int s0 = slot0[h];
    double3 ga, gh, gb;
    double* xs = slots + 3 * ((size_t)s0 + 3);
    if ((MODE == 1 && t.w != 2) || (MODE == 2 && t.w == 2)) return;
    if (MODE == 1)
        e_out[h] = hb_core2<T>(t.x, t.y, t.z, a, x, ga, gh, gb, xs);
    else
        e_out[h] = hb_core<true>(t.x, t.y, t.z, t.w, a, x, ga, gh, gb, xs);
    st3(slots, s0, ga);
    st3(slots, s0 + 1, gh);
    st3(slots, s0 + 2, gb);
}

// per-atom hydrogen-bond gradient, one wavefront per atom: the atom's
// case-1 chunk results in roles A, H, B (chunks [crange[r (n + 1) + i],
// crange[r (n + 1) + i + 1])) and its slots of cases 2-4; lanes stride the
// lists, fixed butterfly
template <int LPA>
__global__ __launch_bounds__(256) void k_hb_gather(int n, const int* __restrict__ crange,
                                                   const double* __restrict__ gout, const int* __restrict__ xoff,
                                                   const int* __restrict__ xidx, const double* __restrict__ xslots,
                                                   double* __restrict__ g) {
    const int lane = threadIdx.x & (LPA - 1);
    const int i = blockIdx.x * (256 / LPA) + threadIdx.x / LPA;
    if (i >= n) return;
    double3 s = make_double3(0, 0, 0);
#pragma unroll
    for (int role = 0; role < 3; ++role) {
        const int* cr = crange + role * (n + 1);
        for (int k = cr[i] + lane; k < cr[i + 1]; k += LPA) s = s + ld3(gout, k);
    }
    for (int k = xoff[i] + lane; k < xoff[i + 1]; k += LPA) s = s + ld3(xslots, xidx[k]);
#pragma unroll
    for (int o = LPA / 2; o > 0; o >>= 1) {
        s.x += __shfl_xor(s.x, o);
        s.y += __shfl_xor(s.y, o);
        s.z += __shfl_xor(s.z, o);
    }
    if (lane == 0) st3(g, i, s);
}

// D4 ATM over bonded triplets (energy +c9, gradient -c9 as in
// ff_workspace_gfnff.cpp:1526-1658) and bonded ATM (:1665-1739)
__device__ inline double c6_pair(int i, int j, int ne, const int* elem, const int* nref, const double* c6ref,
                                 const double* gw) {
    int ei = elem[i], ej = elem[j];
    const double* c = c6ref + ((size_t)ei * ne + ej) * 49;
    double s = 0.0;
    for (int a = 0; a < nref[ei]; ++a)
        for (int b = 0; b < nref[ej]; ++b) s += gw[(size_t)i * 7 + a] * gw[(size_t)j * 7 + b] * c[a * 7 + b];
    return s;
}

__global__ __launch_bounds__(256) void k_atm(int nt, const int* __restrict__ ijk, const double* __restrict__ sc,
                                             const int* __restrict__ slot0, const double* __restrict__ x,
                                             const double* __restrict__ rcov_d3, int ne, const int* __restrict__ elem,
                                             const int* __restrict__ nref, const double* __restrict__ c6ref,
                                             const double* __restrict__ gw, double* __restrict__ slots,
                                             double* __restrict__ e_out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int i = ijk[3 * t], j = ijk[3 * t + 1], k = ijk[3 * t + 2];
    double c9 = sqrt(fabs(c6_pair(i, j, ne, elem, nref, c6ref, gw) * c6_pair(i, k, ne, elem, nref, c6ref, gw) *
                          c6_pair(j, k, ne, elem, nref, c6ref, gw)));
    double3 vij = ld3(x, j) - ld3(x, i), vik = ld3(x, k) - ld3(x, i), vjk = ld3(x, k) - ld3(x, j);
    double rij = nrm(vij), rik = nrm(vik), rjk = nrm(vjk);
    double r2ij = rij * rij, r2ik = rik * rik, r2jk = rjk * rjk;
    double r0ij = 0.58 * sqrt(3.0 * rcov_d3[i] * rcov_d3[j]) + 4.80;
    double r0ik = 0.58 * sqrt(3.0 * rcov_d3[i] * rcov_d3[k]) + 4.80;
    double r0jk = 0.58 * sqrt(3.0 * rcov_d3[j] * rcov_d3[k]) + 4.80;
    double rijk = rij * rik * rjk, r2ijk = r2ij * r2ik * r2jk, r3ijk = rijk * r2ijk, r5ijk = r2ijk * r3ijk;
    double tmp = pow(r0ij * r0ik * r0jk / rijk, 14.0 / 3.0);
    double fdmp = 1.0 / (1.0 + 6.0 * tmp);
    double Aa = r2ij + r2jk - r2ik, Bb = r2ij + r2ik - r2jk, Cc = r2ik + r2jk - r2ij;
    double ang = (0.375 * Aa * Bb * Cc / r2ijk + 1.0) / r3ijk;
    double s = sc[t];
    e_out[t] = ang * fdmp * c9 / 3.0 * s;
    double dfdmp = -2.0 * 14.0 * tmp * fdmp * fdmp;
    double dij = -0.375 * (cube(r2ij) + sq(r2ij) * (r2jk + r2ik) +
                           r2ij * (3.0 * sq(r2jk) + 2.0 * r2jk * r2ik + 3.0 * sq(r2ik)) -
                           5.0 * sq(r2jk - r2ik) * (r2jk + r2ik)) / r5ijk;
    double dik = -0.375 * (cube(r2ik) + sq(r2ik) * (r2jk + r2ij) +
                           r2ik * (3.0 * sq(r2jk) + 2.0 * r2jk * r2ij + 3.0 * sq(r2ij)) -
                           5.0 * sq(r2jk - r2ij) * (r2jk + r2ij)) / r5ijk;
    double djk = -0.375 * (cube(r2jk) + sq(r2jk) * (r2ik + r2ij) +
                           r2jk * (3.0 * sq(r2ik) + 2.0 * r2ik * r2ij + 3.0 * sq(r2ij)) -
                           5.0 * sq(r2ik - r2ij) * (r2ik + r2ij)) / r5ijk;
    double pre = -c9 * s / 3.0;
    double a1 = pre * (-dij * fdmp + ang * dfdmp) / r2ij, a2 = pre * (-dik * fdmp + ang * dfdmp) / r2ik,
           a3 = pre * (-djk * fdmp + ang * dfdmp) / r2jk;
    int q = slot0[t];
    st3(slots, q, (-1.0) * (a1 * vij + a2 * vik));
    st3(slots, q + 1, a1 * vij + (-a3) * vjk);
    st3(slots, q + 2, a2 * vik + a3 * vjk);
}

__global__ __launch_bounds__(256) void k_batm(int nt, const int* __restrict__ ijk, const double* __restrict__ zb,
                                              const int* __restrict__ slot0, const double* __restrict__ x,
                                              const double* __restrict__ qa, double* __restrict__ slots,
                                              double* __restrict__ e_out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int i = ijk[3 * t], j = ijk[3 * t + 1], k = ijk[3 * t + 2];
    double3 vij = ld3(x, j) - ld3(x, i), vik = ld3(x, k) - ld3(x, i), vjk = ld3(x, k) - ld3(x, j);
    double r2ij = dot(vij, vij), r2jk = dot(vjk, vjk), r2ik = dot(vik, vik);
    double rij = sqrt(r2ij), rjk = sqrt(r2jk), rik = sqrt(r2ik);
    double rijk3 = r2ij * r2jk * r2ik;
    double mijk = -r2ij + r2jk + r2ik, imjk = r2ij - r2jk + r2ik, ijmk = r2ij + r2jk - r2ik;
    double ang = 0.375 * ijmk * imjk * mijk / rijk3;
    double rav3 = rijk3 * sqrt(rijk3);
    double angr9 = (ang + 1.0) / rav3;
    double fi = fmin(fmax(1.0 - 3.0 * qa[i], -4.0), 4.0);
    double fj = fmin(fmax(1.0 - 3.0 * qa[j], -4.0), 4.0);
    double fk = fmin(fmax(1.0 - 3.0 * qa[k], -4.0), 4.0);
    double c9 = fi * fj * fk * zb[3 * t] * zb[3 * t + 1] * zb[3 * t + 2];
    e_out[t] = c9 * angr9;
    double dij = -0.375 * (cube(r2ij) + sq(r2ij) * (r2jk + r2ik) +
                           r2ij * (3.0 * sq(r2jk) + 2.0 * r2jk * r2ik + 3.0 * sq(r2ik)) -
                           5.0 * sq(r2jk - r2ik) * (r2jk + r2ik)) / (rij * rijk3 * rav3);
    double djk = -0.375 * (cube(r2jk) + sq(r2jk) * (r2ik + r2ij) +
                           r2jk * (3.0 * sq(r2ik) + 2.0 * r2ik * r2ij + 3.0 * sq(r2ij)) -
                           5.0 * sq(r2ik - r2ij) * (r2ik + r2ij)) / (rjk * rijk3 * rav3);
    double dik = -0.375 * (cube(r2ik) + sq(r2ik) * (r2jk + r2ij) +
                           r2ik * (3.0 * sq(r2jk) + 2.0 * r2jk * r2ij + 3.0 * sq(r2ij)) -
                           5.0 * sq(r2jk - r2ij) * (r2jk + r2ij)) / (rik * rijk3 * rav3);
    double3 gij = (-dij * c9 / rij) * vij, gjk = (-djk * c9 / rjk) * vjk, gik = (-dik * c9 / rik) * vik;
    int q = slot0[t];
    st3(slots, q, gij + gik);                 // i
    st3(slots, q + 1, (-1.0) * gij + gjk);    // j
    st3(slots, q + 2, (-1.0) * (gik + gjk));  // k
}

// ---------------------------------------------------------------------------
// K11: batches of independent molecules (one launch per kernel for the whole
// batch).  Molecule m owns atoms [mol_off[m], mol_off[m+1]); every all-pairs
// loop is restricted to the molecule, the bonded lists are concatenated.  Its
// reduced EEQ system occupies the padded block [vo_m, vo_m + np_m) of the
// concatenated reduced index space, its matrix lies at A + mo_m (np_m x np_m,
// row-major) and the inverses of its diagonal blocks at Linv + vo_m NB.
// ---------------------------------------------------------------------------
constexpr int BATCH_MAX_NP = 512;  // largest padded reduced system of a batched molecule

// One 64 x 64 lower tile (m, bi, bj) of a molecule's reduced matrix per
// workgroup.  When the tile's rows share one last atom la and its columns
// one last atom lb (always within one fragment, so for every tile of a
// one-fragment molecule), the entries A(ia, lb), A(la, ib) and A(la, lb) are
// evaluated once per row, column and tile into LDS and every element needs
// one fresh A(ia, ib); the sum keeps the order of the direct four-term form,
// so R is the same bit for bit.  Other tiles use the direct form.  A32 (the
// mixed path) receives the same entries rounded to FP32.
__global__ __launch_bounds__(256) void k_eeq_reduced_batch(const int4* __restrict__ tiles, const int* __restrict__ mnr,
                                                           const int* __restrict__ mnp,
                                                           const long long* __restrict__ mo,
                                                           const int* __restrict__ vo, const int* __restrict__ ra,
                                                           const int* __restrict__ rl, const double* __restrict__ x,
                                                           const double* __restrict__ alp,
                                                           const double* __restrict__ diag, double* __restrict__ A,
                                                           float* __restrict__ A32) {
    __shared__ double urow[NB], wcol[NB];
    __shared__ double sll;
    const int4 tt = tiles[blockIdx.x];
    const int m = tt.x, bi = tt.y, bj = tt.z;
    const int nr = mnr[m], np = mnp[m], base = vo[m];
    const int tid = threadIdx.x;
    double* Am = A + mo[m];
    // shared last atoms of the tile's valid rows and columns (-1 when none is valid)
    const int r0 = bi * NB, c0 = bj * NB;
    const int la0 = r0 < nr ? rl[base + r0] : -1, lb0 = c0 < nr ? rl[base + c0] : -1;
    bool same = true;
    if (tid < NB) same = !(r0 + tid < nr) || rl[base + r0 + tid] == la0;
    else if (tid < 2 * NB) same = !(c0 + tid - NB < nr) || rl[base + c0 + tid - NB] == lb0;
    const bool uni = __syncthreads_and(same) && la0 >= 0 && lb0 >= 0;
    if (uni) {
        if (tid < NB) {
            const int r = r0 + tid;
            urow[tid] = r < nr ? eeq_a(ra[base + r], lb0, x, alp, diag) : 0.0;
        } else if (tid < 2 * NB) {
            const int c = c0 + tid - NB;
            wcol[tid - NB] = c < nr ? eeq_a(la0, ra[base + c], x, alp, diag) : 0.0;
        } else if (tid == 2 * NB) {
            sll = eeq_a(la0, lb0, x, alp, diag);
        }
        __syncthreads();
    }
    for (int e = tid; e < NB * NB; e += blockDim.x) {
        const int r = r0 + e / NB, c = c0 + e % NB;
        if (c > r) continue;
        double v;
        if (r >= nr || c >= nr) {
            v = (r == c) ? 1.0 : 0.0;
        } else if (uni) {
            v = eeq_a(ra[base + r], ra[base + c], x, alp, diag) - urow[e / NB] - wcol[e % NB] + sll;
        } else {
            const int ia = ra[base + r], la = rl[base + r], ib = ra[base + c], lb = rl[base + c];
            v = eeq_a(ia, ib, x, alp, diag) - eeq_a(ia, lb, x, alp, diag) - eeq_a(la, ib, x, alp, diag) +
                eeq_a(la, lb, x, alp, diag);
        }
        Am[(size_t)r * np + c] = v;
        if (A32) A32[mo[m] + (size_t)r * np + c] = (float)v;  // the mixed path factors this copy
    }
}

// ---- per-molecule factorisation: tile phases shared out over wavefronts ----
// One 64 x 64 output tile per wavefront on MFMA (v_mfma_f64_16x16x4f64 or
// v_mfma_f32_16x16x4f32), in two 32-row halves, operands read from memory
// (L2).  Lane layout: A(i = lane%16, k = lane/16), B(k = lane/16,
// j = lane%16); result element v of a 16 x 16 block at row 4 v + lane/16
// (FP64) or 4 (lane/16) + v (FP32), column lane%16.  Operands are read DEPTH 4-column steps
// ahead (the L2 latency exceeds one step's eight MFMAs); DEPTH 1 keeps the
// register footprint of three workgroups per CU.
//   SUB = true : A(ri.., rj..) -= P(ri.., 0..63) Q(rj.., 0..63)^T (trailing
//                update; P = Q = the panel, leading dimension ldq = np;
//                diagonal tiles keep the lower triangle)
//   SUB = false: A(ri.., rj..)  = P(ri.., 0..63) Q(0.., 0..63)^T (panel
//                solve, Q = the inverse of the diagonal block, ldq = 64)
template <typename T, bool SUB, int DEPTH>
__device__ __forceinline__ void batch_tile_mfma(T* __restrict__ Am, int np, int ri, int rj,
                                                const T* __restrict__ P, const T* __restrict__ Q, int ldq) {
    using M = MfmaT<T>;
    using V = typename M::V;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const bool lower_only = SUB && ri == rj;
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
        const T* Pi = P + (size_t)(32 * h) * np;
        V acc[2][4];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x][y] = V{0, 0, 0, 0};
        if constexpr (DEPTH == 1) {
            // the next step's operands are loaded before the current step's MFMAs
            T fa[2], fb[4];
#pragma unroll
            for (int x = 0; x < 2; ++x) fa[x] = Pi[(size_t)(16 * x + li) * np + lk];
#pragma unroll
            for (int y = 0; y < 4; ++y) fb[y] = Q[(size_t)(16 * y + li) * ldq + lk];
#pragma unroll 2
            for (int kk = 0; kk < NB; kk += 4) {
                T na[2], nb[4];
                if (kk + 4 < NB) {
#pragma unroll
                    for (int x = 0; x < 2; ++x) na[x] = Pi[(size_t)(16 * x + li) * np + kk + 4 + lk];
#pragma unroll
                    for (int y = 0; y < 4; ++y) nb[y] = Q[(size_t)(16 * y + li) * ldq + kk + 4 + lk];
                }
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y)
                        acc[x][y] = M::op(fa[x], fb[y], acc[x][y]);
                if (kk + 4 < NB) {
#pragma unroll
                    for (int x = 0; x < 2; ++x) fa[x] = na[x];
#pragma unroll
                    for (int y = 0; y < 4; ++y) fb[y] = nb[y];
                }
            }
        } else {
            // operands of DEPTH 4-column steps in flight (ring of registers)
            T fa[DEPTH][2], fb[DEPTH][4];
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
#pragma unroll
                for (int x = 0; x < 2; ++x) fa[d][x] = Pi[(size_t)(16 * x + li) * np + 4 * d + lk];
#pragma unroll
                for (int y = 0; y < 4; ++y) fb[d][y] = Q[(size_t)(16 * y + li) * ldq + 4 * d + lk];
            }
#pragma unroll
            for (int st = 0; st < NB / 4; ++st) {
                const int sl = st % DEPTH;
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y)
                        acc[x][y] = M::op(fa[sl][x], fb[sl][y], acc[x][y]);
                if (st + DEPTH < NB / 4) {
                    const int kk = 4 * (st + DEPTH);
#pragma unroll
                    for (int x = 0; x < 2; ++x) fa[sl][x] = Pi[(size_t)(16 * x + li) * np + kk + lk];
#pragma unroll
                    for (int y = 0; y < 4; ++y) fb[sl][y] = Q[(size_t)(16 * y + li) * ldq + kk + lk];
                }
            }
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = 32 * h + 16 * x + (M::kRowFromReg ? 4 * v + lk : 4 * lk + v), c = 16 * y + li;
                    T* p = Am + (size_t)(ri + r) * np + rj + c;
                    if (!SUB) *p = acc[x][y][v];
                    else if (!lower_only || c <= r) *p = *p - acc[x][y][v];
                }
    }
}

// The diagonal-block routine as a call of its own, so that its register
// footprint is not added to the state of the surrounding phases.
template <typename T>
__device__ __noinline__ void batch_diag_block(int np, int k0, T* __restrict__ A, T* __restrict__ Linv,
                                              int* __restrict__ flags, T (*S0)[NB + 1], T (*col)[NB], T* rdiag) {
    chol_diag_block<T>(np, k0, A, Linv, flags, S0, col, rdiag);
}

// z = L^-T L^-1 v for one molecule (blocked substitutions with the stored
// inverses of the diagonal blocks): v in vec (LDS, type T), result in vec.
template <typename T>
__device__ __forceinline__ void batch_substitute(int np, const T* __restrict__ Am, const T* __restrict__ Lm, T* vec,
                                                 T* tmp) {
    const int tid = threadIdx.x, t = tid >> 2, q = tid & 3, nb = np / NB;
    // forward: y_kb = Linv_kb (v_kb - sum_{k < kb} L(kb, k) y_k)
    for (int kb = 0; kb < nb; ++kb) {
        T acc = (q == 0) ? vec[kb * NB + t] : T(0);
        for (int k = 0; k < kb; ++k)
#pragma unroll
            for (int u = 0; u < NB / 4; ++u)
                acc -= Am[(size_t)(kb * NB + t) * np + k * NB + 4 * u + q] * vec[k * NB + 4 * u + q];
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        if (q == 0) tmp[t] = acc;
        __syncthreads();
        const T* Li = Lm + (size_t)kb * NB * NB;
        T s = 0;
#pragma unroll
        for (int u = 0; u < NB / 4; ++u) s += Li[t * NB + 4 * u + q] * tmp[4 * u + q];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (q == 0) vec[kb * NB + t] = s;
        __syncthreads();
    }
    // backward: x_kb = Linv_kb^T (y_kb - sum_{k > kb} L(k, kb)^T x_k)
    for (int kb = nb - 1; kb >= 0; --kb) {
        T acc = (q == 0) ? vec[kb * NB + t] : T(0);
        for (int k = nb - 1; k > kb; --k)
#pragma unroll
            for (int u = 0; u < NB / 4; ++u)
                acc -= Am[(size_t)(k * NB + 4 * u + q) * np + kb * NB + t] * vec[k * NB + 4 * u + q];
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        if (q == 0) tmp[t] = acc;
        __syncthreads();
        const T* Li = Lm + (size_t)kb * NB * NB;
        T s = 0;
#pragma unroll
        for (int u = 0; u < NB / 4; ++u) s += Li[(4 * u + q) * NB + t] * tmp[4 * u + q];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (q == 0) vec[kb * NB + t] = s;
        __syncthreads();
    }
}

// Factorisation of one molecule's reduced system per workgroup of four
// wavefronts, right-looking by 64-column blocks.  Step k: the workgroup
// factors the diagonal block and forms its inverse (chol_diag_block, a call
// of its own so that its registers do not add to the other phases'); the
// panel tiles L(i, k) = A(i, k) L(k, k)^-T and then the trailing tiles
// A(i, j) -= L(i, k) L(j, k)^T are computed one tile per wavefront on MFMA,
// dealt round-robin, with one barrier after each phase.  T = double factors
// the FP64 matrix in place; T = float (mixed path, batch_fp32 = 1) factors
// the FP32 copy.  The factorisation needs only the matrix, so it runs on the
// side stream right after the matrix build, alongside CN and the bonded
// terms; the substitutions (k_bsolve_subst) follow once the right-hand side
// is known.  Registers and LDS let MINB workgroups share a CU, so the
// diagonal-block phases of one molecule overlap the MFMA phases of others.
constexpr int BSOLVE_T = 256;
constexpr int BSOLVE_NREF = 2;  // refinement steps of the FP32 factor
template <typename T, int MINB, int DEPTH>
__global__ __launch_bounds__(BSOLVE_T, MINB) void k_eeq_batch_factor(const int* __restrict__ order,
                                                                  const int* __restrict__ mnp,
                                                                  const long long* __restrict__ mo,
                                                                  const int* __restrict__ vo, T* __restrict__ A,
                                                                  T* __restrict__ Linv, int* __restrict__ flags) {
    __shared__ T S0[NB][NB + 1];
    __shared__ T col[2][NB];
    __shared__ T rdiag[NB];
    constexpr int NW = BSOLVE_T / 64;
    const int m = order[blockIdx.x], np = mnp[m], nb = np / NB, base = vo[m];  // largest systems first
    T* Am = A + mo[m];
    T* Lm = Linv + (size_t)base * NB;
    const int wave = threadIdx.x >> 6;
    for (int kb = 0; kb < nb; ++kb) {
        const int k0 = kb * NB;
        batch_diag_block<T>(np, k0, Am, Lm, flags, S0, col, rdiag);
        __syncthreads();
        for (int bi = kb + 1 + wave; bi < nb; bi += NW)
            batch_tile_mfma<T, false, DEPTH>(Am, np, bi * NB, k0, Am + (size_t)bi * NB * np + k0,
                                             Lm + (size_t)kb * NB * NB, NB);
        __syncthreads();
        int tile = 0;
        for (int bi = kb + 1; bi < nb; ++bi)
            for (int bj = kb + 1; bj <= bi; ++bj, ++tile)
                if (tile % NW == wave)
                    batch_tile_mfma<T, true, DEPTH>(Am, np, bi * NB, bj * NB, Am + (size_t)bi * NB * np + k0,
                                                    Am + (size_t)bj * NB * np + k0, np);
        __syncthreads();
    }
}

// Substitutions of one molecule's factored system (workgroup of BSOLVE_T
// threads): x = L^-T L^-1 b, and for T = float BSOLVE_NREF FP64 refinement
// steps against Rm (FP64, lower triangle).
template <typename T>
__device__ __forceinline__ void batch_solve_tail(int np, const T* __restrict__ Am, const T* __restrict__ Lm,
                                                 const double* __restrict__ Rm, const double* __restrict__ b,
                                                 double* __restrict__ x, T* vec, T* tmp, double* xs) {
    const int tid = threadIdx.x, t = tid >> 2, q = tid & 3;
    for (int r = tid; r < np; r += BSOLVE_T) vec[r] = (T)b[r];
    __syncthreads();
    batch_substitute<T>(np, Am, Lm, vec, tmp);
    if constexpr (sizeof(T) == 4) {
        for (int r = tid; r < np; r += BSOLVE_T) xs[r] = (double)vec[r];
        __syncthreads();
        for (int it = 0; it < BSOLVE_NREF; ++it) {
            // r_i = b_i - (sum_{j <= i} R(i, j) x_j + sum_{j > i} R(j, i) x_j), four lanes per row
            for (int i0 = 0; i0 < np; i0 += NB) {
                const int i = i0 + t;
                double acc = 0.0;
                for (int j = q; j <= i; j += 4) acc += Rm[(size_t)i * np + j] * xs[j];
                for (int j = i + 1 + q; j < np; j += 4) acc += Rm[(size_t)j * np + i] * xs[j];
                acc += __shfl_xor(acc, 1);
                acc += __shfl_xor(acc, 2);
                if (q == 0) vec[i] = (T)(b[i] - acc);
            }
            __syncthreads();
            batch_substitute<T>(np, Am, Lm, vec, tmp);
            for (int r = tid; r < np; r += BSOLVE_T) xs[r] += (double)vec[r];
            __syncthreads();
        }
        for (int r = tid; r < np; r += BSOLVE_T) x[r] = xs[r];
    } else {
        for (int r = tid; r < np; r += BSOLVE_T) x[r] = (double)vec[r];
    }
}

// ---- per-molecule factorisation by steps over the whole batch (batch_solver 3)
// Step k is three launches over every molecule with more than k column
// blocks: the diagonal blocks (one workgroup each), the panel tiles and the
// trailing tiles (one wavefront per tile, four per workgroup, any molecule).
// Every tile of a phase is independent, so a launch holds thousands of
// independent tiles and needs no barrier inside a molecule.  The
// substitutions follow in one launch (k_bsolve_subst) on the evaluation
// stream.
template <typename T>
__global__ __launch_bounds__(256, 3) void k_bstep_diag(const int* __restrict__ mols, int k, const int* __restrict__ mnp,
                                                    const long long* __restrict__ mo, const int* __restrict__ vo,
                                                    T* __restrict__ A, T* __restrict__ Linv, int* __restrict__ flags) {
    __shared__ T S0[NB][NB + 1];
    __shared__ T col[2][NB];
    __shared__ T rdiag[NB];
    const int m = mols[blockIdx.x];
    chol_diag_block<T>(mnp[m], k * NB, A + mo[m], Linv + (size_t)vo[m] * NB, flags, S0, col, rdiag);
}

// FP64 panel tile over the whole 64-row tile in one product (the column
// panel of the inverse read once per tile instead of once per half):
// A(ri.., rj..) = P(ri.., 0..63) Q(0..63, 0..63)^T, P the tile itself (leading
// dimension np), Q the inverse of the diagonal block (64 x 64, zero above the
// diagonal); operands eight columns at a time, the next eight in flight.
__device__ __forceinline__ void panel_tile_whole(double* __restrict__ Am, int np, int ri, int rj,
                                                 const double* __restrict__ P, const double* __restrict__ Q) {
    using M = MfmaT<double>;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    dbl4 acc[4][4];
#pragma unroll
    for (int xx = 0; xx < 4; ++xx)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[xx][y] = dbl4{0.0, 0.0, 0.0, 0.0};
    dbl2 fa[4], fb[4];
#pragma unroll
    for (int xx = 0; xx < 4; ++xx) fa[xx] = *(const dbl2*)(P + (size_t)(16 * xx + li) * np + 2 * lk);
#pragma unroll
    for (int y = 0; y < 4; ++y) fb[y] = *(const dbl2*)(Q + (16 * y + li) * NB + 2 * lk);
#pragma unroll 1
    for (int kc = 0; kc < NB; kc += 8) {
        const bool more = kc + 8 < NB;
        dbl2 na[4], nb[4];
        if (more) {
#pragma unroll
            for (int xx = 0; xx < 4; ++xx) na[xx] = *(const dbl2*)(P + (size_t)(16 * xx + li) * np + kc + 8 + 2 * lk);
#pragma unroll
            for (int y = 0; y < 4; ++y) nb[y] = *(const dbl2*)(Q + (16 * y + li) * NB + kc + 8 + 2 * lk);
        }
#pragma unroll
        for (int st = 0; st < 2; ++st)
#pragma unroll
            for (int xx = 0; xx < 4; ++xx)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[xx][y] = M::op(fa[xx][st], fb[y][st], acc[xx][y]);
        if (more) {
#pragma unroll
            for (int xx = 0; xx < 4; ++xx) fa[xx] = na[xx];
#pragma unroll
            for (int y = 0; y < 4; ++y) fb[y] = nb[y];
        }
    }
#pragma unroll
    for (int xx = 0; xx < 4; ++xx)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) Am[(size_t)(ri + 16 * xx + 4 * v + lk) * np + rj + 16 * y + li] = acc[xx][y][v];
}

// tiles {m, bi, bj}: panel tiles (SUB = false, bj unused) or trailing tiles of step k
template <typename T, bool SUB>
__global__ __launch_bounds__(256, 2) void k_bstep_tiles(const int4* __restrict__ tl, int ntl, int k,
                                                     const int* __restrict__ mnp, const long long* __restrict__ mo,
                                                     const int* __restrict__ vo, T* __restrict__ A,
                                                     const T* __restrict__ Linv) {
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= ntl) return;  // wavefront-local work: no barrier follows
    const int4 tt = tl[w];
    const int m = tt.x, bi = tt.y, bj = tt.z, np = mnp[m], k0 = k * NB;
    T* Am = A + mo[m];
    if constexpr (!SUB && sizeof(T) == 8)
        panel_tile_whole(Am, np, bi * NB, k0, Am + (size_t)bi * NB * np + k0,
                         Linv + (size_t)vo[m] * NB + (size_t)k * NB * NB);
    else if (!SUB)
        batch_tile_mfma<T, false, 4>(Am, np, bi * NB, k0, Am + (size_t)bi * NB * np + k0,
                                     Linv + (size_t)vo[m] * NB + (size_t)k * NB * NB, NB);
    else
        batch_tile_mfma<T, true, 4>(Am, np, bi * NB, bj * NB, Am + (size_t)bi * NB * np + k0,
                                    Am + (size_t)bj * NB * np + k0, np);
}

template <typename T>
__global__ __launch_bounds__(BSOLVE_T) void k_bsolve_subst(const int* __restrict__ order, const int* __restrict__ mnp,
                                                           const long long* __restrict__ mo,
                                                           const int* __restrict__ vo, const T* __restrict__ A,
                                                           const T* __restrict__ Linv,
                                                           const double* __restrict__ A64,
                                                           const double* __restrict__ b, double* __restrict__ x) {
    __shared__ T vec[BATCH_MAX_NP];
    __shared__ T tmp[NB];
    __shared__ double xs[sizeof(T) == 4 ? BATCH_MAX_NP : 1];
    const int m = order[blockIdx.x], base = vo[m];
    batch_solve_tail<T>(mnp[m], A + mo[m], Linv + (size_t)base * NB, A64 ? A64 + mo[m] : nullptr, b + base, x + base,
                        vec, tmp, xs);
}

// ---- per-molecule left-looking factorisation with the matrix formed in place (batch_solver 4)
// Tile (i, j), i >= j, of column block j on one wavefront:
//   C = R(i, j) - L(i, 0:K) L(j, 0:K)^T,  K = 64 j,
// stored at the tile, in two 32-row halves.  First the half's entries of R
// are evaluated from the coordinates with the operations of
// k_eeq_reduced_batch, lane = column, one row at a time: one-fragment
// molecules (uni) as A(ia, ib) - u_a - u_b + A(L, L) with u_r = A(ia_r, L),
// the row's data (coordinates, alpha, u) read from the lane of the half that
// loaded it (readlane, no dependent gather per row); other molecules in the
// direct four-term form; padded rows and columns as the identity.  They are
// stored at the tile (the lines stay in L2) and read back in the MFMA
// accumulator layout as -R; the product adds L(i, :) L(j, :)^T over the whole
// K on FP64 MFMA, and C = -acc is stored.  The operands are read eight
// columns at a time as 16-byte vectors (lane group g holds columns 2 g and
// 2 g + 1 of the eight, MFMA step s takes column 2 g + s: a permutation of
// the columns that both operands share), the next eight columns in flight
// while the current ones are multiplied.
__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// C = -(R - L(r0:r0+16XX, 0:K) L(j, 0:K)^T) over 16 XX rows from r0: element v
// of block (xx, y) at row r0 + 16 xx + 4 v + lk, column 64 j + 16 y + li, read
// in the MFMA accumulator layout after the wavefront's own entry stores;
// the operands eight columns at a time, the next eight in flight.
template <int XX>
__device__ __forceinline__ void fused_product(double* __restrict__ Am, int np, int r0, int j) {
    using M = MfmaT<double>;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int K = j * NB;
    if (K == 0) return;  // column block 0: C = R, already stored by the entry pass
    const double* Qj = Am + (size_t)(j * NB) * np;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    double* pt = Am + (size_t)(r0 + lk) * np + j * NB + li;  // this lane's element (0, 0, 0)
    dbl4 acc[XX][4];
#pragma unroll
    for (int xx = 0; xx < XX; ++xx)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[xx][y][v] = -pt[(size_t)(16 * xx + 4 * v) * np + 16 * y];
    if (K > 0) {
        const double* Pi = Am + (size_t)r0 * np;
        dbl2 fa[XX], fb[4];
#pragma unroll
        for (int xx = 0; xx < XX; ++xx) fa[xx] = *(const dbl2*)(Pi + (size_t)(16 * xx + li) * np + 2 * lk);
#pragma unroll
        for (int y = 0; y < 4; ++y) fb[y] = *(const dbl2*)(Qj + (size_t)(16 * y + li) * np + 2 * lk);
#pragma unroll 1
        for (int kc = 0; kc < K; kc += 8) {
            const bool more = kc + 8 < K;
            dbl2 na[XX], nb[4];
            if (more) {
#pragma unroll
                for (int xx = 0; xx < XX; ++xx)
                    na[xx] = *(const dbl2*)(Pi + (size_t)(16 * xx + li) * np + kc + 8 + 2 * lk);
#pragma unroll
                for (int y = 0; y < 4; ++y) nb[y] = *(const dbl2*)(Qj + (size_t)(16 * y + li) * np + kc + 8 + 2 * lk);
            }
#pragma unroll
            for (int st = 0; st < 2; ++st)
#pragma unroll
                for (int xx = 0; xx < XX; ++xx)
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[xx][y] = M::op(fa[xx][st], fb[y][st], acc[xx][y]);
            if (more) {
#pragma unroll
                for (int xx = 0; xx < XX; ++xx) fa[xx] = na[xx];
#pragma unroll
                for (int y = 0; y < 4; ++y) fb[y] = nb[y];
            }
        }
    }
    // the element addresses are formed again here rather than kept in
    // registers across the product (an empty asm makes pt opaque)
    asm volatile("" : "+v"(pt));
#pragma unroll
    for (int xx = 0; xx < XX; ++xx)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) pt[(size_t)(16 * xx + 4 * v) * np + 16 * y] = -acc[xx][y][v];
}

// MODE 2 of fused_tile: the tile is formed transposed, D = C^T (rows the
// column block's 64 rows cj, columns the tile's rows ri), so that the
// accumulator is directly the B operand of the panel product
//   L(i, j)^T = Linv(j) C^T,   i.e.  L(i, j) = C(i, j) Linv(j)^T,
// the product of the panel launch of the other column orders, without the
// round trip of C through memory.  Lj: the inverse of the diagonal block,
// 64 x 64 row-major, zero above the diagonal.
__device__ __forceinline__ void fused_product_panel(double* __restrict__ Am, int np, int r0, int j,
                                                    const double* __restrict__ Lj) {
    using M = MfmaT<double>;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int K = j * NB;
    const double* Qj = Am + (size_t)(j * NB) * np;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // element v of block (yy, xx): D(16 yy + 4 v + lk, 16 xx + li) = C(r0 + 16 xx + li, 64 j + 16 yy + 4 v + lk)
    double* pr = Am + (size_t)(r0 + li) * np + j * NB + lk;
    dbl4 acc[4][4];
#pragma unroll
    for (int yy = 0; yy < 4; ++yy)
#pragma unroll
        for (int xx = 0; xx < 4; ++xx)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[yy][xx][v] = -pr[(size_t)(16 * xx) * np + 16 * yy + 4 * v];
    if (K > 0) {
        const double* Pi = Am + (size_t)r0 * np;
        dbl2 fa[4], fb[4];
#pragma unroll
        for (int xx = 0; xx < 4; ++xx) fa[xx] = *(const dbl2*)(Pi + (size_t)(16 * xx + li) * np + 2 * lk);
#pragma unroll
        for (int y = 0; y < 4; ++y) fb[y] = *(const dbl2*)(Qj + (size_t)(16 * y + li) * np + 2 * lk);
#pragma unroll 1
        for (int kc = 0; kc < K; kc += 8) {
            const bool more = kc + 8 < K;
            dbl2 na[4], nb[4];
            if (more) {
#pragma unroll
                for (int xx = 0; xx < 4; ++xx)
                    na[xx] = *(const dbl2*)(Pi + (size_t)(16 * xx + li) * np + kc + 8 + 2 * lk);
#pragma unroll
                for (int y = 0; y < 4; ++y) nb[y] = *(const dbl2*)(Qj + (size_t)(16 * y + li) * np + kc + 8 + 2 * lk);
            }
#pragma unroll
            for (int st = 0; st < 2; ++st)
#pragma unroll
                for (int yy = 0; yy < 4; ++yy)
#pragma unroll
                    for (int xx = 0; xx < 4; ++xx) acc[yy][xx] = M::op(fb[yy][st], fa[xx][st], acc[yy][xx]);
            if (more) {
#pragma unroll
                for (int xx = 0; xx < 4; ++xx) fa[xx] = na[xx];
#pragma unroll
                for (int y = 0; y < 4; ++y) fb[y] = nb[y];
            }
        }
    }
    // acc = -C^T.  Per column block xx of the tile's rows: out(mm) = sum over
    // the 16 four-row chunks (yy, v) of Linv(16 mm + li, 16 yy + 4 v + lk)
    // times chunk (yy, v) of acc, = -L(i, j)^T block (mm, xx)
    asm volatile("" : "+v"(pr));
#pragma unroll
    for (int xx = 0; xx < 4; ++xx) {
        dbl4 out[4];
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) out[mm] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int yy = 0; yy < 4; ++yy)
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int mm = 0; mm < 4; ++mm)
                    out[mm] = M::op(Lj[(16 * mm + li) * NB + 16 * yy + 4 * v + lk], acc[yy][xx][v], out[mm]);
#pragma unroll
        for (int mm = 0; mm < 4; ++mm)
#pragma unroll
            for (int v = 0; v < 4; ++v) pr[(size_t)(16 * xx) * np + 16 * mm + 4 * v] = -out[mm][v];
    }
}

// MODE 0: product per 32-row half; 1: one product over the 64-row tile, so
// that the column's own row panel L(j, 0:K) is read once per tile (needs the
// registers of two wavefronts per SIMD); 2: as 1, transposed, followed by
// the panel product with Lj (fused_product_panel)
template <int MODE>
__device__ __forceinline__ void fused_tile(double* __restrict__ Am, int np, int nr, int i, int j,
                                           const int* __restrict__ rab, const int* __restrict__ rlb,
                                           const double* __restrict__ x, const double* __restrict__ alp,
                                           const double* __restrict__ diag, const double* su, double sll, bool uni,
                                           const double* __restrict__ Lj = nullptr) {
    const int lane = threadIdx.x & 63;
    // this lane's column of the entry pass
    const int Cc = j * NB + lane;
    const bool cin = Cc < nr;
    const int ib = cin ? rab[Cc] : 0, lb = (cin && !uni) ? rlb[Cc] : 0;
    const double3 cx = (cin && uni) ? ld3(x, ib) : make_double3(0.0, 0.0, 0.0);
    const double ca = (cin && uni) ? alp[ib] : 0.0, cu = (cin && uni) ? su[Cc] : 0.0;
    const double cdg = (cin && uni) ? diag[ib] : 0.0;
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
        const int r0 = i * NB + 32 * h;
        if (uni) {
            // row r0 + (lane & 31) loaded by this lane
            const int Rl = r0 + (lane & 31);
            const bool rinl = Rl < nr;
            const int ial = rinl ? rab[Rl] : 0;
            const double3 rxl = rinl ? ld3(x, ial) : make_double3(0.0, 0.0, 0.0);
            const double ral = rinl ? alp[ial] : 0.0, rul = rinl ? su[Rl] : 0.0;
#pragma unroll 4
            for (int rr = 0; rr < 32; ++rr) {
                const int R = r0 + rr;
                const double3 rx = make_double3(readlane_d(rxl.x, rr), readlane_d(rxl.y, rr), readlane_d(rxl.z, rr));
                const double ra = readlane_d(ral, rr), ru = readlane_d(rul, rr);
                double a;
                if (R >= nr || !cin) {
                    a = (R == Cc) ? 1.0 : 0.0;
                } else {
                    const double3 d = rx - cx;
                    const double aij = (R == Cc) ? cdg : eeq_a_r2(dot(d, d), ra, ca);
                    a = aij - ru - cu + sll;
                }
                Am[(size_t)R * np + Cc] = a;
            }
        } else {
#pragma unroll 2
            for (int rr = 0; rr < 32; ++rr) {
                const int R = r0 + rr;
                double a;
                if (R >= nr || !cin) {
                    a = (R == Cc) ? 1.0 : 0.0;
                } else {
                    const int ia = rab[R], la = rlb[R];
                    a = eeq_a(ia, ib, x, alp, diag) - eeq_a(ia, lb, x, alp, diag) - eeq_a(la, ib, x, alp, diag) +
                        eeq_a(la, lb, x, alp, diag);
                }
                Am[(size_t)R * np + Cc] = a;
            }
        }
        if (MODE == 0) fused_product<2>(Am, np, r0, j);
    }
    if (MODE == 1) fused_product<4>(Am, np, i * NB, j);
    if (MODE == 2) fused_product_panel(Am, np, i * NB, j, Lj);
}

// One workgroup of four wavefronts per molecule (largest systems first),
// left-looking by 64-column blocks.  Column block j: the tiles (i, j), i >= j,
// are formed (fused_tile, one per wavefront, dealt round-robin); the diagonal
// block is factored and inverted (chol_diag_block); the tiles below it become
// L(i, j) = C(i, j) L(j, j)^-T (MFMA, one per wavefront).  Against the
// right-looking form every tile is written once, the factored panels are read
// once per column block with the product accumulated in registers over the
// whole K, and the matrix is never built as a separate pass (no R in HBM):
// the only matrix traffic is L, so the kernel reads and writes about a third
// of the bytes.  The substitutions follow in k_bsolve_subst<double>.
template <int MINB>
__global__ __launch_bounds__(BSOLVE_T, MINB) void k_eeq_batch_fused(
    const int* __restrict__ order, const int* __restrict__ mnr, const int* __restrict__ mnp,
    const long long* __restrict__ mo, const int* __restrict__ vo, const int* __restrict__ ra,
    const int* __restrict__ rl, const double* __restrict__ x, const double* __restrict__ alp,
    const double* __restrict__ diag, double* __restrict__ A, double* __restrict__ Linv, int* __restrict__ flags) {
    __shared__ double S0[NB][NB + 1];
    __shared__ double col[2][NB];
    __shared__ double rdiag[NB];
    __shared__ double su[BATCH_MAX_NP];
    constexpr int NW = BSOLVE_T / 64;
    const int m = order[blockIdx.x], nr = mnr[m], np = mnp[m], nb = np / NB, base = vo[m];
    const int tid = threadIdx.x, wave = tid >> 6;
    double* Am = A + mo[m];
    double* Lm = Linv + (size_t)base * NB;
    const int* rab = ra + base;
    const int* rlb = rl + base;
    // one fragment: every row's last atom is row 0's
    const int L0 = nr > 0 ? rlb[0] : 0;
    bool same = true;
    for (int r = tid; r < nr; r += BSOLVE_T) same = same && rlb[r] == L0;
    const bool uni = __syncthreads_and(same) && nr > 0;
    if (uni)
        for (int r = tid; r < nr; r += BSOLVE_T) su[r] = eeq_a(rab[r], L0, x, alp, diag);
    const double sll = uni ? eeq_a(L0, L0, x, alp, diag) : 0.0;
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        for (int i = j + wave; i < nb; i += NW) fused_tile<0>(Am, np, nr, i, j, rab, rlb, x, alp, diag, su, sll, uni);
        __syncthreads();
        batch_diag_block<double>(np, j * NB, Am, Lm, flags, S0, col, rdiag);
        __syncthreads();
        for (int i = j + 1 + wave; i < nb; i += NW)
            batch_tile_mfma<double, false, 1>(Am, np, i * NB, j * NB, Am + (size_t)i * NB * np + j * NB,
                                              Lm + (size_t)j * NB * NB, NB);
        __syncthreads();
    }
}

// ---- left-looking factorisation by column blocks over the whole batch (batch_solver 5)
// Column block j is three launches over every molecule with more than j
// column blocks: the tiles (i, j), i >= j, formed by fused_tile (one
// wavefront per tile, any molecule), the diagonal blocks (k_bstep_diag) and
// the panel tiles L(i, j) = C(i, j) L(j, j)^-T (k_bstep_tiles<double,
// false>).  The per-molecule form's parallelism (four wavefronts) becomes
// thousands of independent tiles per launch, and the entries of R are still
// evaluated once, inside the tile that needs them.
// u_v = A(ia_v, last atom of v's fragment) for every reduced row (0 on padding)
__global__ __launch_bounds__(256) void k_red_u(int nv, const int* __restrict__ ra, const int* __restrict__ rl,
                                               const double* __restrict__ x, const double* __restrict__ alp,
                                               const double* __restrict__ diag, double* __restrict__ u) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    u[v] = (ra[v] >= 0) ? eeq_a(ra[v], rl[v], x, alp, diag) : 0.0;
}

// tiles {m, i, one-fragment flag, 0} of column block j
__global__ __launch_bounds__(256, 2) void k_lstep_tiles(const int4* __restrict__ tl, int ntl, int j,
                                                     const int* __restrict__ mnr, const int* __restrict__ mnp,
                                                     const long long* __restrict__ mo, const int* __restrict__ vo,
                                                     const int* __restrict__ ra, const int* __restrict__ rl,
                                                     const double* __restrict__ x, const double* __restrict__ alp,
                                                     const double* __restrict__ diag, const double* __restrict__ u,
                                                     double* __restrict__ A) {
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= ntl) return;  // wavefront-local work: no barrier follows
    const int4 tt = tl[w];
    const int m = tt.x, base = vo[m];
    const bool uni = tt.z != 0;
    const double sll = uni ? eeq_a(rl[base], rl[base], x, alp, diag) : 0.0;
    fused_tile<1>(A + mo[m], mnp[m], mnr[m], tt.y, j, ra + base, rl + base, x, alp, diag, u + base, sll, uni);
}

// batch_solver 6: the tiles below the diagonal of column block j after its
// diagonal block is factored: formed and multiplied by Linv(j) in one pass,
// L(i, j) stored directly (fused_tile MODE 2)
__global__ __launch_bounds__(256, 2) void k_lpanel_tiles(const int4* __restrict__ tl, int ntl, int j,
                                                      const int* __restrict__ mnr, const int* __restrict__ mnp,
                                                      const long long* __restrict__ mo, const int* __restrict__ vo,
                                                      const int* __restrict__ ra, const int* __restrict__ rl,
                                                      const double* __restrict__ x, const double* __restrict__ alp,
                                                      const double* __restrict__ diag, const double* __restrict__ u,
                                                      double* __restrict__ A, const double* __restrict__ Linv) {
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= ntl) return;  // wavefront-local work: no barrier follows
    const int4 tt = tl[w];
    const int m = tt.x, base = vo[m];
    const bool uni = tt.z != 0;
    const double sll = uni ? eeq_a(rl[base], rl[base], x, alp, diag) : 0.0;
    fused_tile<2>(A + mo[m], mnp[m], mnr[m], tt.y, j, ra + base, rl + base, x, alp, diag, u + base, sll, uni,
                  Linv + (size_t)base * NB + (size_t)j * NB * NB);
}

// Previous per-molecule solve (whole workgroup per tile step), kept for
// comparison: the blocked right-looking Cholesky (diag / panel / MFMA tile
// update, the routines of the single-molecule path) and both substitutions,
// all inside the workgroup.  b, x: concatenated padded vectors.
__global__ __launch_bounds__(256) void k_eeq_batch_solve_tiles(const int* __restrict__ order,
                                                         const int* __restrict__ mnp, const long long* __restrict__ mo,
                                                         const int* __restrict__ vo, double* __restrict__ A,
                                                         double* __restrict__ Linv, const double* __restrict__ b,
                                                         double* __restrict__ x, int* __restrict__ flags) {
    __shared__ double S0[NB][NB + 1];
    __shared__ double S1[NB][NB + 1];
    __shared__ double col[2][NB];
    __shared__ double rdiag[NB];
    __shared__ double vec[BATCH_MAX_NP];
    __shared__ double tmp[NB];
    const int m = order[blockIdx.x], np = mnp[m], nb = np / NB, base = vo[m];  // largest systems first
    double* Am = A + mo[m];
    double* Lm = Linv + (size_t)base * NB;
    const int tid = threadIdx.x, t = tid >> 2, q = tid & 3;
    for (int kb = 0; kb < nb; ++kb) {
        chol_diag_block<double>(np, kb * NB, Am, Lm, flags, S0, col, rdiag);
        __syncthreads();
        for (int bi = kb + 1; bi < nb; ++bi) {
            chol_panel_tile<double>(np, kb * NB, bi * NB, Am, Lm, S0, S1);
            __syncthreads();
        }
        for (int bi = kb + 1; bi < nb; ++bi)
            for (int bj = kb + 1; bj <= bi; ++bj) {
                chol_update_tile<double>(np, bi * NB, bj * NB, kb * NB, NB, Am, S0, S1);
                __syncthreads();
            }
    }
    for (int r = tid; r < np; r += 256) vec[r] = b[base + r];
    __syncthreads();
    // forward: y_kb = Linv_kb (b_kb - sum_{k < kb} L(kb, k) y_k), y kept in vec
    for (int kb = 0; kb < nb; ++kb) {
        double acc = (q == 0) ? vec[kb * NB + t] : 0.0;
        for (int k = 0; k < kb; ++k)
#pragma unroll
            for (int u = 0; u < NB / 4; ++u)
                acc -= Am[(size_t)(kb * NB + t) * np + k * NB + 4 * u + q] * vec[k * NB + 4 * u + q];
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        if (q == 0) tmp[t] = acc;
        __syncthreads();
        const double* Li = Lm + (size_t)kb * NB * NB;
        double s = 0.0;
#pragma unroll
        for (int u = 0; u < NB / 4; ++u) s += Li[t * NB + 4 * u + q] * tmp[4 * u + q];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (q == 0) vec[kb * NB + t] = s;
        __syncthreads();
    }
    // backward: x_kb = Linv_kb^T (y_kb - sum_{k > kb} L(k, kb)^T x_k)
    for (int kb = nb - 1; kb >= 0; --kb) {
        double acc = (q == 0) ? vec[kb * NB + t] : 0.0;
        for (int k = nb - 1; k > kb; --k)
#pragma unroll
            for (int u = 0; u < NB / 4; ++u)
                acc -= Am[(size_t)(k * NB + 4 * u + q) * np + kb * NB + t] * vec[k * NB + 4 * u + q];
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        if (q == 0) tmp[t] = acc;
        __syncthreads();
        const double* Li = Lm + (size_t)kb * NB * NB;
        double s = 0.0;
#pragma unroll
        for (int u = 0; u < NB / 4; ++u) s += Li[(4 * u + q) * NB + t] * tmp[4 * u + q];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (q == 0) vec[kb * NB + t] = s;
        __syncthreads();
    }
    for (int r = tid; r < np; r += 256) x[base + r] = vec[r];
}

// Per-molecule total energy: one wavefront per molecule sums the per-atom
// terms over its atoms and the per-entry terms over its list segments
// (seg[t * (nmol + 1) + m]) in a fixed order.
struct MolEnergyArgs {
    int nmol;
    const int* mol_off;
    const int* seg;        // [11][nmol + 1]: bond, angle, torsion, brep, hb case 1 (role-A groups), atm, batm,
                           // inversion, stors, xb, hb cases 2-4
    const double* atom_terms[4];   // coul, disp, rep, self
    const double* list_terms[11];
    double* e_mol;
};
__global__ __launch_bounds__(256) void k_mol_energy(MolEnergyArgs a) {
    const int lane = threadIdx.x & 63;
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (m >= a.nmol) return;
    double acc = 0.0;
    for (int k = 0; k < 4; ++k)
        for (int i = a.mol_off[m] + lane; i < a.mol_off[m + 1]; i += 64) acc += a.atom_terms[k][i];
    for (int k = 0; k < 11; ++k) {
        const int* sg = a.seg + k * (a.nmol + 1);
        for (int i = sg[m] + lane; i < sg[m + 1]; i += 64) acc += a.list_terms[k][i];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) a.e_mol[m] = acc;
}

// ---------------------------------------------------------------------------
// K10: parameter adjoint.  For per-element scale factors theta of the
// repulsion charge (repz), the repulsion exponent (rep_anb / repa) and the D4
// sqrt(Z r4r2), and a weight field w (w_i = dLoss/dg_i of a loss on the
// gradient g, e.g. 2 (g - g_ref) for force matching), the kernels accumulate
//   dLoss/dln(theta_p) = sum_i w_i . d g_i / d ln(theta_p).
// Only the D4 pair term (energy gradient and its dE/dCN chain) and the bonded
// and non-bonded repulsion depend on these parameters; ATM uses rcov_d3, the
// EEQ and CN are parameter-free here.  Output layout: [p * ne + e] with
// p = 0 repz, 1 exponent, 2 sqrt(Z r4r2).
// ---------------------------------------------------------------------------
constexpr int ADJ_P = 3;

// u_a = dlog_a sum_{b in nb(a)} dd(r_ab) (w_a - w_b) . (x_a - x_b) / r_ab:
// the contraction of w with d CN_a / dx (same arithmetic as k_finish).
__global__ __launch_bounds__(256) void k_adj_u(int n, const double* __restrict__ x, const double* __restrict__ rcov,
                                               const double* __restrict__ dlog, const int* __restrict__ nbn,
                                               const int* __restrict__ nbl, const double* __restrict__ w,
                                               double* __restrict__ u) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const double3 xa = ld3(x, a), wa = ld3(w, a);
    double acc = 0.0;
    for (int k = 0; k < nbn[a]; ++k) {
        const int b = nbl[(size_t)a * MAXNB_CN + k];
        const double3 d = xa - ld3(x, b);
        const double r = nrm(d);
        const double rs = rcov[a] + rcov[b];
        const double dr = (r - rs) / rs;
        const double dd = (KN / SQRTPI_TRUNC) * exp(-KN * KN * dr * dr) / rs;
        acc += dd * dot(wa - ld3(w, b), d) / r;
    }
    u[a] = dlog[a] * acc;
}

// derivatives of the repulsion pair term e = ab exp(-al r^1.5) / r:
//   dE'/d ln ab = E',  dE'/d al = -0.5 sqrt(r) e + 1.5 al r^2 e
__device__ inline void rep_adj(double ab, double al, double r, double& dp_lnab, double& dp_al) {
    const double sr = sqrt(r);
    const double e = ab * exp(-al * r * sr) / r;
    dp_lnab = -e / r - 1.5 * al * sr * e;
    dp_al = -0.5 * sr * e + 1.5 * al * r * r * e;
}

// D4 pair, derivatives with respect to q = r4r2 of E' (=dE/dr) and of the row
// atom's dE/dCN share -dci sum zeta (ff_workspace_gfnff.cpp:715-767):
__device__ inline void d4_adj(double c6, double dci, double zc, double q, double r2, double r, double& dEp_dq,
                              double& ddcn_dq) {
    const double r0 = 0.58 * sqrt(q) + 4.80;
    const double dr0 = 0.29 / sqrt(q);
    const double r0s = r0 * r0;
    const double r6 = r2 * r2 * r2, r06 = r0s * r0s * r0s;
    const double t6 = 1.0 / (r6 + r06), t8 = 1.0 / (r6 * r2 + r06 * r0s);
    const double dt6 = -t6 * t6 * 6.0 * r06 / r0 * dr0;        // d t6 / dq
    const double dt8 = -t8 * t8 * 8.0 * r06 * r0 * dr0;        // d t8 / dq
    const double d8 = -8.0 * r2 * r2 * r2 * t8 * t8;
    const double dd6 = -6.0 * r2 * r2 * 2.0 * t6 * dt6;
    const double dd8 = -8.0 * r2 * r2 * r2 * 2.0 * t8 * dt8;
    dEp_dq = -c6 * zc * (dd6 + 2.0 * d8 + 2.0 * q * dd8) * r;
    ddcn_dq = -dci * zc * (dt6 + 2.0 * t8 + 2.0 * q * dt8);
}

// Pair adjoint over full rows in j chunks (the pair kernel's decomposition).
// Row i collects sum_j (w_i . r_hat_ij) dE'_ij/dln(theta) + u_i d(dEdcn_i)/dln(theta)
// for its own element and the partner's; part[chunk][k * n + i], k = p * ne + e.
template <int NE>
__global__ __launch_bounds__(TILE) void k_adj_pairs(PairArgs a, const double* __restrict__ w,
                                                    const double* __restrict__ u, int cj,
                                                    double* __restrict__ part) {
    __shared__ double sx[TILE], sy[TILE], sz[TILE], ssz[TILE], san[TILE], srz[TILE], szt[TILE];
    __shared__ double sgw[TILE * 7];
    __shared__ int sZ[TILE], sel[TILE], smt[TILE];
    const int n = a.n;
    const int i = blockIdx.x * TILE + threadIdx.x;
    const bool act = i < n;
    int j0, j1;
    pair_range(a, cj, j0, j1);
    const int mlo = (a.mol_of && act) ? a.mol_off[a.mol_of[i]] : 0;
    const int mhi = (a.mol_of && act) ? a.mol_off[a.mol_of[i] + 1] : n;
    const double3 xi = act ? ld3(a.x, i) : make_double3(0, 0, 0);
    const double3 wi = act ? ld3(w, i) : make_double3(0, 0, 0);
    const double ui = act ? u[i] : 0.0, zti = act ? a.zeta[i] : 0.0, szi = act ? a.szr[i] : 0.0;
    const double ani = act ? a.anb[i] : 0.0, rzi = act ? a.repz[i] : 0.0;
    const int zi = act ? a.Z[i] : 0, mi = act ? a.metal[i] : 0, ei = act ? a.elem[i] : 0;
    double hr[NE * 7], dhr[NE * 7];
#pragma unroll
    for (int k = 0; k < NE * 7; ++k) {
        hr[k] = act ? a.h[(size_t)i * NE * 7 + k] : 0.0;
        dhr[k] = act ? a.dh[(size_t)i * NE * 7 + k] : 0.0;
    }
    double own[ADJ_P] = {0.0, 0.0, 0.0};    // contributions through the row atom's element
    double oth[ADJ_P][NE];                  // through the partner's element
#pragma unroll
    for (int p = 0; p < ADJ_P; ++p)
#pragma unroll
        for (int e = 0; e < NE; ++e) oth[p][e] = 0.0;
    for (int t0 = j0; t0 < j1; t0 += TILE) {
        __syncthreads();
        const int j = t0 + threadIdx.x;
        if (j < j1) {
            sx[threadIdx.x] = a.x[3 * j];
            sy[threadIdx.x] = a.x[3 * j + 1];
            sz[threadIdx.x] = a.x[3 * j + 2];
            ssz[threadIdx.x] = a.szr[j];
            san[threadIdx.x] = a.anb[j];
            srz[threadIdx.x] = a.repz[j];
            szt[threadIdx.x] = a.zeta[j];
            sZ[threadIdx.x] = a.Z[j];
            sel[threadIdx.x] = a.elem[j];
            smt[threadIdx.x] = a.metal[j];
#pragma unroll
            for (int r = 0; r < 7; ++r) sgw[threadIdx.x * 7 + r] = a.gw[(size_t)j * 7 + r];
        }
        __syncthreads();
        if (!act) continue;
        const int m = min(TILE, j1 - t0);
        for (int jj = 0; jj < m; ++jj) {
            if (t0 + jj == i || t0 + jj < mlo || t0 + jj >= mhi) continue;
            const double dx = xi.x - sx[jj], dy = xi.y - sy[jj], dz = xi.z - sz[jj];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 2500.0) continue;  // beyond the D4 and repulsion cut-offs
            const double r = sqrt(r2);
            if (r < 1e-10) continue;
            const double sw = (wi.x * dx + wi.y * dy + wi.z * dz) / r;
            const int ej = sel[jj];
            double dd4 = 0.0, dz_ = 0.0, da_ = 0.0;
            {   // D4 (r <= 50)
                double c6 = 0.0, dci = 0.0;
#pragma unroll
                for (int e = 0; e < NE; ++e)
                    if (e == ej) {
#pragma unroll
                        for (int bb = 0; bb < 7; ++bb) {
                            const double gj = sgw[jj * 7 + bb];
                            c6 += hr[e * 7 + bb] * gj;
                            dci += dhr[e * 7 + bb] * gj;
                        }
                    }
                const double q = 3.0 * szi * ssz[jj];
                double dEp_dq, ddcn_dq;
                d4_adj(c6, dci, zti * szt[jj], q, r2, r, dEp_dq, ddcn_dq);
                dd4 = (sw * dEp_dq + ui * ddcn_dq) * q;  // d / d ln(theta) of one factor of q
            }
            if (r <= 20.0 && r >= 1e-8) {   // non-bonded repulsion, generic pair factor
                const double ff = rep_ff(zi, sZ[jj], mi, smt[jj]);
                const double al = sqrt(ani * san[jj]) * ff;
                const double ab = rzi * srz[jj] * 0.4270;
                double dp_lnab, dp_al;
                rep_adj(ab, al, r, dp_lnab, dp_al);
                dz_ = sw * dp_lnab;
                da_ = sw * dp_al * 0.5 * al;
            }
            own[0] += dz_;
            own[1] += da_;
            own[2] += dd4;
#pragma unroll
            for (int e = 0; e < NE; ++e)
                if (e == ej) {
                    oth[0][e] += dz_;
                    oth[1][e] += da_;
                    oth[2][e] += dd4;
                }
        }
    }
    if (!act) return;
    double* pp = part + (size_t)blockIdx.y * ADJ_P * NE * n + i;
#pragma unroll
    for (int p = 0; p < ADJ_P; ++p)
#pragma unroll
        for (int e = 0; e < NE; ++e) pp[(size_t)(p * NE + e) * n] = oth[p][e] + (e == ei ? own[p] : 0.0);
}

// Topological corrections of the non-bonded repulsion (bonded pairs removed,
// H-H 1,3 / 1,4 factors), mirroring nbrep_fix_atom; row-wise, appended as one
// more chunk of partials.
template <int NE>
__global__ __launch_bounds__(256) void k_adj_nbrep_fix(int n, const double* __restrict__ x,
                                                       const int* __restrict__ Z, const int* __restrict__ metal,
                                                       const int* __restrict__ elem, const double* __restrict__ anb,
                                                       const double* __restrict__ repz, const int* __restrict__ toff,
                                                       const int* __restrict__ tidx, const int* __restrict__ tdist,
                                                       const double* __restrict__ w, double* __restrict__ part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double3 xi = ld3(x, i), wi = ld3(w, i);
    double acc[ADJ_P - 1][NE];
#pragma unroll
    for (int p = 0; p < ADJ_P - 1; ++p)
#pragma unroll
        for (int e = 0; e < NE; ++e) acc[p][e] = 0.0;
    const int ei = elem[i];
    for (int k = toff[i]; k < toff[i + 1]; ++k) {
        const int j = tidx[k], td = tdist[k];
        const int zi = Z[i], zj = Z[j];
        const double fgen = rep_ff(zi, zj, metal[i], metal[j]);
        double fnew;
        if (td == 1) fnew = 0.0;
        else if (zi == 1 && zj == 1 && td == 2) fnew = 0.6290 * 1.4580;
        else if (zi == 1 && zj == 1 && td == 3) fnew = 0.6290 * 0.7080;
        else continue;
        const double3 d = xi - ld3(x, j);
        const double r = nrm(d);
        if (r > 20.0 || r < 1e-8) continue;
        const double sw = dot(wi, d) / r;
        const double ab = repz[i] * repz[j] * 0.4270;
        const double base = sqrt(anb[i] * anb[j]);
        double dz_ = 0.0, da_ = 0.0, dp_lnab, dp_al;
        rep_adj(ab, base * fgen, r, dp_lnab, dp_al);
        dz_ -= sw * dp_lnab;
        da_ -= sw * dp_al * 0.5 * base * fgen;
        if (fnew != 0.0) {
            rep_adj(ab, base * fnew, r, dp_lnab, dp_al);
            dz_ += sw * dp_lnab;
            da_ += sw * dp_al * 0.5 * base * fnew;
        }
        const int ej = elem[j];
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const double c = (e == ei ? 1.0 : 0.0) + (e == ej ? 1.0 : 0.0);
            acc[0][e] += c * dz_;
            acc[1][e] += c * da_;
        }
    }
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        part[(size_t)(0 * NE + e) * n + i] = acc[0][e];
        part[(size_t)(1 * NE + e) * n + i] = acc[1][e];
        part[(size_t)(2 * NE + e) * n + i] = 0.0;
    }
}

// Bonded repulsion list: entry b couples atoms i, j with al = sqrt(repa_i repa_j),
// ab = repz_i repz_j 1.7583; its gradient contributes (w_i - w_j) . r_hat dE'/dtheta.
// One row of partials per entry: bpart[k * nb + b].
template <int NE>
__global__ __launch_bounds__(256) void k_adj_brep(int nb, const int* __restrict__ ij, const double* __restrict__ par,
                                                  const int* __restrict__ elem, const double* __restrict__ x,
                                                  const double* __restrict__ w, double* __restrict__ bpart) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int i = ij[2 * b], j = ij[2 * b + 1];
    const double3 d = ld3(x, i) - ld3(x, j);
    const double r = nrm(d);
    double dz_ = 0.0, da_ = 0.0;
    if (!(r > 20.0 || r < 1e-8)) {
        const double al = par[2 * b], ab = par[2 * b + 1];
        const double sw = dot(ld3(w, i) - ld3(w, j), d) / r;
        double dp_lnab, dp_al;
        rep_adj(ab, al, r, dp_lnab, dp_al);
        dz_ = sw * dp_lnab;
        da_ = sw * dp_al * 0.5 * al;
    }
    const int ei = elem[i], ej = elem[j];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const double c = (e == ei ? 1.0 : 0.0) + (e == ej ? 1.0 : 0.0);
        bpart[(size_t)(0 * NE + e) * nb + b] = c * dz_;
        bpart[(size_t)(1 * NE + e) * nb + b] = c * da_;
        bpart[(size_t)(2 * NE + e) * nb + b] = 0.0;
    }
}

// Fixed-order reduction of the adjoint partials: one wavefront per output
// component k; lanes stride the rows, then a fixed butterfly.
__global__ __launch_bounds__(64) void k_adj_reduce(int n, int nchunk, const double* __restrict__ part, int nb,
                                                   const double* __restrict__ bpart, int nk,
                                                   double* __restrict__ out) {
    const int k = blockIdx.x, lane = threadIdx.x;
    double acc = 0.0;
    for (int c = 0; c < nchunk; ++c)
        for (int i = lane; i < n; i += 64) acc += part[((size_t)c * nk + k) * n + i];
    for (int b = lane; b < nb; b += 64) acc += bpart[(size_t)k * nb + b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) out[k] = acc;
}

// ---------------------------------------------------------------------------
// K9: gradient assembly: slot gather, bond dE/dCN gather, CN chain rule with
// Coulomb term 1b (ff_workspace.cpp:459-507), fixed order per atom.
// ---------------------------------------------------------------------------
__device__ inline void dcn_total_atom(int i, const int* __restrict__ doff, const int* __restrict__ didx,
                                      const double* __restrict__ dcn_slot, double* __restrict__ dEdcn,
                                      const double* __restrict__ qtmp, double* __restrict__ v) {
    double s = dEdcn[i];
    for (int k = doff[i]; k < doff[i + 1]; ++k) s += dcn_slot[didx[k]];
    dEdcn[i] = s;
    v[i] = s - qtmp[i];
}

// Per-atom tail after the pair terms, one thread per atom: the repulsion
// corrections of the atom's topological neighbourhood (its gradient and
// e_rep) and its dE/dCN total with the chain vector, which touch disjoint data.
struct AtomTailArgs {
    const double* x;
    const int *Z, *metal;
    const double *anb, *repz;
    const int *toff, *tidx, *tdist;
    double *g, *e_rep;
    const int *doff, *didx;
    const double* dcn_slot;
    double* dEdcn;
    const double* qtmp;
    double* v;
};
__global__ __launch_bounds__(256) void k_atom_tail(int n, AtomTailArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    nbrep_fix_atom(i, a.x, a.Z, a.metal, a.anb, a.repz, a.toff, a.tidx, a.tdist, a.g, a.e_rep);
    dcn_total_atom(i, a.doff, a.didx, a.dcn_slot, a.dEdcn, a.qtmp, a.v);
}

// LPA lanes per atom (64: one wavefront; 16 for short lists, four atoms per
// wavefront): lanes stride the atom's slot list and its CN neighbour list
// (chain rule), fixed butterfly over the atom's lanes
template <int W>
__device__ inline double group_sum(double v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
// Per-atom sum of the bonded-list gradient slots (every slot writer runs
// before the EEQ solve, so this gather runs there too, off the solve's path).
template <int LPA>
__global__ __launch_bounds__(256) void k_slot_gather(int n, const int* __restrict__ soff, const int* __restrict__ sidx,
                                                     const double* __restrict__ slots, double* __restrict__ gslot) {
    const int lane = threadIdx.x & (LPA - 1);
    const int i = blockIdx.x * (256 / LPA) + threadIdx.x / LPA;
    if (i >= n) return;
    double3 g = make_double3(0, 0, 0);
    for (int k = soff[i] + lane; k < soff[i + 1]; k += LPA) g = g + ld3(slots, sidx[k]);
    g.x = group_sum<LPA>(g.x);
    g.y = group_sum<LPA>(g.y);
    g.z = group_sum<LPA>(g.z);
    if (lane == 0) st3(gslot, i, g);
}

// CN chain rule over the atom's neighbour list plus the gathered slot, pair
// and hydrogen-bond gradients
template <int LPA>
__global__ __launch_bounds__(256) void k_finish(int n, const double* __restrict__ x, const double* __restrict__ rcov,
                                                const double* __restrict__ dlog, const double* __restrict__ v,
                                                const int* __restrict__ nbn, const int* __restrict__ nbl,
                                                const double* __restrict__ gslot, const double* __restrict__ gpair,
                                                const double* __restrict__ ghb, double* __restrict__ grad) {
    const int lane = threadIdx.x & (LPA - 1);
    const int i = blockIdx.x * (256 / LPA) + threadIdx.x / LPA;
    if (i >= n) return;
    double3 g = make_double3(0, 0, 0);
    const double3 xi = ld3(x, i);
    const double ai = dlog[i] * v[i];
    const int nn = nbn[i];
    for (int k = lane; k < nn; k += LPA) {
        int j = nbl[(size_t)i * MAXNB_CN + k];
        double3 d = ld3(x, j) - xi;
        double r = nrm(d);
        double rs = rcov[i] + rcov[j];
        double dr = (r - rs) / rs;
        double dd = (KN / SQRTPI_TRUNC) * exp(-KN * KN * dr * dr) / rs;
        double s = dd * (ai + dlog[j] * v[j]);
        g = g + (-s / r) * d;
    }
    g.x = group_sum<LPA>(g.x);
    g.y = group_sum<LPA>(g.y);
    g.z = group_sum<LPA>(g.z);
    if (lane == 0) {
        g = g + ld3(gslot, i) + ld3(gpair, i) + ld3(ghb, i);
        st3(grad, i, g);
    }
}

// fixed-shape reduction of up to 16 segments into out[seg]; one workgroup
// per segment, strided partial sums then a tree in LDS
struct Segs {
    const double* p[16];
    int n[16];
    int dst[16];
};
constexpr int RED_BLOCKS = 64;
constexpr int NPARTS = 16;  // energy segments reduced per evaluation
// stage 1: segment sg, block b reduces the fixed chunk b of the segment
__global__ __launch_bounds__(256) void k_reduce1(Segs s, double* __restrict__ part) {
    __shared__ double red[256];
    int sg = blockIdx.y, b = blockIdx.x;
    const double* p = s.p[sg];
    int n = s.n[sg];
    int chunk = (n + RED_BLOCKS - 1) / RED_BLOCKS;
    int lo = b * chunk, hi = min(n, lo + chunk);
    // eight independent partial sums per thread keep eight loads in flight
    double a8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = lo + threadIdx.x; k < hi; k += 8 * 256) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double v = p[min(k + u * 256, hi - 1)];  // unconditional load, masked add
            a8[u] += (k + u * 256 < hi) ? v : 0.0;
        }
    }
    double acc = ((a8[0] + a8[1]) + (a8[2] + a8[3])) + ((a8[4] + a8[5]) + (a8[6] + a8[7]));
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[sg * RED_BLOCKS + b] = red[0];
}
// stage 2 and the term table, one workgroup: wavefront sg adds segment sg's
// block partials with a fixed tree; thread 0 then forms out[0] = total and
// out[1..15] = terms (term index + 1) in a fixed summation order
__global__ __launch_bounds__(NPARTS * 64) void k_reduce_final(Segs s, const double* __restrict__ part,
                                                               double* __restrict__ parts, double* __restrict__ out,
                                                               const int* __restrict__ flags, int* __restrict__ sticky) {
    __shared__ double sp[NPARTS];
    const int sg = threadIdx.x >> 6, l = threadIdx.x & 63;
    double v = part[sg * RED_BLOCKS + l];
    for (int w = 32; w > 0; w >>= 1) v += __shfl_xor(v, w, 64);
    if (l == 0) {
        sp[s.dst[sg]] = v;
        parts[s.dst[sg]] = v;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double t[GFNFF_NTERMS];
    for (int k = 0; k < GFNFF_NTERMS; ++k) t[k] = 0.0;
    // parts: 0 bond,1 angle,2 tors,3 inv,4 disp(pairs),5 brep,6 nbrep(pairs incl fix),7 coul pairs,
    // 8 self,9 hb case 1,10 atm,11 batm,12 triple-bond torsions,13 xb,14 hb cases 2-4,15 solvation
    t[GFNFF_E_BOND] = sp[0];
    t[GFNFF_E_ANGLE] = sp[1];
    t[GFNFF_E_TORSION] = sp[2];
    t[GFNFF_E_INVERSION] = sp[3];
    t[GFNFF_E_DISPERSION] = sp[4];
    t[GFNFF_E_BONDED_REP] = sp[5];
    t[GFNFF_E_NONBONDED_REP] = sp[6];
    t[GFNFF_E_COULOMB] = sp[7] + sp[8];
    t[GFNFF_E_HBOND] = sp[9] + sp[14];
    t[GFNFF_E_ATM] = sp[10];
    t[GFNFF_E_BATM] = sp[11];
    t[GFNFF_E_STORS] = sp[12];
    t[GFNFF_E_XBOND] = sp[13];
    t[GFNFF_E_SOLVATION] = sp[15];
    double tot = 0.0;
    for (int k = 0; k < GFNFF_NTERMS; ++k) tot += t[k];
    // a status flag raised in this evaluation (CN list overflow, EEQ pivot,
    // solver timeout) makes the total NaN and is kept in the sticky word
    // until the host reads and clears it
    const int f = __hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (f) {
        tot = __builtin_nan("");
        atomicOr(sticky, f);
    }
    out[0] = tot;
    for (int k = 0; k < GFNFF_NTERMS; ++k) out[1 + k] = t[k];
}

__global__ void k_copy(int n, const double* __restrict__ a, double* __restrict__ b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}
// coordinates of one evaluation into the engine's buffer; thread 0 also
// clears the status flags (one launch at the head of the evaluation)
__global__ void k_copy_start(int n, const double* __restrict__ a, double* __restrict__ b, int* __restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) flags[0] = 0;
    if (i < n) b[i] = a[i];
}
__global__ void k_fill_i(int n, int* __restrict__ a, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_fill_u64(int n, unsigned long long* __restrict__ a, unsigned long long v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
// K12: device-resident molecular dynamics.  Velocity Verlet as in
// SimpleMD::Verlet (simplemd.cpp:2580-2673) with the optional Berendsen bath
// of SimpleMD::Berendson (:3967-3975), in atomic units (Bohr, electron
// masses, atomic time units; gradient in Eh/Bohr).  One step:
//   k_md_drift   x += dt v - dt^2/2 g/m,  v -= dt/2 g/m   (g of the old x)
//   evaluation   g, Epot at the new x
//   k_md_kick    v -= dt/2 g/m; per-block partial sums of m v^2 and of the
//                displacement |x - x_ref|^2 since the last HB/XB detection
//   k_md_finish  Ekin, T = 2 Ekin / (kB dof), Berendsen factor, the step's
//                record (Epot, Ekin, T, Epot + Ekin), and GFNFF flag 8 when
//                sqrt(sum dx^2) / N exceeds the re-detection threshold
//                (gfnff_method.cpp:2243-2251)
//   k_md_scale   v *= lambda (bath only)
// Sums are formed in a fixed order; nothing is copied to the host.
// ---------------------------------------------------------------------------
constexpr int FLAG_HB_STALE = 8;
constexpr int MD_T = 256;

__global__ __launch_bounds__(MD_T) void k_md_drift(int n3, double dt, const double* __restrict__ im3,
                                                   const double* __restrict__ g, double* __restrict__ x,
                                                   double* __restrict__ v) {
    const int k = blockIdx.x * MD_T + threadIdx.x;
    if (k >= n3) return;
    const double a = g[k] * im3[k];
    x[k] = x[k] + dt * v[k] - 0.5 * a * dt * dt;
    v[k] = v[k] - 0.5 * dt * a;
}

__global__ __launch_bounds__(MD_T) void k_md_kick(int n3, double dt, const double* __restrict__ im3,
                                                  const double* __restrict__ m3, const double* __restrict__ g,
                                                  const double* __restrict__ x, const double* __restrict__ xref,
                                                  double* __restrict__ v, double* __restrict__ part) {
    __shared__ double sk[MD_T / 64], sd[MD_T / 64];
    const int k = blockIdx.x * MD_T + threadIdx.x;
    double ek = 0.0, dd = 0.0;
    if (k < n3) {
        const double vk = v[k] - 0.5 * dt * g[k] * im3[k];
        v[k] = vk;
        ek = m3[k] * vk * vk;
        const double dx = x[k] - xref[k];
        dd = dx * dx;
    }
    ek = wave_sum(ek);
    dd = wave_sum(dd);
    if ((threadIdx.x & 63) == 0) {
        sk[threadIdx.x >> 6] = ek;
        sd[threadIdx.x >> 6] = dd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (sk[0] + sk[1]) + (sk[2] + sk[3]);
        part[2 * blockIdx.x + 1] = (sd[0] + sd[1]) + (sd[2] + sd[3]);
    }
}

__global__ __launch_bounds__(64) void k_md_finish(int nb, const double* __restrict__ part,
                                                  const double* __restrict__ epot, double kb_dof, double dt, double T0,
                                                  double tau, double disp2_max, double* __restrict__ lambda,
                                                  double* __restrict__ rec, int* __restrict__ sticky) {
    const int lane = threadIdx.x;
    double ek = 0.0, dd = 0.0;
    for (int b = lane; b < nb; b += 64) {
        ek += part[2 * b];
        dd += part[2 * b + 1];
    }
    ek = 0.5 * wave_sum(ek);
    dd = wave_sum(dd);
    if (lane != 0) return;
    const double T = 2.0 * ek / kb_dof;
    double lam = 1.0;
    if (T0 > 0.0 && T > 0.0) lam = sqrt(1.0 + (dt / 2.0 * (T0 - T)) / (T * tau));
    *lambda = lam;
    rec[0] = epot[0];
    rec[1] = ek;
    rec[2] = T;
    rec[3] = epot[0] + ek;
    if (dd > disp2_max) atomicOr(sticky, FLAG_HB_STALE);
}

__global__ __launch_bounds__(MD_T) void k_md_scale(int n3, const double* __restrict__ lambda, double* __restrict__ v) {
    const int k = blockIdx.x * MD_T + threadIdx.x;
    if (k < n3) v[k] *= *lambda;
}

// ---------------------------------------------------------------------------
// K13: ALPB implicit solvation of one molecule, FP64 on both precision
// paths.  Born radii from the OBC-II transform of the pairwise descreening
// sum (alpb_solvation.cpp:502-690), surface areas on the 230-point angular
// grid with the smooth neighbour switching (:696-801), the H-bond diagonal
// (:873-881) and the shape descriptor (:908-938).  The Born matrix
// (P16 kernel, :810-867) is never stored: its entries are formed where they
// are used, added to the reduced EEQ matrix and right-hand side (the
// reference's A + B, eeq_solver.cpp:3774-3784), and contracted with the
// solved charges for the energy and the gradient at fixed charges
// (:1016-1041, :1139-1201, :1265-1300).  Every sum runs in a fixed order.
// ---------------------------------------------------------------------------
constexpr int SOLV_NB = 512;     // surface neighbours per atom (LDS list)
constexpr int SOLV_NANG = 230;   // angular grid points per atom
constexpr int FLAG_SOLV_NB = 16; // more than SOLV_NB surface neighbours
constexpr double ALPB_W = 0.3 / 0.52917726;
constexpr double ALPB_AH0 = 0.5;
constexpr double ALPB_AH1 = 3.0 / (4.0 * ALPB_W);
constexpr double ALPB_AH3 = -1.0 / (4.0 * ALPB_W * ALPB_W * ALPB_W);
constexpr double ALPB_ZETA = 1.028;
constexpr double ALPB_TOLSESP = 1.0e-6;

struct SolvArgs {
    int n;
    double keps, alpbet, c1, gshift, lr2, sr2;
    const double* x;
    const double *vdwr, *rho, *svdw, *vdwsa, *wrp, *trjlo, *trjhi, *gam, *hbmag, *grid;
    double *brad, *dbr, *sasa, *hbw;  // Born radius, d radius / d psi, surface area, H-bond diagonal
    double* sv;                       // [n][SOLV_NB][3] d sasa_i / d x_i per neighbour slot
    int *sj, *sn;                     // [n][SOLV_NB] neighbour of each slot (ascending), [n] count
    double* shape;                    // adet, totrad3, centre[3], inertia[9], gradient scale, cofactors[9]
    double *w, *c, *eat, *gs;         // radius weight, surface weight, atom energy, direct gradient [3n]
    double *e_solv, *parts;           // total, {gborn, ghb, gsasa, gshift}
    int* flags;
};

// One descreening contribution to psi of atom s from atom o at distance r
// (rho_o: descreening radius of o, rv: van der Waals radius of s), the
// non-overlapping or overlapping integral; dg = (d g / d r) / r.
__device__ inline double alpb_psi(double r, double rho_o, double rv, double& dg) {
    if (r >= rv + rho_o) {
        const double r1 = 1.0 / r, ap = r + rho_o, am = r - rho_o, ab = ap * am, rhab = rho_o / ab;
        const double lnab = 0.5 * log(am / ap) * r1;
        dg = -2.0 * rhab / ab + (rhab - lnab) * r1 * r1;
        return rhab + lnab;
    }
    if (r + rho_o <= rv) {
        dg = 0.0;
        return 0.0;
    }
    const double r1 = 1.0 / r, r12 = 0.5 * r1, r24 = r12 * r12, ap = r + rho_o, am = r - rho_o;
    const double rh1 = 1.0 / rv, rhr1 = 1.0 / ap, aprh1 = ap * rh1, lnab = log(aprh1);
    dg = (rhr1 * rhr1 * (1.0 - 0.25 * am * r1 * (1.0 + aprh1 * aprh1)) + rho_o * r24 * (rhr1 - rh1 * aprh1) +
          r12 * (r1 * lnab - rhr1)) *
         r1;
    return rh1 - rhr1 + r12 * (0.5 * am * (rhr1 - rh1 * aprh1) - lnab);
}

// P16 interaction a16 = (ab / (ab + zeta r / 16))^16; returns a1 = the base
__device__ inline double p16_pow(double ab, double r, double& a16) {
    const double a1 = ab / (ab + (ALPB_ZETA / 16.0) * r);
    double t = a1 * a1;
    t = t * t;
    t = t * t;
    a16 = t * t;
    return a1;
}

// Born matrix entry without the shape constant (which cancels in the
// reduced EEQ system: every reduced row and column sums entries with total
// weight zero)
__device__ inline double born_b(const SolvArgs& a, int i, int j) {
    if (i == j) return a.keps / a.brad[i] + 2.0 * a.hbw[i];
    const double3 d = ld3(a.x, i) - ld3(a.x, j);
    const double r = sqrt(dot(d, d)), ab = sqrt(a.brad[i] * a.brad[j]);
    double a16;
    (void)p16_pow(ab, r, a16);
    return a.keps / (r + ab * a16);
}

// Born radii: one wavefront per atom over all partners within lrcut
__global__ __launch_bounds__(256) void k_solv_born(SolvArgs a) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
    if (i >= a.n) return;  // wavefront-uniform
    const double3 xi = ld3(a.x, i);
    const double rvi = a.vdwr[i];
    double psi = 0.0;
    for (int j = l; j < a.n; j += 64) {
        if (j == i) continue;
        const double3 d = xi - ld3(a.x, j);
        const double r2 = dot(d, d);
        if (r2 >= a.lr2) continue;
        double dg;
        psi += alpb_psi(sqrt(r2), a.rho[j], rvi, dg);
    }
    psi = wave_sum(psi);
    if (l != 0) return;
    // OBC-II: R = c1 / (1/svdw - tanh(arg) / vdwr), arg a cubic in psi svdw / 2
    const double s1 = 1.0 / a.svdw[i], v1 = 1.0 / rvi, s2 = 0.5 * a.svdw[i];
    const double br = psi * s2;
    double arg2 = br * (4.85 * br - 0.8);
    const double arg = br * (1.0 + arg2);
    arg2 = 2.0 * arg2 + 1.0 + 4.85 * br * br;
    const double th = tanh(arg), ch = cosh(arg);
    a.brad[i] = a.c1 / (s1 - v1 * th);
    const double dp = ch * (s1 - v1 * th);
    a.dbr[i] = a.c1 * s2 * v1 * arg2 / (dp * dp);
}

// Surface area of atom i (one workgroup): the neighbours within srcut are
// compacted into LDS in ascending order; one thread per grid point forms
// the switched product; one thread per neighbour then sums its derivative
// over the grid points (no atomics).
__global__ __launch_bounds__(256) void k_solv_sasa(SolvArgs a) {
    __shared__ double3 nx[SOLV_NB];
    __shared__ double nlo[SOLV_NB], nhi[SOLV_NB], nsa[SOLV_NB];
    __shared__ int nj[SOLV_NB];
    __shared__ double3 px[256];
    __shared__ double pw[256];
    __shared__ int wcnt[4];
    __shared__ double red[4];
    const int i = blockIdx.x, tid = threadIdx.x, l = tid & 63, wv = tid >> 6;
    const double3 xi = ld3(a.x, i);
    int cnt = 0;
    for (int j0 = 0; j0 < a.n; j0 += 256) {
        const int j = j0 + tid;
        bool in = false;
        if (j < a.n && j != i) {
            const double3 d = xi - ld3(a.x, j);
            const double r2 = dot(d, d);
            in = r2 < a.lr2 && r2 < a.sr2;
        }
        const unsigned long long b = __ballot(in);
        const int pre = __popcll(b & ((1ull << l) - 1ull));
        if (l == 0) wcnt[wv] = __popcll(b);
        __syncthreads();
        int off = cnt;
        for (int w = 0; w < wv; ++w) off += wcnt[w];
        const int tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        if (in && off + pre < SOLV_NB) {
            const int s = off + pre;
            nx[s] = ld3(a.x, j);
            nlo[s] = a.trjlo[j];
            nhi[s] = a.trjhi[j];
            nsa[s] = a.vdwsa[j];
            nj[s] = j;
        }
        cnt += tot;
        __syncthreads();
    }
    if (cnt > SOLV_NB) {
        if (tid == 0) atomicOr(a.flags, FLAG_SOLV_NB);
        cnt = SOLV_NB;
    }
    const double rsa = a.vdwsa[i];
    double wsa = 0.0;
    double3 xp = xi;
    if (tid < SOLV_NANG) {
        const double* g = a.grid + 4 * tid;
        xp = make_double3(xi.x + rsa * g[0], xi.y + rsa * g[1], xi.z + rsa * g[2]);
        double sp = 1.0;
        for (int s = 0; s < cnt; ++s) {
            const double3 t = xp - nx[s];
            const double t2 = dot(t, t);
            if (t2 < nhi[s]) {
                if (t2 <= nlo[s]) {
                    sp = 0.0;
                    break;
                }
                const double u = sqrt(t2) - nsa[s], ah3u2 = ALPB_AH3 * u * u;
                sp *= ALPB_AH0 + (ALPB_AH1 + ah3u2) * u;
            }
        }
        if (sp > ALPB_TOLSESP) wsa = g[3] * a.wrp[i] * sp;
    }
    px[tid] = xp;
    pw[tid] = wsa;
    const double sasa = wg_sum256(wsa, red);  // its barriers also publish px and pw
    for (int s = tid; s < cnt; s += 256) {
        const double3 xj = nx[s];
        const double hi = nhi[s], sa = nsa[s];
        double vx = 0.0, vy = 0.0, vz = 0.0;
        for (int p = 0; p < SOLV_NANG; ++p) {
            const double w = pw[p];
            if (w == 0.0) continue;
            const double3 t = px[p] - xj;
            const double t2 = dot(t, t);
            if (t2 < hi) {
                const double sq = sqrt(t2), u = sq - sa, ah3u2 = ALPB_AH3 * u * u;
                const double f = ALPB_AH0 + (ALPB_AH1 + ah3u2) * u;
                const double df = w * (ALPB_AH1 + 3.0 * ah3u2) / (f * sq);
                vx += df * t.x;
                vy += df * t.y;
                vz += df * t.z;
            }
        }
        const size_t o = (size_t)i * SOLV_NB + s;
        a.sv[3 * o] = vx;
        a.sv[3 * o + 1] = vy;
        a.sv[3 * o + 2] = vz;
        a.sj[o] = nj[s];
    }
    if (tid == 0) {
        a.sasa[i] = sasa;
        a.hbw[i] = a.hbmag[i] * sasa / (rsa * rsa);
        a.sn[i] = cnt;
    }
}

// Shape descriptor (one workgroup): the inertia tensor of spheres with the
// van der Waals radii about their volume-weighted centre, and
// adet = sqrt(det^(1/3) / (2/5 sum r^3)).
__global__ __launch_bounds__(256) void k_solv_shape(SolvArgs a) {
    __shared__ double red[4];
    const int tid = threadIdx.x;
    double tr = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int i = tid; i < a.n; i += 256) {
        const double r3 = a.vdwr[i] * a.vdwr[i] * a.vdwr[i];
        tr += r3;
        cx += a.x[3 * i] * r3;
        cy += a.x[3 * i + 1] * r3;
        cz += a.x[3 * i + 2] * r3;
    }
    tr = wg_sum256(tr, red);
    cx = wg_sum256(cx, red) / tr;
    cy = wg_sum256(cy, red) / tr;
    cz = wg_sum256(cz, red) / tr;
    double I[6] = {0, 0, 0, 0, 0, 0};  // xx, yy, zz, xy, xz, yz
    for (int i = tid; i < a.n; i += 256) {
        const double r2v = a.vdwr[i] * a.vdwr[i], r3 = r2v * a.vdwr[i];
        const double vx = a.x[3 * i] - cx, vy = a.x[3 * i + 1] - cy, vz = a.x[3 * i + 2] - cz;
        const double d = vx * vx + vy * vy + vz * vz + 0.4 * r2v;
        I[0] += r3 * (d - vx * vx);
        I[1] += r3 * (d - vy * vy);
        I[2] += r3 * (d - vz * vz);
        I[3] -= r3 * vx * vy;
        I[4] -= r3 * vx * vz;
        I[5] -= r3 * vy * vz;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) I[k] = wg_sum256(I[k], red);
    if (tid != 0) return;
    const double xx = I[0], yy = I[1], zz = I[2], xy = I[3], xz = I[4], yz = I[5];
    const double det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    double* s = a.shape;
    s[0] = sqrt(cbrt(det) / (0.4 * tr));
    s[1] = tr;
    s[2] = cx;
    s[3] = cy;
    s[4] = cz;
    const double full[9] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
    for (int k = 0; k < 9; ++k) s[5 + k] = full[k];
}

// Reduced Born matrix added to the reduced EEQ matrix, in the storage of
// k_eeq_reduced (full == 0: lower triangle of the row-major matrix, 1: also
// the upper triangle, 2: packed tiles with full diagonal tiles)
__global__ __launch_bounds__(256) void k_solv_red_add(int nr, int np, const int* __restrict__ ra,
                                                      const int* __restrict__ rl, SolvArgs a, double* __restrict__ A,
                                                      int full) {
    const int t = blockIdx.x;
    int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    while (bi * (bi + 1) / 2 > t) --bi;
    const int bj = t - bi * (bi + 1) / 2;
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        const int rr = e >> 6, cc = e & 63, r = bi * NB + rr, c = bj * NB + cc;
        if (c > r || r >= nr || c >= nr) continue;
        const int ia = ra[r], la = rl[r], ib = ra[c], lb = rl[c];
        const double v = born_b(a, ia, ib) - born_b(a, ia, lb) - born_b(a, la, ib) + born_b(a, la, lb);
        if (full == 2) {
            A[(size_t)t * NB * NB + rr * NB + cc] += v;
            if (bi == bj && cc != rr) A[(size_t)t * NB * NB + cc * NB + rr] += v;
        } else {
            A[(size_t)r * np + c] += v;
            if (full == 1 && c != r) A[(size_t)c * np + r] += v;
        }
    }
}

// Charged fragments: the Born part of A qp leaves the right-hand side
__global__ __launch_bounds__(256) void k_solv_rhs(SolvArgs a, const double* __restrict__ qp, double* __restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double s = 0.0;
    for (int j = 0; j < a.n; ++j)
        if (qp[j] != 0.0) s += born_b(a, i, j) * qp[j];
    r[i] -= s;
}

// Born pair terms at the solved charges (one wavefront per atom i over all
// partners): atom energy, direct gradient, and d E / d R_i folded with
// d R_i / d psi_i; the surface weight gamma_i + d(H-bond)/d sasa_i.
__global__ __launch_bounds__(256) void k_solv_pairs(SolvArgs a, const double* __restrict__ q) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
    if (i >= a.n) return;
    const double3 xi = ld3(a.x, i);
    const double Ri = a.brad[i], qi = q[i];
    double e = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, dr = 0.0;
    for (int j = l; j < a.n; j += 64) {
        if (j == i) continue;
        const double3 d = xi - ld3(a.x, j);
        const double r = sqrt(dot(d, d)), Rj = a.brad[j], ab = sqrt(Ri * Rj), qj = q[j];
        double a16;
        const double a1 = p16_pow(ab, r, a16);
        const double df = 1.0 / (r + ab * a16), df2 = df * df;
        e += qj * df;
        const double ap = (1.0 - ALPB_ZETA * a1 * a16) * df2 * qj / r;
        gx -= ap * d.x;
        gy -= ap * d.y;
        gz -= ap * d.z;
        dr += Rj * (-0.5 * (r * ALPB_ZETA / ab * a1 + 1.0) / ab * a16 * df2) * qj;
    }
    e = wave_sum(e);
    gx = wave_sum(gx);
    gy = wave_sum(gy);
    gz = wave_sum(gz);
    dr = wave_sum(dr);
    if (l != 0) return;
    const double kq = a.keps * qi;
    a.eat[i] = 0.5 * kq * e + 0.5 * qi * qi * (a.keps / Ri + 2.0 * a.hbw[i]);
    a.gs[3 * i] = kq * gx;
    a.gs[3 * i + 1] = kq * gy;
    a.gs[3 * i + 2] = kq * gz;
    a.w[i] = (kq * dr - 0.5 * kq * qi / (Ri * Ri)) * a.dbr[i];
    a.c[i] = a.gam[i] + a.hbmag[i] / (a.vdwsa[i] * a.vdwsa[i]) * qi * qi;
}

// Solvation energy (one workgroup, fixed order) and the shape gradient's scale
__global__ __launch_bounds__(256) void k_solv_energy(SolvArgs a, const double* __restrict__ q) {
    __shared__ double red[4];
    double eb = 0.0, ehb = 0.0, es = 0.0, qt = 0.0;
    for (int i = threadIdx.x; i < a.n; i += 256) {
        eb += a.eat[i];
        ehb += a.hbw[i] * q[i] * q[i];
        es += a.gam[i] * a.sasa[i];
        qt += q[i];
    }
    eb = wg_sum256(eb, red);
    ehb = wg_sum256(ehb, red);
    es = wg_sum256(es, red);
    qt = wg_sum256(qt, red);
    if (threadIdx.x != 0) return;
    double* s = a.shape;
    const double ad = s[0], tr = s[1];
    double esh = 0.0;
    s[14] = 0.0;
    if (a.alpbet > 0.0) {
        esh = 0.5 * a.keps * a.alpbet * qt * qt / ad;
        s[14] = (250.0 / (48.0 * tr * tr * tr * pow(ad, 5.0))) * (-0.5 * a.keps * a.alpbet * qt * qt / (ad * ad));
    }
    // cofactor-type matrix of the inertia tensor (gbsa.f90:769-779)
    const double* I = s + 5;
    double* D = s + 15;
    D[0] = I[0] * (I[4] + I[8]) - I[1] * I[1] - I[2] * I[2];
    D[1] = I[1] * I[8] - I[2] * I[5];
    D[2] = I[2] * I[4] - I[1] * I[7];
    D[3] = D[1];
    D[4] = I[4] * (I[0] + I[8]) - I[1] * I[1] - I[5] * I[5];
    D[5] = I[0] * I[5] - I[1] * I[2];
    D[6] = D[2];
    D[7] = D[5];
    D[8] = I[8] * (I[0] + I[4]) - I[2] * I[2] - I[5] * I[5];
    a.parts[0] = eb - ehb + esh;
    a.parts[1] = ehb;
    a.parts[2] = es;
    a.parts[3] = a.gshift;
    a.e_solv[0] = ((eb + esh) + es) + a.gshift;
}

// Solvation gradient of atom k (one wavefront): the Born-radius chain rule
// over all partners within lrcut, the surface terms of k's own sphere and of
// every neighbour's sphere (k's slot in the neighbour's ascending list by
// bisection), the direct pair gradient and the shape term; added to grad.
__global__ __launch_bounds__(256) void k_solv_grad(SolvArgs a, double* __restrict__ grad) {
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
    if (k >= a.n) return;
    const double3 xk = ld3(a.x, k);
    const double wk = a.w[k], rvk = a.vdwr[k], rhok = a.rho[k], ck = a.c[k];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int j = l; j < a.n; j += 64) {
        if (j == k) continue;
        const double3 d = xk - ld3(a.x, j);
        const double r2 = dot(d, d);
        if (r2 >= a.lr2) continue;
        const double r = sqrt(r2);
        double d1, d2;
        (void)alpb_psi(r, a.rho[j], rvk, d1);
        (void)alpb_psi(r, rhok, a.vdwr[j], d2);
        const double f = wk * d1 + a.w[j] * d2;
        gx += f * d.x;
        gy += f * d.y;
        gz += f * d.z;
    }
    const int nk = a.sn[k];
    for (int s = l; s < nk; s += 64) {
        const size_t o = (size_t)k * SOLV_NB + s;
        gx += ck * a.sv[3 * o];
        gy += ck * a.sv[3 * o + 1];
        gz += ck * a.sv[3 * o + 2];
        const int j = a.sj[o];
        const int* lj = a.sj + (size_t)j * SOLV_NB;
        int lo = 0, hi = a.sn[j];
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (lj[mid] < k)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo < a.sn[j] && lj[lo] == k) {
            const size_t oj = (size_t)j * SOLV_NB + lo;
            const double cj = a.c[j];
            gx -= cj * a.sv[3 * oj];
            gy -= cj * a.sv[3 * oj + 1];
            gz -= cj * a.sv[3 * oj + 2];
        }
    }
    gx = wave_sum(gx);
    gy = wave_sum(gy);
    gz = wave_sum(gz);
    if (l != 0) return;
    const double* s = a.shape;
    const double r3 = rvk * rvk * rvk * s[14];
    const double vx = xk.x - s[2], vy = xk.y - s[3], vz = xk.z - s[4];
    const double* D = s + 15;
    grad[3 * k] += (gx + a.gs[3 * k]) + r3 * (D[0] * vx + D[1] * vy + D[2] * vz);
    grad[3 * k + 1] += (gy + a.gs[3 * k + 1]) + r3 * (D[3] * vx + D[4] * vy + D[5] * vz);
    grad[3 * k + 2] += (gz + a.gs[3 * k + 2]) + r3 * (D[6] * vx + D[7] * vy + D[8] * vz);
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Kernel classes that can be timed with HIP events (one class at a time).
enum KernelClass {
    KC_CN, KC_EEQ_BUILD, KC_CHOL_DIAG, KC_CHOL_PANEL, KC_CHOL_UPDATE, KC_TRSV, KC_CHARGES, KC_PAIRS, KC_PAIRS_AUX,
    KC_NBREP, KC_REFINE, KC_BONDS, KC_ANGLES, KC_TORS, KC_BREP, KC_HB, KC_ATM, KC_BATM, KC_CHAIN, KC_REDUCE, KC_PCG,
    KC_BATCH_SOLVE, KC_SOLV, KC_COUNT
};
const char* const kKernelClassNames[KC_COUNT] = {
    "cn",    "eeq_build", "chol_diag", "chol_panel", "chol_update", "trsv", "charges", "pairs", "pairs_aux", "nbrep", "refine",
    "bonds", "angles",    "torsions",  "brep",       "hb",          "atm",  "batm",    "chain", "reduce", "pcg",
    "batch_solve", "solvation"};

// Start/stop event pairs around every launch of the selected class, recorded
// on the stream the kernels run on.
struct KernelTimer {
    int cls = -1;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    void mark(int c, hipStream_t s) {
        if (c != cls) return;
        if (used == ev.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) {
                cls = -1;  // timing silently off when no event can be created; read() reports it
                return;
            }
            ev.push_back(e);
        }
        (void)hipEventRecord(ev[used++], s);
    }
    ~KernelTimer() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

// Side stream and fork/join events for the factorisation lookahead.
struct Lookahead {
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

#define LAUNCH(cls, kern, grid, block, stream, ...)                           \
    do {                                                                      \
        E->timer.mark(cls, stream);                                           \
        hipLaunchKernelGGL(kern, grid, block, 0, stream, __VA_ARGS__);        \
        E->timer.mark(cls, stream);                                           \
    } while (0)
#define LAUNCH_SHM(cls, kern, grid, block, shm, stream, ...)                  \
    do {                                                                      \
        E->timer.mark(cls, stream);                                           \
        hipLaunchKernelGGL(kern, grid, block, shm, stream, __VA_ARGS__);      \
        E->timer.mark(cls, stream);                                           \
    } while (0)

struct DeviceEngine {
    KernelTimer timer;
    int n = 0, nelem = 0, nfrag = 0;
    int nr = 0, np = 0;                 // reduced EEQ size, padded size
    int nbond = 0, nangle = 0, ntors = 0, nbrep = 0, nhb = 0, natm = 0, nbatm = 0, ninv = 0, nstors = 0, nxb = 0;
    int nslots = 0;
    int pair_cj = 0, pair_nchunk = 0;   // j-chunk size and count of the pair kernel
    int ps_nb = 0;                      // 64-atom blocks of the symmetric pair kernel (single molecules)
    int ps_ntiles = 0, ps_nblocks = 0;  // its tiles and 64-atom blocks (0 tiles: batches on the row kernel)
    const int4* ps_tiles = nullptr;     // batches: tiles {a0, n, R, C} of every molecule
    const int4* ps_blocks = nullptr;    // batches: blocks {a0, n, A, first tile of the molecule}
    double* tbox = nullptr;             // bounding boxes of the 256-atom tiles (CN partner pruning)
    hipStream_t side = nullptr;         // builds the reduced EEQ matrix alongside the evaluation stream
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_fact = nullptr;       // batches: the factorisations on the side stream are done
    // batch_streams 2: the second half of the batch's factorisation on a third stream, forked from
    // and joined back to the side stream; per half h the column lists of batch_solver 5
    hipStream_t side2 = nullptr;
    hipEvent_t ev_h2f = nullptr, ev_h2j = nullptr;
    int batch_streams = 1;
    std::vector<int> h2_doff[2], h2_poff[2], h2_loff[2];
    int* h2_dmols[2] = {nullptr, nullptr};
    int4 *h2_ptiles[2] = {nullptr, nullptr}, *h2_ltiles[2] = {nullptr, nullptr};
    double* ps_part = nullptr;          // its per-tile row and column partials
    double* ps_part2 = nullptr;         // pair_split: the partials of dispersion and repulsion
    bool pair_split = false;            // dispersion and repulsion before the EEQ solve, Coulomb after
    bool mixed = false;                 // FP32 pair terms + FP32 factorisation with FP64 refinement
    Lookahead la;                       // side stream of the factorisation
    bool use_lookahead = false;
    bool use_cg = false;                // iterative EEQ solve (single molecules)
    int batch_solver = 5;               // batches: 0 / 1 variants of k_eeq_batch_factor, 2 k_eeq_batch_solve_tiles,
                                        // 3 factorisation by steps over the batch, 4 k_eeq_batch_fused,
                                        // 5 left-looking by column blocks over the batch
    bool batch_fp32 = false;            // mixed batches: FP32 factor with FP64 refinement
    bool solv = false;                  // ALPB solvation (single molecules)
    SolvArgs sa{};                      // its parameters and buffers
    int cg_groups = 0;                  // workgroups of the persistent PCG kernel
    double cg_tol2 = 0.0;               // squared relative residual at which the PCG solve stops
    double *cg_z = nullptr, *cg_p = nullptr;
    double* cg_part = nullptr;          // block-product partials [nbk][nbk][64]
    unsigned long long* cg_mb = nullptr;  // all-reduce mailboxes of the persistent PCG
    int* cg_iters = nullptr;
    int* ready = nullptr;               // per-block ready flags of the solve sweeps
    // batches of molecules (nmol > 1)
    int nmol = 1, nbtiles = 0;
    size_t mat_size = 0;                // doubles of the reduced EEQ matrix (or matrices)
    int *mol_of = nullptr, *mol_off = nullptr, *mol_seg = nullptr, *mnr = nullptr, *mnp = nullptr, *mvo = nullptr;
    int* morder = nullptr;  // batch: molecules by decreasing reduced size (solve launch order)
    bool cn_by_mol = false;  // batch of molecules of at most CNMOL_MAX atoms: k_cn_mol
    int cn_cap = 0;          // its LDS capacity in atoms (the largest molecule, rounded up to 64)
    long long* mmo = nullptr;
    int4* btiles = nullptr;
    // batch_solver 3: work of step k over the batch, at [off[k], off[k + 1]) of each list
    std::vector<int> bs_doff, bs_poff, bs_uoff;
    int* bs_dmols = nullptr;            // molecules with more than k column blocks
    int4* bs_ptiles = nullptr;          // panel tiles {m, bi, 0, 0}
    int4* bs_utiles = nullptr;          // trailing tiles {m, bi, bj, 0}
    // batch_solver 5: tiles {m, i, one-fragment flag, 0} of column block j at [ls_off[j], ls_off[j + 1])
    std::vector<int> ls_off;
    int4* ls_tiles = nullptr;
    // batch_solver 6: diagonal tiles and the tiles below them, per column block
    std::vector<int> l6d_off, l6o_off;
    int4 *l6d_tiles = nullptr, *l6o_tiles = nullptr;
    double* ured = nullptr;             // u of every reduced row (k_red_u)
    double* e_mol = nullptr;
    double *adj_w = nullptr, *adj_u = nullptr, *adj_part = nullptr, *adj_bpart = nullptr, *adj_out = nullptr;
    float *Ared32 = nullptr, *Linv32 = nullptr, *b32 = nullptr, *y32 = nullptr, *d32 = nullptr;
    double *rvec = nullptr, *symv = nullptr;
    bool charged = false;
    hipStream_t own = nullptr;
    std::vector<void*> bufs;
    // parameters (device)
    int *Z, *elem, *nref, *metal, *frag, *toff, *tidx, *tdist;
    double *chi_base, *gam, *alp, *cnf, *rcov_cn, *rcov_d3, *hb_rad, *hb_lpd, *qa, *zeta, *szr, *anb, *repz, *refcn, *c6ref;
    int *red_atom, *red_last, *red_of, *frag_off, *frag_atoms;
    double* fsum = nullptr;  // per fragment: the reduced solution summed over its atoms
    int *red_lidx, *red_ul, *red_nul;  // distinct last atoms per 64-block of reduced variables
    double* qp;
    int *bond_ij, *bond_H, *bond_hboff, *bond_hbB, *bond_slot;
    int *bsel_hb, *bsel_plain, nb_hb, nb_plain;  // bond indices with / without HB coordination
    int fin_lpa, hbg_lpa;                          // lanes per atom of k_finish and k_hb_gather
    double* bond_par;
    int *ang_ijk, *ang_slot;
    double* ang_par;
    int *tor_ijkl, *tor_slot;
    double* tor_par;
    int *inv_ijkl, *inv_slot, *st_ijkl, *st_slot, *xb_axb, *xb_slot;
    double *inv_par, *st_par, *xb_par;
    int *brep_ij, *brep_slot;
    double* brep_par;
    // hydrogen bonds by role (k_hb_roles): per-acceptor lists (CSR by atom),
    // role-grouped entries, group results, extra slots and the gather maps
    int *hb_nboff, *hb_nbidx, *hb_parc, *hb_nbcoff, *hb_nbcidx;
    int *hb_goff, *hb_gatom, *hb_crange, *hb_s0, *hb_xoff, *hb_xidx;
    int2* hb_ent;
    int4* hb_ahbc;
    int hb_ng0, hb_ng1, hb_ngt, nhb2, nhb2_c2;  // nhb2_c2: case-2 entries of the cases 2-4 list
    int* hb_ord = nullptr;                      // entries of the cases 2-4 list, case 2 first
    double *hb_cb, *hb_ca, *hb_Qh, *hb_Qab, *hb_gout, *hb_xslots, *e_hb2;
    float4* hb_pk;
    float* hb_qh32;
    int *atm_ijk, *atm_slot;
    double* atm_sc;
    int *batm_ijk, *batm_slot;
    double* batm_zb;
    int *slot_off, *slot_idx, *dcn_off, *dcn_idx;
    // work
    double *x, *cn, *dlog, *gw, *dgw, *rhs, *Ared, *Linv, *b, *rres, *q, *eself, *qtmp, *gpair, *dEdcn, *vchain,
        *slots, *dcn_slot, *grad, *e_coul, *e_disp, *e_rep, *e_bond, *e_ang, *e_tor, *e_brep, *e_hb, *e_atm,
        *e_batm, *e_inv, *e_st, *e_xb, *parts, *rpart, *out, *diag, *ghb, *gslot, *ybuf, *xbuf, *h6, *dh6, *ppart;
    int *nbn, *nbl, *flags;
    int* sticky = nullptr;  // status flags of every evaluation since the last device_status read
    double *lu_w = nullptr, *lu_src = nullptr;  // LU fallback of a non-SPD reduced system (nr <= LU_MAX)

    PairArgs pair_args() const {
        PairArgs a;
        a.n = n;
        a.nelem = nelem;
        a.x = x;
        a.q = q;
        a.alp = alp;
        a.gw = gw;
        a.dgw = dgw;
        a.zeta = zeta;
        a.szr = szr;
        a.anb = anb;
        a.repz = repz;
        a.Z = Z;
        a.elem = elem;
        a.nref = nref;
        a.metal = metal;
        a.c6ref = c6ref;
        a.h = h6;
        a.dh = dh6;
        a.mol_of = mol_of;
        a.mol_off = mol_off;
        a.g = gpair;
        a.dEdcn = dEdcn;
        a.e_coul = e_coul;
        a.e_disp = e_disp;
        a.e_rep = e_rep;
        return a;
    }
    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if (hipMalloc(&p, bytes) != hipSuccess) throw std::bad_alloc();
        bufs.push_back(p);
        return (T*)p;
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* d = alloc<T>(v.size());
        if (!v.empty() && hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("upload failed");
        return d;
    }
    double *md_im3 = nullptr, *md_m3 = nullptr, *md_xref = nullptr, *md_part = nullptr, *md_e = nullptr,
           *md_lambda = nullptr;        // molecular dynamics state (device_md_steps)
    hipGraph_t graph = nullptr;         // captured device-resident evaluation (gfnff_graph_capture)
    hipGraphExec_t gexec = nullptr;
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
    ~DeviceEngine() {
        drop_graph();
        if (la.side) {
            (void)hipStreamSynchronize(la.side);
            (void)hipStreamDestroy(la.side);
        }
        if (la.fork) (void)hipEventDestroy(la.fork);
        if (la.join) (void)hipEventDestroy(la.join);
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
        }
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_fact) (void)hipEventDestroy(ev_fact);
        if (side2) {
            (void)hipStreamSynchronize(side2);
            (void)hipStreamDestroy(side2);
        }
        if (ev_h2f) (void)hipEventDestroy(ev_h2f);
        if (ev_h2j) (void)hipEventDestroy(ev_h2j);
        if (own) {
            (void)hipStreamSynchronize(own);
            (void)hipStreamDestroy(own);
        }
        for (void* p : bufs) (void)hipFree(p);
    }
};

DeviceEngine* device_create(const ParamSet& P, int precision, const Options& opt, int& status, std::string& msg) {
    const bool lookahead = opt.cholesky_lookahead;
    const int solver = opt.eeq_solver;
    const double pcg_tol = opt.pcg_tol;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        status = kErrDevice;
        msg = "no HIP device available";
        return nullptr;
    }
    if (precision != 0 && precision != 1) {
        status = kErrArgs;
        msg = "unknown precision mode";
        return nullptr;
    }
    for (size_t h = 0; h < P.hb_case.size(); ++h)
        if (P.hb_case[h] == 3 && P.hb_parentC[h] >= 0 &&
            (P.hb_nbB_off[h + 1] - P.hb_nbB_off[h] != 1 || P.hb_nbB_idx[P.hb_nbB_off[h]] != P.hb_parentC[h])) {
            status = kErrArgs;
            msg = "hydrogen-bond case 3 needs the carbonyl carbon as the acceptor's only neighbour";
            return nullptr;
        }
    auto* E = new DeviceEngine();
    E->mixed = precision == 1;
    E->use_lookahead = lookahead;
    E->use_cg = solver == 0 && P.nmol == 1;
    E->batch_solver = opt.batch_solver;
    E->batch_streams = (opt.batch_streams == 2 && opt.batch_solver == 5 && P.nmol > 1) ? 2 : 1;
    E->batch_fp32 = opt.batch_fp32 != 0;
    {
        const double tol = pcg_tol > 0.0 ? pcg_tol : (E->mixed ? PCG_TOL_MIXED : PCG_TOL_FP64);
        E->cg_tol2 = tol * tol;
    }
    try {
        const int n = P.n;
        E->n = n;
        E->nelem = P.nelem;
        E->nfrag = P.nfrag;
        if (P.nelem * P.nelem * 49 > 16 * 49) {
            status = kErrUnsupported;
            msg = "more than 4 distinct elements are not supported by the pair kernel yet";
            delete E;
            return nullptr;
        }
        E->Z = E->upload(P.Z);
        E->elem = E->upload(P.elem);
        E->nref = E->upload(P.nref);
        E->metal = E->upload(P.metal);
        E->frag = E->upload(P.frag);
        E->toff = E->upload(P.topo_off);
        E->tidx = E->upload(P.topo_idx);
        E->tdist = E->upload(P.topo_dist);
        E->chi_base = E->upload(P.chi_base);
        E->gam = E->upload(P.gam_eff);
        E->alp = E->upload(P.alpeeq);
        E->cnf = E->upload(P.cnf);
        E->rcov_cn = E->upload(P.rcov_cn);
        E->rcov_d3 = E->upload(P.rcov_d3);
        E->hb_rad = E->upload(P.hb_rad);
        E->hb_lpd = E->upload(P.hb_lpd);
        E->qa = E->upload(P.qa);
        E->zeta = E->upload(P.zeta);
        E->szr = E->upload(P.sqrt_zr4r2);
        E->anb = E->upload(P.rep_anb);
        E->repz = E->upload(P.repz);
        E->refcn = E->upload(P.refcn);
        E->c6ref = E->upload(P.c6ref);

        // fragment null-space reduction
        std::vector<std::vector<int>> fa(P.nfrag);
        for (int i = 0; i < n; ++i) fa[P.frag[i]].push_back(i);
        std::vector<int> red_atom, red_last, red_of(n, -1), foff(P.nfrag + 1, 0), fatoms;
        std::vector<double> qp(n, 0.0);
        for (int f = 0; f < P.nfrag; ++f) {
            int m = (int)fa[f].size();
            int last = fa[f].back();
            for (int k = 0; k + 1 < m; ++k) {
                red_of[fa[f][k]] = (int)red_atom.size();
                red_atom.push_back(fa[f][k]);
                red_last.push_back(last);
            }
            for (int a : fa[f]) {
                qp[a] = P.qfrag[f] / m;
                fatoms.push_back(a);
            }
            foff[f + 1] = (int)fatoms.size();
            if (P.qfrag[f] != 0.0) E->charged = true;
        }
        E->nr = (int)red_atom.size();
        E->np = std::max(NB, cdiv(std::max(E->nr, 1), NB) * NB);
        E->mat_size = (size_t)E->np * E->np;
        if (P.nmol > 1) {
            // batch: each molecule's fragments form their own padded block
            E->nmol = P.nmol;
            std::vector<int> ra2, rl2, mnr(P.nmol), mnp(P.nmol), vo(P.nmol);
            std::vector<long long> mo(P.nmol);
            std::vector<int> fr_first(P.nmol + 1, 0);
            for (int f = 0; f < P.nfrag; ++f) fr_first[P.mol_of_atom(fa[f][0]) + 1] = f + 1;
            for (int m = 0; m < P.nmol; ++m) fr_first[m + 1] = std::max(fr_first[m + 1], fr_first[m]);
            long long off = 0;
            for (int m = 0; m < P.nmol; ++m) {
                vo[m] = (int)ra2.size();
                int loc = 0;
                for (int f = fr_first[m]; f < fr_first[m + 1]; ++f) {
                    const int cnt = (int)fa[f].size(), last = fa[f].back();
                    for (int k = 0; k + 1 < cnt; ++k) {
                        red_of[fa[f][k]] = (int)ra2.size();
                        ra2.push_back(fa[f][k]);
                        rl2.push_back(last);
                        ++loc;
                    }
                }
                mnr[m] = loc;
                mnp[m] = std::max(NB, cdiv(std::max(loc, 1), NB) * NB);
                if (mnp[m] > BATCH_MAX_NP) throw std::runtime_error("batched molecule too large");
                while ((int)ra2.size() < vo[m] + mnp[m]) {
                    ra2.push_back(-1);
                    rl2.push_back(-1);
                }
                mo[m] = off;
                off += (long long)mnp[m] * mnp[m];
            }
            red_atom.swap(ra2);
            red_last.swap(rl2);
            E->nr = (int)red_atom.size();
            E->np = E->nr;
            E->mat_size = (size_t)off;
            std::vector<int4> tiles;
            for (int m = 0; m < P.nmol; ++m)
                for (int bi = 0; bi < mnp[m] / NB; ++bi)
                    for (int bj = 0; bj <= bi; ++bj) tiles.push_back(make_int4(m, bi, bj, 0));
            E->nbtiles = (int)tiles.size();
            E->btiles = E->upload(tiles);
            {
                int maxnb = 0;
                for (int m = 0; m < P.nmol; ++m) maxnb = std::max(maxnb, mnp[m] / NB);
                std::vector<std::vector<int>> dm(maxnb);
                std::vector<std::vector<int4>> pt(maxnb), ut(maxnb);
                for (int m = 0; m < P.nmol; ++m) {
                    const int nbm = mnp[m] / NB;
                    for (int k = 0; k < nbm; ++k) {
                        dm[k].push_back(m);
                        for (int bi = k + 1; bi < nbm; ++bi) {
                            pt[k].push_back(make_int4(m, bi, 0, 0));
                            for (int bj = k + 1; bj <= bi; ++bj) ut[k].push_back(make_int4(m, bi, bj, 0));
                        }
                    }
                }
                std::vector<int> dflat;
                std::vector<int4> pflat, uflat;
                E->bs_doff.assign(1, 0);
                E->bs_poff.assign(1, 0);
                E->bs_uoff.assign(1, 0);
                for (int k = 0; k < maxnb; ++k) {
                    dflat.insert(dflat.end(), dm[k].begin(), dm[k].end());
                    pflat.insert(pflat.end(), pt[k].begin(), pt[k].end());
                    uflat.insert(uflat.end(), ut[k].begin(), ut[k].end());
                    E->bs_doff.push_back((int)dflat.size());
                    E->bs_poff.push_back((int)pflat.size());
                    E->bs_uoff.push_back((int)uflat.size());
                }
                E->bs_dmols = E->upload(dflat);
                if (!pflat.empty()) E->bs_ptiles = E->upload(pflat);
                if (!uflat.empty()) E->bs_utiles = E->upload(uflat);
                // left-looking column tiles; a molecule is one-fragment when
                // all its reduced rows share the last atom
                std::vector<int4> lflat;
                E->ls_off.assign(1, 0);
                std::vector<int> uni(P.nmol, 1);
                for (int m = 0; m < P.nmol; ++m)
                    for (int r = 1; r < mnr[m]; ++r)
                        if (red_last[vo[m] + r] != red_last[vo[m]]) uni[m] = 0;
                for (int j = 0; j < maxnb; ++j) {
                    for (int m = 0; m < P.nmol; ++m)
                        for (int i = j; i < mnp[m] / NB; ++i) lflat.push_back(make_int4(m, i, uni[m] && mnr[m] > 0, 0));
                    E->ls_off.push_back((int)lflat.size());
                }
                E->ls_tiles = E->upload(lflat);
                // batch_solver 6: per column block the diagonal tiles, then the tiles below them
                std::vector<int4> dl, ol;
                E->l6d_off.assign(1, 0);
                E->l6o_off.assign(1, 0);
                for (int j = 0; j < maxnb; ++j) {
                    for (int k = E->ls_off[j]; k < E->ls_off[j + 1]; ++k) (lflat[k].y == j ? dl : ol).push_back(lflat[k]);
                    E->l6d_off.push_back((int)dl.size());
                    E->l6o_off.push_back((int)ol.size());
                }
                E->l6d_tiles = E->upload(dl);
                if (!ol.empty()) E->l6o_tiles = E->upload(ol);
                // batch_streams 2: molecules dealt alternately, by decreasing size, to two halves
                if (E->batch_streams == 2) {
                    std::vector<int> ordm(P.nmol), half(P.nmol);
                    for (int m = 0; m < P.nmol; ++m) ordm[m] = m;
                    std::stable_sort(ordm.begin(), ordm.end(), [&](int a, int b) { return mnp[a] > mnp[b]; });
                    for (int k = 0; k < P.nmol; ++k) half[ordm[k]] = k & 1;
                    for (int h = 0; h < 2; ++h) {
                        std::vector<int> d2;
                        std::vector<int4> p2, l2;
                        E->h2_doff[h].assign(1, 0);
                        E->h2_poff[h].assign(1, 0);
                        E->h2_loff[h].assign(1, 0);
                        for (int j = 0; j < maxnb; ++j) {
                            for (int m : dm[j])
                                if (half[m] == h) d2.push_back(m);
                            for (const int4& t : pt[j])
                                if (half[t.x] == h) p2.push_back(t);
                            for (int k = E->ls_off[j]; k < E->ls_off[j + 1]; ++k)
                                if (half[lflat[k].x] == h) l2.push_back(lflat[k]);
                            E->h2_doff[h].push_back((int)d2.size());
                            E->h2_poff[h].push_back((int)p2.size());
                            E->h2_loff[h].push_back((int)l2.size());
                        }
                        if (!d2.empty()) E->h2_dmols[h] = E->upload(d2);
                        if (!p2.empty()) E->h2_ptiles[h] = E->upload(p2);
                        if (!l2.empty()) E->h2_ltiles[h] = E->upload(l2);
                    }
                }
            }
            E->mnr = E->upload(mnr);
            E->mnp = E->upload(mnp);
            {
                // workgroup order of the per-molecule solves: largest reduced
                // systems first, so the long factorizations start early
                std::vector<int> order(P.nmol);
                std::iota(order.begin(), order.end(), 0);
                std::stable_sort(order.begin(), order.end(), [&](int a2, int b2) { return mnp[a2] > mnp[b2]; });
                E->morder = E->upload(order);
            }
            int maxat = 0;
            for (int m2 = 0; m2 < P.nmol; ++m2) maxat = std::max(maxat, P.mol_off[m2 + 1] - P.mol_off[m2]);
            E->cn_by_mol = maxat <= CNMOL_MAX;
            E->cn_cap = std::max(64, (maxat + 63) / 64 * 64);
            E->mvo = E->upload(vo);
            E->mmo = E->upload(mo);
            E->mol_off = E->upload(P.mol_off);
            std::vector<int> mof(n);
            for (int m = 0; m < P.nmol; ++m)
                for (int i = P.mol_off[m]; i < P.mol_off[m + 1]; ++i) mof[i] = m;
            E->mol_of = E->upload(mof);
            E->mol_seg = E->upload(P.mol_seg);
            E->e_mol = E->alloc<double>(P.nmol);
        }
        E->red_atom = E->upload(red_atom);
        E->red_last = E->upload(red_last);
        {
            // per 64-block of reduced variables: its distinct last atoms (in
            // order of appearance) and each variable's index among them
            const int nbk0 = E->np / NB;
            std::vector<int> lidx(E->np, 0), ul((size_t)E->np, -1), nul(std::max(nbk0, 1), 0);
            for (int bk = 0; bk < nbk0; ++bk) {
                int cnt = 0;
                for (int r = bk * NB; r < std::min((bk + 1) * NB, E->nr); ++r) {
                    const int la = red_last[r];
                    int u = 0;
                    while (u < cnt && ul[(size_t)bk * NB + u] != la) ++u;
                    if (u == cnt) ul[(size_t)bk * NB + cnt++] = la;
                    lidx[r] = u;
                }
                nul[bk] = cnt;
            }
            E->red_lidx = E->upload(lidx);
            E->red_ul = E->upload(ul);
            E->red_nul = E->upload(nul);
        }
        E->red_of = E->upload(red_of);
        E->frag_off = E->upload(foff);
        E->frag_atoms = E->upload(fatoms);
        E->qp = E->upload(qp);
        E->fsum = E->alloc<double>(std::max(P.nfrag, 1));

        // term lists and gradient slots
        int slot = 0;
        std::vector<int> slot_atom;
        auto add_slot = [&](int atom) {
            slot_atom.push_back(atom);
            return slot++;
        };
        // bonds: i, j, (H, B...) for HB-modulated bonds
        E->nbond = (int)P.bond_hbH.size();
        std::vector<int> bhoff(E->nbond + 1, 0), bhB, bslot(E->nbond);
        {
            std::vector<std::vector<int>> byH(n);
            for (size_t k = 0; k + 1 < P.hbcn_pair.size(); k += 2) byH[P.hbcn_pair[k]].push_back(P.hbcn_pair[k + 1]);
            for (int b = 0; b < E->nbond; ++b) {
                int H = P.bond_hbH[b];
                bslot[b] = add_slot(P.bond_ij[2 * b]);
                add_slot(P.bond_ij[2 * b + 1]);
                if (H >= 0) {
                    add_slot(H);
                    for (int B : byH[H]) {
                        bhB.push_back(B);
                        add_slot(B);
                    }
                }
                bhoff[b + 1] = (int)bhB.size();
            }
        }
        E->bond_ij = E->upload(P.bond_ij);
        E->bond_par = E->upload(P.bond_par);
        E->bond_H = E->upload(P.bond_hbH);
        E->bond_hboff = E->upload(bhoff);
        E->bond_hbB = E->upload(bhB);
        E->bond_slot = E->upload(bslot);
        {
            std::vector<int> bh, bp;
            for (int b = 0; b < (int)P.bond_hbH.size(); ++b)
                (P.bond_hbH[b] >= 0 && bhoff[b + 1] > bhoff[b] ? bh : bp).push_back(b);
            E->nb_hb = (int)bh.size();
            E->nb_plain = (int)bp.size();
            E->bsel_hb = E->upload(bh);
            E->bsel_plain = E->upload(bp);
        }
        E->nangle = (int)(P.angle_par.size() / 2);
        std::vector<int> aslot(E->nangle);
        for (int t = 0; t < E->nangle; ++t) {
            aslot[t] = add_slot(P.angle_ijk[3 * t]);
            add_slot(P.angle_ijk[3 * t + 1]);
            add_slot(P.angle_ijk[3 * t + 2]);
        }
        E->ang_ijk = E->upload(P.angle_ijk);
        E->ang_par = E->upload(P.angle_par);
        E->ang_slot = E->upload(aslot);
        E->ntors = (int)(P.tors_par.size() / 4);
        std::vector<int> tslot(E->ntors);
        for (int t = 0; t < E->ntors; ++t) {
            tslot[t] = add_slot(P.tors_ijkl[4 * t]);
            for (int a = 1; a < 4; ++a) add_slot(P.tors_ijkl[4 * t + a]);
        }
        E->tor_ijkl = E->upload(P.tors_ijkl);
        E->tor_par = E->upload(P.tors_par);
        E->tor_slot = E->upload(tslot);
        E->ninv = (int)(P.inv_par.size() / 3);
        std::vector<int> islot(E->ninv);
        for (int t = 0; t < E->ninv; ++t) {
            islot[t] = add_slot(P.inv_ijkl[4 * t]);
            for (int a = 1; a < 4; ++a) add_slot(P.inv_ijkl[4 * t + a]);
        }
        E->inv_ijkl = E->upload(P.inv_ijkl);
        E->inv_par = E->upload(P.inv_par);
        E->inv_slot = E->upload(islot);
        E->nstors = (int)P.stors_par.size();
        std::vector<int> sslot(E->nstors);
        for (int t = 0; t < E->nstors; ++t) {
            sslot[t] = add_slot(P.stors_ijkl[4 * t]);
            for (int a = 1; a < 4; ++a) add_slot(P.stors_ijkl[4 * t + a]);
        }
        E->st_ijkl = E->upload(P.stors_ijkl);
        E->st_par = E->upload(P.stors_par);
        E->st_slot = E->upload(sslot);
        E->nxb = (int)(P.xb_par.size() / 3);
        std::vector<int> xslot(E->nxb);
        for (int t = 0; t < E->nxb; ++t) {
            xslot[t] = add_slot(P.xb_axb[3 * t]);
            add_slot(P.xb_axb[3 * t + 1]);
            add_slot(P.xb_axb[3 * t + 2]);
        }
        E->xb_axb = E->upload(P.xb_axb);
        E->xb_par = E->upload(P.xb_par);
        E->xb_slot = E->upload(xslot);
        E->nbrep = (int)(P.brep_par.size() / 2);
        std::vector<int> rslot(E->nbrep);
        for (int b = 0; b < E->nbrep; ++b) {
            rslot[b] = add_slot(P.brep_ij[2 * b]);
            add_slot(P.brep_ij[2 * b + 1]);
        }
        E->brep_ij = E->upload(P.brep_ij);
        E->brep_par = E->upload(P.brep_par);
        E->brep_slot = E->upload(rslot);
        E->nhb = (int)P.hb_case.size();
        {
            // Hydrogen bonds.  The reference's per-triplet parameters are
            // per-atom values and its neighbour lists depend on the acceptor
            // only; both are checked while they are gathered.  Case-1 triplets
            // go to the role lists, cases 2-4 to a compact list with slots.
            const int nh = E->nhb;
            if (nh && n >= (1 << 28)) throw std::runtime_error("hydrogen-bond lists need atom indices below 2^28");
            std::vector<double> cb(n, 0.0), ca(n, 0.0), qt(n, 0.0);
            std::vector<char> s_cb(n, 0), s_ca(n, 0), s_qt(n, 0);
            auto put = [](std::vector<double>& v, std::vector<char>& set, int i, double val) {
                if (set[i] && v[i] != val) throw std::runtime_error("hydrogen-bond parameter is not a per-atom value");
                v[i] = val;
                set[i] = 1;
            };
            std::vector<std::vector<int>> bnb(n), bD(n);
            std::vector<int> bC(n, -1);
            std::vector<char> s_nb(n, 0), s_c(n, 0);
            for (int h = 0; h < nh; ++h) {
                const int A = P.hb_ahb[3 * h], H = P.hb_ahb[3 * h + 1], B = P.hb_ahb[3 * h + 2], cs = P.hb_case[h];
                const double* q = &P.hb_par[7 * (size_t)h];
                put(cb, s_cb, A, q[0]);
                put(cb, s_cb, B, q[1]);
                put(ca, s_ca, A, q[2]);
                put(ca, s_ca, B, q[3]);
                put(qt, s_qt, H, q[4]);
                put(qt, s_qt, A, q[5]);
                put(qt, s_qt, B, q[6]);
                if (cs < 1 || cs > 4) throw std::runtime_error("unknown hydrogen-bond case");
                if (cs >= 2) {
                    std::vector<int> l(P.hb_nbB_idx.begin() + P.hb_nbB_off[h], P.hb_nbB_idx.begin() + P.hb_nbB_off[h + 1]);
                    if (s_nb[B] && bnb[B] != l) throw std::runtime_error("acceptor neighbour lists differ between hydrogen bonds");
                    bnb[B] = l;
                    s_nb[B] = 1;
                }
                if (cs == 3) {
                    std::vector<int> l(P.hb_nbC_idx.begin() + P.hb_nbC_off[h], P.hb_nbC_idx.begin() + P.hb_nbC_off[h + 1]);
                    if (s_c[B] && (bC[B] != P.hb_parentC[h] || bD[B] != l))
                        throw std::runtime_error("carbonyl lists differ between hydrogen bonds");
                    bC[B] = P.hb_parentC[h];
                    bD[B] = l;
                    s_c[B] = 1;
                }
            }
            // charge switches Q(+q) for H and Q(-q) for A and B
            std::vector<double> Qh(n), Qab(n);
            for (int i = 0; i < n; ++i) {
                const double eh = std::exp(15.0 * qt[i]), em = std::exp(15.0 * -qt[i]);
                Qh[i] = eh / (eh + 1.0);
                Qab[i] = em / (em + 1.0);
            }
            auto csr = [&](const std::vector<std::vector<int>>& L, std::vector<int>& off, std::vector<int>& idx) {
                off.assign(n + 1, 0);
                for (int i = 0; i < n; ++i) off[i + 1] = off[i] + (int)L[i].size();
                idx.clear();
                for (int i = 0; i < n; ++i) idx.insert(idx.end(), L[i].begin(), L[i].end());
            };
            std::vector<int> nboff, nbidx, nbcoff, nbcidx;
            csr(bnb, nboff, nbidx);
            csr(bD, nbcoff, nbcidx);
            // case 1: role lists (counting sort by the role's atom, triplet order within a group)
            std::vector<int> c1;
            for (int h = 0; h < nh; ++h)
                if (P.hb_case[h] == 1) c1.push_back(h);
            std::vector<int2> ent((size_t)3 * c1.size());
            std::vector<int> goff(1, 0), gatom, crange((size_t)3 * (n + 1), 0);
            int ngr[3] = {0, 0, 0};
            size_t eo = 0;
            for (int role = 0; role < 3; ++role) {
                std::vector<int> cnt(n + 1, 0);
                for (int h : c1) cnt[P.hb_ahb[3 * h + role] + 1]++;
                for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
                std::vector<int> order(c1.size()), fill(cnt.begin(), cnt.end() - 1);
                for (int h : c1) order[fill[P.hb_ahb[3 * h + role]]++] = h;
                for (int i = 0; i < n; ++i) {
                    crange[(size_t)role * (n + 1) + i] = (int)gatom.size();
                    for (int k = cnt[i]; k < cnt[i + 1]; ++k) {
                        const int h = order[k];
                        const int A = P.hb_ahb[3 * h], H = P.hb_ahb[3 * h + 1], B = P.hb_ahb[3 * h + 2];
                        ent[eo++] = role == 0 ? make_int2(H, B) : (role == 1 ? make_int2(A, B) : make_int2(A, H));
                        if ((k - cnt[i]) % HB_CHUNK == HB_CHUNK - 1 || k + 1 == cnt[i + 1]) {
                            gatom.push_back(i);
                            goff.push_back((int)eo);
                        }
                    }
                }
                crange[(size_t)role * (n + 1) + n] = (int)gatom.size();
                ngr[role] = (int)gatom.size();
            }
            E->hb_ng0 = ngr[0];
            E->hb_ng1 = ngr[1];
            E->hb_ngt = ngr[2];
            // cases 2-4: compact list, slots triplet-major (A, H, B, neighbours
            // of B, carbonyl D atoms) and the per-atom CSR over them
            std::vector<int4> ahbc;
            std::vector<int> s0v, xatom, c2mol;
            for (int h = 0; h < nh; ++h) {
                const int cs = P.hb_case[h];
                if (cs == 1) continue;
                const int A = P.hb_ahb[3 * h], H = P.hb_ahb[3 * h + 1], B = P.hb_ahb[3 * h + 2];
                ahbc.push_back(make_int4(A, H, B, cs));
                s0v.push_back((int)xatom.size());
                xatom.push_back(A);
                xatom.push_back(H);
                xatom.push_back(B);
                for (int a2 : bnb[B]) xatom.push_back(a2);
                if (cs == 3)
                    for (int a2 : bD[B]) xatom.push_back(a2);
                c2mol.push_back(h);
                if (xatom.size() > (size_t)INT32_MAX / 4)
                    throw std::runtime_error("hydrogen-bond slot count exceeds the index range");
            }
            E->nhb2 = (int)ahbc.size();
            E->nhb2_c2 = 0;
            std::vector<int> ord;
            for (size_t h = 0; h < ahbc.size(); ++h)
                if (ahbc[h].w == 2) ord.push_back((int)h);
            E->nhb2_c2 = (int)ord.size();
            for (size_t h = 0; h < ahbc.size(); ++h)
                if (ahbc[h].w != 2) ord.push_back((int)h);
            E->hb_ord = E->upload(ord);
            std::vector<int> xoff(n + 1, 0), xidx(xatom.size());
            for (int at : xatom) xoff[at + 1]++;
            for (int i = 0; i < n; ++i) xoff[i + 1] += xoff[i];
            {
                std::vector<int> fill(xoff.begin(), xoff.end() - 1);
                for (size_t k = 0; k < xatom.size(); ++k) xidx[fill[xatom[k]]++] = (int)k;
            }
            E->hb_cb = E->upload(cb);
            E->hb_ca = E->upload(ca);
            E->hb_Qh = E->upload(Qh);
            E->hb_Qab = E->upload(Qab);
            {
                std::vector<float4> pk(n);
                std::vector<float> qh32(n);
                for (int i = 0; i < n; ++i) {
                    pk[i] = make_float4((float)cb[i], (float)ca[i], (float)Qab[i], (float)P.hb_rad[i]);
                    qh32[i] = (float)Qh[i];
                }
                E->hb_pk = E->upload(pk);
                E->hb_qh32 = E->upload(qh32);
            }
            E->hb_nboff = E->upload(nboff);
            E->hb_nbidx = E->upload(nbidx);
            E->hb_parc = E->upload(bC);
            E->hb_nbcoff = E->upload(nbcoff);
            E->hb_nbcidx = E->upload(nbcidx);
            E->hb_ent = E->upload(ent);
            E->hb_goff = E->upload(goff);
            E->hb_gatom = E->upload(gatom);
            E->hb_crange = E->upload(crange);
            E->hb_ahbc = E->upload(ahbc);
            E->hb_s0 = E->upload(s0v);
            E->hb_xoff = E->upload(xoff);
            E->hb_xidx = E->upload(xidx);
            // lanes per atom of the gather: 16 when the atoms' lists are short on average
            E->hbg_lpa = ((long long)crange[(size_t)2 * (n + 1) + n] + xoff[n] < 192LL * n) ? 16 : 64;
            if (opt.gather_lanes) E->hbg_lpa = opt.gather_lanes;
            E->hb_gout = E->alloc<double>(3 * (size_t)std::max(E->hb_ngt, 1));
            E->hb_xslots = E->alloc<double>(3 * std::max<size_t>(xatom.size(), 1));
            E->e_hb2 = E->alloc<double>(std::max(E->nhb2, 1));
            if (E->nmol > 1) {
                // per-molecule segments: row 4 the case-1 role-A chunks of the
                // molecule's atoms (chunks ascend in atom index), row 10 its
                // triplets of cases 2-4 (in triplet order, molecule-major)
                const int nm1 = E->nmol + 1;
                std::vector<int> ms(P.mol_seg.begin(), P.mol_seg.begin() + 10 * nm1);
                ms.resize(11 * nm1);
                for (int m = 0; m <= E->nmol; ++m) {
                    const int a0 = (m < E->nmol) ? P.mol_off[m] : n;
                    ms[4 * nm1 + m] = (int)(std::lower_bound(gatom.begin(), gatom.begin() + ngr[0], a0) - gatom.begin());
                    const int h0 = P.mol_seg[4 * nm1 + m];
                    ms[10 * nm1 + m] = (int)(std::lower_bound(c2mol.begin(), c2mol.end(), h0) - c2mol.begin());
                }
                E->mol_seg = E->upload(ms);
            }
        }
        E->natm = (int)P.atm_scale.size();
        std::vector<int> atslot(E->natm);
        for (int t = 0; t < E->natm; ++t) {
            atslot[t] = add_slot(P.atm_ijk[3 * t]);
            add_slot(P.atm_ijk[3 * t + 1]);
            add_slot(P.atm_ijk[3 * t + 2]);
        }
        E->atm_ijk = E->upload(P.atm_ijk);
        E->atm_sc = E->upload(P.atm_scale);
        E->atm_slot = E->upload(atslot);
        E->nbatm = (int)(P.batm_par.size() / 3);
        std::vector<int> btslot(E->nbatm);
        for (int t = 0; t < E->nbatm; ++t) {
            btslot[t] = add_slot(P.batm_ijk[3 * t]);
            add_slot(P.batm_ijk[3 * t + 1]);
            add_slot(P.batm_ijk[3 * t + 2]);
        }
        E->batm_ijk = E->upload(P.batm_ijk);
        E->batm_zb = E->upload(P.batm_par);
        E->batm_slot = E->upload(btslot);
        E->nslots = slot;
        // per-atom CSR over slots (ascending slot order)
        std::vector<int> soff(n + 1, 0), sidx(slot);
        for (int a : slot_atom) soff[a + 1]++;
        for (int i = 0; i < n; ++i) soff[i + 1] += soff[i];
        {
            std::vector<int> fill(soff.begin(), soff.end() - 1);
            for (int s = 0; s < slot; ++s) sidx[fill[slot_atom[s]]++] = s;
        }
        E->slot_off = E->upload(soff);
        E->fin_lpa = ((long long)soff[n] < 192LL * n) ? 16 : 64;  // lanes per atom of k_finish
        if (opt.gather_lanes) E->fin_lpa = opt.gather_lanes;
        E->slot_idx = E->upload(sidx);
        std::vector<int> doff(n + 1, 0), didx(2 * E->nbond);
        for (int b = 0; b < E->nbond; ++b) {
            doff[P.bond_ij[2 * b] + 1]++;
            doff[P.bond_ij[2 * b + 1] + 1]++;
        }
        for (int i = 0; i < n; ++i) doff[i + 1] += doff[i];
        {
            std::vector<int> fill(doff.begin(), doff.end() - 1);
            for (int b = 0; b < E->nbond; ++b) {
                didx[fill[P.bond_ij[2 * b]]++] = 2 * b;
                didx[fill[P.bond_ij[2 * b + 1]]++] = 2 * b + 1;
            }
        }
        E->dcn_off = E->upload(doff);
        E->dcn_idx = E->upload(didx);

        // work buffers
        E->x = E->alloc<double>(3 * (size_t)n);
        E->cn = E->alloc<double>(n);
        E->dlog = E->alloc<double>(n);
        E->gw = E->alloc<double>(7 * (size_t)n);
        E->dgw = E->alloc<double>(7 * (size_t)n);
        E->rhs = E->alloc<double>(n);
        E->diag = E->alloc<double>(n);
        if (E->use_cg) {
            // every workgroup of the persistent kernel must be resident at once
            const int nbk = E->np / NB;
            int dev = 0, ncu = 0, per_cu = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_eeq_pcg, PCG_THREADS, 0) != hipSuccess) {
                status = kErrDevice;
                msg = "PCG occupancy query failed";
                delete E;
                return nullptr;
            }
            const long T = (long)nbk * (nbk + 1) / 2;
            E->cg_groups = (int)std::min<long>({(long)ncu * per_cu, 256L, T});
            if (opt.pcg_groups > 0) E->cg_groups = std::min(E->cg_groups, opt.pcg_groups);
            if (E->cg_groups < 1 || cdiv(nbk, E->cg_groups) > PCG_MAXB) E->use_cg = false;
            if (E->use_cg) E->mat_size = (size_t)T * NB * NB;  // packed symmetric storage
        }
        E->Ared = E->alloc<double>(E->mat_size);
        E->Linv = E->alloc<double>((size_t)E->np * NB);
        E->b = E->alloc<double>(E->np);
        E->ybuf = E->alloc<double>(E->np);
        E->xbuf = E->alloc<double>(E->np);
        E->ready = E->alloc<int>(E->np / NB);
        if (E->nmol > 1) E->ured = E->alloc<double>(E->np);
        if (E->nmol == 1 && E->nr <= LU_MAX) {
            E->lu_w = E->alloc<double>((size_t)std::max(E->nr, 1) * (std::max(E->nr, 1) + 1));
            // the FP64 Cholesky factors R in place: keep a copy for the fallback
            if (!E->use_cg && !E->mixed) E->lu_src = E->alloc<double>(E->mat_size);
        }
        if (E->use_cg) {
            E->cg_z = E->alloc<double>(E->np);
            E->cg_p = E->alloc<double>(2 * (size_t)E->np);
            E->cg_part = E->alloc<double>((size_t)(E->np / NB) * E->np);
            E->cg_mb = E->alloc<unsigned long long>(MB_WORDS);
            E->cg_iters = E->alloc<int>(1);
        } else if (E->mixed && E->nmol > 1 && E->batch_fp32 && E->batch_solver != 2 && E->batch_solver < 4) {
            E->Ared32 = E->alloc<float>(E->mat_size);  // FP32 copies of the molecules' matrices and factors
            E->Linv32 = E->alloc<float>((size_t)E->np * NB);
        } else if (E->mixed && E->nmol == 1) {
            E->Ared32 = E->alloc<float>((size_t)E->np * E->np);
            E->Linv32 = E->alloc<float>((size_t)E->np * NB);
            E->b32 = E->alloc<float>(E->np);
            E->y32 = E->alloc<float>(E->np);
            E->d32 = E->alloc<float>(E->np);
            E->rvec = E->alloc<double>(E->np);
            E->symv = E->alloc<double>((size_t)SYMV_SPLIT * E->np);
        }
        E->rres = E->alloc<double>(n);
        E->q = E->alloc<double>(n);
        E->eself = E->alloc<double>(n);
        E->qtmp = E->alloc<double>(n);
        E->gpair = E->alloc<double>(3 * (size_t)n);
        E->dEdcn = E->alloc<double>(n);
        E->vchain = E->alloc<double>(n);
        E->slots = E->alloc<double>(3 * (size_t)std::max(slot, 1));
        E->dcn_slot = E->alloc<double>(2 * (size_t)std::max(E->nbond, 1));
        E->grad = E->alloc<double>(3 * (size_t)n);
        E->e_coul = E->alloc<double>(n);
        E->e_disp = E->alloc<double>(n);
        E->e_rep = E->alloc<double>(n);
        E->h6 = E->alloc<double>((size_t)n * E->nelem * 7);
        E->dh6 = E->alloc<double>((size_t)n * E->nelem * 7);
        {
            // j chunks: the grid (row blocks x chunks) should fill whole
            // rounds of the resident-workgroup capacity of the pair kernel in
            // use (at most two rounds, at least 128 partners per chunk; ties
            // to fewer chunks)
            const int rowblk = cdiv(n, TILE);
            const void* fn = nullptr;
            switch (E->mixed ? std::min(P.nelem, 5) : 10 + std::min(P.nelem, 5)) {
                case 1: fn = (const void*)k_pairs_part32<1>; break;
                case 2: fn = (const void*)k_pairs_part32<2>; break;
                case 3: fn = (const void*)k_pairs_part32<3>; break;
                case 4: fn = (const void*)k_pairs_part32<4>; break;
                case 5: fn = (const void*)k_pairs_part32<0>; break;
                case 11: fn = (const void*)k_pairs_part<1>; break;
                case 12: fn = (const void*)k_pairs_part<2>; break;
                case 13: fn = (const void*)k_pairs_part<3>; break;
                case 14: fn = (const void*)k_pairs_part<4>; break;
                default: fn = (const void*)k_pairs_part<0>; break;
            }
            int dev = 0, ncu = 0, per_cu = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TILE, 0) != hipSuccess || per_cu < 1)
                throw std::runtime_error("pair kernel occupancy query failed");
            const long cap = (long)ncu * per_cu;
            int best = 1;
            double best_eff = -1.0;
            for (int nck = 1; nck <= std::min(n, (int)(2 * cap)); ++nck) {
                const int cj = cdiv(n, nck), real = cdiv(n, cj);
                const long wg = (long)rowblk * real, rounds = (wg + cap - 1) / cap;
                if (rounds > 2 || cj < std::min(n, 128)) break;  // at least 128 partners per chunk
                const double eff = (double)wg / (double)(rounds * cap);
                if (eff > best_eff + 1e-9) {
                    best_eff = eff;
                    best = real;
                }
            }
            E->pair_cj = cdiv(n, best);
            E->pair_nchunk = cdiv(n, E->pair_cj);
            if (E->nmol > 1) {  // rows see only their molecule: one chunk
                E->pair_cj = n;
                E->pair_nchunk = 1;
            }
        }
        E->ppart = E->alloc<double>((size_t)E->pair_nchunk * PART_K * n);
        if (E->nmol == 1) {
            // each unordered pair once: 64 x 64 tiles of the lower triangle
            E->ps_nb = cdiv(n, PS_B);
            E->ps_ntiles = E->ps_nb * (E->ps_nb + 1) / 2;
            E->ps_nblocks = E->ps_nb;
            E->tbox = E->alloc<double>(6 * (size_t)cdiv(n, 256));
            E->ps_part = E->alloc<double>((size_t)E->ps_ntiles * 2 * PART_K * PS_B);
        } else if (opt.batch_pairs == 0) {
            // batches: the same tiles per molecule, listed {a0, n, R, C}, and the
            // molecules' 64-atom blocks {a0, n, A, first tile} for the reduction
            std::vector<int4> tl, bl;
            for (int m = 0; m < P.nmol; ++m) {
                const int a0 = P.mol_off[m], nm = P.mol_off[m + 1] - a0, nbm = cdiv(nm, PS_B);
                const int base = (int)tl.size();
                for (int R = 0; R < nbm; ++R)
                    for (int C = 0; C <= R; ++C) tl.push_back(make_int4(a0, nm, R, C));
                for (int A = 0; A < nbm; ++A) bl.push_back(make_int4(a0, nm, A, base));
            }
            E->ps_ntiles = (int)tl.size();
            E->ps_nblocks = (int)bl.size();
            E->ps_tiles = E->upload(tl);
            E->ps_blocks = E->upload(bl);
            E->ps_part = E->alloc<double>((size_t)E->ps_ntiles * 2 * PART_K * PS_B);
        }
        E->pair_split = opt.pair_split != 0 && E->ps_ntiles > 0;
        if (E->pair_split) E->ps_part2 = E->alloc<double>((size_t)E->ps_ntiles * 2 * PART_K * PS_B);
        E->e_bond = E->alloc<double>(std::max(E->nbond, 1));
        E->e_ang = E->alloc<double>(std::max(E->nangle, 1));
        E->e_tor = E->alloc<double>(std::max(E->ntors, 1));
        E->e_inv = E->alloc<double>(std::max(E->ninv, 1));
        E->e_st = E->alloc<double>(std::max(E->nstors, 1));
        E->e_xb = E->alloc<double>(std::max(E->nxb, 1));
        E->e_brep = E->alloc<double>(std::max(E->nbrep, 1));
        E->e_hb = E->alloc<double>(std::max(E->hb_ng0, 1));
        E->e_atm = E->alloc<double>(std::max(E->natm, 1));
        E->e_batm = E->alloc<double>(std::max(E->nbatm, 1));
        E->parts = E->alloc<double>(16);
        if (P.solv.n > 0) {
            if (P.nmol != 1 || P.solv.n != n) {
                status = kErrUnsupported;
                msg = "ALPB solvation is built for single molecules";
                delete E;
                return nullptr;
            }
            const AlpbSet& A = P.solv;
            SolvArgs& a = E->sa;
            E->solv = true;
            a.n = n;
            a.keps = A.keps;
            a.alpbet = A.alpbet;
            a.c1 = A.born_scale;
            a.gshift = A.gshift;
            a.lr2 = A.lrcut * A.lrcut;
            a.sr2 = A.srcut * A.srcut;
            a.x = E->x;
            a.vdwr = E->upload(A.vdwr);
            a.rho = E->upload(A.rho);
            a.svdw = E->upload(A.svdw);
            a.vdwsa = E->upload(A.vdwsa);
            a.wrp = E->upload(A.wrp);
            a.trjlo = E->upload(A.trj_lo);
            a.trjhi = E->upload(A.trj_hi);
            a.gam = E->upload(A.gamsasa);
            a.hbmag = E->upload(A.hbmag);
            if (A.nang != SOLV_NANG) throw std::runtime_error("ALPB grid size");
            a.grid = E->upload(A.grid);
            a.brad = E->alloc<double>(n);
            a.dbr = E->alloc<double>(n);
            a.sasa = E->alloc<double>(n);
            a.hbw = E->alloc<double>(n);
            a.sv = E->alloc<double>((size_t)n * SOLV_NB * 3);
            a.sj = E->alloc<int>((size_t)n * SOLV_NB);
            a.sn = E->alloc<int>(n);
            a.shape = E->alloc<double>(32);
            a.w = E->alloc<double>(n);
            a.c = E->alloc<double>(n);
            a.eat = E->alloc<double>(n);
            a.gs = E->alloc<double>(3 * (size_t)n);
            a.e_solv = E->alloc<double>(1);
            a.parts = E->alloc<double>(4);
        }
        E->rpart = E->alloc<double>(16 * RED_BLOCKS);
        E->ghb = E->alloc<double>(3 * (size_t)n);
        E->gslot = E->alloc<double>(3 * (size_t)n);
        E->out = E->alloc<double>(GFNFF_NTERMS + 1);
        E->nbn = E->alloc<int>(n);
        E->nbl = E->alloc<int>((size_t)n * MAXNB_CN);
        E->flags = E->alloc<int>(1);
        E->sticky = E->alloc<int>(1);
        E->sa.flags = E->flags;
        if (hipStreamCreate(&E->own) != hipSuccess || hipStreamCreate(&E->la.side) != hipSuccess ||
            hipEventCreateWithFlags(&E->la.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&E->la.join, hipEventDisableTiming) != hipSuccess ||
            hipStreamCreateWithFlags(&E->side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&E->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&E->ev_join, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&E->ev_fact, hipEventDisableTiming) != hipSuccess)
            throw std::runtime_error("stream");
        if (E->batch_streams == 2 &&
            (hipStreamCreateWithFlags(&E->side2, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreateWithFlags(&E->ev_h2f, hipEventDisableTiming) != hipSuccess ||
             hipEventCreateWithFlags(&E->ev_h2j, hipEventDisableTiming) != hipSuccess))
            throw std::runtime_error("stream");
        hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, E->own, 1, E->flags, 0);
        hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, E->own, 1, E->sticky, 0);
        hipLaunchKernelGGL(k_eeq_diag, dim3(cdiv(n, 256)), dim3(256), 0, E->own, n, E->gam, E->alp, E->diag);
        if (hipStreamSynchronize(E->own) != hipSuccess) throw std::runtime_error("init sync");
    } catch (const std::bad_alloc&) {
        status = kErrDevice;
        msg = "device allocation failed";
        delete E;
        return nullptr;
    } catch (const std::exception& ex) {
        status = kErrDevice;
        msg = ex.what();
        delete E;
        return nullptr;
    }
    status = kOk;
    return E;
}

void device_destroy(DeviceEngine* e) { delete e; }

// Queues the two-level blocked factorisation of the lower triangle of A
// (np x np, np a multiple of NB) on stream s; Linv receives the inverses of the
// diagonal blocks, flags FLAG_CHOL on a non-positive pivot.  tm (optional)
// times the kernel classes.
//
// Lookahead (la != nullptr): once super-panel K is factored, the columns of the
// next super-panel receive their rank-(SP NB) update first on s; the rest of
// the trailing matrix is updated on the side stream while s factors the next
// super-panel, which touches only its own columns.  The side stream is joined
// before the next trailing update and at the end.
template <typename T>
static void queue_cholesky(KernelTimer* tm, hipStream_t s, int np, T* A, T* Linv, int* flags,
                           const Lookahead* la = nullptr) {
    auto mark = [&](int c, hipStream_t st) {
        if (tm) tm->mark(c, st);
    };
    const int nbk = np / NB;
    bool pending = false;  // side-stream update not yet joined
    for (int K = 0; K < nbk; K += SP) {
        const int kend = std::min(K + SP, nbk);
        for (int k = K; k < kend; ++k) {
            const int k0 = k * NB;
            mark(KC_CHOL_DIAG, s);
            hipLaunchKernelGGL(k_chol_diag<T>, dim3(1), dim3(256), 0, s, np, k0, A, Linv, flags);
            mark(KC_CHOL_DIAG, s);
            const int rem = nbk - k - 1;
            if (rem == 0) continue;
            mark(KC_CHOL_PANEL, s);
            hipLaunchKernelGGL(k_chol_panel<T>, dim3(rem), dim3(256), 0, s, np, k0, A, Linv);
            mark(KC_CHOL_PANEL, s);
            const int w = kend - k - 1;
            if (w > 0) {
                const int ntiles = w * (w + 1) / 2 + (rem - w) * w;
                mark(KC_CHOL_UPDATE, s);
                hipLaunchKernelGGL(k_chol_update<T>, dim3(ntiles), dim3(256), 0, s, np, k0, NB, k + 1, w, nbk, A);
                mark(KC_CHOL_UPDATE, s);
            }
        }
        if (kend >= nbk) break;
        if (pending) {  // the previous side update wrote the tiles updated next
            (void)hipStreamWaitEvent(s, la->join, 0);
            pending = false;
        }
        const int M = nbk - kend;
        const int kend2 = std::min(kend + SP, nbk);
        if (!la || kend2 >= nbk) {
            mark(KC_CHOL_UPDATE, s);
            hipLaunchKernelGGL(k_chol_update<T>, dim3(M * (M + 1) / 2), dim3(256), 0, s, np, K * NB,
                               (kend - K) * NB, kend, M, nbk, A);
            mark(KC_CHOL_UPDATE, s);
            continue;
        }
        // rest of the trailing matrix (column blocks >= kend2) on the side stream
        const int M2 = nbk - kend2;
        (void)hipEventRecord(la->fork, s);
        (void)hipStreamWaitEvent(la->side, la->fork, 0);
        mark(KC_CHOL_UPDATE, la->side);
        hipLaunchKernelGGL(k_chol_update<T>, dim3(M2 * (M2 + 1) / 2), dim3(256), 0, la->side, np, K * NB,
                           (kend - K) * NB, kend2, M2, nbk, A);
        mark(KC_CHOL_UPDATE, la->side);
        (void)hipEventRecord(la->join, la->side);
        pending = true;
        // columns of the next super-panel (column blocks [kend, kend2), all rows) on s
        const int w = kend2 - kend;
        mark(KC_CHOL_UPDATE, s);
        hipLaunchKernelGGL(k_chol_update<T>, dim3(w * (w + 1) / 2 + (M - w) * w), dim3(256), 0, s, np, K * NB,
                           (kend - K) * NB, kend, w, nbk, A);
        mark(KC_CHOL_UPDATE, s);
    }
    if (pending) (void)hipStreamWaitEvent(s, la->join, 0);
}

// Forward and backward substitution with the blocked factor, one launch per
// sweep (k_trsv_*_sweep); the ready flags are cleared by a fill kernel, not a
// memset, so that the sequence can be captured in a graph.
template <typename T>
static void queue_solve(DeviceEngine* E, hipStream_t s, const T* A, const T* Linv, const T* b, T* y, T* x) {
    const int np = E->np, nbk = np / NB;
    LAUNCH(KC_TRSV, k_fill_i, dim3(cdiv(nbk, 256)), dim3(256), s, nbk, E->ready, 0);
    LAUNCH(KC_TRSV, k_trsv_fwd_sweep<T>, dim3(nbk), dim3(256), s, np, A, Linv, b, y, E->ready, E->flags);
    LAUNCH(KC_TRSV, k_fill_i, dim3(cdiv(nbk, 256)), dim3(256), s, nbk, E->ready, 0);
    LAUNCH(KC_TRSV, k_trsv_bwd_sweep<T>, dim3(nbk), dim3(256), s, np, A, Linv, y, x, E->ready, E->flags);
}

// batch_solver 3: the per-molecule factorisations advanced one column block at
// a time over the whole batch (three launches per step), then the
// substitutions (and, for FP32 factors, the FP64 refinement) in one launch.
template <typename T>
static void queue_batch_steps(DeviceEngine* E, hipStream_t s, T* A, T* Linv) {
    const int nsteps = (int)E->bs_doff.size() - 1;
    for (int k = 0; k < nsteps; ++k) {
        const int nd = E->bs_doff[k + 1] - E->bs_doff[k];
        const int npn = E->bs_poff[k + 1] - E->bs_poff[k];
        const int nu = E->bs_uoff[k + 1] - E->bs_uoff[k];
        LAUNCH(KC_BATCH_SOLVE, k_bstep_diag<T>, dim3(nd), dim3(256), s, E->bs_dmols + E->bs_doff[k], k, E->mnp, E->mmo,
               E->mvo, A, Linv, E->flags);
        if (npn)
            LAUNCH(KC_BATCH_SOLVE, (k_bstep_tiles<T, false>), dim3(cdiv(npn, 4)), dim3(256), s,
                   E->bs_ptiles + E->bs_poff[k], npn, k, E->mnp, E->mmo, E->mvo, A, Linv);
        if (nu)
            LAUNCH(KC_BATCH_SOLVE, (k_bstep_tiles<T, true>), dim3(cdiv(nu, 4)), dim3(256), s,
                   E->bs_utiles + E->bs_uoff[k], nu, k, E->mnp, E->mmo, E->mvo, A, Linv);
    }
}

// The batch factorisation (every variant but the whole-workgroup solve, which
// includes its substitutions) on stream s.
static void queue_batch_factor(DeviceEngine* E, hipStream_t s) {
    const dim3 g(E->nmol), b(BSOLVE_T);
    if (E->batch_solver == 4) {  // left-looking, matrix formed in place (FP64 on both paths)
        LAUNCH(KC_BATCH_SOLVE, k_eeq_batch_fused<3>, g, b, s, E->morder, E->mnr, E->mnp, E->mmo, E->mvo,
               E->red_atom, E->red_last, E->x, E->alp, E->diag, E->Ared, E->Linv, E->flags);
    } else if (E->batch_solver == 5 && E->batch_streams == 2) {  // two halves of the batch on two streams
        LAUNCH(KC_BATCH_SOLVE, k_red_u, dim3(cdiv(E->np, 256)), dim3(256), s, E->np, E->red_atom, E->red_last, E->x,
               E->alp, E->diag, E->ured);
        (void)hipEventRecord(E->ev_h2f, s);
        (void)hipStreamWaitEvent(E->side2, E->ev_h2f, 0);
        const int ncol = (int)E->h2_doff[0].size() - 1;
        for (int j = 0; j < ncol; ++j)
            for (int h = 0; h < 2; ++h) {
                hipStream_t st = h ? E->side2 : s;
                const int nt = E->h2_loff[h][j + 1] - E->h2_loff[h][j];
                const int nd = E->h2_doff[h][j + 1] - E->h2_doff[h][j];
                const int npn = E->h2_poff[h][j + 1] - E->h2_poff[h][j];
                if (nt)
                    LAUNCH(KC_BATCH_SOLVE, k_lstep_tiles, dim3(cdiv(nt, 4)), dim3(256), st,
                           E->h2_ltiles[h] + E->h2_loff[h][j], nt, j, E->mnr, E->mnp, E->mmo, E->mvo, E->red_atom,
                           E->red_last, E->x, E->alp, E->diag, E->ured, E->Ared);
                if (nd)
                    LAUNCH(KC_BATCH_SOLVE, k_bstep_diag<double>, dim3(nd), dim3(256), st,
                           E->h2_dmols[h] + E->h2_doff[h][j], j, E->mnp, E->mmo, E->mvo, E->Ared, E->Linv, E->flags);
                if (npn)
                    LAUNCH(KC_BATCH_SOLVE, (k_bstep_tiles<double, false>), dim3(cdiv(npn, 4)), dim3(256), st,
                           E->h2_ptiles[h] + E->h2_poff[h][j], npn, j, E->mnp, E->mmo, E->mvo, E->Ared, E->Linv);
            }
        (void)hipEventRecord(E->ev_h2j, E->side2);
        (void)hipStreamWaitEvent(s, E->ev_h2j, 0);
    } else if (E->batch_solver == 5) {  // left-looking by column blocks over the batch (FP64 on both paths)
        LAUNCH(KC_BATCH_SOLVE, k_red_u, dim3(cdiv(E->np, 256)), dim3(256), s, E->np, E->red_atom, E->red_last, E->x,
               E->alp, E->diag, E->ured);
        const int ncol = (int)E->ls_off.size() - 1;
        for (int j = 0; j < ncol; ++j) {
            const int nt = E->ls_off[j + 1] - E->ls_off[j];
            const int nd = E->bs_doff[j + 1] - E->bs_doff[j];
            const int npn = E->bs_poff[j + 1] - E->bs_poff[j];
            LAUNCH(KC_BATCH_SOLVE, k_lstep_tiles, dim3(cdiv(nt, 4)), dim3(256), s, E->ls_tiles + E->ls_off[j], nt, j,
                   E->mnr, E->mnp, E->mmo, E->mvo, E->red_atom, E->red_last, E->x, E->alp, E->diag, E->ured, E->Ared);
            LAUNCH(KC_BATCH_SOLVE, k_bstep_diag<double>, dim3(nd), dim3(256), s, E->bs_dmols + E->bs_doff[j], j,
                   E->mnp, E->mmo, E->mvo, E->Ared, E->Linv, E->flags);
            if (npn)
                LAUNCH(KC_BATCH_SOLVE, (k_bstep_tiles<double, false>), dim3(cdiv(npn, 4)), dim3(256), s,
                       E->bs_ptiles + E->bs_poff[j], npn, j, E->mnp, E->mmo, E->mvo, E->Ared, E->Linv);
        }
    } else if (E->batch_solver == 6) {  // as 5, the panel product inside the tiles below the diagonal
        LAUNCH(KC_BATCH_SOLVE, k_red_u, dim3(cdiv(E->np, 256)), dim3(256), s, E->np, E->red_atom, E->red_last, E->x,
               E->alp, E->diag, E->ured);
        const int ncol = (int)E->l6d_off.size() - 1;
        for (int j = 0; j < ncol; ++j) {
            const int ndt = E->l6d_off[j + 1] - E->l6d_off[j];
            const int not_ = E->l6o_off[j + 1] - E->l6o_off[j];
            const int nd = E->bs_doff[j + 1] - E->bs_doff[j];
            LAUNCH(KC_BATCH_SOLVE, k_lstep_tiles, dim3(cdiv(ndt, 4)), dim3(256), s, E->l6d_tiles + E->l6d_off[j], ndt,
                   j, E->mnr, E->mnp, E->mmo, E->mvo, E->red_atom, E->red_last, E->x, E->alp, E->diag, E->ured,
                   E->Ared);
            LAUNCH(KC_BATCH_SOLVE, k_bstep_diag<double>, dim3(nd), dim3(256), s, E->bs_dmols + E->bs_doff[j], j,
                   E->mnp, E->mmo, E->mvo, E->Ared, E->Linv, E->flags);
            if (not_)
                LAUNCH(KC_BATCH_SOLVE, k_lpanel_tiles, dim3(cdiv(not_, 4)), dim3(256), s,
                       E->l6o_tiles + E->l6o_off[j], not_, j, E->mnr, E->mnp, E->mmo, E->mvo, E->red_atom,
                       E->red_last, E->x, E->alp, E->diag, E->ured, E->Ared, E->Linv);
        }
    } else if (E->mixed && E->batch_fp32) {
        if (E->batch_solver == 3)
            queue_batch_steps<float>(E, s, E->Ared32, E->Linv32);
        else if (E->batch_solver == 1)  // two workgroups per CU, operands four steps ahead
            LAUNCH(KC_BATCH_SOLVE, (k_eeq_batch_factor<float, 2, 4>), g, b, s, E->morder, E->mnp, E->mmo, E->mvo,
                   E->Ared32, E->Linv32, E->flags);
        else  // three workgroups per CU, operands one step ahead
            LAUNCH(KC_BATCH_SOLVE, (k_eeq_batch_factor<float, 3, 1>), g, b, s, E->morder, E->mnp, E->mmo, E->mvo,
                   E->Ared32, E->Linv32, E->flags);
    } else {
        if (E->batch_solver == 3)
            queue_batch_steps<double>(E, s, E->Ared, E->Linv);
        else if (E->batch_solver == 1)
            LAUNCH(KC_BATCH_SOLVE, (k_eeq_batch_factor<double, 2, 4>), g, b, s, E->morder, E->mnp, E->mmo, E->mvo,
                   E->Ared, E->Linv, E->flags);
        else
            LAUNCH(KC_BATCH_SOLVE, (k_eeq_batch_factor<double, 3, 1>), g, b, s, E->morder, E->mnp, E->mmo, E->mvo,
                   E->Ared, E->Linv, E->flags);
    }
}

// k_pairs_sym over the engine's tiles for the selected pair terms
static void queue_pairs_sym(DeviceEngine* E, hipStream_t s, const PairArgs& pa, double* part, int terms) {
    const dim3 grid(E->ps_ntiles);
    if (!E->mixed) {
        switch (E->nelem) {
            case 1: LAUNCH(KC_PAIRS, (k_pairs_sym<double, 1>), grid, dim3(PS_B), s, pa, part, E->ps_tiles, terms); break;
            case 2: LAUNCH(KC_PAIRS, (k_pairs_sym<double, 2>), grid, dim3(PS_B), s, pa, part, E->ps_tiles, terms); break;
            case 3: LAUNCH(KC_PAIRS, (k_pairs_sym<double, 3>), grid, dim3(PS_B), s, pa, part, E->ps_tiles, terms); break;
            case 4: LAUNCH(KC_PAIRS, (k_pairs_sym<double, 4>), grid, dim3(PS_B), s, pa, part, E->ps_tiles, terms); break;
            default: LAUNCH(KC_PAIRS, (k_pairs_sym<double, 0>), grid, dim3(PS_B), s, pa, part, E->ps_tiles, terms); break;
        }
    } else {
        switch (E->nelem) {
            case 1: LAUNCH(KC_PAIRS, (k_pairs_sym<float, 1>), grid, dim3(PS_B), s, pa, part, E->ps_tiles, terms); break;
            case 2: LAUNCH(KC_PAIRS, (k_pairs_sym<float, 2>), grid, dim3(PS_B), s, pa, part, E->ps_tiles, terms); break;
            case 3: LAUNCH(KC_PAIRS, (k_pairs_sym<float, 3>), grid, dim3(PS_B), s, pa, part, E->ps_tiles, terms); break;
            case 4: LAUNCH(KC_PAIRS, (k_pairs_sym<float, 4>), grid, dim3(PS_B), s, pa, part, E->ps_tiles, terms); break;
            default: LAUNCH(KC_PAIRS, (k_pairs_sym<float, 0>), grid, dim3(PS_B), s, pa, part, E->ps_tiles, terms); break;
        }
    }
}

static int launch_eval(DeviceEngine* E, const double* xin, hipStream_t s, std::string& msg) {
    const int n = E->n;
    hipLaunchKernelGGL(k_copy_start, dim3(cdiv(3 * n, 256)), dim3(256), 0, s, 3 * n, xin, E->x, E->flags);
    // The reduced EEQ matrix needs only the coordinates: it is built on the
    // side stream while this stream computes CN, the EEQ right-hand side and
    // every term that does not depend on the EEQ charges; the solve waits
    // for both (and runs alone, as the persistent solver requires).
    if (E->solv) {
        // Born radii, surface areas and the shape descriptor: geometry only,
        // ahead of the fork so that the reduced matrix can take the Born part
        LAUNCH(KC_SOLV, k_solv_born, dim3(cdiv(n, 4)), dim3(256), s, E->sa);
        LAUNCH(KC_SOLV, k_solv_sasa, dim3(n), dim3(256), s, E->sa);
        LAUNCH(KC_SOLV, k_solv_shape, dim3(1), dim3(256), s, E->sa);
    }
    HIP_OK(hipEventRecord(E->ev_fork, s));
    HIP_OK(hipStreamWaitEvent(E->side, E->ev_fork, 0));
    // EEQ
    const int np = E->np, nbk = np / NB;
    if (E->nmol == 1) {
        LAUNCH(KC_EEQ_BUILD, k_eeq_reduced, dim3(nbk * (nbk + 1) / 2), dim3(RED_T), E->side, E->nr, np,
               E->red_atom, E->red_last, E->red_lidx, E->red_ul, E->red_nul, E->x, E->alp, E->diag, E->Ared,
               E->use_cg ? 2 : 0);
        if (E->solv)
            LAUNCH(KC_SOLV, k_solv_red_add, dim3(nbk * (nbk + 1) / 2), dim3(256), E->side, E->nr, np, E->red_atom,
                   E->red_last, E->sa, E->Ared, E->use_cg ? 2 : 0);
        if (E->use_cg) {
            // block-Jacobi factors of R's diagonal blocks and the solver's
            // mailboxes, also off the evaluation stream
            LAUNCH(KC_CHOL_DIAG, k_bjacobi_setup, dim3(nbk), dim3(256), E->side, E->Ared, E->Linv, E->flags);
            hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(MB_WORDS, 256)), dim3(256), 0, E->side, MB_WORDS, E->cg_mb,
                               MB_EMPTY);
        }
    }
    else if (E->batch_solver < 4)  // batch_solver 4 and 5 form the matrix inside their factorisations
        LAUNCH(KC_EEQ_BUILD, k_eeq_reduced_batch, dim3(E->nbtiles), dim3(256), E->side, E->btiles, E->mnr, E->mnp,
               E->mmo, E->mvo, E->red_atom, E->red_last, E->x, E->alp, E->diag, E->Ared,
               E->Ared32);  // the FP32 copy exists only when the mixed path factors in FP32
    const bool split = E->nmol > 1 && E->batch_solver != 2;
    if (split) {
        queue_batch_factor(E, E->side);
        HIP_OK(hipEventRecord(E->ev_fact, E->side));
    }
    if (!E->mol_of) LAUNCH(KC_CN, k_tile_box, dim3(cdiv(n, 256)), dim3(256), s, n, E->x, E->tbox);
    if (E->cn_by_mol)
        LAUNCH_SHM(KC_CN, k_cn_mol, dim3(E->nmol), dim3(256), (size_t)E->cn_cap * CNMOL_LDS_BYTES, s, E->cn_cap,
                   E->mol_off, E->x, E->rcov_cn, E->elem, E->nref,
               E->refcn, E->chi_base, E->cnf, E->cn, E->dlog, E->gw, E->dgw, E->rhs, E->nbn, E->nbl, E->flags);
    else
        LAUNCH(KC_CN, k_cn, dim3(cdiv(n, 4)), dim3(256), s, n, E->x, E->rcov_cn, E->elem, E->nref, E->refcn,
               E->chi_base, E->cnf, E->cn, E->dlog, E->gw, E->dgw, E->rhs, E->nbn, E->nbl, E->flags, E->mol_of,
               E->mol_off, E->tbox);
    const double* rsrc = E->rhs;
    if (E->charged) {
        LAUNCH(KC_EEQ_BUILD, k_eeq_rhs_qp, dim3(cdiv(n, 256)), dim3(256), s, n, E->x, E->alp, E->diag, E->qp,
                           E->rhs, E->rres, E->mol_of, E->mol_off);
        if (E->solv) LAUNCH(KC_SOLV, k_solv_rhs, dim3(cdiv(n, 256)), dim3(256), s, E->sa, E->qp, E->rres);
        rsrc = E->rres;
    }
    LAUNCH(KC_EEQ_BUILD, k_eeq_rhs_reduce, dim3(cdiv(np, 256)), dim3(256), s, E->nr, np, E->red_atom, E->red_last,
                       rsrc, E->b);
    // D4 half contractions need only CN: evaluated before the solve
    LAUNCH(KC_PAIRS_AUX, k_d4_half, dim3(cdiv(n * E->nelem, 256)), dim3(256), s, n, E->nelem, E->elem, E->nref,
           E->c6ref, E->gw, E->dgw, E->h6, E->dh6);
    const PairArgs pa = E->pair_args();
    // bonded lists
    if (E->nbond) {
        if (E->nb_hb)
            LAUNCH(KC_BONDS, k_bonds, dim3(cdiv(E->nb_hb, 4)), dim3(256), s, E->nb_hb, E->bsel_hb, E->bond_ij,
                   E->bond_par, E->bond_hboff, E->bond_hbB, E->bond_slot, E->x, E->cn, E->rcov_d3, E->bond_H,
                   E->slots, E->dcn_slot, E->e_bond);
        if (E->nb_plain)
            LAUNCH(KC_BONDS, k_bonds_plain, dim3(cdiv(E->nb_plain, 256)), dim3(256), s, E->nb_plain,
                   E->bsel_plain, E->bond_ij, E->bond_par, E->bond_slot, E->bond_H, E->x, E->cn, E->slots,
                   E->dcn_slot, E->e_bond);
    }
    if (E->nangle)
        LAUNCH(KC_ANGLES, k_angles, dim3(cdiv(E->nangle, 256)), dim3(256), s, E->nangle, E->ang_ijk, E->ang_par,
                           E->ang_slot, E->x, E->rcov_d3, E->slots, E->e_ang);
    if (E->ntors)
        LAUNCH(KC_TORS, k_tors, dim3(cdiv(E->ntors, 256)), dim3(256), s, E->ntors, E->tor_ijkl, E->tor_par,
                           E->tor_slot, E->x, E->rcov_d3, E->slots, E->e_tor);
    if (E->ninv)
        LAUNCH(KC_TORS, k_inv, dim3(cdiv(E->ninv, 256)), dim3(256), s, E->ninv, E->inv_ijkl, E->inv_par, E->inv_slot,
               E->x, E->rcov_d3, E->slots, E->e_inv);
    if (E->nstors)
        LAUNCH(KC_TORS, k_stors, dim3(cdiv(E->nstors, 256)), dim3(256), s, E->nstors, E->st_ijkl, E->st_par,
               E->st_slot, E->x, E->slots, E->e_st);
    if (E->nxb)
        LAUNCH(KC_HB, k_xb, dim3(cdiv(E->nxb, 256)), dim3(256), s, E->nxb, E->xb_axb, E->xb_par, E->xb_slot, E->x,
               E->hb_rad, E->slots, E->e_xb);
    if (E->nbrep)
        LAUNCH(KC_BREP, k_brep, dim3(cdiv(E->nbrep, 256)), dim3(256), s, E->nbrep, E->brep_ij, E->brep_par,
                           E->brep_slot, E->x, E->slots, E->e_brep);
    if (E->natm)
        LAUNCH(KC_ATM, k_atm, dim3(cdiv(E->natm, 256)), dim3(256), s, E->natm, E->atm_ijk, E->atm_sc,
                           E->atm_slot, E->x, E->rcov_d3, E->nelem, E->elem, E->nref, E->c6ref, E->gw, E->slots,
                           E->e_atm);
    if (E->nbatm)
        LAUNCH(KC_BATM, k_batm, dim3(cdiv(E->nbatm, 256)), dim3(256), s, E->nbatm, E->batm_ijk, E->batm_zb,
                           E->batm_slot, E->x, E->qa, E->slots, E->e_batm);
    if (E->fin_lpa == 64)
        LAUNCH(KC_CHAIN, k_slot_gather<64>, dim3(cdiv(n, 4)), dim3(256), s, n, E->slot_off, E->slot_idx, E->slots,
               E->gslot);
    else
        LAUNCH(KC_CHAIN, k_slot_gather<16>, dim3(cdiv(n, 16)), dim3(256), s, n, E->slot_off, E->slot_idx, E->slots,
               E->gslot);
    // pair_split: the charge-independent pair terms (dispersion, repulsion)
    // last before the solve, beside the factorisation's smaller column blocks
    if (E->pair_split) queue_pairs_sym(E, s, pa, E->ps_part2, PT_DISP | PT_REP);
    // Hydrogen bonds (all kernels and their gather) need only the coordinates
    // and set-up data: a single molecule evaluates them on this stream while
    // the side stream builds the EEQ matrix; a batch appends them to the side
    // stream behind its matrix build and factorisation (run beside the
    // factorisation, they slowed it, and it is the longer chain), and they
    // join this stream before k_finish.
    hipStream_t hs = (E->nmol > 1) ? E->side : s;
    if (E->nhb) {
        HBArgs ha{E->hb_cb, E->hb_ca, E->hb_Qh, E->hb_Qab, E->hb_nboff, E->hb_nbidx, E->hb_parc,
                  E->hb_nbcoff, E->hb_nbcidx, E->hb_rad, E->hb_lpd, E->Z, E->hb_pk, E->hb_qh32};
        HBRoles hr{E->hb_ent, E->hb_goff, E->hb_gatom, E->hb_ng0, E->hb_ng1, E->hb_ngt};
        // case-1 role kernels (FP32 terms on the mixed path, FP64 otherwise)
        const int gr0[3] = {0, E->hb_ng0, E->hb_ng1}, gr1[3] = {E->hb_ng0, E->hb_ng1, E->hb_ngt};
        for (int role = 0; role < 3; ++role) {
            const dim3 grid(cdiv(gr1[role] - gr0[role], 4));
            if (gr1[role] <= gr0[role]) continue;
            if (E->mixed) {
                if (role == 0) LAUNCH(KC_HB, (k_hb_role<0, float>), grid, dim3(256), hs, hr, ha, E->x, E->hb_gout, E->e_hb);
                if (role == 1) LAUNCH(KC_HB, (k_hb_role<1, float>), grid, dim3(256), hs, hr, ha, E->x, E->hb_gout, E->e_hb);
                if (role == 2) LAUNCH(KC_HB, (k_hb_role<2, float>), grid, dim3(256), hs, hr, ha, E->x, E->hb_gout, E->e_hb);
            } else {
                if (role == 0) LAUNCH(KC_HB, (k_hb_role<0, double>), grid, dim3(256), hs, hr, ha, E->x, E->hb_gout, E->e_hb);
                if (role == 1) LAUNCH(KC_HB, (k_hb_role<1, double>), grid, dim3(256), hs, hr, ha, E->x, E->hb_gout, E->e_hb);
                if (role == 2) LAUNCH(KC_HB, (k_hb_role<2, double>), grid, dim3(256), hs, hr, ha, E->x, E->hb_gout, E->e_hb);
            }
        }
        // cases 2-4: on the mixed path the case-2 entries by the lean FP32
        // kernel and those of cases 3 and 4 by the FP64 one
        const int n34 = E->nhb2 - E->nhb2_c2;
        if (E->nhb2 && !E->mixed)
            LAUNCH(KC_HB, (k_hb_full<double, 0>), dim3(cdiv(E->nhb2, 256)), dim3(256), hs, E->nhb2, E->hb_ord,
                   E->hb_ahbc, E->hb_s0, ha, E->x, E->hb_xslots, E->e_hb2);
        if (E->mixed && E->nhb2_c2)
            LAUNCH(KC_HB, (k_hb_full<float, 1>), dim3(cdiv(E->nhb2_c2, 256)), dim3(256), hs, E->nhb2_c2, E->hb_ord,
                   E->hb_ahbc, E->hb_s0, ha, E->x, E->hb_xslots, E->e_hb2);
        if (E->mixed && n34 > 0)
            LAUNCH(KC_HB, (k_hb_full<double, 2>), dim3(cdiv(n34, 256)), dim3(256), hs, n34, E->hb_ord + E->nhb2_c2,
                   E->hb_ahbc, E->hb_s0, ha, E->x, E->hb_xslots, E->e_hb2);
    }
    if (E->hbg_lpa == 64)
        LAUNCH(KC_HB, k_hb_gather<64>, dim3(cdiv(n, 4)), dim3(256), hs, n, E->hb_crange, E->hb_gout, E->hb_xoff,
               E->hb_xidx, E->hb_xslots, E->ghb);
    else
        LAUNCH(KC_HB, k_hb_gather<16>, dim3(cdiv(n, 16)), dim3(256), hs, n, E->hb_crange, E->hb_gout, E->hb_xoff,
               E->hb_xidx, E->hb_xslots, E->ghb);
    HIP_OK(hipEventRecord(E->ev_join, E->side));
    // batches wait here for the factorisation only; the hydrogen-bond work
    // queued behind it on the side stream joins before k_finish
    HIP_OK(hipStreamWaitEvent(s, split ? E->ev_fact : E->ev_join, 0));
    if (E->nmol > 1) {
        // batch: every molecule's system factored (side stream, above) and
        // solved by one workgroup per molecule
        if (E->batch_solver == 2)
            LAUNCH(KC_BATCH_SOLVE, k_eeq_batch_solve_tiles, dim3(E->nmol), dim3(256), s, E->morder, E->mnp, E->mmo,
                   E->mvo, E->Ared, E->Linv, E->b, E->xbuf, E->flags);
        else if (E->mixed && E->batch_fp32 && E->batch_solver < 4)
            LAUNCH(KC_BATCH_SOLVE, k_bsolve_subst<float>, dim3(E->nmol), dim3(BSOLVE_T), s, E->morder, E->mnp,
                   E->mmo, E->mvo, E->Ared32, E->Linv32, E->Ared, E->b, E->xbuf);
        else
            LAUNCH(KC_BATCH_SOLVE, k_bsolve_subst<double>, dim3(E->nmol), dim3(BSOLVE_T), s, E->morder, E->mnp,
                   E->mmo, E->mvo, E->Ared, E->Linv, (const double*)nullptr, E->b, E->xbuf);
    } else if (E->use_cg) {
        // the persistent PCG (block-Jacobi factors and mailboxes were set up on the side stream)
        LAUNCH(KC_PCG, k_eeq_pcg, dim3(E->cg_groups), dim3(PCG_THREADS), s, np, E->Ared, E->Linv, E->b, E->xbuf,
               E->cg_z, E->cg_p, E->cg_part, E->cg_mb, E->cg_iters, E->cg_tol2, PCG_MAXIT, E->flags);
    } else if (!E->mixed) {
        if (E->lu_src)
            LAUNCH(KC_CHOL_DIAG, k_copy, dim3(cdiv((long)E->mat_size, 256)), dim3(256), s, (int)E->mat_size, E->Ared,
                   E->lu_src);
        queue_cholesky(&E->timer, s, np, E->Ared, E->Linv, E->flags, E->use_lookahead ? &E->la : nullptr);
        queue_solve<double>(E, s, E->Ared, E->Linv, E->b, E->ybuf, E->xbuf);
    } else {
        LAUNCH(KC_EEQ_BUILD, k_lower_to_f32, dim3(nbk * (nbk + 1) / 2), dim3(256), s, np, E->Ared, E->Ared32);
        queue_cholesky(&E->timer, s, np, E->Ared32, E->Linv32, E->flags, E->use_lookahead ? &E->la : nullptr);
        LAUNCH(KC_REFINE, k_to_f32, dim3(cdiv(np, 256)), dim3(256), s, np, E->b, E->b32);
        for (int it = 0; it < REFINE_STEPS; ++it) {
            queue_solve<float>(E, s, E->Ared32, E->Linv32, E->b32, E->y32, E->d32);
            LAUNCH(KC_REFINE, k_refine_add, dim3(cdiv(np, 256)), dim3(256), s, np, it == 0 ? 1 : 0, E->d32,
                   E->xbuf);
            if (it + 1 < REFINE_STEPS) {
                LAUNCH(KC_REFINE, k_symv_part, dim3(nbk, SYMV_SPLIT), dim3(256), s, np, SYMV_SPLIT, E->Ared,
                       E->xbuf, E->symv);
                LAUNCH(KC_REFINE, k_eeq_resid, dim3(cdiv(np, 256)), dim3(256), s, np, SYMV_SPLIT, E->symv, E->b,
                       E->rvec, E->b32);
            }
        }
    }
    if (E->lu_w)  // no work unless the system was found not positive definite
        LAUNCH(KC_CHARGES, k_eeq_lu_fallback, dim3(1), dim3(LU_T), s, np, E->nr, E->use_cg ? 0 : 1,
               E->use_cg ? E->Ared : (E->lu_src ? E->lu_src : E->Ared), E->b, E->lu_w, E->xbuf, E->flags);
    LAUNCH(KC_CHARGES, k_frag_sum, dim3(cdiv(E->nfrag, 4)), dim3(256), s, E->nfrag, E->frag_off, E->frag_atoms,
           E->red_of, E->xbuf, E->fsum);
    LAUNCH(KC_CHARGES, k_charges, dim3(cdiv(n, 256)), dim3(256), s, n, E->red_of, E->fsum, E->frag, E->qp, E->xbuf,
           E->chi_base, E->cnf, E->gam, E->alp, E->cn, E->q, E->eself, E->qtmp);
    // pair terms
    if (E->ps_ntiles > 0) {  // each unordered pair once (single molecules; batches unless batch_pairs = 1)
        queue_pairs_sym(E, s, pa, E->ps_part, E->pair_split ? PT_COUL : PT_ALL);
        LAUNCH(KC_PAIRS_AUX, k_pairs_sym_reduce, dim3(E->ps_nblocks), dim3(256), s, n, E->ps_nb, E->ps_part,
               E->pair_split ? (const double*)E->ps_part2 : nullptr,
               E->gpair, E->dEdcn, E->e_coul, E->e_disp, E->e_rep, E->ps_blocks);
    } else {
        const dim3 grid(cdiv(n, TILE), E->pair_nchunk);
        if (!E->mixed) {
            switch (E->nelem) {
                case 1: LAUNCH(KC_PAIRS, k_pairs_part<1>, grid, dim3(TILE), s, pa, E->pair_cj, E->ppart); break;
                case 2: LAUNCH(KC_PAIRS, k_pairs_part<2>, grid, dim3(TILE), s, pa, E->pair_cj, E->ppart); break;
                case 3: LAUNCH(KC_PAIRS, k_pairs_part<3>, grid, dim3(TILE), s, pa, E->pair_cj, E->ppart); break;
                case 4: LAUNCH(KC_PAIRS, k_pairs_part<4>, grid, dim3(TILE), s, pa, E->pair_cj, E->ppart); break;
                default: LAUNCH(KC_PAIRS, k_pairs_part<0>, grid, dim3(TILE), s, pa, E->pair_cj, E->ppart); break;
            }
        } else {
            switch (E->nelem) {
                case 1: LAUNCH(KC_PAIRS, k_pairs_part32<1>, grid, dim3(TILE), s, pa, E->pair_cj, E->ppart); break;
                case 2: LAUNCH(KC_PAIRS, k_pairs_part32<2>, grid, dim3(TILE), s, pa, E->pair_cj, E->ppart); break;
                case 3: LAUNCH(KC_PAIRS, k_pairs_part32<3>, grid, dim3(TILE), s, pa, E->pair_cj, E->ppart); break;
                case 4: LAUNCH(KC_PAIRS, k_pairs_part32<4>, grid, dim3(TILE), s, pa, E->pair_cj, E->ppart); break;
                default: LAUNCH(KC_PAIRS, k_pairs_part32<0>, grid, dim3(TILE), s, pa, E->pair_cj, E->ppart); break;
            }
        }
        LAUNCH(KC_PAIRS_AUX, k_pairs_reduce, dim3(cdiv(n, 256)), dim3(256), s, n, E->pair_nchunk, E->ppart, E->gpair,
               E->dEdcn, E->e_coul, E->e_disp, E->e_rep);
    }
    {
        const AtomTailArgs ta{E->x,    E->Z,       E->metal,   E->anb,      E->repz,  E->toff,
                              E->tidx, E->tdist,   E->gpair,   E->e_rep,    E->dcn_off, E->dcn_idx,
                              E->dcn_slot, E->dEdcn, E->qtmp, E->vchain};
        LAUNCH(KC_CHAIN, k_atom_tail, dim3(cdiv(n, 256)), dim3(256), s, n, ta);
    }
    if (split) HIP_OK(hipStreamWaitEvent(s, E->ev_join, 0));
    if (E->fin_lpa == 64)
        LAUNCH(KC_CHAIN, k_finish<64>, dim3(cdiv(n, 4)), dim3(256), s, n, E->x, E->rcov_cn, E->dlog, E->vchain,
                       E->nbn, E->nbl, E->gslot, E->gpair, E->ghb, E->grad);
    else
        LAUNCH(KC_CHAIN, k_finish<16>, dim3(cdiv(n, 16)), dim3(256), s, n, E->x, E->rcov_cn, E->dlog, E->vchain,
                       E->nbn, E->nbl, E->gslot, E->gpair, E->ghb, E->grad);
    if (E->solv) {
        LAUNCH(KC_SOLV, k_solv_pairs, dim3(cdiv(n, 4)), dim3(256), s, E->sa, E->q);
        LAUNCH(KC_SOLV, k_solv_energy, dim3(1), dim3(256), s, E->sa, E->q);
        LAUNCH(KC_SOLV, k_solv_grad, dim3(cdiv(n, 4)), dim3(256), s, E->sa, E->grad);
    }
    Segs sg;
    const double* ptrs[NPARTS] = {E->e_bond, E->e_ang,  E->e_tor, E->e_inv, E->e_disp, E->e_brep,
                                  E->e_rep,  E->e_coul, E->eself, E->e_hb,  E->e_atm,  E->e_batm,
                                  E->e_st,   E->e_xb,   E->e_hb2, E->solv ? E->sa.e_solv : E->e_bond};
    int cnts[NPARTS] = {E->nbond, E->nangle, E->ntors, E->ninv, n, E->nbrep, n, n, n, E->hb_ng0, E->natm, E->nbatm,
                        E->nstors, E->nxb, E->nhb2, E->solv ? 1 : 0};
    for (int k = 0; k < NPARTS; ++k) {
        sg.p[k] = ptrs[k];
        sg.n[k] = cnts[k];
        sg.dst[k] = k;
    }
    LAUNCH(KC_REDUCE, k_reduce1, dim3(RED_BLOCKS, NPARTS), dim3(256), s, sg, E->rpart);
    LAUNCH(KC_REDUCE, k_reduce_final, dim3(1), dim3(NPARTS * 64), s, sg, E->rpart, E->parts, E->out, E->flags,
           E->sticky);
    if (E->nmol > 1) {
        MolEnergyArgs ma;
        ma.nmol = E->nmol;
        ma.mol_off = E->mol_off;
        ma.seg = E->mol_seg;
        ma.atom_terms[0] = E->e_coul;
        ma.atom_terms[1] = E->e_disp;
        ma.atom_terms[2] = E->e_rep;
        ma.atom_terms[3] = E->eself;
        const double* lists[11] = {E->e_bond, E->e_ang,  E->e_tor, E->e_brep, E->e_hb, E->e_atm,
                                   E->e_batm, E->e_inv, E->e_st,  E->e_xb,   E->e_hb2};
        for (int k = 0; k < 11; ++k) ma.list_terms[k] = lists[k];
        ma.e_mol = E->e_mol;
        LAUNCH(KC_REDUCE, k_mol_energy, dim3(cdiv(E->nmol, 4)), dim3(256), s, ma);
    }
    if (hipGetLastError() != hipSuccess) {
        msg = "kernel launch failed";
        return kErrDevice;
    }
    return kOk;
}

int device_eval_host(DeviceEngine* E, const double* xyz_bohr, bool want_grad, double* e_terms, double* grad,
                     double* q, double* cn, std::string& msg) {
    const int n = E->n;
    HIP_OK(hipMemcpyAsync(E->gpair, xyz_bohr, 3 * sizeof(double) * n, hipMemcpyHostToDevice, E->own));
    // gpair is overwritten by k_pairs_reduce after the copy kernel has consumed it
    int st = launch_eval(E, E->gpair, E->own, msg);
    if (st) return st;
    int flags = 0;
    HIP_OK(hipMemcpyAsync(e_terms, E->out, sizeof(double) * (GFNFF_NTERMS + 1), hipMemcpyDeviceToHost, E->own));
    HIP_OK(hipMemcpyAsync(&flags, E->flags, sizeof(int), hipMemcpyDeviceToHost, E->own));
    if (want_grad && grad) HIP_OK(hipMemcpyAsync(grad, E->grad, 3 * sizeof(double) * n, hipMemcpyDeviceToHost, E->own));
    if (q) HIP_OK(hipMemcpyAsync(q, E->q, sizeof(double) * n, hipMemcpyDeviceToHost, E->own));
    if (cn) HIP_OK(hipMemcpyAsync(cn, E->cn, sizeof(double) * n, hipMemcpyDeviceToHost, E->own));
    HIP_OK(hipStreamSynchronize(E->own));
    if (flags & FLAG_CHOL) {
        msg = "EEQ matrix is not positive definite on the constraint space";
        return kErrEEQ;
    }
    if (flags & FLAG_CNOVF) {
        msg = "more than 128 atoms within the 6 Bohr CN cut-off of one atom";
        return kErrGeometry;
    }
    if (flags & FLAG_SOLVE) {
        msg = "EEQ triangular solve did not complete (a block waited beyond its bound)";
        return kErrDevice;
    }
    return kOk;
}

int device_eval(DeviceEngine* E, const double* xyz_bohr_dev, double* energy_dev, double* grad_dev, void* stream,
                std::string& msg) {
    hipStream_t s = (hipStream_t)stream;
    int st = launch_eval(E, xyz_bohr_dev, s, msg);
    if (st) return st;
    hipLaunchKernelGGL(k_copy, dim3(1), dim3(64), 0, s, GFNFF_NTERMS + 1, E->out, energy_dev);
    if (grad_dev) hipLaunchKernelGGL(k_copy, dim3(cdiv(3 * E->n, 256)), dim3(256), 0, s, 3 * E->n, E->grad, grad_dev);
    if (hipGetLastError() != hipSuccess) {
        msg = "kernel launch failed";
        return kErrDevice;
    }
    return kOk;
}

// Records one device-resident evaluation with fixed buffers into a HIP graph
// (both streams: the fork and join events become graph dependencies) and
// instantiates it; nothing runs during the capture.  The evaluation queues
// no memset or host synchronisation, and the kernel-class timer is off
// while recording.
int device_graph_capture(DeviceEngine* E, const double* xyz_bohr_dev, double* energy_dev, double* grad_dev,
                         void* stream, std::string& msg) {
    hipStream_t s = (hipStream_t)stream;
    if (!s) {
        msg = "graph capture needs a created stream (not the default stream)";
        return kErrArgs;
    }
    if (E->batch_streams == 2) {
        // its capture ended in a crash inside the runtime whose cause is not
        // established: the two-stream factorisation runs eagerly only
        msg = "graph capture is not available with batch_streams 2";
        return kErrArgs;
    }
    E->drop_graph();
    const int cls = E->timer.cls;
    E->timer.cls = -1;
    HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int st = device_eval(E, xyz_bohr_dev, energy_dev, grad_dev, stream, msg);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(s, &g);
    E->timer.cls = cls;
    if (st != kOk || ce != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        if (st == kOk) msg = std::string("hipStreamEndCapture: ") + hipGetErrorString(ce);
        return st != kOk ? st : kErrDevice;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        msg = std::string("hipGraphInstantiate: ") + hipGetErrorString(ie);
        return kErrDevice;
    }
    E->graph = g;
    E->gexec = ex;
    return kOk;
}

int device_graph_launch(DeviceEngine* E, void* stream, std::string& msg) {
    if (!E->gexec) {
        msg = "no captured evaluation (the engine was rebuilt, or gfnff_graph_capture was not called)";
        return kErrArgs;
    }
    HIP_OK(hipGraphLaunch(E->gexec, (hipStream_t)stream));
    return kOk;
}

int device_md_steps(DeviceEngine* E, double* x_dev, double* v_dev, double* grad_dev, const double* mass_amu,
                    const double* xref_bohr, double dt_fs, double T0, double tau_fs, double disp_thresh, int nsteps,
                    double* rec_dev, void* stream, std::string& msg) {
    const int n = E->n, n3 = 3 * n;
    hipStream_t s = (hipStream_t)stream;
    constexpr double kAmuToMe = 1822.8884850;    // electron masses per amu (global.h:46)
    constexpr double kFsToAu = 41.34137314;      // atomic time units per fs (global.h:50)
    constexpr double kBoltzEh = 3.166811e-6;     // Hartree per Kelvin (global.h:47)
    if (!E->md_im3) {
        try {
            E->md_im3 = E->alloc<double>(n3);
            E->md_m3 = E->alloc<double>(n3);
            E->md_xref = E->alloc<double>(n3);
            E->md_part = E->alloc<double>(2 * (size_t)cdiv(n3, MD_T));
            E->md_e = E->alloc<double>(GFNFF_NTERMS + 1);
            E->md_lambda = E->alloc<double>(1);
        } catch (const std::bad_alloc&) {
            msg = "device allocation failed";
            return kErrDevice;
        }
    }
    std::vector<double> m3(n3), im3(n3);
    for (int i = 0; i < n; ++i) {
        if (!(mass_amu[i] > 0.0)) {
            msg = "masses must be positive";
            return kErrArgs;
        }
        for (int c = 0; c < 3; ++c) {
            m3[3 * i + c] = mass_amu[i] * kAmuToMe;
            im3[3 * i + c] = 1.0 / (mass_amu[i] * kAmuToMe);
        }
    }
    // once per call (not per step): masses and the coordinates of the last HB/XB detection
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(E->md_m3, m3.data(), sizeof(double) * n3, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(E->md_im3, im3.data(), sizeof(double) * n3, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(E->md_xref, xref_bohr, sizeof(double) * n3, hipMemcpyHostToDevice));
    const double dt = dt_fs * kFsToAu, tau = tau_fs * kFsToAu;
    const int dof = n > 2 ? 3 * n - 6 : (n == 2 ? 1 : 3);
    const double disp2_max = (disp_thresh * n) * (disp_thresh * n);
    const int nb = cdiv(n3, MD_T);
    for (int step = 0; step < nsteps; ++step) {
        hipLaunchKernelGGL(k_md_drift, dim3(nb), dim3(MD_T), 0, s, n3, dt, E->md_im3, grad_dev, x_dev, v_dev);
        const int st = device_eval(E, x_dev, E->md_e, grad_dev, stream, msg);
        if (st) return st;
        hipLaunchKernelGGL(k_md_kick, dim3(nb), dim3(MD_T), 0, s, n3, dt, E->md_im3, E->md_m3, grad_dev, x_dev,
                           E->md_xref, v_dev, E->md_part);
        hipLaunchKernelGGL(k_md_finish, dim3(1), dim3(64), 0, s, nb, E->md_part, E->md_e, kBoltzEh * dof, dt, T0, tau,
                           disp2_max, E->md_lambda, rec_dev + 4 * (size_t)step, E->sticky);
        if (T0 > 0.0) hipLaunchKernelGGL(k_md_scale, dim3(nb), dim3(MD_T), 0, s, n3, E->md_lambda, v_dev);
    }
    HIP_OK(hipGetLastError());
    return kOk;
}

int device_status(DeviceEngine* E, int* flags, bool clear, std::string& msg) {
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(flags, E->sticky, sizeof(int), hipMemcpyDeviceToHost));
    if (clear) {
        hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, E->own, 1, E->sticky, 0);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(E->own));
    }
    return kOk;
}

int device_solver_iterations(DeviceEngine* E) {
    if (!E->use_cg) return 0;
    int it = -1;
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&it, E->cg_iters, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    return it;
}

int device_cn_neighbours(DeviceEngine* E, int* counts, int* idx, int maxnb, std::string& msg) {
    const int n = E->n;
    std::vector<int> nbl((size_t)n * MAXNB_CN);
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(counts, E->nbn, sizeof(int) * n, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(nbl.data(), E->nbl, sizeof(int) * nbl.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < maxnb; ++k)
            idx[(size_t)i * maxnb + k] = (k < counts[i] && k < MAXNB_CN) ? nbl[(size_t)i * MAXNB_CN + k] : -1;
    return kOk;
}

int device_batch_energies(DeviceEngine* E, double* e_mol, bool device_dst, void* stream, std::string& msg) {
    if (E->nmol < 2) {
        msg = "not a batch";
        return kErrArgs;
    }
    if (device_dst) {
        hipLaunchKernelGGL(k_copy, dim3(cdiv(E->nmol, 256)), dim3(256), 0, (hipStream_t)stream, E->nmol, E->e_mol,
                           e_mol);
        HIP_OK(hipGetLastError());
    } else {
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipMemcpy(e_mol, E->e_mol, sizeof(double) * E->nmol, hipMemcpyDeviceToHost));
    }
    return kOk;
}

int device_update_params(DeviceEngine* E, const ParamSet& P, std::string& msg) {
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(E->szr, P.sqrt_zr4r2.data(), sizeof(double) * P.n, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(E->anb, P.rep_anb.data(), sizeof(double) * P.n, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(E->repz, P.repz.data(), sizeof(double) * P.n, hipMemcpyHostToDevice));
    if (!P.brep_par.empty())
        HIP_OK(hipMemcpy(E->brep_par, P.brep_par.data(), sizeof(double) * P.brep_par.size(),
                         hipMemcpyHostToDevice));
    return kOk;
}

template <int NE>
static void queue_adjoint(DeviceEngine* E, hipStream_t s, const PairArgs& pa, double* part, double* bpart) {
    const int n = E->n;
    const dim3 grid(cdiv(n, TILE), E->pair_nchunk);
    hipLaunchKernelGGL(k_adj_pairs<NE>, grid, dim3(TILE), 0, s, pa, E->adj_w, E->adj_u, E->pair_cj, part);
    hipLaunchKernelGGL(k_adj_nbrep_fix<NE>, dim3(cdiv(n, 256)), dim3(256), 0, s, n, E->x, E->Z, E->metal, E->elem,
                       E->anb, E->repz, E->toff, E->tidx, E->tdist, E->adj_w,
                       part + (size_t)E->pair_nchunk * ADJ_P * NE * n);
    if (E->nbrep)
        hipLaunchKernelGGL(k_adj_brep<NE>, dim3(cdiv(E->nbrep, 256)), dim3(256), 0, s, E->nbrep, E->brep_ij,
                           E->brep_par, E->elem, E->x, E->adj_w, bpart);
    hipLaunchKernelGGL(k_adj_reduce, dim3(ADJ_P * NE), dim3(64), 0, s, n, E->pair_nchunk + 1, part, E->nbrep, bpart,
                       ADJ_P * NE, E->adj_out);
}

// dLoss/dln(theta) = sum_i w_i . dg_i/dln(theta) at the state of the last
// evaluation (coordinates, charges, CN and D4 weights on the device).
int device_param_adjoint(DeviceEngine* E, const double* w, double* out, std::string& msg) {
    const int n = E->n, ne = E->nelem;
    if (ne < 1 || ne > 4) {
        msg = "the parameter adjoint is built for up to 4 elements";
        return kErrUnsupported;
    }
    hipStream_t s = E->own;
    HIP_OK(hipDeviceSynchronize());
    if (!E->adj_w) {
        try {
            E->adj_w = E->alloc<double>(3 * (size_t)n);
            E->adj_u = E->alloc<double>(n);
            E->adj_part = E->alloc<double>((size_t)(E->pair_nchunk + 1) * ADJ_P * ne * n);
            E->adj_bpart = E->alloc<double>((size_t)ADJ_P * ne * std::max(E->nbrep, 1));
            E->adj_out = E->alloc<double>(ADJ_P * ne);
        } catch (const std::bad_alloc&) {
            msg = "device allocation failed";
            return kErrDevice;
        }
    }
    HIP_OK(hipMemcpyAsync(E->adj_w, w, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_adj_u, dim3(cdiv(n, 256)), dim3(256), 0, s, n, E->x, E->rcov_cn, E->dlog, E->nbn, E->nbl,
                       E->adj_w, E->adj_u);
    PairArgs pa = E->pair_args();
    switch (ne) {
        case 1: queue_adjoint<1>(E, s, pa, E->adj_part, E->adj_bpart); break;
        case 2: queue_adjoint<2>(E, s, pa, E->adj_part, E->adj_bpart); break;
        case 3: queue_adjoint<3>(E, s, pa, E->adj_part, E->adj_bpart); break;
        default: queue_adjoint<4>(E, s, pa, E->adj_part, E->adj_bpart); break;
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, E->adj_out, sizeof(double) * ADJ_P * ne, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return kOk;
}

int device_profile_select(DeviceEngine* E, const char* name, std::string& msg) {
    E->timer.used = 0;
    E->timer.cls = -1;
    if (!name || !*name) return kOk;
    for (int c = 0; c < KC_COUNT; ++c)
        if (std::strcmp(name, kKernelClassNames[c]) == 0) {
            E->timer.cls = c;
            return kOk;
        }
    msg = std::string("unknown kernel class: ") + name;
    return kErrArgs;
}

int device_profile_read(DeviceEngine* E, double* total_ms, int* launches, std::string& msg) {
    KernelTimer& T = E->timer;
    *total_ms = 0.0;
    *launches = 0;
    for (size_t i = 0; i + 1 < T.used; i += 2) {
        float ms = 0.0f;
        HIP_OK(hipEventSynchronize(T.ev[i + 1]));  // launches of a class may sit on different streams
        HIP_OK(hipEventElapsedTime(&ms, T.ev[i], T.ev[i + 1]));
        *total_ms += ms;
        ++*launches;
    }
    T.used = 0;
    return kOk;
}

int device_solvation_parts(DeviceEngine* E, double* parts, std::string& msg) {
    if (!E->solv) {
        msg = "no solvent";
        return kErrArgs;
    }
    if (hipMemcpy(parts, E->sa.parts, 4 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        msg = "download failed";
        return kErrDevice;
    }
    return kOk;
}

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// Blocked Cholesky of a dense SPD matrix (the EEQ factorisation kernels on an
// arbitrary input).  a: n x n row-major; l_out receives the lower factor.
template <typename T>
static int dense_cholesky(int n, const T* a, T* l_out, std::string& msg) {
    if (n <= 0) {
        msg = "n must be positive";
        return kErrArgs;
    }
    const int np = cdiv(n, NB) * NB;
    std::vector<T> h((size_t)np * np, 0.0);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) h[(size_t)i * np + j] = (i < n && j < n) ? a[(size_t)i * n + j] : (i == j ? T(1) : T(0));
    T *dA = nullptr, *dL = nullptr;
    int* dflag = nullptr;
    hipStream_t s = nullptr;
    int st = kOk;
    auto cleanup = [&]() {
        if (s) (void)hipStreamDestroy(s);
        if (dA) (void)hipFree(dA);
        if (dL) (void)hipFree(dL);
        if (dflag) (void)hipFree(dflag);
    };
    if (hipStreamCreate(&s) != hipSuccess || hipMalloc(&dA, h.size() * sizeof(T)) != hipSuccess ||
        hipMalloc(&dL, (size_t)np * NB * sizeof(T)) != hipSuccess || hipMalloc(&dflag, sizeof(int)) != hipSuccess) {
        cleanup();
        msg = "device allocation failed";
        return kErrDevice;
    }
    int zero = 0, flag = 0;
    (void)hipMemcpyAsync(dA, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(dflag, &zero, sizeof(int), hipMemcpyHostToDevice, s);
    queue_cholesky(nullptr, s, np, dA, dL, dflag);
    (void)hipMemcpyAsync(h.data(), dA, h.size() * sizeof(T), hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) {
        msg = "kernel failure";
        st = kErrDevice;
    } else if (flag & FLAG_CHOL) {
        msg = "matrix is not positive definite";
        st = kErrEEQ;
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) l_out[(size_t)i * n + j] = (j <= i) ? h[(size_t)i * np + j] : T(0);
    cleanup();
    return st;
}

int device_dense_cholesky(int n, const double* a, double* l_out, std::string& msg) {
    return dense_cholesky<double>(n, a, l_out, msg);
}

int device_dense_cholesky_f32(int n, const float* a, float* l_out, std::string& msg) {
    return dense_cholesky<float>(n, a, l_out, msg);
}

// ---------------------------------------------------------------------------
// Allocation-free variant of the dense Cholesky: the padded working matrix,
// the inverses of its diagonal blocks and the status word sit in one
// file-scope device array (8.5 MB on every device the library runs on, held
// until the process ends).  Host code never uses the array's name as an
// address: the device address comes from hipGetSymbolAddress, taken once per
// device at the first call there, and the kernels receive that address.
// Calls that overlap on one device share the array and race.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_MAX_N = 1024;
constexpr size_t NOALLOC_MAT = (size_t)NOALLOC_MAX_N * NOALLOC_MAX_N;  // padded np x np matrix
constexpr size_t NOALLOC_LINV = (size_t)NOALLOC_MAX_N * NB;           // np / 64 inverses of 64 x 64 blocks
constexpr size_t NOALLOC_DOUBLES = NOALLOC_MAT + NOALLOC_LINV + 32;    // + status word (256-byte line)
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
// A = a padded to np x np with the identity; clears the status word (a store,
// not a memset, so that the sequence can be captured in a graph).
__global__ __launch_bounds__(256) void k_noalloc_pad(int n, int np, const double* __restrict__ a,
                                                     double* __restrict__ A, int* __restrict__ flag) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e == 0) flag[0] = 0;
    if (e >= (size_t)np * np) return;
    const int i = (int)(e / np), j = (int)(e % np);
    A[e] = (i < n && j < n) ? a[(size_t)i * n + j] : (i == j ? 1.0 : 0.0);
}

// l = lower triangle of the factored A (zeros above the diagonal); status
// GFNFF_STATUS_EEQ (3) when a pivot was not positive, else 0.
__global__ __launch_bounds__(256) void k_noalloc_out(int n, int np, const double* __restrict__ A,
                                                     double* __restrict__ l, const int* __restrict__ flag,
                                                     int* __restrict__ status) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e == 0) status[0] = (flag[0] & FLAG_CHOL) ? kErrEEQ : kOk;
    if (e >= (size_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    l[e] = (j <= i) ? A[(size_t)i * np + j] : 0.0;
}

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_base[NOALLOC_MAX_DEVICES] = {};
}  // namespace

int device_dense_cholesky_noalloc(int n, const double* a_dev, double* l_dev, int* status_dev, void* stream,
                                  std::string& msg) {
    if (n <= 0 || n > NOALLOC_MAX_N) {
        msg = "n must be in 1.." + std::to_string(NOALLOC_MAX_N);
        return kErrArgs;
    }
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        msg = "device index out of range";
        return kErrArgs;
    }
    double* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (!g_noalloc_base[dev]) {
            void* p = nullptr;
            HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            if (!p) {
                msg = "no device address for the scratch array";
                return kErrDevice;
            }
            g_noalloc_base[dev] = static_cast<double*>(p);
        }
        base = g_noalloc_base[dev];
    }
    const int np = cdiv(n, NB) * NB;
    double* A = base;
    double* Linv = base + NOALLOC_MAT;
    int* flag = reinterpret_cast<int*>(base + NOALLOC_MAT + NOALLOC_LINV);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_noalloc_pad, dim3(cdiv((long)np * np, 256)), dim3(256), 0, s, n, np, a_dev, A, flag);
    queue_cholesky<double>(nullptr, s, np, A, Linv, flag);
    hipLaunchKernelGGL(k_noalloc_out, dim3(cdiv((long)n * n, 256)), dim3(256), 0, s, n, np, A, l_dev, flag,
                       status_dev);
    HIP_OK(hipGetLastError());
    return kOk;
}

}  // namespace gfnff
