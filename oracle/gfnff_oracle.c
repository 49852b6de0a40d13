/*
 * gfnff_oracle.c -- CPU FP64 restatement of the reference's GFN-FF
 * per-evaluation path.  TEST INFRASTRUCTURE ONLY: it is the checker that the
 * HIP engine is compared against (tests/, __graft_entry__.smoke() and the
 * bench.py cpu_baseline leg).  Nothing in the shipped engine links or calls it.
 *
 * Input is the engine's flat parameter set (include/gfnff_params.h), built by
 * the engine's host-side topology generator; this file restates only the
 * per-evaluation arithmetic:
 *   CN + dlogdcn     cn_calculator.cpp:367-428, gfnff_method.cpp:5837-6258
 *   D4 weights       d4param_generator.cpp:1079-1213, 1692-1768, 1770-1904
 *   EEQ Phase 2      eeq_solver.cpp:3201-4006 (+ Schur solve :2087-2240)
 *   term kernels     ff_workspace_gfnff.cpp:46-1739
 *   self energy and CN chain rule   ff_workspace.cpp:415-536
 *   ALPB solvation   alpb_solvation.cpp:410-1300, eeq_solver.cpp:3774-3784
 * Units: Bohr, Hartree.  Every function cites the reference lines it follows.
 *
 * Parity status per function (see tests/test_oracle_fixtures.py): pinned by the
 * reference's Fortran-derived fixtures (test_cases/reference_data/<name>.ref.json)
 * for total and per-term energies, charges and per-term gradients; gradients
 * additionally by central finite differences of this file's own energy.
 * ALPB: total energies pinned by the reference's committed xtb GFN-FF+ALPB
 * energies (tests/test_alpb.py), the angular grid by exact sphere moments,
 * the solvation gradient by central finite differences.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/gfnff_params.h"

#ifdef _OPENMP
#include <omp.h>
#define THREAD_ID() omp_get_thread_num()
#else
#define THREAD_ID() 0
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define KN (-7.5)
#define CNMAX 4.4
#define SQRTPI 1.77245385091
#define TSQRT2PI 0.797884560802866

typedef struct {
    double e[GFNFF_NTERMS];
} terms_t;

/* per-component gradients (optional), indices */
enum { GB = 0, GA, GT, GR, GRN, GC, GD, GH, GX, GBA, GAT, NGC };

static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void sub3(const double* a, const double* b, double* o) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
static inline void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double norm3(const double* a) { return sqrt(dot3(a, a)); }

typedef struct {
    int n;
    double* g;           /* total gradient [3n] */
    double* gc[NGC];     /* optional component gradients */
} gacc_t;

static inline void gadd(gacc_t* G, int comp, int atom, double x, double y, double z) {
    G->g[3 * atom] += x;
    G->g[3 * atom + 1] += y;
    G->g[3 * atom + 2] += z;
    if (G->gc[comp]) {
        G->gc[comp][3 * atom] += x;
        G->gc[comp][3 * atom + 1] += y;
        G->gc[comp][3 * atom + 2] += z;
    }
}

/* ------------------------------------------------------------------------- */
/* CN with 6 Bohr neighbour list: calculateGFNFFCNWithNeighbors,             */
/* cn_calculator.cpp:367-428 (strict d^2 < cut^2)                            */
/* ------------------------------------------------------------------------- */
static void oracle_cn(const gfnff_params_view* P, const double* x, double cut, double* cn_raw, double* cn,
                      double* dlogdcn, int nthr) {
    int n = P->natoms;
    double cut2 = cut * cut;
    memset(cn_raw, 0, sizeof(double) * n);
    if (nthr > 1) {
        /* threaded: striped rows, each atom sums its own row in ascending j */
#pragma omp parallel for schedule(static, 1) num_threads(nthr)
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) {
                if (j == i) continue;
                double d[3];
                sub3(x + 3 * i, x + 3 * j, d);
                double d2 = dot3(d, d);
                if (!(d2 < cut2)) continue;
                double r = sqrt(d2);
                double rij = P->rcov_cn[i] + P->rcov_cn[j];
                s += 0.5 * (1.0 + erf(KN * (r - rij) / rij));
            }
            cn_raw[i] = s;
        }
    } else
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double d[3];
            sub3(x + 3 * i, x + 3 * j, d);
            double d2 = dot3(d, d);
            if (!(d2 < cut2)) continue;
            double r = sqrt(d2);
            double rij = P->rcov_cn[i] + P->rcov_cn[j];
            double v = 0.5 * (1.0 + erf(KN * (r - rij) / rij));
            cn_raw[i] += v;
            cn_raw[j] += v;
        }
    for (int i = 0; i < n; ++i) {
        cn[i] = log(1.0 + exp(CNMAX)) - log(1.0 + exp(CNMAX - cn_raw[i]));
        dlogdcn[i] = exp(CNMAX) / (exp(CNMAX) + exp(cn_raw[i]));
    }
}

/* ------------------------------------------------------------------------- */
/* D4 Gaussian weights and derivatives: d4param_generator.cpp:1079-1213 and   */
/* :1692-1768 (wf = 4, CN-only weights)                                       */
/* ------------------------------------------------------------------------- */
static void oracle_gw(const gfnff_params_view* P, const double* cn, double* gw, double* dgw) {
    for (int i = 0; i < P->natoms; ++i) {
        int e = P->elem[i], nr = P->nref[e];
        double ex[GFNFF_MAXREF], dex[GFNFF_MAXREF], norm = 0.0, dnorm = 0.0;
        for (int r = 0; r < GFNFF_MAXREF; ++r) {
            gw[i * 7 + r] = 0.0;
            dgw[i * 7 + r] = 0.0;
        }
        for (int r = 0; r < nr; ++r) {
            double df = cn[i] - P->refcn[e * 7 + r];
            ex[r] = exp(-4.0 * df * df);
            dex[r] = -8.0 * df * ex[r];
            norm += ex[r];
            dnorm += dex[r];
        }
        if (norm > 1e-10) {
            for (int r = 0; r < nr; ++r) {
                gw[i * 7 + r] = ex[r] / norm;
                dgw[i * 7 + r] = (dex[r] * norm - ex[r] * dnorm) / (norm * norm);
            }
        } else if (nr > 0) {
            gw[i * 7] = 1.0;
        }
    }
}

/* C6 and dC6/dCN for a pair from the reference C6 grid (getChargeWeightedC6
 * d4param_generator.cpp:1325-1420, computeDC6DCN :1770-1904). */
static void oracle_c6(const gfnff_params_view* P, const double* gw, const double* dgw, int i, int j, double* c6,
                      double* dci, double* dcj) {
    int ei = P->elem[i], ej = P->elem[j], ne = P->nelem;
    const double* c = P->c6ref + ((size_t)ei * ne + ej) * 49;
    double s = 0.0, si = 0.0, sj = 0.0;
    for (int a = 0; a < P->nref[ei]; ++a)
        for (int b = 0; b < P->nref[ej]; ++b) {
            double r = c[a * 7 + b];
            s += gw[i * 7 + a] * gw[j * 7 + b] * r;
            si += dgw[i * 7 + a] * gw[j * 7 + b] * r;
            sj += gw[i * 7 + a] * dgw[j * 7 + b] * r;
        }
    *c6 = s;
    *dci = si;
    *dcj = sj;
}

/* ------------------------------------------------------------------------- */
/* EEQ Phase 2: calculateFinalCharges eeq_solver.cpp:3201-4006; bordered      */
/* system solved by Cholesky of A and the Schur complement over fragments    */
/* (solveWithSchurCholesky :2087-2240).                                       */
/* ------------------------------------------------------------------------- */
static int chol(int n, double* A, int nthr) {
    for (int k = 0; k < n; ++k) {
        double* Ak = A + (size_t)k * n;
        double d = Ak[k];
        for (int p = 0; p < k; ++p) d -= Ak[p] * Ak[p];
        if (!(d > 0.0)) return 0;
        d = sqrt(d);
        Ak[k] = d;
        /* the rows below the pivot are independent: threaded over rows */
#pragma omp parallel for schedule(static) num_threads(nthr) if (nthr > 1 && n - k > 128)
        for (int i = k + 1; i < n; ++i) {
            double* Ai = A + (size_t)i * n;
            double s = Ai[k];
            for (int p = 0; p < k; ++p) s -= Ai[p] * Ak[p];
            Ai[k] = s / d;
        }
    }
    return 1;
}
static void chol_solve(int n, const double* L, double* b) {
    for (int i = 0; i < n; ++i) {
        const double* Li = L + (size_t)i * n;
        double s = b[i];
        for (int p = 0; p < i; ++p) s -= Li[p] * b[p];
        b[i] = s / Li[i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int p = i + 1; p < n; ++p) s -= L[(size_t)p * n + i] * b[p];
        b[i] = s / L[(size_t)i * n + i];
    }
}
static int lu_solve(int m, double* M, double* b) {
    for (int k = 0; k < m; ++k) {
        int p = k;
        double best = fabs(M[(size_t)k * m + k]);
        for (int i = k + 1; i < m; ++i)
            if (fabs(M[(size_t)i * m + k]) > best) {
                best = fabs(M[(size_t)i * m + k]);
                p = i;
            }
        if (best == 0.0) return 0;
        if (p != k) {
            for (int j = 0; j < m; ++j) {
                double t = M[(size_t)k * m + j];
                M[(size_t)k * m + j] = M[(size_t)p * m + j];
                M[(size_t)p * m + j] = t;
            }
            double t = b[k];
            b[k] = b[p];
            b[p] = t;
        }
        for (int i = k + 1; i < m; ++i) {
            double f = M[(size_t)i * m + k] / M[(size_t)k * m + k];
            for (int j = k + 1; j < m; ++j) M[(size_t)i * m + j] -= f * M[(size_t)k * m + j];
            b[i] -= f * b[k];
        }
    }
    for (int i = m - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < m; ++j) s -= M[(size_t)i * m + j] * b[j];
        b[i] = s / M[(size_t)i * m + i];
    }
    return 1;
}

/* B: optional n x n reaction-field matrix added to A before the solve
 * (eeq_solver.cpp:3774-3784); NULL in the gas phase. */
static int oracle_eeq(const gfnff_params_view* P, const double* x, const double* cn, double* q, int nthr,
                      const double* B) {
    int n = P->natoms, nf = P->nfrag;
    double* A = (double*)malloc(sizeof(double) * (size_t)n * n);
    double* rhs = (double*)malloc(sizeof(double) * n);
    double* Z2 = (double*)calloc((size_t)nf * n, sizeof(double));
    double* S = (double*)calloc((size_t)nf * nf, sizeof(double));
    double* lam = (double*)calloc(nf, sizeof(double));
    int* fo = (int*)calloc(nf + 1, sizeof(int));  /* fragment atom lists, ascending (CSR) */
    int* fa = (int*)calloc(n > 0 ? n : 1, sizeof(int));
    int ok = 0, coincident = 0;
    if (!A || !rhs || !Z2 || !S || !lam || !fo || !fa) goto out;
    /* matrix rows striped over threads */
#pragma omp parallel for schedule(static, 1) num_threads(nthr) if (nthr > 1) reduction(| : coincident)
    for (int i = 0; i < n; ++i) {
        A[(size_t)i * n + i] = P->gam_eff[i] + TSQRT2PI / sqrt(P->alpeeq[i]);
        for (int j = 0; j < i; ++j) {
            double d[3];
            sub3(x + 3 * i, x + 3 * j, d);
            double r = norm3(d);
            if (r < 1e-10) { /* reference: zero distance -> fallback charges */
                coincident = 1;
                continue;
            }
            double g = 1.0 / sqrt(P->alpeeq[i] + P->alpeeq[j]);
            double v = erf(g * r) / r;
            A[(size_t)i * n + j] = v;
            A[(size_t)j * n + i] = v;
        }
        rhs[i] = P->chi_base[i] + P->cnf[i] * sqrt(cn[i]);
    }
    if (coincident) goto out;
    if (B)
        for (size_t k = 0; k < (size_t)n * n; ++k) A[k] += B[k];
    for (int i = 0; i < n; ++i) fo[P->frag[i] + 1]++;
    for (int f = 0; f < nf; ++f) fo[f + 1] += fo[f];
    {
        int* fill = (int*)calloc(nf > 0 ? nf : 1, sizeof(int));
        if (!fill) goto out;
        for (int i = 0; i < n; ++i) fa[fo[P->frag[i]] + fill[P->frag[i]]++] = i;
        free(fill);
    }
    if (!chol(n, A, nthr)) {
        /* not positive definite: LU on the bordered system, as the reference's
           dispatchSolve falls back to LU (eeq_solver.cpp:2143-2165) */
        int m = n + nf;
        double* M = (double*)calloc((size_t)m * m, sizeof(double));
        double* b = (double*)calloc(m, sizeof(double));
        if (!M || !b) {
            free(M);
            free(b);
            goto out;
        }
        for (int i = 0; i < n; ++i) {
            M[(size_t)i * m + i] = P->gam_eff[i] + TSQRT2PI / sqrt(P->alpeeq[i]);
            for (int j = 0; j < i; ++j) {
                double d[3];
                sub3(x + 3 * i, x + 3 * j, d);
                double r = norm3(d);
                double g = 1.0 / sqrt(P->alpeeq[i] + P->alpeeq[j]);
                double v = erf(g * r) / r;
                M[(size_t)i * m + j] = v;
                M[(size_t)j * m + i] = v;
            }
            if (B)
                for (int j = 0; j <= i; ++j) {
                    M[(size_t)i * m + j] += B[(size_t)i * n + j];
                    if (j < i) M[(size_t)j * m + i] += B[(size_t)j * n + i];
                }
            M[(size_t)i * m + n + P->frag[i]] = 1.0;
            M[(size_t)(n + P->frag[i]) * m + i] = 1.0;
            b[i] = P->chi_base[i] + P->cnf[i] * sqrt(cn[i]);
        }
        for (int f = 0; f < nf; ++f) b[n + f] = P->qfrag[f];
        int okl = lu_solve(m, M, b);
        if (okl)
            for (int i = 0; i < n; ++i) q[i] = b[i];
        free(M);
        free(b);
        ok = okl;
        goto out;
    }
    chol_solve(n, A, rhs);
    /* one solve per fragment constraint column, threaded over fragments */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthr) if (nthr > 1)
    for (int f = 0; f < nf; ++f) {
        double* c = Z2 + (size_t)f * n;
        for (int i = 0; i < n; ++i) c[i] = (P->frag[i] == f) ? 1.0 : 0.0;
        chol_solve(n, A, c);
    }
    /* Schur complement S(f, g) = sum over the atoms of f (ascending) of Z2(g) */
#pragma omp parallel for schedule(static) num_threads(nthr) if (nthr > 1)
    for (int f = 0; f < nf; ++f) {
        for (int g = 0; g < nf; ++g) {
            double s = 0.0;
            const double* c = Z2 + (size_t)g * n;
            for (int k = fo[f]; k < fo[f + 1]; ++k) s += c[fa[k]];
            S[(size_t)f * nf + g] = s;
        }
        double s = 0.0;
        for (int k = fo[f]; k < fo[f + 1]; ++k) s += rhs[fa[k]];
        lam[f] = s - P->qfrag[f];
    }
    if (!lu_solve(nf, S, lam)) goto out;
    for (int i = 0; i < n; ++i) {
        double s = rhs[i];
        for (int f = 0; f < nf; ++f) s -= Z2[(size_t)f * n + i] * lam[f];
        q[i] = s;
    }
    ok = 1;
out:
    free(A);
    free(rhs);
    free(Z2);
    free(S);
    free(lam);
    free(fo);
    free(fa);
    return ok;
}

/* ------------------------------------------------------------------------- */
/* geometry helpers                                                           */
/* ------------------------------------------------------------------------- */

/* dphi/dr for the signed dihedral i-j-k-l (phi = atan2, the angle of
 * gfnff_geometry.h:60-200); derivative from the standard decomposition
 * dphi/dri = -|b|/|n1|^2 n1, dphi/drl = |b|/|n2|^2 n2. */
static double dihedral_grad(const double* ri, const double* rj, const double* rk, const double* rl, double g[4][3],
                            int* ok) {
    double v1[3], v2[3], v3[3], n1[3], n2[3];
    sub3(rj, ri, v1);
    sub3(rk, rj, v2);
    sub3(rl, rk, v3);
    cross3(v1, v2, n1);
    cross3(v2, v3, n2);
    double nn1 = norm3(n1), nn2 = norm3(n2), nv2 = norm3(v2);
    *ok = 1;
    if (nn1 < 1e-10 || nn2 < 1e-10 || nv2 < 1e-10) {
        memset(g, 0, sizeof(double) * 12);
        *ok = 0;
        return 0.0;
    }
    double c = dot3(n1, n2) / (nn1 * nn2);
    double s = nv2 * dot3(n1, v3) / nn1 / nn2;
    if (s > 1.0) s = 1.0;
    if (s < -1.0) s = -1.0;
    double phi = atan2(s, c);
    double n1s = nn1 * nn1, n2s = nn2 * nn2, b2 = nv2 * nv2;
    double f1 = -nv2 / n1s, f4 = nv2 / n2s;
    double a = dot3(v1, v2) / b2, b = dot3(v3, v2) / b2;
    for (int d = 0; d < 3; ++d) {
        double gi = f1 * n1[d], gl = f4 * n2[d];
        g[0][d] = gi;
        g[3][d] = gl;
        g[1][d] = -gi - a * gi + b * gl;
        g[2][d] = -gl + a * gi - b * gl;
    }
    /* the reference's sign convention for the torsion derivative: the energy
       terms are even in phi, so dE/dr is convention independent */
    return phi;
}

/* out-of-plane angle of the centre i against the plane through its first
 * neighbour j (calculateOutOfPlaneAngle, gfnff_geometry.h:207-300):
 * n = (ri - rj) x (rk - rj), v = rl - ri, omega = asin(n.v / (|n||v|)).
 * Gradient from d(n.v/(|n||v|)) = gn.dn + gv.dv with gn = v/(|n||v|) - s n/|n|^2,
 * gv = n/(|n||v|) - s v/|v|^2 and d(a.(e x d)) = (d x a).de + (a x e).dd. */
static double oop_angle(const double* ri, const double* rj, const double* rk, const double* rl, double g[4][3]) {
    double re[3], rd[3], rv[3], nn[3];
    sub3(ri, rj, re);
    sub3(rk, rj, rd);
    sub3(rl, ri, rv);
    cross3(re, rd, nn);
    double N = norm3(nn), V = norm3(rv);
    memset(g, 0, sizeof(double) * 12);
    if (N < 1e-10 || V < 1e-10) return 0.0;
    double s = dot3(nn, rv) / (N * V);
    if (s > 1.0) s = 1.0;
    if (s < -1.0) s = -1.0;
    double om = asin(s), co = cos(om);
    if (fabs(N * V * co) < 1e-14) return om;
    double gn[3], gv[3], dre[3], drd[3];
    for (int c = 0; c < 3; ++c) {
        gn[c] = rv[c] / (N * V) - s * nn[c] / (N * N);
        gv[c] = nn[c] / (N * V) - s * rv[c] / (V * V);
    }
    cross3(rd, gn, dre);
    cross3(gn, re, drd);
    for (int c = 0; c < 3; ++c) {
        g[0][c] = (dre[c] - gv[c]) / co;
        g[1][c] = (-dre[c] - drd[c]) / co;
        g[2][c] = drd[c] / co;
        g[3][c] = gv[c] / co;
    }
    return om;
}

/* ------------------------------------------------------------------------- */
/* the evaluation                                                             */
/* ------------------------------------------------------------------------- */
/* D4 ATM over bonded triplets: calcATM / calcATMGradient
 * ff_workspace_gfnff.cpp:1526-1658, with the C6 of the current D4 weights.
 * The energy uses +c9 and the gradient is assembled from -c9 with the
 * reference's -(dgij + dgik) orientation.  Worksharing over the triplets when
 * called inside a parallel region (linear per-thread ranges). */
static void atm_terms(const gfnff_params_view* P, const double* x, const double* gw, const double* dgw,
                      int want_grad, terms_t* Tp, gacc_t* Gp) {
#define T (*Tp)
#define G (*Gp)
#pragma omp for schedule(static) nowait
    for (int t = 0; t < P->natm; ++t) {
        int i = P->atm_ijk[3 * t], j = P->atm_ijk[3 * t + 1], k = P->atm_ijk[3 * t + 2];
        double c6ij, c6ik, c6jk, d1, d2;
        oracle_c6(P, gw, dgw, i, j, &c6ij, &d1, &d2);
        oracle_c6(P, gw, dgw, i, k, &c6ik, &d1, &d2);
        oracle_c6(P, gw, dgw, j, k, &c6jk, &d1, &d2);
        double c9 = 1.0 * sqrt(fabs(c6ij * c6ik * c6jk));
        double vij[3], vik[3], vjk[3];
        sub3(x + 3 * j, x + 3 * i, vij);
        sub3(x + 3 * k, x + 3 * i, vik);
        sub3(x + 3 * k, x + 3 * j, vjk);
        double rij = norm3(vij), rik = norm3(vik), rjk = norm3(vjk);
        double r2ij = rij * rij, r2ik = rik * rik, r2jk = rjk * rjk;
        double r0ij = 0.58 * sqrt(3.0 * P->rcov_d3[i] * P->rcov_d3[j]) + 4.80;
        double r0ik = 0.58 * sqrt(3.0 * P->rcov_d3[i] * P->rcov_d3[k]) + 4.80;
        double r0jk = 0.58 * sqrt(3.0 * P->rcov_d3[j] * P->rcov_d3[k]) + 4.80;
        double rijk = rij * rik * rjk, r2ijk = r2ij * r2ik * r2jk, r3ijk = rijk * r2ijk, r5ijk = r2ijk * r3ijk;
        double tmp = pow(r0ij * r0ik * r0jk / rijk, 14.0 / 3.0);
        double fdmp = 1.0 / (1.0 + 6.0 * tmp);
        double Aa = r2ij + r2jk - r2ik, Bb = r2ij + r2ik - r2jk, Cc = r2ik + r2jk - r2ij;
        double ang = (0.375 * Aa * Bb * Cc / r2ijk + 1.0) / r3ijk;
        double sc = P->atm_scale[t];
        T.e[GFNFF_E_ATM] += ang * fdmp * c9 / 3.0 * sc;
        if (want_grad) {
            double dfdmp = -2.0 * 14.0 * tmp * fdmp * fdmp;
            double dij = -0.375 * (pow(r2ij, 3) + pow(r2ij, 2) * (r2jk + r2ik) +
                                   r2ij * (3.0 * pow(r2jk, 2) + 2.0 * r2jk * r2ik + 3.0 * pow(r2ik, 2)) -
                                   5.0 * pow(r2jk - r2ik, 2) * (r2jk + r2ik)) / r5ijk;
            double dik = -0.375 * (pow(r2ik, 3) + pow(r2ik, 2) * (r2jk + r2ij) +
                                   r2ik * (3.0 * pow(r2jk, 2) + 2.0 * r2jk * r2ij + 3.0 * pow(r2ij, 2)) -
                                   5.0 * pow(r2jk - r2ij, 2) * (r2jk + r2ij)) / r5ijk;
            double djk = -0.375 * (pow(r2jk, 3) + pow(r2jk, 2) * (r2ik + r2ij) +
                                   r2jk * (3.0 * pow(r2ik, 2) + 2.0 * r2ik * r2ij + 3.0 * pow(r2ij, 2)) -
                                   5.0 * pow(r2ik - r2ij, 2) * (r2ik + r2ij)) / r5ijk;
            double pre = -c9 * sc / 3.0;
            double a1 = pre * (-dij * fdmp + ang * dfdmp) / r2ij, a2 = pre * (-dik * fdmp + ang * dfdmp) / r2ik,
                   a3 = pre * (-djk * fdmp + ang * dfdmp) / r2jk;
            double gi[3], gj[3], gk[3];
            for (int c = 0; c < 3; ++c) {
                gi[c] = -(a1 * vij[c] + a2 * vik[c]);
                gj[c] = a1 * vij[c] - a3 * vjk[c];
                gk[c] = a2 * vik[c] + a3 * vjk[c];
            }
            gadd(&G, GAT, i, gi[0], gi[1], gi[2]);
            gadd(&G, GAT, j, gj[0], gj[1], gj[2]);
            gadd(&G, GAT, k, gk[0], gk[1], gk[2]);
        }
    }
#undef T
#undef G
}

/* One evaluation.  nthr > 1 runs the FFWorkspace threading scheme (no
 * per-component gradients then); q_in, when given, replaces the EEQ solve by
 * the caller's charges (used to check terms at sizes whose dense solve the
 * CPU cannot afford; the charges are then checked separately). */
static int eval_impl(const gfnff_params_view* P, const double* x, int want_grad, int nthr, const double* q_in,
                     double* E_total, double* terms_out, double* grad, double* q_out, double* cn_out,
                     double* gcomp /* [NGC*3n] or NULL */) {
    const int n = P->natoms;
    int status = 0;
    if (nthr > 1 && gcomp) return 1;
    double *cn_raw = calloc(n, 8), *cn = calloc(n, 8), *dlog = calloc(n, 8), *gw = calloc(7 * (size_t)n, 8),
           *dgw = calloc(7 * (size_t)n, 8), *q = calloc(n, 8), *dEdcn_m = calloc(n, 8), *g = calloc(3 * (size_t)n, 8);
    terms_t Tm;
    memset(&Tm, 0, sizeof(Tm));
    gacc_t Gm;
    Gm.n = n;
    Gm.g = g;
    for (int c = 0; c < NGC; ++c) Gm.gc[c] = gcomp ? gcomp + (size_t)c * 3 * n : NULL;
    if (gcomp) memset(gcomp, 0, sizeof(double) * NGC * 3 * (size_t)n);
    double* dEdcn_bond_m = calloc(n, 8);
    double* hbcn = NULL;
    if (!cn_raw || !cn || !dlog || !gw || !dgw || !q || !dEdcn_m || !g || !dEdcn_bond_m) {
        status = 1;
        goto done;
    }

    oracle_cn(P, x, 6.0, cn_raw, cn, dlog, nthr);
    oracle_gw(P, cn, gw, dgw);
    if (q_in) {
        memcpy(q, q_in, sizeof(double) * n);
    } else if (!oracle_eeq(P, x, cn, q, nthr, NULL)) {
        status = 3;
        goto done;
    }

    /* --- HB coordination numbers for X-H bonds: ff_workspace_gfnff.cpp:46-97 */
    hbcn = calloc(n, 8);
    if (!hbcn) {
        status = 1;
        goto done;
    }
    for (int e = 0; e < P->nhbcn; ++e) {
        int H = P->hbcn_pair[2 * e], B = P->hbcn_pair[2 * e + 1];
        double d[3];
        sub3(x + 3 * B, x + 3 * H, d);
        double r2 = dot3(d, d);
        if (r2 > 900.0) continue;
        double r = sqrt(r2);
        double rc = 1.78 * (4.0 / 3.0) * (P->rcov_d3[H] + P->rcov_d3[B]);
        double arg = -27.5 * (r - rc) / rc;
        hbcn[H] += 0.5 * (1.0 + erf(arg));
    }

    /* Term section: FFWorkspace's scheme (ff_workspace.cpp:127-162, 369-413).
       Every term list is cut into linear per-thread ranges (schedule(static)),
       the O(N^2) pair loops into striped rows; each thread accumulates its own
       energies, gradient and dE/dCN, and the accumulators are reduced in thread
       order after the region.  With one thread the accumulators are the
       outputs themselves. */
    for (int h = 0; h < P->nhb; ++h)
        if (P->hb_nbC_off[h + 1] - P->hb_nbC_off[h] > 16) {
            status = 1;
            goto done;
        }
    const int NT = nthr > 1 ? nthr : 1;
    double* tl = NULL;
    terms_t* tlT = NULL;
    if (NT > 1) {
        tl = calloc((size_t)NT * 5 * n, 8);
        tlT = calloc(NT, sizeof(terms_t));
        if (!tl || !tlT) {
            free(tl);
            free(tlT);
            status = 1;
            goto done;
        }
    }
#pragma omp parallel num_threads(NT) if (NT > 1)
    {
    const int tid = THREAD_ID();
    terms_t T;
    memset(&T, 0, sizeof(T));
    gacc_t G = Gm;
    double* dEdcn = dEdcn_m;
    double* dEdcn_bond = dEdcn_bond_m;
    if (NT > 1) {
        G.g = tl + (size_t)tid * 5 * n;
        for (int c = 0; c < NGC; ++c) G.gc[c] = NULL;
        dEdcn = G.g + 3 * (size_t)n;
        dEdcn_bond = dEdcn + n;
    }
    /* --- bonds: calcBonds ff_workspace_gfnff.cpp:104-180 */
#pragma omp for schedule(static) nowait
    for (int b = 0; b < P->nbond; ++b) {
        int i = P->bond_ij[2 * b], j = P->bond_ij[2 * b + 1];
        const double* p = P->bond_par + 8 * b;
        double d[3];
        sub3(x + 3 * i, x + 3 * j, d);
        double r = norm3(d);
        double r0 = (p[0] + p[2] * cn[i] + p[1] + p[3] * cn[j] + p[4]) * p[5];
        double dr = r - r0;
        double a0 = p[6], a = a0;
        int H = P->bond_hbH[b];
        if (H >= 0) a = (-(1.0 - 0.9) * hbcn[H] + 1.0) * a0;
        double e = p[7] * exp(-a * dr * dr);
        T.e[GFNFF_E_BOND] += e;
        if (want_grad) {
            double dEdr = -2.0 * a * dr * e;
            gadd(&G, GB, i, dEdr * d[0] / r, dEdr * d[1] / r, dEdr * d[2] / r);
            gadd(&G, GB, j, -dEdr * d[0] / r, -dEdr * d[1] / r, -dEdr * d[2] / r);
            if (H >= 0) {
                double zz = 0.1 * a0 * dr * dr * e;
                for (int k = 0; k < P->nhbcn; ++k) {
                    if (P->hbcn_pair[2 * k] != H) continue;
                    int B = P->hbcn_pair[2 * k + 1];
                    double v[3];
                    sub3(x + 3 * B, x + 3 * H, v);
                    double rr2 = dot3(v, v);
                    if (rr2 > 900.0) continue;
                    double rr = sqrt(rr2);
                    double rc = 1.78 * (4.0 / 3.0) * (P->rcov_d3[H] + P->rcov_d3[B]);
                    double arg = -27.5 * (rr - rc) / rc;
                    double dcn = 0.5641895835477563 * (-27.5 / rc) * exp(-arg * arg) / rr;
                    gadd(&G, GB, H, -zz * dcn * v[0], -zz * dcn * v[1], -zz * dcn * v[2]);
                    gadd(&G, GB, B, zz * dcn * v[0], zz * dcn * v[1], zz * dcn * v[2]);
                }
            }
            double yy = -dEdr;
            dEdcn[i] += yy * p[5] * p[2];
            dEdcn[j] += yy * p[5] * p[3];
            dEdcn_bond[i] += yy * p[5] * p[2];
            dEdcn_bond[j] += yy * p[5] * p[3];
        }
    }

    /* --- angles: calcAngles ff_workspace_gfnff.cpp:187-275 */
#pragma omp for schedule(static) nowait
    for (int a = 0; a < P->nangle; ++a) {
        int i = P->angle_ijk[3 * a], j = P->angle_ijk[3 * a + 1], k = P->angle_ijk[3 * a + 2];
        double fc = P->angle_par[2 * a], th0 = P->angle_par[2 * a + 1];
        double u[3], v[3];
        sub3(x + 3 * i, x + 3 * j, u);
        sub3(x + 3 * k, x + 3 * j, v);
        double ru = norm3(u), rv = norm3(v);
        double cth = (ru < 1e-10 || rv < 1e-10) ? 1.0 : dot3(u, v) / (ru * rv);
        if (cth > 1.0) cth = 1.0;
        if (cth < -1.0) cth = -1.0;
        double th = acos(cth), e, dEdcos;
        if (fabs(M_PI - th0) < 1e-6) {
            double dt = th - th0;
            e = fc * dt * dt;
            double s = sin(th);
            dEdcos = (s > 1e-14) ? -2.0 * fc * dt / s : 0.0;
        } else {
            double c0 = cos(th0);
            e = fc * (cth - c0) * (cth - c0);
            dEdcos = 2.0 * fc * (cth - c0);
        }
        double rci = P->rcov_d3[i] * (4.0 / 3.0), rcj = P->rcov_d3[j] * (4.0 / 3.0), rck = P->rcov_d3[k] * (4.0 / 3.0);
        double rcut_ij = 0.595 * (rci + rcj) * (rci + rcj);
        double rcut_jk = 0.595 * (rcj + rck) * (rcj + rck);
        double rij2 = ru * ru, rjk2 = rv * rv;
        double rr1 = rij2 / rcut_ij, rr2 = rjk2 / rcut_jk;
        rr1 *= rr1;
        rr2 *= rr2;
        double d1 = 1.0 / (1.0 + rr1), d2 = 1.0 / (1.0 + rr2), damp = d1 * d2;
        T.e[GFNFF_E_ANGLE] += e * damp;
        if (want_grad) {
            double dd1 = (rij2 > 1e-8) ? -4.0 * rr1 / (rij2 * (1.0 + rr1) * (1.0 + rr1)) : 0.0;
            double dd2 = (rjk2 > 1e-8) ? -4.0 * rr2 / (rjk2 * (1.0 + rr2) * (1.0 + rr2)) : 0.0;
            double gi[3], gk[3];
            double sinth = sqrt(fmax(0.0, 1.0 - cth * cth));
            double pre = (sinth < 1e-14 || ru < 1e-10 || rv < 1e-10) ? 0.0 : dEdcos * damp;
            for (int c = 0; c < 3; ++c) {
                /* dcos/dri = (v/|v| - cos u/|u|)/|u| */
                gi[c] = pre * (v[c] / rv - cth * u[c] / ru) / ru + e * dd1 * d2 * u[c];
                gk[c] = pre * (u[c] / ru - cth * v[c] / rv) / rv + e * dd2 * d1 * v[c];
            }
            gadd(&G, GA, i, gi[0], gi[1], gi[2]);
            gadd(&G, GA, k, gk[0], gk[1], gk[2]);
            gadd(&G, GA, j, -gi[0] - gk[0], -gi[1] - gk[1], -gi[2] - gk[2]);
        }
    }

    /* --- torsions (primary + extra): ff_workspace_gfnff.cpp:282-482 */
#pragma omp for schedule(static) nowait
    for (int t = 0; t < P->ntors; ++t) {
        const int* id = P->tors_ijkl + 4 * t;
        const double* p = P->tors_par + 4 * t;
        double gphi[4][3];
        int ok;
        double phi = dihedral_grad(x + 3 * id[0], x + 3 * id[1], x + 3 * id[2], x + 3 * id[3], gphi, &ok);
        double V = p[0], nn = p[1], phi0 = p[2], atc = p[3];
        double c1 = nn * (phi - phi0) + M_PI;
        double e = V * (1.0 + cos(c1));
        double r2[3], v[3][3];
        sub3(x + 3 * id[0], x + 3 * id[1], v[0]);
        sub3(x + 3 * id[1], x + 3 * id[2], v[1]);
        sub3(x + 3 * id[2], x + 3 * id[3], v[2]);
        double dmp[3], rr[3];
        for (int s = 0; s < 3; ++s) {
            r2[s] = dot3(v[s], v[s]);
            double rs = P->rcov_d3[id[s]] * (4.0 / 3.0) + P->rcov_d3[id[s + 1]] * (4.0 / 3.0);
            double rc = atc * rs * rs;
            rr[s] = r2[s] / rc;
            rr[s] *= rr[s];
            dmp[s] = 1.0 / (1.0 + rr[s]);
        }
        double damp = dmp[0] * dmp[1] * dmp[2];
        T.e[GFNFF_E_TORSION] += e * damp;
        if (want_grad) {
            double dEdphi = ok ? -V * nn * sin(c1) * damp : 0.0;
            for (int a = 0; a < 4; ++a)
                gadd(&G, GT, id[a], dEdphi * gphi[a][0], dEdphi * gphi[a][1], dEdphi * gphi[a][2]);
            double dd[3];
            for (int s = 0; s < 3; ++s)
                dd[s] = (r2[s] > 1e-8) ? -4.0 * rr[s] / (r2[s] * (1.0 + rr[s]) * (1.0 + rr[s])) : 0.0;
            double f0 = e * dd[0] * dmp[1] * dmp[2], f1 = e * dmp[0] * dd[1] * dmp[2], f2 = e * dmp[0] * dmp[1] * dd[2];
            gadd(&G, GT, id[0], f0 * v[0][0], f0 * v[0][1], f0 * v[0][2]);
            gadd(&G, GT, id[1], -f0 * v[0][0] + f1 * v[1][0], -f0 * v[0][1] + f1 * v[1][1], -f0 * v[0][2] + f1 * v[1][2]);
            gadd(&G, GT, id[2], -f1 * v[1][0] + f2 * v[2][0], -f1 * v[1][1] + f2 * v[2][1], -f1 * v[1][2] + f2 * v[2][2]);
            gadd(&G, GT, id[3], -f2 * v[2][0], -f2 * v[2][1], -f2 * v[2][2]);
        }
    }
    /* --- inversions: calcInversions ff_workspace_gfnff.cpp:489-604 (damping
       over the three distances from the first neighbour j) */
#pragma omp for schedule(static) nowait
    for (int v = 0; v < P->ninv; ++v) {
        const int* id = P->inv_ijkl + 4 * v;
        const double* p = P->inv_par + 3 * v;
        double gom[4][3];
        double om = oop_angle(x + 3 * id[0], x + 3 * id[1], x + 3 * id[2], x + 3 * id[3], gom);
        double rc[4];
        for (int a = 0; a < 4; ++a) rc[a] = P->rcov_d3[id[a]] * (4.0 / 3.0);
        double dv[3][3], r2[3], dmp[3], rr[3];
        const int other[3] = {0, 2, 3};
        for (int s2 = 0; s2 < 3; ++s2) {
            int o = other[s2];
            sub3(x + 3 * id[1], x + 3 * id[o], dv[s2]);
            r2[s2] = dot3(dv[s2], dv[s2]);
            double rcut = 0.505 * (rc[1] + rc[o]) * (rc[1] + rc[o]);
            rr[s2] = (r2[s2] / rcut) * (r2[s2] / rcut);
            dmp[s2] = 1.0 / (1.0 + rr[s2]);
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
        T.e[GFNFF_E_INVERSION] += et * damp;
        if (want_grad) {
            for (int a = 0; a < 4; ++a)
                gadd(&G, GT, id[a], dEdom * gom[a][0], dEdom * gom[a][1], dEdom * gom[a][2]);
            double f[3];
            for (int s2 = 0; s2 < 3; ++s2) {
                double dd = (r2[s2] < 1e-8) ? 0.0 : -4.0 * rr[s2] / (r2[s2] * (1.0 + rr[s2]) * (1.0 + rr[s2]));
                f[s2] = et * dd * dmp[(s2 + 1) % 3] * dmp[(s2 + 2) % 3];
            }
            double gj[3] = {0, 0, 0};
            for (int s2 = 0; s2 < 3; ++s2) {
                int o = other[s2];
                gadd(&G, GT, id[o], -f[s2] * dv[s2][0], -f[s2] * dv[s2][1], -f[s2] * dv[s2][2]);
                for (int c = 0; c < 3; ++c) gj[c] += f[s2] * dv[s2][c];
            }
            gadd(&G, GT, id[1], gj[0], gj[1], gj[2]);
        }
    }

    /* --- triple-bond torsions: calcSTorsions ff_workspace_gfnff.cpp:611-648 (undamped) */
#pragma omp for schedule(static) nowait
    for (int t = 0; t < P->nstors; ++t) {
        const int* id = P->stors_ijkl + 4 * t;
        double eh = P->stors_par[t];
        double gphi[4][3];
        int ok;
        double phi = dihedral_grad(x + 3 * id[0], x + 3 * id[1], x + 3 * id[2], x + 3 * id[3], gphi, &ok);
        T.e[GFNFF_E_STORS] += eh - eh * cos(2.0 * phi);
        if (want_grad && ok) {
            double d = 2.0 * eh * sin(2.0 * phi);
            for (int a = 0; a < 4; ++a) gadd(&G, GT, id[a], d * gphi[a][0], d * gphi[a][1], d * gphi[a][2]);
        }
    }

    /* --- D4 dispersion pairs: calcD4Dispersion ff_workspace_gfnff.cpp:715-767 */
#pragma omp for schedule(static, 1) nowait
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double d[3];
            sub3(x + 3 * i, x + 3 * j, d);
            double r = norm3(d);
            if (r > 50.0 || r < 1e-10) continue;
            double c6, dci, dcj;
            oracle_c6(P, gw, dgw, i, j, &c6, &dci, &dcj);
            double r4r2 = 3.0 * P->sqrt_zr4r2[i] * P->sqrt_zr4r2[j];
            double r0 = 0.58 * sqrt(r4r2) + 4.80;
            double r0s = r0 * r0;
            double zc = P->zeta[i] * P->zeta[j];
            double r2 = r * r, r6 = r2 * r2 * r2, r06 = r0s * r0s * r0s;
            double t6 = 1.0 / (r6 + r06), t8 = 1.0 / (r6 * r2 + r06 * r0s);
            double sum = t6 + 2.0 * r4r2 * t8;
            T.e[GFNFF_E_DISPERSION] += -c6 * sum * zc;
            if (want_grad) {
                double d6 = -6.0 * r2 * r2 * t6 * t6, d8 = -8.0 * r2 * r2 * r2 * t8 * t8;
                double dEdr = -c6 * zc * (d6 + 2.0 * r4r2 * d8) * r;
                gadd(&G, GD, i, dEdr * d[0] / r, dEdr * d[1] / r, dEdr * d[2] / r);
                gadd(&G, GD, j, -dEdr * d[0] / r, -dEdr * d[1] / r, -dEdr * d[2] / r);
                dEdcn[i] -= dci * sum * zc;
                dEdcn[j] -= dcj * sum * zc;
            }
        }

    /* --- repulsion: bonded ff_workspace_gfnff.cpp:774-807, non-bonded :814-842
       (non-bonded pair set = all non-bonded pairs within 20 Bohr,
        gfnff_method.cpp:9781-9872) */
#pragma omp for schedule(static) nowait
    for (int b = 0; b < P->nbrep; ++b) {
        int i = P->brep_ij[2 * b], j = P->brep_ij[2 * b + 1];
        double al = P->brep_par[2 * b], ab = P->brep_par[2 * b + 1];
        double d[3];
        sub3(x + 3 * i, x + 3 * j, d);
        double r = norm3(d);
        if (r > 20.0 || r < 1e-8) continue;
        double e = ab * exp(-al * r * sqrt(r)) / r;
        T.e[GFNFF_E_BONDED_REP] += e;
        if (want_grad) {
            double dEdr = -e / r - 1.5 * al * sqrt(r) * e;
            gadd(&G, GR, i, dEdr * d[0] / r, dEdr * d[1] / r, dEdr * d[2] / r);
            gadd(&G, GR, j, -dEdr * d[0] / r, -dEdr * d[1] / r, -dEdr * d[2] / r);
        }
    }
#pragma omp for schedule(static, 1) nowait
    for (int i = 0; i < n; ++i) {
        int zi = P->Z[i];
        for (int j = i + 1; j < n; ++j) {
            double d[3];
            sub3(x + 3 * i, x + 3 * j, d);
            double r2 = dot3(d, d);
            if (r2 > 400.0) continue;
            int td = 999;
            for (int k = P->topo_off[i]; k < P->topo_off[i + 1]; ++k)
                if (P->topo_idx[k] == j) {
                    td = P->topo_dist[k];
                    break;
                }
            if (td == 1) continue;
            double r = sqrt(r2);
            if (r > 20.0 || r < 1e-8) continue;
            int zj = P->Z[j];
            double ff = 1.0;
            if (zi == 1 && zj == 1) {
                ff = 0.6290;
                if (td == 2) ff *= 1.4580;
                else if (td == 3) ff *= 0.7080;
            } else if ((zi == 1 && P->metal[j] > 0) || (zj == 1 && P->metal[i] > 0)) {
                ff = 0.85;
            } else if ((zi == 1 && zj == 6) || (zj == 1 && zi == 6)) {
                ff = 0.91;
            } else if ((zi == 1 && zj == 8) || (zj == 1 && zi == 8)) {
                ff = 1.04;
            }
            double al = sqrt(P->rep_anb[i] * P->rep_anb[j]) * ff;
            double ab = P->repz[i] * P->repz[j] * 0.4270;
            double e = ab * exp(-al * r * sqrt(r)) / r;
            T.e[GFNFF_E_NONBONDED_REP] += e;
            if (want_grad) {
                double dEdr = -e / r - 1.5 * al * sqrt(r) * e;
                gadd(&G, GRN, i, dEdr * d[0] / r, dEdr * d[1] / r, dEdr * d[2] / r);
                gadd(&G, GRN, j, -dEdr * d[0] / r, -dEdr * d[1] / r, -dEdr * d[2] / r);
            }
        }
    }

    /* --- Coulomb pairs: calcCoulomb ff_workspace_gfnff.cpp:850-898 (r_cut 100) */
#pragma omp for schedule(static, 1) nowait
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double d[3];
            sub3(x + 3 * i, x + 3 * j, d);
            double r = norm3(d);
            if (r > 100.0 || r < 1e-10) continue;
            double gij = 1.0 / sqrt(P->alpeeq[i] + P->alpeeq[j]);
            double gr = gij * r;
            double ef = erf(gr);
            T.e[GFNFF_E_COULOMB] += q[i] * q[j] * ef / r;
            if (want_grad) {
                double dEdr = q[i] * q[j] * (gij * exp(-gr * gr) * (2.0 / 1.772453850905516) / r - ef / (r * r));
                gadd(&G, GC, i, dEdr * d[0] / r, dEdr * d[1] / r, dEdr * d[2] / r);
                gadd(&G, GC, j, -dEdr * d[0] / r, -dEdr * d[1] / r, -dEdr * d[2] / r);
            }
        }
    /* self energy: ff_workspace.cpp:426-449 */
    if (tid == 0) {
        double Een = 0.0, Es = 0.0;
        for (int i = 0; i < n; ++i) {
            double chi = P->chi_base[i] + P->cnf[i] * sqrt(fmax(cn[i], 0.0));
            Een -= q[i] * chi;
            Es += 0.5 * q[i] * q[i] * (P->gam_eff[i] + 0.797884560802865 / sqrt(P->alpeeq[i]));
        }
        T.e[GFNFF_E_COULOMB] += Een + Es;
    }

    /* --- hydrogen bonds: calcHydrogenBonds ff_workspace_gfnff.cpp:938-1404 */
#pragma omp for schedule(static) nowait
    for (int h = 0; h < P->nhb; ++h) {
        int A = P->hb_ahb[3 * h], Hh = P->hb_ahb[3 * h + 1], B = P->hb_ahb[3 * h + 2];
        int cs = P->hb_case[h];
        const double* hp = P->hb_par + 7 * h;
        double basA = hp[0], basB = hp[1], aciA = hp[2], aciB = hp[3];
        const double *pA = x + 3 * A, *pH = x + 3 * Hh, *pB = x + 3 * B;
        double vAH[3], vHB[3], vAB[3];
        sub3(pH, pA, vAH);
        sub3(pB, pH, vHB);
        sub3(pB, pA, vAB);
        double rAH = norm3(vAH), rHB = norm3(vHB), rAB = norm3(vAB);
        if (rAB > 50.0) continue;
        double rAH4 = rAH * rAH * rAH * rAH, rHB4 = rHB * rHB * rHB * rHB;
        double den = 1.0 / (rAH4 + rHB4);
        double eH = exp(15.0 * hp[4]), eA = exp(15.0 * -hp[5]), eB = exp(15.0 * -hp[6]);
        double QH = eH / (eH + 1.0), QA = eA / (eA + 1.0), QB = eB / (eB + 1.0);
        double bas = (QA * basA * rAH4 + QB * basB * rHB4) * den;
        double aci = (aciB * rAH4 + aciA * rHB4) * den;
        /* radab: sum of the tabulated Pyykko radii (Angstrom values used as is,
           ff_workspace_gfnff.cpp:982) */
        double radab = P->hb_rad[A] + P->hb_rad[B];
        double dshort = 1.0 / (1.0 + pow(22.0 * radab / (rAB * rAB), 6.0));
        double dlong = 1.0 / (1.0 + pow(rAB * rAB / 85.0, 6.0));
        double denv = dshort * dlong;
        double ratio = (rAH + rHB) / rAB;
        double expo0 = (49.0 / radab) * (ratio - 1.0);
        double doutl = (expo0 > 15.0) ? 0.0 : 2.0 / (1.0 + exp(expo0));
        double outl_nb = 1.0;
        int nb0 = P->hb_nbB_off[h], nb1 = P->hb_nbB_off[h + 1];
        double nbcut = (P->Z[B] == 7 && nb1 - nb0 == 1) ? 2.0 : 11.20;
        if (cs >= 2)
            for (int k = nb0; k < nb1; ++k) {
                int nb = P->hb_nbB_idx[k];
                double u[3], w[3];
                sub3(x + 3 * nb, pA, u);
                sub3(x + 3 * nb, pB, w);
                double ex = (nbcut / radab) * ((norm3(u) + norm3(w)) / rAB - 1.0);
                outl_nb *= (2.0 / (1.0 + exp(-ex)) - 1.0);
            }
        double rdamp = (cs >= 2) ? denv * (1.8 / (rHB * rHB * rHB) - 0.8 / (rAB * rAB * rAB)) : denv / (rAB * rAB * rAB);
        /* case 4 (sp2 N acceptor): virtual lone pair opposite the neighbour sum,
           ff_workspace_gfnff.cpp:1002-1028 */
        double outl_lp = 1.0, lpd = 0.0, lpv[3] = {0, 0, 0}, lp[3] = {pB[0], pB[1], pB[2]}, xlp = 0.0;
        int nlp = 0;
        if (cs == 4) {
            lpd = P->hb_lpd[B];
            for (int k = nb0; k < nb1; ++k)
                for (int c = 0; c < 3; ++c) lpv[c] += x[3 * P->hb_nbB_idx[k] + c] - pB[c];
            double vn = norm3(lpv);
            nlp = nb1 - nb0;
            if (vn > 1e-10 && nlp > 0) {
                for (int c = 0; c < 3; ++c) lp[c] = pB[c] - lpd * lpv[c] / vn;
                double d[3];
                sub3(pA, lp, d);
                xlp = (56.0 / radab) * ((norm3(d) + lpd + 1e-12) / rAB - 1.0);
                outl_lp = 2.0 / (1.0 + exp(xlp));
            } else {
                nlp = 0;
            }
        }
        /* case 3 (carbonyl O acceptor): H...O=C bend and D-C=O...H torsion
           factors, ff_workspace_gfnff.cpp:1046-1135 */
        double eangl = 1.0, etors = 1.0, ga3[3][3];
        int C3 = -1, c0i = 0, c1i = 0;
        double et_k[16], gt_k[16][4][3];
        if (cs == 3 && P->hb_parentC[h] >= 0) {
            C3 = P->hb_parentC[h];
            const double cc0 = cos(120.0 * M_PI / 180.0);
            const double kb = (1.0 - 0.20) / ((1.0 - cc0) * (1.0 - cc0));
            double u[3], w[3];
            sub3(x + 3 * C3, pB, u);  /* C - B */
            sub3(pH, pB, w);          /* H - B */
            double ru = norm3(u), rw = norm3(w);
            double ca = dot3(u, w) / (ru * rw + 1e-14);
            if (ca > 1.0) ca = 1.0;
            if (ca < -1.0) ca = -1.0;
            double ea = kb * (ca - cc0) * (ca - cc0);
            eangl = 1.0 - ea;
            double dea = 2.0 * kb * (ca - cc0); /* d ea / d cos */
            for (int c = 0; c < 3; ++c) {
                double dcu = (w[c] / rw - ca * u[c] / ru) / ru, dcw = (u[c] / ru - ca * w[c] / rw) / rw;
                ga3[1][c] = -dea * dcu;             /* d eangl / d C */
                ga3[2][c] = -dea * dcw;             /* d eangl / d H */
                ga3[0][c] = -ga3[1][c] - ga3[2][c]; /* d eangl / d B */
            }
            c0i = P->hb_nbC_off[h];
            c1i = P->hb_nbC_off[h + 1];
            /* list lengths were checked (<= 16) before the term section */
            for (int k = c0i; k < c1i; ++k) {
                int D = P->hb_nbC_idx[k];
                int ok;
                double phi = dihedral_grad(x + 3 * D, pB, x + 3 * C3, pH, gt_k[k - c0i], &ok);
                double c1 = 2.0 * (phi - M_PI / 2.0) + M_PI, fct = (1.0 - 0.94) / 2.0;
                et_k[k - c0i] = (1.0 + cos(c1)) * fct + 0.94;
                double dij = ok ? -2.0 * sin(c1) * fct : 0.0;
                for (int a = 0; a < 4; ++a)
                    for (int c = 0; c < 3; ++c) gt_k[k - c0i][a][c] *= dij;
                etors *= et_k[k - c0i];
            }
        }
        double qhoutl = QH * doutl * outl_nb * outl_lp;
        double gscale = (cs == 2 || cs == 4) ? 0.268 : (cs == 3 ? 0.350 : 1.0);
        double EHB;
        double cval = aciA * basB * QA * QB * gscale;
        if (cs >= 2) EHB = -rdamp * qhoutl * cval * eangl * etors;
        else EHB = -bas * aci * rdamp * qhoutl;
        T.e[GFNFF_E_HBOND] += EHB;
        if (!want_grad) continue;
        double drab[3], drah[3], drbh[3];
        sub3(pA, pB, drab);
        sub3(pA, pH, drah);
        sub3(pB, pH, drbh);
        double rab2 = rAB * rAB, rbh2 = rHB * rHB, rah2 = rAH * rAH;
        double rahprbh = rAH + rHB + 1e-12;
        double ratio1 = pow(rab2 / 85.0, 6.0);
        double ratio3 = pow(22.0 * radab / rab2, 6.0);
        double ddamp = (-2.0 * 6.0 * ratio1 / (1.0 + ratio1)) + (2.0 * 6.0 * ratio3 / (1.0 + ratio3));
        double expo = (49.0 / radab) * (rahprbh / rAB - 1.0);
        if (expo > 15.0) continue;
        double ratio2 = exp(expo);
        double ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0}, gh[3] = {0, 0, 0};
        if (cs >= 2) {
            double rbhdamp = denv * 1.8 / (rbh2 * rHB), rabdamp = denv * -0.8 / (rab2 * rAB);
            double ct = cval * eangl * etors;
            double dterm = -qhoutl * ct, aterm = -rdamp * QH * outl_nb * outl_lp * ct;
            double nbterm = -rdamp * QH * doutl * outl_lp * ct;
            double gi = ((rabdamp + rbhdamp) * ddamp - 3.0 * rabdamp) / rab2 * dterm;
            for (int c = 0; c < 3; ++c) {
                ga[c] = gi * drab[c];
                gb[c] = -gi * drab[c];
            }
            gi = -3.0 * rbhdamp / rbh2 * dterm;
            for (int c = 0; c < 3; ++c) {
                gb[c] += gi * drbh[c];
                gh[c] = -gi * drbh[c];
            }
            double tmp1 = -2.0 * aterm * ratio2 * expo / ((1.0 + ratio2) * (1.0 + ratio2)) / (rahprbh - rAB);
            gi = -tmp1 * rahprbh / rab2;
            for (int c = 0; c < 3; ++c) {
                ga[c] += gi * drab[c];
                gb[c] -= gi * drab[c];
            }
            for (int c = 0; c < 3; ++c) {
                double dga = tmp1 / rAH * drah[c], dgb = tmp1 / rHB * drbh[c];
                ga[c] += dga;
                gb[c] += dgb;
                gh[c] += -dga - dgb;
            }
            for (int k = nb0; k < nb1; ++k) {
                int nb = P->hb_nbB_idx[k];
                double dranb[3], drbnb[3];
                sub3(pA, x + 3 * nb, dranb);
                sub3(pB, x + 3 * nb, drbnb);
                double ranb = norm3(dranb), rbnb = norm3(drbnb), s = ranb + rbnb + 1e-12;
                double ex = (nbcut / radab) * (s / rAB - 1.0);
                double r2n = exp(-ex);
                double o_i = 2.0 / (1.0 + r2n) - 1.0;
                double others = (fabs(o_i) > 1e-12) ? outl_nb / o_i : 1.0;
                double tmp2 = 2.0 * nbterm * others * r2n * ex / ((1.0 + r2n) * (1.0 + r2n)) / (s - rAB);
                double gi2 = -tmp2 * s / rab2;
                double gnb[3];
                for (int c = 0; c < 3; ++c) {
                    ga[c] += gi2 * drab[c];
                    gb[c] -= gi2 * drab[c];
                    double dga = tmp2 / ranb * dranb[c], dgb = tmp2 / rbnb * drbnb[c];
                    ga[c] += dga;
                    gb[c] += dgb;
                    gnb[c] = -dga - dgb;
                }
                gadd(&G, GH, nb, gnb[0], gnb[1], gnb[2]);
            }
            if (cs == 4 && nlp > 0) {
                /* d outl_lp through |A - lp| + lpd over rAB; lp = B - lpd v/|v|,
                   v = sum_nb (x_nb - x_B) */
                double lpterm = -rdamp * QH * doutl * outl_nb * cval;
                double dal[3];
                sub3(pA, lp, dal);
                double ral = norm3(dal), sl = ral + lpd + 1e-12;
                double el = exp(xlp);
                double dEds = -2.0 * lpterm * el / ((1.0 + el) * (1.0 + el)) * (56.0 / radab) / rAB;
                double dEdrab = 2.0 * lpterm * el / ((1.0 + el) * (1.0 + el)) * (56.0 / radab) * sl / (rAB * rAB);
                double glp[3];
                for (int c = 0; c < 3; ++c) {
                    ga[c] += dEdrab * drab[c] / rAB + dEds * dal[c] / (ral + 1e-12);
                    gb[c] -= dEdrab * drab[c] / rAB;
                    glp[c] = -dEds * dal[c] / (ral + 1e-12); /* dE/dlp */
                }
                /* lp = B - lpd v/|v|: dlp/dB = I + lpd nlp P/|v|, dlp/dx_nb = -lpd P/|v|,
                   P = I - v v^T/|v|^2 */
                double vn = norm3(lpv), vg = dot3(lpv, glp);
                double pg[3];
                for (int c = 0; c < 3; ++c) pg[c] = (glp[c] - lpv[c] * vg / (vn * vn)) * lpd / vn;
                for (int c = 0; c < 3; ++c) gb[c] += glp[c] + nlp * pg[c];
                for (int k = nb0; k < nb1; ++k) gadd(&G, GH, P->hb_nbB_idx[k], -pg[0], -pg[1], -pg[2]);
            }
            if (C3 >= 0) {
                double bterm = -rdamp * qhoutl * etors * cval, tterm = -rdamp * qhoutl * eangl * cval;
                for (int c = 0; c < 3; ++c) gb[c] += bterm * ga3[0][c];
                gadd(&G, GH, C3, bterm * ga3[1][0], bterm * ga3[1][1], bterm * ga3[1][2]);
                for (int c = 0; c < 3; ++c) gh[c] += bterm * ga3[2][c];
                for (int k = c0i; k < c1i; ++k) {
                    double f = 1.0;
                    for (int k2 = c0i; k2 < c1i; ++k2)
                        if (k2 != k) f *= et_k[k2 - c0i];
                    f *= tterm;
                    const double(*gk)[3] = gt_k[k - c0i];
                    gadd(&G, GH, P->hb_nbC_idx[k], f * gk[0][0], f * gk[0][1], f * gk[0][2]);
                    for (int c = 0; c < 3; ++c) gb[c] += f * gk[1][c];
                    gadd(&G, GH, C3, f * gk[2][0], f * gk[2][1], f * gk[2][2]);
                    for (int c = 0; c < 3; ++c) gh[c] += f * gk[3][c];
                }
            }
        } else {
            double caa = QA * basA, cbb = QB * basB;
            double rterm = -aci * rdamp * qhoutl, dterm = -aci * bas * qhoutl, sterm = -rdamp * bas * qhoutl;
            double aterm = -aci * bas * rdamp * QH;
            double tmp = den * den * 4.0;
            double dd24a = rah2 * rHB4 * tmp, dd24b = rbh2 * rAH4 * tmp;
            for (int c = 0; c < 3; ++c) {
                ga[c] = (caa - cbb) * dd24a * rterm * drah[c];
                gb[c] = (cbb - caa) * dd24b * rterm * drbh[c];
                gh[c] = -ga[c] - gb[c];
                double dga = (aciB - aciA) * dd24a * drah[c] * sterm;
                double dgb = (aciA - aciB) * dd24b * drbh[c] * sterm;
                ga[c] += dga;
                gb[c] += dgb;
                gh[c] += -dga - dgb;
            }
            double gi = rdamp * (ddamp - 3.0) / rab2;
            for (int c = 0; c < 3; ++c) {
                ga[c] += gi * drab[c] * dterm;
                gb[c] -= gi * drab[c] * dterm;
            }
            gi = aterm * 2.0 * ratio2 * expo * rahprbh / ((1.0 + ratio2) * (1.0 + ratio2)) / (rahprbh - rAB) / rab2;
            for (int c = 0; c < 3; ++c) {
                ga[c] += gi * drab[c];
                gb[c] -= gi * drab[c];
            }
            double to = -2.0 * aterm * ratio2 * expo / ((1.0 + ratio2) * (1.0 + ratio2)) / (rahprbh - rAB);
            for (int c = 0; c < 3; ++c) {
                double dga = drah[c] * to / rAH, dgb = drbh[c] * to / rHB;
                ga[c] += dga;
                gb[c] += dgb;
                gh[c] += -dga - dgb;
            }
        }
        gadd(&G, GH, A, ga[0], ga[1], ga[2]);
        gadd(&G, GH, B, gb[0], gb[1], gb[2]);
        gadd(&G, GH, Hh, gh[0], gh[1], gh[2]);
    }

    /* --- halogen bonds A-X...B: calcHalogenBonds ff_workspace_gfnff.cpp:1411-1519.
       E = -acid_X Q_X Q_B f_s(r_XB) f_l(r_XB) f_o / r_XB^3 with
       f_s = 1/(1 + (5 radab / r^2)^6), f_l = 1/(1 + (r^2/70)^6),
       f_o = 2/(1 + exp(70 ((r_AX + r_XB)/r_AB - 1))) (skipped beyond exponent 15),
       Q = e^{15 q}/(e^{15 q} + 0.03).  The gradient is the reference's
       (:1459-1509), whose short- and long-range damping derivatives are
       -2*6 f_s t_s/(r (1 + t_s)) and -2*6 r f_l t_l/(70 (1 + t_l)) (t = the
       powered ratios); these are not the derivatives of f_s and f_l, so this
       gradient differs from finite differences of the energy along r_XB,
       exactly as the reference's does.  The out-of-line part is exact. */
#pragma omp for schedule(static) nowait
    for (int b = 0; b < P->nxb; ++b) {
        int A = P->xb_axb[3 * b], X = P->xb_axb[3 * b + 1], B = P->xb_axb[3 * b + 2];
        const double* p = P->xb_par + 3 * b;
        double vax[3], vxb[3], vab[3];
        sub3(x + 3 * X, x + 3 * A, vax);
        sub3(x + 3 * B, x + 3 * X, vxb);
        sub3(x + 3 * B, x + 3 * A, vab);
        double rax = norm3(vax), rxb = norm3(vxb), rab = norm3(vab);
        if (rxb > 20.0) continue;
        double radab = P->hb_rad[A] + P->hb_rad[B];
        double ts = pow(5.0 * radab / (rxb * rxb), 6.0), tl = pow(rxb * rxb / 70.0, 6.0);
        double fs = 1.0 / (1.0 + ts), fl = 1.0 / (1.0 + tl);
        double ex = 70.0 * ((rax + rxb) / rab - 1.0);
        if (ex > 15.0) continue;
        double fo = 2.0 / (1.0 + exp(ex));
        double eX = exp(15.0 * p[1]), eB = exp(15.0 * -p[2]);
        double QX = eX / (eX + 0.03), QB = eB / (eB + 0.03);
        double r3 = rxb * rxb * rxb;
        double Rd = fs * fl * fo / r3;
        double pre = -QB * p[0] * QX;
        T.e[GFNFF_E_XBOND] += pre * Rd;
        if (!want_grad) continue;
        double dfs = -2.0 * 6.0 * fs * ts / (rxb * (1.0 + ts));
        double dfl = -2.0 * 6.0 * rxb * fl * tl / (70.0 * (1.0 + tl));
        double et = exp(ex), de = (1.0 + et) * (1.0 + et);
        double dfo_ax = -2.0 * et * 70.0 / (rab * de);
        double dfo_ab = 2.0 * et * 70.0 * (rax + rxb) / (rab * rab * de);
        double dR_xb = (dfs * fl * fo + fs * dfl * fo + fs * fl * dfo_ax) / r3 - 3.0 * Rd / rxb;
        double dR_ax = fs * fl * dfo_ax / r3, dR_ab = fs * fl * dfo_ab / r3;
        double gxb = pre * dR_xb / (rxb + 1e-14), gax = pre * dR_ax / (rax + 1e-14), gab = pre * dR_ab / (rab + 1e-14);
        gadd(&G, GX, A, -gax * vax[0] - gab * vab[0], -gax * vax[1] - gab * vab[1], -gax * vax[2] - gab * vab[2]);
        gadd(&G, GX, X, gax * vax[0] - gxb * vxb[0], gax * vax[1] - gxb * vxb[1], gax * vax[2] - gxb * vxb[2]);
        gadd(&G, GX, B, gxb * vxb[0] + gab * vab[0], gxb * vxb[1] + gab * vab[1], gxb * vxb[2] + gab * vab[2]);
    }

    /* --- D4 ATM over bonded triplets (atm_terms) */
    atm_terms(P, x, gw, dgw, want_grad, &T, &G);

    /* --- bonded ATM: calcBATM ff_workspace_gfnff.cpp:1665-1739 */
#pragma omp for schedule(static) nowait
    for (int t = 0; t < P->nbatm; ++t) {
        int i = P->batm_ijk[3 * t], j = P->batm_ijk[3 * t + 1], k = P->batm_ijk[3 * t + 2];
        const double* zb = P->batm_par + 3 * t;
        double vij[3], vik[3], vjk[3];
        sub3(x + 3 * j, x + 3 * i, vij);
        sub3(x + 3 * k, x + 3 * i, vik);
        sub3(x + 3 * k, x + 3 * j, vjk);
        double r2ij = dot3(vij, vij), r2jk = dot3(vjk, vjk), r2ik = dot3(vik, vik);
        double rij = sqrt(r2ij), rjk = sqrt(r2jk), rik = sqrt(r2ik);
        double rijk3 = r2ij * r2jk * r2ik;
        double mijk = -r2ij + r2jk + r2ik, imjk = r2ij - r2jk + r2ik, ijmk = r2ij + r2jk - r2ik;
        double ang = 0.375 * ijmk * imjk * mijk / rijk3;
        double rav3 = pow(rijk3, 1.5);
        double angr9 = (ang + 1.0) / rav3;
        double fi = fmin(fmax(1.0 - 3.0 * P->qa[i], -4.0), 4.0);
        double fj = fmin(fmax(1.0 - 3.0 * P->qa[j], -4.0), 4.0);
        double fk = fmin(fmax(1.0 - 3.0 * P->qa[k], -4.0), 4.0);
        double c9 = fi * fj * fk * zb[0] * zb[1] * zb[2];
        T.e[GFNFF_E_BATM] += c9 * angr9;
        if (want_grad) {
            double dij = -0.375 * (pow(r2ij, 3) + pow(r2ij, 2) * (r2jk + r2ik) +
                                   r2ij * (3.0 * pow(r2jk, 2) + 2.0 * r2jk * r2ik + 3.0 * pow(r2ik, 2)) -
                                   5.0 * pow(r2jk - r2ik, 2) * (r2jk + r2ik)) / (rij * rijk3 * rav3);
            double djk = -0.375 * (pow(r2jk, 3) + pow(r2jk, 2) * (r2ik + r2ij) +
                                   r2jk * (3.0 * pow(r2ik, 2) + 2.0 * r2ik * r2ij + 3.0 * pow(r2ij, 2)) -
                                   5.0 * pow(r2ik - r2ij, 2) * (r2ik + r2ij)) / (rjk * rijk3 * rav3);
            double dik = -0.375 * (pow(r2ik, 3) + pow(r2ik, 2) * (r2jk + r2ij) +
                                   r2ik * (3.0 * pow(r2jk, 2) + 2.0 * r2jk * r2ij + 3.0 * pow(r2ij, 2)) -
                                   5.0 * pow(r2jk - r2ij, 2) * (r2jk + r2ij)) / (rik * rijk3 * rav3);
            double gij[3], gjk[3], gik[3];
            for (int c = 0; c < 3; ++c) {
                gij[c] = -dij * c9 * vij[c] / rij;
                gjk[c] = -djk * c9 * vjk[c] / rjk;
                gik[c] = -dik * c9 * vik[c] / rik;
            }
            gadd(&G, GBA, j, -gij[0] + gjk[0], -gij[1] + gjk[1], -gij[2] + gjk[2]);
            gadd(&G, GBA, k, -gik[0] - gjk[0], -gik[1] - gjk[1], -gik[2] - gjk[2]);
            gadd(&G, GBA, i, gij[0] + gik[0], gij[1] + gik[1], gij[2] + gik[2]);
        }
    }

    if (NT > 1) {
        tlT[tid] = T;
    } else {
        Tm = T;
    }
    } /* end of the parallel term section */
    if (NT > 1) {
        /* fixed-order reduce over the thread accumulators */
        for (int t = 0; t < NT; ++t) {
            for (int k = 0; k < GFNFF_NTERMS; ++k) Tm.e[k] += tlT[t].e[k];
            const double* a = tl + (size_t)t * 5 * n;
            for (size_t k = 0; k < 3 * (size_t)n; ++k) g[k] += a[k];
            for (int k = 0; k < n; ++k) {
                dEdcn_m[k] += a[3 * (size_t)n + k];
                dEdcn_bond_m[k] += a[4 * (size_t)n + k];
            }
        }
        free(tl);
        free(tlT);
    }

    /* --- CN chain rule + Coulomb term 1b: ff_workspace.cpp:459-507 */
    if (want_grad) {
        double* v = calloc(n, 8);
        double* vq = calloc(n, 8);
        for (int i = 0; i < n; ++i) {
            double c = fmax(cn[i], 0.0);
            vq[i] = q[i] * P->cnf[i] / (2.0 * sqrt(c) + 1e-16);
            v[i] = dEdcn_m[i] - vq[i];
        }
        double cut2 = 36.0;
        double* tg = (NT > 1) ? calloc((size_t)NT * 3 * n, 8) : NULL;
        if (NT > 1 && !tg) {
            free(v);
            free(vq);
            status = 1;
            goto done;
        }
        /* pairs in striped rows, per-thread gradients reduced in thread order */
#pragma omp parallel num_threads(NT) if (NT > 1)
        {
        const int tid = THREAD_ID();
        gacc_t G = Gm;
        const double* dEdcn = dEdcn_m;
        const double* dEdcn_bond = dEdcn_bond_m;
        if (NT > 1) {
            G.g = tg + (size_t)tid * 3 * n;
            for (int c = 0; c < NGC; ++c) G.gc[c] = NULL;
        }
#pragma omp for schedule(static, 1) nowait
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) {
                double d[3];
                sub3(x + 3 * j, x + 3 * i, d); /* r_j - r_i */
                double d2 = dot3(d, d);
                if (!(d2 < cut2)) continue;
                double r = sqrt(d2);
                double rs = P->rcov_cn[i] + P->rcov_cn[j];
                double dr = (r - rs) / rs;
                double dd = (KN / SQRTPI) * exp(-KN * KN * dr * dr) / rs;
                /* dCN_i/dr_i = -dlog_i*dd*u, dCN_j/dr_i = -dlog_j*dd*u, dCN/dr_j opposite */
                double s = dd * (dlog[i] * v[i] + dlog[j] * v[j]);
                double sb = dd * (dlog[i] * dEdcn_bond[i] + dlog[j] * dEdcn_bond[j]);
                double sd = dd * (dlog[i] * (dEdcn[i] - dEdcn_bond[i]) + dlog[j] * (dEdcn[j] - dEdcn_bond[j]));
                double sq = -dd * (dlog[i] * vq[i] + dlog[j] * vq[j]);
                double ux = d[0] / r, uy = d[1] / r, uz = d[2] / r;
                G.g[3 * i] -= s * ux;
                G.g[3 * i + 1] -= s * uy;
                G.g[3 * i + 2] -= s * uz;
                G.g[3 * j] += s * ux;
                G.g[3 * j + 1] += s * uy;
                G.g[3 * j + 2] += s * uz;
                if (G.gc[GB]) {
                    G.gc[GB][3 * i] -= sb * ux, G.gc[GB][3 * i + 1] -= sb * uy, G.gc[GB][3 * i + 2] -= sb * uz;
                    G.gc[GB][3 * j] += sb * ux, G.gc[GB][3 * j + 1] += sb * uy, G.gc[GB][3 * j + 2] += sb * uz;
                    G.gc[GD][3 * i] -= sd * ux, G.gc[GD][3 * i + 1] -= sd * uy, G.gc[GD][3 * i + 2] -= sd * uz;
                    G.gc[GD][3 * j] += sd * ux, G.gc[GD][3 * j + 1] += sd * uy, G.gc[GD][3 * j + 2] += sd * uz;
                    G.gc[GC][3 * i] -= sq * ux, G.gc[GC][3 * i + 1] -= sq * uy, G.gc[GC][3 * i + 2] -= sq * uz;
                    G.gc[GC][3 * j] += sq * ux, G.gc[GC][3 * j + 1] += sq * uy, G.gc[GC][3 * j + 2] += sq * uz;
                }
            }
        }
        if (NT > 1) {
            for (int t = 0; t < NT; ++t)
                for (size_t k = 0; k < 3 * (size_t)n; ++k) g[k] += tg[(size_t)t * 3 * n + k];
            free(tg);
        }
        free(v);
        free(vq);
    }

    {
        double tot = 0.0;
        for (int k = 0; k < GFNFF_NTERMS; ++k) tot += Tm.e[k];
        if (E_total) *E_total = tot;
        if (terms_out) memcpy(terms_out, Tm.e, sizeof(Tm.e));
        if (grad) memcpy(grad, g, sizeof(double) * 3 * (size_t)n);
        if (q_out) memcpy(q_out, q, sizeof(double) * n);
        if (cn_out) memcpy(cn_out, cn, sizeof(double) * n);
    }
done:
    free(cn_raw);
    free(cn);
    free(dlog);
    free(gw);
    free(dgw);
    free(q);
    free(dEdcn_m);
    free(dEdcn_bond_m);
    free(g);
    free(hbcn);
    return status;
}

/* The reference's serial evaluation (one thread). */
int gfnff_oracle_eval(const gfnff_params_view* P, const double* x, int want_grad, double* E_total, double* terms_out,
                      double* grad, double* q_out, double* cn_out, double* gcomp /* [NGC*3n] or NULL */) {
    return eval_impl(P, x, want_grad, 1, NULL, E_total, terms_out, grad, q_out, cn_out, gcomp);
}

/* Threaded evaluation (FFWorkspace scheme, nthr threads); q_in optional (see
 * eval_impl).  The CPU baseline and large-size checks use it. */
int gfnff_oracle_eval_mt(const gfnff_params_view* P, const double* x, int want_grad, int nthr, const double* q_in,
                         double* E_total, double* terms_out, double* grad, double* q_out, double* cn_out) {
    return eval_impl(P, x, want_grad, nthr, q_in, E_total, terms_out, grad, q_out, cn_out, NULL);
}

/* The ATM term alone at geometry x: with gw_in (natoms x 7 D4 weights) held
 * fixed, or with the weights of the CN at x when gw_in is NULL.  E receives
 * the ATM energy, grad (natoms x 3, may be NULL) the reference's analytic ATM
 * gradient, which does not include the CN dependence of C6. */
int gfnff_oracle_atm(const gfnff_params_view* P, const double* x, const double* gw_in, double* E, double* grad) {
    const int n = P->natoms;
    double *cn_raw = calloc(n, 8), *cn = calloc(n, 8), *dlog = calloc(n, 8), *gw = calloc(7 * (size_t)n, 8),
           *dgw = calloc(7 * (size_t)n, 8), *g = calloc(3 * (size_t)n, 8);
    int status = 0;
    if (!cn_raw || !cn || !dlog || !gw || !dgw || !g) {
        status = 1;
    } else {
        oracle_cn(P, x, 6.0, cn_raw, cn, dlog, 1);
        oracle_gw(P, cn, gw, dgw);
        if (gw_in) memcpy(gw, gw_in, sizeof(double) * 7 * (size_t)n);
        terms_t T;
        memset(&T, 0, sizeof(T));
        gacc_t G;
        G.n = n;
        G.g = g;
        for (int c = 0; c < NGC; ++c) G.gc[c] = NULL;
        atm_terms(P, x, gw, dgw, grad != NULL, &T, &G);
        *E = T.e[GFNFF_E_ATM];
        if (grad) memcpy(grad, g, sizeof(double) * 3 * (size_t)n);
    }
    free(cn_raw);
    free(cn);
    free(dlog);
    free(gw);
    free(dgw);
    free(g);
    return status;
}

/* CN (log-capped) and D4 Gaussian weights (natoms x 7) at geometry x. */
int gfnff_oracle_cn_gw(const gfnff_params_view* P, const double* x, double* cn_out, double* gw_out) {
    const int n = P->natoms;
    double *cn_raw = calloc(n, 8), *cn = calloc(n, 8), *dlog = calloc(n, 8), *gw = calloc(7 * (size_t)n, 8),
           *dgw = calloc(7 * (size_t)n, 8);
    int status = 0;
    if (!cn_raw || !cn || !dlog || !gw || !dgw) {
        status = 1;
    } else {
        oracle_cn(P, x, 6.0, cn_raw, cn, dlog, 1);
        oracle_gw(P, cn, gw, dgw);
        if (cn_out) memcpy(cn_out, cn, sizeof(double) * n);
        if (gw_out) memcpy(gw_out, gw, sizeof(double) * 7 * (size_t)n);
    }
    free(cn_raw);
    free(cn);
    free(dlog);
    free(gw);
    free(dgw);
    return status;
}

/* Relative residual of the constrained EEQ system at given charges q (a check
 * of the engine's solves, not part of an evaluation).  The bordered system
 * A q + C lambda = rhs, C^T q = qfrag (calculateFinalCharges,
 * eeq_solver.cpp:3571-3710) restricted to the null space of C^T, spanned by
 * e_a - e_last for every atom a of a fragment but its last (highest-index)
 * atom, has the residual r_a = v_a - v_last with v = rhs - A q.  Returns
 * |r|_2 / |b|_2 with b the same projection of rhs - A q_p, q_p_i =
 * qfrag_f / m_f (the reduced right-hand side), and the largest
 * |sum_{i in f} q_i - qfrag_f|.  cn: the log-capped coordination numbers of
 * the evaluation; rows of A are striped over nthr threads. */
int gfnff_oracle_eeq_residual(const gfnff_params_view* P, const double* x, const double* cn, const double* q,
                              int nthr, double* rel, double* qsum_err) {
    const int n = P->natoms, nf = P->nfrag;
    double* v = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    double* b = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    double* qp = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    int* cnt = (int*)calloc(nf > 0 ? nf : 1, sizeof(int));
    int* last = (int*)malloc(sizeof(int) * (nf > 0 ? nf : 1));
    double* qs = (double*)calloc(nf > 0 ? nf : 1, sizeof(double));
    int status = 1;
    if (!v || !b || !qp || !cnt || !last || !qs) goto out;
    for (int i = 0; i < n; ++i) {
        cnt[P->frag[i]]++;
        last[P->frag[i]] = i;
        qs[P->frag[i]] += q[i];
    }
    for (int i = 0; i < n; ++i) qp[i] = P->qfrag[P->frag[i]] / cnt[P->frag[i]];
#pragma omp parallel for schedule(static, 16) num_threads(nthr) if (nthr > 1)
    for (int i = 0; i < n; ++i) {
        double s = 0.0, sp = 0.0;
        for (int j = 0; j < n; ++j) {
            double a;
            if (j == i) {
                a = P->gam_eff[i] + TSQRT2PI / sqrt(P->alpeeq[i]);
            } else {
                double d[3];
                sub3(x + 3 * i, x + 3 * j, d);
                double r = norm3(d);
                a = erf(r / sqrt(P->alpeeq[i] + P->alpeeq[j])) / r;
            }
            s += a * q[j];
            sp += a * qp[j];
        }
        double rhs = P->chi_base[i] + P->cnf[i] * sqrt(cn[i]);
        v[i] = rhs - s;
        b[i] = rhs - sp;
    }
    {
        double rr = 0.0, bb = 0.0, qe = 0.0;
        for (int i = 0; i < n; ++i) {
            int l = last[P->frag[i]];
            if (i == l) continue;
            rr += (v[i] - v[l]) * (v[i] - v[l]);
            bb += (b[i] - b[l]) * (b[i] - b[l]);
        }
        for (int f = 0; f < nf; ++f) qe = fmax(qe, fabs(qs[f] - P->qfrag[f]));
        *rel = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
        *qsum_err = qe;
    }
    status = 0;
out:
    free(v);
    free(b);
    free(qp);
    free(cnt);
    free(last);
    free(qs);
    return status;
}

/* ------------------------------------------------------------------------- */
/* ALPB implicit solvation (alpb_solvation.cpp), serial FP64                  */
/* ------------------------------------------------------------------------- */
/* Constants of alpb_parameters.h:30-54. */
#define ALPB_W (0.3 / 0.52917726)
#define ALPB_AH0 0.5
#define ALPB_AH1 (3.0 / (4.0 * ALPB_W))
#define ALPB_AH3 (-1.0 / (4.0 * ALPB_W * ALPB_W * ALPB_W))
#define ALPB_ZETA 1.028
#define ALPB_TOLSESP 1.0e-6
#define OBC_ALPHA 1.0
#define OBC_BETA 0.8
#define OBC_GAMMA 4.85

typedef struct {
    int n;
    double* brad;  /* Born radii */
    double* brdr;  /* [n][3n]: d brad_i / d x_k at brdr[i*3n + 3k + c] */
    double* sasa;  /* surface area per atom (the grid-weighted switching) */
    double* dsdr;  /* [n][3n]: d sasa_i / d x_k */
    double* hbw;   /* H-bond diagonal hbmag sasa / vdwsa^2 */
    double* B;     /* [n][n] Born matrix incl. self, H-bond and shape terms */
    double adet, totrad3, center[3], inertia[9];
} alpb_state;

static void alpb_free(alpb_state* st) {
    free(st->brad);
    free(st->brdr);
    free(st->sasa);
    free(st->dsdr);
    free(st->hbw);
    free(st->B);
}

/* One descreening contribution to psi of atom s from atom o at distance r:
 * the non-overlapping or overlapping integral of born.f90 as restated in
 * alpb_solvation.cpp:530-589 (rho_o: descreening radius of o, rv: van der
 * Waals radius of s).  *dg: (d g / d r) / r. */
static double alpb_psi_term(double r, double rho_o, double rv, double* dg) {
    if (r >= rv + rho_o) {
        const double r1 = 1.0 / r, ap = r + rho_o, am = r - rho_o, ab = ap * am, rhab = rho_o / ab;
        const double lnab = 0.5 * log(am / ap) * r1;
        *dg = -2.0 * rhab / ab + (rhab - lnab) * r1 * r1;
        return rhab + lnab;
    }
    if (r + rho_o <= rv) { /* o lies inside s */
        *dg = 0.0;
        return 0.0;
    }
    const double r1 = 1.0 / r, r12 = 0.5 * r1, r24 = r12 * r12, ap = r + rho_o, am = r - rho_o;
    const double rh1 = 1.0 / rv, rhr1 = 1.0 / ap, aprh1 = ap * rh1, lnab = log(aprh1);
    double d = rhr1 * rhr1 * (1.0 - 0.25 * am * r1 * (1.0 + aprh1 * aprh1)) + rho_o * r24 * (rhr1 - rh1 * aprh1) +
               r12 * (r1 * lnab - rhr1);
    *dg = d * r1;
    return rh1 - rhr1 + r12 * (0.5 * am * (rhr1 - rh1 * aprh1) - lnab);
}

/* Born radii, surface areas, H-bond weights, shape descriptor and the Born
 * matrix at geometry x: ALPBSolvation::update (alpb_solvation.cpp:410-448)
 * with computePsi / computeBornRadii (:502-690), computeSASA (:696-801),
 * computeHBCorrection (:873-881), computeADet (:908-938) and
 * buildBornMatrix (:810-867, P16 kernel). */
static int alpb_update(const gfnff_alpb_view* S, const double* x, alpb_state* st) {
    const int n = S->natoms, n3 = 3 * n;
    memset(st, 0, sizeof(*st));
    st->n = n;
    st->brad = calloc(n, 8);
    st->brdr = calloc((size_t)n * n3, 8);
    st->sasa = calloc(n, 8);
    st->dsdr = calloc((size_t)n * n3, 8);
    st->hbw = calloc(n, 8);
    st->B = calloc((size_t)n * n, 8);
    if (!st->brad || !st->brdr || !st->sasa || !st->dsdr || !st->hbw || !st->B) return 0;
    const double lr2 = S->lrcut * S->lrcut, sr2 = S->srcut * S->srcut;
    /* psi and its derivative, pairs i > j within lrcut */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            double d[3];
            sub3(x + 3 * i, x + 3 * j, d);
            const double r2 = dot3(d, d);
            if (r2 >= lr2) continue;
            const double r = sqrt(r2);
            double dgi, dgj;
            st->brad[i] += alpb_psi_term(r, S->rho[j], S->vdwr[i], &dgi);
            st->brad[j] += alpb_psi_term(r, S->rho[i], S->vdwr[j], &dgj);
            for (int c = 0; c < 3; ++c) {
                /* s = i: v = dg (x_i - x_j); s = j: v = dg (x_j - x_i) */
                st->brdr[(size_t)i * n3 + 3 * i + c] += dgi * d[c];
                st->brdr[(size_t)i * n3 + 3 * j + c] -= dgi * d[c];
                st->brdr[(size_t)j * n3 + 3 * j + c] -= dgj * d[c];
                st->brdr[(size_t)j * n3 + 3 * i + c] += dgj * d[c];
            }
        }
    /* OBC-II transform psi -> Born radius (born.f90:79-112) */
    for (int i = 0; i < n; ++i) {
        const double s1 = 1.0 / S->svdw[i], v1 = 1.0 / S->vdwr[i], s2 = 0.5 * S->svdw[i];
        const double br = st->brad[i] * s2;
        double arg2 = br * (OBC_GAMMA * br - OBC_BETA);
        const double arg = br * (OBC_ALPHA + arg2);
        arg2 = 2.0 * arg2 + OBC_ALPHA + OBC_GAMMA * br * br;
        const double th = tanh(arg), ch = cosh(arg);
        st->brad[i] = S->born_scale / (s1 - v1 * th);
        double dp = ch * (s1 - v1 * th);
        dp = S->born_scale * s2 * v1 * arg2 / (dp * dp);
        for (int k = 0; k < n3; ++k) st->brdr[(size_t)i * n3 + k] *= dp;
    }
    /* surface areas: grid points on each atom's probe sphere, switched by
       every neighbour within srcut */
    int* idx = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    double* gv = (double*)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    if (!idx || !gv) {
        free(idx);
        free(gv);
        return 0;
    }
    for (int i = 0; i < n; ++i) {
        for (int p = 0; p < S->nang; ++p) {
            const double* gp = S->grid + 4 * p;
            double xp[3];
            for (int c = 0; c < 3; ++c) xp[c] = x[3 * i + c] + S->vdwsa[i] * gp[c];
            double sasap = 1.0;
            int nni = 0;
            for (int j = 0; j < n && sasap != 0.0; ++j) {
                if (j == i) continue;
                double dij[3];
                sub3(x + 3 * i, x + 3 * j, dij);
                if (!(dot3(dij, dij) < lr2 && dot3(dij, dij) < sr2)) continue;
                double t[3];
                sub3(xp, x + 3 * j, t);
                const double t2 = dot3(t, t);
                if (t2 >= S->trj_hi[j]) continue;
                if (t2 <= S->trj_lo[j]) {
                    sasap = 0.0;
                    break;
                }
                const double sq = sqrt(t2), uj = sq - S->vdwsa[j], ah3u2 = ALPB_AH3 * uj * uj;
                const double f = ALPB_AH0 + (ALPB_AH1 + ah3u2) * uj;
                const double df = (ALPB_AH1 + 3.0 * ah3u2) / (f * sq);
                sasap *= f;
                idx[nni] = j;
                for (int c = 0; c < 3; ++c) gv[3 * nni + c] = df * t[c];
                nni++;
            }
            if (sasap > ALPB_TOLSESP) {
                const double wsa = gp[3] * S->wrp[i] * sasap;
                st->sasa[i] += wsa;
                for (int k = 0; k < nni; ++k)
                    for (int c = 0; c < 3; ++c) {
                        st->dsdr[(size_t)i * n3 + 3 * i + c] += wsa * gv[3 * k + c];
                        st->dsdr[(size_t)i * n3 + 3 * idx[k] + c] -= wsa * gv[3 * k + c];
                    }
            }
        }
        if (S->hbond) st->hbw[i] = S->hbmag[i] * st->sasa[i] / (S->vdwsa[i] * S->vdwsa[i]);
    }
    free(idx);
    free(gv);
    /* shape descriptor: inertia of spheres of the van der Waals radii */
    {
        const double tof = 0.4;
        double c[3] = {0, 0, 0}, tr = 0.0, I[9] = {0};
        for (int i = 0; i < n; ++i) {
            const double r3 = S->vdwr[i] * S->vdwr[i] * S->vdwr[i];
            tr += r3;
            for (int k = 0; k < 3; ++k) c[k] += x[3 * i + k] * r3;
        }
        for (int k = 0; k < 3; ++k) c[k] /= tr;
        for (int i = 0; i < n; ++i) {
            const double r2v = S->vdwr[i] * S->vdwr[i], r3 = r2v * S->vdwr[i];
            double v[3];
            sub3(x + 3 * i, c, v);
            const double vv = dot3(v, v);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) I[3 * a + b] += r3 * ((a == b ? vv + tof * r2v : 0.0) - v[a] * v[b]);
        }
        const double det = I[0] * (I[4] * I[8] - I[5] * I[7]) - I[1] * (I[3] * I[8] - I[5] * I[6]) +
                           I[2] * (I[3] * I[7] - I[4] * I[6]);
        st->adet = sqrt(cbrt(det) / (tof * tr));
        st->totrad3 = tr;
        memcpy(st->center, c, sizeof(c));
        memcpy(st->inertia, I, sizeof(I));
    }
    /* Born matrix, P16 kernel over all pairs */
    const double shape = S->alpbet > 0.0 ? S->keps * S->alpbet / st->adet : 0.0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < i; ++j) {
            double d[3];
            sub3(x + 3 * i, x + 3 * j, d);
            const double r = norm3(d), ab = sqrt(st->brad[i] * st->brad[j]);
            double a = ab / (ab + ALPB_ZETA / 16.0 * r);
            a = a * a;
            a = a * a;
            a = a * a;
            a = a * a;
            const double v = S->keps / (r + ab * a) + shape;
            st->B[(size_t)i * n + j] = v;
            st->B[(size_t)j * n + i] = v;
        }
        st->B[(size_t)i * n + i] = S->keps / st->brad[i] + 2.0 * st->hbw[i] + shape;
    }
    return 1;
}

/* Solvation energy at charges q (ALPBSolvation::getEnergyParts,
 * alpb_solvation.cpp:1016-1041): parts = {gborn, ghb, gsasa, gshift}; and,
 * when grad != NULL, its gradient at fixed q added to grad
 * (ALPBSolvation::addGradient, :1139-1201 and :1265-1300). */
static double alpb_energy_grad(const gfnff_alpb_view* S, const double* x, const alpb_state* st, const double* q,
                               double* parts, double* grad) {
    const int n = S->natoms, n3 = 3 * n;
    double eb = 0.0, ehb = 0.0, es = 0.0, qt = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += st->B[(size_t)i * n + j] * q[j];
        eb += 0.5 * q[i] * s;
        ehb += st->hbw[i] * q[i] * q[i];
        es += S->gamsasa[i] * st->sasa[i];
        qt += q[i];
    }
    if (parts) {
        parts[0] = eb - ehb;
        parts[1] = ehb;
        parts[2] = es;
        parts[3] = S->gshift;
    }
    if (grad) {
        double* dEdbr = calloc(n, 8);
        if (!dEdbr) return NAN;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) {
                double d[3];
                sub3(x + 3 * i, x + 3 * j, d);
                const double r = norm3(d), ab = sqrt(st->brad[i] * st->brad[j]), qq = q[i] * q[j];
                const double a1 = ab / (ab + ALPB_ZETA / 16.0 * r);
                double a16 = a1 * a1;
                a16 = a16 * a16;
                a16 = a16 * a16;
                a16 = a16 * a16;
                const double fgb = r + ab * a16, df = 1.0 / fgb, df2 = df * df;
                const double ap = (1.0 - ALPB_ZETA * a1 * a16) * df2;
                for (int c = 0; c < 3; ++c) {
                    const double g = ap * d[c] * S->keps / r * qq;
                    grad[3 * i + c] -= g;
                    grad[3 * j + c] += g;
                }
                const double bp = -0.5 * (r * ALPB_ZETA / ab * a1 + 1.0) / ab * a16 * df2;
                dEdbr[i] += st->brad[j] * bp * S->keps * qq;
                dEdbr[j] += st->brad[i] * bp * S->keps * qq;
            }
        for (int i = 0; i < n; ++i) dEdbr[i] -= 0.5 * S->keps * q[i] * q[i] / (st->brad[i] * st->brad[i]);
        for (int i = 0; i < n; ++i) {
            /* surface tension and the H-bond term's surface dependence */
            const double w = S->gamsasa[i] + (S->hbond ? S->hbmag[i] / (S->vdwsa[i] * S->vdwsa[i]) * q[i] * q[i] : 0.0);
            for (int k = 0; k < n3; ++k)
                grad[k] += st->brdr[(size_t)i * n3 + k] * dEdbr[i] + st->dsdr[(size_t)i * n3 + k] * w;
        }
        free(dEdbr);
        if (S->alpbet > 0.0) {
            /* shape term 0.5 keps alpbet qt^2 / adet (gbsa.f90:722-789) */
            const double* I = st->inertia;
            double D[9];
            D[0] = I[0] * (I[4] + I[8]) - I[1] * I[1] - I[2] * I[2];
            D[1] = I[1] * I[8] - I[2] * I[5];
            D[2] = I[2] * I[4] - I[1] * I[7];
            D[3] = D[1];
            D[4] = I[4] * (I[0] + I[8]) - I[1] * I[1] - I[5] * I[5];
            D[5] = I[0] * I[5] - I[1] * I[2];
            D[6] = D[2];
            D[7] = D[5];
            D[8] = I[8] * (I[0] + I[4]) - I[2] * I[2] - I[5] * I[5];
            const double tr = st->totrad3, ad = st->adet;
            const double sc = (250.0 / (48.0 * tr * tr * tr * pow(ad, 5.0))) *
                              (-0.5 * S->keps * S->alpbet * qt * qt / (ad * ad));
            for (int i = 0; i < n; ++i) {
                const double r3 = S->vdwr[i] * S->vdwr[i] * S->vdwr[i];
                double v[3];
                sub3(x + 3 * i, st->center, v);
                for (int a = 0; a < 3; ++a) grad[3 * i + a] += r3 * sc * (D[3 * a] * v[0] + D[3 * a + 1] * v[1] + D[3 * a + 2] * v[2]);
            }
        }
    }
    return eb + es + S->gshift;
}

/* GFN-FF with ALPB solvation at geometry x (gfnff_method.cpp:1129-1140 and
 * 1805-1854): the Phase-2 EEQ solved with the Born matrix added to its
 * Coulomb matrix, the gas-phase terms at those charges, plus the solvation
 * energy; the gradient is the gas-phase gradient at those charges plus the
 * solvation gradient at fixed charges.  terms_out[GFNFF_E_SOLVATION] holds
 * the solvation energy; parts (4, may be NULL) = {gborn, ghb, gsasa, gshift}. */
int gfnff_oracle_eval_alpb(const gfnff_params_view* P, const gfnff_alpb_view* S, const double* x, int want_grad,
                           int nthr, double* E_total, double* terms_out, double* grad, double* q_out,
                           double* parts) {
    const int n = P->natoms;
    if (S->natoms != n) return 1;
    alpb_state st;
    memset(&st, 0, sizeof(st));
    double *cn_raw = calloc(n, 8), *cn = calloc(n, 8), *dlog = calloc(n, 8), *q = calloc(n, 8);
    double terms[GFNFF_NTERMS];
    int status = 1;
    if (!cn_raw || !cn || !dlog || !q || !alpb_update(S, x, &st)) goto out;
    oracle_cn(P, x, 6.0, cn_raw, cn, dlog, nthr);
    if (!oracle_eeq(P, x, cn, q, nthr, st.B)) {
        status = 3;
        goto out;
    }
    status = eval_impl(P, x, want_grad, nthr, q, E_total, terms, grad, q_out, NULL, NULL);
    if (status) goto out;
    {
        const double es = alpb_energy_grad(S, x, &st, q, parts, (want_grad && grad) ? grad : NULL);
        terms[GFNFF_E_SOLVATION] = es;
        *E_total += es;
    }
    if (terms_out) memcpy(terms_out, terms, sizeof(terms));
out:
    alpb_free(&st);
    free(cn_raw);
    free(cn);
    free(dlog);
    free(q);
    return status;
}
