// Host-side GFN-FF topology and parameter generation.  See gfnff_topology.hpp.
//
// Each block cites the reference function whose behaviour it reproduces; the
// arithmetic is kept in the reference's order where the result feeds a
// parameter that the reference's fixtures pin to 1e-12.

#include "gfnff_topology.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <numeric>
#include <queue>
#include <set>
#include <thread>
#include <tuple>

#include "gfnff_tables.inc"

namespace gfnff {

namespace {

using namespace gfnff_tab;

inline double sq(double x) { return x * x; }

struct Mol {
    int n;
    const int32_t* Z;
    const double* x;  // Bohr, row-major
    double dist(int i, int j) const {
        double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    double dist2(int i, int j) const {
        double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
        return dx * dx + dy * dy + dz * dz;
    }
};

// Topology gathered during the build (reference TopologyInfo, gfnff.h:295-1003).
struct Topo {
    std::vector<std::pair<int, int>> bonds;  // cached bond list, i<j, scan order
    std::vector<std::vector<int>> adj;       // neighbour lists in bond-list order
    std::vector<int> frag;                   // 1-based fragment ids
    int nfrag = 0;
    std::vector<double> qfrag;
    std::vector<double> cn;                  // topology CN (40 Bohr threshold)
    std::vector<int> hyb;
    std::vector<int> pi;                     // pi fragment id, 0 = none
    std::vector<std::vector<int>> rings;     // sorted atom sets, size 3..6
    std::vector<std::vector<int>> atom_rings;
    std::vector<int> ring_size;              // smallest ring per atom
    std::vector<bool> is_metal;
    std::vector<double> qa, dxi, dgam, alpeeq;
    std::vector<bool> amide_h;
    std::vector<int> btype;                  // per cached bond
    std::vector<std::set<std::pair<int, int>>> dummy;
    std::set<std::pair<int, int>> bond_set;
    // sparse topological distances up to 3 bonds
    std::vector<std::vector<std::pair<int, int>>> tnb;
    bool bonded(int a, int b) const { return bond_set.count({std::min(a, b), std::max(a, b)}) > 0; }
    int tdist(int a, int b) const {
        for (auto& p : tnb[a])
            if (p.first == b) return p.second;
        return a == b ? 0 : 999;
    }
};

int periodic_row(int z) {
    if (z <= 0) return 0;
    if (z <= 2) return 1;
    if (z <= 10) return 2;
    if (z <= 18) return 3;
    if (z <= 36) return 4;
    if (z <= 54) return 5;
    return 6;
}

double rad_bohr(int z) {  // GFNFF::getCovalentRadius, gfnff_method.cpp:4129
    return (z >= 1 && z <= 86 ? rad[z - 1] : 1.0) * kAngToBohrGeom;
}

// CN radius in Bohr: 4/3 * r_cov(Angstrom) * 1.8897259886 (cn_calculator.cpp:284,375)
double rcov_cn_of(int z) {
    if (z < 1 || z > 86) return 0.0;
    return 4.0 / 3.0 * rcov_cn[z - 1] * kAngToBohrCN;
}

// ---------------------------------------------------------------------------
// Bond detection: getCachedBondList, gfnff_method.cpp:940-1000
// ---------------------------------------------------------------------------
void detect_bonds(const Mol& m, const int32_t* forced, int nforced, Topo& t) {
    t.bonds.clear();
    if (nforced > 0) {
        for (int b = 0; b < nforced; ++b) {
            int a = forced[2 * b], c = forced[2 * b + 1];
            auto p = a < c ? std::make_pair(a, c) : std::make_pair(c, a);
            if (p.first >= 0 && p.second < m.n && p.first != p.second) t.bonds.push_back(p);
        }
    } else {
        std::vector<double> rc(m.n), fv(m.n);
        for (int i = 0; i < m.n; ++i) {
            rc[i] = rad_bohr(m.Z[i]);
            fv[i] = (m.Z[i] >= 0 && m.Z[i] <= 86) ? fat[m.Z[i]] : 1.0;
        }
        for (int i = 0; i < m.n; ++i)
            for (int j = i + 1; j < m.n; ++j) {
                double thr = 1.3 * (rc[i] + rc[j]) * fv[i] * fv[j];
                if (m.dist(i, j) < thr) t.bonds.emplace_back(i, j);
            }
    }
    t.adj.assign(m.n, {});
    for (auto& b : t.bonds) {
        t.adj[b.first].push_back(b.second);
        t.adj[b.second].push_back(b.first);
        t.bond_set.insert(b);
    }
}

// detectMolecularFragments, gfnff_method.cpp:8772-8811 (BFS, 1-based ids)
void detect_fragments(int n, Topo& t) {
    t.frag.assign(n, 0);
    t.nfrag = 0;
    for (int i = 0; i < n; ++i) {
        if (t.frag[i]) continue;
        ++t.nfrag;
        std::queue<int> q;
        q.push(i);
        t.frag[i] = t.nfrag;
        while (!q.empty()) {
            int c = q.front();
            q.pop();
            for (int nb : t.adj[c])
                if (!t.frag[nb]) {
                    t.frag[nb] = t.nfrag;
                    q.push(nb);
                }
        }
    }
}

// CNCalculator::calculateGFNFFCN threshold variant, cn_calculator.cpp:120-160
void topo_cn(const Mol& m, Topo& t) {
    const double kn = -7.5, cnmax = 4.4, thr = 1600.0;
    t.cn.assign(m.n, 0.0);
    std::vector<double> rc(m.n);
    for (int i = 0; i < m.n; ++i) rc[i] = rcov_cn_of(m.Z[i]);
    for (int i = 0; i < m.n; ++i) {
        if (rc[i] == 0.0) continue;
        double raw = 0.0;
        for (int j = 0; j < m.n; ++j) {
            if (i == j || rc[j] == 0.0) continue;
            double d2 = m.dist2(i, j);
            if (d2 > thr) continue;
            double d = std::sqrt(d2);
            double rij = rc[i] + rc[j];
            double dr = (d - rij) / rij;
            raw += 0.5 * (1.0 + std::erf(kn * dr));
        }
        t.cn[i] = std::log(1.0 + std::exp(cnmax)) - std::log(1.0 + std::exp(cnmax - raw));
    }
}

// determineHybridization, gfnff_method.cpp:6446-6586
void hybridization(const Mol& m, Topo& t) {
    t.hyb.assign(m.n, 3);
    for (int i = 0; i < m.n; ++i) {
        int z = m.Z[i];
        const auto& nb = t.adj[i];
        std::vector<std::array<double, 3>> bv;
        for (int j : nb) {
            double v[3] = {m.x[3 * j] - m.x[3 * i], m.x[3 * j + 1] - m.x[3 * i + 1], m.x[3 * j + 2] - m.x[3 * i + 2]};
            double nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            bv.push_back({v[0] / nv, v[1] / nv, v[2] / nv});
        }
        auto dot = [](const std::array<double, 3>& a, const std::array<double, 3>& b) {
            return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
        };
        int nc = (int)nb.size();
        if (nc == 0) {
            t.hyb[i] = 3;
        } else if (nc == 1) {
            if (z == 8 || z == 16) {
                t.hyb[i] = (t.adj[nb[0]].size() == 1) ? 1 : 2;
            } else if (z == 1 || z == 9 || z == 17 || z == 35 || z == 53) {
                t.hyb[i] = 0;
            } else {
                t.hyb[i] = 1;
            }
        } else if (nc == 2) {
            if (z == 8 || z == 16) {
                t.hyb[i] = 3;
            } else if (z == 9 || z == 17 || z == 35 || z == 53) {
                t.hyb[i] = 1;
            } else {
                double ang = std::acos(std::max(-1.0, std::min(1.0, dot(bv[0], bv[1]))));
                t.hyb[i] = ang > 2.8 ? 1 : 2;
            }
        } else if (nc == 3) {
            if (z == 7) {
                t.hyb[i] = 3;
            } else {
                double s = 0.0;
                for (int a = 0; a < 3; ++a) s += std::acos(std::max(-1.0, std::min(1.0, dot(bv[a], bv[(a + 1) % 3]))));
                t.hyb[i] = s > 6.0 ? 2 : 3;
            }
        } else {
            t.hyb[i] = 3;
        }
    }
}

// detectPiSystems, gfnff_method.cpp:6588-6696
void pi_systems(const Mol& m, Topo& t) {
    auto hoff = [](int z) { return z == 5 || z == 6 || z == 7 || z == 8 || z == 9 || z == 16 || z == 17; };
    auto picon = [](int z) { return z == 5 || z == 7 || z == 8 || z == 9 || z == 16 || z == 17; };
    std::vector<std::vector<int>> pn(m.n);
    for (int i = 0; i < m.n; ++i) {
        if (!hoff(m.Z[i])) continue;
        for (int j : t.adj[i]) {
            if (j <= i || !hoff(m.Z[j])) continue;
            bool hi = t.hyb[i] == 1 || t.hyb[i] == 2, hj = t.hyb[j] == 1 || t.hyb[j] == 2;
            bool pb = (hi && hj) || (hi && picon(m.Z[j])) || (hj && picon(m.Z[i]));
            if (pb) {
                pn[i].push_back(j);
                pn[j].push_back(i);
            }
        }
    }
    t.pi.assign(m.n, 0);
    std::vector<bool> vis(m.n, false);
    int id = 1;
    for (int s = 0; s < m.n; ++s) {
        if (pn[s].empty() || vis[s]) continue;
        std::vector<int> st{s};
        vis[s] = true;
        while (!st.empty()) {
            int c = st.back();
            st.pop_back();
            t.pi[c] = id;
            for (int nb : pn[c])
                if (!vis[nb]) {
                    vis[nb] = true;
                    st.push_back(nb);
                }
        }
        ++id;
    }
}

// findSmallestRings, gfnff_method.cpp:6698-6796 (all simple rings of size 3..6)
void rings(int n, Topo& t) {
    std::set<std::vector<int>> uniq;
    std::vector<int> path;
    for (int u = 0; u < n; ++u) {
        for (int v : t.adj[u]) {
            if (v <= u) continue;
            path.assign({u, v});
            std::function<void(int, int)> dfs = [&](int cur, int depth) {
                if (depth >= 6) return;
                for (int nx : t.adj[cur]) {
                    if (nx == u && depth >= 2) {
                        std::vector<int> r = path;
                        std::sort(r.begin(), r.end());
                        uniq.insert(r);
                        continue;
                    }
                    if (nx == u) continue;
                    if (std::find(path.begin(), path.end(), nx) != path.end()) continue;
                    path.push_back(nx);
                    dfs(nx, depth + 1);
                    path.pop_back();
                }
            };
            dfs(v, 1);
        }
    }
    t.rings.assign(uniq.begin(), uniq.end());
    t.atom_rings.assign(n, {});
    for (int r = 0; r < (int)t.rings.size(); ++r)
        for (int a : t.rings[r]) t.atom_rings[a].push_back(r);
    t.ring_size.assign(n, 0);
    for (int a = 0; a < n; ++a)
        for (int r : t.atom_rings[a]) {
            int s = (int)t.rings[r].size();
            if (t.ring_size[a] == 0 || s < t.ring_size[a]) t.ring_size[a] = s;
        }
}

bool in_ring(const std::vector<int>& r, int a) { return std::binary_search(r.begin(), r.end(), a); }

// areAtomsInSameRing, gfnff_method.cpp:6798-6837
int same_ring(const Topo& t, int i, int j) {
    int s = 0;
    for (int r : t.atom_rings[i])
        if (in_ring(t.rings[r], j)) {
            int z = (int)t.rings[r].size();
            if (s == 0 || z < s) s = z;
        }
    return s;
}

// smallestRingContainingBend, gfnff_method.cpp:6870-6936
int ring_bend(const Topo& t, int i, int j, int k) {
    if (t.atom_rings[i].empty() || t.atom_rings[j].empty() || t.atom_rings[k].empty()) return 0;
    int s = 0;
    for (int r : t.atom_rings[i])
        if (in_ring(t.rings[r], j) && in_ring(t.rings[r], k)) {
            int z = (int)t.rings[r].size();
            if (s == 0 || z < s) s = z;
        }
    return s;
}

// smallest/largestRingContainingAll, gfnff_method.cpp:6839-6966
int ring_all(const Topo& t, int i, int j, int k, int l, bool largest) {
    int s = 0;
    for (int r : t.atom_rings[j]) {
        const auto& R = t.rings[r];
        if (in_ring(R, i) && in_ring(R, k) && in_ring(R, l)) {
            int z = (int)R.size();
            if (largest) {
                if (z > s) s = z;
            } else if (s == 0 || z < s) {
                s = z;
            }
        }
    }
    return s;
}

// EEQSolver::detectPiSystem, eeq_solver.cpp:4483-4516
std::vector<bool> eeq_pi_atoms(const Mol& m, const Topo& t) {
    std::vector<bool> p(m.n, false);
    auto pe = [](int z) { return z == 6 || z == 7 || z == 8 || z == 9 || z == 16; };
    for (int i = 0; i < m.n; ++i) {
        if ((t.hyb[i] == 1 || t.hyb[i] == 2) && pe(m.Z[i])) p[i] = true;
        if (t.hyb[i] == 3 && (m.Z[i] == 7 || m.Z[i] == 8 || m.Z[i] == 9))
            for (int j : t.adj[i])
                if (t.hyb[j] == 1 || t.hyb[j] == 2) {
                    p[i] = true;
                    break;
                }
    }
    return p;
}

// EEQSolver::detectAmideNitrogens / detectAmideHydrogens, eeq_solver.cpp:4518-4603.
// The index comparisons against npiall reproduce the reference exactly.
std::vector<bool> eeq_amide_n(const Mol& m, const Topo& t, const std::vector<bool>& pia) {
    std::vector<bool> am(m.n, false);
    int npiall = 0;
    for (bool b : pia) npiall += b ? 1 : 0;
    for (int i = 0; i < m.n; ++i) {
        if (m.Z[i] != 7 || i >= npiall) continue;
        if (t.hyb[i] != 3) continue;
        int nc = 0, ic = -1;
        for (int nb : t.adj[i])
            if (m.Z[nb] == 6 && nb < npiall) {
                ++nc;
                ic = nb;
            }
        if (nc != 1) continue;
        int no = 0;
        for (int n2 : t.adj[ic])
            if (m.Z[n2] == 8 && n2 < npiall && t.adj[n2].size() == 1) ++no;
        if (no == 1) am[i] = true;
    }
    return am;
}

std::vector<bool> eeq_amide_h(const Mol& m, const Topo& t, const std::vector<bool>& am) {
    std::vector<bool> ah(m.n, false);
    for (int i = 0; i < m.n; ++i) {
        if (m.Z[i] != 1) continue;
        for (int nb : t.adj[i]) {
            if (m.Z[nb] == 7 && am[nb]) {
                int c = 0;
                for (int n2 : t.adj[nb])
                    if (m.Z[n2] == 6 && t.hyb[n2] == 3) ++c;
                if (c == 1) ah[i] = true;
                break;
            }
        }
    }
    return ah;
}

// EEQSolver::calculateDxi, eeq_solver.cpp:4059-4299 (Pauling EN average is
// diagnostic only in the reference and not reproduced).
std::vector<double> calc_dxi(const Mol& m, const Topo& t) {
    std::vector<double> dxi(m.n, 0.0);
    std::vector<bool> pia = eeq_pi_atoms(m, t);
    for (int i = 0; i < m.n; ++i) {
        int z = m.Z[i];
        double d = 0.0;
        int nh = 0, nn = (int)t.adj[i].size(), nm = 0;
        for (int j : t.adj[i]) {
            int zj = m.Z[j];
            if (zj == 1) ++nh;
            if (zj > 20 && (zj <= 30 || (zj >= 39 && zj <= 48) || (zj >= 72 && zj <= 80))) ++nm;
        }
        if (z == 5 && nh > 0) d += nh * 0.015;
        if (z == 6 && nn == 2) {
            int a = t.adj[i][0], b = t.adj[i][1];
            double v1[3], v2[3];
            for (int c = 0; c < 3; ++c) {
                v1[c] = m.x[3 * a + c] - m.x[3 * i + c];
                v2[c] = m.x[3 * b + c] - m.x[3 * i + c];
            }
            double r1 = std::sqrt(sq(v1[0]) + sq(v1[1]) + sq(v1[2]));
            double r2 = std::sqrt(sq(v2[0]) + sq(v2[1]) + sq(v2[2]));
            if (r1 > 1e-10 && r2 > 1e-10) {
                double ca = (v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]) / (r1 * r2);
                ca = std::max(-1.0, std::min(1.0, ca));
                if (std::acos(ca) * 180.0 / M_PI < 150.0) d += -0.15;
            }
        }
        if (z == 8) {
            if (nn == 1 && pia[i])
                for (int j : t.adj[i])
                    if (m.Z[j] == 7 && pia[j]) d += 0.05;
            if (nn == 1)
                for (int j : t.adj[i])
                    if (m.Z[j] == 6 && t.adj[j].size() == 1) d += 0.15;
            if (nn == 2 && nh == 2) d += -0.02;
            if (nh > 0) d += -nh * 0.005;
            if (nn > 2) d += nn * 0.005;
        }
        if (z == 16) {
            if (nh > 0) d += -nh * 0.005;
            if (nn > 2) d += nn * 0.005;
        }
        int grp = (z >= 1 && z <= 86) ? periodic_group[z - 1] : 0;
        if (z > 9 && grp == 7 && nn > 1) d += (nm == 0) ? -nn * 0.021 : nn * 0.05;
        dxi[i] = d;
    }
    return dxi;
}

// EEQSolver::calculateDgam, eeq_solver.cpp:4301-4387
std::vector<double> calc_dgam(const Mol& m, const Topo& t, const std::vector<double>& qa) {
    std::vector<bool> pia = eeq_pi_atoms(m, t);
    std::vector<bool> am = eeq_amide_n(m, t, pia);
    int npiall = 0;
    for (bool b : pia) npiall += b ? 1 : 0;
    std::vector<double> dg(m.n, 0.0);
    for (int i = 0; i < m.n; ++i) {
        int Z = m.Z[i];
        double ff = 0.0;
        if (Z == 1) ff = -0.08;
        else if (Z == 5) ff = -0.05;
        else if (Z == 6) {
            ff = -0.27;
            if (t.hyb[i] < 3) ff = -0.45;
            if (t.hyb[i] < 2) ff = -0.34;
        } else if (Z == 7) {
            ff = -0.13;
            if (i < npiall) ff = -0.14;
            if (am[i]) ff = -0.16;
        } else if (Z == 8) {
            ff = -0.15;
            if (t.hyb[i] < 3) ff = -0.08;
        } else if (Z == 9) {
            ff = 0.10;
        } else if (Z > 10) {
            ff = -0.02;
            if (Z == 17) ff = -0.02;
            if (Z == 35) ff = -0.11;
            if (Z == 53) ff = -0.07;
            if (Z <= 86) {
                if (metal_type[Z - 1] == 1) ff = -0.08;
                if (metal_type[Z - 1] == 2) ff = -0.9;
                if (periodic_group[Z - 1] == 8) ff = 0.0;
            }
        }
        dg[i] = qa[i] * ff;
    }
    return dg;
}

// calculateAlpeeq, gfnff_method.cpp:11182-11276
std::vector<double> calc_alpeeq(const Mol& m, const Topo& t) {
    std::vector<double> a(m.n);
    for (int i = 0; i < m.n; ++i) {
        int z = m.Z[i];
        double base = (z >= 1 && z <= 86) ? alp[z - 1] : 0.903430;
        double ff = 0.0;
        if (z == 6) ff = 0.09;
        else if (z == 7) ff = -0.21;
        else if (z >= 1 && z <= 86) {
            int g = periodic_group[z - 1];
            if (g == 6) ff = -0.03;
            else if (g == 7) ff = 0.50;
            else if (t.is_metal[i]) {
                if (metal_type[z - 1] == 2) ff = -0.1;
                else if (metal_type[z - 1] == 1) ff = 0.3;
            }
        }
        a[i] = std::pow(base + ff * t.qa[i], 2);
    }
    return a;
}

// EEQ Phase 1 with topological distances: calculateTopologyCharges,
// eeq_solver.cpp:2684-3003, distances by computeTopologicalDistancesSparse
// (:3093-3199, float32 Dijkstra over Pyykko radii).
bool phase1_charges(const Mol& m, Topo& t, int threads, std::string& msg) {
    const int n = m.n;
    const double TSQRT2PI = 0.797884560802866;
    const float RABD_CUT = 13.0f, TDIST_THR = 12.0f;
    const double RFGOED1 = 1.175, BOHR = 0.52917726;
    std::vector<double> dxi = calc_dxi(m, t);
    std::vector<double> chi_v(n), gam_v(n), alp_v(n);
    for (int i = 0; i < n; ++i) {
        int z = m.Z[i];
        double nb = std::min((double)t.adj[i].size(), 4.4);
        double c = (z >= 1 && z <= 86) ? chi[z - 1] : 1.0;
        double cf = (z >= 1 && z <= 86) ? cnf[z - 1] : 0.0;
        chi_v[i] = -c + dxi[i] + cf * std::sqrt(nb);
        gam_v[i] = (z >= 1 && z <= 86) ? gam[z - 1] : 0.0;
        double a = (z >= 1 && z <= 86) ? alp[z - 1] : 1.0;
        alp_v[i] = a * a;
    }
    std::vector<float> radf(n);
    for (int i = 0; i < n; ++i) radf[i] = (float)((m.Z[i] >= 1 && m.Z[i] <= 86) ? rad[m.Z[i] - 1] : 1.0);
    // Matrix rows over an atom subset `sub` (local indices); Dijkstra from each
    // source over the bond graph, capped as in the reference.
    auto build = [&](const std::vector<int>& sub, std::vector<double>& A) {
        const int ns = (int)sub.size();
        A.assign((size_t)ns * ns, 0.0);
        std::vector<int> loc(n, -1);
        for (int k = 0; k < ns; ++k) loc[sub[k]] = k;
        auto row_worker = [&](int t0, int T) {
            std::vector<float> dist(n, RABD_CUT);
            std::vector<int> touched;
            for (int ls = t0; ls < ns; ls += T) {
                int src = sub[ls];
                for (int v : touched) dist[v] = RABD_CUT;
                touched.assign(1, src);
                dist[src] = 0.0f;
                using E = std::pair<float, int>;
                std::priority_queue<E, std::vector<E>, std::greater<E>> pq;
                pq.push({0.0f, src});
                while (!pq.empty()) {
                    auto [d, u] = pq.top();
                    pq.pop();
                    if (d > dist[u]) continue;
                    if (d > TDIST_THR) continue;
                    for (int v : t.adj[u]) {
                        float w = radf[u] + radf[v];
                        float nd = d + w;
                        if (nd < dist[v]) {
                            if (dist[v] == RABD_CUT) touched.push_back(v);
                            dist[v] = nd;
                            pq.push({nd, v});
                        }
                    }
                }
                double* Ai = &A[(size_t)ls * ns];
                Ai[ls] = gam_v[src] + TSQRT2PI / std::sqrt(alp_v[src]);
                for (int lj = 0; lj < ls; ++lj) {
                    int j = sub[lj];
                    float rij = dist[j];
                    double val = (rij > TDIST_THR) ? (double)RABD_CUT : (double)rij;
                    double r = RFGOED1 * val / BOHR;
                    double g = 1.0 / std::sqrt(alp_v[src] + alp_v[j]);
                    Ai[lj] = std::erf(g * r) / r;
                }
            }
        };
        int T = std::max(1, std::min(threads, ns));
        std::vector<std::thread> th;
        for (int k = 1; k < T; ++k) th.emplace_back(row_worker, k, T);
        row_worker(0, T);
        for (auto& x : th) x.join();
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < i; ++j) A[(size_t)j * ns + i] = A[(size_t)i * ns + j];
    };
    // Atoms of different fragments are at the capped topological distance
    // r_cap = 1.175*13/0.52917726 Bohr.  When erf(gamma_ij r_cap) is exactly 1
    // for every pair, the inter-fragment coupling is the constant 1/r_cap,
    // which within a fragment is a constant shift absorbed by that fragment's
    // Lagrange multiplier (every fragment's charge is fixed), so the bordered
    // system decouples into independent per-fragment systems exactly.
    const double rcap = RFGOED1 * (double)RABD_CUT / BOHR;
    double amax = 0.0;
    for (double a : alp_v) amax = std::max(amax, a);
    bool decoupled = t.nfrag > 1 && std::erf(rcap / std::sqrt(2.0 * amax)) == 1.0;
    if (decoupled) {
        std::vector<std::vector<int>> fa(t.nfrag);
        for (int i = 0; i < n; ++i) fa[t.frag[i] - 1].push_back(i);
        t.qa.assign(n, 0.0);
        for (int f = 0; f < t.nfrag; ++f) {
            const auto& sub = fa[f];
            std::vector<double> A, xs(sub.size()), qs;
            for (size_t k = 0; k < sub.size(); ++k) xs[k] = chi_v[sub[k]];
            build(sub, A);
            std::vector<int32_t> fr(sub.size(), 0);
            std::vector<double> qf{t.qfrag[f]};
            if (!eeq_solve_host((int)sub.size(), 1, A, xs, fr, qf, qs, 1)) {
                msg = "Phase-1 EEQ solve failed";
                return false;
            }
            for (size_t k = 0; k < sub.size(); ++k) t.qa[sub[k]] = qs[k];
        }
    } else {
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        std::vector<double> A;
        build(all, A);
        std::vector<int32_t> fr(n);
        for (int i = 0; i < n; ++i) fr[i] = t.frag[i] - 1;
        if (!eeq_solve_host(n, t.nfrag, A, chi_v, fr, t.qfrag, t.qa, threads)) {
            msg = "Phase-1 EEQ solve failed";
            return false;
        }
    }
    for (double q : t.qa)
        if (!std::isfinite(q)) {
            msg = "Phase-1 EEQ produced non-finite charges";
            return false;
        }
    return true;
}

// ---------------------------------------------------------------------------
// Hueckel pi bond orders (HuckelSolver, huckel_solver.cpp:32-565)
// ---------------------------------------------------------------------------

// Symmetric eigenproblem by cyclic Jacobi rotations (Golub & Van Loan, Matrix
// Computations, sec. 8.5, sym.schur2 + cyclic sweep).  a: n x n row-major,
// destroyed.  Returns eigenvalues ascending in w and the matching
// eigenvectors as the columns of v (row-major n x n).
void sym_eig(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& v) {
    v.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
    auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int i = 0; i < n; ++i) {
            dia += A(i, i) * A(i, i);
            for (int j = i + 1; j < n; ++j) off += A(i, j) * A(i, j);
        }
        if (off <= 1e-32 * (dia + 1e-300)) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A(p, q);
                if (apq == 0.0) continue;
                double tau = (A(q, q) - A(p, p)) / (2.0 * apq);
                double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
                double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < n; ++k) {  // A <- A J
                    double akp = A(k, p), akq = A(k, q);
                    A(k, p) = c * akp - s * akq;
                    A(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // A <- J^T A
                    double apk = A(p, k), aqk = A(q, k);
                    A(p, k) = c * apk - s * aqk;
                    A(q, k) = s * apk + c * aqk;
                }
                A(p, q) = A(q, p) = 0.0;
                for (int k = 0; k < n; ++k) {  // V <- V J
                    double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
                    v[(size_t)k * n + p] = c * vkp - s * vkq;
                    v[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return A(x, x) < A(y, y); });
    w.resize(n);
    std::vector<double> vs((size_t)n * n);
    for (int c = 0; c < n; ++c) {
        w[c] = A(ord[c], ord[c]);
        for (int r = 0; r < n; ++r) vs[(size_t)r * n + c] = v[(size_t)r * n + ord[c]];
    }
    v.swap(vs);
}

// pi electrons an atom brings into its pi system (countPiElectrons,
// huckel_solver.cpp:197-266; the reference passes itag = 0 for every atom,
// gfnff_method.cpp:7397-7398)
int pi_electrons(int z, int hyb) {
    switch (z) {
        case 5: return hyb == 1 ? 1 : 0;
        case 6: return 1;
        case 7: return hyb <= 2 ? 1 : (hyb == 3 ? 2 : 0);
        case 8:
        case 16: return (hyb == 1 || hyb == 2) ? 1 : (hyb == 3 ? 2 : 0);
        case 9: return hyb != 1 ? 2 : 3;
        case 17: return hyb == 0 ? 2 : (hyb == 1 ? 3 : 0);
        default: return 0;
    }
}

// Closed-shell occupations of one pi system: Fermi smearing at 4000 K per spin
// channel, doubled, with the integer fallback for a degenerate HOMO/LUMO
// (solveAndBuildDensity + fermiSmear, huckel_solver.cpp:374-514).  e in eV,
// ascending.  Returns sum(occ * e).
double pi_occupations(const std::vector<double>& e, int nel, std::vector<double>& occ) {
    const int no = (int)e.size();
    const int na = nel / 2;
    const double bkt = 8.617333262e-5 * 4000.0;
    occ.assign(no, 0.0);
    if (na > 0) {
        int ih = std::min(na, no) - 1, il = std::min(na, no - 1);
        double ef = 0.5 * (e[ih] + e[il]);
        for (int cyc = 0; cyc < 200; ++cyc) {
            double tot = 0.0, der = 0.0;
            for (int i = 0; i < no; ++i) {
                double x = (e[i] - ef) / bkt, f = 0.0, d = 0.0;
                if (x < 50.0) {
                    double ex = std::exp(x);
                    f = 1.0 / (ex + 1.0);
                    d = ex / (bkt * (ex + 1.0) * (ex + 1.0));
                }
                occ[i] = f;
                tot += f;
                der += d;
            }
            ef += ((double)na - tot) / der;
            if (std::abs((double)na - tot) < 1e-9) break;
        }
    }
    for (double& o : occ) o *= 2.0;
    if (na > 0 && na < no && std::abs(occ[na - 1] - occ[na]) < 1e-4) {
        std::fill(occ.begin(), occ.end(), 0.0);
        for (int i = 0; i < na; ++i) occ[i] = 2.0;
    }
    double el = 0.0;
    for (int i = 0; i < no; ++i) el += occ[i] * e[i];
    return el;
}

// Iterative Hueckel per pi system (calculatePiBondOrders, huckel_solver.cpp:32-196;
// Hamiltonian :271-369).  Returns the pi bond order of every cached bond
// (0 for bonds outside a pi system).  ipis[s] = charge localised on pi system s.
std::vector<double> hueckel_pibo(const Mol& m, const Topo& t, const std::vector<int>& ipis) {
    static const double hdiag[18] = {0, 0, 0, 0, 0, -0.5, 0, 0.14, -0.38, -0.29, 0, 0, 0, 0, 0, 0, -0.30, -0.30};
    static const double hoff[18] = {0, 0, 0, 0, 0, 0.5, 1.0, 0.66, 1.1, 0.23, 0, 0, 0, 0, 0, 0, 0.6, 1.0};
    const double hiter = 0.700, htriple = 1.45, hp3 = -0.24, pilpf = 0.530, ev = 0.1 * 27.2113957;
    std::vector<double> pibo(t.bonds.size(), 0.0);
    int maxid = 0;
    for (int p : t.pi) maxid = std::max(maxid, p);
    std::vector<int> loc(m.n, -1);
    for (int s = 1; s <= maxid; ++s) {
        std::vector<int> at;
        int nel = 0;
        for (int k = 0; k < m.n; ++k)
            if (t.pi[k] == s) {
                loc[k] = (int)at.size();
                at.push_back(k);
                nel += pi_electrons(m.Z[k], t.hyb[k]);
            }
        if (s < (int)ipis.size()) nel -= ipis[s];
        const int np = (int)at.size();
        std::vector<int> bl;  // cached bonds inside this pi system
        for (size_t b = 0; b < t.bonds.size(); ++b)
            if (loc[t.bonds[b].first] >= 0 && loc[t.bonds[b].second] >= 0) bl.push_back((int)b);
        if (np >= 2 && nel >= 1) {
            std::vector<double> Pm((size_t)np * np, 2.0 / 3.0), H, w, C, occ;
            double eold = 0.0;
            for (int it = 0; it < 5; ++it) {
                H.assign((size_t)np * np, 0.0);
                for (int a = 0; a < np; ++a) {
                    int k = at[a], z = m.Z[k];
                    int ne = std::min(pi_electrons(z, t.hyb[k]), 2);
                    H[(size_t)a * np + a] = (z < 18 ? hdiag[z] : 0.0) + t.qa[k] * hp3 - (double)(ne - 1) * pilpf;
                }
                for (int b : bl) {
                    int ii = t.bonds[b].first, jj = t.bonds[b].second;
                    int ia = loc[ii], ja = loc[jj];
                    int zi = m.Z[ii], zj = m.Z[jj];
                    double beta = std::sqrt((zi < 18 ? hoff[zi] : 0.0) * (zj < 18 ? hoff[zj] : 0.0));
                    beta -= 1.0e-9 * m.dist(ii, jj);
                    double dmp = hiter;
                    if (t.hyb[ii] == 1) dmp *= htriple;
                    if (t.hyb[jj] == 1) dmp *= htriple;
                    double h = -beta * (1.0 - dmp * (2.0 / 3.0 - Pm[(size_t)ja * np + ia]));
                    H[(size_t)ja * np + ia] = h;
                    H[(size_t)ia * np + ja] = h;
                }
                sym_eig(np, H, w, C);
                for (double& x : w) x *= ev;
                double enew = pi_occupations(w, nel, occ);
                for (int r = 0; r < np; ++r)
                    for (int c = 0; c < np; ++c) {
                        double acc = 0.0;
                        for (int o = 0; o < np; ++o) acc += C[(size_t)r * np + o] * C[(size_t)c * np + o] * occ[o];
                        Pm[(size_t)r * np + c] = acc;
                    }
                if (std::abs(enew - eold) < 1e-4) break;
                eold = enew;
            }
            for (int b : bl) pibo[b] = Pm[(size_t)loc[t.bonds[b].second] * np + loc[t.bonds[b].first]];
        }
        for (int k : at) loc[k] = -1;
    }
    return pibo;
}

// Charge localised on each pi system (gfnff_method.cpp:9311-9373): the change
// of the H-condensed topology charges of its atoms when its fragment is made
// neutral, times 1.1, rounded.  Zero for pi systems in neutral fragments.
bool pi_system_charges(const Mol& m, const Topo& t, int threads, std::vector<int>& ipis, std::string& msg) {
    int maxid = 0;
    for (int p : t.pi) maxid = std::max(maxid, p);
    ipis.assign(maxid + 1, 0);
    auto qheavy = [&](const std::vector<double>& q) {
        std::vector<double> qh = q;
        for (int i = 0; i < m.n; ++i) {
            if (m.Z[i] != 1 || t.adj[i].empty()) continue;
            double share = q[i] / (double)t.adj[i].size();
            for (int nb : t.adj[i]) qh[nb] += share;
            qh[i] = 0.0;
        }
        return qh;
    };
    std::vector<double> qah;
    for (int s = 1; s <= maxid; ++s) {
        int first = -1;
        for (int k = 0; k < m.n && first < 0; ++k)
            if (t.pi[k] == s) first = k;
        if (first < 0) continue;
        int f = t.frag[first] - 1;
        if (t.qfrag[f] == 0.0) continue;  // neutralising changes nothing
        if (qah.empty()) qah = qheavy(t.qa);
        Topo tn = t;
        tn.qfrag[f] = 0.0;
        if (!phase1_charges(m, tn, threads, msg)) return false;
        std::vector<double> qnh = qheavy(tn.qa);
        double dum = 0.0;
        for (int k = 0; k < m.n; ++k)
            if (t.pi[k] == s) dum += qah[k] - qnh[k];
        ipis[s] = (int)std::lround(dum * 1.1);
    }
    return true;
}

// classifyBondType, gfnff_method.cpp:4021-4098, plus the sp re-classification
// of calculateTopologyInfo :9430-9440 (pibo = 0 without a Hueckel pi system).
void bond_types(const Mol& m, Topo& t, const std::vector<double>& pibo_of_bond) {
    t.btype.assign(t.bonds.size(), 1);
    auto halogen = [](int z) { return z == 9 || z == 17 || z == 35 || z == 53 || z == 85; };
    for (size_t b = 0; b < t.bonds.size(); ++b) {
        int i = t.bonds[b].first, j = t.bonds[b].second;
        int hi = t.hyb[i], hj = t.hyb[j];
        if (hi == 3 && t.pi[i] > 0) hi = 2;
        if (hj == 3 && t.pi[j] > 0) hj = 2;
        int zi = m.Z[i], zj = m.Z[j];
        int bt = 1;
        if (hi == 2 && hj == 2) bt = 2;
        if ((hi == 3 && hj == 2 && zi == 7) || (hj == 3 && hi == 2 && zj == 7)) bt = 2;
        if (hi == 1 || hj == 1) bt = 3;
        if ((halogen(zi) || zi == 1) && hi == 1) bt = 3;
        if ((halogen(zj) || zj == 1) && hj == 1) bt = 3;
        if (hi == 5 || hj == 5) bt = 4;
        if (t.is_metal[i] || t.is_metal[j]) bt = 5;
        if (t.is_metal[i] && t.is_metal[j]) bt = 7;
        if (bt == 3) {
            int bb = 3;
            if (hi == 0 || hj == 0 || hi == 3 || hj == 3) bb = 1;
            else if (hi == 2 || hj == 2) bb = 2;
            if (bb != 3 && pibo_of_bond[b] > 0.1) bt = 2;
        }
        t.btype[b] = bt;
    }
}

// Sparse topological distances up to 3 bonds (calculateTopologyDistances,
// gfnff_method.cpp:8693-8770, truncated at the depth the terms read).
void topo_neighbourhood(int n, Topo& t) {
    t.tnb.assign(n, {});
    std::vector<int> depth(n, -1);
    std::vector<int> touched;
    for (int s = 0; s < n; ++s) {
        std::queue<int> q;
        q.push(s);
        depth[s] = 0;
        touched.assign(1, s);
        while (!q.empty()) {
            int c = q.front();
            q.pop();
            if (depth[c] >= 3) continue;
            for (int nb : t.adj[c])
                if (depth[nb] < 0) {
                    depth[nb] = depth[c] + 1;
                    touched.push_back(nb);
                    q.push(nb);
                }
        }
        for (int a : touched)
            if (a != s) t.tnb[s].emplace_back(a, depth[a]);
        std::sort(t.tnb[s].begin(), t.tnb[s].end());
        for (int a : touched) depth[a] = -1;
    }
}

// ---------------------------------------------------------------------------
// Bond parameters: getGFNFFBondParameters, gfnff_method.cpp:4184-5081
// ---------------------------------------------------------------------------
struct BondPar {
    double r0b_i, r0b_j, cnfak_i, cnfak_j, rabshift, ff, alpha, fc;
};

int count_nb20(const Mol& m, int a) {
    int c = 0;
    for (int j = 0; j < m.n; ++j)
        if (j != a && m.dist2(a, j) < 400.0) ++c;
    return c;
}

BondPar bond_params(const Mol& m, const Topo& t, int a1, int a2, double pibo) {
    int z1 = m.Z[a1], z2 = m.Z[a2];
    double bp1 = (z1 >= 1 && z1 <= 86) ? bond[z1 - 1] : 0.15;
    double bp2 = (z2 >= 1 && z2 <= 86) ? bond[z2 - 1] : 0.15;
    double cn1 = t.cn[a1], cn2 = t.cn[a2];
    double r01 = (z1 >= 1 && z1 <= 86) ? r0[z1 - 1] : 2.0;
    double r02 = (z2 >= 1 && z2 <= 86) ? r0[z2 - 1] : 2.0;
    double ck1 = (z1 >= 1 && z1 <= 86) ? cnfak[z1 - 1] : 0.0;
    double ck2 = (z2 >= 1 && z2 <= 86) ? cnfak[z2 - 1] : 0.0;
    double ra = r01 + ck1 * cn1, rb = r02 + ck2 * cn2;
    (void)ra;
    (void)rb;
    double en1 = (z1 >= 1 && z1 <= 86) ? en_rab[z1 - 1] : 2.2;
    double en2 = (z2 >= 1 && z2 <= 86) ? en_rab[z2 - 1] : 2.2;
    int row1 = periodic_row(z1), row2 = periodic_row(z2);
    double k1 = 0.005 * (p_enpoly[2 * (row1 - 1)] + p_enpoly[2 * (row2 - 1)]);
    double k2 = 0.005 * (p_enpoly[2 * (row1 - 1) + 1] + p_enpoly[2 * (row2 - 1) + 1]);
    double end = std::abs(en1 - en2);
    double ff = 1.0 - k1 * end - k2 * end * end;

    double shift = 0.0;
    int h1 = t.hyb[a1], h2 = t.hyb[a2];
    int mi = (h1 < 1 || h1 > 3) ? 3 : h1, mj = (h2 < 1 || h2 > 3) ? 3 : h2;
    int hybi = std::max(mi, mj), hybj = std::min(mi, mj);
    int bbtyp = 1;
    if (hybi == 5 || hybj == 5) bbtyp = 4;
    else if (hybi == 1 || hybj == 1) bbtyp = 3;
    else if (hybi == 2 && hybj == 2) bbtyp = 2;
    else if (hybi == 3 && hybj == 2 && (z1 == 7 || z2 == 7)) bbtyp = 2;
    if (bbtyp == 4) shift = 0.030;
    else if (z1 == 1 || z2 == 1) shift = -0.050;
    if (z1 == 9 && z2 == 9) shift += 0.22;
    if ((h1 == 3 && h2 == 0) || (h1 == 0 && h2 == 3)) shift -= 0.022;
    if ((h1 == 1 && h2 == 0) || (h1 == 0 && h2 == 1)) shift += 0.14;
    if (z1 > 10 && z2 > 10) {
        shift += -0.110;
        if (z1 > 18) shift += -0.110;
        if (z2 > 18) shift += -0.110;
        if (z1 > 36) shift += -0.060;
        if (z2 > 36) shift += -0.060;
    }
    double rabshift = -0.110 + shift;

    static const double bsm[4][4] = {{1.0000, 1.3234, 1.0792, 1.0000},
                                     {1.3234, 1.9800, 1.4842, 1.3234},
                                     {1.0792, 1.4842, 1.2400, 1.0792},
                                     {1.0000, 1.3234, 1.0792, 1.0000}};
    hybi = std::max(mi, mj);
    hybj = std::min(mi, mj);
    double bstr;
    if (z1 == 1 && z2 == 1) bstr = bstren[1];
    else if (hybi == 5 || hybj == 5) bstr = bstren[4];
    else bstr = bsm[hybi][hybj];
    if (hybi == 3 && hybj == 2 && (z1 == 7 || z2 == 7)) bstr = bstren[2] * 1.04;
    if (bbtyp == 3 && ((z1 == 6 && z2 == 8) || (z1 == 8 && z2 == 6))) bstr = bstren[3] * 0.90;
    if (bbtyp == 3) {
        if (h1 == 0 || h2 == 0) bbtyp = 1;
        else if (h1 == 3 || h2 == 3) bbtyp = 1;
        else if (h1 == 2 || h2 == 2) bbtyp = 2;
    }
    {
        int g1 = (z1 >= 1 && z1 <= 86) ? periodic_group[z1 - 1] : 0;
        int g2 = (z2 >= 1 && z2 <= 86) ? periodic_group[z2 - 1] : 0;
        bool bridge = false;
        if ((g1 == 7 || z1 == 1) && h1 == 1) {
            bbtyp = 3;
            bridge = true;
        }
        if ((g2 == 7 || z2 == 1) && h2 == 1) {
            bbtyp = 3;
            bridge = true;
        }
        if (bridge) {
            if (g1 == 7) bstr = bstren[1] * 0.50;
            if (g2 == 7) bstr = bstren[1] * 0.50;
            if (z1 == 1 || z1 == 9) bstr = bstren[1] * 0.30;
            if (z2 == 1 || z2 == 9) bstr = bstren[1] * 0.30;
        }
    }
    double qa1 = t.qa[a1], qa2 = t.qa[a2];
    double qafac = qa1 * qa2 * 70.0;
    double ex = std::exp(-15.0 * qafac);
    double fqq = 1.0 + 0.047 * ex / (1.0 + ex);
    double ringf = 1.0;
    int rs = same_ring(t, a1, a2);
    if (rs > 0) ringf = 1.0 + 0.020 * std::pow(6.0 - rs, 2);
    double fxh = 1.0;
    if (z1 == 1 || z2 == 1) {
        int hz = (z1 == 1) ? z2 : z1;
        int hidx = (z1 == 1) ? a2 : a1;
        bool r3 = hz == 6 && t.ring_size[hidx] == 3;
        bool ald = false;
        if (hz == 6 && t.pi[hidx] > 0) {
            int c = 0;
            for (int nb : t.adj[hidx])
                if (m.Z[nb] == 8 && t.pi[nb] > 0) ++c;
            ald = (c == 1);
        }
        if (hz == 6 && r3) fxh = 1.05;
        else if (hz == 6 && ald) fxh = 0.95;
        else if (hz == 5) fxh = 1.10;
        else if (hz == 7) fxh = 1.06;
        else if (hz == 8) fxh = 0.93;
    }
    double fcn = 1.0;
    int nb201 = 0, nb202 = 0;
    if (z1 > 10 && z2 > 10) {
        nb201 = count_nb20(m, a1);
        nb202 = count_nb20(m, a2);
        fcn /= (1.0 + 0.007 * nb201 * nb201);
        fcn /= (1.0 + 0.007 * nb202 * nb202);
    }
    double fheavy = 1.0, fpi = 1.0, metal_shift = 0.0, pi_shift = 0.0;
    if (pibo > 0.0) {
        pi_shift = 0.340 * (0.370 - pibo);
        fpi = 1.0 - 1.00 * (0.315 - pibo);
    }
    int im1 = (z1 >= 1 && z1 <= 86) ? metal_type[z1 - 1] : 0;
    int im2 = (z2 >= 1 && z2 <= 86) ? metal_type[z2 - 1] : 0;
    int g1 = (z1 >= 1 && z1 <= 86) ? periodic_group[z1 - 1] : 0;
    int g2 = (z2 >= 1 && z2 <= 86) ? periodic_group[z2 - 1] : 0;
    int mt1 = 0, mt2 = 0;
    if (z1 > 1 && g1 == 1) mt1 = 1;
    else if (z1 > 1 && g1 == 2) mt1 = 2;
    else if (g1 > 2 && im1 == 1) mt1 = 3;
    else if (im1 == 2) mt1 = 4;
    if (z2 > 1 && g2 == 1) mt2 = 1;
    else if (z2 > 1 && g2 == 2) mt2 = 2;
    else if (g2 > 2 && im2 == 1) mt2 = 3;
    else if (im2 == 2) mt2 = 4;
    if (mt1 == 4 || mt2 == 4) {
        if (z1 > 10 || z2 > 10) {
            if (nb201 == 0) nb201 = count_nb20(m, a1);
            if (nb202 == 0) nb202 = count_nb20(m, a2);
        }
        if (im1 == 2 && z2 > 10) fheavy = 0.65;
        if (im2 == 2 && z1 > 10) fheavy = 0.65;
        if (im1 == 2 && z2 == 15) fheavy = 1.60;
        if (im2 == 2 && z1 == 15) fheavy = 1.60;
        if (im1 == 2 && g2 == 6) fheavy = 0.85;
        if (im2 == 2 && g1 == 6) fheavy = 0.85;
        if (im1 == 2 && g2 == 7) fheavy = 1.30;
        if (im2 == 2 && g1 == 7) fheavy = 1.30;
        int nb1 = count_nb20(m, a1), nb2 = count_nb20(m, a2);
        if (im2 == 2 && h1 == 1) {
            if (z1 == 6) {
                fpi = 1.5;
                metal_shift = -0.45;
            }
            if (z1 == 7 && nb1 != 1) {
                fpi = 0.4;
                metal_shift = 0.47;
            }
        }
        if (im1 == 2 && h2 == 1) {
            if (z2 == 6) {
                fpi = 1.5;
                metal_shift = -0.45;
            }
            if (z2 == 7 && nb2 != 1) {
                fpi = 0.4;
                metal_shift = 0.47;
            }
        }
        if (im1 == 2 && z2 == 1) fxh = (row1 <= 5) ? 0.80 : 1.00;
        if (im2 == 2 && z1 == 1) fxh = (row2 <= 5) ? 0.80 : 1.00;
    }
    if ((im1 == 1 && z2 == 1) || (im2 == 1 && z1 == 1)) fxh = 1.20;
    if (im1 == 2) metal_shift += 0.15;
    if (im2 == 2) metal_shift += 0.15;
    if (im1 == 1 && g1 <= 2) metal_shift += 0.2;
    if (im2 == 1 && g2 <= 2) metal_shift += 0.2;
    if (mt1 == 3) metal_shift += 0.05;
    if (mt2 == 3) metal_shift += 0.05;
    double total_rabshift = rabshift + pi_shift + metal_shift;
    double fc = -(bp1 * bp2 * bstr * fqq * ringf * fheavy * fpi * fxh * fcn);
    const double srb1 = 0.3731, srb2 = 0.3171, srb3 = 0.2538;
    double fsrb2;
    if (mt1 == 4 || mt2 == 4) fsrb2 = -srb2 * 0.22;
    else if (mt1 > 0 || mt2 > 0) fsrb2 = srb2 * 0.28;
    else fsrb2 = srb2;
    double ea1 = (z1 >= 1 && z1 <= 86) ? en[z1 - 1] : 2.2;
    double ea2 = (z2 >= 1 && z2 <= 86) ? en[z2 - 1] : 2.2;
    double ed = std::abs(ea1 - ea2);
    double asum = 1.0 + fsrb2 * ed * ed + srb3 * bstr;
    BondPar p;
    p.r0b_i = r01;
    p.r0b_j = r02;
    p.cnfak_i = ck1;
    p.cnfak_j = ck2;
    p.rabshift = total_rabshift;
    p.ff = ff;
    p.alpha = srb1 * asum;
    p.fc = fc;
    return p;
}

// ---------------------------------------------------------------------------
// Angle parameters: getGFNFFAngleParameters, gfnff_method.cpp:5083-5785
// ---------------------------------------------------------------------------
bool has_carbonyl_o(const Mol& m, const Topo& t, int c) {  // functional_groups.cpp hasCarbonylOxygen
    for (int nb : t.adj[c])
        if (m.Z[nb] == 8 && t.pi[nb] != 0 && t.adj[nb].size() == 1) return true;
    return false;
}
bool fg_amide_n(const Mol& m, const Topo& t, int a) {  // functional_groups.cpp:112-140
    if (m.Z[a] != 7 || t.hyb[a] != 3 || t.pi[a] == 0 || t.adj[a].empty()) return false;
    int pc = 0, co = 0;
    for (int nb : t.adj[a])
        if (m.Z[nb] == 6 && t.pi[nb] != 0) {
            ++pc;
            if (has_carbonyl_o(m, t, nb)) ++co;
        }
    return pc == 1 && co == 1;
}

bool angle_params(const Mol& m, const Topo& t, const std::map<std::pair<int, int>, double>& pibo, int ai,
                  int aj, int ak, double cur, double& fc_out, double& th0_out) {
    int zc = m.Z[aj], zi = m.Z[ai], zk = m.Z[ak];
    int hyb = t.hyb[aj];
    double ap = (zc >= 1 && zc <= 86) ? angle[zc - 1] : 0.1;
    double a2i = (zi >= 1 && zi <= 86) ? angl2[zi - 1] : 0.1;
    double a2k = (zk >= 1 && zk <= 86) ? angl2[zk - 1] : 0.1;
    double fijk = ap * a2i * a2k;
    double fqq = 1.0;
    {
        double qc = t.qa[aj], qi = t.qa[ai], qk = t.qa[ak];
        if (std::abs(qc) < 1.0 && std::abs(qi) < 1.0 && std::abs(qk) < 1.0) {
            double cp = qc * qi + qc * qk;
            bool hm = t.is_metal[ai] || t.is_metal[aj] || t.is_metal[ak];
            fqq = 1.0 - cp * -0.54 * (hm ? 2.5 : 1.0);
        }
    }
    double f2 = 1.0;
    int nh = 0, no = 0, nc = 0, nnn = 0, nsi = 0, nmet = 0, npi = 0;
    for (int a : {ai, ak}) {
        if (m.Z[a] == 1) ++nh;
        if (m.Z[a] == 6) ++nc;
        if (m.Z[a] == 7) ++nnn;
        if (m.Z[a] == 8) ++no;
        if (m.Z[a] == 14) ++nsi;
        if (t.is_metal[a]) ++nmet;
        if (t.pi[a] != 0) ++npi;
    }
    int nn_c = (int)t.adj[aj].size();
    double nn = std::max(1.0, (double)nn_c);
    double fn = std::max(0.05, 1.0 - 2.36 / (nn * nn));
    double feta = 1.0;
    if (zc >= 1 && zc <= 86 && metal_type[zc - 1] == 2) {
        if (t.hyb[ai] == 1 || t.hyb[ai] == 2) feta *= 0.3;
        if (t.hyb[ak] == 1 || t.hyb[ak] == 2) feta *= 0.3;
    }
    double r0d = 100.0;
    switch (hyb) {
        case 1: r0d = 180.0; break;
        case 2: r0d = 120.0; break;
        case 3: r0d = 109.5; break;
        case 5: r0d = 90.0; break;
        default: r0d = 100.0;
    }
    int gc = (zc >= 1 && zc <= 86) ? periodic_group[zc - 1] : 0;
    double curd = cur * 180.0 / M_PI;
    if (hyb == 3 && zc > 10) {
        if (nn_c <= 3) r0d = 89.0;
        if (nn_c >= 4) r0d = 100.0;
        if (nn_c == 4 && gc == 5) r0d = 109.5;
        if (nn_c == 4 && gc == 4 && zc > 49) r0d = 109.5;
        if (gc == 4) r0d = r0d - nh * 5.0;
        if (gc == 5) r0d = r0d - nh * 5.0;
        if (gc == 6) r0d = r0d - nh * 5.0;
    }
    bool co2 = false;
    if (hyb == 5) {
        f2 = 0.11;
        if (curd > 160.0) r0d = 180.0;
    } else if (zc == 5) {
        if (hyb == 3 || hyb == 2) r0d = 115.0;
    } else if (zc == 6) {
        if (hyb == 3) {
            if (nh == 2) r0d = 108.6;
            if (no == 1) r0d = 108.5;
            if (nn_c > 4 && curd > 160.0) r0d = 180.0;
        } else if (hyb == 2) {
            if (no == 2) r0d = 122.0;
            if (no == 1) f2 = 0.7;
        } else if (hyb == 1) {
            if (no == 2) {
                co2 = true;
                f2 = 2.0;
            }
        }
    } else if (zc == 7) {
        int rc = ring_bend(t, aj, ai, ak);
        if (nn_c == 2) {
            f2 = 1.4;
            r0d = 115.0;
            if (rc != 0) r0d = 105.0;
            if (no >= 1) r0d = 103.0;
            int nf = (m.Z[ai] == 9) + (m.Z[ak] == 9);
            if (nf >= 1) r0d = 102.0;
            if (hyb == 1) r0d = 180.0;
            bool tmn = false;
            if (t.is_metal[ai] && metal_type[m.Z[ai] - 1] == 2) tmn = true;
            if (t.is_metal[ak] && metal_type[m.Z[ak] - 1] == 2) tmn = true;
            if (tmn && hyb == 1 && nnn >= 1) r0d = 135.0;
        } else if (hyb == 3) {
            bool sp2n = t.hyb[ai] == 1 || t.hyb[ai] == 2 || t.hyb[ak] == 1 || t.hyb[ak] == 2;
            if (npi > 0 || sp2n) {
                if (fg_amide_n(m, t, aj)) {
                    r0d = 115.0;
                    f2 = 1.2;
                } else {
                    r0d = 113.0;
                    auto pb = [&](int a, int b) {
                        auto it = pibo.find({std::min(a, b), std::max(a, b)});
                        return it == pibo.end() ? 0.0 : it->second;
                    };
                    double s = pb(aj, ai) + pb(aj, ak);
                    f2 = 1.0 - s * 0.7;
                }
            } else {
                r0d = 104.0;
                f2 = 0.40;
                f2 += nh * 0.19;
                f2 += no * 0.25;
                f2 += nc * 0.01;
            }
        }
    }
    int rc = ring_bend(t, aj, ai, ak);
    if (rc == 3) r0d = 82.0;
    else if (rc == 4) r0d = 96.0;
    else if (rc == 5 && zc == 6) r0d = 109.0;
    if (rc == 0 && t.ring_size[aj] == 3) {
        int ri = t.ring_size[ai], rk = t.ring_size[ak];
        if ((ri == 3 && rk == 0) || (ri == 0 && rk == 3)) r0d += 4.0;
    }
    bool triple = !co2 && (hyb == 1 || t.hyb[ai] == 1 || t.hyb[ak] == 1);
    if (triple) {
        f2 = 0.60;
        if (nnn >= 1) f2 = 1.00;
        int atj = m.Z[ai], atk = m.Z[ak];
        int imi = (atj >= 1 && atj <= 86) ? metal_type[atj - 1] : 0;
        int imk = (atk >= 1 && atk <= 86) ? metal_type[atk - 1] : 0;
        if ((imi == 2 || imk == 2) && curd > 160.0) {
            if (zc == 6 && atj == 6) f2 = 3.0;
            if (zc == 6 && atk == 6) f2 = 3.0;
            if (zc == 6 && atj == 7) f2 = 3.0;
            if (zc == 6 && atk == 7) f2 = 3.0;
            int gj = (atj >= 1 && atj <= 86) ? periodic_group[atj - 1] : 0;
            int gk = (atk >= 1 && atk <= 86) ? periodic_group[atk - 1] : 0;
            if (zc == 6 && gj == 6) f2 = 14.0;
            if (zc == 6 && gk == 6) f2 = 14.0;
            if (zc == 7 && atj == 7) f2 = 10.0;
            if (zc == 7 && atj == 6) f2 = 10.0;
            if (zc == 7 && atk == 6) f2 = 10.0;
            if (zc == 7 && atj == 8) {
                r0d = 180.0;
                f2 = 12.0;
            }
            if (zc == 7 && atk == 8) {
                r0d = 180.0;
                f2 = 12.0;
            }
        }
    }
    if (gc == 6 && nn_c == 4 && no >= 1) r0d = 115.0;
    if (gc == 7 && hyb == 1) {
        r0d = 90.0;
        if (zc > 9 && curd > 160.0) r0d = 180.0;
        f2 = 0.6 / std::pow((double)zc, 0.15);
    }
    if (hyb == 3 && gc == 4 && zc > 32 && t.qa[aj] > 0.4) {
        if (curd > 140.0) r0d = 180.0;
        if (curd < 100.0) r0d = 90.0;
        f2 = 1.0;
    }
    int imc = (zc >= 1 && zc <= 86) ? metal_type[zc - 1] : 0;
    if (imc > 0) {
        if (hyb == 0) {
            r0d = 90.0;
            f2 = 1.35;
        }
        if (hyb == 1) r0d = 180.0;
        if (hyb == 2) r0d = 120.0;
        if (hyb == 3) r0d = 109.5;
        if (curd > 160.0) r0d = 180.0;
    }
    if (zc == 8 && nn_c == 2) {
        r0d = 104.5;
        if (nh == 2) {
            r0d = 100.0;
            f2 = 1.20;
        }
        r0d = r0d + 7.0 * nsi;
        r0d = r0d + 14.0 * nmet;
        if (npi == 2) r0d = 109.0;
        if (nmet > 0 && curd > 160.0) {
            r0d = 180.0;
            f2 = 0.3;
        }
    }
    double th0 = r0d * M_PI / 180.0;
    double fbsmall = 1.0 - 0.50 * std::exp(-0.64 * (th0 - M_PI) * (th0 - M_PI));
    double fc = (fijk < 0.001) ? 0.0 : fijk * fqq * f2 * fn * fbsmall * feta;
    fc_out = fc;
    th0_out = th0;
    return true;
}

// ---------------------------------------------------------------------------
// Torsion parameters: getGFNFFTorsionParameters, gfnff_torsions.cpp:510-1314
// ---------------------------------------------------------------------------
struct TorsPar {
    double V = 0.0;
    int n = 1;
    double phi0 = M_PI;
    double fij = 0.0, fkl = 0.0, fqq = 1.0;
};

bool fg_terminal_carbonyl_o(const Mol& m, const Topo& t, int c) {  // gfnff_torsions.cpp:339-362
    int no = 0;
    for (int j : t.adj[c])
        if (m.Z[j] == 8 && t.pi[j] != 0 && t.adj[j].size() == 1) ++no;
    return no == 1;
}
bool tors_alpha_co(const Mol& m, const Topo& t, int a, int b) {  // :364-385
    if (t.pi[a] && t.hyb[b] == 3 && m.Z[a] == 6 && m.Z[b] == 6 && fg_terminal_carbonyl_o(m, t, a)) return true;
    if (t.pi[b] && t.hyb[a] == 3 && m.Z[b] == 6 && m.Z[a] == 6 && fg_terminal_carbonyl_o(m, t, b)) return true;
    return false;
}
bool tors_amide(const Mol& m, const Topo& t, int a) {  // :425-448
    if (!t.pi[a] || t.hyb[a] != 3 || m.Z[a] != 7) return false;
    int nc = 0, ic = -1;
    for (int j : t.adj[a])
        if (m.Z[j] == 6 && t.pi[j] != 0) {
            ++nc;
            ic = j;
        }
    return nc == 1 && fg_terminal_carbonyl_o(m, t, ic);
}

TorsPar tors_params(const Mol& m, const Topo& t, const std::map<std::pair<int, int>, double>& pibo, int ia,
                    int ja, int ka, int la, bool inring, int ringsz, int btyp) {
    TorsPar p;
    int zi = m.Z[ia], zj = m.Z[ja], zk = m.Z[ka], zl = m.Z[la];
    int hj = t.hyb[ja], hk = t.hyb[ka];
    double qaj = t.qa[ja], qak = t.qa[ka];
    double cni = (double)t.adj[ia].size(), cnl = (double)t.adj[la].size();
    bool jpi = t.pi[ja] > 0, kpi = t.pi[ka] > 0;
    if (jpi && kpi) {
        p.n = 2;
        p.phi0 = M_PI;
    } else if (hj == 3 && hk == 3) {
        p.n = 3;
    } else if (hj == 2 && hk == 2) {
        p.n = 2;
    } else if ((hj == 2 && hk == 3) || (hj == 3 && hk == 2)) {
        p.n = 3;
    } else {
        p.n = 1;
    }
    const double torsf_pi = 1.18, fcthr = 1.0e-3;
    if (zi < 1 || zi > 86 || zj < 1 || zj > 86 || zk < 1 || zk > 86 || zl < 1 || zl > 86) return p;
    double fij = tors[zj - 1] * tors[zk - 1];
    if (fij < fcthr || tors[zj - 1] < 0.0 || tors[zk - 1] < 0.0) return p;
    double fkl_base = tors2[zi - 1] * tors2[zl - 1];
    double fkl = fkl_base;
    int hi = t.hyb[ia], hl = t.hyb[la];
    auto pb = [&](int a, int b) {
        auto it = pibo.find({std::min(a, b), std::max(a, b)});
        return it == pibo.end() ? 0.0 : it->second;
    };
    auto n_not_pi = [&](int a) {
        if (m.Z[a] != 7) return false;
        for (auto& kv : pibo)
            if ((kv.first.first == a || kv.first.second == a) && kv.second > 0.1) return false;
        for (int nb : t.adj[a])
            if (t.hyb[nb] == 1 || t.hyb[nb] == 2) return false;
        return true;
    };
    if (n_not_pi(ia)) fkl = fkl_base * 0.5;
    if (n_not_pi(la)) fkl = fkl_base * 0.5;
    if (fkl < fcthr || tors[zi - 1] < 0.0 || tors[zl - 1] < 0.0) return p;
    double cnp = cni * cnl;
    if (cnp > 0.01) fkl *= std::pow(cnp, -0.14);
    if (hi == 5 || hl == 5) fkl *= 1.5;

    double f1 = 1.00;
    int ehj = hj, ehk = hk;
    if (t.pi[ja] > 0 && ehj == 3) ehj = 2;
    if (t.pi[ka] > 0 && ehk == 3) ehk = 2;
    if (ehj == 3 && ehk == 3) {
        int gj = (zj == 7 || zj == 15) ? 5 : (zj == 8 || zj == 16) ? 6 : 0;
        int gk = (zk == 7 || zk == 15) ? 5 : (zk == 8 || zk == 16) ? 6 : 0;
        if (gj == 6 && gk == 6) {
            f1 = 5.0;
            p.n = 2;
            p.phi0 = 90.0 * M_PI / 180.0;
            if (zj >= 16 && zk >= 16) f1 = 25.0;
        } else if (gj == 5 && gk == 5) {
            f1 = 3.0;
            p.n = 3;
            p.phi0 = 60.0 * M_PI / 180.0;
        } else if ((gj == 5 && gk == 6) || (gj == 6 && gk == 5)) {
            f1 = 1.0;
            p.n = 2;
            p.phi0 = 90.0 * M_PI / 180.0;
            if (zj >= 15 && zk >= 15) f1 = 20.0;
        }
    } else if (!inring && ((ehj == 2 && ehk == 3) || (ehj == 3 && ehk == 2))) {
        f1 = 0.5;
        if (zj == 7 || zk == 7) f1 = 0.2;
    } else if (ehj == 2 && ehk == 2) {
        f1 = 1.00;
    }
    double f2 = 0.0, central_pibo = 0.0;
    {
        double pbv = pb(ja, ka);
        central_pibo = pbv;
        if (pbv > 0.0) {
            double diff = 1.24 - pbv;
            f2 = pbv * std::exp(-2.5 * std::pow(diff, 14));
            if (zi > 10 && !(t.pi[ia] > 0)) f2 *= 1.3;
            if (zl > 10 && !(t.pi[la] > 0)) f2 *= 1.3;
            f1 *= 0.55;
        }
    }
    int nhi = 1, nhj = 1;
    for (int nb : t.adj[ja])
        if (m.Z[nb] == 1) ++nhi;
    for (int nb : t.adj[ka])
        if (m.Z[nb] == 1) ++nhj;
    fij *= std::pow((double)(nhi * nhj), 0.07);
    if (tors_alpha_co(m, t, ja, ka)) fij *= 1.3;
    if (tors_amide(m, t, ja) && zk == 6 && hk == 3) fij *= 1.3;
    if (tors_amide(m, t, ka) && zj == 6 && hj == 3) fij *= 1.3;
    if (t.is_metal[ja] && t.adj[ja].size() > 4) return p;
    if (t.is_metal[ka] && t.adj[ka].size() > 4) return p;
    if (inring) {
        p.n = 1;
        p.phi0 = 0.0;
        if (btyp == 2) p.n = 2;
        int rings4 = 0, ringl = 0;
        if (ringsz == 3) rings4 = 3;
        else if (ringsz > 3) rings4 = ring_all(t, ia, ja, ka, la, false);
        bool notpicon = !(t.pi[ia] > 0 || t.pi[la] > 0);
        if (rings4 > 0 && btyp == 1) {
            if (rings4 >= 4) ringl = ring_all(t, ia, ja, ka, la, true);
            if (rings4 == 3 && notpicon) {
                p.n = 1;
                p.phi0 = 0.0;
                f1 = 0.3;
            }
            if (rings4 == 4 && ringl == rings4 && notpicon) {
                p.n = 6;
                p.phi0 = 30.0 * M_PI / 180.0;
                f1 = 1.0;
            }
            if (rings4 == 5 && ringl == rings4 && notpicon) {
                p.n = 6;
                p.phi0 = 30.0 * M_PI / 180.0;
                f1 = 1.5;
            }
            if (rings4 == 6 && ringl == rings4 && notpicon) {
                p.n = 3;
                p.phi0 = 60.0 * M_PI / 180.0;
                f1 = 5.7;
            }
        }
        if (rings4 == 0 && btyp == 1 && (int)cni == 1 && (int)cnl == 1) {
            p.n = 6;
            p.phi0 = 30.0 * M_PI / 180.0;
            f1 = 0.30;
        }
        if (btyp == 2 && ringsz == 5 && zj * zk == 42) {
            if (tors_amide(m, t, ja) || tors_amide(m, t, ka)) {
                f1 = 5.0;
                if (central_pibo > 0.0) f1 *= 0.55;
            }
        }
    }
    if (!inring && jpi && !kpi && hk == 3) {
        f1 = 0.5;
        if (zj == 7) f1 = 0.2;
        p.phi0 = M_PI;
        p.n = 3;
    } else if (!inring && kpi && !jpi && hj == 3) {
        f1 = 0.5;
        if (zk == 7) f1 = 0.2;
        p.phi0 = M_PI;
        p.n = 3;
    }
    double fqq = 1.0 + std::abs(qaj * qak) * 12.0;
    double fctot = (f1 + 10.0 * torsf_pi * f2) * fqq * fij * fkl;
    if (fctot < fcthr) return p;
    p.V = fctot;
    p.fij = fij;
    p.fkl = fkl;
    p.fqq = fqq;
    return p;
}

double dihedral(const Mol& m, int i, int j, int k, int l) {  // gfnff_geometry.h:60-110
    double v1[3], v2[3], v3[3];
    for (int c = 0; c < 3; ++c) {
        v1[c] = m.x[3 * j + c] - m.x[3 * i + c];
        v2[c] = m.x[3 * k + c] - m.x[3 * j + c];
        v3[c] = m.x[3 * l + c] - m.x[3 * k + c];
    }
    auto cross = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double n1[3], n2[3];
    cross(v1, v2, n1);
    cross(v2, v3, n2);
    double nn1 = std::sqrt(sq(n1[0]) + sq(n1[1]) + sq(n1[2])), nn2 = std::sqrt(sq(n2[0]) + sq(n2[1]) + sq(n2[2]));
    double nv2 = std::sqrt(sq(v2[0]) + sq(v2[1]) + sq(v2[2]));
    if (nn1 < 1e-10 || nn2 < 1e-10 || nv2 < 1e-10) return 0.0;
    double c = (n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) / (nn1 * nn2);
    double s = nv2 * (n1[0] * v3[0] + n1[1] * v3[1] + n1[2] * v3[2]) / nn1 / nn2;
    s = std::max(-1.0, std::min(1.0, s));
    return std::atan2(s, c);
}

double bond_angle(const Mol& m, int a, int b, int c) {
    double u[3], v[3];
    for (int d = 0; d < 3; ++d) {
        u[d] = m.x[3 * a + d] - m.x[3 * b + d];
        v[d] = m.x[3 * c + d] - m.x[3 * b + d];
    }
    double nu = std::sqrt(sq(u[0]) + sq(u[1]) + sq(u[2])), nv = std::sqrt(sq(v[0]) + sq(v[1]) + sq(v[2]));
    double d = (u[0] * v[0] + u[1] * v[1] + u[2] * v[2]) / (nu * nv);
    return std::acos(std::max(-1.0, std::min(1.0, d)));
}

double zeta_scale(int Z, double q) {  // gfnff_par.h:941-960
    if (Z < 1 || Z > 86) return 1.0;
    double zf = zeta_zeff[Z - 1], c = zeta_c[Z - 1];
    double qm = zf + q;
    if (qm < 0.0) return std::exp(3.0);
    return std::exp(3.0 * (1.0 - std::exp(c * (1.0 - zf / qm))));
}

// computeC6Reference, d4param_generator.cpp:975-1060 (trapezoid over 23
// frequencies; effective alphas are tabulated with ascale/hcount folded in).
double c6_reference(int ei, int ri, int ej, int rj) {
    const double THREE_OVER_PI = 3.0 / 3.14159265358979323846;
    const double* ai = &d4_alpha_eff[((size_t)ei * kMaxRef + ri) * kNFreq];
    const double* aj = &d4_alpha_eff[((size_t)ej * kMaxRef + rj) * kNFreq];
    double c6 = 0.0;
    for (int w = 0; w < kNFreq - 1; ++w) {
        double p0 = ai[w] * aj[w], p1 = ai[w + 1] * aj[w + 1];
        double dw = d4_freq[w + 1] - d4_freq[w];
        c6 += 0.5 * (p0 + p1) * dw;
    }
    return c6 * THREE_OVER_PI;
}

}  // namespace

// ---------------------------------------------------------------------------
// Dense EEQ solve for set-up (Cholesky of A + Schur complement over fragment
// constraints; LU with partial pivoting on the bordered matrix as fallback).
// Mirrors EEQSolver::solveWithSchurCholesky, eeq_solver.cpp:2087-2240.
// ---------------------------------------------------------------------------
namespace {
bool cholesky_inplace(int n, double* A, int threads) {
    const int NB = 64;
    for (int k0 = 0; k0 < n; k0 += NB) {
        int kb = std::min(NB, n - k0);
        for (int k = k0; k < k0 + kb; ++k) {
            double d = A[(size_t)k * n + k];
            for (int p = k0; p < k; ++p) d -= A[(size_t)k * n + p] * A[(size_t)k * n + p];
            if (!(d > 0.0)) return false;
            d = std::sqrt(d);
            A[(size_t)k * n + k] = d;
            for (int i = k + 1; i < k0 + kb; ++i) {
                double s = A[(size_t)i * n + k];
                for (int p = k0; p < k; ++p) s -= A[(size_t)i * n + p] * A[(size_t)k * n + p];
                A[(size_t)i * n + k] = s / d;
            }
        }
        int rest0 = k0 + kb;
        if (rest0 >= n) break;
        auto panel = [&](int t0, int T) {
            for (int i = rest0 + t0; i < n; i += T) {
                double* Ai = &A[(size_t)i * n];
                for (int k = k0; k < k0 + kb; ++k) {
                    double s = Ai[k];
                    const double* Ak = &A[(size_t)k * n];
                    for (int p = k0; p < k; ++p) s -= Ai[p] * Ak[p];
                    Ai[k] = s / Ak[k];
                }
            }
        };
        auto update = [&](int t0, int T) {
            for (int i = rest0 + t0; i < n; i += T) {
                double* Ai = &A[(size_t)i * n];
                for (int j = rest0; j <= i; ++j) {
                    const double* Aj = &A[(size_t)j * n];
                    double s = 0.0;
                    for (int p = k0; p < k0 + kb; ++p) s += Ai[p] * Aj[p];
                    Ai[j] -= s;
                }
            }
        };
        int T = std::max(1, threads);
        for (auto fn : {std::function<void(int, int)>(panel), std::function<void(int, int)>(update)}) {
            std::vector<std::thread> th;
            for (int q = 1; q < T; ++q) th.emplace_back(fn, q, T);
            fn(0, T);
            for (auto& x : th) x.join();
        }
    }
    return true;
}

void chol_solve(int n, const double* L, double* b) {  // L row-major lower
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        const double* Li = &L[(size_t)i * n];
        for (int p = 0; p < i; ++p) s -= Li[p] * b[p];
        b[i] = s / Li[i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int p = i + 1; p < n; ++p) s -= L[(size_t)p * n + i] * b[p];
        b[i] = s / L[(size_t)i * n + i];
    }
}

bool lu_solve(int m, std::vector<double>& M, std::vector<double>& b) {
    std::vector<int> piv(m);
    for (int k = 0; k < m; ++k) {
        int p = k;
        double best = std::abs(M[(size_t)k * m + k]);
        for (int i = k + 1; i < m; ++i)
            if (std::abs(M[(size_t)i * m + k]) > best) {
                best = std::abs(M[(size_t)i * m + k]);
                p = i;
            }
        if (best == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < m; ++j) std::swap(M[(size_t)k * m + j], M[(size_t)p * m + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < m; ++i) {
            double f = M[(size_t)i * m + k] / M[(size_t)k * m + k];
            M[(size_t)i * m + k] = f;
            for (int j = k + 1; j < m; ++j) M[(size_t)i * m + j] -= f * M[(size_t)k * m + j];
            b[i] -= f * b[k];
        }
    }
    for (int i = m - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < m; ++j) s -= M[(size_t)i * m + j] * b[j];
        b[i] = s / M[(size_t)i * m + i];
    }
    return true;
}
}  // namespace

bool eeq_solve_host(int n, int nfrag, std::vector<double>& A, const std::vector<double>& x,
                    const std::vector<int32_t>& frag, const std::vector<double>& qfrag, std::vector<double>& q,
                    int threads) {
    std::vector<double> Acopy;
    bool small = (size_t)(n + nfrag) * (n + nfrag) < 4000000;
    if (small) Acopy = A;
    if (cholesky_inplace(n, A.data(), threads)) {
        std::vector<double> z1 = x;
        chol_solve(n, A.data(), z1.data());
        // Z2 = A^{-1} C^T, one column per fragment
        std::vector<double> Z2((size_t)nfrag * n, 0.0);  // fragment-major
        {
            auto work = [&](int t0, int T) {
                for (int f = t0; f < nfrag; f += T) {
                    double* c = &Z2[(size_t)f * n];
                    for (int i = 0; i < n; ++i) c[i] = (frag[i] == f) ? 1.0 : 0.0;
                    chol_solve(n, A.data(), c);
                }
            };
            int T = std::max(1, std::min(threads, nfrag));
            std::vector<std::thread> th;
            for (int k = 1; k < T; ++k) th.emplace_back(work, k, T);
            work(0, T);
            for (auto& t : th) t.join();
        }
        std::vector<double> S((size_t)nfrag * nfrag, 0.0), rhs(nfrag, 0.0);
        for (int f = 0; f < nfrag; ++f) {
            for (int g = 0; g < nfrag; ++g) {
                double s = 0.0;
                const double* c = &Z2[(size_t)g * n];
                for (int i = 0; i < n; ++i)
                    if (frag[i] == f) s += c[i];
                S[(size_t)f * nfrag + g] = s;
            }
            double s = 0.0;
            for (int i = 0; i < n; ++i)
                if (frag[i] == f) s += z1[i];
            rhs[f] = s - qfrag[f];
        }
        if (!lu_solve(nfrag, S, rhs)) return false;
        q.assign(n, 0.0);
        for (int i = 0; i < n; ++i) {
            double s = z1[i];
            for (int f = 0; f < nfrag; ++f) s -= Z2[(size_t)f * n + i] * rhs[f];
            q[i] = s;
        }
        return true;
    }
    if (!small) return false;
    int m = n + nfrag;
    std::vector<double> M((size_t)m * m, 0.0), b(m, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) M[(size_t)i * m + j] = Acopy[(size_t)i * n + j];
        M[(size_t)i * m + n + frag[i]] = 1.0;
        M[(size_t)(n + frag[i]) * m + i] = 1.0;
        b[i] = x[i];
    }
    for (int f = 0; f < nfrag; ++f) b[n + f] = qfrag[f];
    if (!lu_solve(m, M, b)) return false;
    q.assign(b.begin(), b.begin() + n);
    return true;
}

// ---------------------------------------------------------------------------
// Parameter-set assembly: generateGFNFFParameterSet, gfnff_method.cpp:3557-3747
// ---------------------------------------------------------------------------
// D4 element tables for the elements in P.present_Z: reference counts, the
// reference coordination numbers and the C6 reference grid (element-pair data
// of the D4 model; d4param_generator.cpp:1079-1213).
void set_element_tables(ParamSet& P) {
    P.nelem = (int)P.present_Z.size();
    const int ne = P.nelem;
    P.nref.assign(ne, 0);
    P.refcn.assign((size_t)ne * 7, 0.0);
    P.c6ref.assign((size_t)ne * ne * 49, 0.0);
    for (int e = 0; e < ne; ++e) {
        const int z = P.present_Z[e];
        P.nref[e] = std::min(d4_refn[z - 1], 7);
        for (int r = 0; r < 7; ++r) P.refcn[e * 7 + r] = d4_refcn[(z - 1) * 7 + r];
    }
    for (int a = 0; a < ne; ++a)
        for (int b = 0; b < ne; ++b)
            for (int ra = 0; ra < P.nref[a]; ++ra)
                for (int rb = 0; rb < P.nref[b]; ++rb) {
                    const int za = P.present_Z[a] - 1, zb = P.present_Z[b] - 1;
                    // the reference stores the lower triangle and mirrors it
                    const double c6 = (za > zb || (za == zb && ra >= rb)) ? c6_reference(za, ra, zb, rb)
                                                                          : c6_reference(zb, rb, za, ra);
                    P.c6ref[(((size_t)a * ne + b) * 7 + ra) * 7 + rb] = c6;
                }
}

int ParamSet::mol_of_atom(int i) const {
    if (nmol <= 1) return 0;
    return (int)(std::upper_bound(mol_off.begin(), mol_off.end(), i) - mol_off.begin()) - 1;
}

namespace {
template <class T>
void append(std::vector<T>& dst, const std::vector<T>& src) {
    dst.insert(dst.end(), src.begin(), src.end());
}
void append_shift(std::vector<int32_t>& dst, const std::vector<int32_t>& src, int shift, bool keep_negative = false) {
    for (int32_t v : src) dst.push_back((keep_negative && v < 0) ? v : v + shift);
}
void append_csr(std::vector<int32_t>& off, const std::vector<int32_t>& src_off, int base) {
    for (size_t k = 1; k < src_off.size(); ++k) off.push_back(src_off[k] + base);
}
}  // namespace

int merge_batch(const std::vector<ParamSet>& mols, ParamSet& B, std::string& msg) {
    B = ParamSet();
    const int nm = (int)mols.size();
    if (nm < 1) {
        msg = "empty batch";
        return kErrArgs;
    }
    B.nmol = nm;
    B.mol_off.assign(nm + 1, 0);
    B.mol_seg.assign(10 * (size_t)(nm + 1), 0);
    std::vector<int> seen(87, -1);
    for (const auto& P : mols)
        for (int z : P.present_Z)
            if (seen[z] < 0) {
                seen[z] = (int)B.present_Z.size();
                B.present_Z.push_back(z);
            }
    set_element_tables(B);
    // reserve every concatenated list once (avoids repeated regrowth of the
    // multi-GB lists of large batches)
    {
        auto reserve = [&](auto member) {
            size_t tot = 1;
            for (const auto& P : mols) tot += (P.*member).size();
            (B.*member).reserve(tot);
        };
        reserve(&ParamSet::Z);
        reserve(&ParamSet::alpeeq);
        reserve(&ParamSet::angle_ijk);
        reserve(&ParamSet::angle_par);
        reserve(&ParamSet::atm_ijk);
        reserve(&ParamSet::atm_scale);
        reserve(&ParamSet::batm_ijk);
        reserve(&ParamSet::batm_par);
        reserve(&ParamSet::bond_hbH);
        reserve(&ParamSet::bond_ij);
        reserve(&ParamSet::bond_par);
        reserve(&ParamSet::brep_ij);
        reserve(&ParamSet::brep_par);
        reserve(&ParamSet::chi_base);
        reserve(&ParamSet::cn_topo);
        reserve(&ParamSet::cnf);
        reserve(&ParamSet::frag);
        reserve(&ParamSet::gam_eff);
        reserve(&ParamSet::hb_ahb);
        reserve(&ParamSet::hb_case);
        reserve(&ParamSet::hb_nbB_idx);
        reserve(&ParamSet::hb_nbB_off);
        reserve(&ParamSet::hb_nbC_idx);
        reserve(&ParamSet::hb_nbC_off);
        reserve(&ParamSet::hb_par);
        reserve(&ParamSet::hb_parentC);
        reserve(&ParamSet::hb_rad);
        reserve(&ParamSet::hbcn_pair);
        reserve(&ParamSet::hyb);
        reserve(&ParamSet::inv_ijkl);
        reserve(&ParamSet::inv_par);
        reserve(&ParamSet::metal);
        reserve(&ParamSet::pi_frag);
        reserve(&ParamSet::qa);
        reserve(&ParamSet::qfrag);
        reserve(&ParamSet::rcov_cn);
        reserve(&ParamSet::rcov_d3);
        reserve(&ParamSet::rep_anb);
        reserve(&ParamSet::repz);
        reserve(&ParamSet::hb_lpd);
        reserve(&ParamSet::hb_bas);
        reserve(&ParamSet::hb_aci);
        reserve(&ParamSet::sqrt_zr4r2);
        reserve(&ParamSet::stors_ijkl);
        reserve(&ParamSet::stors_par);
        reserve(&ParamSet::topo_bonds);
        reserve(&ParamSet::topo_dist);
        reserve(&ParamSet::topo_idx);
        reserve(&ParamSet::topo_off);
        reserve(&ParamSet::tors_ijkl);
        reserve(&ParamSet::tors_par);
        reserve(&ParamSet::xb_axb);
        reserve(&ParamSet::xb_par);
        reserve(&ParamSet::zeta);
    }
    B.topo_off.push_back(0);
    B.hb_nbB_off.push_back(0);
    B.hb_nbC_off.push_back(0);
    int aoff = 0, foff = 0;
    auto seg = [&](int k, int m) -> int32_t& { return B.mol_seg[(size_t)k * (nm + 1) + m]; };
    for (int m = 0; m < nm; ++m) {
        const ParamSet& P = mols[m];
        B.mol_off[m] = aoff;
        seg(0, m) = (int32_t)B.bond_hbH.size();
        seg(1, m) = (int32_t)(B.angle_par.size() / GFNFF_ANGLE_NPAR);
        seg(2, m) = (int32_t)(B.tors_par.size() / GFNFF_TORS_NPAR);
        seg(3, m) = (int32_t)(B.brep_par.size() / GFNFF_REP_NPAR);
        seg(4, m) = (int32_t)B.hb_case.size();
        seg(5, m) = (int32_t)B.atm_scale.size();
        seg(6, m) = (int32_t)(B.batm_par.size() / GFNFF_BATM_NPAR);
        seg(7, m) = (int32_t)(B.inv_par.size() / GFNFF_INV_NPAR);
        seg(8, m) = (int32_t)B.stors_par.size();
        seg(9, m) = (int32_t)(B.xb_par.size() / GFNFF_XB_NPAR);
        append(B.Z, P.Z);
        for (int i = 0; i < P.n; ++i) B.elem.push_back(seen[P.Z[i]]);
        append_shift(B.frag, P.frag, foff);
        append(B.chi_base, P.chi_base);
        append(B.gam_eff, P.gam_eff);
        append(B.alpeeq, P.alpeeq);
        append(B.cnf, P.cnf);
        append(B.rcov_cn, P.rcov_cn);
        append(B.rcov_d3, P.rcov_d3);
        append(B.hb_rad, P.hb_rad);
        append(B.metal, P.metal);
        append(B.qa, P.qa);
        append(B.zeta, P.zeta);
        append(B.sqrt_zr4r2, P.sqrt_zr4r2);
        append(B.rep_anb, P.rep_anb);
        append(B.repz, P.repz);
        append(B.hb_lpd, P.hb_lpd);
        append(B.hb_bas, P.hb_bas);
        append(B.hb_aci, P.hb_aci);
        append(B.qfrag, P.qfrag);
        append(B.hyb, P.hyb);
        append(B.pi_frag, P.pi_frag);
        append(B.cn_topo, P.cn_topo);
        append_shift(B.topo_bonds, P.topo_bonds, aoff);
        append_csr(B.topo_off, P.topo_off, (int)B.topo_idx.size());
        append_shift(B.topo_idx, P.topo_idx, aoff);
        append(B.topo_dist, P.topo_dist);
        append_shift(B.bond_ij, P.bond_ij, aoff);
        append(B.bond_par, P.bond_par);
        append_shift(B.bond_hbH, P.bond_hbH, aoff, true);
        append_shift(B.hbcn_pair, P.hbcn_pair, aoff);
        append_shift(B.angle_ijk, P.angle_ijk, aoff);
        append(B.angle_par, P.angle_par);
        append_shift(B.tors_ijkl, P.tors_ijkl, aoff);
        append(B.tors_par, P.tors_par);
        append_shift(B.inv_ijkl, P.inv_ijkl, aoff);
        append(B.inv_par, P.inv_par);
        append_shift(B.stors_ijkl, P.stors_ijkl, aoff);
        append(B.stors_par, P.stors_par);
        append_shift(B.brep_ij, P.brep_ij, aoff);
        append(B.brep_par, P.brep_par);
        append_shift(B.hb_ahb, P.hb_ahb, aoff);
        append(B.hb_case, P.hb_case);
        append(B.hb_par, P.hb_par);
        append_csr(B.hb_nbB_off, P.hb_nbB_off, (int)B.hb_nbB_idx.size());
        append_shift(B.hb_nbB_idx, P.hb_nbB_idx, aoff);
        append_shift(B.hb_parentC, P.hb_parentC, aoff, true);
        append_csr(B.hb_nbC_off, P.hb_nbC_off, (int)B.hb_nbC_idx.size());
        append_shift(B.hb_nbC_idx, P.hb_nbC_idx, aoff);
        append_shift(B.xb_axb, P.xb_axb, aoff);
        append(B.xb_par, P.xb_par);
        append_shift(B.atm_ijk, P.atm_ijk, aoff);
        append(B.atm_scale, P.atm_scale);
        append_shift(B.batm_ijk, P.batm_ijk, aoff);
        append(B.batm_par, P.batm_par);
        aoff += P.n;
        foff += P.nfrag;
        B.charge += P.charge;
    }
    B.mol_off[nm] = aoff;
    seg(0, nm) = (int32_t)B.bond_hbH.size();
    seg(1, nm) = (int32_t)(B.angle_par.size() / GFNFF_ANGLE_NPAR);
    seg(2, nm) = (int32_t)(B.tors_par.size() / GFNFF_TORS_NPAR);
    seg(3, nm) = (int32_t)(B.brep_par.size() / GFNFF_REP_NPAR);
    seg(4, nm) = (int32_t)B.hb_case.size();
    seg(5, nm) = (int32_t)B.atm_scale.size();
    seg(6, nm) = (int32_t)(B.batm_par.size() / GFNFF_BATM_NPAR);
    seg(7, nm) = (int32_t)(B.inv_par.size() / GFNFF_INV_NPAR);
    seg(8, nm) = (int32_t)B.stors_par.size();
    seg(9, nm) = (int32_t)(B.xb_par.size() / GFNFF_XB_NPAR);
    B.n = aoff;
    B.nfrag = foff;
    return kOk;
}

namespace {

// Hydrogen-bond and halogen-bond lists of the current geometry
// (detectHydrogenBondsNative, gfnff_method.cpp:7936-8494;
// detectHalogenBondsNative :8520-8662).  Replaces P's hb_* and xb_* lists;
// with cross_ref also derives the bond/HB cross reference (bond_hbH,
// hbcn_pair) that the set-up keeps for the whole run (gfnff_method.cpp:3705-3731).
void detect_hbxb(const Mol& m, const Topo& t, const Options& opt, ParamSet& P, bool cross_ref) {
    const int n = m.n;
    const int32_t* Z = m.Z;
    const double* xyz = m.x;
    P.hb_ahb.clear();
    P.hb_case.clear();
    P.hb_par.clear();
    P.hb_nbB_off.clear();
    P.hb_nbB_idx.clear();
    P.hb_parentC.clear();
    P.hb_nbC_off.clear();
    P.hb_nbC_idx.clear();
    P.xb_axb.clear();
    P.xb_par.clear();
    {
        const auto& q = t.qa;
        std::vector<double> cb(n), ca(n);
        for (int i = 0; i < n; ++i) {
            cb[i] = hb_basicity[Z[i]];
            ca[i] = hb_acidity[Z[i]];
            if (Z[i] == 8) {
                bool carbonyl = t.adj[i].size() == 1 && t.pi[i] != 0 && Z[t.adj[i][0]] == 6 && t.pi[t.adj[i][0]] != 0;
                bool nitro = t.adj[i].size() == 1 && Z[t.adj[i][0]] == 7;
                if (carbonyl) cb[i] = 0.68;
                else if (nitro) cb[i] = 0.47;
            } else if (Z[i] == 6 && t.adj[i].size() == 2) {
                double ph = bond_angle(m, t.adj[i][0], i, t.adj[i][1]) * 180.0 / M_PI;
                if (ph < 150.0 && q[i] > -0.4) cb[i] = 1.46;
            }
        }
        for (int i = 0; i < n; ++i) {  // amide H lowers the N acidity (FunctionalGroupDetector::isAmideHydrogen)
            if (Z[i] != 1 || t.adj[i].size() != 1) continue;
            int nn = t.adj[i][0];
            if (!fg_amide_n(m, t, nn)) continue;
            int c = 0;
            for (int nb : t.adj[nn])
                if (Z[nb] == 6 && t.hyb[nb] == 3) ++c;
            if (c == 1) ca[nn] *= 0.80;
        }
        P.hb_bas = cb;
        P.hb_aci = ca;
        std::vector<int> hbh;
        std::vector<char> is_hbh(n, 0);
        for (int h = 0; h < n; ++h) {
            if (Z[h] != 1 || t.hyb[h] == 1) continue;
            int A = -1;
            for (auto& b : t.bonds) {
                if (b.first == h) A = b.second;
                else if (b.second == h) A = b.first;
                if (A != -1) break;
            }
            if (A == -1) continue;
            double thr = 0.01;
            if (Z[A] > 10) thr -= 0.20;
            if (t.hyb[A] == 3 && Z[A] == 6) thr += 0.05;
            if (q[h] > thr) {
                hbh.push_back(h);
                is_hbh[h] = 1;
            }
        }
        const double hbthr1 = 200.0 - std::log10(opt.hb_accuracy) * 50.0;
        const double hbthr2 = 400.0 - std::log10(opt.hb_accuracy) * 50.0;
        auto ab_ok = [&](int a) {
            if (Z[a] == 6 && t.pi[a] == 0) return false;
            double thr = 0.10 + (Z[a] > 10 ? 0.2 : 0.0);
            return !(q[a] > thr);
        };
        // uniform grid over candidate A/B atoms and HB hydrogens
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) {
                lo[c] = std::min(lo[c], xyz[3 * i + c]);
                hi[c] = std::max(hi[c], xyz[3 * i + c]);
            }
        const double cell = std::sqrt(hbthr2);
        int nc[3];
        for (int c = 0; c < 3; ++c) nc[c] = std::max(1, (int)std::ceil((hi[c] - lo[c]) / cell) + 1);
        auto cid = [&](int i, int* o) {
            for (int c = 0; c < 3; ++c) o[c] = std::min(nc[c] - 1, (int)((xyz[3 * i + c] - lo[c]) / cell));
        };
        std::map<long, std::vector<int>> hcell;
        for (int h : hbh) {
            int o[3];
            cid(h, o);
            hcell[((long)o[0] * nc[1] + o[1]) * nc[2] + o[2]].push_back(h);
        }
        std::vector<int> hb_A, hb_H, hb_B, hb_cs;
        // A-B candidate pairs: (i, j) with j < i, legacy order
        std::vector<int> cand;
        for (int i = 0; i < n; ++i)
            if (ab_ok(i)) cand.push_back(i);
        std::map<long, std::vector<int>> acell;
        for (int i : cand) {
            int o[3];
            cid(i, o);
            acell[((long)o[0] * nc[1] + o[1]) * nc[2] + o[2]].push_back(i);
        }
        std::vector<std::vector<int>> hyd_on(n);
        for (auto& b : t.bonds) {
            if (is_hbh[b.first]) hyd_on[b.second].push_back(b.first);
            if (is_hbh[b.second]) hyd_on[b.first].push_back(b.second);
        }
        for (int i : cand) {
            int oi[3];
            cid(i, oi);
            std::vector<int> js;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        int c0 = oi[0] + dx, c1 = oi[1] + dy, c2 = oi[2] + dz;
                        if (c0 < 0 || c1 < 0 || c2 < 0 || c0 >= nc[0] || c1 >= nc[1] || c2 >= nc[2]) continue;
                        auto it = acell.find(((long)c0 * nc[1] + c1) * nc[2] + c2);
                        if (it == acell.end()) continue;
                        for (int j : it->second)
                            if (j < i) js.push_back(j);
                    }
            std::sort(js.begin(), js.end());
            for (int j : js) {
                double s1 = cb[i] * ca[j], s2 = cb[j] * ca[i];
                if (s1 < 1e-6 && s2 < 1e-6) continue;
                double rab2 = m.dist2(i, j);
                if (rab2 > hbthr1) continue;
                bool nonb = !t.bonded(i, j);
                std::vector<int> hs;
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dz = -1; dz <= 1; ++dz) {
                            int c0 = oi[0] + dx, c1 = oi[1] + dy, c2 = oi[2] + dz;
                            if (c0 < 0 || c1 < 0 || c2 < 0 || c0 >= nc[0] || c1 >= nc[1] || c2 >= nc[2]) continue;
                            auto it = hcell.find(((long)c0 * nc[1] + c1) * nc[2] + c2);
                            if (it == hcell.end()) continue;
                            hs.insert(hs.end(), it->second.begin(), it->second.end());
                        }
                for (int H : hyd_on[i])
                    if (std::find(hs.begin(), hs.end(), H) == hs.end()) hs.push_back(H);
                for (int H : hyd_on[j])
                    if (std::find(hs.begin(), hs.end(), H) == hs.end()) hs.push_back(H);
                std::sort(hs.begin(), hs.end());
                for (int H : hs) {
                    bool bi = t.bonded(i, H), bj = t.bonded(j, H);
                    if (bi && nonb) {
                        hb_A.push_back(i), hb_H.push_back(H), hb_B.push_back(j), hb_cs.push_back(2);
                    } else if (bj && nonb) {
                        hb_A.push_back(j), hb_H.push_back(H), hb_B.push_back(i), hb_cs.push_back(2);
                    } else if (!bi && !bj) {
                        if (rab2 + m.dist2(i, H) + m.dist2(j, H) < hbthr2)
                            hb_A.push_back(i), hb_H.push_back(H), hb_B.push_back(j), hb_cs.push_back(1);
                    }
                }
            }
        }
        P.hb_nbB_off.push_back(0);
        P.hb_nbC_off.push_back(0);
        std::map<std::pair<int, int>, std::vector<int>> ah_b;
        for (size_t e = 0; e < hb_A.size(); ++e) {
            int A = hb_A[e], H = hb_H[e], B = hb_B[e], cs = hb_cs[e];
            int parent = -1;
            if (cs == 2) {
                if (Z[B] == 8 && t.adj[B].size() == 1) {
                    int pa = t.adj[B][0];
                    if (Z[pa] == 6 || Z[pa] == 7) {
                        cs = 3;
                        parent = pa;
                    }
                } else if (Z[B] == 7 && t.adj[B].size() == 2) {
                    cs = 4;
                }
            }
            P.hb_ahb.insert(P.hb_ahb.end(), {A, H, B});
            P.hb_case.push_back(cs);
            P.hb_par.insert(P.hb_par.end(), {cb[A], cb[B], ca[A], ca[B], q[H], q[A], q[B]});
            if (cs >= 2)
                for (int nb : t.adj[B]) P.hb_nbB_idx.push_back(nb);
            P.hb_nbB_off.push_back((int)P.hb_nbB_idx.size());
            P.hb_parentC.push_back(parent);
            if (cs == 3)
                for (int nb : t.adj[parent])
                    if (nb != B) P.hb_nbC_idx.push_back(nb);
            P.hb_nbC_off.push_back((int)P.hb_nbC_idx.size());
            if (cs >= 2 && (Z[B] == 7 || Z[B] == 8)) ah_b[{A, H}].push_back(B);
        }
        // bond/HB cross reference (gfnff_method.cpp:3705-3731)
        for (size_t b = 0; cross_ref && b < P.bond_hbH.size(); ++b) {
            int i = P.bond_ij[2 * b], j = P.bond_ij[2 * b + 1];
            int H = -1, A = -1;
            if (Z[i] == 1) {
                H = i;
                A = j;
            } else if (Z[j] == 1) {
                H = j;
                A = i;
            } else {
                continue;
            }
            if (Z[A] != 7 && Z[A] != 8) continue;
            auto it = ah_b.find({A, H});
            if (it == ah_b.end() || it->second.empty()) continue;
            P.bond_hbH[b] = H;
            for (int B : it->second) {
                P.hbcn_pair.push_back(H);
                P.hbcn_pair.push_back(B);
            }
        }

        // --- halogen (and chalcogen/pnictogen) bonds A-X...B
        // (detectHalogenBondsNative, gfnff_method.cpp:8520-8662): X in
        // {P, S (at most two neighbours), Cl, As, Se, Br, Sb, Te, I} bonded to A,
        // acceptors B within 10 Bohr of X in ascending order with the HB
        // basicities, carbon acceptors only in pi systems
        auto xlike = [](int z) {
            return z == 17 || z == 35 || z == 53 || z == 16 || z == 34 || z == 52 || z == 15 || z == 33 || z == 51;
        };
        for (auto& b : t.bonds) {
            int X = -1, A = -1;
            if (xlike(Z[b.first])) {
                X = b.first;
                A = b.second;
            } else if (xlike(Z[b.second])) {
                X = b.second;
                A = b.first;
            }
            if (X < 0) continue;
            if (Z[X] == 16 && t.adj[X].size() > 2) continue;
            for (int B = 0; B < n; ++B) {
                // within 10 Bohr of X: strict below the cell-list size, inclusive
                // from it on (forEachNeighbor's r2 <= cutoff^2, spatial_cell_list.h:171)
                const double rxb2 = m.dist2(X, B);
                if (n >= kCellListMinAtoms ? rxb2 > 100.0 : rxb2 >= 100.0) continue;
                if (B == A || B == X || cb[B] < 1e-6) continue;
                if (Z[B] == 6) {
                    if (t.pi[B] == 0 || q[B] >= 0.05) continue;
                } else if (q[B] > 0.05) {
                    continue;
                }
                if (t.bonded(X, B)) continue;
                P.xb_axb.insert(P.xb_axb.end(), {A, X, B});
                P.xb_par.insert(P.xb_par.end(), {xb_acidity[Z[X]], q[X], q[B]});
            }
        }
    }
}

// EEQ electrostatic energy of charges q with the element parameters
// (GFNFF::calculateEEQEnergy, gfnff_method.cpp:7505-7553):
//   sum_{i<j} q_i q_j erf(g r) / r,  g = 1 / sqrt(a_i + a_j)
//   + sum_i [-q_i (chi_i + cnf_i sqrt(cn_i)) + q_i^2 (gam_i + sqrt(2/pi) / sqrt(a_i)) / 2],
// a = alpha^2; used to place the charge of a two-fragment system.
double eeq_energy(const Mol& m, const std::vector<double>& q, const std::vector<double>& cn) {
    const double TSQRT2PI = 0.79788456080287;
    double e = 0.0;
    for (int i = 0; i < m.n; ++i) {
        const int zi = m.Z[i];
        const double ai = alp[zi - 1] * alp[zi - 1];
        for (int j = 0; j < i; ++j) {
            const int zj = m.Z[j];
            const double aj = alp[zj - 1] * alp[zj - 1];
            const double r = std::sqrt(m.dist2(i, j));
            const double g = 1.0 / std::sqrt(ai + aj);
            e += q[i] * q[j] * (std::erf(g * r) / r);
        }
        const double self = gam[zi - 1] + TSQRT2PI / std::sqrt(ai);
        const double en = chi[zi - 1] + cnf[zi - 1] * std::sqrt(cn[i]);
        e += -q[i] * en + 0.5 * q[i] * q[i] * self;
    }
    return e;
}

}  // namespace

int build_params(int n, const int32_t* Z, const double* xyz, int charge, const int32_t* forced, int nforced,
                 const Options& opt, ParamSet& P, std::string& msg) {
    if (n <= 0 || !Z || !xyz) {
        msg = "empty molecule";
        return kErrArgs;
    }
    for (int i = 0; i < n; ++i)
        if (Z[i] < 1 || Z[i] > 86) {
            msg = "element outside H..Rn";
            return kErrUnsupported;
        }
    for (int i = 0; i < 3 * n; ++i)
        if (!std::isfinite(xyz[i])) {
            msg = "non-finite coordinates";
            return kErrGeometry;
        }
    // Two atoms closer than 1e-10 Bohr: the reference's Phase-1 EEQ falls back to
    // placeholder charges (eeq_solver.cpp:2843-2846) and setMolecule refuses the
    // molecule (qm_methods/gfnff_method.cpp:43-51).  Sweep over x-sorted atoms.
    {
        std::vector<int> ord(n);
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return xyz[3 * a] < xyz[3 * b]; });
        for (int u = 0; u < n; ++u)
            for (int w = u + 1; w < n && xyz[3 * ord[w]] - xyz[3 * ord[u]] < 1e-10; ++w) {
                const int i = ord[u], j = ord[w];
                const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                             dz = xyz[3 * i + 2] - xyz[3 * j + 2];
                if (std::sqrt(dx * dx + dy * dy + dz * dz) < 1e-10) {
                    msg = "atoms " + std::to_string(std::min(i, j)) + " and " + std::to_string(std::max(i, j)) +
                          " coincide (zero distance): the EEQ charges would be placeholders";
                    return kErrEEQ;
                }
            }
    }
    Mol m{n, Z, xyz};
    Topo t;
    detect_bonds(m, forced, nforced, t);
    detect_fragments(n, t);
    // molecular charge over the fragments (gfnff_method.cpp:9062-9101): one
    // fragment takes it; more than two put it on fragment 0; two fragments try
    // both placements below and keep the lower EEQ energy
    t.qfrag.assign(t.nfrag, 0.0);
    if (t.nfrag == 1 || charge != 0) t.qfrag[0] = charge;
    topo_cn(m, t);
    hybridization(m, t);
    pi_systems(m, t);
    rings(n, t);
    t.is_metal.assign(n, false);
    for (int i = 0; i < n; ++i) {
        int z = Z[i];
        t.is_metal[i] = metal_type[z - 1] > 0 || (z >= 21 && z <= 30) || (z >= 39 && z <= 48) || (z >= 57 && z <= 80);
    }
    if (!phase1_charges(m, t, opt.threads, msg)) return kErrEEQ;
    if (t.nfrag == 2 && charge != 0) {
        const double ea = eeq_energy(m, t.qa, t.cn);
        const std::vector<double> qa_a = t.qa;
        t.qfrag = {0.0, (double)charge};
        if (!phase1_charges(m, t, opt.threads, msg)) return kErrEEQ;
        const double eb = eeq_energy(m, t.qa, t.cn);
        if (ea <= eb) {
            t.qfrag = {(double)charge, 0.0};
            t.qa = qa_a;
        }
    }
    t.dxi = calc_dxi(m, t);
    t.alpeeq = calc_alpeeq(m, t);
    t.dgam = calc_dgam(m, t, t.qa);
    {
        std::vector<bool> pia = eeq_pi_atoms(m, t);
        t.amide_h = eeq_amide_h(m, t, eeq_amide_n(m, t, pia));
    }
    std::vector<int> ipis;
    if (!pi_system_charges(m, t, opt.threads, ipis, msg)) return kErrEEQ;
    std::vector<double> pibo_b = hueckel_pibo(m, t, ipis);
    std::map<std::pair<int, int>, double> pibo;  // bonds of the pi systems only
    for (size_t b = 0; b < t.bonds.size(); ++b)
        if (pibo_b[b] != 0.0) pibo[{std::min(t.bonds[b].first, t.bonds[b].second), std::max(t.bonds[b].first, t.bonds[b].second)}] = pibo_b[b];
    bond_types(m, t, pibo_b);
    topo_neighbourhood(n, t);

    P = ParamSet();
    P.n = n;
    P.charge = charge;
    P.nfrag = t.nfrag;
    P.Z.assign(Z, Z + n);
    P.frag.resize(n);
    for (int i = 0; i < n; ++i) P.frag[i] = t.frag[i] - 1;
    P.qfrag = t.qfrag;
    P.qa = t.qa;
    P.hyb.assign(t.hyb.begin(), t.hyb.end());
    P.pi_frag.assign(t.pi.begin(), t.pi.end());
    P.cn_topo = t.cn;
    for (auto& b : t.bonds) {
        P.topo_bonds.push_back(b.first);
        P.topo_bonds.push_back(b.second);
    }
    // present elements
    {
        std::vector<int> seen(87, -1);
        P.elem.resize(n);
        for (int i = 0; i < n; ++i) {
            if (seen[Z[i]] < 0) {
                seen[Z[i]] = (int)P.present_Z.size();
                P.present_Z.push_back(Z[i]);
            }
            P.elem[i] = seen[Z[i]];
        }
        set_element_tables(P);
    }
    P.chi_base.resize(n);
    P.gam_eff.resize(n);
    P.alpeeq = t.alpeeq;
    P.cnf.resize(n);
    P.rcov_cn.resize(n);
    P.rcov_d3.resize(n);
    P.hb_rad.resize(n);
    P.metal.resize(n);
    P.zeta.resize(n);
    P.sqrt_zr4r2.resize(n);
    P.rep_anb.resize(n);
    P.repz.resize(n);
    P.hb_lpd.resize(n);
    for (int i = 0; i < n; ++i) {
        int z = Z[i];
        P.chi_base[i] = -chi[z - 1] + t.dxi[i];
        if (t.amide_h[i]) P.chi_base[i] -= 0.02;
        P.gam_eff[i] = gam[z - 1] + t.dgam[i];
        P.cnf[i] = cnf[z - 1];
        P.rcov_cn[i] = rcov_cn_of(z);
        P.rcov_d3[i] = rcov_d3[z - 1];
        P.hb_rad[i] = rad[z - 1];
        P.metal[i] = metal_type[z - 1];
        P.zeta[i] = zeta_scale(z, t.qa[i]);
        P.sqrt_zr4r2[i] = std::sqrt(0.5 * r4r2[z - 1] * std::sqrt((double)z));
        double nbc = (double)t.adj[i].size();
        double fn = 1.0 + (-0.1270) / (1.0 + nbc * nbc);
        P.rep_anb[i] = repan[z - 1] * (1.0 + t.qa[i] * 0.3480) * fn;
        P.repz[i] = repz[z - 1];
        // case-4 lone-pair distance from the static element table
        // (ff_workspace_gfnff.cpp:1011-1014), not from the per-atom repz
        P.hb_lpd[i] = 0.50 - 0.018 * repz[z - 1];
    }
    // topological neighbourhood CSR
    P.topo_off.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        P.topo_off[i + 1] = P.topo_off[i] + (int)t.tnb[i].size();
        for (auto& pr : t.tnb[i]) {
            P.topo_idx.push_back(pr.first);
            P.topo_dist.push_back(pr.second);
        }
    }

    // --- bonds (generateBondsNative, gfnff_method.cpp:7557-7607: 1.3*(r_i+r_j), no fat)
    std::vector<std::pair<int, int>> pbonds;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (m.dist(i, j) < 1.3 * (rad_bohr(Z[i]) + rad_bohr(Z[j]))) pbonds.emplace_back(i, j);
    for (auto& b : pbonds) {
        auto pit = pibo.find(b);
        BondPar bp = bond_params(m, t, b.first, b.second, pit == pibo.end() ? 0.0 : pit->second);
        P.bond_ij.push_back(b.first);
        P.bond_ij.push_back(b.second);
        double v[8] = {bp.r0b_i, bp.r0b_j, bp.cnfak_i, bp.cnfak_j, bp.rabshift, bp.ff, bp.alpha, bp.fc};
        P.bond_par.insert(P.bond_par.end(), v, v + 8);
        P.bond_hbH.push_back(-1);
    }

    // --- angles (generateAnglesNative, gfnff_method.cpp:7658-7743)
    {
        struct A3 {
            int i, j, k;
            double fc, th0;
        };
        std::vector<A3> av;
        for (int c = 0; c < n; ++c) {
            const auto& nb = t.adj[c];
            if (nb.size() <= 1 || nb.size() > 6) continue;
            for (size_t a = 0; a < nb.size(); ++a)
                for (size_t b = a + 1; b < nb.size(); ++b) {
                    int ai = nb[a], ak = nb[b];
                    double v1 = m.dist(ai, c), v2 = m.dist(ak, c);
                    if (v1 < 1e-10 || v2 < 1e-10) continue;
                    double cur = bond_angle(m, ai, c, ak);
                    double fc, th0;
                    angle_params(m, t, pibo, ai, c, ak, cur, fc, th0);
                    if (fc < 1e-10) continue;
                    if (metal_type[Z[c] - 1] > 0 && cur * 180.0 / M_PI < 60.0) continue;
                    av.push_back({ai, c, ak, fc, th0});
                }
        }
        std::sort(av.begin(), av.end(), [](const A3& x, const A3& y) {
            if (x.j != y.j) return x.j < y.j;
            if (x.i != y.i) return x.i < y.i;
            return x.k < y.k;
        });
        for (auto& a : av) {
            P.angle_ijk.push_back(a.i);
            P.angle_ijk.push_back(a.j);
            P.angle_ijk.push_back(a.k);
            P.angle_par.push_back(a.fc);
            P.angle_par.push_back(a.th0);
        }
    }

    // --- torsions (generateTorsionsNative, gfnff_torsions.cpp:1561-2050)
    {
        std::vector<int32_t> extra_ijkl;
        std::vector<double> extra_par;
        std::set<std::array<int, 4>> done;
        for (size_t bi = 0; bi < t.bonds.size(); ++bi) {
            int j = t.bonds[bi].first, k = t.bonds[bi].second;
            if (j >= k) continue;
            if (t.adj[j].size() < 2 || t.adj[k].size() < 2) continue;
            int bt = t.btype[bi];
            if (bt == 3 || bt == 6) continue;
            for (int i : t.adj[j]) {
                if (i == k) continue;
                for (int l : t.adj[k]) {
                    if (l == j || i == l) continue;
                    std::array<int, 4> f = {i, j, k, l}, r = {l, k, j, i};
                    auto key = f < r ? f : r;
                    if (done.count(key)) continue;
                    done.insert(key);
                    double aijk = bond_angle(m, i, j, k), ajkl = bond_angle(m, j, k, l);
                    const double athr = 170.0 * M_PI / 180.0;
                    if (aijk > athr || ajkl > athr) continue;
                    int rsz = same_ring(t, j, k);
                    bool inr = rsz > 0;
                    TorsPar tp = tors_params(m, t, pibo, i, j, k, l, inr, rsz, bt);
                    if (std::abs(tp.V) >= 1.0e-3) {
                        int q4[4] = {l, j, k, i};
                        P.tors_ijkl.insert(P.tors_ijkl.end(), q4, q4 + 4);
                        double v[4] = {tp.V, (double)tp.n, tp.phi0, 0.505};
                        P.tors_par.insert(P.tors_par.end(), v, v + 4);
                    }
                    bool sp3c = t.hyb[j] == 3 && t.hyb[k] == 3, sp3o = t.hyb[i] == 3 && t.hyb[l] == 3;
                    if (sp3c && sp3o && !inr && bt < 5 && tp.fij > 0.0 && tp.fkl > 0.0) {
                        double ff = -0.90;
                        if (Z[j] == 7 || Z[k] == 7) ff = 0.70;
                        if (Z[j] == 8 || Z[k] == 8) ff = -2.00;
                        double eb = ff * tp.fij * tp.fkl * tp.fqq;
                        if (std::isfinite(eb)) {
                            int q4[4] = {l, j, k, i};
                            extra_ijkl.insert(extra_ijkl.end(), q4, q4 + 4);
                            double v[4] = {eb, 1.0, M_PI, 0.505};
                            extra_par.insert(extra_par.end(), v, v + 4);
                        }
                    }
                }
            }
        }
        P.tors_ijkl.insert(P.tors_ijkl.end(), extra_ijkl.begin(), extra_ijkl.end());
        P.tors_par.insert(P.tors_par.end(), extra_par.begin(), extra_par.end());
    }

    // --- inversions (generateInversionsNative, gfnff_inversions.cpp:756-870): one
    // term per three-coordinate pi atom or non-pi nitrogen.  The neighbours are
    // ordered by the reference's partial selection sort on the centre distances
    // (Fortran ssort): position p takes the last entry of the non-increasing run
    // starting at p, and the pass stops at the first p whose run has length one.
    for (int i = 0; i < n; ++i) {
        if (t.adj[i].size() != 3) continue;
        bool pi = t.pi[i] > 0, satn = !pi && Z[i] == 7;
        if (!pi && !satn) continue;
        int nb[3] = {t.adj[i][0], t.adj[i][1], t.adj[i][2]};
        double sd[3] = {m.dist(i, nb[0]), m.dist(i, nb[1]), m.dist(i, nb[2])};
        for (int p = 0; p < 2; ++p) {
            int k = p;
            double pp = sd[p];
            for (int j = p + 1; j < 3 && !(sd[j] > pp); ++j) {
                k = j;
                pp = sd[j];
            }
            if (k == p) break;
            sd[k] = sd[p];
            sd[p] = pp;
            std::swap(nb[p], nb[k]);
        }
        double fc, om0, type;
        if (satn) {
            fc = 0.0;
            for (int q = 0; q < 3; ++q) fc += 0.60 * std::sqrt(repz[Z[nb[q]] - 1]);
            om0 = 80.0 * M_PI / 180.0;
            type = -1.0;
        } else {
            int ncarbo = 0, nf = 0;
            for (int q = 0; q < 3; ++q) {
                int zq = Z[nb[q]];
                if (zq == 8 || zq == 16) ++ncarbo;
                if (zq == 9 || zq == 17 || zq == 35 || zq == 53) ++nf;
            }
            double fqq = 1.0 + t.qa[i] * 5.0;
            double sumppi = 0.0;
            for (int q = 0; q < 3; ++q) {
                auto it = pibo.find({std::min(i, nb[q]), std::max(i, nb[q])});
                if (it != pibo.end()) sumppi += it->second;
            }
            double f2 = 1.0 - sumppi * 0.50;
            fc = 1.05 * f2 * fqq;
            if ((Z[i] == 5 || Z[i] == 6) && ncarbo > 0) fc *= 38.0;
            if (Z[i] == 6 && nf > 0 && ncarbo == 0) fc *= 10.0;
            if (Z[i] == 7 && ncarbo > 0) fc *= 10.0 / f2;
            om0 = 0.0;
            type = 0.0;
        }
        P.inv_ijkl.insert(P.inv_ijkl.end(), {i, nb[0], nb[1], nb[2]});
        P.inv_par.insert(P.inv_par.end(), {fc, om0, type});
    }

    // --- triple-bond torsions (generateSTorsionsNative, gfnff_torsions.cpp:2070-2161):
    // C(sp2)-C(sp2)-C#C-C(sp2)-C(sp2) chains, the outer sp2 carbons taken as the
    // largest-index three-coordinate sp2 carbon neighbour
    for (int i = 0; i < n; ++i) {
        if (Z[i] != 6 || t.adj[i].size() != 2) continue;
        for (int nbi : t.adj[i]) {
            if (nbi <= i || Z[nbi] != 6 || t.adj[nbi].size() != 2) continue;
            if (m.dist(i, nbi) > 2.37) continue;
            int jj = -1, kk = -1;
            for (int a : t.adj[i])
                if (a != nbi && Z[a] == 6) jj = a;
            for (int a : t.adj[nbi])
                if (a != i && Z[a] == 6) kk = a;
            if (jj < 0 || kk < 0 || t.hyb[jj] != 2 || t.hyb[kk] != 2) continue;
            auto outer = [&](int c, int excl) {
                int best = -1;
                for (int a : t.adj[c])
                    if (a != excl && t.hyb[a] == 2 && Z[a] == 6 && t.adj[a].size() == 3) best = std::max(best, a);
                return best;
            };
            int ii = outer(jj, i), ll = outer(kk, nbi);
            if (ii < 0 || ll < 0) continue;
            P.stors_ijkl.insert(P.stors_ijkl.end(), {ii, jj, kk, ll});
            P.stors_par.push_back(3.75e-4);
        }
    }

    // --- bonded repulsion over the cached bond list (gfnff_method.cpp:9760-9780)
    for (auto& b : t.bonds) {
        int zi = Z[b.first] - 1, zj = Z[b.second] - 1;
        P.brep_ij.push_back(b.first);
        P.brep_ij.push_back(b.second);
        P.brep_par.push_back(std::sqrt(repa[zi] * repa[zj]));
        P.brep_par.push_back(repz[zi] * repz[zj] * 1.7583);
    }

    // --- ATM over bonded triplets (gfnff_method.cpp:9960-10015)
    {
        std::set<std::tuple<int, int, int>> tri;
        for (int i = 0; i < n; ++i)
            for (int j : t.adj[i]) {
                if (j <= i) continue;
                for (int k : t.adj[i])
                    if (k != j) {
                        std::array<int, 3> a = {i, j, k};
                        std::sort(a.begin(), a.end());
                        tri.insert({a[0], a[1], a[2]});
                    }
                for (int k : t.adj[j])
                    if (k != i) {
                        std::array<int, 3> a = {i, j, k};
                        std::sort(a.begin(), a.end());
                        tri.insert({a[0], a[1], a[2]});
                    }
            }
        for (auto& tr : tri) {
            int a = std::get<0>(tr), b = std::get<1>(tr), c = std::get<2>(tr);
            P.atm_ijk.push_back(a);
            P.atm_ijk.push_back(b);
            P.atm_ijk.push_back(c);
            double s = (a == b) ? ((a == c) ? 1.0 / 6.0 : 0.5) : ((a != c && b != c) ? 1.0 : 0.5);
            P.atm_scale.push_back(s);
        }
    }

    // --- BATM over 1,4 pairs (gfnff_method.cpp:9520-9547 and :3634-3650)
    {
        const double cr = std::pow(0.30, 1.0 / 3.0);
        auto zb = [&](int z) { return z == 1 ? -0.25 * cr : -(double)z * cr; };
        for (int i = 0; i < n; ++i) {
            for (auto& pr : t.tnb[i]) {
                int j = pr.first;
                if (j >= i || pr.second != 3) continue;
                for (int k : t.adj[j]) {
                    P.batm_ijk.insert(P.batm_ijk.end(), {i, j, k});
                    P.batm_par.insert(P.batm_par.end(), {zb(Z[i]), zb(Z[j]), zb(Z[k])});
                }
                for (int k : t.adj[i]) {
                    P.batm_ijk.insert(P.batm_ijk.end(), {i, j, k});
                    P.batm_par.insert(P.batm_par.end(), {zb(Z[i]), zb(Z[j]), zb(Z[k])});
                }
            }
        }
    }

    detect_hbxb(m, t, opt, P, true);
    return kOk;
}

int refresh_hbxb(ParamSet& P, const double* xyz_bohr, const Options& opt, bool& changed, std::string& msg) {
    changed = false;
    if (P.nmol != 1) {
        msg = "hydrogen/halogen-bond lists of a batch are fixed at set-up";
        return kErrArgs;
    }
    Mol m{P.n, P.Z.data(), xyz_bohr};
    Topo t;
    t.adj.assign(P.n, {});
    for (size_t k = 0; k + 1 < P.topo_bonds.size(); k += 2) {
        std::pair<int, int> b(P.topo_bonds[k], P.topo_bonds[k + 1]);
        t.bonds.push_back(b);
        t.adj[b.first].push_back(b.second);
        t.adj[b.second].push_back(b.first);
        t.bond_set.insert(b);
    }
    t.hyb.assign(P.hyb.begin(), P.hyb.end());
    t.pi.assign(P.pi_frag.begin(), P.pi_frag.end());
    t.qa = P.qa;
    ParamSet Q;
    Q.n = P.n;
    detect_hbxb(m, t, opt, Q, false);
    changed = Q.hb_ahb != P.hb_ahb || Q.hb_case != P.hb_case || Q.hb_par != P.hb_par ||
              Q.hb_nbB_off != P.hb_nbB_off || Q.hb_nbB_idx != P.hb_nbB_idx || Q.hb_parentC != P.hb_parentC ||
              Q.hb_nbC_off != P.hb_nbC_off || Q.hb_nbC_idx != P.hb_nbC_idx || Q.xb_axb != P.xb_axb ||
              Q.xb_par != P.xb_par;
    P.hb_bas.swap(Q.hb_bas);  // the carbene basicity depends on the geometry
    P.hb_aci.swap(Q.hb_aci);
    if (changed) {
        P.hb_ahb.swap(Q.hb_ahb);
        P.hb_case.swap(Q.hb_case);
        P.hb_par.swap(Q.hb_par);
        P.hb_nbB_off.swap(Q.hb_nbB_off);
        P.hb_nbB_idx.swap(Q.hb_nbB_idx);
        P.hb_parentC.swap(Q.hb_parentC);
        P.hb_nbC_off.swap(Q.hb_nbC_off);
        P.hb_nbC_idx.swap(Q.hb_nbC_idx);
        P.xb_axb.swap(Q.xb_axb);
        P.xb_par.swap(Q.xb_par);
    }
    return kOk;
}

gfnff_params_view ParamSet::view() const {
    gfnff_params_view v;
    std::memset(&v, 0, sizeof(v));
    v.natoms = n;
    v.nfrag = nfrag;
    v.nelem = nelem;
    v.total_charge = charge;
    v.Z = Z.data();
    v.elem = elem.data();
    v.frag = frag.data();
    v.chi_base = chi_base.data();
    v.gam_eff = gam_eff.data();
    v.alpeeq = alpeeq.data();
    v.cnf = cnf.data();
    v.rcov_cn = rcov_cn.data();
    v.rcov_d3 = rcov_d3.data();
    v.hb_rad = hb_rad.data();
    v.metal = metal.data();
    v.qa = qa.data();
    v.zeta = zeta.data();
    v.sqrt_zr4r2 = sqrt_zr4r2.data();
    v.rep_anb = rep_anb.data();
    v.repz = repz.data();
    v.hb_lpd = hb_lpd.data();
    v.hb_bas = hb_bas.data();
    v.hb_aci = hb_aci.data();
    v.qfrag = qfrag.data();
    v.present_Z = present_Z.data();
    v.nref = nref.data();
    v.refcn = refcn.data();
    v.c6ref = c6ref.data();
    v.topo_off = topo_off.data();
    v.topo_idx = topo_idx.data();
    v.topo_dist = topo_dist.data();
    v.nbond = (int32_t)bond_hbH.size();
    v.bond_ij = bond_ij.data();
    v.bond_par = bond_par.data();
    v.bond_hbH = bond_hbH.data();
    v.nhbcn = (int32_t)(hbcn_pair.size() / 2);
    v.hbcn_pair = hbcn_pair.data();
    v.nangle = (int32_t)(angle_par.size() / GFNFF_ANGLE_NPAR);
    v.angle_ijk = angle_ijk.data();
    v.angle_par = angle_par.data();
    v.ntors = (int32_t)(tors_par.size() / GFNFF_TORS_NPAR);
    v.tors_ijkl = tors_ijkl.data();
    v.tors_par = tors_par.data();
    v.ninv = (int32_t)(inv_par.size() / GFNFF_INV_NPAR);
    v.inv_ijkl = inv_ijkl.data();
    v.inv_par = inv_par.data();
    v.nxb = (int32_t)(xb_par.size() / GFNFF_XB_NPAR);
    v.xb_axb = xb_axb.data();
    v.xb_par = xb_par.data();
    v.nstors = (int32_t)stors_par.size();
    v.stors_ijkl = stors_ijkl.data();
    v.stors_par = stors_par.data();
    v.nbrep = (int32_t)(brep_par.size() / GFNFF_REP_NPAR);
    v.brep_ij = brep_ij.data();
    v.brep_par = brep_par.data();
    v.nhb = (int32_t)hb_case.size();
    v.hb_ahb = hb_ahb.data();
    v.hb_case = hb_case.data();
    v.hb_par = hb_par.data();
    v.hb_nbB_off = hb_nbB_off.data();
    v.hb_nbB_idx = hb_nbB_idx.data();
    v.hb_parentC = hb_parentC.data();
    v.hb_nbC_off = hb_nbC_off.data();
    v.hb_nbC_idx = hb_nbC_idx.data();
    v.natm = (int32_t)atm_scale.size();
    v.atm_ijk = atm_ijk.data();
    v.atm_scale = atm_scale.data();
    v.nbatm = (int32_t)(batm_par.size() / GFNFF_BATM_NPAR);
    v.batm_ijk = batm_ijk.data();
    v.batm_par = batm_par.data();
    return v;
}

}  // namespace gfnff
