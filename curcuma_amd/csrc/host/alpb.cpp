// ALPB solvation set-up for GFN-FF (host side).
//
// Follows alpb_solvation.cpp:196-403 of the reference (loadSolventParameters,
// setupRadii, the Lebedev grid) with the constants of alpb_parameters.h:20-56.
// The parameter tables are data (host/alpb_tables.inc).

#include "alpb.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>

#include "gfnff_topology.hpp"

namespace {
#include "alpb_tables.inc"

// alpb_parameters.h:21-55 (the solvation code's own Bohr and kcal conversions)
constexpr double kAng = 1.0 / 0.52917726;
constexpr double kKcal = 1.0 / 627.50947428;
constexpr double kPi = 3.1415926535897932384626433832795029;
constexpr double kWSmooth = 0.3 * kAng;               // surface switching half-width
constexpr double kAh3 = -1.0 / (4.0 * kWSmooth * kWSmooth * kWSmooth);
constexpr double kSurfaceTension = 1.0e-5;
constexpr double kLrcut = 35.0 * kAng;
constexpr double kROffset = 2.0 * kAng;
constexpr double kAlpbFactor = 0.571412;

std::string lower(const char* s) {
    std::string r = s ? s : "";
    for (auto& c : r) c = (char)std::tolower((unsigned char)c);
    return r;
}

const alpb_tab::Solvent* find_solvent(const std::string& name) {
    for (const auto& s : alpb_tab::solvents)
        if (name == s.name) return &s;
    return nullptr;
}

// Octahedral orbits of a Lebedev grid: code 1 (+-1, 0, 0), 3 (+-a, +-a, +-a)
// with a = 1/sqrt 3, 4 (+-a, +-a, +-b), 5 (+-a, +-b, 0), 6 (+-a, +-b, +-c);
// every permutation of the coordinates and every sign, each point weight v.
void orbit(std::vector<double>& g, int code, double a, double b, double v) {
    std::vector<std::array<double, 3>> base;
    switch (code) {
        case 1: base = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}; break;
        case 3: {
            const double t = 1.0 / std::sqrt(3.0);
            base = {{t, t, t}};
            break;
        }
        case 4: {
            const double c = std::sqrt(1.0 - 2.0 * a * a);
            base = {{a, a, c}, {a, c, a}, {c, a, a}};
            break;
        }
        case 5: {
            const double c = std::sqrt(1.0 - a * a);
            base = {{a, c, 0}, {c, a, 0}, {a, 0, c}, {c, 0, a}, {0, a, c}, {0, c, a}};
            break;
        }
        default: {
            const double c = std::sqrt(1.0 - a * a - b * b);
            base = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
            break;
        }
    }
    for (const auto& p : base)
        for (int s = 0; s < 8; ++s) {
            const double x = (s & 1) ? -p[0] : p[0], y = (s & 2) ? -p[1] : p[1], z = (s & 4) ? -p[2] : p[2];
            // a zero coordinate takes one sign only
            if (((s & 1) && p[0] == 0.0) || ((s & 2) && p[1] == 0.0) || ((s & 4) && p[2] == 0.0)) continue;
            g.insert(g.end(), {x, y, z, v});
        }
}
}  // namespace

namespace gfnff {

std::vector<double> lebedev230() {
    // orbit parameters of the 230-point rule (Lebedev and Laikov, Doklady
    // Mathematics 59 (1999) 477; the values of the reference's
    // alpb_solvation.cpp:121-173)
    std::vector<double> g;
    g.reserve(230 * 4);
    orbit(g, 1, 0, 0, -0.5522639919727325e-1);
    orbit(g, 3, 0, 0, 0.4450274607445226e-2);
    orbit(g, 4, 0.4492044687397611e+0, 0, 0.4496841067921404e-2);
    orbit(g, 4, 0.2520419490210201e+0, 0, 0.5049153450478750e-2);
    orbit(g, 4, 0.6981906658447242e+0, 0, 0.3976408018051883e-2);
    orbit(g, 4, 0.6587405243460960e+0, 0, 0.4401400650381014e-2);
    orbit(g, 4, 0.4038544050097660e-1, 0, 0.1724544350544401e-1);
    orbit(g, 5, 0.5823842309715585e+0, 0, 0.4231083095357343e-2);
    orbit(g, 5, 0.3545877390518688e+0, 0, 0.5198069864064399e-2);
    orbit(g, 6, 0.2272181808998187e+0, 0.4864661535886647e+0, 0.4695720972568883e-2);
    return g;
}

std::string alpb_solvent_name(const char* name) {
    const std::string s = lower(name);
    if (find_solvent(s)) return s;
    for (const auto& a : alpb_tab::aliases)
        if (s == a.alias) return a.name;
    return std::string();
}

gfnff_alpb_view AlpbSet::view() const {
    gfnff_alpb_view v{};
    v.natoms = n;
    v.hbond = hbond;
    v.nang = nang;
    v.keps = keps;
    v.alpbet = alpbet;
    v.born_scale = born_scale;
    v.gshift = gshift;
    v.lrcut = lrcut;
    v.srcut = srcut;
    v.vdwr = vdwr.data();
    v.rho = rho.data();
    v.svdw = svdw.data();
    v.vdwsa = vdwsa.data();
    v.wrp = wrp.data();
    v.trj_lo = trj_lo.data();
    v.trj_hi = trj_hi.data();
    v.gamsasa = gamsasa.data();
    v.hbmag = hbmag.data();
    v.grid = grid.data();
    return v;
}

// alpb_solvation.cpp:267-403 (GFN-FF parameter branch, setupRadii)
int alpb_setup(const char* solvent, int n, const int32_t* Z, AlpbSet& S, std::string& msg) {
    S = AlpbSet();
    const std::string name = alpb_solvent_name(solvent);
    const alpb_tab::Solvent* p = name.empty() ? nullptr : find_solvent(name);
    if (!p) {
        msg = std::string("ALPB: unknown solvent '") + (solvent ? solvent : "") + "'";
        return kErrArgs;
    }
    S.solvent = name;
    S.n = n;
    S.alpbet = kAlpbFactor / p->eps;
    S.keps = (1.0 / p->eps - 1.0) / (1.0 + S.alpbet);
    S.born_scale = p->born_scale;
    S.gshift = p->shift_kcal * kKcal;
    S.lrcut = kLrcut;
    const double offset = p->born_offset * 0.1 * kAng, probe = p->probe_ang * kAng;
    for (auto* v : {&S.vdwr, &S.rho, &S.svdw, &S.vdwsa, &S.wrp, &S.trj_lo, &S.trj_hi, &S.gamsasa, &S.hbmag})
        v->assign(n, 0.0);
    double maxsa = 0.0;
    for (int i = 0; i < n; ++i) {
        const int z = Z[i];
        const bool known = z >= 1 && z <= alpb_tab::kMaxZ;
        const double r = (known ? alpb_tab::vdw_d3_ang[z - 1] : 2.0) * kAng;
        S.vdwr[i] = r;
        S.rho[i] = r * (known ? p->descreening[z - 1] : 0.8);
        S.svdw[i] = r - offset;
        const double sa = r + probe;
        S.vdwsa[i] = sa;
        maxsa = std::max(maxsa, sa);
        S.trj_lo[i] = (sa - kWSmooth) * (sa - kWSmooth);
        S.trj_hi[i] = (sa + kWSmooth) * (sa + kWSmooth);
        // radial weight: the switching function integrated over the shell
        auto shell = [&](double rr) {
            return (0.25 / kWSmooth + 3.0 * kAh3 * (0.2 * rr * rr - 0.5 * rr * sa + sa * sa / 3.0)) * rr * rr * rr;
        };
        S.wrp[i] = shell(sa + kWSmooth) - shell(sa - kWSmooth);
        S.gamsasa[i] = known ? p->surface_tension[z - 1] * 4.0 * kPi * kSurfaceTension : 0.0;
        const double hb = known ? p->hbond[z - 1] : 0.0;
        S.hbmag[i] = std::abs(hb) > 1.0e-3 ? -(hb * hb) * kKcal : 0.0;
        if (S.hbmag[i] < 0.0) S.hbond = 1;
    }
    S.srcut = 2.0 * (kWSmooth + maxsa) + kROffset;
    S.grid = lebedev230();
    S.nang = (int)(S.grid.size() / 4);
    return kOk;
}

}  // namespace gfnff
