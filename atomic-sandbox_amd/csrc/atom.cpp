// Free-atom densities for the Fermi–Dirac cold guess of the SCF (qm/src/nuclear.rs:153-187 free_atom_density,
// qm/src/scf.rs:948-965 cold_guess): the spherical, spin-restricted LSDA Kohn–Sham atom (the reference solves it in
// qm/src/radial/atom.rs), then its density convolved with the nuclear Gaussian of width sigma (nuclear.rs:233-256).
//
// Method. Log grid r_i = r0 e^{i dx}; u(r) = r R(r) = sqrt(r) w(x) turns the radial equation into
// w'' = [2 r^2 (V - eps) + (l + 1/2)^2] w, integrated by Numerov. Each (n, l) level is found by bisection on the node
// count of the outward solution (the count rises by one as eps crosses the eigenvalue); the orbital is the outward
// solution up to the outermost classical turning point joined to the decaying inward solution. Occupations are the
// aufbau of the level energies (each (n, l) holds 2(2l+1) electrons, the last one fractionally, spread over m). The
// potential -Z/r + V_H + v_xc is iterated with linear mixing of the density until the density change is below 1e-8.
// The profile only seeds the SCF: the converged answer does not depend on it.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "lira_internal.hpp"
#include "lsda.hpp"

namespace lira {

namespace {

constexpr double R_MIN = 1e-6, R_MAX = 60.0, DX = 0.01;

struct LogGrid {
    std::vector<double> r;
    double x0, dx;
    explicit LogGrid(double zf) : x0(std::log(R_MIN / zf)), dx(DX) {
        const int n = (int)std::ceil((std::log(R_MAX) - x0) / dx) + 1;
        r.resize(n);
        for (int i = 0; i < n; i++) r[i] = std::exp(x0 + i * dx);
    }
    int size() const { return (int)r.size(); }
    // trapezoid of f(r) dr = f r dx
    double integrate(const std::vector<double> &f) const {
        double s = 0.0;
        for (int i = 0; i + 1 < size(); i++) s += 0.5 * (f[i] * r[i] + f[i + 1] * r[i + 1]) * dx;
        return s;
    }
};

// Numerov outward with the node count; returns the number of sign changes over the grid
int shoot_nodes(const LogGrid &g, const std::vector<double> &V, int l, double eps) {
    const int n = g.size();
    const double h2 = g.dx * g.dx / 12.0, lh = (l + 0.5) * (l + 0.5);
    auto G = [&](int i) { return 2.0 * g.r[i] * g.r[i] * (V[i] - eps) + lh; };
    double w0 = std::pow(g.r[0], l + 0.5), w1 = std::pow(g.r[1], l + 0.5);
    double g0 = G(0), g1 = G(1);
    int nodes = 0;
    for (int i = 1; i + 1 < n; i++) {
        const double g2 = G(i + 1);
        // deep in the forbidden region (h^2 g / 12 > 1/2 only far outside every turning point) the solution just
        // grows: no node can follow, and Numerov would lose stability there
        if (h2 * g2 > 0.5 && g.r[i + 1] > 1.0) break;
        const double w2 = (2.0 * w1 * (1.0 + 5.0 * h2 * g1) - w0 * (1.0 - h2 * g0)) / (1.0 - h2 * g2);
        if ((w2 < 0.0) != (w1 < 0.0) && w2 != 0.0) nodes++;
        w0 = w1;
        w1 = w2;
        g0 = g1;
        g1 = g2;
        if (std::fabs(w1) > 1e150) {
            w0 *= 1e-150;
            w1 *= 1e-150;
        }
    }
    return nodes;
}

// the normalised orbital w (u = sqrt(r) w, int u^2 dr = 1) of level (n, l) at eigenvalue eps
std::vector<double> orbital(const LogGrid &g, const std::vector<double> &V, int l, double eps) {
    const int n = g.size();
    const double h2 = g.dx * g.dx / 12.0, lh = (l + 0.5) * (l + 0.5);
    std::vector<double> Gv(n), w(n, 0.0);
    for (int i = 0; i < n; i++) Gv[i] = 2.0 * g.r[i] * g.r[i] * (V[i] - eps) + lh;
    int turn = n / 2;
    for (int i = n - 2; i > 0; i--)
        if (Gv[i] < 0.0) {
            turn = i;
            break;
        }
    w[0] = std::pow(g.r[0], l + 0.5);
    w[1] = std::pow(g.r[1], l + 0.5);
    for (int i = 1; i < turn; i++) {
        w[i + 1] = (2.0 * w[i] * (1.0 + 5.0 * h2 * Gv[i]) - w[i - 1] * (1.0 - h2 * Gv[i - 1])) / (1.0 - h2 * Gv[i + 1]);
        if (std::fabs(w[i + 1]) > 1e150)
            for (int j = 0; j <= i + 1; j++) w[j] *= 1e-150;
    }
    const double wt = w[turn];
    // the inward solution starts where Numerov is still stable (h^2 g / 12 < 1/2); the orbital is 0 beyond
    int last = n - 1;
    for (int i = turn + 2; i < n; i++)
        if (h2 * Gv[i] > 0.5) {
            last = i;
            break;
        }
    std::vector<double> in(n, 0.0);
    in[last] = 0.0;
    in[last - 1] = 1e-300;
    for (int i = last - 1; i > turn; i--) {
        in[i - 1] = (2.0 * in[i] * (1.0 + 5.0 * h2 * Gv[i]) - in[i + 1] * (1.0 - h2 * Gv[i + 1])) / (1.0 - h2 * Gv[i - 1]);
        if (std::fabs(in[i - 1]) > 1e150)
            for (int j = i - 1; j < n; j++) in[j] *= 1e-150;
    }
    const double scale = in[turn] != 0.0 ? wt / in[turn] : 0.0;
    for (int i = turn + 1; i < n; i++) w[i] = i <= last ? in[i] * scale : 0.0;
    // normalise: int u^2 dr = sum w^2 r^2 dx
    std::vector<double> f(n);
    for (int i = 0; i < n; i++) f[i] = w[i] * w[i] * g.r[i];
    const double norm = g.integrate(f);
    const double s = norm > 0.0 ? 1.0 / std::sqrt(norm) : 0.0;
    for (double &x : w) x *= s;
    return w;
}

struct Level {
    int n, l;
    double eps;
};

// the lowest n - l - 1 node level of angular momentum l, or false if unbound
bool eigen(const LogGrid &g, const std::vector<double> &V, int n, int l, double z, double &eps) {
    const int want = n - l - 1;
    double lo = -z * z - 10.0, hi = -1e-10;
    if (shoot_nodes(g, V, l, hi) <= want) return false;
    for (int it = 0; it < 200 && hi - lo > 1e-13 * std::max(1.0, std::fabs(lo)); it++) {
        const double mid = 0.5 * (lo + hi);
        if (shoot_nodes(g, V, l, mid) > want) hi = mid;
        else lo = mid;
    }
    eps = 0.5 * (lo + hi);
    return true;
}

// spherical restricted LSDA atom: the total density rho(r) on the grid, or empty if it does not solve
std::vector<double> solve_atom(const LogGrid &g, int z) {
    const int n = g.size();
    const double zf = z;
    std::vector<double> rho(n, 0.0), V(n);
    const double mix = 0.3;
    for (int iter = 0; iter < 400; iter++) {
        // potential: -Z/r + V_H + v_xc (restricted: rho_up = rho_dn = rho/2)
        std::vector<double> qin(n), f4(n);
        for (int i = 0; i < n; i++) f4[i] = 4.0 * M_PI * g.r[i] * g.r[i] * rho[i];
        // Q(r) cumulative and the outer integral of 4 pi r rho
        std::vector<double> Q(n, 0.0), outer(n, 0.0);
        Q[0] = 4.0 * M_PI / 3.0 * rho[0] * g.r[0] * g.r[0] * g.r[0];
        for (int i = 1; i < n; i++) Q[i] = Q[i - 1] + 0.5 * (f4[i - 1] * g.r[i - 1] + f4[i] * g.r[i]) * g.dx;
        for (int i = n - 2; i >= 0; i--)
            outer[i] = outer[i + 1] + 0.5 * (f4[i] + f4[i + 1]) * g.dx;  // 4 pi r rho dr = 4 pi r^2 rho dx
        for (int i = 0; i < n; i++) {
            double e, vu, vd;
            exc_point(0.5 * rho[i], 0.5 * rho[i], e, vu, vd);
            V[i] = -zf / g.r[i] + Q[i] / g.r[i] + outer[i] + vu;
        }
        // levels n <= 4, l <= 2 (every shell occupied up to Z = 36) sorted by energy, filled by aufbau
        std::vector<Level> lv;
        for (int l = 0; l <= 2; l++)
            for (int nn = l + 1; nn <= 4; nn++) {
                double eps;
                if (eigen(g, V, nn, l, zf, eps)) lv.push_back({nn, l, eps});
            }
        std::sort(lv.begin(), lv.end(), [](const Level &a, const Level &b) {
            return a.eps < b.eps || (a.eps == b.eps && a.l < b.l);
        });
        double left = zf;
        std::vector<double> out(n, 0.0);
        for (const Level &L : lv) {
            if (left <= 1e-12) break;
            const double occ = std::min(left, 2.0 * (2 * L.l + 1));
            left -= occ;
            const std::vector<double> w = orbital(g, V, L.l, L.eps);
            for (int i = 0; i < n; i++) out[i] += occ * w[i] * w[i] / (4.0 * M_PI * g.r[i]);
        }
        if (left > 1e-9) return {};  // not enough bound levels
        std::vector<double> d(n);
        for (int i = 0; i < n; i++) d[i] = std::fabs(out[i] - rho[i]) * 4.0 * M_PI * g.r[i] * g.r[i];
        const double change = g.integrate(d);
        for (int i = 0; i < n; i++) rho[i] = iter == 0 ? out[i] * mix : rho[i] + mix * (out[i] - rho[i]);
        if (iter > 0 && change < 1e-8) return out;
    }
    return rho;
}

// rho convolved with the normalised 3-D Gaussian of width sigma (spherical; nuclear.rs:239-256): trapezoid in r'
// over |r - r'| <= 10 sigma, the bracket written e^{-(r-r')^2/2s^2} (-expm1(-2 r r'/s^2))
std::vector<double> smear(const LogGrid &g, const std::vector<double> &rho, double sigma) {
    const int n = g.size();
    const double c = 1.0 / (sigma * std::sqrt(2.0 * M_PI)), s2 = sigma * sigma;
    std::vector<double> out(n, 0.0);
    for (int i = 0; i < n; i++) {
        const double ri = g.r[i];
        const int lo = (int)(std::lower_bound(g.r.begin(), g.r.end(), ri - 10.0 * sigma) - g.r.begin());
        const int hi = std::min(n, (int)(std::upper_bound(g.r.begin(), g.r.end(), ri + 10.0 * sigma) - g.r.begin()));
        if (hi <= lo + 1) continue;
        auto f = [&](int j) {
            const double d = ri - g.r[j];
            return g.r[j] * rho[j] * std::exp(-d * d / (2.0 * s2)) * -std::expm1(-2.0 * ri * g.r[j] / s2);
        };
        double acc = 0.0, prev = f(lo);
        for (int j = lo + 1; j < hi; j++) {
            const double cur = f(j);
            acc += 0.5 * (prev + cur) * (g.r[j] - g.r[j - 1]);
            prev = cur;
        }
        out[i] = c * acc / ri;
    }
    return out;
}

// the Kohn-Sham potential -Z/r + V_H + v_xc of a restricted density on the grid
std::vector<double> potential_of(const LogGrid &g, int z, const std::vector<double> &rho) {
    const int n = g.size();
    std::vector<double> f(n), Q(n, 0.0), outer(n, 0.0), V(n);
    for (int i = 0; i < n; i++) f[i] = 4.0 * M_PI * g.r[i] * g.r[i] * rho[i];
    Q[0] = 4.0 * M_PI / 3.0 * rho[0] * g.r[0] * g.r[0] * g.r[0];
    for (int i = 1; i < n; i++) Q[i] = Q[i - 1] + 0.5 * (f[i - 1] * g.r[i - 1] + f[i] * g.r[i]) * g.dx;
    for (int i = n - 2; i >= 0; i--) outer[i] = outer[i + 1] + 0.5 * (f[i] + f[i + 1]) * g.dx;
    for (int i = 0; i < n; i++) {
        double e, vu, vd;
        exc_point(0.5 * rho[i], 0.5 * rho[i], e, vu, vd);
        V[i] = -(double)z / g.r[i] + Q[i] / g.r[i] + outer[i] + vu;
    }
    return V;
}

}  // namespace

RadialProfile free_atom_profile(int z, double sigma) {
    static std::mutex mu;
    static std::map<std::pair<int, double>, RadialProfile> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({z, sigma});
    if (it != cache.end()) return it->second;
    RadialProfile p;
    LogGrid g(std::max(1.0, (double)z));
    p.x0 = g.x0;
    p.dx = g.dx;
    std::vector<double> rho = z >= 1 && z <= 36 ? solve_atom(g, z) : std::vector<double>();
    p.ok = !rho.empty();
    if (p.ok) {
        p.rho = smear(g, rho, sigma);
        // the bound levels of the converged potential, R(r) = u(r) / r on the same grid, ascending in energy
        const std::vector<double> V = potential_of(g, z, rho);
        std::vector<Level> lv;
        for (int l = 0; l <= 2; l++)
            for (int nn = l + 1; nn <= 4; nn++) {
                double eps;
                if (eigen(g, V, nn, l, z, eps)) lv.push_back({nn, l, eps});
            }
        std::sort(lv.begin(), lv.end(), [](const Level &a, const Level &b) {
            return a.eps < b.eps || (a.eps == b.eps && a.l < b.l);
        });
        // every bound level of the converged potential (the occupied ones and the bound unoccupied ones above them)
        for (const Level &L : lv) {
            const std::vector<double> w = orbital(g, V, L.l, L.eps);
            RadialProfile::Orbital o;
            o.l = L.l;
            o.eps = L.eps;
            o.R.resize(g.size());
            for (int i = 0; i < g.size(); i++) o.R[i] = w[i] / std::sqrt(g.r[i]);
            p.orbitals.push_back(std::move(o));
        }
    }
    cache[{z, sigma}] = p;
    return p;
}

// the unsmeared free atom, for tests: levels' eigenvalues and the density's electron count
bool free_atom_levels(int z, std::vector<double> &eps_out, double &electrons) {
    LogGrid g(std::max(1.0, (double)z));
    std::vector<double> rho = solve_atom(g, z);
    if (rho.empty()) return false;
    std::vector<double> f(g.size());
    for (int i = 0; i < g.size(); i++) f[i] = 4.0 * M_PI * g.r[i] * g.r[i] * rho[i];
    electrons = g.integrate(f);
    // the levels of the self-consistent potential
    std::vector<double> Q(g.size(), 0.0), outer(g.size(), 0.0), V(g.size());
    const int n = g.size();
    Q[0] = 4.0 * M_PI / 3.0 * rho[0] * g.r[0] * g.r[0] * g.r[0];
    for (int i = 1; i < n; i++) Q[i] = Q[i - 1] + 0.5 * (f[i - 1] * g.r[i - 1] + f[i] * g.r[i]) * g.dx;
    for (int i = n - 2; i >= 0; i--) outer[i] = outer[i + 1] + 0.5 * (f[i] + f[i + 1]) * g.dx;
    for (int i = 0; i < n; i++) {
        double e, vu, vd;
        exc_point(0.5 * rho[i], 0.5 * rho[i], e, vu, vd);
        V[i] = -(double)z / g.r[i] + Q[i] / g.r[i] + outer[i] + vu;
    }
    eps_out.clear();
    for (int l = 0; l <= 1; l++)
        for (int nn = l + 1; nn <= 3; nn++) {
            double eps;
            eps_out.push_back(eigen(g, V, nn, l, z, eps) ? eps : 1.0);
        }
    return true;
}

}  // namespace lira
