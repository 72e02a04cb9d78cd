// Self-generated norm-conserving pseudopotentials: Z in, the tables of one pseudised ion filtered for a grid spacing h
// out (qm::radial, crates/qm/src/radial/*). Host C++, f64, atomic units (Hartree, Bohr); V is the electron's
// potential energy. Nothing here runs on the GPU: the generator feeds lira_set_ionic, whose tables the device reads.
//
// Pipeline (crates/qm/src/radial/recipe.rs:140-176):
//   spin-restricted LSDA reference atom on the universal log grid (atom.rs) -> core / valence partition
//   (partition.rs) -> LFC partial core at the ledger's crossover (nlcc.rs) -> for each (f, local) of the ladder:
//   Troullier-Martins channels at r_c = f r_peak (tm.rs), unscreening, the Kleinman-Bylander form with its derived
//   support r_beta (kb.rs), the ghost gates G1 (GSS) and G2 (the radial KB atom's exact level count, kb.rs:611-663),
//   the transferability gates T1-T7 (gates.rs: the KB atom solved self-consistently in the cation, s -> p, anion and
//   Hund configurations, log derivatives, hardness) -> T8 on the accepted entry (the filtered re-run at every ladder
//   spacing: h_req, h_eb) -> the band-limited tables for spacing h (filter.rs: King-Smith-Payne-Lin completion, uniform
//   3-D tables).
//
// The molecular gate on f for Z <= 2 (the reference providers' GridDimerGate) runs through the library's own solver
// when asked for (lira_pseudo_generate_gated); without it H and He come out as the reference's ungated generate()
// makes them.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "lira.h"
#include "lsda.hpp"

namespace lira {
namespace radial {

namespace {

// ---- the recipe's constants (crates/qm/src/radial/ledger.rs:177-197) ----
constexpr double R0 = 1e-5, R_MAX = 60.0, DELTA = 0.008;
constexpr double DELTA_CORE = 2.0;
constexpr int L_MAX_ABOVE_VALENCE = 1;
constexpr double F_LADDER[4] = {1.6, 1.4, 1.2, 1.0};
constexpr double NLCC_CROSSOVER = 0.25;
constexpr double V_ION_TAIL = 1e-8;
constexpr double T5_DE_EQUIV = 5e-3, T5_WINDOW = 0.5, T5_MESH = 0.01, T5_RD_FACTOR = 1.2;
constexpr double Q_C_OVER_NYQUIST = 0.7, GAMMA_R = 1.5, SIGMA_LR_OVER_H = 2.0, V_TAIL_FLOOR = 1e-6;
constexpr double EPS_CORE = 1e-3, R_PEAK_FLOOR = 1e-3;
constexpr double SCF_TOL = 1e-10, MIXING = 0.3;
constexpr int SCF_MAX_ITER = 400, DIIS_HISTORY = 6, BACKTRACK_MAX = 6, KSPL_TERMS = 24;
constexpr double KSPL_STEP_OVER_H = 8.0;
constexpr int TABLE_POINTS_PER_H = 256;
constexpr int Z_MAX = 18;  // the elements whose restricted reference needs no near-degenerate shell scan (atom.rs:18)

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Neumaier-compensated sum (qm::grid::neumaier)
struct Neumaier {
    double s = 0.0, c = 0.0;
    void add(double x) {
        const double t = s + x;
        c += std::fabs(s) >= std::fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    double sum() const { return s + c; }
};

double pw92(double nu, double nd, double *vu = nullptr, double *vd = nullptr) {
    double e, a, b;
    exc_point(nu, nd, e, a, b);
    if (vu) *vu = a;
    if (vd) *vd = b;
    return e;
}

// ---- the universal log grid r_i = r0 e^{i delta} (grid.rs) ----
struct LogGrid {
    std::vector<double> r;
    double x0 = 0.0, delta = DELTA;
    LogGrid() {
        const int n = (int)std::floor(std::log(R_MAX / R0) / DELTA) + 1;
        r.resize(n);
        for (int i = 0; i < n; i++) r[i] = R0 * std::exp(i * DELTA);
        x0 = std::log(R0);
    }
    int size() const { return (int)r.size(); }
    // the sphere inside r0 for f ~ a r^p (grid.rs:34-43)
    double inner(const std::vector<double> &f) const {
        const double f0 = f[0], f1 = f[1];
        if (f0 == 0.0 || f0 * f1 <= 0.0) return 0.0;
        const double p = std::log(f1 / f0) / delta;
        return p > -1.0 ? f0 * r[0] / (p + 1.0) : 0.0;
    }
    // integral of f dr: trapezoid in x with Neumaier summation, plus the inner sphere (grid.rs:24-32)
    double integrate(const std::vector<double> &f) const {
        const int n = size();
        Neumaier s;
        for (int i = 0; i < n; i++) s.add(f[i] * r[i]);
        return inner(f) + (s.sum() - 0.5 * (f[0] * r[0] + f[n - 1] * r[n - 1])) * delta;
    }
    // c_i = integral_0^{r_i} f dr, 4th order in x (grid.rs:45-57)
    std::vector<double> cumulative(const std::vector<double> &f) const {
        const int n = size();
        std::vector<double> g(n), c(n, 0.0);
        for (int i = 0; i < n; i++) g[i] = f[i] * r[i];
        const double w = delta / 24.0;
        c[0] = inner(f);
        c[1] = c[0] + w * (9.0 * g[0] + 19.0 * g[1] - 5.0 * g[2] + g[3]);
        for (int i = 1; i < n - 2; i++) c[i + 1] = c[i] + w * (-g[i - 1] + 13.0 * g[i] + 13.0 * g[i + 1] - g[i + 2]);
        c[n - 1] = c[n - 2] + w * (g[n - 4] - 5.0 * g[n - 3] + 19.0 * g[n - 2] + 9.0 * g[n - 1]);
        return c;
    }
    double partial(const std::vector<double> &f, int i_end) const { return cumulative(f)[i_end]; }
    int nearest(double x) const {
        const double s = std::round((std::log(x) - x0) / delta);
        const int i = s > 0.0 ? (int)s : 0;
        return std::min(i, size() - 1);
    }
    // df/dr: 5-point in x, one-sided at both ends (grid.rs:68-85)
    std::vector<double> derivative(const std::vector<double> &f) const {
        const int n = size();
        const double d = 12.0 * delta;
        std::vector<double> out(n);
        for (int i = 0; i < n; i++) {
            double fx;
            if (i >= 2 && i + 2 < n) fx = (f[i - 2] - 8.0 * f[i - 1] + 8.0 * f[i + 1] - f[i + 2]) / d;
            else if (i == 0) fx = (-25.0 * f[0] + 48.0 * f[1] - 36.0 * f[2] + 16.0 * f[3] - 3.0 * f[4]) / d;
            else if (i == 1) fx = (-3.0 * f[0] - 10.0 * f[1] + 18.0 * f[2] - 6.0 * f[3] + f[4]) / d;
            else {
                const int o = n - 5;
                if (i == n - 1) fx = (3.0 * f[o] - 16.0 * f[o + 1] + 36.0 * f[o + 2] - 48.0 * f[o + 3] + 25.0 * f[o + 4]) / d;
                else fx = (-f[o] + 6.0 * f[o + 1] - 18.0 * f[o + 2] + 10.0 * f[o + 3] + 3.0 * f[o + 4]) / d;
            }
            out[i] = fx / r[i];
        }
        return out;
    }
};

const LogGrid &universal_grid() {
    static const LogGrid g;
    return g;
}

std::vector<double> shell_weight(const LogGrid &g, const std::vector<double> &rho) {
    std::vector<double> out(g.size());
    for (int i = 0; i < g.size(); i++) out[i] = 4.0 * M_PI * g.r[i] * g.r[i] * rho[i];
    return out;
}

// ---- Numerov in x = ln r: y'' = g y, u = e^{x/2} y (numerov.rs) ----
std::vector<double> g_coeff(const LogGrid &grid, const std::vector<double> &v, int l, double e) {
    const double lh = l + 0.5;
    std::vector<double> g(grid.size());
    for (int i = 0; i < grid.size(); i++) g[i] = lh * lh + 2.0 * grid.r[i] * grid.r[i] * (v[i] - e);
    return g;
}

void regular_start(const LogGrid &grid, int l, double z_origin, double &y0, double &y1) {
    auto f = [&](double r) { return std::pow(r, l + 0.5) * (1.0 - z_origin * r / (l + 1.0)); };
    y0 = f(grid.r[0]);
    y1 = f(grid.r[1]);
}

std::vector<double> outward(const std::vector<double> &g, double y0, double y1, int i_end, double delta, int *nodes) {
    const double k = delta * delta / 12.0;
    std::vector<double> y(i_end + 1, 0.0);
    y[0] = y0;
    if (i_end >= 1) y[1] = y1;
    int nn = 0;
    for (int i = 1; i < i_end; i++) {
        y[i + 1] = (2.0 * y[i] * (1.0 + 5.0 * k * g[i]) - y[i - 1] * (1.0 - k * g[i - 1])) / (1.0 - k * g[i + 1]);
        if (y[i] * y[i + 1] < 0.0) nn++;
        if (std::fabs(y[i + 1]) > 1e150)
            for (int j = 0; j <= i + 1; j++) y[j] *= 1e-150;
    }
    if (nodes) *nodes = nn;
    return y;
}

// the outward run's node count only (no stored solution)
int outward_nodes(const std::vector<double> &g, double y0, double y1, int i_end, double delta) {
    const double k = delta * delta / 12.0;
    double ym = y0, y = y1;
    int nn = 0;
    for (int i = 1; i < i_end; i++) {
        double yp = (2.0 * y * (1.0 + 5.0 * k * g[i]) - ym * (1.0 - k * g[i - 1])) / (1.0 - k * g[i + 1]);
        if (y * yp < 0.0) nn++;
        if (std::fabs(yp) > 1e150) {
            yp *= 1e-150;
            y *= 1e-150;
        }
        ym = y;
        y = yp;
    }
    return nn;
}

std::vector<double> outward_sourced(const std::vector<double> &g, const std::vector<double> &s, int i_end, double delta) {
    const double k = delta * delta / 12.0;
    std::vector<double> y(i_end + 1, 0.0);
    for (int i = 1; i < i_end; i++)
        y[i + 1] = (2.0 * y[i] * (1.0 + 5.0 * k * g[i]) - y[i - 1] * (1.0 - k * g[i - 1]) +
                    k * (s[i + 1] + 10.0 * s[i] + s[i - 1])) /
                   (1.0 - k * g[i + 1]);
    return y;
}

std::vector<double> inward(const std::vector<double> &g, int i_match, int i_start, double delta) {
    const double k = delta * delta / 12.0;
    const int m = i_start - i_match;
    std::vector<double> y(m + 1, 0.0);
    y[m] = 1e-20;
    y[m - 1] = 1e-20 * std::exp(std::sqrt(std::max(g[i_start], 0.0)) * delta);
    for (int j = m - 1; j >= 1; j--) {
        const int i = i_match + j;
        y[j - 1] = (2.0 * y[j] * (1.0 + 5.0 * k * g[i]) - y[j + 1] * (1.0 - k * g[i + 1])) / (1.0 - k * g[i - 1]);
        if (std::fabs(y[j - 1]) > 1e150)
            for (int q = j - 1; q <= m; q++) y[q] *= 1e-150;
    }
    return y;
}

int last_negative(const std::vector<double> &g, int fallback) {
    for (int i = (int)g.size() - 1; i >= 0; i--)
        if (g[i] < 0.0) return i;
    return fallback;
}

// beyond the last turning point: where the WKB decay exceeds 40 or Numerov's denominator degrades (numerov.rs:112-123)
int shooting_end(const std::vector<double> &g, double delta) {
    const int n = (int)g.size();
    const int it = last_negative(g, 0);
    const double k = delta * delta / 12.0;
    double acc = 0.0;
    for (int i = it; i < n; i++) {
        if (k * g[i] > 0.1) return i;
        acc += std::sqrt(std::max(g[i], 0.0)) * delta;
        if (acc > 40.0) return i;
    }
    return n - 1;
}

struct BoundState {
    double e = 0.0;
    std::vector<double> u;
    double r_peak = 0.0;
};

BoundState eigenfunction(const LogGrid &grid, const std::vector<double> &v, int l, double e, double z_origin) {
    const int n = grid.size();
    const auto g = g_coeff(grid, v, l, e);
    double y0, y1;
    regular_start(grid, l, z_origin, y0, y1);
    const int it = std::clamp(last_negative(g, n / 2), 2, n - 3);
    const int ie = std::min(std::max(shooting_end(g, grid.delta), it + 2), n - 1);
    const auto yo = outward(g, y0, y1, it, grid.delta, nullptr);
    const auto yi = inward(g, it, ie, grid.delta);
    const double scale = yo[it] / yi[0];
    std::vector<double> u(n, 0.0);
    for (int i = 0; i <= it; i++) u[i] = yo[i] * std::sqrt(grid.r[i]);
    for (int j = 1; j < (int)yi.size(); j++) u[it + j] = scale * yi[j] * std::sqrt(grid.r[it + j]);
    std::vector<double> u2(n);
    for (int i = 0; i < n; i++) u2[i] = u[i] * u[i];
    const double norm = std::sqrt(grid.integrate(u2));
    double umax = 0.0;
    for (auto &x : u) {
        x /= norm;
        umax = std::max(umax, std::fabs(x));
    }
    int ip = 0;
    for (int i = n - 2; i >= 1; i--)
        if (std::fabs(u[i]) >= std::fabs(u[i - 1]) && std::fabs(u[i]) >= std::fabs(u[i + 1]) &&
            std::fabs(u[i]) > R_PEAK_FLOOR * umax) {
            ip = i;
            break;
        }
    if (u[ip] < 0.0)
        for (auto &x : u) x = -x;
    BoundState b;
    b.e = e;
    b.u = std::move(u);
    b.r_peak = grid.r[ip];
    return b;
}

// the state with `nodes` radial nodes by bisection on the node count (numerov.rs:125-147)
std::optional<BoundState> bound_state(const LogGrid &grid, const std::vector<double> &v, int l, int nodes,
                                      double z_origin, std::optional<double> guess) {
    const int n = grid.size();
    double y0, y1;
    regular_start(grid, l, z_origin, y0, y1);
    auto count = [&](double e) {
        const auto g = g_coeff(grid, v, l, e);
        return outward_nodes(g, y0, y1, shooting_end(g, grid.delta), grid.delta);
    };
    double e_hi = 0.0;
    if (count(e_hi) <= nodes) return std::nullopt;
    double e_lo = INFINITY;
    for (int i = 0; i < n; i++) e_lo = std::min(e_lo, v[i] + l * (l + 1.0) / (2.0 * grid.r[i] * grid.r[i]));
    if (guess) {
        const double gs = *guess, w = 0.1 * std::max(std::fabs(gs), 0.1);
        if (count(gs - w) <= nodes) e_lo = gs - w;
        if (gs + w < 0.0 && count(gs + w) > nodes) e_hi = gs + w;
    }
    while (e_hi - e_lo > 1e-13 * std::max(std::fabs(e_lo), 1.0)) {
        const double mid = 0.5 * (e_lo + e_hi);
        if (count(mid) > nodes) e_hi = mid;
        else e_lo = mid;
    }
    return eigenfunction(grid, v, l, 0.5 * (e_lo + e_hi), z_origin);
}

// V_H(r) = Q(r)/r + integral_r^inf 4 pi r' rho dr' (atom.rs:181-187)
std::vector<double> hartree_potential(const LogGrid &grid, const std::vector<double> &rho) {
    const int n = grid.size();
    std::vector<double> a(n), b(n);
    for (int i = 0; i < n; i++) {
        a[i] = 4.0 * M_PI * grid.r[i] * grid.r[i] * rho[i];
        b[i] = 4.0 * M_PI * grid.r[i] * rho[i];
    }
    const auto q = grid.cumulative(a), s = grid.cumulative(b);
    const double s_tot = s[n - 1];
    std::vector<double> v(n);
    for (int i = 0; i < n; i++) v[i] = q[i] / grid.r[i] + (s_tot - s[i]);
    return v;
}

// ---- the spin-restricted reference atom (atom.rs) ----
struct Shell {
    int n, l;
    double occ;  // both spins
    BoundState b;
};

struct AtomResult {
    int z = 0;
    std::vector<Shell> shells;  // occupied, in aufbau order
    std::vector<double> rho;    // total density
    std::vector<double> v;      // the potential of the output density (either spin)
};

struct Unbound : Fail {
    using Fail::Fail;
};

// restricted integer aufbau over the eigenvalues (ties to the lower l), 2(2l+1) per shell (atom.rs:197-226)
std::vector<Shell> aufbau(const LogGrid &grid, const std::vector<double> &v, int z, double electrons,
                          const std::vector<Shell> &prev) {
    auto guess = [&](int n, int l) -> std::optional<double> {
        for (const auto &s : prev)
            if (s.n == n && s.l == l && std::isfinite(s.b.e)) return s.b.e;
        return std::nullopt;
    };
    std::optional<BoundState> next[4];
    int next_n[4];
    for (int l = 0; l < 4; l++) {
        next[l] = bound_state(grid, v, l, 0, z, guess(l + 1, l));
        next_n[l] = l + 1;
    }
    std::vector<Shell> out;
    double left = electrons;
    while (left > 1e-12) {
        int pick = -1;
        for (int l = 0; l < 4; l++)
            if (next[l] && (pick < 0 || next[l]->e < next[pick]->e)) pick = l;
        if (pick < 0) throw Unbound("Z = " + std::to_string(z) + ": no bound shell left for the remaining electrons");
        const int n = next_n[pick];
        const double occ = std::min(left, 2.0 * (2 * pick + 1));
        left -= occ;
        out.push_back({n, pick, occ, std::move(*next[pick])});
        next_n[pick] = n + 1;
        next[pick] = bound_state(grid, v, pick, n - pick, z, guess(n + 1, pick));
    }
    return out;
}

std::vector<double> shell_density(const LogGrid &grid, const std::vector<Shell> &shells) {
    std::vector<double> rho(grid.size(), 0.0);
    for (int i = 0; i < grid.size(); i++) {
        double s = 0.0;
        for (const auto &sh : shells) s += sh.occ * sh.b.u[i] * sh.b.u[i];
        rho[i] = s / (4.0 * M_PI * grid.r[i] * grid.r[i]);
    }
    return rho;
}

std::vector<double> atom_potential(const LogGrid &grid, int z, const std::vector<double> &rho) {
    const auto vh = hartree_potential(grid, rho);
    std::vector<double> v(grid.size());
    for (int i = 0; i < grid.size(); i++) {
        double vu;
        pw92(0.5 * rho[i], 0.5 * rho[i], &vu, nullptr);
        v[i] = -z / grid.r[i] + vh[i] + vu;
    }
    return v;
}

// Pulay/DIIS on the density, inner product integral a b r^2 dr (atom.rs:476-523); in the restricted mode the stacked
// [rho_up, rho_dn] holds two equal halves, so the weights and the mixing act on the total density alike
struct Diis {
    int m;
    std::vector<std::pair<std::vector<double>, std::vector<double>>> hist;
    explicit Diis(int m_) : m(std::max(m_, 1)) {}
    void clear() { hist.clear(); }
    std::vector<double> mix(const std::vector<double> &x, const std::vector<double> &y, const std::vector<double> &w,
                            double alpha) {
        const int n = (int)x.size();
        std::vector<double> r(n);
        for (int i = 0; i < n; i++) r[i] = y[i] - x[i];
        hist.emplace_back(x, r);
        if ((int)hist.size() > m) hist.erase(hist.begin());
        for (;;) {
            const int k = (int)hist.size();
            std::vector<double> b((k + 1) * (k + 1), 0.0), rhs(k + 1, 0.0);
            for (int i = 0; i < k; i++)
                for (int j = 0; j < k; j++) {
                    double s = 0.0;
                    for (int p = 0; p < n; p++) s += hist[i].second[p] * hist[j].second[p] * w[p];
                    b[i * (k + 1) + j] = 2.0 * s;  // the two equal spin halves
                }
            double scale = 0.0;
            for (int i = 0; i < k; i++) scale = std::max(scale, b[i * (k + 1) + i]);
            scale = std::max(scale, 2.2250738585072014e-308);
            for (int i = 0; i < k; i++) {
                for (int j = 0; j < k; j++) b[i * (k + 1) + j] /= scale;
                b[i * (k + 1) + i] += 1e-12;
                b[i * (k + 1) + k] = 1.0;
                b[k * (k + 1) + i] = 1.0;
            }
            rhs[k] = 1.0;
            std::vector<double> c;
            if (lu_solve(b, rhs, k + 1, c)) {
                std::vector<double> out(n, 0.0);
                for (int i = 0; i < k; i++)
                    for (int p = 0; p < n; p++) out[p] += c[i] * (hist[i].first[p] + alpha * hist[i].second[p]);
                for (auto &v : out) v = std::max(v, 0.0);
                return out;
            }
            if (k > 1) hist.erase(hist.begin());
            else throw Fail("DIIS: a one-entry history failed to solve");
        }
    }
    // Gaussian elimination with partial pivoting; false when singular or not finite
    static bool lu_solve(std::vector<double> a, std::vector<double> b, int n, std::vector<double> &x) {
        for (int c = 0; c < n; c++) {
            int p = c;
            for (int r = c + 1; r < n; r++)
                if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
            if (a[p * n + c] == 0.0) return false;
            if (p != c) {
                for (int j = 0; j < n; j++) std::swap(a[c * n + j], a[p * n + j]);
                std::swap(b[c], b[p]);
            }
            for (int r = c + 1; r < n; r++) {
                const double f = a[r * n + c] / a[c * n + c];
                for (int j = c; j < n; j++) a[r * n + j] -= f * a[c * n + j];
                b[r] -= f * b[c];
            }
        }
        x.assign(n, 0.0);
        for (int r = n - 1; r >= 0; r--) {
            double s = b[r];
            for (int j = r + 1; j < n; j++) s -= a[r * n + j] * x[j];
            x[r] = s / a[r * n + r];
        }
        for (double v : x)
            if (!std::isfinite(v)) return false;
        return true;
    }
};

// one SCF pass from rho = 0 (atom.rs:331-384): DIIS with the backtracking linear-mixing fallback, or plain linear
// mixing with beta halved for good at every failure (the confirmation pass)
AtomResult scf_pass(int z, bool use_diis, bool &exhausted) {
    const LogGrid &grid = universal_grid();
    const int n = grid.size();
    std::vector<double> w(n);
    for (int i = 0; i < n; i++) w[i] = grid.r[i] * grid.r[i] * grid.r[i] * grid.delta;
    std::vector<double> rho_in(n, 0.0);
    std::vector<Shell> prev;
    Diis diis(use_diis ? DIIS_HISTORY : 1);
    bool have_good = false;
    std::vector<double> g_in, g_out;
    std::vector<Shell> g_shells;
    int halvings = 0;
    exhausted = false;
    double delta = INFINITY;
    for (int it = 1; it <= SCF_MAX_ITER; it++) {
        const auto v = atom_potential(grid, z, rho_in);
        std::vector<Shell> shells;
        bool ok = true;
        try {
            shells = aufbau(grid, v, z, z, prev);
        } catch (const Unbound &) {
            ok = false;
        }
        if (ok) {
            if (use_diis) halvings = 0;
            const auto rho_out = shell_density(grid, shells);
            delta = 0.0;
            for (int i = 0; i < n; i++) delta += std::fabs(rho_out[i] - rho_in[i]) * w[i];
            if (delta < SCF_TOL) {
                AtomResult a;
                a.z = z;
                a.shells = std::move(shells);
                a.rho = rho_out;
                a.v = atom_potential(grid, z, rho_out);
                return a;
            }
            have_good = true;
            g_in = rho_in;
            g_out = rho_out;
            g_shells = shells;
            prev = shells;
            // the halves of the stacked spin densities are mixed as one total density
            std::vector<double> half_in(n), half_out(n);
            for (int i = 0; i < n; i++) {
                half_in[i] = 0.5 * rho_in[i];
                half_out[i] = 0.5 * rho_out[i];
            }
            const auto mixed = diis.mix(half_in, half_out, w, MIXING / (double)(1ull << halvings));
            for (int i = 0; i < n; i++) rho_in[i] = 2.0 * mixed[i];
        } else {
            if (!have_good) throw Fail("Z = " + std::to_string(z) + ": the bare nucleus binds no shell");
            if (halvings == BACKTRACK_MAX) {
                exhausted = true;
                throw Unbound("Z = " + std::to_string(z) + ": an occupied shell is unbound after backtracking");
            }
            halvings++;
            const double beta = MIXING / (double)(1ull << halvings);
            diis.clear();
            for (int i = 0; i < n; i++) rho_in[i] = g_in[i] + beta * (g_out[i] - g_in[i]);
            prev = g_shells;
        }
    }
    throw Fail("Z = " + std::to_string(z) + ": the radial SCF did not converge (delta rho " + std::to_string(delta) + ")");
}

AtomResult ground_state_restricted(int z) {
    bool exhausted = false;
    try {
        return scf_pass(z, true, exhausted);
    } catch (const Unbound &) {
        if (!exhausted) throw;
    }
    return scf_pass(z, false, exhausted);
}

// ---- the partition (partition.rs) ----
struct Channel {
    int l, n;
    double eps, occ, r_peak;
    std::vector<double> u;
    bool occupied;
};

struct Partition {
    int z = 0;
    std::vector<const Shell *> core, valence;
    int n_core = 0;
    double z_v = 0.0;
    int l_max = 0;
    std::vector<Channel> channels;
};

Partition partition(const AtomResult &a) {
    const LogGrid &grid = universal_grid();
    Partition p;
    p.z = a.z;
    int n_max = 0;
    for (const auto &s : a.shells) n_max = std::max(n_max, s.n);
    double eps_v_min = INFINITY;
    for (const auto &s : a.shells)
        if (s.n == n_max) {
            p.valence.push_back(&s);
            eps_v_min = std::min(eps_v_min, s.b.e);
        }
    for (int n = n_max - 1; n >= 1; n--) {
        std::vector<const Shell *> sh;
        bool semicore = false;
        for (const auto &s : a.shells)
            if (s.n == n) {
                sh.push_back(&s);
                if (s.b.e > eps_v_min - DELTA_CORE) semicore = true;
            }
        if (!semicore) break;
        p.valence.insert(p.valence.end(), sh.begin(), sh.end());
    }
    double n_core = 0.0;
    int lv_max = 0;
    for (const auto &s : a.shells) {
        const bool val = std::find(p.valence.begin(), p.valence.end(), &s) != p.valence.end();
        if (val) {
            p.z_v += s.occ;
            lv_max = std::max(lv_max, s.l);
        } else {
            p.core.push_back(&s);
            n_core += s.occ;
        }
    }
    p.n_core = (int)std::lround(n_core);
    p.l_max = lv_max + L_MAX_ABOVE_VALENCE;
    for (int l = 0; l <= p.l_max; l++) {
        std::vector<const Shell *> of_l;
        for (const Shell *s : p.valence)
            if (s->l == l) of_l.push_back(s);
        if (of_l.empty()) {
            // the lowest unoccupied state of l (atom.rs:525-530)
            int filled = 0;
            for (const auto &s : a.shells)
                if (s.l == l) filled = std::max(filled, s.n);
            const int n = filled ? filled + 1 : l + 1;
            auto b = bound_state(grid, a.v, l, n - l - 1, a.z, std::nullopt);
            if (b) p.channels.push_back({l, n, b->e, 0.0, b->r_peak, b->u, false});
        } else if (of_l.size() == 1) {
            const Shell *s = of_l[0];
            p.channels.push_back({l, s->n, s->b.e, s->occ, s->b.r_peak, s->b.u, true});
        } else {
            throw Fail("Z = " + std::to_string(a.z) + ": two valence states of l = " + std::to_string(l) +
                       " -- a one-projector TM/KB potential cannot carry both");
        }
    }
    return p;
}

std::vector<double> density_of(const std::vector<const Shell *> &shells) {
    const LogGrid &grid = universal_grid();
    std::vector<double> rho(grid.size(), 0.0);
    for (int i = 0; i < grid.size(); i++) {
        double s = 0.0;
        for (const Shell *sh : shells) s += sh->occ * sh->b.u[i] * sh->b.u[i];
        rho[i] = s / (4.0 * M_PI * grid.r[i] * grid.r[i]);
    }
    return rho;
}

// ---- the LFC partial core and the core radius (nlcc.rs) ----
struct PartialCore {
    std::vector<double> rho_pc;
    double r_nlc = 0.0;  // 0: no core
    double r_core = 0.0, n_core = 0.0;
};

double core_radius(const LogGrid &grid, const std::vector<double> &rho_c) {
    const auto c = grid.cumulative(shell_weight(grid, rho_c));
    const double n_c = c[grid.size() - 1];
    if (n_c <= 0.0) return 0.0;
    for (int i = 0; i < grid.size(); i++)
        if (n_c - c[i] <= EPS_CORE * n_c) return grid.r[i];
    return grid.r[grid.size() - 1];
}

PartialCore partial_core(const LogGrid &grid, const std::vector<double> &rho_c, const std::vector<double> &rho_v) {
    const int n = grid.size();
    PartialCore pc;
    pc.n_core = grid.integrate(shell_weight(grid, rho_c));
    pc.r_core = core_radius(grid, rho_c);
    pc.rho_pc.assign(n, 0.0);
    if (pc.n_core <= 0.0) return pc;
    int i = -1;
    for (int j = n - 1; j >= 0; j--)
        if (rho_c[j] > 0.0 && rho_c[j] >= NLCC_CROSSOVER * rho_v[j]) {
            i = j;
            break;
        }
    if (i < 0) throw Fail("no LFC partial core: the core never dominates the valence density");
    const double r = grid.r[i];
    const double kappa = 1.0 + r * grid.derivative(rho_c)[i] / rho_c[i];
    if (kappa >= 1.0) throw Fail("no LFC partial core: the core density is not falling at r_nlc");
    double lo = 1e-12, hi = M_PI - 1e-12;
    for (int k = 0; k < 200; k++) {
        const double m = 0.5 * (lo + hi);
        if (m / std::tan(m) > kappa) lo = m;
        else hi = m;
    }
    const double x = 0.5 * (lo + hi), b = x / r, a = rho_c[i] * r / std::sin(x);
    for (int j = 0; j < n; j++) pc.rho_pc[j] = j < i ? a * std::sin(b * grid.r[j]) / grid.r[j] : rho_c[j];
    pc.r_nlc = r;
    return pc;
}

// ---- Troullier-Martins channels and the unscreening (tm.rs) ----
struct TmChannel {
    int l;
    double eps, r_c;
    int i_c;
    double c[7];
    std::vector<double> u, v_scr;
};

double dpow(int m, int k, double r) {
    double f = 1.0;
    for (int q = 0; q < k; q++) f *= (double)(m - q);
    return f == 0.0 ? 0.0 : f * std::pow(r, m - k);
}

double poly_derivative(const double *c, int k, double r) {
    double s = 0.0;
    for (int j = 0; j < 7; j++) s += c[j] * dpow(2 * j, k, r);
    return s;
}

// p = ln(u / r^{l+1}) and its first four derivatives at r_c from the AE solution (tm.rs:119-131)
void tm_targets(const LogGrid &grid, int l, double eps, const std::vector<double> &u_ae, const std::vector<double> &v_ae,
                int ic, double t[5]) {
    const double lp = l + 1.0, rc = grid.r[ic];
    const auto du = grid.derivative(u_ae);
    const auto dv = grid.derivative(v_ae);
    const auto d2v = grid.derivative(dv);
    const double p0 = std::log(u_ae[ic] / std::pow(rc, lp));
    const double p1 = du[ic] / u_ae[ic] - lp / rc;
    const double p2 = 2.0 * (v_ae[ic] - eps) - 2.0 * lp * p1 / rc - p1 * p1;
    const double p3 = 2.0 * dv[ic] + 2.0 * lp * p1 / (rc * rc) - 2.0 * lp * p2 / rc - 2.0 * p1 * p2;
    const double p4 = 2.0 * d2v[ic] - 4.0 * lp * p1 / std::pow(rc, 3) + 4.0 * lp * p2 / (rc * rc) - 2.0 * lp * p3 / rc -
                      2.0 * p2 * p2 - 2.0 * p1 * p3;
    t[0] = p0, t[1] = p1, t[2] = p2, t[3] = p3, t[4] = p4;
}

// c4 = -c2^2 / (2l + 5), the 5x5 solve for (c0, c6, c8, c10, c12), Newton on c2 from 0 for norm conservation
// (tm.rs:164-200)
void tm_solve(const LogGrid &grid, int l, const std::vector<double> &u_ae, int ic, const double target[5], double out[7]) {
    const int n = grid.size();
    const double rc = grid.r[ic];
    std::vector<double> sq(n);
    for (int i = 0; i < n; i++) sq[i] = u_ae[i] * u_ae[i];
    const double n_ae = grid.partial(sq, ic);
    const int ms[5] = {0, 6, 8, 10, 12};
    std::vector<double> m(25);
    for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++) m[k * 5 + j] = dpow(ms[j], k, rc);
    auto coeffs = [&](double c2, double c[7]) {
        const double c4 = -c2 * c2 / (2.0 * l + 5.0);
        std::vector<double> b(5), x;
        for (int k = 0; k < 5; k++) b[k] = target[k] - c2 * dpow(2, k, rc) - c4 * dpow(4, k, rc);
        if (!Diis::lu_solve(m, b, 5, x)) return false;
        c[0] = x[0], c[1] = c2, c[2] = c4, c[3] = x[1], c[4] = x[2], c[5] = x[3], c[6] = x[4];
        for (int j = 0; j < 7; j++)
            if (!std::isfinite(c[j])) return false;
        return true;
    };
    auto u_of = [&](const double *c, double r) { return std::pow(r, l + 1.0) * std::exp(poly_derivative(c, 0, r)); };
    auto resid = [&](double c2, double &f) {
        double c[7];
        if (!coeffs(c2, c)) return false;
        std::vector<double> g(n);
        for (int i = 0; i < n; i++) {
            const double u = i <= ic ? u_of(c, grid.r[i]) : 0.0;
            g[i] = i <= ic ? u * u : sq[i];
        }
        f = std::log(grid.partial(g, ic)) - std::log(n_ae);
        return std::isfinite(f);
    };
    const std::string where = "l = " + std::to_string(l) + ", r_c = " + std::to_string(rc);
    double c2 = 0.0, f;
    if (!resid(c2, f)) throw Fail(where + ": TM matching system singular or overflowing");
    for (int it = 0; it < 100; it++) {
        if (std::fabs(f) < 1e-14) break;
        const double h = 1e-6 * std::max(std::fabs(c2), 1.0);
        double fp, fm;
        if (!resid(c2 + h, fp) || !resid(c2 - h, fm)) throw Fail(where + ": TM matching system singular or overflowing");
        const double df = (fp - fm) / (2.0 * h);
        if (!std::isfinite(df) || df == 0.0) throw Fail(where + ": TM norm equation has a flat derivative");
        c2 += std::clamp(-f / df, -1.0, 1.0);
        if (!resid(c2, f)) throw Fail(where + ": TM matching system singular or overflowing");
    }
    if (!(std::fabs(f) < 1e-12)) throw Fail(where + ": norm conservation has no TM solution");
    if (!coeffs(c2, out)) throw Fail(where + ": TM matching system singular or overflowing");
}

TmChannel tm_channel(const LogGrid &grid, int l, double eps, const std::vector<double> &u_ae,
                     const std::vector<double> &v_ae, double r_c) {
    const int n = grid.size();
    const int ic = grid.nearest(r_c);
    const double rc = grid.r[ic];
    if (ic < 2 || ic + 3 > n) throw Fail("l = " + std::to_string(l) + ": r_c outside the grid's interior");
    bool node = u_ae[ic] <= 0.0;
    for (int i = ic; i < n - 1 && !node; i++)
        if (u_ae[i] * u_ae[i + 1] < 0.0) node = true;
    if (node) throw Fail("l = " + std::to_string(l) + ": r_c is inside the outermost node of u_AE");
    double target[5];
    tm_targets(grid, l, eps, u_ae, v_ae, ic, target);
    TmChannel t;
    t.l = l, t.eps = eps, t.r_c = rc, t.i_c = ic;
    tm_solve(grid, l, u_ae, ic, target, t.c);
    t.u = u_ae;
    t.v_scr = v_ae;
    for (int i = 0; i <= ic; i++) {
        const double r = grid.r[i], r2 = r * r;
        t.u[i] = std::pow(r, l + 1.0) * std::exp(poly_derivative(t.c, 0, r));
        const double p1_over_r =
            2.0 * t.c[1] +
            r2 * (4.0 * t.c[2] + r2 * (6.0 * t.c[3] + r2 * (8.0 * t.c[4] + r2 * (10.0 * t.c[5] + r2 * 12.0 * t.c[6]))));
        const double p1 = r * p1_over_r, p2 = poly_derivative(t.c, 2, r);
        t.v_scr[i] = eps + (l + 1.0) * p1_over_r + 0.5 * (p1 * p1 + p2);
    }
    return t;
}

std::vector<TmChannel> tm_channels(const AtomResult &ref, const Partition &part, double f) {
    const LogGrid &g = universal_grid();
    double cap = 0.0;
    for (const auto &c : part.channels)
        if (c.occupied) cap = std::max(cap, g.r[g.nearest(f * c.r_peak)]);
    std::vector<TmChannel> out;
    for (const auto &c : part.channels) {
        const double r_c = c.occupied ? f * c.r_peak : std::min(f * c.r_peak, cap);
        out.push_back(tm_channel(g, c.l, c.eps, c.u, ref.v, r_c));
    }
    return out;
}

struct Unscreened {
    std::vector<std::vector<double>> v_ion;
    std::vector<double> rho_v;
    double r_cmp = 0.0;
};

Unscreened unscreen(const LogGrid &grid, double z_v, const std::vector<TmChannel> &ch, const std::vector<double> &occ,
                    const std::vector<double> &rho_pc) {
    const int n = grid.size();
    Unscreened un;
    un.rho_v.assign(n, 0.0);
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (size_t k = 0; k < ch.size(); k++) s += occ[k] * ch[k].u[i] * ch[k].u[i];
        un.rho_v[i] = s / (4.0 * M_PI * grid.r[i] * grid.r[i]);
    }
    const auto vh = hartree_potential(grid, un.rho_v);
    std::vector<double> vxc(n);
    for (int i = 0; i < n; i++) {
        const double half = 0.5 * (un.rho_v[i] + rho_pc[i]);
        pw92(half, half, &vxc[i], nullptr);
    }
    for (const auto &c : ch) {
        std::vector<double> v(n);
        for (int i = 0; i < n; i++) v[i] = c.v_scr[i] - vh[i] - vxc[i];
        un.v_ion.push_back(std::move(v));
    }
    int bad = -1;
    for (int i = n - 1; i >= 0 && bad < 0; i--)
        for (const auto &v : un.v_ion)
            if (std::fabs(v[i] + z_v / grid.r[i]) >= V_ION_TAIL) {
                bad = i;
                break;
            }
    un.r_cmp = bad < 0 ? grid.r[0] : grid.r[std::min(bad + 1, n - 1)];
    return un;
}

// ---- the Kleinman-Bylander form (kb.rs) ----
struct Projector {
    int l;
    double e_kb, gss;
    std::vector<double> beta, dv, u;
    int i_beta;
};

struct KbPotential {
    int local;  // LIRA_PSEUDO_LOCAL_*
    int local_l = -1;
    std::vector<double> v_loc;
    std::vector<Projector> projectors;
    double r_cmp = 0.0;
};

// tau_c = V_H[rho_c] - N_c / r, the full core's Hartree tail (kb.rs:201-204)
std::vector<double> core_tail(const LogGrid &grid, const std::vector<double> &rho_c, double n_core) {
    auto v = hartree_potential(grid, rho_c);
    for (int i = 0; i < grid.size(); i++) v[i] -= n_core / grid.r[i];
    return v;
}

// -Z_v / r + tau_c from r_loc outwards, the even quartic matched C^2 inside (kb.rs:208-224)
std::vector<double> polynomial_local(const LogGrid &grid, double z_v, double r_loc, const std::vector<double> &tau) {
    const int i = grid.nearest(r_loc);
    const double r = grid.r[i];
    const auto d = grid.derivative(tau);
    const auto d2 = grid.derivative(d);
    const double t0 = -z_v / r + tau[i], t1 = z_v / (r * r) + d[i], t2 = -2.0 * z_v / (r * r * r) + d2[i];
    const double a4 = (t2 - t1 / r) / (8.0 * r * r);
    const double a2 = (t1 - 4.0 * a4 * r * r * r) / (2.0 * r);
    const double a0 = t0 - a2 * r * r - a4 * r * r * r * r;
    std::vector<double> v(grid.size());
    for (int k = 0; k < grid.size(); k++) {
        const double x = grid.r[k];
        v[k] = k >= i ? -z_v / x + tau[k] : a0 + x * x * (a2 + x * x * a4);
    }
    return v;
}

KbPotential kb_form(const LogGrid &grid, double z_v, const std::vector<TmChannel> &ch, const Unscreened &un, int local,
                    const std::vector<double> &tau) {
    const int n = grid.size();
    KbPotential kb;
    kb.local = local;
    kb.r_cmp = un.r_cmp;
    const int top = (int)ch.size() - 1;
    double r_loc = INFINITY;
    for (const auto &c : ch) r_loc = std::min(r_loc, c.r_c);
    if (local == LIRA_PSEUDO_LOCAL_HIGHEST) {
        kb.v_loc = un.v_ion[top];
        kb.local_l = ch[top].l;
    } else {
        kb.v_loc = polynomial_local(grid, z_v, r_loc, tau);
    }
    // the derived support r_beta: the first radius from which every V_l^ion is -Z_v/r + tau_c to v_ion_tail/100 for
    // good (kb.rs:235-241)
    const double tol = V_ION_TAIL / 100.0;
    int i_beta = 0;
    for (int i = n - 1; i >= 0; i--) {
        bool hit = false;
        for (const auto &v : un.v_ion)
            if (std::fabs(v[i] + z_v / grid.r[i] - tau[i]) >= tol) hit = true;
        if (hit) {
            i_beta = i + 1;
            break;
        }
    }
    i_beta = std::min(i_beta, n - 1);
    for (size_t k = 0; k < ch.size(); k++) {
        if (ch[k].l == kb.local_l) continue;
        Projector p;
        p.l = ch[k].l;
        p.u = ch[k].u;
        p.dv.resize(n);
        p.beta.resize(n);
        double worst = 0.0;
        for (int i = 0; i < n; i++) {
            p.dv[i] = un.v_ion[k][i] - kb.v_loc[i];
            if (i >= i_beta) {
                worst = std::max(worst, std::fabs(p.dv[i]));
                p.dv[i] = 0.0;
            }
            p.beta[i] = p.dv[i] * p.u[i];
        }
        if (!(worst < tol)) throw Fail("the local potential does not carry the ionic tail beyond r_beta (l = " +
                                       std::to_string(p.l) + ")");
        std::vector<double> ub(n), bb(n);
        for (int i = 0; i < n; i++) {
            ub[i] = p.u[i] * p.beta[i];
            bb[i] = p.beta[i] * p.beta[i];
        }
        const double denom = grid.integrate(ub), num = grid.integrate(bb);
        p.e_kb = 1.0 / denom;
        p.gss = num / denom;
        p.i_beta = i_beta;
        kb.projectors.push_back(std::move(p));
    }
    return kb;
}

// G1 (Gonze-Stumpf-Scheffler, kb.rs:611-634): ghost-free iff (E_KB < 0 and eps < eps0_loc) or (E_KB > 0 and
// eps0_loc < eps < eps1_loc), an unbound local level counting as the continuum edge 0
void ghost_g1(const LogGrid &grid, const KbPotential &kb, const std::vector<double> &screen, const Partition &part) {
    std::vector<double> v_ref(grid.size());
    for (int i = 0; i < grid.size(); i++) v_ref[i] = kb.v_loc[i] + screen[i];
    for (const auto &p : kb.projectors) {
        double eps = 0.0;
        for (const auto &c : part.channels)
            if (c.l == p.l) eps = c.eps;
        const auto b0 = bound_state(grid, v_ref, p.l, 0, 0.0, std::nullopt);
        const auto b1 = bound_state(grid, v_ref, p.l, 1, 0.0, std::nullopt);
        const double e0 = b0 ? b0->e : 0.0, e1 = b1 ? b1->e : 0.0;
        const bool ok = p.e_kb < 0.0 ? eps < e0 : (e0 < eps && eps < e1);
        if (!ok) throw Fail("G1: ghost in l = " + std::to_string(p.l));
    }
}

// the regular KB solution at energy e out to i_end, pole-free: u = (1 - E_KB<b|u_p>) u_h + E_KB<b|u_h> u_p
// (kb.rs:443-462)
std::vector<double> kb_regular(const LogGrid &grid, const std::vector<double> &v, const Projector *proj, int l, double e,
                               int i_end) {
    const int n = grid.size();
    const auto g = g_coeff(grid, v, l, e);
    double y0, y1;
    regular_start(grid, l, 0.0, y0, y1);
    int reach = i_end;
    if (proj) {
        double m = 0.0;
        for (double b : proj->beta) m = std::max(m, std::fabs(b));
        int sup = 0;
        for (int i = n - 1; i >= 0; i--)
            if (std::fabs(proj->beta[i]) > 1e-12 * m) {
                sup = i;
                break;
            }
        reach = std::max(sup + 3, i_end);
    }
    const auto yh = outward(g, y0, y1, reach, grid.delta, nullptr);
    std::vector<double> uh(reach + 1);
    for (int i = 0; i <= reach; i++) uh[i] = yh[i] * std::sqrt(grid.r[i]);
    if (!proj) return std::vector<double>(uh.begin(), uh.begin() + i_end + 1);
    std::vector<double> s(reach + 1);
    for (int i = 0; i <= reach; i++) s[i] = 2.0 * std::pow(grid.r[i], 1.5) * proj->beta[i];
    const auto yp = outward_sourced(g, s, reach, grid.delta);
    std::vector<double> up(reach + 1);
    for (int i = 0; i <= reach; i++) up[i] = yp[i] * std::sqrt(grid.r[i]);
    auto dot = [&](const std::vector<double> &u) {
        std::vector<double> f(n, 0.0);
        for (int i = 0; i <= reach; i++) f[i] = proj->beta[i] * u[i];
        return grid.partial(f, reach);
    };
    const double a = 1.0 - proj->e_kb * dot(up), c = proj->e_kb * dot(uh);
    std::vector<double> u(i_end + 1);
    for (int i = 0; i <= i_end; i++) u[i] = a * uh[i] + c * up[i];
    return u;
}

// T5 (gates.rs:182-217): delta E_equiv(E) = 1/2 W(u_AE, u_PP)(r_d), each regular solution normalised inside r_d,
// r_d = 1.2 max r_c; the maximum over eps_l +- 0.5 Ha on a 10 mHa mesh with both end points exact
double t5_channel(const LogGrid &g, const AtomResult &ref, const std::vector<TmChannel> &ch, const KbPotential &kb,
                  const std::vector<double> &screen, int l, double eps, const std::vector<double> &extra) {
    double rmax = 0.0;
    for (const auto &c : ch) rmax = std::max(rmax, c.r_c);
    const int i_d = g.nearest(T5_RD_FACTOR * rmax);
    std::vector<double> v_pp(g.size());
    for (int i = 0; i < g.size(); i++) v_pp[i] = kb.v_loc[i] + screen[i];
    const Projector *proj = nullptr;
    for (const auto &p : kb.projectors)
        if (p.l == l) proj = &p;
    auto at_rd = [&](const std::vector<double> &u, double &val, double &der) {
        std::vector<double> sq(g.size(), 0.0);
        for (size_t i = 0; i < u.size(); i++) sq[i] = u[i] * u[i];
        const double norm = std::sqrt(g.partial(sq, i_d));
        const double du = (u[i_d - 2] - 8.0 * u[i_d - 1] + 8.0 * u[i_d + 1] - u[i_d + 2]) / (12.0 * g.delta * g.r[i_d]);
        val = u[i_d] / norm;
        der = du / norm;
    };
    auto de_equiv = [&](double e) {
        const auto gc = g_coeff(g, ref.v, l, e);
        double y0, y1;
        regular_start(g, l, ref.z, y0, y1);
        auto yae = outward(gc, y0, y1, i_d + 2, g.delta, nullptr);
        for (size_t i = 0; i < yae.size(); i++) yae[i] *= std::sqrt(g.r[i]);
        double ae, dae, pp, dpp;
        at_rd(yae, ae, dae);
        at_rd(kb_regular(g, v_pp, proj, l, e, i_d + 2), pp, dpp);
        return 0.5 * (ae * dpp - dae * pp);
    };
    const double lo = eps - T5_WINDOW, hi = eps + T5_WINDOW;
    const int n = (int)std::ceil((hi - lo) / T5_MESH);
    double worst = 0.0;
    for (int k = 0; k <= n; k++) {
        const double e = k == n ? hi : std::min(lo + k * T5_MESH, hi);
        worst = std::max(worst, std::fabs(de_equiv(e)));
    }
    for (double e : extra) worst = std::max(worst, std::fabs(de_equiv(e)));
    return worst;
}


// ---- the general all-electron atom: spin-polarised and fixed occupations, the total energy (atom.rs) ----
struct ShellOcc {
    int n, l;
    double occ[2];
};

struct FullOrbital {
    int n, l;
    double eps[2] = {NAN, NAN};
    double occ[2] = {0.0, 0.0};
    std::vector<double> u[2];
};

struct FullAtom {
    int z = 0;
    bool restricted = true;
    int multiplicity = 1;
    std::vector<FullOrbital> orbitals;
    std::vector<double> rho[2], v[2];
    double total = 0.0;
};

struct AtomOpts {
    bool restricted = true;
    bool aufbau = true;
    int multiplicity = 1;
    std::vector<ShellOcc> fixed;
};

// one spin channel's aufbau, cap(l) = factor (2l + 1) (atom.rs:197-226)
std::vector<Shell> aufbau_cap(const LogGrid &grid, const std::vector<double> &v, int z, double electrons, double factor,
                              const std::vector<FullOrbital> &prev, int spin) {
    auto guess = [&](int n, int l) -> std::optional<double> {
        for (const auto &o : prev)
            if (o.n == n && o.l == l && std::isfinite(o.eps[spin])) return o.eps[spin];
        return std::nullopt;
    };
    std::optional<BoundState> next[4];
    int next_n[4];
    for (int l = 0; l < 4; l++) {
        next[l] = bound_state(grid, v, l, 0, z, guess(l + 1, l));
        next_n[l] = l + 1;
    }
    std::vector<Shell> out;
    double left = electrons;
    while (left > 1e-12) {
        int pick = -1;
        for (int l = 0; l < 4; l++)
            if (next[l] && (pick < 0 || next[l]->e < next[pick]->e)) pick = l;
        if (pick < 0) throw Unbound("Z = " + std::to_string(z) + ": no bound shell left for the remaining electrons");
        const int n = next_n[pick];
        const double occ = std::min(left, factor * (2 * pick + 1));
        left -= occ;
        out.push_back({n, pick, occ, std::move(*next[pick])});
        next_n[pick] = n + 1;
        next[pick] = bound_state(grid, v, pick, n - pick, z, guess(n + 1, pick));
    }
    return out;
}

std::vector<FullOrbital> occupy_full(const LogGrid &grid, int z, const std::vector<double> *v, const AtomOpts &o,
                                     const std::vector<FullOrbital> &prev, std::vector<double> *rho) {
    std::vector<Shell> ch[2];
    if (o.aufbau && o.restricted) {
        ch[0] = aufbau_cap(grid, v[0], z, z, 2.0, prev, 0);
        for (auto &s : ch[0]) s.occ *= 0.5;
        ch[1] = ch[0];
    } else if (o.aufbau) {
        const int m = o.multiplicity;
        if (m < 1 || m - 1 > z || (z - (m - 1)) % 2 != 0) throw Fail("multiplicity impossible for the electron count");
        const int up = (z + m - 1) / 2, dn = (z + 1 - m) / 2;
        ch[0] = aufbau_cap(grid, v[0], z, up, 1.0, prev, 0);
        ch[1] = aufbau_cap(grid, v[1], z, dn, 1.0, prev, 1);
    } else {
        for (int sp = 0; sp < 2; sp++)
            for (const auto &sh : o.fixed) {
                if (sh.occ[sp] <= 0.0) continue;
                const int src = o.restricted ? 0 : sp;
                std::optional<double> g;
                for (const auto &q : prev)
                    if (q.n == sh.n && q.l == sh.l && std::isfinite(q.eps[src])) g = q.eps[src];
                auto b = bound_state(grid, v[src], sh.l, sh.n - sh.l - 1, z, g);
                if (!b) throw Unbound("Z = " + std::to_string(z) + ": a fixed shell is unbound");
                ch[sp].push_back({sh.n, sh.l, sh.occ[sp], std::move(*b)});
            }
    }
    for (int sp = 0; sp < 2; sp++) rho[sp] = shell_density(grid, ch[sp]);
    std::vector<FullOrbital> out;
    for (int sp = 0; sp < 2; sp++)
        for (auto &sh : ch[sp]) {
            FullOrbital *o2 = nullptr;
            for (auto &q : out)
                if (q.n == sh.n && q.l == sh.l) o2 = &q;
            if (!o2) {
                out.push_back(FullOrbital{sh.n, sh.l});
                o2 = &out.back();
            }
            o2->eps[sp] = sh.b.e;
            o2->occ[sp] = sh.occ;
            o2->u[sp] = std::move(sh.b.u);
        }
    std::sort(out.begin(), out.end(), [](const FullOrbital &a, const FullOrbital &b) {
        return a.n != b.n ? a.n < b.n : a.l < b.l;
    });
    return out;
}

void full_potentials(const LogGrid &grid, int z, const std::vector<double> *rho, std::vector<double> *v) {
    const int n = grid.size();
    std::vector<double> total(n);
    for (int i = 0; i < n; i++) total[i] = rho[0][i] + rho[1][i];
    const auto vh = hartree_potential(grid, total);
    v[0].resize(n);
    v[1].resize(n);
    for (int i = 0; i < n; i++) {
        double vu, vd;
        pw92(rho[0][i], rho[1][i], &vu, &vd);
        const double c = -z / grid.r[i] + vh[i];
        v[0][i] = c + vu;
        v[1][i] = c + vd;
    }
}

double shell_integral(const LogGrid &g, const std::vector<double> &f) {
    std::vector<double> w(g.size());
    for (int i = 0; i < g.size(); i++) w[i] = 4.0 * M_PI * g.r[i] * g.r[i] * f[i];
    return g.integrate(w);
}

// the energy from the output density and orbitals: T_s = sum f eps - sum_s int rho_s V_s^in (atom.rs:387-417)
double full_energy(const LogGrid &grid, int z, const std::vector<FullOrbital> &orb, const std::vector<double> *rho,
                   const std::vector<double> *v_in) {
    const int n = grid.size();
    Neumaier band;
    for (const auto &o : orb)
        for (int s = 0; s < 2; s++)
            if (o.occ[s] > 0.0) band.add(o.occ[s] * o.eps[s]);
    Neumaier pot;
    for (int s = 0; s < 2; s++) {
        std::vector<double> f(n);
        for (int i = 0; i < n; i++) f[i] = rho[s][i] * v_in[s][i];
        pot.add(shell_integral(grid, f));
    }
    const double kinetic = band.sum() - pot.sum();
    std::vector<double> total(n), f(n);
    for (int i = 0; i < n; i++) total[i] = rho[0][i] + rho[1][i];
    for (int i = 0; i < n; i++) f[i] = 4.0 * M_PI * grid.r[i] * total[i];
    const double nuclear = -z * grid.integrate(f);
    const auto vh = hartree_potential(grid, total);
    for (int i = 0; i < n; i++) f[i] = total[i] * vh[i];
    const double hartree = 0.5 * shell_integral(grid, f);
    for (int i = 0; i < n; i++) f[i] = pw92(rho[0][i], rho[1][i]);
    const double xc = shell_integral(grid, f);
    Neumaier e;
    e.add(kinetic);
    e.add(nuclear);
    e.add(hartree);
    e.add(xc);
    return e.sum();
}

// one SCF pass from rho = 0 (atom.rs:331-384) on the stacked spin densities
FullAtom full_pass(int z, const AtomOpts &o, bool use_diis, bool &exhausted) {
    const LogGrid &grid = universal_grid();
    const int n = grid.size();
    std::vector<double> w(2 * n);
    for (int i = 0; i < n; i++) w[i] = w[n + i] = grid.r[i] * grid.r[i] * grid.r[i] * grid.delta;
    std::vector<double> rho_in[2] = {std::vector<double>(n, 0.0), std::vector<double>(n, 0.0)};
    std::vector<FullOrbital> prev, g_orb;
    Diis diis(use_diis ? DIIS_HISTORY : 1);
    bool have_good = false;
    std::vector<double> g_in[2], g_out[2];
    int halvings = 0;
    exhausted = false;
    double delta = INFINITY;
    for (int it = 1; it <= SCF_MAX_ITER; it++) {
        std::vector<double> v[2], rho_out[2];
        full_potentials(grid, z, rho_in, v);
        std::vector<FullOrbital> orb;
        bool ok = true;
        try {
            orb = occupy_full(grid, z, v, o, prev, rho_out);
        } catch (const Unbound &) {
            ok = false;
        }
        if (ok) {
            if (use_diis) halvings = 0;
            delta = 0.0;
            for (int s = 0; s < 2; s++)
                for (int i = 0; i < n; i++) delta += std::fabs(rho_out[s][i] - rho_in[s][i]) * w[i];
            if (delta < SCF_TOL) {
                FullAtom a;
                a.z = z;
                a.restricted = o.restricted;
                a.total = full_energy(grid, z, orb, rho_out, v);
                double up = 0.0, dn = 0.0;
                for (const auto &q : orb) up += q.occ[0], dn += q.occ[1];
                a.multiplicity = (int)std::lround(std::fabs(up - dn)) + 1;
                a.orbitals = std::move(orb);
                a.rho[0] = rho_out[0];
                a.rho[1] = rho_out[1];
                full_potentials(grid, z, rho_out, a.v);
                return a;
            }
            have_good = true;
            for (int s = 0; s < 2; s++) {
                g_in[s] = rho_in[s];
                g_out[s] = rho_out[s];
            }
            g_orb = orb;
            prev = orb;
            std::vector<double> x(2 * n), y(2 * n);
            for (int i = 0; i < n; i++) {
                x[i] = rho_in[0][i], x[n + i] = rho_in[1][i];
                y[i] = rho_out[0][i], y[n + i] = rho_out[1][i];
            }
            const auto mixed = diis.mix(x, y, w, MIXING / (double)(1ull << halvings));
            for (int i = 0; i < n; i++) {
                rho_in[0][i] = mixed[i];
                rho_in[1][i] = mixed[n + i];
            }
        } else {
            if (!have_good) throw Unbound("Z = " + std::to_string(z) + ": the first step binds no shell");
            if (halvings == BACKTRACK_MAX) {
                exhausted = true;
                throw Unbound("Z = " + std::to_string(z) + ": an occupied shell is unbound after backtracking");
            }
            halvings++;
            const double beta = MIXING / (double)(1ull << halvings);
            diis.clear();
            for (int s = 0; s < 2; s++)
                for (int i = 0; i < n; i++) rho_in[s][i] = g_in[s][i] + beta * (g_out[s][i] - g_in[s][i]);
            prev = g_orb;
        }
    }
    throw Fail("Z = " + std::to_string(z) + ": the radial SCF did not converge");
}

FullAtom solve_atom_full(int z, const AtomOpts &o) {
    bool exhausted = false;
    try {
        return full_pass(z, o, true, exhausted);
    } catch (const Unbound &) {
        if (!exhausted) throw;
    }
    return full_pass(z, o, false, exhausted);
}

// the spin-polarised ground state: the lowest total energy over the allowed multiplicities (the parity of N, up to 5,
// while M - 1 <= N; atom.rs:172-175, 541-575)
FullAtom ground_state_polarised(int z) {
    std::optional<FullAtom> best;
    std::string last;
    for (int m = 1 + z % 2; m <= 5; m += 2) {
        if (m - 1 > z) break;
        AtomOpts o;
        o.restricted = false;
        o.multiplicity = m;
        try {
            FullAtom a = solve_atom_full(z, o);
            if (!best || a.total < best->total) best = std::move(a);
        } catch (const Fail &e) {
            last = e.what();
        }
    }
    if (!best) throw Fail("Hund atom: " + last);
    return std::move(*best);
}

// ---- the radial KB atom (kb.rs) ----
struct KbProj {
    int l;
    double e_kb;
    std::vector<double> beta;
};

struct KbInput {
    int z = 0;
    std::vector<double> v_loc, rho_pc;
    std::vector<KbProj> proj;
    const KbProj *of(int l) const {
        for (const auto &p : proj)
            if (p.l == l) return &p;
        return nullptr;
    }
};

struct KbAtom {
    std::vector<std::array<double, 2>> eps;
    std::vector<std::array<std::vector<double>, 2>> u;
    std::vector<double> rho[2], v[2];
    double energy = 0.0;
};

int support_index(const std::vector<double> &beta) {
    double m = 0.0;
    for (double b : beta) m = std::max(m, std::fabs(b));
    for (int i = (int)beta.size() - 1; i >= 0; i--)
        if (std::fabs(beta[i]) > 1e-12 * m) return i;
    return 0;
}

int matching_index(const std::vector<double> &beta) { return support_index(beta) + 3; }

// the solution vanishing at r_{i_start}, run inward to the first point (kb.rs:330-338)
std::vector<double> inward_dirichlet(const std::vector<double> &g, int i_start, double delta) {
    const double k = delta * delta / 12.0;
    std::vector<double> y(i_start + 1, 0.0);
    y[i_start - 1] = 1e-20;
    for (int i = i_start - 1; i >= 1; i--) {
        y[i - 1] = (2.0 * y[i] * (1.0 + 5.0 * k * g[i]) - y[i + 1] * (1.0 - k * g[i + 1])) / (1.0 - k * g[i - 1]);
        if (std::fabs(y[i - 1]) > 1e150)
            for (int q = i - 1; q <= i_start; q++) y[q] *= 1e-150;
    }
    return y;
}

int kb_wall(const std::vector<double> &g, double delta, int i_m) {
    return std::min(std::max(shooting_end(g, delta), i_m + 4), (int)g.size() - 1);
}

struct Green {
    std::vector<double> uh, ui, ch;
    double w = 0.0, w_scale = 0.0;
};

// the pieces of the Green-function form at energy e (kb.rs:295-328)
Green green(const LogGrid &grid, const std::vector<double> &v, const std::vector<double> &beta, int l, double e, int i_m) {
    const int n = grid.size();
    const auto g = g_coeff(grid, v, l, e);
    double y0, y1;
    regular_start(grid, l, 0.0, y0, y1);
    const auto yh = outward(g, y0, y1, i_m + 2, grid.delta, nullptr);
    const int ie = kb_wall(g, grid.delta, i_m);
    const auto yi = inward_dirichlet(g, ie, grid.delta);
    Green gr;
    gr.uh.resize(i_m + 3);
    for (int i = 0; i <= i_m + 2; i++) gr.uh[i] = yh[i] * std::sqrt(grid.r[i]);
    gr.ui.assign(n, 0.0);
    for (int i = 0; i <= ie; i++) gr.ui[i] = yi[i] * std::sqrt(grid.r[i]);
    auto dx = [&](const std::vector<double> &f, int i) {
        return (f[i - 2] - 8.0 * f[i - 1] + 8.0 * f[i + 1] - f[i + 2]) / (12.0 * grid.delta);
    };
    int iw = 2;
    for (int i = 2; i <= i_m - 2; i++)
        if (std::fabs(g[i]) < std::fabs(g[iw])) iw = i;
    const double a = gr.uh[iw] * dx(gr.ui, iw), b = dx(gr.uh, iw) * gr.ui[iw];
    std::vector<double> f(n, 0.0);
    for (int i = 0; i <= i_m; i++) f[i] = beta[i] * gr.uh[i];
    gr.ch = grid.cumulative(f);
    gr.w = (a - b) / grid.r[iw];
    gr.w_scale = (std::fabs(a) + std::fabs(b)) / grid.r[iw];
    return gr;
}

// E_KB W s(e) = -W + 4 E_KB int_0^{r_m} beta u_i C_h dr, and W (kb.rs:341-346)
std::pair<double, double> kb_green(const LogGrid &grid, const std::vector<double> &v, double e_kb,
                                   const std::vector<double> &beta, int l, double e, int i_m) {
    const Green gr = green(grid, v, beta, l, e, i_m);
    std::vector<double> f(grid.size(), 0.0);
    for (int i = 0; i <= i_m; i++) f[i] = beta[i] * gr.ui[i] * gr.ch[i];
    const double j = grid.partial(f, i_m);
    return {-gr.w + 4.0 * e_kb * j, gr.w};
}

double kb_lower_bound(const LogGrid &grid, const std::vector<double> &v, int l, const KbProj *p) {
    double floor = INFINITY;
    for (int i = 0; i < grid.size(); i++) floor = std::min(floor, v[i] + l * (l + 1.0) / (2.0 * grid.r[i] * grid.r[i]));
    if (!p) return floor;
    std::vector<double> bb(grid.size());
    for (int i = 0; i < grid.size(); i++) bb[i] = p->beta[i] * p->beta[i];
    return floor + std::min(p->e_kb * grid.integrate(bb), 0.0);
}

int local_count(const LogGrid &grid, const std::vector<double> &g, int l, int i_end) {
    double y0, y1;
    regular_start(grid, l, 0.0, y0, y1);
    return outward_nodes(g, y0, y1, i_end, grid.delta);
}

// the exact number of KB eigenvalues of channel l below e: n_loc(e) + [s(e) < 0] - [E_KB > 0] (kb.rs:371-383)
int kb_count_below(const LogGrid &grid, const std::vector<double> &v, const KbProj *p, int l, double e) {
    const auto g = g_coeff(grid, v, l, e);
    if (!p) return local_count(grid, g, l, shooting_end(g, grid.delta));
    const int i_m = matching_index(p->beta);
    const int n_loc = local_count(grid, g, l, kb_wall(g, grid.delta, i_m));
    const auto dw = kb_green(grid, v, p->e_kb, p->beta, l, e, i_m);
    const double sv = dw.first / (p->e_kb * dw.second);
    return std::max(n_loc + (sv < 0.0 ? 1 : 0) - (p->e_kb > 0.0 ? 1 : 0), 0);
}

// every eigenvalue of channel l in [e_lo, e_hi] (kb.rs:390-418)
std::vector<double> kb_eigenvalues(const LogGrid &grid, const std::vector<double> &v, const KbProj *p, int l, double e_lo,
                                   double e_hi, double mesh) {
    std::vector<double> roots;
    if (!p) {
        for (int k = 0;; k++) {
            auto b = bound_state(grid, v, l, k, 0.0, std::nullopt);
            if (!b || b->e >= e_hi) break;
            if (b->e >= e_lo) roots.push_back(b->e);
        }
        return roots;
    }
    const int i_m = matching_index(p->beta);
    auto det = [&](double e) { return kb_green(grid, v, p->e_kb, p->beta, l, e, i_m).first; };
    const int steps = std::max(1, (int)std::ceil((e_hi - e_lo) / mesh));
    double a = e_lo, fa = det(e_lo);
    for (int k = 1; k <= steps; k++) {
        const double b = std::min(e_lo + k * mesh, e_hi), fb = det(b);
        if (fa * fb < 0.0) {
            double lo = a, hi = b, flo = fa;
            while (hi - lo > 1e-13 * std::max(std::fabs(lo), 1.0)) {
                const double mid = 0.5 * (lo + hi), fm = det(mid);
                if (flo * fm <= 0.0) hi = mid;
                else {
                    lo = mid;
                    flo = fm;
                }
            }
            roots.push_back(0.5 * (lo + hi));
        }
        a = b;
        fa = fb;
    }
    return roots;
}

// the normalised eigenfunction at an eigenvalue, u proportional to G beta (kb.rs:429-446)
std::vector<double> kb_state(const LogGrid &grid, const std::vector<double> &v, const KbProj *p, int l, double e) {
    if (!p) return eigenfunction(grid, v, l, e, 0.0).u;
    const int n = grid.size();
    const int i_m = matching_index(p->beta);
    const Green gr = green(grid, v, p->beta, l, e, i_m);
    if (std::fabs(gr.w) < 1e-6 * gr.w_scale) return eigenfunction(grid, v, l, e, 0.0).u;
    std::vector<double> f(n, 0.0);
    for (int i = 0; i <= i_m; i++) f[i] = gr.ui[i] * p->beta[i];
    const auto di = grid.cumulative(f);
    const double d_tot = di[i_m];
    std::vector<double> u(n);
    for (int i = 0; i < n; i++)
        u[i] = i <= i_m ? gr.ui[i] * gr.ch[i] + gr.uh[i] * (d_tot - di[i]) : gr.ui[i] * gr.ch[i_m];
    std::vector<double> sq(n);
    double umax = 0.0;
    for (int i = 0; i < n; i++) {
        sq[i] = u[i] * u[i];
        umax = std::max(umax, std::fabs(u[i]));
    }
    const double norm = std::sqrt(grid.integrate(sq));
    int ip = 0;
    for (int i = n - 2; i >= 1; i--)
        if (std::fabs(u[i]) >= std::fabs(u[i - 1]) && std::fabs(u[i]) >= std::fabs(u[i + 1]) &&
            std::fabs(u[i]) > R_PEAK_FLOOR * umax) {
            ip = i;
            break;
        }
    const double sgn = u[ip] < 0.0 ? -1.0 : 1.0;
    for (auto &x : u) x = sgn * x / norm;
    return u;
}

// the lowest eigenvalue of channel l by bisection on the exact count (kb.rs:497-516)
std::optional<double> kb_lowest(const LogGrid &grid, const std::vector<double> &v, const KbProj *p, int l, double prev) {
    const std::optional<double> guess = std::isfinite(prev) ? std::optional<double>(prev) : std::nullopt;
    if (!p) {
        auto b = bound_state(grid, v, l, 0, 0.0, guess);
        return b ? std::optional<double>(b->e) : std::nullopt;
    }
    auto count = [&](double e) { return kb_count_below(grid, v, p, l, e); };
    double hi = 0.0;
    if (count(hi) == 0) return std::nullopt;
    double lo = kb_lower_bound(grid, v, l, p);
    if (guess) {
        const double gs = *guess, w = 0.1 * std::max(std::fabs(gs), 0.1);
        if (gs - w > lo && count(gs - w) == 0) lo = gs - w;
        if (gs + w < hi && count(gs + w) >= 1) hi = gs + w;
    }
    while (hi - lo > 1e-13 * std::max(std::fabs(lo), 1.0)) {
        const double mid = 0.5 * (lo + hi);
        if (count(mid) >= 1) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

void kb_potential(const LogGrid &grid, const KbInput &in, const std::vector<double> *rho, std::vector<double> *v) {
    const int n = grid.size();
    std::vector<double> total(n);
    for (int i = 0; i < n; i++) total[i] = rho[0][i] + rho[1][i];
    const auto vh = hartree_potential(grid, total);
    v[0].resize(n);
    v[1].resize(n);
    for (int i = 0; i < n; i++) {
        const double c = 0.5 * in.rho_pc[i];
        double vu, vd;
        pw92(rho[0][i] + c, rho[1][i] + c, &vu, &vd);
        v[0][i] = in.v_loc[i] + vh[i] + vu;
        v[1][i] = in.v_loc[i] + vh[i] + vd;
    }
}

using ValenceConfig = std::vector<std::array<double, 2>>;

// one SCF pass of the KB atom (kb.rs:532-609)
KbAtom kb_scf_pass(const LogGrid &grid, const KbInput &in, const ValenceConfig &cfg, bool restricted, bool linear,
                   bool &exhausted) {
    const int n = grid.size(), nl = (int)cfg.size();
    std::vector<double> w(2 * n);
    for (int i = 0; i < n; i++) w[i] = w[n + i] = grid.r[i] * grid.r[i] * grid.r[i] * grid.delta;
    std::vector<std::array<double, 2>> eps(nl, {NAN, NAN});
    std::vector<std::array<std::vector<double>, 2>> us(nl);
    auto occupy = [&](const std::vector<double> *v, std::vector<double> *rho) {
        rho[0].assign(n, 0.0);
        rho[1].assign(n, 0.0);
        for (int l = 0; l < nl; l++) {
            const KbProj *p = in.of(l);
            for (int s = 0; s < 2; s++) {
                if (cfg[l][s] <= 0.0) continue;
                if (restricted && s == 1 && cfg[l][0] > 0.0) {
                    eps[l][1] = eps[l][0];
                    us[l][1] = us[l][0];
                } else {
                    const int src = restricted ? 0 : s;
                    const auto e = kb_lowest(grid, v[src], p, l, eps[l][s]);
                    if (!e) throw Unbound("the KB atom does not bind l = " + std::to_string(l));
                    us[l][s] = kb_state(grid, v[src], p, l, *e);
                    eps[l][s] = *e;
                }
                for (int i = 0; i < n; i++)
                    rho[s][i] += cfg[l][s] * us[l][s][i] * us[l][s][i] / (4.0 * M_PI * grid.r[i] * grid.r[i]);
            }
        }
    };
    std::vector<double> rho_in[2] = {std::vector<double>(n, 0.0), std::vector<double>(n, 0.0)};
    Diis diis(linear ? 1 : DIIS_HISTORY);
    bool have_good = false;
    std::vector<double> g_in[2], g_out[2];
    std::vector<std::array<double, 2>> g_eps;
    std::vector<std::array<std::vector<double>, 2>> g_us;
    int halvings = 0;
    exhausted = false;
    for (int it = 0; it < SCF_MAX_ITER; it++) {
        std::vector<double> v[2], rho_out[2];
        kb_potential(grid, in, rho_in, v);
        try {
            occupy(v, rho_out);
        } catch (const Unbound &) {
            if (!have_good) throw;
            if (halvings == BACKTRACK_MAX) {
                exhausted = true;
                throw;
            }
            halvings++;
            const double beta = MIXING / (double)(1ull << halvings);
            diis.clear();
            for (int s = 0; s < 2; s++)
                for (int i = 0; i < n; i++) rho_in[s][i] = g_in[s][i] + beta * (g_out[s][i] - g_in[s][i]);
            eps = g_eps;
            us = g_us;
            continue;
        }
        if (!linear) halvings = 0;
        double delta = 0.0;
        for (int s = 0; s < 2; s++)
            for (int i = 0; i < n; i++) delta += std::fabs(rho_out[s][i] - rho_in[s][i]) * w[i];
        if (delta < SCF_TOL) {
            Neumaier band, pot;
            for (int l = 0; l < nl; l++)
                for (int s = 0; s < 2; s++)
                    if (cfg[l][s] > 0.0) band.add(cfg[l][s] * eps[l][s]);
            std::vector<double> f(n), total(n);
            for (int s = 0; s < 2; s++) {
                for (int i = 0; i < n; i++) f[i] = rho_out[s][i] * v[s][i];
                pot.add(shell_integral(grid, f));
            }
            for (int i = 0; i < n; i++) total[i] = rho_out[0][i] + rho_out[1][i];
            const auto vh = hartree_potential(grid, total);
            for (int i = 0; i < n; i++) f[i] = total[i] * in.v_loc[i];
            const double e_loc = shell_integral(grid, f);
            for (int i = 0; i < n; i++) f[i] = total[i] * vh[i];
            const double e_h = 0.5 * shell_integral(grid, f);
            for (int i = 0; i < n; i++) {
                const double c = 0.5 * in.rho_pc[i];
                f[i] = pw92(rho_out[0][i] + c, rho_out[1][i] + c);
            }
            const double e_xc = shell_integral(grid, f);
            Neumaier e;
            e.add(band.sum());
            e.add(-pot.sum());
            e.add(e_loc);
            e.add(e_h);
            e.add(e_xc);
            KbAtom a;
            a.eps = eps;
            a.u = us;
            a.rho[0] = rho_out[0];
            a.rho[1] = rho_out[1];
            kb_potential(grid, in, rho_out, a.v);
            a.energy = e.sum();
            return a;
        }
        have_good = true;
        for (int s = 0; s < 2; s++) {
            g_in[s] = rho_in[s];
            g_out[s] = rho_out[s];
        }
        g_eps = eps;
        g_us = us;
        std::vector<double> x(2 * n), y(2 * n);
        for (int i = 0; i < n; i++) {
            x[i] = rho_in[0][i], x[n + i] = rho_in[1][i];
            y[i] = rho_out[0][i], y[n + i] = rho_out[1][i];
        }
        const auto mixed = diis.mix(x, y, w, MIXING / (double)(1ull << halvings));
        for (int i = 0; i < n; i++) {
            rho_in[0][i] = mixed[i];
            rho_in[1][i] = mixed[n + i];
        }
    }
    throw Fail("the KB atom's SCF did not converge");
}

KbAtom solve_kb_atom(const KbInput &in, const ValenceConfig &cfg, bool restricted) {
    const LogGrid &grid = universal_grid();
    bool exhausted = false;
    try {
        return kb_scf_pass(grid, in, cfg, restricted, false, exhausted);
    } catch (const Unbound &) {
        if (!exhausted) throw;
    }
    return kb_scf_pass(grid, in, cfg, restricted, true, exhausted);
}

// ---- the band-limiting filter (filter.rs) ----
double sph_bessel(int l, double x);

// j_l(x) / x^l, finite at 0 (filter.rs:174-188)
double sph_bessel_scaled(int l, double x) {
    if (std::fabs(x) < 1.0) {
        double df = 1.0;
        for (int k = 1; k <= l; k++) df *= 2.0 * k + 1.0;
        double term = 1.0 / df, sum = term;
        for (int k = 1; k < 10; k++) {
            term *= -x * x / (2.0 * k * (2.0 * l + 2.0 * k + 1.0));
            sum += term;
        }
        return sum;
    }
    return sph_bessel(l, x) / std::pow(x, l);
}

double sph_bessel(int l, double x) {
    if (std::fabs(x) < 1.0) return sph_bessel_scaled(l, x) * std::pow(x, l);
    const double s = std::sin(x), c = std::cos(x);
    switch (l) {
    case 0: return s / x;
    case 1: return s / (x * x) - c / x;
    case 2: return (3.0 / (x * x) - 1.0) * s / x - 3.0 * c / (x * x);
    case 3: return (15.0 / (x * x * x) - 6.0 / x) * s / x - (15.0 / (x * x) - 1.0) * c / x;
    default: {
        double a = sph_bessel(2, x), b = sph_bessel(3, x);
        for (int k = 3; k < l; k++) {
            const double nb = (2.0 * k + 1.0) / x * b - a;
            a = b;
            b = nb;
        }
        return b;
    }
    }
}

// the C^inf step psi(1-t) / (psi(t) + psi(1-t)), psi(t) = e^{-1/t} (filter.rs:206-217)
double mask(double r, double r_in, double r_out) {
    if (r <= r_in) return 1.0;
    if (r >= r_out) return 0.0;
    const double t = (r - r_in) / (r_out - r_in);
    auto psi = [](double x) { return x > 0.0 ? std::exp(-1.0 / x) : 0.0; };
    return psi(1.0 - t) / (psi(t) + psi(1.0 - t));
}

// uniform Simpson mesh on [a, b], an even number of intervals, dq <= pi / (2 r_max) (filter.rs:219-226)
void q_mesh(double a, double b, std::vector<double> &q, std::vector<double> &w) {
    const int n = 2 * std::max((int)std::ceil((b - a) * R_MAX / M_PI), 1);
    const double dq = (b - a) / n;
    q.resize(n + 1);
    w.resize(n + 1);
    for (int k = 0; k <= n; k++) {
        q[k] = a + k * dq;
        w[k] = dq / 3.0 * (k == 0 || k == n ? 1.0 : (k % 2 == 1 ? 4.0 : 2.0));
    }
}

// The completed function as a finite Bessel sum times the outer step: F(r) / r^l = (1 / 2 pi^2) sum_k c_k q_k^{l+2}
// j_l(q_k r) / (q_k r)^l (filter.rs:245-254)
struct Completion {
    int l = 0;
    std::vector<double> q, c;
    double r_cut = 0.0, width = 1.0;
    double value_at(double r) const {
        const double m = mask(r, r_cut, r_cut + width);
        if (m == 0.0) return 0.0;
        Neumaier s;
        for (size_t k = 0; k < q.size(); k++) s.add(c[k] * std::pow(q[k], l + 2) * sph_bessel_scaled(l, q[k] * r));
        return s.sum() / (2.0 * M_PI * M_PI) * m;
    }
};

// One-sided Jacobi SVD of the rows x cols matrix a (row-major): a = U diag(s) V^T with U's columns returned in u
// (rows x cols, column k = u[i * cols + k]) and V (cols x cols, column k = v[j * cols + k]).
void jacobi_svd(std::vector<double> a, int rows, int cols, std::vector<double> &u, std::vector<double> &s,
                std::vector<double> &v) {
    v.assign((size_t)cols * cols, 0.0);
    for (int j = 0; j < cols; j++) v[j * cols + j] = 1.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (int p = 0; p < cols - 1; p++)
            for (int q = p + 1; q < cols; q++) {
                double alpha = 0.0, beta = 0.0, gamma = 0.0;
                for (int i = 0; i < rows; i++) {
                    const double x = a[(size_t)i * cols + p], y = a[(size_t)i * cols + q];
                    alpha += x * x;
                    beta += y * y;
                    gamma += x * y;
                }
                if (gamma == 0.0) continue;
                const double rel = std::fabs(gamma) / std::sqrt(alpha * beta);
                off = std::max(off, rel);
                if (rel < 1e-15) continue;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                for (int i = 0; i < rows; i++) {
                    const double x = a[(size_t)i * cols + p], y = a[(size_t)i * cols + q];
                    a[(size_t)i * cols + p] = c * x - sn * y;
                    a[(size_t)i * cols + q] = sn * x + c * y;
                }
                for (int i = 0; i < cols; i++) {
                    const double x = v[i * cols + p], y = v[i * cols + q];
                    v[i * cols + p] = c * x - sn * y;
                    v[i * cols + q] = sn * x + c * y;
                }
            }
        if (off < 1e-15) break;
    }
    s.assign(cols, 0.0);
    u.assign((size_t)rows * cols, 0.0);
    for (int k = 0; k < cols; k++) {
        double nrm = 0.0;
        for (int i = 0; i < rows; i++) nrm += a[(size_t)i * cols + k] * a[(size_t)i * cols + k];
        nrm = std::sqrt(nrm);
        s[k] = nrm;
        if (nrm > 0.0)
            for (int i = 0; i < rows; i++) u[(size_t)i * cols + k] = a[(size_t)i * cols + k] / nrm;
    }
}

// King-Smith, Payne & Lin's band completion (filter.rs:256-295): the transform kept on [0, q_c]; on (q_c, q_max] a
// Chebyshev band minimising the tail integral_{r_cut}^{r_max} |F|^2 r^2 dr (least squares, discrete Picard truncation);
// the outer step over [r_cut, r_cut + width]
Completion complete(const LogGrid &grid, const std::vector<double> &f, int l, double q_c, double q_max, double r_cut,
                    double width) {
    const int n = grid.size();
    std::vector<double> qk, wk, qb, wb;
    q_mesh(0.0, q_c, qk, wk);
    q_mesh(q_c, q_max, qb, wb);
    std::vector<double> fk(qk.size()), tmp(n);
    for (size_t k = 0; k < qk.size(); k++) {
        for (int i = 0; i < n; i++) tmp[i] = f[i] * sph_bessel(l, qk[k] * grid.r[i]) * grid.r[i] * grid.r[i];
        fk[k] = 4.0 * M_PI * grid.integrate(tmp);
    }
    const int m = KSPL_TERMS;
    std::vector<double> cheb(qb.size() * m);
    for (size_t k = 0; k < qb.size(); k++) {
        const double s = 2.0 * (qb[k] - q_c) / (q_max - q_c) - 1.0;
        double *t = &cheb[k * m];
        t[0] = 1.0;
        if (m > 1) t[1] = s;
        for (int c = 2; c < m; c++) t[c] = 2.0 * s * t[c - 1] - t[c - 2];
    }
    const double dr = M_PI / (4.0 * q_max);
    std::vector<double> rs;
    for (int j = 0;; j++) {
        const double r = r_cut + j * dr;
        if (r > R_MAX) break;
        rs.push_back(r);
    }
    for (int i = 0; i < n; i++) tmp[i] = std::fabs(f[i]) * grid.r[i] * grid.r[i];
    const double e_f = 2.220446049250313e-16 * 4.0 * M_PI * grid.integrate(tmp);
    const int rows = (int)rs.size();
    std::vector<double> a((size_t)rows * m, 0.0), b(rows, 0.0);
    double noise2 = 0.0;
    for (int j = 0; j < rows; j++) {
        const double r = rs[j], sw = std::sqrt(r * r * dr) / (2.0 * M_PI * M_PI);
        Neumaier acc;
        double nz = 0.0;
        for (size_t k = 0; k < qk.size(); k++) {
            const double jl = sph_bessel(l, qk[k] * r);
            acc.add(wk[k] * fk[k] * qk[k] * qk[k] * jl);
            nz += wk[k] * qk[k] * qk[k] * std::fabs(jl);
        }
        b[j] = -sw * acc.sum();
        noise2 += (e_f * sw * nz) * (e_f * sw * nz);
        for (size_t k = 0; k < qb.size(); k++) {
            const double jb = sw * wb[k] * qb[k] * qb[k] * sph_bessel(l, qb[k] * r);
            for (int c = 0; c < m; c++) a[(size_t)j * m + c] += jb * cheb[k * m + c];
        }
    }
    std::vector<double> uu, sv, vv;
    jacobi_svd(a, rows, m, uu, sv, vv);
    double smax = 0.0;
    for (double s : sv) smax = std::max(smax, s);
    const double rank_floor = 2.220446049250313e-16 * std::max(rows, m) * smax, noise = std::sqrt(noise2);
    std::vector<double> x(m, 0.0);
    for (int k = 0; k < m; k++) {
        double ub = 0.0;
        for (int j = 0; j < rows; j++) ub += uu[(size_t)j * m + k] * b[j];
        if (sv[k] > rank_floor && std::fabs(ub) > noise)
            for (int c = 0; c < m; c++) x[c] += vv[c * m + k] * (ub / sv[k]);
    }
    Completion out;
    out.l = l, out.r_cut = r_cut, out.width = width;
    for (size_t k = 0; k < qk.size(); k++) {
        out.q.push_back(qk[k]);
        out.c.push_back(wk[k] * fk[k]);
    }
    for (size_t k = 0; k < qb.size(); k++) {
        double fb = 0.0;
        for (int c = 0; c < m; c++) fb += cheb[k * m + c] * x[c];
        out.q.push_back(qb[k]);
        out.c.push_back(wb[k] * fb);
    }
    return out;
}

std::vector<double> uniform_fill(double du, double radius, const Completion &c) {
    const int n = (int)std::ceil(radius / du) + 4;
    std::vector<double> t(n);
    for (int k = 0; k < n; k++) t[k] = c.value_at(k * du);
    return t;
}

}  // namespace

// ---- a generated, filtered pseudopotential ----
struct Generated;

struct Table {
    lira_pseudo_record rec{};
    std::shared_ptr<const Generated> gen;  // the accepted entry the tables were filtered from
    double du = 0.0, sigma_lr = 0.0, r_sr = 0.0, pc_radius = 0.0, r_core = 0.0;
    std::vector<double> v_sr_u, rho_pc_u;
    std::vector<std::vector<double>> t_u;
    std::vector<double> proj_radius;
    // the same band-limited functions at the log grid's nodes (filled for the filtered re-run of T8)
    std::vector<double> v_sr_log, rho_pc_log;
    std::vector<std::vector<double>> t_log;
};

namespace {

struct T3Config {
    std::string name;
    ValenceConfig pp;
    bool restricted;
    double de_ae;
};

struct Candidate {
    std::vector<TmChannel> channels;
    Unscreened un;
    KbPotential kb;
    std::vector<double> screen;
    double t5[4] = {0, 0, 0, 0};
    KbAtom ref_atom;              // the reference KB atom (unfiltered)
    std::vector<T3Config> t3;     // the transfer configurations both atoms bind, with the all-electron dE
};

// What the gates need of an element, independent of the ladder entry: the all-electron bound spectrum above the core
// per l (partition.rs:118-121), the restricted reference's total energy, and the all-electron atoms of the transfer
// configurations and of T6's shifted occupations, solved once and kept.
struct ElementData {
    const AtomResult &ref;
    const Partition &part;
    const PartialCore &pc;
    std::vector<std::vector<double>> spectrum;
    double ref_energy = 0.0;
    std::optional<FullAtom> hund;
    std::string hund_error;
    bool hund_tried = false;
    std::deque<std::pair<std::string, std::optional<FullAtom>>> ae_cache;  // key -> atom (nullopt: unbound); a deque
                                                                            // keeps references to its entries valid
    std::vector<std::pair<std::string, std::string>> ae_errors;             // key -> a failure that is not unbound
};

ValenceConfig reference_config(const Partition &part) {
    ValenceConfig c(part.l_max + 1, std::array<double, 2>{0.0, 0.0});
    for (const auto &ch : part.channels)
        if (ch.occupied) c[ch.l] = {0.5 * ch.occ, 0.5 * ch.occ};
    return c;
}

// the core full in both spins plus the valence channels' (n, l) at the configuration's occupations (gates.rs:120-129)
std::vector<ShellOcc> ae_shells(const Partition &part, const ValenceConfig &val) {
    std::vector<ShellOcc> v;
    for (const Shell *s : part.core) v.push_back({s->n, s->l, {0.5 * s->occ, 0.5 * s->occ}});
    for (int l = 0; l < (int)val.size(); l++) {
        if (val[l][0] + val[l][1] <= 0.0) continue;
        int n = -1;
        for (const auto &c : part.channels)
            if (c.l == l) n = c.n;
        if (n < 0) throw Fail("a configuration occupies an l without a channel");
        v.push_back({n, l, {val[l][0], val[l][1]}});
    }
    return v;
}

std::string config_key(const ValenceConfig &c) {
    std::string k;
    for (const auto &o : c) k += std::to_string(o[0]) + "/" + std::to_string(o[1]) + ";";
    return k;
}

// the restricted all-electron atom at fixed occupations, cached per configuration; nullopt when it is unbound
const std::optional<FullAtom> &ae_fixed(ElementData &d, const ValenceConfig &val) {
    const std::string key = config_key(val);
    for (auto &e : d.ae_cache)
        if (e.first == key) return e.second;
    for (auto &e : d.ae_errors)
        if (e.first == key) throw Fail(e.second);
    AtomOpts o;
    o.restricted = true;
    o.aufbau = false;
    o.fixed = ae_shells(d.part, val);
    try {
        d.ae_cache.emplace_back(key, solve_atom_full(d.part.z, o));
    } catch (const Unbound &) {
        d.ae_cache.emplace_back(key, std::nullopt);
    } catch (const Fail &e) {
        d.ae_errors.emplace_back(key, e.what());
        throw;
    }
    return d.ae_cache.back().second;
}

const FullAtom &hund_atom(ElementData &d) {
    if (!d.hund_tried) {
        d.hund_tried = true;
        try {
            d.hund = ground_state_polarised(d.part.z);
        } catch (const Fail &e) {
            d.hund_error = e.what();
        }
    }
    if (!d.hund) throw Fail("Hund atom: " + d.hund_error);
    return *d.hund;
}

KbInput kb_input(const KbPotential &kb, const PartialCore &pc, int z) {
    KbInput in;
    in.z = z;
    in.v_loc = kb.v_loc;
    in.rho_pc = pc.rho_pc;
    for (const auto &p : kb.projectors) in.proj.push_back({p.l, p.e_kb, p.beta});
    return in;
}

struct Missing {
    int l;
    std::vector<double> ae, roots;
};

// G2 (kb.rs:635-663): the reference KB atom reproduces the occupied eigenvalues to 1 mHa, and per l <= l_max the exact
// count of KB levels below top = min(0, eps_v,max + 0.5) neither exceeds the all-electron atom's levels of l above its
// core (more: a ghost) nor disagrees with the roots located on the 1 mHa mesh; fewer is recorded for T5
std::vector<Missing> ghost_g2(const KbInput &in, const KbAtom &atom, const ElementData &d) {
    const LogGrid &g = universal_grid();
    double err = 0.0, eps_v_max = -INFINITY;
    for (const auto &c : d.part.channels)
        if (c.occupied) {
            err = std::max(err, std::fabs(atom.eps[c.l][0] - c.eps));
            eps_v_max = std::max(eps_v_max, c.eps);
        }
    if (err >= 1e-3) throw Fail("G2: KB eigenvalues miss the reference by " + std::to_string(err * 1e3) + " mHa");
    const double top = std::min(eps_v_max + 0.5, 0.0);
    std::vector<Missing> missing;
    for (int l = 0; l <= d.part.l_max; l++) {
        const KbProj *p = in.of(l);
        const int count = kb_count_below(g, atom.v[0], p, l, top);
        const auto roots = kb_eigenvalues(g, atom.v[0], p, l, kb_lower_bound(g, atom.v[0], l, p), top, 1e-3);
        if ((int)roots.size() != count)
            throw Fail("G2: l = " + std::to_string(l) + ": the mesh located " + std::to_string(roots.size()) +
                       " roots but the inertia count is " + std::to_string(count));
        std::vector<double> ae;
        for (double e : d.spectrum[l])
            if (e < top) ae.push_back(e);
        if (count > (int)ae.size()) throw Fail("G2: l = " + std::to_string(l) + " has an extra bound state (ghost)");
        if (count < (int)ae.size()) missing.push_back({l, ae, roots});
    }
    return missing;
}

// T7 in one configuration (gates.rs:221-245): per spin and l <= l_max the KB count below top must not exceed the
// all-electron levels of l above the core there
void t7_counts(const KbInput &in, const ElementData &d, const FullAtom &ae, const KbAtom &pp, bool restricted,
               const std::string &name) {
    const LogGrid &g = universal_grid();
    double eps_v_max = -INFINITY;
    for (const auto &o : ae.orbitals) {
        bool valence = false;
        for (const Shell *v : d.part.valence)
            if (v->n == o.n && v->l == o.l) valence = true;
        if (!valence) continue;
        for (int s = 0; s < 2; s++)
            if (o.occ[s] > 0.0) eps_v_max = std::max(eps_v_max, o.eps[s]);
    }
    const double top = std::min(eps_v_max + 0.5, 0.0);
    for (int s = 0; s < (restricted ? 1 : 2); s++)
        for (int l = 0; l <= d.part.l_max; l++) {
            int n_core_l = 0;
            for (const Shell *c : d.part.core)
                if (c->l == l) n_core_l++;
            int levels = 0;
            for (int k = n_core_l;; k++) {
                auto b = bound_state(g, ae.v[s], l, k, d.part.z, std::nullopt);
                if (!b || b->e >= top) break;
                levels++;
            }
            const int n_kb = kb_count_below(g, pp.v[s], in.of(l), l, top);
            if (n_kb > levels)
                throw Fail("T7 " + name + " spin " + std::to_string(s) + ", l = " + std::to_string(l) +
                           ": an extra KB bound state (ghost)");
        }
}

// T1-T7 on a candidate that passed G1/G2 (gates.rs:251-371); T5 on the channels with G2's missing AE states
void run_gates(Candidate &c, const KbInput &in, const KbAtom &pp_ref, const std::vector<Missing> &missing,
               ElementData &d) {
    const LogGrid &g = universal_grid();
    const Partition &part = d.part;
    // T1: every channel with a reference state, an unoccupied one through its lowest KB level
    double t1 = 0.0;
    for (const auto &ch : part.channels) {
        double e;
        if (ch.occupied) e = pp_ref.eps[ch.l][0];
        else {
            const auto lo = kb_lowest(g, pp_ref.v[0], in.of(ch.l), ch.l, NAN);
            if (!lo) throw Fail("T1: the unoccupied l = " + std::to_string(ch.l) + " reference state is not bound");
            e = *lo;
        }
        t1 = std::max(t1, std::fabs(e - ch.eps));
    }
    if (t1 >= 1e-3) throw Fail("T1: " + std::to_string(t1 * 1e3) + " mHa");
    // T2: norm conservation as a regression
    for (size_t k = 0; k < c.channels.size(); k++) {
        const TmChannel &tm = c.channels[k];
        std::vector<double> a(g.size()), b(g.size());
        for (int i = 0; i < g.size(); i++) {
            a[i] = tm.u[i] * tm.u[i];
            b[i] = part.channels[k].u[i] * part.channels[k].u[i];
        }
        if (std::fabs(g.partial(a, tm.i_c) - g.partial(b, tm.i_c)) >= 1e-10) throw Fail("T2: norm error");
    }
    // T3 / T4 / T7 over the transfer configurations: cation, s -> p, anion, Hund
    const ValenceConfig base = reference_config(part);
    auto total = [](const ValenceConfig &v, int l) { return v[l][0] + v[l][1]; };
    auto cap = [](int l) { return 2.0 * (2 * l + 1); };
    auto moved = [&](int from, int to) {
        ValenceConfig v = base;
        if (from >= 0) v[from][0] -= 0.5, v[from][1] -= 0.5;
        if (to >= 0) v[to][0] += 0.5, v[to][1] += 0.5;
        return v;
    };
    auto has = [&](int l) {
        for (const auto &ch : part.channels)
            if (ch.l == l) return true;
        return false;
    };
    auto eps_of = [&](int l) -> double {
        for (const auto &ch : part.channels)
            if (ch.l == l) return ch.eps;
        return INFINITY;
    };
    int homo = -1;
    double homo_e = -INFINITY;
    for (const auto &ch : part.channels)
        if (ch.occupied && ch.eps > homo_e) homo_e = ch.eps, homo = ch.l;
    struct Config {
        std::string name;
        ValenceConfig pp;
        bool restricted;
    };
    std::vector<Config> configs{{"cation", moved(homo, -1), true}};
    if (has(0) && has(1) && total(base, 0) >= 1.0 && total(base, 1) <= cap(1) - 1.0)
        configs.push_back({"s->p", moved(0, 1), true});
    int room = -1;
    for (int l = 0; l < (int)base.size(); l++)
        if (has(l) && total(base, l) < cap(l) && (room < 0 || eps_of(l) < eps_of(room))) room = l;
    if (room >= 0) configs.push_back({"anion", moved(-1, room), true});
    const FullAtom &hund = hund_atom(d);
    ValenceConfig hund_pp(base.size(), std::array<double, 2>{0.0, 0.0});
    for (const auto &o : hund.orbitals)
        for (const Shell *v : part.valence)
            if (v->n == o.n && v->l == o.l) hund_pp[o.l] = {o.occ[0], o.occ[1]};
    configs.push_back({"Hund", hund_pp, false});
    for (const auto &cfg : configs) {
        if (cfg.name == "Hund") {
            double n_v = 0.0;
            for (const auto &o : cfg.pp) n_v += o[0] + o[1];
            if (std::fabs(n_v - part.z_v) > 1e-9) throw Fail("T3 Hund: the polarised ground state's valence count differs");
        }
        const FullAtom *ae = nullptr;
        if (cfg.name == "Hund") ae = &hund;
        else {
            const auto &r = ae_fixed(d, cfg.pp);
            if (r) ae = &*r;
        }
        std::optional<KbAtom> pp;
        try {
            pp = solve_kb_atom(in, cfg.pp, cfg.restricted);
        } catch (const Unbound &) {
        }
        if (!ae || !pp) {
            if (cfg.name == "anion") continue;  // not bound in both: an LSDA property, recorded by the reference
            throw Fail("T3 " + cfg.name + ": an atom of the configuration is unbound");
        }
        const double de_ae = ae->total - d.ref_energy, de_pp = pp->energy - pp_ref.energy;
        double worst = 0.0;
        for (const auto &o : ae->orbitals) {
            bool valence = false;
            for (const Shell *v : part.valence)
                if (v->n == o.n && v->l == o.l) valence = true;
            if (!valence) continue;
            for (int s = 0; s < 2; s++)
                if (o.occ[s] > 0.0) worst = std::max(worst, std::fabs(o.eps[s] - pp->eps[o.l][s]));
        }
        if (std::fabs(de_ae - de_pp) >= 5e-3)
            throw Fail("T3 " + cfg.name + ": dE_AE " + std::to_string(de_ae) + " vs dE_PP " + std::to_string(de_pp));
        c.t3.push_back({cfg.name, cfg.pp, cfg.restricted, de_ae});
        if (worst >= 5e-3) throw Fail("T4 " + cfg.name + ": eigenvalue error " + std::to_string(worst * 1e3) + " mHa");
        t7_counts(in, d, *ae, *pp, cfg.restricted, cfg.name);
    }
    // T5: every generated channel on eps_l +- 0.5 Ha plus the AE states of its l the KB atom does not bind
    for (size_t k = 0; k < part.channels.size() && k < 4; k++) {
        const int l = part.channels[k].l;
        std::vector<double> extra;
        for (const auto &m : missing)
            if (m.l == l)
                for (size_t q = m.roots.size(); q < m.ae.size(); q++) extra.push_back(m.ae[q]);
        c.t5[k] = t5_channel(g, d.ref, c.channels, c.kb, c.screen, l, part.channels[k].eps, extra);
        if (c.t5[k] >= T5_DE_EQUIV)
            throw Fail("T5 l = " + std::to_string(l) + ": max delta E_equiv " + std::to_string(c.t5[k] * 1e3) + " mHa");
    }
    // T6: d eps_l / d f_l' by +-0.25 e (restricted), AE vs PP; a full shell one-sided over [f - 0.5, f]
    std::vector<int> occupied;
    for (const auto &ch : part.channels)
        if (ch.occupied) occupied.push_back(ch.l);
    for (int lp : occupied) {
        const bool full = total(base, lp) + 0.25 > cap(lp) + 1e-12;
        const double hi = full ? 0.0 : 0.25, lo = full ? -0.5 : -0.25;
        ValenceConfig cp = base, cm = base;
        cp[lp][0] += 0.5 * hi, cp[lp][1] += 0.5 * hi;
        cm[lp][0] += 0.5 * lo, cm[lp][1] += 0.5 * lo;
        const auto &ae_p = ae_fixed(d, cp);
        const auto &ae_m = ae_fixed(d, cm);
        if (!ae_p || !ae_m) throw Fail("T6: an all-electron atom is unbound");
        KbAtom pp_p, pp_m;
        try {
            pp_p = solve_kb_atom(in, cp, true);
            pp_m = solve_kb_atom(in, cm, true);
        } catch (const Unbound &) {
            throw Fail("T6: a KB atom is unbound");
        }
        for (int l : occupied) {
            int n = 0;
            for (const auto &ch : part.channels)
                if (ch.l == l) n = ch.n;
            auto ae_eps = [&](const FullAtom &a) -> double {
                for (const auto &o : a.orbitals)
                    if (o.n == n && o.l == l) return o.eps[0];
                return NAN;
            };
            const double slope_ae = (ae_eps(*ae_p) - ae_eps(*ae_m)) / (hi - lo);
            const double slope_pp = (pp_p.eps[l][0] - pp_m.eps[l][0]) / (hi - lo);
            if (!(std::fabs(slope_ae - slope_pp) < 10e-3))
                throw Fail("T6 d eps_" + std::to_string(l) + " / d f_" + std::to_string(lp));
        }
    }
}

Candidate candidate(ElementData &d, const std::vector<double> &tau, double f, int local) {
    const LogGrid &g = universal_grid();
    const Partition &part = d.part;
    Candidate c;
    c.channels = tm_channels(d.ref, part, f);
    std::vector<double> occ;
    for (const auto &ch : part.channels) occ.push_back(ch.occ);
    c.un = unscreen(g, part.z_v, c.channels, occ, d.pc.rho_pc);
    c.screen.resize(g.size());
    for (int i = 0; i < g.size(); i++) c.screen[i] = c.channels[0].v_scr[i] - c.un.v_ion[0][i];
    c.kb = kb_form(g, part.z_v, c.channels, c.un, local, tau);
    ghost_g1(g, c.kb, c.screen, part);
    const KbInput in = kb_input(c.kb, d.pc, part.z);
    KbAtom pp_ref;
    try {
        pp_ref = solve_kb_atom(in, reference_config(part), true);
    } catch (const Fail &e) {
        throw Fail(std::string("G2: ") + e.what());
    }
    const auto missing = ghost_g2(in, pp_ref, d);
    run_gates(c, in, pp_ref, missing, d);
    c.ref_atom = std::move(pp_ref);
    return c;
}

// the filtered tables at spacing h (filter.rs:384-478): V_SR = V_loc + Z_v erf(d/a)/d on R_SR = max(r_cmp, r_floor), every
// beta_l / r on r_beta, rho_pc on its XC-weight support; each completed and sampled on the uniform 3-D table
void filter_tables(Table &t, const Partition &part, const PartialCore &pc, const Candidate &c, double h,
                   bool with_log = false) {
    const LogGrid &g = universal_grid();
    const int n = g.size();
    const double a = SIGMA_LR_OVER_H * h * std::sqrt(2.0);
    const double q_c = Q_C_OVER_NYQUIST * M_PI / h, q_max = M_PI / h, width = KSPL_STEP_OVER_H * h;
    t.du = h / TABLE_POINTS_PER_H;
    t.sigma_lr = SIGMA_LR_OVER_H * h;
    // V_SR and its support
    std::vector<double> v_sr(n);
    for (int i = 0; i < n; i++) v_sr[i] = c.kb.v_loc[i] + part.z_v * std::erf(g.r[i] / a) / g.r[i];
    double r_floor = g.r[0];
    for (int i = n - 1; i >= 0; i--)
        if (part.z_v * std::erfc(g.r[i] / a) / g.r[i] >= V_TAIL_FLOOR) {
            r_floor = g.r[i];
            break;
        }
    const double r_sr = std::max(c.kb.r_cmp, r_floor);
    const Completion cv = complete(g, v_sr, 0, q_c, q_max, GAMMA_R * r_sr, width);
    t.r_sr = GAMMA_R * r_sr + width;
    t.v_sr_u = uniform_fill(t.du, t.r_sr, cv);
    auto on_log = [&](const Completion &cc) {
        std::vector<double> out(n);
        for (int i = 0; i < n; i++) out[i] = cc.value_at(g.r[i]);
        return out;
    };
    if (with_log) t.v_sr_log = on_log(cv);
    // projectors: F_l = beta_u / r on r_beta
    for (const auto &p : c.kb.projectors) {
        std::vector<double> fl(n);
        for (int i = 0; i < n; i++) fl[i] = p.beta[i] / g.r[i];
        const double rb = g.r[p.i_beta];
        const Completion cb = complete(g, fl, p.l, q_c, q_max, GAMMA_R * rb, width);
        t.proj_radius.push_back(GAMMA_R * rb + width);
        t.t_u.push_back(uniform_fill(t.du, t.proj_radius.back(), cb));
        if (with_log) t.t_log.push_back(on_log(cb));
    }
    // the partial core on its XC-weight support: the smallest r with 4 pi int_r^inf |v_xc| rho_pc r'^2 dr' <= v_tail_floor,
    // v_xc the mean spin potential on rho_v + rho_pc of the reference KB atom (filter.rs:400-415)
    bool core = false;
    for (double x : pc.rho_pc)
        if (x != 0.0) core = true;
    if (core) {
        std::vector<double> wv(n);
        for (int i = 0; i < n; i++) {
            const double cc = 0.5 * pc.rho_pc[i];
            double vu, vd;
            pw92(c.ref_atom.rho[0][i] + cc, c.ref_atom.rho[1][i] + cc, &vu, &vd);
            wv[i] = 4.0 * M_PI * g.r[i] * g.r[i] * std::fabs(0.5 * (vu + vd)) * pc.rho_pc[i];
        }
        const auto cum = g.cumulative(wv);
        const double total = cum[n - 1];
        double r_pc = g.r[n - 1];
        for (int i = 0; i < n; i++)
            if (total - cum[i] <= V_TAIL_FLOOR) {
                r_pc = g.r[i];
                break;
            }
        const Completion cp = complete(g, pc.rho_pc, 0, q_c, q_max, GAMMA_R * r_pc, width);
        t.pc_radius = GAMMA_R * r_pc + width;
        t.rho_pc_u = uniform_fill(t.du, t.pc_radius, cp);
        t.rec.r_pc_support = r_pc;
        if (with_log) t.rho_pc_log = on_log(cp);
    } else {
        t.rho_pc_u.assign(4, 0.0);
        t.pc_radius = 0.0;
        if (with_log) t.rho_pc_log.assign(n, 0.0);
    }
    t.r_core = pc.r_core;
}

// ---- T8: the filtered re-run at every ladder spacing, h_req and h_eb (filter.rs:480-676) ----
constexpr double H_LADDER[5] = {0.35, 0.30, 0.25, 0.20, 0.15};
constexpr double H_REQ_TOL = 1e-3, H_REQ_FORCE = 2e-6, T3_DELTA_E = 5e-3;

// 4-point Lagrange on a uniform table of an even function, index -1 mirroring index 1: (value, d/dd); 0 beyond it
std::pair<double, double> uniform_lagrange(const std::vector<double> &f, double du, double d) {
    const double s = d / du;
    const long i0 = (long)std::floor(s) - 1;
    if (i0 + 3 >= (long)f.size()) return {0.0, 0.0};
    auto at = [&](long i) { return f[(size_t)std::labs(i)]; };
    const double t = s - (double)i0;
    const double a = at(i0), b = at(i0 + 1), c = at(i0 + 2), e = at(i0 + 3);
    const double l0 = -(t - 1.0) * (t - 2.0) * (t - 3.0) / 6.0, l1 = t * (t - 2.0) * (t - 3.0) / 2.0,
                 l2 = -t * (t - 1.0) * (t - 3.0) / 2.0, l3 = t * (t - 1.0) * (t - 2.0) / 6.0;
    const double d0 = -((t - 2.0) * (t - 3.0) + (t - 1.0) * (t - 3.0) + (t - 1.0) * (t - 2.0)) / 6.0,
                 d1 = ((t - 2.0) * (t - 3.0) + t * (t - 3.0) + t * (t - 2.0)) / 2.0,
                 d2 = -((t - 1.0) * (t - 3.0) + t * (t - 3.0) + t * (t - 1.0)) / 2.0,
                 d3 = ((t - 1.0) * (t - 2.0) + t * (t - 2.0) + t * (t - 1.0)) / 6.0;
    return {a * l0 + b * l1 + c * l2 + e * l3, (a * d0 + b * d1 + c * d2 + e * d3) / du};
}

std::pair<double, double> table_read(const std::vector<double> &f, double du, double radius, double d) {
    if (d > radius) return {0.0, 0.0};
    return uniform_lagrange(f, du, d);
}

// the hard cut at q_max (no completion, no step): the band-limited weight fields of the egg-box probe
Completion hard_cut(const LogGrid &grid, const std::vector<double> &f, int l, double q_max) {
    const int n = grid.size();
    std::vector<double> q, w, tmp(n);
    q_mesh(0.0, q_max, q, w);
    Completion c;
    c.l = l, c.r_cut = INFINITY, c.width = 1.0;
    for (size_t k = 0; k < q.size(); k++) {
        for (int i = 0; i < n; i++) tmp[i] = f[i] * sph_bessel(l, q[k] * grid.r[i]) * grid.r[i] * grid.r[i];
        c.q.push_back(q[k]);
        c.c.push_back(w[k] * 4.0 * M_PI * grid.integrate(tmp));
    }
    return c;
}

std::vector<double> hard_cut_table(const std::vector<double> &f, int l, double h, double du, double radius) {
    return uniform_fill(du, radius, hard_cut(universal_grid(), f, l, M_PI / h));
}

// the lattice points r_p - R (spacing h) within radius of R
template <typename F> void for_lattice(double h, double radius, const double *r, F &&fn) {
    const long k = (long)std::ceil(radius / h) + 1;
    for (long i = -k; i <= k; i++)
        for (long j = -k; j <= k; j++)
            for (long l = -k; l <= k; l++) {
                const double d[3] = {i * h - r[0], j * h - r[1], l * h - r[2]};
                if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= radius * radius) fn(d);
            }
}

// max over R in {+-h/4}^3 of |grad_R E(R)| by central differences of step 2e-4 Bohr
template <typename E> double probe(double h, E &&e) {
    const double q = h / 4.0, step = 2e-4;
    double worst = 0.0;
    for (double sx : {-q, q})
        for (double sy : {-q, q})
            for (double sz : {-q, q}) {
                const double r[3] = {sx, sy, sz};
                double g2 = 0.0;
                for (int a = 0; a < 3; a++) {
                    double p[3] = {r[0], r[1], r[2]}, m[3] = {r[0], r[1], r[2]};
                    p[a] += step;
                    m[a] -= step;
                    const double ga = (e(p) - e(m)) / (2.0 * step);
                    g2 += ga * ga;
                }
                worst = std::max(worst, std::sqrt(g2));
            }
    return worst;
}

double norm3(const double *d) { return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]); }

// h^3 sum_p w(|r_p - R|) g(|r_p - R|) for a spherical table g with a spherical weight w on a uniform table
template <typename G> double spherical_probe(double h, double du, const std::vector<double> &w, double radius, G &&g) {
    const double h3 = h * h * h;
    return probe(h, [&](const double *r) {
        Neumaier s;
        for_lattice(h, radius, r, [&](const double *d) {
            const double x = norm3(d);
            s.add(uniform_lagrange(w, du, x).first * g(x));
        });
        return h3 * s.sum();
    });
}

std::vector<double> mean_vxc(const std::vector<double> *rho, const std::vector<double> &rho_pc) {
    std::vector<double> v(rho_pc.size());
    for (size_t i = 0; i < rho_pc.size(); i++) {
        const double c = 0.5 * rho_pc[i];
        double vu, vd;
        pw92(rho[0][i] + c, rho[1][i] + c, &vu, &vd);
        v[i] = 0.5 * (vu + vd);
    }
    return v;
}

int xc_domain_violations(const std::vector<double> *rho, const std::vector<double> &rho_pc) {
    int n = 0;
    for (size_t i = 0; i < rho_pc.size(); i++)
        if (rho_pc[i] < 0.0)
            for (int s = 0; s < 2; s++)
                if (rho[s][i] + 0.5 * rho_pc[i] < 0.0) n++;
    return n;
}

// the KB atom's input on the filtered tables at the log nodes: V_loc = V_LR + V_SR, beta_u = r^{l+1} t_l, rho_pc
KbInput filtered_input(const Table &t, int z) {
    const LogGrid &g = universal_grid();
    const int n = g.size();
    const double a = t.sigma_lr * std::sqrt(2.0), z_v = t.rec.z_v;
    KbInput in;
    in.z = z;
    in.v_loc.resize(n);
    for (int i = 0; i < n; i++) {
        const double d = g.r[i];
        const double v_lr = d < 1e-10 ? -z_v * 2.0 / (std::sqrt(M_PI) * a) : -z_v * std::erf(d / a) / d;
        in.v_loc[i] = v_lr + t.v_sr_log[i];
    }
    for (int k = 0; k < t.rec.n_projectors; k++) {
        KbProj p{t.rec.proj_l[k], t.rec.e_kb[k], std::vector<double>(n)};
        for (int i = 0; i < n; i++) p.beta[i] = std::pow(g.r[i], p.l + 1) * t.t_log[k][i];
        in.proj.push_back(std::move(p));
    }
    in.rho_pc = t.rho_pc_log;
    return in;
}

// T8(d) (filter.rs:573-611): [F_SR, F_NL, F_pc] of the filtered KB atom `atom`
std::array<double, 3> table_eggbox(const Table &t, const Partition &part, const Candidate &c, const KbAtom &atom,
                                   double h) {
    const LogGrid &g = universal_grid();
    const int n = g.size();
    const double du = t.du, h3 = h * h * h;
    std::vector<double> rho_v(n);
    for (int i = 0; i < n; i++) rho_v[i] = atom.rho[0][i] + atom.rho[1][i];
    const auto w_sr = hard_cut_table(rho_v, 0, h, du, t.r_sr);
    const double f_sr =
        spherical_probe(h, du, w_sr, t.r_sr, [&](double x) { return table_read(t.v_sr_u, du, t.r_sr, x).first; });
    double f_pc = 0.0;
    if (t.pc_radius > 0.0) {
        const auto w_pc = hard_cut_table(mean_vxc(atom.rho, t.rho_pc_log), 0, h, du, t.pc_radius);
        f_pc = spherical_probe(h, du, w_pc, t.pc_radius,
                               [&](double x) { return table_read(t.rho_pc_u, du, t.pc_radius, x).first; });
    }
    // E_NL = sum_k sum_m (f_l / (2l+1)) E_KB c_km^2, c_km = h^3 sum_p beta_klm phi_lm
    struct P {
        int k, l;
        double f, e_kb;
        std::vector<double> phi;
    };
    std::vector<P> projs;
    const ValenceConfig base = reference_config(part);
    for (int k = 0; k < t.rec.n_projectors; k++) {
        const int l = t.rec.proj_l[k];
        if (l > 1) throw Fail("T8: projectors above l = 1 are not probed");
        const double occ = base[l][0] + base[l][1];
        const std::vector<double> *u = nullptr;
        double f = 1.0;
        if (occ > 0.0) {
            u = &atom.u[l][0];
            f = occ;
        } else {
            for (const auto &ch : c.channels)
                if (ch.l == l) u = &ch.u;
        }
        std::vector<double> ur(n);
        for (int i = 0; i < n; i++) ur[i] = (*u)[i] / g.r[i];
        projs.push_back({k, l, f, t.rec.e_kb[k], hard_cut_table(ur, l, h, du, t.proj_radius[k])});
    }
    double f_nl = 0.0;
    if (!projs.empty())
        f_nl = probe(h, [&](const double *r) {
            Neumaier e;
            for (const P &p : projs) {
                const double radius = t.proj_radius[p.k];
                const double nrm2 = (2.0 * p.l + 1.0) / (4.0 * M_PI);
                Neumaier cm[3];
                for_lattice(h, radius, r, [&](const double *d) {
                    const double x = norm3(d);
                    const double tp = table_read(t.t_u[p.k], du, radius, x).first * uniform_lagrange(p.phi, du, x).first * nrm2;
                    if (p.l == 0) cm[0].add(tp);
                    else
                        for (int m = 0; m < 3; m++) cm[m].add(tp * d[m] * d[m]);  // real solid harmonics of l = 1: x, y, z
                });
                const double a = p.f / (2.0 * p.l + 1.0) * p.e_kb;
                for (int m = 0; m < 2 * p.l + 1; m++) {
                    const double sm = h3 * cm[m].sum();
                    e.add(a * sm * sm);
                }
            }
            return e.sum();
        });
    return {f_sr, f_nl, f_pc};
}

// T8(e), reported (filter.rs:528-561): the nonlinear partial-core force F_pc(R) = h^3 sum_p vbar(d) rho_pc'(|d|) d/|d|
double pc_alias(const Table &t, const KbAtom &atom, double h) {
    if (t.pc_radius <= 0.0) return 0.0;
    const double du = t.du, radius = t.pc_radius, q = h / 4.0, h3 = h * h * h;
    const auto ru = hard_cut_table(atom.rho[0], 0, h, du, radius), rd = hard_cut_table(atom.rho[1], 0, h, du, radius);
    double worst = 0.0;
    for (double sx : {-q, q})
        for (double sy : {-q, q})
            for (double sz : {-q, q}) {
                const double r[3] = {sx, sy, sz};
                Neumaier f[3];
                for_lattice(h, radius, r, [&](const double *d) {
                    const double x = norm3(d);
                    if (x == 0.0) return;
                    const double u = uniform_lagrange(ru, du, x).first, dn = uniform_lagrange(rd, du, x).first;
                    const auto pc = table_read(t.rho_pc_u, du, radius, x);
                    double vu, vd;
                    pw92(u + 0.5 * pc.first, dn + 0.5 * pc.first, &vu, &vd);
                    const double w = 0.5 * (vu + vd) * pc.second / x;
                    for (int a = 0; a < 3; a++) f[a].add(w * d[a]);
                });
                const double fx = h3 * f[0].sum(), fy = h3 * f[1].sum(), fz = h3 * f[2].sum();
                worst = std::max(worst, std::sqrt(fx * fx + fy * fy + fz * fz));
            }
    return worst;
}

struct FilterRow {
    double h, de, max_eig, t3_max;
    int clamped;
    std::array<double, 3> f_table;
    double f_pc_nl;
    bool passes() const {
        return std::fabs(de) < H_REQ_TOL && max_eig < H_REQ_TOL && t3_max < T3_DELTA_E && clamped == 0 &&
               f_table[0] + f_table[1] + f_table[2] < H_REQ_FORCE;
    }
    bool eggbox_certified() const { return f_table[0] + f_table[1] + f_pc_nl < H_REQ_FORCE; }
};

// the filtered re-run at one spacing (filter.rs:630-658)
FilterRow filter_row(const Partition &part, const PartialCore &pc, const Candidate &c, double h) {
    Table t;
    t.rec.z = part.z;
    t.rec.z_v = part.z_v;
    t.rec.n_projectors = (int)c.kb.projectors.size();
    for (int k = 0; k < t.rec.n_projectors; k++) {
        t.rec.proj_l[k] = c.kb.projectors[k].l;
        t.rec.e_kb[k] = c.kb.projectors[k].e_kb;
    }
    filter_tables(t, part, pc, c, h, true);
    const KbInput in = filtered_input(t, part.z);
    FilterRow row{h, INFINITY, INFINITY, INFINITY, 0, {INFINITY, INFINITY, INFINITY}, INFINITY};
    KbAtom fa;
    try {
        fa = solve_kb_atom(in, reference_config(part), true);
    } catch (const Fail &) {
        return row;
    }
    row.de = fa.energy - c.ref_atom.energy;
    row.max_eig = 0.0;
    for (const auto &ch : part.channels)
        if (ch.occupied) row.max_eig = std::max(row.max_eig, std::fabs(fa.eps[ch.l][0] - c.ref_atom.eps[ch.l][0]));
    row.clamped = xc_domain_violations(fa.rho, t.rho_pc_log);
    row.t3_max = 0.0;
    std::optional<KbAtom> hund;
    for (const auto &cfg : c.t3) {
        try {
            KbAtom pp = solve_kb_atom(in, cfg.pp, cfg.restricted);
            row.clamped += xc_domain_violations(pp.rho, t.rho_pc_log);
            row.t3_max = std::max(row.t3_max, std::fabs((pp.energy - fa.energy) - cfg.de_ae));
            if (cfg.name == "Hund") hund = std::move(pp);
        } catch (const Fail &) {
            row.t3_max = INFINITY;
        }
    }
    row.f_pc_nl = pc_alias(t, fa, h);
    if (hund) row.f_pc_nl = std::max(row.f_pc_nl, pc_alias(t, *hund, h));
    row.f_table = table_eggbox(t, part, c, fa, h);
    return row;
}

// h_req (filter.rs:661-669) and h_eb (671-676): the coarsest ladder spacing from which every finer row passes; NaN
// when even the finest fails
void h_req_of(const Partition &part, const PartialCore &pc, const Candidate &c, double &h_req, double &h_eb) {
    std::vector<FilterRow> rows;
    for (double h : H_LADDER) rows.push_back(filter_row(part, pc, c, h));
    h_req = h_eb = NAN;
    for (size_t k = 0; k < rows.size(); k++) {
        bool all = true, all_eb = true;
        for (size_t j = k; j < rows.size(); j++) {
            all = all && rows[j].passes();
            all_eb = all_eb && rows[j].eggbox_certified();
        }
        if (all && std::isnan(h_req)) h_req = rows[k].h;
        if (all_eb && std::isnan(h_eb)) h_eb = rows[k].h;
    }
}

thread_local std::string g_pseudo_error;  // per calling thread, as lira_last_error is per solver

}  // namespace

// ---- the dimer gate on f for Z <= 2 (crates/providers/src/pseudo.rs:798-1100, recipe.rs:152-176) ----
constexpr double DIMER_H = 0.15, DIMER_BOX = 12.0, DIMER_RE = 0.02, DIMER_OMEGA_REL = 0.05, DIMER_WALL_REL = 0.05;
constexpr double DIMER_BRACKET[2] = {1.0, 4.0}, DIMER_SIGMAS[4] = {1.0, 1.2, 1.4, 1.6};
constexpr double AMU_IN_ME = 1822.888486209;  // the unified atomic mass unit in electron masses (CODATA 2018)

// the most abundant isotope's mass (u) of the elements the gate applies to: 1H, 4He (AME2020)
double isotope_mass_amu(int z) { return z == 1 ? 1.00782503223 : 4.00260325413; }

// the lira_species view of a filtered table (pointers into t)
void species_of(const Table &t, lira_species &sp, lira_projector *proj) {
    std::memset(&sp, 0, sizeof(sp));
    sp.z = t.rec.z;
    sp.z_v = t.rec.z_v;
    sp.sigma_lr = t.sigma_lr;
    sp.du = t.du;
    sp.v_sr_u = t.v_sr_u.data();
    sp.n_v_sr = (int)t.v_sr_u.size();
    sp.r_sr = t.r_sr;
    const bool core = t.pc_radius > 0.0;
    sp.rho_pc_u = core ? t.rho_pc_u.data() : nullptr;
    sp.n_rho_pc = core ? (int)t.rho_pc_u.size() : 0;
    sp.pc_radius = t.pc_radius;
    for (int k = 0; k < t.rec.n_projectors; k++) {
        proj[k].l = t.rec.proj_l[k];
        proj[k].e_kb = t.rec.e_kb[k];
        proj[k].radius = t.proj_radius[k];
        proj[k].t_u = t.t_u[k].data();
        proj[k].n_t = (int)t.t_u[k].size();
    }
    sp.projectors = t.rec.n_projectors > 0 ? proj : nullptr;
    sp.n_projectors = t.rec.n_projectors;
    sp.r_core = t.r_core;
}

// One gate's solver: the bond force 1/2 (F_1,x - F_0,x) of a cold dimer solve (nucleus 1 at +R/2) through the C ABI
struct DimerSolver {
    lira_solver *s = nullptr;
    int solves = 0;
    DimerSolver() {
        const int n = 2 * (int)std::lround(DIMER_BOX / (2.0 * DIMER_H)) + 1;
        if (lira_solver_create(n, DIMER_H, LIRA_MIXED, &s) != LIRA_OK) throw Fail("dimer gate: no GPU solver");
    }
    ~DimerSolver() {
        if (s) lira_solver_destroy(s);
    }
    DimerSolver(const DimerSolver &) = delete;
    DimerSolver &operator=(const DimerSolver &) = delete;
    // electrons per spin: a singlet of 2 Z (all-electron) or 2 Z_v (pseudised) electrons
    double force(int z, double r, const lira_species *sp, double sigma) {
        if (lira_set_ionic(s, sp, sp ? 1 : 0) != LIRA_OK) throw Fail("dimer gate: lira_set_ionic failed");
        lira_nucleus nuc[2] = {{{-0.5 * r, 0.0, 0.0}, z}, {{0.5 * r, 0.0, 0.0}, z}};
        const int per_spin = sp ? (int)std::lround(sp->z_v) : z;
        lira_system sys{nuc, 2, per_spin, per_spin};
        lira_settings st;
        lira_default_settings(DIMER_H, &st);
        st.sigma = sigma;
        st.tol = 1e-6;
        st.eig_tol = 1e-7;
        st.occupations = LIRA_FERMI_DIRAC;
        st.kt = 5e-3;
        lira_scf_stats stats;
        const lira_status status = lira_solve(s, &sys, &st, 0, nullptr, nullptr, &stats);
        solves++;
        if (status != LIRA_OK)
            throw Fail(std::string("dimer gate: the solve at R = ") + std::to_string(r) + " failed: " + lira_last_error(s));
        double f[6];
        if (lira_forces(s, nuc, 2, sigma, f) != LIRA_OK) throw Fail("dimer gate: forces failed");
        return 0.5 * (f[3] - f[0]);
    }
    // the all-electron force at sigma -> 0: a + b s^2 + c s^3 + d s^4 through the four widths (qm::extrapolate_sigma)
    double all_electron(int z, double r) {
        std::vector<double> a(16), b(4), x;
        for (int k = 0; k < 4; k++) {
            const double sg = DIMER_SIGMAS[k] * DIMER_H;
            b[k] = force(z, r, nullptr, sg);
            a[k * 4 + 0] = 1.0;
            a[k * 4 + 1] = sg * sg;
            a[k * 4 + 2] = sg * sg * sg;
            a[k * 4 + 3] = sg * sg * sg * sg;
        }
        if (!Diis::lu_solve(a, b, 4, x)) throw Fail("dimer gate: the sigma extrapolation is singular");
        return x[0];
    }
};

struct DimerFit {
    bool bracketed = false;
    double r_e = NAN, omega = NAN;
    double lo = 0.0, hi = 0.0, f_lo = 0.0, f_hi = 0.0;
};

// R_e by bisection of the force's sign to 0.1 x 0.02 Bohr, then linear interpolation; k = -dF/dR by +-h (pseudo.rs:
// 912-948)
DimerFit dimer_fit(int z, double lo, double hi, const std::function<double(double)> &force) {
    DimerFit out;
    out.lo = lo, out.hi = hi;
    double a = lo, b = hi, fa = force(a), fb = force(b);
    out.f_lo = fa, out.f_hi = fb;
    if (!(fa > 0.0 && fb < 0.0)) return out;
    while (b - a > 0.1 * DIMER_RE) {
        const double m = 0.5 * (a + b), fm = force(m);
        if (fm > 0.0) a = m, fa = fm;
        else b = m, fb = fm;
    }
    out.bracketed = true;
    out.r_e = a + (b - a) * fa / (fa - fb);
    const double d = DIMER_H;
    const double k = -(force(out.r_e + d) - force(out.r_e - d)) / (2.0 * d);
    if (!(k > 0.0)) throw Fail("dimer gate: k <= 0 at R_e");
    const double mu = 0.5 * isotope_mass_amu(z) * AMU_IN_ME;
    out.omega = std::sqrt(k / mu);
    return out;
}

struct AeDimer {
    bool done = false;
    DimerFit fit;
    std::string error;
};

// The accepted ladder entry of one element, independent of h: the reference atom, its partition and partial core, the
// candidate and its record. Kept for the process (a generation runs every gate: seconds of host time), keyed by
// (Z, f_index, local); a refusal is kept as its reason.
struct Generated {
    AtomResult ref;
    std::unique_ptr<Partition> part;
    PartialCore pc;
    Candidate c;
    lira_pseudo_record rec{};
};

namespace {
std::mutex g_gen_mutex;
std::map<std::tuple<int, int, int, bool>, std::shared_ptr<Generated>> g_generated;
std::map<std::tuple<int, int, int, bool>, std::string> g_refused;
AeDimer g_ae_dimer[3];  // per Z <= 2: the all-electron dimer does not depend on the ladder entry
}  // namespace

// gate (optional): run on every entry that passes the atomic gates; false refuses the entry with the reason in why
using EntryGate = std::function<bool(Generated &, std::string &why)>;

std::shared_ptr<Generated> generate_entry(int z, int f_index, int local, const EntryGate &gate = nullptr) {
    const LogGrid &g = universal_grid();
    auto gen = std::make_shared<Generated>();
    gen->ref = ground_state_restricted(z);
    gen->part = std::make_unique<Partition>(partition(gen->ref));
    const Partition &part = *gen->part;
    const auto rho_c = density_of(part.core), rho_v = density_of(part.valence);
    gen->pc = partial_core(g, rho_c, rho_v);
    const auto tau = core_tail(g, rho_c, part.n_core);
    ElementData d{gen->ref, part, gen->pc};
    for (int l = 0; l <= part.l_max; l++) {
        int n_core_l = 0;
        for (const Shell *c : part.core)
            if (c->l == l) n_core_l++;
        std::vector<double> levels;
        for (int k = n_core_l;; k++) {
            auto b = bound_state(g, gen->ref.v, l, k, z, std::nullopt);
            if (!b) break;
            levels.push_back(b->e);
        }
        d.spectrum.push_back(std::move(levels));
    }
    d.ref_energy = solve_atom_full(z, AtomOpts{}).total;
    std::string tried;
    for (int fi = 0; fi < 4; fi++) {
        if (f_index >= 0 && fi != f_index) continue;
        for (int lo = 0; lo < 2; lo++) {
            if (local >= 0 && lo != local) continue;
            try {
                gen->c = candidate(d, tau, F_LADDER[fi], lo);
            } catch (const Fail &e) {
                tried += std::string(tried.empty() ? "" : "; ") + "f = " + std::to_string(F_LADDER[fi]).substr(0, 3) +
                         (lo == 0 ? " highest-channel local: " : " polynomial local: ") + e.what();
                continue;
            }
            const Candidate &c = gen->c;
            lira_pseudo_record &r = gen->rec;
            r.z = z;
            r.f = F_LADDER[fi];
            r.local = lo;
            r.local_l = c.kb.local_l;
            r.z_v = part.z_v;
            r.n_core = part.n_core;
            r.n_channels = (int)c.channels.size();
            for (int k = 0; k < r.n_channels && k < 4; k++) {
                r.channel_l[k] = c.channels[k].l;
                r.r_c[k] = c.channels[k].r_c;
                r.eps[k] = c.channels[k].eps;
                r.t5[k] = c.t5[k];
            }
            r.n_projectors = (int)c.kb.projectors.size();
            for (int k = 0; k < r.n_projectors && k < 3; k++) {
                r.proj_l[k] = c.kb.projectors[k].l;
                r.e_kb[k] = c.kb.projectors[k].e_kb;
                r.gss[k] = c.kb.projectors[k].gss;
                r.r_beta[k] = g.r[c.kb.projectors[k].i_beta];
            }
            r.r_nlc = gen->pc.r_nlc;
            r.r_core = gen->pc.r_core;
            r.r_cmp = c.kb.r_cmp;
            r.dimer_pp[0] = r.dimer_pp[1] = r.dimer_ae[0] = r.dimer_ae[1] = NAN;
            r.provisional = z <= 2 ? 1 : 0;
            r.h_req = r.h_eb = NAN;
            if (gate && z <= 2) {
                std::string why;
                if (!gate(*gen, why)) {
                    tried += std::string(tried.empty() ? "" : "; ") + "f = " + std::to_string(F_LADDER[fi]).substr(0, 3) +
                             (lo == 0 ? " highest-channel local: " : " polynomial local: ") + "dimer gate " + why;
                    continue;
                }
                r.provisional = 0;
            }
            h_req_of(part, gen->pc, gen->c, r.h_req, r.h_eb);  // T8 on the accepted entry (recipe.rs:166-170)
            return gen;
        }
    }
    throw Fail("Z = " + std::to_string(z) + ": no ladder entry passes: " + tried);
}

// the dimer gate of one ladder entry (recipe.rs:152-176 dimer_verdict, pseudo.rs:1018-1098 GridDimerGate::dimers)
bool dimer_gate(DimerSolver &solver, Generated &gen, std::string &why) {
    const Partition &part = *gen.part;
    const int z = part.z;
    double r_peak = 0.0;
    for (const auto &c : part.channels)
        if (c.occupied) r_peak = std::max(r_peak, c.r_peak);
    const double lo = DIMER_BRACKET[0] * r_peak, hi = DIMER_BRACKET[1] * r_peak;
    AeDimer &ae = g_ae_dimer[z];
    if (!ae.done) {
        ae.done = true;
        try {
            ae.fit = dimer_fit(z, lo, hi, [&](double r) { return solver.all_electron(z, r); });
        } catch (const Fail &e) {
            ae.error = e.what();
        }
    }
    if (!ae.error.empty()) {
        why = "could not run: " + ae.error;
        return false;
    }
    Table t;
    t.rec = gen.rec;
    filter_tables(t, part, gen.pc, gen.c, DIMER_H);
    lira_species sp;
    lira_projector proj[3];
    species_of(t, sp, proj);
    lira_pseudo_record &r = gen.rec;
    r.dimer_ae[0] = ae.fit.r_e;
    r.dimer_ae[1] = ae.fit.omega;
    try {
        if (ae.fit.bracketed) {
            const DimerFit pp = dimer_fit(z, lo, hi, [&](double x) { return solver.force(z, x, &sp, DIMER_H); });
            r.dimer_pp[0] = pp.r_e;
            r.dimer_pp[1] = pp.omega;
            if (!pp.bracketed) {
                why = "fail: the pseudised dimer is unbound where the all-electron one is bound";
                return false;
            }
            const double dr = pp.r_e - ae.fit.r_e, dw = pp.omega / ae.fit.omega - 1.0;
            if (!(std::fabs(dr) < DIMER_RE && std::fabs(dw) < DIMER_OMEGA_REL)) {
                why = "fail: PP R_e " + std::to_string(pp.r_e) + " Bohr, omega " + std::to_string(pp.omega) +
                      " a.u. vs AE R_e " + std::to_string(ae.fit.r_e) + ", omega " + std::to_string(ae.fit.omega);
                return false;
            }
            r.dimer = 1;
            return true;
        }
        // no all-electron force zero in the bracket: not applicable if the pseudised repulsive wall agrees at the outer end
        solver.force(z, lo, &sp, DIMER_H);
        const double p_hi = solver.force(z, hi, &sp, DIMER_H);
        const double wall = p_hi / ae.fit.f_hi - 1.0;
        if (std::fabs(wall) > DIMER_WALL_REL) {
            why = "fail: no AE minimum in the bracket and the pseudised repulsive wall differs by " +
                  std::to_string(100.0 * wall) + " %";
            return false;
        }
        r.dimer = 2;
        return true;
    } catch (const Fail &e) {
        why = std::string("could not run: ") + e.what();
        return false;
    }
}

Table *generate(int z, double h, int f_index, int local, bool gated = false) {
    if (z < 1 || z > Z_MAX) throw Fail("Z = " + std::to_string(z) + " is outside 1..18");
    gated = gated && z <= 2;
    const auto key = std::make_tuple(z, f_index, local, gated);
    std::shared_ptr<Generated> gen;
    {
        std::lock_guard<std::mutex> lock(g_gen_mutex);
        auto it = g_generated.find(key);
        if (it != g_generated.end()) gen = it->second;
        else {
            auto bad = g_refused.find(key);
            if (bad != g_refused.end()) throw Fail(bad->second);
            try {
                if (gated) {
                    DimerSolver solver;
                    gen = generate_entry(z, f_index, local,
                                         [&](Generated &g2, std::string &why) { return dimer_gate(solver, g2, why); });
                    gen->rec.dimer_solves = solver.solves;
                } else {
                    gen = generate_entry(z, f_index, local);
                }
            } catch (const Fail &e) {
                g_refused[key] = e.what();
                throw;
            }
            g_generated[key] = gen;
        }
    }
    auto t = std::make_unique<Table>();
    t->rec = gen->rec;
    t->rec.h = h;
    t->gen = gen;
    filter_tables(*t, *gen->part, gen->pc, gen->c, h);
    return t.release();
}

// ionic::check (crates/qm/src/ionic.rs:84-102): the table may be used at its own spacing h when it is not provisional,
// h <= h_req (or, with no h_req on the ladder, h finer than the whole ladder), and its filtered re-run at h passes T8
// (a)-(d). Returns the refusal, empty when the table passes.
std::string check_table(const Table &t) {
    const lira_pseudo_record &r = t.rec;
    const std::string el = "Z = " + std::to_string(r.z);
    if (r.provisional) return el + ": the pseudopotential is provisional (generated without the dimer gate)";
    const double finest = H_LADDER[4];
    const bool fine_enough = std::isnan(r.h_req) ? r.h < finest - 1e-12 : r.h <= r.h_req + 1e-12;
    if (!fine_enough)
        return el + ": h = " + std::to_string(r.h) + " Bohr is coarser than the pseudopotential allows (h_req = " +
               (std::isnan(r.h_req) ? std::string("none on the ladder") : std::to_string(r.h_req) + " Bohr") + ")";
    const FilterRow row = filter_row(*t.gen->part, t.gen->pc, t.gen->c, r.h);
    if (!row.passes())
        return el + ": the filtered pseudopotential fails the filtered gates at h = " + std::to_string(r.h) + ": dE " +
               std::to_string(row.de * 1e3) + " mHa, d eps " + std::to_string(row.max_eig * 1e3) + " mHa, T3 " +
               std::to_string(row.t3_max * 1e3) + " mHa, XC clamp at " + std::to_string(row.clamped) +
               " points, table egg-box force " + std::to_string(row.f_table[0] + row.f_table[1] + row.f_table[2]);
    return "";
}

const std::string &last_error() { return g_pseudo_error; }
void set_error(const std::string &s) { g_pseudo_error = s; }

}  // namespace radial
}  // namespace lira

struct lira_pseudo : lira::radial::Table {};

using namespace lira;

// the spacings a table is filtered for: its uniform tables hold radius * 256 / h points
static bool spacing_ok(double h) { return h >= 0.02 && h <= 1.0; }

extern "C" lira_status lira_pseudo_generate(int z, double h, int f_index, int local, lira_pseudo **out) {
    if (!out || !spacing_ok(h) || f_index < -1 || f_index > 3 || local < -1 || local > 1) return LIRA_INVALID;
    *out = nullptr;
    try {
        radial::Table *t = radial::generate(z, h, f_index, local);
        *out = static_cast<lira_pseudo *>(t);
        radial::set_error("");
        return LIRA_OK;
    } catch (const std::exception &e) {
        radial::set_error(e.what());
        return LIRA_NUMERICAL;
    }
}

extern "C" lira_status lira_pseudo_generate_gated(int z, double h, lira_pseudo **out) {
    if (!out || !spacing_ok(h)) return LIRA_INVALID;
    *out = nullptr;
    try {
        radial::Table *t = radial::generate(z, h, -1, -1, true);
        *out = static_cast<lira_pseudo *>(t);
        radial::set_error("");
        return LIRA_OK;
    } catch (const std::exception &e) {
        radial::set_error(e.what());
        return LIRA_NUMERICAL;
    }
}

extern "C" const char *lira_pseudo_error(void) { return radial::last_error().c_str(); }

extern "C" lira_status lira_pseudo_check(const lira_pseudo *p) {
    if (!p) return LIRA_INVALID;
    try {
        const std::string why = radial::check_table(*p);
        radial::set_error(why);
        return why.empty() ? LIRA_OK : LIRA_NUMERICAL;
    } catch (const std::exception &e) {
        radial::set_error(e.what());
        return LIRA_NUMERICAL;
    }
}

extern "C" lira_status lira_pseudo_record_of(const lira_pseudo *p, lira_pseudo_record *rec) {
    if (!p || !rec) return LIRA_INVALID;
    *rec = p->rec;
    return LIRA_OK;
}

extern "C" lira_status lira_pseudo_species(const lira_pseudo *p, lira_species *sp, lira_projector *proj, int proj_cap) {
    if (!p || !sp || (p->rec.n_projectors > 0 && (!proj || proj_cap < p->rec.n_projectors))) return LIRA_INVALID;
    std::memset(sp, 0, sizeof(*sp));
    sp->z = p->rec.z;
    sp->z_v = p->rec.z_v;
    sp->sigma_lr = p->sigma_lr;
    sp->du = p->du;
    sp->v_sr_u = p->v_sr_u.data();
    sp->n_v_sr = (int)p->v_sr_u.size();
    sp->r_sr = p->r_sr;
    const bool core = p->pc_radius > 0.0;
    sp->rho_pc_u = core ? p->rho_pc_u.data() : nullptr;
    sp->n_rho_pc = core ? (int)p->rho_pc_u.size() : 0;
    sp->pc_radius = p->pc_radius;
    for (int k = 0; k < p->rec.n_projectors; k++) {
        proj[k].l = p->rec.proj_l[k];
        proj[k].e_kb = p->rec.e_kb[k];
        proj[k].radius = p->proj_radius[k];
        proj[k].t_u = p->t_u[k].data();
        proj[k].n_t = (int)p->t_u[k].size();
    }
    sp->projectors = p->rec.n_projectors > 0 ? proj : nullptr;
    sp->n_projectors = p->rec.n_projectors;
    sp->r_core = p->r_core;
    return LIRA_OK;
}

extern "C" void lira_pseudo_free(lira_pseudo *p) { delete static_cast<radial::Table *>(p); }
