// spimc_oracle.cpp -- CPU restatement of simpimc's action-evaluation and
// sampling hot path.
//
// TEST INFRASTRUCTURE.  This file is the checker, never the product: only the
// tests, the smoke check and the CPU-baseline leg of the benchmark load it.
// The product library must not include, link or call anything here.
//
// Every routine restates the reference routine cited beside it (paths relative
// to the simpimc source tree) with the same loop order, minimum-image
// convention, window/level conventions, Metropolis draw order and (single-
// thread) summation order.  Departures, all result-preserving:
//   * Per-bead rho_k caches (bead_class.h:20-21,125-133) are not stored; the OLD
//     phase factor of a bead is recomputed by the same CalcC recurrence from the
//     same position, which yields the same bits as the cached copy.
//   * Splines: einspline's NATURAL cubic B-splines are replaced by the natural
//     cubic spline through the same knots (the interpolant is unique), evaluated
//     in second-derivative form.  Outside the knot range the boundary interval's
//     cubic is extended (einspline's out-of-range behaviour is parity unpinned).
//   * The scaffold RNG is unavailable; draws come either from a replay tape
//     (post-transform values, reference call order) or from this file's own
//     xoshiro256** stream.
//   * One species, homologous pair action, max_level = 0 (the configurations
//     this engine targets).
#include "../include/spimc.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cplx = std::complex<double>;

thread_local std::string g_err;
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// ---------------------------------------------------------------------------
// Natural cubic spline (stands in for einspline create_*_1d_d with
// BCtype_d{NATURAL, NATURAL}; call sites free_spline_class.h:65-80,
// bare_pair_action_class.h:47-67, david_pair_action_class.h:260-293).
// ---------------------------------------------------------------------------
struct NatSpline {
    std::vector<double> x, y, m2;  // knots, values, second derivatives
    bool uniform = false;
    double x0 = 0., inv_dx = 0.;

    void build(const double* xs, const double* ys, uint32_t n, bool is_uniform) {
        x.assign(xs, xs + n);
        y.assign(ys, ys + n);
        m2.assign(n, 0.);
        uniform = is_uniform;
        x0 = xs[0];
        inv_dx = (n > 1) ? (n - 1) / (xs[n - 1] - xs[0]) : 0.;
        if (n < 3) return;
        // Thomas algorithm on the interior second derivatives, m2[0]=m2[n-1]=0
        std::vector<double> c(n, 0.), d(n, 0.);
        for (uint32_t i = 1; i + 1 < n; ++i) {
            double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
            double a = h0, b = 2. * (h0 + h1), cc = h1;
            double rhs = 6. * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
            double denom = b - a * c[i - 1];
            c[i] = cc / denom;
            d[i] = (rhs - a * d[i - 1]) / denom;
        }
        for (uint32_t i = n - 2; i >= 1; --i) {
            m2[i] = d[i] - c[i] * m2[i + 1];
            if (i == 1) break;
        }
    }

    // einspline general_grid_reverse_map: largest i with x[i] <= t (binary search);
    // uniform grids index directly.  Clamped to the last interval.
    uint32_t locate(double t) const {
        const uint32_t n = (uint32_t)x.size();
        long i;
        if (uniform) {
            i = (long)std::floor((t - x0) * inv_dx);
        } else if (t <= x[0]) {
            i = 0;
        } else if (t >= x[n - 1]) {
            i = n - 1;
        } else {
            long lo = 0, hi = n - 1;
            while (hi - lo >= 2) {
                long mid = (hi + lo) >> 1;
                if (x[mid] > t)
                    hi = mid;
                else
                    lo = mid;
            }
            i = lo;
        }
        if (i < 0) i = 0;
        if (i > (long)n - 2) i = (long)n - 2;
        return (uint32_t)i;
    }

    double eval(double t) const {
        uint32_t i = locate(t);
        double h = x[i + 1] - x[i];
        double a = (x[i + 1] - t) / h, b = (t - x[i]) / h;
        return a * y[i] + b * y[i + 1] + ((a * a * a - a) * m2[i] + (b * b * b - b) * m2[i + 1]) * (h * h) / 6.;
    }

    double deriv(double t) const {
        uint32_t i = locate(t);
        double h = x[i + 1] - x[i];
        double a = (x[i + 1] - t) / h, b = (t - x[i]) / h;
        return (y[i + 1] - y[i]) / h - (3. * a * a - 1.) / 6. * h * m2[i] + (3. * b * b - 1.) / 6. * h * m2[i + 1];
    }
};

// ---------------------------------------------------------------------------
// Tensor-product natural cubic spline on general grids (stands in for einspline
// create_NUBspline_2d_d with BCtype_d{NATURAL, NATURAL} on both axes,
// ilkka_pair_action_class.h:276-278, 322-324).  Evaluated as a spline of
// splines: one natural spline along x per y knot, whose values at x are then
// interpolated by a natural spline along y.
// ---------------------------------------------------------------------------
struct NatSpline2 {
    std::vector<double> x, y;
    std::vector<NatSpline> rows;  // rows[j]: along x through vals[j][*]

    void build(const double* xs, uint32_t nx, const double* ys, uint32_t ny, const double* vals) {
        x.assign(xs, xs + nx);
        y.assign(ys, ys + ny);
        rows.resize(ny);
        for (uint32_t j = 0; j < ny; ++j) rows[j].build(xs, vals + (size_t)j * nx, nx, false);
    }
    // value and, when gx/gy are non-null, the partial derivatives at (xv, yv)
    double eval(double xv, double yv, double* gx = nullptr, double* gy = nullptr) const {
        const uint32_t ny = (uint32_t)y.size();
        std::vector<double> g(ny), gd(ny);
        for (uint32_t j = 0; j < ny; ++j) {
            g[j] = rows[j].eval(xv);
            if (gx) gd[j] = rows[j].deriv(xv);
        }
        NatSpline col;
        col.build(y.data(), g.data(), ny, false);
        if (gy) *gy = col.deriv(yv);
        if (gx) {
            NatSpline dcol;
            dcol.build(y.data(), gd.data(), ny, false);
            *gx = dcol.eval(yv);
        }
        return col.eval(yv);
    }
};

// ---------------------------------------------------------------------------
// FreeSpline (actions/free_spline_class.h:25-122)
// ---------------------------------------------------------------------------
struct FreeSpline {
    double i4lt = 0., i4ltt = 0.;
    NatSpline image, dimage;
    bool has_dtau = false;

    void build(double L, int n_images, double lambda, double tau, bool use_tau_derivative) {
        // free_spline_class.h:27-63
        i4lt = 1. / (4. * lambda * tau);
        i4ltt = 1. / (4. * lambda * tau * tau);
        double t_l = (L == 0.) ? 1000. : L;
        const uint32_t num = 10000;
        double start = -t_l / 2., end = t_l / 2.;
        double dr = (end - start) / (num - 1);
        std::vector<double> xs(num), ia(num, 0.), dia(num, 0.);
        for (uint32_t i = 0; i < num; ++i) {
            double r = start + i * dr;
            xs[i] = r;
            double r_2 = r * r;
            double r_2_i4lt = r_2 * i4lt;
            for (int image_i = 1; image_i <= n_images; ++image_i) {
                double r_p = r + image_i * t_l;
                double r_p_2_i4lt = r_p * r_p * i4lt;
                double d_p = r_2_i4lt - r_p_2_i4lt;
                double e_p = std::exp(d_p);
                double r_m = r - image_i * t_l;
                double r_m_2_i4lt = r_m * r_m * i4lt;
                double d_m = r_2_i4lt - r_m_2_i4lt;
                double e_m = std::exp(d_m);
                ia[i] += e_p + e_m;
                if (use_tau_derivative) dia[i] += (d_p * e_p + d_m * e_m) / tau;
            }
            if (use_tau_derivative) dia[i] = dia[i] / (1. + ia[i]);
            ia[i] = -std::log1p(ia[i]);
        }
        image.build(xs.data(), ia.data(), num, true);
        has_dtau = use_tau_derivative;
        if (use_tau_derivative) dimage.build(xs.data(), dia.data(), num, true);
    }

    // GetLogRhoFree (free_spline_class.h:76-84)
    double log_rho_free(const double* r) const {
        double tot = 0.;
        for (int d = 0; d < 3; ++d) tot -= image.eval(r[d]);
        return tot - (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) * i4lt;
    }
    // GetDLogRhoFreeDTau (free_spline_class.h:92-100)
    double dlog_rho_free_dtau(const double* r) const {
        double tot = 0.;
        for (int d = 0; d < 3; ++d) tot -= dimage.eval(r[d]);
        return tot - (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) * i4ltt;
    }
    // gradient of -log rho_free wrt r (primitive part + image spline slope)
    void grad_neg_log(const double* r, double* g) const {
        for (int d = 0; d < 3; ++d) g[d] = image.deriv(r[d]) + 2. * r[d] * i4lt;
    }
};

// ---------------------------------------------------------------------------
// xoshiro256** for free-running chains (the reference's scaffold RNG is not
// available; transforms below are this engine's own definitions).
// ---------------------------------------------------------------------------
struct Xoshiro {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& z) {
        z += 0x9e3779b97f4a7c15ull;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
        return x ^ (x >> 31);
    }
    void seed(uint64_t v) {
        for (auto& w : s) w = splitmix(v);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double unif() { return ((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }  // (0,1]
    uint32_t index(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    double normal() {
        double u1 = unif(), u2 = unif();
        return std::sqrt(-2. * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// ---------------------------------------------------------------------------
// One replica: Path + Species + Kinetic + PairAction + Bisect + DisplaceParticle
// + Energy / StructureFactor / PairCorrelation.
// ---------------------------------------------------------------------------
struct Oracle {
    // Path (path_class.h:25-69)
    uint32_t N = 0, M = 0;
    double L = 0., iL = 0., vol = 1., beta = 0., tau = 0., lambda = 0.;
    int mode = SPIMC_MODE_NEW;
    std::vector<double> r, rc;  // [M][N][3] NEW and OLD positions (Bead::r_, r_c_)

    // KSpace (k_space_class.h:33-94)
    uint32_t nk = 0;
    std::vector<int> kn;  // [nk][3]
    std::vector<double> kmag;
    int max_index = 0;
    double box = 0.;
    std::vector<cplx> rho, rhoc;  // [M][nk] Species::rho_k_, rho_k_c_
    bool need_update_rho_k = true;

    // Kinetic (kinetic_class.h:16-32)
    int kin_n_images = 0;
    double kin_i4lt = 0., kin_dbeta_const = 0.;
    std::vector<FreeSpline> kin_splines;

    // PairAction (pair_action_class.h:205-238)
    int pair_type = SPIMC_PAIR_NONE;
    bool use_long_range = false;
    uint32_t max_level = 0;
    // Bare
    bool is_coulomb = false;
    NatSpline v_r, v_long_r;
    double r_v_min = 0., r_v_max = 0., r_v_long_min = 0., r_v_long_max = 0.;
    // David
    uint32_t n_order = 0, n_val = 0;
    std::vector<NatSpline> u_kj, du_kj;  // n_val+1 components each
    double g_start = 0., g_end = 0.;
    // Ilkka (ilkka_pair_action_class.h:9-30): U and dU/dbeta tensor splines, the
    // long-range parts of U and dU/dbeta on their r grids; v_r and v_long_r above
    // hold its potential
    NatSpline2 u_xy, du_xy;
    NatSpline u_long_r_s, du_long_r_s;
    double r_u_min = 0., r_u_max = 0., r_du_min = 0., r_du_max = 0.;
    // long range
    std::vector<double> v_long_k, u_long_k, du_long_k;
    double v_long_k_0 = 0., v_long_r_0 = 0., du_long_k_0 = 0., du_long_r_0 = 0.;

    // Bisect (bisect_class.h:167-190) and DisplaceParticle (displace_particle_class.h:93)
    uint32_t n_level = 0, n_bisect_beads = 0;
    int bis_n_images = 0;
    std::vector<FreeSpline> bis_splines;
    double step_size = 0.;
    uint64_t counters[4] = {0, 0, 0, 0};
    uint64_t bead_updates = 0;

    // draws
    Xoshiro rng;
    std::vector<double> tape;
    uint64_t cursor = 0;
    bool use_tape = false;

    // proposal bookkeeping for the interface entry points
    bool have_proposal = false;
    uint32_t prop_p = 0, prop_first = 0, prop_n = 0;

    // observables
    int energy_estimator = SPIMC_ENERGY_THERMAL;
    uint32_t virial_window = 0;
    double sk_k_cut = 0.;
    double gr_r_min = 0., gr_r_max = 0., gr_dr = 0., gr_d_ir = 0.;
    uint32_t gr_n_r = 100;
    double e_acc[2] = {0., 0.};
    std::vector<double> sk_acc, gr_acc;
    uint64_t n_measure[3] = {0, 0, 0};

    // Path topology.  sigma[p] is the particle whose slice-0 bead follows the
    // slice-(M-1) bead of particle p: Bead::next_ across the wrap, where
    // PermBisect::AssignParticleLabels (perm_bisect_class.h:47-56) leaves every
    // permutation.  Within slices 0..M-1 a bead is followed by the same label.
    // During a permuting attempt the NEW-mode path is also relinked at slice
    // att_e: bead (x, att_e) is followed by the bead that followed (cyc[x],
    // att_e) (PermBisect::PermuteBeads, perm_bisect_class.h:32-44).
    std::vector<uint32_t> sigma, isigma;
    bool att_on = false;
    uint32_t att_e = 0;
    std::vector<uint32_t> cyc;
    // PermBisect (perm_bisect_class.h:20-29, perm_bisect_iterative_class.h)
    double perm_log_epsilon = std::log(1.e-100);
    double perm_epsilon = 1.e-100;
    // PermBisectTable's n_perm_part (perm_bisect_table_class.h:235); 0 = not configured
    uint32_t table_k = 0;
    // the final table check: 0 the reference's perm_tot_0 / perm_tot_1
    // (perm_bisect_table_class.h:223); 1 perm_tot_0 / (perm_tot_1 w), w the
    // selected entry's weight, which detailed balance asks for when n_perm_part
    // < N (the reverse move selects the inverse entry, of weight 1 / w, from the
    // new table)
    int table_balance = 0;
    uint32_t perm_n_level = 0;  // the permuting move's own level (its Bisect base, bisect_class.h:168)
    std::vector<uint64_t> perm_attempt, perm_accept;  // per cycle length - 1
    uint64_t perm_counters[2] = {0, 0};               // Move::n_attempt, n_accept

    // ---- helpers ------------------------------------------------------------
    uint32_t bead_loop(uint32_t b) const { return b >= M ? b - M : b; }  // species_class.h:101
    struct Bd {
        uint32_t p, b;  // label, slice
        bool operator==(const Bd& o) const { return p == o.p && b == o.b; }
        bool operator!=(const Bd& o) const { return !(*this == o); }
    };
    // Bead::GetNextBead(1) in the current mode
    Bd next1(Bd x) const {
        const uint32_t y = (att_on && mode == SPIMC_MODE_NEW && x.b == att_e) ? cyc[x.p] : x.p;
        return x.b + 1 == M ? Bd{sigma[y], 0} : Bd{y, x.b + 1};
    }
    Bd next(Bd x, uint32_t n) const {
        while (n--) x = next1(x);
        return x;
    }
    // Bead::GetPrevBead(1) outside an attempt
    Bd prev1(Bd x) const { return x.b == 0 ? Bd{isigma[x.p], M - 1} : Bd{x.p, x.b - 1}; }
    double* R(Bd x) { return R(x.p, x.b); }
    double* R(uint32_t p, uint32_t b) { return (mode ? r.data() : rc.data()) + 3 * ((size_t)bead_loop(b) * N + p); }
    double* Rm(int m, uint32_t p, uint32_t b) { return (m ? r.data() : rc.data()) + 3 * ((size_t)bead_loop(b) * N + p); }
    std::vector<cplx>& rho_mode() { return mode ? rho : rhoc; }  // species_class.h:428

    // Path::PutInBox (path_class.h:128-131)
    void put_in_box(double* v) const {
        for (int d = 0; d < 3; ++d) v[d] -= std::nearbyint(v[d] * iL) * L;
    }
    // Path::Dr (path_class.h:108-112): r0 - r1 put in the box
    void dr(const double* r0, const double* r1, double* out) const {
        for (int d = 0; d < 3; ++d) out[d] = r0[d] - r1[d];
        put_in_box(out);
    }
    static double mag(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

    // Path::DrDrpDrrp (path_class.h:137-150)
    void dr_drp_drrp(uint32_t b_i, uint32_t b_j, uint32_t p0, uint32_t p1, double& r_mag, double& rp_mag,
                     double& rrp_mag) {
        double rv[3], rpv[3], rrp[3];
        const double *a0 = R(p0, b_i), *a1 = R(p1, b_i), *c0 = R(p0, b_j), *c1 = R(p1, b_j);
        for (int d = 0; d < 3; ++d) {
            rv[d] = a1[d] - a0[d];
            rpv[d] = c1[d] - c0[d];
        }
        for (int d = 0; d < 3; ++d) {
            rv[d] -= std::nearbyint(rv[d] * iL) * L;
            rpv[d] += std::nearbyint((rv[d] - rpv[d]) * iL) * L;
        }
        for (int d = 0; d < 3; ++d) rrp[d] = rv[d] - rpv[d];
        for (int d = 0; d < 3; ++d) rrp[d] -= std::nearbyint(rrp[d] * iL) * L;
        r_mag = mag(rv);
        rp_mag = mag(rpv);
        rrp_mag = mag(rrp);
    }

    // KSpace::CalcC + Bead::CalcRhoK (k_space_class.h:83-94, bead_class.h:125-133)
    void calc_c(const double* pos, std::vector<cplx>& c) const {
        const int w = 2 * max_index + 1;
        c.assign(3 * w, cplx(0., 0.));
        for (int d = 0; d < 3; ++d) {
            double phi = pos[d] * box;
            cplx t(std::cos(phi), std::sin(phi));
            cplx* cd = &c[d * w];
            cd[max_index] = 1.;
            for (int k = 1; k <= max_index; ++k) {
                cd[max_index + k] = t * cd[max_index + k - 1];
                cd[max_index - k] = std::conj(cd[max_index + k]);
            }
        }
    }
    void calc_rho_k(const double* pos, std::vector<cplx>& out) const {
        std::vector<cplx> c;
        calc_c(pos, c);
        const int w = 2 * max_index + 1;
        out.resize(nk);
        for (uint32_t k = 0; k < nk; ++k) {
            cplx f = 1.;
            for (int d = 0; d < 3; ++d) f *= c[d * w + max_index + kn[3 * k + d]];
            out[k] = f;
        }
    }

    // Species::InitRhoK (species_class.h:380-403), from the NEW positions
    void init_rho_k() {
        rho.assign((size_t)M * nk, cplx(0., 0.));
        std::vector<cplx> bk;
        for (uint32_t b = 0; b < M; ++b)
            for (uint32_t p = 0; p < N; ++p) {
                calc_rho_k(Rm(1, p, b), bk);
                for (uint32_t k = 0; k < nk; ++k) rho[(size_t)b * nk + k] += bk[k];
            }
        rhoc = rho;
    }

    // Species::UpdateRhoK (species_class.h:406-425)
    void update_rho_k(uint32_t b0, uint32_t b1, const std::vector<uint32_t>& parts, uint32_t level) {
        uint32_t skip = 1u << level;
        for (uint32_t b = b0; b < b1; b += skip)
            for (uint32_t k = 0; k < nk; ++k) rho[(size_t)bead_loop(b) * nk + k] = rhoc[(size_t)bead_loop(b) * nk + k];
        std::vector<cplx> nw, od;
        for (uint32_t p : parts)
            for (uint32_t b = b0; b < b1; b += skip) {
                calc_rho_k(Rm(1, p, b), nw);
                calc_rho_k(Rm(0, p, b), od);  // the bead's cached OLD copy
                cplx* rb = &rho[(size_t)bead_loop(b) * nk];
                for (uint32_t k = 0; k < nk; ++k) rb[k] += nw[k] - od[k];
            }
        need_update_rho_k = false;
    }

    static double cmag2(const cplx& a, const cplx& b) { return a.real() * b.real() + a.imag() * b.imag(); }

    // ---- pair action --------------------------------------------------------
    static void set_limits(double r_min, double r_max, double& r0, double& r1) {  // pair_action_class.h:32-42
        if (r0 > r_max)
            r0 = r_max;
        else if (r0 < r_min)
            r0 = r_min;
        if (r1 > r_max)
            r1 = r_max;
        else if (r1 < r_min)
            r1 = r_min;
    }

    // BarePairAction::CalcV (bare_pair_action_class.h:99-123)
    double bare_calc_v(double r0, double r1) const {
        set_limits(r_v_min, r_v_max, r0, r1);
        double v = 0.;
        if (is_coulomb) {
            v += (0.5 / r0) + (0.5 / r1);
        } else {
            v += 0.5 * v_r.eval(r0);
            v += 0.5 * v_r.eval(r1);
        }
        if (use_long_range) {
            set_limits(r_v_long_min, r_v_long_max, r0, r1);
            v -= 0.5 * v_long_r.eval(r0);
            v -= 0.5 * v_long_r.eval(r1);
        }
        return v;
    }

    // DavidPairAction::CalcV (david_pair_action_class.h:42-57)
    double david_calc_v(double r0, double r1) const {
        set_limits(g_start, g_end, r0, r1);
        return 0.5 * (u_kj[0].eval(r0) + u_kj[0].eval(r1));
    }

    // DavidPairAction::CalcU / CalcdUdBeta (david_pair_action_class.h:63-102,126-168)
    double david_calc_u(double r0, double r1, double s, const std::vector<NatSpline>& tab, bool add_v) const {
        double q = 0.5 * (r0 + r1);
        double z = r0 - r1;
        double r_min = g_start, r_max = g_end;
        set_limits(r_min, r_max, r0, r1);
        double u = 0.5 * (tab[1].eval(r0) + tab[1].eval(r1));
        if (add_v) u += 0.5 * (tab[0].eval(r0) + tab[0].eval(r1));
        if (s > 0.0 && q < r_max) {
            double z_2 = z * z, s_2 = s * s, i_s_2 = 1. / s_2, s_2_k = s_2;
            for (uint32_t k = 1; k <= n_order; ++k) {
                double z_2_j = 1., cur_s = s_2_k;
                for (uint32_t j = 0; j <= k; ++j) {
                    double cof = tab[k * (k + 1) / 2 + (j + 1)].eval(q);
                    u += cof * z_2_j * cur_s;
                    z_2_j *= z_2;
                    cur_s *= i_s_2;
                }
                s_2_k *= s_2;
            }
        }
        return u;
    }

    // IlkkaPairAction::CalcU / CalcdUdBeta (ilkka_pair_action_class.h:83-107,126-150)
    double ilkka_calc_u(double r0, double r1, double s, const NatSpline2& t, const NatSpline& lr, double lo,
                        double hi) const {
        double q = 0.5 * (r0 + r1);
        double x = q + 0.5 * s;
        double y = q - 0.5 * s;
        double u = t.eval(x, y);
        if (use_long_range) {
            set_limits(lo, hi, r0, r1);
            u -= 0.5 * lr.eval(r0);
            u -= 0.5 * lr.eval(r1);
        }
        return u;
    }

    // CalcU of the selected family at `level` (max_level = 0, so level = 0 here)
    double calc_u(double r0, double r1, double s, uint32_t level) const {
        if (pair_type == SPIMC_PAIR_BARE) {
            uint32_t skip = 1u >> level;  // bare_pair_action_class.h:147 (sic)
            double level_tau = skip * tau;
            return level_tau * bare_calc_v(r0, r1);
        }
        if (pair_type == SPIMC_PAIR_ILKKA) return ilkka_calc_u(r0, r1, s, u_xy, u_long_r_s, r_u_min, r_u_max);
        return david_calc_u(r0, r1, s, u_kj, false);
    }
    double calc_du_dbeta(double r0, double r1, double s) const {
        if (pair_type == SPIMC_PAIR_BARE) return bare_calc_v(r0, r1);  // bare_pair_action_class.h:175-177
        if (pair_type == SPIMC_PAIR_ILKKA) return ilkka_calc_u(r0, r1, s, du_xy, du_long_r_s, r_du_min, r_du_max);
        // david_pair_action_class.h:126-168: du = 0.5(du_kj[1](r)+du_kj[1](r')) + V + off-diagonal
        double q = 0.5 * (r0 + r1), z = r0 - r1;
        double r_max = g_end;
        double a = r0, b = r1;
        set_limits(g_start, g_end, a, b);
        double v = 0.5 * (du_kj[0].eval(a) + du_kj[0].eval(b));
        double du = 0.5 * (du_kj[1].eval(a) + du_kj[1].eval(b));
        du += v;
        if (s > 0.0 && q < r_max) {
            double z_2 = z * z, s_2 = s * s, i_s_2 = 1. / s_2, s_2_k = s_2;
            for (uint32_t k = 1; k <= n_order; ++k) {
                double z_2_j = 1., cur_s = s_2_k;
                for (uint32_t j = 0; j <= k; ++j) {
                    double cof = du_kj[k * (k + 1) / 2 + j + 1].eval(q);
                    du += cof * z_2_j * cur_s;
                    z_2_j *= z_2;
                    cur_s *= i_s_2;
                }
                s_2_k *= s_2;
            }
        }
        return du;
    }
    // CalcV; IlkkaPairAction::CalcV (ilkka_pair_action_class.h:34-54) is the Bare form
    double calc_v(double r0, double r1) const {
        return (pair_type == SPIMC_PAIR_BARE || pair_type == SPIMC_PAIR_ILKKA) ? bare_calc_v(r0, r1) : david_calc_v(r0, r1);
    }

    // CalcULong (bare_pair_action_class.h:153-172, david_pair_action_class.h:105-123)
    double calc_u_long(uint32_t b0, uint32_t b1, uint32_t level) {
        std::vector<cplx>& rk = rho_mode();
        uint32_t skip = 1u << level;
        double tot = 0.;
        const std::vector<double>& coef = (pair_type == SPIMC_PAIR_BARE) ? v_long_k : u_long_k;
        const uint32_t nb = (b1 > b0) ? (b1 - b0 + skip - 1) / skip : 0;
        const long nkl = nk;
#pragma omp parallel for collapse(2) reduction(+ : tot)
        for (long k = 0; k < nkl; ++k)
            for (uint32_t t = 0; t < nb; ++t) {
                uint32_t b = bead_loop(b0 + t * skip);
                const cplx& a = rk[(size_t)b * nk + k];
                tot += coef[k] * cmag2(a, a);
            }
        if (pair_type == SPIMC_PAIR_BARE) return skip * tau * tot;
        return tot;
    }
    // CalcVLong / CalcdUdBetaLong (bare_pair_action_class.h:126-143,180-182;
    // david_pair_action_class.h:42-60,171-188)
    double calc_whole_long(bool dbeta) {
        std::vector<cplx>& rk = rho_mode();
        const bool du = (pair_type == SPIMC_PAIR_DAVID || pair_type == SPIMC_PAIR_ILKKA) && dbeta;
        const std::vector<double>& coef = du ? du_long_k : v_long_k;
        double tot = 0.;
        const long nkl = nk;
#pragma omp parallel for collapse(2) reduction(+ : tot)
        for (long k = 0; k < nkl; ++k)
            for (uint32_t b = 0; b < M; ++b) {
                const cplx& a = rk[(size_t)b * nk + k];
                tot += coef[k] * cmag2(a, a);
            }
        if (du) return tot + du_long_k_0 + du_long_r_0;
        return tot + v_long_k_0 + v_long_r_0;
    }

    // PairAction::GetAction (pair_action_class.h:267-302), homologous, one species
    double pair_get_action(uint32_t b0, uint32_t b1, const std::vector<uint32_t>& parts, uint32_t level) {
        if (level > max_level || pair_type == SPIMC_PAIR_NONE) return 0.;
        if (N == 1 || lambda == 0.) return 0.;  // is_constant (pair_action_class.h:225)
        // GenerateParticlePairs (pair_action_class.h:62-114)
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        std::vector<uint32_t> others;
        for (uint32_t p = 0; p < N; ++p)
            if (std::find(parts.begin(), parts.end(), p) == parts.end()) others.push_back(p);
        for (uint32_t p : parts)
            for (uint32_t q : others) pairs.emplace_back(p, q);
        for (size_t i = 0; i + 1 < parts.size(); ++i)
            for (size_t j = i + 1; j < parts.size(); ++j) pairs.emplace_back(parts[i], parts[j]);
        if (pairs.empty()) return 0.;
        uint32_t skip = 1u << level;
        double tot = 0.;
        for (uint32_t b_i = b0; b_i < b1; b_i += skip) {
            uint32_t b_j = b_i + skip;
            for (auto& pq : pairs) {
                double rm, rpm, rrpm;
                dr_drp_drrp(b_i, b_j, pq.first, pq.second, rm, rpm, rrpm);
                tot += calc_u(rm, rpm, rrpm, level);
            }
        }
        last_pair_short = tot;
        last_pair_long = 0.;
        if (use_long_range) {
            if (need_update_rho_k && mode == SPIMC_MODE_NEW) update_rho_k(b0, b1, parts, level);
            last_pair_long = calc_u_long(b0, b1, level);
            tot += last_pair_long;
        }
        return tot;
    }

    // PairAction::CalcLaplacianU (pair_action_class.h:167-199): the central second
    // difference (eps 1e-4) of the link action (b_i, b_j) of (p_i, p_j) in each
    // coordinate of bead (p_i, b_i)
    double calc_laplacian_u(uint32_t b_i, uint32_t b_j, uint32_t p_i, uint32_t p_j, uint32_t level) {
        double* b = R(p_i, b_i);
        const double r0[3] = {b[0], b[1], b[2]};
        const double eps = 1.e-4;
        double rm, rpm, rrpm, tot = 0.;
        dr_drp_drrp(b_i, b_j, p_i, p_j, rm, rpm, rrpm);
        const double f0 = calc_u(rm, rpm, rrpm, level);
        for (int d = 0; d < 3; ++d) {
            b[d] = r0[d] + eps;
            dr_drp_drrp(b_i, b_j, p_i, p_j, rm, rpm, rrpm);
            const double fp1 = calc_u(rm, rpm, rrpm, level);
            b[d] = r0[d] - eps;
            dr_drp_drrp(b_i, b_j, p_i, p_j, rm, rpm, rrpm);
            const double fm1 = calc_u(rm, rpm, rrpm, level);
            tot += (fp1 + fm1 - 2 * f0) / (eps * eps);
            b[d] = r0[d];
        }
        return tot;
    }
    // PairAction::GetActionLaplacian (pair_action_class.h:336-363): both links of
    // every window bead b_i, (b_i, b_i + skip) and (b_i, b_i - skip + M), of every
    // pair of GenerateParticlePairs; the long-range part is left out, as there.
    // The link slices are taken mod M.  The reference's bead_loop subtracts M
    // once (species_class.h:101), so for a window bead b_i >= M + skip its index
    // b_i - skip + M reads past the last slice; mod M is the same slice wherever
    // the reference's index is in range.
    double pair_get_action_laplacian(uint32_t b0, uint32_t b1, const std::vector<uint32_t>& parts, uint32_t level) {
        if (level > max_level || pair_type == SPIMC_PAIR_NONE) return 0.;
        if (N == 1 || lambda == 0.) return 0.;
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        std::vector<uint32_t> others;
        for (uint32_t p = 0; p < N; ++p)
            if (std::find(parts.begin(), parts.end(), p) == parts.end()) others.push_back(p);
        for (uint32_t p : parts)
            for (uint32_t q : others) pairs.emplace_back(p, q);
        for (size_t i = 0; i + 1 < parts.size(); ++i)
            for (size_t j = i + 1; j < parts.size(); ++j) pairs.emplace_back(parts[i], parts[j]);
        if (pairs.empty()) return 0.;
        const uint32_t skip = 1u << level;
        double tot = 0.;
        for (uint32_t b_i = b0; b_i < b1; b_i += skip) {
            const uint32_t b_j = (b_i + skip) % M, b_k = (b_i - skip + M) % M;
            for (auto& pq : pairs)
                tot += calc_laplacian_u(b_i, b_j, pq.first, pq.second, level) +
                       calc_laplacian_u(b_i, b_k, pq.first, pq.second, level);
        }
        return tot;
    }

    std::vector<std::pair<uint32_t, uint32_t>> all_pairs() const {  // pair_action_class.h:45-59
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        for (uint32_t p = 0; p + 1 < N; ++p)
            for (uint32_t q = p + 1; q < N; ++q) pairs.emplace_back(p, q);
        return pairs;
    }

    // PairAction::DActionDBeta (pair_action_class.h:241-264)
    double pair_d_action_d_beta() {
        if (pair_type == SPIMC_PAIR_NONE) return 0.;
        auto pairs = all_pairs();
        double tot = 0.;
        const long np = (long)pairs.size();
#pragma omp parallel for collapse(2) reduction(+ : tot)
        for (long i = 0; i < np; ++i)
            for (uint32_t b = 0; b < M; ++b) {
                double rm, rpm, rrpm;
                dr_drp_drrp(b, b + 1, pairs[i].first, pairs[i].second, rm, rpm, rrpm);
                tot += calc_du_dbeta(rm, rpm, rrpm);
            }
        if (use_long_range) tot += calc_whole_long(true);
        return tot;
    }

    // PairAction::Potential (pair_action_class.h:369-395)
    double pair_potential() {
        if (pair_type == SPIMC_PAIR_NONE) return 0.;
        auto pairs = all_pairs();
        double tot = 0.;
        const long np = (long)pairs.size();
#pragma omp parallel for collapse(2) reduction(+ : tot)
        for (long i = 0; i < np; ++i)
            for (uint32_t b = 0; b < M; ++b) {
                double d0[3], d1[3];
                dr(R(pairs[i].first, b), R(pairs[i].second, b), d0);
                dr(R(pairs[i].first, b + 1), R(pairs[i].second, b + 1), d1);
                tot += calc_v(mag(d0), mag(d1));
            }
        if (use_long_range) tot += calc_whole_long(false);
        return tot;
    }

    // PairAction::CalcGradientU (pair_action_class.h:135-155); Ilkka's analytic
    // override (ilkka_pair_action_class.h:172-222)
    void calc_gradient_u(uint32_t b_i, uint32_t b_j, uint32_t p_i, uint32_t p_j, double* tot) {
        if (pair_type == SPIMC_PAIR_ILKKA) {
            // DrDrpDrrp with vectors (path_class.h:153-166): r = r_{p_j} - r_{p_i}
            double rv[3], rpv[3], rrp[3];
            const double *a0 = R(p_i, b_i), *a1 = R(p_j, b_i), *c0 = R(p_i, b_j), *c1 = R(p_j, b_j);
            for (int d = 0; d < 3; ++d) {
                rv[d] = a1[d] - a0[d];
                rpv[d] = c1[d] - c0[d];
            }
            for (int d = 0; d < 3; ++d) {
                rv[d] -= std::nearbyint(rv[d] * iL) * L;
                rpv[d] += std::nearbyint((rv[d] - rpv[d]) * iL) * L;
            }
            for (int d = 0; d < 3; ++d) rrp[d] = rv[d] - rpv[d];
            for (int d = 0; d < 3; ++d) rrp[d] -= std::nearbyint(rrp[d] * iL) * L;
            double r_mag = mag(rv), r_p_mag = mag(rpv), r_r_p_mag = mag(rrp);
            double q = 0.5 * (r_mag + r_p_mag);
            double x = q + 0.5 * r_r_p_mag;
            double y = q - 0.5 * r_r_p_mag;
            double ux = 0., uy = 0.;
            u_xy.eval(x, y, &ux, &uy);
            double rh[3];
            for (int d = 0; d < 3; ++d) {
                rh[d] = r_mag == 0. ? 0. : rv[d] / r_mag;
                double sh = r_r_p_mag == 0. ? 0. : rrp[d] / r_r_p_mag;
                tot[d] = -0.5 * (ux * (rh[d] + sh) + uy * (rh[d] - sh));
            }
            if (use_long_range) {
                set_limits(r_u_min, r_u_max, r_mag, r_p_mag);
                double du_dr = u_long_r_s.deriv(r_mag);
                for (int d = 0; d < 3; ++d) tot[d] -= 0.5 * du_dr * rh[d];
            }
            return;
        }
        double* b = R(p_i, b_i);
        double r0[3] = {b[0], b[1], b[2]};
        const double eps = 1.e-4;
        for (int d = 0; d < 3; ++d) {
            double rm, rpm, rrpm;
            b[d] = r0[d] + eps;
            dr_drp_drrp(b_i, b_j, p_i, p_j, rm, rpm, rrpm);
            double f1 = calc_u(rm, rpm, rrpm, 0);
            b[d] = r0[d] - eps;
            dr_drp_drrp(b_i, b_j, p_i, p_j, rm, rpm, rrpm);
            double f2 = calc_u(rm, rpm, rrpm, 0);
            tot[d] = (f1 - f2) / (2. * eps);
            b[d] = r0[d];
        }
    }

    // PairAction::GetActionGradient(b, b+1, {p}, 0) (pair_action_class.h:305-337);
    // terms selects the link terms (SPIMC_GRAD_PAIR_SHORT) and the long-range
    // term (SPIMC_GRAD_PAIR_LONG)
    void pair_gradient_window(uint32_t b0, uint32_t b1, uint32_t p, double* out, uint32_t terms) {
        out[0] = out[1] = out[2] = 0.;
        if (pair_type == SPIMC_PAIR_NONE || N == 1) return;
        for (uint32_t b_i = b0; b_i < b1 && (terms & SPIMC_GRAD_PAIR_SHORT); ++b_i) {
            uint32_t b_j = b_i + 1;
            uint32_t b_k = b_i - 1 + M;
            for (uint32_t q = 0; q < N; ++q) {
                if (q == p) continue;
                double g[3];
                calc_gradient_u(b_i, b_j, p, q, g);
                for (int d = 0; d < 3; ++d) out[d] += g[d];
                calc_gradient_u(b_i, b_k, p, q, g);
                for (int d = 0; d < 3; ++d) out[d] += g[d];
            }
        }
        // long range: CalcGradientULong of the base class is zero (pair_action_class.h:158-165);
        // Ilkka's (ilkka_pair_action_class.h:231-249) is added once per particle pair
        // (pair_action_class.h:331-334)
        if (pair_type == SPIMC_PAIR_ILKKA && use_long_range && (terms & SPIMC_GRAD_PAIR_LONG)) {
            double g[3] = {0., 0., 0.};
            std::vector<cplx> ri;
            const std::vector<cplx>& rk = rho_mode();
            for (uint32_t b_i = b0; b_i < b1; ++b_i) {
                calc_rho_k(R(p, b_i), ri);
                const cplx* rb = &rk[(size_t)bead_loop(b_i) * nk];
                for (uint32_t k = 0; k < nk; ++k) {
                    double w = u_long_k[k] * (ri[k].real() * rb[k].imag() - ri[k].imag() * rb[k].real());
                    for (int d = 0; d < 3; ++d) g[d] += w * kn[3 * k + d] * box;
                }
            }
            for (uint32_t q = 0; q < N; ++q)
                if (q != p)
                    for (int d = 0; d < 3; ++d) out[d] += g[d];
        }
    }

    // ---- kinetic -----------------------------------------------------------
    // Kinetic::GetAction (kinetic_class.h:105-122): walks the path from bead
    // (p, b0) to the bead b1 - b0 links later; the walk compares beads, so it is
    // empty when it returns to its start.
    double kinetic_get_action(uint32_t b0, uint32_t b1, const std::vector<uint32_t>& parts, uint32_t level) {
        uint32_t skip = 1u << level;
        double tot = 0.;
        for (uint32_t p : parts) {
            Bd a{p, bead_loop(b0)};
            const Bd end = next(a, b1 - b0);
            uint32_t guard = 0;
            while (a != end && guard++ <= N * M) {
                Bd nx = next(a, skip);
                double d[3];
                dr(R(a), R(nx), d);
                tot -= kin_splines[skip - 1].log_rho_free(d);
                a = nx;
            }
        }
        return tot;
    }
    // Kinetic::GetActionLaplacian (kinetic_class.h:147-165): n_d 4 / (4 lambda tau
    // skip) per step of the walk from (p, b0) to the bead b1 - b0 links later
    // (i_4_lambda_tau, single_action_class.h:18)
    double kinetic_get_action_laplacian(uint32_t b0, uint32_t b1, const std::vector<uint32_t>& parts, uint32_t level) {
        const uint32_t skip = 1u << level;
        const double i4lt_l = (1. / (4. * lambda * tau)) / skip;
        double tot = 0.;
        for (uint32_t p : parts) {
            Bd a{p, bead_loop(b0)};
            const Bd end = next(a, b1 - b0);
            uint32_t guard = 0;
            while (a != end && guard++ <= N * M) {
                tot += 3 * 4. * i4lt_l;
                a = next(a, skip);
            }
        }
        return tot;
    }
    // Kinetic::DActionDBeta (kinetic_class.h:35-45)
    double kinetic_d_action_d_beta() {
        double tot = kin_dbeta_const;
#pragma omp parallel for collapse(2) reduction(+ : tot)
        for (uint32_t b = 0; b < M; ++b)
            for (uint32_t p = 0; p < N; ++p) {
                double d[3];
                dr(R(p, b), R(next1(Bd{p, b})), d);
                tot += kin_splines[0].dlog_rho_free_dtau(d);
            }
        return tot;
    }
    // Kinetic::VirialEnergy (kinetic_class.h:48-102)
    double kinetic_virial_energy(uint32_t W) {
        double tot = kin_dbeta_const / W;
        for (uint32_t p = 0; p < N; ++p) {
            Bd b_0{p, 0}, b_1 = next1(b_0), b_2 = next1(b_1);
            double dr_i_1[3], dr_i_L[3], t[3];
            dr(R(b_1), R(b_0), dr_i_1);
            for (int d = 0; d < 3; ++d) dr_i_L[d] = dr_i_1[d];
            for (uint32_t skip = 1; skip < W; ++skip) {
                dr(R(b_2), R(b_1), t);
                for (int d = 0; d < 3; ++d) dr_i_L[d] += t[d];
                b_1 = b_2;
                b_2 = next1(b_2);
            }
            tot -= (dr_i_L[0] * dr_i_1[0] + dr_i_L[1] * dr_i_1[1] + dr_i_L[2] * dr_i_1[2]) * kin_i4lt / (W * tau);
            for (uint32_t b_i = 1; b_i < M; ++b_i) {
                for (int d = 0; d < 3; ++d) dr_i_L[d] -= dr_i_1[d];
                dr(R(b_2), R(b_1), t);
                for (int d = 0; d < 3; ++d) dr_i_L[d] += t[d];
                b_0 = next1(b_0);
                dr(R(next1(b_0)), R(b_0), dr_i_1);
                b_1 = b_2;
                b_2 = next1(b_2);
                tot -= (dr_i_L[0] * dr_i_1[0] + dr_i_L[1] * dr_i_1[1] + dr_i_L[2] * dr_i_1[2]) * kin_i4lt / (W * tau);
            }
        }
        return tot;
    }

    // ---- whole-path action totals, for finite-difference gradient checks ----
    double total_kinetic() {
        double tot = 0.;
        for (uint32_t b = 0; b < M; ++b)
            for (uint32_t p = 0; p < N; ++p) {
                double d[3];
                dr(R(p, b), R(next1(Bd{p, b})), d);
                tot -= kin_splines[0].log_rho_free(d);
            }
        return tot;
    }
    double total_pair_short() {
        double tot = 0.;
        if (pair_type == SPIMC_PAIR_NONE) return 0.;
        for (auto& pq : all_pairs())
            for (uint32_t b = 0; b < M; ++b) {
                double rm, rpm, rrpm;
                dr_drp_drrp(b, b + 1, pq.first, pq.second, rm, rpm, rrpm);
                tot += calc_u(rm, rpm, rrpm, 0);
            }
        return tot;
    }
    double slice_long(uint32_t /* slice b, whose rho_k row is rb */, const cplx* rb) const {
        const std::vector<double>& coef = (pair_type == SPIMC_PAIR_BARE) ? v_long_k : u_long_k;
        double tot = 0.;
        for (uint32_t k = 0; k < nk; ++k) tot += coef[k] * cmag2(rb[k], rb[k]);
        return pair_type == SPIMC_PAIR_BARE ? tau * tot : tot;
    }

    // Central-difference gradient of the whole-path action terms for every bead,
    // using local re-evaluation of the terms a bead enters (test reference for
    // the adjoint; it restates no reference routine).
    // Central difference of the terms bead (b, p) enters, in mode NEW.
    void fd_bead(uint32_t terms, double h, uint32_t b, uint32_t p, double* g) {
        std::vector<cplx> ph0, ph1, rb(nk);
        double* pos = R(p, b);
        for (int d = 0; d < 3; ++d) {
            double x0 = pos[d], f[2];
            for (int sgn = 0; sgn < 2; ++sgn) {
                pos[d] = x0 + (sgn ? -h : h);
                double s = 0.;
                uint32_t bm = (b + M - 1) % M;
                if (terms & SPIMC_GRAD_KINETIC) {
                    double dd[3];
                    dr(R(p, b), R(next1(Bd{p, b})), dd);
                    s -= kin_splines[0].log_rho_free(dd);
                    dr(R(prev1(Bd{p, b})), R(p, b), dd);
                    s -= kin_splines[0].log_rho_free(dd);
                }
                if ((terms & SPIMC_GRAD_PAIR_SHORT) && pair_type != SPIMC_PAIR_NONE) {
                    for (uint32_t q = 0; q < N; ++q) {
                        if (q == p) continue;
                        uint32_t lo = std::min(p, q), hi = std::max(p, q);
                        double rm, rpm, rrpm;
                        dr_drp_drrp(b, b + 1, lo, hi, rm, rpm, rrpm);
                        s += calc_u(rm, rpm, rrpm, 0);
                        dr_drp_drrp(bm, bm + 1, lo, hi, rm, rpm, rrpm);
                        s += calc_u(rm, rpm, rrpm, 0);
                    }
                }
                if ((terms & SPIMC_GRAD_PAIR_LONG) && use_long_range && pair_type != SPIMC_PAIR_NONE) {
                    pos[d] = x0;
                    calc_rho_k(pos, ph0);
                    pos[d] = x0 + (sgn ? -h : h);
                    calc_rho_k(pos, ph1);
                    for (uint32_t k = 0; k < nk; ++k) rb[k] = rho[(size_t)b * nk + k] + (ph1[k] - ph0[k]);
                    s += slice_long(b, rb.data());
                }
                f[sgn] = s;
            }
            pos[d] = x0;
            g[d] = (f[0] - f[1]) / (2. * h);
        }
    }
    void gradient_fd(uint32_t terms, double h, double* out) {
        int saved = mode;
        mode = SPIMC_MODE_NEW;
        for (uint32_t b = 0; b < M; ++b)
            for (uint32_t p = 0; p < N; ++p) fd_bead(terms, h, b, p, out + 3 * ((size_t)b * N + p));
        mode = saved;
    }

    // ---- action trace: the level-0 decision's terms per attempt --------------
    // kinetic OLD, NEW; short-range pair OLD, NEW; long-range pair OLD, NEW;
    // log sample ratio; previous level's change; the long-range action of the
    // window's slices where no bead moves (in both OLD and NEW) (NaN rows: no
    // level-0 decision)
    static constexpr int kTraceWords = 9;
    bool tracing = false;
    std::vector<double> trace;
    double trace_row[kTraceWords];
    bool trace_row_set = false;
    double last_pair_short = 0., last_pair_long = 0.;
    void trace_begin() { trace_row_set = false; }
    void trace_end() {
        if (!tracing) return;
        for (int i = 0; i < kTraceWords; ++i) trace.push_back(trace_row_set ? trace_row[i] : std::nan(""));
    }
    // CalcULong at level 0 over the slices of [b0, b1) whose beads all keep their
    // OLD positions
    double unmoved_long(uint32_t b0, uint32_t b1) {
        if (!use_long_range || pair_type == SPIMC_PAIR_NONE || N == 1 || lambda == 0.) return 0.;
        const int saved = mode;
        mode = SPIMC_MODE_OLD;
        double tot = 0.;
        for (uint32_t b = b0; b < b1; ++b) {
            bool same = true;
            for (uint32_t p = 0; p < N && same; ++p) same = std::memcmp(Rm(0, p, b), Rm(1, p, b), 3 * sizeof(double)) == 0;
            if (same) tot += calc_u_long(b, b + 1, 0);
        }
        mode = saved;
        return tot;
    }

    // ---- draws (post-transform values, reference call order) ---------------
    bool recording = false;
    std::vector<double> rec;
    double from_tape() {
        if (cursor >= tape.size()) {
            g_err = "replay tape exhausted";
            return 0.5;
        }
        return tape[cursor++];
    }
    double keep(double v) {
        if (recording) rec.push_back(v);
        return v;
    }
    double draw() { return keep(use_tape ? from_tape() : rng.unif()); }
    uint32_t draw_index(uint32_t n) { return (uint32_t)keep(use_tape ? from_tape() : (double)rng.index(n)); }
    double draw_normal() { return keep(use_tape ? from_tape() : rng.normal()); }
    double draw_step() { return keep(use_tape ? from_tape() : step_size * (rng.unif() - 0.5)); }

    // ---- moves -------------------------------------------------------------
    void store_r_all_beads_of(uint32_t p, uint32_t first, uint32_t n) {
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t b = (first + i) % M;
            std::memcpy(Rm(0, p, b), Rm(1, p, b), 3 * sizeof(double));
        }
    }
    void restore_r_all_beads_of(uint32_t p, uint32_t first, uint32_t n) {
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t b = (first + i) % M;
            std::memcpy(Rm(1, p, b), Rm(0, p, b), 3 * sizeof(double));
        }
    }
    void store_rho(uint32_t b0, uint32_t b1) {  // Species::StoreRhoK (species_class.h:431)
        for (uint32_t b = b0; b < b1; ++b)
            std::memcpy(&rhoc[(size_t)bead_loop(b) * nk], &rho[(size_t)bead_loop(b) * nk], nk * sizeof(cplx));
    }
    void restore_rho(uint32_t b0, uint32_t b1) {  // Species::RestoreRhoK (species_class.h:434)
        for (uint32_t b = b0; b < b1; ++b)
            std::memcpy(&rho[(size_t)bead_loop(b) * nk], &rhoc[(size_t)bead_loop(b) * nk], nk * sizeof(cplx));
    }

    // The beads the last attempt moved (Bisect::affected_beads) and its bead0
    std::vector<Bd> affected;
    uint32_t last_bead0 = 0;

    // Bisection down to level 0 of the chains that start at the beads `start`
    // and end at the bead n_bisect_beads links later in each mode: the level
    // loop of Bisect::Attempt (bisect_class.h:65-122) and of
    // PermBisectIterative::Attempt (perm_bisect_iterative_class.h:159-217),
    // which samples the chains one after the other.  The chain ends are fixed;
    // the beads between are resampled top-down.  parts is the particle list the
    // actions see; prev_action_change is carried into the first level.
    bool bisect_levels(const std::vector<Bd>& start, uint32_t bead0, uint32_t bead1, const std::vector<uint32_t>& parts,
                       double prev_action_change) {
        const size_t n = start.size();
        std::vector<Bd> end_new(n);
        mode = SPIMC_MODE_NEW;
        for (size_t i = 0; i < n; ++i) end_new[i] = next(start[i], n_bisect_beads);
        for (int level = (int)n_level - 1; level >= 0; --level) {
            uint32_t skip = 1u << level;
            double level_tau = tau * skip;
            double sigma_l = std::sqrt(lambda * level_tau);
            double old_lsp = 0., new_lsp = 0.;
            for (size_t i = 0; i < n; ++i) {
                Bd a = start[i];
                uint32_t guard = 0;
                while (a != end_new[i] && guard++ <= N * M) {
                    double rbar[3], t[3], delta_old[3], delta_new[3];
                    // old sampling, OLD mode: RBar(c, a) = a + 0.5 Dr(c, a)
                    mode = SPIMC_MODE_OLD;
                    Bd bb = next(a, skip), cc = next(bb, skip);
                    dr(R(cc), R(a), t);
                    for (int d = 0; d < 3; ++d) rbar[d] = R(a)[d] + 0.5 * t[d];
                    dr(R(bb), rbar, delta_old);
                    old_lsp += bis_splines[level].log_rho_free(delta_old);
                    // new sampling, NEW mode
                    mode = SPIMC_MODE_NEW;
                    bb = next(a, skip);
                    cc = next(bb, skip);
                    for (int d = 0; d < 3; ++d) delta_new[d] = sigma_l * draw_normal();
                    put_in_box(delta_new);
                    dr(R(cc), R(a), t);
                    for (int d = 0; d < 3; ++d) rbar[d] = R(a)[d] + 0.5 * t[d];
                    double* nb = R(bb);
                    for (int d = 0; d < 3; ++d) nb[d] = rbar[d] + delta_new[d];
                    new_lsp += bis_splines[level].log_rho_free(delta_new);
                    a = cc;
                }
            }
            double old_action = 0., new_action = 0.;
            // action list order: Kinetic, pair action
            mode = SPIMC_MODE_OLD;
            const double k_old = kinetic_get_action(bead0, bead1, parts, level);
            old_action += k_old;
            mode = SPIMC_MODE_NEW;
            const double k_new = kinetic_get_action(bead0, bead1, parts, level);
            new_action += k_new;
            mode = SPIMC_MODE_OLD;
            last_pair_short = last_pair_long = 0.;
            old_action += pair_get_action(bead0, bead1, parts, level);
            const double ps_old = last_pair_short, pl_old = last_pair_long;
            mode = SPIMC_MODE_NEW;
            last_pair_short = last_pair_long = 0.;
            new_action += pair_get_action(bead0, bead1, parts, level);
            double log_sample_ratio = -new_lsp + old_lsp;
            if (level == 0 && tracing) {
                const double row[kTraceWords] = {k_old, k_new, ps_old, last_pair_short, pl_old, last_pair_long,
                                                 log_sample_ratio, prev_action_change, unmoved_long(bead0, bead1)};
                std::memcpy(trace_row, row, sizeof row);
                trace_row_set = true;
            }
            double current_action_change = new_action - old_action;
            double log_accept = log_sample_ratio - current_action_change + prev_action_change;
            if (log_accept < std::log(draw())) return false;
            prev_action_change = current_action_change;
        }
        return true;
    }

    // Bisect::Attempt (bisect_class.h:39-124)
    bool bisect_attempt(uint32_t& p_out, uint32_t& b0_out) {
        uint32_t p = draw_index(N);
        uint32_t bead0 = draw_index(M);
        uint32_t bead1 = bead0 + n_bisect_beads;
        p_out = p;
        b0_out = bead0;
        mode = SPIMC_MODE_NEW;
        const Bd bead_i{p, bead0};
        affected.clear();
        Bd bead_f = bead_i;
        affected.push_back(bead_i);
        for (uint32_t i = 0; i < n_bisect_beads; ++i) {
            bead_f = next1(bead_f);
            affected.push_back(bead_f);
        }
        std::vector<uint32_t> parts{p};
        if (bead_f.p != p) parts.push_back(bead_f.p);
        return bisect_levels({bead_i}, bead0, bead1, parts, 0.);
    }
    void store_affected() {
        for (const Bd& x : affected) std::memcpy(Rm(0, x.p, x.b), Rm(1, x.p, x.b), 3 * sizeof(double));
    }
    void restore_affected() {
        for (const Bd& x : affected) std::memcpy(Rm(1, x.p, x.b), Rm(0, x.p, x.b), 3 * sizeof(double));
    }
    void bisect_accept(uint32_t /* p */, uint32_t bead0) {  // bisect_class.h:24-36
        store_affected();
        store_rho(bead0, bead0 + n_bisect_beads);
        need_update_rho_k = true;
    }
    void bisect_reject(uint32_t /* p */, uint32_t bead0) {  // bisect_class.h:127-139
        restore_affected();
        restore_rho(bead0, bead0 + n_bisect_beads);
        need_update_rho_k = true;
    }

    // PermBisectIterative::UpdatePermTable (perm_bisect_iterative_class.h:10-29),
    // row p only: t(p, j) from the bead (p, bead0) to the bead n_bisect_beads
    // links after (j, bead0)
    void perm_table_row(uint32_t p, uint32_t bead0, std::vector<double>& row) {
        row.resize(N);
        const double i4ltn = 1. / (4. * lambda * tau) / n_bisect_beads;
        const double* a = R(p, bead0);
        for (uint32_t j = 0; j < N; ++j) {
            double d[3];
            dr(a, R(next(Bd{j, bead0}, n_bisect_beads)), d);
            double exponent = (-(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])) * i4ltn;
            row[j] = exponent > perm_log_epsilon ? std::exp(exponent) : 0.;
        }
    }

    // PermBisectIterative::SelectCycleIterative (perm_bisect_iterative_class.h:32-110):
    // false when the walk stops before it closes; else the cycle's particles
    // (ps) and weight
    bool select_cycle(uint32_t bead0, std::vector<uint32_t>& ps, double& weight) {
        mode = SPIMC_MODE_NEW;
        ps.clear();
        uint32_t p0 = draw_index(N);
        uint32_t p = p0;
        std::vector<double> t_row, t_c(N);
        weight = 1.;
        std::vector<double> w_num, w_den;
        do {
            ps.push_back(p);
            perm_table_row(p, bead0, t_row);
            // a row is taken from t with the visited particles removed and the
            // closing link restored (the reference's t_c rows)
            for (uint32_t i = 0; i < N; ++i) t_c[i] = t_row[i];
            for (uint32_t q : ps) t_c[q] = 0.;
            t_c[p0] = t_row[p0];
            double Q_p = 0., Q_p_c = 0.;
            for (uint32_t i = 0; i < N; ++i) {
                Q_p += t_row[i];
                Q_p_c += t_c[i];
            }
            if (Q_p_c / Q_p < draw()) return false;
            double x = draw();
            double t_Q = 0.;
            const uint32_t from = p;
            bool found = false;
            int last = -1;  // the last candidate with a nonzero weight
            for (uint32_t i = 0; i < N; ++i) {
                if (t_c[i] > 0.) last = (int)i;
                t_Q += t_c[i] / Q_p_c;
                if (t_Q > x) {
                    p = i;
                    found = true;
                    break;
                }
            }
            // a row whose running sum never passes x selects no particle; the
            // reference would then revisit the same particle.  Free-running, this
            // engine's rule: the running sum fell an ulp short (the last candidate
            // is the selection), or no candidate is left (the walk stops).  On a
            // replay tape it is an error.
            if (!found) {
                if (use_tape) {
                    g_err = "permutation walk selected no particle";
                    return false;
                }
                if (last < 0) return false;
                p = (uint32_t)last;
            }
            // the weight's factors t(ps[i], ps[i+1]) / t(ps[i], ps[i]), in order
            w_num.push_back(t_row[p]);
            w_den.push_back(t_row[from]);
            if (ps.size() > (size_t)N + 1) {
                g_err = "permutation walk did not close";
                return false;
            }
        } while (p != p0);
        for (size_t i = 0; i < w_num.size(); ++i) weight *= w_num[i] / w_den[i];
        return true;
    }

    // PermBisectIterative::Attempt (perm_bisect_iterative_class.h:113-223).
    // n_chain_out = 0 when the cycle walk stopped (no bisection).
    bool perm_bisect_attempt(uint32_t& n_chain_out) {
        uint32_t bead0 = draw_index(M);
        last_bead0 = bead0;
        uint32_t bead1 = bead0 + n_bisect_beads;
        bool roll_over = bead1 > M - 1;
        std::vector<uint32_t> ps;
        double weight = 1.;
        n_chain_out = 0;
        if (!select_cycle(bead0, ps, weight)) return false;
        const uint32_t n = (uint32_t)ps.size();
        n_chain_out = n;
        mode = SPIMC_MODE_NEW;
        std::vector<Bd> start(n), fm1(n);
        std::vector<uint32_t> parts;
        for (uint32_t i = 0; i < n; ++i) {
            start[i] = Bd{ps[i], bead0};
            fm1[i] = next(start[i], n_bisect_beads - 1);
            parts.push_back(ps[i]);
        }
        if (roll_over) {
            for (uint32_t i = 0; i < n; ++i) parts.push_back(next1(fm1[i]).p);
            std::sort(parts.begin(), parts.end());
            parts.erase(std::unique(parts.begin(), parts.end()), parts.end());
        }
        // PermuteBeads: bead fm1(i) is followed by the old successor of fm1(i + 1)
        cyc.resize(N);
        for (uint32_t x = 0; x < N; ++x) cyc[x] = x;
        for (uint32_t i = 0; i < n; ++i) cyc[fm1[i].p] = fm1[(i + 1) % n].p;
        att_e = fm1[0].b;
        att_on = true;
        affected.clear();
        for (uint32_t i = 0; i < n; ++i) {
            Bd a = next1(start[i]);
            for (uint32_t j = 1; j < n_bisect_beads; ++j, a = next1(a)) affected.push_back(a);
        }
        return bisect_levels(start, bead0, bead1, parts, -std::log(weight));
    }
    // PermBisect::Accept / AssignParticleLabels (perm_bisect_class.h:47-73) then
    // Bisect::Accept: slices after the relinked one take the labels of the path
    // they now continue, which moves the permutation to the wrap.
    void perm_bisect_accept(uint32_t bead0, uint32_t n) {
        store_affected();
        store_rho(bead0, bead0 + n_bisect_beads);
        need_update_rho_k = true;
        if (n > 1) {
            std::vector<double> tmp(3 * (size_t)N);
            for (uint32_t b = att_e + 1; b < M; ++b)
                for (int m = 0; m < 2; ++m) {
                    for (uint32_t x = 0; x < N; ++x) std::memcpy(&tmp[3 * x], Rm(m, x, b), 3 * sizeof(double));
                    for (uint32_t x = 0; x < N; ++x) std::memcpy(Rm(m, x, b), &tmp[3 * cyc[x]], 3 * sizeof(double));
                }
            std::vector<uint32_t> s2(N);
            for (uint32_t x = 0; x < N; ++x) s2[x] = sigma[cyc[x]];
            sigma = s2;
            for (uint32_t x = 0; x < N; ++x) isigma[sigma[x]] = x;
        }
        att_on = false;
    }
    // ---- PermBisectTable (perm_bisect_table_class.h:10-228) ------------------
    // The cycle list (BuildCycles, :54-102) is every permutation type of k =
    // n_perm_part particles (scaffold's GenPerm) crossed with every k-subset of
    // the particles (GenCombPermK without repetition or order), type-major.  The
    // scaffold submodule is not in the reference tree, so the order within both
    // lists is this code's: permutations of 0..k-1 and subsets (sorted members)
    // in lexicographic order.  The order decides which entry a draw selects, not
    // which distribution is sampled.
    static bool next_subset(std::vector<uint32_t>& s, uint32_t n) {
        const int k = (int)s.size();
        int i = k - 1;
        while (i >= 0 && s[i] == n - (uint32_t)k + (uint32_t)i) --i;
        if (i < 0) return false;
        ++s[i];
        for (int j = i + 1; j < k; ++j) s[j] = s[j - 1] + 1;
        return true;
    }
    // UpdatePermTable (:31-51) in the current mode: t(i, j) from the bead
    // (i, bead0) to the bead n_bisect_beads links after (j, bead0)
    void table_update(uint32_t bead0, std::vector<double>& t) {
        t.assign((size_t)N * N, 0.);
        const double i4ltn = 1. / (4. * lambda * tau) / n_bisect_beads;
        std::vector<Bd> b1(N);
        for (uint32_t p = 0; p < N; ++p) b1[p] = next(Bd{p, bead0}, n_bisect_beads);
        for (uint32_t i = 0; i < N; ++i) {
            const double* a = R(i, bead0);
            double d[3];
            dr(a, R(b1[i]), d);
            const double s_ii = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            for (uint32_t j = 0; j < N; ++j) {
                dr(a, R(b1[j]), d);
                const double s_ij = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const double exponent = (-s_ij + s_ii) * i4ltn;
                t[(size_t)i * N + j] = exponent > perm_log_epsilon ? std::exp(exponent) : 0.;
            }
        }
    }
    // ConstructPermTable (:10-28): the running total of the cycle weights above
    // epsilon, in list order.  With select, the entry whose running total first
    // exceeds x (SelectCycle, :105-119, a bisection over the same running
    // totals): its subset and permutation type; false if none does.
    double table_total(const std::vector<double>& t, bool select, double x, std::vector<uint32_t>& part,
                       std::vector<uint32_t>& type, double& weight, bool& found) {
        const uint32_t k = table_k;
        std::vector<uint32_t> pt(k), sub(k);
        for (uint32_t i = 0; i < k; ++i) pt[i] = i;
        double total = 0.;
        found = false;
        std::vector<uint32_t> last_part, last_type;
        double last_w = 0.;
        do {
            for (uint32_t i = 0; i < k; ++i) sub[i] = i;
            do {
                double w = 1.;
                for (uint32_t i = 0; i < k; ++i) w *= t[(size_t)sub[i] * N + sub[pt[i]]];
                if (w > perm_epsilon) {
                    total += w;
                    if (select) {
                        last_part = sub;
                        last_type = pt;
                        last_w = w;
                        if (total > x) {
                            part = sub;
                            type = pt;
                            weight = w;
                            found = true;
                            return total;
                        }
                    }
                }
            } while (next_subset(sub, N));
        } while (std::next_permutation(pt.begin(), pt.end()));
        if (select && !found && !last_part.empty()) {
            // x at or past the total: the reference's bisection would index past
            // its list.  This engine takes the last entry (free-running rounding);
            // a replay tape reports an error.
            part = last_part;
            type = last_type;
            weight = last_w;
        }
        return total;
    }
    // PermBisectTable::Attempt (:122-228).  The particle list given to the
    // actions is that of PermBisectIterative (perm_bisect_iterative_class.h:139-146):
    // on a window that crosses the wrap it also lists the labels the chains carry
    // past it.  The reference's table move lists the cycle's own labels only
    // (:133-141), which leaves the moved beads under the other labels out of the
    // pair action and of Species::UpdateRhoK (species_class.h:406-424).
    bool perm_table_attempt(uint32_t& n_chain_out) {
        uint32_t bead0 = draw_index(M);
        last_bead0 = bead0;
        uint32_t bead1 = bead0 + n_bisect_beads;
        bool roll_over = bead1 > M - 1;
        const uint32_t k = table_k;
        n_chain_out = k;
        mode = SPIMC_MODE_NEW;
        std::vector<double> t;
        table_update(bead0, t);
        std::vector<uint32_t> part, type;
        double weight = 0.;
        bool found = false;
        const double tot0 = table_total(t, false, 0., part, type, weight, found);
        const double x = tot0 * draw();  // UnifRand(0, perm_tot_0)
        table_total(t, true, x, part, type, weight, found);
        if (!found) {
            if (use_tape || part.empty()) {
                g_err = "permutation table selected no cycle";
                return false;
            }
        }
        std::vector<Bd> start(k), fm1(k);
        std::vector<uint32_t> parts;
        for (uint32_t i = 0; i < k; ++i) {
            start[i] = Bd{part[i], bead0};
            fm1[i] = next(start[i], n_bisect_beads - 1);
            parts.push_back(part[i]);
        }
        if (roll_over) {
            for (uint32_t i = 0; i < k; ++i) parts.push_back(next1(fm1[i]).p);
            std::sort(parts.begin(), parts.end());
            parts.erase(std::unique(parts.begin(), parts.end()), parts.end());
        }
        // PermuteBeads: bead fm1(i) is followed by the old successor of fm1(type(i))
        cyc.resize(N);
        for (uint32_t q = 0; q < N; ++q) cyc[q] = q;
        for (uint32_t i = 0; i < k; ++i) cyc[fm1[i].p] = fm1[type[i]].p;
        att_e = fm1[0].b;
        att_on = true;
        affected.clear();
        for (uint32_t i = 0; i < k; ++i) {
            Bd a = next1(start[i]);
            for (uint32_t j = 1; j < n_bisect_beads; ++j, a = next1(a)) affected.push_back(a);
        }
        const double w_sel = weight;
        if (!bisect_levels(start, bead0, bead1, parts, -std::log(w_sel))) return false;
        // the neighbourhood changed: the table of the new path (:216-225)
        if (k < N) {
            mode = SPIMC_MODE_NEW;
            table_update(bead0, t);
            const double tot1 = table_total(t, false, 0., part, type, weight, found);
            const double ratio = table_balance ? tot0 / (tot1 * w_sel) : tot0 / tot1;
            if (ratio < draw()) return false;
        }
        return true;
    }

    void perm_bisect_reject(uint32_t bead0, uint32_t n) {  // PermBisect::Reject (perm_bisect_class.h:76-90)
        att_on = false;
        if (n > 0) restore_affected();
        restore_rho(bead0, bead0 + n_bisect_beads);
        need_update_rho_k = true;
    }
    // Species::SetCycleCount (species_class.h:28-44)
    std::vector<uint32_t> cycle_lengths() const {
        std::vector<uint32_t> out;
        std::vector<char> seen(N, 0);
        for (uint32_t p = 0; p < N; ++p) {
            if (seen[p]) continue;
            uint32_t len = 1, q = sigma[p];
            seen[p] = 1;
            while (q != p) {
                seen[q] = 1;
                ++len;
                q = sigma[q];
            }
            out.push_back(len);
        }
        return out;
    }

    // DisplaceParticle::Attempt (displace_particle_class.h:28-72)
    bool displace_attempt(uint32_t& p_out) {
        uint32_t p = draw_index(N);
        p_out = p;
        std::vector<uint32_t> parts{p};
        mode = SPIMC_MODE_NEW;
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = draw_step();
        for (uint32_t b = 0; b < M; ++b) {
            double* x = R(p, b);
            for (int k = 0; k < 3; ++k) x[k] = x[k] + d[k];
        }
        double old_action = 0., new_action = 0.;
        mode = SPIMC_MODE_OLD;
        const double k_old = kinetic_get_action(0, M, parts, 0);
        old_action += k_old;
        mode = SPIMC_MODE_NEW;
        const double k_new = kinetic_get_action(0, M, parts, 0);
        new_action += k_new;
        mode = SPIMC_MODE_OLD;
        last_pair_short = last_pair_long = 0.;
        old_action += pair_get_action(0, M, parts, 0);
        const double ps_old = last_pair_short, pl_old = last_pair_long;
        mode = SPIMC_MODE_NEW;
        last_pair_short = last_pair_long = 0.;
        new_action += pair_get_action(0, M, parts, 0);
        if (tracing) {
            const double row[kTraceWords] = {k_old, k_new, ps_old, last_pair_short, pl_old, last_pair_long, 0., 0.,
                                             unmoved_long(0, M)};
            std::memcpy(trace_row, row, sizeof row);
            trace_row_set = true;
        }
        double log_accept = old_action - new_action;
        return !(log_accept < std::log(draw()));
    }
    void displace_accept(uint32_t p) {  // displace_particle_class.h:15-26
        store_r_all_beads_of(p, 0, M);
        store_rho(0, M);
        need_update_rho_k = true;
    }
    void displace_reject(uint32_t p) {  // displace_particle_class.h:75-86
        restore_r_all_beads_of(p, 0, M);
        restore_rho(0, M);
        need_update_rho_k = true;
    }

    // ---- observables -------------------------------------------------------
    // Energy::ThermalEstimator / VirialEstimator (energy_class.h:22-116)
    void energy_accumulate() {
        mode = SPIMC_MODE_NEW;
        if (energy_estimator == SPIMC_ENERGY_THERMAL) {
            e_acc[0] += kinetic_d_action_d_beta();
            e_acc[1] += pair_d_action_d_beta();
        } else {
            uint32_t W = virial_window ? virial_window : M;
            double kinetic_energy = kinetic_virial_energy(W);
            double centroid = 0.;
            std::vector<double> g;
            if (energy_estimator == SPIMC_ENERGY_VIRIAL_ADJOINT) {
                g.resize((size_t)3 * M * N);
                gradient_fd(SPIMC_GRAD_PAIR_SHORT | SPIMC_GRAD_PAIR_LONG, 1.e-5, g.data());
            }
            for (uint32_t p = 0; p < N; ++p)
                for (uint32_t b = 0; b < M; ++b) {
                    double d[3], grad[3];
                    dr(R(p, b), R(p, 0), d);
                    if (energy_estimator == SPIMC_ENERGY_VIRIAL_COMPAT) {
                        pair_gradient_window(b, b + 1, p, grad, SPIMC_GRAD_PAIR_SHORT | SPIMC_GRAD_PAIR_LONG);
                    } else {
                        const double* gg = &g[3 * ((size_t)b * N + p)];
                        grad[0] = gg[0];
                        grad[1] = gg[1];
                        grad[2] = gg[2];
                    }
                    centroid += grad[0] * d[0] + grad[1] * d[1] + grad[2] * d[2];
                }
            kinetic_energy += centroid / (2. * tau);
            e_acc[0] += kinetic_energy;
            e_acc[1] += pair_d_action_d_beta();
        }
        n_measure[0] += 1;
    }
    // StructureFactor::Accumulate (structure_factor_class.h:15-32)
    void sk_accumulate() {
        mode = SPIMC_MODE_NEW;
        for (uint32_t k = 0; k < nk; ++k)
            if (kmag[k] < sk_k_cut)
                for (uint32_t b = 0; b < M; ++b) {
                    const cplx& a = rho[(size_t)b * nk + k];
                    sk_acc[k] += cmag2(a, a);
                }
        n_measure[1] += 1;
    }
    // PairCorrelation::Accumulate (pair_correlation_class.h:15-28), LinearGrid::ReverseMap
    // (observable_class.h:56-62)
    void gr_accumulate() {
        mode = SPIMC_MODE_NEW;
        auto pairs = all_pairs();
        for (uint32_t b = 0; b < M; ++b)
            for (auto& pq : pairs) {
                double d[3];
                dr(R(pq.first, b), R(pq.second, b), d);
                // (uint32_t)nearbyint(x): negative x wraps past n_r and is dropped
                double f = std::nearbyint((mag(d) - gr_r_min) * gr_d_ir - 0.5);
                if (f >= 0. && f < (double)gr_n_r) {
                    uint32_t i = (uint32_t)f;
                    gr_acc[i] = gr_acc[i] + 1.;
                }
            }
        n_measure[2] += 1;
    }
};

}  // namespace

extern "C" {

const char* spimc_oracle_last_error(void) { return g_err.c_str(); }

void spimc_oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int spimc_oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// KSpace::Setup / Include (k_space_class.h:17-80).  Enumerates every integer
// triple in [-max_index, max_index]^3, max_index = ceil(1.1 k_cut / (2 pi / L)),
// keeps the strict-cutoff half space.  Order: n0, then n1, then n2 ascending.
int spimc_oracle_kspace_build(double L, double k_cut, int32_t* k_n, uint32_t cap, uint32_t* n_k) {
    if (!(L > 0.) || !(k_cut >= 0.) || !n_k) return fail(SPIMC_ERR_ARG, "bad kspace arguments");
    double box = 2. * M_PI / L;
    int mx = (int)(uint32_t)std::ceil(1.1 * k_cut / box);
    uint32_t cnt = 0;
    for (int a = -mx; a <= mx; ++a)
        for (int b = -mx; b <= mx; ++b)
            for (int c = -mx; c <= mx; ++c) {
                double k[3] = {a * box, b * box, c * box};
                double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                bool inc = false;
                if (k2 < k_cut * k_cut && k2 != 0.) {
                    if (k[0] > 0.)
                        inc = true;
                    else if (k[0] == 0. && k[1] > 0.)
                        inc = true;
                    else if (k[0] == 0. && k[1] == 0. && k[2] > 0.)
                        inc = true;
                }
                if (!inc) continue;
                if (k_n && cnt < cap) {
                    k_n[3 * cnt] = a;
                    k_n[3 * cnt + 1] = b;
                    k_n[3 * cnt + 2] = c;
                }
                ++cnt;
            }
    *n_k = cnt;
    return SPIMC_OK;
}

int spimc_oracle_create(const spimc_system_desc* s, uint64_t seed, void** out) {
    if (!s || !out) return fail(SPIMC_ERR_ARG, "null argument");
    if (s->n_d != 3) return fail(SPIMC_ERR_UNSUPPORTED, "n_d must be 3");
    if (s->n_bead < 2 || s->n_part < 1) return fail(SPIMC_ERR_ARG, "bad sizes");
    if (s->pair_max_level != 0) return fail(SPIMC_ERR_UNSUPPORTED, "max_level must be 0");
    Oracle* o = new Oracle();
    o->N = s->n_part;
    o->M = s->n_bead;
    o->L = s->L;
    o->iL = s->L > 0. ? 1. / s->L : 0.;
    o->vol = s->L > 0. ? s->L * s->L * s->L : 1.;
    o->beta = s->beta;
    o->tau = s->beta / (1. * s->n_bead);  // path_class.h:43
    o->lambda = s->lambda;
    o->r.assign((size_t)3 * o->M * o->N, 0.);
    o->rc = o->r;
    o->sigma.resize(o->N);
    for (uint32_t p = 0; p < o->N; ++p) o->sigma[p] = p;
    o->isigma = o->sigma;
    o->perm_attempt.assign(o->N, 0);
    o->perm_accept.assign(o->N, 0);
    // k-space
    o->nk = s->n_k;
    o->kn.assign(s->k_n, s->k_n + 3 * (size_t)s->n_k);
    o->box = s->L > 0. ? 2. * M_PI / s->L : 0.;
    o->max_index = 0;
    o->kmag.resize(o->nk);
    for (uint32_t k = 0; k < o->nk; ++k) {
        double k2 = 0.;
        for (int d = 0; d < 3; ++d) {
            o->max_index = std::max(o->max_index, std::abs(o->kn[3 * k + d]));
            double kv = o->kn[3 * k + d] * o->box;
            k2 += kv * kv;
        }
        o->kmag[k] = std::sqrt(k2);
    }
    // Kinetic (kinetic_class.h:16-32)
    o->kin_n_images = s->kinetic_n_images;
    o->kin_i4lt = 1. / (4. * o->lambda * o->tau);
    o->kin_dbeta_const = o->N * o->M * 3 / (2. * o->tau);
    uint32_t n_spline = o->M / 2 + (o->M % 2) + 1;
    o->kin_splines.resize(n_spline);
    o->kin_splines[0].build(o->L, o->kin_n_images, o->lambda, o->tau, true);
    for (uint32_t i = 1; i < n_spline; ++i)
        o->kin_splines[i].build(o->L, o->kin_n_images, o->lambda, o->tau * (i + 1), false);
    // Bisect (bisect_class.h:158-190)
    o->n_level = s->bisect_n_level;
    o->n_bisect_beads = 1u << o->n_level;
    o->perm_n_level = o->n_level;
    o->bis_n_images = s->bisect_n_images;
    uint32_t n_bs = (uint32_t)std::floor(std::log2((double)o->M));
    if (o->n_level < 1 || o->n_level > n_bs) {
        delete o;
        return fail(SPIMC_ERR_ARG, "bisect n_level must be in [1, floor(log2 n_bead)]");
    }
    o->bis_splines.resize(n_bs);
    for (uint32_t i = 0; i < n_bs; ++i)
        o->bis_splines[i].build(o->L, o->bis_n_images, o->lambda, 0.5 * o->tau * (1u << i), false);
    o->step_size = s->displace_step;
    // pair action
    o->pair_type = s->pair_type;
    o->use_long_range = s->pair_use_long_range != 0;
    o->max_level = s->pair_max_level;
    double nn = (double)o->N;
    if (o->pair_type == SPIMC_PAIR_BARE) {
        // BarePairAction ctor (bare_pair_action_class.h:22-96)
        o->is_coulomb = s->bare_is_coulomb != 0;
        o->v_r.build(s->bare_r, s->bare_v_r, s->bare_n_r, false);  // create_general_grid
        o->r_v_min = s->bare_r[0];
        o->r_v_max = s->bare_r[s->bare_n_r - 1];
        if (o->use_long_range) {
            o->v_long_r.build(s->bare_r_long, s->bare_v_long_r, s->bare_n_r_long, false);
            o->r_v_long_min = s->bare_r_long[0];
            o->r_v_long_max = s->bare_r_long[s->bare_n_r_long - 1];
            o->v_long_k.assign(s->pair_v_long_k, s->pair_v_long_k + o->nk);
            o->v_long_k_0 = s->pair_v_long_k_0 * 0.5 * nn * nn * o->M;
            o->v_long_r_0 = s->pair_v_long_r_0 * (-0.5) * nn * o->M;
        }
    } else if (o->pair_type == SPIMC_PAIR_DAVID) {
        // DavidPairAction ctor (david_pair_action_class.h:192-360)
        o->n_order = s->david_n_order;
        o->n_val = 1;
        for (uint32_t i = 1; i <= o->n_order; ++i) o->n_val += 1 + i;
        uint32_t ng = s->david_n_grid;
        o->g_start = s->david_grid[0];
        o->g_end = s->david_grid[ng - 1];
        o->u_kj.resize(o->n_val + 1);
        o->du_kj.resize(o->n_val + 1);
        for (uint32_t c = 0; c <= o->n_val; ++c) {
            o->u_kj[c].build(s->david_grid, s->david_u_kj + (size_t)c * ng, ng, false);
            o->du_kj[c].build(s->david_grid, s->david_du_kj + (size_t)c * ng, ng, false);
        }
        if (o->use_long_range) {
            o->v_long_k.assign(s->pair_v_long_k, s->pair_v_long_k + o->nk);
            o->u_long_k.resize(o->nk);
            o->du_long_k.resize(o->nk);
            for (uint32_t k = 0; k < o->nk; ++k) {
                o->u_long_k[k] = o->v_long_k[k] * o->tau;
                o->du_long_k[k] = o->v_long_k[k];
            }
            double vk0 = s->pair_v_long_k_0, vr0 = s->pair_v_long_r_0;
            o->du_long_k_0 = vk0 * 0.5 * nn * nn * o->M;
            o->du_long_r_0 = vr0 * (-0.5) * nn * o->M;
            o->v_long_k_0 = vk0 * 0.5 * nn * nn * o->M;
            o->v_long_r_0 = vr0 * (-0.5) * nn * o->M;
        }
    } else if (o->pair_type == SPIMC_PAIR_ILKKA) {
        // IlkkaPairAction ctor (ilkka_pair_action_class.h:252-431)
        o->u_xy.build(s->ilkka_x_u, s->ilkka_n_x_u, s->ilkka_y_u, s->ilkka_n_y_u, s->ilkka_u_xy);
        o->du_xy.build(s->ilkka_x_du, s->ilkka_n_x_du, s->ilkka_y_du, s->ilkka_n_y_du, s->ilkka_du_xy);
        o->is_coulomb = false;
        o->v_r.build(s->ilkka_r_v, s->ilkka_v_r, s->ilkka_n_r_v, false);
        o->r_v_min = s->ilkka_r_v[0];
        o->r_v_max = s->ilkka_r_v[s->ilkka_n_r_v - 1];
        if (o->use_long_range) {
            o->u_long_r_s.build(s->ilkka_r_u_long, s->ilkka_u_long_r, s->ilkka_n_r_u_long, false);
            o->r_u_min = s->ilkka_r_u_long[0];
            o->r_u_max = s->ilkka_r_u_long[s->ilkka_n_r_u_long - 1];
            o->du_long_r_s.build(s->ilkka_r_du_long, s->ilkka_du_long_r, s->ilkka_n_r_du_long, false);
            o->r_du_min = s->ilkka_r_du_long[0];
            o->r_du_max = s->ilkka_r_du_long[s->ilkka_n_r_du_long - 1];
            o->v_long_r.build(s->ilkka_r_v_long, s->ilkka_v_long_r, s->ilkka_n_r_v_long, false);
            o->r_v_long_min = s->ilkka_r_v_long[0];
            o->r_v_long_max = s->ilkka_r_v_long[s->ilkka_n_r_v_long - 1];
            o->u_long_k.assign(s->ilkka_u_long_k, s->ilkka_u_long_k + o->nk);
            o->du_long_k.assign(s->ilkka_du_long_k, s->ilkka_du_long_k + o->nk);
            o->v_long_k.assign(s->pair_v_long_k, s->pair_v_long_k + o->nk);
            // homologous constants (ilkka_pair_action_class.h:418-423)
            o->du_long_k_0 = s->ilkka_du_long_k_0 * 0.5 * nn * nn * o->M;
            o->du_long_r_0 = s->ilkka_du_long_r_0 * (-0.5) * nn * o->M;
            o->v_long_k_0 = s->pair_v_long_k_0 * 0.5 * nn * nn * o->M;
            o->v_long_r_0 = s->pair_v_long_r_0 * (-0.5) * nn * o->M;
        }
    } else if (o->pair_type != SPIMC_PAIR_NONE) {
        delete o;
        return fail(SPIMC_ERR_ARG, "unknown pair type");
    }
    o->rng.seed(seed);
    o->virial_window = o->M;
    o->sk_k_cut = 1e300;
    o->gr_n_r = 100;
    o->gr_r_min = 0.;
    o->gr_r_max = o->L / 2.;
    o->gr_dr = (o->gr_r_max - o->gr_r_min) / (o->gr_n_r - 1.);
    o->gr_d_ir = 1. / o->gr_dr;
    o->sk_acc.assign(o->nk, 0.);
    o->gr_acc.assign(o->gr_n_r, 0.);
    o->init_rho_k();
    *out = o;
    return SPIMC_OK;
}

int spimc_oracle_destroy(void* h) {
    delete (Oracle*)h;
    return SPIMC_OK;
}

// positions r[d][b][p]
int spimc_oracle_path_upload(void* h, const double* r) {
    Oracle* o = (Oracle*)h;
    for (uint32_t d = 0; d < 3; ++d)
        for (uint32_t b = 0; b < o->M; ++b)
            for (uint32_t p = 0; p < o->N; ++p) o->r[3 * ((size_t)b * o->N + p) + d] = r[((size_t)d * o->M + b) * o->N + p];
    o->rc = o->r;
    o->init_rho_k();
    o->need_update_rho_k = true;
    o->have_proposal = false;
    return SPIMC_OK;
}

int spimc_oracle_path_download(void* h, int mode, double* r) {
    Oracle* o = (Oracle*)h;
    const std::vector<double>& src = mode ? o->r : o->rc;
    for (uint32_t d = 0; d < 3; ++d)
        for (uint32_t b = 0; b < o->M; ++b)
            for (uint32_t p = 0; p < o->N; ++p) r[((size_t)d * o->M + b) * o->N + p] = src[3 * ((size_t)b * o->N + p) + d];
    return SPIMC_OK;
}

int spimc_oracle_rho_k_download(void* h, int mode, double* out) {
    Oracle* o = (Oracle*)h;
    const std::vector<cplx>& src = mode ? o->rho : o->rhoc;
    for (size_t i = 0; i < src.size(); ++i) {
        out[2 * i] = src[i].real();
        out[2 * i + 1] = src[i].imag();
    }
    return SPIMC_OK;
}

int spimc_oracle_rho_k_rebuild(void* h) {
    Oracle* o = (Oracle*)h;
    o->r = o->rc;
    o->init_rho_k();
    return SPIMC_OK;
}

int spimc_oracle_set_proposal(void* h, uint32_t p, uint32_t b_first, uint32_t n, const double* r_new) {
    Oracle* o = (Oracle*)h;
    if (p >= o->N || n > o->M) return fail(SPIMC_ERR_ARG, "bad proposal");
    for (uint32_t i = 0; i < n; ++i) {
        double* x = o->Rm(1, p, (b_first + i) % o->M);
        for (int d = 0; d < 3; ++d) x[d] = r_new[3 * i + d];
    }
    o->have_proposal = true;
    o->prop_p = p;
    o->prop_first = b_first % o->M;
    o->prop_n = n;
    o->need_update_rho_k = true;
    return SPIMC_OK;
}

int spimc_oracle_accept_proposal(void* h) {
    Oracle* o = (Oracle*)h;
    if (!o->have_proposal) return fail(SPIMC_ERR_STATE, "no proposal");
    o->store_r_all_beads_of(o->prop_p, o->prop_first, o->prop_n);
    o->store_rho(0, o->M);
    o->need_update_rho_k = true;
    o->have_proposal = false;
    return SPIMC_OK;
}

int spimc_oracle_reject_proposal(void* h) {
    Oracle* o = (Oracle*)h;
    if (!o->have_proposal) return fail(SPIMC_ERR_STATE, "no proposal");
    o->restore_r_all_beads_of(o->prop_p, o->prop_first, o->prop_n);
    o->restore_rho(0, o->M);
    o->need_update_rho_k = true;
    o->have_proposal = false;
    return SPIMC_OK;
}

int spimc_oracle_get_action(void* h, int action, int mode, uint32_t b0, uint32_t b1, const uint32_t* parts,
                            uint32_t n_parts, uint32_t level, double* out) {
    Oracle* o = (Oracle*)h;
    std::vector<uint32_t> pv(parts, parts + n_parts);
    o->mode = mode;
    if (action == SPIMC_ACTION_KINETIC)
        *out = o->kinetic_get_action(b0, b1, pv, level);
    else
        *out = o->pair_get_action(b0, b1, pv, level);
    o->mode = SPIMC_MODE_NEW;
    return SPIMC_OK;
}

// Action::GetActionLaplacian (action_class.h:49) of the Kinetic or pair action
int spimc_oracle_get_action_laplacian(void* h, int action, int mode, uint32_t b0, uint32_t b1, const uint32_t* parts,
                                      uint32_t n_parts, uint32_t level, double* out) {
    Oracle* o = (Oracle*)h;
    std::vector<uint32_t> pv(parts, parts + n_parts);
    o->mode = mode;
    if (action == SPIMC_ACTION_KINETIC)
        *out = o->kinetic_get_action_laplacian(b0, b1, pv, level);
    else
        *out = o->pair_get_action_laplacian(b0, b1, pv, level);
    o->mode = SPIMC_MODE_NEW;
    return SPIMC_OK;
}

int spimc_oracle_d_action_d_beta(void* h, int action, double* out) {
    Oracle* o = (Oracle*)h;
    o->mode = SPIMC_MODE_NEW;
    *out = action == SPIMC_ACTION_KINETIC ? o->kinetic_d_action_d_beta() : o->pair_d_action_d_beta();
    return SPIMC_OK;
}

int spimc_oracle_potential(void* h, int action, double* out) {
    Oracle* o = (Oracle*)h;
    o->mode = SPIMC_MODE_NEW;
    *out = action == SPIMC_ACTION_KINETIC ? 0. : o->pair_potential();
    return SPIMC_OK;
}

int spimc_oracle_kinetic_virial_energy(void* h, uint32_t window, double* out) {
    Oracle* o = (Oracle*)h;
    o->mode = SPIMC_MODE_NEW;
    *out = o->kinetic_virial_energy(window ? window : o->M);
    return SPIMC_OK;
}

// compat=1: the reference's GetActionGradient per bead (pair_action_class.h:305-337),
// link terms and/or long-range term per `terms`;
// compat=0: central difference (step h) of the selected whole-path terms.
int spimc_oracle_action_gradient(void* h, uint32_t terms, int compat, double fd_h, double* out) {
    Oracle* o = (Oracle*)h;
    o->mode = SPIMC_MODE_NEW;
    if (compat) {
        for (uint32_t b = 0; b < o->M; ++b)
            for (uint32_t p = 0; p < o->N; ++p)
                o->pair_gradient_window(b, b + 1, p, out + 3 * ((size_t)b * o->N + p), terms);
        return SPIMC_OK;
    }
    o->gradient_fd(terms, fd_h, out);
    return SPIMC_OK;
}

// The same central difference for a list of beads (slice, particle) only: out[n][3].
int spimc_oracle_action_gradient_beads(void* h, uint32_t terms, double fd_h, uint32_t n, const uint32_t* beads,
                                       double* out) {
    Oracle* o = (Oracle*)h;
    for (uint32_t i = 0; i < n; ++i)
        if (beads[2 * i] >= o->M || beads[2 * i + 1] >= o->N) {
            g_err = "bead out of range";
            return SPIMC_ERR_ARG;
        }
    int saved = o->mode;
    o->mode = SPIMC_MODE_NEW;
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        // each thread perturbs its own bead on a private copy of the path (the
        // difference formulas read the neighbours' positions)
        Oracle local = *o;
        local.fd_bead(terms, fd_h, beads[2 * i], beads[2 * i + 1], out + 3 * i);
    }
    o->mode = saved;
    return SPIMC_OK;
}

// whole-path totals: out[0] kinetic, out[1] pair short range, out[2] long range (per-slice sum, no constants)
int spimc_oracle_action_totals(void* h, double* out) {
    Oracle* o = (Oracle*)h;
    o->mode = SPIMC_MODE_NEW;
    out[0] = o->total_kinetic();
    out[1] = o->total_pair_short();
    double l = 0.;
    if (o->use_long_range && o->pair_type != SPIMC_PAIR_NONE)
        for (uint32_t b = 0; b < o->M; ++b) l += o->slice_long(b, &o->rho[(size_t)b * o->nk]);
    out[2] = l;
    return SPIMC_OK;
}

int spimc_oracle_tape_set(void* h, const double* draws, uint64_t n) {
    Oracle* o = (Oracle*)h;
    if (!draws) {
        o->use_tape = false;
        o->tape.clear();
    } else {
        o->use_tape = true;
        o->tape.assign(draws, draws + n);
    }
    o->cursor = 0;
    return SPIMC_OK;
}

// Action trace of the move stages: 9 doubles per attempt (see Oracle::trace).
// Enabling clears it; read copies up to cap attempts and returns their count.
int spimc_oracle_trace(void* h, int enable) {
    Oracle* o = (Oracle*)h;
    o->tracing = enable != 0;
    o->trace.clear();
    return SPIMC_OK;
}
int spimc_oracle_trace_read(void* h, double* out, uint64_t cap, uint64_t* n) {
    Oracle* o = (Oracle*)h;
    const uint64_t w = Oracle::kTraceWords, have = o->trace.size() / w, k = have < cap ? have : cap;
    if (k) std::memcpy(out, o->trace.data(), k * w * sizeof(double));
    *n = have;
    return SPIMC_OK;
}

// Records every consumed draw (post-transform) while enabled; read with spimc_oracle_recorded.
int spimc_oracle_record(void* h, int enable) {
    Oracle* o = (Oracle*)h;
    o->recording = enable != 0;
    if (enable) o->rec.clear();
    return SPIMC_OK;
}

int spimc_oracle_recorded(void* h, double* out, uint64_t cap, uint64_t* n) {
    Oracle* o = (Oracle*)h;
    *n = o->rec.size();
    if (out) std::copy(o->rec.begin(), o->rec.begin() + std::min<uint64_t>(cap, o->rec.size()), out);
    return SPIMC_OK;
}

int spimc_oracle_tape_cursor(void* h, uint64_t* c) {
    *c = ((Oracle*)h)->cursor;
    return SPIMC_OK;
}

int spimc_oracle_bisect_stage(void* h, uint32_t n_attempts, uint8_t* accepted) {
    Oracle* o = (Oracle*)h;
    g_err.clear();
    for (uint32_t i = 0; i < n_attempts; ++i) {
        uint32_t p, b0;
        o->counters[0]++;
        o->trace_begin();
        bool ok = o->bisect_attempt(p, b0);
        o->trace_end();
        if (ok) {
            o->counters[1]++;
            o->bisect_accept(p, b0);
        } else {
            o->bisect_reject(p, b0);
        }
        o->mode = SPIMC_MODE_NEW;
        o->bead_updates += o->n_bisect_beads - 1;
        if (accepted) accepted[i] = ok ? 1 : 0;
    }
    return g_err.empty() ? SPIMC_OK : SPIMC_ERR_STATE;
}

int spimc_oracle_displace_stage(void* h, uint32_t n_attempts, uint8_t* accepted) {
    Oracle* o = (Oracle*)h;
    g_err.clear();
    for (uint32_t i = 0; i < n_attempts; ++i) {
        uint32_t p;
        o->counters[2]++;
        o->trace_begin();
        bool ok = o->displace_attempt(p);
        o->trace_end();
        if (ok) {
            o->counters[3]++;
            o->displace_accept(p);
        } else {
            o->displace_reject(p);
        }
        o->mode = SPIMC_MODE_NEW;
        o->bead_updates += o->M;
        if (accepted) accepted[i] = ok ? 1 : 0;
    }
    return g_err.empty() ? SPIMC_OK : SPIMC_ERR_STATE;
}

int spimc_oracle_perm_bisect_stage(void* h, uint32_t n_attempts, uint8_t* accepted) {
    Oracle* o = (Oracle*)h;
    g_err.clear();
    // the permuting move bisects at its own level
    const uint32_t saved_level = o->n_level;
    o->n_level = o->perm_n_level;
    o->n_bisect_beads = 1u << o->n_level;
    for (uint32_t i = 0; i < n_attempts; ++i) {
        uint32_t n = 0;
        o->perm_counters[0]++;
        o->trace_begin();
        bool ok = o->perm_bisect_attempt(n);
        o->trace_end();
        if (ok) {
            o->perm_counters[1]++;
            o->perm_attempt[n - 1]++;
            o->perm_accept[n - 1]++;
            o->perm_bisect_accept(o->last_bead0, n);
        } else {
            if (n > 0) o->perm_attempt[n - 1]++;
            o->perm_bisect_reject(o->last_bead0, n);
        }
        o->mode = SPIMC_MODE_NEW;
        o->bead_updates += (uint64_t)n * (o->n_bisect_beads - 1);
        if (accepted) accepted[i] = ok ? 1 : 0;
        if (!g_err.empty()) break;
    }
    o->n_level = saved_level;
    o->n_bisect_beads = 1u << o->n_level;
    return g_err.empty() ? SPIMC_OK : SPIMC_ERR_STATE;
}

// PermBisectTable's stage: the same counters and commits as the iterative move;
// every attempt reaches the bisection with all k chains (n_perm_part =
// c->part.size(), perm_bisect_table_class.h:134), so perm_attempt counts at k - 1.
int spimc_oracle_perm_table_stage(void* h, uint32_t n_attempts, uint8_t* accepted) {
    Oracle* o = (Oracle*)h;
    g_err.clear();
    if (o->table_k == 0) return fail(SPIMC_ERR_STATE, "permutation table not configured");
    const uint32_t saved_level = o->n_level;
    o->n_level = o->perm_n_level;
    o->n_bisect_beads = 1u << o->n_level;
    for (uint32_t i = 0; i < n_attempts; ++i) {
        uint32_t n = 0;
        o->perm_counters[0]++;
        o->trace_begin();
        bool ok = o->perm_table_attempt(n);
        o->trace_end();
        if (ok) {
            o->perm_counters[1]++;
            o->perm_attempt[n - 1]++;
            o->perm_accept[n - 1]++;
            o->perm_bisect_accept(o->last_bead0, n);
        } else {
            o->perm_attempt[n - 1]++;
            o->perm_bisect_reject(o->last_bead0, n);
        }
        o->mode = SPIMC_MODE_NEW;
        o->bead_updates += (uint64_t)n * (o->n_bisect_beads - 1);
        if (accepted) accepted[i] = ok ? 1 : 0;
        if (!g_err.empty()) break;
    }
    o->n_level = saved_level;
    o->n_bisect_beads = 1u << o->n_level;
    return g_err.empty() ? SPIMC_OK : SPIMC_ERR_STATE;
}

// PermBisectTable's n_perm_part (perm_bisect_table_class.h:235) and the rule of
// its final table check (0 the reference's, 1 detailed balance)
int spimc_oracle_perm_table_configure(void* h, uint32_t n_perm_part, uint32_t balance) {
    Oracle* o = (Oracle*)h;
    if (n_perm_part < 1 || n_perm_part > o->N || n_perm_part > 6)
        return fail(SPIMC_ERR_ARG, "n_perm_part must be in 1 .. min(n_part, 6)");
    if (balance > 1) return fail(SPIMC_ERR_ARG, "balance must be 0 or 1");
    o->table_k = n_perm_part;
    o->table_balance = (int)balance;
    return SPIMC_OK;
}

// out[0] attempts, out[1] accepts, out[2 + c] perm_attempt and out[2 + N + c]
// perm_accept of cycle length c + 1 (perm_bisect_class.h:25-26,66-87)
int spimc_oracle_perm_counters(void* h, uint64_t* out) {
    Oracle* o = (Oracle*)h;
    out[0] = o->perm_counters[0];
    out[1] = o->perm_counters[1];
    for (uint32_t c = 0; c < o->N; ++c) {
        out[2 + c] = o->perm_attempt[c];
        out[2 + o->N + c] = o->perm_accept[c];
    }
    return SPIMC_OK;
}

int spimc_oracle_perm_configure(void* h, uint32_t n_level, double epsilon) {
    Oracle* o = (Oracle*)h;
    if (!(epsilon > 0.)) return fail(SPIMC_ERR_ARG, "epsilon must be > 0");
    if (n_level < 1 || n_level > (uint32_t)std::floor(std::log2((double)o->M)))
        return fail(SPIMC_ERR_ARG, "permuting bisection n_level out of range");
    o->perm_log_epsilon = std::log(epsilon);
    o->perm_epsilon = epsilon;
    o->perm_n_level = n_level;
    return SPIMC_OK;
}

// sigma[p]: the particle whose slice-0 bead follows particle p's last bead
int spimc_oracle_perm_get(void* h, uint32_t* sigma) {
    Oracle* o = (Oracle*)h;
    for (uint32_t p = 0; p < o->N; ++p) sigma[p] = o->sigma[p];
    return SPIMC_OK;
}

int spimc_oracle_perm_set(void* h, const uint32_t* sigma) {
    Oracle* o = (Oracle*)h;
    std::vector<char> seen(o->N, 0);
    for (uint32_t p = 0; p < o->N; ++p) {
        if (sigma[p] >= o->N || seen[sigma[p]]) return fail(SPIMC_ERR_ARG, "not a permutation");
        seen[sigma[p]] = 1;
    }
    for (uint32_t p = 0; p < o->N; ++p) {
        o->sigma[p] = sigma[p];
        o->isigma[sigma[p]] = p;
    }
    return SPIMC_OK;
}

// cycle lengths of the path's permutation in particle order (Species::SetCycleCount)
int spimc_oracle_cycle_lengths(void* h, uint32_t* out, uint32_t* n) {
    Oracle* o = (Oracle*)h;
    auto c = o->cycle_lengths();
    for (size_t i = 0; i < c.size(); ++i) out[i] = c[i];
    *n = (uint32_t)c.size();
    return SPIMC_OK;
}

int spimc_oracle_move_counters(void* h, uint64_t* out) {
    Oracle* o = (Oracle*)h;
    for (int i = 0; i < 4; ++i) out[i] = o->counters[i];
    return SPIMC_OK;
}

int spimc_oracle_bead_updates(void* h, uint64_t* out) {
    *out = ((Oracle*)h)->bead_updates;
    return SPIMC_OK;
}

int spimc_oracle_obs_configure(void* h, int energy_estimator, uint32_t virial_window, double sk_k_cut, double gr_r_min,
                               double gr_r_max, uint32_t gr_n_r) {
    Oracle* o = (Oracle*)h;
    if (gr_n_r < 2) return fail(SPIMC_ERR_ARG, "n_r must be >= 2");
    o->energy_estimator = energy_estimator;
    o->virial_window = virial_window ? virial_window : o->M;
    o->sk_k_cut = sk_k_cut;
    o->gr_r_min = gr_r_min;
    o->gr_r_max = gr_r_max > 0. ? gr_r_max : o->L / 2.;
    o->gr_n_r = gr_n_r;
    // LinearGrid::CreateGrid (observable_class.h:29-38)
    o->gr_dr = (o->gr_r_max - o->gr_r_min) / (gr_n_r - 1.);
    o->gr_d_ir = 1. / o->gr_dr;
    o->gr_acc.assign(gr_n_r, 0.);
    o->sk_acc.assign(o->nk, 0.);
    o->e_acc[0] = o->e_acc[1] = 0.;
    for (auto& m : o->n_measure) m = 0;
    return SPIMC_OK;
}

int spimc_oracle_obs_accumulate(void* h, uint32_t which) {
    Oracle* o = (Oracle*)h;
    if (which & SPIMC_OBS_ENERGY) o->energy_accumulate();
    if (which & SPIMC_OBS_SK) o->sk_accumulate();
    if (which & SPIMC_OBS_GR) o->gr_accumulate();
    return SPIMC_OK;
}

int spimc_oracle_obs_read(void* h, uint32_t which, double* out, uint64_t* n_measure) {
    Oracle* o = (Oracle*)h;
    if (which == SPIMC_OBS_ENERGY) {
        out[0] = o->e_acc[0];
        out[1] = o->e_acc[1];
        *n_measure = o->n_measure[0];
    } else if (which == SPIMC_OBS_SK) {
        std::copy(o->sk_acc.begin(), o->sk_acc.end(), out);
        *n_measure = o->n_measure[1];
    } else if (which == SPIMC_OBS_GR) {
        std::copy(o->gr_acc.begin(), o->gr_acc.end(), out);
        *n_measure = o->n_measure[2];
    } else {
        return fail(SPIMC_ERR_ARG, "read one observable at a time");
    }
    return SPIMC_OK;
}

// One sweep = N*M/2^n_level bisect attempts, N displace attempts, then estimators.
int spimc_oracle_sweep(void* h, uint32_t n_sweeps, uint32_t flags) {
    Oracle* o = (Oracle*)h;
    uint32_t n_bis = o->N * o->M / o->n_bisect_beads;
    for (uint32_t s = 0; s < n_sweeps; ++s) {
        int rc = spimc_oracle_bisect_stage(h, n_bis, nullptr);
        if (rc) return rc;
        rc = spimc_oracle_displace_stage(h, o->N, nullptr);
        if (rc) return rc;
        uint32_t which = 0;
        if (flags & SPIMC_SWEEP_ENERGY) which |= SPIMC_OBS_ENERGY;
        if (flags & SPIMC_SWEEP_SK) which |= SPIMC_OBS_SK;
        if (flags & SPIMC_SWEEP_GR) which |= SPIMC_OBS_GR;
        spimc_oracle_obs_accumulate(h, which);
    }
    return SPIMC_OK;
}

// Natural spline evaluation helper for golden-vector tests of the spline itself.
int spimc_oracle_spline_eval(const double* x, const double* y, uint32_t n, const double* t, uint32_t nt, double* out,
                             double* dout) {
    NatSpline s;
    s.build(x, y, n, false);
    for (uint32_t i = 0; i < nt; ++i) {
        out[i] = s.eval(t[i]);
        if (dout) dout[i] = s.deriv(t[i]);
    }
    return SPIMC_OK;
}

// Bisect::Reset with adaptive="1" (bisect_class.h:143-151): the next level.
int spimc_oracle_set_bisect_level(void* h, uint32_t n_level) {
    Oracle* o = (Oracle*)h;
    if (n_level < 1 || n_level > (uint32_t)std::floor(std::log2((double)o->M))) {
        g_err = "bisect n_level out of range";
        return SPIMC_ERR_ARG;
    }
    o->n_level = n_level;
    o->n_bisect_beads = 1u << n_level;
    return SPIMC_OK;
}

// Tensor natural spline through vals[ny][nx] at n points: values and partials.
int spimc_oracle_spline2_eval(const double* x, uint32_t nx, const double* y, uint32_t ny, const double* vals,
                              const double* px, const double* py, uint32_t n, double* out, double* gx, double* gy) {
    NatSpline2 s;
    s.build(x, nx, y, ny, vals);
    for (uint32_t i = 0; i < n; ++i) out[i] = s.eval(px[i], py[i], gx ? gx + i : nullptr, gy ? gy + i : nullptr);
    return SPIMC_OK;
}

// One link of the pair action from its lengths: which 0 CalcU, 1 CalcdUdBeta, 2 CalcV.
int spimc_oracle_pair_link(void* h, int which, double r0, double r1, double s, double* out) {
    Oracle* o = (Oracle*)h;
    *out = which == 0 ? o->calc_u(r0, r1, s, 0) : which == 1 ? o->calc_du_dbeta(r0, r1, s) : o->calc_v(r0, r1);
    return SPIMC_OK;
}

// FreeSpline::GetLogRhoFree helper for kinetic tests.
int spimc_oracle_free_spline_log_rho(double L, int n_images, double lambda, double tau, const double* r, uint32_t n,
                                     double* out) {
    FreeSpline f;
    f.build(L, n_images, lambda, tau, false);
    for (uint32_t i = 0; i < n; ++i) out[i] = f.log_rho_free(r + 3 * i);
    return SPIMC_OK;
}

}  // extern "C"
