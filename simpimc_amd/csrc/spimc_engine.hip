// spimc_engine.hip -- context management and C ABI (include/spimc.h).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/spimc.h"
#include "spimc_perm.h"
#include "spimc_stage.h"

using namespace spimc;

// Device scratch of spimc_get_action_noalloc (long-range partials, the result
// and the work items of one query), so that the call allocates nothing.  Every
// device the library runs on holds its own copy until the process ends; calls
// on one device share it.  In host code the name is only a stand-in: the
// device address comes from hipGetSymbolAddress (noalloc_scratch_ptr).
constexpr size_t kNoAllocScratchBytes = SPIMC_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// Per-device address of `scratch`, looked up once per device with that device
// current (at context creation).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(SPIMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int noalloc_scratch_ptr(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return fail(SPIMC_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(SPIMC_ERR_HIP, "hipGetSymbolAddress returned no address for the device scratch");
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return SPIMC_OK;
}

constexpr int kRhoKpt = 8;        // k-vectors per thread in k_rho (default; SPIMC_RHO_KPT 2, 4, 8, 16)
constexpr int kRhoKptMin = 2;     // smallest chunking: bounds the number of k chunks
constexpr int kNumCUs = 256;      // MI355X compute units (for workgroups-per-replica choices)
// fewest waves for which the pair-once adjoint beats the per-particle one
constexpr long long kAdjPairsMinWaves = 4096;
// PermBisectTable: particles of a replica (two N x N tables) and cycle-list entries
constexpr int kTabMaxN = 1024;
constexpr double kTabMaxEntries = 268435456.;
constexpr int kBuildKpt = 2;
constexpr int kRoundsPerGraph = 16;
constexpr int kAdjMaxIdx = 40;

uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// Natural cubic spline through (x_i, y_i) in per-interval polynomial form
// c[i][comp][0..3] (stride ncomp*4), s(t) = c0 + dt (c1 + dt (c2 + dt c3)).
void spline_coefs(const double* x, const double* y, int n, int ncomp, int comp, double* c) {
    std::vector<double> m(n, 0.);
    if (n >= 3) {
        std::vector<double> cp(n, 0.), dp(n, 0.);
        for (int i = 1; i < n - 1; ++i) {
            const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
            const double rhs = 6. * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
            const double den = 2. * (h0 + h1) - h0 * cp[i - 1];
            cp[i] = h1 / den;
            dp[i] = (rhs - h0 * dp[i - 1]) / den;
        }
        for (int i = n - 2; i >= 1; --i) m[i] = dp[i] - cp[i] * m[i + 1];
    }
    for (int i = 0; i < n - 1; ++i) {
        const double h = x[i + 1] - x[i];
        double* ci = c + ((size_t)i * ncomp + comp) * 4;
        ci[0] = y[i];
        ci[1] = (y[i + 1] - y[i]) / h - h * (2. * m[i] + m[i + 1]) / 6.;
        ci[2] = 0.5 * m[i];
        ci[3] = (m[i + 1] - m[i]) / (6. * h);
    }
}

// B-spline coefficients of the natural cubic spline through (x_i, y_i) on a
// uniform grid: c_{i+1} = y_i - h^2 m_i / 6 (m = second derivatives, m_0 =
// m_{n-1} = 0), c_0 = 2 c_1 - c_2, c_{n+1} = 2 c_n - c_{n-1}.
void bspline_coefs(const double* x, const double* y, int n, double* c) {
    std::vector<double> m(n, 0.);
    if (n >= 3) {
        std::vector<double> cp(n, 0.), dp(n, 0.);
        for (int i = 1; i < n - 1; ++i) {
            const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
            const double rhs = 6. * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
            const double den = 2. * (h0 + h1) - h0 * cp[i - 1];
            cp[i] = h1 / den;
            dp[i] = (rhs - h0 * dp[i - 1]) / den;
        }
        for (int i = n - 2; i >= 1; --i) m[i] = dp[i] - cp[i] * m[i + 1];
    }
    const double h = (x[n - 1] - x[0]) / (n - 1);
    for (int i = 0; i < n; ++i) c[i + 1] = y[i] - h * h * m[i] / 6.;
    c[0] = 2. * c[1] - c[2];
    c[n + 1] = 2. * c[n] - c[n - 1];
}

bool grid_is_uniform(const double* x, int n) {
    if (n < 3) return true;
    const double h = (x[n - 1] - x[0]) / (n - 1);
    for (int i = 0; i < n; ++i)
        if (std::fabs(x[i] - (x[0] + i * h)) > 1e-9 * std::fabs(h)) return false;
    return true;
}

// Host-side spline staged for upload.
struct SplHost {
    std::vector<double> x, c;
    int n = 0, ncomp = 0;
    bool uniform = false;
};

SplHost make_spline(const double* x, int n, int ncomp, const std::vector<const double*>& ys) {
    SplHost s;
    s.n = n;
    s.ncomp = ncomp;
    s.x.assign(x, x + n);
    s.c.assign((size_t)(n - 1) * ncomp * 4, 0.);
    for (int k = 0; k < ncomp; ++k) spline_coefs(x, ys[k], n, ncomp, k, s.c.data());
    s.uniform = grid_is_uniform(x, n);
    return s;
}

// Tensor-product natural cubic spline through vals[ny][nx] on x[nx] x y[ny]
// (einspline create_NUBspline_2d_d with NATURAL conditions on both axes,
// ilkka_pair_action_class.h:276-278): a natural spline along x through each
// row gives, per x interval, four polynomial coefficients at every y knot; a
// natural spline along y through each of them gives the cell polynomials.
struct Spl2Host {
    std::vector<double> x, y, c;
    int nx = 0, ny = 0;
};

Spl2Host make_spline2(const double* x, int nx, const double* y, int ny, const double* vals) {
    Spl2Host s;
    s.nx = nx;
    s.ny = ny;
    s.x.assign(x, x + nx);
    s.y.assign(y, y + ny);
    s.c.assign((size_t)(nx - 1) * (ny - 1) * 16, 0.);
    std::vector<double> ax((size_t)ny * (nx - 1) * 4);  // [row j][interval i][power a]
    for (int j = 0; j < ny; ++j) spline_coefs(x, vals + (size_t)j * nx, nx, 1, 0, ax.data() + (size_t)j * (nx - 1) * 4);
    std::vector<double> col(ny), cy((size_t)(ny - 1) * 4);
    for (int i = 0; i < nx - 1; ++i)
        for (int a = 0; a < 4; ++a) {
            for (int j = 0; j < ny; ++j) col[j] = ax[((size_t)j * (nx - 1) + i) * 4 + a];
            spline_coefs(y, col.data(), ny, 1, 0, cy.data());
            for (int j = 0; j < ny - 1; ++j)
                for (int b = 0; b < 4; ++b) s.c[((size_t)j * (nx - 1) + i) * 16 + 4 * a + b] = cy[(size_t)j * 4 + b];
        }
    return s;
}

// FreeSpline tables (free_spline_class.h:25-63): image action or its tau derivative
// on 10000 uniform points over [-L/2, L/2].
SplHost make_free_spline(double L, int n_images, double lambda, double tau, bool dtau) {
    const int num = 10000;
    const double t_l = (L == 0.) ? 1000. : L;
    const double start = -t_l / 2., end = t_l / 2., dr = (end - start) / (num - 1);
    const double i4lt = 1. / (4. * lambda * tau);
    std::vector<double> xs(num), ia(num, 0.), dia(num, 0.);
    for (int i = 0; i < num; ++i) {
        const double r = start + i * dr;
        xs[i] = r;
        const double r2l = r * r * i4lt;
        for (int im = 1; im <= n_images; ++im) {
            const double rp = r + im * t_l, rm = r - im * t_l;
            const double dp = r2l - rp * rp * i4lt, dm = r2l - rm * rm * i4lt;
            const double ep = std::exp(dp), em = std::exp(dm);
            ia[i] += ep + em;
            dia[i] += (dp * ep + dm * em) / tau;
        }
        dia[i] = dia[i] / (1. + ia[i]);
        ia[i] = -std::log1p(ia[i]);
    }
    return make_spline(xs.data(), num, 1, {dtau ? dia.data() : ia.data()});
}

// Bump allocator over one device arena (sizing pass with base == nullptr).
struct Carver {
    char* base = nullptr;
    size_t off = 0;
    template <class T>
    T* take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T* p = base ? (T*)(base + off) : nullptr;
        off += n * sizeof(T);
        return p;
    }
};

struct BTabHost {
    std::vector<double> c;
    int knots = 0, ncomp = 0;
    double x0 = 0., ih = 0.;
};

BTabHost make_btab(const double* x, int n, const std::vector<const double*>& ys) {
    BTabHost t;
    t.knots = n;
    t.ncomp = (int)ys.size();
    t.x0 = x[0];
    t.ih = (n - 1) / (x[n - 1] - x[0]);
    t.c.assign((size_t)(n + 2) * t.ncomp, 0.);
    for (int k = 0; k < t.ncomp; ++k) bspline_coefs(x, ys[k], n, t.c.data() + (size_t)k * (n + 2));
    for (double& v : t.c) v *= 1. / 6.;  // the device basis weights omit the 1/6
    return t;
}

constexpr size_t kPairLdsMaxBytes = 150 * 1024;

struct RepHost {
    spimc_system_desc d{};
    BTabHost bt_a, bt_b, bt_du;
    std::vector<int32_t> k_n;
    std::vector<int> kidx;
    std::vector<int4> k_runs;  // runs of consecutive n2 at fixed (n0, n1), ordered by (n0, n1)
    std::vector<int> k_order;  // k index of each k-vector in run order
    std::vector<double> kmag, coef_s, coef_b, coef_v;
    SplHost v_r, v_long_r, u, du, u_long_r, du_long_r;
    Spl2Host u2, du2;
    std::vector<SplHost> kin, bis;
    SplHost kin_dtau;
    Rep dev{};
    int const_owner = -1;  // replica whose arena holds this replica's constant tables
    uint64_t const_hash = 0;
    char* arena = nullptr;
    unsigned char* accepted = nullptr;
    double* tape = nullptr;
    unsigned long long* gr_cnt = nullptr;
    uint32_t accepted_cap = 0;
    double* trace = nullptr;        // action trace, [accepted_cap][kTraceWords] (when on)
    uint32_t stage_attempts = 0;    // attempts of this replica in the last move stage
    int n_chunks = 1;
    char* perm_arena = nullptr;    // perm, iperm
    char* perm_scratch = nullptr;  // the permuting move's chain arrays and counters
    double* perm_t = nullptr;      // PermBisectTable's two N x N tables
};

// Replicas with identical constant tables (k-set, Ewald coefficients, pair and
// free-particle splines) share one device copy, so a batch of clones of one
// system keeps a single set of tables hot in L2 instead of one per replica.
void hash_bytes(uint64_t& h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001B3ull;
}
template <class T>
void hash_vec(uint64_t& h, const std::vector<T>& v) {
    const size_t n = v.size();
    hash_bytes(h, &n, sizeof(n));
    if (n) hash_bytes(h, v.data(), n * sizeof(T));
}
void hash_spline(uint64_t& h, const SplHost& s) {
    hash_vec(h, s.x);
    hash_vec(h, s.c);
    const int m[3] = {s.n, s.ncomp, s.uniform ? 1 : 0};
    hash_bytes(h, m, sizeof(m));
}
void hash_btab(uint64_t& h, const BTabHost& t) {
    hash_vec(h, t.c);
    const int m[2] = {t.knots, t.ncomp};
    hash_bytes(h, m, sizeof(m));
    hash_bytes(h, &t.x0, sizeof(double));
    hash_bytes(h, &t.ih, sizeof(double));
}
bool same_spline(const SplHost& a, const SplHost& b) {
    return a.n == b.n && a.ncomp == b.ncomp && a.uniform == b.uniform && a.x == b.x && a.c == b.c;
}
bool same_btab(const BTabHost& a, const BTabHost& b) {
    return a.knots == b.knots && a.ncomp == b.ncomp && a.x0 == b.x0 && a.ih == b.ih && a.c == b.c;
}
uint64_t constants_hash(const RepHost& h) {
    uint64_t x = 0xCBF29CE484222325ull;
    hash_vec(x, h.kidx);
    hash_vec(x, h.kmag);
    hash_vec(x, h.coef_s);
    hash_vec(x, h.coef_b);
    hash_vec(x, h.coef_v);
    for (const SplHost* s : {&h.v_r, &h.v_long_r, &h.u, &h.du, &h.u_long_r, &h.du_long_r, &h.kin_dtau}) hash_spline(x, *s);
    for (const Spl2Host* s : {&h.u2, &h.du2}) {
        hash_vec(x, s->x);
        hash_vec(x, s->y);
        hash_vec(x, s->c);
    }
    for (const auto& s : h.kin) hash_spline(x, s);
    for (const auto& s : h.bis) hash_spline(x, s);
    for (const BTabHost* t : {&h.bt_a, &h.bt_b, &h.bt_du}) hash_btab(x, *t);
    return x;
}
bool same_constants(const RepHost& a, const RepHost& b) {
    if (a.const_hash != b.const_hash || a.kidx != b.kidx || a.kmag != b.kmag || a.coef_s != b.coef_s ||
        a.coef_b != b.coef_b || a.coef_v != b.coef_v || a.kin.size() != b.kin.size() || a.bis.size() != b.bis.size())
        return false;
    if (!same_spline(a.v_r, b.v_r) || !same_spline(a.v_long_r, b.v_long_r) || !same_spline(a.u, b.u) ||
        !same_spline(a.du, b.du) || !same_spline(a.u_long_r, b.u_long_r) || !same_spline(a.du_long_r, b.du_long_r) ||
        !same_spline(a.kin_dtau, b.kin_dtau))
        return false;
    for (auto pr : {std::make_pair(&a.u2, &b.u2), std::make_pair(&a.du2, &b.du2)})
        if (pr.first->x != pr.second->x || pr.first->y != pr.second->y || pr.first->c != pr.second->c) return false;
    for (size_t l = 0; l < a.kin.size(); ++l)
        if (!same_spline(a.kin[l], b.kin[l])) return false;
    for (size_t l = 0; l < a.bis.size(); ++l)
        if (!same_spline(a.bis[l], b.bis[l])) return false;
    return same_btab(a.bt_a, b.bt_a) && same_btab(a.bt_b, b.bt_b) && same_btab(a.bt_du, b.bt_du);
}

struct TimedLaunch {
    int kernel;
    hipEvent_t a, b;
};

}  // namespace

// A contiguous range of replicas whose move rounds run on a stream of their
// own, so that one group's VALU-bound pair kernel and another group's
// HBM-bound rho_k kernel (and the one-wavefront propose/decide kernels) share
// the GPU at the same time.  Group 0 runs on the context stream.
struct Group {
    hipStream_t stream = nullptr;
    int r0 = 0, n = 0;
    int max_bis_links = 0, max_M = 0;  // largest bisection window and slice count in the group
    // pair work lists (absolute replica indices; global-table pair kernel)
    uint2 *d_pair_bis = nullptr, *d_pair_dis = nullptr;
    int n_pair_bis = 0, n_pair_dis = 0;
    // rho_k work items, written by the proposal kernels: the group's replicas' items in order
    RhoItem *d_ritem_bis = nullptr, *d_ritem_dis = nullptr;
    int n_ritem_bis = 0, n_ritem_dis = 0;
    // captured round graphs and whether each was captured with the device clocks on
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    bool graph_clock[2] = {false, false};
    unsigned long long* d_kt = nullptr;  // device clock slots of this group's rounds
    hipEvent_t done = nullptr;
    // second stream of the group: the rho_k kernel of a round runs on it beside
    // the pair kernel (both need only the proposal), joined before the decision
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // permuting rounds: the proposals for the label-general kernels (2 counters, 2 lists of n)
    unsigned int* d_multi = nullptr;
    // one-launch bisection rounds: pair workgroups per replica, rho_k items per
    // workgroup, the rho_k workgroups' items (first, count, group-local replica),
    // and the replicas' arrival counters (zero between rounds)
    int round_bpr[2] = {1, 1}, round_ipb[2] = {1, 1}, n_rgrp[2] = {0, 0};  // per move
    int4* d_rgrp[2] = {nullptr, nullptr};
    unsigned int* d_arrive = nullptr;
};

struct spimc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<RepHost> reps;
    Rep* d_reps = nullptr;
    std::vector<Group> groups;
    hipEvent_t stage_ready = nullptr;
    uint2* d_items_tmp = nullptr;
    size_t items_tmp_cap = 0;
    unsigned int* d_counts = nullptr;
    double* d_tmp = nullptr;  // interface partials + result
    size_t tmp_cap = 0;
    bool any_long = false;
    int max_M = 0, max_N = 0, max_nk = 0;
    // interface proposal
    bool prop_set = false, rho_updated = false;
    int prop_rep = 0, prop_p = 0, prop_first = 0, prop_count = 0;
    // observables
    int energy_estimator = SPIMC_ENERGY_THERMAL;
    uint32_t virial_window = 0;
    double sk_k_cut = 1e300, gr_r_min = 0., gr_r_max = 0.;
    uint32_t gr_n_r = 100;
    double* d_gr_rmax = nullptr;
    bool adjoint_fresh = false;
    // LDS bytes/8 of the largest pair table when every replica's tables stage through LDS, else 0
    int lds_doubles_max = 0;
    // the LDS pair kernel specialised for the replicas' pair type (mixed types: generic)
    void (*pair_lds_fn)(const Rep*, int, int, unsigned long long*) = nullptr;
    int pair_block = kPairLdsBlock;
    int pair_variant = 3;  // 0 Bare, 1 David order 2, 2 David other order, 3 mixed
    void (*adj_lds_fn)(const Rep*, unsigned int, int) = nullptr;
    // the pair-once adjoint (two launches), used when its link ranges give the
    // GPU enough waves; SPIMC_ADJ_PAIRS = 0 / 1 forces the choice
    void (*adj_pairs_fn)(const Rep*, unsigned int, int, int) = nullptr;
    int adj_pairs_env = -1;
    bool perm_table = false;  // the running permuting stage is PermBisectTable's
    bool adj_pairs_ok = false;  // its tables and slots fit in LDS
    int groups_wanted = 1;
    bool fork = false;  // rho_k beside the pair kernel on a second stream per group
    int perm_stage = kPermStage;  // largest permuting proposal (chain beads) bisected on LDS
    // k_rho variant: k-vectors per thread and non-temporal rho_k traffic
    int rho_kpt = kRhoKpt;
    bool rho_nt = true;  // measured: slightly faster, and keeps the pair tables in L2
    void (*rho_fn)(const Rep*, const uint2*, int, int, unsigned long long*) = nullptr;
    // looping variant with rho_bpc workgroups per CU (0: one workgroup per item)
    int rho_bpc = 0;
    void (*rho_loop_fn)(const RhoItem*, int, unsigned long long*) = nullptr;
    // LDS doubles of the adjoint kernel (pair tables) and of the whole-path
    // estimator kernel (Bare: pair tables, David: du table); 0 = global tables
    int lds_adj_max = 0, lds_whole_max = 0;
    // dynamic LDS bytes of the run-ordered long-range adjoint (0: the per-particle table kernel)
    size_t adj_runs_lds = 0;
    // timing: per-launch dispatch events (every kernel) and device clocks (pair, rho_k)
    bool timing = false, clock = false;
    double clock_hz = 0.;
    unsigned long long kt_sum[2] = {0, 0}, kt_launches[2] = {0, 0};
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> pool;
    double t_ms[SPIMC_KERNEL_COUNT] = {0};
    uint64_t t_n[SPIMC_KERNEL_COUNT] = {0};
    double t_bytes[SPIMC_KERNEL_COUNT] = {0};
    unsigned long long* d_work = nullptr;
    unsigned int* d_error = nullptr;
    unsigned long long work_seen[WORK_COUNT] = {0, 0};
    // a replica's path may be permuted: rounds also launch the label-general pair kernel
    bool perm_on = false;
    // a round's decision runs in the next round's proposal launch (SPIMC_FUSE_ROUNDS=0: its own launch)
    bool fuse_rounds = true;
    // the move stages record each attempt's level-0 action terms (spimc_action_trace_enable)
    bool trace_on = false;
    // device address of this device's `scratch` array (spimc_get_action_noalloc)
    void* noalloc_scratch = nullptr;
    // persistent stage kernel for small replicas: its workgroup size (0: stages run in rounds)
    int pers_block = 0;
    void (*pers_fn)(const Rep*, int, int, int, int) = nullptr;
    hipGraphExec_t pers_graph[2] = {nullptr, nullptr};  // stage begin + persistent stage, per move
    // a round in one launch (k_round: pair links, rho_k items, then the replica's
    // decision and next proposal by its last workgroup), for small batches
    bool one_launch = false;
    void (*round_fn[2])(const Rep*, int, int, const RhoItem*, const int4*, int, unsigned int*) = {nullptr, nullptr};
    int one_moves = 1;  // one-launch rounds for bisection (1), or for both moves (2)
    std::string noalloc_error;  // why the address lookup failed, if it did
};

namespace {

hipEvent_t take_event(spimc_ctx* c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Launch helper.  With timing on, the start and stop events are written by the
// dispatch packet itself (hipExtLaunchKernel), so an event pair measures the
// kernel's own execution, as a kernel trace does, not the gaps around it.
template <class... P, class... A>
int launch(spimc_ctx* c, int kernel, void (*k)(P...), dim3 grid, dim3 block, size_t shmem, hipStream_t st,
           A&&... a) {
    std::tuple<std::decay_t<P>...> args{static_cast<std::decay_t<P>>(a)...};
    void* ptrs[sizeof...(P) > 0 ? sizeof...(P) : 1];
    std::apply([&](auto&... x) {
        int i = 0;
        ((ptrs[i++] = static_cast<void*>(&x)), ...);
    }, args);
    hipEvent_t ea = nullptr, eb = nullptr;
    if (c->timing) {
        ea = take_event(c);
        eb = take_event(c);
    }
    hipError_t e;
    if (ea && eb)
        e = hipExtLaunchKernel(reinterpret_cast<const void*>(k), grid, block, ptrs, shmem, st, ea, eb, 0);
    else
        e = hipLaunchKernel(reinterpret_cast<const void*>(k), grid, block, ptrs, shmem, st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return fail(SPIMC_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    if (ea && eb) c->pending.push_back({kernel, ea, eb});
    return SPIMC_OK;
}

// Work done by the move kernels so far, summed over replicas on the device.
int work_totals(spimc_ctx* c, unsigned long long* w) {
    k_work_sum<<<1, 64, 0, c->stream>>>(c->d_reps, (int)c->reps.size(), c->d_work);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(w, c->d_work, WORK_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return SPIMC_OK;
}

int sync_and_resolve(spimc_ctx* c) {
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& t : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            c->t_ms[t.kernel] += ms;
            c->t_n[t.kernel] += 1;
        }
        c->pool.push_back(t.a);
        c->pool.push_back(t.b);
    }
    c->pending.clear();
    if (c->clock) {
        for (int k = 0; k < KT_CLASSES; ++k) c->kt_sum[k] = c->kt_launches[k] = 0;
        for (auto& G : c->groups) {
            unsigned long long kt[KT_END];
            HIPCHK(hipMemcpy(kt, G.d_kt, sizeof(kt), hipMemcpyDeviceToHost));
            for (int k = 0; k < KT_CLASSES; ++k) {
                c->kt_sum[k] += kt[2 + 2 * k];
                c->kt_launches[k] += kt[3 + 2 * k];
            }
        }
    }
    if ((c->timing || c->clock) && c->d_work) {
        unsigned long long w[WORK_COUNT];
        int rc = work_totals(c, w);
        if (rc) return rc;
        c->t_bytes[SPIMC_KERNEL_RHO] += 32.0 * (double)(w[WORK_RHO_SLICEK] - c->work_seen[WORK_RHO_SLICEK]);
        c->t_bytes[SPIMC_KERNEL_PAIR] += (double)(w[WORK_PAIR_LINKS] - c->work_seen[WORK_PAIR_LINKS]);
        for (int i = 0; i < WORK_COUNT; ++i) c->work_seen[i] = w[i];
    }
    return SPIMC_OK;
}

int check_rep(spimc_ctx* c, uint32_t r) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    if (r >= c->reps.size()) return fail(SPIMC_ERR_ARG, "replica index out of range");
    return SPIMC_OK;
}

int validate(const spimc_system_desc& s, uint32_t r) {
    char buf[160];
    auto bad = [&](int code, const char* what) {
        snprintf(buf, sizeof buf, "replica %u: %s", r, what);
        return fail(code, buf);
    };
    if (s.n_d != 3) return bad(SPIMC_ERR_UNSUPPORTED, "n_d must be 3");
    if (s.n_bead < 2 || s.n_part < 1) return bad(SPIMC_ERR_ARG, "n_bead >= 2 and n_part >= 1 required");
    if (!(s.L > 0.)) return bad(SPIMC_ERR_UNSUPPORTED, "periodic box with L > 0 required");
    if (!(s.beta > 0.) || !(s.lambda > 0.)) return bad(SPIMC_ERR_ARG, "beta and lambda must be positive");
    int max_lvl = (int)std::floor(std::log2((double)s.n_bead));
    if (s.bisect_n_level < 1 || (int)s.bisect_n_level > std::min(max_lvl, kMaxLevel))
        return bad(SPIMC_ERR_UNSUPPORTED, "bisect n_level must be in [1, min(6, floor(log2 n_bead))]");
    if (s.pair_max_level != 0) return bad(SPIMC_ERR_UNSUPPORTED, "pair action max_level must be 0");
    if (s.kinetic_n_images < 0 || s.bisect_n_images < 0) return bad(SPIMC_ERR_ARG, "n_images must be >= 0");
    if (s.pair_type < SPIMC_PAIR_NONE || s.pair_type > SPIMC_PAIR_ILKKA) return bad(SPIMC_ERR_ARG, "unknown pair type");
    if (s.pair_type == SPIMC_PAIR_ILKKA) {
        if (!s.ilkka_x_u || !s.ilkka_y_u || !s.ilkka_u_xy || s.ilkka_n_x_u < 2 || s.ilkka_n_y_u < 2 || !s.ilkka_x_du ||
            !s.ilkka_y_du || !s.ilkka_du_xy || s.ilkka_n_x_du < 2 || s.ilkka_n_y_du < 2)
            return bad(SPIMC_ERR_ARG, "ilkka u/off_diag or du/off_diag table missing");
        if (!s.ilkka_r_v || !s.ilkka_v_r || s.ilkka_n_r_v < 3) return bad(SPIMC_ERR_ARG, "ilkka v/diag table missing");
        if (s.pair_use_long_range &&
            (!s.ilkka_r_v_long || !s.ilkka_v_long_r || s.ilkka_n_r_v_long < 3 || !s.ilkka_r_u_long || !s.ilkka_u_long_r ||
             s.ilkka_n_r_u_long < 3 || !s.ilkka_r_du_long || !s.ilkka_du_long_r || s.ilkka_n_r_du_long < 3 ||
             !s.ilkka_u_long_k || !s.ilkka_du_long_k))
            return bad(SPIMC_ERR_ARG, "ilkka long-range tables missing");
        auto increasing = [](const double* x, uint32_t n) {
            for (uint32_t i = 1; i < n; ++i)
                if (!(x[i] > x[i - 1])) return false;
            return true;
        };
        if (!increasing(s.ilkka_x_u, s.ilkka_n_x_u) || !increasing(s.ilkka_y_u, s.ilkka_n_y_u) ||
            !increasing(s.ilkka_x_du, s.ilkka_n_x_du) || !increasing(s.ilkka_y_du, s.ilkka_n_y_du) ||
            !increasing(s.ilkka_r_v, s.ilkka_n_r_v))
            return bad(SPIMC_ERR_ARG, "ilkka grids must increase strictly");
    }
    if (s.pair_type == SPIMC_PAIR_BARE) {
        if (!s.bare_r || !s.bare_v_r || s.bare_n_r < 3) return bad(SPIMC_ERR_ARG, "bare v/diag table missing");
        if (s.pair_use_long_range && (!s.bare_r_long || !s.bare_v_long_r || s.bare_n_r_long < 3))
            return bad(SPIMC_ERR_ARG, "bare long-range table missing");
    }
    if (s.pair_type == SPIMC_PAIR_DAVID) {
        if (!s.david_grid || !s.david_u_kj || !s.david_du_kj || s.david_n_grid < 3)
            return bad(SPIMC_ERR_ARG, "david table missing");
        if (s.david_n_order > 3) return bad(SPIMC_ERR_UNSUPPORTED, "david n_order must be <= 3");
    }
    if (s.pair_type != SPIMC_PAIR_NONE && s.pair_use_long_range) {
        if (s.n_k == 0 || !s.k_n || !s.pair_v_long_k) return bad(SPIMC_ERR_ARG, "k-vectors and v_long_k required");
        for (uint32_t k = 0; k < 3 * s.n_k; ++k)
            if (std::abs(s.k_n[k]) > kMaxIdx) return bad(SPIMC_ERR_UNSUPPORTED, "|n| of a k-vector must be <= 63");
        if ((s.n_k + kRhoBlock * kRhoKptMin - 1) / (kRhoBlock * kRhoKptMin) > 255)
            return bad(SPIMC_ERR_UNSUPPORTED, "too many k-vectors");
    }
    return SPIMC_OK;
}

Spl carve_spline(Carver& cv, const SplHost& h) {
    Spl s{};
    s.n = h.n;
    s.ncomp = h.ncomp;
    s.uniform = h.uniform ? 1 : 0;
    if (h.n > 1) {
        s.x0 = h.x[0];
        s.inv_h = (h.n - 1) / (h.x[h.n - 1] - h.x[0]);
    }
    s.x = cv.take<double>(h.x.size());
    s.c = cv.take<double>(h.c.size());
    return s;
}

int upload_spline(const Spl& s, const SplHost& h) {
    if (!h.n) return SPIMC_OK;
    HIPCHK(hipMemcpy((void*)s.x, h.x.data(), h.x.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((void*)s.c, h.c.data(), h.c.size() * sizeof(double), hipMemcpyHostToDevice));
    return SPIMC_OK;
}

// Builds the host tables of one replica and lays out its arena.
int build_replica(spimc_ctx* c, uint32_t r, const spimc_system_desc& s, uint64_t seed) {
    RepHost& h = c->reps[r];
    h.d = s;
    const int N = s.n_part, M = s.n_bead;
    const bool lr = s.pair_type != SPIMC_PAIR_NONE && s.pair_use_long_range;
    const int nk = lr ? (int)s.n_k : 0;
    const double tau = s.beta / (1. * M), L = s.L, box = 2. * M_PI / L;
    // k-space
    h.k_n.assign(s.k_n ? s.k_n : nullptr, s.k_n ? s.k_n + 3 * (size_t)nk : nullptr);
    h.kidx.resize(nk);
    h.kmag.resize(nk);
    h.coef_s.resize(nk);
    h.coef_b.resize(nk);
    h.coef_v.resize(nk);
    int mx = 0;
    for (int k = 0; k < nk; ++k) {
        int n0 = h.k_n[3 * k], n1 = h.k_n[3 * k + 1], n2 = h.k_n[3 * k + 2];
        mx = std::max(mx, std::max(std::abs(n0), std::max(std::abs(n1), std::abs(n2))));
        h.kidx[k] = (n0 + 128) | ((n1 + 128) << 8) | ((n2 + 128) << 16);
        double kv[3] = {n0 * box, n1 * box, n2 * box};
        h.kmag[k] = std::sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
        // Bare: CalcULong multiplies the k sum by level_tau (bare_pair_action_class.h:170);
        // David: u_long_k = v_long_k * tau (david_pair_action_class.h:334).
        // Ilkka: u_long_k, du_long_k and v_long_k as read (ilkka_pair_action_class.h:63-80, 109-124, 152-168).
        h.coef_s[k] = (s.pair_type == SPIMC_PAIR_DAVID) ? s.pair_v_long_k[k] * tau : s.pair_v_long_k[k];
        h.coef_b[k] = s.pair_v_long_k[k];
        h.coef_v[k] = s.pair_v_long_k[k];
        if (s.pair_type == SPIMC_PAIR_ILKKA) {
            h.coef_s[k] = s.ilkka_u_long_k[k];
            h.coef_b[k] = s.ilkka_du_long_k[k];
        }
    }
    // runs of the long-range adjoint: group by (n0, n1), split each group's sorted
    // n2 into maximal consecutive sequences, and walk each from the n2 nearest 0
    {
        std::vector<int> idx(nk);
        for (int k = 0; k < nk; ++k) idx[k] = k;
        auto key = [&](int k) { return std::array<int, 3>{h.k_n[3 * k], h.k_n[3 * k + 1], h.k_n[3 * k + 2]}; };
        std::sort(idx.begin(), idx.end(), [&](int a, int b) { return key(a) < key(b); });
        h.k_runs.clear();
        h.k_order.clear();
        auto emit = [&](int n0, int n1, const std::vector<int>& ks, const std::vector<int>& n2s, int lo, int hi) {
            // ks[lo..hi] have consecutive n2s; split at 0 into an ascending and a descending run
            int z = lo;
            while (z <= hi && n2s[z] < 0) ++z;
            if (z <= hi) {  // ascending from n2s[z] (>= 0)
                h.k_runs.push_back(make_int4(n0, n1, n2s[z], hi - z + 1));
                for (int i = z; i <= hi; ++i) h.k_order.push_back(ks[i]);
            }
            if (z > lo) {  // descending from n2s[z-1] (< 0)
                h.k_runs.push_back(make_int4(n0, n1, n2s[z - 1], -(z - lo)));
                for (int i = z - 1; i >= lo; --i) h.k_order.push_back(ks[i]);
            }
        };
        for (int a = 0; a < nk;) {
            const int n0 = h.k_n[3 * idx[a]], n1 = h.k_n[3 * idx[a] + 1];
            std::vector<int> ks, n2s;
            int e = a;
            while (e < nk && h.k_n[3 * idx[e]] == n0 && h.k_n[3 * idx[e] + 1] == n1) {
                ks.push_back(idx[e]);
                n2s.push_back(h.k_n[3 * idx[e] + 2]);
                ++e;
            }
            for (size_t lo = 0; lo < ks.size();) {
                size_t hi = lo;
                while (hi + 1 < ks.size() && n2s[hi + 1] == n2s[hi] + 1) ++hi;
                emit(n0, n1, ks, n2s, (int)lo, (int)hi);
                lo = hi + 1;
            }
            a = e;
        }
    }
    // pair tables
    int n_val = 0;
    if (s.pair_type == SPIMC_PAIR_BARE) {
        h.v_r = make_spline(s.bare_r, s.bare_n_r, 1, {s.bare_v_r});
        if (lr) h.v_long_r = make_spline(s.bare_r_long, s.bare_n_r_long, 1, {s.bare_v_long_r});
    } else if (s.pair_type == SPIMC_PAIR_DAVID) {
        n_val = 1;
        for (uint32_t i = 1; i <= s.david_n_order; ++i) n_val += 1 + i;
        std::vector<const double*> yu, ydu;
        for (int k = 0; k <= n_val; ++k) {
            yu.push_back(s.david_u_kj + (size_t)k * s.david_n_grid);
            ydu.push_back(s.david_du_kj + (size_t)k * s.david_n_grid);
        }
        h.u = make_spline(s.david_grid, s.david_n_grid, n_val + 1, yu);
        h.du = make_spline(s.david_grid, s.david_n_grid, n_val + 1, ydu);
    } else if (s.pair_type == SPIMC_PAIR_ILKKA) {
        h.u2 = make_spline2(s.ilkka_x_u, s.ilkka_n_x_u, s.ilkka_y_u, s.ilkka_n_y_u, s.ilkka_u_xy);
        h.du2 = make_spline2(s.ilkka_x_du, s.ilkka_n_x_du, s.ilkka_y_du, s.ilkka_n_y_du, s.ilkka_du_xy);
        h.v_r = make_spline(s.ilkka_r_v, s.ilkka_n_r_v, 1, {s.ilkka_v_r});
        if (lr) {
            h.v_long_r = make_spline(s.ilkka_r_v_long, s.ilkka_n_r_v_long, 1, {s.ilkka_v_long_r});
            h.u_long_r = make_spline(s.ilkka_r_u_long, s.ilkka_n_r_u_long, 1, {s.ilkka_u_long_r});
            h.du_long_r = make_spline(s.ilkka_r_du_long, s.ilkka_n_r_du_long, 1, {s.ilkka_du_long_r});
        }
    }
    // LDS-staged B-spline tables for the pair kernel (uniform grids only)
    h.bt_a = BTabHost();
    h.bt_b = BTabHost();
    if (s.pair_type == SPIMC_PAIR_BARE) {
        const bool ua = grid_is_uniform(s.bare_r, s.bare_n_r);
        const bool ub = !lr || grid_is_uniform(s.bare_r_long, s.bare_n_r_long);
        if (ua && ub) {
            if (!s.bare_is_coulomb) h.bt_a = make_btab(s.bare_r, s.bare_n_r, {s.bare_v_r});
            if (lr) h.bt_b = make_btab(s.bare_r_long, s.bare_n_r_long, {s.bare_v_long_r});
        }
    } else if (s.pair_type == SPIMC_PAIR_DAVID && grid_is_uniform(s.david_grid, s.david_n_grid)) {
        std::vector<const double*> comps, dcomps;
        for (int k = 1; k <= n_val; ++k) comps.push_back(s.david_u_kj + (size_t)k * s.david_n_grid);
        for (int k = 0; k <= n_val; ++k) dcomps.push_back(s.david_du_kj + (size_t)k * s.david_n_grid);
        h.bt_a = make_btab(s.david_grid, s.david_n_grid, comps);
        h.bt_du = make_btab(s.david_grid, s.david_n_grid, dcomps);
    }
    // kinetic and sampling image splines
    h.kin.clear();
    h.bis.clear();
    if (s.kinetic_n_images > 0) {
        for (int l = 0; l < kMaxKinSplines && (1 << l) <= M / 2 + 1; ++l)
            h.kin.push_back(make_free_spline(L, s.kinetic_n_images, s.lambda, tau * (double)(1 << l), false));
        h.kin_dtau = make_free_spline(L, s.kinetic_n_images, s.lambda, tau, true);
    }
    // the sampling splines of every level the bisection may use (Bisect::SetupSpline,
    // bisect_class.h:155-162: floor(log2 M) of them; at most kMaxLevel here), so that
    // spimc_set_bisect_level needs no new tables
    if (s.bisect_n_images > 0)
        for (int l = 0; l < std::min((int)std::floor(std::log2((double)M)), kMaxLevel); ++l)
            h.bis.push_back(make_free_spline(L, s.bisect_n_images, s.lambda, 0.5 * tau * (double)(1 << l), false));
    h.n_chunks = std::max(1, (nk + kRhoBlock * c->rho_kpt - 1) / (kRhoBlock * c->rho_kpt));

    // constant tables: shared with an earlier replica that has the same ones
    h.const_hash = constants_hash(h);
    h.const_owner = (int)r;
    for (uint32_t o = 0; o < r; ++o)
        if (c->reps[o].const_owner == (int)o && same_constants(c->reps[o], h)) {
            h.const_owner = (int)o;
            break;
        }
    const bool own = h.const_owner == (int)r;
    // lay out the arena (sizing pass, allocation, carving pass)
    Rep& R = h.dev;
    for (int pass = 0; pass < 2; ++pass) {
        Carver cv;
        if (pass == 1) cv.base = h.arena;
        R.pos = cv.take<double>((size_t)3 * M * N);
        R.rho = cv.take<double2>((size_t)2 * M * std::max(nk, 1));
        R.sel = cv.take<unsigned char>(M);
        if (own) {
            R.kidx = cv.take<int>(std::max(nk, 1));
            R.kmag = cv.take<double>(std::max(nk, 1));
            R.coef_s = cv.take<double>(std::max(nk, 1));
            R.coef_b = cv.take<double>(std::max(nk, 1));
            R.coef_v = cv.take<double>(std::max(nk, 1));
            R.k_runs = cv.take<int4>(std::max<size_t>(h.k_runs.size(), 1));
            R.k_order = cv.take<int>(std::max(nk, 1));
            if (h.v_r.n) R.v_r = carve_spline(cv, h.v_r);
            if (h.v_long_r.n) R.v_long_r = carve_spline(cv, h.v_long_r);
            if (h.u.n) R.u = carve_spline(cv, h.u);
            if (h.du.n) R.du = carve_spline(cv, h.du);
            if (h.u_long_r.n) R.u_long_r = carve_spline(cv, h.u_long_r);
            if (h.du_long_r.n) R.du_long_r = carve_spline(cv, h.du_long_r);
            for (auto pr : {std::make_pair(&R.u2, &h.u2), std::make_pair(&R.du2, &h.du2)})
                if (pr.second->nx) {
                    pr.first->nx = pr.second->nx;
                    pr.first->ny = pr.second->ny;
                    pr.first->x = cv.take<double>(pr.second->x.size());
                    pr.first->y = cv.take<double>(pr.second->y.size());
                    pr.first->c = cv.take<double>(pr.second->c.size());
                }
            for (size_t l = 0; l < h.kin.size(); ++l) R.kin[l] = carve_spline(cv, h.kin[l]);
            if (h.kin_dtau.n) R.kin_dtau = carve_spline(cv, h.kin_dtau);
            for (size_t l = 0; l < h.bis.size(); ++l) R.bis[l] = carve_spline(cv, h.bis[l]);
            R.bt_a.c = cv.take<double>(std::max<size_t>(h.bt_a.c.size(), 1));
            R.bt_b.c = cv.take<double>(std::max<size_t>(h.bt_b.c.size(), 1));
            R.bt_du.c = cv.take<double>(std::max<size_t>(h.bt_du.c.size(), 1));
        }
        R.phase = cv.take<double2>((size_t)M * 6 * (mx + 1));
        R.rho_rows = cv.take<int2>(M);
        R.prop = cv.take<Prop>(1);
        R.prop_new = cv.take<double>((size_t)3 * M);
        R.pair_part = cv.take<double2>(M);
        R.rho_part = cv.take<double2>((size_t)M * h.n_chunks);
        R.kin_part = cv.take<double2>(1);
        R.cnt = cv.take<Counters>(1);
        R.tape_cursor = cv.take<long long>(1);
        R.e_acc = cv.take<double>(2);
        R.sk_acc = cv.take<double>(std::max(nk, 1));
        R.scratch = cv.take<double>(scratch_size(M, std::max(nk, 1)));
        R.grad = cv.take<double>((size_t)3 * M * N);
        R.grad_est = cv.take<double>((size_t)3 * M * N);
        R.n_measure = cv.take<unsigned long long>(3);
        if (pass == 0) {
            hipError_t e = hipMalloc((void**)&h.arena, cv.off + 256);
            if (e != hipSuccess) return fail(SPIMC_ERR_NOMEM, "device arena allocation failed");
            HIPCHK(hipMemset(h.arena, 0, cv.off + 256));
        }
    }
    // scalars
    R.N = N;
    R.M = M;
    R.nk = nk;
    R.max_idx = mx;
    R.n_level = s.bisect_n_level;
    R.perm_n_level = s.bisect_n_level;
    R.perm_log_eps = std::log(1.e-100);  // PermBisect epsilon default (perm_bisect_class.h:104)
    R.perm_eps = 1.e-100;
    R.pair_type = s.pair_type;
    R.use_long_range = lr ? 1 : 0;
    R.is_coulomb = (s.pair_type == SPIMC_PAIR_BARE && s.bare_is_coulomb) ? 1 : 0;
    R.n_order = s.david_n_order;
    R.n_val = n_val;
    R.kin_images = s.kinetic_n_images;
    R.bis_images = s.bisect_n_images;
    R.n_links_max = M;
    R.n_chunks = h.n_chunks;
    R.n_runs = (int)h.k_runs.size();
    R.rho_kpt = c->rho_kpt;
    R.index = r;
    R.L = L;
    R.iL = 1. / L;
    R.tau = tau;
    R.lambda = s.lambda;
    R.step = s.displace_step;
    R.box = box;
    R.kin_i4lt = 1. / (4. * s.lambda * tau);
    R.kin_i4ltt = 1. / (4. * s.lambda * tau * tau);
    R.kin_dbeta_const = N * M * 3 / (2. * tau);
    R.long_scale = (s.pair_type == SPIMC_PAIR_BARE) ? tau : 1.;
    const double nn = (double)N;
    // homologous constants (bare_pair_action_class.h:89-95, david_pair_action_class.h:340-361)
    R.const_v = lr ? s.pair_v_long_k_0 * 0.5 * nn * nn * M + s.pair_v_long_r_0 * (-0.5) * nn * M : 0.;
    R.const_du = R.const_v;
    // Ilkka: du/diag constants for dU/dbeta (ilkka_pair_action_class.h:418-423)
    if (s.pair_type == SPIMC_PAIR_ILKKA)
        R.const_du = lr ? s.ilkka_du_long_k_0 * 0.5 * nn * nn * M + s.ilkka_du_long_r_0 * (-0.5) * nn * M : 0.;
    if (s.pair_type == SPIMC_PAIR_BARE) {
        R.r_v_min = s.bare_r[0];
        R.r_v_max = s.bare_r[s.bare_n_r - 1];
        if (lr) {
            R.r_l_min = s.bare_r_long[0];
            R.r_l_max = s.bare_r_long[s.bare_n_r_long - 1];
        }
    } else if (s.pair_type == SPIMC_PAIR_DAVID) {
        R.g_min = s.david_grid[0];
        R.g_max = s.david_grid[s.david_n_grid - 1];
    } else if (s.pair_type == SPIMC_PAIR_ILKKA) {
        R.r_v_min = s.ilkka_r_v[0];
        R.r_v_max = s.ilkka_r_v[s.ilkka_n_r_v - 1];
        if (lr) {
            R.r_l_min = s.ilkka_r_v_long[0];
            R.r_l_max = s.ilkka_r_v_long[s.ilkka_n_r_v_long - 1];
            R.r_u_min = s.ilkka_r_u_long[0];
            R.r_u_max = s.ilkka_r_u_long[s.ilkka_n_r_u_long - 1];
            R.r_du_min = s.ilkka_r_du_long[0];
            R.r_du_max = s.ilkka_r_du_long[s.ilkka_n_r_du_long - 1];
        }
    }
    R.key = splitmix64(seed ^ (0x9E3779B97F4A7C15ull * (r + 1)));
    auto set_bt = [](BTab& d, const BTabHost& hh) {
        d.knots = hh.knots;
        d.ncoef = hh.knots ? hh.knots + 2 : 0;
        d.ncomp = hh.ncomp;
        d.x0 = hh.x0;
        d.ih = hh.ih;
    };
    set_bt(R.bt_a, h.bt_a);
    set_bt(R.bt_b, h.bt_b);
    set_bt(R.bt_du, h.bt_du);
    {
        const size_t nd = h.bt_a.c.size() + h.bt_b.c.size();
        const bool have = (s.pair_type == SPIMC_PAIR_DAVID && h.bt_a.knots) ||
                          (s.pair_type == SPIMC_PAIR_BARE && (s.bare_is_coulomb || h.bt_a.knots) && (!lr || h.bt_b.knots));
        R.lds_doubles = (have && nd * sizeof(double) <= kPairLdsMaxBytes) ? (int)std::max<size_t>(nd, 1) : 0;
        R.lds_doubles_du = 0;
        if (s.pair_type == SPIMC_PAIR_DAVID && h.bt_du.knots && h.bt_du.c.size() * sizeof(double) <= kPairLdsMaxBytes)
            R.lds_doubles_du = (int)h.bt_du.c.size();
    }
    R.tape = nullptr;
    R.tape_len = 0;
    R.accepted = nullptr;
    R.accepted_cap = 0;
    R.gr_cnt = nullptr;
    if (!own) {
        const Rep& O = c->reps[h.const_owner].dev;
        R.kidx = O.kidx;
        R.kmag = O.kmag;
        R.coef_s = O.coef_s;
        R.coef_b = O.coef_b;
        R.coef_v = O.coef_v;
        R.k_runs = O.k_runs;
        R.k_order = O.k_order;
        R.v_r = O.v_r;
        R.v_long_r = O.v_long_r;
        R.u = O.u;
        R.du = O.du;
        R.u2 = O.u2;
        R.du2 = O.du2;
        R.u_long_r = O.u_long_r;
        R.du_long_r = O.du_long_r;
        for (int l = 0; l < kMaxKinSplines; ++l) R.kin[l] = O.kin[l];
        R.kin_dtau = O.kin_dtau;
        for (int l = 0; l < kMaxKinSplines; ++l) R.bis[l] = O.bis[l];
        R.bt_a.c = O.bt_a.c;
        R.bt_b.c = O.bt_b.c;
        R.bt_du.c = O.bt_du.c;
        // the host copies are only compared against owners; drop the large ones
        h.u = SplHost();
        h.du = SplHost();
        h.u2.c.clear();
        h.du2.c.clear();
        h.bt_a.c.clear();
        h.bt_b.c.clear();
        h.bt_du.c.clear();
        return SPIMC_OK;
    }
    // uploads
    if (!h.bt_a.c.empty()) HIPCHK(hipMemcpy((void*)R.bt_a.c, h.bt_a.c.data(), h.bt_a.c.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!h.bt_b.c.empty()) HIPCHK(hipMemcpy((void*)R.bt_b.c, h.bt_b.c.data(), h.bt_b.c.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!h.bt_du.c.empty()) HIPCHK(hipMemcpy((void*)R.bt_du.c, h.bt_du.c.data(), h.bt_du.c.size() * sizeof(double), hipMemcpyHostToDevice));
    if (nk) {
        HIPCHK(hipMemcpy((void*)R.kidx, h.kidx.data(), nk * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void*)R.kmag, h.kmag.data(), nk * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void*)R.coef_s, h.coef_s.data(), nk * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void*)R.coef_b, h.coef_b.data(), nk * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void*)R.coef_v, h.coef_v.data(), nk * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void*)R.k_runs, h.k_runs.data(), h.k_runs.size() * sizeof(int4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void*)R.k_order, h.k_order.data(), nk * sizeof(int), hipMemcpyHostToDevice));
    }
    int rc;
    if ((rc = upload_spline(R.v_r, h.v_r)) || (rc = upload_spline(R.v_long_r, h.v_long_r)) ||
        (rc = upload_spline(R.u, h.u)) || (rc = upload_spline(R.du, h.du)) || (rc = upload_spline(R.kin_dtau, h.kin_dtau)) ||
        (rc = upload_spline(R.u_long_r, h.u_long_r)) || (rc = upload_spline(R.du_long_r, h.du_long_r)))
        return rc;
    for (auto pr : {std::make_pair(&R.u2, &h.u2), std::make_pair(&R.du2, &h.du2)})
        if (pr.second->nx) {
            HIPCHK(hipMemcpy((void*)pr.first->x, pr.second->x.data(), pr.second->x.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy((void*)pr.first->y, pr.second->y.data(), pr.second->y.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy((void*)pr.first->c, pr.second->c.data(), pr.second->c.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    for (size_t l = 0; l < h.kin.size(); ++l)
        if ((rc = upload_spline(R.kin[l], h.kin[l]))) return rc;
    for (size_t l = 0; l < h.bis.size(); ++l)
        if ((rc = upload_spline(R.bis[l], h.bis[l]))) return rc;
    return SPIMC_OK;
}

int upload_reps(spimc_ctx* c) {
    std::vector<Rep> v;
    for (auto& h : c->reps) v.push_back(h.dev);
    HIPCHK(hipMemcpy(c->d_reps, v.data(), v.size() * sizeof(Rep), hipMemcpyHostToDevice));
    return SPIMC_OK;
}

int upload_items(const std::vector<uint2>& v, uint2** d, int* n) {
    *n = (int)v.size();
    if (v.empty()) return SPIMC_OK;
    HIPCHK(hipMalloc((void**)d, v.size() * sizeof(uint2)));
    HIPCHK(hipMemcpy(*d, v.data(), v.size() * sizeof(uint2), hipMemcpyHostToDevice));
    return SPIMC_OK;
}

int build_work_lists(spimc_ctx* c, Group& G) {
    std::vector<uint2> pb, pd;
    int nb_items = 0, nd_items = 0;
    for (uint32_t r = (uint32_t)G.r0; r < (uint32_t)(G.r0 + G.n); ++r) {
        const Rep& R = c->reps[r].dev;
        const int nb = 1 << R.n_level;
        for (int j = 0; j < nb; ++j) pb.push_back(make_uint2(r, j));
        for (int j = 0; j < R.M; ++j) pd.push_back(make_uint2(r, j));
        if (R.use_long_range) {
            nb_items += (nb - 1) * R.n_chunks;
            nd_items += R.M * R.n_chunks;
        }
    }
    int rc;
    if ((rc = upload_items(pb, &G.d_pair_bis, &G.n_pair_bis)) || (rc = upload_items(pd, &G.d_pair_dis, &G.n_pair_dis)))
        return rc;
    // rho_k items: zeroed (idle) until a proposal kernel fills its replica's range
    for (int mv = 0; mv < 2; ++mv) {
        const int n = mv == 0 ? nb_items : nd_items;
        RhoItem*& d = mv == 0 ? G.d_ritem_bis : G.d_ritem_dis;
        (mv == 0 ? G.n_ritem_bis : G.n_ritem_dis) = n;
        if (!n) continue;
        HIPCHK(hipMalloc((void**)&d, n * sizeof(RhoItem)));
        HIPCHK(hipMemset(d, 0, n * sizeof(RhoItem)));
    }
    int ob = 0, od = 0;
    for (int r = G.r0; r < G.r0 + G.n; ++r) {
        Rep& R = c->reps[r].dev;
        if (!R.use_long_range) continue;
        const int nb = ((1 << R.n_level) - 1) * R.n_chunks, nd = R.M * R.n_chunks;
        R.ritems_bis = G.d_ritem_bis + ob;
        R.ritems_dis = G.d_ritem_dis + od;
        R.n_ritems_bis = nb;
        R.n_ritems_dis = nd;
        ob += nb;
        od += nd;
    }
    if (c->round_fn[0] && c->lds_doubles_max > 0) {
        // one-launch rounds, per move: pair workgroups per replica (fewer links
        // each when a replica of more than 256 particles splits its links'
        // partners), and as many rho_k items per workgroup as keep the launch
        // within what the GPU holds at once (a deciding workgroup would otherwise
        // hold a slot that later workgroups wait for)
        const int waves = kRoundBlock / 64;
        int max_N = 1;
        for (int r = G.r0; r < G.r0 + G.n; ++r) max_N = std::max(max_N, c->reps[r].dev.N);
        const int lpb = waves / std::min(waves, (max_N + 255) / 256);
        for (int mv = 0; mv < 2; ++mv) {
            const int links = mv == 0 ? G.max_bis_links : G.max_M, n_items = mv == 0 ? ob : od;
            G.round_bpr[mv] = std::max(1, std::min((links + lpb - 1) / lpb, (2 * kNumCUs + G.n - 1) / G.n));
            int occ = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)c->round_fn[mv], kRoundBlock,
                                                             (size_t)c->lds_doubles_max * sizeof(double)) != hipSuccess)
                occ = 1;
            // at least half of what the GPU holds goes to rho_k workgroups, also
            // when the pair workgroups alone would fill it
            const int cap = std::max(1, occ) * kNumCUs;
            const int room = std::max(cap / 2, cap - G.n * G.round_bpr[mv]);
            G.round_ipb[mv] = std::max(1, (n_items + room - 1) / room);
            if (const char* f = std::getenv("SPIMC_ROUND_IPB")) G.round_ipb[mv] = std::max(1, std::atoi(f));
            std::vector<int4> grp;
            int o = 0;
            for (int r = G.r0; r < G.r0 + G.n; ++r) {
                const Rep& R = c->reps[r].dev;
                if (!R.use_long_range) continue;
                const int n = mv == 0 ? R.n_ritems_bis : R.n_ritems_dis;
                for (int f = 0; f < n; f += G.round_ipb[mv])
                    grp.push_back(make_int4(o + f, std::min(G.round_ipb[mv], n - f), r - G.r0, 0));
                o += n;
            }
            G.n_rgrp[mv] = (int)grp.size();
            if (!grp.empty()) {
                HIPCHK(hipMalloc((void**)&G.d_rgrp[mv], grp.size() * sizeof(int4)));
                HIPCHK(hipMemcpy(G.d_rgrp[mv], grp.data(), grp.size() * sizeof(int4), hipMemcpyHostToDevice));
            }
        }
    }
    if (!G.d_arrive) {
        HIPCHK(hipMalloc((void**)&G.d_arrive, G.n * sizeof(unsigned int)));
        HIPCHK(hipMemset(G.d_arrive, 0, G.n * sizeof(unsigned int)));
    }
    return SPIMC_OK;
}

// Splits the replicas into n_groups contiguous ranges of about equal bead count.
int build_groups(spimc_ctx* c, int n_groups) {
    const int n_reps = (int)c->reps.size();
    n_groups = std::max(1, std::min(n_groups, n_reps));
    double total = 0.;
    for (auto& h : c->reps) total += (double)h.dev.N * h.dev.M;
    c->groups.resize(n_groups);
    int r = 0;
    double acc = 0.;
    for (int g = 0; g < n_groups; ++g) {
        Group& G = c->groups[g];
        G.r0 = r;
        const double target = total * (g + 1) / n_groups;
        while (r < n_reps && (r == G.r0 || acc + 0.5 * c->reps[r].dev.N * c->reps[r].dev.M <= target) &&
               n_reps - r > n_groups - g - 1) {
            acc += (double)c->reps[r].dev.N * c->reps[r].dev.M;
            ++r;
        }
        if (g == n_groups - 1) r = n_reps;
        G.n = r - G.r0;
        for (int q = G.r0; q < r; ++q) {
            G.max_bis_links = std::max(G.max_bis_links, 1 << c->reps[q].dev.n_level);
            G.max_M = std::max(G.max_M, c->reps[q].dev.M);
        }
        if (g == 0)
            G.stream = c->stream;
        else if (hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking) != hipSuccess)
            return fail(SPIMC_ERR_HIP, "stream creation failed");
        if (hipEventCreateWithFlags(&G.done, hipEventDisableTiming) != hipSuccess) return fail(SPIMC_ERR_HIP, "event");
        // the side stream exists whether or not the fork is on (the autotuner may turn it on)
        if (hipStreamCreateWithFlags(&G.side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&G.ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&G.ev_join, hipEventDisableTiming) != hipSuccess)
            return fail(SPIMC_ERR_HIP, "fork stream");
        HIPCHK(hipMalloc((void**)&G.d_kt, KT_WORDS * sizeof(unsigned long long)));
        HIPCHK(hipMemset(G.d_kt, 0, KT_WORDS * sizeof(unsigned long long)));
        int rc = build_work_lists(c, G);
        if (rc) return rc;
    }
    HIPCHK(hipEventCreateWithFlags(&c->stage_ready, hipEventDisableTiming));
    return SPIMC_OK;
}

int ensure_tmp(spimc_ctx* c, size_t n_doubles, size_t n_items) {
    if (n_doubles > c->tmp_cap) {
        if (c->d_tmp) (void)hipFree(c->d_tmp);
        c->d_tmp = nullptr;
        HIPCHK(hipMalloc((void**)&c->d_tmp, n_doubles * sizeof(double)));
        c->tmp_cap = n_doubles;
    }
    if (n_items > c->items_tmp_cap) {
        if (c->d_items_tmp) (void)hipFree(c->d_items_tmp);
        c->d_items_tmp = nullptr;
        HIPCHK(hipMalloc((void**)&c->d_items_tmp, n_items * sizeof(uint2)));
        c->items_tmp_cap = n_items;
    }
    return SPIMC_OK;
}

// One round of a move stage for every replica.
// Selects the LDS pair kernel of workgroup size bs (256, 512, 768 or 1024).
int set_pair_block(spimc_ctx* c, int bs) {
    using PairFn = void (*)(const Rep*, int, int, unsigned long long*);
    const PairFn fns[4][4] = {
        {k_pair_lds<1, 0, 256>, k_pair_lds<2, 2, 256>, k_pair_lds<2, 0, 256>, k_pair_lds<0, 0, 256>},
        {k_pair_lds<1, 0, 512>, k_pair_lds<2, 2, 512>, k_pair_lds<2, 0, 512>, k_pair_lds<0, 0, 512>},
        {k_pair_lds<1, 0, 768>, k_pair_lds<2, 2, 768>, k_pair_lds<2, 0, 768>, k_pair_lds<0, 0, 768>},
        {k_pair_lds<1, 0, 1024>, k_pair_lds<2, 2, 1024>, k_pair_lds<2, 0, 1024>, k_pair_lds<0, 0, 1024>}};
    const int bi = bs == 256 ? 0 : bs == 512 ? 1 : bs == 768 ? 2 : bs == 1024 ? 3 : -1;
    if (bi < 0) return fail(SPIMC_ERR_ARG, "pair workgroup size must be 256, 512, 768 or 1024");
    c->pair_block = bs;
    c->pair_lds_fn = fns[bi][c->pair_variant];
    if (c->lds_doubles_max > 0 &&
        hipFuncSetAttribute((const void*)c->pair_lds_fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            c->lds_doubles_max * (int)sizeof(double)) != hipSuccess)
        return fail(SPIMC_ERR_HIP, "pair kernel LDS size");
    return SPIMC_OK;
}

// Dynamic LDS of the persistent kernel: the pair tables (rounded to 16 bytes),
// then the phase chains of a stage's moved beads (up to M of them).
constexpr size_t kPersLdsBytes = 160 * 1024 - 24 * 1024;  // the CU's LDS less the kernel's static arrays
size_t pers_tables(const spimc_ctx* c) { return ((size_t)c->lds_doubles_max + 1) & ~(size_t)1; }
size_t pers_chain_bytes(const spimc_ctx* c) {
    size_t b = 0;
    for (const auto& h : c->reps) b = std::max(b, (size_t)h.dev.M * 6 * (h.dev.max_idx + 1) * sizeof(double2));
    return b;
}
size_t pers_shmem(const spimc_ctx* c) { return pers_tables(c) * sizeof(double) + pers_chain_bytes(c); }

// Replicas small enough for the persistent stage kernel (a whole stage of one
// replica in one workgroup, its per-attempt sums in LDS), on paths that are not
// permuted.
constexpr int kPersMaxBeads = 16384;  // N M of the largest replica
bool pers_eligible(const spimc_ctx* c) {
    if (c->perm_on) return false;
    for (const auto& h : c->reps)
        if ((long long)h.dev.N * h.dev.M > kPersMaxBeads || h.dev.M > kStageMaxM || h.dev.n_chunks > 1 ||
            stage_draws(0, h.dev.n_level) > kStageMaxDraws)
            return false;
    return pers_shmem(c) <= kPersLdsBytes;
}

// One-launch rounds: the LDS pair kernel, the default rho_k chunking, rounds
// decided in the next proposal's launch, and unpermuted paths.
bool one_launch_eligible(const spimc_ctx* c) {
    return c->round_fn[0] && c->round_fn[1] && c->lds_doubles_max > 0 && c->fuse_rounds && !c->perm_on &&
           c->rho_kpt == kRhoKpt;
}
// Whether a stage of this move runs one-launch rounds: bisection, and with
// one_moves 2 displacement too; not while the device clocks time the pair and
// rho_k kernels (k_round has no separate rho_k kernel to time, so those rounds
// run as three launches).
bool one_launch_now(const spimc_ctx* c, int move) {
    return c->one_launch && move < c->one_moves && !c->clock && one_launch_eligible(c);
}



// Selects the persistent stage kernel of workgroup size bs (256, 512 or 1024;
// 0: stages run in rounds).
int set_pers_block(spimc_ctx* c, int bs) {
    using StageFn = void (*)(const Rep*, int, int, int, int);
    const StageFn fns[3][4] = {
        {k_stage_persistent<1, 0, 256>, k_stage_persistent<2, 2, 256>, k_stage_persistent<2, 0, 256>,
         k_stage_persistent<0, 0, 256>},
        {k_stage_persistent<1, 0, 512>, k_stage_persistent<2, 2, 512>, k_stage_persistent<2, 0, 512>,
         k_stage_persistent<0, 0, 512>},
        {k_stage_persistent<1, 0, 1024>, k_stage_persistent<2, 2, 1024>, k_stage_persistent<2, 0, 1024>,
         k_stage_persistent<0, 0, 1024>}};
    if (bs == 0) {
        c->pers_block = 0;
        c->pers_fn = nullptr;
        return SPIMC_OK;
    }
    const int bi = bs == 256 ? 0 : bs == 512 ? 1 : bs == 1024 ? 2 : -1;
    if (bi < 0) return fail(SPIMC_ERR_ARG, "persistent stage workgroup size must be 0, 256, 512 or 1024");
    const StageFn fn = fns[bi][c->pair_variant];
    if (pers_shmem(c) > kPersLdsBytes)
        return fail(SPIMC_ERR_UNSUPPORTED, "persistent stage kernel: the tables and phase chains exceed its LDS");
    if (hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pers_shmem(c)) !=
        hipSuccess)
        return fail(SPIMC_ERR_HIP, "persistent stage kernel LDS size");
    c->pers_block = bs;
    c->pers_fn = fn;
    return SPIMC_OK;
}

// Drops the captured round and stage graphs (the launch shape changed).
void drop_graphs(spimc_ctx* c) {
    for (auto& G : c->groups)
        for (auto& g : G.graph)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    for (auto& g : c->pers_graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// A stage in one launch of the persistent kernel, after the stage's counts are
// set (k_stage_begin reads them from d_counts).
int launch_pers_stage(spimc_ctx* c, int move) {
    const int R = (int)c->reps.size();
    int rc = launch(c, SPIMC_KERNEL_DECIDE, k_stage_begin, (R + 63) / 64, 64, 0, c->stream, c->d_reps, R, c->d_counts);
    if (rc) return rc;
    return launch(c, SPIMC_KERNEL_PROPOSE, c->pers_fn, R, c->pers_block, pers_shmem(c), c->stream, c->d_reps, R, move,
                  c->lds_doubles_max > 0 ? 1 : 0, (int)pers_tables(c));
}

int capture_pers_graph(spimc_ctx* c, int move) {
    if (c->pers_graph[move]) return SPIMC_OK;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const bool saved = c->timing;
    c->timing = false;
    int rc = launch_pers_stage(c, move);
    c->timing = saved;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(SPIMC_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->pers_graph[move], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(SPIMC_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    return SPIMC_OK;
}

// The rho_k kernel of a round: rho_bpc workgroups per CU looping over the items
// (0: one workgroup per item).
int launch_rho(spimc_ctx* c, hipStream_t st, const RhoItem* ri, int nr, unsigned long long* kt) {
    const int g = c->rho_bpc > 0 ? std::min(nr, c->rho_bpc * kNumCUs) : nr;
    return launch(c, SPIMC_KERNEL_RHO, c->rho_loop_fn, g, kRhoBlock, 0, st, ri, nr, kt);
}

// Label-general pair kernel of a round: workgroups per replica that loop over
// the window links (only proposals marked generic do work).
constexpr int kPairMultiBpr = 16;
int launch_pair_multi(spimc_ctx* c, const Group& G, hipStream_t st) {
    return launch(c, SPIMC_KERNEL_PAIR, k_pair_multi, G.n * kPairMultiBpr, kPairMultiBlock, 0, st, c->d_reps + G.r0, G.n,
                  kPairMultiBpr);
}

// A permuting bisection round.  One-chain proposals (most of them) run on the
// Boltzmann bisection's kernels: the LDS pair kernel and the rho_k items;
// several chains, or a chain whose label changes inside the window, on the
// label-general pair kernel and k_rho_multi.  rho_k runs beside the pair
// kernels on the group's second stream.
int launch_perm_round(spimc_ctx* c, const Group& G) {
    const int R = G.n;
    Rep* reps = c->d_reps + G.r0;
    hipStream_t st = G.stream;
    int rc;
    const bool tuned = c->lds_doubles_max > 0;
    const size_t shmem = (size_t)c->max_N * (sizeof(double) + sizeof(int));
    if ((rc = launch(c, SPIMC_KERNEL_PROPOSE, c->perm_table ? k_propose_perm<true> : k_propose_perm<false>, R,
                     kPermBlock, shmem, st, reps, R, (int)tuned,
                     c->perm_stage)))
        return rc;
    const bool fork = c->any_long && G.side;
    hipStream_t rst = fork ? G.side : st;
    if (fork) {
        HIPCHK(hipEventRecord(G.ev_fork, st));
        HIPCHK(hipStreamWaitEvent(G.side, G.ev_fork, 0));
    }
    if (c->any_long) {
        if (G.n_ritem_bis && (rc = launch_rho(c, rst, G.d_ritem_bis, G.n_ritem_bis, nullptr))) return rc;
        int max_chunks = 1, max_slots = 1;
        for (int r = G.r0; r < G.r0 + G.n; ++r) {
            max_chunks = std::max(max_chunks, c->reps[r].dev.n_chunks);
            max_slots = std::max(max_slots, (1 << c->reps[r].dev.perm_n_level) - 1);
        }
        // a fixed grid over the listed proposals' (replica, slot, chunk) items
        auto fn = c->rho_kpt == 2    ? k_rho_multi_list<2>
                  : c->rho_kpt == 4  ? k_rho_multi_list<4>
                  : c->rho_kpt == 16 ? k_rho_multi_list<16>
                                     : k_rho_multi_list<8>;
        const int* list = reinterpret_cast<const int*>(G.d_multi + 2) + G.n;
        if ((rc = launch(c, SPIMC_KERNEL_RHO, fn, 4 * kNumCUs, kRhoBlock, 0, rst, reps, list, G.d_multi, max_slots,
                         max_chunks)))
            return rc;
    }
    if (fork) HIPCHK(hipEventRecord(G.ev_join, G.side));
    if (tuned) {
        const size_t lds = (size_t)c->lds_doubles_max * sizeof(double);
        int max_links = 1;
        for (int r = G.r0; r < G.r0 + G.n; ++r) max_links = std::max(max_links, 1 << c->reps[r].dev.perm_n_level);
        const int waves = c->pair_block / 64;
        const int bpr = std::max(1, std::min((max_links + waves - 1) / waves, (2 * kNumCUs + R - 1) / R));
        if ((rc = launch(c, SPIMC_KERNEL_PAIR, c->pair_lds_fn, R * bpr, c->pair_block, lds, st, reps, R, bpr,
                         (unsigned long long*)nullptr)))
            return rc;
    }
    // as many workgroups as a full grid: a listed item each, the others return at once
    if ((rc = launch(c, SPIMC_KERNEL_PAIR, k_pair_multi_list, R * kPairMultiBpr, kPairMultiBlock, 0, st, reps,
                     reinterpret_cast<const int*>(G.d_multi + 2), (const unsigned int*)G.d_multi, kPairMultiBpr)))
        return rc;
    if (fork) HIPCHK(hipStreamWaitEvent(st, G.ev_join, 0));
    return launch(c, SPIMC_KERNEL_DECIDE, k_decide_perm, R, 64, 0, st, reps, R, G.d_multi);
}

// A bisection round of a small batch in one launch (k_round); the group's
// first proposals come from a k_decide_propose_bisect launch at the stage's
// start.  (Displace rounds stay three launches: their many rho_k items exceed
// what the GPU holds at once, and the workgroups that decide then delay the
// rest; measured slower.)
int launch_one_round(spimc_ctx* c, const Group& G, int move) {
    const int R = G.n;
    const size_t shmem = (size_t)c->lds_doubles_max * sizeof(double);
    return launch(c, SPIMC_KERNEL_PAIR, c->round_fn[move], R * G.round_bpr[move] + G.n_rgrp[move], kRoundBlock, shmem,
                  G.stream, c->d_reps + G.r0, R, G.round_bpr[move],
                  (const RhoItem*)(move == 0 ? G.d_ritem_bis : G.d_ritem_dis), (const int4*)G.d_rgrp[move],
                  G.round_ipb[move], G.d_arrive);
}

int launch_round(spimc_ctx* c, const Group& G, int move) {
    if (move == 3) return launch_perm_round(c, G);
    if (one_launch_now(c, move)) return launch_one_round(c, G, move);
    const int R = G.n;
    unsigned long long* kt = c->clock ? G.d_kt : nullptr;
    Rep* reps = c->d_reps + G.r0;  // kernels indexed by block: replicas of this group
    Rep* all = c->d_reps;          // kernels driven by work items: absolute indices
    hipStream_t st = G.stream;
    int rc;
    if (c->fuse_rounds) {
        // the previous round's decision and this round's proposal in one launch
        rc = move == 0 ? launch(c, SPIMC_KERNEL_PROPOSE, k_decide_propose_bisect, R, kBisBlock, 0, st, reps, R, kt)
                       : launch(c, SPIMC_KERNEL_PROPOSE, k_decide_propose_displace, R, kDisBlock, 0, st, reps, R, kt);
    } else if (move == 0) {
        rc = launch(c, SPIMC_KERNEL_PROPOSE, k_propose_bisect, R, 64, 0, st, reps, R);
    } else {
        rc = launch(c, SPIMC_KERNEL_PROPOSE, k_propose_displace, R, 64, 0, st, reps, R);
    }
    if (rc) return rc;
    // a bisection chain that crosses the wrap of a permuted path changes label:
    // the label-general pair kernel takes those proposals, after the tuned pair
    // kernel on the same stream (beside rho_k when that is forked)
    const bool multi = move == 0 && c->perm_on;
    const uint2* pi = move == 0 ? G.d_pair_bis : G.d_pair_dis;
    const int np = move == 0 ? G.n_pair_bis : G.n_pair_dis;
    const RhoItem* ri = move == 0 ? G.d_ritem_bis : G.d_ritem_dis;
    const int nr = move == 0 ? G.n_ritem_bis : G.n_ritem_dis;
    if (c->fork && c->lds_doubles_max > 0 && np && G.side && nr) {
        // fork: rho_k on the side stream beside the pair kernel, joined before the decision
        HIPCHK(hipEventRecord(G.ev_fork, st));
        HIPCHK(hipStreamWaitEvent(G.side, G.ev_fork, 0));
        if ((rc = launch_rho(c, G.side, ri, nr, kt))) return rc;
        HIPCHK(hipEventRecord(G.ev_join, G.side));
        const size_t shmem = (size_t)c->lds_doubles_max * sizeof(double);
        const int waves = c->pair_block / 64, max_links = move == 0 ? G.max_bis_links : G.max_M;
        const int bpr = std::max(1, std::min((max_links + waves - 1) / waves, (2 * kNumCUs + R - 1) / R));
        rc = launch(c, SPIMC_KERNEL_PAIR, c->pair_lds_fn, R * bpr, c->pair_block, shmem, st, reps, R, bpr, kt);
        if (rc) return rc;
        if (multi && (rc = launch_pair_multi(c, G, st))) return rc;
        HIPCHK(hipStreamWaitEvent(st, G.ev_join, 0));
    } else if (c->lds_doubles_max > 0 && np) {
        const size_t shmem = (size_t)c->lds_doubles_max * sizeof(double);
        // workgroups per replica: one when the replicas fill the GPU, more (up to one
        // per BS/64 links) when they are few
        const int waves = c->pair_block / 64, max_links = move == 0 ? G.max_bis_links : G.max_M;
        const int bpr = std::max(1, std::min((max_links + waves - 1) / waves, (2 * kNumCUs + R - 1) / R));
        rc = launch(c, SPIMC_KERNEL_PAIR, c->pair_lds_fn, R * bpr, c->pair_block, shmem, st, reps, R, bpr, kt);
        if (rc) return rc;
        if (multi && (rc = launch_pair_multi(c, G, st))) return rc;
        if (nr) rc = launch_rho(c, st, ri, nr, kt);
        if (rc) return rc;
    } else {
        if (np) {
            const int g = (np + 3) / 4;
            rc = launch(c, SPIMC_KERNEL_PAIR, k_pair, g, kPairBlock, 0, st, all, pi, np, 0, kt);
            if (rc) return rc;
        }
        if (multi && (rc = launch_pair_multi(c, G, st))) return rc;
        if (nr) {
            if ((rc = launch_rho(c, st, ri, nr, kt))) return rc;
        }
    }
    if (c->fuse_rounds) return SPIMC_OK;  // decided by the next round's first launch or the stage's end
    return launch(c, SPIMC_KERNEL_DECIDE, k_decide, R, 64, 0, st, reps, R, kt);
}

// The decision of a stage's last round when rounds are fused (one-launch rounds
// decide their own).
int launch_final_decide(spimc_ctx* c, const Group& G) {
    if (!c->fuse_rounds) return SPIMC_OK;
    unsigned long long* kt = c->clock ? G.d_kt : nullptr;
    return launch(c, SPIMC_KERNEL_DECIDE, k_decide, G.n, 64, 0, G.stream, c->d_reps + G.r0, G.n, kt);
}

int ensure_accepted(spimc_ctx* c, uint32_t need) {
    bool changed = false;
    for (auto& h : c->reps) {
        if (h.accepted_cap < need) {
            if (h.accepted) (void)hipFree(h.accepted);
            if (h.trace) (void)hipFree(h.trace);
            h.accepted = nullptr;
            h.trace = nullptr;
            uint32_t cap = std::max<uint32_t>(need, 1024);
            HIPCHK(hipMalloc((void**)&h.accepted, cap));
            h.accepted_cap = cap;
            h.dev.accepted = h.accepted;
            h.dev.accepted_cap = cap;
            h.dev.act_trace = nullptr;
            changed = true;
        }
        if (c->trace_on && !h.trace) {
            HIPCHK(hipMalloc((void**)&h.trace, (size_t)h.accepted_cap * kTraceWords * sizeof(double)));
            h.dev.act_trace = h.trace;
            changed = true;
        } else if (c->trace_on && h.dev.act_trace != h.trace) {
            h.dev.act_trace = h.trace;
            changed = true;
        } else if (!c->trace_on && h.dev.act_trace) {
            h.dev.act_trace = nullptr;
            changed = true;
        }
    }
    if (changed) {
        int rc = upload_reps(c);
        if (rc) return rc;
        // captured graphs hold no pointers from the Rep table, so they stay valid
    }
    return SPIMC_OK;
}

int capture_round_graph(spimc_ctx* c, Group& G, int move) {
    if (G.graph[move] && G.graph_clock[move] != c->clock) {
        (void)hipGraphExecDestroy(G.graph[move]);
        G.graph[move] = nullptr;
    }
    if (G.graph[move]) return SPIMC_OK;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(G.stream, hipStreamCaptureModeThreadLocal));
    bool saved = c->timing;
    c->timing = false;
    int rc = SPIMC_OK;
    for (int i = 0; i < kRoundsPerGraph && rc == SPIMC_OK; ++i) rc = launch_round(c, G, move);
    c->timing = saved;
    G.graph_clock[move] = c->clock;
    hipError_t e = hipStreamEndCapture(G.stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SPIMC_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    HIPCHK(hipGraphInstantiate(&G.graph[move], g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    return SPIMC_OK;
}

// The path topology arrays of every replica (the identity at first) and, with
// scratch, the permuting move's chain arrays.  Once a path may be permuted the
// rounds also launch the label-general pair kernel.
int ensure_perm(spimc_ctx* c, bool scratch) {
    bool changed = false;
    for (auto& h : c->reps) {
        Rep& R = h.dev;
        if (!h.perm_arena) {
            if (!changed) HIPCHK(hipDeviceSynchronize());
            std::vector<int> id(2 * (size_t)R.N);
            for (int p = 0; p < R.N; ++p) id[p] = id[R.N + p] = p;
            HIPCHK(hipMalloc((void**)&h.perm_arena, id.size() * sizeof(int)));
            HIPCHK(hipMemcpy(h.perm_arena, id.data(), id.size() * sizeof(int), hipMemcpyHostToDevice));
            R.perm = reinterpret_cast<int*>(h.perm_arena);
            R.iperm = R.perm + R.N;
            changed = true;
        }
        if (scratch && !h.perm_scratch) {
            if (R.N > kPermMaxN) return fail(SPIMC_ERR_UNSUPPORTED, "permuting moves support at most 4096 particles");
            if (!changed) HIPCHK(hipDeviceSynchronize());
            const size_t N = (size_t)R.N, M = (size_t)R.M;
            auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
            const size_t b_int = up(N * sizeof(int)), b_cnt = up((2 + 2 * N) * sizeof(unsigned long long));
            const size_t b_new = up(N * (M - 1) * 3 * sizeof(double));
            HIPCHK(hipMalloc((void**)&h.perm_scratch, 6 * b_int + b_cnt + b_new));
            char* q = h.perm_scratch;
            R.ch_lbl = reinterpret_cast<int*>(q);
            R.ch_lbl2 = reinterpret_cast<int*>(q + b_int);
            R.mv_pre = reinterpret_cast<int*>(q + 2 * b_int);
            R.mv_post = reinterpret_cast<int*>(q + 3 * b_int);
            R.ch_nxt = reinterpret_cast<int*>(q + 4 * b_int);
            R.ch_lead = reinterpret_cast<int*>(q + 5 * b_int);
            R.perm_cnt = reinterpret_cast<unsigned long long*>(q + 6 * b_int);
            R.ch_new = reinterpret_cast<double*>(q + 6 * b_int + b_cnt);
            HIPCHK(hipMemset(q, 0, 6 * b_int + b_cnt));
            HIPCHK(hipMemset(R.mv_pre, 0xff, 2 * b_int));  // -1: no chain
            changed = true;
        }
    }
    if (scratch)
        for (auto& G : c->groups) {
            if (G.d_multi) continue;
            HIPCHK(hipMalloc((void**)&G.d_multi, (2 + 2 * (size_t)G.n) * sizeof(unsigned int)));
            HIPCHK(hipMemset(G.d_multi, 0, (2 + 2 * (size_t)G.n) * sizeof(unsigned int)));
            for (int r = G.r0; r < G.r0 + G.n; ++r) {
                Rep& R = c->reps[r].dev;
                R.multi_cnt = G.d_multi;
                R.multi_list = reinterpret_cast<int*>(G.d_multi + 2);
                R.multi_cap = G.n;
            }
            changed = true;
        }
    if (!changed) return SPIMC_OK;
    c->perm_on = true;
    drop_graphs(c);
    return upload_reps(c);
}

int finish_stage(spimc_ctx* c, const std::vector<unsigned int>& cnt, uint32_t mx, uint8_t* accepted);

int run_stage(spimc_ctx* c, int move, const uint32_t* counts, uint32_t n_uniform, uint8_t* accepted, bool use_graph) {
    if (c->prop_set) return fail(SPIMC_ERR_STATE, "an interface proposal is pending; accept or reject it first");
    const int R = (int)c->reps.size();
    std::vector<unsigned int> cnt(R);
    uint32_t mx = 0;
    for (int r = 0; r < R; ++r) {
        cnt[r] = counts ? counts[r] : n_uniform;
        mx = std::max(mx, cnt[r]);
    }
    if (mx == 0) return SPIMC_OK;
    int rc = ensure_accepted(c, mx);
    if (rc) return rc;
    if (c->pers_fn && move < 2 && pers_eligible(c)) {
        // the whole stage in one launch (from a captured graph with use_graph)
        const bool pgraph = use_graph && !c->timing;
        if (pgraph && (rc = capture_pers_graph(c, move))) return rc;
        HIPCHK(hipMemcpyAsync(c->d_counts, cnt.data(), R * sizeof(unsigned int), hipMemcpyHostToDevice, c->stream));
        if (pgraph)
            HIPCHK(hipGraphLaunch(c->pers_graph[move], c->stream));
        else if ((rc = launch_pers_stage(c, move)))
            return rc;
        return finish_stage(c, cnt, mx, accepted);
    }
    const bool graph = use_graph && !c->timing && move < 2;
    if (graph)
        for (auto& G : c->groups)
            if ((rc = capture_round_graph(c, G, move))) return rc;
    HIPCHK(hipMemcpyAsync(c->d_counts, cnt.data(), R * sizeof(unsigned int), hipMemcpyHostToDevice, c->stream));
    rc = launch(c, SPIMC_KERNEL_DECIDE, k_stage_begin, (R + 63) / 64, 64, 0, c->stream, c->d_reps, R, c->d_counts);
    if (rc) return rc;
    const int n_groups = (int)c->groups.size();
    std::vector<uint32_t> gmx(n_groups, 0);
    for (int g = 0; g < n_groups; ++g)
        for (int r = c->groups[g].r0; r < c->groups[g].r0 + c->groups[g].n; ++r) gmx[g] = std::max(gmx[g], cnt[r]);
    if (n_groups > 1) {
        HIPCHK(hipEventRecord(c->stage_ready, c->stream));
        for (int g = 1; g < n_groups; ++g) HIPCHK(hipStreamWaitEvent(c->groups[g].stream, c->stage_ready, 0));
    }
    // one-launch rounds: each group's first proposals, then rounds that decide
    // and propose on their own (the last round's proposals find no attempt left)
    const bool one = one_launch_now(c, move);
    if (one)
        for (int g = 0; g < n_groups; ++g) {
            const Group& G = c->groups[g];
            if (!gmx[g]) continue;
            rc = move == 0 ? launch(c, SPIMC_KERNEL_PROPOSE, k_decide_propose_bisect, G.n, kBisBlock, 0, G.stream,
                                    c->d_reps + G.r0, G.n, (unsigned long long*)nullptr)
                           : launch(c, SPIMC_KERNEL_PROPOSE, k_decide_propose_displace, G.n, kDisBlock, 0, G.stream,
                                    c->d_reps + G.r0, G.n, (unsigned long long*)nullptr);
            if (rc) return rc;
        }
    // rounds issued group by group in turn, so the streams advance together
    const uint32_t step = graph ? kRoundsPerGraph : 1;
    for (uint32_t t = 0; t < mx; t += step)
        for (int g = 0; g < n_groups; ++g) {
            Group& G = c->groups[g];
            if (t >= gmx[g]) continue;
            if (graph)
                HIPCHK(hipGraphLaunch(G.graph[move], G.stream));
            else if ((rc = launch_round(c, G, move)))
                return rc;
        }
    if (move != 3 && !one)
        for (int g = 0; g < n_groups; ++g)
            if (gmx[g] && (rc = launch_final_decide(c, c->groups[g]))) return rc;
    for (int g = 1; g < n_groups; ++g) {
        HIPCHK(hipEventRecord(c->groups[g].done, c->groups[g].stream));
        HIPCHK(hipStreamWaitEvent(c->stream, c->groups[g].done, 0));
    }
    return finish_stage(c, cnt, mx, accepted);
}

// After a stage's launches: wait, report replay errors, copy the accept flags.
int finish_stage(spimc_ctx* c, const std::vector<unsigned int>& cnt, uint32_t mx, uint8_t* accepted) {
    const int R = (int)c->reps.size();
    int rc;
    if ((rc = sync_and_resolve(c))) return rc;
    c->adjoint_fresh = false;
    // replay errors: one context-wide flag, then the replica that raised it
    unsigned int any = 0;
    HIPCHK(hipMemcpy(&any, c->d_error, sizeof any, hipMemcpyDeviceToHost));
    if (any) {
        HIPCHK(hipMemset(c->d_error, 0, sizeof any));
        for (int r = 0; r < R; ++r) {
            Counters k;
            HIPCHK(hipMemcpy(&k, c->reps[r].dev.cnt, sizeof(Counters), hipMemcpyDeviceToHost));
            if (k.error) {
                unsigned long long z = 0;
                HIPCHK(hipMemcpy(&c->reps[r].dev.cnt->error, &z, sizeof z, hipMemcpyHostToDevice));
                char b[128];
                snprintf(b, sizeof b, "replica %d: replay tape exhausted or invalid draw", r);
                return fail(SPIMC_ERR_STATE, b);
            }
        }
    }
    for (int r = 0; r < R; ++r) c->reps[r].stage_attempts = cnt[r];
    if (accepted)
        for (int r = 0; r < R; ++r)
            HIPCHK(hipMemcpy(accepted + (size_t)r * mx, c->reps[r].accepted, cnt[r], hipMemcpyDeviceToHost));
    return SPIMC_OK;
}

// Rebuild rho_k of one replica from its committed positions (buffer sel[b] = 0).
int rho_rebuild(spimc_ctx* c, uint32_t r) {
    const Rep& R = c->reps[r].dev;
    if (!R.use_long_range) return SPIMC_OK;
    HIPCHK(hipMemsetAsync(R.sel, 0, R.M, c->stream));
    const int chunks = (R.nk + 256 * kBuildKpt - 1) / (256 * kBuildKpt);
    dim3 g(R.M, chunks);
    return launch(c, SPIMC_KERNEL_ESTIMATORS, k_rho_build<kBuildKpt>, g, 256, 0, c->stream, c->d_reps, (int)r);
}

// ---- estimator drivers (launch over all replicas or one) -------------------
// `base` offsets the replica table so that gridDim.y counts replicas from it.
int est_kin_dbeta(spimc_ctx* c, Rep* base, int nrep) {
    dim3 g(c->max_M, nrep);
    return launch(c, SPIMC_KERNEL_ESTIMATORS, k_kin_dbeta, g, 256, 0, c->stream, base, kSlotKin);
}
int est_pair_whole(spimc_ctx* c, Rep* base, int nrep, int which) {
    int rc;
    if (c->lds_whole_max > 0) {
        const size_t shmem = (size_t)c->lds_whole_max * sizeof(double);
        // workgroups per replica, each over every spb-th slice: enough to fill the
        // GPU when the replicas are few
        const int spb = std::max(1, std::min(c->max_M, (2 * kNumCUs + nrep - 1) / nrep));
        rc = launch(c, SPIMC_KERNEL_ESTIMATORS, k_pair_whole_lds, dim3(nrep, spb), kAdjLdsBlock, shmem, c->stream, base,
                    which, kSlotPair);
    } else {
        dim3 g(c->max_M, nrep);
        rc = launch(c, SPIMC_KERNEL_ESTIMATORS, k_pair_whole, g, 256, 0, c->stream, base, which, kSlotPair);
    }
    if (rc || !c->any_long) return rc;
    dim3 g2((c->max_nk + 255) / 256, nrep);
    return launch(c, SPIMC_KERNEL_ESTIMATORS, k_long_whole, g2, 256, 0, c->stream, base, which);
}
int est_finalize(spimc_ctx* c, Rep* base, int nrep, int what, int acc) {
    return launch(c, SPIMC_KERNEL_ESTIMATORS, k_finalize, (nrep + 63) / 64, 64, 0, c->stream, base, nrep, what, acc);
}
int est_gradient(spimc_ctx* c, Rep* base, int nrep, uint32_t terms, int compat, int out_est) {
    dim3 g((c->max_N + 63) / 64, c->max_M, nrep);
    int rc;
    if (compat) {
        rc = launch(c, SPIMC_KERNEL_ADJOINT, k_grad_compat, g, 64, 0, c->stream, base, out_est,
                    (terms & SPIMC_GRAD_PAIR_SHORT) ? 1 : 0);
        bool ilkka_long = false;
        for (int r = 0; r < nrep; ++r) {
            const Rep& Rr = c->reps[(base - c->d_reps) + r].dev;
            ilkka_long |= Rr.pair_type == SPIMC_PAIR_ILKKA && Rr.use_long_range;
        }
        if (!ilkka_long) return rc;
    } else if (c->lds_adj_max > 0) {
        const size_t shmem = (size_t)c->lds_adj_max * sizeof(double);
        // workgroups per replica: enough to fill the GPU when the replicas are few
        const int bpr = std::max(1, std::min(c->max_M, (2 * kNumCUs + nrep - 1) / nrep));
        // pair-once kernel: at most max_M / 2 link ranges of 16 waves' worth per
        // replica; chosen when they give at least kAdjPairsMinWaves waves
        const int rng_max = std::max(1, c->max_M / 2);
        const int bpr2 = std::max(1, std::min((rng_max + kAdjPairWaves - 1) / kAdjPairWaves,
                                              (int)((kAdjPairsMinWaves + kAdjPairWaves * nrep - 1) / (kAdjPairWaves * nrep))));
        const long long waves2 = (long long)nrep * std::min(kAdjPairWaves * bpr2, rng_max);
        const bool pairs = c->adj_pairs_ok &&
                           (c->adj_pairs_env >= 0 ? c->adj_pairs_env != 0 : waves2 >= kAdjPairsMinWaves);
        if (pairs) {
            const size_t shmem2 = shmem + (size_t)kAdjPairWaves * kAdjSlotDoubles * sizeof(double);
            rc = launch(c, SPIMC_KERNEL_ADJOINT, c->adj_pairs_fn, dim3(nrep, bpr2), kAdjLdsBlock, shmem2, c->stream,
                        base, terms, out_est, 0);
            if (!rc)
                rc = launch(c, SPIMC_KERNEL_ADJOINT, c->adj_pairs_fn, dim3(nrep, bpr2), kAdjLdsBlock, shmem2, c->stream,
                            base, terms, out_est, 1);
        } else {
            rc = launch(c, SPIMC_KERNEL_ADJOINT, c->adj_lds_fn, dim3(nrep, bpr), kAdjLdsBlock, shmem, c->stream, base,
                        terms, out_est);
        }
    } else {
        rc = launch(c, SPIMC_KERNEL_ADJOINT, k_adjoint_local, g, 64, 0, c->stream, base, terms, out_est);
    }
    if (rc || !(terms & SPIMC_GRAD_PAIR_LONG) || !c->any_long) return rc;
    if (c->adj_runs_lds > 0) {
        dim3 gr(c->max_M, nrep);
        return launch(c, SPIMC_KERNEL_ADJOINT, k_adjoint_long_runs, gr, kAdjRunBlock, c->adj_runs_lds, c->stream, base,
                      out_est, compat);
    }
    int mx = 0;
    for (auto& h : c->reps) mx = std::max(mx, h.dev.max_idx);
    if (mx > kAdjMaxIdx) return fail(SPIMC_ERR_UNSUPPORTED, "adjoint long range supports |n| <= 40");
    const size_t shmem = (size_t)64 * 3 * (mx + 1) * sizeof(double2) + kAdjTile * (sizeof(double2) + sizeof(int) + sizeof(double));
    return launch(c, SPIMC_KERNEL_ADJOINT, k_adjoint_long, g, 64, shmem, c->stream, base, out_est, compat);
}

int energy_accumulate(spimc_ctx* c, Rep* base, int nrep) {
    int rc;
    if (c->energy_estimator == SPIMC_ENERGY_THERMAL) {
        if ((rc = est_kin_dbeta(c, base, nrep)) || (rc = est_finalize(c, base, nrep, 0, 0)) ||
            (rc = est_pair_whole(c, base, nrep, 0)) || (rc = est_finalize(c, base, nrep, 1, 1)))
            return rc;
        return SPIMC_OK;
    }
    const int W = (int)c->virial_window;
    rc = launch(c, SPIMC_KERNEL_ESTIMATORS, k_kin_virial, nrep, 256, 0, c->stream, base, W, kSlotVirial);
    if (rc) return rc;
    int from_sweep = 0;
    if (c->energy_estimator == SPIMC_ENERGY_VIRIAL_COMPAT) {
        rc = est_gradient(c, base, nrep, SPIMC_GRAD_PAIR_SHORT | SPIMC_GRAD_PAIR_LONG, 1, 1);
    } else if (c->adjoint_fresh) {
        from_sweep = 1;  // dS/dr of the sweep holds the kinetic term too; the centroid kernel removes it
    } else {
        rc = est_gradient(c, base, nrep, SPIMC_GRAD_PAIR_SHORT | SPIMC_GRAD_PAIR_LONG, 0, 1);
    }
    if (rc) return rc;
    dim3 g(c->max_M, nrep);
    rc = launch(c, SPIMC_KERNEL_ESTIMATORS, k_centroid, g, 256, 0, c->stream, base, kSlotCentroid, from_sweep);
    if (rc) return rc;
    if ((rc = est_finalize(c, base, nrep, 2, 0)) || (rc = est_pair_whole(c, base, nrep, 0)) ||
        (rc = est_finalize(c, base, nrep, 1, 1)))
        return rc;
    return SPIMC_OK;
}

int obs_accumulate_all(spimc_ctx* c, uint32_t which) {
    const int R = (int)c->reps.size();
    int rc;
    if (which & SPIMC_OBS_ENERGY)
        if ((rc = energy_accumulate(c, c->d_reps, R))) return rc;
    if ((which & SPIMC_OBS_SK) && c->any_long) {
        dim3 g((c->max_nk + 255) / 256, R);
        const double kc = c->sk_k_cut;
        rc = launch(c, SPIMC_KERNEL_ESTIMATORS, k_sk, g, 256, 0, c->stream, c->d_reps, kc);
        if (rc) return rc;
    }
    if (which & SPIMC_OBS_GR) {
        dim3 g(c->max_M, R);
        const double rmin = c->gr_r_min;
        const int nr = (int)c->gr_n_r;
        rc = launch(c, SPIMC_KERNEL_ESTIMATORS, k_gr, g, 256, 0, c->stream, c->d_reps, rmin, 0., nr, c->d_gr_rmax);
        if (rc) return rc;
    }
    return SPIMC_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int spimc_abi_version(void) { return SPIMC_ABI_VERSION; }
const char* spimc_last_error(void) { return g_err.c_str(); }

int spimc_kspace_build(double L, double k_cut, int32_t* k_n, uint32_t cap, uint32_t* n_k) {
    // KSpace::Setup / Include (k_space_class.h:17-80): floating-point test on k = n * 2pi/L
    if (!(L > 0.) || !(k_cut >= 0.) || !n_k) return fail(SPIMC_ERR_ARG, "bad k-space arguments");
    const double box = 2. * M_PI / L;
    const int mx = (int)(uint32_t)std::ceil(1.1 * k_cut / box);
    uint32_t cnt = 0;
    for (int a = -mx; a <= mx; ++a)
        for (int b = -mx; b <= mx; ++b)
            for (int z = -mx; z <= mx; ++z) {
                const double kx = a * box, ky = b * box, kz = z * box;
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (!(k2 < k_cut * k_cut) || k2 == 0.) continue;
                const bool half = (kx > 0.) || (kx == 0. && ky > 0.) || (kx == 0. && ky == 0. && kz > 0.);
                if (!half) continue;
                if (k_n && cnt < cap) {
                    k_n[3 * cnt] = a;
                    k_n[3 * cnt + 1] = b;
                    k_n[3 * cnt + 2] = z;
                }
                ++cnt;
            }
    *n_k = cnt;
    return SPIMC_OK;
}

int spimc_ctx_create(const spimc_system_desc* descs, uint32_t n_replicas, uint64_t seed, spimc_ctx** out) {
    if (!descs || !out || n_replicas == 0) return fail(SPIMC_ERR_ARG, "null descriptors or zero replicas");
    *out = nullptr;
    for (uint32_t r = 0; r < n_replicas; ++r) {
        int rc = validate(descs[r], r);
        if (rc) return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SPIMC_ERR_NODEVICE, "no HIP device");
    spimc_ctx* c = new spimc_ctx();
    auto bail = [&](int rc) {
        spimc_ctx_destroy(c);
        return rc;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(SPIMC_ERR_HIP, "hipGetDevice"));
    // the device scratch of spimc_get_action_noalloc; a failed lookup disables only that entry point
    if (noalloc_scratch_ptr(c->device, &c->noalloc_scratch) != SPIMC_OK) {
        c->noalloc_error = g_err;
        c->noalloc_scratch = nullptr;
    }
    if (const char* f = std::getenv("SPIMC_RHO_KPT")) {
        const int v = std::atoi(f);
        if (v == 2 || v == 4 || v == 8 || v == 16) c->rho_kpt = v;
    }
    if (const char* f = std::getenv("SPIMC_FUSE_ROUNDS")) c->fuse_rounds = std::atoi(f) != 0;
    if (const char* f = std::getenv("SPIMC_RHO_NT")) c->rho_nt = std::atoi(f) != 0;
    // Launch shape (measured best on MI355X at 1024 replicas of N=256, M=128):
    // with enough replicas to fill the GPU, each round's rho_k kernel runs on a
    // second stream beside its pair kernel, looping over its items with 2
    // workgroups per CU so that every CU keeps room for a pair workgroup.
    // SPIMC_STREAMS (replica groups), SPIMC_FORK and SPIMC_RHO_BPC override.
    c->groups_wanted = 1;
    if (const char* f = std::getenv("SPIMC_STREAMS")) c->groups_wanted = std::max(1, std::atoi(f));
    c->fork = n_replicas >= 64;
    if (const char* f = std::getenv("SPIMC_FORK")) c->fork = std::atoi(f) != 0;
    if (const char* f = std::getenv("SPIMC_PERM_STAGE")) c->perm_stage = std::max(0, std::min(kPermStage, std::atoi(f)));
    c->rho_bpc = c->fork ? 2 : c->groups_wanted > 1 ? 3 : 0;
    if (const char* f = std::getenv("SPIMC_RHO_BPC")) c->rho_bpc = std::max(0, std::atoi(f));
    {
        using RhoFn = void (*)(const Rep*, const uint2*, int, int, unsigned long long*);
        const RhoFn fns[2][4] = {{k_rho<2, false>, k_rho<4, false>, k_rho<8, false>, k_rho<16, false>},
                                 {k_rho<2, true>, k_rho<4, true>, k_rho<8, true>, k_rho<16, true>}};
        const int ki = c->rho_kpt == 2 ? 0 : c->rho_kpt == 4 ? 1 : c->rho_kpt == 8 ? 2 : 3;
        c->rho_fn = fns[c->rho_nt ? 1 : 0][ki];
        using RhoLoopFn = void (*)(const RhoItem*, int, unsigned long long*);
        const RhoLoopFn lfns[2][4] = {
            {k_rho_loop<2, false>, k_rho_loop<4, false>, k_rho_loop<8, false>, k_rho_loop<16, false>},
            {k_rho_loop<2, true>, k_rho_loop<4, true>, k_rho_loop<8, true>, k_rho_loop<16, true>}};
        c->rho_loop_fn = lfns[c->rho_nt ? 1 : 0][ki];
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SPIMC_ERR_HIP, "stream creation failed"));
    c->reps.resize(n_replicas);
    for (uint32_t r = 0; r < n_replicas; ++r) {
        int rc = build_replica(c, r, descs[r], seed);
        if (rc) return bail(rc);
        const Rep& R = c->reps[r].dev;
        c->any_long = c->any_long || R.use_long_range;
        c->max_M = std::max(c->max_M, R.M);
        c->max_N = std::max(c->max_N, R.N);
        c->max_nk = std::max(c->max_nk, R.nk);
    }
    if (hipMalloc((void**)&c->d_reps, n_replicas * sizeof(Rep)) != hipSuccess) return bail(fail(SPIMC_ERR_NOMEM, "rep table"));
    if (hipMalloc((void**)&c->d_counts, n_replicas * sizeof(unsigned int)) != hipSuccess)
        return bail(fail(SPIMC_ERR_NOMEM, "counts"));
    if (hipMalloc((void**)&c->d_work, WORK_COUNT * sizeof(unsigned long long)) != hipSuccess)
        return bail(fail(SPIMC_ERR_NOMEM, "work counters"));
    if (hipMemset(c->d_work, 0, WORK_COUNT * sizeof(unsigned long long)) != hipSuccess)
        return bail(fail(SPIMC_ERR_HIP, "memset"));
    {
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            return bail(fail(SPIMC_ERR_HIP, "wall clock rate"));
        c->clock_hz = 1e3 * khz;
    }
    if (hipMalloc((void**)&c->d_gr_rmax, n_replicas * sizeof(double)) != hipSuccess)
        return bail(fail(SPIMC_ERR_NOMEM, "g(r) table"));
    if (hipMalloc((void**)&c->d_error, sizeof(unsigned int)) != hipSuccess ||
        hipMemset(c->d_error, 0, sizeof(unsigned int)) != hipSuccess)
        return bail(fail(SPIMC_ERR_NOMEM, "error flag"));
    for (auto& h : c->reps) h.dev.ctx_error = c->d_error;
    {
        bool all = true;
        for (auto& h : c->reps) {
            if (h.dev.pair_type != SPIMC_PAIR_NONE && h.dev.lds_doubles == 0) all = false;
            c->lds_doubles_max = std::max(c->lds_doubles_max, h.dev.lds_doubles);
        }
        const char* f = std::getenv("SPIMC_PAIR_LDS");
        if (!all || (f && std::atoi(f) == 0)) c->lds_doubles_max = 0;
        // one pair type across the replicas selects the specialised instantiation
        int pt = -1, nord = -1;
        for (auto& h : c->reps)
            if (h.dev.pair_type != SPIMC_PAIR_NONE) {
                pt = (pt < 0 || pt == h.dev.pair_type) ? h.dev.pair_type : 0;
                nord = (nord < 0 || nord == h.dev.n_order) ? h.dev.n_order : 0;
            }
        // 512-thread pair workgroups leave each CU room for the rho_k workgroups of
        // the other stream
        int bs = (c->groups_wanted > 1 || c->fork) ? 512 : kPairLdsBlock;
        if (const char* fb = std::getenv("SPIMC_PAIR_BLOCK")) {
            const int v = std::atoi(fb);
            if (v == 256 || v == 512 || v == 768 || v == 1024) bs = v;
        }
        const int vi = pt == SPIMC_PAIR_BARE ? 0 : pt == SPIMC_PAIR_DAVID ? (nord == 2 ? 1 : 2) : 3;
        c->pair_variant = vi;
        if (set_pair_block(c, bs) != SPIMC_OK) c->lds_doubles_max = 0;
        {
            // one-launch rounds (SPIMC_ONE_LAUNCH or spimc_autotune): the pair type's
            // instantiation with the default rho_k chunking
            using RoundFn = void (*)(const Rep*, int, int, const RhoItem*, const int4*, int, unsigned int*);
            const RoundFn rf[2][2][4] = {
                {{k_round<1, 0, kRhoKpt, false, 0>, k_round<2, 2, kRhoKpt, false, 0>, k_round<2, 0, kRhoKpt, false, 0>,
                  k_round<0, 0, kRhoKpt, false, 0>},
                 {k_round<1, 0, kRhoKpt, true, 0>, k_round<2, 2, kRhoKpt, true, 0>, k_round<2, 0, kRhoKpt, true, 0>,
                  k_round<0, 0, kRhoKpt, true, 0>}},
                {{k_round<1, 0, kRhoKpt, false, 1>, k_round<2, 2, kRhoKpt, false, 1>, k_round<2, 0, kRhoKpt, false, 1>,
                  k_round<0, 0, kRhoKpt, false, 1>},
                 {k_round<1, 0, kRhoKpt, true, 1>, k_round<2, 2, kRhoKpt, true, 1>, k_round<2, 0, kRhoKpt, true, 1>,
                  k_round<0, 0, kRhoKpt, true, 1>}}};
            for (int mv = 0; mv < 2; ++mv) c->round_fn[mv] = rf[mv][c->rho_nt ? 1 : 0][vi];
            if (c->lds_doubles_max > 0)
                for (int mv = 0; mv < 2; ++mv)
                    if (hipFuncSetAttribute((const void*)c->round_fn[mv], hipFuncAttributeMaxDynamicSharedMemorySize,
                                            c->lds_doubles_max * (int)sizeof(double)) != hipSuccess)
                        c->round_fn[0] = c->round_fn[1] = nullptr;
            if (const char* fo = std::getenv("SPIMC_ONE_LAUNCH")) {
                c->one_launch = std::atoi(fo) != 0;
                c->one_moves = std::atoi(fo) >= 2 ? 2 : 1;
            }
        }
        // stages of small replicas in one persistent launch: off by default, set by
        // spimc_autotune or SPIMC_PERSISTENT (workgroup size 256, 512 or 1024)
        if (const char* fp = std::getenv("SPIMC_PERSISTENT")) {
            const int v = std::atoi(fp);
            if (v && set_pers_block(c, v) != SPIMC_OK) return bail(SPIMC_ERR_ARG);
        }
        c->lds_adj_max = c->lds_doubles_max;
        c->adj_lds_fn = vi == 0 ? k_adjoint_local_lds<1, 0> : vi == 1 ? k_adjoint_local_lds<2, 2>
                                                        : vi == 2 ? k_adjoint_local_lds<2, 0> : k_adjoint_local_lds<0, 0>;
        c->adj_pairs_fn = vi == 0 ? k_adjoint_pairs_lds<1, 0> : vi == 1 ? k_adjoint_pairs_lds<2, 2>
                                                            : vi == 2 ? k_adjoint_pairs_lds<2, 0> : k_adjoint_pairs_lds<0, 0>;
        if (const char* fa = std::getenv("SPIMC_ADJ_PAIRS")) c->adj_pairs_env = std::atoi(fa) != 0;
        if (c->lds_adj_max > 0 &&
            hipFuncSetAttribute((const void*)c->adj_lds_fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                c->lds_adj_max * (int)sizeof(double)) != hipSuccess)
            c->lds_adj_max = 0;
        {
            // the pair-once kernel adds its waves' partner slots to the tables
            const size_t b2 = ((size_t)c->lds_adj_max + (size_t)kAdjPairWaves * kAdjSlotDoubles) * sizeof(double);
            c->adj_pairs_ok = c->lds_adj_max > 0 && b2 <= kPairLdsMaxBytes &&
                              hipFuncSetAttribute((const void*)c->adj_pairs_fn,
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)b2) == hipSuccess;
        }
        bool whole = c->lds_doubles_max > 0;
        for (auto& h : c->reps) {
            const int need = h.dev.pair_type == SPIMC_PAIR_DAVID ? h.dev.lds_doubles_du : h.dev.lds_doubles;
            if (h.dev.pair_type != SPIMC_PAIR_NONE && need == 0) whole = false;
            c->lds_whole_max = std::max(c->lds_whole_max, need);
        }
        if (!whole) c->lds_whole_max = 0;
        for (auto& h : c->reps)
            if (h.dev.use_long_range)
                c->adj_runs_lds = std::max(c->adj_runs_lds, (size_t)h.dev.nk * sizeof(double2) + (size_t)h.dev.n_runs * sizeof(int4));
        const char* fr = std::getenv("SPIMC_ADJ_RUNS");
        if (c->adj_runs_lds > kPairLdsMaxBytes || (fr && std::atoi(fr) == 0)) c->adj_runs_lds = 0;
        if (c->adj_runs_lds > 0 &&
            hipFuncSetAttribute((const void*)k_adjoint_long_runs, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)c->adj_runs_lds) != hipSuccess)
            c->adj_runs_lds = 0;
        if (c->lds_whole_max > 0 &&
            hipFuncSetAttribute((const void*)k_pair_whole_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                c->lds_whole_max * (int)sizeof(double)) != hipSuccess)
            c->lds_whole_max = 0;
    }
    if (std::getenv("SPIMC_VERBOSE"))
        fprintf(stderr,
                "spimc: %u replica(s), pair kernel %s (%d LDS doubles, %d threads), rho_k kernel %s (%d k/thread, %d "
                "workgroups/CU), long-range adjoint %s, adjoint %s, estimators %s\n",
                n_replicas, c->lds_doubles_max ? "k_pair_lds" : "k_pair",
                c->lds_doubles_max, c->pair_block,
                "k_rho_loop",
                c->rho_kpt, c->rho_bpc, c->adj_runs_lds ? "k_adjoint_long_runs" : "k_adjoint_long",
                c->lds_adj_max ? "k_adjoint_local_lds" : "k_adjoint_local", c->lds_whole_max ? "k_pair_whole_lds" : "k_pair_whole");
    int rc = upload_reps(c);
    if (rc) return bail(rc);
    {
        if ((rc = build_groups(c, c->groups_wanted))) return bail(rc);
        if ((rc = upload_reps(c))) return bail(rc);  // the replicas now point at their rho_k items
        if (std::getenv("SPIMC_VERBOSE"))
            fprintf(stderr, "spimc: %d replica group(s), rho_k beside the pair kernel on a second stream: %s\n",
                    (int)c->groups.size(), c->fork ? "yes" : "no");
    }
    if ((rc = spimc_obs_configure(c, SPIMC_ENERGY_THERMAL, 0, 1e300, 0., 0., 100))) return bail(rc);
    *out = c;
    return SPIMC_OK;
}

int spimc_ctx_destroy(spimc_ctx* c) {
    if (!c) return SPIMC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& G : c->groups) {
        if (G.stream) (void)hipStreamSynchronize(G.stream);
        for (auto& g : G.graph)
            if (g) (void)hipGraphExecDestroy(g);
        for (void* p : {(void*)G.d_pair_bis, (void*)G.d_pair_dis, (void*)G.d_ritem_bis, (void*)G.d_ritem_dis, (void*)G.d_kt,
                        (void*)G.d_multi, (void*)G.d_rgrp[0], (void*)G.d_rgrp[1],
                        (void*)G.d_arrive})
            if (p) (void)hipFree(p);
        if (G.done) (void)hipEventDestroy(G.done);
        if (G.side) {
            (void)hipStreamSynchronize(G.side);
            (void)hipStreamDestroy(G.side);
        }
        if (G.ev_fork) (void)hipEventDestroy(G.ev_fork);
        if (G.ev_join) (void)hipEventDestroy(G.ev_join);
        if (G.stream && G.stream != c->stream) (void)hipStreamDestroy(G.stream);
    }
    if (c->stage_ready) (void)hipEventDestroy(c->stage_ready);
    for (auto& h : c->reps) {
        if (h.arena) (void)hipFree(h.arena);
        if (h.accepted) (void)hipFree(h.accepted);
        if (h.trace) (void)hipFree(h.trace);
        if (h.tape) (void)hipFree(h.tape);
        if (h.gr_cnt) (void)hipFree(h.gr_cnt);
        if (h.perm_arena) (void)hipFree(h.perm_arena);
        if (h.perm_scratch) (void)hipFree(h.perm_scratch);
        if (h.perm_t) (void)hipFree(h.perm_t);
    }
    for (void* p : {(void*)c->d_reps, (void*)c->d_counts, (void*)c->d_work, (void*)c->d_gr_rmax, (void*)c->d_error,
                    (void*)c->d_tmp, (void*)c->d_items_tmp})
        if (p) (void)hipFree(p);
    for (auto& t : c->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto& e : c->pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SPIMC_OK;
}

int spimc_ctx_num_replicas(const spimc_ctx* c, uint32_t* out) {
    if (!c || !out) return fail(SPIMC_ERR_ARG, "null argument");
    *out = (uint32_t)c->reps.size();
    return SPIMC_OK;
}

int spimc_path_upload(spimc_ctx* c, uint32_t r, const double* pos) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!pos) return fail(SPIMC_ERR_ARG, "null positions");
    const Rep& R = c->reps[r].dev;
    HIPCHK(hipMemcpyAsync(R.pos, pos, (size_t)3 * R.M * R.N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = rho_rebuild(c, r))) return rc;
    c->adjoint_fresh = false;
    return sync_and_resolve(c);
}

int spimc_path_download(spimc_ctx* c, uint32_t r, double* pos) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!pos) return fail(SPIMC_ERR_ARG, "null positions");
    const Rep& R = c->reps[r].dev;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(pos, R.pos, (size_t)3 * R.M * R.N * sizeof(double), hipMemcpyDeviceToHost));
    return SPIMC_OK;
}

int spimc_rho_k_download(spimc_ctx* c, uint32_t r, double* rho) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!rho) return fail(SPIMC_ERR_ARG, "null output");
    const Rep& R = c->reps[r].dev;
    if (!R.use_long_range) return fail(SPIMC_ERR_STATE, "replica has no long-range action");
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<unsigned char> sel(R.M);
    HIPCHK(hipMemcpy(sel.data(), R.sel, R.M, hipMemcpyDeviceToHost));
    for (int b = 0; b < R.M; ++b)
        HIPCHK(hipMemcpy(rho + (size_t)2 * b * R.nk, R.rho + ((size_t)sel[b] * R.M + b) * R.nk,
                         (size_t)R.nk * sizeof(double2), hipMemcpyDeviceToHost));
    return SPIMC_OK;
}

int spimc_rho_k_rebuild(spimc_ctx* c, uint32_t r) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if ((rc = rho_rebuild(c, r))) return rc;
    return sync_and_resolve(c);
}

// ---- Action interface --------------------------------------------------------

int spimc_set_proposal(spimc_ctx* c, uint32_t r, uint32_t particle, uint32_t b_first, uint32_t n_beads, const double* r_new) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    const Rep& R = c->reps[r].dev;
    if (particle >= (uint32_t)R.N || n_beads == 0 || n_beads > (uint32_t)R.M || !r_new)
        return fail(SPIMC_ERR_ARG, "bad proposal");
    if (c->prop_set && c->prop_rep != (int)r) return fail(SPIMC_ERR_STATE, "a proposal on another replica is pending");
    HIPCHK(hipMemcpy(R.prop_new, r_new, (size_t)3 * n_beads * sizeof(double), hipMemcpyHostToDevice));
    c->prop_set = true;
    c->rho_updated = false;
    c->prop_rep = r;
    c->prop_p = particle;
    c->prop_first = b_first % R.M;
    c->prop_count = n_beads;
    return SPIMC_OK;
}

static int write_iface_prop(spimc_ctx* c, uint32_t r, int p, int b0, int n_links) {
    Prop pr{};
    pr.status = 0;
    pr.move = 2;
    pr.p = p;
    pr.p2 = p;
    pr.b0 = b0;
    pr.n_links = n_links;
    const bool mine = c->prop_set && c->prop_rep == (int)r && c->prop_p == p;
    pr.first = mine ? c->prop_first : 0;
    pr.count = mine ? c->prop_count : 0;
    HIPCHK(hipMemcpy(c->reps[r].dev.prop, &pr, sizeof(Prop), hipMemcpyHostToDevice));
    return SPIMC_OK;
}

static int finish_proposal(spimc_ctx* c, uint32_t r, int accept) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!c->prop_set || c->prop_rep != (int)r) return fail(SPIMC_ERR_STATE, "no proposal pending on this replica");
    if ((rc = write_iface_prop(c, r, c->prop_p, c->prop_first, 0))) return rc;
    const int upd = c->rho_updated ? 1 : 0;
    rc = launch(c, SPIMC_KERNEL_DECIDE, k_commit, 1, 256, 0, c->stream, c->d_reps, (int)r, accept, upd);
    if (rc) return rc;
    c->prop_set = false;
    c->rho_updated = false;
    c->adjoint_fresh = false;
    return sync_and_resolve(c);
}

int spimc_accept_proposal(spimc_ctx* c, uint32_t r) { return finish_proposal(c, r, 1); }
int spimc_reject_proposal(spimc_ctx* c, uint32_t r) { return finish_proposal(c, r, 0); }

namespace {

// Device memory of one GetAction query beyond its inputs and outputs: per-slice
// long-range partials and the result (tmp, doubles) and the pair / rho_k work
// items (items).  spimc_get_action grows context buffers for them;
// spimc_get_action_noalloc carves them from the device scratch array.
struct IfaceScratch {
    double* tmp;
    uint2* items;
    size_t tmp_cap, items_cap;
};

int get_action_impl(spimc_ctx* c, uint32_t r, int action, int mode, uint32_t b0, uint32_t b1, const uint32_t* parts,
                    uint32_t n_parts, uint32_t level, double* out, const IfaceScratch* fixed) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!out || (n_parts && !parts)) return fail(SPIMC_ERR_ARG, "null argument");
    if (n_parts > 1) return fail(SPIMC_ERR_UNSUPPORTED, "at most one moved particle per query");
    if (b1 < b0) return fail(SPIMC_ERR_ARG, "b1 < b0");
    const Rep& R = c->reps[r].dev;
    if (n_parts == 1 && parts[0] >= (uint32_t)R.N) return fail(SPIMC_ERR_ARG, "particle index out of range");
    if (b0 >= (uint32_t)R.M || b1 - b0 > (uint32_t)R.M) return fail(SPIMC_ERR_ARG, "window out of range");
    *out = 0.;
    if (n_parts == 0) return SPIMC_OK;
    const int p = (int)parts[0];
    const uint32_t skip = 1u << level;
    const size_t need_tmp = (size_t)R.M + 8, need_items = (size_t)R.M * std::max(R.n_chunks, 1);
    IfaceScratch sc;
    if (fixed) {
        sc = *fixed;
        if (need_tmp > sc.tmp_cap || need_items > sc.items_cap)
            return fail(SPIMC_ERR_ARG, "query too large for the fixed scratch: n_bead + 8 partials and n_bead * k chunks "
                                       "work items must fit spimc_get_action_noalloc's scratch (SPIMC_NOALLOC_SCRATCH_BYTES)");
    } else {
        if ((rc = ensure_tmp(c, need_tmp, need_items))) return rc;
        sc = IfaceScratch{c->d_tmp, c->d_items_tmp, c->tmp_cap, c->items_tmp_cap};
    }
    double* d_out = sc.tmp + R.M;
    if (action == SPIMC_ACTION_KINETIC) {
        if (level >= kMaxKinSplines || (R.kin_images > 0 && level >= c->reps[r].kin.size()))
            return fail(SPIMC_ERR_UNSUPPORTED, "kinetic level out of range");
        uint32_t span = (b1 - b0) % (uint32_t)R.M;
        // a whole-particle walk returns to its first bead only when the particle's
        // path closes on itself (kinetic_class.h:111-118)
        if (span == 0 && b1 > b0 && R.perm) {
            int nxt = p;
            HIPCHK(hipDeviceSynchronize());
            HIPCHK(hipMemcpy(&nxt, R.perm + p, sizeof(int), hipMemcpyDeviceToHost));
            if (nxt != p) span = (uint32_t)R.M;
        }
        if (span % skip) return fail(SPIMC_ERR_ARG, "window not a multiple of the level step");
        const int nl = (int)(span / skip);
        if ((rc = write_iface_prop(c, r, p, (int)b0, 0))) return rc;
        rc = launch(c, SPIMC_KERNEL_ESTIMATORS, k_kinetic_window, 1, 64, 0, c->stream, c->d_reps, (int)r, mode, p, (int)b0, nl, (int)level, d_out);
    } else {
        // PairAction::GetAction: 0 above max_level (= 0), and for is_constant
        if (level > 0 || R.pair_type == SPIMC_PAIR_NONE || R.N == 1) return SPIMC_OK;
        const int nl = (int)(b1 - b0);
        if ((rc = write_iface_prop(c, r, p, (int)b0, nl))) return rc;
        if (nl > 0) {
            // work items (r, link j), written on the device
            rc = launch(c, SPIMC_KERNEL_PAIR, k_iface_items, (nl + 255) / 256, 256, 0, c->stream, sc.items, (int)r, nl, 0);
            if (rc) return rc;
            const int g = (nl + 3) / 4;
            rc = launch(c, SPIMC_KERNEL_PAIR, k_pair, g, kPairBlock, 0, c->stream, c->d_reps, (const uint2*)sc.items, nl, 1, nullptr);
            if (rc) return rc;
        }
        int n_long = 0;
        if (R.use_long_range) {
            const bool mine = c->prop_set && c->prop_rep == (int)r && c->prop_p == p;
            if (mode == SPIMC_MODE_NEW && mine && !c->rho_updated) {
                // work items (r, moved bead m << 8 | k chunk), after the pair kernel has read its items
                const int nri = c->prop_count * R.n_chunks;
                rc = launch(c, SPIMC_KERNEL_RHO, k_iface_items, (nri + 255) / 256, 256, 0, c->stream, sc.items, (int)r, nri,
                            R.n_chunks);
                if (rc) return rc;
                const int gph = (6 * c->prop_count + 63) / 64;
                rc = launch(c, SPIMC_KERNEL_RHO, k_phase_proposal, gph, 64, 0, c->stream, c->d_reps, (int)r);
                if (rc) return rc;
                rc = launch(c, SPIMC_KERNEL_RHO, c->rho_fn, nri, kRhoBlock, 0, c->stream, c->d_reps, (const uint2*)sc.items, nri, 1, nullptr);
                if (rc) return rc;
                c->rho_updated = true;
            }
            n_long = (int)((b1 - b0 + skip - 1) / skip);
            const int upd = (mine && c->rho_updated) ? 1 : 0;
            if (n_long > 0) {
                rc = launch(c, SPIMC_KERNEL_RHO, k_long_window, n_long, 256, 0, c->stream, c->d_reps, (int)r, mode, (int)b0, upd, sc.tmp);
                if (rc) return rc;
            }
        }
        rc = launch(c, SPIMC_KERNEL_PAIR, k_action_sum, 1, 64, 0, c->stream, c->d_reps, (int)r, mode, nl, n_long, (const double*)sc.tmp, d_out);
    }
    if (rc) return rc;
    if ((rc = sync_and_resolve(c))) return rc;
    HIPCHK(hipMemcpy(out, d_out, sizeof(double), hipMemcpyDeviceToHost));
    return SPIMC_OK;
}

}  // namespace

int spimc_get_action(spimc_ctx* c, uint32_t r, int action, int mode, uint32_t b0, uint32_t b1, const uint32_t* parts,
                     uint32_t n_parts, uint32_t level, double* out) {
    return get_action_impl(c, r, action, mode, b0, b1, parts, n_parts, level, out, nullptr);
}

int spimc_get_action_noalloc(spimc_ctx* c, uint32_t r, int action, int mode, uint32_t b0, uint32_t b1,
                             const uint32_t* parts, uint32_t n_parts, uint32_t level, double* out) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    if (c->timing) return fail(SPIMC_ERR_STATE, "per-launch timing events are on; spimc_get_action_noalloc records none");
    if (!c->noalloc_scratch)
        return fail(SPIMC_ERR_HIP, "device scratch of this context's device is not available: " + c->noalloc_error);
    // the scratch array: long-range partials and result first, then the work items
    constexpr size_t kTmpBytes = kNoAllocScratchBytes / 4;
    IfaceScratch sc{reinterpret_cast<double*>(c->noalloc_scratch),
                    reinterpret_cast<uint2*>(static_cast<char*>(c->noalloc_scratch) + kTmpBytes),
                    kTmpBytes / sizeof(double), (kNoAllocScratchBytes - kTmpBytes) / sizeof(uint2)};
    return get_action_impl(c, r, action, mode, b0, b1, parts, n_parts, level, out, &sc);
}

int spimc_get_action_laplacian(spimc_ctx* c, uint32_t r, int action, int mode, uint32_t b0, uint32_t b1,
                               const uint32_t* parts, uint32_t n_parts, uint32_t level, double* out) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!out || (n_parts && !parts)) return fail(SPIMC_ERR_ARG, "null argument");
    if (n_parts > 1) return fail(SPIMC_ERR_UNSUPPORTED, "at most one particle per query");
    if (b1 < b0) return fail(SPIMC_ERR_ARG, "b1 < b0");
    if (level > (uint32_t)kMaxLevel) return fail(SPIMC_ERR_ARG, "level out of range");
    const Rep& R = c->reps[r].dev;
    if (n_parts == 1 && parts[0] >= (uint32_t)R.N) return fail(SPIMC_ERR_ARG, "particle index out of range");
    if (b0 >= (uint32_t)R.M || b1 - b0 > (uint32_t)R.M) return fail(SPIMC_ERR_ARG, "window out of range");
    *out = 0.;
    if (n_parts == 0) return SPIMC_OK;
    const int p = (int)parts[0];
    const uint32_t skip = 1u << level;
    if (action == SPIMC_ACTION_KINETIC) {
        // Kinetic::GetActionLaplacian (kinetic_class.h:147-165): n_d 4 / (4 lambda
        // tau skip) per step of the walk from (p, b0) to the bead b1 - b0 links
        // later, a closed form of the walk's length added in the reference's order
        uint32_t span = (b1 - b0) % (uint32_t)R.M;
        if (span == 0 && b1 > b0 && R.perm) {
            int nxt = p;
            HIPCHK(hipDeviceSynchronize());
            HIPCHK(hipMemcpy(&nxt, R.perm + p, sizeof(int), hipMemcpyDeviceToHost));
            if (nxt != p) span = (uint32_t)R.M;
        }
        if (span % skip) return fail(SPIMC_ERR_ARG, "window not a multiple of the level step");
        const double i4lt_l = (1. / (4. * R.lambda * R.tau)) / skip;
        double tot = 0.;
        for (uint32_t i = 0; i < span / skip; ++i) tot += 3 * 4. * i4lt_l;
        *out = tot;
        return SPIMC_OK;
    }
    // the pair action: 0 above max_level (= 0) and for is_constant
    if (level > 0 || R.pair_type == SPIMC_PAIR_NONE || R.N == 1 || b1 == b0) return SPIMC_OK;
    if ((rc = write_iface_prop(c, r, p, (int)b0, 0))) return rc;
    if ((rc = ensure_tmp(c, 1, 0))) return rc;
    rc = launch(c, SPIMC_KERNEL_ESTIMATORS, k_pair_laplacian, 1, kLapBlock, 0, c->stream, c->d_reps, (int)r, mode, (int)b0,
                (int)b1, p, c->d_tmp);
    if (rc) return rc;
    if ((rc = sync_and_resolve(c))) return rc;
    HIPCHK(hipMemcpy(out, c->d_tmp, sizeof(double), hipMemcpyDeviceToHost));
    return SPIMC_OK;
}

int spimc_d_action_d_beta(spimc_ctx* c, uint32_t r, int action, double* out) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!out) return fail(SPIMC_ERR_ARG, "null output");
    Rep* base = c->d_reps + r;
    const Rep& R = c->reps[r].dev;
    if (action == SPIMC_ACTION_KINETIC) {
        if ((rc = est_kin_dbeta(c, base, 1)) || (rc = est_finalize(c, base, 1, 0, 0))) return rc;
    } else {
        if ((rc = est_pair_whole(c, base, 1, 0)) || (rc = est_finalize(c, base, 1, 1, 0))) return rc;
    }
    if ((rc = sync_and_resolve(c))) return rc;
    double res[2];
    HIPCHK(hipMemcpy(res, R.scratch + scratch_res(R.M, R.nk), sizeof res, hipMemcpyDeviceToHost));
    *out = action == SPIMC_ACTION_KINETIC ? res[0] : (R.pair_type == SPIMC_PAIR_NONE ? 0. : res[1]);
    return SPIMC_OK;
}

int spimc_potential(spimc_ctx* c, uint32_t r, int action, double* out) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!out) return fail(SPIMC_ERR_ARG, "null output");
    *out = 0.;
    const Rep& R = c->reps[r].dev;
    if (action == SPIMC_ACTION_KINETIC || R.pair_type == SPIMC_PAIR_NONE) return SPIMC_OK;
    Rep* base = c->d_reps + r;
    if ((rc = est_pair_whole(c, base, 1, 1)) || (rc = est_finalize(c, base, 1, 1, 0))) return rc;
    if ((rc = sync_and_resolve(c))) return rc;
    double res[2];
    HIPCHK(hipMemcpy(res, R.scratch + scratch_res(R.M, R.nk), sizeof res, hipMemcpyDeviceToHost));
    *out = res[1];
    return SPIMC_OK;
}

int spimc_kinetic_virial_energy(spimc_ctx* c, uint32_t r, uint32_t window, double* out) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!out) return fail(SPIMC_ERR_ARG, "null output");
    const Rep& R = c->reps[r].dev;
    const int W = window ? (int)window : 0;
    rc = launch(c, SPIMC_KERNEL_ESTIMATORS, k_kin_virial, 1, 256, 0, c->stream, c->d_reps + r, W, kSlotVirial);
    if (rc || (rc = sync_and_resolve(c))) return rc;
    HIPCHK(hipMemcpy(out, R.scratch + kSlotVirial * R.M, sizeof(double), hipMemcpyDeviceToHost));
    return SPIMC_OK;
}

int spimc_action_gradient(spimc_ctx* c, uint32_t r, uint32_t terms, int compat, double* out) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!out) return fail(SPIMC_ERR_ARG, "null output");
    if (terms & ~SPIMC_GRAD_ALL) return fail(SPIMC_ERR_ARG, "unknown gradient terms");
    if (compat && (terms & SPIMC_GRAD_KINETIC))
        return fail(SPIMC_ERR_ARG, "compat gradient covers the pair action only");
    const Rep& R = c->reps[r].dev;
    Rep* base = c->d_reps + r;
    dim3 g((R.N + 63) / 64, R.M, 1);
    // Long range.  Adjoint: d/dr of sum_k coef_s |rho_k|^2 per slice.  Compat:
    // IlkkaPairAction::CalcGradientULong (ilkka_pair_action_class.h:231-249),
    // sum_k u_long_k k Im(conj(rho_k) e^{i k.r}) -- half the adjoint term -- added
    // once per particle pair (pair_action_class.h:331-334), i.e. (N - 1) times;
    // zero for the Bare and David actions (pair_action_class.h:158-165).
    const bool long_on = (terms & SPIMC_GRAD_PAIR_LONG) && R.use_long_range &&
                         (!compat || R.pair_type == SPIMC_PAIR_ILKKA);
    if (compat) {
        rc = launch(c, SPIMC_KERNEL_ADJOINT, k_grad_compat, g, 64, 0, c->stream, base, 0,
                    (terms & SPIMC_GRAD_PAIR_SHORT) ? 1 : 0);
    } else {
        rc = launch(c, SPIMC_KERNEL_ADJOINT, k_adjoint_local, g, 64, 0, c->stream, base, terms, 0);
    }
    if (!rc && long_on && c->adj_runs_lds > 0) {
        rc = launch(c, SPIMC_KERNEL_ADJOINT, k_adjoint_long_runs, dim3(R.M, 1), kAdjRunBlock, c->adj_runs_lds,
                    c->stream, base, 0, compat);
    } else if (!rc && long_on) {
        if (R.max_idx > kAdjMaxIdx) return fail(SPIMC_ERR_UNSUPPORTED, "adjoint long range supports |n| <= 40");
        const size_t shmem = (size_t)64 * 3 * (R.max_idx + 1) * sizeof(double2) +
                             kAdjTile * (sizeof(double2) + sizeof(int) + sizeof(double));
        rc = launch(c, SPIMC_KERNEL_ADJOINT, k_adjoint_long, g, 64, shmem, c->stream, base, 0, compat);
    }
    if (rc || (rc = sync_and_resolve(c))) return rc;
    HIPCHK(hipMemcpy(out, R.grad, (size_t)3 * R.M * R.N * sizeof(double), hipMemcpyDeviceToHost));
    if (!compat && terms == SPIMC_GRAD_ALL) c->adjoint_fresh = true;
    return SPIMC_OK;
}

// Bisect::Reset with adaptive="1" (bisect_class.h:143-151): the replica's
// bisection level for the following stages; the work lists are rebuilt and the
// captured graphs dropped.
int spimc_set_bisect_level(spimc_ctx* c, uint32_t r, uint32_t n_level) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    RepHost& h = c->reps[r];
    const int max_lvl = std::min((int)std::floor(std::log2((double)h.dev.M)), kMaxLevel);
    if (n_level < 1 || (int)n_level > max_lvl)
        return fail(SPIMC_ERR_UNSUPPORTED, "bisect n_level must be in [1, min(6, floor(log2 n_bead))]");
    if ((int)n_level == h.dev.n_level) return SPIMC_OK;
    HIPCHK(hipDeviceSynchronize());
    h.dev.n_level = (int)n_level;
    h.d.bisect_n_level = n_level;
    for (auto& G : c->groups) {
        if ((int)r < G.r0 || (int)r >= G.r0 + G.n) continue;
        for (void** p : {(void**)&G.d_pair_bis, (void**)&G.d_pair_dis, (void**)&G.d_ritem_bis, (void**)&G.d_ritem_dis,
                         (void**)&G.d_rgrp[0], (void**)&G.d_rgrp[1]})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
        G.max_bis_links = 0;
        for (int q = G.r0; q < G.r0 + G.n; ++q)
            G.max_bis_links = std::max(G.max_bis_links, 1 << c->reps[q].dev.n_level);
        if ((rc = build_work_lists(c, G))) return rc;
    }
    drop_graphs(c);
    return upload_reps(c);
}

// ---- Move interface ----------------------------------------------------------

int spimc_tape_set(spimc_ctx* c, uint32_t r, const double* draws, uint64_t n) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    RepHost& h = c->reps[r];
    HIPCHK(hipStreamSynchronize(c->stream));
    if (h.tape) (void)hipFree(h.tape);
    h.tape = nullptr;
    h.dev.tape = nullptr;
    h.dev.tape_len = 0;
    if (draws && n) {
        HIPCHK(hipMalloc((void**)&h.tape, n * sizeof(double)));
        HIPCHK(hipMemcpy(h.tape, draws, n * sizeof(double), hipMemcpyHostToDevice));
        h.dev.tape = h.tape;
        h.dev.tape_len = (long long)n;
    }
    long long z = 0;
    HIPCHK(hipMemcpy(h.dev.tape_cursor, &z, sizeof z, hipMemcpyHostToDevice));
    // graphs captured earlier read the Rep table through its device pointer, so
    // they see the new tape without re-capture
    return upload_reps(c);
}

int spimc_tape_cursor(spimc_ctx* c, uint32_t r, uint64_t* cursor) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    long long v = 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&v, c->reps[r].dev.tape_cursor, sizeof v, hipMemcpyDeviceToHost));
    *cursor = (uint64_t)v;
    return SPIMC_OK;
}

int spimc_bisect_stage(spimc_ctx* c, const uint32_t* n_attempts, uint32_t n_uniform, uint8_t* accepted) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    return run_stage(c, 0, n_attempts, n_uniform, accepted, false);
}

int spimc_displace_stage(spimc_ctx* c, const uint32_t* n_attempts, uint32_t n_uniform, uint8_t* accepted) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    return run_stage(c, 1, n_attempts, n_uniform, accepted, false);
}

int spimc_perm_configure(spimc_ctx* c, uint32_t r, uint32_t n_level, double epsilon) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    RepHost& h = c->reps[r];
    const int max_lvl = std::min((int)std::floor(std::log2((double)h.dev.M)), kMaxLevel);
    if (n_level < 1 || (int)n_level > max_lvl)
        return fail(SPIMC_ERR_UNSUPPORTED, "permuting bisection n_level must be in [1, min(6, floor(log2 n_bead))]");
    if (!(epsilon > 0.)) return fail(SPIMC_ERR_ARG, "epsilon must be > 0");
    HIPCHK(hipDeviceSynchronize());
    h.dev.perm_n_level = (int)n_level;
    h.dev.perm_log_eps = std::log(epsilon);
    h.dev.perm_eps = epsilon;
    return upload_reps(c);
}

int spimc_perm_bisect_stage(spimc_ctx* c, const uint32_t* n_attempts, uint32_t n_uniform, uint8_t* accepted) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    int rc = ensure_perm(c, true);
    if (rc) return rc;
    c->perm_table = false;
    return run_stage(c, 3, n_attempts, n_uniform, accepted, false);
}

int spimc_perm_table_configure(spimc_ctx* c, uint32_t r, uint32_t n_perm_part, uint32_t balance) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    RepHost& h = c->reps[r];
    const int N = h.dev.N;
    if (n_perm_part < 1 || (int)n_perm_part > std::min(N, kTabMaxK))
        return fail(SPIMC_ERR_UNSUPPORTED, "n_perm_part must be in [1, min(n_part, 6)]");
    if (balance > 1) return fail(SPIMC_ERR_ARG, "balance must be 0 or 1");
    if (N > kTabMaxN) return fail(SPIMC_ERR_UNSUPPORTED, "the permutation table supports at most 1024 particles");
    // the cycle list: k! types x C(N, k) subsets
    double entries = 1.;
    for (uint32_t i = 0; i < n_perm_part; ++i) entries *= (double)(N - (int)i);
    if (entries > (double)kTabMaxEntries)
        return fail(SPIMC_ERR_UNSUPPORTED, "the permutation table's cycle list exceeds 2^28 entries");
    HIPCHK(hipDeviceSynchronize());
    if (!h.perm_t) HIPCHK(hipMalloc((void**)&h.perm_t, 2 * (size_t)N * N * sizeof(double)));
    h.dev.perm_t = h.perm_t;
    h.dev.tab_k = (int)n_perm_part;
    h.dev.tab_balance = (int)balance;
    return upload_reps(c);
}

int spimc_perm_table_stage(spimc_ctx* c, const uint32_t* n_attempts, uint32_t n_uniform, uint8_t* accepted) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    for (size_t r = 0; r < c->reps.size(); ++r)
        if ((n_attempts ? n_attempts[r] : n_uniform) > 0 && c->reps[r].dev.tab_k == 0)
            return fail(SPIMC_ERR_STATE, "spimc_perm_table_configure was not called for a replica with attempts");
    int rc = ensure_perm(c, true);
    if (rc) return rc;
    c->perm_table = true;
    rc = run_stage(c, 3, n_attempts, n_uniform, accepted, false);
    c->perm_table = false;
    return rc;
}

int spimc_perm_download(spimc_ctx* c, uint32_t r, uint32_t* sigma) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!sigma) return fail(SPIMC_ERR_ARG, "null argument");
    const Rep& R = c->reps[r].dev;
    HIPCHK(hipDeviceSynchronize());
    if (!R.perm) {
        for (int p = 0; p < R.N; ++p) sigma[p] = (uint32_t)p;
        return SPIMC_OK;
    }
    std::vector<int> v(R.N);
    HIPCHK(hipMemcpy(v.data(), R.perm, R.N * sizeof(int), hipMemcpyDeviceToHost));
    for (int p = 0; p < R.N; ++p) sigma[p] = (uint32_t)v[p];
    return SPIMC_OK;
}

int spimc_perm_upload(spimc_ctx* c, uint32_t r, const uint32_t* sigma) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!sigma) return fail(SPIMC_ERR_ARG, "null argument");
    const int N = c->reps[r].dev.N;
    std::vector<int> v(2 * (size_t)N, -1);
    bool ident = true;
    for (int p = 0; p < N; ++p) {
        if (sigma[p] >= (uint32_t)N || v[N + sigma[p]] >= 0) return fail(SPIMC_ERR_ARG, "not a permutation");
        v[p] = (int)sigma[p];
        v[N + sigma[p]] = p;
        ident = ident && sigma[p] == (uint32_t)p;
    }
    if (ident && !c->reps[r].dev.perm) return SPIMC_OK;
    if ((rc = ensure_perm(c, false))) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(c->reps[r].dev.perm, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
    c->adjoint_fresh = false;
    return SPIMC_OK;
}

int spimc_perm_counters(spimc_ctx* c, uint32_t r, uint64_t* out) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!out) return fail(SPIMC_ERR_ARG, "null argument");
    const Rep& R = c->reps[r].dev;
    const size_t n = 2 + 2 * (size_t)R.N;
    HIPCHK(hipDeviceSynchronize());
    if (!R.perm_cnt) {
        for (size_t i = 0; i < n; ++i) out[i] = 0;
        return SPIMC_OK;
    }
    std::vector<unsigned long long> v(n);
    HIPCHK(hipMemcpy(v.data(), R.perm_cnt, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) out[i] = v[i];
    return SPIMC_OK;
}

int spimc_move_counters(spimc_ctx* c, uint32_t r, uint64_t* out) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    Counters k;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&k, c->reps[r].dev.cnt, sizeof k, hipMemcpyDeviceToHost));
    out[0] = k.bis_att;
    out[1] = k.bis_acc;
    out[2] = k.dis_att;
    out[3] = k.dis_acc;
    return SPIMC_OK;
}

int spimc_action_trace_enable(spimc_ctx* c, int enable) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    if (c->prop_set) return fail(SPIMC_ERR_STATE, "an interface proposal is pending; accept or reject it first");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->trace_on = enable != 0;
    uint32_t cap = 0;
    for (auto& h : c->reps) cap = std::max(cap, h.accepted_cap);
    int rc = ensure_accepted(c, std::max<uint32_t>(cap, 1));
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    return SPIMC_OK;
}

int spimc_action_trace_read(spimc_ctx* c, uint32_t r, uint32_t n_attempts, double* out) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!out && n_attempts) return fail(SPIMC_ERR_ARG, "null output");
    const RepHost& h = c->reps[r];
    if (!c->trace_on || !h.trace) return fail(SPIMC_ERR_STATE, "the action trace is off");
    if (n_attempts > h.stage_attempts) return fail(SPIMC_ERR_ARG, "more attempts than the last stage made");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(out, h.trace, (size_t)n_attempts * kTraceWords * sizeof(double), hipMemcpyDeviceToHost));
    return SPIMC_OK;
}

// ---- Observable interface ------------------------------------------------------

int spimc_obs_configure(spimc_ctx* c, int energy_estimator, uint32_t virial_window, double sk_k_cut, double gr_r_min,
                        double gr_r_max, uint32_t gr_n_r) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    if (energy_estimator < SPIMC_ENERGY_THERMAL || energy_estimator > SPIMC_ENERGY_VIRIAL_ADJOINT)
        return fail(SPIMC_ERR_ARG, "unknown energy estimator");
    if (gr_n_r < 2 || gr_n_r > (uint32_t)kGrBins) return fail(SPIMC_ERR_ARG, "n_r must be in [2, 4096]");
    for (auto& h : c->reps)
        if (virial_window > (uint32_t)h.dev.M) return fail(SPIMC_ERR_ARG, "virial window larger than n_bead");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->energy_estimator = energy_estimator;
    c->virial_window = virial_window;
    c->sk_k_cut = sk_k_cut;
    c->gr_r_min = gr_r_min;
    c->gr_r_max = gr_r_max;
    c->gr_n_r = gr_n_r;
    std::vector<double> rmax;
    for (auto& h : c->reps) {
        rmax.push_back(gr_r_max > 0. ? gr_r_max : h.dev.L / 2.);
        if (h.gr_cnt) (void)hipFree(h.gr_cnt);
        h.gr_cnt = nullptr;
        HIPCHK(hipMalloc((void**)&h.gr_cnt, gr_n_r * sizeof(unsigned long long)));
        h.dev.gr_cnt = h.gr_cnt;
    }
    HIPCHK(hipMemcpy(c->d_gr_rmax, rmax.data(), rmax.size() * sizeof(double), hipMemcpyHostToDevice));
    int rc = upload_reps(c);
    if (rc) return rc;
    return spimc_obs_reset(c, SPIMC_OBS_ENERGY | SPIMC_OBS_SK | SPIMC_OBS_GR);
}

int spimc_obs_reset(spimc_ctx* c, uint32_t which) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& h : c->reps) {
        const Rep& R = h.dev;
        if (which & SPIMC_OBS_ENERGY) {
            HIPCHK(hipMemset(R.e_acc, 0, 2 * sizeof(double)));
            HIPCHK(hipMemset(R.n_measure, 0, sizeof(unsigned long long)));
        }
        if (which & SPIMC_OBS_SK) {
            HIPCHK(hipMemset(R.sk_acc, 0, std::max(R.nk, 1) * sizeof(double)));
            HIPCHK(hipMemset(R.n_measure + 1, 0, sizeof(unsigned long long)));
        }
        if (which & SPIMC_OBS_GR) {
            HIPCHK(hipMemset(R.gr_cnt, 0, c->gr_n_r * sizeof(unsigned long long)));
            HIPCHK(hipMemset(R.n_measure + 2, 0, sizeof(unsigned long long)));
        }
    }
    return SPIMC_OK;
}

int spimc_obs_accumulate(spimc_ctx* c, uint32_t which) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    if (c->prop_set) return fail(SPIMC_ERR_STATE, "an interface proposal is pending");
    int rc = obs_accumulate_all(c, which);
    if (rc) return rc;
    return sync_and_resolve(c);
}

int spimc_obs_read(spimc_ctx* c, uint32_t r, uint32_t which, double* out, uint64_t* n_measure) {
    int rc = check_rep(c, r);
    if (rc) return rc;
    if (!out || !n_measure) return fail(SPIMC_ERR_ARG, "null output");
    const Rep& R = c->reps[r].dev;
    HIPCHK(hipStreamSynchronize(c->stream));
    unsigned long long nm[3];
    HIPCHK(hipMemcpy(nm, R.n_measure, sizeof nm, hipMemcpyDeviceToHost));
    if (which == SPIMC_OBS_ENERGY) {
        HIPCHK(hipMemcpy(out, R.e_acc, 2 * sizeof(double), hipMemcpyDeviceToHost));
        *n_measure = nm[0];
    } else if (which == SPIMC_OBS_SK) {
        if (R.nk) HIPCHK(hipMemcpy(out, R.sk_acc, R.nk * sizeof(double), hipMemcpyDeviceToHost));
        *n_measure = nm[1];
    } else if (which == SPIMC_OBS_GR) {
        std::vector<unsigned long long> v(c->gr_n_r);
        HIPCHK(hipMemcpy(v.data(), R.gr_cnt, v.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < v.size(); ++i) out[i] = (double)v[i];
        *n_measure = nm[2];
    } else {
        return fail(SPIMC_ERR_ARG, "read one observable at a time");
    }
    return SPIMC_OK;
}

// ---- batched sweep ------------------------------------------------------------

int spimc_sweep(spimc_ctx* c, uint32_t n_sweeps, uint32_t flags) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    const int R = (int)c->reps.size();
    std::vector<uint32_t> nb(R), nd(R);
    for (int r = 0; r < R; ++r) {
        const Rep& x = c->reps[r].dev;
        nb[r] = (uint32_t)(x.N * x.M) >> x.n_level;
        nd[r] = (uint32_t)x.N;
    }
    const bool graph = (flags & SPIMC_SWEEP_GRAPH) != 0;
    for (uint32_t s = 0; s < n_sweeps; ++s) {
        int rc = run_stage(c, 0, nb.data(), 0, nullptr, graph);
        if (rc) return rc;
        if ((rc = run_stage(c, 1, nd.data(), 0, nullptr, graph))) return rc;
        if (flags & SPIMC_SWEEP_ADJOINT) {
            if ((rc = est_gradient(c, c->d_reps, R, SPIMC_GRAD_ALL, 0, 0))) return rc;
            c->adjoint_fresh = true;
        }
        uint32_t which = 0;
        if (flags & SPIMC_SWEEP_ENERGY) which |= SPIMC_OBS_ENERGY;
        if (flags & SPIMC_SWEEP_SK) which |= SPIMC_OBS_SK;
        if (flags & SPIMC_SWEEP_GR) which |= SPIMC_OBS_GR;
        if (which && (rc = obs_accumulate_all(c, which))) return rc;
        if ((rc = sync_and_resolve(c))) return rc;
        c->adjoint_fresh = (flags & SPIMC_SWEEP_ADJOINT) != 0;
    }
    return SPIMC_OK;
}

int spimc_autotune(spimc_ctx* c, uint32_t n_rounds, int32_t* chosen) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    if (n_rounds == 0) return fail(SPIMC_ERR_ARG, "n_rounds must be positive");
    if (c->prop_set) return fail(SPIMC_ERR_STATE, "an interface proposal is pending; accept or reject it first");
    struct Shape {
        int fork, bpc, bs, pers, one;
    };
    // launch shapes tried: rho_k beside the pair kernel with 1-3 looping workgroups
    // per CU and 256/512-thread pair workgroups, or the two kernels one after the
    // other; for small replicas also the persistent stage kernel with 256-1024 threads
    std::vector<Shape> shapes = {{1, 2, 512, 0, 0}, {1, 3, 512, 0, 0}, {1, 1, 512, 0, 0}, {1, 2, 256, 0, 0},
                                 {1, 3, 256, 0, 0}, {0, 0, 1024, 0, 0}, {0, 2, 512, 0, 0}};
    if (!c->any_long || c->lds_doubles_max == 0)
        shapes = {{0, 0, 1024, 0, 0}, {0, 0, 512, 0, 0}, {0, 0, 256, 0, 0}};
    if (pers_eligible(c))
        for (int pb : {256, 512, 1024}) shapes.push_back({c->fork ? 1 : 0, c->rho_bpc, c->pair_block, pb, 0});
    // batches of at most one replica per CU: also one-launch bisection rounds,
    // with the displace rounds' rho_k after or beside their pair kernel
    if (one_launch_eligible(c) && (int)c->reps.size() <= kNumCUs) {
        shapes.push_back({0, 0, c->pair_block, 0, 1});
        if (c->any_long) shapes.push_back({1, 2, 512, 0, 1});
        shapes.push_back({0, 0, c->pair_block, 0, 2});
    }
    const Shape saved{c->fork ? 1 : 0, c->rho_bpc, c->pair_block, c->pers_block, c->one_launch ? c->one_moves : 0};
    // the sweep's mix: N M / 2^l bisection attempts per N displace attempts
    const Rep& R0 = c->reps[0].dev;
    const uint32_t n_dis = std::max<uint32_t>(1, (uint32_t)((uint64_t)n_rounds * (1u << R0.n_level) / R0.M));
    double best_t = 1e300;
    Shape best = saved;
    int tried = 0;
    for (const Shape& sh : shapes) {
        c->fork = sh.fork != 0;
        c->rho_bpc = sh.bpc;
        c->one_launch = sh.one != 0;
        c->one_moves = sh.one >= 2 ? 2 : 1;
        int rc = c->lds_doubles_max > 0 ? set_pair_block(c, sh.bs) : SPIMC_OK;
        if (rc == SPIMC_OK) rc = set_pers_block(c, sh.pers);
        if (rc) continue;
        drop_graphs(c);
        // one untimed round of each move under the new shape, then the timed mix
        if ((rc = run_stage(c, 0, nullptr, 1, nullptr, false)) == SPIMC_OK) rc = run_stage(c, 1, nullptr, 1, nullptr, false);
        const auto t0 = std::chrono::steady_clock::now();
        if (rc == SPIMC_OK) rc = run_stage(c, 0, nullptr, n_rounds, nullptr, false);  // returns after the device work
        if (rc == SPIMC_OK) rc = run_stage(c, 1, nullptr, n_dis, nullptr, false);
        const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc) {
            c->fork = saved.fork != 0;
            c->rho_bpc = saved.bpc;
            c->one_launch = saved.one != 0;
            c->one_moves = saved.one >= 2 ? 2 : 1;
            if (c->lds_doubles_max > 0) (void)set_pair_block(c, saved.bs);
            (void)set_pers_block(c, saved.pers);
            return rc;
        }
        ++tried;
        if (t < best_t) {
            best_t = t;
            best = sh;
        }
    }
    c->fork = best.fork != 0;
    c->rho_bpc = best.bpc;
    c->one_launch = best.one != 0;
    c->one_moves = best.one >= 2 ? 2 : 1;
    if (c->lds_doubles_max > 0) {
        int rc = set_pair_block(c, best.bs);
        if (rc) return rc;
    }
    if (int rc = set_pers_block(c, best.pers)) return rc;
    drop_graphs(c);
    if (chosen) {
        chosen[0] = best.fork;
        chosen[1] = best.bpc;
        chosen[2] = c->pair_block;
        chosen[3] = tried;
        chosen[4] = (int32_t)n_dis;
        chosen[5] = c->pers_block;
        chosen[6] = c->one_launch ? c->one_moves : 0;
    }
    return SPIMC_OK;
}

int spimc_bead_updates(spimc_ctx* c, uint64_t* out) {
    if (!c || !out) return fail(SPIMC_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    uint64_t t = 0;
    for (auto& h : c->reps) {
        Counters k;
        HIPCHK(hipMemcpy(&k, h.dev.cnt, sizeof k, hipMemcpyDeviceToHost));
        t += k.bead_updates;
    }
    *out = t;
    return SPIMC_OK;
}

// ---- measurement ---------------------------------------------------------------

int spimc_timing_enable(spimc_ctx* c, int mode) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    int rc = sync_and_resolve(c);
    if (rc) return rc;
    if (mode & ~(SPIMC_TIMING_CLOCK | SPIMC_TIMING_EVENTS)) return fail(SPIMC_ERR_ARG, "unknown timing mode");
    c->timing = (mode & SPIMC_TIMING_EVENTS) != 0;
    c->clock = (mode & SPIMC_TIMING_CLOCK) != 0;
    rc = work_totals(c, c->work_seen);
    if (rc) return rc;
    return SPIMC_OK;
}

int spimc_kernel_timing(spimc_ctx* c, int kernel, double* total_ms, uint64_t* launches, double* bytes) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    if (kernel < 0 || kernel >= SPIMC_KERNEL_COUNT) return fail(SPIMC_ERR_ARG, "unknown kernel class");
    int rc = sync_and_resolve(c);
    if (rc) return rc;
    const int kc = kernel == SPIMC_KERNEL_RHO ? KT_RHO : kernel == SPIMC_KERNEL_PAIR ? KT_PAIR : -1;
    if (c->clock && kc >= 0) {
        if (total_ms) *total_ms = 1e3 * (double)c->kt_sum[kc] / c->clock_hz;
        if (launches) *launches = c->kt_launches[kc];
    } else {
        if (total_ms) *total_ms = c->t_ms[kernel];
        if (launches) *launches = c->t_n[kernel];
    }
    if (bytes) *bytes = c->t_bytes[kernel];
    return SPIMC_OK;
}

int spimc_timing_reset(spimc_ctx* c) {
    if (!c) return fail(SPIMC_ERR_ARG, "null context");
    int rc = sync_and_resolve(c);
    if (rc) return rc;
    for (int i = 0; i < SPIMC_KERNEL_COUNT; ++i) {
        c->t_ms[i] = 0.;
        c->t_n[i] = 0;
        c->t_bytes[i] = 0.;
    }
    for (auto& G : c->groups) HIPCHK(hipMemset(G.d_kt, 0, KT_WORDS * sizeof(unsigned long long)));
    for (int k = 0; k < KT_CLASSES; ++k) c->kt_sum[k] = c->kt_launches[k] = 0;
    return SPIMC_OK;
}

}  // extern "C"
