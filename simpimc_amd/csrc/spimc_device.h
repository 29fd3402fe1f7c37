// spimc_device.h -- device data layout and per-element math of the engine.
//
// Layout in HBM (one replica):
//   pos   : FP64 [3][M][N]      committed bead positions, slice-major SoA
//   rho   : double2 [2][M][nk]  two buffers of per-slice rho_k; sel[b] names the
//                               committed one, the other receives proposals
//   sel   : uint8 [M]
//   tables: splines in per-interval polynomial form, components interleaved
//           so that one interval of all components is contiguous.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spimc {

constexpr int kMaxLevel = 6;      // bisection levels supported (window of 64 links)
constexpr int kWave = 64;         // CDNA wavefront
constexpr int kMaxKinSplines = 8; // kinetic / bisect image splines at tau * 2^l
// action trace record of one attempt: kinetic OLD, NEW; short-range pair OLD,
// NEW; long-range pair OLD, NEW; log sample ratio; previous level's change
constexpr int kTraceWords = 8;

// Natural cubic spline: s(t) = c0 + dt (c1 + dt (c2 + dt c3)), dt = t - x[i].
struct Spl {
    const double* x;  // [n] knots
    const double* c;  // [n-1][ncomp][4]
    int n;
    int ncomp;
    int uniform;
    int pad;
    double x0;
    double inv_h;
};

// Tensor-product natural cubic spline on general grids x[nx] x y[ny]: cell
// (i, j) holds s(x, y) = sum_{a,b} c[a][b] dx^a dy^b, dx = x - x[i], dy = y - y[j],
// its 16 coefficients contiguous at c[(j * (nx - 1) + i) * 16 + 4 a + b].
struct Spl2 {
    const double* x;
    const double* y;
    const double* c;
    int nx, ny;
};

// Uniform cubic B-spline (the natural cubic spline of a uniform grid in B-spline
// form): coefficient j of component c at c[c * ncoef + j], ncoef = knots + 2;
// s(t) = sum_{j=0..3} c[i + j] B_j(f) with u = (t - x0) / h, i = clamp(floor u), f = u - i.
struct BTab {
    const double* c;  // global copy; kernels stage it through LDS
    int knots;
    int ncoef;
    int ncomp;
    int pad;
    double x0;
    double ih;
};

// Per-attempt proposal of one replica.
struct Prop {
    int status;   // 0 idle, 1 pending level-0 decision, 2 rejected at a coarse level
    int move;     // 0 bisect, 1 displace, 2 interface
    int p;        // moved particle
    int b0;       // first slice of the link window
    int n_links;  // links (b, b+1) in the window
    int first;    // first moved slice
    int count;    // moved slices
    int stage_i;  // attempt index within the current stage (for the accepted flags)
    double lsr;   // log sample ratio at level 0
    double k_old; // kinetic action at level 0, OLD / NEW
    double k_new;
    double prev;  // action change of the previous (coarser) level
    double u0;    // level-0 uniform
    int p2;       // label of the moved chain after the wrap (bisect; else p)
    int generic;  // 1: the label-general pair kernel evaluates this proposal
    long long tape_next;  // tape cursor after this attempt
    int n_chain;  // chains of a permuting bisection (0: the cycle walk stopped)
    int e_slice;  // slice after which a permuting bisection relinks the chains
    int n_level;  // bisection level of a permuting bisection
    int rho_multi;  // 1: k_rho_multi updates rho_k (several beads per slice); 0: the k_rho items
    // PermBisectTable's final table check, after a level-0 accept (tab_on): reject
    // if tab_ratio < u, u the next tape value or u_tab (device RNG)
    int tab_on;
    double tab_ratio;
    double u_tab;
};

struct Counters {
    unsigned long long bis_att, bis_acc, dis_att, dis_acc;
    unsigned long long seq;           // attempts issued (device RNG counter)
    unsigned long long bead_updates;
    unsigned int remaining;           // attempts left in the running stage
    unsigned int stage_i;             // attempts done in the running stage
    unsigned long long error;         // replay tape overrun or invalid draw
    unsigned long long pair_links;    // link evaluations of the pair kernel (moves)
    unsigned long long rho_slicek;    // (moved slice, k) updates of the rho_k kernel
};

// One rho_k work item of a move round: a (moved bead, k chunk) of one replica,
// described in full so that the rho_k kernel reads one 64-byte record per item.
struct __align__(64) RhoItem {
    const double2* rc;    // committed rho_k row of the bead's slice
    double2* rn;          // spare row: receives the proposal's rho_k
    const double2* ph;    // the bead's six phase chains (OLD x3, NEW x3)
    const int* kidx;      // packed k-vectors
    const double* coef;   // long-range coefficients per k
    double2* part;        // (OLD, NEW) partial sums of this item
    int k0, k1;           // k range
    int stride, valid;    // phase-chain stride; 0 = idle item
};

struct Rep {
    // sizes and flags
    int N, M, nk, max_idx;
    int n_level, pair_type, use_long_range, is_coulomb;
    int n_order, n_val, kin_images, bis_images;
    int n_links_max, n_chunks, index, pad;
    // scalars
    double L, iL, tau, lambda;
    double step, box;
    double kin_i4lt, kin_i4ltt, kin_dbeta_const;
    double long_scale;  // per-slice long-range action = long_scale * sum_k coef_s |rho|^2
    double const_v, const_du;
    double r_v_min, r_v_max, r_l_min, r_l_max, g_min, g_max;
    uint64_t key;       // device RNG key
    // bead storage
    double* pos;
    double2* rho;
    unsigned char* sel;
    // k-space
    const int* kidx;       // packed (n0+128) | (n1+128) << 8 | (n2+128) << 16
    const double* kmag;    // |k|
    const int4* k_runs;    // [n_runs] runs of consecutive n2 at fixed (n0, n1): {n0, n1, n2 first, +-length}
    const int* k_order;    // [nk] k index of the i-th k-vector in run order
    int n_runs, rho_kpt;   // rho_kpt: k-vectors per thread of the rho_k kernel (sets the k chunks)
    RhoItem* ritems_bis;   // this replica's rho_k work items of a bisection round
    RhoItem* ritems_dis;   // ... and of a displace round
    int n_ritems_bis, n_ritems_dis;
    const double* coef_s;  // per-slice action coefficient per k
    const double* coef_b;  // per-slice beta-derivative coefficient per k
    const double* coef_v;  // per-slice potential coefficient per k
    // pair tables
    Spl v_r, v_long_r;     // Bare
    Spl u, du;             // David, ncomp = n_val + 1
    Spl2 u2, du2;          // Ilkka U(x, y), dU/dbeta(x, y); v_r and v_long_r hold its potential
    Spl u_long_r, du_long_r;  // Ilkka long-range parts of U and dU/dbeta
    double r_u_min, r_u_max, r_du_min, r_du_max;
    // LDS-staged B-spline form of the link action tables (uniform grids):
    // Bare: bt_a = v_r, bt_b = v_long_r; David: bt_a = u components 1..n_val.
    BTab bt_a, bt_b;
    BTab bt_du;            // David du_kj_dbeta components 0..n_val (estimators)
    int lds_doubles;       // doubles of LDS the pair kernel stages (0: tables stay in global memory)
    int lds_doubles_du;    // doubles of LDS the beta-derivative estimator stages (0: global)
    int pad4;
    // kinetic and sampling image splines (n_images > 0 only)
    Spl kin[kMaxKinSplines];
    Spl kin_dtau;
    Spl bis[kMaxKinSplines];
    // per-attempt state
    double2* phase;        // [M][2][3][max_idx+1] e^{i m phi} of the moved beads, OLD then NEW
    int2* rho_rows;        // [M] (slice, committed buffer) of each moved bead, written with the proposal
    Prop* prop;
    double* prop_new;      // [M][3] NEW positions of the moved slices
    double2* pair_part;    // [n_links_max] (old, new)
    double2* rho_part;     // [M][n_chunks] (old, new)
    double2* kin_part;     // interface kinetic (old, new)
    Counters* cnt;
    unsigned int* ctx_error;  // shared by all replicas of a context
    unsigned char* accepted; // [accepted_cap]
    unsigned int accepted_cap;
    unsigned int pad3;
    // action trace of the move stages (nullptr: off): per attempt of the running
    // stage, kTraceWords doubles of what the level-0 Metropolis decision adds up
    double* act_trace;       // [accepted_cap][kTraceWords]
    // replay tape
    const double* tape;
    long long tape_len;
    long long* tape_cursor;
    // estimators
    double* e_acc;          // [2]
    double* sk_acc;         // [nk]
    unsigned long long* gr_cnt;  // [gr_n_r]
    double* scratch;        // [M * 8 + nk] reduction partials
    double* grad;           // [M][N][3] dS/dr of the last adjoint
    double* grad_est;       // [M][N][3] estimator scratch gradient
    unsigned long long* n_measure;  // [3]
    // path topology: perm[p] is the particle whose slice-0 bead follows the
    // slice-(M-1) bead of particle p, iperm its inverse (nullptr: the identity)
    int* perm;
    int* iperm;
    // permuting bisection (nullptr until the first permuting stage)
    int* ch_lbl;         // [N] chain labels before the wrap, in cycle order
    int* ch_lbl2;        // [N] chain labels after the wrap
    int* ch_nxt;         // [N] NEW mode: chain i ends at the old end of chain ch_nxt[i]
    int* ch_lead;        // [N] 1: chain i is the first of its cycle under ch_nxt
    int* mv_pre;         // [N] chain of a label before the wrap, -1 if none
    int* mv_post;        // [N] chain of a label after the wrap, -1 if none
    double* ch_new;      // [N][M - 1][3] NEW positions of the chains' moved beads
    unsigned long long* perm_cnt;  // [2 + 2N] attempts, accepts, perm_attempt[N], perm_accept[N]
    double perm_log_eps;            // log of the table cutoff epsilon
    double perm_eps;                // the cutoff epsilon
    double* perm_t;                 // [2][N][N] PermBisectTable's table (old path, new path)
    int tab_k;                      // PermBisectTable's n_perm_part (0: not configured)
    int tab_balance;                // its final check: 0 the reference's, 1 detailed balance
    int perm_n_level;               // bisection level of the permuting move
    int multi_cap;                  // replicas of the group: the second list's offset
    // per group: the proposals of a permuting round that the label-general pair
    // kernel (cnt[0], list[0..]) and k_rho_multi (cnt[1], list[multi_cap..]) take,
    // as replica indices within the group
    unsigned int* multi_cnt;
    int* multi_list;
};

// Path topology (Bead::next_ / prev_ across the wrap, which is where
// PermBisect::AssignParticleLabels leaves every permutation,
// perm_bisect_class.h:47-56); within slices 0..M-1 a bead is followed by the
// same label.
__device__ __forceinline__ int label_after_wrap(const Rep& R, int p) { return R.perm ? R.perm[p] : p; }
__device__ __forceinline__ int label_before_wrap(const Rep& R, int p) { return R.iperm ? R.iperm[p] : p; }
// label of the bead after (p, s) and before (p, s)
__device__ __forceinline__ int next_label(const Rep& R, int s, int p) { return s == R.M - 1 ? label_after_wrap(R, p) : p; }
__device__ __forceinline__ int prev_label(const Rep& R, int s, int p) { return s == 0 ? label_before_wrap(R, p) : p; }

// ---------------------------------------------------------------------------
// splines
// ---------------------------------------------------------------------------
__device__ __forceinline__ int spl_locate(const Spl& s, double t) {
    int i;
    if (s.uniform) {
        double f = floor((t - s.x0) * s.inv_h);
        i = (f < 0.) ? 0 : (f > (double)(s.n - 2) ? s.n - 2 : (int)f);
    } else {
        const double* x = s.x;
        if (t <= x[0]) {
            i = 0;
        } else if (t >= x[s.n - 1]) {
            i = s.n - 2;
        } else {
            int lo = 0, hi = s.n - 1;
            while (hi - lo >= 2) {
                int mid = (hi + lo) >> 1;
                if (x[mid] > t)
                    hi = mid;
                else
                    lo = mid;
            }
            i = lo;
        }
    }
    return i;
}

// The four coefficients of one interval and component are 32-byte aligned and
// read as two 16-byte loads.
__device__ __forceinline__ double spl_val(const Spl& s, int i, int comp, double dt) {
    const double2* c = reinterpret_cast<const double2*>(s.c + ((size_t)i * s.ncomp + comp) * 4);
    const double2 a = c[0], b = c[1];
    return a.x + dt * (a.y + dt * (b.x + dt * b.y));
}

__device__ __forceinline__ double spl_der(const Spl& s, int i, int comp, double dt) {
    const double2* c = reinterpret_cast<const double2*>(s.c + ((size_t)i * s.ncomp + comp) * 4);
    const double2 a = c[0], b = c[1];
    return a.y + dt * (2. * b.x + dt * (3. * b.y));
}

__device__ __forceinline__ double spl_eval1(const Spl& s, double t) {
    int i = spl_locate(s, t);
    return spl_val(s, i, 0, t - s.x[i]);
}

// B-spline basis at fraction f and the interval of t
// clamp to [lo, hi] (lo <= hi, no NaN inputs): the same values as the
// comparisons of the reference's SetLimits, as two min/max instructions
__device__ __forceinline__ double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

struct BBasis {
    int i;
    double w0, w1, w2, w3;
};
__device__ __forceinline__ BBasis bt_basis(const BTab& T, double t) {
    const double u = (t - T.x0) * T.ih;
    const double fl = clampd(floor(u), 0., (double)(T.knots - 2));
    const double f = u - fl, g = 1. - f, f2 = f * f, f3 = f2 * f;
    BBasis b;
    b.i = (int)fl;
    // the cubic B-spline weights times 6 (the tables hold coefficients / 6)
    b.w0 = g * g * g;
    b.w1 = 3. * f3 - 6. * f2 + 4.;
    b.w2 = -3. * f3 + 3. * f2 + 3. * f + 1.;
    b.w3 = f3;
    return b;
}
__device__ __forceinline__ double bt_apply(const double* c, const BBasis& b) {
    return c[b.i] * b.w0 + c[b.i + 1] * b.w1 + c[b.i + 2] * b.w2 + c[b.i + 3] * b.w3;
}
// basis weights of the first derivative d/dt (includes 1/h)
__device__ __forceinline__ BBasis bt_basis_d(const BTab& T, double t) {
    const double u = (t - T.x0) * T.ih;
    const double fl = clampd(floor(u), 0., (double)(T.knots - 2));
    const double f = u - fl, g = 1. - f, f2 = f * f, h = 3. * T.ih;
    BBasis b;
    b.i = (int)fl;
    b.w0 = -g * g * h;
    b.w1 = (3. * f2 - 4. * f) * h;
    b.w2 = (-3. * f2 + 2. * f + 1.) * h;
    b.w3 = f2 * h;
    return b;
}

// ---------------------------------------------------------------------------
// minimum image (path_class.h:108-150); contraction off so that positions and
// distances round exactly like the host restatement.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double minimg(double d, double L, double iL) {
#pragma clang fp contract(off)
    return d - rint(d * iL) * L;
}

struct LinkGeom {
    double rv[3], rpv[3], sv[3];
    double r, rp, s;
};

// Path::DrDrpDrrp: r = rq(b) - rp(b) imaged; r' = rq(b') - rp(b') shifted by the
// image nearest r; s = minimum image of r - r'.
__device__ __forceinline__ void link_geom(const double* a0, const double* c0, const double* a1, const double* c1,
                                          double L, double iL, LinkGeom& g) {
#pragma clang fp contract(off)
    for (int d = 0; d < 3; ++d) {
        double rv = c0[d] - a0[d];
        double rpv = c1[d] - a1[d];
        rv -= rint(rv * iL) * L;
        rpv += rint((rv - rpv) * iL) * L;
        double sv = rv - rpv;
        sv -= rint(sv * iL) * L;
        g.rv[d] = rv;
        g.rpv[d] = rpv;
        g.sv[d] = sv;
    }
    g.r = sqrt(g.rv[0] * g.rv[0] + g.rv[1] * g.rv[1] + g.rv[2] * g.rv[2]);
    g.rp = sqrt(g.rpv[0] * g.rpv[0] + g.rpv[1] * g.rpv[1] + g.rpv[2] * g.rpv[2]);
    g.s = sqrt(g.sv[0] * g.sv[0] + g.sv[1] * g.sv[1] + g.sv[2] * g.sv[2]);
}

// The same link geometry for the estimator kernels (compared at 1e-6 and
// 1e-10): contraction allowed, s without its identity minimum image, lengths by
// sqrt_pos.
__device__ __forceinline__ double sqrt_pos(double x);
__device__ __forceinline__ void link_geom_fast(const double* a0, const double* c0, const double* a1, const double* c1,
                                               double L, double iL, LinkGeom& g) {
    double rr = 0., pp = 0., ss = 0.;
    for (int d = 0; d < 3; ++d) {
        double rv = c0[d] - a0[d];
        double rpv = c1[d] - a1[d];
        rv -= rint(rv * iL) * L;
        rpv += rint((rv - rpv) * iL) * L;
        const double sv = rv - rpv;
        g.rv[d] = rv;
        g.rpv[d] = rpv;
        g.sv[d] = sv;
        rr += rv * rv;
        pp += rpv * rpv;
        ss += sv * sv;
    }
    g.r = sqrt_pos(rr);
    g.rp = sqrt_pos(pp);
    g.s = sqrt_pos(ss);
}

// Square root of a positive argument well inside the normal range (squared
// link lengths): v_rsq_f64, one Goldschmidt step and one Newton correction,
// without the range scaling of the general sqrt.  0 gives 0.
__device__ __forceinline__ double sqrt_pos(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    const double d = fma(-g, g, x);
    g = fma(d, h, g);
    return x > 0. ? g : 0.;
}

// DrDrpDrrp magnitudes |r|, |r'| and the squared |s|^2 (the link action needs
// only even powers of s).  r' is shifted to the image nearest r, so s = r - r'
// already lies in [-L/2, L/2] and its own minimum image is the identity (up to
// the sign at exactly L/2, which leaves |s| unchanged).  Contraction is allowed:
// the image choices are the same rint() of the same products, the lengths
// differ from the uncontracted form in the last bits only.
__device__ __forceinline__ void link_rr_s2(const double* a0, const double* c0, const double* a1, const double* c1,
                                           double L, double iL, double& r, double& rp, double& s2) {
    double rr = 0., pp = 0., ss = 0.;
    for (int d = 0; d < 3; ++d) {
        double rv = c0[d] - a0[d];
        double rpv = c1[d] - a1[d];
        rv -= rint(rv * iL) * L;
        rpv += rint((rv - rpv) * iL) * L;
        const double sv = rv - rpv;
        rr += rv * rv;
        pp += rpv * rpv;
        ss += sv * sv;
    }
    r = sqrt_pos(rr);
    rp = sqrt_pos(pp);
    s2 = ss;
}

__device__ __forceinline__ void clamp2(double lo, double hi, double& a, double& b) {
    a = clampd(a, lo, hi);
    b = clampd(b, lo, hi);
}

// BarePairAction::CalcV (bare_pair_action_class.h:99-123)
__device__ __forceinline__ double bare_v(const Rep& R, double r0, double r1) {
    clamp2(R.r_v_min, R.r_v_max, r0, r1);
    double v = 0.;
    if (R.is_coulomb) {
        v += (0.5 / r0) + (0.5 / r1);
    } else {
        v += 0.5 * spl_eval1(R.v_r, r0);
        v += 0.5 * spl_eval1(R.v_r, r1);
    }
    if (R.use_long_range) {
        clamp2(R.r_l_min, R.r_l_max, r0, r1);
        v -= 0.5 * spl_eval1(R.v_long_r, r0);
        v -= 0.5 * spl_eval1(R.v_long_r, r1);
    }
    return v;
}

// DavidPairAction off-diagonal sum: sum_{k=1..n} sum_{j=0..k} c_kj(q) z^2j s^2(k-j)
// (david_pair_action_class.h:80-99), components 2.. of the table at interval iq.
__device__ __forceinline__ double david_offdiag(const Spl& T, int n_order, int iq, double dq, double z, double s) {
    double z_2 = z * z, s_2 = s * s, i_s_2 = 1. / s_2, s_2_k = s_2;
    double u = 0.;
    for (int k = 1; k <= n_order; ++k) {
        double z_2_j = 1., cur_s = s_2_k;
        for (int j = 0; j <= k; ++j) {
            u += spl_val(T, iq, k * (k + 1) / 2 + j + 1, dq) * z_2_j * cur_s;
            z_2_j *= z_2;
            cur_s *= i_s_2;
        }
        s_2_k *= s_2;
    }
    return u;
}

// DavidPairAction::CalcU (david_pair_action_class.h:63-102), table T (u or du)
// add_v: CalcdUdBeta adds the potential (component 0).
__device__ __forceinline__ double david_u(const Rep& R, const Spl& T, double r0, double r1, double s, bool add_v) {
    double q = 0.5 * (r0 + r1);
    double z = r0 - r1;
    clamp2(R.g_min, R.g_max, r0, r1);
    int i0 = spl_locate(T, r0), i1 = spl_locate(T, r1);
    double d0 = r0 - T.x[i0], d1 = r1 - T.x[i1];
    double u = 0.5 * (spl_val(T, i0, 1, d0) + spl_val(T, i1, 1, d1));
    if (add_v) u += 0.5 * (spl_val(T, i0, 0, d0) + spl_val(T, i1, 0, d1));
    if (s > 0.0 && q < R.g_max) {
        int iq = spl_locate(T, q);
        u += david_offdiag(T, R.n_order, iq, q - T.x[iq], z, s);
    }
    return u;
}

// Interval of t on the general grid x[n]: the largest i <= n - 2 with x[i] <= t,
// 0 below the grid (binary search).
__device__ __forceinline__ int grid_locate(const double* x, int n, double t) {
    if (t <= x[0]) return 0;
    if (t >= x[n - 1]) return n - 2;
    int lo = 0, hi = n - 1;
    while (hi - lo >= 2) {
        const int mid = (hi + lo) >> 1;
        if (x[mid] > t)
            hi = mid;
        else
            lo = mid;
    }
    return lo;
}

// s(x, y) of a tensor spline; outside the grid the edge cell's polynomial
// continues.  With gx/gy non-null also ds/dx, ds/dy.
__device__ __forceinline__ double spl2_eval(const Spl2& S, double x, double y, double* gx = nullptr,
                                            double* gy = nullptr) {
    const int i = grid_locate(S.x, S.nx, x), j = grid_locate(S.y, S.ny, y);
    const double dx = x - S.x[i], dy = y - S.y[j];
    const double2* c = reinterpret_cast<const double2*>(S.c + ((size_t)j * (S.nx - 1) + i) * 16);
    double a[4], ay[4];
    for (int k = 0; k < 4; ++k) {
        const double2 lo = c[2 * k], hi = c[2 * k + 1];
        a[k] = lo.x + dy * (lo.y + dy * (hi.x + dy * hi.y));
        ay[k] = lo.y + dy * (2. * hi.x + dy * (3. * hi.y));
    }
    if (gx) *gx = a[1] + dx * (2. * a[2] + dx * (3. * a[3]));
    if (gy) *gy = ay[0] + dx * (ay[1] + dx * (ay[2] + dx * ay[3]));
    return a[0] + dx * (a[1] + dx * (a[2] + dx * a[3]));
}

// IlkkaPairAction::CalcU / CalcdUdBeta (ilkka_pair_action_class.h:83-107,
// 126-150): the tensor spline at x = q + s/2, y = q - s/2, less half the
// long-range spline at r and r' (clamped to its grid) when use_long_range.
__device__ __forceinline__ double ilkka_u(const Rep& R, const Spl2& T, const Spl& lr, double lo, double hi, double r0,
                                          double r1, double s) {
    const double q = 0.5 * (r0 + r1);
    double u = spl2_eval(T, q + 0.5 * s, q - 0.5 * s);
    if (R.use_long_range) {
        clamp2(lo, hi, r0, r1);
        u -= 0.5 * spl_eval1(lr, r0);
        u -= 0.5 * spl_eval1(lr, r1);
    }
    return u;
}

// CalcU of the pair family at level 0 (bare_pair_action_class.h:146-150)
__device__ __forceinline__ double pair_u(const Rep& R, double r0, double r1, double s) {
    if (R.pair_type == 1) return R.tau * bare_v(R, r0, r1);
    if (R.pair_type == 3) return ilkka_u(R, R.u2, R.u_long_r, R.r_u_min, R.r_u_max, r0, r1, s);
    return david_u(R, R.u, r0, r1, s, false);
}

// CalcdUdBeta (bare_pair_action_class.h:175-177, david_pair_action_class.h:126-168)
__device__ __forceinline__ double pair_du_dbeta(const Rep& R, double r0, double r1, double s) {
    if (R.pair_type == 1) return bare_v(R, r0, r1);
    if (R.pair_type == 3) return ilkka_u(R, R.du2, R.du_long_r, R.r_du_min, R.r_du_max, r0, r1, s);
    return david_u(R, R.du, r0, r1, s, true);
}

// CalcV (bare_pair_action_class.h:99-123, david_pair_action_class.h:42-57,
// ilkka_pair_action_class.h:34-54: the Bare form with a tabulated v_r)
__device__ __forceinline__ double pair_v(const Rep& R, double r0, double r1) {
    if (R.pair_type == 1 || R.pair_type == 3) return bare_v(R, r0, r1);
    clamp2(R.g_min, R.g_max, r0, r1);
    int i0 = spl_locate(R.u, r0), i1 = spl_locate(R.u, r1);
    return 0.5 * (spl_val(R.u, i0, 0, r0 - R.u.x[i0]) + spl_val(R.u, i1, 0, r1 - R.u.x[i1]));
}

// Link action U(r, r', s) at level 0 from LDS-staged B-spline tables `lds`
// (Bare: [v_r][v_long_r], David: u components 1..n_val).
__device__ __forceinline__ double pair_u_lds(const Rep& R, const double* lds, double r0, double r1, double s) {
    if (R.pair_type == 1) {
        clamp2(R.r_v_min, R.r_v_max, r0, r1);
        double v = 0.;
        if (R.is_coulomb) {
            v += (0.5 / r0) + (0.5 / r1);
        } else {
            v += 0.5 * bt_apply(lds, bt_basis(R.bt_a, r0));
            v += 0.5 * bt_apply(lds, bt_basis(R.bt_a, r1));
        }
        if (R.use_long_range) {
            const double* vl = lds + (R.is_coulomb ? 0 : R.bt_a.ncoef);
            clamp2(R.r_l_min, R.r_l_max, r0, r1);
            v -= 0.5 * bt_apply(vl, bt_basis(R.bt_b, r0));
            v -= 0.5 * bt_apply(vl, bt_basis(R.bt_b, r1));
        }
        return R.tau * v;
    }
    const BTab& T = R.bt_a;
    const double q = 0.5 * (r0 + r1), z = r0 - r1;
    clamp2(R.g_min, R.g_max, r0, r1);
    double u = 0.5 * (bt_apply(lds, bt_basis(T, r0)) + bt_apply(lds, bt_basis(T, r1)));
    if (s > 0.0 && q < R.g_max) {
        const BBasis bq = bt_basis(T, q);
        double z_2 = z * z, s_2 = s * s, i_s_2 = 1. / s_2, s_2_k = s_2;
        for (int k = 1; k <= R.n_order; ++k) {
            double z_2_j = 1., cur_s = s_2_k;
            for (int j = 0; j <= k; ++j) {
                // table component k(k+1)/2 + j + 1 is LDS component k(k+1)/2 + j
                u += bt_apply(lds + (size_t)(k * (k + 1) / 2 + j) * T.ncoef, bq) * z_2_j * cur_s;
                z_2_j *= z_2;
                cur_s *= i_s_2;
            }
            s_2_k *= s_2;
        }
    }
    return u;
}

// dU/dbeta of a link (CalcdUdBeta) from LDS tables: Bare uses the pair tables
// [v_r][v_long_r]; David uses du_kj_dbeta components 0..n_val at `ldu`.
__device__ __forceinline__ double pair_du_dbeta_lds(const Rep& R, const double* lds, const double* ldu, double r0,
                                                    double r1, double s) {
    if (R.pair_type == 1) return pair_u_lds(R, lds, r0, r1, s) / R.tau;
    const BTab& T = R.bt_du;
    const double q = 0.5 * (r0 + r1), z = r0 - r1;
    clamp2(R.g_min, R.g_max, r0, r1);
    const BBasis b0 = bt_basis(T, r0), b1 = bt_basis(T, r1);
    const double v = 0.5 * (bt_apply(ldu, b0) + bt_apply(ldu, b1));
    double du = 0.5 * (bt_apply(ldu + T.ncoef, b0) + bt_apply(ldu + T.ncoef, b1));
    du += v;
    if (s > 0.0 && q < R.g_max) {
        const BBasis bq = bt_basis(T, q);
        double z_2 = z * z, s_2 = s * s, i_s_2 = 1. / s_2, s_2_k = s_2;
        for (int k = 1; k <= R.n_order; ++k) {
            double z_2_j = 1., cur_s = s_2_k;
            for (int j = 0; j <= k; ++j) {
                du += bt_apply(ldu + (size_t)(k * (k + 1) / 2 + j + 1) * T.ncoef, bq) * z_2_j * cur_s;
                z_2_j *= z_2;
                cur_s *= i_s_2;
            }
            s_2_k *= s_2;
        }
    }
    return du;
}

// CalcV from LDS tables (Bare: [v_r][v_long_r]; David: component 0 of du).
__device__ __forceinline__ double pair_v_lds(const Rep& R, const double* lds, const double* ldu, double r0, double r1) {
    if (R.pair_type == 1) return pair_u_lds(R, lds, r0, r1, 0.) / R.tau;
    clamp2(R.g_min, R.g_max, r0, r1);
    return 0.5 * (bt_apply(ldu, bt_basis(R.bt_du, r0)) + bt_apply(ldu, bt_basis(R.bt_du, r1)));
}

// dU/dr, dU/dr', dU/ds of the level-0 link action from LDS tables.
// PT, NORD as for pair_u_lds_s2
template <int PT = 0, int NORD = 0>
__device__ __forceinline__ void pair_u_grad_lds(const Rep& R, const double* lds, double r0, double r1, double s,
                                                double& ur, double& urp, double& us) {
    ur = urp = us = 0.;
    if (PT == 1 || (PT == 0 && R.pair_type == 1)) {
        double a = r0, b = r1;
        const bool ina = (a >= R.r_v_min && a <= R.r_v_max), inb = (b >= R.r_v_min && b <= R.r_v_max);
        clamp2(R.r_v_min, R.r_v_max, a, b);
        double ga, gb;
        if (R.is_coulomb) {
            ga = -0.5 / (a * a);
            gb = -0.5 / (b * b);
        } else {
            ga = 0.5 * bt_apply(lds, bt_basis_d(R.bt_a, a));
            gb = 0.5 * bt_apply(lds, bt_basis_d(R.bt_a, b));
        }
        ga = ina ? ga : 0.;
        gb = inb ? gb : 0.;
        if (R.use_long_range) {
            const double* vl = lds + (R.is_coulomb ? 0 : R.bt_a.ncoef);
            const bool la = ina && (a >= R.r_l_min && a <= R.r_l_max), lb = inb && (b >= R.r_l_min && b <= R.r_l_max);
            clamp2(R.r_l_min, R.r_l_max, a, b);
            ga -= la ? 0.5 * bt_apply(vl, bt_basis_d(R.bt_b, a)) : 0.;
            gb -= lb ? 0.5 * bt_apply(vl, bt_basis_d(R.bt_b, b)) : 0.;
        }
        ur = R.tau * ga;
        urp = R.tau * gb;
        return;
    }
    const BTab& T = R.bt_a;
    const double q = 0.5 * (r0 + r1), z = r0 - r1;
    double a = r0, b = r1;
    const bool ina = (a >= R.g_min && a <= R.g_max), inb = (b >= R.g_min && b <= R.g_max);
    clamp2(R.g_min, R.g_max, a, b);
    ur = ina ? 0.5 * bt_apply(lds, bt_basis_d(T, a)) : 0.;
    urp = inb ? 0.5 * bt_apply(lds, bt_basis_d(T, b)) : 0.;
    if (s > 0.0 && q < R.g_max) {
        const BBasis bq = bt_basis(T, q), bqd = bt_basis_d(T, q);
        const int n_order = NORD ? NORD : R.n_order;
#pragma unroll
        for (int k = 1; k <= (NORD ? NORD : 3); ++k) {
            if (k > n_order) break;
#pragma unroll
            for (int j = 0; j <= k; ++j) {
                const double* cc = lds + (size_t)(k * (k + 1) / 2 + j) * T.ncoef;
                const double c = bt_apply(cc, bq), cd = bt_apply(cc, bqd);
                const int e = k - j;
                double zp = 1., zm = 0., sp = 1., sm = 0.;
                for (int t = 0; t < 2 * j; ++t) {
                    zm = zp;
                    zp *= z;
                }
                for (int t = 0; t < 2 * e; ++t) {
                    sm = sp;
                    sp *= s;
                }
                const double dqterm = 0.5 * cd * zp * sp;
                const double dz = c * (2. * j) * zm * sp;
                ur += dqterm + dz;
                urp += dqterm - dz;
                us += c * zp * (2. * e) * sm;
            }
        }
    }
}

// The same link action from |s|^2: sum_{k,j} c_kj(q) (z^2)^j (s^2)^(k-j) with the
// powers formed directly (the reference forms s^2k (1/s^2)^j; equal in exact
// arithmetic).
// PT: 1 Bare, 2 David, 0 read from the replica; NORD: David expansion order
// (0 = read from the replica)
template <int PT = 0, int NORD = 0>
__device__ __forceinline__ double pair_u_lds_s2(const Rep& R, const double* lds, double r0, double r1, double s2) {
    if (PT == 1 || (PT == 0 && R.pair_type == 1)) return pair_u_lds(R, lds, r0, r1, 0.);
    const BTab& T = R.bt_a;
    const double q = 0.5 * (r0 + r1), z = r0 - r1;
    clamp2(R.g_min, R.g_max, r0, r1);
    double u = 0.5 * (bt_apply(lds, bt_basis(T, r0)) + bt_apply(lds, bt_basis(T, r1)));
    if (s2 > 0.0 && q < R.g_max) {
        const BBasis bq = bt_basis(T, q);
        const double z2 = z * z;
        double sp[4], zp[4];
        sp[0] = 1.;
        zp[0] = 1.;
#pragma unroll
        for (int e = 1; e <= (NORD ? NORD : 3); ++e) {
            sp[e] = sp[e - 1] * s2;
            zp[e] = zp[e - 1] * z2;
        }
        const int n_order = NORD ? NORD : R.n_order;
#pragma unroll
        for (int k = 1; k <= (NORD ? NORD : 3); ++k) {
            if (k > n_order) break;
#pragma unroll
            for (int j = 0; j <= k; ++j)
                u += bt_apply(lds + (size_t)(k * (k + 1) / 2 + j) * T.ncoef, bq) * zp[j] * sp[k - j];
        }
    }
    return u;
}

// Partial derivatives of the level-0 link action U(r, r', s) of pair_u.
__device__ __forceinline__ void pair_u_grad(const Rep& R, double r0, double r1, double s, double& ur, double& urp,
                                            double& us) {
    ur = urp = us = 0.;
    if (R.pair_type == 3) {
        // x = q + s/2, y = q - s/2, q = (r + r')/2
        const double q = 0.5 * (r0 + r1);
        double sx, sy;
        spl2_eval(R.u2, q + 0.5 * s, q - 0.5 * s, &sx, &sy);
        ur = urp = 0.5 * (sx + sy);
        us = 0.5 * (sx - sy);
        if (R.use_long_range) {
            double a = r0, b = r1;
            const bool la = (a >= R.r_u_min && a <= R.r_u_max), lb = (b >= R.r_u_min && b <= R.r_u_max);
            clamp2(R.r_u_min, R.r_u_max, a, b);
            const int ia = spl_locate(R.u_long_r, a), ib = spl_locate(R.u_long_r, b);
            ur -= la ? 0.5 * spl_der(R.u_long_r, ia, 0, a - R.u_long_r.x[ia]) : 0.;
            urp -= lb ? 0.5 * spl_der(R.u_long_r, ib, 0, b - R.u_long_r.x[ib]) : 0.;
        }
        return;
    }
    if (R.pair_type == 1) {
        double a = r0, b = r1;
        bool ina = (a >= R.r_v_min && a <= R.r_v_max), inb = (b >= R.r_v_min && b <= R.r_v_max);
        clamp2(R.r_v_min, R.r_v_max, a, b);
        double ga = 0., gb = 0.;
        if (R.is_coulomb) {
            ga = -0.5 / (a * a);
            gb = -0.5 / (b * b);
        } else {
            int ia = spl_locate(R.v_r, a), ib = spl_locate(R.v_r, b);
            ga = 0.5 * spl_der(R.v_r, ia, 0, a - R.v_r.x[ia]);
            gb = 0.5 * spl_der(R.v_r, ib, 0, b - R.v_r.x[ib]);
        }
        ga = ina ? ga : 0.;
        gb = inb ? gb : 0.;
        if (R.use_long_range) {
            bool la = ina && (a >= R.r_l_min && a <= R.r_l_max), lb = inb && (b >= R.r_l_min && b <= R.r_l_max);
            clamp2(R.r_l_min, R.r_l_max, a, b);
            int ia = spl_locate(R.v_long_r, a), ib = spl_locate(R.v_long_r, b);
            double la_d = 0.5 * spl_der(R.v_long_r, ia, 0, a - R.v_long_r.x[ia]);
            double lb_d = 0.5 * spl_der(R.v_long_r, ib, 0, b - R.v_long_r.x[ib]);
            ga -= la ? la_d : 0.;
            gb -= lb ? lb_d : 0.;
        }
        ur = R.tau * ga;
        urp = R.tau * gb;
        return;
    }
    const Spl& T = R.u;
    double q = 0.5 * (r0 + r1), z = r0 - r1;
    double a = r0, b = r1;
    bool ina = (a >= R.g_min && a <= R.g_max), inb = (b >= R.g_min && b <= R.g_max);
    clamp2(R.g_min, R.g_max, a, b);
    int ia = spl_locate(T, a), ib = spl_locate(T, b);
    ur = ina ? 0.5 * spl_der(T, ia, 1, a - T.x[ia]) : 0.;
    urp = inb ? 0.5 * spl_der(T, ib, 1, b - T.x[ib]) : 0.;
    if (s > 0.0 && q < R.g_max) {
        int iq = spl_locate(T, q);
        double dq = q - T.x[iq];
        for (int k = 1; k <= R.n_order; ++k)
            for (int j = 0; j <= k; ++j) {
                int comp = k * (k + 1) / 2 + j + 1;
                double c = spl_val(T, iq, comp, dq), cd = spl_der(T, iq, comp, dq);
                int e = k - j;  // power of s^2
                // zp = z^2j, zm = z^(2j-1); sp = s^2e, sm = s^(2e-1)
                double zp = 1., zm = 0., sp = 1., sm = 0.;
                for (int t = 0; t < 2 * j; ++t) {
                    zm = zp;
                    zp *= z;
                }
                for (int t = 0; t < 2 * e; ++t) {
                    sm = sp;
                    sp *= s;
                }
                double dqterm = 0.5 * cd * zp * sp;         // q = (r + r') / 2
                double dz = c * (2. * j) * zm * sp;          // z = r - r'
                ur += dqterm + dz;
                urp += dqterm - dz;
                us += c * zp * (2. * e) * sm;
            }
    }
}

// ---------------------------------------------------------------------------
// kinetic link action -log rho_free (free_spline_class.h:76-84) and its slope
// ---------------------------------------------------------------------------
__device__ __forceinline__ double free_neg_log(const Spl* img, int images, double i4lt, const double* d) {
    double tot = 0.;
    if (images > 0)
        for (int k = 0; k < 3; ++k) tot -= spl_eval1(*img, d[k]);
    return -(tot - (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * i4lt);
}

// ---------------------------------------------------------------------------
// phase factors e^{i m phi} by the recurrence of KSpace::CalcC (k_space_class.h:83-94)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// table[0..mx] = e^{i m phi}, m = 0..mx
__device__ __forceinline__ void phase_chain(double phi, int mx, double2* table) {
    double sn, cs;
    sincos(phi, &sn, &cs);
    double2 t = make_double2(cs, sn);
    double2 c = make_double2(1., 0.);
    table[0] = c;
    for (int m = 1; m <= mx; ++m) {
        c = cmul(t, c);
        table[m] = c;
    }
}

__device__ __forceinline__ double2 phase_at(const double2* table, int m) {
    double2 v = table[m < 0 ? -m : m];
    if (m < 0) v.y = -v.y;
    return v;
}

// e^{i k.r} = c0(n0) c1(n1) c2(n2), tables laid out [3][stride]
__device__ __forceinline__ double2 phase_k(const double2* tab, int stride, int packed) {
    int n0 = (packed & 255) - 128, n1 = ((packed >> 8) & 255) - 128, n2 = ((packed >> 16) & 255) - 128;
    double2 f = phase_at(tab, n0);
    f = cmul(f, phase_at(tab + stride, n1));
    f = cmul(f, phase_at(tab + 2 * stride, n2));
    return f;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) for the device RNG
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox(uint4 c, uint2 k) {
    for (int i = 0; i < 10; ++i) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {  // (0, 1]
    uint64_t v = (((uint64_t)a << 32) | b) >> 11;
    return (double)(v + 1) * (1.0 / 9007199254740992.0);
}

struct Rng {
    uint2 key;
    uint32_t tag;
    uint64_t seq;
    __device__ uint4 call(uint32_t i) const {
        return philox(make_uint4(i, tag, (uint32_t)seq, (uint32_t)(seq >> 32)), key);
    }
    __device__ uint32_t index(uint32_t i, int word, uint32_t n) const {
        uint4 w = call(i);
        uint32_t x = word == 0 ? w.x : (word == 1 ? w.y : (word == 2 ? w.z : w.w));
        return (uint32_t)(((uint64_t)x * n) >> 32);
    }
    __device__ double unif(uint32_t i, int half) const {
        uint4 w = call(i);
        return half ? u53(w.z, w.w) : u53(w.x, w.y);
    }
    __device__ double normal(uint32_t i, int which) const {
        uint4 w = call(i);
        double u1 = u53(w.x, w.y), u2 = u53(w.z, w.w);
        double r = sqrt(-2. * log(u1));
        double sn, cs;
        sincospi(2. * u2, &sn, &cs);
        return which ? r * sn : r * cs;
    }
    // normal(i, 0) and normal(i, 1) from one call
    __device__ void normal2(uint32_t i, double* z0, double* z1) const {
        uint4 w = call(i);
        double u1 = u53(w.x, w.y), u2 = u53(w.z, w.w);
        double r = sqrt(-2. * log(u1));
        double sn, cs;
        sincospi(2. * u2, &sn, &cs);
        *z0 = r * cs;
        *z1 = r * sn;
    }
};

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.;
    if (threadIdx.x == 0)
        for (int i = 0; i < NT / kWave; ++i) t += sh[i];
    return t;  // valid on thread 0
}

}  // namespace spimc
