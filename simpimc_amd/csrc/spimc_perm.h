// spimc_perm.h -- permuting bisection (PermBisectIterative) and the kernels
// that evaluate a proposal whose moved beads carry more than one particle
// label: the chains of a permutation cycle, or a bisection chain whose label
// changes where it crosses the wrap of a permuted path.
//
//   k_propose_perm  one workgroup per replica: bead0, the iterative cycle walk,
//                   the bisection of every chain down to level 0
//   k_pair_multi    workgroups per replica loop over the window links: the
//                   label-based pair action of every label with a moved bead
//                   on the link, OLD and NEW
//   k_rho_multi     one workgroup per (replica, moved slice, k chunk): the
//                   incremental rho_k of a slice with several moved beads
// The decision and commit (relabelling included) stay in k_decide.
#pragma once

#include "spimc_kernels.h"

namespace spimc {

constexpr int kPermBlock = 256;
constexpr int kPermMaxN = 4096;  // particles of a replica that runs permuting moves

// Window position t (0..nb) of chain i of a permuting proposal: the chain is
// walked from (ch_lbl[i], b0) and carries label ch_lbl2[i] past the wrap.
__device__ __forceinline__ int chain_label(const Rep& R, int b0, int t, int i) {
    return (b0 + t >= R.M) ? R.ch_lbl2[i] : R.ch_lbl[i];
}
__device__ __forceinline__ double chain_old(const Rep& R, int b0, int t, int i, int d) {
    return pos_at(R, d, (b0 + t) % R.M, chain_label(R, b0, t, i));
}
// NEW mode: the start is shared; the end of chain i is the old end of chain
// ch_nxt[i] (i + 1 of a cycle walk; PermBisect::PermuteBeads,
// perm_bisect_class.h:32-44); the beads between are the proposal's.
__device__ __forceinline__ double chain_new(const Rep& R, int b0, int nb, int n, int t, int i, int d) {
    if (t == 0) return chain_old(R, b0, 0, i, d);
    if (t == nb) return chain_old(R, b0, nb, R.ch_nxt[i], d);
    return R.ch_new[((size_t)i * (nb - 1) + t - 1) * 3 + d];
}

constexpr int kPermStage = 256;  // beads of a proposal bisected on LDS by one wavefront

// Orders the LDS traffic of one wavefront (the compiler may not move loads or
// stores across it).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The Metropolis step closing level `level` of a permuting proposal
// (perm_bisect_iterative_class.h:206-215), on one thread: returns 1 if the
// proposal is rejected.  A coarse level returns its action change in *change;
// level 0 fills the proposal for the pair, rho_k and decision kernels, and
// *status (1, or 3 after a replay-tape overrun).
__device__ __forceinline__ int perm_level_step(const Rep& R, Prop* pr, int level, double lspo, double lspn, double ko,
                                               double kn, double prev, double u, int bad, int b0, int nb, int n,
                                               int n_lev, long long cur, int* status, double* change) {
    const int M = R.M, nm = nb - 1;
    const double lsr = -lspn + lspo;
    int rej = 0;
    if (level > 0) {
        const double ch = (0. + kn) - (0. + ko);
        const double la = lsr - ch + prev;
        rej = la < log(u);
        *change = ch;
    } else {
        pr->status = bad ? 3 : 1;
        *status = bad ? 3 : 1;
        pr->move = 3;
        pr->p = R.ch_lbl[0];
        pr->p2 = R.ch_lbl2[0];
        pr->generic = 1;
        pr->b0 = b0;
        pr->n_links = R.pair_type ? nb : 0;
        pr->first = (b0 + 1) % M;
        pr->count = nm;
        pr->lsr = lsr;
        pr->k_old = ko;
        pr->k_new = kn;
        pr->prev = prev;
        pr->u0 = u;
        pr->tape_next = cur + 1;
        pr->n_chain = n;
        pr->e_slice = (b0 + nb - 1) % M;
        pr->n_level = n_lev;
    }
    if (rej) {
        pr->status = 2;
        pr->move = 3;
        pr->p = R.ch_lbl[0];
        pr->p2 = R.ch_lbl2[0];
        pr->generic = 1;
        pr->b0 = b0;
        pr->count = 0;
        pr->n_chain = n;
        pr->n_level = n_lev;
        pr->tape_next = cur + 1;
    }
    return rej;
}

// ---------------------------------------------------------------------------
// PermBisectTable (perm_bisect_table_class.h:10-228): the cycle list is every
// permutation type of k = n_perm_part particles crossed with every k-subset of
// the particles, type-major (BuildCycles, :54-102).  scaffold's GenPerm and
// GenCombPermK are not in the reference tree; both lists are taken here in
// lexicographic order (permutations of 0..k-1; subsets by their sorted
// members), as in the oracle.  An entry is never stored: thread t of the
// proposal's workgroup walks entries [E t / 256, E (t + 1) / 256) of the list.
// ---------------------------------------------------------------------------
constexpr int kTabMaxK = 6;

__device__ __forceinline__ long long tab_binom(int n, int r) {
    if (r < 0 || r > n) return 0;
    long long v = 1;
    for (int i = 1; i <= r; ++i) v = v * (n - r + i) / i;
    return v;
}
// element j of a k-array held in registers (no dynamic register indexing)
__device__ __forceinline__ int tab_pick(const int (&a)[kTabMaxK], int j) {
    int v = a[0];
#pragma unroll
    for (int i = 1; i < kTabMaxK; ++i) v = j == i ? a[i] : v;
    return v;
}
// the rank-th k-subset of 0..n-1 in lexicographic order
__device__ __forceinline__ void tab_subset(long long rank, int n, int k, int (&s)[kTabMaxK]) {
    int x = 0;
#pragma unroll
    for (int i = 0; i < kTabMaxK; ++i) {
        if (i < k) {
            for (;;) {
                const long long c = tab_binom(n - x - 1, k - i - 1);
                if (rank < c) break;
                rank -= c;
                ++x;
            }
            s[i] = x++;
        } else {
            s[i] = 0;
        }
    }
}
// the next k-subset in lexicographic order; false after the last
__device__ __forceinline__ bool tab_next_subset(int n, int k, int (&s)[kTabMaxK]) {
    int top = -1;  // the last position below its maximum n - k + i
#pragma unroll
    for (int i = 0; i < kTabMaxK; ++i)
        if (i < k && s[i] != n - k + i) top = i;
    if (top < 0) return false;
#pragma unroll
    for (int i = 0; i < kTabMaxK; ++i) {
        if (i == top) s[i] += 1;
        if (i > top && i < k) s[i] = s[i - 1] + 1;
    }
    return true;
}
// the rank-th permutation of 0..k-1 in lexicographic order
__device__ __forceinline__ void tab_perm(int rank, int k, int (&p)[kTabMaxK]) {
    int fact = 1;
    for (int i = 2; i < k; ++i) fact *= i;  // (k - 1)!
    unsigned used = 0;
#pragma unroll
    for (int i = 0; i < kTabMaxK; ++i) {
        if (i < k) {
            const int d = rank / fact;
            rank -= d * fact;
            if (k - 1 - i > 0) fact /= (k - 1 - i);
            int c = 0, seen = -1;
            for (int v = 0; v < k; ++v)
                if (!((used >> v) & 1u) && ++seen == d) {
                    c = v;
                    break;
                }
            p[i] = c;
            used |= 1u << c;
        } else {
            p[i] = i;
        }
    }
}
// ConstructPermTable's weight of entry (subset s, type p): the product of
// t(part(i), part(perm(i))) in order (perm_bisect_table_class.h:18-20)
__device__ __forceinline__ double tab_weight(const double* __restrict__ t, int N, int k, const int (&s)[kTabMaxK],
                                             const int (&p)[kTabMaxK]) {
    double w = 1.;
#pragma unroll
    for (int i = 0; i < kTabMaxK; ++i)
        if (i < k) w *= t[(size_t)s[i] * N + tab_pick(s, p[i])];
    return w;
}
// UpdatePermTable (perm_bisect_table_class.h:31-51) into t: the path ends n links
// after slice b0; pi (LDS, or null for the identity) maps particle j to the
// particle whose old end j's chain now reaches.
__device__ __forceinline__ void tab_build(const Rep& R, int b0, int nb, const int* pi, double* __restrict__ t,
                                          int tid) {
#pragma clang fp contract(off)
    const int N = R.N, e_end = (b0 + nb) % R.M;
    const bool end_wrapped = b0 + nb >= R.M;
    const double i4ltn = 1. / (4. * R.lambda * R.tau) / nb;
    for (int idx = tid; idx < N * N; idx += kPermBlock) {
        const int i = idx / N, j = idx % N;
        const int ei = pi ? pi[i] : i, ej = pi ? pi[j] : j;
        const int li = end_wrapped ? label_after_wrap(R, ei) : ei, lj = end_wrapped ? label_after_wrap(R, ej) : ej;
        double di[3], dj[3];
        for (int d = 0; d < 3; ++d) {
            const double a = pos_at(R, d, b0, i);
            di[d] = minimg(a - pos_at(R, d, e_end, li), R.L, R.iL);
            dj[d] = minimg(a - pos_at(R, d, e_end, lj), R.L, R.iL);
        }
        const double s_ii = di[0] * di[0] + di[1] * di[1] + di[2] * di[2];
        const double s_ij = dj[0] * dj[0] + dj[1] * dj[1] + dj[2] * dj[2];
        const double e = (-s_ij + s_ii) * i4ltn;
        t[idx] = e > R.perm_log_eps ? exp(e) : 0.;
    }
}
// One pass of the workgroup over the cycle list of table t: returns the
// running total of the weights above epsilon (each thread sums its entries in
// order, the thread sums are added in thread order: the association differs
// from the reference's single running sum in the last bits).  With select, also
// the first entry whose running total exceeds x and the last entry above
// epsilon (-1 if none) in *sel, *last.  sh: kPermBlock doubles of LDS.
__device__ __forceinline__ double tab_pass(const Rep& R, const double* __restrict__ t, int tid, bool select, double x,
                                           double* sh, long long* shl, long long* sel, long long* last) {
    const int N = R.N, k = R.tab_k;
    const long long n_comb = tab_binom(N, k);
    long long fk = 1;
    for (int i = 2; i <= k; ++i) fk *= i;
    const long long E = n_comb * fk;
    const long long lo = E * tid / kPermBlock, hi = E * (tid + 1) / kPermBlock;
    const double eps = R.perm_eps;
    int s[kTabMaxK], p[kTabMaxK];
    auto walk = [&](double run, bool pick, long long* found, long long* lastp) {
        if (lo >= hi) return run;
        int type = (int)(lo / n_comb);
        tab_perm(type, k, p);
        tab_subset(lo % n_comb, N, k, s);
        for (long long e = lo; e < hi; ++e) {
            const double w = tab_weight(t, N, k, s, p);
            if (w > eps) {
                run += w;
                if (pick) {
                    *lastp = e;
                    if (*found < 0 && run > x) *found = e;
                }
            }
            if (!tab_next_subset(N, k, s)) {
                ++type;
                if (type < (int)fk) {
                    tab_perm(type, k, p);
                    tab_subset(0, N, k, s);
                }
            }
        }
        return run;
    };
    sh[tid] = walk(0., false, nullptr, nullptr);
    __syncthreads();
    if (tid == 0) {
        double acc = 0.;
        for (int i = 0; i < kPermBlock; ++i) {
            const double v = sh[i];
            sh[i] = acc;  // exclusive
            acc += v;
        }
        shl[0] = __double_as_longlong(acc);
    }
    __syncthreads();
    const double total = __longlong_as_double(shl[0]);
    if (select) {
        long long f = -1, l = -1;
        walk(sh[tid], true, &f, &l);
        __syncthreads();
        shl[1 + tid] = f;
        shl[1 + kPermBlock + tid] = l;
        __syncthreads();
        if (tid == 0) {
            long long fb = -1, lb = -1;
            for (int i = 0; i < kPermBlock; ++i) {
                if (fb < 0 && shl[1 + i] >= 0) fb = shl[1 + i];
                if (shl[1 + kPermBlock + i] >= 0) lb = shl[1 + kPermBlock + i];
            }
            shl[1] = fb;
            shl[2] = lb;
        }
        __syncthreads();
        *sel = shl[1];
        *last = shl[2];
        __syncthreads();
    }
    return total;
}

// ---------------------------------------------------------------------------
// PermBisectIterative::Attempt (perm_bisect_iterative_class.h:113-223) down to
// level 0.  Draws in the reference's order: bead0, the first particle, per
// walk step the continue and select uniforms (SelectCycleIterative, :32-110),
// then per level the chains' normals chain by chain and one uniform.  The coarse
// levels carry only the kinetic action (pair_action_class.h:269).  The
// permutation-table row of the current particle is rebuilt at each step: row p
// of the reference's t_c is t(p, .) with the visited particles removed and the
// closing link t(p, p0) restored (:48-53).
// Dynamic LDS: N doubles (the row) and N ints (visited step, -1 if not).
// ---------------------------------------------------------------------------
// tuned: the round also runs the LDS pair kernel, which takes a one-chain
// proposal whose label does not change inside the window
// stage_max: the largest proposal, in chain beads, bisected on LDS (at most
// kPermStage)
// TABLE: PermBisectTable::Attempt (perm_bisect_table_class.h:122-228) instead:
// bead0, the table's selection uniform, then the same bisection; the final
// table check after level 0 is k_decide's (Prop::tab_*).
template <bool TABLE>
__global__ __launch_bounds__(kPermBlock) void k_propose_perm(const Rep* __restrict__ reps, int n_reps, int tuned,
                                                             int stage_max) {
    const int r = blockIdx.x;
    if (r >= n_reps) return;
    const Rep& R = reps[r];
    const int tid = threadIdx.x;
    extern __shared__ double lds[];
    double* t_row = lds;
    int* vis = reinterpret_cast<int*>(lds + R.N);
    __shared__ double red[kPermBlock / kWave];
    __shared__ int s_i[8];
    __shared__ double s_w[2];
    __shared__ unsigned long long s_seq;
    __shared__ long long s_cur;
    Prop* pr = R.prop;
    if (tid == 0) {
        Counters* c = R.cnt;
        unsigned int rem = c->remaining;
        s_i[0] = rem > 0;
        if (rem > 0) {
            c->remaining = rem - 1;
            s_seq = c->seq;
            c->seq = s_seq + 1;
        }
    }
    __syncthreads();
    // an attempt that does not reach level 0 leaves the bisection's rho_k items idle
    if (!s_i[0]) {
        if (tid == 0) pr->status = 0;
        if (tid < kWave) rho_items_write(R, 0, 0, 0, 0, tid);
        return;
    }
    const int N = R.N, M = R.M, n_lev = R.perm_n_level, nb = 1 << n_lev;
    const Rng rng{make_uint2((uint32_t)R.key, (uint32_t)(R.key >> 32)), 3u, s_seq};
    const bool tape = R.tape != nullptr;
    // device draw counters: 0 bead0 and p0; 1 + step the walk; N + 3 + level the
    // level uniforms; N + 3 + n_level + g / 2 the normals
    const uint32_t c_lev = (uint32_t)N + 3u, c_norm = c_lev + (uint32_t)n_lev;
    for (int i = tid; i < N; i += kPermBlock) vis[i] = -1;
    if (tid == 0) {
        int bad = 0, b0, p0;
        long long cur = 0;
        if (tape) {
            cur = *R.tape_cursor;
            b0 = (int)tape_get(R, cur, &bad);
            p0 = TABLE ? 0 : (int)tape_get(R, cur + 1, &bad);
            cur += TABLE ? 1 : 2;
            if (b0 < 0 || b0 >= M || p0 < 0 || p0 >= N) {
                bad = 1;
                b0 = 0;
                p0 = 0;
            }
        } else {
            b0 = (int)rng.index(0, 0, M);
            p0 = TABLE ? 0 : (int)rng.index(0, 1, N);
        }
        pr->tab_on = 0;
        s_i[1] = b0;
        s_i[2] = p0;
        s_i[3] = bad;
        s_cur = cur;
        s_w[0] = 1.;
    }
    __syncthreads();
    const int b0 = s_i[1], p0 = s_i[2];
    const int e_end = (b0 + nb) % M;
    const bool end_wrapped = b0 + nb >= M;
    const double i4ltn = 1. / (4. * R.lambda * R.tau) / nb;
    // ---- the cycle walk ----
    int p = p0, n = 0, stop = 0;  // stop: 1 closed, 2 walk stopped, 3 error
    __shared__ int s_tp[kTabMaxK];  // the table entry's permutation type
    if constexpr (TABLE) {
        // ---- the table of the old path, its running total, the selection
        // (ConstructPermTable, SelectCycle: perm_bisect_table_class.h:10-28,105-119)
        // and the table of the new path (:216-225) ----
        __shared__ double sh[kPermBlock];
        __shared__ long long shl[1 + 2 * kPermBlock];
        const int k = R.tab_k;
        double* t0 = R.perm_t;
        double* t1 = R.perm_t + (size_t)N * N;
        tab_build(R, b0, nb, nullptr, t0, tid);
        __syncthreads();
        long long sel = -1, last = -1;
        const double tot0 = tab_pass(R, t0, tid, false, 0., sh, shl, &sel, &last);
        if (tid == 0) {
            int bad = s_i[3];
            long long cur = s_cur;
            const double u = tape ? tape_get(R, cur++, &bad) : rng.unif(1u, 0);
            s_w[1] = tot0 * u;  // UnifRand(0, perm_tot_0)
            s_i[3] = bad;
            s_cur = cur;
        }
        __syncthreads();
        tab_pass(R, t0, tid, true, s_w[1], sh, shl, &sel, &last);
        // x at or past the total (rounding): the last entry free-running, an error on a tape
        if (sel < 0 && !tape) sel = last;
        if (sel < 0) {
            stop = 3;
        } else {
            stop = 1;
            n = k;
            const long long n_comb = tab_binom(N, k);
            int sb[kTabMaxK], tp[kTabMaxK];
            tab_perm((int)(sel / n_comb), k, tp);
            tab_subset(sel % n_comb, N, k, sb);
            if (tid == 0) {
                for (int i = 0; i < k; ++i) {
                    R.ch_lbl[i] = tab_pick(sb, i);
                    s_tp[i] = tab_pick(tp, i);
                }
                s_w[0] = tab_weight(t0, N, k, sb, tp);
            }
            // the new path's ends: chain part(i) reaches the old end of part(type(i))
            for (int j = tid; j < N; j += kPermBlock) vis[j] = j;
            __syncthreads();
            if (tid == 0)
                for (int i = 0; i < k; ++i) vis[tab_pick(sb, i)] = tab_pick(sb, tab_pick(tp, i));
            __syncthreads();
            double tot1 = tot0;
            if (k < N) {
                tab_build(R, b0, nb, vis, t1, tid);
                __syncthreads();
                tot1 = tab_pass(R, t1, tid, false, 0., sh, shl, &sel, &last);
            }
            if (tid == 0) {
                pr->tab_on = k < N;
                pr->tab_ratio = R.tab_balance ? tot0 / (tot1 * s_w[0]) : tot0 / tot1;
                pr->u_tab = tape ? 0. : rng.unif(2u, 0);
            }
            __syncthreads();
            for (int j = tid; j < N; j += kPermBlock) vis[j] = -1;
        }
        __syncthreads();
    }
    for (int step = 0; !TABLE && !stop; ++step) {
        if (tid == 0) {
            R.ch_lbl[n] = p;
            vis[p] = n;
        }
        ++n;
        {
#pragma clang fp contract(off)
            double a[3];
            for (int d = 0; d < 3; ++d) a[d] = pos_at(R, d, b0, p);
            for (int j = tid; j < N; j += kPermBlock) {
                const int lj = end_wrapped ? label_after_wrap(R, j) : j;
                double dd[3];
                for (int d = 0; d < 3; ++d) dd[d] = minimg(a[d] - pos_at(R, d, e_end, lj), R.L, R.iL);
                const double e = (-(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2])) * i4ltn;
                t_row[j] = e > R.perm_log_eps ? exp(e) : 0.;
            }
        }
        __syncthreads();
        // the first wavefront takes the row: its sums, the continue draw and the
        // selection; the terms are summed in a different order than the
        // reference's running sums (a difference in the last bits)
        if (tid < kWave) {
            const int lane = tid;
            double q_p = 0., q_c = 0.;
            for (int i = lane; i < N; i += kWave) {
                const double ti = t_row[i];
                q_p += ti;
                q_c += (i == p0) ? ti : (vis[i] >= 0 ? 0. : ti);
            }
            q_p = wave_sum(q_p);
            q_c = wave_sum(q_c);
            int bad = s_i[3];
            long long cur = s_cur;
            double u = 0., x = 0.;
            if (lane == 0) u = tape ? tape_get(R, cur++, &bad) : rng.unif(1u + (uint32_t)step, 0);
            u = __shfl(u, 0, kWave);
            const bool go = !(q_c / q_p < u);
            if (go && lane == 0) x = tape ? tape_get(R, cur++, &bad) : rng.unif(1u + (uint32_t)step, 1);
            x = __shfl(x, 0, kWave);
            int found = -1;
            if (go) {
                // t_Q of the reference's scan, lane by lane over contiguous blocks:
                // the first index whose running sum of t_c / Q_p_c exceeds x
                const int c = (N + kWave - 1) / kWave, lo = min(N, lane * c), hi = min(N, lo + c);
                double part = 0.;
                for (int i = lo; i < hi; ++i) {
                    const double tc = (i == p0) ? t_row[i] : (vis[i] >= 0 ? 0. : t_row[i]);
                    part += tc / q_c;
                }
                double incl = part;  // inclusive scan over the lanes
                for (int o = 1; o < kWave; o <<= 1) {
                    const double v = __shfl_up(incl, o, kWave);
                    if (lane >= o) incl += v;
                }
                double t_q = incl - part;
                int mine = 1 << 30, last = -1;
                for (int i = lo; i < hi; ++i) {
                    const double tc = (i == p0) ? t_row[i] : (vis[i] >= 0 ? 0. : t_row[i]);
                    if (tc > 0.) last = i;
                    t_q += tc / q_c;
                    if (t_q > x && mine == (1 << 30)) mine = i;
                }
                for (int o = kWave / 2; o > 0; o >>= 1) {
                    mine = min(mine, __shfl_xor(mine, o, kWave));
                    last = max(last, __shfl_xor(last, o, kWave));
                }
                found = mine < N ? mine : -1;
                // Without a replay tape a scan that selects nothing is rounding (its
                // total an ulp short of x: the last candidate is the selection) or a
                // row without candidates (every weight below epsilon: the walk
                // cannot continue, the attempt is rejected).  With a tape it stays
                // an error, as in the oracle.
                if (found < 0 && !tape) found = last >= 0 ? last : -2;
            }
            if (lane == 0) {
                int st = 0, nx = p;
                if (!go || found == -2) {
                    st = 2;
                } else if (found < 0) {
                    st = 3;  // no particle selected (a degenerate row on a replay tape)
                } else {
                    nx = found;
                    // the cycle weight's factors t(ps[i], ps[i+1]) / t(ps[i], ps[i]), in order
                    s_w[0] = s_w[0] * (t_row[nx] / t_row[p]);
                    if (nx == p0)
                        st = 1;
                    else if (n >= N)
                        st = 3;
                }
                s_i[3] = bad;
                s_cur = cur;
                s_i[4] = st;
                s_i[5] = nx;
            }
        }
        __syncthreads();
        stop = s_i[4];
        p = s_i[5];
    }
    long long cur = s_cur;
    int bad = s_i[3];
    // the visited map becomes the chain map before the wrap; chain labels after
    // it; the chain-end map of NEW mode and the first chain of each of its cycles
    for (int i = tid; i < n; i += kPermBlock) {
        const int l = R.ch_lbl[i];
        R.mv_pre[l] = i;
        const int l2 = label_after_wrap(R, l);
        R.ch_lbl2[i] = l2;
        R.mv_post[l2] = i;
        if (TABLE) {
            R.ch_nxt[i] = s_tp[i];
            int lead = 1;
            for (int j = s_tp[i]; j != i; j = s_tp[j]) lead &= j > i;
            R.ch_lead[i] = lead;
        } else {
            R.ch_nxt[i] = (i + 1) % n;
            R.ch_lead[i] = i == 0;
        }
    }
    __syncthreads();  // chain_old reads the labels of every chain across the wrap
    if (stop != 1) {
        if (tid == 0) {
            pr->status = stop == 3 ? 3 : 2;
            pr->move = 3;
            pr->p = p0;
            pr->p2 = p0;
            pr->generic = 1;
            pr->b0 = b0;
            pr->count = 0;
            pr->n_chain = 0;
            pr->n_level = n_lev;
            pr->tape_next = cur;
        }
        __syncthreads();
        for (int i = tid; i < n; i += kPermBlock) {
            R.mv_pre[R.ch_lbl[i]] = -1;
            R.mv_post[R.ch_lbl2[i]] = -1;
        }
        if (tid < kWave) rho_items_write(R, 0, 0, 0, 0, tid);
        return;
    }
    const double prev0 = -log(s_w[0]);
    const int nm = nb - 1;
    // A window of the whole path (2^n_level == M) can close a chain on itself:
    // its end, (label_after_wrap(ps[i + 1]), b0) in NEW mode or
    // (label_after_wrap(ps[i]), b0) in OLD mode, is its first bead.  The
    // reference's bead walks then stop at once (perm_bisect_iterative_class.h:163-185,
    // kinetic_class.h:111-118): a chain closed in NEW mode has no midpoint and
    // draws no normal, and a closed chain adds no kinetic link in that mode.
    // The list of open chains (NEW mode) reuses the walk's visited map.
    __shared__ int s_open;
    if (tid == 0) {
        int k = 0;
        for (int i = 0; i < n; ++i)
            if (nb != M || R.ch_lbl2[R.ch_nxt[i]] != R.ch_lbl[i]) vis[k++] = i;
        s_open = k;
    }
    // A proposal of at most kPermStage beads (a few chains) is bisected by the
    // first wavefront on its beads staged in LDS, [chain][t = 0..nb][d]; a
    // larger one by the whole workgroup on global memory.  Both do the same
    // arithmetic in the same order per bead, with the draws in the same order.
    const bool fast = n * (nb + 1) <= min(stage_max, kPermStage);
    __shared__ double st_old[kPermStage * 3], st_new[kPermStage * 3];
    // the LDS form's normals (normal g is normal(c_norm + g / 2, g & 1), 3 per
    // midpoint of the open chains in level order) and level uniforms, drawn by
    // the whole workgroup before the levels (device RNG)
    __shared__ double st_z[kPermStage * 3 + 1], st_u[kMaxLevel];
    const int per = (nb + 1) * 3;
    if (fast) {
        for (int idx = tid; idx < n * per; idx += kPermBlock) {
            const int i = idx / per, t = (idx % per) / 3, d = idx % 3;
            const double x = chain_old(R, b0, t, i, d);
            st_old[idx] = x;
            // NEW mode: the end of chain i is the old end of chain ch_nxt[i]
            st_new[idx] = t == nb ? chain_old(R, b0, nb, R.ch_nxt[i], d) : x;
        }
        if (!tape) {
            const int n_norm = 3 * n * (nb - 1);  // at least the open chains' count
            for (int q = tid; 2 * q < n_norm; q += kPermBlock) {
                double z0, z1;
                rng.normal2(c_norm + (uint32_t)q, &z0, &z1);
                st_z[2 * q] = z0;
                st_z[2 * q + 1] = z1;
            }
            for (int l = tid; l < n_lev; l += kPermBlock) st_u[l] = rng.unif(c_lev + (uint32_t)l, 0);
        }
    } else {
        // NEW positions start as the old ones
        for (int idx = tid; idx < n * nm; idx += kPermBlock) {
            const int i = idx / nm, t = idx % nm + 1;
            for (int d = 0; d < 3; ++d) R.ch_new[(size_t)idx * 3 + d] = chain_old(R, b0, t, i, d);
        }
    }
    __syncthreads();
    const int n_open = s_open;
    if (fast) {
        if (tid < kWave) {
            const int lane = tid;
            double prev = prev0;
            uint32_t g_base = 0;
            int rej = 0;
            for (int level = n_lev - 1; level >= 0 && !rej; --level) {
                const int skip = 1 << level;
                const int nmid = nb >> (level + 1);
                const double level_tau = R.tau * skip;
                const double sigma = sqrt(R.lambda * level_tau);
                const double i4lt_b = 1. / (4. * R.lambda * (0.5 * R.tau * (double)skip));
                double lspo = 0., lspn = 0.;
                for (int it = lane; it < n_open * nmid; it += kWave) {
#pragma clang fp contract(off)
                    const int i = vis[it / nmid], k = it % nmid;
                    const int a = 2 * k * skip, b = a + skip, c = a + 2 * skip;
                    const double* so = st_old + i * per;
                    double* sn = st_new + i * per;
                    double t[3], rbar[3], dold[3], dnew[3], nbv[3];
                    for (int d = 0; d < 3; ++d) {
                        const double xa = so[3 * a + d];
                        t[d] = minimg(so[3 * c + d] - xa, R.L, R.iL);
                        rbar[d] = xa + 0.5 * t[d];
                        dold[d] = minimg(so[3 * b + d] - rbar[d], R.L, R.iL);
                    }
                    for (int d = 0; d < 3; ++d) {
                        double z;
                        if (tape)
                            z = tape_get(R, cur + 3 * it + d, &bad);
                        else
                            z = st_z[g_base + 3u * (uint32_t)it + (uint32_t)d];
                        dnew[d] = sigma * z;
                        dnew[d] = minimg(dnew[d], R.L, R.iL);
                        const double xa = sn[3 * a + d];
                        t[d] = minimg(sn[3 * c + d] - xa, R.L, R.iL);
                        rbar[d] = xa + 0.5 * t[d];
                        nbv[d] = rbar[d] + dnew[d];
                    }
                    for (int d = 0; d < 3; ++d) sn[3 * b + d] = nbv[d];
                    lspo -= free_neg_log(&R.bis[level], R.bis_images, i4lt_b, dold);
                    lspn -= free_neg_log(&R.bis[level], R.bis_images, i4lt_b, dnew);
                }
                wave_lds_sync();
                const int nlinks = nb >> level;
                const double i4lt_k = 1. / (4. * R.lambda * level_tau);
                double ko = 0., kn = 0.;
                for (int it = lane; it < n * nlinks; it += kWave) {
                    const int i = it / nlinks, a = (it % nlinks) * skip, b = a + skip;
                    const bool ring = nb == M;
                    const double* so = st_old + i * per;
                    const double* sn = st_new + i * per;
                    double dd[3];
                    if (!ring || R.ch_lbl2[i] != R.ch_lbl[i]) {
                        for (int d = 0; d < 3; ++d) dd[d] = minimg(so[3 * a + d] - so[3 * b + d], R.L, R.iL);
                        ko += free_neg_log(&R.kin[level], R.kin_images, i4lt_k, dd);
                    }
                    if (!ring || R.ch_lbl2[R.ch_nxt[i]] != R.ch_lbl[i]) {
                        for (int d = 0; d < 3; ++d) dd[d] = minimg(sn[3 * a + d] - sn[3 * b + d], R.L, R.iL);
                        kn += free_neg_log(&R.kin[level], R.kin_images, i4lt_k, dd);
                    }
                }
                lspo = wave_sum(lspo);
                lspn = wave_sum(lspn);
                ko = wave_sum(ko);
                kn = wave_sum(kn);
                bad = __any(bad) ? 1 : 0;  // a tape overrun in any lane's draws
                cur += 3 * n_open * nmid;
                g_base += 3u * (uint32_t)(n_open * nmid);
                double change = 0.;
                if (lane == 0) {
                    const double u = tape ? tape_get(R, cur, &bad) : st_u[level];
                    rej = perm_level_step(R, pr, level, lspo, lspn, ko, kn, prev, u, bad, b0, nb, n, n_lev, cur, &s_i[7],
                                          &change);
                }
                rej = __shfl(rej, 0, kWave);
                prev = __shfl(change, 0, kWave);
                cur += 1;
                wave_lds_sync();
            }
            // the NEW beads for the round's other kernels and the decision
            if (!rej)
                for (int idx = lane; idx < n * nm; idx += kWave) {
                    const int i = idx / nm, t = idx % nm + 1;
                    for (int d = 0; d < 3; ++d) R.ch_new[(size_t)idx * 3 + d] = st_new[i * per + 3 * t + d];
                }
            if (lane == 0) s_i[6] = rej;
        }
        __syncthreads();
        if (s_i[6]) {
            for (int i = tid; i < n; i += kPermBlock) {
                R.mv_pre[R.ch_lbl[i]] = -1;
                R.mv_post[R.ch_lbl2[i]] = -1;
            }
            if (tid < kWave) rho_items_write(R, 0, 0, 0, 0, tid);
            return;
        }
    } else {
        double prev = prev0;
        uint32_t g_base = 0;
        for (int level = n_lev - 1; level >= 0; --level) {
            const int skip = 1 << level;
            const int nmid = nb >> (level + 1);
            const double level_tau = R.tau * skip;
            const double sigma = sqrt(R.lambda * level_tau);
            const double i4lt_b = 1. / (4. * R.lambda * (0.5 * R.tau * (double)skip));
            double lspo = 0., lspn = 0.;
            for (int it = tid; it < n_open * nmid; it += kPermBlock) {
#pragma clang fp contract(off)
                const int i = vis[it / nmid], k = it % nmid;
                const int a = 2 * k * skip, b = a + skip, c = a + 2 * skip;
                double t[3], rbar[3], dold[3], dnew[3], nbv[3];
                for (int d = 0; d < 3; ++d) {
                    const double xa = chain_old(R, b0, a, i, d);
                    t[d] = minimg(chain_old(R, b0, c, i, d) - xa, R.L, R.iL);
                    rbar[d] = xa + 0.5 * t[d];
                    dold[d] = minimg(chain_old(R, b0, b, i, d) - rbar[d], R.L, R.iL);
                }
                for (int d = 0; d < 3; ++d) {
                    double z;
                    if (tape) {
                        z = tape_get(R, cur + 3 * it + d, &bad);
                    } else {
                        const uint32_t g = g_base + 3u * (uint32_t)it + (uint32_t)d;
                        z = rng.normal(c_norm + g / 2u, (int)(g & 1u));
                    }
                    dnew[d] = sigma * z;
                    dnew[d] = minimg(dnew[d], R.L, R.iL);
                    const double xa = chain_new(R, b0, nb, n, a, i, d);
                    t[d] = minimg(chain_new(R, b0, nb, n, c, i, d) - xa, R.L, R.iL);
                    rbar[d] = xa + 0.5 * t[d];
                    nbv[d] = rbar[d] + dnew[d];
                }
                for (int d = 0; d < 3; ++d) R.ch_new[((size_t)i * nm + b - 1) * 3 + d] = nbv[d];
                lspo -= free_neg_log(&R.bis[level], R.bis_images, i4lt_b, dold);
                lspn -= free_neg_log(&R.bis[level], R.bis_images, i4lt_b, dnew);
            }
            __syncthreads();
            const int nlinks = nb >> level;
            const double i4lt_k = 1. / (4. * R.lambda * level_tau);
            double ko = 0., kn = 0.;
            for (int it = tid; it < n * nlinks; it += kPermBlock) {
                const int i = it / nlinks, a = (it % nlinks) * skip, b = a + skip;
                const bool ring = nb == M;
                double dd[3];
                if (!ring || R.ch_lbl2[i] != R.ch_lbl[i]) {
                    for (int d = 0; d < 3; ++d)
                        dd[d] = minimg(chain_old(R, b0, a, i, d) - chain_old(R, b0, b, i, d), R.L, R.iL);
                    ko += free_neg_log(&R.kin[level], R.kin_images, i4lt_k, dd);
                }
                if (!ring || R.ch_lbl2[R.ch_nxt[i]] != R.ch_lbl[i]) {
                    for (int d = 0; d < 3; ++d)
                        dd[d] = minimg(chain_new(R, b0, nb, n, a, i, d) - chain_new(R, b0, nb, n, b, i, d), R.L, R.iL);
                    kn += free_neg_log(&R.kin[level], R.kin_images, i4lt_k, dd);
                }
            }
            lspo = block_sum<kPermBlock>(lspo, red);
            lspn = block_sum<kPermBlock>(lspn, red);
            ko = block_sum<kPermBlock>(ko, red);
            kn = block_sum<kPermBlock>(kn, red);
            bad = __syncthreads_or(bad);  // a tape overrun in any thread's draws
            cur += 3 * n_open * nmid;
            g_base += 3u * (uint32_t)(n_open * nmid);
            if (tid == 0) {
                const double u = tape ? tape_get(R, cur, &bad) : rng.unif(c_lev + (uint32_t)level, 0);
                double change = 0.;
                s_i[6] = perm_level_step(R, pr, level, lspo, lspn, ko, kn, prev, u, bad, b0, nb, n, n_lev, cur, &s_i[7],
                                         &change);
                s_w[1] = change;
            }
            cur += 1;
            __syncthreads();
            if (s_i[6]) {
                for (int i = tid; i < n; i += kPermBlock) {
                    R.mv_pre[R.ch_lbl[i]] = -1;
                    R.mv_post[R.ch_lbl2[i]] = -1;
                }
                if (tid < kWave) rho_items_write(R, 0, 0, 0, 0, tid);
                return;
            }
            prev = s_w[1];
        }
    }
    // Level 0 reached.  A one-chain proposal (a 1-cycle: plain bisection of p0,
    // whose label may change at the wrap) goes to the kernels of the Boltzmann
    // bisection: its beads in prop_new, its phase chains and the bisection's
    // rho_k items, and the LDS pair kernel unless its label changes in the
    // window.  Several chains go to k_pair_multi and k_rho_multi.
    const bool single = n == 1 && (!R.use_long_range || (R.ritems_bis && nm * R.n_chunks <= R.n_ritems_bis));
    if (single) {
        const int l1 = R.ch_lbl[0], l2 = R.ch_lbl2[0];
        for (int idx = tid; idx < nm * 3; idx += kPermBlock) R.prop_new[idx] = R.ch_new[idx];
        if (R.use_long_range)
            for (int t = tid; t < 6 * nm; t += kPermBlock) {
                const int m = t / 6;
                double xo[3], xn[3];
                for (int d = 0; d < 3; ++d) {
                    xo[d] = chain_old(R, b0, m + 1, 0, d);
                    xn[d] = R.ch_new[(size_t)m * 3 + d];
                }
                moved_phase_chain(R, m, t % 6, xo, xn);
            }
        if (tid == 0) {
            const int gen = !tuned || ((l2 != l1) && (b0 + nb > M));
            pr->generic = gen;
            pr->rho_multi = 0;
            if (gen && s_i[7] == 1 && R.multi_cnt) R.multi_list[atomicAdd(&R.multi_cnt[0], 1u)] = r;
        }
        if (tid < kWave) rho_items_write(R, 0, s_i[7], nm, (b0 + 1) % M, tid);
    } else {
        if (tid == 0) {
            pr->rho_multi = 1;
            // the label-general kernels of this round take the proposal
            if (s_i[7] == 1 && R.multi_cnt) {
                R.multi_list[atomicAdd(&R.multi_cnt[0], 1u)] = r;
                if (R.use_long_range) R.multi_list[R.multi_cap + atomicAdd(&R.multi_cnt[1], 1u)] = r;
            }
        }
        if (tid < kWave) rho_items_write(R, 0, 0, 0, 0, tid);
        if (R.use_long_range)
            for (int m = tid; m < nm; m += kPermBlock) moved_row(R, m, (b0 + 1 + m) % M);
    }
}

// ---------------------------------------------------------------------------
// Label-based pair action change of a proposal whose moved beads carry several
// labels (PairAction::GetAction, pair_action_class.h:267-302, with the particle
// list of Bisect::Attempt, bisect_class.h:60-63, or
// PermBisectIterative::Attempt, perm_bisect_iterative_class.h:133-146).  The
// reference sums every pair of its particle list over every window link; the
// pairs and links without a moved bead are the same in OLD and NEW and are
// left out here.  On link j the labels with a moved bead on it are listed
// (chains before the wrap, then chains after it not already listed); each
// listed label is paired with every particle not listed before it.  A bead's
// NEW position is the proposal's where it moved, else its committed one.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int pre_chain(const Rep& R, const Prop& pr, int y) {
    return pr.move == 3 ? R.mv_pre[y] : (y == pr.p ? 0 : -1);
}
__device__ __forceinline__ int post_chain(const Rep& R, const Prop& pr, int y) {
    return pr.move == 3 ? R.mv_post[y] : (y == pr.p2 ? 0 : -1);
}
// chain whose moved bead label y is at window position t, or -1
__device__ __forceinline__ int moved_chain(const Rep& R, const Prop& pr, int nb, int y, int t) {
    if (t < 1 || t > nb - 1) return -1;
    return (pr.b0 + t >= R.M) ? post_chain(R, pr, y) : pre_chain(R, pr, y);
}
__device__ __forceinline__ double moved_new(const Rep& R, const Prop& pr, int nb, int c, int t, int d) {
    return pr.move == 3 ? R.ch_new[((size_t)c * (nb - 1) + t - 1) * 3 + d] : R.prop_new[3 * (t - 1) + d];
}

constexpr int kPairMultiBlock = 256;
// part of bpr workgroups of replica r: links part, part + bpr, ...
__device__ __forceinline__ void pair_multi_item(const Rep* __restrict__ reps, int r, int part, int bpr) {
    const Rep& R = reps[r];
    const Prop pr = *R.prop;
    if (pr.status != 1 || !pr.generic) return;
    __shared__ double sh_o[kPairMultiBlock / kWave], sh_n[kPairMultiBlock / kWave];
    const int M = R.M, N = R.N, nb = pr.count + 1, b0 = pr.b0;
    const int nch = pr.move == 3 ? pr.n_chain : 1;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    constexpr int kWaves = kPairMultiBlock / kWave;
    for (int j = part; j < pr.n_links; j += bpr) {
        const int t0 = j, t1 = j + 1;
        const int s0 = (b0 + j) % M, s1 = (s0 + 1) % M;
        const bool post0 = b0 + t0 >= M, post1 = b0 + t1 >= M;
        const bool in0 = t0 >= 1 && t0 <= nb - 1, in1 = t1 <= nb - 1;
        const bool pre_on = (in0 && !post0) || (in1 && !post1);
        const bool post_on = (in0 && post0) || (in1 && post1);
        const int n_items = (pre_on ? nch : 0) + (post_on ? nch : 0);
        double uo = 0., un = 0.;
        // the link's listed labels one after the other; the workgroup's threads
        // share the partners of each
        for (int it = 0; it < n_items; ++it) {
            const bool is_post = !pre_on || it >= nch;
            const int ci = it < nch ? it : it - nch;
            int x;
            if (pr.move == 3)
                x = is_post ? R.ch_lbl2[ci] : R.ch_lbl[ci];
            else
                x = is_post ? pr.p2 : pr.p;
            if (is_post && pre_on && pre_chain(R, pr, x) >= 0) continue;  // listed before the wrap already
            double xo0[3], xo1[3], xn0[3], xn1[3];
            const int cx0 = moved_chain(R, pr, nb, x, t0), cx1 = moved_chain(R, pr, nb, x, t1);
            for (int d = 0; d < 3; ++d) {
                xo0[d] = pos_at(R, d, s0, x);
                xo1[d] = pos_at(R, d, s1, x);
                xn0[d] = cx0 >= 0 ? moved_new(R, pr, nb, cx0, t0, d) : xo0[d];
                xn1[d] = cx1 >= 0 ? moved_new(R, pr, nb, cx1, t1, d) : xo1[d];
            }
            for (int q = threadIdx.x; q < N; q += kPairMultiBlock) {
                if (q == x) continue;
                // a particle listed on this link before x was paired with x already
                int iq = 1 << 30;
                const int qa = pre_on ? pre_chain(R, pr, q) : -1;
                if (qa >= 0) {
                    iq = qa;
                } else if (post_on) {
                    const int qb = post_chain(R, pr, q);
                    if (qb >= 0) iq = (pre_on ? nch : 0) + qb;
                }
                if (iq < it) continue;
                const int cq0 = moved_chain(R, pr, nb, q, t0), cq1 = moved_chain(R, pr, nb, q, t1);
                double qo0[3], qo1[3], qn0[3], qn1[3];
                for (int d = 0; d < 3; ++d) {
                    qo0[d] = pos_at(R, d, s0, q);
                    qo1[d] = pos_at(R, d, s1, q);
                    qn0[d] = cq0 >= 0 ? moved_new(R, pr, nb, cq0, t0, d) : qo0[d];
                    qn1[d] = cq1 >= 0 ? moved_new(R, pr, nb, cq1, t1, d) : qo1[d];
                }
                LinkGeom g;
                link_geom(xo0, qo0, xo1, qo1, R.L, R.iL, g);
                uo += pair_u(R, g.r, g.rp, g.s);
                link_geom(xn0, qn0, xn1, qn1, R.L, R.iL, g);
                un += pair_u(R, g.r, g.rp, g.s);
            }
        }
        uo = wave_sum(uo);
        un = wave_sum(un);
        if (lane == 0) {
            sh_o[w] = uo;
            sh_n[w] = un;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0., b = 0.;
            for (int i = 0; i < kWaves; ++i) {
                a += sh_o[i];
                b += sh_n[i];
            }
            R.pair_part[j] = make_double2(a, b);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kPairMultiBlock) void k_pair_multi(const Rep* __restrict__ reps, int n_reps, int bpr) {
    const int r = blockIdx.x / bpr;
    if (r < n_reps) pair_multi_item(reps, r, blockIdx.x % bpr, bpr);
}
// The same over the replicas a permuting round listed (cnt[0] of them in
// list[0..]): a fixed grid loops over their (replica, part) items.
__global__ __launch_bounds__(kPairMultiBlock) void k_pair_multi_list(const Rep* __restrict__ reps,
                                                                     const int* __restrict__ list,
                                                                     const unsigned int* __restrict__ cnt, int bpr) {
    const int n = (int)cnt[0] * bpr;
    for (int it = blockIdx.x; it < n; it += gridDim.x) pair_multi_item(reps, list[it / bpr], it % bpr, bpr);
}

// ---------------------------------------------------------------------------
// Species::UpdateRhoK (species_class.h:406-425) for a permuting proposal: the
// slice of moved slot m holds one moved bead per chain; its rho_k row gets the
// sum of their phase changes, fused with the long-range sums of CalcULong like
// k_rho.  The beads' phase chains (KSpace::CalcC recurrence) are built in LDS a
// few chains at a time.  Grid (k chunk, slot, replica).
// ---------------------------------------------------------------------------
constexpr int kPermPhaseBatch = 4;
// KPT: k-vectors per thread, the rho_kpt of the replicas' k chunks
template <int KPT>
__device__ __forceinline__ void rho_multi_item(const Rep* __restrict__ reps, int r, int m, int chunk) {
    const Rep& R = reps[r];
    const Prop pr = *R.prop;
    if (pr.status != 1 || pr.move != 3 || !pr.rho_multi || !R.use_long_range) return;
    if (m >= pr.count || chunk >= R.n_chunks) return;
    __shared__ double2 tab[kPermPhaseBatch * 6 * (kMaxIdx + 1)];
    __shared__ double red[kRhoBlock / kWave];
    const int M = R.M, nk = R.nk, stride = R.max_idx + 1, n = pr.n_chain, nb = pr.count + 1;
    const int2 row = R.rho_rows[m];
    const int s = row.x, cur = row.y;
    const double2* rc = R.rho + ((size_t)cur * M + s) * nk;
    double2* rn = R.rho + ((size_t)(1 - cur) * M + s) * nk;
    const int ck = kRhoBlock * KPT, k_lo = chunk * ck, k_hi = min(nk, k_lo + ck);
    // this thread's k-vectors: k_lo + threadIdx.x + i * kRhoBlock, i < KPT; the row
    // and k-table loads are issued before the phase chains are built
    double dx[KPT], dy[KPT];
    double2 rv[KPT];
    int kp[KPT];
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        dx[i] = dy[i] = 0.;
        const int k = k_lo + threadIdx.x + i * kRhoBlock;
        const bool ok = k < k_hi;
        rv[i] = ok ? ld2<true>(rc + k) : make_double2(0., 0.);
        kp[i] = ok ? gld(R.kidx + k) : 0;
    }
    for (int i0 = 0; i0 < n; i0 += kPermPhaseBatch) {
        const int nbat = min(kPermPhaseBatch, n - i0);
        __syncthreads();
        for (int c = threadIdx.x; c < nbat * 6; c += kRhoBlock) {
            const int i = i0 + c / 6, o = (c % 6) / 3, d = c % 3;
            const double x = o ? chain_new(R, pr.b0, nb, n, m + 1, i, d) : chain_old(R, pr.b0, m + 1, i, d);
            phase_chain(x * R.box, R.max_idx, tab + (size_t)c * stride);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < KPT; ++i) {
            const int k = k_lo + threadIdx.x + i * kRhoBlock;
            if (k >= k_hi) continue;
            for (int b = 0; b < nbat; ++b) {
                const double2 en = phase_k(tab + (size_t)(b * 6 + 3) * stride, stride, kp[i]);
                const double2 eo = phase_k(tab + (size_t)(b * 6) * stride, stride, kp[i]);
                dx[i] += en.x - eo.x;
                dy[i] += en.y - eo.y;
            }
        }
    }
    double ao = 0., an = 0.;
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        const int k = k_lo + threadIdx.x + i * kRhoBlock;
        if (k >= k_hi) continue;
        double2 nv;
        nv.x = rv[i].x + dx[i];
        nv.y = rv[i].y + dy[i];
        st2<true>(rn + k, nv);
        const double cf = gld(R.coef_s + k);
        ao += cf * (rv[i].x * rv[i].x + rv[i].y * rv[i].y);
        an += cf * (nv.x * nv.x + nv.y * nv.y);
    }
    ao = block_sum<kRhoBlock>(ao, red);
    an = block_sum<kRhoBlock>(an, red);
    if (threadIdx.x == 0) R.rho_part[m * R.n_chunks + chunk] = make_double2(ao, an);
    __syncthreads();  // the LDS tables and partials are reused by the block's next item
}
template <int KPT>
__global__ __launch_bounds__(kRhoBlock) void k_rho_multi(const Rep* __restrict__ reps) {
    rho_multi_item<KPT>(reps, blockIdx.z, blockIdx.y, blockIdx.x);
}
// The same over the replicas a permuting round listed (cnt[1] of them in
// list[0..]): a fixed grid loops over their (replica, slot, chunk) items.
template <int KPT>
__global__ __launch_bounds__(kRhoBlock) void k_rho_multi_list(const Rep* __restrict__ reps, const int* __restrict__ list,
                                                              const unsigned int* __restrict__ cnt, int slots,
                                                              int chunks) {
    const int per = slots * chunks, n = (int)cnt[1] * per;
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const int q = it % per;
        rho_multi_item<KPT>(reps, list[it / per], q / chunks, q % chunks);
    }
}

// k_decide for a permuting round, which also empties the group's lists for the
// next round
__global__ __launch_bounds__(64) void k_decide_perm(const Rep* __restrict__ reps, int n_reps,
                                                    unsigned int* __restrict__ multi_cnt) {
    decide_body(reps, n_reps, nullptr);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        multi_cnt[0] = 0u;
        multi_cnt[1] = 0u;
    }
}

}  // namespace spimc
