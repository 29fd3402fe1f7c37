// spimc_kernels.h -- HIP kernels of the engine (included once, by spimc_engine.hip).
//
// Move stages run in rounds: every replica with attempts left performs one
// attempt per round, through four launches
//   k_propose_*  one wavefront per replica: draws, bisection sampling of the
//                coarse levels with their kinetic Metropolis decisions, and the
//                level-0 NEW beads
//   k_pair       one wavefront per (replica, link): short-range pair action of
//                the moved particle against all others, OLD and NEW
//   k_rho        one workgroup per (replica, moved slice, k chunk): incremental
//                rho_k update into the spare buffer fused with the long-range
//                action sums, OLD and NEW
//   k_decide     one wavefront per replica: level-0 Metropolis decision, commit
// A single chain (the plain method) is the same launches with one replica.
#pragma once

#include "spimc_device.h"

namespace spimc {

constexpr int kPairBlock = 256;       // 4 wavefronts, one link each
constexpr int kRhoBlock = 256;
constexpr int kMaxIdx = 63;           // largest |n_d| of a k-vector
constexpr int kGrBins = 4096;

enum { WORK_PAIR_LINKS = 0, WORK_RHO_SLICEK = 1, WORK_COUNT = 2 };

// Per-replica estimator scratch: slots [slot][M] for slots 0..3, then the
// long-range k-chunk partials, then two results.
constexpr int kSlotKin = 0, kSlotPair = 1, kSlotCentroid = 2, kSlotVirial = 3;
__host__ __device__ inline size_t scratch_long(int M) { return (size_t)4 * M; }
__host__ __device__ inline size_t scratch_res(int M, int nk) { return (size_t)4 * M + (nk + 255) / 256 + 1; }
__host__ __device__ inline size_t scratch_size(int M, int nk) { return scratch_res(M, nk) + 4; }

// Blocks are dispatched round-robin over the 8 XCDs; this remap hands each XCD
// a contiguous range of logical blocks, so the items of one replica (sorted
// together) share an L2.
__device__ __forceinline__ int xcd_remap(int hw, int G) {
    const int X = 8;
    int q = G / X, rem = G % X;
    int xcd = hw % X, local = hw / X;
    return (xcd < rem) ? xcd * (q + 1) + local : rem * (q + 1) + (xcd - rem) * q + local;
}

// Loads through the global address space (replica arrays live in device
// memory; a generic pointer would compile to flat loads, which also count
// against the LDS wait counter).
template <class T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}

// Partial sums passed between the workgroups of one launch (k_round): 64-bit
// atomic stores and loads at agent scope, which the memory model performs at
// the device's coherence point, so a workgroup on another XCD reads them
// without any L2 write-back or invalidation of the caches.
__device__ __forceinline__ void coh_store2(double2* p, double2 v) {
    __hip_atomic_store(&p->x, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&p->y, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double2 coh_load2(const double2* p) {
    double2* q = const_cast<double2*>(p);
    return make_double2(__hip_atomic_load(&q->x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                        __hip_atomic_load(&q->y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// 16-byte global loads and stores of one rho_k value; NT marks them
// non-temporal (streamed through, not kept in the caches)
typedef double d2v __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ double2 ld2(const double2* p) {
    const __attribute__((address_space(1))) d2v* q = (const __attribute__((address_space(1))) d2v*)p;
    const d2v v = NT ? __builtin_nontemporal_load(q) : *q;
    return make_double2(v.x, v.y);
}
template <bool NT>
__device__ __forceinline__ void st2(double2* p, double2 v) {
    __attribute__((address_space(1))) d2v* q = (__attribute__((address_space(1))) d2v*)p;
    const d2v w = {v.x, v.y};
    if (NT)
        __builtin_nontemporal_store(w, q);
    else
        *q = w;
}

__device__ __forceinline__ double pos_at(const Rep& R, int d, int s, int p) {
    return gld(R.pos + ((size_t)d * R.M + s) * R.N + p);
}

// NEW-mode position of the proposal's particle at slice s
__device__ __forceinline__ double pos_new(const Rep& R, const Prop& pr, int d, int s) {
    int m = s - pr.first;
    if (m < 0) m += R.M;
    return (m < pr.count) ? R.prop_new[3 * m + d] : pos_at(R, d, s, pr.p);
}

__device__ __forceinline__ double tape_get(const Rep& R, long long i, int* bad) {
    if (i < 0 || i >= R.tape_len) {
        *bad = 1;
        return 0.5;
    }
    return R.tape[i];
}

// Phase tables of moved bead m (OLD position xo, NEW position xn) for k_rho:
// chain c = o * 3 + d of length max_idx + 1, by the KSpace::CalcC recurrence.
__device__ __forceinline__ void moved_phase_chain(const Rep& R, int m, int c, const double* xo, const double* xn) {
    const int stride = R.max_idx + 1, o = c / 3, d = c % 3;
    phase_chain((o ? xn[d] : xo[d]) * R.box, R.max_idx, R.phase + ((size_t)m * 6 + c) * stride);
}

// Row descriptor of moved bead m at slice s: k_rho streams rho[sel][s] -> rho[1-sel][s].
__device__ __forceinline__ void moved_row(const Rep& R, int m, int s) { R.rho_rows[m] = make_int2(s, R.sel[s]); }

// The rho_k work items of a replica's round, written by its proposal kernel
// (one wavefront): item t = (moved bead t / n_chunks, k chunk t % n_chunks),
// valid when the proposal goes to level 0 and moves that bead (bead m sits at
// slice first + m).  Idle items (no attempt left, rejected at a coarse level,
// fewer moved beads) are marked invalid.
__device__ __forceinline__ void rho_items_write(const Rep& R, int move, int status, int count, int first, int lane,
                                                int nthr = kWave) {
    RhoItem* items = move == 0 ? R.ritems_bis : R.ritems_dis;
    const int n = move == 0 ? R.n_ritems_bis : R.n_ritems_dis;
    if (!items) return;
    const int stride = R.max_idx + 1, nch = R.n_chunks, ck = kRhoBlock * R.rho_kpt;
    for (int t = lane; t < n; t += nthr) {
        const int m = t / nch, ch = t % nch;
        RhoItem d{};
        if (status == 1 && m < count) {
            const int s = (first + m) % R.M, cur = R.sel[s];
            d.rc = R.rho + ((size_t)cur * R.M + s) * R.nk;
            d.rn = R.rho + ((size_t)(1 - cur) * R.M + s) * R.nk;
            d.ph = R.phase + (size_t)m * 6 * stride;
            d.kidx = R.kidx;
            d.coef = R.coef_s;
            d.part = R.rho_part + m * nch + ch;
            d.k0 = ch * ck;
            d.k1 = min(R.nk, d.k0 + ck);
            d.stride = stride;
            d.valid = 1;
        }
        items[t] = d;
    }
}

// ---------------------------------------------------------------------------
// K3: rho_k of every slice (Species::InitRhoK, species_class.h:380-403);
// particle sums in particle order per k.
// ---------------------------------------------------------------------------
constexpr int kBuildTile = 16;
template <int KPT>
__global__ __launch_bounds__(256) void k_rho_build(const Rep* __restrict__ reps, int r) {
    const Rep& R = reps[r];
    const int b = blockIdx.x, chunk = blockIdx.y;
    const int mx = R.max_idx, stride = mx + 1;
    __shared__ double2 tab[kBuildTile * 3 * (kMaxIdx + 1)];
    double2 acc[KPT];
    int kk[KPT];
    for (int i = 0; i < KPT; ++i) {
        acc[i] = make_double2(0., 0.);
        int k = chunk * 256 * KPT + i * 256 + threadIdx.x;
        kk[i] = (k < R.nk) ? R.kidx[k] : -1;
    }
    for (int p0 = 0; p0 < R.N; p0 += kBuildTile) {
        __syncthreads();
        if (threadIdx.x < kBuildTile * 3) {
            int pp = threadIdx.x / 3, d = threadIdx.x % 3;
            if (p0 + pp < R.N) phase_chain(pos_at(R, d, b, p0 + pp) * R.box, mx, &tab[(pp * 3 + d) * stride]);
        }
        __syncthreads();
        const int np = min(kBuildTile, R.N - p0);
        for (int pp = 0; pp < np; ++pp)
            for (int i = 0; i < KPT; ++i)
                if (kk[i] >= 0) {
                    double2 e = phase_k(&tab[pp * 3 * stride], stride, kk[i]);
                    acc[i].x += e.x;
                    acc[i].y += e.y;
                }
    }
    double2* out = R.rho + ((size_t)R.sel[b] * R.M + b) * R.nk;
    for (int i = 0; i < KPT; ++i) {
        int k = chunk * 256 * KPT + i * 256 + threadIdx.x;
        if (k < R.nk) out[k] = acc[i];
    }
}

// ---------------------------------------------------------------------------
// K4: Bisect::Attempt down to level 0 (bisect_class.h:39-124), one wavefront
// per replica.  Coarse levels carry only the kinetic action (the pair action
// returns 0 above max_level = 0, pair_action_class.h:269).
// ---------------------------------------------------------------------------
// Synchronisation of the one-wavefront bodies below (the whole 64-thread
// workgroup of the round kernels).
struct BlockSync {
    __device__ void operator()() const { __syncthreads(); }
};
// Ordering point of the one-wavefront bodies when they run in one wavefront of
// a larger workgroup: the wavefront's memory operations complete, and no code
// moves across it.
struct WaveSync {
    __device__ void operator()() const {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
};

// LDS of one bisection proposal (one wavefront): the window's beads, OLD and
// NEW; the attempt's normals, 3 per midpoint in level order, and level
// uniforms (device RNG), drawn before the level loop across the lanes; the
// attempt's counter step and first draws.
struct BisectLds {
    double x_old[(1 << kMaxLevel) + 1][3];
    double x_new[(1 << kMaxLevel) + 1][3];
    double z[3 * ((1 << kMaxLevel) - 1) + 1];
    double u[kMaxLevel];
    int info[4];  // an attempt is left, p, b0, bad draw
    long long cur;
    unsigned long long seq;
};

// The counter step of an attempt (one lane): whether one is left, and its
// sequence number (the device RNG's counter).
__device__ __forceinline__ void bisect_take(const Rep& R, BisectLds& S) {
    Counters* c = R.cnt;
    const unsigned int rem = c->remaining;
    S.info[0] = rem > 0;
    if (rem > 0) {
        c->remaining = rem - 1;
        S.seq = c->seq;
        c->seq = S.seq + 1;
    }
}

// The window's beads of (p, b0) into S, and on the device RNG (draws) the
// attempt's normals and level uniforms, made while the loads are in flight
// (the window holds nb + 1 <= 65 beads: two per lane at most).
__device__ __forceinline__ void bisect_window(const Rep& R, int lane, BisectLds& S, int p, int p2, int b0,
                                              bool draws) {
    const int n = R.n_level, nb = 1 << n, M = R.M;
    double xw[2][3];
    for (int k = 0; k < 2; ++k) {
        const int j = lane + k * kWave;
        if (j > nb) break;
        const int s = (b0 + j) % M;
        const int lbl = (b0 + j >= M) ? p2 : p;
        for (int d = 0; d < 3; ++d) xw[k][d] = pos_at(R, d, s, lbl);
    }
    if (draws) {
        Rng rng{make_uint2((uint32_t)R.key, (uint32_t)(R.key >> 32)), 1u, S.seq};
        // normal g is normal(1 + n + g / 2, g & 1): one call per pair of normals
        const int n_norm = 3 * (nb - 1);
        for (int q = lane; 2 * q < n_norm; q += kWave) {
            double z0, z1;
            rng.normal2(1 + n + q, &z0, &z1);
            S.z[2 * q] = z0;
            S.z[2 * q + 1] = z1;
        }
        for (int l = lane; l < n; l += kWave) S.u[l] = rng.unif(1 + l, 0);
    }
    for (int k = 0; k < 2; ++k) {
        const int j = lane + k * kWave;
        if (j > nb) break;
        for (int d = 0; d < 3; ++d) {
            S.x_old[j][d] = xw[k][d];
            S.x_new[j][d] = xw[k][d];
        }
    }
}

// The start of an attempt on the device RNG, ahead of its proposal and by
// another wavefront than the one that finishes it (k_round): the counter
// step, particle and first bead, window, normals and level uniforms into S.
// The finishing wavefront reloads the window if the decision before it
// committed moves of the same particle.
template <class Sync>
__device__ __forceinline__ void bisect_start_ahead(const Rep& R, int lane, Sync sync, BisectLds& S) {
    if (lane == 0) bisect_take(R, S);
    sync();
    if (!S.info[0]) return;
    if (lane == 0) {
        // rng.index(0, 0, N) and rng.index(0, 1, M) from one call
        Rng rng{make_uint2((uint32_t)R.key, (uint32_t)(R.key >> 32)), 1u, S.seq};
        const uint4 w = rng.call(0);
        S.info[1] = (int)(((uint64_t)w.x * (uint32_t)R.N) >> 32);
        S.info[2] = (int)(((uint64_t)w.y * (uint32_t)R.M) >> 32);
        S.info[3] = 0;
        S.cur = 0;
    }
    sync();
    const int p = S.info[1];
    bisect_window(R, lane, S, p, label_after_wrap(R, p), S.info[2], true);
}

// Bisect::Attempt up to the level-0 decision, one wavefront (lane 0..63) on
// replica R; items: also write the round's rho_k work items.  ahead: the
// attempt was started by bisect_start_ahead (device RNG only).
template <class Sync>
__device__ __forceinline__ void propose_bisect_impl(const Rep& R, int lane, Sync sync, bool items, BisectLds& S,
                                                    bool ahead) {
    Prop* pr = R.prop;
    if (!ahead) {
        if (lane == 0) bisect_take(R, S);
        sync();
    }
    if (!S.info[0]) {
        if (lane == 0) pr->status = 0;
        if (items) rho_items_write(R, 0, 0, 0, 0, lane);
        return;
    }
    const int n = R.n_level, nb = 1 << n, N = R.N, M = R.M;
    const bool tape = R.tape != nullptr;
    int bad = 0;
    if (!ahead) {
        if (lane == 0) {
            int p, b0;
            long long cur = 0;
            if (tape) {
                cur = *R.tape_cursor;
                p = (int)tape_get(R, cur, &bad);
                b0 = (int)tape_get(R, cur + 1, &bad);
                cur += 2;
                if (p < 0 || p >= N || b0 < 0 || b0 >= M) {
                    bad = 1;
                    p = 0;
                    b0 = 0;
                }
            } else {
                // rng.index(0, 0, N) and rng.index(0, 1, M) from one call
                Rng rng{make_uint2((uint32_t)R.key, (uint32_t)(R.key >> 32)), 1u, S.seq};
                const uint4 w = rng.call(0);
                p = (int)(((uint64_t)w.x * (uint32_t)N) >> 32);
                b0 = (int)(((uint64_t)w.y * (uint32_t)M) >> 32);
            }
            S.info[1] = p;
            S.info[2] = b0;
            S.info[3] = bad;
            S.cur = cur;
        }
        sync();
    }
    const int p = S.info[1], b0 = S.info[2];
    // the chain walked from (p, b0) carries label p2 after the wrap
    const int p2 = label_after_wrap(R, p);
    // 2^n_level == M and the chain closes on itself: the window's last bead is
    // its first one, so the reference's bead walks visit no midpoint and no
    // kinetic link (bisect_class.h:51-58,78)
    const bool wraps = (nb % M) == 0 && p2 == p;
    if (!ahead) bisect_window(R, lane, S, p, p2, b0, !tape);
    sync();
    long long cur = S.cur;
    double prev = 0.;
    int g_base = 0;  // normals drawn so far (device RNG indexing)
    for (int level = n - 1; level >= 0; --level) {
        const int skip = 1 << level;
        const int nmid = wraps ? 0 : nb >> (level + 1);
        const double level_tau = R.tau * skip;
        const double sigma = sqrt(R.lambda * level_tau);
        const double i4lt_b = 1. / (4. * R.lambda * (0.5 * R.tau * (double)skip));
        double lspo = 0., lspn = 0.;
        if (lane < nmid) {
#pragma clang fp contract(off)
            const int a = 2 * lane * skip, b = a + skip, c = a + 2 * skip;
            double t[3], rbar[3], dold[3], dnew[3];
            for (int d = 0; d < 3; ++d) {
                t[d] = minimg(S.x_old[c][d] - S.x_old[a][d], R.L, R.iL);
                rbar[d] = S.x_old[a][d] + 0.5 * t[d];
                dold[d] = minimg(S.x_old[b][d] - rbar[d], R.L, R.iL);
            }
            for (int d = 0; d < 3; ++d) {
                double z;
                if (tape)
                    z = tape_get(R, cur + 3 * lane + d, &bad);
                else
                    z = S.z[g_base + 3 * lane + d];
                dnew[d] = sigma * z;
                dnew[d] = minimg(dnew[d], R.L, R.iL);
                t[d] = minimg(S.x_new[c][d] - S.x_new[a][d], R.L, R.iL);
                rbar[d] = S.x_new[a][d] + 0.5 * t[d];
                S.x_new[b][d] = rbar[d] + dnew[d];
            }
            lspo = -free_neg_log(&R.bis[level], R.bis_images, i4lt_b, dold);
            lspn = -free_neg_log(&R.bis[level], R.bis_images, i4lt_b, dnew);
        }
        sync();
        const int nlinks = wraps ? 0 : nb >> level;
        const double i4lt_k = 1. / (4. * R.lambda * level_tau);
        double ko = 0., kn = 0.;
        for (int i = lane; i < nlinks; i += kWave) {
            const int a = i * skip, b = a + skip;
            double dd[3];
            for (int d = 0; d < 3; ++d) dd[d] = minimg(S.x_old[a][d] - S.x_old[b][d], R.L, R.iL);
            ko += free_neg_log(&R.kin[level], R.kin_images, i4lt_k, dd);
            for (int d = 0; d < 3; ++d) dd[d] = minimg(S.x_new[a][d] - S.x_new[b][d], R.L, R.iL);
            kn += free_neg_log(&R.kin[level], R.kin_images, i4lt_k, dd);
        }
        lspo = wave_sum(lspo);
        lspn = wave_sum(lspn);
        ko = wave_sum(ko);
        kn = wave_sum(kn);
        const double lsr = -lspn + lspo;
        cur += 3 * nmid;
        g_base += 3 * nmid;
        double u;
        if (tape)
            u = tape_get(R, cur, &bad);
        else
            u = S.u[level];
        cur += 1;
        if (level > 0) {
            const double old_action = 0. + ko, new_action = 0. + kn;
            const double change = new_action - old_action;
            const double la = lsr - change + prev;
            if (la < log(u)) {
                if (lane == 0) {
                    pr->status = 2;
                    pr->move = 0;
                    pr->p = p;
                    pr->p2 = p2;
                    pr->generic = 0;
                    pr->b0 = b0;
                    pr->count = 0;
                    pr->tape_next = cur;
                }
                if (items) rho_items_write(R, 0, 2, 0, 0, lane);
                return;
            }
            prev = change;
        } else {
            const int any_bad = __any(bad);
            const int status = (any_bad || S.info[3]) ? 3 : 1;
            if (lane == 0) {
                pr->status = status;
                pr->move = 0;
                pr->p = p;
                pr->p2 = p2;
                // moved beads on both sides of the wrap under two labels: the
                // label-general pair kernel takes the proposal
                pr->generic = (p2 != p) && (b0 + nb > M);
                pr->b0 = b0;
                pr->n_links = R.pair_type ? nb : 0;  // no pair action: no links to evaluate
                pr->first = (b0 + 1) % M;
                pr->count = wraps ? 0 : nb - 1;
                pr->lsr = lsr;
                pr->k_old = ko;
                pr->k_new = kn;
                pr->prev = prev;
                pr->u0 = u;
                pr->tape_next = cur;
            }
            for (int j = lane; j < (wraps ? 0 : nb - 1); j += kWave)
                for (int d = 0; d < 3; ++d) R.prop_new[3 * j + d] = S.x_new[j + 1][d];
            if (R.use_long_range) {
                for (int t = lane; t < (wraps ? 0 : 6 * (nb - 1)); t += kWave)
                    moved_phase_chain(R, t / 6, t % 6, S.x_old[t / 6 + 1], S.x_new[t / 6 + 1]);
                if (items) rho_items_write(R, 0, status, wraps ? 0 : nb - 1, (b0 + 1) % M, lane);
            }
        }
    }
}

__device__ __forceinline__ void propose_bisect_body(const Rep* __restrict__ reps, int n_reps) {
    const int r = blockIdx.x;
    if (r >= n_reps) return;
    __shared__ BisectLds S;
    propose_bisect_impl(reps[r], (int)threadIdx.x, BlockSync{}, true, S, false);
}

// DisplaceParticle::Attempt (displace_particle_class.h:28-72): NEW beads only.
// Kinetic::GetAction(0, M, {p}) walks from (p, 0) to the bead M links later:
// the walk is empty when the particle's path closes on itself, else it holds
// the wrap link (p, M-1) -> (perm[p], 0), the one link of the walk a rigid shift
// changes.  The link into (p, 0) from (iperm[p], M-1), which the shift also
// changes, is outside the walk and not counted, as in the reference.
// lane: the thread's index among the nthr that run the proposal (its draws and
// record on thread 0; the moved slices' positions, phase chains and rho_k
// items over all nthr).
template <class Sync>
__device__ __forceinline__ void propose_displace_impl(const Rep& R, int lane, Sync sync, bool items,
                                                      int nthr = kWave) {
    __shared__ int s_act, s_p, s_status;
    __shared__ double s_dr[3];
    Prop* pr = R.prop;
    if (lane == 0) {
        Counters* c = R.cnt;
        unsigned int rem = c->remaining;
        s_act = rem > 0;
        if (rem > 0) {
            c->remaining = rem - 1;
            unsigned long long seq = c->seq;
            c->seq = seq + 1;
            int bad = 0, p;
            double u0;
            long long cur = 0;
            if (R.tape) {
                cur = *R.tape_cursor;
                p = (int)tape_get(R, cur, &bad);
                for (int d = 0; d < 3; ++d) s_dr[d] = tape_get(R, cur + 1 + d, &bad);
                u0 = tape_get(R, cur + 4, &bad);
                cur += 5;
                if (p < 0 || p >= R.N) {
                    bad = 1;
                    p = 0;
                }
            } else {
                Rng rng{make_uint2((uint32_t)R.key, (uint32_t)(R.key >> 32)), 2u, seq};
                p = (int)rng.index(0, 0, R.N);
                s_dr[0] = R.step * (rng.unif(1, 0) - 0.5);
                s_dr[1] = R.step * (rng.unif(1, 1) - 0.5);
                s_dr[2] = R.step * (rng.unif(2, 0) - 0.5);
                u0 = rng.unif(2, 1);
            }
            s_p = p;
            double k_old = 0., k_new = 0.;
            const int pn = label_after_wrap(R, p);
            if (pn != p) {
                const double i4lt = 1. / (4. * R.lambda * R.tau);
                double dd[3];
                for (int d = 0; d < 3; ++d) dd[d] = minimg(pos_at(R, d, R.M - 1, p) - pos_at(R, d, 0, pn), R.L, R.iL);
                k_old = free_neg_log(&R.kin[0], R.kin_images, i4lt, dd);
                for (int d = 0; d < 3; ++d)
                    dd[d] = minimg((pos_at(R, d, R.M - 1, p) + s_dr[d]) - pos_at(R, d, 0, pn), R.L, R.iL);
                k_new = free_neg_log(&R.kin[0], R.kin_images, i4lt, dd);
            }
            pr->status = bad ? 3 : 1;
            pr->move = 1;
            pr->p = p;
            pr->p2 = p;
            pr->generic = 0;
            pr->b0 = 0;
            pr->n_links = R.pair_type ? R.M : 0;
            pr->first = 0;
            pr->count = R.M;
            pr->lsr = 0.;
            pr->k_old = k_old;
            pr->k_new = k_new;
            pr->prev = 0.;
            pr->u0 = u0;
            pr->tape_next = cur;
            s_status = bad ? 3 : 1;
        } else {
            pr->status = 0;
        }
    }
    sync();
    if (!s_act) {
        if (items) rho_items_write(R, 1, 0, 0, 0, lane, nthr);
        return;
    }
    const int p = s_p;
    for (int s = lane; s < R.M; s += nthr)
        for (int d = 0; d < 3; ++d) R.prop_new[3 * s + d] = pos_at(R, d, s, p) + s_dr[d];
    if (R.use_long_range)
        for (int t = lane; t < 6 * R.M; t += nthr) {
            const int s = t / 6;
            double xo[3], xn[3];
            for (int d = 0; d < 3; ++d) {
                xo[d] = pos_at(R, d, s, p);
                xn[d] = xo[d] + s_dr[d];
            }
            moved_phase_chain(R, s, t % 6, xo, xn);
        }
    if (R.use_long_range) if (items) rho_items_write(R, 1, s_status, R.M, 0, lane, nthr);
}

__device__ __forceinline__ void propose_displace_body(const Rep* __restrict__ reps, int n_reps) {
    const int r = blockIdx.x;
    if (r >= n_reps) return;
    propose_displace_impl(reps[r], (int)threadIdx.x, BlockSync{}, true);
}

// Phase tables of an interface proposal's moved beads (before k_rho in interface mode).
__global__ __launch_bounds__(64) void k_phase_proposal(const Rep* __restrict__ reps, int r) {
    const Rep& R = reps[r];
    const Prop pr = *R.prop;
    for (int t = blockIdx.x * 64 + threadIdx.x; t < 6 * pr.count; t += gridDim.x * 64) {
        const int m = t / 6, s = (pr.first + m) % R.M;
        double xo[3], xn[3];
        for (int d = 0; d < 3; ++d) {
            xo[d] = pos_at(R, d, s, pr.p);
            xn[d] = R.prop_new[3 * m + d];
        }
        moved_phase_chain(R, m, t % 6, xo, xn);
        if (t % 6 == 0) moved_row(R, m, s);
    }
}

// ---------------------------------------------------------------------------
// K1: short-range pair action of the moved particle on the window links,
// OLD and NEW (pair_action_class.h:279-290), one wavefront per (replica, link).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pair_block(const Rep* __restrict__ reps, const uint2* __restrict__ items, int n_items,
                                           int interface_mode, int blk) {
    const int item = blk * (kPairBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (item >= n_items) return;
    const uint2 it = items[item];
    const Rep& R = reps[it.x];
    const Prop pr = *R.prop;
    if (!interface_mode && (pr.status != 1 || pr.generic)) return;
    const int j = (int)it.y;
    if (j >= pr.n_links) return;
    const int M = R.M, N = R.N, p = pr.p;
    const int s0 = (pr.b0 + j) % M, s1 = (s0 + 1) % M;
    double po0[3], po1[3], pn0[3], pn1[3];
    for (int d = 0; d < 3; ++d) {
        po0[d] = pos_at(R, d, s0, p);
        po1[d] = pos_at(R, d, s1, p);
        pn0[d] = pos_new(R, pr, d, s0);
        pn1[d] = pos_new(R, pr, d, s1);
    }
    double uo = 0., un = 0.;
    for (int q = lane; q < N; q += kWave) {
        if (q == p) continue;
        double q0[3], q1[3];
        for (int d = 0; d < 3; ++d) {
            q0[d] = pos_at(R, d, s0, q);
            q1[d] = pos_at(R, d, s1, q);
        }
        LinkGeom g;
        link_geom(po0, q0, po1, q1, R.L, R.iL, g);
        uo += pair_u(R, g.r, g.rp, g.s);
        link_geom(pn0, q0, pn1, q1, R.L, R.iL, g);
        un += pair_u(R, g.r, g.rp, g.s);
    }
    uo = wave_sum(uo);
    un = wave_sum(un);
    if (lane == 0) {
        R.pair_part[j] = make_double2(uo, un);
    }
}

// Device-side kernel clocks.  Block 0 of a timed launch (the first dispatched)
// stores its start on the constant-rate wall clock; every block stores its end
// in a slot of its own (plain stores; grids above KT_SLOTS blocks fold several
// blocks into a slot with atomicMax).  The round's decide kernel (its first
// wavefront) takes the latest end of each class, adds end - start to a running
// sum and clears the slots: a launch is timed from its first block's start to
// its last block's end, whatever runs beside it on other streams.
// kt[c]: open start of class c (0 = none); kt[2 + 2c], kt[3 + 2c]: summed
// ticks and launches; kt[KT_END + c * KT_SLOTS + slot]: block ends.
enum { KT_RHO = 0, KT_PAIR = 1, KT_CLASSES = 2, KT_SLOTS = 1024, KT_END = 8, KT_WORDS = KT_END + KT_CLASSES * KT_SLOTS };
__device__ __forceinline__ void kt_begin(unsigned long long* kt, int cls) {
    if (kt && blockIdx.x == 0 && threadIdx.x == 0) kt[cls] = (unsigned long long)wall_clock64();
}
__device__ __forceinline__ void kt_end(unsigned long long* kt, int cls) {
    if (!kt) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long* slot = kt + KT_END + cls * KT_SLOTS + blockIdx.x % KT_SLOTS;
        const unsigned long long now = (unsigned long long)wall_clock64();
        if (gridDim.x <= KT_SLOTS)
            *slot = now;
        else
            atomicMax(slot, now);
    }
}
// called by one full wavefront
__device__ __forceinline__ void kt_close(unsigned long long* kt) {
    const int lane = threadIdx.x & (kWave - 1);
    for (int c = 0; c < KT_CLASSES; ++c) {
        unsigned long long e = 0;
        for (int t = lane; t < KT_SLOTS; t += kWave) {
            unsigned long long* slot = kt + KT_END + c * KT_SLOTS + t;
            e = max(e, *slot);
            *slot = 0ull;
        }
        for (int o = kWave / 2; o > 0; o >>= 1) e = max(e, (unsigned long long)__shfl_xor(e, o));
        if (lane == 0) {
            const unsigned long long b = kt[c];
            if (b && e >= b) {
                kt[2 + 2 * c] += e - b;
                kt[3 + 2 * c] += 1ull;
            }
            kt[c] = 0ull;
        }
    }
}

// K1 with the link-action tables staged in LDS: one 1024-thread workgroup per
// replica loads its B-spline tables once, then each wavefront takes window links
// j = w, w + 16, ...; lanes run over the partners q.
constexpr int kPairLdsBlock = 1024;
// The window links j = j0, j0 + step, ... of a proposal, one wavefront: lanes
// run over the partners q; the OLD and NEW sums of link j go to pair_part[j]
// (COH: as agent-scope atomic stores).
// One lane's OLD and NEW sums of window link j over the partners q_lo + lane,
// q_lo + lane + 64, ... below q_hi (q_lo a multiple of 128).
template <int PT, int NORD>
__device__ __forceinline__ void pair_link_range(const Rep& R, const double* lds, const Prop& pr, int j, int q_lo,
                                                int q_hi, int lane, double& uo, double& un) {
    const int M = R.M, p = pr.p;
    const int s0 = (pr.b0 + j) % M, s1 = (s0 + 1) % M;
    double po0[3], po1[3], pn0[3], pn1[3];
    for (int d = 0; d < 3; ++d) {
        po0[d] = pos_at(R, d, s0, p);
        po1[d] = pos_at(R, d, s1, p);
        pn0[d] = pos_new(R, pr, d, s0);
        pn1[d] = pos_new(R, pr, d, s1);
    }
    uo = 0.;
    un = 0.;
    {
        const int N = q_hi;
        // two partners per lane per step: independent evaluation chains for the scheduler
        for (int q = q_lo + lane; q < N; q += 2 * kWave) {
            const int q2 = q + kWave;
            const bool ok1 = q != p, ok2 = (q2 < N) && (q2 != p);
            const int qq = (q2 < N) ? q2 : q;
            double a0[3], a1[3], b0[3], b1[3];
            for (int d = 0; d < 3; ++d) {
                a0[d] = pos_at(R, d, s0, q);
                a1[d] = pos_at(R, d, s1, q);
                b0[d] = pos_at(R, d, s0, qq);
                b1[d] = pos_at(R, d, s1, qq);
            }
            double r, rp, s2;
            link_rr_s2(po0, a0, po1, a1, R.L, R.iL, r, rp, s2);
            const double uo1 = pair_u_lds_s2<PT, NORD>(R, lds, r, rp, s2);
            link_rr_s2(pn0, a0, pn1, a1, R.L, R.iL, r, rp, s2);
            const double un1 = pair_u_lds_s2<PT, NORD>(R, lds, r, rp, s2);
            link_rr_s2(po0, b0, po1, b1, R.L, R.iL, r, rp, s2);
            const double uo2 = pair_u_lds_s2<PT, NORD>(R, lds, r, rp, s2);
            link_rr_s2(pn0, b0, pn1, b1, R.L, R.iL, r, rp, s2);
            const double un2 = pair_u_lds_s2<PT, NORD>(R, lds, r, rp, s2);
            uo += (ok1 ? uo1 : 0.) + (ok2 ? uo2 : 0.);
            un += (ok1 ? un1 : 0.) + (ok2 ? un2 : 0.);
        }
    }
}

template <int PT, int NORD, bool COH = false>
__device__ __forceinline__ void pair_links_lds(const Rep& R, const double* lds, const Prop& pr, int j0, int step,
                                               int lane) {
    for (int j = j0; j < pr.n_links; j += step) {
        double uo, un;
        pair_link_range<PT, NORD>(R, lds, pr, j, 0, R.N, lane, uo, un);
        uo = wave_sum(uo);
        un = wave_sum(un);
        if (lane == 0) {
            if (COH)
                coh_store2(R.pair_part + j, make_double2(uo, un));
            else
                R.pair_part[j] = make_double2(uo, un);
        }
    }
}

// Wavefronts per window link of a W-wavefront workgroup: a replica of more
// than 256 particles splits each link's partners over up to W wavefronts.
template <int W>
__device__ __forceinline__ int pair_split(int N) {
    return min(W, (N + 4 * kWave - 1) / (4 * kWave));
}

// The window links of a proposal on a W-wavefront workgroup (part of bpr for
// the replica), ns = pair_split<W>(N) wavefronts per link: wavefront w takes
// link jb + w / ns and partner range w % ns, ranges of q_len partners (a
// multiple of 128).  A link's sum adds its ranges' wavefront sums in range
// order; with ns = 1 it is pair_links_lds's.  Every wavefront of the
// workgroup calls it (it holds workgroup barriers); wpart: W entries of LDS.
template <int PT, int NORD, bool COH, int W>
__device__ __forceinline__ void pair_links_split(const Rep& R, const double* lds, const Prop& pr, int part, int bpr,
                                                 int w, int lane, double2* wpart) {
    const int ns = pair_split<W>(R.N), lpb = W / ns;
    const int q_len = ((R.N + ns - 1) / ns + 2 * kWave - 1) / (2 * kWave) * (2 * kWave);
    for (int jb = part * lpb; jb < pr.n_links; jb += bpr * lpb) {
        const int j = jb + w / ns, k = w % ns;
        double uo = 0., un = 0.;
        if (w < lpb * ns && j < pr.n_links) {
            pair_link_range<PT, NORD>(R, lds, pr, j, k * q_len, min(R.N, (k + 1) * q_len), lane, uo, un);
            uo = wave_sum(uo);
            un = wave_sum(un);
        }
        if (lane == 0) wpart[w] = make_double2(uo, un);
        __syncthreads();
        if ((int)threadIdx.x < lpb && jb + (int)threadIdx.x < pr.n_links) {
            double so = wpart[threadIdx.x * ns].x, sn = wpart[threadIdx.x * ns].y;
            for (int kk = 1; kk < ns; ++kk) {
                so += wpart[threadIdx.x * ns + kk].x;
                sn += wpart[threadIdx.x * ns + kk].y;
            }
            if (COH)
                coh_store2(R.pair_part + jb + threadIdx.x, make_double2(so, sn));
            else
                R.pair_part[jb + threadIdx.x] = make_double2(so, sn);
        }
        __syncthreads();
    }
}

// Copies values [t_lo, n) of the tables a | b (a holds the first na) into lds
// with nthr threads, 16 values per thread at a time, every load issued before
// the stores that wait on it (a large table then costs a few load latencies
// instead of one per value).
__device__ __forceinline__ void stage_batched(double* lds, const double* a, const double* b, int na, int n, int t_lo,
                                              int nthr) {
    for (int t0 = t_lo + (int)threadIdx.x; t0 < n; t0 += 16 * nthr) {
        double v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int t = t0 + k * nthr;
            v[k] = t < n ? (t < na ? a[t] : b[t - na]) : 0.;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int t = t0 + k * nthr;
            if (t < n) lds[t] = v[k];
        }
    }
}

template <int PT, int NORD, int BS>
__device__ __forceinline__ void pair_lds_block(const Rep* __restrict__ reps, int n_reps, int bpr) {
    extern __shared__ double lds[];
    // bpr workgroups per replica (when replicas are few): part takes every bpr-th
    // group of BS/64 links
    const int idx = xcd_remap(blockIdx.x, gridDim.x), r = idx / bpr, part = idx % bpr;
    if (r >= n_reps) return;
    const Rep R = reps[r];  // by value: the fields used stay in registers across the loops
    const Prop pr = *R.prop;
    if (pr.status != 1 || pr.generic) return;
    stage_batched(lds, R.bt_a.c, R.bt_b.c, R.bt_a.ncoef * R.bt_a.ncomp, R.lds_doubles, 0, BS);
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (part * (BS / kWave) >= pr.n_links) return;
    pair_links_lds<PT, NORD>(R, lds, pr, part * (BS / kWave) + w, bpr * (BS / kWave), lane);
}


// PT: 1 all replicas Bare, 2 all David, 0 mixed (type read per replica);
// NORD: the David expansion order when all replicas share it, else 0
// BS: workgroup size; with BS < 1024 a CU keeps room for other kernels' waves
// (at most 128 VGPRs per wave whatever BS is)
template <int PT, int NORD, int BS = kPairLdsBlock>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) void k_pair_lds(
    const Rep* __restrict__ reps, int n_reps, int bpr, unsigned long long* __restrict__ kt) {
    kt_begin(kt, KT_PAIR);
    pair_lds_block<PT, NORD, BS>(reps, n_reps, bpr);
    kt_end(kt, KT_PAIR);
}

__global__ __launch_bounds__(kPairBlock) void k_pair(const Rep* __restrict__ reps, const uint2* __restrict__ items,
                                                     int n_items, int interface_mode,
                                                     unsigned long long* __restrict__ kt) {
    kt_begin(kt, KT_PAIR);
    pair_block(reps, items, n_items, interface_mode, xcd_remap(blockIdx.x, gridDim.x));
    kt_end(kt, KT_PAIR);
}

// ---------------------------------------------------------------------------
// K2: incremental rho_k of the moved slices (Species::UpdateRhoK,
// species_class.h:406-425) written to the spare buffer, fused with the
// long-range action sums of CalcULong (bare_pair_action_class.h:153-172,
// david_pair_action_class.h:105-123) for OLD and NEW.
// ---------------------------------------------------------------------------
template <int KPT, bool NT = false>
__device__ __forceinline__ void rho_block(const Rep* __restrict__ reps, const uint2* __restrict__ items, int n_items,
                                          int interface_mode, int blk) {
    if (blk >= n_items) return;
    const uint2 it = items[blk];
    const Rep& R = reps[it.x];
    const Prop pr = *R.prop;
    if (!interface_mode && pr.status != 1) return;
    const int m = (int)(it.y >> 8), chunk = (int)(it.y & 255);
    if (m >= pr.count) return;
    const int M = R.M, nk = R.nk, stride = R.max_idx + 1;
    const int2 row = R.rho_rows[m];
    const int s = row.x, cur = row.y;
    __shared__ double2 tab[2 * 3 * (kMaxIdx + 1)];
    __shared__ double red[2][kRhoBlock / kWave];
    // the six phase chains of this moved bead (OLD x3, NEW x3)
    const double2* src = R.phase + (size_t)m * 6 * stride;
    for (int t = threadIdx.x; t < 6 * stride; t += kRhoBlock) tab[t] = src[t];
    const double2* __restrict__ rc = R.rho + ((size_t)cur * M + s) * nk;
    double2* __restrict__ rn = R.rho + ((size_t)(1 - cur) * M + s) * nk;
    const int k0 = chunk * kRhoBlock * KPT + threadIdx.x;
    // issue the rho_k and k-table loads before waiting on the tables
    double2 rv[KPT];
    int kp[KPT];
    double cf[KPT];
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        const int k = k0 + i * kRhoBlock;
        const bool ok = k < nk;
        rv[i] = ok ? ld2<NT>(rc + k) : make_double2(0., 0.);
        kp[i] = ok ? gld(R.kidx + k) : (128 | (128 << 8) | (128 << 16));
        cf[i] = ok ? gld(R.coef_s + k) : 0.;
    }
    __syncthreads();
    double ao = 0., an = 0.;
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        const int k = k0 + i * kRhoBlock;
        const double2 en = phase_k(&tab[3 * stride], stride, kp[i]);
        const double2 eo = phase_k(tab, stride, kp[i]);
        double2 nv;
        nv.x = rv[i].x + (en.x - eo.x);
        nv.y = rv[i].y + (en.y - eo.y);
        if (k < nk) st2<NT>(rn + k, nv);
        ao += cf[i] * (rv[i].x * rv[i].x + rv[i].y * rv[i].y);
        an += cf[i] * (nv.x * nv.x + nv.y * nv.y);
    }
    ao = wave_sum(ao);
    an = wave_sum(an);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) {
        red[0][w] = ao;
        red[1][w] = an;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double to = 0., tn = 0.;
        for (int i = 0; i < kRhoBlock / kWave; ++i) {
            to += red[0][i];
            tn += red[1][i];
        }
        R.rho_part[m * R.n_chunks + chunk] = make_double2(to, tn);
    }
}

template <int KPT, bool NT>
__global__ __launch_bounds__(kRhoBlock) void k_rho(const Rep* __restrict__ reps, const uint2* __restrict__ items,
                                                   int n_items, int interface_mode,
                                                   unsigned long long* __restrict__ kt) {
    kt_begin(kt, KT_RHO);
    rho_block<KPT, NT>(reps, items, n_items, interface_mode, xcd_remap(blockIdx.x, gridDim.x));
    kt_end(kt, KT_RHO);
}

// One rho_k work item on a whole kRhoBlock-thread workgroup: its rho_k, k-table
// and phase-chain loads are all issued before the first wait; the item's
// (OLD, NEW) sums go to *d.part.  The caller separates items that reuse tab and
// red with a workgroup barrier.  COH: the sums go out as agent-scope atomic
// stores.
template <int KPT, bool NT, bool COH = false>
__device__ __forceinline__ void rho_item(const RhoItem& d, double2* tab, double (*red)[kRhoBlock / kWave]) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int k0 = d.k0 + threadIdx.x;
    double2 rv[KPT];
    int kp[KPT];
    double cf[KPT];
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        const int k = k0 + i * kRhoBlock;
        const bool ok = k < d.k1;
        rv[i] = ok ? ld2<NT>(d.rc + k) : make_double2(0., 0.);
        kp[i] = ok ? gld(d.kidx + k) : (128 | (128 << 8) | (128 << 16));
        cf[i] = ok ? gld(d.coef + k) : 0.;
    }
    const int nt = 6 * d.stride, tx = (int)threadIdx.x;
    const double2 t0 = tx < nt ? ld2<false>(d.ph + tx) : make_double2(0., 0.);
    const double2 t1 = tx + kRhoBlock < nt ? ld2<false>(d.ph + tx + kRhoBlock) : make_double2(0., 0.);
    if (tx < nt) tab[tx] = t0;
    if (tx + kRhoBlock < nt) tab[tx + kRhoBlock] = t1;
    __syncthreads();
    double ao = 0., an = 0.;
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
        const int k = k0 + i * kRhoBlock;
        const double2 en = phase_k(&tab[3 * d.stride], d.stride, kp[i]);
        const double2 eo = phase_k(tab, d.stride, kp[i]);
        double2 nv;
        nv.x = rv[i].x + (en.x - eo.x);
        nv.y = rv[i].y + (en.y - eo.y);
        if (k < d.k1) st2<NT>(d.rn + k, nv);
        ao += cf[i] * (rv[i].x * rv[i].x + rv[i].y * rv[i].y);
        an += cf[i] * (nv.x * nv.x + nv.y * nv.y);
    }
    ao = wave_sum(ao);
    an = wave_sum(an);
    if (lane == 0) {
        red[0][w] = ao;
        red[1][w] = an;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double to = 0., tn = 0.;
        for (int i = 0; i < kRhoBlock / kWave; ++i) {
            to += red[0][i];
            tn += red[1][i];
        }
        if (COH)
            coh_store2(d.part, make_double2(to, tn));
        else
            *d.part = make_double2(to, tn);
    }
}

// K2 with a fixed number of workgroups that loop over the work items: the
// launch then holds at most that many workgroups per CU, so a concurrent pair
// kernel on another stream always finds room on every CU.
// Each item is one descriptor (scalar loads, prefetched while the previous
// item computes).
template <int KPT, bool NT>
__global__ __launch_bounds__(kRhoBlock) void k_rho_loop(const RhoItem* __restrict__ items, int n_items,
                                                        unsigned long long* __restrict__ kt) {
    kt_begin(kt, KT_RHO);
    __shared__ double2 tab[6 * (kMaxIdx + 1)];
    __shared__ double red[2][kRhoBlock / kWave];
    int it = blockIdx.x;
    RhoItem d;
    if (it < n_items) d = items[it];
    for (; it < n_items; it += gridDim.x) {
        RhoItem nd;
        const int nx = it + gridDim.x;
        if (nx < n_items) nd = items[nx];
        if (d.valid) {
            rho_item<KPT, NT>(d, tab, red);
            __syncthreads();  // tab and red are reused by the next item
        }
        d = nd;
    }
    kt_end(kt, KT_RHO);
}

// ---------------------------------------------------------------------------
// K5: level-0 Metropolis decision and commit (bisect_class.h:100-123,24-36,127-139;
// displace_particle_class.h:68-86), one wavefront per replica.
// ---------------------------------------------------------------------------
template <class Sync>
__device__ __forceinline__ void commit_impl(const Rep& R, const Prop& pr, int lane, int acc, const double* pre,
                                            unsigned char pre_sel, Sync sync);
// The counters, accept flag and action trace of one decided attempt (one lane):
// status as decided (3: replay error), po/pn/lo/ln the summed level-0 terms.
__device__ __forceinline__ void decide_book(const Rep& R, const Prop& pr, int acc, int status, double po, double pn,
                                            double lo, double ln) {
    Counters* c = R.cnt;
    if (pr.status == 1 && pr.move != 3) {  // the round's pair and rho_k work, for the roofline figures
        c->rho_slicek += (unsigned long long)pr.count * (unsigned long long)R.nk;
        if (R.pair_type) c->pair_links += 2ull * (unsigned long long)pr.n_links * (unsigned long long)(R.N - 1);
    }
    if (pr.move == 0) {
        c->bis_att += 1;
        c->bis_acc += acc;
        c->bead_updates += (1ull << R.n_level) - 1;
    } else if (pr.move == 3) {
        // Move::DoEvent counts every attempt; PermBisect counts per cycle length
        // the attempts that reached the bisection (perm_bisect_class.h:66-87)
        unsigned long long* pc = R.perm_cnt;
        pc[0] += 1;
        pc[1] += acc;
        if (pr.n_chain > 0) {
            pc[2 + pr.n_chain - 1] += 1;
            pc[2 + R.N + pr.n_chain - 1] += acc;
        }
        c->bead_updates += (unsigned long long)pr.n_chain * ((1ull << pr.n_level) - 1);
    } else {
        c->dis_att += 1;
        c->dis_acc += acc;
        c->bead_updates += R.M;
    }
    unsigned int si = c->stage_i;
    if (si < R.accepted_cap) R.accepted[si] = (unsigned char)acc;
    if (R.act_trace && si < R.accepted_cap) {
        // the terms of the level-0 decision as summed (NaN: no level-0 decision)
        double* t = R.act_trace + (size_t)si * kTraceWords;
        const bool l0 = pr.status == 1;
        const double nan = __builtin_nan("");
        t[0] = l0 ? pr.k_old : nan;
        t[1] = l0 ? pr.k_new : nan;
        t[2] = l0 ? po : nan;
        t[3] = l0 ? pn : nan;
        t[4] = l0 ? R.long_scale * lo : nan;
        t[5] = l0 ? R.long_scale * ln : nan;
        t[6] = l0 ? pr.lsr : nan;
        t[7] = l0 ? pr.prev : nan;
    }
    c->stage_i = si + 1;
    if (status == 3) {  // replay tape overrun / bad draw
        c->error = 1;
        atomicOr(R.ctx_error, 1u);
    }
}

// What decide_impl hands to a later decide_book when the caller defers it.
struct DecideOut {
    Prop pr;
    double po, pn, lo, ln;
    int acc, status;  // status 0: nothing was decided
};

// COH: the pair and rho_k sums are read with agent-scope atomic loads (they were
// written in the same launch, by workgroups on any XCD).  hint_links /
// hint_rho >= 0: the pair and rho_k sums the proposal is expected to hold,
// loaded before the proposal itself (summed again if it holds others).  out:
// the bookkeeping is left to the caller (decide_book), which receives what it
// needs there.  Returns the particle whose beads an accepted one-particle move
// committed, else -1.
template <bool COH = false, class Sync>
__device__ __forceinline__ int decide_impl(const Rep& R, int lane, Sync sync, int hint_links = -1, int hint_rho = -1,
                                           DecideOut* out = nullptr) {
    Prop* prp = R.prop;
    double po = 0., pn = 0., lo = 0., ln = 0.;
    if (hint_links >= 0) {
        for (int j = lane; j < hint_links; j += kWave) {
            const double2 v = COH ? coh_load2(R.pair_part + j) : R.pair_part[j];
            po += v.x;
            pn += v.y;
        }
        for (int j = lane; j < hint_rho; j += kWave) {
            const double2 v = COH ? coh_load2(R.rho_part + j) : R.rho_part[j];
            lo += v.x;
            ln += v.y;
        }
    }
    const Prop pr = *prp;
    if (pr.status == 0) return -1;
    // the commit's loads of a one-particle move, issued before the sums: the
    // lane's first new bead and the selector of its slice
    double pre[3] = {0., 0., 0.};
    unsigned char pre_sel = 0;
    if (pr.status == 1 && pr.move != 3 && lane < pr.count) {
        for (int d = 0; d < 3; ++d) pre[d] = R.prop_new[3 * lane + d];
        if (R.use_long_range) pre_sel = R.sel[(pr.first + lane) % R.M];
    }
    if (pr.status == 1) {
        const int np = R.use_long_range ? pr.count * R.n_chunks : 0;
        if (pr.n_links != hint_links || np != hint_rho) {
            po = pn = lo = ln = 0.;
            for (int j = lane; j < pr.n_links; j += kWave) {
                const double2 v = COH ? coh_load2(R.pair_part + j) : R.pair_part[j];
                po += v.x;
                pn += v.y;
            }
            for (int j = lane; j < np; j += kWave) {
                const double2 v = COH ? coh_load2(R.rho_part + j) : R.rho_part[j];
                lo += v.x;
                ln += v.y;
            }
        }
        po = wave_sum(po);
        pn = wave_sum(pn);
        lo = wave_sum(lo);
        ln = wave_sum(ln);
    }
    __shared__ int s_acc;
    int acc = 0, status = pr.status;
    long long tape_next = pr.tape_next;
    if (lane == 0) {
        if (pr.status == 1) {
            const double pair_old = po + R.long_scale * lo, pair_new = pn + R.long_scale * ln;
            const double old_action = (0. + pr.k_old) + pair_old;
            const double new_action = (0. + pr.k_new) + pair_new;
            const double change = new_action - old_action;
            const double la = pr.lsr - change + pr.prev;
            acc = !(la < log(pr.u0));
        }
        if (acc && pr.move == 3 && pr.tab_on) {
            // PermBisectTable: the table of the new path (perm_bisect_table_class.h:216-225)
            int bad = 0;
            const double u = R.tape ? tape_get(R, tape_next++, &bad) : pr.u_tab;
            if (bad) status = 3;
            acc = !(pr.tab_ratio < u);
        }
        s_acc = acc;
    }
    sync();
    commit_impl(R, pr, lane, s_acc, pre, pre_sel, sync);
    // bookkeeping after the commit's stores are issued
    if (lane == 0) {
        if (R.tape) *R.tape_cursor = tape_next;
        if (out) {
            out->pr = pr;
            out->acc = acc;
            out->status = status;
            out->po = po;
            out->pn = pn;
            out->lo = lo;
            out->ln = ln;
        } else {
            decide_book(R, pr, acc, status, po, pn, lo, ln);
        }
        prp->status = 0;
    }
    return (pr.status == 1 && s_acc && pr.move != 3) ? pr.p : -1;
}

// The commit of a decided proposal (part of decide_impl): Bisect::Accept /
// DisplaceParticle::Accept of its beads, or of every chain of a permuting
// proposal and its relabelling.  pre / pre_sel: the lane's first new bead and
// selector, loaded before the decision.
template <class Sync>
__device__ __forceinline__ void commit_impl(const Rep& R, const Prop& pr, int lane, int acc, const double* pre,
                                            unsigned char pre_sel, Sync sync) {
    const int M = R.M, N = R.N;
    const int s_acc = acc;
    if (s_acc && pr.move == 3) {
        // Bisect::Accept of every chain, then PermBisect::AssignParticleLabels
        // (perm_bisect_class.h:47-56): the slices after the relinked slice e take
        // the labels of the path they now continue, so chain label a_i there
        // holds what a_{ch_nxt[i]} held (each cycle of ch_nxt followed from its
        // first chain), and the wrap permutation follows.
        const int n = pr.n_chain, nm = pr.count, nb = nm + 1;
        for (int idx = lane; idx < n * nm; idx += kWave) {
            const int i = idx / nm, t = idx % nm + 1;
            const int s = (pr.b0 + t) % M, l = (pr.b0 + t >= M) ? R.ch_lbl2[i] : R.ch_lbl[i];
            for (int d = 0; d < 3; ++d) R.pos[((size_t)d * M + s) * N + l] = R.ch_new[(size_t)idx * 3 + d];
        }
        if (R.use_long_range)
            for (int m = lane; m < nm; m += kWave) R.sel[(pr.first + m) % M] ^= 1;
        sync();
        if (n > 1) {
            const bool wrapped = pr.b0 + nb - 1 >= M;
            const int e = pr.e_slice;
            const int* lbl = wrapped ? R.ch_lbl2 : R.ch_lbl;
            for (int idx = lane; idx < (M - 1 - e) * 3; idx += kWave) {
                const int s = e + 1 + idx / 3, d = idx % 3;
                double* col = R.pos + ((size_t)d * M + s) * N;
                for (int i0 = 0; i0 < n; ++i0) {
                    if (!R.ch_lead[i0]) continue;
                    const double keep = col[lbl[i0]];
                    int ai = lbl[i0];
                    for (int i = R.ch_nxt[i0]; i != i0; i = R.ch_nxt[i]) {
                        col[ai] = col[lbl[i]];
                        ai = lbl[i];
                    }
                    col[ai] = keep;
                }
            }
            if (lane == 0) {
                for (int i0 = 0; i0 < n; ++i0) {
                    if (!R.ch_lead[i0]) continue;
                    const int keep = R.perm[lbl[i0]];
                    int ai = lbl[i0];
                    for (int i = R.ch_nxt[i0]; i != i0; i = R.ch_nxt[i]) {
                        R.perm[ai] = R.perm[lbl[i]];
                        ai = lbl[i];
                    }
                    R.perm[ai] = keep;
                }
                for (int i = 0; i < n; ++i) {
                    const int a = wrapped ? R.ch_lbl2[i] : R.ch_lbl[i];
                    R.iperm[R.perm[a]] = a;
                }
            }
        }
    } else if (s_acc) {
        for (int m = lane; m < pr.count; m += kWave) {
            const int s = (pr.first + m) % M;
            // a bisection chain carries label p2 past the wrap
            const int l = (pr.move == 0 && pr.b0 + m + 1 >= M) ? pr.p2 : pr.p;
            const bool first = m == lane;
            for (int d = 0; d < 3; ++d) R.pos[((size_t)d * M + s) * N + l] = first ? pre[d] : R.prop_new[3 * m + d];
            if (R.use_long_range) R.sel[s] = (unsigned char)((first ? pre_sel : R.sel[s]) ^ 1);
        }
    }
    if (pr.move == 3 && pr.status == 1) {
        sync();
        for (int i = lane; i < pr.n_chain; i += kWave) {
            R.mv_pre[R.ch_lbl[i]] = -1;
            R.mv_post[R.ch_lbl2[i]] = -1;
        }
    }
}

__device__ __forceinline__ void decide_body(const Rep* __restrict__ reps, int n_reps,
                                            unsigned long long* __restrict__ kt) {
    if (kt && blockIdx.x == 0) kt_close(kt);
    const int r = blockIdx.x;
    if (r >= n_reps) return;
    decide_impl(reps[r], (int)threadIdx.x, BlockSync{});
}

__global__ __launch_bounds__(64) void k_propose_bisect(const Rep* __restrict__ reps, int n_reps) {
    propose_bisect_body(reps, n_reps);
}
__global__ __launch_bounds__(64) void k_propose_displace(const Rep* __restrict__ reps, int n_reps) {
    propose_displace_body(reps, n_reps);
}
__global__ __launch_bounds__(64) void k_decide(const Rep* __restrict__ reps, int n_reps,
                                               unsigned long long* __restrict__ kt) {
    decide_body(reps, n_reps, kt);
}
// K5 of one round and K4 of the next in one launch: a replica's pending
// proposal is decided and committed (nothing is pending at the first round of
// a stage), then its next proposal is made.  Saves one dependent launch per
// round; the stage's last decision is a k_decide of its own.  Two wavefronts
// per replica: the first decides while, on the device RNG of an unpermuted
// path, the second starts the next attempt (counter step, draws, window) and
// afterwards does the decision's bookkeeping, as in k_round.
constexpr int kBisBlock = 2 * kWave;
__global__ __launch_bounds__(kBisBlock) void k_decide_propose_bisect(const Rep* __restrict__ reps, int n_reps,
                                                                     unsigned long long* __restrict__ kt) {
    __shared__ BisectLds S;
    __shared__ DecideOut s_out;
    __shared__ int s_moved;
    const int r = blockIdx.x, w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    if (kt && blockIdx.x == 0 && w == 0) kt_close(kt);
    if (r >= n_reps) return;
    const Rep& R = reps[r];
    // a window that crosses the wrap of a permuted path holds a second label,
    // which the window check after the decision does not follow
    const bool ahead = R.tape == nullptr && R.perm == nullptr;
    if (w == 0) {
        if (lane == 0) s_out.status = 0;
        const int nb = 1 << R.n_level;
        const int moved = decide_impl(R, lane, WaveSync{}, R.pair_type ? nb : 0,
                                      R.use_long_range ? (nb - 1) * R.n_chunks : 0, &s_out);
        if (lane == 0) s_moved = moved;
    } else if (ahead) {
        bisect_start_ahead(R, lane, WaveSync{}, S);
    }
    __syncthreads();
    if (w == 1) {
        if (lane == 0 && s_out.status != 0)
            decide_book(R, s_out.pr, s_out.acc, s_out.status, s_out.po, s_out.pn, s_out.lo, s_out.ln);
        return;
    }
    if (ahead && S.info[0] && S.info[1] == s_moved) {
        // the decision just committed beads of the attempt's particle: its window again
        const int p = S.info[1];
        bisect_window(R, lane, S, p, p, S.info[2], false);
        WaveSync{}();
    }
    propose_bisect_impl(R, lane, WaveSync{}, true, S, ahead);
}
// Displacement: the decision on the first wavefront, then the proposal, whose
// M moved slices (positions, phase chains, rho_k items) take all kDisBlock
// threads.
constexpr int kDisBlock = 256;
__global__ __launch_bounds__(kDisBlock) void k_decide_propose_displace(const Rep* __restrict__ reps, int n_reps,
                                                                       unsigned long long* __restrict__ kt) {
    const int r = blockIdx.x;
    if (threadIdx.x < kWave) {
        if (kt && blockIdx.x == 0) kt_close(kt);
        if (r < n_reps) decide_impl(reps[r], (int)threadIdx.x, WaveSync{});
    }
    __syncthreads();
    if (r >= n_reps) return;
    propose_displace_impl(reps[r], (int)threadIdx.x, BlockSync{}, true, kDisBlock);
}


// A whole round (MOVE 0 bisection, 1 displacement) in one launch, for batches
// too small to fill the GPU (a round is then bound by latency, not by
// throughput).  Workgroups
// [0, n_reps * bpr) (after the XCD remap) run the window links of the LDS pair
// kernel, the others a group of up to ipb rho_k items of one replica each
// (rgrp: first item, items, replica; ipb > 1 only where one item per
// workgroup would exceed what the GPU holds at once).  Every workgroup then
// counts itself in on its replica's arrival counter; the last of the
// replica's bpr + ceil(n_items / ipb) workgroups decides and commits the replica's
// proposal (K5) and makes its next one (K4), with its rho_k items, in its
// first wavefront.  A replica's round thus ends with its own last workgroup:
// no launch boundary separates the sums from the decision, and nothing waits
// (the counter is only read by the atomic that increments it).  The sums pass
// between workgroups as agent-scope atomic stores and loads, each waited for
// before its workgroup arrives; agent-scope fences would instead write back a
// whole L2 per workgroup, which measured slower than three launches per round.
// The decision and proposal are the same code as in k_decide_propose_*, so the
// chains are the same bit for bit.  The last workgroup resets the counter for
// the next round.
constexpr int kRoundBlock = kRhoBlock;
template <int PT, int NORD, int KPT, bool NT, int MOVE>
__global__ __launch_bounds__(kRoundBlock) void k_round(const Rep* __restrict__ reps, int n_reps, int bpr,
                                                       const RhoItem* __restrict__ items,
                                                       const int4* __restrict__ rgrp, int ipb,
                                                       unsigned int* __restrict__ arrive) {
    extern __shared__ double lds[];
    __shared__ double2 tab[6 * (kMaxIdx + 1)];
    __shared__ double red[2][kRhoBlock / kWave];
    __shared__ double2 wpart[kRoundBlock / kWave];
    __shared__ int s_last;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int n_pair = n_reps * bpr, idx = xcd_remap(blockIdx.x, gridDim.x);
    int r;
    if (idx < n_pair) {
        r = idx / bpr;
        const int part = idx % bpr;
        const Rep R = reps[r];
        // the tables' first kRoundBlock * 4 values are loaded while the proposal is
        const int na = R.bt_a.ncoef * R.bt_a.ncomp, nd = R.lds_doubles;
        double tv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int t = threadIdx.x + k * kRoundBlock;
            tv[k] = t < nd ? (t < na ? R.bt_a.c[t] : R.bt_b.c[t - na]) : 0.;
        }
        const Prop pr = *R.prop;
        constexpr int W = kRoundBlock / kWave;
        if (pr.status == 1 && !pr.generic && part * (W / pair_split<W>(R.N)) < pr.n_links) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int t = threadIdx.x + k * kRoundBlock;
                if (t < nd) lds[t] = tv[k];
            }
            // the rest (large tables: 96 KB for David's)
            stage_batched(lds, R.bt_a.c, R.bt_b.c, na, nd, 4 * kRoundBlock, kRoundBlock);
            __syncthreads();
            pair_links_split<PT, NORD, true, W>(R, lds, pr, part, bpr, w, lane, wpart);
        }
    } else {
        // a group of up to ipb rho_k items of one replica, one after the other
        const int4 g = rgrp[idx - n_pair];  // first item, items, replica
        r = g.z;
        RhoItem d = items[g.x];
        for (int i = 0; i < g.y; ++i) {
            RhoItem nd;
            if (i + 1 < g.y) nd = items[g.x + i + 1];
            if (d.valid) {
                rho_item<KPT, NT, true>(d, tab, red);
                __syncthreads();  // tab and red are reused by the next item
            }
            d = nd;
        }
    }
    // every wavefront's sums are performed before the workgroup arrives
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const Rep& R = reps[r];
        const int n_items = MOVE == 0 ? R.n_ritems_bis : R.n_ritems_dis;
        const unsigned int expect = (unsigned int)bpr + (unsigned int)((n_items + ipb - 1) / ipb);
        s_last = atomicAdd(arrive + r, 1u) + 1u == expect;
    }
    __syncthreads();
    if (!s_last) return;
    // the replica's last workgroup: its first wavefront decides while, for a
    // bisection on the device RNG, its second starts the next attempt (the draws
    // depend on the attempt's sequence number only, not on the decision)
    __shared__ BisectLds S;
    __shared__ DecideOut s_out;
    __shared__ int s_moved;
    const Rep& R = reps[r];
    const bool ahead = MOVE == 0 && R.tape == nullptr;
    if (w == 0) {
        if (lane == 0) {
            __hip_atomic_store(arrive + r, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_out.status = 0;
        }
        // the pair and rho_k sums of a proposal that reached level 0 (bisection:
        // 2^l links and 2^l - 1 moved slices; displacement: M of each), loaded
        // before the proposal itself
        const int nl = MOVE == 0 ? 1 << R.n_level : R.M, nm = MOVE == 0 ? nl - 1 : R.M;
        const int moved = decide_impl<true>(R, lane, WaveSync{}, R.pair_type ? nl : 0,
                                            R.use_long_range ? nm * R.n_chunks : 0, &s_out);
        if (lane == 0) s_moved = moved;
    } else if (w == 1 && ahead) {
        bisect_start_ahead(R, lane, WaveSync{}, S);
    }
    __syncthreads();
    if (w == 2) {
        // the decision's counters and flags, off the proposal's path
        if (lane == 0 && s_out.status != 0)
            decide_book(R, s_out.pr, s_out.acc, s_out.status, s_out.po, s_out.pn, s_out.lo, s_out.ln);
        return;
    }
    if (w != 0) return;
    if (MOVE == 1) {
        propose_displace_impl(R, lane, WaveSync{}, true);
        return;
    }
    if (ahead && S.info[0] && S.info[1] == s_moved) {
        // the decision just committed beads of the attempt's particle: its window again
        const int p = S.info[1];
        bisect_window(R, lane, S, p, label_after_wrap(R, p), S.info[2], false);
        WaveSync{}();
    }
    propose_bisect_impl(R, lane, WaveSync{}, true, S, ahead);
}

// Sums the per-replica work counters into work[WORK_PAIR_LINKS], work[WORK_RHO_SLICEK].
__global__ __launch_bounds__(64) void k_work_sum(const Rep* __restrict__ reps, int n_reps,
                                                 unsigned long long* __restrict__ work) {
    unsigned long long a = 0, b = 0;
    for (int r = threadIdx.x; r < n_reps; r += kWave) {
        a += reps[r].cnt->pair_links;
        b += reps[r].cnt->rho_slicek;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    if (threadIdx.x == 0) {
        work[WORK_PAIR_LINKS] = a;
        work[WORK_RHO_SLICEK] = b;
    }
}

// Stage set-up on the device (kept out of hipMemset so captured graphs stay plain kernels)
__global__ void k_stage_begin(const Rep* __restrict__ reps, int n_reps, const unsigned int* __restrict__ counts) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_reps) return;
    reps[r].cnt->remaining = counts[r];
    reps[r].cnt->stage_i = 0;
}

// ---------------------------------------------------------------------------
// Interface kernels (Action::GetAction in a chosen mode)
// ---------------------------------------------------------------------------
// Kinetic::GetAction (kinetic_class.h:105-122) for the proposal's particle p.
__global__ __launch_bounds__(64) void k_kinetic_window(const Rep* __restrict__ reps, int r, int mode, int p, int b0,
                                                        int n_links, int level, double* __restrict__ out) {
    const Rep& R = reps[r];
    const Prop pr = *R.prop;
    const int skip = 1 << level, M = R.M;
    const double i4lt = 1. / (4. * R.lambda * (R.tau * (double)skip));
    double tot = 0.;
    // the walk from (p, b0) carries label perm[p] past the wrap; NEW mode reads
    // the proposal's beads where the label is the proposal's particle
    const int p2 = label_after_wrap(R, p);
    for (int i = threadIdx.x; i < n_links; i += kWave) {
        const int ta = i * skip, tb = ta + skip;
        const int a = (b0 + ta) % M, b = (b0 + tb) % M;
        const int la = (b0 + ta >= M) ? p2 : p, lb = (b0 + tb >= M) ? p2 : p;
        double dd[3];
        for (int d = 0; d < 3; ++d) {
            double xa = (mode && pr.p == la) ? pos_new(R, pr, d, a) : pos_at(R, d, a, la);
            double xb = (mode && pr.p == lb) ? pos_new(R, pr, d, b) : pos_at(R, d, b, lb);
            dd[d] = minimg(xa - xb, R.L, R.iL);
        }
        tot += free_neg_log(&R.kin[level], R.kin_images, i4lt, dd);
    }
    tot = wave_sum(tot);
    if (threadIdx.x == 0) *out = tot;
}

// Long-range sum of CalcULong over window slices in a mode; NEW mode reads the
// spare buffer for slices the proposal moved (after k_rho in interface mode).
__global__ __launch_bounds__(256) void k_long_window(const Rep* __restrict__ reps, int r, int mode, int b0,
                                                     int rho_updated, double* __restrict__ part) {
    const Rep& R = reps[r];
    const Prop pr = *R.prop;
    const int w = blockIdx.x, M = R.M;
    const int s = (b0 + w) % M;
    int m = s - pr.first;
    if (m < 0) m += M;
    const bool moved = mode && rho_updated && (m < pr.count);
    const int buf = moved ? 1 - R.sel[s] : R.sel[s];
    const double2* rb = R.rho + ((size_t)buf * M + s) * R.nk;
    double t = 0.;
    for (int k = threadIdx.x; k < R.nk; k += 256) {
        double2 v = rb[k];
        t += R.coef_s[k] * (v.x * v.x + v.y * v.y);
    }
    __shared__ double sh[4];
    t = block_sum<256>(t, sh);
    if (threadIdx.x == 0) part[w] = t;
}

// Work items of an interface query: (r, t) for the pair kernel's links
// (n_chunks == 0), or (r, moved bead << 8 | k chunk) for the rho_k kernel.
__global__ __launch_bounds__(256) void k_iface_items(uint2* __restrict__ items, int r, int n, int n_chunks) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    items[t] = make_uint2((unsigned)r, n_chunks ? (unsigned)(((t / n_chunks) << 8) | (t % n_chunks)) : (unsigned)t);
}

// Sums a pair action query: links (mode column) + long-range window partials.
// PairAction::GetActionLaplacian (pair_action_class.h:336-363) of particle p at
// level 0: for the window beads b0 <= b < b1, the links (b, b + 1) and
// (b, b - 1 + M) with every other particle, each through CalcLaplacianU
// (:167-199), the central second difference (eps 1e-4) of the link action in
// each coordinate of bead (p, b).  The long-range part is left out, as in the
// reference.  NEW mode reads the pending interface proposal's beads.  One
// workgroup; the thread sums are added in thread order.
constexpr int kLapBlock = 256;
__global__ __launch_bounds__(kLapBlock) void k_pair_laplacian(const Rep* __restrict__ reps, int r, int mode, int b0,
                                                              int b1, int p, double* __restrict__ out) {
    const Rep& R = reps[r];
    const Prop pr = *R.prop;
    const int M = R.M, N = R.N, nb = b1 - b0;
    const double eps = 1.e-4;
    __shared__ double red[kLapBlock / kWave];
    auto at = [&](int d, int s, int lbl) { return (mode && pr.p == lbl) ? pos_new(R, pr, d, s) : pos_at(R, d, s, lbl); };
    double tot = 0.;
    for (int it = threadIdx.x; it < nb * (N - 1); it += kLapBlock) {
        const int bi = b0 + it / (N - 1), qi = it % (N - 1), q = qi < p ? qi : qi + 1;
        const int si = bi % M;
        double lap = 0.;
        for (int side = 0; side < 2; ++side) {
            const int sj = side == 0 ? (bi + 1) % M : (bi - 1 + M) % M;
            double a0[3], c0[3], a1[3], c1[3];
            for (int d = 0; d < 3; ++d) {
                a0[d] = at(d, si, p);
                c0[d] = at(d, si, q);
                a1[d] = at(d, sj, p);
                c1[d] = at(d, sj, q);
            }
            LinkGeom g;
            link_geom(a0, c0, a1, c1, R.L, R.iL, g);
            const double f0 = pair_u(R, g.r, g.rp, g.s);
            double part = 0.;
            for (int d = 0; d < 3; ++d) {
                const double x0 = a0[d];
                a0[d] = x0 + eps;
                link_geom(a0, c0, a1, c1, R.L, R.iL, g);
                const double fp1 = pair_u(R, g.r, g.rp, g.s);
                a0[d] = x0 - eps;
                link_geom(a0, c0, a1, c1, R.L, R.iL, g);
                const double fm1 = pair_u(R, g.r, g.rp, g.s);
                part += (fp1 + fm1 - 2 * f0) / (eps * eps);
                a0[d] = x0;
            }
            lap += part;
        }
        tot += lap;
    }
    tot = wave_sum(tot);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.;
        for (int w = 0; w < kLapBlock / kWave; ++w) t += red[w];
        *out = t;
    }
}

__global__ void k_action_sum(const Rep* __restrict__ reps, int r, int mode, int n_links, int n_long,
                             const double* __restrict__ long_part, double* __restrict__ out) {
    const Rep& R = reps[r];
    double s = 0., l = 0.;
    for (int j = threadIdx.x; j < n_links; j += kWave) s += mode ? R.pair_part[j].y : R.pair_part[j].x;
    for (int j = threadIdx.x; j < n_long; j += kWave) l += long_part[j];
    s = wave_sum(s);
    l = wave_sum(l);
    if (threadIdx.x == 0) *out = s + R.long_scale * l;
}

// Move::Accept / Reject of an interface proposal.
__global__ void k_commit(const Rep* __restrict__ reps, int r, int accept, int rho_updated) {
    const Rep& R = reps[r];
    const Prop pr = *R.prop;
    if (accept)
        for (int m = threadIdx.x; m < pr.count; m += blockDim.x) {
            const int s = (pr.first + m) % R.M;
            for (int d = 0; d < 3; ++d) R.pos[((size_t)d * R.M + s) * R.N + pr.p] = R.prop_new[3 * m + d];
            if (rho_updated && R.use_long_range) R.sel[s] ^= 1;
        }
}

// ---------------------------------------------------------------------------
// K6: whole-path estimators (thermal energy, potential)
// ---------------------------------------------------------------------------
// Kinetic::DActionDBeta summand per slice (kinetic_class.h:35-45)
__global__ __launch_bounds__(256) void k_kin_dbeta(const Rep* __restrict__ reps, int slot) {
    const Rep& R = reps[blockIdx.y];
    const int b = blockIdx.x;
    if (b >= R.M) return;
    const int b1 = (b + 1) % R.M;
    double t = 0.;
    for (int p = threadIdx.x; p < R.N; p += 256) {
        const int pn = next_label(R, b, p);
        double dd[3];
        for (int d = 0; d < 3; ++d) dd[d] = minimg(pos_at(R, d, b, p) - pos_at(R, d, b1, pn), R.L, R.iL);
        double img = 0.;
        if (R.kin_images > 0)
            for (int d = 0; d < 3; ++d) img -= spl_eval1(R.kin_dtau, dd[d]);
        t += img - (dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]) * R.kin_i4ltt;
    }
    __shared__ double sh[4];
    t = block_sum<256>(t, sh);
    if (threadIdx.x == 0) R.scratch[slot * R.M + b] = t;
}

// Pair sums over all pairs on link (b, b+1): which 0 = CalcdUdBeta
// (pair_action_class.h:241-264), 1 = CalcV on plain minimum images
// (pair_action_class.h:369-395).  Pairs (i, i+d mod N), d = 1..N/2.
__global__ __launch_bounds__(256) void k_pair_whole(const Rep* __restrict__ reps, int which, int slot) {
    const Rep& R = reps[blockIdx.y];
    const int b = blockIdx.x;
    if (b >= R.M || R.pair_type == 0) {
        if (b < R.M && threadIdx.x == 0) R.scratch[slot * R.M + b] = 0.;
        return;
    }
    const int N = R.N, M = R.M, b1 = (b + 1) % M;
    double t = 0.;
    const int half = N / 2;
    for (int i = threadIdx.x; i < N; i += 256) {
        double a0[3], a1[3];
        for (int d = 0; d < 3; ++d) {
            a0[d] = pos_at(R, d, b, i);
            a1[d] = pos_at(R, d, b1, i);
        }
        for (int dd = 1; dd <= half; ++dd) {
            if ((N % 2 == 0) && dd == half && i >= half) continue;
            const int j = (i + dd) % N;
            double c0[3], c1[3];
            for (int d = 0; d < 3; ++d) {
                c0[d] = pos_at(R, d, b, j);
                c1[d] = pos_at(R, d, b1, j);
            }
            // pair (lo, hi): r = r_hi - r_lo; magnitudes do not depend on the orientation
            LinkGeom g;
            if (i < j)
                link_geom(a0, c0, a1, c1, R.L, R.iL, g);
            else
                link_geom(c0, a0, c1, a1, R.L, R.iL, g);
            if (which == 0) {
                t += pair_du_dbeta(R, g.r, g.rp, g.s);
            } else {
                double v0[3], v1[3];
                for (int d = 0; d < 3; ++d) {
                    v0[d] = minimg(a0[d] - c0[d], R.L, R.iL);
                    v1[d] = minimg(a1[d] - c1[d], R.L, R.iL);
                }
                t += pair_v(R, sqrt(v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]),
                            sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]));
            }
        }
    }
    __shared__ double sh[4];
    t = block_sum<256>(t, sh);
    if (threadIdx.x == 0) R.scratch[slot * R.M + b] = t;
}

// sum_k coef(k) sum_b |rho(b,k)|^2 per k chunk, plus (chunk 0) the k=0 and r=0
// constants: which 0 CalcdUdBetaLong (coef_b, const_du), 1 CalcVLong (coef_v,
// const_v)
__global__ __launch_bounds__(256) void k_long_whole(const Rep* __restrict__ reps, int which) {
    const Rep& R = reps[blockIdx.y];
    const int k = (int)(blockIdx.x * 256 + threadIdx.x);
    if ((int)blockIdx.x * 256 >= R.nk) return;
    double t = 0.;
    if (k < R.nk && R.use_long_range) {
        for (int b = 0; b < R.M; ++b) {
            const double2 v = R.rho[((size_t)R.sel[b] * R.M + b) * R.nk + k];
            t += v.x * v.x + v.y * v.y;
        }
        t *= which ? R.coef_v[k] : R.coef_b[k];
    }
    __shared__ double sh[4];
    t = block_sum<256>(t, sh);
    if (threadIdx.x == 0) R.scratch[scratch_long(R.M) + blockIdx.x] = t + (blockIdx.x == 0 ? (which ? R.const_v : R.const_du) : 0.);
}

// Kinetic::VirialEnergy (kinetic_class.h:48-102), thread per particle
__global__ __launch_bounds__(256) void k_kin_virial(const Rep* __restrict__ reps, int W_in, int slot) {
    const Rep& R = reps[blockIdx.x];
    const int M = R.M;
    const int W = W_in > 0 ? W_in : M;
    const double wt = W * R.tau;
    double tot = 0.;
    // the walks follow the path: past slice M - 1 a bead is followed by slice 0
    // of the particle perm[label]
    for (int p = threadIdx.x; p < R.N; p += 256) {
        int b_0 = 0, b_1 = 1 % M, b_2 = 2 % M;
        int l_0 = p, l_1 = next_label(R, 0, p), l_2 = next_label(R, b_1, l_1);
        double d1[3], dL[3], t[3];
        for (int d = 0; d < 3; ++d) d1[d] = minimg(pos_at(R, d, b_1, l_1) - pos_at(R, d, b_0, l_0), R.L, R.iL);
        for (int d = 0; d < 3; ++d) dL[d] = d1[d];
        for (int sk = 1; sk < W; ++sk) {
            for (int d = 0; d < 3; ++d) t[d] = minimg(pos_at(R, d, b_2, l_2) - pos_at(R, d, b_1, l_1), R.L, R.iL);
            for (int d = 0; d < 3; ++d) dL[d] += t[d];
            b_1 = b_2;
            l_1 = l_2;
            l_2 = next_label(R, b_2, l_2);
            b_2 = (b_2 + 1) % M;
        }
        tot -= (dL[0] * d1[0] + dL[1] * d1[1] + dL[2] * d1[2]) * R.kin_i4lt / wt;
        for (int bi = 1; bi < M; ++bi) {
            for (int d = 0; d < 3; ++d) dL[d] -= d1[d];
            for (int d = 0; d < 3; ++d) t[d] = minimg(pos_at(R, d, b_2, l_2) - pos_at(R, d, b_1, l_1), R.L, R.iL);
            for (int d = 0; d < 3; ++d) dL[d] += t[d];
            l_0 = next_label(R, b_0, l_0);
            b_0 = (b_0 + 1) % M;
            const int bn = (b_0 + 1) % M, ln = next_label(R, b_0, l_0);
            for (int d = 0; d < 3; ++d) d1[d] = minimg(pos_at(R, d, bn, ln) - pos_at(R, d, b_0, l_0), R.L, R.iL);
            b_1 = b_2;
            l_1 = l_2;
            l_2 = next_label(R, b_2, l_2);
            b_2 = (b_2 + 1) % M;
            tot -= (dL[0] * d1[0] + dL[1] * d1[1] + dL[2] * d1[2]) * R.kin_i4lt / wt;
        }
    }
    __shared__ double sh[4];
    tot = block_sum<256>(tot, sh);
    if (threadIdx.x == 0) R.scratch[slot * M] = R.kin_dbeta_const / W + tot;
}

// ---------------------------------------------------------------------------
// K7: gradients.  Thread per bead (b, p); output grad[b][p][3].
// ---------------------------------------------------------------------------
// Reference-compatible PairAction::GetActionGradient(b, b+1, {p}, 0):
// CalcGradientU by central difference, eps = 1e-4 (pair_action_class.h:135-155,305-337)
__global__ __launch_bounds__(64) void k_grad_compat(const Rep* __restrict__ reps, int out_est, int short_on) {
    const Rep& R = reps[blockIdx.z];
    const int p = blockIdx.x * 64 + threadIdx.x, b = blockIdx.y;
    if (p >= R.N || b >= R.M) return;
    const int M = R.M, N = R.N;
    const int bj = (b + 1) % M, bk = (b + M - 1) % M;
    double r0[3], pj[3], pk[3], g[3] = {0., 0., 0.};
    for (int d = 0; d < 3; ++d) {
        r0[d] = pos_at(R, d, b, p);
        pj[d] = pos_at(R, d, bj, p);
        pk[d] = pos_at(R, d, bk, p);
    }
    const double eps = 1.e-4;
    if (!short_on) {
    } else if (R.pair_type == 3 && N > 1) {
        // IlkkaPairAction::CalcGradientU (ilkka_pair_action_class.h:172-222) per link,
        // as the reference forms it from DrDrpDrrp's r = r_q - r_p (path_class.h:153-166):
        // -(1/2)[U_x (r/|r| + (r - r')/|r - r'|) + U_y (r/|r| - (r - r')/|r - r'|)]
        // - (1/2) u_long'(|r|) r/|r|, |r| and |r'| clamped to the u/diag grid.  The
        // first term is dU/dr_p; the second has the opposite sign of d/dr_p of the
        // -(1/2) u_long(|r|) term of CalcU.
        for (int q = 0; q < N; ++q) {
            if (q == p) continue;
            double q0[3];
            for (int d = 0; d < 3; ++d) q0[d] = pos_at(R, d, b, q);
            for (int lk = 0; lk < 2; ++lk) {
                const double* a1 = lk ? pk : pj;
                double c1[3];
                for (int d = 0; d < 3; ++d) c1[d] = pos_at(R, d, lk ? bk : bj, q);
                LinkGeom gg;
                link_geom(r0, q0, a1, c1, R.L, R.iL, gg);
                const double qq = 0.5 * (gg.r + gg.rp);
                double ux, uy;
                spl2_eval(R.u2, qq + 0.5 * gg.s, qq - 0.5 * gg.s, &ux, &uy);
                double dl = 0.;
                if (R.use_long_range) {
                    double a = gg.r, bb = gg.rp;
                    clamp2(R.r_u_min, R.r_u_max, a, bb);
                    const int ia = spl_locate(R.u_long_r, a);
                    dl = spl_der(R.u_long_r, ia, 0, a - R.u_long_r.x[ia]);
                }
                for (int d = 0; d < 3; ++d) {
                    const double rh = gg.r == 0. ? 0. : gg.rv[d] / gg.r, sh = gg.s == 0. ? 0. : gg.sv[d] / gg.s;
                    g[d] += -0.5 * (ux * (rh + sh) + uy * (rh - sh)) - 0.5 * dl * rh;
                }
            }
        }
    } else if (R.pair_type != 0 && N > 1)
        for (int q = 0; q < N; ++q) {
            if (q == p) continue;
            double q0[3], qj[3], qk[3];
            for (int d = 0; d < 3; ++d) {
                q0[d] = pos_at(R, d, b, q);
                qj[d] = pos_at(R, d, bj, q);
                qk[d] = pos_at(R, d, bk, q);
            }
            for (int lk = 0; lk < 2; ++lk) {
                const double* a1 = lk ? pk : pj;
                const double* c1 = lk ? qk : qj;
                for (int d = 0; d < 3; ++d) {
                    double a0[3] = {r0[0], r0[1], r0[2]};
                    LinkGeom gg;
                    a0[d] = r0[d] + eps;
                    link_geom(a0, q0, a1, c1, R.L, R.iL, gg);
                    const double f1 = pair_u(R, gg.r, gg.rp, gg.s);
                    a0[d] = r0[d] - eps;
                    link_geom(a0, q0, a1, c1, R.L, R.iL, gg);
                    const double f2 = pair_u(R, gg.r, gg.rp, gg.s);
                    g[d] += (f1 - f2) / (2. * eps);
                }
            }
        }
    double* o = (out_est ? R.grad_est : R.grad) + 3 * ((size_t)b * N + p);
    o[0] = g[0];
    o[1] = g[1];
    o[2] = g[2];
}

// Analytic adjoint, local terms: kinetic links (b-1,b), (b,b+1) and short-range
// pair links of every other particle.  Writes (overwrites) grad.
__global__ __launch_bounds__(64) void k_adjoint_local(const Rep* __restrict__ reps, unsigned int terms, int out_est) {
    const Rep& R = reps[blockIdx.z];
    const int i = blockIdx.x * 64 + threadIdx.x, b = blockIdx.y;
    if (i >= R.N || b >= R.M) return;
    const int M = R.M, N = R.N;
    const int bn = (b + 1) % M, bp = (b + M - 1) % M;
    double ri[3], rn[3], rp[3], g[3] = {0., 0., 0.};
    for (int d = 0; d < 3; ++d) {
        ri[d] = pos_at(R, d, b, i);
        rn[d] = pos_at(R, d, bn, i);
        rp[d] = pos_at(R, d, bp, i);
    }
    if (terms & 1u) {
        // K(d) = sum_x img(d_x) + |d|^2 i4lt, d = r_b - r_{b+1}: dK/dr_b = K'(d_fwd) - K'(d_bwd);
        // the kinetic links follow the path across the wrap
        const int kn = next_label(R, b, i), kp = prev_label(R, b, i);
        double df[3], db[3];
        for (int d = 0; d < 3; ++d) {
            df[d] = minimg(ri[d] - (kn == i ? rn[d] : pos_at(R, d, bn, kn)), R.L, R.iL);
            db[d] = minimg((kp == i ? rp[d] : pos_at(R, d, bp, kp)) - ri[d], R.L, R.iL);
        }
        for (int d = 0; d < 3; ++d) {
            double kf = 2. * df[d] * R.kin_i4lt, kb = 2. * db[d] * R.kin_i4lt;
            if (R.kin_images > 0) {
                int ia = spl_locate(R.kin[0], df[d]), ib = spl_locate(R.kin[0], db[d]);
                kf += spl_der(R.kin[0], ia, 0, df[d] - R.kin[0].x[ia]);
                kb += spl_der(R.kin[0], ib, 0, db[d] - R.kin[0].x[ib]);
            }
            g[d] += kf - kb;
        }
    }
    if ((terms & 2u) && R.pair_type != 0)
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            double j0[3], jn[3], jp[3];
            for (int d = 0; d < 3; ++d) {
                j0[d] = pos_at(R, d, b, j);
                jn[d] = pos_at(R, d, bn, j);
                jp[d] = pos_at(R, d, bp, j);
            }
            LinkGeom gg;
            double ur, urp, us;
            // link (b, b+1), i first: r = r_j(b) - r_i(b); dU/dr_i(b) = -(U_r rhat + U_s shat)
            link_geom(ri, j0, rn, jn, R.L, R.iL, gg);
            pair_u_grad(R, gg.r, gg.rp, gg.s, ur, urp, us);
            for (int d = 0; d < 3; ++d)
                g[d] -= ur * gg.rv[d] / gg.r + ((gg.s > 0.) ? us * gg.sv[d] / gg.s : 0.);
            // link (b-1, b), i first: r' = r_j(b) - r_i(b); dU/dr_i(b) = -(U_r' rphat - U_s shat)
            link_geom(rp, jp, ri, j0, R.L, R.iL, gg);
            pair_u_grad(R, gg.r, gg.rp, gg.s, ur, urp, us);
            for (int d = 0; d < 3; ++d)
                g[d] -= urp * gg.rpv[d] / gg.rp - ((gg.s > 0.) ? us * gg.sv[d] / gg.s : 0.);
        }
    double* o = (out_est ? R.grad_est : R.grad) + 3 * ((size_t)b * N + i);
    o[0] = g[0];
    o[1] = g[1];
    o[2] = g[2];
}

// Loads a replica's link-action tables (du_only: the David du_kj_dbeta table
// instead) into LDS.
__device__ __forceinline__ void stage_tables(const Rep& R, double* lds, bool du_only) {
    if (du_only)
        stage_batched(lds, R.bt_du.c, R.bt_du.c, R.lds_doubles_du, R.lds_doubles_du, 0, (int)blockDim.x);
    else
        stage_batched(lds, R.bt_a.c, R.bt_b.c, R.bt_a.ncoef * R.bt_a.ncomp, R.lds_doubles, 0, (int)blockDim.x);
    __syncthreads();
}

// Analytic adjoint, local terms with LDS tables: one 1024-thread workgroup per
// replica; thread t owns particle i = t % N over a contiguous slice range and
// walks it in order, so each link (b, b+1) of (i, j) is evaluated once for i:
// its r_i(b) part goes to bead b, its r_i(b+1) part is carried to bead b+1.
constexpr int kAdjLdsBlock = 1024;
// PT, NORD: the pair type and David order shared by all replicas (0: per replica)
template <int PT, int NORD>
__global__ __launch_bounds__(kAdjLdsBlock) void k_adjoint_local_lds(const Rep* __restrict__ reps, unsigned int terms,
                                                                    int out_est) {
    extern __shared__ double lds[];
    // gridDim.y workgroups per replica (when replicas are few), each taking a
    // contiguous block of the slice ranges
    const Rep R = reps[blockIdx.x];  // by value: the fields used stay in registers
    stage_tables(R, lds, false);
    const int M = R.M, N = R.N, bpr = gridDim.y;
    const int per_blk = kAdjLdsBlock / N > 0 ? kAdjLdsBlock / N : 1;  // slice ranges per particle in one workgroup
    const int per = min(M, per_blk * bpr);                               // ... over the replica's workgroups
    const int p_lo = (int)((long long)per * blockIdx.y / bpr), p_hi = (int)((long long)per * (blockIdx.y + 1) / bpr);
    const int n_units = N * (p_hi - p_lo);
    for (int unit = threadIdx.x; unit < n_units; unit += kAdjLdsBlock) {
        const int i = unit % N, part = p_lo + unit / N;
        const int b_lo = (int)((long long)M * part / per), b_hi = (int)((long long)M * (part + 1) / per);
        if (b_lo >= b_hi) continue;
        double carry[3] = {0., 0., 0.};  // d/dr_i(b) from link (b-1, b)
        // the link entering the range
        {
            const int bp = (b_lo + M - 1) % M;
            double rp_[3], ri_[3];
            for (int d = 0; d < 3; ++d) {
                rp_[d] = pos_at(R, d, bp, i);
                ri_[d] = pos_at(R, d, b_lo, i);
            }
            if ((terms & 2u) && R.pair_type != 0)
                for (int j = 0; j < N; ++j) {
                    if (j == i) continue;
                    double jp[3], j0[3];
                    for (int d = 0; d < 3; ++d) {
                        jp[d] = pos_at(R, d, bp, j);
                        j0[d] = pos_at(R, d, b_lo, j);
                    }
                    LinkGeom gg;
                    double ur, urp, us;
                    link_geom_fast(rp_, jp, ri_, j0, R.L, R.iL, gg);
                    pair_u_grad_lds<PT, NORD>(R, lds, gg.r, gg.rp, gg.s, ur, urp, us);
                    const double frp = urp / gg.rp, so = (gg.s > 0.) ? us / gg.s : 0.;
                    for (int d = 0; d < 3; ++d) carry[d] -= frp * gg.rpv[d] - so * gg.sv[d];
                }
        }
        for (int b = b_lo; b < b_hi; ++b) {
            const int bn = (b + 1) % M, bp = (b + M - 1) % M;
            double ri[3], rn[3], rpv[3], g[3], nxt[3] = {0., 0., 0.};
            for (int d = 0; d < 3; ++d) {
                ri[d] = pos_at(R, d, b, i);
                rn[d] = pos_at(R, d, bn, i);
                rpv[d] = pos_at(R, d, bp, i);
                g[d] = carry[d];
            }
            if (terms & 1u) {
                // the kinetic links follow the path across the wrap
                const int kn = next_label(R, b, i), kp = prev_label(R, b, i);
                for (int d = 0; d < 3; ++d) {
                    const double xn = kn == i ? rn[d] : pos_at(R, d, bn, kn), xp = kp == i ? rpv[d] : pos_at(R, d, bp, kp);
                    const double df = minimg(ri[d] - xn, R.L, R.iL), db = minimg(xp - ri[d], R.L, R.iL);
                    double kf = 2. * df * R.kin_i4lt, kb = 2. * db * R.kin_i4lt;
                    if (R.kin_images > 0) {
                        int ia = spl_locate(R.kin[0], df), ib = spl_locate(R.kin[0], db);
                        kf += spl_der(R.kin[0], ia, 0, df - R.kin[0].x[ia]);
                        kb += spl_der(R.kin[0], ib, 0, db - R.kin[0].x[ib]);
                    }
                    g[d] += kf - kb;
                }
            }
            if ((terms & 2u) && R.pair_type != 0)
                for (int j = 0; j < N; ++j) {
                    if (j == i) continue;
                    double j0[3], jn[3];
                    for (int d = 0; d < 3; ++d) {
                        j0[d] = pos_at(R, d, b, j);
                        jn[d] = pos_at(R, d, bn, j);
                    }
                    LinkGeom gg;
                    double ur, urp, us;
                    link_geom_fast(ri, j0, rn, jn, R.L, R.iL, gg);
                    pair_u_grad_lds<PT, NORD>(R, lds, gg.r, gg.rp, gg.s, ur, urp, us);
                    const double fr = ur / gg.r, frp = urp / gg.rp, so = (gg.s > 0.) ? us / gg.s : 0.;
                    for (int d = 0; d < 3; ++d) {
                        g[d] -= fr * gg.rv[d] + so * gg.sv[d];
                        nxt[d] -= frp * gg.rpv[d] - so * gg.sv[d];
                    }
                }
            double* o = (out_est ? R.grad_est : R.grad) + 3 * ((size_t)b * N + i);
            o[0] = g[0];
            o[1] = g[1];
            o[2] = g[2];
            carry[0] = nxt[0];
            carry[1] = nxt[1];
            carry[2] = nxt[2];
        }
    }
}

// dK/dr_i(b) of the two kinetic links at bead (b, i); the links follow the path
// across the wrap.
__device__ __forceinline__ void adj_kinetic(const Rep& R, int b, int i, double* g) {
    const int M = R.M, bn = (b + 1) % M, bp = (b + M - 1) % M;
    const int kn = next_label(R, b, i), kp = prev_label(R, b, i);
    for (int d = 0; d < 3; ++d) {
        const double ri = pos_at(R, d, b, i);
        const double df = minimg(ri - pos_at(R, d, bn, kn), R.L, R.iL), db = minimg(pos_at(R, d, bp, kp) - ri, R.L, R.iL);
        double kf = 2. * df * R.kin_i4lt, kb = 2. * db * R.kin_i4lt;
        if (R.kin_images > 0) {
            int ia = spl_locate(R.kin[0], df), ib = spl_locate(R.kin[0], db);
            kf += spl_der(R.kin[0], ia, 0, df - R.kin[0].x[ia]);
            kb += spl_der(R.kin[0], ib, 0, db - R.kin[0].x[ib]);
        }
        g[d] = kf - kb;
    }
}

// Adds the short-range pair terms of link (b, b+1) to dS/dr at beads b and b+1,
// one wave, each pair-link evaluated once for both of its particles.  The
// particles are taken in tiles of 64.  For a tile pair (I, J >= I), lane l holds
// particle p = 64I + l and at step s evaluates its link with q = 64J + (l + s)
// mod 64 (I = J: (l + s) mod n for s = 1 .. n/2, the last step on half the lanes
// when n is even).  p's share stays in registers; q's, the same vector with the
// opposite sign, is added to q's slot in the wave's LDS slots (6 x 64 doubles,
// component-major): at one step the lanes touch distinct slots, and one wave's
// LDS operations complete in order, so every sum is formed in a fixed order.
template <int PT, int NORD>
__device__ void adj_link_pairs(const Rep& R, const double* lds, double* slot, double* __restrict__ o, int b, int lane) {
    const int M = R.M, N = R.N, b1 = (b + 1) % M, T = (N + kWave - 1) / kWave;
    for (int I = 0; I < T; ++I) {
        const int nI = min(kWave, N - I * kWave), p = I * kWave + lane;
        const bool pv = lane < nI;
        double a0[3], a1[3], acc[6] = {0., 0., 0., 0., 0., 0.};
        for (int d = 0; d < 3; ++d) {
            a0[d] = pv ? pos_at(R, d, b, p) : 0.;
            a1[d] = pv ? pos_at(R, d, b1, p) : 0.;
        }
        for (int J = I; J < T; ++J) {
            const bool diag = J == I;
            const int nJ = min(kWave, N - J * kWave);
            const int s0 = diag ? 1 : 0, s1 = diag ? nI / 2 : kWave - 1;
            for (int k = 0; k < 6; ++k) slot[k * kWave + lane] = 0.;
            for (int s = s0; s <= s1; ++s) {
                const int ql = diag ? (pv ? (lane + s) % nI : lane) : ((lane + s) & (kWave - 1));
                const bool act = pv && (diag ? !((nI % 2 == 0) && s == nI / 2 && lane >= nI / 2) : ql < nJ);
                if (!act) continue;
                const int q = J * kWave + ql;
                double c0[3], c1[3];
                for (int d = 0; d < 3; ++d) {
                    c0[d] = pos_at(R, d, b, q);
                    c1[d] = pos_at(R, d, b1, q);
                }
                // the geometry of link_geom_fast, s = r - r' formed again where it is used
                double rv[3], rpv[3], rr = 0., pp = 0., ss = 0.;
                for (int d = 0; d < 3; ++d) {
                    rv[d] = c0[d] - a0[d];
                    rpv[d] = c1[d] - a1[d];
                    rv[d] -= rint(rv[d] * R.iL) * R.L;
                    rpv[d] += rint((rv[d] - rpv[d]) * R.iL) * R.L;
                    const double sv = rv[d] - rpv[d];
                    rr += rv[d] * rv[d];
                    pp += rpv[d] * rpv[d];
                    ss += sv * sv;
                }
                const double r = sqrt_pos(rr), rp = sqrt_pos(pp), sl = sqrt_pos(ss);
                double ur, urp, us;
                // r = r_q(b) - r_p(b): dU/dr_p(b) = -(U_r rhat + U_s shat), dU/dr_p(b+1) = -(U_r' rphat - U_s shat)
                pair_u_grad_lds<PT, NORD>(R, lds, r, rp, sl, ur, urp, us);
                const double fr = ur / r, frp = urp / rp, so = (sl > 0.) ? us / sl : 0.;
                for (int d = 0; d < 3; ++d) {
                    const double sv = rv[d] - rpv[d];
                    const double e0 = fr * rv[d] + so * sv, e1 = frp * rpv[d] - so * sv;
                    acc[d] -= e0;
                    acc[3 + d] -= e1;
                    slot[d * kWave + ql] += e0;
                    slot[(3 + d) * kWave + ql] += e1;
                }
            }
            if (diag) {
                for (int k = 0; k < 6; ++k) acc[k] += slot[k * kWave + lane];
            } else if (lane < nJ) {
                double* o0 = o + 3 * ((size_t)b * N + J * kWave + lane);
                double* o1 = o + 3 * ((size_t)b1 * N + J * kWave + lane);
                for (int d = 0; d < 3; ++d) {
                    o0[d] += slot[d * kWave + lane];
                    o1[d] += slot[(3 + d) * kWave + lane];
                }
            }
        }
        if (pv) {
            double* o0 = o + 3 * ((size_t)b * N + p);
            double* o1 = o + 3 * ((size_t)b1 * N + p);
            for (int d = 0; d < 3; ++d) {
                o0[d] += acc[d];
                o1[d] += acc[3 + d];
            }
        }
    }
}

// Analytic adjoint, local terms, each pair-link evaluated once: one 1024-thread
// workgroup per replica (gridDim.y per replica when replicas are few); wave w
// owns links [b_lo, b_hi) of n_rng contiguous ranges of at least two links.
// Phase 0 writes the kinetic term of beads b_lo+1 .. b_hi and adds the pair
// terms of links b_lo+1 .. b_hi-1, all of whose beads it owns; phase 1, the
// next launch, adds those of link b_lo, whose bead b_lo the previous range
// wrote.  Every bead's sum is formed in a fixed order.  Dynamic LDS: the
// replica's tables, then kAdjSlotDoubles per wave.
constexpr int kAdjPairWaves = kAdjLdsBlock / kWave;
constexpr int kAdjSlotDoubles = 6 * kWave;  // one wave's partner slots
template <int PT, int NORD>
__global__ __launch_bounds__(kAdjLdsBlock) void k_adjoint_pairs_lds(const Rep* __restrict__ reps, unsigned int terms,
                                                                    int out_est, int phase) {
    extern __shared__ double lds[];
    const Rep R = reps[blockIdx.x];
    const bool pairs = (terms & 2u) && R.pair_type != 0 && R.N > 1;
    if (phase == 1 && !pairs) return;
    if (pairs) stage_tables(R, lds, false);
    const int M = R.M, N = R.N;
    const int n_rng = max(1, min(kAdjPairWaves * (int)gridDim.y, M / 2));
    const int w = blockIdx.y * kAdjPairWaves + threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    if (w >= n_rng) return;
    const int b_lo = (int)((long long)M * w / n_rng), b_hi = (int)((long long)M * (w + 1) / n_rng);
    double* o = out_est ? R.grad_est : R.grad;
    // the wave's partner slots follow the replica's tables
    double* slot = lds + R.lds_doubles + (threadIdx.x / kWave) * kAdjSlotDoubles;
    if (phase == 1) {
        adj_link_pairs<PT, NORD>(R, lds, slot, o, b_lo, lane);
        return;
    }
    for (int b = b_lo + 1; b <= b_hi; ++b) {
        const int bb = b % M;
        for (int p = lane; p < N; p += kWave) {
            double g[3] = {0., 0., 0.};
            if (terms & 1u) adj_kinetic(R, bb, p, g);
            double* op = o + 3 * ((size_t)bb * N + p);
            op[0] = g[0];
            op[1] = g[1];
            op[2] = g[2];
        }
    }
    if (pairs)
        for (int b = b_lo + 1; b < b_hi; ++b) adj_link_pairs<PT, NORD>(R, lds, slot, o, b, lane);
}

// Whole-path pair sums with LDS tables (which 0: CalcdUdBeta, 1: CalcV on plain
// minimum images): gridDim.y 1024-thread workgroups per replica (blockIdx.x),
// each over every gridDim.y-th slice, per-slice partials.
__global__ __launch_bounds__(kAdjLdsBlock) void k_pair_whole_lds(const Rep* __restrict__ reps, int which, int slot) {
    extern __shared__ double lds[];
    const Rep& R = reps[blockIdx.x];
    // David: the du_kj_dbeta table (potential in component 0); Bare: the pair tables
    stage_tables(R, lds, R.pair_type == 2);
    const double* ldu = lds;
    __shared__ double red[kAdjLdsBlock / kWave];
    const int N = R.N, M = R.M, half = N / 2;
    // slices b = blockIdx.y, + gridDim.y, ...: each slice's sum is made by one
    // workgroup in the same order whatever the number of workgroups
    for (int b = blockIdx.y; b < M; b += gridDim.y) {
        const int b1 = (b + 1) % M;
        double t = 0.;
        // pairs (i, i + d mod N), d = 1..N/2: thread handles i = tid % N and a slice of d
        const int per = kAdjLdsBlock / N > 0 ? kAdjLdsBlock / N : 1;
        for (int unit = threadIdx.x; unit < N * per; unit += kAdjLdsBlock) {
            const int i = unit % N, part = unit / N;
            const int d_lo = 1 + (int)((long long)half * part / per), d_hi = 1 + (int)((long long)half * (part + 1) / per);
            double a0[3], a1[3];
            for (int d = 0; d < 3; ++d) {
                a0[d] = pos_at(R, d, b, i);
                a1[d] = pos_at(R, d, b1, i);
            }
            for (int dd = d_lo; dd < d_hi; ++dd) {
                if ((N % 2 == 0) && dd == half && i >= half) continue;
                const int j = (i + dd) % N;
                double c0[3], c1[3];
                for (int d = 0; d < 3; ++d) {
                    c0[d] = pos_at(R, d, b, j);
                    c1[d] = pos_at(R, d, b1, j);
                }
                if (which == 0) {
                    LinkGeom g;
                    if (i < j)
                        link_geom_fast(a0, c0, a1, c1, R.L, R.iL, g);
                    else
                        link_geom_fast(c0, a0, c1, a1, R.L, R.iL, g);
                    t += pair_du_dbeta_lds(R, lds, ldu, g.r, g.rp, g.s);
                } else {
                    double v0[3], v1[3];
                    for (int d = 0; d < 3; ++d) {
                        v0[d] = minimg(a0[d] - c0[d], R.L, R.iL);
                        v1[d] = minimg(a1[d] - c1[d], R.L, R.iL);
                    }
                    t += pair_v_lds(R, lds, ldu, sqrt(v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]),
                                    sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]));
                }
            }
        }
        t = wave_sum(t);
        if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tt = 0.;
            for (int w = 0; w < kAdjLdsBlock / kWave; ++w) tt += red[w];
            R.scratch[slot * M + b] = tt;
        }
        __syncthreads();
    }
}

// Factor on the long-range gradient: 1 for the adjoint; with compat, the
// reference's GetActionGradient (pair_action_class.h:331-334) adds
// IlkkaPairAction::CalcGradientULong (ilkka_pair_action_class.h:231-249), half
// the adjoint term, once per particle pair, i.e. (N - 1) / 2; the Bare and David
// actions add nothing (pair_action_class.h:158-165).
__device__ __forceinline__ double long_grad_scale(const Rep& R, int compat) {
    if (!compat) return 1.;
    return R.pair_type == 3 ? 0.5 * (R.N - 1.) : 0.;
}

// Analytic adjoint, Ewald term: d/dr_i of long_scale sum_k coef_s |rho(b,k)|^2
// = -2 long_scale sum_k coef_s k Im(conj(rho) e^{i k.r_i}).  Block = 64 beads of
// one slice; rho(b,:) staged through LDS in k tiles; per-bead phase tables in LDS.
constexpr int kAdjTile = 128;
__global__ __launch_bounds__(64) void k_adjoint_long(const Rep* __restrict__ reps, int out_est, int compat) {
    const Rep& R = reps[blockIdx.z];
    const int b = blockIdx.y;
    if (b >= R.M || (int)blockIdx.x * 64 >= R.N || !R.use_long_range || R.pair_type == 0) return;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int mx = R.max_idx, stride = mx + 1;
    extern __shared__ double2 dyn[];
    double2* tab = dyn;                            // [64][3][stride]
    double2* rho_t = dyn + 64 * 3 * stride;        // [kAdjTile]
    int* kp_t = (int*)(rho_t + kAdjTile);          // [kAdjTile]
    double* cf_t = (double*)(kp_t + kAdjTile);     // [kAdjTile]
    const bool valid = i < R.N;
    double2* my = tab + threadIdx.x * 3 * stride;
    if (valid)
        for (int d = 0; d < 3; ++d) phase_chain(pos_at(R, d, b, i) * R.box, mx, my + d * stride);
    double g[3] = {0., 0., 0.};
    const double2* rb = R.rho + ((size_t)R.sel[b] * R.M + b) * R.nk;
    for (int k0 = 0; k0 < R.nk; k0 += kAdjTile) {
        __syncthreads();
        for (int t = threadIdx.x; t < kAdjTile; t += 64) {
            const int k = k0 + t;
            if (k < R.nk) {
                rho_t[t] = rb[k];
                kp_t[t] = R.kidx[k];
                cf_t[t] = R.coef_s[k];
            }
        }
        __syncthreads();
        if (valid) {
            const int nt = min(kAdjTile, R.nk - k0);
            for (int t = 0; t < nt; ++t) {
                const int kp = kp_t[t];
                const double2 e = phase_k(my, stride, kp);
                const double2 rv = rho_t[t];
                // Im(conj(rho) e) = rho.x e.y - rho.y e.x
                const double im = cf_t[t] * (rv.x * e.y - rv.y * e.x);
                g[0] += im * ((kp & 255) - 128);
                g[1] += im * (((kp >> 8) & 255) - 128);
                g[2] += im * (((kp >> 16) & 255) - 128);
            }
        }
    }
    if (valid) {
        double* o = (out_est ? R.grad_est : R.grad) + 3 * ((size_t)b * R.N + i);
        const double f = -2. * long_grad_scale(R, compat) * R.long_scale * R.box;
        o[0] += f * g[0];
        o[1] += f * g[1];
        o[2] += f * g[2];
    }
}

// Long-range adjoint by runs: dS_long/dr_p(b) = -2 s box sum_k c_k n_k Im(conj(rho_k) e^{ik.r}).
// One workgroup per (slice, replica), a thread per particle.  The k-vectors are
// walked in runs of consecutive n2 at fixed (n0, n1), so each phase is the
// previous one times e^{+-i phi_2} (one complex product), and the weights
// w_k = c_k conj(rho_k) sit in LDS in run order: every lane reads the same
// address (a broadcast).  Run starts step running phases e^{i n0 phi_0},
// e^{i n1 phi_1}; the runs are ordered by (n0, n1) so those steps are short.
constexpr int kAdjRunBlock = 256;
__global__ __launch_bounds__(kAdjRunBlock) void k_adjoint_long_runs(const Rep* __restrict__ reps, int out_est, int compat) {
    const Rep& R = reps[blockIdx.y];
    const int b = blockIdx.x;
    if (b >= R.M || !R.use_long_range || R.pair_type == 0) return;
    extern __shared__ double2 wsh[];
    int4* runs = (int4*)(wsh + R.nk);
    const int nk = R.nk, n_runs = R.n_runs;
    const double2* rb = R.rho + ((size_t)R.sel[b] * R.M + b) * nk;
    for (int t = threadIdx.x; t < nk; t += kAdjRunBlock) {
        const int k = R.k_order[t];
        const double2 rv = rb[k];
        const double cf = R.coef_s[k];
        wsh[t] = make_double2(cf * rv.x, -cf * rv.y);
    }
    for (int t = threadIdx.x; t < n_runs; t += kAdjRunBlock) runs[t] = R.k_runs[t];
    __syncthreads();
    const double f = -2. * long_grad_scale(R, compat) * R.long_scale * R.box;
    for (int p = threadIdx.x; p < R.N; p += kAdjRunBlock) {
        double2 st[3];
        for (int d = 0; d < 3; ++d) {
            double sn, cs;
            sincos(pos_at(R, d, b, p) * R.box, &sn, &cs);
            st[d] = make_double2(cs, sn);
        }
        const double2 cst0 = make_double2(st[0].x, -st[0].y), cst1 = make_double2(st[1].x, -st[1].y);
        const double2 cst2 = make_double2(st[2].x, -st[2].y);
        double2 e0 = make_double2(1., 0.), e1 = make_double2(1., 0.);
        int c0 = 0, c1 = 0, off = 0;
        double g0 = 0., g1 = 0., g2 = 0.;
        for (int u = 0; u < n_runs; ++u) {
            const int4 h = runs[u];
            for (; c0 < h.x; ++c0) e0 = cmul(e0, st[0]);
            for (; c0 > h.x; --c0) e0 = cmul(e0, cst0);
            for (; c1 < h.y; ++c1) e1 = cmul(e1, st[1]);
            for (; c1 > h.y; --c1) e1 = cmul(e1, cst1);
            double2 e = cmul(e0, e1);
            const int n2 = h.z, len = h.w > 0 ? h.w : -h.w;
            for (int m = 0; m < n2; ++m) e = cmul(e, st[2]);
            for (int m = 0; m > n2; --m) e = cmul(e, cst2);
            const double2 step = h.w > 0 ? st[2] : cst2;
            const double dn = h.w > 0 ? 1. : -1.;
            double sr = 0., tr = 0., n2d = (double)n2;
            for (int i = 0; i < len; ++i) {
                const double2 w = wsh[off + i];
                const double im = w.x * e.y + w.y * e.x;  // Im(w e)
                sr += im;
                tr = fma(im, n2d, tr);
                n2d += dn;
                e = cmul(e, step);
            }
            off += len;
            g0 = fma((double)h.x, sr, g0);
            g1 = fma((double)h.y, sr, g1);
            g2 += tr;
        }
        double* o = (out_est ? R.grad_est : R.grad) + 3 * ((size_t)b * R.N + p);
        o[0] += f * g0;
        o[1] += f * g1;
        o[2] += f * g2;
    }
}

// Centroid term of Energy::VirialEstimator (energy_class.h:79-94):
// sum_p grad_nonkinetic(b,p) . Dr(r_p(b), r_p(0)), per slice.  from_sweep: the
// gradient is the sweep's full dS/dr, whose kinetic term is removed here.
__global__ __launch_bounds__(256) void k_centroid(const Rep* __restrict__ reps, int slot, int from_sweep) {
    const Rep& R = reps[blockIdx.y];
    const int b = blockIdx.x;
    if (b >= R.M) return;
    const int bn = (b + 1) % R.M, bp = (b + R.M - 1) % R.M;
    double t = 0.;
    for (int p = threadIdx.x; p < R.N; p += 256) {
        const double* g = (from_sweep ? R.grad : R.grad_est) + 3 * ((size_t)b * R.N + p);
        for (int d = 0; d < 3; ++d) {
            double gd = g[d];
            if (from_sweep) {
                const double x = pos_at(R, d, b, p);
                const double df = minimg(x - pos_at(R, d, bn, next_label(R, b, p)), R.L, R.iL);
                const double db = minimg(pos_at(R, d, bp, prev_label(R, b, p)) - x, R.L, R.iL);
                double kf = 2. * df * R.kin_i4lt, kb = 2. * db * R.kin_i4lt;
                if (R.kin_images > 0) {
                    int ia = spl_locate(R.kin[0], df), ib = spl_locate(R.kin[0], db);
                    kf += spl_der(R.kin[0], ia, 0, df - R.kin[0].x[ia]);
                    kb += spl_der(R.kin[0], ib, 0, db - R.kin[0].x[ib]);
                }
                gd -= kf - kb;
            }
            t += gd * minimg(pos_at(R, d, b, p) - pos_at(R, d, 0, p), R.L, R.iL);
        }
    }
    __shared__ double sh[4];
    t = block_sum<256>(t, sh);
    if (threadIdx.x == 0) R.scratch[slot * R.M + b] = t;
}

// Combines estimator partials (fixed order) into results res[0..1] of a replica.
//   what 0: kinetic dbeta -> res[0]; 1: pair dbeta or potential -> res[1];
//   what 2: virial kinetic (virial slot + centroid / 2 tau) -> res[0]
//   acc: add res[0..1] into the energy accumulators
// Scratch layout per replica: [slot][M] for slots 0..3, then long-range chunk
// partials, then the two results.

__global__ void k_finalize(const Rep* __restrict__ reps, int n_reps, int what, int acc) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_reps) return;
    const Rep& R = reps[r];
    const int M = R.M;
    const double* S = R.scratch;
    double* res = R.scratch + scratch_res(M, R.nk);
    if (what == 0) {
        double t = R.kin_dbeta_const;
        for (int b = 0; b < M; ++b) t += S[kSlotKin * M + b];
        res[0] = t;
    } else if (what == 1) {
        double t = 0.;
        for (int b = 0; b < M; ++b) t += S[kSlotPair * M + b];
        if (R.use_long_range && R.pair_type != 0) {
            const double* lw = R.scratch + scratch_long(M);
            double l = 0.;
            for (int c = 0; c < (R.nk + 255) / 256; ++c) l += lw[c];
            t += l;
        }
        res[1] = t;
    } else if (what == 2) {
        double c = 0.;
        for (int b = 0; b < M; ++b) c += S[kSlotCentroid * M + b];
        res[0] = S[kSlotVirial * M] + c / (2. * R.tau);
    }
    if (acc) {
        R.e_acc[0] += res[0];
        R.e_acc[1] += res[1];
        R.n_measure[0] += 1;
    }
}

// ---------------------------------------------------------------------------
// K8: StructureFactor::Accumulate (structure_factor_class.h:15-32)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sk(const Rep* __restrict__ reps, double k_cut) {
    const Rep& R = reps[blockIdx.y];
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= R.nk) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) R.n_measure[1] += 1;
    if (!(R.kmag[k] < k_cut)) return;
    double t = 0.;
    for (int b = 0; b < R.M; ++b) {
        const double2 v = R.rho[((size_t)R.sel[b] * R.M + b) * R.nk + k];
        t += v.x * v.x + v.y * v.y;
    }
    R.sk_acc[k] += t;
}

// ---------------------------------------------------------------------------
// K9: PairCorrelation::Accumulate (pair_correlation_class.h:15-28) with
// LinearGrid::ReverseMap (observable_class.h:56-62); LDS histogram per slice.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gr(const Rep* __restrict__ reps, double r_min, double d_ir, int n_r,
                                            const double* __restrict__ r_max_per_rep) {
    const Rep& R = reps[blockIdx.y];
    const int b = blockIdx.x;
    if (b >= R.M) return;
    __shared__ unsigned int hist[kGrBins];
    for (int i = threadIdx.x; i < n_r; i += 256) hist[i] = 0;
    __syncthreads();
    const int N = R.N, half = N / 2;
    double dir = d_ir;
    if (r_max_per_rep) dir = 1. / ((r_max_per_rep[blockIdx.y] - r_min) / (n_r - 1.));
    for (int i = threadIdx.x; i < N; i += 256) {
        double a[3];
        for (int d = 0; d < 3; ++d) a[d] = pos_at(R, d, b, i);
        for (int dd = 1; dd <= half; ++dd) {
            if ((N % 2 == 0) && dd == half && i >= half) continue;
            const int j = (i + dd) % N;
            const int lo = min(i, j), hi = max(i, j);
            double v[3];
            {
#pragma clang fp contract(off)
                for (int d = 0; d < 3; ++d) {
                    const double x = (lo == i) ? a[d] : pos_at(R, d, b, lo);
                    const double y = (hi == i) ? a[d] : pos_at(R, d, b, hi);
                    v[d] = minimg(x - y, R.L, R.iL);
                }
                const double rr = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                const double f = rint((rr - r_min) * dir - 0.5);
                const unsigned int bin = (unsigned int)(long long)f;
                if (f >= 0. && bin < (unsigned int)n_r) atomicAdd(&hist[bin], 1u);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n_r; i += 256)
        if (hist[i]) atomicAdd(&R.gr_cnt[i], (unsigned long long)hist[i]);
    if (threadIdx.x == 0 && b == 0) R.n_measure[2] += 1;
}

}  // namespace spimc
