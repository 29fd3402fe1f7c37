// spimc_stage.h -- a whole move stage of small replicas in one launch
// (included once, by spimc_engine.hip, after spimc_kernels.h).
//
// The batched uneven replicas (simpimc's MPI clones, framework_class.h:44-53)
// are latency-bound: a replica's attempts are serial, and each round of the
// four-launch design costs several dependent kernel launches.  Here one
// workgroup per replica loops over all of that replica's attempts of the
// stage (Bisect::Attempt, bisect_class.h:39-124, or DisplaceParticle::Attempt,
// displace_particle_class.h:28-72, then Accept / Reject), with workgroup
// barriers only:
//   A  first wavefront: draws (the attempt's tape draws staged in LDS in one
//      pass), the window beads (LDS), coarse levels and the level-0 beads
//   B  the first wavefront prepares the next attempt's draws (uniforms and
//      their logs, device-RNG normals) while the others compute the moved
//      beads' phase chains and the window's pair links (per-link sums in LDS);
//      then every wavefront takes groups of moved slices of rho_k (per-slice
//      sums in LDS; the committed-buffer flags of the slices are kept in LDS)
//   C  first wavefront: the level-0 Metropolis decision, commit, counters
// The stage's counters, draw cursor and the proposal live in LDS and are
// written back once at the end; the arithmetic is that of the round kernels.
#pragma once

#include "spimc_kernels.h"

namespace spimc {

constexpr int kStageMaxM = 256;        // slices of a replica this kernel takes
constexpr int kStageMaxDraws = 256;    // tape draws of one attempt staged in LDS

// draws one attempt consumes at most
__host__ __device__ inline int stage_draws(int move, int n_level) {
    return move == 0 ? 2 + 3 * ((1 << n_level) - 1) + n_level : 5;
}

struct StageCounters {
    unsigned int remaining, stage_i;
    unsigned long long seq;
    long long cursor;
    unsigned long long att, acc, beads, rho_slicek, pair_links;
    int error, pad;
};

// The attempt's draws from the tape, lanes of one wavefront: s_draw[i] =
// tape[cur + i] for i < need; *avail = how many exist.
__device__ __forceinline__ void stage_load_draws(const Rep& R, long long cur, int need, double* s_draw, int* avail,
                                                 int lane) {
    const long long left = R.tape_len - cur;
    const int n = (int)(left < 0 ? 0 : (left < need ? left : need));
    for (int i = lane; i < n; i += kWave) s_draw[i] = R.tape[cur + i];
    if (lane == 0) *avail = n;
}

// draw i of the attempt (tape_get semantics: a missing draw marks the attempt bad)
__device__ __forceinline__ double stage_draw(const double* s_draw, int avail, int i, int* bad) {
    if (i < 0 || i >= avail) {
        *bad = 1;
        return 0.5;
    }
    return s_draw[i];
}

// The pair links j_lo <= j < j_hi of the proposal, one wavefront, lanes over
// the partners q (two per lane per pass); a partner's beads at a link's second
// slice are reused as the next link's first, and the moved particle's OLD /
// NEW beads come from old_at / new_at.  OLD and NEW sums of link j to s_pair[j].
template <int PT, int NORD, class OldAt, class NewAt>
__device__ __forceinline__ void stage_pair_links(const Rep& R, const double* lds, bool tables, const Prop& pr,
                                                 OldAt old_at, NewAt new_at, int j_lo, int j_hi, int lane,
                                                 double2* s_pair) {
    const int M = R.M, N = R.N, p = pr.p;
    if (j_lo >= j_hi) return;
    for (int q0 = 0; q0 < N; q0 += 2 * kWave) {
        const int q = q0 + lane, q2 = q + kWave;
        const bool ok1 = q < N && q != p, ok2 = q2 < N && q2 != p;
        const int qa = q < N ? q : 0, qb = q2 < N ? q2 : qa;
        double a0[3], b0[3], po0[3], pn0[3];
        int s0 = (pr.b0 + j_lo) % M;
        for (int d = 0; d < 3; ++d) {
            a0[d] = pos_at(R, d, s0, qa);
            b0[d] = pos_at(R, d, s0, qb);
            po0[d] = old_at(j_lo, d);
            pn0[d] = new_at(j_lo, d);
        }
        // the beads at the link's second slice, loaded one link ahead
        double a1[3], b1[3], po1[3], pn1[3];
        {
            const int s1 = s0 + 1 == M ? 0 : s0 + 1;
            for (int d = 0; d < 3; ++d) {
                a1[d] = pos_at(R, d, s1, qa);
                b1[d] = pos_at(R, d, s1, qb);
                po1[d] = old_at(j_lo + 1, d);
                pn1[d] = new_at(j_lo + 1, d);
            }
        }
        for (int j = j_lo; j < j_hi; ++j) {
            const int s1 = s0 + 1 == M ? 0 : s0 + 1;
            double a2[3] = {0., 0., 0.}, b2[3] = {0., 0., 0.}, po2[3] = {0., 0., 0.}, pn2[3] = {0., 0., 0.};
            if (j + 1 < j_hi) {
                const int s2 = s1 + 1 == M ? 0 : s1 + 1;
                for (int d = 0; d < 3; ++d) {
                    a2[d] = pos_at(R, d, s2, qa);
                    b2[d] = pos_at(R, d, s2, qb);
                    po2[d] = old_at(j + 2, d);
                    pn2[d] = new_at(j + 2, d);
                }
            }
            double uo = 0., un = 0.;
            if (tables) {
                double r, rp, s2;
                link_rr_s2(po0, a0, po1, a1, R.L, R.iL, r, rp, s2);
                const double uo1 = pair_u_lds_s2<PT, NORD>(R, lds, r, rp, s2);
                link_rr_s2(pn0, a0, pn1, a1, R.L, R.iL, r, rp, s2);
                const double un1 = pair_u_lds_s2<PT, NORD>(R, lds, r, rp, s2);
                link_rr_s2(po0, b0, po1, b1, R.L, R.iL, r, rp, s2);
                const double uo2 = pair_u_lds_s2<PT, NORD>(R, lds, r, rp, s2);
                link_rr_s2(pn0, b0, pn1, b1, R.L, R.iL, r, rp, s2);
                const double un2 = pair_u_lds_s2<PT, NORD>(R, lds, r, rp, s2);
                uo = (ok1 ? uo1 : 0.) + (ok2 ? uo2 : 0.);
                un = (ok1 ? un1 : 0.) + (ok2 ? un2 : 0.);
            } else {
                LinkGeom g;
                if (ok1) {
                    link_geom(po0, a0, po1, a1, R.L, R.iL, g);
                    uo += pair_u(R, g.r, g.rp, g.s);
                    link_geom(pn0, a0, pn1, a1, R.L, R.iL, g);
                    un += pair_u(R, g.r, g.rp, g.s);
                }
                if (ok2) {
                    link_geom(po0, b0, po1, b1, R.L, R.iL, g);
                    uo += pair_u(R, g.r, g.rp, g.s);
                    link_geom(pn0, b0, pn1, b1, R.L, R.iL, g);
                    un += pair_u(R, g.r, g.rp, g.s);
                }
            }
            uo = wave_sum(uo);
            un = wave_sum(un);
            if (lane == 0) {
                const double2 v = q0 ? s_pair[j] : make_double2(0., 0.);
                s_pair[j] = make_double2(v.x + uo, v.y + un);
            }
            for (int d = 0; d < 3; ++d) {
                a0[d] = a1[d];
                b0[d] = b1[d];
                po0[d] = po1[d];
                pn0[d] = pn1[d];
                a1[d] = a2[d];
                b1[d] = b2[d];
                po1[d] = po2[d];
                pn1[d] = pn2[d];
            }
            s0 = s1;
        }
    }
}

// The six phase chains (OLD x3, NEW x3) of moved beads m = t / 6 for t = i0,
// i0 + step, ... < 6 count, into tab[m][6][max_idx + 1] (KSpace::CalcC's recurrence).
template <class OldAt, class NewAt>
__device__ __forceinline__ void stage_phase_chains(const Rep& R, int count, OldAt old_at, NewAt new_at, double2* tab,
                                                   int i0, int step) {
    const int stride = R.max_idx + 1;
    for (int t = i0; t < 6 * count; t += step) {
        const int m = t / 6, c = t % 6, d = c % 3;
        const double x = c < 3 ? old_at(m, d) : new_at(m, d);
        phase_chain(x * R.box, R.max_idx, tab + (size_t)t * stride);
    }
}

// rho_k of up to IPW moved beads m0, m0 + 1, ... (< count) of one wavefront,
// from their phase chains (tab[m][6][max_idx + 1], OLD x3 then NEW x3):
// rho_NEW = rho_OLD + e^{ik.r_new} - e^{ik.r_old} into the spare rows and the
// OLD / NEW sums of coef_s |rho|^2 into s_rho[m] (Species::UpdateRhoK,
// species_class.h:406-425, with CalcULong's per-slice sum).  A pass issues
// every load of its KPL k-vectors per lane for all the beads at once; the k
// indices and coefficients are shared by the beads.
template <int KPL, int IPW>
__device__ __forceinline__ void stage_rho_group(const Rep& R, const Prop& pr, int m0, const double2* tab,
                                                const unsigned char* s_sel, double2* s_rho, int lane) {
    const int stride = R.max_idx + 1, nk = R.nk, M = R.M;
    const double2* rc[IPW];
    double2* rn[IPW];
    bool on[IPW];
#pragma unroll
    for (int i = 0; i < IPW; ++i) {
        const int m = m0 + i;
        on[i] = m < pr.count;
        const int s = (pr.first + (on[i] ? m : 0)) % M, cur = s_sel[s];
        rc[i] = R.rho + ((size_t)cur * M + s) * nk;
        rn[i] = R.rho + ((size_t)(1 - cur) * M + s) * nk;
    }
    double ao[IPW], an[IPW];
#pragma unroll
    for (int i = 0; i < IPW; ++i) ao[i] = an[i] = 0.;
    for (int kb = 0; kb < nk; kb += KPL * kWave) {
        int kp[KPL];
        double cf[KPL];
        double2 rv[IPW][KPL];
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
            const int k = kb + j * kWave + lane;
            const bool ok = k < nk;
            kp[j] = ok ? gld(R.kidx + k) : (128 | (128 << 8) | (128 << 16));
            cf[j] = ok ? gld(R.coef_s + k) : 0.;
#pragma unroll
            for (int i = 0; i < IPW; ++i) rv[i][j] = (ok && on[i]) ? ld2<false>(rc[i] + k) : make_double2(0., 0.);
        }
#pragma unroll
        for (int i = 0; i < IPW; ++i) {
            if (!on[i]) continue;
            const double2* t = tab + (size_t)(m0 + i) * 6 * stride;
#pragma unroll
            for (int j = 0; j < KPL; ++j) {
                const int k = kb + j * kWave + lane;
                const double2 en = phase_k(t + 3 * stride, stride, kp[j]);
                const double2 eo = phase_k(t, stride, kp[j]);
                double2 nv;
                nv.x = rv[i][j].x + (en.x - eo.x);
                nv.y = rv[i][j].y + (en.y - eo.y);
                if (k < nk) st2<false>(rn[i] + k, nv);
                ao[i] += cf[j] * (rv[i][j].x * rv[i][j].x + rv[i][j].y * rv[i][j].y);
                an[i] += cf[j] * (nv.x * nv.x + nv.y * nv.y);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < IPW; ++i) {
        const double a = wave_sum(ao[i]), b = wave_sum(an[i]);
        if (lane == 0 && on[i]) s_rho[m0 + i] = make_double2(a, b);
    }
}

// In-wavefront ordering point for LDS data exchanged between lanes.
__device__ __forceinline__ void stage_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The draws of one attempt, prepared by one wavefront ahead of its use:
// tape: the attempt's draws staged in LDS (avail of them exist); device RNG:
// the bisection's normals in consumption order.  Both: the level uniforms (or
// the displacement's) and their logs, computed in parallel over the lanes.
struct StageDraws {
    double v[kStageMaxDraws];
    double u[kMaxLevel + 1], lu[kMaxLevel + 1];
    double dr[3];
    int p, b0, avail, bad;
};

// Bisection draw layout on the tape: p, b0, then per level l = n-1 .. 0 its
// 3 nmid_l normals and one uniform; nmid_l = 2^n >> (l + 1) (none when the
// window is the whole closed path).
__device__ __forceinline__ int stage_uniform_at(int n, int level, bool wraps) {
    int o = 2;
    for (int l = n - 1; l > level; --l) o += 3 * (wraps ? 0 : ((1 << n) >> (l + 1))) + 1;
    return o + 3 * (wraps ? 0 : ((1 << n) >> (level + 1)));
}

template <bool TAPE>
__device__ __forceinline__ void stage_prepare(const Rep& R, int move, unsigned long long seq, long long cur,
                                              StageDraws& D, int lane) {
    const int n = R.n_level, N = R.N, M = R.M;
    if (TAPE) {
        stage_load_draws(R, cur, stage_draws(move, n), D.v, &D.avail, lane);
        stage_wave_sync();
        const int avail = D.avail;
        if (move == 0) {
            const bool wraps = ((1 << n) % M) == 0;
            if (lane <= n && lane >= 1) {
                const int level = lane - 1;
                int bad = 0;
                const double u = stage_draw(D.v, avail, stage_uniform_at(n, level, wraps), &bad);
                D.u[level] = u;
                D.lu[level] = log(u);
            }
            if (lane == 0) {
                int bad = 0;
                int p = (int)stage_draw(D.v, avail, 0, &bad);
                int b0 = (int)stage_draw(D.v, avail, 1, &bad);
                if (p < 0 || p >= N || b0 < 0 || b0 >= M) {
                    bad = 1;
                    p = 0;
                    b0 = 0;
                }
                D.p = p;
                D.b0 = b0;
                D.bad = bad;
            }
        } else if (lane == 0) {
            int bad = 0;
            int p = (int)stage_draw(D.v, avail, 0, &bad);
            for (int d = 0; d < 3; ++d) D.dr[d] = stage_draw(D.v, avail, 1 + d, &bad);
            const double u0 = stage_draw(D.v, avail, 4, &bad);
            if (p < 0 || p >= N) {
                bad = 1;
                p = 0;
            }
            D.p = p;
            D.u[0] = u0;
            D.lu[0] = log(u0);
            D.bad = bad;
        }
    } else {
        if (move == 0) {
            Rng rng{make_uint2((uint32_t)R.key, (uint32_t)(R.key >> 32)), 1u, seq};
            const int n_norm = 3 * ((1 << n) - 1);
            // Box-Muller pairs: normals 2i and 2i + 1 come from Philox counter 1 + n + i
            for (int i = lane; 2 * i < n_norm; i += kWave) {
                uint4 w4 = rng.call(1 + n + i);
                const double u1 = u53(w4.x, w4.y), u2 = u53(w4.z, w4.w);
                const double rr = sqrt(-2. * log(u1));
                double sn, cs;
                sincospi(2. * u2, &sn, &cs);
                D.v[2 * i] = rr * cs;
                if (2 * i + 1 < n_norm) D.v[2 * i + 1] = rr * sn;
            }
            if (lane >= kWave - kMaxLevel - 1 && lane - (kWave - kMaxLevel - 1) < n) {
                const int level = lane - (kWave - kMaxLevel - 1);
                const double u = rng.unif(1 + level, 0);
                D.u[level] = u;
                D.lu[level] = log(u);
            }
            if (lane == 0) {
                D.p = (int)rng.index(0, 0, N);
                D.b0 = (int)rng.index(0, 1, M);
                D.bad = 0;
                D.avail = 0;
            }
        } else if (lane == 0) {
            Rng rng{make_uint2((uint32_t)R.key, (uint32_t)(R.key >> 32)), 2u, seq};
            D.p = (int)rng.index(0, 0, N);
            D.dr[0] = R.step * (rng.unif(1, 0) - 0.5);
            D.dr[1] = R.step * (rng.unif(1, 1) - 0.5);
            D.dr[2] = R.step * (rng.unif(2, 0) - 0.5);
            const double u0 = rng.unif(2, 1);
            D.u[0] = u0;
            D.lu[0] = log(u0);
            D.bad = 0;
            D.avail = 0;
        }
    }
}

// The coarse levels and the level-0 beads of Bisect::Attempt from prepared
// draws, one wavefront (propose_bisect_body's arithmetic and draw order).
// Writes the proposal to sp; returns nothing else.
template <bool TAPE>
__device__ __forceinline__ void stage_bisect_levels(const Rep& R, const StageDraws& D, long long cur0, Prop& sp,
                                                    double (*s_old)[3], double (*s_new)[3], int lane) {
    const int n = R.n_level, nb = 1 << n, M = R.M;
    const int p = D.p, b0 = D.b0, avail = D.avail;
    const bool wraps = (nb % M) == 0;
    for (int j = lane; j <= nb; j += kWave) {
        const int s = (b0 + j) % M;
        for (int d = 0; d < 3; ++d) {
            const double x = pos_at(R, d, s, p);
            s_old[j][d] = x;
            s_new[j][d] = x;
        }
    }
    stage_wave_sync();
    int ci = 2;       // tape position within the attempt
    int g_base = 0;   // normals drawn so far
    int bad = 0;
    double prev = 0.;
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
                t[d] = minimg(s_old[c][d] - s_old[a][d], R.L, R.iL);
                rbar[d] = s_old[a][d] + 0.5 * t[d];
                dold[d] = minimg(s_old[b][d] - rbar[d], R.L, R.iL);
            }
            for (int d = 0; d < 3; ++d) {
                const double z = TAPE ? stage_draw(D.v, avail, ci + 3 * lane + d, &bad) : D.v[g_base + 3 * lane + d];
                dnew[d] = sigma * z;
                dnew[d] = minimg(dnew[d], R.L, R.iL);
                t[d] = minimg(s_new[c][d] - s_new[a][d], R.L, R.iL);
                rbar[d] = s_new[a][d] + 0.5 * t[d];
                s_new[b][d] = rbar[d] + dnew[d];
            }
            lspo = -free_neg_log(&R.bis[level], R.bis_images, i4lt_b, dold);
            lspn = -free_neg_log(&R.bis[level], R.bis_images, i4lt_b, dnew);
        }
        stage_wave_sync();
        const int nlinks = wraps ? 0 : nb >> level;
        const double i4lt_k = 1. / (4. * R.lambda * level_tau);
        double ko = 0., kn = 0.;
        for (int i = lane; i < nlinks; i += kWave) {
            const int a = i * skip, b = a + skip;
            double dd[3];
            for (int d = 0; d < 3; ++d) dd[d] = minimg(s_old[a][d] - s_old[b][d], R.L, R.iL);
            ko += free_neg_log(&R.kin[level], R.kin_images, i4lt_k, dd);
            for (int d = 0; d < 3; ++d) dd[d] = minimg(s_new[a][d] - s_new[b][d], R.L, R.iL);
            kn += free_neg_log(&R.kin[level], R.kin_images, i4lt_k, dd);
        }
        lspo = wave_sum(lspo);
        lspn = wave_sum(lspn);
        ko = wave_sum(ko);
        kn = wave_sum(kn);
        const double lsr = -lspn + lspo;
        ci += 3 * nmid;
        g_base += 3 * nmid;
        if (TAPE && ci >= avail) bad = 1;  // the level's uniform is missing
        ci += 1;
        if (level > 0) {
            const double change = (0. + kn) - (0. + ko);
            const double la = lsr - change + prev;
            if (la < D.lu[level]) {
                if (lane == 0) {
                    sp.status = 2;
                    sp.move = 0;
                    sp.p = p;
                    sp.b0 = b0;
                    sp.n_links = 0;
                    sp.count = 0;
                    sp.tape_next = cur0 + ci;
                }
                return;
            }
            prev = change;
        } else {
            const int any_bad = __any(bad) || D.bad;  // every lane votes
            if (lane == 0) {
                sp.status = any_bad ? 3 : 1;
                sp.move = 0;
                sp.p = p;
                sp.p2 = p;
                sp.generic = 0;
                sp.b0 = b0;
                sp.n_links = R.pair_type ? nb : 0;
                sp.first = (b0 + 1) % M;
                sp.count = wraps ? 0 : nb - 1;
                sp.lsr = lsr;
                sp.k_old = ko;
                sp.k_new = kn;
                sp.prev = prev;
                sp.u0 = D.u[0];
                sp.tape_next = cur0 + ci;
            }
        }
    }
}

// PT, NORD as for k_pair_lds; use_lds: the pair tables stage into LDS (else the
// global-table arithmetic); move 0 Bisect, 1 DisplaceParticle.  Dynamic LDS:
// the pair tables (lds_tables doubles, when staged), then the moved beads'
// phase chains [moved bead][6][max_idx + 1] (double2).
template <int PT, int NORD, int BS, bool TAPE>
__device__ __forceinline__ void stage_body(const Rep& R, int move, bool tables, double* lds, double2* chains) {
    __shared__ StageCounters st;
    __shared__ Prop sp;
    __shared__ double s_old[(1 << kMaxLevel) + 1][3];
    __shared__ double s_new[(1 << kMaxLevel) + 1][3];
    __shared__ double2 s_pair[kStageMaxM];
    __shared__ double2 s_rho[kStageMaxM];
    __shared__ unsigned char s_sel[kStageMaxM];  // committed rho_k buffer of each slice
    __shared__ StageDraws s_d[2];
    __shared__ int s_go[2];
    __shared__ int s_acc;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    constexpr int kWaves = BS / kWave;
    // a private copy for the evaluation and the decision: its fields stay in
    // registers across the stores of the loop (never indexed at run time, which
    // the proposal's per-level spline arrays are: those go through R)
    const Rep Rv = R;
    // with several wavefronts the first one prepares the next attempt's draws
    // while the others evaluate the current attempt
    constexpr int kEval0 = kWaves > 1 ? 1 : 0, kEvalN = kWaves > 1 ? kWaves - 1 : 1;
    const int M = R.M, N = R.N, n = R.n_level;
    if (threadIdx.x == 0) {
        Counters* c = R.cnt;
        st.remaining = c->remaining;
        st.stage_i = c->stage_i;
        st.seq = c->seq;
        st.cursor = TAPE ? *R.tape_cursor : 0;
        st.att = st.acc = st.beads = st.rho_slicek = st.pair_links = 0;
        st.error = 0;
        s_go[0] = st.remaining > 0;
    }
    if (R.use_long_range)
        for (int t = threadIdx.x; t < M; t += BS) s_sel[t] = R.sel[t];
    __syncthreads();
    if (w == 0 && s_go[0]) stage_prepare<TAPE>(R, move, st.seq, st.cursor, s_d[0], lane);
    __syncthreads();
    for (int t = 0;; ++t) {
        const int b = t & 1;
        // ---- A: proposal (first wavefront) ----------------------------------
        if (w == 0) {
            if (!s_go[b]) {
                if (lane == 0) sp.status = 0;
            } else {
                const StageDraws& D = s_d[b];
                if (move == 0) {
                    stage_bisect_levels<TAPE>(R, D, st.cursor, sp, s_old, s_new, lane);
                } else if (lane == 0) {
                    sp.status = D.bad ? 3 : 1;
                    sp.move = 1;
                    sp.p = D.p;
                    sp.p2 = D.p;
                    sp.generic = 0;
                    sp.b0 = 0;
                    sp.n_links = R.pair_type ? M : 0;
                    sp.first = 0;
                    sp.count = M;
                    sp.lsr = 0.;
                    sp.k_old = 0.;  // the walk over [0, M) of an unpermuted path holds no link
                    sp.k_new = 0.;
                    sp.prev = 0.;
                    sp.u0 = D.u[0];
                    sp.tape_next = st.cursor + (TAPE ? 5 : 0);
                }
                if (lane == 0) {
                    st.remaining -= 1;
                    s_go[b ^ 1] = st.remaining > 0;
                }
            }
        }
        __syncthreads();
        const Prop pr = sp;
        if (pr.status == 0) break;  // no attempt left: the same in every wavefront
        const StageDraws& D = s_d[b];
        // ---- B1: the next attempt's draws (first wavefront); the phase chains
        //      of the moved beads and the pair links (the others) -----------
        if (w == 0 && kWaves > 1) {
            if (s_go[b ^ 1]) stage_prepare<TAPE>(R, move, st.seq + 1, TAPE ? pr.tape_next : 0, s_d[b ^ 1], lane);
        } else if (pr.status == 1) {
            const int p = pr.p, we = w - kEval0;
            // moved bead m: window bead m + 1 of a bisection, slice m of a displacement
            if (move == 0) {
                if (R.use_long_range)
                    stage_phase_chains(
                        Rv, pr.count, [&](int m, int d) { return s_old[m + 1][d]; },
                        [&](int m, int d) { return s_new[m + 1][d]; }, chains, we * kWave + lane, kEvalN * kWave);
                const int per = (pr.n_links + kEvalN - 1) / kEvalN;
                stage_pair_links<PT, NORD>(
                    Rv, lds, tables, pr, [&](int j, int d) { return s_old[j][d]; },
                    [&](int j, int d) { return s_new[j][d]; }, we * per, min(pr.n_links, (we + 1) * per), lane,
                    s_pair);
            } else {
                if (R.use_long_range)
                    stage_phase_chains(
                        Rv, pr.count, [&](int m, int d) { return pos_at(Rv, d, m, p); },
                        [&](int m, int d) { return pos_at(Rv, d, m, p) + D.dr[d]; }, chains, we * kWave + lane,
                        kEvalN * kWave);
                const int per = (pr.n_links + kEvalN - 1) / kEvalN;
                stage_pair_links<PT, NORD>(
                    Rv, lds, tables, pr, [&](int j, int d) { return pos_at(Rv, d, j % M, p); },
                    [&](int j, int d) { return pos_at(Rv, d, j % M, p) + D.dr[d]; }, we * per,
                    min(pr.n_links, (we + 1) * per), lane, s_pair);
            }
        }
        __syncthreads();
        // ---- B2: rho_k of the moved slices (every wavefront) -----------------
        if (pr.status == 1 && R.use_long_range) {
            constexpr int kIpw = 4;
            for (int m0 = w * kIpw; m0 < pr.count; m0 += kWaves * kIpw)
                stage_rho_group<5, kIpw>(Rv, pr, m0, chains, s_sel, s_rho, lane);
        }
        __syncthreads();
        // ---- C: level-0 decision, commit, counters (first wavefront) --------
        if (w == 0) {
            double po = 0., pn = 0., lo = 0., ln = 0.;
            if (pr.status == 1) {
                for (int j = lane; j < pr.n_links; j += kWave) {
                    po += s_pair[j].x;
                    pn += s_pair[j].y;
                }
                if (R.use_long_range)
                    for (int j = lane; j < pr.count; j += kWave) {
                        lo += s_rho[j].x;
                        ln += s_rho[j].y;
                    }
                po = wave_sum(po);
                pn = wave_sum(pn);
                lo = wave_sum(lo);
                ln = wave_sum(ln);
            }
            if (lane == 0) {
                int acc = 0;
                if (pr.status == 1) {
                    const double pair_old = po + R.long_scale * lo, pair_new = pn + R.long_scale * ln;
                    const double old_action = (0. + pr.k_old) + pair_old;
                    const double new_action = (0. + pr.k_new) + pair_new;
                    const double change = new_action - old_action;
                    const double la = pr.lsr - change + pr.prev;
                    acc = !(la < D.lu[0]);
                    st.rho_slicek += (unsigned long long)pr.count * (unsigned long long)R.nk;
                    if (R.pair_type) st.pair_links += 2ull * (unsigned long long)pr.n_links * (unsigned long long)(N - 1);
                }
                s_acc = acc;
                st.att += 1;
                st.acc += acc;
                st.beads += move == 0 ? (1ull << n) - 1 : (unsigned long long)M;
                const unsigned int si = st.stage_i;
                if (si < R.accepted_cap) R.accepted[si] = (unsigned char)acc;
                if (R.act_trace && si < R.accepted_cap) {
                    double* tr = R.act_trace + (size_t)si * kTraceWords;
                    const bool l0 = pr.status == 1;
                    const double nan = __builtin_nan("");
                    tr[0] = l0 ? pr.k_old : nan;
                    tr[1] = l0 ? pr.k_new : nan;
                    tr[2] = l0 ? po : nan;
                    tr[3] = l0 ? pn : nan;
                    tr[4] = l0 ? R.long_scale * lo : nan;
                    tr[5] = l0 ? R.long_scale * ln : nan;
                    tr[6] = l0 ? pr.lsr : nan;
                    tr[7] = l0 ? pr.prev : nan;
                }
                st.stage_i = si + 1;
                st.seq += 1;
                if (TAPE) st.cursor = pr.tape_next;
                if (pr.status == 3) st.error = 1;
            }
            stage_wave_sync();
            if (s_acc) {
                const int p = pr.p;
                for (int m = lane; m < pr.count; m += kWave) {
                    const int s = (pr.first + m) % M;
                    for (int d = 0; d < 3; ++d) {
                        double* x = R.pos + ((size_t)d * M + s) * N + p;
                        *x = move == 0 ? s_new[m + 1][d] : *x + D.dr[d];
                    }
                    if (R.use_long_range) {
                        s_sel[s] ^= 1;
                        R.sel[s] = s_sel[s];
                    }
                }
            }
        }
        // without a spare wavefront the next attempt's draws are prepared here
        if (kWaves == 1 && s_go[b ^ 1]) stage_prepare<TAPE>(R, move, st.seq, st.cursor, s_d[b ^ 1], lane);
        __syncthreads();
    }
    // ---- the stage's counters, once -----------------------------------------
    if (threadIdx.x == 0) {
        Counters* c = R.cnt;
        c->remaining = 0;
        c->stage_i = st.stage_i;
        c->seq = st.seq;
        if (move == 0) {
            c->bis_att += st.att;
            c->bis_acc += st.acc;
        } else {
            c->dis_att += st.att;
            c->dis_acc += st.acc;
        }
        c->bead_updates += st.beads;
        c->rho_slicek += st.rho_slicek;
        c->pair_links += st.pair_links;
        if (TAPE) *R.tape_cursor = st.cursor;
        if (st.error) {
            c->error = 1;
            atomicOr(R.ctx_error, 1u);
        }
        R.prop->status = 0;
    }
}

template <int PT, int NORD, int BS>
__global__ __launch_bounds__(BS) void k_stage_persistent(const Rep* __restrict__ reps, int n_reps, int move,
                                                         int use_lds, int lds_tables) {
    extern __shared__ double lds[];
    const int r = blockIdx.x;
    if (r >= n_reps) return;
    // by reference: the proposal indexes the per-level spline arrays with a
    // run-time level, which a private copy would send to scratch memory
    const Rep& R = reps[r];
    const bool tables = use_lds && R.lds_doubles > 0;
    double2* chains = reinterpret_cast<double2*>(lds + lds_tables);
    if (tables) {
        const int na = R.bt_a.ncoef * R.bt_a.ncomp;
        for (int t = threadIdx.x; t < R.lds_doubles; t += BS) lds[t] = t < na ? R.bt_a.c[t] : R.bt_b.c[t - na];
    }
    if (R.tape)
        stage_body<PT, NORD, BS, true>(R, move, tables, lds, chains);
    else
        stage_body<PT, NORD, BS, false>(R, move, tables, lds, chains);
}

}  // namespace spimc
