"""GPU parity of permutation sampling (PermBisectIterative,
perm_bisect_iterative_class.h:10-233, perm_bisect_class.h:7-135) and of the
moves and estimators on a permuted path, against the CPU oracle.

* Replay: the oracle's draws are fed to the engine; the accept/reject
  sequences, positions (relabelled slices included), the wrap permutation and
  the move counters must match bit for bit, rho_k to 1e-11 relative.  Stages of
  permuting bisections are interleaved with Bisect and DisplaceParticle, which
  then run on permuted paths (a bisection chain that changes label at the wrap,
  the one-sided wrap link of a displaced particle).
* Estimators on a fixed permuted path: Kinetic::GetAction windows across the
  wrap, DActionDBeta, VirialEnergy and the adjoint's kinetic term, to 1e-10
  relative (1e-6 against the central difference for the gradient).
* Free running: free bosons sampled with the device RNG reproduce the ideal
  Bose gas closed form (energy and mean k-cycle counts), like the oracle's
  CPU test, within 3.5 standard errors over independent replicas.
"""
import numpy as np
import pytest

from oracle_ffi import Oracle
from simpimc_amd import abi, system
from test_permutation_cpu import BETA, LAM, L, M, N, free_system, ideal_bose

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    return float(np.max(np.abs(a - b))) / scale if a.size else 0.0


def dense_bare():
    # 8 electrons at r_s = 1, beta = 0.8: neighbours within a few thermal lengths
    return system.config_c1(n_part=8, n_bead=16, r_s=1.0, tau=0.05, n_level=3)


def small_david():
    return system.config_c2(n_part=16, n_bead=16, r_s=1.0, tau=0.02, n_level=3, k_shell=20.5, n_grid=400, seed=7)


SYSTEMS = {
    "bare_dense": dense_bare,
    "david_small": small_david,
    "ilkka_c1": lambda: system.config_ilkka(),
    "free_bose": free_system,
}


REPLAY = [(name, "lds") for name in sorted(SYSTEMS)] + [("bare_dense", "global"), ("david_small", "global")]


# "lds": proposals of a few chains are bisected by one wavefront on LDS-staged
# beads; "global": SPIMC_PERM_STAGE=0 sends every proposal to the workgroup-wide
# form on global memory, which otherwise serves only proposals of more than 256
# chain beads
@pytest.mark.parametrize("name,form", REPLAY)
def test_replay_with_permutations_bit_exact(name, form, monkeypatch):
    if form == "global":
        monkeypatch.setenv("SPIMC_PERM_STAGE", "0")
    s = SYSTEMS[name]()
    pos = system.random_path(s, seed=3, spread=0.3)
    o = Oracle(s, seed=11)
    o.upload(pos)
    o.record(True)
    n_b = max(1, s.n_part * s.n_bead >> s.bisect_n_level)
    ref = []
    for _ in range(4):
        ref.append(o.perm_bisect(n_b))
        ref.append(o.bisect(n_b))
        ref.append(o.displace(s.n_part))
    tape = o.recorded()
    print(f"{name}: {len(tape)} draws", flush=True)
    with abi.Engine(s, seed=99) as e:
        e.upload(pos)
        e.tape_set(tape)
        got = []
        for _ in range(4):
            got.append(e.perm_bisect(n_b)[0])
            got.append(e.bisect(n_b)[0])
            got.append(e.displace(s.n_part)[0])
        assert e.tape_cursor() == len(tape)
        for i, (g, r) in enumerate(zip(got, ref)):
            assert np.array_equal(g, r), (i, g, r)
        assert np.array_equal(e.perm(), o.perm())
        assert np.array_equal(e.download(), o.download())
        if s.n_k:
            assert rel_err(e.rho_k(), o.rho_k()) < 1e-11 * s.n_part
        assert np.array_equal(e.counters(), o.counters())
        eg, og = e.perm_counters(), o.perm_counters()
        assert eg[0] == og[0] and eg[1] == og[1]
        assert np.array_equal(eg[2], og[2]) and np.array_equal(eg[3], og[3])
        assert e.bead_updates() == o.bead_updates()
        sigma = o.perm()
        print(f"{name}: perm accepts per cycle length {og[3].tolist()}, sigma {sigma.tolist()}", flush=True)
    if name == "bare_dense":
        # the run relinked the path: multi-particle cycles were accepted
        assert og[3][1:].sum() > 0


PERMS = {4: [2, 0, 1, 3], 8: [1, 0, 3, 4, 2, 5, 7, 6], 16: list(range(1, 16)) + [0]}


def permuted(name):
    s = SYSTEMS[name]()
    pos = system.random_path(s, seed=5, spread=0.3)
    sigma = np.array(PERMS[s.n_part] if s.n_part in PERMS else np.roll(np.arange(s.n_part), 1), dtype=np.uint32)
    return s, pos, sigma


@pytest.mark.parametrize("name", ["bare_dense", "david_small"])
def test_estimators_on_a_permuted_path(name):
    s, pos, sigma = permuted(name)
    o = Oracle(s)
    o.upload(pos)
    o.set_perm(sigma)
    with abi.Engine(s) as e:
        e.upload(pos)
        e.set_perm(sigma)
        assert np.array_equal(e.perm(), sigma)
        for p in range(s.n_part):
            for b0, span, lvl in ((s.n_bead - 3, 4, 0), (s.n_bead - 4, 8, 1), (0, s.n_bead, 0), (5, s.n_bead, 0)):
                a = e.get_action(abi.ACTION_KINETIC, abi.MODE_NEW, b0, b0 + span, [p], lvl)
                b = o.get_action(abi.ACTION_KINETIC, abi.MODE_NEW, b0, b0 + span, [p], lvl)
                assert abs(a - b) <= 1e-10 * max(1.0, abs(b)), (p, b0, span, a, b)
        a, b = e.d_action_d_beta(abi.ACTION_KINETIC), o.d_action_d_beta(abi.ACTION_KINETIC)
        assert abs(a - b) <= 1e-10 * max(1.0, abs(b)), (a, b)
        for w in (0, 4):
            a, b = e.kinetic_virial_energy(w), o.kinetic_virial_energy(w)
            assert abs(a - b) <= 1e-10 * max(1.0, abs(b)), (w, a, b)
        g = e.gradient(abi.GRAD_ALL)
        ref = o.gradient(abi.GRAD_ALL, compat=False, h=1e-5)
        assert rel_err(g, ref) < 1e-6
        # the same path with the identity permutation differs on the wrap links
        g_id = e.gradient(abi.GRAD_KINETIC)
        e.set_perm(np.arange(s.n_part))
        assert rel_err(g_id, e.gradient(abi.GRAD_KINETIC)) > 1e-3


def test_free_bosons_match_the_closed_form():
    s = free_system()
    n_rep, n_eq, n_meas, every = 64, 300, 2500, 5
    e_exact, nk_exact = ideal_bose(N, BETA, L, LAM)
    with abi.Engine([s] * n_rep, seed=4242) as e:
        for r in range(n_rep):
            e.upload(system.random_path(s, seed=300 + r), r=r)
        e.obs_configure(energy=abi.ENERGY_THERMAL)
        for _ in range(n_eq):
            e.perm_bisect(3)
            e.bisect(3)
        e.obs_reset(abi.OBS_ENERGY)
        cyc = np.zeros((n_rep, N))
        n_cyc = 0
        for it in range(n_meas):
            e.perm_bisect(3)
            e.bisect(3)
            e.obs_accumulate(abi.OBS_ENERGY)
            if it % every == 0:
                n_cyc += 1
                for r in range(n_rep):
                    sig = e.perm(r=r)
                    seen = np.zeros(N, dtype=bool)
                    for p in range(N):
                        if seen[p]:
                            continue
                        ln, q = 0, p
                        while not seen[q]:
                            seen[q] = True
                            q = sig[q]
                            ln += 1
                        cyc[r, ln - 1] += 1
            if it % 500 == 0:
                print(f"free bosons: sample {it} of {n_meas}", flush=True)
        en = []
        for r in range(n_rep):
            v, nm = e.obs_read(abi.OBS_ENERGY, r=r)
            assert nm == n_meas
            en.append((v[0] + v[1]) / (M * nm))
    en = np.array(en)
    em, ese = en.mean(), en.std(ddof=1) / np.sqrt(n_rep)
    print(f"energy {em:.5f} +- {ese:.5f}, exact {e_exact:.5f}", flush=True)
    assert abs(em - e_exact) < 3.5 * ese
    nk = cyc / n_cyc
    nkm, nkse = nk.mean(axis=0), nk.std(axis=0, ddof=1) / np.sqrt(n_rep)
    print(f"cycles {nkm} +- {nkse}, exact {nk_exact}", flush=True)
    for k in range(N):
        assert abs(nkm[k] - nk_exact[k]) < 3.5 * nkse[k] + 1e-3, (k, nkm, nkse, nk_exact)


def test_uneven_replicas_with_permutations():
    """Replicas of different systems and sizes in one context, each with its own
    number of attempts, replay their single chains' tapes bit for bit through
    permuting, bisect and displace stages."""
    systems = [dense_bare(), small_david(), system.config_ilkka(n_part=24, n_bead=8, n_level=2), free_system()]
    counts = np.array([30, 20, 25, 12], dtype=np.uint32)
    tapes, refs, finals, perms = [], [], [], []
    for i, s in enumerate(systems):
        pos = system.random_path(s, seed=70 + i, spread=0.3)
        o = Oracle(s, seed=300 + i)
        o.upload(pos)
        o.record(True)
        refs.append((o.perm_bisect(int(counts[i])), o.bisect(int(counts[i])), o.displace(int(counts[i]) // 2)))
        tapes.append((pos, o.recorded()))
        finals.append(o.download())
        perms.append(o.perm())
    with abi.Engine(systems, seed=5) as e:
        for i, (pos, tape) in enumerate(tapes):
            e.upload(pos, r=i)
            e.tape_set(tape, r=i)
        ap = e.perm_bisect(counts)
        ab = e.bisect(counts)
        ad = e.displace(counts // 2)
        for i in range(len(systems)):
            n = int(counts[i])
            assert np.array_equal(ap[i, :n], refs[i][0]), i
            assert np.array_equal(ab[i, :n], refs[i][1]), i
            assert np.array_equal(ad[i, :n // 2], refs[i][2]), i
            assert np.array_equal(e.download(r=i), finals[i]), i
            assert np.array_equal(e.perm(r=i), perms[i]), i
            assert e.tape_cursor(r=i) == len(tapes[i][1]), i


def test_graph_sweep_on_permuted_paths():
    """On permuted paths (after permuting stages on the device RNG) a sweep
    replayed from captured graphs, which then also hold the label-general pair
    kernel, equals the eager sweep of an identical context."""
    s = dense_bare()
    R = 6
    flags = abi.SWEEP_ENERGY | abi.SWEEP_SK | abi.SWEEP_GR | abi.SWEEP_ADJOINT
    out = []
    for graph in (False, True):
        with abi.Engine([s] * R, seed=77) as e:
            for r in range(R):
                e.upload(system.random_path(s, seed=90 + r, spread=0.3), r=r)
            e.obs_configure(energy=abi.ENERGY_VIRIAL_ADJOINT, sk_k_cut=1e300, gr_n_r=50)
            for _ in range(3):
                e.perm_bisect(40)
            e.sweep(2, flags | (abi.SWEEP_GRAPH if graph else 0))
            out.append(([e.download(r=r) for r in range(R)], [e.perm(r=r) for r in range(R)],
                        [e.counters(r=r) for r in range(R)], [e.obs_read(abi.OBS_ENERGY, r=r)[0] for r in range(R)]))
    (p0, s0, c0, e0), (p1, s1, c1, e1) = out
    assert any(not np.array_equal(x, np.arange(s.n_part)) for x in s0), "no replica was permuted"
    for r in range(R):
        assert np.array_equal(p0[r], p1[r]), r
        assert np.array_equal(s0[r], s1[r]), r
        assert np.array_equal(c0[r], c1[r]), r
        assert np.array_equal(e0[r], e1[r]), r


EDGE = {
    # one particle: every cycle is a 1-cycle and the pair action is constant
    "one_particle": (lambda: system.config_c1(n_part=1, n_bead=8, r_s=1.0, tau=0.1, n_level=2), None, None),
    # 2^n_level = M on a permuted path: the window runs once around into the next particle
    "window_is_path": (lambda: system.config_c1(n_part=4, n_bead=8, r_s=1.0, tau=0.1, n_level=3), None, [1, 2, 3, 0]),
    # the permuting level above the bisection's: one-chain proposals outgrow the
    # bisection's rho_k items and take the label-general rho_k kernel
    "perm_level_above_bisect": (lambda: system.config_c1(n_part=8, n_bead=16, r_s=1.0, tau=0.05, n_level=2), 3,
                                [1, 0, 2, 3, 5, 4, 6, 7]),
    # two particles: the only cycles are the identity and the swap
    "two_particles": (lambda: system.config_c1(n_part=2, n_bead=16, r_s=1.0, tau=0.1, n_level=2), None, None),
}


@pytest.mark.parametrize("case", sorted(EDGE))
def test_permutation_edge_cases(case):
    make, perm_level, sigma = EDGE[case]
    s = make()
    pos = system.random_path(s, seed=17, spread=0.3)
    o = Oracle(s, seed=23)
    o.upload(pos)
    if sigma is not None:
        o.set_perm(sigma)
    if perm_level is not None:
        o.perm_configure(perm_level)
    o.record(True)
    n = max(16, s.n_part * s.n_bead >> s.bisect_n_level)
    ref = []
    for _ in range(3):
        ref.append(o.perm_bisect(n))
        ref.append(o.bisect(n))
        ref.append(o.displace(max(1, s.n_part)))
    tape = o.recorded()
    print(f"{case}: {len(tape)} draws, perm accepts {o.perm_counters()[3].tolist()}", flush=True)
    with abi.Engine(s, seed=1) as e:
        e.upload(pos)
        if sigma is not None:
            e.set_perm(sigma)
        if perm_level is not None:
            e.perm_configure(perm_level)
        e.tape_set(tape)
        got = []
        for _ in range(3):
            got.append(e.perm_bisect(n)[0])
            got.append(e.bisect(n)[0])
            got.append(e.displace(max(1, s.n_part))[0])
        assert e.tape_cursor() == len(tape)
        for i, (g, r) in enumerate(zip(got, ref)):
            assert np.array_equal(g, r), (i, g, r)
        assert np.array_equal(e.perm(), o.perm())
        assert np.array_equal(e.download(), o.download())
        if s.n_part > 1:
            assert rel_err(e.rho_k(), o.rho_k()) < 1e-11 * s.n_part
        else:
            # one particle: the pair action is constant, so the reference never
            # updates rho_k (pair_action_class.h:225,269) and the oracle's stays at
            # the start; the engine keeps it in step with the beads.  |rho_k|^2 = 1,
            # all that S(k) reads, is the same in both.
            x = e.download()
            want = np.exp(1j * np.einsum("kd,dm->mk", 2 * np.pi * s.k_n / s.L, x[:, :, 0]))
            assert np.abs(e.rho_k() - want).max() < 1e-11
            assert np.allclose(np.abs(o.rho_k()), 1.0, atol=1e-13)
        assert np.array_equal(e.counters(), o.counters())
        eg, og = e.perm_counters(), o.perm_counters()
        assert eg[:2] == og[:2] and np.array_equal(eg[2], og[2]) and np.array_equal(eg[3], og[3])
        assert e.bead_updates() == o.bead_updates()
        a, b = e.d_action_d_beta(abi.ACTION_KINETIC), o.d_action_d_beta(abi.ACTION_KINETIC)
        assert abs(a - b) <= 1e-10 * max(1.0, abs(b)), (a, b)
    if case == "two_particles":
        assert og[3][1] > 0  # swaps were accepted


def test_walk_without_candidates_is_rejected_free_running():
    """Free-running (device RNG), a cycle walk whose row has no candidate above
    epsilon rejects the attempt instead of failing the stage; replaying the
    oracle's draws, the same walk is an error on both sides."""
    from test_permutation_cpu import degenerate_rows_system

    s, pos = degenerate_rows_system()
    with abi.Engine([s, s], seed=21) as e:
        for r in range(2):
            e.upload(pos, r=r)
            e.perm_configure(3, epsilon=0.5, r=r)
        acc = e.perm_bisect(200)
        for r in range(2):
            att, n_acc, _, _ = e.perm_counters(r)
            assert att == 200 and n_acc == int(acc[r].sum())
    o = Oracle(s, seed=5)
    o.upload(pos)
    o.perm_configure(3, epsilon=0.5)
    o.record(True)
    o.perm_bisect(200)
    tape = o.recorded()
    with abi.Engine(s, seed=1) as e:
        e.upload(pos)
        e.perm_configure(3, epsilon=0.5)
        e.tape_set(tape)
        with pytest.raises(abi.SpimcError):
            e.perm_bisect(200)


@pytest.mark.parametrize("table", [False, True])
def test_lds_and_global_forms_agree_on_the_device_rng(table, monkeypatch):
    """The permuting proposal's two forms (a few chains bisected by one
    wavefront on LDS-staged beads, with the attempt's normals and level
    uniforms drawn up front; the workgroup-wide form on global memory, which
    draws each where it is used, SPIMC_PERM_STAGE=0) take the same draws in
    the same order: free-running on the device RNG they give the same chains
    bit for bit, for PermBisectIterative and for PermBisectTable."""
    s = dense_bare()
    out = []
    for stage in ("256", "0"):
        monkeypatch.setenv("SPIMC_PERM_STAGE", stage)
        with abi.Engine([s] * 4, seed=606) as e:
            for r in range(4):
                e.upload(system.random_path(s, seed=60 + r, spread=0.3), r=r)
                if table:
                    e.perm_table_configure(3, 1, r=r)
            for _ in range(6):
                (e.perm_table if table else e.perm_bisect)(20)
                e.bisect(16)
            out.append(([e.download(r=r) for r in range(4)], [e.perm(r=r) for r in range(4)],
                        [e.perm_counters(r=r) for r in range(4)]))
    (p0, s0, c0), (p1, s1, c1) = out
    moved = 0
    for r in range(4):
        assert np.array_equal(p0[r], p1[r]), r
        assert np.array_equal(s0[r], s1[r]), r
        assert c0[r][0] == c1[r][0] and c0[r][1] == c1[r][1], r
        assert np.array_equal(c0[r][3], c1[r][3]), r
        moved += int(c0[r][3][1:].sum())
    assert moved > 0  # multi-particle cycles were accepted
