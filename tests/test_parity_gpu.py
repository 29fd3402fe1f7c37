"""GPU parity: the HIP engine against the CPU oracle on the same seeded inputs.

Tolerances: positions, accept/reject sequences and g(r) counts are compared
bit for bit; actions, action differences, rho_k, gradients and estimators to
1e-10 relative to the magnitude of the compared quantity (the engine and the
oracle sum in different orders and use different libm/device sin, cos, exp).
The analytic adjoint is compared with a central difference (h = 1e-5) to 1e-6.
"""
import numpy as np
import pytest

from simpimc_amd import abi, system
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    return float(np.max(np.abs(a - b))) / scale if a.size else 0.0


def small_david(seed=7):
    return system.config_c2(n_part=16, n_bead=16, r_s=1.0, tau=0.02, n_level=3, k_shell=20.5, n_grid=400, seed=seed)


SYSTEMS = {
    "bare_c1": lambda: system.config_c1(),
    "david_small": small_david,
    "ilkka_c1": lambda: system.config_ilkka(),
}


@pytest.fixture(params=sorted(SYSTEMS))
def sysname(request):
    return request.param


def make(sysname, seed=3, spread=0.3):
    s = SYSTEMS[sysname]()
    pos = system.random_path(s, seed=seed, spread=spread)
    return s, pos


def test_rho_k_build(sysname):
    s, pos = make(sysname)
    o = Oracle(s)
    o.upload(pos)
    with abi.Engine(s) as e:
        e.upload(pos)
        got = e.rho_k()
    ref = o.rho_k()
    assert got.shape == ref.shape
    assert rel_err(got, ref) < 1e-12 * s.n_part


def test_get_action_interface(sysname):
    s, pos = make(sysname)
    o = Oracle(s)
    o.upload(pos)
    rng = np.random.default_rng(5)
    with abi.Engine(s) as e:
        e.upload(pos)
        p, b0 = 5, 3
        nb = 1 << s.bisect_n_level
        new = pos[:, (b0 + 1 + np.arange(nb - 1)) % s.n_bead, p].T + rng.normal(0, 0.2, size=(nb - 1, 3))
        for eng in (e, o):
            eng.set_proposal(p, b0 + 1, new)
        for action in (abi.ACTION_KINETIC, abi.ACTION_PAIR):
            for level in range(s.bisect_n_level):
                for mode in (abi.MODE_OLD, abi.MODE_NEW):
                    a = e.get_action(action, mode, b0, b0 + nb, [p], level)
                    b = o.get_action(action, mode, b0, b0 + nb, [p], level)
                    assert abs(a - b) <= RTOL * max(1.0, abs(b)), (action, level, mode, a, b)
        # displacement window: all slices
        for mode in (abi.MODE_OLD, abi.MODE_NEW):
            a = e.get_action(abi.ACTION_PAIR, mode, 0, s.n_bead, [p], 0)
            b = o.get_action(abi.ACTION_PAIR, mode, 0, s.n_bead, [p], 0)
            assert abs(a - b) <= RTOL * max(1.0, abs(b))
        e.accept()
        o.accept()
        assert np.array_equal(e.download(), o.download())
        assert rel_err(e.rho_k(), o.rho_k()) < 1e-12 * s.n_part


def test_whole_path_estimators(sysname):
    s, pos = make(sysname)
    o = Oracle(s)
    o.upload(pos)
    with abi.Engine(s) as e:
        e.upload(pos)
        for action in (abi.ACTION_KINETIC, abi.ACTION_PAIR):
            a, b = e.d_action_d_beta(action), o.d_action_d_beta(action)
            assert abs(a - b) <= RTOL * max(1.0, abs(b)), (action, a, b)
        a, b = e.potential(abi.ACTION_PAIR), o.potential(abi.ACTION_PAIR)
        assert abs(a - b) <= RTOL * max(1.0, abs(b))
        for w in (0, 4):
            a, b = e.kinetic_virial_energy(w), o.kinetic_virial_energy(w)
            assert abs(a - b) <= RTOL * max(1.0, abs(b)), (w, a, b)


@pytest.mark.parametrize("one", ["0", "1", "2"])
def test_replay_accept_reject_bit_exact(sysname, one, monkeypatch):
    """Bisection and displace stages replay the oracle's tape bit for bit, with
    rounds of three launches and with rounds of one launch (SPIMC_ONE_LAUNCH 1:
    bisection, 2: both moves; the global-table replica keeps three)."""
    monkeypatch.setenv("SPIMC_ONE_LAUNCH", one)
    s, pos = make(sysname, spread=0.05)
    o = Oracle(s, seed=11)
    o.upload(pos)
    o.record(True)
    n_b = s.n_part * s.n_bead >> s.bisect_n_level
    ref = []
    for _ in range(3):
        ref.append(o.bisect(n_b))
        ref.append(o.displace(s.n_part))
    tape = o.recorded()
    with abi.Engine(s, seed=99) as e:
        e.upload(pos)
        e.tape_set(tape)
        got = []
        for _ in range(3):
            got.append(e.bisect(n_b)[0])
            got.append(e.displace(s.n_part)[0])
        assert e.tape_cursor() == len(tape)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)
        assert np.array_equal(e.download(), o.download())
        assert rel_err(e.rho_k(), o.rho_k()) < 1e-11 * s.n_part
        assert np.array_equal(e.counters(), o.counters())


@pytest.mark.parametrize("block", [256, 1024])
def test_persistent_stage_replay_bit_exact(sysname, block, monkeypatch):
    """Whole stages in one launch of the persistent stage kernel (one workgroup
    per replica looping over its attempts) replay the oracle's chain bit for
    bit, three single replicas and the three batched together."""
    monkeypatch.setenv("SPIMC_PERSISTENT", str(block))
    s, pos = make(sysname, spread=0.05)
    o = Oracle(s, seed=11)
    o.upload(pos)
    o.record(True)
    n_b = s.n_part * s.n_bead >> s.bisect_n_level
    ref = []
    for _ in range(2):
        ref.append(o.bisect(n_b))
        ref.append(o.displace(s.n_part))
    tape = o.recorded()
    with abi.Engine(s, seed=99) as e:
        e.upload(pos)
        e.tape_set(tape)
        for i in range(2):
            assert np.array_equal(e.bisect(n_b)[0], ref[2 * i])
            assert np.array_equal(e.displace(s.n_part)[0], ref[2 * i + 1])
        assert e.tape_cursor() == len(tape)
        assert np.array_equal(e.download(), o.download())
        assert rel_err(e.rho_k(), o.rho_k()) < 1e-11 * s.n_part
        assert np.array_equal(e.counters(), o.counters())


def test_persistent_stage_mixed_batch(monkeypatch):
    """Bare, David and Ilkka replicas in one persistent launch per stage (the
    global-table arithmetic, pair type read per replica) replay their chains."""
    monkeypatch.setenv("SPIMC_PERSISTENT", "512")
    systems = mixed_systems()
    tapes, refs, finals = [], [], []
    for i, s in enumerate(systems):
        pos = system.random_path(s, seed=60 + i, spread=0.05)
        o = Oracle(s, seed=200 + i)
        o.upload(pos)
        o.record(True)
        refs.append((o.bisect(40 + 8 * i), o.displace(8)))
        tapes.append((pos, o.recorded()))
        finals.append(o.download())
    with abi.Engine(systems, seed=1) as e:
        for i, (pos, tape) in enumerate(tapes):
            e.upload(pos, r=i)
            e.tape_set(tape, r=i)
        ab = e.bisect([40 + 8 * i for i in range(len(systems))])
        ad = e.displace(8)
        for i in range(len(systems)):
            assert np.array_equal(ab[i, :40 + 8 * i], refs[i][0]), i
            assert np.array_equal(ad[i], refs[i][1]), i
            assert np.array_equal(e.download(r=i), finals[i]), i


def test_gradient_compat(sysname):
    s, pos = make(sysname)
    o = Oracle(s)
    o.upload(pos)
    with abi.Engine(s) as e:
        e.upload(pos)
        got = e.gradient(abi.GRAD_PAIR_SHORT, compat=True)
    ref = o.gradient(abi.GRAD_PAIR_SHORT, compat=True)
    # the reference's central difference (eps 1e-4) amplifies rounding by 1/eps
    assert rel_err(got, ref) < 1e-9


@pytest.mark.parametrize("terms", [abi.GRAD_PAIR_LONG, abi.GRAD_PAIR_SHORT | abi.GRAD_PAIR_LONG])
def test_gradient_compat_long_range(sysname, terms):
    """The reference's GetActionGradient with its long-range term: zero for Bare
    and David, Ilkka's CalcGradientULong added once per particle pair."""
    s, pos = make(sysname)
    o = Oracle(s)
    o.upload(pos)
    with abi.Engine(s) as e:
        e.upload(pos)
        got = e.gradient(terms, compat=True)
    ref = o.gradient(terms, compat=True)
    assert rel_err(got, ref) < 1e-9
    if terms == abi.GRAD_PAIR_LONG:
        assert (np.abs(ref).max() > 0) == (s.pair_type == abi.PAIR_ILKKA)


@pytest.mark.parametrize("terms", [abi.GRAD_KINETIC, abi.GRAD_PAIR_SHORT, abi.GRAD_PAIR_LONG, abi.GRAD_ALL])
def test_adjoint_matches_central_difference(sysname, terms):
    s, pos = make(sysname, spread=0.1)
    o = Oracle(s)
    o.upload(pos)
    with abi.Engine(s) as e:
        e.upload(pos)
        got = e.gradient(terms)
    ref = o.gradient(terms, compat=False, h=1e-5)
    assert rel_err(got, ref) < 1e-6


def test_action_laplacian(sysname):
    """Action::GetActionLaplacian (action_class.h:49): the kinetic closed form
    bit for bit; the pair action's second differences (eps 1e-4) against the
    oracle's.  A second difference divides the rounding of the link action by
    eps^2 = 1e-8, so the two sides agree to n_terms * 8 * 2.2e-16 * max|U| / eps^2,
    with max|U| the largest link action, OLD and NEW (an interface proposal)."""
    s, pos = make(sysname)
    o = Oracle(s)
    o.upload(pos)
    rng = np.random.default_rng(9)
    with abi.Engine(s) as e:
        e.upload(pos)
        p, b0 = 5, 3
        nb = 1 << s.bisect_n_level
        new = pos[:, (b0 + 1 + np.arange(nb - 1)) % s.n_bead, p].T + rng.normal(0, 0.2, size=(nb - 1, 3))
        for eng in (e, o):
            eng.set_proposal(p, b0 + 1, new)
        for mode in (abi.MODE_OLD, abi.MODE_NEW):
            for level in range(s.bisect_n_level):
                a = e.get_action_laplacian(abi.ACTION_KINETIC, mode, b0, b0 + nb, [p], level)
                b = o.get_action_laplacian(abi.ACTION_KINETIC, mode, b0, b0 + nb, [p], level)
                assert a == b, (mode, level, a, b)
            for w0, w1 in ((b0, b0 + nb), (0, s.n_bead), (s.n_bead - 2, s.n_bead + 2)):
                a = e.get_action_laplacian(abi.ACTION_PAIR, mode, w0, w1, [p], 0)
                b = o.get_action_laplacian(abi.ACTION_PAIR, mode, w0, w1, [p], 0)
                u_max = max(abs(o.get_action(abi.ACTION_PAIR, mode, w, w + 1, [p], 0)) for w in range(w0, w1))
                n_terms = 2 * 3 * 3 * (w1 - w0) * (s.n_part - 1)
                tol = n_terms * 8 * 2.2e-16 * max(u_max, 1e-3) / 1e-8
                print(f"{sysname} mode {mode} window {w0}..{w1}: {a:.12g} vs {b:.12g}, tol {tol:.3g}", flush=True)
                assert abs(a - b) <= tol, (mode, w0, w1, a, b, tol)
                assert abs(b) > 10 * tol or s.pair_type == abi.PAIR_NONE
            assert e.get_action_laplacian(abi.ACTION_PAIR, mode, b0, b0 + nb, [p], 1) == 0.0
        e.reject()
        o.reject()


def test_adjoint_long_range_on_a_ragged_k_set():
    """The long-range adjoint walks the k-set in runs of consecutive n2; a k-set
    with holes (every third vector dropped, shuffled order) splits the runs at
    the holes and at n2 = 0 in every way."""
    s = system.config_c1()
    rng = np.random.default_rng(1)
    keep = np.array([i for i in range(s.n_k) if i % 3 != 1])
    keep = keep[rng.permutation(len(keep))]
    s.k_n = s.k_n[keep]
    s.v_long_k = s.v_long_k[keep]
    pos = system.random_path(s, seed=6, spread=0.1)
    o = Oracle(s)
    o.upload(pos)
    with abi.Engine(s) as e:
        e.upload(pos)
        got = e.gradient(abi.GRAD_PAIR_LONG)
    ref = o.gradient(abi.GRAD_PAIR_LONG, compat=False, h=1e-5)
    assert rel_err(got, ref) < 1e-6


def test_observables_after_replay(sysname):
    s, pos = make(sysname, spread=0.05)
    o = Oracle(s, seed=21)
    o.upload(pos)
    o.record(True)
    o.bisect(64)
    o.displace(8)
    tape = o.recorded()
    kc = 0.8 * float(system.kmag(s.L, s.k_n).max())
    with abi.Engine(s) as e:
        e.upload(pos)
        e.tape_set(tape)
        e.bisect(64)
        e.displace(8)
        for est in (abi.ENERGY_THERMAL, abi.ENERGY_VIRIAL_COMPAT):
            for eng in (e, o):
                eng.obs_configure(energy=est, sk_k_cut=kc, gr_n_r=50)
                eng.obs_accumulate(abi.OBS_ENERGY | abi.OBS_SK | abi.OBS_GR)
            (ge, gn), (re, rn) = e.obs_read(abi.OBS_ENERGY), o.obs_read(abi.OBS_ENERGY)
            assert gn == rn == 1
            assert rel_err(ge, re) < RTOL, (est, ge, re)
            gs, _ = e.obs_read(abi.OBS_SK)
            rs, _ = o.obs_read(abi.OBS_SK)
            assert rel_err(gs, rs) < RTOL
            gg, _ = e.obs_read(abi.OBS_GR)
            rg, _ = o.obs_read(abi.OBS_GR)
            assert np.array_equal(gg, rg)


def test_virial_adjoint_estimator(sysname):
    s, pos = make(sysname, spread=0.05)
    o = Oracle(s)
    o.upload(pos)
    with abi.Engine(s) as e:
        e.upload(pos)
        for eng in (e, o):
            eng.obs_configure(energy=abi.ENERGY_VIRIAL_ADJOINT)
            eng.obs_accumulate(abi.OBS_ENERGY)
        ge, _ = e.obs_read(abi.OBS_ENERGY)
    re, _ = o.obs_read(abi.OBS_ENERGY)
    assert rel_err(ge, re) < 1e-6


@pytest.mark.parametrize("streams,fork,one", [("1", "0", "0"), ("3", "0", "0"), ("2", "1", "0"), ("1", "0", "1"),
                                              ("3", "0", "1")])
def test_batched_replicas_match_single_chains(streams, fork, one, monkeypatch):
    """Replicas advanced together (on one stream, split into groups on three
    streams, or two groups whose rho_k kernels run on a second stream beside the
    pair kernels; rounds of three launches or of one) replay each single chain's
    tape bit for bit."""
    monkeypatch.setenv("SPIMC_STREAMS", streams)
    monkeypatch.setenv("SPIMC_FORK", fork)
    monkeypatch.setenv("SPIMC_ONE_LAUNCH", one)
    monkeypatch.setenv("SPIMC_RHO_BPC", "3" if fork == "1" else "0")
    systems = [system.config_c1(n_part=n, n_bead=m, r_s=rs) for n, m, rs in [(16, 16, 1.5), (32, 32, 2.0), (8, 16, 3.0), (24, 8, 1.2)]]
    tapes, refs, finals = [], [], []
    for i, s in enumerate(systems):
        pos = system.random_path(s, seed=40 + i, spread=0.05)
        o = Oracle(s, seed=100 + i)
        o.upload(pos)
        o.record(True)
        refs.append((o.bisect(50), o.displace(10)))
        tapes.append((pos, o.recorded()))
        finals.append(o.download())
    with abi.Engine(systems, seed=1) as e:
        for i, (pos, tape) in enumerate(tapes):
            e.upload(pos, r=i)
            e.tape_set(tape, r=i)
        ab = e.bisect(50)
        ad = e.displace(10)
        for i in range(len(systems)):
            assert np.array_equal(ab[i], refs[i][0])
            assert np.array_equal(ad[i], refs[i][1])
            assert np.array_equal(e.download(r=i), finals[i])


def mixed_systems():
    return [system.config_c1(n_part=16, n_bead=16), small_david(), system.config_ilkka(n_part=24, n_bead=8, n_level=2)]


def test_mixed_pair_actions_in_one_context():
    """Bare, David and Ilkka replicas of different sizes batched in one context
    (the global-table kernels, pair type read per replica) replay each single
    chain's tape bit for bit, and their estimators match the oracle."""
    systems = mixed_systems()
    tapes, refs, finals, energies = [], [], [], []
    for i, s in enumerate(systems):
        pos = system.random_path(s, seed=60 + i, spread=0.05)
        o = Oracle(s, seed=200 + i)
        o.upload(pos)
        o.record(True)
        refs.append((o.bisect(40), o.displace(8)))
        tapes.append((pos, o.recorded()))
        finals.append(o.download())
        energies.append((o.d_action_d_beta(abi.ACTION_PAIR), o.potential(abi.ACTION_PAIR)))
    with abi.Engine(systems, seed=1) as e:
        for i, (pos, tape) in enumerate(tapes):
            e.upload(pos, r=i)
            e.tape_set(tape, r=i)
        ab = e.bisect(40)
        ad = e.displace(8)
        for i in range(len(systems)):
            assert np.array_equal(ab[i], refs[i][0]), i
            assert np.array_equal(ad[i], refs[i][1]), i
            assert np.array_equal(e.download(r=i), finals[i]), i
            a, b = e.d_action_d_beta(abi.ACTION_PAIR, r=i), energies[i][0]
            assert abs(a - b) <= RTOL * max(1.0, abs(b)), (i, a, b)
            a, b = e.potential(abi.ACTION_PAIR, r=i), energies[i][1]
            assert abs(a - b) <= RTOL * max(1.0, abs(b)), (i, a, b)


def test_mixed_pair_actions_graph_sweep():
    """The same mixed context swept eagerly and from captured graphs (device RNG)
    ends in the same state, with the sweep's estimators and adjoint."""
    systems = mixed_systems()
    flags = abi.SWEEP_ADJOINT | abi.SWEEP_ENERGY | abi.SWEEP_SK | abi.SWEEP_GR
    with abi.Engine(systems, seed=5) as e, abi.Engine(systems, seed=5) as g:
        for i, s in enumerate(systems):
            pos = system.random_path(s, seed=70 + i, spread=0.05)
            e.upload(pos, r=i)
            g.upload(pos, r=i)
        for eng in (e, g):
            eng.obs_configure(energy=abi.ENERGY_VIRIAL_ADJOINT, gr_n_r=50)
        e.sweep(2, flags)
        g.sweep(2, flags | abi.SWEEP_GRAPH)
        for i in range(len(systems)):
            assert np.array_equal(e.download(r=i), g.download(r=i)), i
            assert np.array_equal(e.counters(r=i), g.counters(r=i)), i
            assert np.array_equal(e.obs_read(abi.OBS_ENERGY, r=i)[0], g.obs_read(abi.OBS_ENERGY, r=i)[0]), i


@pytest.mark.parametrize("streams,fork,one", [("1", "0", "0"), ("2", "0", "0"), ("2", "1", "0"), ("2", "0", "1")])
def test_uneven_counts_and_graph_sweep(streams, fork, one, monkeypatch):
    systems = system.config_c4(seed=3, n_replicas=6)
    with abi.Engine(systems, seed=5) as e:
        monkeypatch.setenv("SPIMC_STREAMS", streams)
        monkeypatch.setenv("SPIMC_FORK", fork)
        monkeypatch.setenv("SPIMC_ONE_LAUNCH", one)
        g = abi.Engine(systems, seed=5)
        for i, s in enumerate(systems):
            pos = system.random_path(s, seed=i, spread=0.05)
            e.upload(pos, r=i)
            g.upload(pos, r=i)
        e.sweep(2, 0)
        g.sweep(2, abi.SWEEP_GRAPH)
        for i, s in enumerate(systems):
            assert np.array_equal(e.download(r=i), g.download(r=i))
            c = e.counters(r=i)
            assert c[0] == 2 * ((s.n_part * s.n_bead) >> s.bisect_n_level)
            assert c[2] == 2 * s.n_part
            assert np.array_equal(c, g.counters(r=i))
        g.close()


def test_c2_full_size_properties():
    """At the full HEG N=256, M=128 size: the incrementally maintained rho_k after a
    move stage equals a rebuild from the final positions, and the short-range
    pair gradient is translation invariant (sum over particles vanishes)."""
    s = system.config_c2()
    pos = system.random_path(s, seed=2, spread=0.05)
    with abi.Engine(s, seed=3) as e:
        e.upload(pos)
        e.bisect(256)
        e.displace(16)
        inc = e.rho_k()
        e.rho_k_rebuild()
        full = e.rho_k()
        assert rel_err(inc, full) < 1e-11 * s.n_part
        g = e.gradient(abi.GRAD_PAIR_SHORT)
        tot = g.sum(axis=1)
        assert np.max(np.abs(tot)) < 1e-9 * max(1.0, np.max(np.abs(g)))
        c = e.counters()
        assert c[0] == 256 and c[2] == 16 and 0 < c[1] < 256


def test_large_replica_spread_over_workgroups():
    """N=1024, M=256, 4012 k-vectors (the large HEG case): a single replica's
    pair links and adjoint slice ranges are spread over many workgroups.  The
    replayed chain matches the oracle bit for bit; the incremental rho_k equals
    a rebuild; short- and long-range pair gradients sum to zero over the
    particles of every slice (translation invariance)."""
    import os
    import oracle_ffi

    s = system.config_c5()
    assert s.n_k == 4012
    pos = system.random_path(s, seed=31, spread=0.05)
    oracle_ffi.set_threads(min(16, len(os.sched_getaffinity(0))))
    try:
        o = Oracle(s, seed=5)
        o.upload(pos)
        o.record(True)
        ref = np.concatenate([o.bisect(24), o.displace(3)])
        tape = o.recorded()
        final = o.download()
        ref_rho = o.rho_k()
    finally:
        oracle_ffi.set_threads(1)
    with abi.Engine(s) as e:
        e.upload(pos)
        e.tape_set(tape)
        got = np.concatenate([e.bisect(24)[0], e.displace(3)[0]])
        assert np.array_equal(got, ref)
        assert np.array_equal(e.download(), final)
        inc = e.rho_k()
        assert rel_err(inc, ref_rho) < 1e-11 * s.n_part
        e.rho_k_rebuild()
        assert rel_err(inc, e.rho_k()) < 1e-11 * s.n_part
        for terms in (abi.GRAD_PAIR_SHORT, abi.GRAD_PAIR_LONG):
            g = e.gradient(terms)  # [M][N][3]
            tot = g.sum(axis=1)
            assert np.max(np.abs(tot)) < 1e-9 * max(1.0, float(np.max(np.abs(g)))), terms


def test_global_table_kernels_match_lds_kernels(monkeypatch):
    """The kernels that read the spline tables from global memory (taken for
    non-uniform grids) give the oracle's results too: replay, adjoint, energy."""
    monkeypatch.setenv("SPIMC_PAIR_LDS", "0")
    s = small_david()
    pos = system.random_path(s, seed=13, spread=0.05)
    o = Oracle(s, seed=17)
    o.upload(pos)
    o.record(True)
    ref = np.concatenate([o.bisect(64), o.displace(8)])
    tape = o.recorded()
    with abi.Engine(s) as e:
        e.upload(pos)
        e.tape_set(tape)
        got = np.concatenate([e.bisect(64)[0], e.displace(8)[0]])
        assert np.array_equal(got, ref)
        assert np.array_equal(e.download(), o.download())
        g = e.gradient(abi.GRAD_ALL)
        assert rel_err(g, o.gradient(abi.GRAD_ALL, compat=False, h=1e-5)) < 1e-6
        for eng in (e, o):
            eng.obs_configure(energy=abi.ENERGY_THERMAL)
            eng.obs_accumulate(abi.OBS_ENERGY)
        assert rel_err(e.obs_read(abi.OBS_ENERGY)[0], o.obs_read(abi.OBS_ENERGY)[0]) < RTOL


def test_autotune_keeps_the_chain():
    """The autotuner's stages are ordinary move stages under different launch
    shapes: afterwards the chains equal those of an engine that ran the same
    sequence of bisection and displace stages with its default shape, bit for
    bit."""
    s = system.config_c1()
    n_rep, rounds = 64, 4
    a = abi.Engine([s] * n_rep, seed=31)
    b = abi.Engine([s] * n_rep, seed=31)
    try:
        for e in (a, b):
            for r in range(n_rep):
                e.upload(system.random_path(s, seed=700 + r, spread=0.2), r=r)
        chosen = a.autotune(rounds)
        assert chosen["shapes_tried"] >= 2 and chosen["pair_workgroup"] in (256, 512, 768, 1024)
        for _ in range(chosen["shapes_tried"]):
            b.bisect(1)
            b.displace(1)
            b.bisect(rounds)
            b.displace(chosen["displace_rounds"])
        for e in (a, b):
            e.sweep(1, abi.SWEEP_GRAPH)
        for r in (0, 17, n_rep - 1):
            assert np.array_equal(a.download(r=r), b.download(r=r))
            assert np.array_equal(a.counters(r=r), b.counters(r=r))
    finally:
        a.close()
        b.close()


def test_adjoint_on_many_configurations():
    """The analytic dS/dr of every bead (kinetic + short- and long-range pair)
    against the oracle's central difference (h = 1e-5) on 20 seeded paths of the
    N=32, M=32 Bare/Ewald system."""
    s = system.config_c1()
    worst = 0.0
    with abi.Engine(s) as e:
        for seed in range(20):
            pos = system.random_path(s, seed=1000 + seed, spread=0.05 + 0.01 * seed)
            o = Oracle(s)
            o.upload(pos)
            e.upload(pos)
            worst = max(worst, rel_err(e.gradient(abi.GRAD_ALL), o.gradient(abi.GRAD_ALL, compat=False, h=1e-5)))
    assert worst < 1e-6, worst


def test_adjoint_on_1000_configurations():
    """The differentiation configuration at its stated size: 1000 seeded paths
    of the N=32, M=32 Bare/Ewald system (batched as 1000 replicas of one
    context), the analytic dS/dr of every bead computed on each, and 8 random
    beads per path checked against the oracle's central difference (h = 1e-5)
    of the kinetic, short-range and Ewald action, to 1e-6 relative."""
    s = system.config_c1()
    n_cfg, n_beads = 1000, 8
    rng = np.random.default_rng(12)
    paths = [system.random_path(s, seed=20000 + i, spread=0.03 + 0.1 * rng.random()) for i in range(n_cfg)]
    beads = [np.stack([rng.integers(0, s.n_bead, n_beads), rng.integers(0, s.n_part, n_beads)], axis=1)
             for _ in range(n_cfg)]
    o = Oracle(s)
    ref = []
    for pos, b in zip(paths, beads):
        o.upload(pos)
        ref.append(o.gradient_beads(b, abi.GRAD_ALL, 1e-5))
    worst = 0.0
    with abi.Engine([s] * n_cfg, seed=1) as e:
        for i, pos in enumerate(paths):
            e.upload(pos, r=i)
        for i, b in enumerate(beads):
            g = e.gradient(abi.GRAD_ALL, r=i)
            worst = max(worst, rel_err(g[b[:, 0], b[:, 1]], ref[i]))
    print(f"worst relative error over {n_cfg} configurations: {worst:.3g}", flush=True)
    assert worst < 1e-6, worst


def test_adjoint_at_full_size_on_sampled_beads():
    """At N=256, M=128 (David pair action, 1979 k-vectors) the analytic dS/dr of
    96 randomly chosen beads against the oracle's central difference of those
    beads (h = 1e-5), after a few moves."""
    s = system.config_c2()
    pos = system.random_path(s, seed=41, spread=0.05)
    rng = np.random.default_rng(7)
    beads = np.stack([rng.integers(0, s.n_bead, 96), rng.integers(0, s.n_part, 96)], axis=1)
    o = Oracle(s, seed=3)
    o.upload(pos)
    o.record(True)
    o.bisect(64)
    o.displace(4)
    tape = o.recorded()
    ref = o.gradient_beads(beads, abi.GRAD_ALL, 1e-5)
    with abi.Engine(s) as e:
        e.upload(pos)
        e.tape_set(tape)
        e.bisect(64)
        e.displace(4)
        assert np.array_equal(e.download(), o.download())
        g = e.gradient(abi.GRAD_ALL)
    got = g[beads[:, 0], beads[:, 1]]
    assert rel_err(got, ref) < 1e-6


def _edge(name):
    """Small and degenerate inputs the reference accepts."""
    if name == "one_particle":
        return system.config_c1(n_part=1, n_bead=8, n_level=2)
    if name == "two_slices":
        return system.config_c1(n_part=4, n_bead=2, n_level=1)
    if name == "window_is_path":  # 2^n_level = M: the bead walks visit nothing
        return system.config_c1(n_part=6, n_bead=8, n_level=3)
    if name == "no_long_range":
        s = system.config_c1(n_part=8, n_bead=16, n_level=2)
        s.use_long_range = False
        return s
    if name == "images":
        s = system.config_c1(n_part=6, n_bead=8, n_level=2)
        s.kinetic_n_images = 1
        s.bisect_n_images = 1
        return s
    if name == "three_k_vectors":
        s = system.config_c1(n_part=8, n_bead=8, n_level=2)
        s.k_n = system.kspace(s.L, (1.0 + 1e-9) * 2 * np.pi / s.L)
        return system.bare_coulomb_ewald(s, alpha=5.0 / s.L)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["one_particle", "two_slices", "window_is_path", "no_long_range", "images",
                                  "three_k_vectors"])
@pytest.mark.parametrize("one", ["0", "1", "2"])
def test_edge_configurations(name, one, monkeypatch):
    """Replay, estimators and gradients on edge inputs: one particle (no pairs),
    two slices, a bisection window as long as the path, no long range, image
    sums in the kinetic action and the sampling, and a three-vector k-set;
    rounds of three launches or of one (SPIMC_ONE_LAUNCH)."""
    monkeypatch.setenv("SPIMC_ONE_LAUNCH", one)
    s = _edge(name)
    if name == "three_k_vectors":
        assert s.n_k == 3
    pos = system.random_path(s, seed=17, spread=0.1)
    o = Oracle(s, seed=23)
    o.upload(pos)
    o.record(True)
    n_b = max(1, (s.n_part * s.n_bead) >> s.bisect_n_level)
    ref = [o.bisect(n_b), o.displace(s.n_part), o.bisect(n_b)]
    tape = o.recorded()
    with abi.Engine(s, seed=3) as e:
        e.upload(pos)
        e.tape_set(tape)
        got = [e.bisect(n_b)[0], e.displace(s.n_part)[0], e.bisect(n_b)[0]]
        for g, r in zip(got, ref):
            assert np.array_equal(g, r), name
        assert np.array_equal(e.download(), o.download())
        assert np.array_equal(e.counters(), o.counters())
        e.tape_set(None)
        for action in (abi.ACTION_KINETIC, abi.ACTION_PAIR):
            a, b = e.d_action_d_beta(action), o.d_action_d_beta(action)
            assert abs(a - b) <= RTOL * max(1.0, abs(b)), (name, action, a, b)
        a, b = e.potential(abi.ACTION_PAIR), o.potential(abi.ACTION_PAIR)
        assert abs(a - b) <= RTOL * max(1.0, abs(b)), (name, a, b)
        a, b = e.kinetic_virial_energy(0), o.kinetic_virial_energy(0)
        assert abs(a - b) <= RTOL * max(1.0, abs(b)), (name, a, b)
        assert rel_err(e.gradient(abi.GRAD_PAIR_SHORT, compat=True), o.gradient(abi.GRAD_PAIR_SHORT, compat=True)) < 1e-9
        # with one particle the pair action is constant, so the reference never
        # updates rho_k during moves (pair_action_class.h:267-272); the central
        # difference is taken on rho_k of the current positions
        o.rho_k_rebuild()
        assert rel_err(e.gradient(abi.GRAD_ALL), o.gradient(abi.GRAD_ALL, h=1e-5)) < 1e-6


@pytest.mark.parametrize("images,one", [(0, "0"), (1, "0"), (0, "1")])
def test_bisect_level_changes(images, one, monkeypatch):
    """Adaptive bisection (bisect_class.h:143-151): stages at levels 3, 2, 4, 1
    and 3 of two replicas, each replaying its oracle's tape, with the sampling
    image splines of every level; the sweep's attempt count follows the level.
    With one-launch rounds (one = "1") every level change rebuilds their
    workgroup tables."""
    monkeypatch.setenv("SPIMC_ONE_LAUNCH", one)
    systems = [system.config_c1(n_part=16, n_bead=32), system.config_c1(n_part=8, n_bead=16)]
    for s in systems:
        s.bisect_n_images = images
    levels = [3, 2, 4, 1, 3]
    tapes, refs, finals = [], [], []
    for i, s in enumerate(systems):
        pos = system.random_path(s, seed=80 + i, spread=0.05)
        o = Oracle(s, seed=300 + i)
        o.upload(pos)
        o.record(True)
        acc = []
        for lv in levels:
            o.set_bisect_level(lv)
            acc.append(o.bisect(12))
            acc.append(o.displace(4))
        refs.append(acc)
        tapes.append((pos, o.recorded()))
        finals.append(o.download())
    with abi.Engine(systems, seed=1) as e:
        for i, (pos, tape) in enumerate(tapes):
            e.upload(pos, r=i)
            e.tape_set(tape, r=i)
        got = [[] for _ in systems]
        for lv in levels:
            for i in range(len(systems)):
                e.set_bisect_level(lv, r=i)
            ab, ad = e.bisect(12), e.displace(4)
            for i in range(len(systems)):
                got[i] += [ab[i], ad[i]]
        for i in range(len(systems)):
            for g, r in zip(got[i], refs[i]):
                assert np.array_equal(g, r), i
            assert np.array_equal(e.download(r=i), finals[i]), i
        with pytest.raises(abi.SpimcError):
            e.set_bisect_level(5, r=1)  # beyond floor(log2 16)
    with abi.Engine(systems[0], seed=4) as e:
        e.upload(system.random_path(systems[0], seed=1, spread=0.05))
        e.set_bisect_level(2)
        c0 = e.counters()
        e.sweep(1, abi.SWEEP_GRAPH)
        e.set_bisect_level(4)
        e.sweep(1, abi.SWEEP_GRAPH)
        c = e.counters() - c0
        s = systems[0]
        assert c[0] == ((s.n_part * s.n_bead) >> 2) + ((s.n_part * s.n_bead) >> 4)


def _random_system(rng):
    """A small system drawn at random: size, level, pair action (none, Bare
    Coulomb, tabulated Bare, David, Ilkka), long range, images and k cut."""
    n = int(rng.integers(1, 40))
    m = int(rng.choice([2, 4, 8, 16, 32], p=[0.1, 0.15, 0.25, 0.25, 0.25]))
    lv = int(rng.integers(1, min(int(np.log2(m)), 4) + 1))
    kind = rng.choice(["none", "coulomb", "tabulated", "david", "ilkka"])
    rs = float(rng.uniform(1.0, 3.0))
    if kind == "david":
        s = system.config_c2(n_part=n, n_bead=m, r_s=rs, tau=0.02, n_level=lv, k_shell=float(rng.uniform(3.0, 12.0)),
                             n_grid=300, seed=int(rng.integers(1, 100)))
    elif kind == "ilkka":
        s = system.config_ilkka(n_part=n, n_bead=m, r_s=rs, n_level=lv)
    else:
        s = system.config_c1(n_part=n, n_bead=m, r_s=rs, n_level=lv)
        s.k_n = system.kspace(s.L, float(rng.uniform(1.0, 4.0)) * 2 * np.pi / s.L + 1e-9)
        system.bare_coulomb_ewald(s, alpha=5.0 / s.L)
        if kind == "tabulated":
            s.bare_is_coulomb = False
            s.bare_v_r = 1.0 / np.sqrt(s.bare_r ** 2 + 0.1)
        elif kind == "none":
            s.pair_type = abi.PAIR_NONE
    if kind != "none" and rng.random() < 0.25:
        s.use_long_range = False
    if rng.random() < 0.3:
        s.kinetic_n_images = 1
    if rng.random() < 0.3:
        s.bisect_n_images = 1
    s.name = f"{kind}-N{n}-M{m}-l{lv}-lr{int(s.use_long_range)}-img{s.kinetic_n_images}{s.bisect_n_images}"
    return s


@pytest.mark.parametrize("seed", range(16))
def test_randomized_configurations(seed):
    """Seeded random small systems: a replayed bisect/displace chain bit for bit,
    rho_k, the whole-path estimators, the reference gradient and the adjoint."""
    rng = np.random.default_rng(1000 + seed)
    s = _random_system(rng)
    print(s.name, flush=True)
    pos = system.random_path(s, seed=seed, spread=float(rng.uniform(0.02, 0.3)))
    o = Oracle(s, seed=seed)
    o.upload(pos)
    o.record(True)
    n_b = max(1, (s.n_part * s.n_bead) >> s.bisect_n_level)
    ref = [o.bisect(n_b), o.displace(s.n_part), o.bisect(n_b)]
    tape = o.recorded()
    with abi.Engine(s, seed=seed) as e:
        e.upload(pos)
        e.tape_set(tape)
        got = [e.bisect(n_b)[0], e.displace(s.n_part)[0], e.bisect(n_b)[0]]
        for g, r in zip(got, ref):
            assert np.array_equal(g, r), s.name
        assert np.array_equal(e.download(), o.download()), s.name
        e.tape_set(None)
        o.rho_k_rebuild()  # one particle: the reference leaves rho_k stale (constant pair action)
        e.rho_k_rebuild()
        if s.use_long_range and s.pair_type != abi.PAIR_NONE:
            assert rel_err(e.rho_k(), o.rho_k()) < 1e-12 * max(1, s.n_part), s.name
        for action in (abi.ACTION_KINETIC, abi.ACTION_PAIR):
            a, b = e.d_action_d_beta(action), o.d_action_d_beta(action)
            assert abs(a - b) <= RTOL * max(1.0, abs(b)), (s.name, action, a, b)
        a, b = e.potential(abi.ACTION_PAIR), o.potential(abi.ACTION_PAIR)
        assert abs(a - b) <= RTOL * max(1.0, abs(b)), (s.name, a, b)
        a, b = e.kinetic_virial_energy(0), o.kinetic_virial_energy(0)
        assert abs(a - b) <= RTOL * max(1.0, abs(b)), (s.name, a, b)
        terms = abi.GRAD_PAIR_SHORT | abi.GRAD_PAIR_LONG
        assert rel_err(e.gradient(terms, compat=True), o.gradient(terms, compat=True)) < 1e-9, s.name
        # The link action jumps where a pair vector crosses half a box: DrDrpDrrp
        # takes r' at the image nearest r, so when r changes sides |r'| changes
        # (path_class.h:137-150).  A central difference that straddles such a
        # point is not a derivative; beads whose differences at h and h/5 disagree
        # are left out (there are at most a few).
        g, f5, f1 = e.gradient(abi.GRAD_ALL), o.gradient(abi.GRAD_ALL, h=1e-5), o.gradient(abi.GRAD_ALL, h=2e-6)
        scale = max(1.0, float(np.abs(f5).max()))
        smooth = np.all(np.abs(f5 - f1) < 1e-6 * scale, axis=2)
        assert smooth.sum() >= smooth.size - 3, (s.name, int(smooth.size - smooth.sum()))
        assert float(np.abs(g - f5)[smooth].max()) < 1e-6 * scale, s.name
