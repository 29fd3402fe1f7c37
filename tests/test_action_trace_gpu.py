"""GPU parity of the action differences that decide Metropolis on the hot path.

Every move attempt's level-0 decision adds up the kinetic action, the
short-range pair action (the LDS pair kernel, or the global-table and
label-general kernels) and the Ewald k-space action (the incremental rho_k
kernel's fused sums), each in OLD and NEW mode (bisect_class.h:100-114,
displace_particle_class.h:52-68, perm_bisect_iterative_class.h:194-213).  The
engine's action trace records exactly the terms its decide kernel sums; the
oracle records the values its GetAction calls return for the same proposal.
The oracle's draws are replayed, so both see the same proposals.

Tolerance: every term, and every OLD - NEW difference, within 1e-10 relative
to max(1, |oracle value|).  The reference's long-range terms also sum the
window's slices where no bead moves (the first slice of a bisection window);
they are equal in OLD and NEW, and the engine's rho_k kernel leaves them out
of both, so the oracle reports that part separately and the comparison takes
it off.  Attempts rejected at a coarser level (or whose
cycle walk stopped) carry no level-0 terms on either side (NaN rows), and must
coincide.
"""
import numpy as np
import pytest

from oracle_ffi import Oracle
from simpimc_amd import abi, system

pytestmark = pytest.mark.gpu

RTOL = 1e-10
TERMS = ["kinetic old", "kinetic new", "pair short old", "pair short new", "pair long old", "pair long new",
         "log sample ratio", "previous change"]


def small_david():
    return system.config_c2(n_part=16, n_bead=16, r_s=1.0, tau=0.02, n_level=3, k_shell=20.5, n_grid=400, seed=7)


def dense_bare():
    return system.config_c1(n_part=8, n_bead=16, r_s=1.0, tau=0.05, n_level=3)


def c5_replica():
    return system.config_c5()


# name: (system, bisect attempts per stage, displace attempts per stage, stages)
CASES = {
    "bare_c1": (lambda: system.config_c1(), 128, 32, 2),
    "david_small": (small_david, 64, 16, 2),
    "ilkka_c1": (lambda: system.config_ilkka(), 128, 32, 2),
    "c2_full": (lambda: system.config_c2(), 192, 24, 1),
    "c5_replica": (c5_replica, 48, 6, 1),
}


def compare(got, ref_full, label, absolute=True):
    """ref_full carries a 9th column: the long-range action of the window's
    slices where no bead moves.  It enters OLD and NEW alike, so the engine's
    rho_k kernel, which visits moved slices only, leaves it out of both.
    absolute=False compares only the OLD -> NEW differences (and the sample
    ratio and previous change): on a permuted path the reference's particle
    list also holds labels with no moved bead, whose kinetic links and pairs
    enter OLD and NEW alike and which the engine leaves out."""
    assert got.shape == (ref_full.shape[0], 8), (label, got.shape, ref_full.shape)
    nan_g, nan_r = np.isnan(got[:, 0]), np.isnan(ref_full[:, 0])
    assert np.array_equal(nan_g, nan_r), (label, "level-0 attempts differ", np.flatnonzero(nan_g != nan_r)[:8])
    ok = ~nan_r
    g, r = got[ok], ref_full[ok, :8].copy()
    fixed = ref_full[ok, 8]
    scale = np.maximum(1.0, np.abs(r))
    r[:, 4] -= fixed
    r[:, 5] -= fixed
    assert np.all(np.isfinite(g)), label
    cols = slice(None) if absolute else slice(6, 8)
    err = np.abs(g[:, cols] - r[:, cols]) / scale[:, cols]
    worst = np.unravel_index(np.argmax(err), err.shape) if err.size else None
    assert err.size == 0 or err.max() <= RTOL, (label, TERMS[worst[1] + (0 if absolute else 6)], int(worst[0]),
                                                err.max())
    # the differences that enter log A: kinetic, short-range and long-range, NEW - OLD
    for a, b in ((0, 1), (2, 3), (4, 5)):
        dg, dr = g[:, b] - g[:, a], r[:, b] - r[:, a]
        e = np.abs(dg - dr) / np.maximum(1.0, np.abs(r[:, a]))
        assert e.size == 0 or e.max() <= RTOL, (label, TERMS[a], e.max())
    return int(ok.sum())


RUNS = ([(name, 0, 0) for name in sorted(CASES)] + [("bare_c1", 256, 0), ("david_small", 256, 0), ("ilkka_c1", 512, 0)]
        + [("bare_c1", 0, 1), ("david_small", 0, 1), ("c5_replica", 0, 1), ("david_small", 0, 2), ("c5_replica", 0, 2)])


# persistent: 0 stages as rounds of launches, else the persistent stage kernel's
# workgroup size; one: rounds of one launch each (1 bisection, 2 both moves)
@pytest.mark.parametrize("name,persistent,one", RUNS)
def test_hot_path_action_terms_match_oracle(name, persistent, one, monkeypatch):
    monkeypatch.setenv("SPIMC_PERSISTENT", str(persistent))
    monkeypatch.setenv("SPIMC_ONE_LAUNCH", str(one))
    make, n_b, n_d, stages = CASES[name]
    s = make()
    pos = system.random_path(s, seed=3, spread=0.05)
    print(f"{name}: N={s.n_part} M={s.n_bead} nk={s.n_k}", flush=True)
    o = Oracle(s, seed=11)
    o.upload(pos)
    o.record(True)
    ref = []
    for _ in range(stages):
        for kind, n in (("bisect", n_b), ("displace", n_d)):
            o.trace(True)
            getattr(o, kind)(n)
            ref.append((kind, o.trace_read()))
    tape = o.recorded()
    with abi.Engine(s, seed=99) as e:
        e.upload(pos)
        e.tape_set(tape)
        e.trace(True)
        n_level0 = 0
        for i, (kind, r) in enumerate(ref):
            getattr(e, kind)(len(r))
            n_level0 += compare(e.trace_read(len(r)), r, f"{name} stage {i} {kind}")
        assert e.tape_cursor() == len(tape)
        assert np.array_equal(e.download(), o.download())
    print(f"{name}: {n_level0} level-0 decisions compared", flush=True)
    assert n_level0 > 0


PERM_CASES = {"bare_dense": dense_bare, "david_small": small_david, "ilkka_c1": lambda: system.config_ilkka()}


@pytest.mark.parametrize("name", sorted(PERM_CASES))
def test_permuting_action_terms_match_oracle(name):
    """PermBisectIterative's level-0 action changes, kinetic sums over several
    chains included, and those of Bisect / DisplaceParticle on the permuted
    paths that follow."""
    s = PERM_CASES[name]()
    pos = system.random_path(s, seed=3, spread=0.3)
    n_b = max(1, s.n_part * s.n_bead >> s.bisect_n_level)
    o = Oracle(s, seed=11)
    o.upload(pos)
    o.record(True)
    ref = []
    for _ in range(3):
        for kind, n in (("perm_bisect", n_b), ("bisect", n_b), ("displace", s.n_part)):
            o.trace(True)
            getattr(o, kind)(n)
            ref.append((kind, o.trace_read()))
    tape = o.recorded()
    with abi.Engine(s, seed=99) as e:
        e.upload(pos)
        e.tape_set(tape)
        e.trace(True)
        for i, (kind, r) in enumerate(ref):
            getattr(e, kind)(len(r))
            compare(e.trace_read(len(r)), r, f"{name} stage {i} {kind}", absolute=False)
        assert np.array_equal(e.perm(), o.perm())
        assert np.array_equal(e.download(), o.download())


def test_trace_is_off_by_default_and_graph_sweeps_fill_it():
    """The trace must be enabled before it can be read, and the sweep's
    graph-replayed rounds record it as the eager rounds do."""
    s = system.config_c1(n_part=16, n_bead=16)
    pos = system.random_path(s, seed=4, spread=0.1)
    with abi.Engine(s, seed=5) as e:
        e.upload(pos)
        e.bisect(8)
        with pytest.raises(abi.SpimcError):
            e.trace_read(8)
    rows = {}
    for graph in (False, True):
        with abi.Engine(s, seed=5) as e:
            e.upload(pos)
            e.trace(True)
            e.sweep(1, abi.SWEEP_GRAPH if graph else 0)
            rows[graph] = e.trace_read(s.n_part)  # the sweep's last stage: displace
    assert np.array_equal(rows[False], rows[True], equal_nan=True)
    assert np.isfinite(rows[False][:, 2]).all()
