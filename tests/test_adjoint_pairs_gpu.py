"""The pair-once adjoint (each short-range pair-link evaluated once for both of
its particles, k_adjoint_pairs_lds) against the per-particle adjoint
(k_adjoint_local_lds, every link evaluated from both sides) and against the
oracle's central difference of the action (h = 1e-5, 1e-6 relative).

The two kernels evaluate the same analytic dS/dr and differ only in the order
of the sums, so they agree to 1e-10 relative.  The pair-once kernel forms every
sum in a fixed order, so two calls give the same bits.  SPIMC_ADJ_PAIRS = 1 / 0
forces the kernel (read when a context is created); the cases cover particle
counts that are not multiples of the 64-particle tile, odd and even tiles, one
and two particles, two and three slices, several workgroups per replica,
replicas of different sizes in one context, and permuted paths.
"""
import numpy as np
import pytest

from simpimc_amd import abi, system
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    return float(np.max(np.abs(a - b))) / scale if a.size else 0.0


def _small_david(n_part=16, n_bead=16, seed=7):
    return system.config_c2(n_part=n_part, n_bead=n_bead, r_s=1.0, tau=0.02, n_level=3, k_shell=20.5, n_grid=400,
                            seed=seed)


CASES = {
    "bare_32x32": lambda: system.config_c1(),
    "bare_65x8": lambda: system.config_c1(n_part=65, n_bead=8, n_level=2),
    "bare_100x5": lambda: system.config_c1(n_part=100, n_bead=5, n_level=2),
    "bare_130x4": lambda: system.config_c1(n_part=130, n_bead=4, n_level=2),
    "bare_2x8": lambda: system.config_c1(n_part=2, n_bead=8, n_level=2),
    "bare_1x8": lambda: system.config_c1(n_part=1, n_bead=8, n_level=2),
    "bare_7x2": lambda: system.config_c1(n_part=7, n_bead=2, n_level=1),
    "bare_9x3": lambda: system.config_c1(n_part=9, n_bead=3, n_level=1),
    "bare_33x128": lambda: system.config_c1(n_part=33, n_bead=128, n_level=3),
    "david_16x16": _small_david,
    "david_64x16": lambda: _small_david(n_part=64),
    "ilkka_32x32": lambda: system.config_ilkka(),
}


def _gradients(monkeypatch, systems, paths, pairs, terms=abi.GRAD_ALL, perms=None, calls=1):
    monkeypatch.setenv("SPIMC_ADJ_PAIRS", "1" if pairs else "0")
    systems = systems if isinstance(systems, list) else [systems]
    out = []
    with abi.Engine(systems) as e:
        for i, pos in enumerate(paths):
            e.upload(pos, r=i)
            if perms is not None:
                e.set_perm(perms[i], r=i)
        for _ in range(calls):
            out.append([e.gradient(terms, r=i) for i in range(len(systems))])
    return out


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("terms", [abi.GRAD_PAIR_SHORT, abi.GRAD_ALL])
def test_pair_once_matches_per_particle_adjoint(monkeypatch, name, terms):
    s = CASES[name]()
    pos = system.random_path(s, seed=17, spread=0.1)
    print(f"case {name} terms {terms}", flush=True)
    new = _gradients(monkeypatch, s, [pos], True, terms, calls=2)
    old = _gradients(monkeypatch, s, [pos], False, terms)
    assert np.array_equal(new[0][0], new[1][0]), "pair-once adjoint is not reproducible"
    assert rel_err(new[0][0], old[0][0]) < 1e-10
    if s.n_part == 1 and terms == abi.GRAD_PAIR_SHORT:
        assert np.abs(new[0][0]).max() == 0.0


@pytest.mark.parametrize("name", ["bare_65x8", "bare_130x4", "bare_9x3", "david_16x16", "ilkka_32x32"])
def test_pair_once_matches_central_difference(monkeypatch, name):
    # (a central difference across a kink of the clamped David table, as on some
    # denser David paths, is no check of the analytic derivative: those paths are
    # covered against the per-particle kernel above)
    s = CASES[name]()
    pos = system.random_path(s, seed=3, spread=0.1)
    print(f"case {name}", flush=True)
    o = Oracle(s)
    o.upload(pos)
    ref = o.gradient(abi.GRAD_ALL, compat=False, h=1e-5)
    got = _gradients(monkeypatch, s, [pos], True)[0][0]
    assert rel_err(got, ref) < 1e-6


def test_pair_once_on_a_permuted_path(monkeypatch):
    """The kinetic links follow the permutation across the wrap; the pair links
    join the same particle slots at slices b and b+1."""
    s = CASES["bare_65x8"]()
    pos = system.random_path(s, seed=5, spread=0.3)
    sigma = np.roll(np.arange(s.n_part), 1).astype(np.uint32)
    o = Oracle(s)
    o.upload(pos)
    o.set_perm(sigma)
    ref = o.gradient(abi.GRAD_ALL, compat=False, h=1e-5)
    got = _gradients(monkeypatch, s, [pos], True, perms=[sigma])[0][0]
    assert rel_err(got, ref) < 1e-6


def test_pair_once_mixed_replicas(monkeypatch):
    """Replicas of different pair actions and sizes in one context (the generic
    kernel instance), each against its own per-particle result."""
    systems = [system.config_c1(), _small_david(), system.config_c1(n_part=70, n_bead=6, n_level=2),
               system.config_ilkka()]
    paths = [system.random_path(s, seed=31 + i, spread=0.1) for i, s in enumerate(systems)]
    new = _gradients(monkeypatch, systems, paths, True)[0]
    old = _gradients(monkeypatch, systems, paths, False)[0]
    for i in range(len(systems)):
        assert rel_err(new[i], old[i]) < 1e-10, i


def test_pair_once_at_full_size(monkeypatch):
    """N=256, M=128 with the David pair action: every bead against the
    per-particle kernel; the short-range pair gradient sums to zero over the
    particles of each slice."""
    s = system.config_c2()
    pos = system.random_path(s, seed=41, spread=0.05)
    new = _gradients(monkeypatch, s, [pos], True, calls=2)
    old = _gradients(monkeypatch, s, [pos], False)[0][0]
    assert np.array_equal(new[0][0], new[1][0])
    assert rel_err(new[0][0], old) < 1e-10
    short = _gradients(monkeypatch, s, [pos], True, abi.GRAD_PAIR_SHORT)[0][0]
    assert np.abs(short.sum(axis=1)).max() < 1e-9 * max(1.0, np.abs(short).max())
