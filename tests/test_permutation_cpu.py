"""Permutation sampling in the oracle (PermBisectIterative,
perm_bisect_iterative_class.h:10-233, perm_bisect_class.h:7-135), pinned by
closed forms.  CPU only.

* Ideal Bose gas: with the kinetic action alone the path-integral weight is the
  exact density matrix of free bosons in a periodic box at any number of
  slices, so the thermal energy and the mean number of k-cycles of the sampled
  permutations must match the canonical recursion
  Z_N = (1/N) sum_k z(k beta) Z_{N-k}, <n_k> = z(k beta) Z_{N-k} / (k Z_N),
  E = -d ln Z_N / d beta, with z the one-particle partition function
  [sum_n exp(-beta lambda (2 pi n / L)^2)]^3.  The box is large enough that
  windings (weight e^{-L^2 / 4 lambda beta} = 3e-4) do not matter.  A dropped
  cycle weight, a wrong relinking or a wrong relabelling biases both.  The same
  run without permutations must give the Boltzmann energy instead, which
  differs by many error bars: the check can tell the two statistics apart.
* The kinetic action walks the permuted path (Kinetic::GetAction,
  kinetic_class.h:105-122): checked against link sums written out in numpy
  for a path with a permutation at the wrap, including the reference's
  one-sided wrap link of a whole-particle window (the DisplaceParticle case).
"""
import math

import numpy as np
import pytest

import oracle_ffi
from simpimc_amd import abi, system

N, M, L, BETA, LAM = 3, 16, 8.0, 4.0, 0.5


def ideal_bose(n, beta, box, lam):
    """(E, <n_k> for k = 1..n) of n free bosons in a periodic cube, canonical."""
    kk = 2.0 * np.pi * np.arange(-80, 81) / box

    def z1(b):
        return float(np.sum(np.exp(-b * lam * kk * kk))) ** 3

    def zn(b):
        z = [z1(k * b) for k in range(1, n + 1)]
        Z = [1.0]
        for m in range(1, n + 1):
            Z.append(sum(z[k - 1] * Z[m - k] for k in range(1, m + 1)) / m)
        return Z, z

    h = 1e-5
    e = -(math.log(zn(beta + h)[0][n]) - math.log(zn(beta - h)[0][n])) / (2 * h)
    Z, z = zn(beta)
    return e, [z[k - 1] * Z[n - k] / (k * Z[n]) for k in range(1, n + 1)]


def boltzmann_energy(n, beta, box, lam):
    kk = 2.0 * np.pi * np.arange(-80, 81) / box
    h = 1e-5
    f = lambda b: 3 * math.log(float(np.sum(np.exp(-b * lam * kk * kk))))  # noqa: E731
    return -n * (f(beta + h) - f(beta - h)) / (2 * h)


def free_system():
    return system.System(n_part=N, n_bead=M, beta=BETA, L=L, lam=LAM, kinetic_n_images=2, bisect_n_level=3,
                         bisect_n_images=2)


def run_chain(permute, seed, n_iter=60000, n_burn=1000, n_blocks=40):
    s = free_system()
    o = oracle_ffi.Oracle(s, seed=seed)
    o.upload(system.random_path(s, seed=seed + 100))
    es, cyc = [], []
    for it in range(n_iter):
        if permute:
            o.perm_bisect(3)
        o.bisect(3)
        if it >= n_burn:
            es.append(o.d_action_d_beta(abi.ACTION_KINETIC) / M)
            c = np.zeros(N)
            for ln in o.cycle_lengths():
                c[ln - 1] += 1
            cyc.append(c)
    es, cyc = np.array(es), np.array(cyc)
    nb = len(es) // n_blocks * n_blocks
    be = es[:nb].reshape(n_blocks, -1).mean(axis=1)
    bc = cyc[:nb].reshape(n_blocks, -1, N).mean(axis=1)
    return (be.mean(), be.std(ddof=1) / math.sqrt(n_blocks), bc.mean(axis=0), bc.std(axis=0, ddof=1) / math.sqrt(n_blocks),
            o)


def test_ideal_bose_gas_energy_and_cycle_counts():
    e_exact, nk_exact = ideal_bose(N, BETA, L, LAM)
    e_boltz = boltzmann_energy(N, BETA, L, LAM)
    e, de, nk, dnk, o = run_chain(True, seed=21)
    assert abs(e - e_exact) < 3.5 * de, (e, de, e_exact)
    # the check has power: the Boltzmann energy is many error bars away
    assert abs(e_boltz - e_exact) > 5 * de, (e_boltz, e_exact, de)
    for k in range(N):
        assert abs(nk[k] - nk_exact[k]) < 3.5 * dnk[k] + 1e-3, (k, nk, dnk, nk_exact)
    assert abs(np.sum(nk * np.arange(1, N + 1)) - N) < 1e-12
    att, acc, perm_att, perm_acc = o.perm_counters()
    assert att == 3 * 60000 and 0 < acc < att
    # every cycle length was attempted and accepted; attempts that reach the
    # bisection are counted per cycle length (perm_bisect_class.h:66-87)
    assert np.all(perm_acc > 0) and np.all(perm_att >= perm_acc)
    assert int(perm_acc.sum()) == acc


def test_boltzmann_chain_without_permutations():
    e_boltz = boltzmann_energy(N, BETA, L, LAM)
    e, de, nk, _, o = run_chain(False, seed=22, n_iter=40000)
    assert abs(e - e_boltz) < 3.5 * de, (e, de, e_boltz)
    assert np.array_equal(o.perm(), np.arange(N))
    assert nk[0] == N


def kinetic_links_numpy(s, pos, sigma, p, b0, n_links, level=0):
    """Kinetic::GetAction from bead (p, b0) over n_links links of 2^level slices,
    written out: after slice M - 1 the walk continues at slice 0 of sigma[label]."""
    skip = 1 << level
    lam_tau = s.lam * s.tau * skip
    tot, lbl, b = 0.0, p, b0
    for _ in range(n_links):
        nl, nb = lbl, b
        for _ in range(skip):
            if nb == s.n_bead - 1:
                nl, nb = sigma[nl], 0
            else:
                nb += 1
        d = pos[:, b, lbl] - pos[:, nb, nl]
        d -= np.rint(d / s.L) * s.L
        tot += float(d @ d) / (4.0 * lam_tau)
        lbl, b = nl, nb
    return tot


@pytest.mark.parametrize("level", [0, 1])
def test_kinetic_action_follows_the_permutation(level):
    s = system.System(n_part=4, n_bead=8, beta=2.0, L=4.0, lam=0.5, bisect_n_level=2)
    pos = system.random_path(s, seed=5, spread=0.4)
    o = oracle_ffi.Oracle(s, seed=1)
    o.upload(pos)
    sigma = np.array([2, 0, 1, 3])  # a 3-cycle and a fixed point
    o.set_perm(sigma)
    assert np.array_equal(o.perm(), sigma)
    for p in range(4):
        for b0 in range(0, 8, 2):
            for span in (4, 8):  # a window that crosses the wrap, and a whole particle
                got = o.get_action(abi.ACTION_KINETIC, abi.MODE_NEW, b0, b0 + span, [p], level)
                want = kinetic_links_numpy(s, pos, sigma, p, b0, span >> level, level)
                if span == 8 and sigma[p] == p:
                    want = 0.0  # the walk returns to its first bead: no link (kinetic_class.h:111-118)
                assert got == pytest.approx(want, rel=1e-13, abs=1e-13), (p, b0, span)
    # the thermal estimator's links also follow the permutation
    lt = s.lam * s.tau
    dbeta = s.n_part * s.n_bead * 3 / (2 * s.tau)
    for p in range(4):
        for b in range(8):
            nl, nb = (sigma[p], 0) if b == 7 else (p, b + 1)
            d = pos[:, b, p] - pos[:, nb, nl]
            d -= np.rint(d / s.L) * s.L
            dbeta -= float(d @ d) / (4.0 * lt * s.tau)
    assert o.d_action_d_beta(abi.ACTION_KINETIC) == pytest.approx(dbeta, rel=1e-12)
    with pytest.raises(RuntimeError):
        o.set_perm([0, 0, 1, 2])


def test_cycle_lengths_of_a_permutation():
    s = system.System(n_part=6, n_bead=4, beta=1.0, L=5.0, bisect_n_level=1)
    o = oracle_ffi.Oracle(s, seed=1)
    o.upload(system.random_path(s, seed=1))
    o.set_perm([1, 2, 0, 4, 3, 5])
    assert list(o.cycle_lengths()) == [3, 2, 1]


def degenerate_rows_system():
    """Free bosons whose beads sit independently at random, with a table cutoff
    epsilon so large that most rows t(p, .) of the permutation table have no
    entry above it: the walk there has no candidate to select."""
    s = free_system()
    rng = np.random.default_rng(9)
    pos = rng.uniform(-L / 2, L / 2, size=(3, M, N))
    return s, pos


def test_walk_without_candidates_rejects_free_running_and_fails_on_a_tape():
    """A cycle walk whose row has no candidate (every weight below epsilon):
    free-running, the attempt is rejected and the chain goes on; replaying the
    same draws from a tape, it is an error (the reference would revisit the
    same particle)."""
    s, pos = degenerate_rows_system()
    o = oracle_ffi.Oracle(s, seed=5)
    o.upload(pos)
    o.perm_configure(3, epsilon=0.5)
    o.record(True)
    acc = o.perm_bisect(200)
    tape = o.recorded()
    c = o.perm_counters()
    # attempts ran without error; most walks stopped at a row without candidates
    assert c[0] == 200 and acc.sum() == c[1]
    assert int(np.sum(c[2])) < 200
    r = oracle_ffi.Oracle(s, seed=6)
    r.upload(pos)
    r.perm_configure(3, epsilon=0.5)
    r.tape_set(tape)
    with pytest.raises(RuntimeError, match="selected no particle"):
        r.perm_bisect(200)
