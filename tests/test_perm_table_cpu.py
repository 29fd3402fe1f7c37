"""PermBisectTable in the oracle (perm_bisect_table_class.h:10-228), pinned by
the ideal Bose gas and by a table written out in numpy.  CPU only.

* Ideal Bose gas (the closed form of test_permutation_cpu): with the kinetic
  action alone the sampled energy and mean k-cycle counts must match the
  canonical recursion.
  - n_perm_part = N (every permutation of the 3 particles in the table; the
    final table check does not run): the reference's move is exact.
  - n_perm_part = 2 < N: the reference's final check accepts with
    perm_tot_0 / perm_tot_1 (:216-225).  Detailed balance asks for
    perm_tot_0 / (perm_tot_1 w), w the selected entry's weight, because the
    reverse move selects the inverse entry, of weight 1 / w, from the new
    table.  With the reference's rule the chain is biased (many error bars off
    the closed form); with the balanced rule (balance=1) it is exact.  The
    oracle keeps the reference's rule by default.
* Table, enumeration and selection: on a 4-particle path the weights
  t(i, j) = exp((-|r_i - r'_j|^2 + |r_i - r'_i|^2) / (4 lambda tau n)) and the
  running totals of the cycle weights over the list (permutation types x
  subsets, both in lexicographic order: scaffold's GenPerm / GenCombPermK order
  is not in the reference tree) are written out in numpy.  A replay tape whose
  uniform falls inside one entry's interval must relink the path by that entry
  (the wrap permutation after the accepted move).
"""
import itertools
import math

import numpy as np
import pytest

import oracle_ffi
from simpimc_amd import abi, system
from test_permutation_cpu import BETA, LAM, L, M, N, boltzmann_energy, free_system, ideal_bose


def run_table_chain(k, balance, seed, n_iter=30000, n_burn=1000, n_blocks=40):
    s = free_system()
    o = oracle_ffi.Oracle(s, seed=seed)
    o.upload(system.random_path(s, seed=seed + 100))
    o.perm_table_configure(k, balance)
    es, cyc = [], []
    for it in range(n_iter):
        o.perm_table(3)
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
    return (be.mean(), be.std(ddof=1) / math.sqrt(n_blocks), bc.mean(axis=0),
            bc.std(axis=0, ddof=1) / math.sqrt(n_blocks), o)


@pytest.mark.parametrize("balance", [0, 1])
def test_table_of_all_particles_samples_the_ideal_bose_gas(balance):
    e_exact, nk_exact = ideal_bose(N, BETA, L, LAM)
    e, de, nk, dnk, o = run_table_chain(N, balance, seed=41)
    assert abs(e - e_exact) < 3.5 * de, (e, de, e_exact)
    assert abs(boltzmann_energy(N, BETA, L, LAM) - e_exact) > 5 * de
    for c in range(N):
        assert abs(nk[c] - nk_exact[c]) < 3.5 * dnk[c] + 1e-3, (c, nk, dnk, nk_exact)
    att, acc, perm_att, perm_acc = o.perm_counters()
    # every attempt reaches the bisection with all N chains (n_perm_part = N)
    assert att == 3 * 30000 and perm_att[N - 1] == att and perm_acc[N - 1] == acc
    assert perm_att[:N - 1].sum() == 0 and 0 < acc < att


def test_balanced_table_of_two_samples_the_ideal_bose_gas():
    e_exact, nk_exact = ideal_bose(N, BETA, L, LAM)
    e, de, nk, dnk, o = run_table_chain(2, 1, seed=43)
    assert abs(e - e_exact) < 3.5 * de, (e, de, e_exact)
    for c in range(N):
        assert abs(nk[c] - nk_exact[c]) < 3.5 * dnk[c] + 1e-3, (c, nk, dnk, nk_exact)
    att, acc, perm_att, perm_acc = o.perm_counters()
    assert perm_att[1] == att and perm_acc[1] == acc


def test_reference_rule_biases_the_table_of_two():
    """The reference's final check with n_perm_part < N: the two-cycle count is
    many error bars below the closed form (and below the balanced rule's)."""
    _, nk_exact = ideal_bose(N, BETA, L, LAM)
    _, _, nk, dnk, _ = run_table_chain(2, 0, seed=43)
    assert nk_exact[1] - nk[1] > 8 * dnk[1], (nk, dnk, nk_exact)


def _table_numpy(s, pos, b0, nbb, eps):
    n = s.n_part
    i4ltn = 1.0 / (4.0 * s.lam * s.tau) / nbb
    r0 = pos[:, b0, :].T
    r1 = pos[:, b0 + nbb, :].T  # no wrap: b0 + nbb < M, identity permutation
    d = r0[:, None, :] - r1[None, :, :]
    d -= np.rint(d / s.L) * s.L
    sq = np.einsum("ijd,ijd->ij", d, d)
    e = (-sq + np.diag(sq)[:, None]) * i4ltn
    return np.where(e > math.log(eps), np.exp(e), 0.0)


def _entries(t, k):
    n = t.shape[0]
    out = []
    for perm in itertools.permutations(range(k)):
        for sub in itertools.combinations(range(n), k):
            w = 1.0
            for i in range(k):
                w *= t[sub[i], sub[perm[i]]]
            out.append((sub, perm, w))
    return out


@pytest.mark.parametrize("k", [2, 3])
def test_table_selection_and_relinking_on_a_tape(k):
    s = system.System(n_part=4, n_bead=16, beta=0.8, L=3.0, lam=0.5, bisect_n_level=2)
    nbb, b0, eps = 4, 5, 1e-100
    pos = system.random_path(s, seed=8, spread=0.25)
    t = _table_numpy(s, pos, b0, nbb, eps)
    ents = _entries(t, k)
    tot = 0.0
    run = []
    for sub, perm, w in ents:
        if w > eps:
            tot += w
            run.append((tot, sub, perm, w))
    assert len(run) > 2
    n_norm = sum(3 * k * (nbb >> (lv + 1)) for lv in range(2))
    checked = 0
    for idx in range(len(run)):
        hi, sub, perm, w = run[idx]
        lo = run[idx - 1][0] if idx else 0.0
        if w < 1e-6 * tot:
            continue
        u = 0.5 * (lo + hi) / tot
        o = oracle_ffi.Oracle(s, seed=1)
        o.upload(pos)
        o.perm_configure(2)
        o.perm_table_configure(k, 1)
        # bead0, the table uniform, zero normals, level uniforms 0 (accept), final 0
        tape = [float(b0), u] + [0.0] * (n_norm + 2 + 1)
        o.tape_set(tape)
        acc = o.perm_table(1)
        assert acc[0] == 1
        sigma = np.arange(4)
        for i in range(k):
            sigma[sub[i]] = sub[perm[i]]
        assert np.array_equal(o.perm(), sigma), (idx, sub, perm, o.perm())
        checked += 1
    assert checked >= 3
