"""GPU parity of PermBisectTable (perm_bisect_table_class.h:10-228) against the
CPU oracle.

* Replay: the oracle's draws are fed to the engine. The accept/reject
  sequences, positions (relabelled slices included), the wrap permutation and
  the move counters must match bit for bit, rho_k to 1e-11 relative. Table
  stages are interleaved with PermBisectIterative, Bisect and
  DisplaceParticle stages, for n_perm_part 2 and 3 and both rules of the final
  table check. The table sums are formed in another association than the
  oracle's single running sum: a selection or check could differ only if a
  uniform fell within rounding of a boundary.
* Batches: replicas of different systems, n_perm_part and attempt counts in one
  context each replay their own chain.
* Free running: free bosons sampled with the device RNG reproduce the ideal
  Bose gas closed form (energy and mean k-cycle counts) with the table of all
  particles and with the balanced table of two.
* Arguments: a stage on an unconfigured replica, n_perm_part out of range and a
  cycle list above 2^28 entries are refused before any device work.
"""
import numpy as np
import pytest

from oracle_ffi import Oracle
from simpimc_amd import abi, system
from test_permutation_cpu import BETA, LAM, L, M, N, free_system, ideal_bose
from test_permutation_gpu import SYSTEMS, dense_bare, rel_err, small_david

pytestmark = pytest.mark.gpu

REPLAY = [("bare_dense", 2, 0, "lds"), ("bare_dense", 3, 1, "lds"), ("bare_dense", 3, 0, "global"),
          ("david_small", 2, 1, "lds"), ("ilkka_c1", 2, 0, "lds"), ("free_bose", 3, 0, "lds"),
          ("free_bose", 2, 1, "lds")]


@pytest.mark.parametrize("name,k,balance,form", REPLAY)
def test_table_replay_bit_exact(name, k, balance, form, monkeypatch):
    if form == "global":
        monkeypatch.setenv("SPIMC_PERM_STAGE", "0")
    s = SYSTEMS[name]()
    pos = system.random_path(s, seed=5, spread=0.3)
    o = Oracle(s, seed=17)
    o.upload(pos)
    o.perm_table_configure(k, balance)
    o.record(True)
    n_b = max(1, s.n_part * s.n_bead >> s.bisect_n_level)
    ref = []
    for _ in range(3):
        ref.append(o.perm_table(n_b))
        ref.append(o.bisect(n_b))
        ref.append(o.perm_bisect(4))
        ref.append(o.displace(s.n_part))
    tape = o.recorded()
    print(f"{name} k={k} balance={balance} {form}: {len(tape)} draws", flush=True)
    with abi.Engine(s, seed=99) as e:
        e.upload(pos)
        e.perm_table_configure(k, balance)
        e.tape_set(tape)
        got = []
        for _ in range(3):
            got.append(e.perm_table(n_b)[0])
            got.append(e.bisect(n_b)[0])
            got.append(e.perm_bisect(4)[0])
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
        print(f"{name}: perm accepts per cycle length {og[3].tolist()}, sigma {o.perm().tolist()}", flush=True)
    assert og[2][k - 1] >= 3 * n_b


def test_table_batch_of_different_replicas():
    systems = [dense_bare(), small_david(), free_system(), system.config_c1(n_part=5, n_bead=8, n_level=2)]
    ks, balances = [2, 3, 3, 2], [0, 1, 0, 1]
    counts = np.array([24, 10, 17, 9], dtype=np.uint32)
    tapes, refs, finals, perms = [], [], [], []
    for i, s in enumerate(systems):
        pos = system.random_path(s, seed=90 + i, spread=0.3)
        o = Oracle(s, seed=500 + i)
        o.upload(pos)
        o.perm_table_configure(ks[i], balances[i])
        o.record(True)
        refs.append(o.perm_table(int(counts[i])))
        tapes.append(o.recorded())
        finals.append(o.download())
        perms.append(o.perm())
    with abi.Engine(systems, seed=1) as e:
        for i, s in enumerate(systems):
            e.upload(system.random_path(s, seed=90 + i, spread=0.3), r=i)
            e.perm_table_configure(ks[i], balances[i], r=i)
            e.tape_set(tapes[i], r=i)
        acc = e.perm_table(counts)
        for i in range(len(systems)):
            n = int(counts[i])
            assert np.array_equal(acc[i, :n], refs[i]), i
            assert np.array_equal(e.download(r=i), finals[i]), i
            assert np.array_equal(e.perm(r=i), perms[i]), i
            assert e.tape_cursor(r=i) == len(tapes[i]), i


@pytest.mark.parametrize("k,balance", [(3, 0), (2, 1)])
def test_table_free_bosons_match_the_closed_form(k, balance):
    s = free_system()
    n_rep, n_eq, n_meas, every = 64, 300, 2500, 5
    e_exact, nk_exact = ideal_bose(N, BETA, L, LAM)
    with abi.Engine([s] * n_rep, seed=777 + k) as e:
        for r in range(n_rep):
            e.upload(system.random_path(s, seed=400 + r), r=r)
            e.perm_table_configure(k, balance, r=r)
        e.obs_configure(energy=abi.ENERGY_THERMAL)
        for _ in range(n_eq):
            e.perm_table(3)
            e.bisect(3)
        e.obs_reset(abi.OBS_ENERGY)
        cyc = np.zeros((n_rep, N))
        n_cyc = 0
        for it in range(n_meas):
            e.perm_table(3)
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
                print(f"free bosons, table of {k}: sample {it} of {n_meas}", flush=True)
        en = []
        for r in range(n_rep):
            v, nm = e.obs_read(abi.OBS_ENERGY, r=r)
            en.append((v[0] + v[1]) / (M * nm))
        att = e.perm_counters(r=0)
    en = np.array(en)
    em, ese = en.mean(), en.std(ddof=1) / np.sqrt(n_rep)
    print(f"energy {em:.5f} +- {ese:.5f}, exact {e_exact:.5f}", flush=True)
    assert abs(em - e_exact) < 3.5 * ese
    nk = cyc / n_cyc
    nkm, nkse = nk.mean(axis=0), nk.std(axis=0, ddof=1) / np.sqrt(n_rep)
    print(f"cycles {nkm} +- {nkse}, exact {nk_exact}", flush=True)
    for c in range(N):
        assert abs(nkm[c] - nk_exact[c]) < 3.5 * nkse[c] + 1e-3, (c, nkm, nkse, nk_exact)
    assert att[2][k - 1] == att[0] and 0 < att[1] < att[0]


def test_table_arguments_are_checked():
    s = dense_bare()
    with abi.Engine([s, free_system()]) as e:
        with pytest.raises(abi.SpimcError):
            e.perm_table(1)  # no replica configured
        with pytest.raises(abi.SpimcError):
            e.perm_table_configure(7)
        with pytest.raises(abi.SpimcError):
            e.perm_table_configure(4, r=1)  # 3 particles
        with pytest.raises(abi.SpimcError):
            e.perm_table_configure(2, balance=2)
    big = system.config_c1(n_part=300, n_bead=4, n_level=1)
    with abi.Engine(big) as e:
        with pytest.raises(abi.SpimcError):
            e.perm_table_configure(4)  # 300 * 299 * 298 * 297 > 2^28 entries
