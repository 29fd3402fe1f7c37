"""Free-running chains: the engine on its own device RNG and the oracle on its own
host RNG sample the same distribution.  For the Bare/Ewald electron gas (N=32,
M=32) the chain averages of the thermal energy and of the bisection and
displace acceptance rates agree within two standard errors of their difference;
each standard error comes from the spread of independent chains (32 replicas
on the GPU, 12 oracle chains), so autocorrelation within a chain does not bias
it.  S(k) on the first three k shells and g(r) in seven coarse bins are
compared together: the sum of their squared differences in units of the
standard error stays below the 95.45 % quantile of chi-square with as many
degrees of freedom (the two-sigma level of a joint test).  Seeds are fixed, so
the test is deterministic."""
import os

import numpy as np
import pytest
import scipy.stats

import oracle_ffi
from oracle_ffi import Oracle
from simpimc_amd import abi, system

pytestmark = pytest.mark.gpu


def sk_cut(s):
    """k cut that keeps the shells |n|^2 <= 3 of the measured S(k)."""
    return (np.sqrt(3.0) + 1e-6) * 2 * np.pi / s.L

N_EQ, N_MEAS = 50, 200
N_ORACLE = 12
GR_BINS = 8
FLAGS = abi.SWEEP_ENERGY | abi.SWEEP_SK | abi.SWEEP_GR


def chain_stats(energies, acc_bis, acc_dis):
    out = {}
    for name, v in (("energy", energies), ("acc_bisect", acc_bis), ("acc_displace", acc_dis)):
        v = np.asarray(v, dtype=np.float64)
        out[name] = (float(v.mean()), float(v.std(ddof=1) / np.sqrt(len(v))))
    return out


def structure(s, sk, gr, nm):
    """S(k) averaged over the shells |n|^2 = 1, 2, 3 and g(r) counts per pair
    and slice in bins 1..7 of 8 over [0, L/2], for one chain."""
    n2 = (np.asarray(s.k_n) ** 2).sum(axis=1)
    shells = [sk[n2 == m].mean() / (nm * s.n_bead * s.n_part) for m in (1, 2, 3)]
    pairs = 0.5 * s.n_part * (s.n_part - 1)
    return np.concatenate([shells, gr[1:GR_BINS] / (nm * s.n_bead * pairs)])


def mean_se(rows):
    a = np.asarray(rows, dtype=np.float64)
    return a.mean(axis=0), a.std(axis=0, ddof=1) / np.sqrt(len(a))


def test_free_running_chains_agree_with_oracle():
    s = system.config_c1()
    n_rep = 32
    with abi.Engine([s] * n_rep, seed=20261) as e:
        for r in range(n_rep):
            e.upload(system.random_path(s, seed=500 + r, spread=0.3), r=r)
        e.obs_configure(energy=abi.ENERGY_THERMAL, sk_k_cut=sk_cut(s), gr_r_max=s.L / 2, gr_n_r=GR_BINS)
        e.sweep(N_EQ, 0)
        c0 = [e.counters(r=r) for r in range(n_rep)]
        e.obs_reset(abi.OBS_ENERGY | abi.OBS_SK | abi.OBS_GR)
        e.sweep(N_MEAS, FLAGS)
        en, ab, ad, st = [], [], [], []
        for r in range(n_rep):
            v, nm = e.obs_read(abi.OBS_ENERGY, r=r)
            assert nm == N_MEAS
            en.append((v[0] + v[1]) / (s.n_bead * nm))
            c = e.counters(r=r) - c0[r]
            ab.append(c[1] / c[0])
            ad.append(c[3] / c[2])
            st.append(structure(s, e.obs_read(abi.OBS_SK, r=r)[0], e.obs_read(abi.OBS_GR, r=r)[0], nm))
    gpu = chain_stats(en, ab, ad)
    gpu_st = mean_se(st)

    oracle_ffi.set_threads(min(16, len(os.sched_getaffinity(0))))
    try:
        en, ab, ad, st = [], [], [], []
        for k in range(N_ORACLE):
            o = Oracle(s, seed=900 + k)
            o.upload(system.random_path(s, seed=800 + k, spread=0.3))
            o.obs_configure(energy=abi.ENERGY_THERMAL, sk_k_cut=sk_cut(s), gr_r_max=s.L / 2, gr_n_r=GR_BINS)
            o.sweep(N_EQ, 0)
            c0 = o.counters()
            o.sweep(N_MEAS, FLAGS)
            v, nm = o.obs_read(abi.OBS_ENERGY)
            assert nm == N_MEAS
            en.append((v[0] + v[1]) / (s.n_bead * nm))
            c = o.counters() - c0
            ab.append(c[1] / c[0])
            ad.append(c[3] / c[2])
            st.append(structure(s, o.obs_read(abi.OBS_SK)[0], o.obs_read(abi.OBS_GR)[0], nm))
            print(f"oracle chain {k} done", flush=True)
    finally:
        oracle_ffi.set_threads(1)
    cpu = chain_stats(en, ab, ad)
    cpu_st = mean_se(st)

    for name in gpu:
        (mg, sg), (mc, sc) = gpu[name], cpu[name]
        sigma = np.hypot(sg, sc)
        print(f"{name}: engine {mg:.6g} +- {sg:.2g}, oracle {mc:.6g} +- {sc:.2g}, |diff| / sigma = {abs(mg - mc) / sigma:.2f}")
        assert abs(mg - mc) < 2.0 * sigma, (name, gpu[name], cpu[name])
    (mg, sg), (mc, sc) = gpu_st, cpu_st
    z = (mg - mc) / np.hypot(sg, sc)
    for i, name in enumerate(["S(k) |n|^2=1", "S(k) |n|^2=2", "S(k) |n|^2=3"] + [f"g(r) bin {b}" for b in range(1, GR_BINS)]):
        print(f"{name}: engine {mg[i]:.6g} +- {sg[i]:.2g}, oracle {mc[i]:.6g} +- {sc[i]:.2g}, z = {z[i]:+.2f}")
    chi2, dof = float((z ** 2).sum()), len(z)
    limit = scipy.stats.chi2.ppf(0.9545, dof)
    print(f"S(k), g(r): chi2 = {chi2:.2f} over {dof} (two-sigma limit {limit:.2f})")
    assert chi2 < limit, (chi2, limit, z)


def test_long_chain_thermal_and_virial_energies_agree():
    """The plain-method configuration at its stated length: 32 replicas of the
    N=32, M=32 Bare/Ewald gas run 10^4 sweeps on the device RNG (after 200 to
    equilibrate), twice from the same seed: once accumulating the thermal energy
    estimator, once the virial one (from the analytic adjoint).  Estimators draw
    nothing, so both runs are the same chains (final paths equal bit for bit);
    the two estimators of the energy must agree on them within 3 standard
    errors of their per-replica difference."""
    s = system.config_c1()
    n_rep, n_eq, n_meas = 32, 200, 10000
    out = {}
    for est in (abi.ENERGY_THERMAL, abi.ENERGY_VIRIAL_ADJOINT):
        with abi.Engine([s] * n_rep, seed=31337) as e:
            for r in range(n_rep):
                e.upload(system.random_path(s, seed=700 + r, spread=0.3), r=r)
            e.obs_configure(energy=est)
            e.sweep(n_eq, abi.SWEEP_GRAPH)
            e.obs_reset(abi.OBS_ENERGY)
            flags = abi.SWEEP_GRAPH | abi.SWEEP_ENERGY | (abi.SWEEP_ADJOINT if est == abi.ENERGY_VIRIAL_ADJOINT else 0)
            for blk in range(10):
                e.sweep(n_meas // 10, flags)
                print(f"estimator {est}: {(blk + 1) * n_meas // 10} of {n_meas} sweeps", flush=True)
            en = []
            for r in range(n_rep):
                v, nm = e.obs_read(abi.OBS_ENERGY, r=r)
                assert nm == n_meas
                en.append((v[0] + v[1]) / (s.n_bead * nm))
            out[est] = (np.array(en), [e.download(r=r) for r in range(n_rep)])
    (e_t, p_t), (e_v, p_v) = out[abi.ENERGY_THERMAL], out[abi.ENERGY_VIRIAL_ADJOINT]
    for r in range(n_rep):
        assert np.array_equal(p_t[r], p_v[r]), r
    d = e_t - e_v
    dm, dse = d.mean(), d.std(ddof=1) / np.sqrt(n_rep)
    print(f"thermal {e_t.mean():.6f} +- {e_t.std(ddof=1) / np.sqrt(n_rep):.2g}, virial {e_v.mean():.6f} +- "
          f"{e_v.std(ddof=1) / np.sqrt(n_rep):.2g}, difference {dm:.3g} +- {dse:.2g}", flush=True)
    assert abs(dm) < 3.0 * dse, (dm, dse)
    # the virial estimator's variance is the smaller (its purpose)
    assert e_v.std() < e_t.std()
