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
