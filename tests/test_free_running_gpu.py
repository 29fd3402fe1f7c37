"""Free-running chains: the engine on its own device RNG and the oracle on its own
host RNG sample the same distribution.  For the Bare/Ewald electron gas (N=32,
M=32) the chain averages of the thermal energy and of the bisection and
displace acceptance rates agree within two standard errors of their difference;
each standard error comes from the spread of independent chains (32 replicas
on the GPU, 6 oracle chains), so autocorrelation within a chain does not bias
it.  Seeds are fixed, so the test is deterministic."""
import os

import numpy as np
import pytest

import oracle_ffi
from oracle_ffi import Oracle
from simpimc_amd import abi, system

pytestmark = pytest.mark.gpu

N_EQ, N_MEAS = 50, 200


def chain_stats(energies, acc_bis, acc_dis):
    out = {}
    for name, v in (("energy", energies), ("acc_bisect", acc_bis), ("acc_displace", acc_dis)):
        v = np.asarray(v, dtype=np.float64)
        out[name] = (float(v.mean()), float(v.std(ddof=1) / np.sqrt(len(v))))
    return out


def test_free_running_chains_agree_with_oracle():
    s = system.config_c1()
    n_rep = 32
    with abi.Engine([s] * n_rep, seed=20261) as e:
        for r in range(n_rep):
            e.upload(system.random_path(s, seed=500 + r, spread=0.3), r=r)
        e.obs_configure(energy=abi.ENERGY_THERMAL)
        e.sweep(N_EQ, 0)
        c0 = [e.counters(r=r) for r in range(n_rep)]
        e.obs_reset(abi.OBS_ENERGY)
        e.sweep(N_MEAS, abi.SWEEP_ENERGY)
        en, ab, ad = [], [], []
        for r in range(n_rep):
            v, nm = e.obs_read(abi.OBS_ENERGY, r=r)
            assert nm == N_MEAS
            en.append((v[0] + v[1]) / (s.n_bead * nm))
            c = e.counters(r=r) - c0[r]
            ab.append(c[1] / c[0])
            ad.append(c[3] / c[2])
    gpu = chain_stats(en, ab, ad)

    oracle_ffi.set_threads(min(16, len(os.sched_getaffinity(0))))
    try:
        en, ab, ad = [], [], []
        for k in range(6):
            o = Oracle(s, seed=900 + k)
            o.upload(system.random_path(s, seed=800 + k, spread=0.3))
            o.obs_configure(energy=abi.ENERGY_THERMAL)
            o.sweep(N_EQ, 0)
            c0 = o.counters()
            o.sweep(N_MEAS, abi.SWEEP_ENERGY)
            v, nm = o.obs_read(abi.OBS_ENERGY)
            assert nm == N_MEAS
            en.append((v[0] + v[1]) / (s.n_bead * nm))
            c = o.counters() - c0
            ab.append(c[1] / c[0])
            ad.append(c[3] / c[2])
    finally:
        oracle_ffi.set_threads(1)
    cpu = chain_stats(en, ab, ad)

    for name in gpu:
        (mg, sg), (mc, sc) = gpu[name], cpu[name]
        sigma = np.hypot(sg, sc)
        print(f"{name}: engine {mg:.6g} +- {sg:.2g}, oracle {mc:.6g} +- {sc:.2g}, |diff| / sigma = {abs(mg - mc) / sigma:.2f}")
        assert abs(mg - mc) < 2.0 * sigma, (name, gpu[name], cpu[name])
