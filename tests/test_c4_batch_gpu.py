"""The uneven batch at its stated size: 64 replicas with N in {16, 32, 64, 128},
M in {16, 32, 64} and r_s in [1, 4], each on its own box and k-set (the
reference's MPI clones, framework_class.h:44-53), advanced together by whole
sweeps in one context, eagerly and from captured hipGraphs, with the stages
run as rounds of launches (three per round, or one whose last workgroup per
replica decides and proposes) or as one persistent launch per stage.

Every replica replays its own oracle chain: a sweep of N M / 2^l bisections
(bisect_class.h:39-124) and N displacements (displace_particle_class.h:28-72),
then the thermal energy, S(k) and g(r) estimators.  Checked per replica:
positions and move counters bit for bit, the whole tape consumed, rho_k within
1e-11 N, the energies and S(k) within 1e-10 relative, g(r) counts exactly.
"""
import numpy as np
import pytest

from oracle_ffi import Oracle
from simpimc_amd import abi, system

pytestmark = pytest.mark.gpu

SWEEPS = 1
FLAGS = abi.SWEEP_ENERGY | abi.SWEEP_SK | abi.SWEEP_GR
GR_BINS = 50

_ref = {}


def oracle_chains():
    if "c4" in _ref:
        return _ref["c4"]
    systems = system.config_c4()
    out = []
    for i, s in enumerate(systems):
        pos = system.random_path(s, seed=1000 + i, spread=0.05)
        o = Oracle(s, seed=77 + i)
        o.upload(pos)
        o.obs_configure(energy=abi.ENERGY_THERMAL, gr_n_r=GR_BINS)
        o.record(True)
        o.sweep(SWEEPS, FLAGS)
        out.append(dict(pos=pos, tape=o.recorded(), final=o.download(), rho=o.rho_k(), counters=o.counters(),
                        energy=o.obs_read(abi.OBS_ENERGY), sk=o.obs_read(abi.OBS_SK), gr=o.obs_read(abi.OBS_GR)))
    _ref["c4"] = (systems, out)
    return systems, out


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b))) / max(1.0, float(np.max(np.abs(b)))) if b.size else 0.0


# stages as rounds of launches (one: each round in one launch), or each stage in
# one persistent launch (workgroup size)
# (one: 1 bisection rounds in one launch, 2 displace rounds too; ipb: rho_k
# items per workgroup of those launches, 0 as sized for the GPU)
@pytest.mark.parametrize("graph,persistent,one,ipb", [(False, 0, 0, 0), (True, 0, 0, 0), (False, 256, 0, 0),
                                                      (True, 256, 0, 0), (False, 1024, 0, 0), (False, 0, 1, 0),
                                                      (True, 0, 1, 0), (False, 0, 1, 3), (False, 0, 2, 0),
                                                      (True, 0, 2, 5)])
def test_c4_batch_replays_every_replica(graph, persistent, one, ipb, monkeypatch):
    monkeypatch.setenv("SPIMC_PERSISTENT", str(persistent))
    monkeypatch.setenv("SPIMC_ONE_LAUNCH", str(one))
    if ipb:
        monkeypatch.setenv("SPIMC_ROUND_IPB", str(ipb))
    systems, ref = oracle_chains()
    assert len(systems) == 64
    assert {s.n_part for s in systems} == {16, 32, 64, 128} and {s.n_bead for s in systems} == {16, 32, 64}
    assert all(s.n_k == 257 for s in systems)
    print(f"C4: {sum(s.n_part * s.n_bead for s in systems)} beads, graph={graph}, persistent={persistent}, one={one}, "
          f"ipb={ipb}", flush=True)
    with abi.Engine(systems, seed=1) as e:
        for i, r in enumerate(ref):
            e.upload(r["pos"], r=i)
            e.tape_set(r["tape"], r=i)
        e.obs_configure(energy=abi.ENERGY_THERMAL, gr_n_r=GR_BINS)
        e.sweep(SWEEPS, FLAGS | (abi.SWEEP_GRAPH if graph else 0))
        for i, (s, r) in enumerate(zip(systems, ref)):
            assert e.tape_cursor(r=i) == len(r["tape"]), i
            assert np.array_equal(e.counters(r=i), r["counters"]), (i, e.counters(r=i), r["counters"])
            assert np.array_equal(e.download(r=i), r["final"]), i
            assert rel_err(e.rho_k(r=i), r["rho"]) < 1e-11 * s.n_part, i
            ge, gn = e.obs_read(abi.OBS_ENERGY, r=i)
            assert gn == r["energy"][1] == SWEEPS
            assert rel_err(ge, r["energy"][0]) < 1e-10, (i, ge, r["energy"][0])
            assert rel_err(e.obs_read(abi.OBS_SK, r=i)[0], r["sk"][0]) < 1e-10, i
            assert np.array_equal(e.obs_read(abi.OBS_GR, r=i)[0], r["gr"][0]), i
        # the batch rejected moves of both kinds somewhere (dilute replicas may accept every bisection)
        c = np.array([r["counters"] for r in ref], dtype=np.int64).sum(axis=0)
        assert 0 < c[1] < c[0] and 0 < c[3] < c[2], c
