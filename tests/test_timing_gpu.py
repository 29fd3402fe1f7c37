"""Device-clock timing of the move kernels: one pair and one rho_k span per round
that had work, in eager launches and in graph replay, and timing leaves the
chain unchanged."""
import numpy as np
import pytest

from simpimc_amd import abi, system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph,fork", [(False, "0"), (True, "0"), (False, "1"), (True, "1")])
def test_clock_spans_per_round(graph, fork, monkeypatch):
    monkeypatch.setenv("SPIMC_FORK", fork)
    s = system.config_c1()
    pos = system.random_path(s, seed=4, spread=0.05)
    n_b = (s.n_part * s.n_bead) >> s.bisect_n_level
    with abi.Engine(s, seed=9) as e, abi.Engine(s, seed=9) as ref:
        e.upload(pos)
        ref.upload(pos)
        e.timing(abi.TIMING_CLOCK)
        e.timing_reset()
        flags = abi.SWEEP_GRAPH if graph else 0
        e.sweep(1, flags)
        ref.sweep(1, flags)
        rounds = n_b + s.n_part
        for k in (abi.KERNEL_PAIR, abi.KERNEL_RHO):
            ms, n, by = e.kernel_timing(k)
            # graph replay runs whole 16-round graphs; rounds past a replica's count do no work
            assert (rounds <= n < rounds + 32) if graph else (n == rounds), (k, n, rounds)
            assert 0.0 < ms < 60e3
            assert by > 0
        _, _, rho_bytes = e.kernel_timing(abi.KERNEL_RHO)
        # 32 B per (slice, k): bisect moves 2^l - 1 slices, displace all M; proposals
        # rejected at a coarse bisection level never reach the rho_k kernel
        slices = n_b * ((1 << s.bisect_n_level) - 1) + s.n_part * s.n_bead
        assert 32.0 * s.n_part * s.n_bead * s.n_k < rho_bytes <= 32.0 * slices * s.n_k
        assert np.array_equal(e.download(), ref.download())
        e.timing_reset()
        assert e.kernel_timing(abi.KERNEL_RHO)[1] == 0
