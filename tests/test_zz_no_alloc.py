"""spimc_get_action_noalloc, called through the C ABI, against spimc_get_action.

The variant runs the same kernels with its scratch in a fixed device array, so
its results must equal the allocating entry point's bit for bit.  Two calls in
a row, the second on a smaller input, check that the array's leftover
contents from the first call do not leak into the second.
"""
import ctypes as C

import numpy as np
import pytest

from simpimc_amd import abi, system

pytestmark = pytest.mark.gpu

SCRATCH = 256 * 1024


def noalloc(e, action, mode, b0, b1, parts, level, r=0):
    arr = (C.c_uint32 * max(1, len(parts)))(*parts)
    out = C.c_double()
    rc = e.lib.spimc_get_action_noalloc(e.h, r, action, mode, b0, b1, arr, len(parts), level, C.byref(out))
    abi.check(rc)
    return out.value


def queries(s, p, b0):
    """(action, mode, b0, b1, level) from the largest window down to smaller ones."""
    nb = 1 << s.bisect_n_level
    out = [(abi.ACTION_PAIR, m, 0, s.n_bead, 0) for m in (abi.MODE_OLD, abi.MODE_NEW)]
    out += [(abi.ACTION_PAIR, m, b0, b0 + nb, 0) for m in (abi.MODE_OLD, abi.MODE_NEW)]
    out += [(abi.ACTION_KINETIC, m, b0, b0 + nb, lvl) for lvl in range(s.bisect_n_level) for m in (abi.MODE_OLD, abi.MODE_NEW)]
    return out


@pytest.mark.parametrize("name", ["bare_c1", "c2"])
def test_noalloc_get_action_equals_get_action(name):
    s = system.config_c1() if name == "bare_c1" else system.config_c2()
    pos = system.random_path(s, seed=3, spread=0.3)
    rng = np.random.default_rng(5)
    p, b0 = 5, 3
    nb = 1 << s.bisect_n_level
    new = pos[:, (b0 + 1 + np.arange(nb - 1)) % s.n_bead, p].T + rng.normal(0, 0.2, size=(nb - 1, 3))
    with abi.Engine(s) as e:
        e.upload(pos)
        # a proposal is pending, so NEW-mode pair queries run the rho_k update too
        e.set_proposal(p, b0 + 1, new)
        for action, mode, a, b, lvl in queries(s, p, b0):
            want = e.get_action(action, mode, a, b, [p], lvl)
            got1 = noalloc(e, action, mode, a, b, [p], lvl)
            # second call, smaller input: a shorter window of the same kind
            small = (a, a + max(1, (b - a) // 2)) if action == abi.ACTION_PAIR else (a, a + (1 << lvl))
            want2 = e.get_action(action, mode, small[0], small[1], [p], lvl)
            got2 = noalloc(e, action, mode, small[0], small[1], [p], lvl)
            assert got1 == want, (action, mode, a, b, lvl, got1, want)
            assert got2 == want2, (action, mode, small, lvl, got2, want2)
        e.accept()
        # after the accept, a fresh proposal on a smaller window
        e.set_proposal(p, b0 + 1, new[:1])
        for mode in (abi.MODE_OLD, abi.MODE_NEW):
            a = noalloc(e, abi.ACTION_PAIR, mode, b0, b0 + 2, [p], 0)
            b = e.get_action(abi.ACTION_PAIR, mode, b0, b0 + 2, [p], 0)
            assert a == b, (mode, a, b)
        e.reject()


def test_noalloc_get_action_rejects_queries_beyond_its_scratch():
    """A replica with more slices than the scratch holds partials for: the
    allocating entry point answers, the fixed-scratch one returns SPIMC_ERR_ARG."""
    M = SCRATCH // 32  # n_bead + 8 partials no longer fit
    s = system.config_c1(n_part=2, n_bead=M, n_level=2)
    s.k_n = s.k_n[:3]
    s.v_long_k = s.v_long_k[:3]
    pos = system.random_path(s, seed=1, spread=0.1)
    with abi.Engine(s) as e:
        e.upload(pos)
        a = e.get_action(abi.ACTION_PAIR, abi.MODE_OLD, 0, M, [0], 0)
        assert np.isfinite(a)
        with pytest.raises(abi.SpimcError) as ei:
            noalloc(e, abi.ACTION_PAIR, abi.MODE_OLD, 0, M, [0], 0)
        assert ei.value.code == abi.ERR_ARG
        # and a small query on another context still works after the refusal
    s2 = system.config_c1(n_part=8, n_bead=8, n_level=2)
    with abi.Engine(s2) as e2:
        e2.upload(system.random_path(s2, seed=2, spread=0.1))
        assert noalloc(e2, abi.ACTION_PAIR, abi.MODE_OLD, 0, 8, [1], 0) == e2.get_action(abi.ACTION_PAIR, abi.MODE_OLD, 0, 8, [1], 0)
