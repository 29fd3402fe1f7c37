"""ctypes wrapper of the CPU oracle (oracle/libspimc_oracle.so).

Test infrastructure: mirrors simpimc_amd.abi.Engine for one replica so that the
parity tests drive the oracle and the HIP engine with the same calls.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from simpimc_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "libspimc_oracle.so")
# the same source built with -O3 -ffast-math: for timing the CPU baseline only
LIB_FAST = os.path.join(ROOT, "oracle", "libspimc_oracle_fast.so")

_dp, _ip = abi._dp, abi._ip
_c, _u32, _i32, _u64, _d = C.c_void_p, C.c_uint32, C.c_int, C.c_uint64, C.c_double
_pu32, _pu64, _pu8 = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint8)

SIG = {
    "spimc_oracle_last_error": ([], C.c_char_p),
    "spimc_oracle_set_threads": ([_i32], None),
    "spimc_oracle_max_threads": ([], _i32),
    "spimc_oracle_kspace_build": ([_d, _d, _ip, _u32, _pu32], _i32),
    "spimc_oracle_create": ([C.POINTER(abi.SystemDesc), _u64, C.POINTER(_c)], _i32),
    "spimc_oracle_destroy": ([_c], _i32),
    "spimc_oracle_path_upload": ([_c, _dp], _i32),
    "spimc_oracle_path_download": ([_c, _i32, _dp], _i32),
    "spimc_oracle_rho_k_download": ([_c, _i32, _dp], _i32),
    "spimc_oracle_rho_k_rebuild": ([_c], _i32),
    "spimc_oracle_set_proposal": ([_c, _u32, _u32, _u32, _dp], _i32),
    "spimc_oracle_accept_proposal": ([_c], _i32),
    "spimc_oracle_reject_proposal": ([_c], _i32),
    "spimc_oracle_get_action": ([_c, _i32, _i32, _u32, _u32, _pu32, _u32, _u32, _dp], _i32),
    "spimc_oracle_d_action_d_beta": ([_c, _i32, _dp], _i32),
    "spimc_oracle_potential": ([_c, _i32, _dp], _i32),
    "spimc_oracle_kinetic_virial_energy": ([_c, _u32, _dp], _i32),
    "spimc_oracle_action_gradient": ([_c, _u32, _i32, _d, _dp], _i32),
    "spimc_oracle_action_gradient_beads": ([_c, _u32, _d, _u32, C.POINTER(C.c_uint32), _dp], _i32),
    "spimc_oracle_action_totals": ([_c, _dp], _i32),
    "spimc_oracle_tape_set": ([_c, _dp, _u64], _i32),
    "spimc_oracle_tape_cursor": ([_c, _pu64], _i32),
    "spimc_oracle_record": ([_c, _i32], _i32),
    "spimc_oracle_recorded": ([_c, _dp, _u64, _pu64], _i32),
    "spimc_oracle_trace": ([_c, _i32], _i32),
    "spimc_oracle_trace_read": ([_c, _dp, _u64, _pu64], _i32),
    "spimc_oracle_bisect_stage": ([_c, _u32, _pu8], _i32),
    "spimc_oracle_displace_stage": ([_c, _u32, _pu8], _i32),
    "spimc_oracle_move_counters": ([_c, _pu64], _i32),
    "spimc_oracle_bead_updates": ([_c, _pu64], _i32),
    "spimc_oracle_obs_configure": ([_c, _i32, _u32, _d, _d, _d, _u32], _i32),
    "spimc_oracle_obs_accumulate": ([_c, _u32], _i32),
    "spimc_oracle_obs_read": ([_c, _u32, _dp, _pu64], _i32),
    "spimc_oracle_sweep": ([_c, _u32, _u32], _i32),
    "spimc_oracle_spline_eval": ([_dp, _dp, _u32, _dp, _u32, _dp, _dp], _i32),
    "spimc_oracle_spline2_eval": ([_dp, _u32, _dp, _u32, _dp, _dp, _dp, _u32, _dp, _dp, _dp], _i32),
    "spimc_oracle_pair_link": ([_c, _i32, _d, _d, _d, _dp], _i32),
    "spimc_oracle_set_bisect_level": ([_c, _u32], _i32),
    "spimc_oracle_free_spline_log_rho": ([_d, _i32, _d, _d, _dp, _u32, _dp], _i32),
    "spimc_oracle_perm_bisect_stage": ([_c, _u32, _pu8], _i32),
    "spimc_oracle_perm_counters": ([_c, _pu64], _i32),
    "spimc_oracle_perm_configure": ([_c, _u32, _d], _i32),
    "spimc_oracle_perm_table_configure": ([_c, _u32, _u32], _i32),
    "spimc_oracle_get_action_laplacian": ([_c, _i32, _i32, _u32, _u32, _pu32, _u32, _u32, _dp], _i32),
    "spimc_oracle_perm_table_stage": ([_c, _u32, _pu8], _i32),
    "spimc_oracle_perm_get": ([_c, _pu32], _i32),
    "spimc_oracle_perm_set": ([_c, _pu32], _i32),
    "spimc_oracle_cycle_lengths": ([_c, _pu32, _pu32], _i32),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle library {LIB} missing; run __graft_entry__.build()")
        L = C.CDLL(LIB)
        for k, (a, r) in SIG.items():
            f = getattr(L, k)
            f.argtypes = a
            f.restype = r
        _lib = L
    return _lib


def use_timing_build():
    """Loads the -O3 -ffast-math build instead of the checker build; must come
    before the first call into the oracle in this process."""
    global LIB
    if _lib is not None and LIB != LIB_FAST:
        raise RuntimeError("the checker build of the oracle is already loaded")
    LIB = LIB_FAST


def check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib().spimc_oracle_last_error().decode()}")


def dp(a):
    return abi.dptr(a)


def set_threads(n: int):
    lib().spimc_oracle_set_threads(n)


def kspace_build(L, k_cut):
    n = C.c_uint32()
    check(lib().spimc_oracle_kspace_build(L, k_cut, None, 0, C.byref(n)))
    out = np.zeros((n.value, 3), dtype=np.int32)
    check(lib().spimc_oracle_kspace_build(L, k_cut, out.ctypes.data_as(_ip), n.value, C.byref(n)))
    return out


def spline2_eval(x, y, vals, px, py):
    """Tensor natural spline through vals[ny][nx]: values and d/dx, d/dy at the points."""
    x, y = np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(y, dtype=np.float64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    px, py = np.ascontiguousarray(px, dtype=np.float64), np.ascontiguousarray(py, dtype=np.float64)
    out, gx, gy = np.zeros(len(px)), np.zeros(len(px)), np.zeros(len(px))
    check(lib().spimc_oracle_spline2_eval(dp(x), len(x), dp(y), len(y), dp(vals), dp(px), dp(py), len(px), dp(out),
                                          dp(gx), dp(gy)))
    return out, gx, gy


def spline_eval(x, y, t):
    x, y, t = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, t))
    out, dout = np.zeros_like(t), np.zeros_like(t)
    check(lib().spimc_oracle_spline_eval(dp(x), dp(y), len(x), dp(t), len(t), dp(out), dp(dout)))
    return out, dout


def free_log_rho(L, n_images, lam, tau, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(r.shape[0])
    check(lib().spimc_oracle_free_spline_log_rho(L, n_images, lam, tau, dp(r), r.shape[0], dp(out)))
    return out


class Oracle:
    def __init__(self, system, seed: int = 1):
        self.L = lib()
        self.s = system
        self.n_part = system.n_part
        self._desc = system.desc()
        h = C.c_void_p()
        check(self.L.spimc_oracle_create(C.byref(self._desc), seed, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.L.spimc_oracle_destroy(self.h)
            self.h = None

    def upload(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        check(self.L.spimc_oracle_path_upload(self.h, dp(pos)))

    def download(self, mode=0):
        out = np.zeros((3, self.s.n_bead, self.s.n_part))
        check(self.L.spimc_oracle_path_download(self.h, mode, dp(out)))
        return out

    def rho_k(self, mode=0):
        out = np.zeros((self.s.n_bead, self.s.n_k, 2))
        check(self.L.spimc_oracle_rho_k_download(self.h, mode, dp(out)))
        return out[..., 0] + 1j * out[..., 1]

    def rho_k_rebuild(self):
        check(self.L.spimc_oracle_rho_k_rebuild(self.h))

    def set_proposal(self, p, b_first, r_new):
        r_new = np.ascontiguousarray(r_new, dtype=np.float64)
        check(self.L.spimc_oracle_set_proposal(self.h, p, b_first, r_new.shape[0], dp(r_new)))

    def accept(self):
        check(self.L.spimc_oracle_accept_proposal(self.h))

    def reject(self):
        check(self.L.spimc_oracle_reject_proposal(self.h))

    def get_action(self, action, mode, b0, b1, parts, level):
        arr = (C.c_uint32 * max(1, len(parts)))(*parts)
        out = C.c_double()
        check(self.L.spimc_oracle_get_action(self.h, action, mode, b0, b1, arr, len(parts), level, C.byref(out)))
        return out.value

    def get_action_laplacian(self, action, mode, b0, b1, parts, level):
        pv = np.asarray(parts, dtype=np.uint32)
        out = np.zeros(1)
        check(self.L.spimc_oracle_get_action_laplacian(self.h, action, mode, b0, b1, pv.ctypes.data_as(_pu32), len(pv),
                                                       level, dp(out)))
        return float(out[0])

    def set_bisect_level(self, n_level):
        check(self.L.spimc_oracle_set_bisect_level(self.h, n_level))

    def pair_link(self, which, r0, r1, s):
        """One link of the pair action from |r|, |r'|, |r - r'|: which 0 CalcU,
        1 CalcdUdBeta, 2 CalcV."""
        out = C.c_double()
        check(self.L.spimc_oracle_pair_link(self.h, which, r0, r1, s, C.byref(out)))
        return out.value

    def d_action_d_beta(self, action):
        out = C.c_double()
        check(self.L.spimc_oracle_d_action_d_beta(self.h, action, C.byref(out)))
        return out.value

    def potential(self, action):
        out = C.c_double()
        check(self.L.spimc_oracle_potential(self.h, action, C.byref(out)))
        return out.value

    def kinetic_virial_energy(self, window=0):
        out = C.c_double()
        check(self.L.spimc_oracle_kinetic_virial_energy(self.h, window, C.byref(out)))
        return out.value

    def gradient(self, terms=abi.GRAD_ALL, compat=False, h=1e-5):
        out = np.zeros((self.s.n_bead, self.s.n_part, 3))
        check(self.L.spimc_oracle_action_gradient(self.h, terms, int(compat), h, dp(out)))
        return out

    def gradient_beads(self, beads, terms=abi.GRAD_ALL, h=1e-5):
        """Central-difference dS/dr of the beads [(slice, particle), ...] only: [n][3]."""
        b = np.ascontiguousarray(np.asarray(beads, dtype=np.uint32).reshape(-1, 2))
        out = np.zeros((len(b), 3))
        check(self.L.spimc_oracle_action_gradient_beads(self.h, terms, h, len(b),
                                                         b.ctypes.data_as(C.POINTER(C.c_uint32)), dp(out)))
        return out

    def action_totals(self):
        out = np.zeros(3)
        check(self.L.spimc_oracle_action_totals(self.h, dp(out)))
        return out

    def tape_set(self, draws):
        if draws is None:
            check(self.L.spimc_oracle_tape_set(self.h, None, 0))
        else:
            d = np.ascontiguousarray(draws, dtype=np.float64)
            check(self.L.spimc_oracle_tape_set(self.h, dp(d), d.size))

    def tape_cursor(self):
        v = C.c_uint64()
        check(self.L.spimc_oracle_tape_cursor(self.h, C.byref(v)))
        return v.value

    def record(self, enable=True):
        check(self.L.spimc_oracle_record(self.h, int(enable)))

    def recorded(self):
        n = C.c_uint64()
        check(self.L.spimc_oracle_recorded(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        check(self.L.spimc_oracle_recorded(self.h, dp(out), n.value, C.byref(n)))
        return out

    def trace(self, enable=True):
        """Records each move attempt's level-0 action terms (cleared here)."""
        check(self.L.spimc_oracle_trace(self.h, int(enable)))

    def trace_read(self):
        n = C.c_uint64()
        check(self.L.spimc_oracle_trace_read(self.h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 9))
        check(self.L.spimc_oracle_trace_read(self.h, dp(out), n.value, C.byref(n)))
        return out

    def bisect(self, n):
        acc = np.zeros(max(n, 1), dtype=np.uint8)
        check(self.L.spimc_oracle_bisect_stage(self.h, n, acc.ctypes.data_as(_pu8)))
        return acc[:n]

    def displace(self, n):
        acc = np.zeros(max(n, 1), dtype=np.uint8)
        check(self.L.spimc_oracle_displace_stage(self.h, n, acc.ctypes.data_as(_pu8)))
        return acc[:n]

    def perm_bisect(self, n):
        acc = np.zeros(max(n, 1), dtype=np.uint8)
        check(self.L.spimc_oracle_perm_bisect_stage(self.h, n, acc.ctypes.data_as(_pu8)))
        return acc[:n]

    def perm_table_configure(self, n_perm_part, balance=0):
        check(self.L.spimc_oracle_perm_table_configure(self.h, n_perm_part, balance))

    def perm_table(self, n):
        acc = np.zeros(max(n, 1), dtype=np.uint8)
        check(self.L.spimc_oracle_perm_table_stage(self.h, n, acc.ctypes.data_as(_pu8)))
        return acc[:n]

    def perm_counters(self):
        """(attempts, accepts, perm_attempt[N], perm_accept[N]) of the permuting bisection"""
        out = np.zeros(2 + 2 * self.n_part, dtype=np.uint64)
        check(self.L.spimc_oracle_perm_counters(self.h, out.ctypes.data_as(_pu64)))
        return int(out[0]), int(out[1]), out[2:2 + self.n_part], out[2 + self.n_part:]

    def perm_configure(self, n_level, epsilon=1e-100):
        check(self.L.spimc_oracle_perm_configure(self.h, n_level, epsilon))

    def perm(self):
        out = np.zeros(self.n_part, dtype=np.uint32)
        check(self.L.spimc_oracle_perm_get(self.h, out.ctypes.data_as(_pu32)))
        return out

    def set_perm(self, sigma):
        s = np.ascontiguousarray(sigma, dtype=np.uint32)
        check(self.L.spimc_oracle_perm_set(self.h, s.ctypes.data_as(_pu32)))

    def cycle_lengths(self):
        out = np.zeros(self.n_part, dtype=np.uint32)
        n = C.c_uint32()
        check(self.L.spimc_oracle_cycle_lengths(self.h, out.ctypes.data_as(_pu32), C.byref(n)))
        return out[:n.value]

    def counters(self):
        out = np.zeros(4, dtype=np.uint64)
        check(self.L.spimc_oracle_move_counters(self.h, out.ctypes.data_as(_pu64)))
        return out

    def bead_updates(self):
        v = C.c_uint64()
        check(self.L.spimc_oracle_bead_updates(self.h, C.byref(v)))
        return v.value

    def obs_configure(self, energy=abi.ENERGY_THERMAL, virial_window=0, sk_k_cut=1e300, gr_r_min=0.0, gr_r_max=0.0,
                      gr_n_r=100):
        self.gr_n_r = gr_n_r
        check(self.L.spimc_oracle_obs_configure(self.h, energy, virial_window, sk_k_cut, gr_r_min, gr_r_max, gr_n_r))

    def obs_accumulate(self, which):
        check(self.L.spimc_oracle_obs_accumulate(self.h, which))

    def obs_read(self, which):
        n = {abi.OBS_ENERGY: 2, abi.OBS_SK: max(self.s.n_k, 1), abi.OBS_GR: getattr(self, "gr_n_r", 100)}[which]
        out = np.zeros(n)
        nm = C.c_uint64()
        check(self.L.spimc_oracle_obs_read(self.h, which, dp(out), C.byref(nm)))
        return out, nm.value

    def sweep(self, n=1, flags=0):
        check(self.L.spimc_oracle_sweep(self.h, n, flags))
