"""ctypes binding of the engine's C ABI (include/spimc.h).

The binding loads ``libspimc.so`` built in-tree next to this file and fails
loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspimc.so")

OK = 0
ERR_ARG, ERR_UNSUPPORTED, ERR_HIP, ERR_STATE, ERR_NOMEM, ERR_NODEVICE = -1, -2, -3, -4, -5, -6

PAIR_NONE, PAIR_BARE, PAIR_DAVID, PAIR_ILKKA = 0, 1, 2, 3
ACTION_KINETIC, ACTION_PAIR = 0, 1
MODE_OLD, MODE_NEW = 0, 1
GRAD_KINETIC, GRAD_PAIR_SHORT, GRAD_PAIR_LONG, GRAD_ALL = 1, 2, 4, 7
ENERGY_THERMAL, ENERGY_VIRIAL_COMPAT, ENERGY_VIRIAL_ADJOINT = 0, 1, 2
OBS_ENERGY, OBS_SK, OBS_GR = 1, 2, 4
SWEEP_GRAPH, SWEEP_ADJOINT, SWEEP_ENERGY, SWEEP_SK, SWEEP_GR = 1, 2, 4, 8, 16
KERNEL_PROPOSE, KERNEL_PAIR, KERNEL_RHO, KERNEL_DECIDE, KERNEL_ADJOINT, KERNEL_ESTIMATORS = range(6)
TIMING_CLOCK, TIMING_EVENTS = 1, 2
TRACE_WORDS = 8  # action trace record of one move attempt

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class SystemDesc(C.Structure):
    """Mirror of spimc_system_desc."""

    _fields_ = [
        ("n_d", C.c_uint32),
        ("n_bead", C.c_uint32),
        ("beta", C.c_double),
        ("L", C.c_double),
        ("n_part", C.c_uint32),
        ("lam", C.c_double),
        ("n_k", C.c_uint32),
        ("k_n", _ip),
        ("kinetic_n_images", C.c_int32),
        ("bisect_n_level", C.c_uint32),
        ("bisect_n_images", C.c_int32),
        ("displace_step", C.c_double),
        ("pair_type", C.c_int32),
        ("pair_use_long_range", C.c_int32),
        ("pair_max_level", C.c_uint32),
        ("bare_is_coulomb", C.c_int32),
        ("bare_n_r", C.c_uint32),
        ("bare_r", _dp),
        ("bare_v_r", _dp),
        ("bare_n_r_long", C.c_uint32),
        ("bare_r_long", _dp),
        ("bare_v_long_r", _dp),
        ("david_n_order", C.c_uint32),
        ("david_n_grid", C.c_uint32),
        ("david_grid", _dp),
        ("david_u_kj", _dp),
        ("david_du_kj", _dp),
        ("pair_v_long_k", _dp),
        ("pair_v_long_k_0", C.c_double),
        ("pair_v_long_r_0", C.c_double),
        ("ilkka_n_x_u", C.c_uint32),
        ("ilkka_n_y_u", C.c_uint32),
        ("ilkka_x_u", _dp),
        ("ilkka_y_u", _dp),
        ("ilkka_u_xy", _dp),
        ("ilkka_n_x_du", C.c_uint32),
        ("ilkka_n_y_du", C.c_uint32),
        ("ilkka_x_du", _dp),
        ("ilkka_y_du", _dp),
        ("ilkka_du_xy", _dp),
        ("ilkka_n_r_v", C.c_uint32),
        ("ilkka_r_v", _dp),
        ("ilkka_v_r", _dp),
        ("ilkka_n_r_v_long", C.c_uint32),
        ("ilkka_r_v_long", _dp),
        ("ilkka_v_long_r", _dp),
        ("ilkka_n_r_u_long", C.c_uint32),
        ("ilkka_r_u_long", _dp),
        ("ilkka_u_long_r", _dp),
        ("ilkka_n_r_du_long", C.c_uint32),
        ("ilkka_r_du_long", _dp),
        ("ilkka_du_long_r", _dp),
        ("ilkka_u_long_k", _dp),
        ("ilkka_du_long_k", _dp),
        ("ilkka_du_long_k_0", C.c_double),
        ("ilkka_du_long_r_0", C.c_double),
    ]


# Every entry point of include/spimc.h with its ctypes signature.
_c = C.c_void_p
_u32, _i32, _u64, _d = C.c_uint32, C.c_int, C.c_uint64, C.c_double
_pu32, _pu64, _pu8 = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint8)
SIGNATURES = {
    "spimc_abi_version": ([], _i32),
    "spimc_last_error": ([], C.c_char_p),
    "spimc_kspace_build": ([_d, _d, _ip, _u32, _pu32], _i32),
    "spimc_ctx_create": ([C.POINTER(SystemDesc), _u32, _u64, C.POINTER(_c)], _i32),
    "spimc_ctx_destroy": ([_c], _i32),
    "spimc_ctx_num_replicas": ([_c, _pu32], _i32),
    "spimc_path_upload": ([_c, _u32, _dp], _i32),
    "spimc_path_download": ([_c, _u32, _dp], _i32),
    "spimc_rho_k_download": ([_c, _u32, _dp], _i32),
    "spimc_rho_k_rebuild": ([_c, _u32], _i32),
    "spimc_set_proposal": ([_c, _u32, _u32, _u32, _u32, _dp], _i32),
    "spimc_accept_proposal": ([_c, _u32], _i32),
    "spimc_reject_proposal": ([_c, _u32], _i32),
    "spimc_get_action": ([_c, _u32, _i32, _i32, _u32, _u32, _pu32, _u32, _u32, _dp], _i32),
    "spimc_get_action_noalloc": ([_c, _u32, _i32, _i32, _u32, _u32, _pu32, _u32, _u32, _dp], _i32),
    "spimc_d_action_d_beta": ([_c, _u32, _i32, _dp], _i32),
    "spimc_potential": ([_c, _u32, _i32, _dp], _i32),
    "spimc_kinetic_virial_energy": ([_c, _u32, _u32, _dp], _i32),
    "spimc_action_gradient": ([_c, _u32, _u32, _i32, _dp], _i32),
    "spimc_tape_set": ([_c, _u32, _dp, _u64], _i32),
    "spimc_set_bisect_level": ([_c, _u32, _u32], _i32),
    "spimc_tape_cursor": ([_c, _u32, _pu64], _i32),
    "spimc_bisect_stage": ([_c, _pu32, _u32, _pu8], _i32),
    "spimc_displace_stage": ([_c, _pu32, _u32, _pu8], _i32),
    "spimc_move_counters": ([_c, _u32, _pu64], _i32),
    "spimc_action_trace_enable": ([_c, _i32], _i32),
    "spimc_action_trace_read": ([_c, _u32, _u32, _dp], _i32),
    "spimc_perm_configure": ([_c, _u32, _u32, _d], _i32),
    "spimc_perm_bisect_stage": ([_c, _pu32, _u32, _pu8], _i32),
    "spimc_perm_table_configure": ([_c, _u32, _u32, _u32], _i32),
    "spimc_get_action_laplacian": ([_c, _u32, _i32, _i32, _u32, _u32, _pu32, _u32, _u32, _dp], _i32),
    "spimc_perm_table_stage": ([_c, _pu32, _u32, _pu8], _i32),
    "spimc_perm_download": ([_c, _u32, _pu32], _i32),
    "spimc_perm_upload": ([_c, _u32, _pu32], _i32),
    "spimc_perm_counters": ([_c, _u32, _pu64], _i32),
    "spimc_obs_configure": ([_c, _i32, _u32, _d, _d, _d, _u32], _i32),
    "spimc_obs_accumulate": ([_c, _u32], _i32),
    "spimc_obs_read": ([_c, _u32, _u32, _dp, _pu64], _i32),
    "spimc_obs_reset": ([_c, _u32], _i32),
    "spimc_sweep": ([_c, _u32, _u32], _i32),
    "spimc_bead_updates": ([_c, _pu64], _i32),
    "spimc_autotune": ([_c, _u32, C.POINTER(C.c_int32)], _i32),
    "spimc_timing_enable": ([_c, _i32], _i32),
    "spimc_kernel_timing": ([_c, _i32, _dp, _pu64, _dp], _i32),
    "spimc_timing_reset": ([_c], _i32),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads the engine library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"engine library {path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


class SpimcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"spimc error {code}: {msg}")
        self.code = code


def check(code: int) -> None:
    if code != OK:
        raise SpimcError(code, load().spimc_last_error().decode())


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def kspace_build(L: float, k_cut: float) -> np.ndarray:
    lib = load()
    n = C.c_uint32(0)
    check(lib.spimc_kspace_build(L, k_cut, None, 0, C.byref(n)))
    out = np.zeros((n.value, 3), dtype=np.int32)
    check(lib.spimc_kspace_build(L, k_cut, out.ctypes.data_as(_ip), n.value, C.byref(n)))
    return out


class Engine:
    """A context of one or more replicas (plain method: one replica)."""

    def __init__(self, systems, seed: int = 1):
        self.lib = load()
        if not isinstance(systems, (list, tuple)):
            systems = [systems]
        self.systems = list(systems)
        descs = (SystemDesc * len(systems))(*[s.desc() for s in systems])
        self._descs = descs
        h = C.c_void_p()
        check(self.lib.spimc_ctx_create(descs, len(systems), seed, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.spimc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- path --------------------------------------------------------------
    def upload(self, pos: np.ndarray, r: int = 0):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        check(self.lib.spimc_path_upload(self.h, r, dptr(pos)))

    def download(self, r: int = 0) -> np.ndarray:
        s = self.systems[r]
        out = np.zeros((3, s.n_bead, s.n_part))
        check(self.lib.spimc_path_download(self.h, r, dptr(out)))
        return out

    def rho_k(self, r: int = 0) -> np.ndarray:
        s = self.systems[r]
        out = np.zeros((s.n_bead, len(s.k_n), 2))
        check(self.lib.spimc_rho_k_download(self.h, r, dptr(out)))
        return out[..., 0] + 1j * out[..., 1]

    def rho_k_rebuild(self, r: int = 0):
        check(self.lib.spimc_rho_k_rebuild(self.h, r))

    # ---- actions -------------------------------------------------------------
    def set_proposal(self, p: int, b_first: int, r_new: np.ndarray, r: int = 0):
        r_new = np.ascontiguousarray(r_new, dtype=np.float64)
        check(self.lib.spimc_set_proposal(self.h, r, p, b_first, r_new.shape[0], dptr(r_new)))

    def accept(self, r: int = 0):
        check(self.lib.spimc_accept_proposal(self.h, r))

    def reject(self, r: int = 0):
        check(self.lib.spimc_reject_proposal(self.h, r))

    def get_action(self, action: int, mode: int, b0: int, b1: int, parts, level: int, r: int = 0) -> float:
        arr = (C.c_uint32 * max(1, len(parts)))(*parts)
        out = C.c_double()
        check(self.lib.spimc_get_action(self.h, r, action, mode, b0, b1, arr, len(parts), level, C.byref(out)))
        return out.value

    def get_action_laplacian(self, action: int, mode: int, b0: int, b1: int, parts, level: int, r: int = 0) -> float:
        arr = (C.c_uint32 * max(1, len(parts)))(*parts)
        out = C.c_double()
        check(self.lib.spimc_get_action_laplacian(self.h, r, action, mode, b0, b1, arr, len(parts), level,
                                                  C.byref(out)))
        return out.value

    def d_action_d_beta(self, action: int, r: int = 0) -> float:
        out = C.c_double()
        check(self.lib.spimc_d_action_d_beta(self.h, r, action, C.byref(out)))
        return out.value

    def potential(self, action: int, r: int = 0) -> float:
        out = C.c_double()
        check(self.lib.spimc_potential(self.h, r, action, C.byref(out)))
        return out.value

    def kinetic_virial_energy(self, window: int = 0, r: int = 0) -> float:
        out = C.c_double()
        check(self.lib.spimc_kinetic_virial_energy(self.h, r, window, C.byref(out)))
        return out.value

    def gradient(self, terms: int = GRAD_ALL, compat: bool = False, r: int = 0) -> np.ndarray:
        s = self.systems[r]
        out = np.zeros((s.n_bead, s.n_part, 3))
        check(self.lib.spimc_action_gradient(self.h, r, terms, int(compat), dptr(out)))
        return out

    # ---- moves ---------------------------------------------------------------
    def set_bisect_level(self, n_level: int, r: int = 0):
        check(self.lib.spimc_set_bisect_level(self.h, r, n_level))

    def tape_set(self, draws, r: int = 0):
        if draws is None:
            check(self.lib.spimc_tape_set(self.h, r, None, 0))
        else:
            d = np.ascontiguousarray(draws, dtype=np.float64)
            check(self.lib.spimc_tape_set(self.h, r, dptr(d), d.size))

    def tape_cursor(self, r: int = 0) -> int:
        v = C.c_uint64()
        check(self.lib.spimc_tape_cursor(self.h, r, C.byref(v)))
        return v.value

    def _stage(self, fn, n):
        R = len(self.systems)
        counts = np.broadcast_to(np.asarray(n, dtype=np.uint32), (R,)).copy()
        mx = int(counts.max()) if R else 0
        acc = np.zeros((R, max(mx, 1)), dtype=np.uint8)
        check(fn(self.h, counts.ctypes.data_as(_pu32), 0, acc.ctypes.data_as(_pu8)))
        return acc[:, :mx]

    def bisect(self, n):
        return self._stage(self.lib.spimc_bisect_stage, n)

    def displace(self, n):
        return self._stage(self.lib.spimc_displace_stage, n)

    def counters(self, r: int = 0) -> np.ndarray:
        out = np.zeros(4, dtype=np.uint64)
        check(self.lib.spimc_move_counters(self.h, r, out.ctypes.data_as(_pu64)))
        return out

    def trace(self, enable: bool = True):
        """Records each move attempt's level-0 action terms from now on."""
        check(self.lib.spimc_action_trace_enable(self.h, int(enable)))

    def trace_read(self, n: int, r: int = 0) -> np.ndarray:
        """[n][8]: kinetic OLD, NEW; short-range pair OLD, NEW; long-range pair
        OLD, NEW; log sample ratio; previous level's change (NaN rows: no
        level-0 decision) of the last stage's first n attempts."""
        out = np.zeros((n, TRACE_WORDS), dtype=np.float64)
        check(self.lib.spimc_action_trace_read(self.h, r, n, dptr(out)))
        return out

    # ---- permutations ----------------------------------------------------------
    def perm_configure(self, n_level: int, epsilon: float = 1e-100, r: int = 0):
        check(self.lib.spimc_perm_configure(self.h, r, n_level, epsilon))

    def perm_bisect(self, n):
        return self._stage(self.lib.spimc_perm_bisect_stage, n)

    def perm_table_configure(self, n_perm_part: int, balance: int = 0, r: int = 0):
        check(self.lib.spimc_perm_table_configure(self.h, r, n_perm_part, balance))

    def perm_table(self, n):
        return self._stage(self.lib.spimc_perm_table_stage, n)

    def perm(self, r: int = 0) -> np.ndarray:
        out = np.zeros(self.systems[r].n_part, dtype=np.uint32)
        check(self.lib.spimc_perm_download(self.h, r, out.ctypes.data_as(_pu32)))
        return out

    def set_perm(self, sigma, r: int = 0):
        s = np.ascontiguousarray(sigma, dtype=np.uint32)
        check(self.lib.spimc_perm_upload(self.h, r, s.ctypes.data_as(_pu32)))

    def perm_counters(self, r: int = 0):
        """(attempts, accepts, perm_attempt[N], perm_accept[N]) of the permuting bisection"""
        n = self.systems[r].n_part
        out = np.zeros(2 + 2 * n, dtype=np.uint64)
        check(self.lib.spimc_perm_counters(self.h, r, out.ctypes.data_as(_pu64)))
        return int(out[0]), int(out[1]), out[2:2 + n], out[2 + n:]

    # ---- observables ---------------------------------------------------------
    def obs_configure(self, energy=ENERGY_THERMAL, virial_window=0, sk_k_cut=1e300, gr_r_min=0.0, gr_r_max=0.0, gr_n_r=100):
        self.gr_n_r = gr_n_r
        check(self.lib.spimc_obs_configure(self.h, energy, virial_window, sk_k_cut, gr_r_min, gr_r_max, gr_n_r))

    def obs_accumulate(self, which: int):
        check(self.lib.spimc_obs_accumulate(self.h, which))

    def obs_read(self, which: int, r: int = 0):
        s = self.systems[r]
        n = {OBS_ENERGY: 2, OBS_SK: max(len(s.k_n), 1), OBS_GR: getattr(self, "gr_n_r", 100)}[which]
        out = np.zeros(n)
        nm = C.c_uint64()
        check(self.lib.spimc_obs_read(self.h, r, which, dptr(out), C.byref(nm)))
        return out, nm.value

    def obs_reset(self, which: int):
        check(self.lib.spimc_obs_reset(self.h, which))

    def sweep(self, n: int = 1, flags: int = 0):
        check(self.lib.spimc_sweep(self.h, n, flags))

    def autotune(self, n_rounds: int = 32) -> dict:
        """Try the launch shapes on n_rounds bisection attempts each (real moves of
        the chains) and keep the fastest; returns the shape chosen."""
        out = (C.c_int32 * 7)()
        check(self.lib.spimc_autotune(self.h, n_rounds, out))
        return {"fork": bool(out[0]), "rho_workgroups_per_cu": int(out[1]), "pair_workgroup": int(out[2]),
                "shapes_tried": int(out[3]), "displace_rounds": int(out[4]), "persistent_workgroup": int(out[5]),
                "one_launch": int(out[6])}

    def bead_updates(self) -> int:
        v = C.c_uint64()
        check(self.lib.spimc_bead_updates(self.h, C.byref(v)))
        return v.value

    # ---- measurement -----------------------------------------------------------
    def timing(self, mode: int):
        """0 off; TIMING_CLOCK (device clocks of the pair and rho_k kernels) and/or
        TIMING_EVENTS (dispatch events on every launch)."""
        check(self.lib.spimc_timing_enable(self.h, int(mode)))

    def kernel_timing(self, kernel: int):
        ms, n, b = C.c_double(), C.c_uint64(), C.c_double()
        check(self.lib.spimc_kernel_timing(self.h, kernel, C.byref(ms), C.byref(n), C.byref(b)))
        return ms.value, n.value, b.value

    def timing_reset(self):
        check(self.lib.spimc_timing_reset(self.h))
