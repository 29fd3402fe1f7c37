"""Systems, pair-action tables and initial paths for the engine.

This module plays the role of simpimc's input side for the hot path: it holds
one replica's <System>/<Species>/<Action>/<Move> parameters, builds the
k-vector set of KSpace::Setup, the Ewald-split pair-action tables in the
layouts the reference loaders read (Bare: v/diag/*, bare_pair_action_class.h:39-86;
David: u_kj_<n>/du_kj_dbeta_<n>/potential/long_range/squarer,
david_pair_action_class.h:192-360; Ilkka: u/off_diag, du/off_diag, u/diag,
du/diag, v/diag, ilkka_pair_action_class.h:253-432), and initial paths
(species_class.h:222-377).
The synthetic tables stand in for pagen output, which is not available.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi


def heg_box_length(n_part: int, r_s: float) -> float:
    """Cubic box side of a homogeneous electron gas: L = (4 pi N / 3)^(1/3) r_s."""
    return (4.0 * math.pi * n_part / 3.0) ** (1.0 / 3.0) * r_s


def kspace(L: float, k_cut: float) -> np.ndarray:
    """k-vector set of KSpace::Setup/Include (k_space_class.h:17-80): integer
    triples n with |2 pi n / L|^2 < k_cut^2, n != 0, in the half space; the
    floating-point test is the reference's.  Order: lexicographic in n."""
    box = 2.0 * math.pi / L
    mx = int(math.ceil(1.1 * k_cut / box))
    rng = np.arange(-mx, mx + 1)
    a, b, c = np.meshgrid(rng, rng, rng, indexing="ij")
    a, b, c = a.ravel(), b.ravel(), c.ravel()
    kx, ky, kz = a * box, b * box, c * box
    k2 = kx * kx + ky * ky + kz * kz
    half = (kx > 0) | ((kx == 0) & (ky > 0)) | ((kx == 0) & (ky == 0) & (kz > 0))
    keep = (k2 < k_cut * k_cut) & (k2 != 0) & half
    return np.stack([a[keep], b[keep], c[keep]], axis=1).astype(np.int32)


def kmag(L: float, k_n: np.ndarray) -> np.ndarray:
    box = 2.0 * math.pi / L
    k = k_n.astype(np.float64) * box
    return np.sqrt(k[:, 0] * k[:, 0] + k[:, 1] * k[:, 1] + k[:, 2] * k[:, 2])


def _erf(x: np.ndarray) -> np.ndarray:
    return np.vectorize(math.erf, otypes=[np.float64])(x)


def _erfc(x: np.ndarray) -> np.ndarray:
    return np.vectorize(math.erfc, otypes=[np.float64])(x)


def ewald_k(L: float, alpha: float, k_n: np.ndarray) -> tuple[np.ndarray, float, float]:
    """Ewald long-range Coulomb pieces for a unit-charge homologous pair action:
    v_k = 4 pi exp(-k^2 / 4 alpha^2) / (V k^2) per stored (half-space) k,
    v_long_k_0 = -pi / (alpha^2 V) (neutralising background), v_long_r_0 = 2 alpha / sqrt(pi)."""
    vol = L ** 3
    k = kmag(L, k_n)
    vk = 4.0 * math.pi * np.exp(-k * k / (4.0 * alpha * alpha)) / (vol * k * k)
    return vk, -math.pi / (alpha * alpha * vol), 2.0 * alpha / math.sqrt(math.pi)


@dataclass
class System:
    """One replica: the fields of spimc_system_desc as numpy data."""

    n_part: int
    n_bead: int
    beta: float
    L: float
    lam: float = 0.5
    k_n: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), dtype=np.int32))
    kinetic_n_images: int = 0
    bisect_n_level: int = 3
    bisect_n_images: int = 0
    displace_step: float = 0.0
    pair_type: int = abi.PAIR_NONE
    use_long_range: bool = False
    bare_is_coulomb: bool = False
    bare_r: Optional[np.ndarray] = None
    bare_v_r: Optional[np.ndarray] = None
    bare_r_long: Optional[np.ndarray] = None
    bare_v_long_r: Optional[np.ndarray] = None
    david_n_order: int = 0
    david_grid: Optional[np.ndarray] = None
    david_u_kj: Optional[np.ndarray] = None
    david_du_kj: Optional[np.ndarray] = None
    v_long_k: Optional[np.ndarray] = None
    v_long_k_0: float = 0.0
    v_long_r_0: float = 0.0
    # IlkkaPairAction knots: keys of ILKKA_KEYS (grids, values, per-k vectors, constants)
    ilkka: Optional[dict] = None
    name: str = ""

    @property
    def tau(self) -> float:
        return self.beta / self.n_bead

    @property
    def n_k(self) -> int:
        return int(len(self.k_n))

    def desc(self) -> abi.SystemDesc:
        """Builds the C descriptor; arrays stay owned by this object."""
        # arrays handed to C stay referenced for the lifetime of this object, also
        # across several descriptors built from it (one per replica)
        keep = self.__dict__.setdefault("_keep", [])

        def arr(a, dt=np.float64):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data_as(abi._dp if dt == np.float64 else abi._ip)

        d = abi.SystemDesc()
        d.n_d = 3
        d.n_bead = self.n_bead
        d.beta = self.beta
        d.L = self.L
        d.n_part = self.n_part
        d.lam = self.lam
        d.n_k = self.n_k
        d.k_n = arr(self.k_n, np.int32) if self.n_k else None
        d.kinetic_n_images = self.kinetic_n_images
        d.bisect_n_level = self.bisect_n_level
        d.bisect_n_images = self.bisect_n_images
        d.displace_step = self.displace_step
        d.pair_type = self.pair_type
        d.pair_use_long_range = int(self.use_long_range)
        d.pair_max_level = 0
        d.bare_is_coulomb = int(self.bare_is_coulomb)
        d.bare_n_r = 0 if self.bare_r is None else len(self.bare_r)
        d.bare_r = arr(self.bare_r)
        d.bare_v_r = arr(self.bare_v_r)
        d.bare_n_r_long = 0 if self.bare_r_long is None else len(self.bare_r_long)
        d.bare_r_long = arr(self.bare_r_long)
        d.bare_v_long_r = arr(self.bare_v_long_r)
        d.david_n_order = self.david_n_order
        d.david_n_grid = 0 if self.david_grid is None else len(self.david_grid)
        d.david_grid = arr(self.david_grid)
        d.david_u_kj = arr(self.david_u_kj)
        d.david_du_kj = arr(self.david_du_kj)
        d.pair_v_long_k = arr(self.v_long_k)
        d.pair_v_long_k_0 = self.v_long_k_0
        d.pair_v_long_r_0 = self.v_long_r_0
        if self.ilkka is not None:
            t = self.ilkka
            for g, x, y, v in (("u", "x_u", "y_u", "u_xy"), ("du", "x_du", "y_du", "du_xy")):
                setattr(d, f"ilkka_n_x_{g}", len(t[x]))
                setattr(d, f"ilkka_n_y_{g}", len(t[y]))
                setattr(d, f"ilkka_{x}", arr(t[x]))
                setattr(d, f"ilkka_{y}", arr(t[y]))
                setattr(d, f"ilkka_{v}", arr(t[v]))  # [n_y][n_x], x fastest
            for r, v in (("r_v", "v_r"), ("r_v_long", "v_long_r"), ("r_u_long", "u_long_r"), ("r_du_long", "du_long_r")):
                if t.get(r) is not None:
                    setattr(d, f"ilkka_n_{r}", len(t[r]))
                    setattr(d, f"ilkka_{r}", arr(t[r]))
                    setattr(d, f"ilkka_{v}", arr(t[v]))
            d.ilkka_u_long_k = arr(t.get("u_long_k"))
            d.ilkka_du_long_k = arr(t.get("du_long_k"))
            d.ilkka_du_long_k_0 = t.get("du_long_k_0", 0.0)
            d.ilkka_du_long_r_0 = t.get("du_long_r_0", 0.0)
        return d


# ---------------------------------------------------------------------------
# pair-action tables
# ---------------------------------------------------------------------------
def bare_coulomb_ewald(sys: System, alpha: float, n_r: int = 1000, n_r_long: int = 1000, r_min: float = 1e-3) -> System:
    """BarePairAction tables for the primitive Coulomb action tau V with an Ewald
    split (v/diag/*): V(r) = 1/r via is_coulomb, v_long_r = erf(alpha r)/r on a
    grid over [0, sqrt(3) L/2], so that the short range is erfc(alpha r)/r."""
    r_max = math.sqrt(3.0) * sys.L / 2.0
    sys.pair_type = abi.PAIR_BARE
    sys.use_long_range = True
    sys.bare_is_coulomb = True
    sys.bare_r = np.linspace(r_min, r_max, n_r)
    sys.bare_v_r = 1.0 / sys.bare_r
    rl = np.linspace(0.0, r_max, n_r_long)
    vl = np.empty_like(rl)
    vl[0] = 2.0 * alpha / math.sqrt(math.pi)
    vl[1:] = _erf(alpha * rl[1:]) / rl[1:]
    sys.bare_r_long = rl
    sys.bare_v_long_r = vl
    sys.v_long_k, sys.v_long_k_0, sys.v_long_r_0 = ewald_k(sys.L, alpha, sys.k_n)
    return sys


def david_n_val(n_order: int) -> int:
    """Number of u_kj components (david_pair_action_class.h:250-253)."""
    return 1 + sum(1 + i for i in range(1, n_order + 1))


def _smooth_perturbation(r: np.ndarray, r_max: float, rng: np.random.Generator, amp: float) -> np.ndarray:
    a = amp * rng.uniform(0.5, 1.0) * rng.choice([-1.0, 1.0])
    w = rng.uniform(0.5, 3.0)
    ph = rng.uniform(0.0, 2.0 * math.pi)
    return a * (1.0 - r / r_max) ** 2 * np.exp(-0.5 * r) * np.cos(w * r + ph)


def david_file_arrays(L: float, alpha: float, tau: float, n_grid: int, n_order: int, seed: int, amp: float = 1e-3,
                      r_min: float = 0.01):
    """Pair-action file contents in DavidPairAction's layout for the primitive
    approximation of the short-range Ewald Coulomb potential V(r) = erfc(alpha r)/r,
    plus a seeded smooth perturbation of amplitude `amp` on every u_kj and
    du_kj_dbeta component: grid, potential/data, u_kj_<n>/data [n_val][n_grid],
    du_kj_dbeta_<n>/data [n_val][n_grid]."""
    r_max = math.sqrt(3.0) * L / 2.0
    grid = np.linspace(r_min, r_max, n_grid)
    V = _erfc(alpha * grid) / grid
    n_val = david_n_val(n_order)
    rng = np.random.default_rng(seed)
    u = np.zeros((n_val, n_grid))
    du = np.zeros((n_val, n_grid))
    u[0] = tau * V  # primitive diagonal action
    # du_kj_dbeta component 0 holds dU/dtau - V, which is 0 for the primitive action
    for c in range(n_val):
        u[c] += _smooth_perturbation(grid, r_max, rng, amp)
        du[c] += _smooth_perturbation(grid, r_max, rng, amp)
    return grid, V, u, du


def david_assemble(grid: np.ndarray, V: np.ndarray, u: np.ndarray, du: np.ndarray):
    """The spline knots DavidPairAction's loader builds (david_pair_action_class.h:262-309):
    component 0 = potential, components 1.. = file data; the loader's fill loop
    stops at n_grid - 1, so the last knot of every component stays 0."""
    n_val, n_grid = u.shape
    uk = np.zeros((n_val + 1, n_grid))
    duk = np.zeros((n_val + 1, n_grid))
    uk[0, : n_grid - 1] = V[: n_grid - 1]
    duk[0, : n_grid - 1] = V[: n_grid - 1]
    uk[1:, : n_grid - 1] = u[:, : n_grid - 1]
    duk[1:, : n_grid - 1] = du[:, : n_grid - 1]
    return uk, duk


def david_ewald(sys: System, alpha: float, n_grid: int = 2000, n_order: int = 2, seed: int = 7, amp: float = 1e-3,
                r_min: float = 0.01) -> System:
    grid, V, u, du = david_file_arrays(sys.L, alpha, sys.tau, n_grid, n_order, seed, amp, r_min)
    uk, duk = david_assemble(grid, V, u, du)
    sys.pair_type = abi.PAIR_DAVID
    sys.use_long_range = True
    sys.david_n_order = n_order
    sys.david_grid = grid
    sys.david_u_kj = uk
    sys.david_du_kj = duk
    # long_range/u_k holds 4 pi e^{-k^2/4a^2}/k^2 unnormalised; the loader divides by the volume
    sys.v_long_k, sys.v_long_k_0, sys.v_long_r_0 = ewald_k(sys.L, alpha, sys.k_n)
    return sys


def _graded_grid(a: float, b: float, n: int, power: float = 1.6) -> np.ndarray:
    """n strictly increasing knots on [a, b], denser near a (a general grid)."""
    t = np.linspace(0.0, 1.0, n)
    return a + (b - a) * (0.5 * t + 0.5 * t ** power)


def _soft(q: np.ndarray, a: float) -> np.ndarray:
    return 1.0 / np.sqrt(q * q + a * a)


def ilkka_ewald(sys: System, alpha: float, n_x: int = 40, n_y: int = 30, n_r: int = 200, soft: float = 0.3,
                eps: float = 1e-3) -> System:
    """IlkkaPairAction tables on general (graded) grids for a soft-core Coulomb
    pair with an Ewald split.  U(x, y) = tau v(q) + eps tau s^2 e^{-q} and
    dU/dbeta(x, y) = v(q) + 2 eps s^2 e^{-q}, with q = (x + y)/2, s = x - y and
    v(q) = 1/sqrt(q^2 + soft^2); x spans [0, sqrt(3) L + 0.1] and y [0,
    sqrt(3) L/2 + 0.1], the range the minimum-image link geometry reaches, so the
    tensor splines are never evaluated outside their grid.  v/diag v_r = v(r);
    the long-range parts are erf(alpha r)/r (u/diag scaled by tau) on [0.01,
    sqrt(3) L/2] with the Ewald k components of ewald_k (u_long_k = tau v_k)."""
    tau = sys.tau
    r_cap = math.sqrt(3.0) * sys.L / 2.0
    x = _graded_grid(0.0, 2.0 * r_cap + 0.1, n_x)
    y = _graded_grid(0.0, r_cap + 0.1, n_y, power=1.3)
    X, Y = np.meshgrid(x, y, indexing="xy")  # [n_y][n_x], x fastest
    q, sd = 0.5 * (X + Y), X - Y
    u_xy = tau * _soft(q, soft) + eps * tau * sd * sd * np.exp(-q)
    du_xy = _soft(q, soft) + 2.0 * eps * sd * sd * np.exp(-q)
    r = _graded_grid(0.01, r_cap, n_r, power=1.2)
    lr = _erf(alpha * r) / r
    vk, vk0, vr0 = ewald_k(sys.L, alpha, sys.k_n)
    sys.pair_type = abi.PAIR_ILKKA
    sys.use_long_range = True
    sys.v_long_k, sys.v_long_k_0, sys.v_long_r_0 = vk, vk0, vr0
    sys.ilkka = {
        "x_u": x, "y_u": y, "u_xy": u_xy, "x_du": x, "y_du": y, "du_xy": du_xy,
        "r_v": r, "v_r": _soft(r, soft), "r_v_long": r, "v_long_r": lr,
        "r_u_long": r, "u_long_r": tau * lr, "r_du_long": r, "du_long_r": lr,
        "u_long_k": tau * vk, "du_long_k": vk, "du_long_k_0": vk0, "du_long_r_0": vr0,
    }
    return sys


def config_ilkka(n_part: int = 32, n_bead: int = 32, r_s: float = 2.0, tau: float = 0.05, n_level: int = 3) -> System:
    """The C1 electron gas with an IlkkaPairAction (soft-core Coulomb, Ewald alpha
    = 5/L, |k| <= 5 2pi/L)."""
    s = config_c1(n_part=n_part, n_bead=n_bead, r_s=r_s, tau=tau, n_level=n_level)
    s.bare_r = s.bare_v_r = s.bare_r_long = s.bare_v_long_r = None
    s.bare_is_coulomb = False
    s.name = "HEG-Ilkka"
    return ilkka_ewald(s, alpha=5.0 / s.L)


# ---------------------------------------------------------------------------
# initial paths (species_class.h:262-300), positions as r[3][M][N]
# ---------------------------------------------------------------------------
def random_path(sys: System, seed: int, spread: float = 0.0, cofactor: float = 1.0) -> np.ndarray:
    """init_type="Random": one uniform point per particle in +-cofactor L/2, on all
    beads; `spread` adds independent Gaussian offsets per bead (test paths)."""
    rng = np.random.default_rng(seed)
    N, M, L = sys.n_part, sys.n_bead, sys.L
    centre = rng.uniform(-cofactor * L / 2.0, cofactor * L / 2.0, size=(N, 3))
    pos = np.broadcast_to(centre.T[:, None, :], (3, M, N)).copy()
    if spread > 0.0:
        pos += rng.normal(0.0, spread, size=pos.shape)
    return pos


def bcc_path(sys: System) -> np.ndarray:
    """init_type="BCC" (species_class.h:275-300)."""
    N, M, L = sys.n_part, sys.n_bead, sys.L
    n = int(math.ceil((0.5 * N) ** (1.0 / 3.0)))
    delta = L / n
    pos = np.zeros((3, M, N))
    for p_i in range(N):
        p = p_i // 2
        t0 = p // (n * n)
        t1 = (p - t0 * n * n) // n
        t2 = p - t0 * n * n - t1 * n
        r = np.array([delta * t0 - 0.5 * L, delta * t1 - 0.5 * L, delta * t2 - 0.5 * L])
        if p_i % 2:
            r += 0.5 * delta
        pos[:, :, p_i] = r[:, None]
    return pos


# ---------------------------------------------------------------------------
# the workloads
# ---------------------------------------------------------------------------
def config_c1(n_part: int = 32, n_bead: int = 32, r_s: float = 2.0, tau: float = 0.05, n_level: int = 3) -> System:
    """Plain method, CPU-runnable: 32 Boltzmann electrons at r_s = 2, M = 32,
    tau = 0.05, primitive Coulomb with Ewald split (alpha = 5/L, |k| <= 5 2pi/L,
    257 stored half-space vectors), bisection level 3 plus displace (step L/10)."""
    L = heg_box_length(n_part, r_s)
    s = System(n_part=n_part, n_bead=n_bead, beta=tau * n_bead, L=L, bisect_n_level=n_level, displace_step=L / 10.0,
               name="HEG-Bare-C1")
    s.k_n = kspace(L, (5.0 + 1e-9) * 2.0 * math.pi / L)
    return bare_coulomb_ewald(s, alpha=5.0 / L)


def config_c2(n_part: int = 256, n_bead: int = 128, r_s: float = 1.0, tau: float = 0.02, n_level: int = 4,
              k_shell: float = 97.5, n_grid: int = 2000, seed: int = 7) -> System:
    """HEG at scale: 256 Boltzmann electrons, r_s = 1, M = 128, tau = 0.02, David
    table (n_order 2, 2000-point linear grid to sqrt(3) L/2, primitive + seeded
    1e-3 perturbation), Ewald alpha = 7/L with |n|^2 <= 97 (1979 stored k),
    bisection level 4."""
    L = heg_box_length(n_part, r_s)
    s = System(n_part=n_part, n_bead=n_bead, beta=tau * n_bead, L=L, bisect_n_level=n_level, displace_step=L / 10.0,
               name="HEG-David-C2")
    s.k_n = kspace(L, math.sqrt(k_shell) * 2.0 * math.pi / L)
    return david_ewald(s, alpha=7.0 / L, n_grid=n_grid, seed=seed)


def config_c4(seed: int = 11, n_replicas: int = 64) -> list[System]:
    """64 independent replicas of uneven N in {16,32,64,128}, M in {16,32,64},
    r_s ~ U[1,4], each with the C1 interaction on its own box and k-set."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_replicas):
        N = int(rng.choice([16, 32, 64, 128]))
        M = int(rng.choice([16, 32, 64]))
        rs = float(rng.uniform(1.0, 4.0))
        s = config_c1(n_part=N, n_bead=M, r_s=rs)
        s.name = f"C4-{i}"
        out.append(s)
    return out


def config_c5(n_part: int = 1024, n_bead: int = 256, r_s: float = 1.0, tau: float = 0.02, k_shell: float = 153.5) -> System:
    """Combined large case: HEG 1024 electrons, r_s = 1, M = 256, |n|^2 <= 153 (4012 k)."""
    s = config_c2(n_part=n_part, n_bead=n_bead, r_s=r_s, tau=tau, k_shell=k_shell)
    s.name = "HEG-David-C5"
    return s


# ---------------------------------------------------------------------------
# pair-action table files
# ---------------------------------------------------------------------------
# Text container of named datasets ("spimc-table 1"): one header line per
# dataset `<name> <f64|i64|str> <count>` followed by its values.  Dataset names
# follow the HDF5 paths the reference loaders read.
def write_table(path: str, datasets: dict) -> None:
    with open(path, "w") as f:
        f.write("spimc-table 1\n")
        for name, val in datasets.items():
            if isinstance(val, str):
                f.write(f"{name} str 1\n{val}\n")
                continue
            a = np.atleast_1d(np.asarray(val))
            kind = "i64" if np.issubdtype(a.dtype, np.integer) else "f64"
            f.write(f"{name} {kind} {a.size}\n")
            fmt = "%d" if kind == "i64" else "%.17g"
            vals = [fmt % v for v in a.ravel()]
            for i in range(0, len(vals), 8):
                f.write(" ".join(vals[i:i + 8]) + "\n")


def bare_table_datasets(sys: System) -> dict:
    """v/diag/* of BarePairAction (bare_pair_action_class.h:39-86); v_long_k keyed
    by the distinct |k| of the system's k-set."""
    mags = kmag(sys.L, sys.k_n)
    uniq, idx = np.unique(np.round(mags, 12), return_index=True)
    return {
        "v/diag/n_r": len(sys.bare_r), "v/diag/r": sys.bare_r, "v/diag/v_r": sys.bare_v_r,
        "v/diag/n_r_long": len(sys.bare_r_long), "v/diag/r_long": sys.bare_r_long,
        "v/diag/v_long_r": sys.bare_v_long_r, "v/diag/v_long_r_0": sys.v_long_r_0,
        "v/diag/n_k": len(uniq), "v/diag/k": mags[idx], "v/diag/v_long_k": sys.v_long_k[idx],
        "v/diag/v_long_k_0": sys.v_long_k_0,
    }


def ilkka_table_datasets(sys: System) -> dict:
    """u/off_diag, du/off_diag, u/diag, du/diag and v/diag of IlkkaPairAction
    (ilkka_pair_action_class.h:262-410); off-diagonal values flat with x
    fastest; the per-k datasets keyed by the distinct |k| of the system's k-set."""
    t = sys.ilkka
    mags = kmag(sys.L, sys.k_n)
    uniq, idx = np.unique(np.round(mags, 12), return_index=True)
    out = {}
    for g, x, y, v in (("u", "x_u", "y_u", "u_xy"), ("du", "x_du", "y_du", "du_xy")):
        out[f"{g}/off_diag/n_x"] = len(t[x])
        out[f"{g}/off_diag/n_y"] = len(t[y])
        out[f"{g}/off_diag/x"] = t[x]
        out[f"{g}/off_diag/y"] = t[y]
        out[f"{g}/off_diag/{v}"] = np.ascontiguousarray(t[v]).ravel()
    for g, r, v, k, k0, r0 in (("u", "r_u_long", "u_long_r", t["u_long_k"], 0.0, 0.0),
                                ("du", "r_du_long", "du_long_r", t["du_long_k"], t["du_long_k_0"], t["du_long_r_0"])):
        out[f"{g}/diag/n_r_long"] = len(t[r])
        out[f"{g}/diag/r_long"] = t[r]
        out[f"{g}/diag/{g}_long_r"] = t[v]
        out[f"{g}/diag/{g}_long_r_0"] = r0
        out[f"{g}/diag/n_k"] = len(uniq)
        out[f"{g}/diag/k"] = mags[idx]
        out[f"{g}/diag/{g}_long_k"] = k[idx]
        out[f"{g}/diag/{g}_long_k_0"] = k0
    out.update({
        "v/diag/n_r": len(t["r_v"]), "v/diag/r": t["r_v"], "v/diag/v_r": t["v_r"],
        "v/diag/n_r_long": len(t["r_v_long"]), "v/diag/r_long": t["r_v_long"], "v/diag/v_long_r": t["v_long_r"],
        "v/diag/v_long_r_0": sys.v_long_r_0, "v/diag/n_k": len(uniq), "v/diag/k": mags[idx],
        "v/diag/v_long_k": sys.v_long_k[idx], "v/diag/v_long_k_0": sys.v_long_k_0,
    })
    return out


def david_table_datasets(L: float, alpha: float, tau: float, k_n: np.ndarray, n_grid: int = 2000, n_order: int = 2,
                         seed: int = 7, amp: float = 1e-3, r_min: float = 0.01) -> dict:
    """u_kj_<n>/*, du_kj_dbeta_<n>/*, potential/data, long_range/*, squarer/v_image of
    DavidPairAction (david_pair_action_class.h:192-360).  data arrays are the
    (n_val, n_grid, n_tau) cubes flattened with the first index fastest."""
    grid, V, u, du = david_file_arrays(L, alpha, tau, n_grid, n_order, seed, amp, r_min)
    vol = L ** 3
    mags = kmag(L, k_n)
    uniq, idx = np.unique(np.round(mags, 12), return_index=True)
    vk, vk0, vr0 = ewald_k(L, alpha, k_n)
    pre = f"u_kj_{n_order}"
    dpre = f"du_kj_dbeta_{n_order}"
    return {
        f"{pre}/grid/start": grid[0], f"{pre}/grid/end": grid[-1], f"{pre}/grid/n_grid_points": n_grid,
        f"{pre}/grid/type": "LINEAR", f"{pre}/taus": np.array([tau]),
        "potential/data": V,
        f"{pre}/data": u.T.ravel(), f"{dpre}/data": du.T.ravel(),
        "long_range/k_cut": float(mags.max()) if len(mags) else 0.0,
        "long_range/n_k": len(uniq) + 1,
        "long_range/k_points": np.concatenate([[0.0], mags[idx]]),
        "long_range/u_k": np.concatenate([[vk0 * vol], vk[idx] * vol]),
        "squarer/v_image": vr0,
    }
