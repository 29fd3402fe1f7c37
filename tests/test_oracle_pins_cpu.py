"""CPU pins of oracle functions that no golden vector covers: each is checked
against a closed form, a library routine or an independent brute force, chosen
so that a plausible slip (a dropped term, a wrong sign, link or window, a
missing normalisation) fails.

* Kinetic::VirialEnergy (kinetic_class.h:47-102): a circular ring path, whose
  window sums have a closed form in cos(2 pi W / M); the whole-path window
  closes exactly on any non-winding path.
* DavidPairAction::CalcdUdBeta and CalcV (david_pair_action_class.h:26-38,
  125-168): the q, s expansion of the du/dbeta table evaluated with scipy's
  natural splines, with the reference's clamps.
* StructureFactor::Accumulate (structure_factor_class.h:15-32): a BCC lattice,
  whose |rho_k|^2 is N^2 on its reciprocal lattice and 0 elsewhere, and a
  brute-force numpy rho_k with the strict k cutoff.
* PairAction::GetActionGradient with CalcGradientU (pair_action_class.h:
  135-155, 305-337): central differences (eps 1e-4) of the link action written
  in numpy (Bare Coulomb) or of the scipy-pinned link action (David), over the
  links (b, b+1) and (b, b-1) read from bead b.
* Energy::VirialEstimator (energy_class.h:35-116): its assembly from the
  pinned parts, the centroid vector Dr(r_pb, r_p0) and 1/(2 tau).
* DisplaceParticle::Attempt and Bisect::Attempt acceptance
  (displace_particle_class.h:28-72, bisect_class.h:39-124): with free-particle
  sampling the bisection's sample ratio cancels the kinetic action change
  level by level, so both moves accept exactly when log u <= -(S_pair,new -
  S_pair,old); S_pair is written in numpy (Bare Coulomb), and replayed uniforms
  just above and below exp(-dS) must flip the decision.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle_ffi import Oracle
from simpimc_amd import abi, system
from test_oracle_cpu import _ref_link_geometry, minimg


# ---------------------------------------------------------------------------
# Kinetic::VirialEnergy
# ---------------------------------------------------------------------------
def _ring_path(s, R, centres):
    M = s.n_bead
    th = 2 * math.pi * np.arange(M) / M
    pos = np.zeros((3, M, s.n_part))
    for p, c in enumerate(centres):
        pos[0, :, p] = c[0] + R * np.cos(th)
        pos[1, :, p] = c[1] + R * np.sin(th)
        pos[2, :, p] = c[2]
    return pos


@pytest.mark.parametrize("W", [1, 2, 3, 5, 12, 16])
def test_kinetic_virial_energy_on_a_ring(W):
    """r_b = c + R (cos b theta, sin b theta, 0), theta = 2 pi / M: the window
    vector r_{b+W} - r_b dotted with the link r_{b+1} - r_b is
    R^2 [1 + cos((W-1) theta) - cos(W theta) - cos(theta)] for every b, so
    VirialEnergy(W) = d N M / (2 tau W) - N M R^2 [...] / (4 lambda tau^2 W)."""
    s = system.config_c1(n_part=3, n_bead=16, n_level=2)
    R = 0.4
    o = Oracle(s)
    o.upload(_ring_path(s, R, [(0.1, 0.2, 0.3), (-2.0, 1.0, 0.5), (3.0, -3.0, -1.0)]))
    th = 2 * math.pi / s.n_bead
    N, M, tau, lam = s.n_part, s.n_bead, s.tau, s.lam
    bracket = 1 + math.cos((W - 1) * th) - math.cos(W * th) - math.cos(th)
    want = 3 * N * M / (2 * tau * W) - N * M * R * R * bracket / (4 * lam * tau * tau * W)
    got = o.kinetic_virial_energy(W)
    assert abs(got - want) < 1e-11 * abs(want), (got, want)


def test_kinetic_virial_energy_whole_window_closes():
    """With W = M the window vector of every bead is the sum of all M links of a
    non-winding path, zero: VirialEnergy(M) = d N / (2 tau) whatever the beads."""
    s = system.config_c1(n_part=5, n_bead=12, n_level=2)
    o = Oracle(s)
    o.upload(system.random_path(s, seed=3, spread=0.3))
    want = 3 * s.n_part / (2 * s.tau)
    assert abs(o.kinetic_virial_energy(s.n_bead) - want) < 1e-12 * want
    # and the window is not the whole path for W < M: the link terms contribute
    assert abs(o.kinetic_virial_energy(4) - 3 * s.n_part * s.n_bead / (2 * s.tau * 4)) > 1e-3


# ---------------------------------------------------------------------------
# DavidPairAction::CalcdUdBeta, CalcV
# ---------------------------------------------------------------------------
def _david_expansion(spl, grid, r, rp, s, diag_extra=None):
    q, z = 0.5 * (r + rp), r - rp
    rc, rpc = np.clip(r, grid[0], grid[-1]), np.clip(rp, grid[0], grid[-1])
    u = 0.5 * (spl[1](rc) + spl[1](rpc))
    if diag_extra is not None:
        u += 0.5 * (diag_extra(rc) + diag_extra(rpc))
    if s > 0 and q < grid[-1]:
        for k in range(1, 3):
            for j in range(k + 1):
                u += spl[k * (k + 1) // 2 + j + 1](q) * z ** (2 * j) * s ** (2 * (k - j))
    return float(u)


def test_david_du_dbeta_and_potential_match_scipy_splines():
    """CalcdUdBeta = (du_0 + du_1)(r, r') / 2 + sum_kj du_kj(q) z^2j s^2(k-j), where
    du_0 is the potential the table's diagonal compensates (:143); CalcV is the
    u_kj table's component 0 averaged over r, r'.  r, r' clamp to the grid, q and
    z come from the unclamped lengths, the expansion needs s > 0 and q < r_max."""
    s = system.config_c2(n_part=2, n_bead=4, n_level=1, k_shell=2.5, n_grid=300, seed=5)
    s.use_long_range = False
    o = Oracle(s)
    g = s.david_grid
    du = [CubicSpline(g, s.david_du_kj[c], bc_type="natural") for c in range(s.david_du_kj.shape[0])]
    u = [CubicSpline(g, s.david_u_kj[c], bc_type="natural") for c in range(s.david_u_kj.shape[0])]
    rng = np.random.default_rng(8)
    n_off = 0
    for i in range(60):
        r0, r1 = rng.uniform(0.0, 1.08 * g[-1], 2)
        sd = 0.0 if i % 10 == 0 else rng.uniform(0.0, 1.0) * abs(r0 - r1) + rng.uniform(0.0, 0.4)
        want_du = _david_expansion(du, g, r0, r1, sd, diag_extra=du[0])
        got_du = o.pair_link(1, r0, r1, sd)
        assert got_du == pytest.approx(want_du, rel=1e-11, abs=1e-12), (r0, r1, sd)
        want_v = 0.5 * (u[0](np.clip(r0, g[0], g[-1])) + u[0](np.clip(r1, g[0], g[-1])))
        assert o.pair_link(2, r0, r1, sd) == pytest.approx(float(want_v), rel=1e-11, abs=1e-12)
        if sd > 0 and 0.5 * (r0 + r1) < g[-1]:
            n_off += 1
    assert n_off > 30  # the off-diagonal expansion was exercised


# ---------------------------------------------------------------------------
# StructureFactor::Accumulate
# ---------------------------------------------------------------------------
def test_structure_factor_bcc_reciprocal_lattice():
    """N = 2 n^3 particles on the reference's BCC lattice (species_class.h:
    275-300), every slice alike: rho_k is N on the BCC reciprocal lattice
    (every n_d a multiple of n, and sum n_d / n even) and 0 elsewhere, so each
    accumulation adds M N^2 there and nothing elsewhere."""
    N, nc = 16, 2
    L = system.heg_box_length(N, 1.0)
    s = system.System(n_part=N, n_bead=3, beta=0.06, L=L, bisect_n_level=1)
    s.k_n = system.kspace(L, math.sqrt(20.5) * 2 * math.pi / L)
    system.bare_coulomb_ewald(s, alpha=5.0 / L)
    o = Oracle(s)
    o.upload(system.bcc_path(s))
    o.obs_configure(sk_k_cut=1e300)
    o.obs_accumulate(abi.OBS_SK)
    o.obs_accumulate(abi.OBS_SK)
    got, nm = o.obs_read(abi.OBS_SK)
    n = s.k_n.astype(np.int64)
    on = np.all(n % nc == 0, axis=1) & ((n.sum(axis=1) // nc) % 2 == 0)
    assert nm == 2 and on.sum() > 3
    np.testing.assert_allclose(got[on], 2 * s.n_bead * N * N, rtol=1e-12)
    assert np.max(np.abs(got[~on])) < 1e-20 * N * N * s.n_bead + 1e-18


def test_structure_factor_brute_force_and_cutoff():
    """Random beads: sk(k) = sum_b |sum_p e^{i k.r_pb}|^2 for |k| < k_cut (strict),
    0 above; the k_cut sits exactly on a shell's magnitude, which is excluded."""
    s = system.config_c1(n_part=7, n_bead=5, n_level=1)
    pos = system.random_path(s, seed=12, spread=0.4)
    o = Oracle(s)
    o.upload(pos)
    k = s.k_n.astype(np.float64) * (2 * math.pi / s.L)
    mags = np.linalg.norm(k, axis=1)
    kc = float(np.sort(np.unique(np.round(mags, 12)))[10])
    kc = float(mags[np.argmin(np.abs(mags - kc))])  # the exact stored magnitude of that shell
    o.obs_configure(sk_k_cut=kc)
    o.obs_accumulate(abi.OBS_SK)
    got, _ = o.obs_read(abi.OBS_SK)
    rho = np.exp(1j * np.einsum("kd,dbp->bkp", k, pos)).sum(axis=2)
    want = np.where(mags < kc, (np.abs(rho) ** 2).sum(axis=0), 0.0)
    assert (mags == kc).sum() > 0 and np.all(got[mags >= kc] == 0)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# GetActionGradient (compat) and the virial estimator
# ---------------------------------------------------------------------------
def _bare_link_u(s, pos, b, b1, p, q):
    """Bare Coulomb CalcU at level 0 without long range: tau (1/(2r) + 1/(2r')),
    r and r' clamped to the v_r grid (bare_pair_action_class.h:99-123,146-150)."""
    r, rp, _ = _ref_link_geometry(pos, s.L, b, b1, p, q)
    r, rp = np.clip([r, rp], s.bare_r[0], s.bare_r[-1])
    return s.tau * (0.5 / r + 0.5 / rp)


def _compat_gradient(s, pos, link_u, eps=1e-4):
    """Sum over partners q of central differences (eps) of the link actions
    (b, b+1) and (b, b-1), both read from bead b, in the moved particle's
    coordinate d."""
    M, N = s.n_bead, s.n_part
    g = np.zeros((M, N, 3))
    for b in range(M):
        for p in range(N):
            for d in range(3):
                f = []
                for sg in (1, -1):
                    x = pos.copy()
                    x[d, b, p] += sg * eps
                    f.append(sum(link_u(x, b, (b + 1) % M, p, q) + link_u(x, b, (b - 1) % M, p, q)
                                 for q in range(N) if q != p))
                g[b, p, d] = (f[0] - f[1]) / (2 * eps)
    return g


def test_compat_gradient_bare_is_central_difference_of_link_action():
    s = system.config_c1(n_part=4, n_bead=6, n_level=1)
    s.use_long_range = False
    pos = system.random_path(s, seed=5, spread=0.3)
    o = Oracle(s)
    o.upload(pos)
    got = o.gradient(abi.GRAD_PAIR_SHORT, compat=True)
    want = _compat_gradient(s, pos, lambda x, b, b1, p, q: _bare_link_u(s, x, b, b1, p, q))
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())
    # the long-range term of Bare is zero in the reference (pair_action_class.h:158-165)
    s.use_long_range = True
    o2 = Oracle(s)
    o2.upload(pos)
    assert np.abs(o2.gradient(abi.GRAD_PAIR_LONG, compat=True)).max() == 0.0


def test_compat_gradient_david_is_central_difference_of_link_action():
    s = system.config_c2(n_part=3, n_bead=4, n_level=1, k_shell=2.5, n_grid=300, seed=5)
    s.use_long_range = False
    pos = system.random_path(s, seed=6, spread=0.3, cofactor=0.4)
    o = Oracle(s)
    o.upload(pos)

    def link_u(x, b, b1, p, q):
        r, rp, sd = _ref_link_geometry(x, s.L, b, b1, p, q)
        return o.pair_link(0, r, rp, sd)

    got = o.gradient(abi.GRAD_PAIR_SHORT, compat=True)
    want = _compat_gradient(s, pos, link_u)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


@pytest.mark.parametrize("W", [0, 3])
def test_virial_estimator_assembly(W):
    """Energy::VirialEstimator: kinetic part = VirialEnergy(W) + sum_{p,b}
    GetActionGradient(b, b+1, {p}) . Dr(r_pb, r_p0) / (2 tau) (the reference's
    centroid vector is the bead minus the particle's slice-0 bead, minimum
    image), interaction part = the pair action's DActionDBeta; W = 0 is M."""
    s = system.config_c1(n_part=4, n_bead=6, n_level=1)
    pos = system.random_path(s, seed=7, spread=0.3)
    o = Oracle(s)
    o.upload(pos)
    g = o.gradient(abi.GRAD_PAIR_SHORT | abi.GRAD_PAIR_LONG, compat=True)
    dr = minimg(pos - pos[:, :1, :], s.L)  # [3][M][N]
    centroid = float(np.einsum("bpd,dbp->", g, dr))
    want_kin = o.kinetic_virial_energy(W) + centroid / (2 * s.tau)
    want_pair = o.d_action_d_beta(abi.ACTION_PAIR)
    o.obs_configure(energy=abi.ENERGY_VIRIAL_COMPAT, virial_window=W)
    o.obs_accumulate(abi.OBS_ENERGY)
    got, nm = o.obs_read(abi.OBS_ENERGY)
    assert nm == 1
    assert abs(centroid) / (2 * s.tau) > 1e-9 * abs(want_kin) * 1e4  # far above the 1e-12 tolerance
    assert got[0] == pytest.approx(want_kin, rel=1e-12)
    assert got[1] == pytest.approx(want_pair, rel=1e-12)


# ---------------------------------------------------------------------------
# Metropolis acceptance of DisplaceParticle and Bisect
# ---------------------------------------------------------------------------
def _coulomb_system(n_part=3, n_bead=8, n_level=2):
    s = system.config_c1(n_part=n_part, n_bead=n_bead, r_s=1.0, n_level=n_level)
    s.use_long_range = False
    return s


def _pair_action_of(s, pos, p):
    """Short-range pair action of particle p against all others, all links (b, b+1)."""
    M = s.n_bead
    return sum(_bare_link_u(s, pos, b, (b + 1) % M, p, q) for b in range(M) for q in range(s.n_part) if q != p)


def _decides(s, pos, tape, kind):
    o = Oracle(s)
    o.upload(pos)
    o.tape_set(tape)
    acc = o.displace(1) if kind == "displace" else o.bisect(1)
    assert o.tape_cursor() == len(tape)
    return bool(acc[0]), o.download()


def test_displace_acceptance_threshold():
    s = _coulomb_system()
    pos = system.random_path(s, seed=4, spread=0.1, cofactor=0.5)
    rng = np.random.default_rng(2)
    flips = 0
    for _ in range(6):
        p = int(rng.integers(s.n_part))
        d = rng.uniform(-0.5, 0.5, 3) * s.displace_step
        ok, new = _decides(s, pos, [p, *d, 1e-300], "displace")
        assert ok
        # the move translates every bead of p, unboxed
        assert np.array_equal(new[:, :, p], pos[:, :, p] + d[:, None])
        dS = _pair_action_of(s, new, p) - _pair_action_of(s, pos, p)
        thr = math.exp(-dS)
        if thr < 1 - 1e-6:
            assert _decides(s, pos, [p, *d, thr * (1 - 1e-9)], "displace")[0]
            assert not _decides(s, pos, [p, *d, thr * (1 + 1e-9)], "displace")[0]
            flips += 1
        else:
            assert _decides(s, pos, [p, *d, 1 - 1e-9], "displace")[0]
    assert flips > 0


@pytest.mark.parametrize("n_level", [1, 2, 3])
def test_bisect_acceptance_threshold(n_level):
    """Free-particle bisection (Levy) sampling: at every coarse level the log
    sample ratio cancels the kinetic change, and at level 0 what is left is
    -(pair action change) over the window."""
    s = _coulomb_system(n_bead=16, n_level=n_level)
    pos = system.random_path(s, seed=9, spread=0.15, cofactor=0.4)
    rng = np.random.default_rng(3)
    nb = 1 << n_level
    flips = 0
    for _ in range(5):
        p, b0 = int(rng.integers(s.n_part)), int(rng.integers(s.n_bead))
        levels = []
        for lvl in range(n_level - 1, -1, -1):
            levels.append(rng.standard_normal(3 * (nb >> (lvl + 1))).tolist())

        def tape(u0):
            t = [p, b0]
            for i, z in enumerate(levels):
                t += z + [u0 if i == len(levels) - 1 else 1 - 1e-9]
            return t

        ok, new = _decides(s, pos, tape(1e-300), "bisect")
        assert ok
        moved = [(b0 + j) % s.n_bead for j in range(1, nb)]
        assert not np.array_equal(new[:, moved, p], pos[:, moved, p])
        dS = _pair_action_of(s, new, p) - _pair_action_of(s, pos, p)
        thr = math.exp(-dS)
        if thr < 1 - 1e-6:
            assert _decides(s, pos, tape(thr * (1 - 1e-8)), "bisect")[0]
            assert not _decides(s, pos, tape(thr * (1 + 1e-8)), "bisect")[0]
            flips += 1
        else:
            assert _decides(s, pos, tape(1 - 1e-8), "bisect")[0]
    assert flips > 0


def test_free_particle_bisection_always_accepts():
    """Without a pair action the bisection samples the free-particle density
    exactly: every attempt accepts, at every level, for any uniform below 1."""
    s = system.config_c1(n_part=4, n_bead=16, n_level=3)
    s.pair_type = abi.PAIR_NONE
    s.use_long_range = False
    pos = system.random_path(s, seed=2, spread=0.2)
    o = Oracle(s)
    o.upload(pos)
    rng = np.random.default_rng(4)
    tape = []
    for _ in range(40):
        tape += [int(rng.integers(4)), int(rng.integers(16))]
        for lvl in range(2, -1, -1):
            tape += rng.standard_normal(3 * (8 >> (lvl + 1))).tolist() + [1 - 1e-9]
    o.tape_set(tape)
    assert o.bisect(40).all()


# ---------------------------------------------------------------------------
# GetActionLaplacian (action_class.h:49)
# ---------------------------------------------------------------------------
def _laplacian_numpy(s, pos, link_u, b0, b1, p, eps=1e-4):
    """Second central differences (eps) in the coordinates of bead (p, b) of the
    link actions (b, b+1) and (b, b-1) summed over the partners q, for the window
    beads b0 <= b < b1 (level 0)."""
    M, N = s.n_bead, s.n_part
    tot = 0.0
    for b in range(b0, b1):
        for q in range(N):
            if q == p:
                continue
            for b1_ in ((b + 1) % M, (b - 1) % M):
                f0 = link_u(pos, b % M, b1_, p, q)
                for d in range(3):
                    x = pos.copy()
                    x[d, b % M, p] += eps
                    fp = link_u(x, b % M, b1_, p, q)
                    x[d, b % M, p] -= 2 * eps
                    fm = link_u(x, b % M, b1_, p, q)
                    tot += (fp + fm - 2 * f0) / (eps * eps)
    return tot


def test_laplacian_of_the_bare_action_is_the_ewald_short_range_laplacian():
    """Bare with the Ewald split: V = 1/r - erf(alpha r)/r, and away from r = 0
    laplacian(erfc(alpha r)/r) = (4 alpha^3 / sqrt(pi)) exp(-alpha^2 r^2); the two
    links of a bead each carry tau V(r) / 2 of its slice, so the pair Laplacian of
    particle p over a window is tau (4 alpha^3 / sqrt(pi)) sum_b sum_q
    exp(-alpha^2 r_pq(b)^2).  Without the long-range split (V = 1/r) it vanishes."""
    s = system.config_c1(n_part=8, n_bead=8, n_level=2)
    alpha = 5.0 / s.L
    pos = system.random_path(s, seed=21, spread=0.2)
    o = Oracle(s)
    o.upload(pos)
    p, b0, b1 = 3, 2, 6
    want = 0.0
    for b in range(b0, b1):
        d = minimg(pos[:, b, :] - pos[:, b, p:p + 1], s.L)
        r2 = np.sum(d * d, axis=0)
        r2 = np.delete(r2, p)
        assert np.all(np.sqrt(r2) > 0.2)
        want += float(np.sum(np.exp(-alpha * alpha * r2)))
    want *= s.tau * 4 * alpha ** 3 / math.sqrt(math.pi)
    got = o.get_action_laplacian(abi.ACTION_PAIR, abi.MODE_NEW, b0, b1, [p], 0)
    assert got == pytest.approx(want, rel=2e-6), (got, want)
    s.use_long_range = False
    o2 = Oracle(s)
    o2.upload(pos)
    assert abs(o2.get_action_laplacian(abi.ACTION_PAIR, abi.MODE_NEW, b0, b1, [p], 0)) < 1e-5 * want
    # above max_level (0) the pair action has no Laplacian
    assert o.get_action_laplacian(abi.ACTION_PAIR, abi.MODE_NEW, b0, b1, [p], 1) == 0.0


def test_laplacian_of_the_david_action_is_a_second_difference_of_the_link_action():
    s = system.config_c2(n_part=3, n_bead=4, n_level=1, k_shell=2.5, n_grid=300, seed=5)
    pos = system.random_path(s, seed=6, spread=0.3, cofactor=0.4)
    o = Oracle(s)
    o.upload(pos)

    def link_u(x, b, b1, p, q):
        r, rp, sd = _ref_link_geometry(x, s.L, b, b1, p, q)
        return o.pair_link(0, r, rp, sd)

    for p, b0, b1 in ((0, 0, 4), (2, 1, 3), (1, 3, 5)):
        got = o.get_action_laplacian(abi.ACTION_PAIR, abi.MODE_NEW, b0, b1, [p], 0)
        want = _laplacian_numpy(s, pos, link_u, b0, b1, p)
        assert got == pytest.approx(want, rel=1e-7, abs=1e-9), (p, b0, b1, got, want)


@pytest.mark.parametrize("level", [0, 1, 2])
def test_kinetic_laplacian_counts_the_walk_steps(level):
    """Kinetic::GetActionLaplacian: n_d * 4 / (4 lambda tau 2^level) per step of the
    walk; a whole-path window of a particle whose path closes on itself has no step."""
    s = system.System(n_part=4, n_bead=8, beta=2.0, L=4.0, lam=0.5, bisect_n_level=2)
    pos = system.random_path(s, seed=5, spread=0.4)
    o = Oracle(s)
    o.upload(pos)
    skip = 1 << level
    step = 3 * 4.0 * (1.0 / (4.0 * s.lam * s.tau)) / skip
    for b0, span in ((1, 4), (6, 4), (0, 8)):
        want = 0.0
        for _ in range(span // skip):
            want += step
        got = o.get_action_laplacian(abi.ACTION_KINETIC, abi.MODE_NEW, b0, b0 + span, [2], level)
        if span == s.n_bead:
            want = 0.0  # the walk returns to its first bead at once
        assert got == want, (b0, span, got, want)
    o.set_perm(np.array([1, 0, 2, 3], dtype=np.uint32))
    got = o.get_action_laplacian(abi.ACTION_KINETIC, abi.MODE_NEW, 0, 8, [0], level)
    assert got == pytest.approx((8 // skip) * step, rel=1e-15)
