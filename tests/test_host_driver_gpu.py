"""The C++ XML driver (simpimc_amd/spimc) on a simpimc input file gives the same
block averages as the same run driven through the Python binding: both reach
the device through the C ABI with the same seed, so the chains are identical."""
import json
import os
import subprocess

import numpy as np
import pytest

import h5dump
from simpimc_amd import abi, system

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "simpimc_amd", "spimc")

INIT_FILE = 'init_type="File" file_name="init.xyz"'
PATH_DUMP = '\n    <Observable name="path_dump" type="PathDump" />'

BARE = ('<Action name="Coulomb" type="BarePairAction" file="{file}" species_a="e" species_b="e" is_coulomb="1" '
        'max_level="0" use_long_range="1" />')
ILKKA = ('<Action name="Ilkka" type="IlkkaPairAction" file="{file}" species_a="e" species_b="e" max_level="0" '
         'use_long_range="1" />')
DAVID = ('<Action name="David" type="DavidPairAction" file="{file}" species_a="e" species_b="e" n_order="2" '
         'max_level="0" use_long_range="1" />')

XML = """<Input>
  <RNG seed="{seed}" />
  <IO output_prefix="{prefix}" />
  <System n_d="3" n_bead="{M}" beta="{beta!r}" L="{L!r}" pbc="1" k_cut="{kc!r}" />
  <Particles>
    <Species name="e" n_part="{N}" lambda="0.5" {init} />
  </Particles>
  <Actions>
    <Action name="Kinetic" type="Kinetic" species="e" />
    {action}
  </Actions>
  <Moves>
    <Move name="Bisect" type="Bisect" species="e" n_level="3" />
    <Move name="Displace" type="DisplaceParticle" species="e" />
  </Moves>
  <Observables>
    <Observable name="Energy" type="Energy" />
    <Observable name="SK" type="StructureFactor" species_a="e" species_b="e" />
    <Observable name="GR" type="PairCorrelation" species_a="e" species_b="e" n_r="60" />{extra}
  </Observables>
  <Algorithm>
    <Loop n_step="2">
      <Loop n_step="3">
        <Loop n_step="{nb}"><Move name="Bisect" /></Loop>
        <Loop n_step="{N}"><Move name="Displace" /></Loop>
        <Observable name="Energy" />
        <Observable name="SK" />
        <Observable name="GR" />
      </Loop>
      <Write />
    </Loop>
  </Algorithm>
</Input>
"""


def test_xml_driver_matches_binding(tmp_path):
    assert os.path.exists(DRIVER), "build the host driver first"
    s = system.config_c1()
    seed = 4242
    pos = system.random_path(s, seed=8)
    with open(tmp_path / "init.xyz", "w") as f:
        for p in range(s.n_part):
            f.write("%.17g %.17g %.17g\n" % tuple(pos[:, 0, p]))
    system.write_table(str(tmp_path / "bare.table"), system.bare_table_datasets(s))
    kc = (5.0 + 1e-9) * 2 * np.pi / s.L
    nb = (s.n_part * s.n_bead) >> s.bisect_n_level
    (tmp_path / "in.xml").write_text(XML.format(seed=seed, M=s.n_bead, beta=s.beta, L=s.L, kc=kc, N=s.n_part, nb=nb,
                                                prefix="run", action=BARE.format(file="bare.table"), init=INIT_FILE,
                                                extra=""))
    res = subprocess.run([DRIVER, "in.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    blocks = [json.loads(line) for line in open(tmp_path / "run.spimc.jsonl")]
    assert len(blocks) == 2

    with abi.Engine(s, seed=seed) as e:
        e.upload(pos)
        e.obs_configure(energy=abi.ENERGY_THERMAL, sk_k_cut=kc, gr_n_r=60)
        for blk in blocks:
            for _ in range(3):
                e.bisect(nb)
                e.displace(s.n_part)
                e.obs_accumulate(abi.OBS_ENERGY | abi.OBS_SK | abi.OBS_GR)
            en, nm = e.obs_read(abi.OBS_ENERGY)
            sk, _ = e.obs_read(abi.OBS_SK)
            e.obs_reset(abi.OBS_ENERGY | abi.OBS_SK | abi.OBS_GR)
            got = blk["Observables"]["Energy"][0]
            assert got["n_measure"] == nm == 3
            assert got["Kinetic"] == pytest.approx(en[0] / (s.n_bead * nm), rel=1e-14)
            assert got["Coulomb"] == pytest.approx(en[1] / (s.n_bead * nm), rel=1e-14)
            y = np.array(blk["Observables"]["SK"]["y"][0])
            np.testing.assert_allclose(y, sk / (nm * s.n_bead * s.n_part ** 2), rtol=1e-14)
            mv = blk["Moves"]["Bisect"][0]
            assert mv["n_attempt"] == 3 * nb


def run_driver(tmp_path, s, action, prefix, replicas, seed=4242, init=None, extra=""):
    kc = (5.0 + 1e-9) * 2 * np.pi / s.L
    nb = (s.n_part * s.n_bead) >> s.bisect_n_level
    xml = tmp_path / f"{prefix}.xml"
    xml.write_text(XML.format(seed=seed, M=s.n_bead, beta=s.beta, L=s.L, kc=kc, N=s.n_part, nb=nb, prefix=prefix,
                              action=action, init=init or INIT_FILE, extra=extra))
    print(f"driver: {prefix} ({replicas} replicas)", flush=True)
    res = subprocess.run([DRIVER, xml.name, "--replicas", str(replicas)], cwd=tmp_path, capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    return [json.loads(line) for line in open(tmp_path / f"{prefix}.spimc.jsonl")]


def convert(tmp_path, src, dst, *extra):
    res = subprocess.run([DRIVER, "--table-to-h5", src, dst, *extra], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr


def write_init(tmp_path, s):
    pos = system.random_path(s, seed=8)
    with open(tmp_path / "init.xyz", "w") as f:
        for p in range(s.n_part):
            f.write("%.17g %.17g %.17g\n" % tuple(pos[:, 0, p]))


@pytest.mark.skipif(h5dump.H5DUMP is None, reason="h5dump not installed")
def test_hdf5_table_and_output(tmp_path):
    """The Bare table as HDF5 gives the run of the text table, and the per-replica
    HDF5 files hold the blocks of the JSON-lines output in the reference's layout."""
    assert os.path.exists(DRIVER), "build the host driver first"
    s = system.config_c1()
    write_init(tmp_path, s)
    system.write_table(str(tmp_path / "bare.table"), system.bare_table_datasets(s))
    convert(tmp_path, "bare.table", "bare.h5")
    R = 2
    txt = run_driver(tmp_path, s, BARE.format(file="bare.table"), "txt", R)
    h5 = run_driver(tmp_path, s, BARE.format(file="bare.h5"), "h5", R)
    assert txt == h5
    for r in range(R):
        f = str(tmp_path / f"h5.{r}.h5")
        ds = lambda name: h5dump.dataset(f, name)[1]  # noqa: E731
        assert ds("/Actions/Coulomb/type") == ["BarePairAction"]
        assert ds("/Actions/Coulomb/file") == ["bare.h5"]
        assert ds("/RNG/seed")[0] == 4242
        assert ds("/Input/file_name") == ["h5.xml"]
        for mv in ("Bisect", "Displace"):
            assert ds(f"/Moves/{mv}/data_type") == ["scalar"]
            att = ds(f"/Moves/{mv}/n_attempt")
            acc = ds(f"/Moves/{mv}/n_accept")
            assert list(att) == [b["Moves"][mv][r]["n_attempt"] for b in h5]
            assert list(acc) == [b["Moves"][mv][r]["n_accept"] for b in h5]
            np.testing.assert_array_equal(ds(f"/Moves/{mv}/x"), acc / att)
        assert ds("/Observables/Energy/type") == ["Energy"]
        for key in ("total", "Kinetic", "Coulomb"):
            assert ds(f"/Observables/Energy/{key}/data_type") == ["scalar"]
            np.testing.assert_allclose(ds(f"/Observables/Energy/{key}/x"),
                                       [b["Observables"]["Energy"][r][key] for b in h5], rtol=1e-14)
        for ob in ("SK", "GR"):
            assert ds(f"/Observables/{ob}/data_type") == ["histogram"]
            assert ds(f"/Observables/{ob}/species_a") == ["e"]
            dims, y = h5dump.dataset(f, f"/Observables/{ob}/y")
            want = np.array([b["Observables"][ob]["y"][r] for b in h5])
            assert dims == want.shape
            np.testing.assert_array_equal(y.reshape(dims), want)
            np.testing.assert_array_equal(ds(f"/Observables/{ob}/x"), h5[0]["Observables"][ob]["x"])
        assert ds("/Observables/GR/n_r")[0] == 60


@pytest.mark.skipif(h5dump.H5DUMP is None, reason="h5dump not installed")
def test_david_hdf5_table_matches_text(tmp_path):
    """A David table whose cubes hold three taus, as text and as HDF5 with the
    transposed cube layout, drives identical runs."""
    from test_tables_h5_cpu import david_three_taus

    assert os.path.exists(DRIVER), "build the host driver first"
    s = system.config_c1()
    write_init(tmp_path, s)
    d, _, _ = david_three_taus(s)
    system.write_table(str(tmp_path / "david.table"), d)
    convert(tmp_path, "david.table", "david.h5", "--transposed")
    txt = run_driver(tmp_path, s, DAVID.format(file="david.table"), "txt", 1)
    h5 = run_driver(tmp_path, s, DAVID.format(file="david.h5"), "h5", 1)
    assert txt == h5
    e = txt[-1]["Observables"]["Energy"][0]
    assert np.isfinite(e["total"]) and e["n_measure"] == 3


@pytest.mark.skipif(h5dump.H5DUMP is None, reason="h5dump not installed")
def test_path_dump_and_restart(tmp_path):
    """PathDump appends each replica's path to its output file at every Write; the
    last dump equals the chain the binding reaches with the same seed; a run with
    init_type Restart (parallel: replica r from file r, else every replica from
    file 0) starts from that dump."""
    assert os.path.exists(DRIVER), "build the host driver first"
    s = system.config_c1()
    write_init(tmp_path, s)
    system.write_table(str(tmp_path / "bare.table"), system.bare_table_datasets(s))
    R = 2
    blocks = run_driver(tmp_path, s, BARE.format(file="bare.table"), "A", R, extra=PATH_DUMP)
    assert [b["Observables"]["path_dump"]["n_dump"] for b in blocks] == [1, 2]
    N, M = s.n_part, s.n_bead
    last = []
    for r in range(R):
        f = str(tmp_path / f"A.{r}.h5")
        assert h5dump.dataset(f, "/Observables/path_dump/e/n_dump")[1][0] == 2
        dims, pos = h5dump.dataset(f, "/Observables/path_dump/e/positions")
        assert dims == (2, N, M, 3)
        pdims, perm = h5dump.dataset(f, "/Observables/path_dump/e/permutation")
        assert pdims == (2, N, 2)
        assert np.array_equal(perm.reshape(pdims)[-1], np.repeat(np.arange(N), 2).reshape(N, 2))
        last.append(pos.reshape(dims)[-1].transpose(2, 1, 0))  # -> [3][M][N]
    # replica 0 of the driver is the binding's chain with the same seed
    pos0 = system.random_path(s, seed=8)
    nb = (N * M) >> s.bisect_n_level
    kc = (5.0 + 1e-9) * 2 * np.pi / s.L
    with abi.Engine([s] * R, seed=4242) as e:
        for r in range(R):
            e.upload(pos0, r=r)
        e.obs_configure(energy=abi.ENERGY_THERMAL, sk_k_cut=kc, gr_n_r=60)
        for _ in range(2 * 3):
            e.bisect(nb)
            e.displace(N)
            e.obs_accumulate(abi.OBS_ENERGY | abi.OBS_SK | abi.OBS_GR)
        for r in range(R):
            assert np.array_equal(e.download(r=r), last[r]), r
    for parallel in (1, 0):
        prefix = f"B{parallel}"
        init = f'init_type="Restart" prefix="A" parallel="{parallel}"'
        run_driver(tmp_path, s, BARE.format(file="bare.table"), prefix, R, init=init, extra=PATH_DUMP)
        # the restarted chain equals the binding's chain started from the dump
        with abi.Engine([s] * R, seed=4242) as e:
            for r in range(R):
                e.upload(last[r if parallel else 0], r=r)
            e.obs_configure(energy=abi.ENERGY_THERMAL, sk_k_cut=kc, gr_n_r=60)
            for _ in range(2 * 3):
                e.bisect(nb)
                e.displace(N)
                e.obs_accumulate(abi.OBS_ENERGY | abi.OBS_SK | abi.OBS_GR)
            for r in range(R):
                dims, pos = h5dump.dataset(str(tmp_path / f"{prefix}.{r}.h5"), "/Observables/path_dump/e/positions")
                assert np.array_equal(e.download(r=r), pos.reshape(dims)[-1].transpose(2, 1, 0)), (parallel, r)


@pytest.mark.skipif(h5dump.H5DUMP is None, reason="h5dump not installed")
def test_ilkka_driver_matches_binding_and_hdf5_table(tmp_path):
    """The IlkkaPairAction loader: the driver on the Ilkka tables (text, and HDF5
    with transposed matrices) gives the chain the binding gives on the same system."""
    assert os.path.exists(DRIVER), "build the host driver first"
    s = system.config_ilkka()
    write_init(tmp_path, s)
    system.write_table(str(tmp_path / "ilkka.table"), system.ilkka_table_datasets(s))
    convert(tmp_path, "ilkka.table", "ilkka.h5", "--transposed")
    txt = run_driver(tmp_path, s, ILKKA.format(file="ilkka.table"), "txt", 1)
    h5 = run_driver(tmp_path, s, ILKKA.format(file="ilkka.h5"), "h5", 1)
    assert txt == h5
    pos = system.random_path(s, seed=8)
    nb = (s.n_part * s.n_bead) >> s.bisect_n_level
    kc = (5.0 + 1e-9) * 2 * np.pi / s.L
    with abi.Engine(s, seed=4242) as e:
        e.upload(pos)
        e.obs_configure(energy=abi.ENERGY_THERMAL, sk_k_cut=kc, gr_n_r=60)
        for blk in txt:
            for _ in range(3):
                e.bisect(nb)
                e.displace(s.n_part)
                e.obs_accumulate(abi.OBS_ENERGY | abi.OBS_SK | abi.OBS_GR)
            en, nm = e.obs_read(abi.OBS_ENERGY)
            e.obs_reset(abi.OBS_ENERGY | abi.OBS_SK | abi.OBS_GR)
            got = blk["Observables"]["Energy"][0]
            assert got["n_measure"] == nm == 3
            assert got["Kinetic"] == pytest.approx(en[0] / (s.n_bead * nm), rel=1e-14)
            assert got["Ilkka"] == pytest.approx(en[1] / (s.n_bead * nm), rel=1e-14)
            assert blk["Moves"]["Bisect"][0]["n_attempt"] == 3 * nb


@pytest.mark.skipif(h5dump.H5DUMP is None, reason="h5dump not installed")
def test_adaptive_bisection(tmp_path):
    """<Move type="Bisect" adaptive="1" target_ratio=...>: from the second write on,
    each replica's level steps down when its block acceptance is below the
    target and up otherwise (bisect_class.h:143-151); the driver's chain equals the
    binding's with the same level sequence, and the levels are written per block."""
    assert os.path.exists(DRIVER), "build the host driver first"
    s = system.config_c1()
    write_init(tmp_path, s)
    system.write_table(str(tmp_path / "bare.table"), system.bare_table_datasets(s))
    target = 0.985
    adapt = lambda x: x.replace('<Move name="Bisect" type="Bisect" species="e" n_level="3" />',  # noqa: E731
                                f'<Move name="Bisect" type="Bisect" species="e" n_level="3" adaptive="1" '
                                f'target_ratio="{target}" />').replace('<Loop n_step="2">', '<Loop n_step="5">')
    kc = (5.0 + 1e-9) * 2 * np.pi / s.L
    nb = (s.n_part * s.n_bead) >> s.bisect_n_level
    xml = tmp_path / "ad.xml"
    xml.write_text(adapt(XML.format(seed=4242, M=s.n_bead, beta=s.beta, L=s.L, kc=kc, N=s.n_part, nb=nb, prefix="ad",
                                    action=BARE.format(file="bare.table"), init=INIT_FILE, extra="")))
    res = subprocess.run([DRIVER, xml.name, "--replicas", "2"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    blocks = [json.loads(line) for line in open(tmp_path / "ad.spimc.jsonl")]
    assert len(blocks) == 5
    pos = system.random_path(s, seed=8)
    with abi.Engine([s, s], seed=4242) as e:
        for r in range(2):
            e.upload(pos, r=r)
        e.obs_configure(energy=abi.ENERGY_THERMAL, sk_k_cut=kc, gr_n_r=60)
        level = [3, 3]
        c_prev = [e.counters(r=r) for r in range(2)]
        for k, blk in enumerate(blocks):
            for _ in range(3):
                e.bisect(nb)
                e.displace(s.n_part)
                e.obs_accumulate(abi.OBS_ENERGY | abi.OBS_SK | abi.OBS_GR)
            for r in range(2):
                en, nm = e.obs_read(abi.OBS_ENERGY, r=r)
                got = blk["Observables"]["Energy"][r]
                assert got["Kinetic"] == pytest.approx(en[0] / (s.n_bead * nm), rel=1e-14), (k, r)
                mv = blk["Moves"]["Bisect"][r]
                assert mv["n_level"] == level[r], (k, r)
                c = e.counters(r=r)
                att, acc = c[0] - c_prev[r][0], c[1] - c_prev[r][1]
                c_prev[r] = c
                assert (mv["n_attempt"], mv["n_accept"]) == (att, acc)
                if k > 0:
                    if acc / att < target and level[r] > 1:
                        level[r] -= 1
                    elif (1 << level[r]) < s.n_bead // 2:
                        level[r] += 1
                    e.set_bisect_level(level[r], r=r)
            e.obs_reset(abi.OBS_ENERGY | abi.OBS_SK | abi.OBS_GR)
    levels = [b["Moves"]["Bisect"][0]["n_level"] for b in blocks]
    assert len(set(levels)) > 1, levels  # the level did change
    for r in range(2):
        got = h5dump.dataset(str(tmp_path / f"ad.{r}.h5"), "/Moves/Bisect/n_level")[1]
        assert list(got) == [b["Moves"]["Bisect"][r]["n_level"] for b in blocks]


@pytest.mark.skipif(h5dump.H5DUMP is None, reason="h5dump not installed")
def test_time_observable(tmp_path):
    """<Observable type="Time">: per block, the wall time and each event's time
    (time_class.h:38-67), as positive scalars in both outputs."""
    assert os.path.exists(DRIVER), "build the host driver first"
    s = system.config_c1()
    write_init(tmp_path, s)
    system.write_table(str(tmp_path / "bare.table"), system.bare_table_datasets(s))
    blocks = run_driver(tmp_path, s, BARE.format(file="bare.table"), "tm", 1,
                        extra='\n    <Observable name="T" type="Time" />')
    for b in blocks:
        t = b["Observables"]["T"]
        assert t["block"] > 0 and t["Bisect"] > 0 and t["Displace"] > 0
        assert t["Bisect"] + t["Displace"] + t["Energy"] + t["SK"] + t["GR"] <= t["block"] * 1.001
    f = str(tmp_path / "tm.0.h5")
    assert h5dump.dataset(f, "/Observables/T/block/data_type")[1] == ["scalar"]
    for ev in ("block", "Bisect", "Displace", "Energy"):
        dims, x = h5dump.dataset(f, f"/Observables/T/{ev}/x")
        assert dims == (2,) and np.all(x >= 0)
    np.testing.assert_allclose(h5dump.dataset(f, "/Observables/T/block/x")[1], [b["Observables"]["T"]["block"] for b in blocks],
                               rtol=1e-12)


PERM_XML = """<Input>
  <RNG seed="{seed}" />
  <IO output_prefix="{prefix}" />
  <System n_d="3" n_bead="{M}" beta="{beta!r}" L="{L!r}" pbc="1" k_cut="{kc!r}" />
  <Particles>
    <Species name="e" n_part="{N}" lambda="0.5" {init} />
  </Particles>
  <Actions>
    <Action name="Kinetic" type="Kinetic" species="e" />
    {action}
  </Actions>
  <Moves>
    <Move name="Perm" type="PermBisectIterative" species="e" n_level="3" epsilon="1e-100" />
    <Move name="Bisect" type="Bisect" species="e" n_level="3" />
    <Move name="Displace" type="DisplaceParticle" species="e" />
  </Moves>
  <Observables>
    <Observable name="Energy" type="Energy" />
    <Observable name="Cycles" type="Permutation" species="e" />
    <Observable name="path_dump" type="PathDump" />
  </Observables>
  <Algorithm>
    <Loop n_step="2">
      <Loop n_step="3">
        <Loop n_step="{nb}"><Move name="Perm" /></Loop>
        <Loop n_step="{nb}"><Move name="Bisect" /></Loop>
        <Loop n_step="{N}"><Move name="Displace" /></Loop>
        <Observable name="Energy" />
        <Observable name="Cycles" />
      </Loop>
      <Write />
    </Loop>
  </Algorithm>
</Input>
"""


def cycle_lengths(sigma):
    seen, out = np.zeros(len(sigma), dtype=bool), []
    for p in range(len(sigma)):
        if not seen[p]:
            n, q = 0, p
            while not seen[q]:
                seen[q] = True
                q = sigma[q]
                n += 1
            out.append(n)
    return out


@pytest.mark.skipif(h5dump.H5DUMP is None, reason="h5dump not installed")
def test_permuting_driver_matches_binding_and_restarts(tmp_path):
    """<Move type="PermBisectIterative"> with a Permutation observable and a
    PathDump: the driver's chain equals the binding's with the same seed (energies,
    per-cycle-length counters, the cycle histogram, the dumped path and its
    permutation rows); a Restart from the dump carries the permutation on."""
    assert os.path.exists(DRIVER), "build the host driver first"
    s = system.config_c1(n_part=8, n_bead=16, r_s=1.0, tau=0.05, n_level=3)
    write_init(tmp_path, s)
    system.write_table(str(tmp_path / "bare.table"), system.bare_table_datasets(s))
    kc = (5.0 + 1e-9) * 2 * np.pi / s.L
    N, M = s.n_part, s.n_bead
    nb = (N * M) >> s.bisect_n_level

    def run(prefix, init):
        (tmp_path / f"{prefix}.xml").write_text(PERM_XML.format(
            seed=4242, M=M, beta=s.beta, L=s.L, kc=kc, N=N, nb=nb, prefix=prefix, init=init,
            action=BARE.format(file="bare.table")))
        res = subprocess.run([DRIVER, f"{prefix}.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
        assert res.returncode == 0, res.stdout + res.stderr
        return [json.loads(line) for line in open(tmp_path / f"{prefix}.spimc.jsonl")]

    def binding(pos, sigma, blocks, h5file):
        with abi.Engine(s, seed=4242) as e:
            e.upload(pos)
            e.set_perm(sigma)
            e.obs_configure(energy=abi.ENERGY_THERMAL, sk_k_cut=kc, gr_n_r=100)
            last = e.perm_counters()
            for blk in blocks:
                cyc = np.zeros(N)
                for _ in range(3):
                    e.perm_bisect(nb)
                    e.bisect(nb)
                    e.displace(N)
                    e.obs_accumulate(abi.OBS_ENERGY)
                    for ln in cycle_lengths(e.perm()):
                        cyc[ln - 1] += 1
                en, nm = e.obs_read(abi.OBS_ENERGY)
                e.obs_reset(abi.OBS_ENERGY)
                got = blk["Observables"]["Energy"][0]
                assert got["Kinetic"] == pytest.approx(en[0] / (M * nm), rel=1e-14)
                assert got["Coulomb"] == pytest.approx(en[1] / (M * nm), rel=1e-14)
                pc = e.perm_counters()
                mv = blk["Moves"]["Perm"][0]
                assert (mv["n_attempt"], mv["n_accept"]) == (pc[0] - last[0], pc[1] - last[1])
                assert mv["perm_attempt"] == (pc[2] - last[2]).tolist()
                assert mv["perm_accept"] == (pc[3] - last[3]).tolist()
                last = pc
                np.testing.assert_allclose(blk["Observables"]["Cycles"][0]["cycles"], cyc / cyc.sum(), rtol=1e-15)
            dims, pdump = h5dump.dataset(h5file, "/Observables/path_dump/e/positions")
            assert np.array_equal(e.download(), pdump.reshape(dims)[-1].transpose(2, 1, 0))
            pdims, perm = h5dump.dataset(h5file, "/Observables/path_dump/e/permutation")
            rows = perm.reshape(pdims)[-1]
            sig = e.perm()
            assert np.array_equal(rows[:, 1], sig)
            assert np.array_equal(rows[sig, 0], np.arange(N))
            return sig

    blocks = run("P", INIT_FILE)
    sig = binding(system.random_path(s, seed=8), np.arange(N), blocks, str(tmp_path / "P.0.h5"))
    assert not np.array_equal(sig, np.arange(N)), "the run left the path unpermuted"
    f = str(tmp_path / "P.0.h5")
    assert h5dump.dataset(f, "/Moves/Perm/n_perm_type")[1][0] == N
    dims, pa = h5dump.dataset(f, "/Moves/Perm/perm_attempt")
    assert dims == (2, N)
    assert pa.reshape(dims)[1].tolist() == blocks[1]["Moves"]["Perm"][0]["perm_attempt"]
    dims, cy = h5dump.dataset(f, "/Observables/Cycles/cycles/y")
    assert dims == (2, N)
    np.testing.assert_allclose(cy.reshape(dims)[1], blocks[1]["Observables"]["Cycles"][0]["cycles"], rtol=1e-15)
    # sector pairs: every measurement of the block lands in one sector
    dims, sy = h5dump.dataset(f, "/Observables/Cycles/sectors/y")
    assert sum(v for _, v in sy.reshape(-1, 2)) == 2 * 3
    # restart from the dump: the permuted path goes on as the binding's does
    dims, pdump = h5dump.dataset(f, "/Observables/path_dump/e/positions")
    last_pos = pdump.reshape(dims)[-1].transpose(2, 1, 0)
    blocks2 = run("Q", 'init_type="Restart" prefix="P"')
    binding(last_pos, sig, blocks2, str(tmp_path / "Q.0.h5"))


@pytest.mark.skipif(h5dump.H5DUMP is None, reason="h5dump not installed")
def test_table_driver_matches_binding(tmp_path):
    """<Move type="PermBisectTable" n_perm_part="2"> beside PermBisectIterative:
    the driver's chain equals the binding's with the same seed (energies, the
    final path and permutation), each move reports its own attempts and
    per-cycle-length counts, and the table move's n_perm_type is 2! = 2."""
    assert os.path.exists(DRIVER), "build the host driver first"
    s = system.config_c1(n_part=8, n_bead=16, r_s=1.0, tau=0.05, n_level=3)
    write_init(tmp_path, s)
    system.write_table(str(tmp_path / "bare.table"), system.bare_table_datasets(s))
    kc = (5.0 + 1e-9) * 2 * np.pi / s.L
    N, M = s.n_part, s.n_bead
    nb = (N * M) >> s.bisect_n_level
    xml = PERM_XML.replace(
        '<Move name="Bisect" type="Bisect" species="e" n_level="3" />',
        '<Move name="Table" type="PermBisectTable" species="e" n_level="2" epsilon="1e-100" n_perm_part="2" '
        'detailed_balance="1" />\n    <Move name="Bisect" type="Bisect" species="e" n_level="3" />').replace(
        '<Loop n_step="{nb}"><Move name="Bisect" /></Loop>',
        '<Loop n_step="{nt}"><Move name="Table" /></Loop>\n        <Loop n_step="{nb}"><Move name="Bisect" /></Loop>')
    nt = 20
    (tmp_path / "T.xml").write_text(xml.format(seed=99, M=M, beta=s.beta, L=s.L, kc=kc, N=N, nb=nb, nt=nt, prefix="T",
                                               init=INIT_FILE, action=BARE.format(file="bare.table")))
    res = subprocess.run([DRIVER, "T.xml"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    blocks = [json.loads(line) for line in open(tmp_path / "T.spimc.jsonl")]
    with abi.Engine(s, seed=99) as e:
        e.upload(system.random_path(s, seed=8))
        e.obs_configure(energy=abi.ENERGY_THERMAL, sk_k_cut=kc, gr_n_r=100)
        e.perm_configure(3)
        e.perm_table_configure(2, 1)
        for blk in blocks:
            own = {"Perm": np.zeros(2 + 2 * N, dtype=np.int64), "Table": np.zeros(2 + 2 * N, dtype=np.int64)}

            def stage(name, fn, n):
                c0 = np.concatenate([np.atleast_1d(x) for x in e.perm_counters()]).astype(np.int64)
                fn(n)
                c1 = np.concatenate([np.atleast_1d(x) for x in e.perm_counters()]).astype(np.int64)
                own[name] += c1 - c0

            for _ in range(3):
                stage("Perm", e.perm_bisect, nb)
                e.perm_configure(2)  # the table move bisects at its own level
                stage("Table", e.perm_table, nt)
                e.perm_configure(3)
                e.bisect(nb)
                e.displace(N)
                e.obs_accumulate(abi.OBS_ENERGY)
            en, nm = e.obs_read(abi.OBS_ENERGY)
            e.obs_reset(abi.OBS_ENERGY)
            got = blk["Observables"]["Energy"][0]
            assert got["Kinetic"] == pytest.approx(en[0] / (M * nm), rel=1e-14)
            for name, n_per in (("Perm", nb), ("Table", nt)):
                mv = blk["Moves"][name][0]
                assert mv["n_attempt"] == 3 * n_per == own[name][0]
                assert mv["n_accept"] == own[name][1]
                assert mv["perm_attempt"] == own[name][2:2 + N].tolist()
                assert mv["perm_accept"] == own[name][2 + N:].tolist()
            assert blk["Moves"]["Table"][0]["perm_attempt"][1] == 3 * nt
        final = e.download()
        sig = e.perm()
    f = str(tmp_path / "T.0.h5")
    assert h5dump.dataset(f, "/Moves/Table/n_perm_type")[1][0] == 2
    dims, pdump = h5dump.dataset(f, "/Observables/path_dump/e/positions")
    assert np.array_equal(final, pdump.reshape(dims)[-1].transpose(2, 1, 0))
    pdims, perm = h5dump.dataset(f, "/Observables/path_dump/e/permutation")
    assert np.array_equal(perm.reshape(pdims)[-1][:, 1], sig)
