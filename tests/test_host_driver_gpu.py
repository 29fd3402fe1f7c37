"""The C++ XML driver (simpimc_amd/spimc) on a simpimc input file gives the same
block averages as the same run driven through the Python binding: both reach
the device through the C ABI with the same seed, so the chains are identical."""
import json
import os
import subprocess

import numpy as np
import pytest

from simpimc_amd import abi, system

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "simpimc_amd", "spimc")

XML = """<Input>
  <RNG seed="{seed}" />
  <IO output_prefix="run" />
  <System n_d="3" n_bead="{M}" beta="{beta!r}" L="{L!r}" pbc="1" k_cut="{kc!r}" />
  <Particles>
    <Species name="e" n_part="{N}" lambda="0.5" init_type="File" file_name="init.xyz" />
  </Particles>
  <Actions>
    <Action name="Kinetic" type="Kinetic" species="e" />
    <Action name="Coulomb" type="BarePairAction" file="bare.table" species_a="e" species_b="e" is_coulomb="1"
            max_level="0" use_long_range="1" />
  </Actions>
  <Moves>
    <Move name="Bisect" type="Bisect" species="e" n_level="3" />
    <Move name="Displace" type="DisplaceParticle" species="e" />
  </Moves>
  <Observables>
    <Observable name="Energy" type="Energy" />
    <Observable name="SK" type="StructureFactor" species_a="e" species_b="e" />
    <Observable name="GR" type="PairCorrelation" species_a="e" species_b="e" n_r="60" />
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
    (tmp_path / "in.xml").write_text(XML.format(seed=seed, M=s.n_bead, beta=s.beta, L=s.L, kc=kc, N=s.n_part, nb=nb))
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
