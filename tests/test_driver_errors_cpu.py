"""Input errors of the XML driver, reported before any device work (the
reference's std::cerr + exit(1) paths: actions.h:32-33, moves.h:30-31,
david_pair_action_class.h:241-244, a missing table file).  CPU only: the
driver rejects these while building the system, before it creates a context."""
import os
import subprocess

import numpy as np
import pytest

from simpimc_amd import system
from test_host_driver_gpu import BARE, DAVID, DRIVER, XML

INIT = 'init_type="File" file_name="init.xyz"'


def run(tmp_path, action, extra="", moves=None, tweak=None):
    s = system.config_c1()
    pos = system.random_path(s, seed=8)
    with open(tmp_path / "init.xyz", "w") as f:
        for p in range(s.n_part):
            f.write("%.17g %.17g %.17g\n" % tuple(pos[:, 0, p]))
    kc = (5.0 + 1e-9) * 2 * np.pi / s.L
    nb = (s.n_part * s.n_bead) >> s.bisect_n_level
    xml = XML.format(seed=1, M=s.n_bead, beta=s.beta, L=s.L, kc=kc, N=s.n_part, nb=nb, prefix="run", action=action,
                     init=INIT, extra=extra)
    if tweak:
        xml = tweak(xml)
    (tmp_path / "in.xml").write_text(xml)
    assert os.path.exists(DRIVER), "build the host driver first"
    return subprocess.run([DRIVER, "in.xml", "--no-h5"], cwd=tmp_path, capture_output=True, text=True, timeout=60)


def test_unknown_action_type(tmp_path):
    res = run(tmp_path, '<Action name="X" type="TrapAction" species="e" />')
    assert res.returncode == 1
    assert "Unrecognized Action, TrapAction" in res.stderr


def test_missing_table_file(tmp_path):
    res = run(tmp_path, BARE.format(file="nowhere.table"))
    assert res.returncode == 1
    assert "cannot open pair-action file nowhere.table" in res.stderr


def test_david_tau_not_in_table(tmp_path):
    s = system.config_c1()
    d = system.david_table_datasets(s.L, 7.0 / s.L, s.tau, s.k_n, n_grid=64)
    d["u_kj_2/taus"] = np.array([2.0 * s.tau])  # the run's tau is missing
    system.write_table(str(tmp_path / "david.table"), d)
    res = run(tmp_path, DAVID.format(file="david.table"))
    assert res.returncode == 1
    assert "tau of" in res.stderr and "not found" in res.stderr


def test_unknown_move_and_observable(tmp_path):
    system.write_table(str(tmp_path / "bare.table"), system.bare_table_datasets(system.config_c1()))
    res = run(tmp_path, BARE.format(file="bare.table"),
              tweak=lambda x: x.replace('type="DisplaceParticle"', 'type="VaryAction"'))
    assert res.returncode == 1
    assert "Unrecognized Move, VaryAction" in res.stderr
    res = run(tmp_path, BARE.format(file="bare.table"), extra='\n    <Observable name="S" type="Sign" />')
    assert res.returncode == 1
    assert "Unrecognized observable, Sign" in res.stderr


@pytest.mark.parametrize("attr", ["fermi=\"1\"", "pbc=\"0\""])
def test_unsupported_system_is_reported(tmp_path, attr):
    system.write_table(str(tmp_path / "bare.table"), system.bare_table_datasets(system.config_c1()))
    if attr.startswith("fermi"):
        tweak = lambda x: x.replace('lambda="0.5"', 'lambda="0.5" ' + attr)  # noqa: E731
    else:
        tweak = lambda x: x.replace('pbc="1"', attr)  # noqa: E731
    res = run(tmp_path, BARE.format(file="bare.table"), tweak=tweak)
    assert res.returncode == 1
    assert "ERROR:" in res.stderr
