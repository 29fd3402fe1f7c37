"""HDF5 pair-action tables: the driver's converter (`spimc --table-to-h5`)
writes every dataset of a table with its name, kind and values, as h5dump reads
them back; a David u_kj cube is stored [n_tau][n_grid][n_val] and an Ilkka
u_xy matrix [n_y][n_x], or [n_val][n_grid][n_tau] and [n_x][n_y] with
--transposed; and the driver's HDF5 reader (used by the converter when its
input is HDF5) restores the same values from either layout.  CPU only: the
converter makes no device call."""
import os
import subprocess

import numpy as np
import pytest

import h5dump
from simpimc_amd import system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "simpimc_amd", "spimc")

pytestmark = pytest.mark.skipif(h5dump.H5DUMP is None, reason="h5dump not installed")


def convert(src, dst, *extra):
    assert os.path.exists(DRIVER), "build the host driver first"
    res = subprocess.run([DRIVER, "--table-to-h5", str(src), str(dst), *extra], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr


def david_three_taus(s):
    """A David table whose cubes hold three taus; tau slices 1 and 2 are seeded noise."""
    d = system.david_table_datasets(s.L, 7.0 / s.L, s.tau, s.k_n, n_grid=64)
    pre, dpre = "u_kj_2", "du_kj_dbeta_2"
    n_grid = d[f"{pre}/grid/n_grid_points"]
    n_val = d[f"{pre}/data"].size // n_grid
    rng = np.random.default_rng(11)
    for name in (f"{pre}/data", f"{dpre}/data"):
        cube = np.zeros((n_val, n_grid, 3), order="F")
        cube[:, :, 0] = d[name].reshape((n_val, n_grid), order="F")
        cube[:, :, 1:] = rng.standard_normal((n_val, n_grid, 2))
        d[name] = cube.ravel(order="F")
    d[f"{pre}/taus"] = np.array([s.tau, 2 * s.tau, 3 * s.tau])
    return d, n_val, n_grid


def test_converter_writes_every_dataset(tmp_path):
    s = system.config_c1()
    ilkka = system.ilkka_table_datasets(system.config_ilkka())
    for label, d in (("bare", system.bare_table_datasets(s)), ("david", david_three_taus(s)[0]), ("ilkka", ilkka)):
        system.write_table(str(tmp_path / f"{label}.table"), d)
        convert(tmp_path / f"{label}.table", tmp_path / f"{label}.h5")
        assert sorted(h5dump.names(str(tmp_path / f"{label}.h5"))) == sorted("/" + k for k in d)
        for name, val in d.items():
            dims, got = h5dump.dataset(str(tmp_path / f"{label}.h5"), "/" + name)
            if isinstance(val, str):
                assert got == [val], name
                continue
            want = np.atleast_1d(np.asarray(val, dtype=np.float64)).ravel()
            assert np.array_equal(got, want), name  # C-order storage = first cube index fastest
            if name.endswith("/data") and name.startswith(("u_kj_", "du_kj_dbeta_")):
                assert dims == (3, 64, want.size // (3 * 64)), (name, dims)
            if name.endswith("_xy"):
                g = name.split("/")[0]
                assert dims == (d[f"{g}/off_diag/n_y"], d[f"{g}/off_diag/n_x"]), (name, dims)


def test_cube_layouts_round_trip(tmp_path):
    s = system.config_c1()
    d, n_val, n_grid = david_three_taus(s)
    system.write_table(str(tmp_path / "d.table"), d)
    convert(tmp_path / "d.table", tmp_path / "a.h5")
    convert(tmp_path / "d.table", tmp_path / "b.h5", "--transposed")
    for name in ("u_kj_2/data", "du_kj_dbeta_2/data"):
        dims, b = h5dump.dataset(str(tmp_path / "b.h5"), "/" + name)
        assert dims == (n_val, n_grid, 3)
        # element [c][g][t] of the transposed cube is component c, knot g, tau t
        ref = d[name].reshape((n_val, n_grid, 3), order="F")
        assert np.array_equal(b.reshape(dims), ref)
    # HDF5 -> HDF5 through the driver's reader: both layouts give file a
    convert(tmp_path / "b.h5", tmp_path / "c.h5")
    convert(tmp_path / "a.h5", tmp_path / "e.h5")
    for out in ("c.h5", "e.h5"):
        assert sorted(h5dump.names(str(tmp_path / out))) == sorted(h5dump.names(str(tmp_path / "a.h5")))
        for name in h5dump.names(str(tmp_path / "a.h5")):
            da, va = h5dump.dataset(str(tmp_path / "a.h5"), name)
            dc, vc = h5dump.dataset(str(tmp_path / out), name)
            assert da == dc, (out, name)
            if isinstance(va, list):
                assert va == vc, (out, name)
            else:
                assert np.array_equal(va, vc), (out, name)


def test_matrix_layouts_round_trip(tmp_path):
    s = system.config_ilkka()
    d = system.ilkka_table_datasets(s)
    system.write_table(str(tmp_path / "i.table"), d)
    convert(tmp_path / "i.table", tmp_path / "a.h5")
    convert(tmp_path / "i.table", tmp_path / "b.h5", "--transposed")
    for g, v in (("u", "u_xy"), ("du", "du_xy")):
        n_x, n_y = d[f"{g}/off_diag/n_x"], d[f"{g}/off_diag/n_y"]
        assert n_x != n_y
        dims, b = h5dump.dataset(str(tmp_path / "b.h5"), f"/{g}/off_diag/{v}")
        assert dims == (n_x, n_y)
        # element [i][j] of the transposed matrix is the value at (x_i, y_j)
        assert np.array_equal(b.reshape(dims), s.ilkka[v].T)
    convert(tmp_path / "b.h5", tmp_path / "c.h5")
    for name in h5dump.names(str(tmp_path / "a.h5")):
        da, va = h5dump.dataset(str(tmp_path / "a.h5"), name)
        dc, vc = h5dump.dataset(str(tmp_path / "c.h5"), name)
        assert da == dc, name
        assert (va == vc) if isinstance(va, list) else np.array_equal(va, vc), name


def test_text_and_hdf5_errors(tmp_path):
    bad = tmp_path / "bad.table"
    bad.write_text("not a table\n")
    res = subprocess.run([DRIVER, "--table-to-h5", str(bad), str(tmp_path / "x.h5")], capture_output=True, text=True)
    assert res.returncode == 1 and "neither HDF5 nor a spimc-table 1 file" in res.stderr
    res = subprocess.run([DRIVER, "--table-to-h5", str(tmp_path / "missing.h5"), str(tmp_path / "y.h5")], capture_output=True,
                         text=True)
    assert res.returncode == 1 and "cannot open" in res.stderr
