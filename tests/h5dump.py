"""Reads datasets of an HDF5 file through the HDF5 distribution's own h5dump
tool, so that files written by the driver are checked by code other than the
driver's reader."""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import numpy as np

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def dataset(path: str, name: str):
    """(dims, values) of dataset `name`: a float64 array in storage order, or a
    list of str for a string dataset."""
    res = subprocess.run([H5DUMP, "-y", "-w", "0", "-m", "%.17g", "-d", name, path], capture_output=True, text=True)
    if res.returncode != 0 or "h5dump error" in res.stdout + res.stderr:
        raise KeyError(f"{path}: {name}: {res.stdout[-300:]} {res.stderr[-300:]}")
    txt = res.stdout
    m = re.search(r"DATASPACE\s+SIMPLE \{ \( ([^)]*) \)", txt)
    dims = tuple(int(x) for x in m.group(1).split(",")) if m else ()
    body = txt[txt.index("DATA {") + len("DATA {"):]
    body = body[: body.index("}")]
    if "H5T_STRING" in txt:
        return dims, re.findall(r'"((?:[^"\\]|\\.)*)"', body)
    vals = [v for v in re.split(r"[,\s]+", body.strip()) if v]
    return dims, np.array([float(v) for v in vals], dtype=np.float64)


def names(path: str) -> list[str]:
    """Every dataset path of the file."""
    res = subprocess.run([H5DUMP, "-n", path], capture_output=True, text=True, check=True)
    return [line.split()[1] for line in res.stdout.splitlines() if line.strip().startswith("dataset ")]
