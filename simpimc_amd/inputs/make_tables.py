"""Writes the pair-action table files the example inputs name.

python simpimc_amd/inputs/make_tables.py [outdir]
  heg_c1_bare.table   BarePairAction v/diag/* for 32 electrons at r_s = 2 (alpha = 5/L)
  heg_c2_david.table  DavidPairAction tables for 256 electrons at r_s = 1, tau = 0.02
"""
import math
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from simpimc_amd import system  # noqa: E402


def main(outdir="."):
    c1 = system.config_c1()
    system.write_table(os.path.join(outdir, "heg_c1_bare.table"), system.bare_table_datasets(c1))
    c2 = system.config_c2()
    system.write_table(os.path.join(outdir, "heg_c2_david.table"),
                       system.david_table_datasets(c2.L, 7.0 / c2.L, c2.tau, c2.k_n, n_grid=2000, n_order=2, seed=7))
    print("tables written to", os.path.abspath(outdir))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else ".")
