#!/usr/bin/env python3
"""Extract the CAMB columns the reference's linear-theory test reads.

delta_tot_table_test.c:231-310 (load_camb_transfer) reads, for each of the 99
scale factors a = 0.01 .. 0.99 of camb_linear/, the first NREAD=200 rows with
k > 2 pi / 512 of the transfer file (columns k, T_nu = col 6, T_tot = col 7,
T_cdm = col 8) and of the matter-power file (k, P).  This script copies those
raw numbers (data only) into camb_linear_200.npz so the test can run without
the 20 MB directory.  It reads the data files under /root/reference/camb_linear.

Run: python3 tests/golden/make_camb_fixture.py
"""
import math
import os

import numpy as np

SRC = "/root/reference/camb_linear"
NREAD = 200
KMIN = 2 * math.pi / 512.


def read_rows(path, ncol):
    rows = []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                continue
            parts = line.split()
            if len(parts) < ncol:
                break
            vals = [float(p) for p in parts[:ncol]]
            if vals[0] > KMIN:
                rows.append(vals)
            if len(rows) == NREAD:
                break
    assert len(rows) == NREAD, path
    return np.array(rows)


def main():
    scalefacts = [0.01 + i * 0.01 for i in range(99)]
    tk, tnu, ttot, tcdm, pk_k, pk = [], [], [], [], [], []
    for a in scalefacts:
        tag = "%g" % a
        t = read_rows(os.path.join(SRC, "ics_transfer_%s.dat" % tag), 8)
        m = read_rows(os.path.join(SRC, "ics_matterpow_%s.dat" % tag), 2)
        tk.append(t[:, 0]); tnu.append(t[:, 5]); ttot.append(t[:, 6]); tcdm.append(t[:, 7])
        pk_k.append(m[:, 0]); pk.append(m[:, 1])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "camb_linear_200.npz")
    np.savez_compressed(out, scalefact=np.array(scalefacts), k_transfer=np.array(tk), T_nu=np.array(tnu),
                        T_tot=np.array(ttot), T_cdm=np.array(tcdm), k_pk=np.array(pk_k), P=np.array(pk))
    print("wrote", out)


if __name__ == "__main__":
    main()
