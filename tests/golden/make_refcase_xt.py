"""Writes tests/golden/refcase_xt_coarse.npz from eduPIC's published XT maps
of the reference case (C/ReferenceCase_RawData/Figure11*_xt.dat: N_G = 400
rows, N_XT = 200 time bins of N_BIN = 20 steps).  Each map is reduced to the
means of blocks of 10 nodes x 5 time bins (40 x 40), small enough to commit.

    python tests/golden/make_refcase_xt.py [/root/reference]

The data files are read as text; nothing of the reference is executed.
"""
import os
import sys

import numpy as np

FILES = {
    "ne": "Figure11a)_ne_xt.dat", "ni": "Figure11b)_ni_xt.dat", "pot": "Figure11c)_pot_xt.dat",
    "efield": "Figure11d)_efield_xt.dat", "je": "Figure11e)_je_xt.dat", "ji": "Figure11f)_ji_xt.dat",
    "powere": "Figure11g)_heate_xt.dat", "poweri": "Figure11h)_heati_xt.dat",
    "meanee": "Figure11i)_meanee_xt.dat", "meanei": "Figure11j)_meanei_xt.dat", "ioniz": "Figure11k)_ioniz_xt.dat",
}
BLOCK_NODES, BLOCK_BINS = 10, 5


def coarse(a):
    n, m = a.shape
    return a.reshape(n // BLOCK_NODES, BLOCK_NODES, m // BLOCK_BINS, BLOCK_BINS).mean(axis=(1, 3))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "C", "ReferenceCase_RawData")
    out = {}
    for name, fname in FILES.items():
        a = np.loadtxt(os.path.join(src, fname))
        assert a.shape == (400, 200), (fname, a.shape)
        out[name] = coarse(a)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "refcase_xt_coarse.npz")
    np.savez_compressed(dst, block=np.array([BLOCK_NODES, BLOCK_BINS]), **out)
    print("wrote", dst)


if __name__ == "__main__":
    main()
