#!/usr/bin/env python3
"""Rebuild the golden fixtures of this directory from the reference's own data files.

Only data is read (``.npy`` arrays and the Love-number text file); no reference code is
imported or run.  Run it in a checkout that has the reference mounted at /root/reference:

    python tests/golden/make_fixtures.py [REFERENCE_ROOT]

Outputs (all small, compressed):
  topography_seed_mm.npz   reference/synthetic/topography_seed.npy as int32 millimetres
                           (the seed is rounded to whole mm, so the conversion is exact;
                           tools/make_synthetic_case.py:105-107)
  tgrid8_change.npz        reference/synthetic/tgrid8.npy minus the seed (the topography
                           change after 8 steps, 512 x 1024, lmax 511; the fixture of
                           tests/test_fortran_synthetic.sh:51-52,91-96); topo + change
                           reproduces tgrid8.npy bit for bit
  prem_512.l96C.ump3.lm1.gz   lovenum/prem_512.l96C.ump3.lm1, gzip-compressed
"""
import gzip
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref="/root/reference"):
    seed = np.load(os.path.join(ref, "reference", "synthetic", "topography_seed.npy"), allow_pickle=False)
    tg8 = np.load(os.path.join(ref, "reference", "synthetic", "tgrid8.npy"), allow_pickle=False)
    mm = np.round(seed * 1000.0).astype(np.int32)
    if not np.array_equal(mm / 1000.0, seed):
        raise SystemExit("topography seed is not whole millimetres")
    change = tg8 - seed
    if not np.array_equal(mm / 1000.0 + change, tg8):
        raise SystemExit("seed + change does not reproduce tgrid8 exactly")
    np.savez_compressed(os.path.join(HERE, "topography_seed_mm.npz"), topo_mm=mm)
    np.savez_compressed(os.path.join(HERE, "tgrid8_change.npz"), change=change)
    src = os.path.join(ref, "lovenum", "prem_512.l96C.ump3.lm1")
    with open(src, "rb") as f, gzip.open(os.path.join(HERE, "prem_512.l96C.ump3.lm1.gz"), "wb", 9) as o:
        shutil.copyfileobj(f, o)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(*sys.argv[1:])
