#!/usr/bin/env python3
"""Oracle vectors for the 1024 x 2048, lmax 1023 workload (two solver steps).

The FP64 NumPy oracle takes about 45 s per step at this size, too slow for a GPU test to run
beside the device path, so its outputs are stored here as a compact fixture:

  oracle_lmax1023.npz
    numiter            inner iterations of steps 1 and 2
    tgrid_sub, R_sub   every 16th row and column (64 x 128) of tgrid and R after each step
    tgrid_sum, R_sum   the full-grid sums (FP64), and tgrid_ss / R_ss the sums of squares,
                       after each step
    mm                 the rotation vector after each step
    gmslc_ocn, gmslc_ocnBeta, ocean_area, oceanBeta_area, ice_volume   degree-0 outputs of each step

Inputs are the seeded "lmax1023" workload (slevolve_amd.synthetic), epsilon1 1e-5, inner "extrap",
rotation and marine check on.  Run:  python tests/golden/make_oracle_lmax1023.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

STEPS = 2
SUB = 16
KEYS = ("gmslc_ocn", "gmslc_ocnBeta", "ocean_area", "oceanBeta_area", "ice_volume")


def main():
    from oracle.love import LoveNumbers
    from oracle.sht import OracleSHT
    from oracle.solver import SeaLevelSolver
    from slevolve_amd import synthetic as syn

    w = syn.workload("lmax1023")
    slv = SeaLevelSolver(OracleSHT(w.nlat, w.nlon, w.lmax), LoveNumbers(w.love_text, w.lmax, text=True), w.topo,
                         w.ice(0), norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner="extrap")
    out = {k: [] for k in ("numiter", "tgrid_sub", "R_sub", "tgrid_sum", "R_sum", "tgrid_ss", "R_ss", "mm") + KEYS}
    for k in range(1, STEPS + 1):
        r = slv.step(w.ice(k))
        print(f"step {k}: numiter {r['numiter']}", flush=True)
        out["numiter"].append(r["numiter"])
        for f in ("tgrid", "R"):
            g = np.asarray(r[f])
            out[f + "_sub"].append(g[::SUB, ::SUB].copy())
            out[f + "_sum"].append(g.sum())
            out[f + "_ss"].append((g * g).sum())
        out["mm"].append(np.asarray(r["mm"]).copy())
        for key in KEYS:
            out[key].append(float(np.real(r[key])))
    np.savez_compressed(os.path.join(HERE, "oracle_lmax1023.npz"), sub=SUB, **{k: np.asarray(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
