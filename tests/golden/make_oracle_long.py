#!/usr/bin/env python3
"""Long-horizon oracle vectors for the two benchmark workloads.

The FP64 NumPy oracle needs ~3-4 s per step of the lmax-512 workload and ~45 s per step of the
lmax-1023 workload on 8 cores, so its outputs over long runs are stored as compact fixtures that
the GPU tests compare with the timed path (SeaLevelSolver.run: the native step loop with the
incremental grounded-ice analyses):

  oracle_lmax512_long.npz    550 steps (the whole 512 x 1024, lmax 512 workload)
  oracle_lmax1023_long.npz   40 steps of the 1024 x 2048, lmax 1023 workload

    numiter, mm, gmslc_ocn, gmslc_ocnBeta, ocean_area, oceanBeta_area, ice_volume   every step
    check                      the checkpoint steps (1-based)
    tgrid_sub, R_sub           every SUB-th row and column of tgrid and R at each checkpoint
    tgrid_sum, R_sum, tgrid_ss, R_ss   full-grid sums and sums of squares at each checkpoint

Inputs: the seeded workloads of slevolve_amd.synthetic, epsilon1 1e-5, inner "extrap", rotation
and marine check on.  A file is rewritten after every checkpoint, so an interrupted run leaves the
checkpoints it reached ("complete" says whether all steps ran).  With --state DIR the whole
generator state (oracle solver included) is pickled into DIR every 10 steps and a restarted run
continues from there.
Run:  python tests/golden/make_oracle_long.py [--state DIR] [lmax512|lmax1023 ...]
"""
import os
import pickle
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

PLAN = {"lmax512": (550, (25, 100, 550), 8), "lmax1023": (40, (10, 20, 40), 16)}
KEYS = ("gmslc_ocn", "gmslc_ocnBeta", "ocean_area", "oceanBeta_area", "ice_volume")


def run(name, state_dir=None):
    from oracle.love import LoveNumbers
    from oracle.sht import OracleSHT
    from oracle.solver import SeaLevelSolver
    from slevolve_amd import synthetic as syn

    steps, check, sub = PLAN[name]
    w = syn.workload(name)
    spath = os.path.join(state_dir, f"oracle_{name}_state.pkl") if state_dir else None
    if spath and os.path.exists(spath):
        with open(spath, "rb") as f:  # written by this script
            slv, per, at, done, k_done = pickle.load(f)
        print(f"{name}: resuming after step {k_done}", flush=True)
    else:
        slv = SeaLevelSolver(OracleSHT(w.nlat, w.nlon, w.lmax), LoveNumbers(w.love_text, w.lmax, text=True), w.topo,
                             w.ice(0), norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner="extrap")
        per = {k: [] for k in ("numiter", "mm") + KEYS}
        at = {k: [] for k in ("tgrid_sub", "R_sub", "tgrid_sum", "R_sum", "tgrid_ss", "R_ss")}
        done, k_done = [], 0
    t0 = time.time()
    for k in range(k_done + 1, steps + 1):
        r = slv.step(w.ice(k))
        per["numiter"].append(r["numiter"])
        per["mm"].append(np.asarray(r["mm"]).copy())
        for key in KEYS:
            per[key].append(float(np.real(r[key])))
        print(f"{name} step {k}: numiter {r['numiter']} ({time.time() - t0:.0f} s)", flush=True)
        if k in check:
            for f in ("tgrid", "R"):
                g = np.asarray(r[f])
                at[f + "_sub"].append(g[::sub, ::sub].copy())
                at[f + "_sum"].append(g.sum())
                at[f + "_ss"].append((g * g).sum())
            done.append(k)
            np.savez_compressed(os.path.join(HERE, f"oracle_{name}_long.npz"), sub=sub, check=np.asarray(done),
                                complete=(k == steps), **{a: np.asarray(v) for a, v in per.items()},
                                **{a: np.asarray(v) for a, v in at.items()})
        if spath and (k % 10 == 0 or k == steps):
            with open(spath + ".tmp", "wb") as f:
                pickle.dump((slv, per, at, done, k), f, protocol=pickle.HIGHEST_PROTOCOL)
            os.replace(spath + ".tmp", spath)


if __name__ == "__main__":
    args = sys.argv[1:]
    state = None
    if args[:1] == ["--state"]:
        state, args = args[1], args[2:]
        os.makedirs(state, exist_ok=True)
    for n in args or list(PLAN):
        run(n, state)
