"""Oracle outer (initial-topography) iteration of run_outer.py, test infrastructure only.

Restates python/slevolve/run_outer.py:174-279 around the oracle solver: pass k runs the whole
forward model from the initial topography T0_k, the residual is r_k = pred_k - obs with pred_k
the topography after the last step, and
  Picard (run_outer.py:270):            T0_{k+1} = T0_k - r_k
  secant (run_outer.py:252-268, from pass 2 on, per cell):
      s = (r_k - r_{k-1}) / (T0_k - T0_{k-1})   where |T0_k - T0_{k-1}| > secant_tol
      s = 1 where the denominator is tiny or s is not finite;  s = clip(s, smin, smax)
      T0_{k+1} = T0_k - r_k / s.
The per-pass record holds the area-weighted RMS of r_k and of T0_k - obs (GL row weights,
run_outer.py:140-142), max |r_k| and the inner-iteration counts.  reuse="warm" shares one
reuse_store across the passes (run_outer.py:198-204).
"""
from __future__ import annotations

import numpy as np

from .solver import SeaLevelSolver


def aw_rms(f, w):
    """Area-weighted RMS with the row weights w (run_outer.py:140-142)."""
    W = np.broadcast_to(w[:, None], f.shape)
    return float(np.sqrt((W * f**2).sum() / W.sum()))


def outer_iteration(sht, love, obs, ice, *, lmax, dt=None, dts=None, iters=3, outer="picard", secant_tol=1e-3,
                    smin=0.5, smax=1.5, reuse="off", topo0=None, solver_cls=SeaLevelSolver, **solver_kw):
    """Returns (records, T0_next, [pred_k]).  ice: sequence of nsteps + 1 grids; dts: per-step
    lengths (or dt for all); solver_cls: the forward model (the oracle solver by default)."""
    if outer not in ("picard", "secant"):
        raise ValueError("outer must be picard or secant")
    obs = np.asarray(obs, dtype=np.float64)
    nsteps = len(ice) - 1
    gw = np.asarray(sht.weights, dtype=np.float64)
    T0 = obs.copy() if topo0 is None else np.asarray(topo0, dtype=np.float64).copy()
    store = {}
    hist = []
    records, preds = [], []
    for it in range(1, iters + 1):
        slv = solver_cls(sht, love, T0, ice[0], norder=lmax, dt=dt if dts is None else float(dts[0]),
                         reuse=reuse, reuse_store=store, **solver_kw)
        nits = []
        for k in range(1, nsteps + 1):
            r = slv.step(ice[k], dt=None if dts is None else float(dts[k - 1]))
            nits.append(r["numiter"])
        pred = r["tgrid"].copy()
        preds.append(pred)
        res = pred - obs
        rec = {"it": it, "rms_pred_minus_obs": aw_rms(res, gw), "max_abs": float(np.abs(res).max()),
               "rms_T0_minus_obs": aw_rms(T0 - obs, gw), "numiter": int(sum(nits)), "numiter_steps": nits}
        hist.append((T0, res))
        if outer == "secant" and len(hist) >= 2:
            (T0p, rp), (T0c, rc) = hist[-2], hist[-1]
            dT = T0c - T0p
            ok = np.abs(dT) > secant_tol
            s = np.ones_like(dT)
            with np.errstate(divide="ignore", invalid="ignore"):
                s[ok] = (rc[ok] - rp[ok]) / dT[ok]
            s[~np.isfinite(s)] = 1.0
            ncl = int(((s < smin) | (s > smax)).sum())
            s = np.clip(s, smin, smax)
            T0 = T0c - rc / s
            rec.update({"secant_cells_picard": int((~ok).sum()), "secant_cells_clamped": ncl})
        else:
            T0 = T0 - res
        hist = hist[-2:]
        records.append(rec)
    return records, T0, preds
