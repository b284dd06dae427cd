"""Outer (initial-topography) iteration on the GPU, as run_outer.py drives it.

Restates python/slevolve/run_outer.py:174-279 around the device solver: pass k runs the whole
forward model from T0_k (every ice grid already resident in HBM), the residual r_k = pred_k - obs
(pred_k = topography after the last step) stays on the device, and T0 is updated there
  Picard (run_outer.py:270):   T0_{k+1} = T0_k - r_k
  secant (run_outer.py:252-268, from pass 2 on, per cell):
      s = (r_k - r_{k-1}) / (T0_k - T0_{k-1}) where |T0_k - T0_{k-1}| > secant_tol, else 1;
      non-finite s -> 1;  s = clip(s, smin, smax);  T0_{k+1} = T0_k - r_k / s.
reuse="warm" keeps one store of per-step converged states across the passes
(solver_opt.py:272-286), on the device: 16 B per coefficient + 8 B per grid point per step.

    recs, T0_next, preds = outer_iteration(sht, love, obs, ice, lmax=511, dt=500.0, iters=5,
                                           outer="secant", reuse="warm", inner="extrap")
"""
from __future__ import annotations

import numpy as np
import torch

from . import _device as D
from .solver import OptSeaLevelSolver


def aw_rms(f: torch.Tensor, w: torch.Tensor) -> float:
    """Area-weighted RMS with GL row weights w (run_outer.py:140-142)."""
    W = w[:, None].expand_as(f)
    return float(torch.sqrt((W * f * f).sum() / W.sum()))


def outer_iteration(sht, love, obs, ice, *, lmax, dt=None, dts=None, iters=3, outer="picard", secant_tol=1e-3,
                    smin=0.5, smax=1.5, reuse="off", topo0=None, solver_cls=OptSeaLevelSolver, **solver_kw):
    """Returns (records, T0_next, [pred_k]) with device tensors; records as run_outer.py's log
    (rms_pred_minus_obs, max_abs, rms_T0_minus_obs, numiter, numiter_steps, secant counters).
    ice: sequence of nsteps + 1 grids (numpy or device tensors); dts: per-step lengths (years) or
    dt for every step."""
    if outer not in ("picard", "secant"):
        raise ValueError("outer must be picard or secant")
    dev = sht.dev
    shape = (sht.nlat, sht.nlon)
    obs_d = D.as_device(obs, torch.float64, dev, shape).clone()
    ice_d = [D.as_device(g, torch.float64, dev, shape) for g in ice]
    nsteps = len(ice_d) - 1
    gw = torch.as_tensor(np.asarray(sht.weights, dtype=np.float64), device=dev)
    T0 = obs_d.clone() if topo0 is None else D.as_device(topo0, torch.float64, dev, shape).clone()
    store = {}
    hist = []
    records, preds = [], []
    for it in range(1, iters + 1):
        kw = dict(solver_kw)
        if reuse != "off":
            kw.update(reuse=reuse, reuse_store=store)
        slv = solver_cls(sht, love, T0, ice_d[0], norder=lmax, dt=dt if dts is None else float(dts[0]), **kw)
        nits = []
        for k in range(1, nsteps + 1):
            r = slv.step(ice_d[k], dt=None if dts is None else float(dts[k - 1]))
            nits.append(int(r["numiter"]))
        pred = r["tgrid"].clone()
        del slv
        preds.append(pred)
        res = pred - obs_d
        rec = {"it": it, "rms_pred_minus_obs": aw_rms(res, gw), "max_abs": float(res.abs().max()),
               "rms_T0_minus_obs": aw_rms(T0 - obs_d, gw), "numiter": int(sum(nits)), "numiter_steps": nits}
        hist.append((T0, res))
        if outer == "secant" and len(hist) >= 2:
            (T0p, rp), (T0c, rc) = hist[-2], hist[-1]
            dT = T0c - T0p
            ok = dT.abs() > secant_tol
            s = torch.ones_like(dT)
            s[ok] = (rc[ok] - rp[ok]) / dT[ok]
            s[~torch.isfinite(s)] = 1.0
            ncl = int(((s < smin) | (s > smax)).sum())
            s = s.clamp(smin, smax)
            T0 = T0c - rc / s
            rec.update({"secant_cells_picard": int((~ok).sum()), "secant_cells_clamped": ncl})
        else:
            T0 = T0 - res
        hist = hist[-2:]
        records.append(rec)
    return records, T0, preds
