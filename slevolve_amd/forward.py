"""The forward-run driver: the command line and outputs of the reference's ``run_ais.py``
(python/slevolve/run_ais.py:21-96 options, :230-468 main; console entry ``slevolve-forward``,
pyproject.toml:29) on the device solver.

    python -m slevolve_amd.forward --steps 550 --dt 1 --icedir ICE --icepfx ice --topo TOPO \\
        --earth LOVE --nglv 512 --lmax 511 --inner extrap --out RUN [--precision fp64|mixed]

The ice history is read once (directly, or through the reference's ice cache with its JSON
sidecar) into HBM; every step runs on the device; per-step snapshots (``--save-fields``) are
written by a writer thread from pinned copies taken on a side stream, so file I/O overlaps the
next steps.  Files: tgrid0/ocean0/beta0.npy, <field><k>.npy per step, numiter.npy,
step_wall_seconds.npy and run_log.json with the reference's keys plus the device, the precision
and the FP32-product sweep counts.  Options the device solver does not provide
(``--axial-rotation`` other than off, ``--horizontal-grid``) are refused; ``--threads``,
``--chunk`` and ``--mem-threads`` are accepted and recorded (they set CPU thread counts in the
reference).  Additional options: ``--precision``, ``--eps-switch``, ``--device``,
``--cache-dir``, ``--native`` (no per-step outputs: all steps through the native step loop).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

DEFAULT_PFX = "AISload_gd_Run85_0_0_"


def parse_args(argv=None):
    ap = argparse.ArgumentParser(prog="slevolve-forward", description=__doc__.split("\n\n")[0])
    ap.add_argument("--steps", type=int, default=50, help="number of time steps N; reads ice files 0..N")
    ap.add_argument("--dt", type=float, default=1.0, help="step length in years (ice file k is at time k dt)")
    ap.add_argument("--times", default=None, help="times file: one time (years) per ice file 0..N; --dt is ignored")
    ap.add_argument("--threads", type=int, default=1, help="CPU threads (recorded; the step runs on the GPU)")
    ap.add_argument("--out", default=None, help="output directory")
    ap.add_argument("--root", default=None, help="root directory with INPUT_FILES/ (default inputs)")
    ap.add_argument("--norder", type=int, default=512, help="spectral truncation")
    ap.add_argument("--lmax", type=int, default=None, help="alias for --norder")
    ap.add_argument("--sht", default="numpy", choices=("numpy", "ducc", "hip"), help="any name maps to the GPU transform")
    ap.add_argument("--nglv", type=int, default=512, help="Gauss-Legendre latitudes (grid nglv x 2 nglv)")
    ap.add_argument("--tpw", default="T", help="T|F rotational feedback")
    ap.add_argument("--checkmarine", default="T", help="T|F marine check")
    ap.add_argument("--calcRG", default="T", help="T|F radial displacement and geoid")
    ap.add_argument("--epsilon1", type=float, default=1.0e-5, help="inner-iteration tolerance")
    ap.add_argument("--profile", action="store_true", help="per-stage timing of the device step (log)")
    ap.add_argument("--cprofile", action="store_true", help="host profile of the run (cprofile.txt)")
    ap.add_argument("--save-fields", default="tgrid,ocean,beta",
                    help="per-step outputs: tgrid,ocean,beta,dS,R,G,dSLdt,dRdt")
    ap.add_argument("--ramp", action="store_true", help="second-order load placement")
    ap.add_argument("--axial-rotation", default="off", choices=("off", "deg2", "full"))
    ap.add_argument("--axial-tau-c", type=float, default=None)
    ap.add_argument("--horizontal", action="store_true", help="write horiz<N>.npy (Psi_lm)")
    ap.add_argument("--horizontal-grid", action="store_true")
    ap.add_argument("--cache-only", action="store_true", help="only build the ice cache")
    ap.add_argument("--rebuild-cache", action="store_true")
    ap.add_argument("--chunk", type=int, default=16, help="ice staging chunk (files per host-to-device copy)")
    ap.add_argument("--mem-threads", type=int, default=1)
    ap.add_argument("--opt", default="", help="raw,diag,dsl,numba,sums or 'all' (validated; the device path is fused)")
    ap.add_argument("--inner", default="off", choices=("off", "extrap", "anderson"))
    ap.add_argument("--anderson-depth", type=int, default=3)
    ap.add_argument("--anderson-freeze", type=int, default=2)
    ap.add_argument("--count-sht", action="store_true", help="record the transform count per step (log)")
    ap.add_argument("--anderson-keep", action="store_true")
    ap.add_argument("--rhoi", type=float, default=None)
    ap.add_argument("--radius", type=float, default=None)
    ap.add_argument("--icedir", default=None, help="ice directory: <icedir>/<icepfx><k>{.npy,.bin,text}")
    ap.add_argument("--icepfx", default=DEFAULT_PFX)
    ap.add_argument("--topo", default=None, help="initial topography (.npy, .bin or text)")
    ap.add_argument("--io-compat", default="fortran", choices=("fortran", "legacy"))
    ap.add_argument("--earth", default=None, help="Love-number file")
    # device-path options
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed"),
                    help="Legendre contractions: FP64, or FP32 products in the early sweeps of each step")
    ap.add_argument("--eps-switch", type=float, default=None, help="mixed: FP64 sweeps once xi <= this")
    ap.add_argument("--device", type=int, default=0, help="GPU index")
    ap.add_argument("--cache-dir", default=None, help="ice cache directory (default <root>/cache without --icedir)")
    ap.add_argument("--native", action="store_true",
                    help="run every step through the native step loop (per-step outputs: numiter and scalars only)")
    return ap.parse_args(argv)


def _uptime():
    try:
        return subprocess.check_output(["uptime"], text=True).strip()
    except Exception:
        return "?"


def _loadavg():
    try:
        return os.getloadavg()
    except Exception:
        return (float("nan"),) * 3


def make_planet(rhoi=None, radius=None):
    """Planet() with the --rhoi / --radius overrides (moments of inertia follow the radius,
    run_ais.py:138-149)."""
    from .solver import Planet

    p = Planet()
    if rhoi is not None:
        p.rhoi = float(rhoi)
    if radius is not None:
        p.radius = float(radius)
        p.moiA = 0.3296145 * p.mass * p.radius ** 2
        p.moiC = 0.3307007 * p.mass * p.radius ** 2
    return p


def main(argv=None):
    a = parse_args(argv)
    import numpy as np
    import torch

    from . import npyio
    from .love import LoveNumbers
    from .sht import make_sht
    from .snapshots import IceStack, SnapshotWriter
    from .solver import OptSeaLevelSolver, SeaLevelSolver, parse_opts

    if a.axial_rotation != "off" or a.horizontal_grid:
        raise SystemExit("--axial-rotation deg2|full and --horizontal-grid are not provided by the device solver")
    root = a.root or os.getcwd()
    icedir = a.icedir or os.path.join(root, "INPUT_FILES", "icemodel")
    topo_path = a.topo or os.path.join(root, "INPUT_FILES", "others", "topo0_archive")
    earth = a.earth or os.path.join(root, "INPUT_FILES", "earthmodel",
                                    "prem_512.l60K2C.sum18p6.dum19p2.tz19p4.lm22")
    nglv, nlon = a.nglv, 2 * a.nglv
    norder = a.lmax if a.lmax is not None else a.norder
    tpw, cm, rg = (v.upper().startswith("T") for v in (a.tpw, a.checkmarine, a.calcRG))
    out = a.out or os.path.join(root, "results", f"slevolve_amd_n{a.steps}")
    os.makedirs(out, exist_ok=True)
    fields = set(f for f in a.save_fields.split(",") if f)
    opts = parse_opts(a.opt)
    torch.cuda.set_device(a.device)
    dev = torch.device("cuda", a.device)
    times = npyio.read_times(a.times, a.steps) if a.times else None
    dts = np.diff(times) if times is not None else None
    log = {"argv": sys.argv if argv is None else ["slevolve-forward", *argv], "threads": a.threads,
           "steps": a.steps, "dt_years": a.dt if times is None else "variable",
           "times_file": os.path.abspath(a.times) if a.times else None,
           "times_years": [float(x) for x in times] if times is not None else None, "norder": norder,
           "sht_backend": a.sht, "mem_threads": a.mem_threads, "opt": sorted(opts), "nglv": nglv, "tpw": tpw,
           "checkmarine": cm, "calcRG": rg, "epsilon1": a.epsilon1, "ramp": bool(a.ramp),
           "axial_rotation": a.axial_rotation, "axial_tau_c": a.axial_tau_c, "numpy": np.__version__,
           "torch": torch.__version__, "device": torch.cuda.get_device_name(dev), "device_index": a.device,
           "precision": a.precision, "uptime_start": _uptime(), "loadavg_start": _loadavg()}
    print(f"[slevolve_amd] lmax={norder} grid {nglv}x{nlon} steps={a.steps} precision={a.precision} "
          f"inner={a.inner} device={log['device']}", flush=True)

    cache_dir = a.cache_dir if a.cache_dir is not None else (None if a.icedir else os.path.join(root, "cache"))
    t0 = time.perf_counter()
    stack = IceStack(icedir, a.icepfx, a.steps, nglv, nlon, device=dev, cache_dir=cache_dir, rebuild=a.rebuild_cache,
                     chunk=a.chunk)
    log.update({"icedir": os.path.abspath(icedir), "ice_cache_path": stack.cache_path,
                "ice_cache_seconds": stack.cache_seconds, "ice_upload_s": time.perf_counter() - t0})
    if a.cache_only:
        return 0
    log.update({"inner": a.inner, "anderson_depth": a.anderson_depth, "rhoi": a.rhoi, "radius": a.radius,
                "topo": topo_path, "earth": earth})
    t0 = time.perf_counter()
    topo0 = npyio.read_grid(topo_path, nglv, nlon)
    t_topo = time.perf_counter() - t0
    t0 = time.perf_counter()
    love = LoveNumbers(earth, norder)
    t_love = time.perf_counter() - t0
    t0 = time.perf_counter()
    sht = make_sht(a.sht if a.sht != "hip" else "ducc", nglv, nlon, norder, workers=a.threads, chunk=a.chunk,
                   raw="raw" in opts, device=dev)
    t_sht = time.perf_counter() - t0
    log.update({"setup_topo_s": t_topo, "setup_love_s": t_love, "setup_sht_s": t_sht,
                "sht_table_bytes": int(sht.table_bytes), "love_max_modes": int(getattr(love, "max_modes", 0))})
    print(f"[slevolve_amd] setup: topo {t_topo:.2f}s  love {t_love:.2f}s  transform {t_sht:.2f}s", flush=True)

    t0 = time.perf_counter()
    acc = a.inner != "off"
    extra = dict(precision=a.precision, eps_switch=a.eps_switch)
    if opts or acc:
        cls = OptSeaLevelSolver
        extra.update(opt=opts, inner=a.inner, anderson_depth=a.anderson_depth, anderson_freeze=a.anderson_freeze,
                     anderson_keep=a.anderson_keep)
    else:
        cls = SeaLevelSolver
    if a.rhoi is not None or a.radius is not None:
        extra["planet"] = make_planet(a.rhoi, a.radius)
    slv = cls(sht, love, topo0, stack[0], norder=norder, dt=a.dt if times is None else float(dts[0]), tpw=tpw,
              checkmarine=cm, calcRG=rg, epsilon1=a.epsilon1, mem_threads=a.mem_threads, ramp=a.ramp, **extra)
    if a.horizontal:
        slv.horizontal = True
        log["horizontal"] = "coef"
    log["solver_init_s"] = time.perf_counter() - t0

    def h(t):
        return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)

    npyio.save_grid(os.path.join(out, "tgrid0.npy"), h(slv.tinit_0))
    npyio.save_mask(os.path.join(out, "ocean0.npy"), h(slv.cxy0), a.io_compat)
    npyio.save_mask(os.path.join(out, "beta0.npy"), h(slv.beta0), a.io_compat)
    scalars = [{"ocean_area": slv.ocean_area0, "oceanBeta_area": slv.oceanBeta_area0,
                "ice_volume": slv.ice_volume0}]
    numiter, walls, sweeps32, prof_tot = [], [], [], {}
    keys = ("ocean_area", "oceanBeta_area", "gmslc_ocn", "gmslc_ocnBeta", "ice_volume")
    want_rates = bool(fields & {"dSLdt", "dRdt"})
    pr = None
    if a.cprofile:
        import cProfile

        pr = cProfile.Profile()
        pr.enable()
    t_run = time.perf_counter()
    if a.native:
        torch.cuda.synchronize(dev)
        rec = slv.run(stack.grids, 1, a.steps + 1, dts=dts)
        torch.cuda.synchronize(dev)
        el = time.perf_counter() - t_run
        numiter = [r["numiter"] for r in rec]
        sweeps32 = [r["sweeps32"] for r in rec]
        walls = [el / max(1, a.steps)] * a.steps
        scalars += [{k: r[k] for k in keys} for r in rec]
        for f in fields & {"tgrid", "ocean", "beta", "R", "G"}:  # the last step's grids
            grid = {"tgrid": slv.topoxy, "ocean": slv.cxy, "beta": slv.beta, "R": slv.rrxy, "G": slv.Gxy}[f]
            if f in ("ocean", "beta"):
                npyio.save_mask(os.path.join(out, f"{f}{a.steps}.npy"), h(grid), a.io_compat)
            else:
                npyio.save_grid(os.path.join(out, f"{f}{a.steps}.npy"), h(grid))
    else:
        with SnapshotWriter(out, fields=fields, io_compat=a.io_compat, device=dev) as wr:
            for k in range(1, a.steps + 1):
                r = slv.step(stack[k], profile=a.profile, dt=None if dts is None else float(dts[k - 1]),
                             rates=want_rates)
                wr.write(k, r)
                numiter.append(r["numiter"])
                walls.append(r["wall"])
                sweeps32.append(slv._last_status.sweeps32)
                if a.profile and r.get("prof"):
                    for kk, vv in r["prof"].items():
                        prof_tot[kk] = prof_tot.get(kk, 0.0) + vv
                scalars.append({kk: r[kk] for kk in keys})
                if k % 10 == 0 or k <= 3:
                    print(f"[slevolve_amd] step {k:4d}  numiter={r['numiter']}  {r['wall'] * 1e3:.2f} ms", flush=True)
    t_run = time.perf_counter() - t_run
    if pr is not None:
        import pstats

        pr.disable()
        with open(os.path.join(out, "cprofile.txt"), "w") as fh:
            pstats.Stats(pr, stream=fh).sort_stats("cumulative").print_stats(40)
    np.save(os.path.join(out, "numiter.npy"), np.asarray(numiter, dtype=np.int32))
    np.save(os.path.join(out, "step_wall_seconds.npy"), np.asarray(walls))
    log.update({"total_solver_seconds": t_run, "per_step_mean": float(np.mean(walls)) if walls else None,
                "per_step_median": float(np.median(walls)) if walls else None,
                "per_step_first10": [float(x) for x in walls[:10]],
                "per_step_last10": [float(x) for x in walls[-10:]],
                "numiter_total": int(np.sum(numiter)), "numiter_mean": float(np.mean(numiter)) if numiter else None,
                "inner_stats": getattr(slv, "inner_stats", None),
                "sht_counts": ({"per_step": "2 + 2 numiter sweeps (fused); see the device kernels"}
                               if a.count_sht else None),
                "profile_seconds": prof_tot, "fp32_sweeps": [int(x) for x in sweeps32],
                "uptime_end": _uptime(), "loadavg_end": _loadavg(), "scalars": scalars,
                "steps_per_second": a.steps / t_run if t_run > 0 else None})
    with open(os.path.join(out, "run_log.json"), "w") as fh:
        json.dump(log, fh, indent=1, default=float)
    print(f"[slevolve_amd] DONE {a.steps} steps in {t_run:.2f} s ({a.steps / max(t_run, 1e-9):.1f} steps/s), "
          f"numiter total {int(np.sum(numiter))}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
