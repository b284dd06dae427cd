#!/usr/bin/env python3
"""Solver steps/sec of the sea-level time step on one MI355X, FP64 and mixed precision, at
lmax 512 and lmax 1023, with the accuracy condition of the metric measured in the same run.

Headline (``value``): workload "lmax512" = Gauss-Legendre 512 x 1024 grid, lmax 512, 30 viscous
modes per degree, 550 steps of 1 yr (seeded synthetic Love numbers, topography to degree 256, a
polar cap and a retreating marine sector), FP64, extrapolated inner start, rotation and marine
check on, epsilon1 = 1e-5.  A "step" is one solver step (modal advance, inner fixed point to
convergence, radial displacement, commit).  All ice grids are uploaded before the timed region;
W warm-up steps (SeaLevelSolver.step), then K timed steps through the native step loop
(SeaLevelSolver.run), each timed segment bracketed by device synchronisation; the run is cut into
segments at the steps where the oracle's long-run fixture stores fields (tests/golden), and the
fields are read there outside the timed region.

Also measured in a default run (every leg in full whatever --steps is, unless --no-also):
  * lmax512 FP64 over the whole workload (when --steps shortens the headline);
  * lmax512 and lmax1023 (1024 x 2048, 240 x 0.5 kyr, three caps) in FP64 and in mixed precision
    (FP32-product Legendre contractions in the early sweeps of each step, FP64 to convergence);
  * accuracy: FP64 runs against the oracle's stored fields and per-step iteration counts; mixed
    runs against the FP64 GPU run of the same steps (max |dSL|, |dR| in m, max |dm|/|m|, steps
    whose iteration count differs);
  * the roofline of the dominant kernel (the per-order Legendre analysis contraction), timed live
    with HIP events on the stream it is launched on, with the HBM traffic per launch from the
    committed rocprofv3 counter summary;
  * the CPU baseline on a bounded sample of the same workload on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lmax512|lmax1023|lmax63]
                    [--precision fp64|mixed] [--no-also] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solver steps/sec at lmax=512/1023, mixed vs FP64 GPU, at max|ΔSL|≤1e-6 m"
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12
GOLDEN = os.path.join(ROOT, "tests", "golden")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's steps - warmup)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="lmax512", choices=("lmax63", "lmax512", "lmax1023"))
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed"), help="precision of the headline")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-also", action="store_true", help="only the headline leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-run ensemble leg")
    ap.add_argument("--ensemble-runs", type=int, default=64)
    ap.add_argument("--inner", default="extrap", choices=("off", "extrap", "anderson"),
                    help="inner start / acceleration (the metric is quoted with extrap)")
    ap.add_argument("--kernels-only", default=None,
                    choices=("ana2", "ana2x32", "syn1", "syn1x32", "modal", "step"),
                    help="only launch one kernel 20x (the 2-field analysis contraction, FP64 or FP32-product; the "
                         "1-field synthesis contraction; the modal advance) or run 4 solver steps through the "
                         "native loop (step), then a 256 MiB device copy, for PMC runs")
    return ap.parse_args()


def measured_traffic(workload, path_kernel, kernel):
    """HBM bytes per launch of `kernel` (e.g. "k_leg_ana_ck<2,2>") on `workload` from the committed
    rocprofv3 PMC summary profiles/pmc_<path_kernel>_<workload>.txt (FETCH_SIZE and WRITE_SIZE,
    each scaled by what the counter reads for a 256 MiB device copy), or None when no summary
    covers this kernel and workload."""
    path = os.path.join(ROOT, "profiles", f"pmc_{path_kernel}_{workload}.txt")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return float(line.split(":", 1)[1]), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(w, budget):
    """The CPU baseline on this host: the oracle's solver step with its transforms and modal advance
    in multi-threaded C (oracle/cpu.py: on-the-fly Legendre recurrence with N-S symmetry, OpenMP over
    orders, scipy.fft with worker threads), or the plain NumPy oracle when that library is not built.
    Set-up, then as many steps as fit in `budget` seconds (at least 1) of the same workload; steps/s
    over the timed steps."""
    from oracle.love import LoveNumbers as OLove

    try:
        from oracle import cpu as ocpu

        ocpu.load()
        nthreads = ocpu.threads()
        slv = ocpu.FastCpuSolver(ocpu.FastCpuSHT(w.nlat, w.nlon, w.lmax), OLove(w.love_text, w.lmax, text=True),
                                 w.topo, w.ice(0), norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner="extrap")
        variant = (f"oracle step with C/OpenMP on-the-fly Legendre transforms and modal advance "
                   f"(oracle/cpu_kernels.c, {nthreads} threads) and scipy.fft ({nthreads} workers)")
    except (ImportError, OSError, RuntimeError):
        from oracle.sht import OracleSHT
        from oracle.solver import SeaLevelSolver as OSolver

        try:
            nthreads = len(os.sched_getaffinity(0))
        except AttributeError:
            nthreads = os.cpu_count()
        nthreads = int(min(nthreads, int(os.environ.get("OMP_NUM_THREADS", nthreads))))
        slv = OSolver(OracleSHT(w.nlat, w.nlon, w.lmax), OLove(w.love_text, w.lmax, text=True), w.topo, w.ice(0),
                      norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner="extrap")
        variant = f"FP64 NumPy/SciPy oracle (stored Legendre tables, BLAS threads {nthreads})"
    t0 = time.perf_counter()
    n = 0
    nit = 0
    while True:
        nit += slv.step(w.ice(n + 1))["numiter"]
        n += 1
        if time.perf_counter() - t0 > budget or n >= w.nsteps:
            break
    el = time.perf_counter() - t0
    return {"value": n / el, "unit": "steps/s", "cores": int(nthreads), "kind": "port",
            "sample": f"first {n} steps of the {w.name} workload ({nit / n:.2f} inner iterations per step), "
                      f"{el:.1f} s; {variant}"}


def fixture(name):
    """The oracle's long-run fixture of workload `name` (tests/golden/oracle_<name>_long.npz) or None."""
    import numpy as np

    path = os.path.join(GOLDEN, f"oracle_{name}_long.npz")
    return dict(np.load(path)) if os.path.exists(path) else None


class Leg:
    """One timed run of a workload: warm-up steps through step(), then the timed steps through the
    native loop in segments that end at the checkpoint steps, where tgrid, R and m are kept."""

    def __init__(self, w, ice_dev, dev, precision, inner, warmup, steps, checkpoints):
        import numpy as np
        import torch

        from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

        last = warmup + steps
        sht = SHT(w.nlat, w.nlon, w.lmax, device=dev)
        slv = OptSeaLevelSolver(sht, LoveNumbers(w.love_text, w.lmax, text=True), w.topo, w.ice(0), norder=w.lmax,
                                dt=w.dt, tpw=True, checkmarine=True, calcRG=True, epsilon1=1e-5, inner=inner,
                                precision=precision)
        self.numiter, self.sweeps32, self.fields = [], [], {}
        cps = sorted({k for k in checkpoints if warmup < k <= last} | {last})
        for k in range(1, warmup + 1):
            r = slv.step(ice_dev[k])
            self.numiter.append(r["numiter"])
            self.sweeps32.append(slv._last_status.sweeps32)
            if k in checkpoints:
                self.fields[k] = (r["tgrid"].clone(), r["R"].clone(), np.array(r["mm"]))
        el, k0 = 0.0, warmup + 1
        for cp in cps:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rec = slv.run(ice_dev, k0, cp + 1)
            torch.cuda.synchronize()
            el += time.perf_counter() - t0
            self.numiter += [x["numiter"] for x in rec]
            self.sweeps32 += [x["sweeps32"] for x in rec]
            self.fields[cp] = (slv.topoxy.clone(), slv.rrxy.clone(), np.array(rec[-1]["mm"]))
            k0 = cp + 1
        self.elapsed, self.steps, self.warmup = el, steps, warmup
        self.sps = steps / el
        self.timed_numiter = self.numiter[warmup:]
        self.sht_kernels = dict(sht.kernels)
        del slv, sht
        torch.cuda.synchronize()
        print(f"[bench] {w.name} {precision}: {steps} steps in {el:.3f} s ({self.sps:.1f} steps/s), numiter mean "
              f"{np.mean(self.timed_numiter):.2f}", file=sys.stderr, flush=True)

    def summary(self, w, precision, inner):
        import numpy as np

        out = {"workload": describe(w, inner, precision), "value": self.sps, "unit": "steps/s", "steps": self.steps,
               "warmup": self.warmup, "ms_per_step": 1e3 * self.elapsed / self.steps,
               "numiter_mean": float(np.mean(self.timed_numiter)), "numiter_total": int(np.sum(self.timed_numiter))}
        if precision == "mixed":
            n = max(1, int(np.sum(self.timed_numiter)))
            out["fp32_sweep_fraction"] = float(np.sum(self.sweeps32[self.warmup:]) / n)
        return out


def vs_oracle(leg, fx):
    """FP64 leg against the oracle's long-run fixture: max |dtgrid|, |dR| on the stored subsample at
    the checkpoints reached, steps whose iteration count differs."""
    import numpy as np

    if fx is None:
        return {"note": "no oracle fixture for this workload"}
    n = min(len(leg.numiter), len(fx["numiter"]))
    res = {"reference": "FP64 NumPy oracle (tests/golden long-run fixture)", "steps_compared": n,
           "numiter_mismatch_steps": int(np.sum(np.asarray(leg.numiter[:n]) != fx["numiter"][:n])),
           "checkpoints": []}
    s = int(fx["sub"])
    for i, k in enumerate(fx["check"]):
        k = int(k)
        if k not in leg.fields:
            continue
        t, r, mm = leg.fields[k]
        dt = float(np.abs(t.cpu().numpy()[::s, ::s] - fx["tgrid_sub"][i]).max())
        dr = float(np.abs(r.cpu().numpy()[::s, ::s] - fx["R_sub"][i]).max())
        dm = float(np.abs(mm - fx["mm"][k - 1]).max() / max(np.abs(fx["mm"][k - 1]).max(), 1e-30))
        res["checkpoints"].append({"step": k, "max_dSL_m": dt, "max_dR_m": dr, "max_dm_rel": dm})
    if res["checkpoints"]:
        res["max_dSL_m"] = max(c["max_dSL_m"] for c in res["checkpoints"])
        res["max_dR_m"] = max(c["max_dR_m"] for c in res["checkpoints"])
        res["max_dm_rel"] = max(c["max_dm_rel"] for c in res["checkpoints"])
    return res


def vs_fp64(mixed, fp64):
    """Mixed leg against the FP64 GPU leg of the same steps (full grids at the common checkpoints)."""
    import numpy as np

    n = min(len(mixed.numiter), len(fp64.numiter))
    res = {"reference": "FP64 GPU run of the same steps", "steps_compared": n,
           "numiter_mismatch_steps": int(np.sum(np.asarray(mixed.numiter[:n]) != np.asarray(fp64.numiter[:n]))),
           "checkpoints": []}
    for k in sorted(set(mixed.fields) & set(fp64.fields)):
        (ta, ra, ma), (tb, rb, mb) = mixed.fields[k], fp64.fields[k]
        res["checkpoints"].append({"step": k, "max_dSL_m": float((ta - tb).abs().max().item()),
                                   "max_dR_m": float((ra - rb).abs().max().item()),
                                   "max_dm_rel": float(np.abs(ma - mb).max() / max(np.abs(mb).max(), 1e-30))})
    for key in ("max_dSL_m", "max_dR_m", "max_dm_rel"):
        res[key] = max((c[key] for c in res["checkpoints"]), default=None)
    res["meets_1e-6_m"] = bool(res["max_dSL_m"] is not None and res["max_dSL_m"] <= 1e-6)
    return res


def ensemble_leg(nruns, dev):
    """The batched configuration: nruns independent runs (lmax 128/256/512, 50-1200 steps, dt 0.2 /
    0.5 / 1.0 kyr, perturbed Love numbers, random caps), every member alone on one stream, then the
    whole ensemble on the autotuned number of concurrent streams with per-lmax tuned kernel
    variants; ensemble steps/s, the members' solo steps/s and the achieved concurrency (sum of the
    members' solo times over the ensemble's wall time)."""
    import numpy as np
    import torch

    from slevolve_amd import ensemble as E

    t0 = time.perf_counter()
    specs = E.make_ensemble(nruns)
    ices = E.prepare(specs, dev)
    var = E.tune_variants({s.lmax for s in specs}, dev)
    prep = time.perf_counter() - t0
    print(f"[bench] ensemble: {nruns} members prepared in {prep:.1f} s, variants {var}", file=sys.stderr, flush=True)
    slot = E._Stream(dev, var)
    solo = {}
    for s in specs:
        solo[s.id] = E.run_one(s, ices[s.id], slot)
    t_solo = sum(v[0] for v in solo.values())
    best, rates = E.autotune_streams(specs, ices, dev=dev, variants=var)
    wall, out, _ = E.run_ensemble(specs, ices, nstreams=best, dev=dev, variants=var)
    nst = sum(s.nsteps for s in specs)
    same = all(torch.equal(solo[s.id][2], out[s.id][2]) for s in specs)
    by_l = {}
    for s in specs:
        d = by_l.setdefault(str(s.lmax), {"members": 0, "steps": 0, "solo_s": 0.0})
        d["members"] += 1
        d["steps"] += s.nsteps
        d["solo_s"] += solo[s.id][0]
    for d in by_l.values():
        d["solo_steps_per_s"] = d["steps"] / d["solo_s"]
    del ices
    torch.cuda.empty_cache()
    print(f"[bench] ensemble: {nst} steps, alone {t_solo:.2f} s, ensemble {wall:.2f} s on {best} streams",
          file=sys.stderr, flush=True)
    return {"members": nruns, "steps_total": nst, "value": nst / wall, "unit": "steps/s",
            "solo_steps_per_s": nst / t_solo, "speedup_vs_alone": t_solo / wall, "streams": best,
            "stream_probe_steps_per_s": {str(k): v for k, v in rates.items()},
            "kernel_variants": {str(k): v for k, v in var.items()}, "achieved_concurrency": t_solo / wall,
            "members_equal_solo_runs": bool(same), "by_lmax": by_l, "prepare_s": prep,
            "numiter_mean": float(np.mean([r["numiter"] for v in out.values() for r in v[1]]))}


def describe(w, inner="extrap", precision="fp64"):
    prec = "FP64" if precision == "fp64" else "mixed (FP32-product contractions in early sweeps, FP64 to convergence)"
    return (f"{w.name}: {w.nlat}x{w.nlon} GL grid, lmax {w.lmax}, {w.nmodes} modes/degree, "
            f"{w.nsteps} x {w.dt:g} yr, {prec}, inner {inner}, tpw+checkmarine on, eps1 1e-5")


def upload(w, nsteps, dev):
    import torch

    t0 = time.perf_counter()
    ice_dev = torch.empty((nsteps + 1, w.nlat, w.nlon), dtype=torch.float64, device=dev)
    for k in range(nsteps + 1):
        ice_dev[k].copy_(torch.from_numpy(w.ice(k)))
    print(f"[bench] {w.name}: inputs on the device ({ice_dev.numel() * 8 / 1e9:.2f} GB, "
          f"{time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    return ice_dev


def kernels_only(w, which, dev, precision="fp64"):
    """Launch one contraction 20x, or run 4 steps of the workload through the native step loop
    (every kernel of the time step), and a 256 MiB copy as the counter scale, for rocprofv3 --pmc."""
    import numpy as np
    import torch

    from slevolve_amd import SHT
    from slevolve_amd import _device as D
    from slevolve_amd import _lib

    if which == "modal":
        from slevolve_amd import LoveNumbers
        from slevolve_amd.memory import ExactMemory

        mem = ExactMemory(LoveNumbers(w.love_text, w.lmax, text=True), w.lmax, w.dt, device=dev)
        mem._p_load.copy_(torch.randn_like(mem._p_load))
        for _ in range(20):
            mem.advance_and_read_packed()
    elif which == "step":
        from slevolve_amd import LoveNumbers, OptSeaLevelSolver

        ice_dev = upload(w, 4, dev)
        slv = OptSeaLevelSolver(SHT(w.nlat, w.nlon, w.lmax, device=dev), LoveNumbers(w.love_text, w.lmax, text=True),
                                w.topo, w.ice(0), norder=w.lmax, dt=w.dt, tpw=True, checkmarine=True, calcRG=True,
                                epsilon1=1e-5, inner="extrap", precision=precision)
        rec = slv.run(ice_dev, 1, 5)
        print(f"[bench] {w.name}: 4 steps, numiter {[r['numiter'] for r in rec]}", file=sys.stderr, flush=True)
    else:
        sht = SHT(w.nlat, w.nlon, w.lmax, device=dev, precision="mixed" if which.endswith("x32") else "fp64")
        grids = torch.from_numpy(np.stack([w.topo, w.ice(1)])).to(dev)
        coef = torch.zeros((2, sht.ncoef), dtype=torch.complex128, device=dev)
        sht.analysis_packed(grids, coef)
        for _ in range(20):
            if which.startswith("ana"):
                sht.analysis_contraction(2, coef)
            else:
                _lib.call("slx_sht_synthesis", sht.handle, coef.data_ptr(), 1, grids.data_ptr(), D.stream_ptr())
    src = torch.ones(32 * 2**20, dtype=torch.float64, device=dev)  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"kernels_only": which, "copy_bytes": src.numel() * 8}))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU sharding of "
              f"this path is a design note, not a measured configuration", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from slevolve_amd import SHT
    from slevolve_amd import synthetic as syn

    w = syn.workload(a.workload)
    dev = torch.device("cuda", 0)
    if a.kernels_only:
        kernels_only(w, a.kernels_only, dev, a.precision)
        return
    steps = a.steps if a.steps is not None else w.nsteps - a.warmup
    if a.warmup + steps > w.nsteps:
        raise SystemExit(f"warmup + steps = {a.warmup + steps} exceeds the workload's {w.nsteps} steps")
    fx = fixture(w.name)
    cps = set(int(k) for k in fx["check"]) if fx is not None else set()
    ice_dev = upload(w, w.nsteps, dev)
    head = Leg(w, ice_dev, dev, a.precision, a.inner, a.warmup, steps, cps)
    # ---- dominant kernel: the per-order analysis contraction of a 2-field sweep, timed with HIP
    # events on the current stream (the stream the library launches on)
    sht = SHT(w.nlat, w.nlon, w.lmax, device=dev)
    grids = torch.from_numpy(np.stack([w.topo, w.ice(1)])).to(dev)
    coef = torch.zeros((2, sht.ncoef), dtype=torch.complex128, device=dev)
    sht.analysis_packed(grids, coef)
    reps = 50
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(5):
        sht.analysis_contraction(2, coef)
    ev0.record(stream)
    for _ in range(reps):
        sht.analysis_contraction(2, coef)
    ev1.record(stream)
    torch.cuda.synchronize()
    t_kernel = ev0.elapsed_time(ev1) / 1e3 / reps
    flops = 2.0 * 2 * sht.ncoef * w.nlat  # 4 flops per (l, m, ring pair) per field, 2 fields
    kname = f"{sht.kernels['analysis']}<2,{sht.kernels['analysis_rings_per_lane']}>"
    traffic, traffic_src = measured_traffic(w.name, "leg_ana_2field", kname)
    roof = {"bound": "mfma", "achieved": flops / t_kernel / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
            "frac": flops / t_kernel / PEAK_FP64, "traffic": traffic, "traffic_source": traffic_src,
            "kernel": f"{kname} (per-order FP64 Legendre analysis contraction on the matrix cores, 2 fields; "
                      f"variant chosen by shape)",
            "kernel_us": t_kernel * 1e6, "flops_per_launch": flops}
    del sht, grids, coef
    out = {
        "metric": METRIC, "value": head.sps, "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * head.elapsed / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f64+f32", "data": "synthetic",
        "config": {"workload": describe(w, a.inner, a.precision),
                   "numiter_mean": float(np.mean(head.timed_numiter)),
                   "numiter_total": int(np.sum(head.timed_numiter)), "ice": w.note},
        "roofline": roof,
    }
    acc = {}
    if a.precision == "fp64":
        acc[f"{w.name}_fp64"] = vs_oracle(head, fx)
    if not a.no_also and a.inner == "extrap" and a.workload in ("lmax512", "lmax1023"):
        also, mixed = [], {}
        legs = {}
        full = w.nsteps - a.warmup
        if a.precision == "fp64" and steps == full:
            legs[(w.name, "fp64")] = head
        elif a.precision == "fp64":
            legs[(w.name, "fp64")] = Leg(w, ice_dev, dev, "fp64", a.inner, a.warmup, full, cps)
            also.append(legs[(w.name, "fp64")].summary(w, "fp64", a.inner))
            acc[f"{w.name}_fp64_full"] = vs_oracle(legs[(w.name, "fp64")], fx)
        else:
            legs[(w.name, "fp64")] = Leg(w, ice_dev, dev, "fp64", a.inner, a.warmup, full, cps)
        legs[(w.name, "mixed")] = Leg(w, ice_dev, dev, "mixed", a.inner, a.warmup, full, cps)
        del ice_dev
        torch.cuda.empty_cache()
        other = "lmax1023" if w.name == "lmax512" else "lmax512"
        w2 = syn.workload(other)
        fx2 = fixture(other)
        cps2 = set(int(k) for k in fx2["check"]) if fx2 is not None else set()
        ice2 = upload(w2, w2.nsteps, dev)
        warm2 = 5 if other == "lmax1023" else 10
        legs[(other, "fp64")] = Leg(w2, ice2, dev, "fp64", a.inner, warm2, w2.nsteps - warm2, cps2)
        legs[(other, "mixed")] = Leg(w2, ice2, dev, "mixed", a.inner, warm2, w2.nsteps - warm2, cps2)
        del ice2
        also.append(legs[(other, "fp64")].summary(w2, "fp64", a.inner))
        acc[f"{other}_fp64"] = vs_oracle(legs[(other, "fp64")], fx2)
        for ww in (w, w2):
            f64, mx = legs[(ww.name, "fp64")], legs[(ww.name, "mixed")]
            s = mx.summary(ww, "mixed", a.inner)
            s["ratio_to_fp64"] = mx.sps / f64.sps
            s["accuracy_vs_fp64"] = vs_fp64(mx, f64)
            mixed[ww.name] = s
        out["also"] = also
        out["mixed"] = mixed
    out["accuracy"] = acc
    if not a.no_ensemble and not a.no_also:
        out["ensemble"] = ensemble_leg(a.ensemble_runs, dev)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
