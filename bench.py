#!/usr/bin/env python3
"""Solver steps/sec of the sea-level time step on one MI355X.

Workload (default "lmax512"): Gauss-Legendre 512 x 1024 grid, lmax 512, 30 viscous modes per
degree, 550 steps of 1 yr (seeded synthetic Love numbers, topography to degree 256, a polar cap and
a retreating marine sector), FP64, extrapolated inner start, rotation and marine check on,
epsilon1 = 1e-5.  A "step" is one solver step (modal advance, inner fixed point to convergence,
radial displacement, commit).  All ice grids are uploaded before the timed region; W warm-up steps
(SeaLevelSolver.step), then K timed steps through the native step loop (SeaLevelSolver.run)
bracketed by device synchronisation.  Prints one JSON line.

Also reported: the roofline of the dominant kernel (the per-order Legendre analysis contraction),
timed live with HIP events on the stream it is launched on, with the HBM traffic per launch from
the committed rocprofv3 counter summary; the CPU oracle's steps/s on a bounded sample of the same
workload on this host; and, in a default run, the 1024 x 2048, lmax 1023 grid (240 x 0.5 kyr,
three ice caps) in FP64 under "also".

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lmax512|lmax1023|lmax63]
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's steps - warmup)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="lmax512", choices=("lmax63", "lmax512", "lmax1023"))
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--inner", default="extrap", choices=("off", "extrap", "anderson"),
                    help="inner start / acceleration (the metric is quoted with extrap)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only launch the dominant kernel (20x) and a 256 MiB device copy (for PMC counter runs)")
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


def cpu_baseline(w, ice, budget):
    """The FP64 NumPy oracle on this host: set-up, then as many steps as fit in `budget` seconds
    (at least 1) of the same workload; steps/s over the timed steps."""
    import numpy as np

    from oracle.love import LoveNumbers as OLove
    from oracle.sht import OracleSHT
    from oracle.solver import SeaLevelSolver as OSolver

    slv = OSolver(OracleSHT(w.nlat, w.nlon, w.lmax), OLove(w.love_text, w.lmax, text=True), w.topo, ice(0),
                  norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner="extrap")
    t0 = time.perf_counter()
    n = 0
    while True:
        slv.step(ice(n + 1))
        n += 1
        if time.perf_counter() - t0 > budget or n >= w.nsteps:
            break
    el = time.perf_counter() - t0
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    return {"value": n / el, "unit": "steps/s", "cores": int(min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))),
            "kind": "port",
            "sample": f"first {n} steps of the {w.name} workload (FP64 NumPy/SciPy oracle, BLAS threads "
                      f"{os.environ.get('OMP_NUM_THREADS', 'default')}), {el:.1f} s"}


def run_workload(w, warmup, steps, dev, inner="extrap"):
    """W warm-up steps, then K timed steps of workload w with every ice grid resident in HBM,
    bracketed by device synchronisation.  Returns (steps/s, seconds, K, numiter list, transform)."""
    import numpy as np
    import torch

    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    steps = steps if steps is not None else w.nsteps - warmup
    nsteps = warmup + steps
    if nsteps > w.nsteps:
        raise SystemExit(f"warmup + steps = {nsteps} exceeds the workload's {w.nsteps} steps")
    t0 = time.perf_counter()
    ice_dev = torch.empty((nsteps + 1, w.nlat, w.nlon), dtype=torch.float64, device=dev)
    for k in range(nsteps + 1):
        ice_dev[k].copy_(torch.from_numpy(w.ice(k)))
    print(f"[bench] {w.name}: inputs on the device ({ice_dev.numel() * 8 / 1e9:.2f} GB, "
          f"{time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    sht = SHT(w.nlat, w.nlon, w.lmax, device=dev)
    love = LoveNumbers(w.love_text, w.lmax, text=True)
    slv = OptSeaLevelSolver(sht, love, w.topo, w.ice(0), norder=w.lmax, dt=w.dt, tpw=True, checkmarine=True,
                            calcRG=True, epsilon1=1e-5, inner=inner)
    for k in range(1, warmup + 1):
        slv.step(ice_dev[k])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec = slv.run(ice_dev, warmup + 1, nsteps + 1)  # the native step loop
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    numiter = [r["numiter"] for r in rec]
    print(f"[bench] {w.name}: {steps} steps in {el:.3f} s, numiter mean {np.mean(numiter):.2f}", file=sys.stderr,
          flush=True)
    del slv, ice_dev
    torch.cuda.synchronize()
    return steps / el, el, steps, numiter, sht


def describe(w, inner="extrap"):
    return (f"{w.name}: {w.nlat}x{w.nlon} GL grid, lmax {w.lmax}, {w.nmodes} modes/degree, "
            f"{w.nsteps} x {w.dt:g} yr, FP64, inner {inner}, tpw+checkmarine on, eps1 1e-5")


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
    if a.kernels_only:
        dev = torch.device("cuda", 0)
        sht = SHT(w.nlat, w.nlon, w.lmax, device=dev)
        grids = torch.from_numpy(np.stack([w.topo, w.ice(1)])).to(dev)
        coef = torch.zeros((2, sht.ncoef), dtype=torch.complex128, device=dev)
        sht.analysis_packed(grids, coef)
        for _ in range(20):
            sht.analysis_contraction(2, coef)
        src = torch.ones(32 * 2**20, dtype=torch.float64, device=dev)  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"kernels_only": True, "copy_bytes": src.numel() * 8}))
        return
    dev = torch.device("cuda", 0)
    sps, el, steps, numiter, sht = run_workload(w, a.warmup, a.steps, dev, a.inner)
    # ---- dominant kernel: the per-order analysis contraction of a 2-field sweep, timed with HIP
    # events on the current stream (the stream the library launches on)
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
    out = {
        "metric": METRIC, "value": sps, "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * el / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": describe(w, a.inner), "numiter_mean": float(np.mean(numiter)),
                   "numiter_total": int(np.sum(numiter)), "ice": w.note},
        "roofline": roof,
        "mixed": "not provided: on this GPU FP32 vector math runs at the FP64 rate and the contraction is "
                 "bound by its FP64 recurrence, so reduced-precision sweeps would not be cheaper",
    }
    del sht
    torch.cuda.synchronize()
    if a.steps is None and a.workload == "lmax512":
        # the metric's second grid (lmax 1023, 1024 x 2048, 240 x 0.5 kyr), FP64, every step but 5
        w3 = syn.workload("lmax1023")
        sps3, el3, steps3, nit3, sht3 = run_workload(w3, 5, None, dev, a.inner)
        del sht3
        out["also"] = [{"workload": describe(w3, a.inner), "value": sps3, "unit": "steps/s", "steps": steps3, "warmup": 5,
                        "ms_per_step": 1e3 * el3 / steps3, "numiter_mean": float(np.mean(nit3)), "ice": w3.note}]
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, w.ice, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
