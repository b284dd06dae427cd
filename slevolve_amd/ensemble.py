"""Ensembles of independent forward runs on one GPU (the batched configuration: 64 runs with
lmax drawn from {128, 256, 512}, 50-1200 steps, dt in {0.2, 0.5, 1.0} kyr, per-run Love numbers
with decay rates scaled by 10^U(-0.3, 0.3), per-run ice caps).  The reference runs them one after
another (run_ais.py:393-438 per run; run_outer.py:148-278 re-runs the same loop).

Design.  A single run of lmax 128-512 does not fill the GPU: its contraction launches have tens
to a few hundred blocks and its step is a chain of short dependent launches.  The ensemble
therefore keeps several runs in flight at once, each run on its own HIP stream with the native
step loop (slx_run_steps, which holds the host only while it waits for its own stream's posted
decisions), so that the device interleaves the launches of different runs.  Every run's inputs
(all its ice grids) are generated on the device before the timed region; runs are dealt to the
streams largest first (lmax^3 x steps).  Per lmax the transform handles are created once per
stream and reused by the runs that stream takes (a handle's work buffers belong to one stream
at a time).  A run's results do not depend on which stream or when it runs: its kernels and
their summation order are those of the run alone.

Autotuning: per lmax, the contraction kernel variants (analysis: one group per block or
persistent, 2 or 4 rings per lane; synthesis: matrix cores with 2 or 4 rings per lane, or
scalar-fed) are timed on the device once (tune_variants) and every handle of that lmax is then
created with the winners, so that all members of one lmax sum in the same order; the number of
concurrent streams is picked from {1, 2, 4, 8} by timing the first steps of a probe batch that
holds the largest members of every lmax (autotune_streams).
"""
from __future__ import annotations

import math
import os
import queue
import threading
import time

import numpy as np
import torch

from . import synthetic as syn

LMAX_CHOICES = (128, 256, 512)
DT_CHOICES = (200.0, 500.0, 1000.0)  # years (0.2, 0.5, 1.0 kyr)


class RunSpec:
    """One member: truncation, grid (lmax x 2 lmax), steps, step length, Love text, topography and
    its ice caps [(lat, lon, radius_deg, H0_m)] growing and melting as sin(pi t / T)."""

    def __init__(self, rid, lmax, nsteps, dt, love_text, topo, caps):
        self.id, self.lmax, self.nsteps, self.dt = rid, int(lmax), int(nsteps), float(dt)
        self.nlat, self.nlon = self.lmax, 2 * self.lmax
        self.love_text, self.topo, self.caps = love_text, topo, caps

    @property
    def cost(self):
        return self.lmax ** 3 * self.nsteps

    def ice_stack(self, dev):
        """All nsteps + 1 ice grids on the device (mm-rounded)."""
        lat, lon = syn.gl_lat_lon(self.nlat)
        LON, LAT = np.meshgrid(lon, lat)
        shapes = []
        for la, lo, R, H in self.caps:
            d = syn.angular_distance(LAT, LON, la, lo)
            shapes.append(H * np.sqrt(np.clip(1.0 - (d / R) ** 2, 0.0, None)))
        sh = torch.from_numpy(np.stack(shapes)).to(dev)  # [ncaps, nlat, nlon]
        f = torch.sin(math.pi * torch.arange(self.nsteps + 1, dtype=torch.float64, device=dev) / self.nsteps)
        stack = torch.einsum("t,cij->tij", f, sh)
        return torch.round(stack * 1000.0) / 1000.0


def perturbed_love_text(lmax, nmodes, seq_love, seq_tidal, seq_jitter):
    """synthetic_love_text with every relaxation rate scaled by 10^U(-0.3, 0.3)."""
    rj = np.random.default_rng(seq_jitter)
    base = np.random.default_rng(seq_love)
    rt = np.random.default_rng(seq_tidal)
    s, el, amp, tel, tamp = [None], [None], [None], [None], [None]
    for l in range(1, lmax + 1):
        e = np.array([-1.0 / (1 + l / 20.0), 0.05 / (1 + l / 30.0), -0.3 / (1 + l / 10.0)])
        rates = 1000.0 * 10.0 ** base.uniform(-5.0, -1.0, nmodes) * 10.0 ** rj.uniform(-0.3, 0.3, nmodes)
        s.append(np.sort(rates))
        el.append(e)
        amp.append(base.uniform(0.0, 1.0, (3, nmodes)) * e[:, None] / nmodes)
        tel.append(e)
        tamp.append(rt.uniform(0.0, 1.0, (3, nmodes)) * e[:, None] / nmodes)
    return syn.love_file_text(s, el, amp, tel, tamp, header=f"ensemble lmax {lmax} modes {nmodes}")


def make_ensemble(n_runs=64, lmax_choices=LMAX_CHOICES, steps=(50, 1200), dt_choices=DT_CHOICES, nmodes=30,
                  seed=20261203 + 4):
    """The seeded ensemble: per run lmax, steps U[steps], dt, Love numbers, 1-3 caps with centres,
    radii U(8, 20) deg and H0 U(1, 4) km; one topography per lmax (seeded, to degree lmax / 4)."""
    seqs = np.random.SeedSequence(seed).spawn(5)  # love, tidal, (unused), members, rate jitter
    rng = np.random.default_rng(seqs[3])
    topo = {}
    for L in sorted(set(lmax_choices)):
        topo[L] = syn.scaled_topography(L, max(8, L // 4), np.random.default_rng(np.random.SeedSequence([seed, L])))
    specs = []
    love_seqs = seqs[0].spawn(n_runs)
    tidal_seqs = seqs[1].spawn(n_runs)
    jit_seqs = seqs[4].spawn(n_runs)
    for r in range(n_runs):
        L = int(rng.choice(lmax_choices))
        n = int(rng.integers(steps[0], steps[1] + 1))
        dt = float(rng.choice(dt_choices))
        caps = []
        for _ in range(int(rng.integers(1, 4))):
            caps.append((float(rng.uniform(-85.0, 85.0)), float(rng.uniform(0.0, 360.0)), float(rng.uniform(8.0, 20.0)),
                         float(rng.uniform(1000.0, 4000.0))))
        love = perturbed_love_text(L, nmodes, love_seqs[r], tidal_seqs[r], jit_seqs[r])
        specs.append(RunSpec(r, L, n, dt, love, topo[L], caps))
    return specs


_ENV_LOCK = threading.Lock()


def _variant_names(sht):
    """The SLX_ANA_VARIANT / SLX_SYN_VARIANT names of the kernels a handle runs."""
    k = sht.kernels
    ana = {"k_leg_ana_mf": "mf", "k_leg_ana_ck": "ck", "k_leg_ana_qw": "qw"}.get(k["analysis"])
    syn_v = {"k_leg_syn_s": "scalar", "k_leg_syn_mf": "mfma", "k_leg_syn_mw": "mw"}[k["synthesis"]]
    if syn_v != "scalar":
        syn_v += str(k["synthesis_rings_per_lane"])
    return (f"{ana}{k['analysis_rings_per_lane']}" if ana else None), syn_v


def tune_variants(lmaxes, dev=None):
    """Per lmax: the contraction variants a handle picks by timing them on the device
    (SLX_ANA_VARIANT = SLX_SYN_VARIANT = auto); returns {lmax: (analysis, synthesis) names}."""
    from .sht import SHT

    dev = dev or torch.device("cuda", torch.cuda.current_device())
    out = {}
    for L in sorted(set(lmaxes)):
        with _ENV_LOCK:
            old = {k: os.environ.get(k) for k in ("SLX_ANA_VARIANT", "SLX_SYN_VARIANT")}
            os.environ["SLX_ANA_VARIANT"] = os.environ["SLX_SYN_VARIANT"] = "auto"
            try:
                out[L] = _variant_names(SHT(L, 2 * L, L, device=dev))
            finally:
                for k, v in old.items():
                    if v is None:
                        os.environ.pop(k, None)
                    else:
                        os.environ[k] = v
    return out


class _Stream:
    """A HIP stream and its transform handles (one per lmax, with the variants given per lmax)."""

    def __init__(self, dev, variants=None):
        self.dev = dev
        self.stream = torch.cuda.Stream(dev)
        self.sht = {}
        self.variants = variants or {}

    def handle(self, lmax):
        from .sht import SHT

        if lmax not in self.sht:
            ana, syn_v = self.variants.get(lmax, (None, None))
            with _ENV_LOCK:
                old = {k: os.environ.get(k) for k in ("SLX_ANA_VARIANT", "SLX_SYN_VARIANT")}
                for k, v in (("SLX_ANA_VARIANT", ana), ("SLX_SYN_VARIANT", syn_v)):
                    if v is None:
                        os.environ.pop(k, None)
                    else:
                        os.environ[k] = v
                try:
                    with torch.cuda.stream(self.stream):
                        self.sht[lmax] = SHT(lmax, 2 * lmax, lmax, device=self.dev)
                finally:
                    for k, v in old.items():
                        if v is None:
                            os.environ.pop(k, None)
                        else:
                            os.environ[k] = v
        return self.sht[lmax]


def run_one(spec, ice, slot, epsilon1=1e-5, inner="extrap", k1=None):
    """One member on the slot's stream through the native loop; (seconds, records, tgrid)."""
    from .love import LoveNumbers
    from .solver import OptSeaLevelSolver

    sht = slot.handle(spec.lmax)
    with torch.cuda.stream(slot.stream):
        slv = OptSeaLevelSolver(sht, LoveNumbers(spec.love_text, spec.lmax, text=True), spec.topo, ice[0],
                                norder=spec.lmax, dt=spec.dt, epsilon1=epsilon1, inner=inner)
        slot.stream.synchronize()
        t0 = time.perf_counter()
        rec = slv.run(ice, 1, (spec.nsteps if k1 is None else k1) + 1)
        slot.stream.synchronize()
        el = time.perf_counter() - t0
        tgrid = slv.topoxy.clone()
        slot.stream.synchronize()
    return el, rec, tgrid


def run_ensemble(specs, ices, nstreams=4, slots=None, dev=None, variants=None):
    """Every member, largest first, on nstreams concurrent streams.  Returns (wall seconds, {id:
    (seconds in its stream, records, tgrid)}, slots)."""
    dev = dev or torch.device("cuda", torch.cuda.current_device())
    slots = slots or [_Stream(dev, variants) for _ in range(nstreams)]
    todo = queue.Queue()
    for s in sorted(specs, key=lambda s: -s.cost):
        todo.put(s)
    out, errors = {}, []

    def work(slot):
        torch.cuda.set_device(dev)
        while True:
            try:
                s = todo.get_nowait()
            except queue.Empty:
                return
            try:
                out[s.id] = run_one(s, ices[s.id], slot)
            except Exception as e:  # reported by the caller
                errors.append((s.id, e))
                return

    for slot in slots[:nstreams]:  # handles first, outside the timed region
        for L in sorted({s.lmax for s in specs}):
            slot.handle(L)
    torch.cuda.synchronize(dev)
    threads = [threading.Thread(target=work, args=(slot,)) for slot in slots[:nstreams]]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    if errors:
        raise RuntimeError(f"ensemble member {errors[0][0]} failed: {errors[0][1]}")
    return wall, out, slots


def autotune_streams(specs, ices, candidates=(1, 2, 4, 8), probe_steps=10, dev=None, variants=None):
    """The stream count, from candidates, with the best throughput on the first probe_steps steps
    of a probe batch (the largest members, up to the largest candidate, of every lmax); returns
    (best count, {count: probe steps/s})."""
    dev = dev or torch.device("cuda", torch.cuda.current_device())
    by_l = {}
    for s in sorted(specs, key=lambda s: -s.cost):
        by_l.setdefault(s.lmax, []).append(s)
    group = [s for L in by_l for s in by_l[L][: max(candidates)]]
    probe = [RunSpec(s.id, s.lmax, min(s.nsteps, probe_steps), s.dt, s.love_text, s.topo, s.caps) for s in group]
    pi = {s.id: ices[s.id][: probe_steps + 1] for s in probe}
    slots = [_Stream(dev, variants) for _ in range(max(candidates))]
    rates = {}
    for n in candidates:
        wall, _, _ = run_ensemble(probe, pi, nstreams=n, slots=slots, dev=dev)
        rates[n] = sum(p.nsteps for p in probe) / wall
    return max(rates, key=rates.get), rates


def prepare(specs, dev=None):
    dev = dev or torch.device("cuda", torch.cuda.current_device())
    return {s.id: s.ice_stack(dev) for s in specs}


__all__ = ["RunSpec", "make_ensemble", "run_ensemble", "run_one", "autotune_streams", "tune_variants", "prepare",
           "perturbed_love_text"]
