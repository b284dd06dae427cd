"""Step I/O around the device solver: the ice history resident in HBM and per-step snapshot files
written off the step loop.

``IceStack`` holds ice files 0..n as one [n+1][nlat][nlon] float64 device tensor (all 551 grids
of a 512 x 1024 run are 2.3 GB), filled from the ice cache of run_ais.py:180-227 (built or
reused through its manifest) or straight from the ice directory, in chunks through one pinned
host buffer.  The step loop then reads no host memory.

``SnapshotWriter`` writes the per-step outputs of run_ais.py:400-425 (tgrid<N>.npy, ocean<N>,
beta<N>, R<N>, G<N>, dS_converged<N>, dSLdt<N>/dRdt<N>/TPWrate<N>.txt, horiz<N>) in the file
conventions of ``npyio``.  The solver overwrites its output buffers at the next step, so each
snapshot is first copied device-to-device into a ring slot on the compute stream (a few
microseconds per 4 MB grid); the slot is copied to pinned host memory on a separate copy stream
and a writer thread saves the files once that copy has completed.  The compute stream never waits
for PCIe or the file system; when all slots are in flight, ``write`` blocks until one is free.
"""
from __future__ import annotations

import os
import queue
import threading

import numpy as np
import torch

from . import npyio

# snapshot field -> (file stem, npyio kind); masks follow io_compat
_FIELDS = {
    "tgrid": ("tgrid", "grid"),
    "ocean": ("ocean", "mask"),
    "beta": ("beta", "mask"),
    "R": ("R", "grid"),
    "G": ("G", "grid"),
    "dS": ("dS_converged", "spectral"),
}


class IceStack:
    """Ice files 0..nsteps of <icedir>/<pfx><k> on the device.  With cache_dir the stack is read
    from (or first written to) the memory-mapped ice cache; without it, file by file."""

    def __init__(self, icedir, pfx, nsteps, nlat=npyio.NLAT, nlon=npyio.NLON, device=None, cache_dir=None,
                 rebuild=False, chunk=16):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.cache_path, self.cache_seconds = None, 0.0
        n = nsteps + 1
        if cache_dir is not None:
            self.cache_path, self.cache_seconds = npyio.build_ice_cache(icedir, pfx, nsteps, nlat, nlon, cache_dir,
                                                                        rebuild)
            src = np.load(self.cache_path, mmap_mode="r")
            get = lambda k: src[k]  # noqa: E731
        else:
            ice = npyio.IceDir(icedir, pfx, nlat, nlon)
            get = ice.__getitem__
        self.grids = torch.empty((n, nlat, nlon), dtype=torch.float64, device=self.device)
        chunk = max(1, min(int(chunk), n))
        stage = torch.empty((chunk, nlat, nlon), dtype=torch.float64, pin_memory=True)
        for k0 in range(0, n, chunk):
            k1 = min(n, k0 + chunk)
            for k in range(k0, k1):
                stage[k - k0].numpy()[...] = get(k)
            self.grids[k0:k1].copy_(stage[: k1 - k0], non_blocking=True)
            torch.cuda.current_stream(self.device).synchronize()  # the staging buffer is reused

    def __len__(self):
        return self.grids.shape[0]

    def __getitem__(self, k):
        return self.grids[k]


class SnapshotWriter:
    """Asynchronous per-step snapshot files.  ``fields`` as run_ais.py --fields: any of tgrid,
    ocean, beta, R, G, dS, dSLdt, dRdt (dSLdt/dRdt also write TPWrate<N>.txt when the step has
    rotational rates); results with "horiz" write horiz<N>.npy.  ``depth`` ring slots."""

    def __init__(self, outdir, fields=("tgrid",), io_compat="fortran", depth=3, device=None):
        self.outdir = outdir
        os.makedirs(outdir, exist_ok=True)
        self.fields = set(fields)
        self.io_compat = io_compat
        self.depth = int(depth)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.copy_stream = torch.cuda.Stream(self.device)
        self._free = queue.Queue()
        for i in range(self.depth):
            self._free.put(i)
        self._slots = [dict() for _ in range(self.depth)]  # name -> (device tensor, pinned host tensor)
        self._todo = queue.Queue()
        self._error = None
        self.files_written = 0
        self._thread = threading.Thread(target=self._run, daemon=True)
        self._thread.start()

    # ------------------------------------------------------------------------ producer
    def _items(self, k, r):
        """(slot key, file name, kind, value) of the snapshot of step k from the step result r."""
        out = []
        for f in ("tgrid", "ocean", "beta", "R", "G", "dS"):
            key = "dS_converged" if f == "dS" else f
            if f in self.fields and r.get(key) is not None:
                stem, kind = _FIELDS[f]
                if kind == "mask" and self.io_compat != "fortran":
                    kind = "mask_legacy"
                if kind == "spectral" and self.io_compat != "fortran":
                    kind = "spectral_c"
                out.append((f, f"{stem}{k}.npy", kind, r[key]))
        rates = r.get("rates")
        if rates is not None:
            for f in ("dSLdt", "dRdt"):
                if f in self.fields:
                    out.append((f, f"{f}{k}.npy", "grid", rates[f]))
            if "mdot" in rates and self.fields & {"dSLdt", "dRdt"}:
                out.append(("tpw", f"TPWrate{k}.txt", "tpw", {q: rates[q] for q in ("ildot", "mdot", "lambdadot")}))
        if r.get("horiz") is not None:
            out.append(("horiz", f"horiz{k}.npy", "spectral", r["horiz"]))
        return out

    def write(self, k, r):
        """Enqueue the snapshot of step k (call right after the step, before the next one)."""
        self._raise()
        items = self._items(k, r)
        if not items:
            return
        slot = self._free.get()
        self._raise()
        bufs = self._slots[slot]
        compute = torch.cuda.current_stream(self.device)
        staged = []
        for key, name, kind, v in items:
            if isinstance(v, torch.Tensor) and v.is_cuda:
                pair = bufs.get(key)
                if pair is None or pair[0].shape != v.shape or pair[0].dtype != v.dtype:
                    pair = (torch.empty_like(v, memory_format=torch.contiguous_format),
                            torch.empty(v.shape, dtype=v.dtype, pin_memory=True))
                    bufs[key] = pair
                pair[0].copy_(v)  # device to device on the compute stream
                staged.append((name, kind, pair))
            else:
                staged.append((name, kind, v if isinstance(v, dict) else np.array(npyio._host(v))))
        ev = torch.cuda.Event()
        ev.record(compute)
        with torch.cuda.stream(self.copy_stream):
            self.copy_stream.wait_event(ev)
            for name, kind, v in staged:
                if isinstance(v, tuple):
                    v[1].copy_(v[0], non_blocking=True)
            done = torch.cuda.Event()
            done.record(self.copy_stream)
        self._todo.put((slot, staged, done))

    # ------------------------------------------------------------------------ consumer
    def _run(self):
        while True:
            job = self._todo.get()
            if job is None:
                return
            slot, staged, done = job
            try:
                done.synchronize()
                for name, kind, v in staged:
                    path = os.path.join(self.outdir, name)
                    host = v[1].numpy() if isinstance(v, tuple) else v
                    if kind == "tpw":
                        tag = name[len("TPWrate"):-len(".txt")]
                        npyio.save_rates(self.outdir, tag, host, fields=())
                    elif kind == "spectral_c":
                        np.save(path, np.ascontiguousarray(host))
                    else:
                        npyio.save(path, kind, host)
                    self.files_written += 1
            except BaseException as e:  # reported to the producer
                self._error = e
            finally:
                self._free.put(slot)

    def _raise(self):
        if self._error is not None:
            e, self._error = self._error, None
            raise RuntimeError(f"snapshot writer failed: {e!r}") from e

    def flush(self):
        """Wait until every enqueued snapshot is on disk."""
        for _ in range(self.depth):
            self._free.get()
        for i in range(self.depth):
            self._free.put(i)
        self._raise()

    def close(self):
        self.flush()
        self._todo.put(None)
        self._thread.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
