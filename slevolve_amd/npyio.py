"""``.npy`` conventions of the reference's solvers (python/slevolve/npyio.py:1-14, 61-103, 119-126,
180-193): grids (nlat, nlon) '<f8' C order, rows north -> south; spectral fields (N+1, N+1) '<c16'
written fortran_order=True; masks '<f8' (Fortran-compatible) or uint8 (legacy)."""
from __future__ import annotations

import os

import numpy as np


def _host(a):
    try:
        import torch

        if isinstance(a, torch.Tensor):
            return a.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(a)


def save_grid(path, a, dtype="f8"):
    a = np.ascontiguousarray(_host(a), dtype=np.float64)
    if dtype == "f4" and np.array_equal(a.astype(np.float32).astype(np.float64), a):
        np.save(path, a.astype("<f4"))
        return "f4"
    np.save(path, a.astype("<f8", copy=False))
    return "f8"


def save_mask(path, m, io_compat="fortran"):
    m = _host(m)
    if io_compat == "fortran":
        np.save(path, np.ascontiguousarray(m, dtype="<f8"))
    else:
        np.save(path, np.asarray(m).astype(np.uint8))


def save_spectral(path, z):
    np.save(path, np.asfortranarray(np.asarray(_host(z), dtype="<c16")))


def read_text_grid(path, nlat, nlon):
    a = np.fromfile(path, sep=" ")
    if a.size != nlat * nlon:
        raise ValueError(f"{path}: got {a.size} values, expected {nlat * nlon}")
    return a.reshape(nlat, nlon)


def read_grid(path, nlat, nlon):
    """.npy (any real dtype), headerless little-endian f8 .bin, or text; without an extension
    <path>.npy, then <path>.bin, then the text file."""
    if not path.endswith((".npy", ".bin")):
        for ext in (".npy", ".bin"):
            if os.path.exists(path + ext):
                return read_grid(path + ext, nlat, nlon)
        return read_text_grid(path, nlat, nlon)
    if path.endswith(".npy"):
        a = np.load(path, allow_pickle=False)
        if a.shape != (nlat, nlon):
            if a.size != nlat * nlon:
                raise ValueError(f"{path}: shape {a.shape}, expected {(nlat, nlon)}")
            a = a.reshape(nlat, nlon)
        return np.ascontiguousarray(a, dtype=np.float64)
    n = os.path.getsize(path)
    if n != 8 * nlat * nlon:
        raise ValueError(f"{path}: {n} bytes, expected {8 * nlat * nlon} (LE f8)")
    return np.fromfile(path, dtype="<f8").reshape(nlat, nlon)


class IceDir:
    """ice[k] read on demand from <dir>/<pfx><k>{.npy,.bin,} (npy preferred)."""

    def __init__(self, d, pfx, nlat, nlon):
        self.d, self.pfx, self.nlat, self.nlon = d, pfx, nlat, nlon

    def __getitem__(self, k):
        return read_grid(os.path.join(self.d, f"{self.pfx}{k}"), self.nlat, self.nlon)


def read_times(path, nsteps=None):
    t = np.loadtxt(path, dtype=np.float64, ndmin=1)
    if t.ndim != 1 or t.size < 2:
        raise ValueError(f"{path}: need one time per line, at least 2 values")
    if nsteps is not None:
        if t.size < nsteps + 1:
            raise ValueError(f"{path}: {t.size} times, need {nsteps + 1} (files 0..{nsteps})")
        t = t[: nsteps + 1]
    if not np.all(np.isfinite(t)) or not np.all(np.diff(t) > 0):
        raise ValueError(f"{path}: times must be finite and strictly increasing")
    return t
