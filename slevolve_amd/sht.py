"""Gauss-Legendre spherical-harmonic transform on the GPU, with the interface of the reference's
``SHT`` (python/slevolve/sht.py:148-343) and ``make_sht`` (python/slevolve/sht_ducc.py:206-217).

``analysis(grid[nlat, nlon] float64, out=None) -> coeffs[lmax+1, lmax+1] complex128`` ([l, m],
zero above the diagonal) and ``synthesis(coeffs, out=None) -> grid``.  ``raw=True`` exchanges
V[l, m] = (-1)^m U[l, m] instead, as the ducc backend's raw mode does (sht_ducc.py:101-113).
``precision="mixed"`` makes these two calls use the FP32-product contractions (FP64 recurrence,
FP32 matrix-core products; about 1e-6 relative); the default is FP64.  Inputs may be NumPy arrays
(results come back as NumPy arrays) or torch tensors (results stay on the device).  Without
``out`` a persistent internal buffer is returned, as in the reference.  The work runs in the HIP
kernels of libslx.so (ring FFTs + per-order FP64 Legendre contractions); there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _device as D
from . import _lib

SQRT2_LEGACY = 1.41421356


class SHT:
    def __init__(self, nlat: int, nlon: int, lmax: int, workers: int = 1, chunk: int = 16,
                 device=None, max_fields: int = 3, raw: bool = False, precision: str = "fp64"):
        if nlat % 2 != 0:
            raise ValueError("nlat must be even")
        if lmax > nlon // 2:
            raise ValueError("lmax must be <= nlon/2")
        if nlon & (nlon - 1):
            raise ValueError("nlon must be a power of two")
        self.nlat, self.nlon, self.lmax = int(nlat), int(nlon), int(lmax)
        self.nm = self.lmax + 1
        self.nh = self.nlat // 2
        self.workers = self.nthreads = int(workers)
        self.chunk = int(chunk)
        self.max_fields = int(max_fields)
        self.raw = bool(raw)
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}, got {precision!r}")
        self.precision = precision
        self.dev = D.device(device)
        lib = _lib.load()
        h = C.c_void_p()
        with torch.cuda.device(self.dev):
            _lib.check(lib.slx_sht_create(self.nlat, self.nlon, self.lmax, _lib.PRECISIONS[precision],
                                          self.max_fields, C.byref(h)), "slx_sht_create")
        self._h = h
        dims = (C.c_int * 9)()
        x = np.zeros(self.nlat)
        w = np.zeros(self.nlat)
        _lib.check(lib.slx_sht_info(h, dims, x.ctypes.data_as(_lib.P_double), w.ctypes.data_as(_lib.P_double)))
        self.ncoef = int(dims[3])
        self.gaulats, self.weights = x, w
        # kernels the handle chose at creation (a fixed rule per shape unless SLX_ANA_VARIANT /
        # SLX_SYN_VARIANT ask otherwise)
        ana = {-1: "k_leg_ana", 0: "k_leg_ana_mf", 1: "k_leg_ana_ck", 2: "k_leg_ana_qw"}[int(dims[5])]
        self.kernels = {"analysis": ana, "analysis_rings_per_lane": int(dims[6]),
                        "synthesis": {0: "k_leg_syn_s", 1: "k_leg_syn_mf", 2: "k_leg_syn_mw"}[int(dims[7])],
                        "synthesis_rings_per_lane": int(dims[8])}
        self._tri = torch.zeros((self.max_fields, self.ncoef), dtype=torch.complex128, device=self.dev)
        self._coef_out = torch.zeros((self.nm, self.nm), dtype=torch.complex128, device=self.dev)
        self._grid_out = torch.zeros((self.nlat, self.nlon), dtype=torch.float64, device=self.dev)
        self._msign = None
        if self.raw:  # (-1)^m per column of the dense [l, m] layout
            sg = torch.ones(self.nm, dtype=torch.complex128, device=self.dev)
            sg[1::2] = -1.0
            self._msign = sg

    @property
    def handle(self):
        return self._h

    @property
    def table_bytes(self) -> int:
        """Bytes of the on-device tables (nodes, weights, P~_mm, recurrence coefficients); the
        Legendre functions themselves are regenerated on the fly."""
        return 8 * (3 * self.nh + self.nm * self.nh + 2 * self.ncoef + self.nm) + 16 * self.nlon

    def mean00_weights(self) -> np.ndarray:
        """Row weights with ``analysis(g)[0, 0] == w @ g.sum(axis=1)`` (sht_ducc.py:190-203)."""
        return self.weights / self.nlon * np.sqrt(0.5) / SQRT2_LEGACY

    # ------------------------------------------------------------------ packed interface
    def analysis_packed(self, grids: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """grids (nf, nlat, nlon) float64 device -> out (nf, ncoef) complex128 packed."""
        nf = grids.shape[0]
        _lib.call("slx_sht_analysis", self._h, grids.data_ptr(), nf, out.data_ptr(), D.stream_ptr())
        return out

    def analysis_contraction(self, nf: int, out: torch.Tensor) -> torch.Tensor:
        """Re-run only the per-order contraction of the last analysis of nf fields."""
        _lib.call("slx_sht_analysis_contraction", self._h, nf, out.data_ptr(), D.stream_ptr())
        return out

    def synthesis_packed(self, coefs: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        nf = coefs.shape[0]
        _lib.call("slx_sht_synthesis", self._h, coefs.data_ptr(), nf, out.data_ptr(), D.stream_ptr())
        return out

    def to_dense(self, tri: torch.Tensor, out: torch.Tensor | None = None, raw: bool | None = None) -> torch.Tensor:
        """Packed U triangle -> dense [l, m] (V = (-1)^m U for a raw-mode handle unless raw=False)."""
        if out is None:
            out = torch.empty((self.nm, self.nm), dtype=torch.complex128, device=self.dev)
        _lib.call("slx_sht_tri_to_dense", self._h, tri.data_ptr(), out.data_ptr(), 1, D.stream_ptr())
        if (self.raw if raw is None else raw) and self._msign is not None:
            out.mul_(self._msign)
        return out

    def to_packed(self, dense: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Dense [l, m] (V for a raw-mode handle) -> packed U triangle."""
        if out is None:
            out = torch.empty((self.ncoef,), dtype=torch.complex128, device=self.dev)
        if self.raw:
            dense = dense * self._msign
        _lib.call("slx_sht_dense_to_tri", self._h, dense.data_ptr(), out.data_ptr(), 1, D.stream_ptr())
        return out

    # --------------------------------------------------------------- reference interface
    def analysis(self, grid, out=None):
        host = not isinstance(grid, torch.Tensor)
        g = D.as_device(grid, torch.float64, self.dev, (self.nlat, self.nlon))
        tri = self._tri[0]
        self.analysis_packed(g.view(1, self.nlat, self.nlon), tri.view(1, -1))
        if host:
            res = self.to_dense(tri).cpu().numpy()
            if out is not None:
                out[...] = res
                return out
            return res
        dst = self._coef_out if out is None else out
        if dst.dtype != torch.complex128 or not dst.is_contiguous() or dst.device != self.dev:
            raise ValueError("out must be a contiguous complex128 tensor on the transform's device")
        return self.to_dense(tri, dst)

    def synthesis(self, coeffs, out=None):
        host = not isinstance(coeffs, torch.Tensor)
        c = D.as_device(coeffs, torch.complex128, self.dev, (self.nm, self.nm))
        tri = self.to_packed(c, self._tri[0])
        if host:
            g = self.synthesis_packed(tri.view(1, -1), self._grid_out.view(1, self.nlat, self.nlon))
            res = g.cpu().numpy()
            if out is not None:
                out[...] = res
                return out
            return res
        dst = self._grid_out if out is None else out
        if dst.dtype != torch.float64 or not dst.is_contiguous() or dst.device != self.dev:
            raise ValueError("out must be a contiguous float64 tensor on the transform's device")
        return self.synthesis_packed(tri.view(1, -1), dst.view(1, self.nlat, self.nlon)).view(self.nlat, self.nlon)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().slx_sht_destroy(h)
            except Exception:
                pass
            self._h = None


def make_sht(backend: str, nlat: int, nlon: int, lmax: int, workers: int = 1, chunk: int = 16,
             raw: bool = False, device=None, precision: str = "fp64"):
    """Factory with the reference's signature (sht_ducc.py:206-217); every backend name maps to the
    GPU transform.  raw=True (the ducc backend's raw mode) exchanges V = (-1)^m U; as in the
    reference it is refused for the "numpy" backend."""
    if backend not in ("numpy", "ducc", "hip", "gpu"):
        raise ValueError(f"unknown SHT backend {backend!r} (expected 'numpy', 'ducc' or 'hip')")
    if raw and backend == "numpy":
        raise ValueError("raw mode needs the ducc backend")
    return SHT(nlat, nlon, lmax, workers=workers, chunk=chunk, device=device, raw=raw, precision=precision)


__all__ = ["SHT", "make_sht", "SQRT2_LEGACY"]
