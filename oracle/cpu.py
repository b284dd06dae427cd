"""Optimised CPU baseline of the oracle: the oracle's solver step with its transforms and modal
advance in multi-threaded C (oracle/cpu_kernels.c, OpenMP) and its ring FFTs in scipy.fft with
worker threads.  Test infrastructure only (see ``oracle/__init__.py``): bench.py times it as the
CPU baseline beside the GPU path; nothing in the product imports it.

It follows the reference's optimised CPU algorithm where that matters for time: on-the-fly
Legendre recurrence with N-S symmetry (python/slevolve/sht_ducc.py:206-217, fortran/src/
otf_kernels.inc:5-16) instead of stored tables, and threads over orders.  The step logic is the
oracle's (oracle/solver.py), so results equal the oracle's to rounding (the sums over rings run in
another order).

Build: ``python -c "import oracle.cpu as c; c.build()"`` (gcc, OpenMP), into oracle/_cpu/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.fft as sfft

from .memory import ExactMemory
from .sht import SQRT2_LEGACY, gauss_legendre, sectoral_starts
from .solver import SeaLevelSolver

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_cpu", "liboracle_cpu.so")
SRC = os.path.join(HERE, "cpu_kernels.c")
_lib = None


def build():
    """Compile oracle/cpu_kernels.c (x86-64-v3: AVX2 and FMA, which the GPU hosts' EPYC CPUs have)."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-o", tmp, SRC, "-lm"],
                   check=True)
    os.replace(tmp, LIB)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (oracle.cpu.build())")
        lib = C.CDLL(LIB)
        P, I = C.c_void_p, C.c_int
        lib.cpu_legendre_analysis.argtypes = [I, I, P, P, P, P, P, P, C.c_double]
        lib.cpu_legendre_synthesis.argtypes = [I, I, I, P, P, P, P, P, P]
        lib.cpu_modal_advance.argtypes = [I, I, P, P, P, P, P, P, P]
        lib.cpu_threads.restype = I
        _lib = lib
    return _lib


def threads() -> int:
    """OpenMP threads the kernels use (OMP_NUM_THREADS, else the cores available)."""
    return int(load().cpu_threads())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class FastCpuSHT:
    """The oracle's transform (OracleSHT conventions: dense [l, m] coefficients, legacy sqrt 2
    factor) with on-the-fly Legendre values."""

    def __init__(self, nlat: int, nlon: int, lmax: int, workers: int | None = None):
        if nlat % 2:
            raise ValueError("nlat must be even")
        if lmax > nlon // 2:
            raise ValueError("lmax must be <= nlon/2")
        self.nlat, self.nlon, self.lmax = int(nlat), int(nlon), int(lmax)
        self.nh = self.nlat // 2
        self.workers = workers or threads()
        self.gaulats, self.weights = gauss_legendre(self.nlat)
        x = np.ascontiguousarray(self.gaulats[: self.nh])
        p0, sc = sectoral_starts(x, self.lmax)
        self._x = x
        self._pmm = np.ascontiguousarray(np.stack(p0))
        self._pms = np.ascontiguousarray(np.stack(sc).astype(np.int32))
        self._wn = self.weights[: self.nh] / self.nlon
        self._ws = self.weights[::-1][: self.nh] / self.nlon
        load()

    def analysis(self, grid, out=None):
        nh, L = self.nh, self.lmax
        F = sfft.rfft(np.asarray(grid, dtype=np.float64), axis=1, workers=self.workers)[:, : L + 1]
        north = F[:nh] * self._wn[:, None]
        south = F[::-1][:nh] * self._ws[:, None]
        S = np.ascontiguousarray((north + south).T)
        D = np.ascontiguousarray((north - south).T)
        U = np.zeros((L + 1, L + 1), dtype=np.complex128)
        _lib.cpu_legendre_analysis(L, nh, _p(self._x), _p(self._pmm), _p(self._pms), _p(S), _p(D), _p(U),
                                   1.0 / SQRT2_LEGACY)
        if out is not None:
            out[...] = U
            return out
        return U

    def synthesis(self, coeffs, out=None):
        nh, L = self.nh, self.lmax
        c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.complex128) * (SQRT2_LEGACY * self.nlon))
        nf = self.nlon // 2 + 1
        FN = np.zeros((nh, nf), dtype=np.complex128)
        FS = np.zeros((nh, nf), dtype=np.complex128)
        _lib.cpu_legendre_synthesis(L, nh, nf, _p(self._x), _p(self._pmm), _p(self._pms), _p(c), _p(FN), _p(FS))
        F = np.empty((self.nlat, nf), dtype=np.complex128)
        F[:nh] = FN
        F[self.nlat - 1 - np.arange(nh)] = FS
        g = sfft.irfft(F, n=self.nlon, axis=1, workers=self.workers)
        if out is not None:
            out[...] = g
            return out
        return g

    def mean00_weights(self) -> np.ndarray:
        return self.weights / self.nlon / np.sqrt(2.0) / SQRT2_LEGACY


class FastExactMemory(ExactMemory):
    """The oracle's modal state with the dense advance-and-read in C (same operations per element);
    the degree-2 rotational states as in ExactMemory."""

    def advance_and_read(self):
        nl = self.lmax + 1
        visc = np.zeros((nl, nl), dtype=np.complex128)
        rad = np.zeros((nl, nl), dtype=np.complex128)
        d = np.ascontiguousarray(self.decay)
        _lib.cpu_modal_advance(self.kmax, nl, _p(self.R), _p(self.p_load), _p(d), _p(np.ascontiguousarray(self.a_sl)),
                               _p(np.ascontiguousarray(self.a_rr)), _p(visc), _p(rad))
        sum_il = np.zeros((3, 3))
        sum_m = np.zeros(3)
        sum_tsl = np.zeros(3, dtype=np.complex128)
        sum_trr = np.zeros(3, dtype=np.complex128)
        for k in range(self.nmod2):
            d2 = self.decay[k, 2]
            self.r_il[k] = d2 * self.r_il[k] + (1.0 - d2) * self.p_il
            self.r_m[k] = d2 * self.r_m[k] + (1.0 - d2) * self.p_m
            self.r_lambda[k] = d2 * self.r_lambda[k] + (1.0 - d2) * self.p_lambda
            sum_il += self.a_il[k] * self.r_il[k]
            sum_m += self.a_m[k] * self.r_m[k]
            sum_tsl += self.a_tsl[k] * self.r_lambda[k]
            sum_trr += self.a_trr[k] * self.r_lambda[k]
        return visc, rad, sum_il, sum_m, sum_tsl, sum_trr


class FastCpuSolver(SeaLevelSolver):
    """The oracle's solver step on FastCpuSHT and FastExactMemory."""

    def __init__(self, sht, *args, **kw):
        super().__init__(sht, *args, **kw)
        self.state.__class__ = FastExactMemory
