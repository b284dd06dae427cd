"""Recursive modal state on the GPU, with the interface of the reference's ``ExactMemory``
(python/slevolve/solver.py:244-506; default step placement).

The state R_k is stored packed on the device ([kmax][ncoef] complex128); the advance runs in the
fused decay-and-accumulate kernel of libslx.so.  The degree-2 rotational states (a few dozen
numbers) follow the same recursion on the host, in the reference's order.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _device as D
from . import _lib


def love_weights(love, lmax):
    """Modal weights of solver.py:252-279: s, a_sl = (r' - r)/s, a_rr = r/s per (k, l) and the
    degree-2 rotational a_il = r'/s, a_m = r'T/s, a_tsl = (r'T - rT)/s, a_trr = rT/s."""
    nl = lmax + 1
    K = int(love.max_modes)
    nmod = np.zeros(nl, dtype=np.int32)
    s = np.zeros((K, nl))
    a_sl = np.zeros((K, nl))
    a_rr = np.zeros((K, nl))
    a2 = np.zeros((4, K))
    for l in range(1, nl):
        n = int(love.nmod[l])
        nmod[l] = n
        sl = love.s[:n, l]
        s[:n, l] = sl
        a_sl[:n, l] = (love.rprime[:n, l] - love.r[:n, l]) / sl
        a_rr[:n, l] = love.r[:n, l] / sl
        if l == 2:
            a2[0, :n] = love.rprime[:n, l] / sl
            a2[1, :n] = love.rprimeT[:n, l] / sl
            a2[2, :n] = (love.rprimeT[:n, l] - love.rT[:n, l]) / sl
            a2[3, :n] = love.rT[:n, l] / sl
    nmod2 = int(love.nmod[2]) if lmax >= 2 else 0
    return nmod, s, a_sl, a_rr, a2, nmod2


def horizontal_weights(love, lmax):
    """Horizontal read-out weights of solver.py:424-434: a_hl = l_k/s_k per (k, l) and the tidal
    a_thl = l_k/s_k at degree 2."""
    nl = lmax + 1
    K = int(love.max_modes)
    a_hl = np.zeros((K, nl))
    a_thl = np.zeros(K)
    for l in range(1, nl):
        n = int(love.nmod[l])
        a_hl[:n, l] = love.rl[:n, l] / love.s[:n, l]
        if l == 2:
            a_thl[:n] = love.rlT[:n, l] / love.s[:n, l]
    return a_hl, a_thl


def decay_factors(s, nmod, dt_years):
    """decay[k, l] = exp(-s dt / 1000) for existing modes, 0 elsewhere; omd = 1 - decay."""
    d = np.zeros_like(s)
    for l in range(1, s.shape[1]):
        n = int(nmod[l])
        d[:n, l] = np.exp(-s[:n, l] * dt_years / 1000.0)
    return d, 1.0 - d


class ExactMemory:
    def __init__(self, love, lmax: int, dt_years: float, tpw: bool = True, nthreads: int = 1, sht=None,
                 device=None):
        from .sht import SHT

        self.lmax, self.tpw = int(lmax), bool(tpw)
        if sht is None:  # index tables only: the smallest grid that holds this truncation
            nlat = max(2, 2 * ((lmax + 2) // 4 + 1))
            nlon = 8
            while nlon < 2 * (lmax + 1):
                nlon *= 2
            sht = SHT(nlat, nlon, lmax, device=device, max_fields=1)
        if sht.lmax != self.lmax:
            raise ValueError("transform and state truncations differ")
        self._sht = sht
        dev = self.dev = sht.dev
        self.nmod, self._s, a_sl, a_rr, a2, self.nmod2 = love_weights(love, lmax)
        self.kmax = K = self._s.shape[0]
        self.a_sl, self.a_rr = a_sl, a_rr
        self.a_il, self.a_m, self.a_tsl, self.a_trr = (a2[i].copy() for i in range(4))
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self._nmod_d = t(self.nmod.astype(np.int32))
        self._asl_d, self._arr_d = t(a_sl), t(a_rr)
        nc = sht.ncoef
        self.R = torch.zeros((max(K, 1), nc), dtype=torch.complex128, device=dev)
        self._p_load = torch.zeros(nc, dtype=torch.complex128, device=dev)
        self._visc = torch.zeros(nc, dtype=torch.complex128, device=dev)
        self._rad = torch.zeros(nc, dtype=torch.complex128, device=dev)
        self.r_il = np.zeros((K, 3, 3))
        self.r_m = np.zeros((K, 3))
        self.r_lambda = np.zeros((K, 3), dtype=np.complex128)
        self.p_il = np.zeros((3, 3))
        self.p_m = np.zeros(3)
        self.p_lambda = np.zeros(3, dtype=np.complex128)
        self.dt_years = None
        self.set_dt(dt_years)

    def set_dt(self, dt_years):
        dt_years = float(dt_years)
        if dt_years == self.dt_years:
            return
        d, o = decay_factors(self._s, self.nmod, dt_years)
        self.decay, self.omd = d, o
        self._dec_d = torch.from_numpy(d).to(self.dev)
        self._omd_d = torch.from_numpy(o).to(self.dev)
        self.dt_years = dt_years

    @property
    def p_load(self):
        """Committed load (dense [l, m] view, a copy)."""
        return self._sht.to_dense(self._p_load)

    def advance_and_read_packed(self):
        _lib.call("slx_modal_advance_read", self._sht.handle, self.kmax, self._nmod_d.data_ptr(),
                  self._dec_d.data_ptr(), self._omd_d.data_ptr(), self._asl_d.data_ptr(), self._arr_d.data_ptr(),
                  self.R.data_ptr(), self._p_load.data_ptr(), self._visc.data_ptr(), self._rad.data_ptr(),
                  D.stream_ptr())
        return self._visc, self._rad

    def _advance_rotation(self):
        sum_il = np.zeros((3, 3))
        sum_m = np.zeros(3)
        sum_tsl = np.zeros(3, dtype=np.complex128)
        sum_trr = np.zeros(3, dtype=np.complex128)
        if self.lmax >= 2:
            for k in range(self.nmod2):
                d = self.decay[k, 2]
                o = 1.0 - d
                self.r_il[k] = d * self.r_il[k] + o * self.p_il
                self.r_m[k] = d * self.r_m[k] + o * self.p_m
                self.r_lambda[k] = d * self.r_lambda[k] + o * self.p_lambda
                sum_il += self.a_il[k] * self.r_il[k]
                sum_m += self.a_m[k] * self.r_m[k]
                sum_tsl += self.a_tsl[k] * self.r_lambda[k]
                sum_trr += self.a_trr[k] * self.r_lambda[k]
        return sum_il, sum_m, sum_tsl, sum_trr

    def advance_and_read(self):
        """Advance every mode one step; returns (viscous, radial, sum_il, sum_m, sum_tsl, sum_trr)
        with viscous/radial as dense [l, m] device tensors."""
        v, r = self.advance_and_read_packed()
        return (self._sht.to_dense(v), self._sht.to_dense(r)) + self._advance_rotation()

    def commit(self, current_load, dil=None, dm=None, dlambda=None):
        """P += current_load (dense [l, m] tensor/array, or packed); rotational states likewise."""
        cl = D.as_device(current_load, torch.complex128, self.dev)
        if cl.dim() == 2:
            cl = self._sht.to_packed(cl)
        self._p_load += cl
        if dil is not None:
            self.p_il = self.p_il + dil
        if dm is not None:
            self.p_m = self.p_m + dm
        if dlambda is not None:
            self.p_lambda = self.p_lambda + dlambda
