"""Recursive modal state on the GPU, with the interface of the reference's ``ExactMemory``
(python/slevolve/solver.py:244-506), including the second-order placement (``ramp``) and the
horizontal and rate read-outs.

The state R_k is stored packed on the device ([kmax][ncoef] complex128).  Every pass over the
states runs in the fused decay-and-accumulate kernel of libslx.so (slx_modal_advance_read):

* advance:        R_k <- d R_k + (1 - d) P,  visc = sum_k a_sl R_k,  rad = sum_k a_rr R_k;
* ramp commit:    d = 1, 1 - d -> 1 - phi_k, P -> L:  R_k <- R_k + (1 - phi_k) L  (weights zero);
* read-outs:      d = 1, 1 - d = 0:  R_k unchanged, visc/rad = sum_k w R_k for the read-out's
                  weights (a_hl for the horizontal read-out, a s / 1000 for the rate read-out).

The kernel evaluates x = r d + p o without contraction, so with d = 1 and o = 0 the states are
rewritten unchanged and the sums run over k in ascending order, as the reference's loops do.
The degree-2 rotational states (a few dozen numbers) follow the same recursions on the host, in
the reference's order.
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


def ramp_factor(x):
    """phi(x) = (1 - e^-x)/x and 1 - phi(x), third-order series where |x| < 1e-4 (solver.py:229-241)."""
    x = np.asarray(x, dtype=np.float64)
    small = np.abs(x) < 1.0e-4
    xs = np.where(small, 1.0, x)
    phi = np.where(small, 1.0 - x / 2.0 + x * x / 6.0 - x**3 / 24.0, -np.expm1(-xs) / xs)
    omp = np.where(small, x / 2.0 - x * x / 6.0 + x**3 / 24.0, (xs + np.expm1(-xs)) / xs)
    return phi, omp


def decay_factors(s, nmod, dt_years):
    """decay[k, l] = exp(-s dt / 1000) for existing modes, 0 elsewhere; omd = 1 - decay."""
    d = np.zeros_like(s)
    for l in range(1, s.shape[1]):
        n = int(nmod[l])
        d[:n, l] = np.exp(-s[:n, l] * dt_years / 1000.0)
    return d, 1.0 - d


class ExactMemory:
    """Hereditary modal state of the reference's ``ExactMemory`` (solver.py:244-506) on the device.

    ``ramp`` (off by default, set by the caller as in the reference) selects the second-order
    placement in commit.  The read-outs need the horizontal weights (``horizontal_weights``) or the
    rate weights a s / 1000 (always present)."""

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
        self.w_sl = a_sl * self._s / 1000.0
        self.w_rr = a_rr * self._s / 1000.0
        self.ramp = False
        self.a_hl = self.a_thl = None
        t = self._t
        self._nmod_d = torch.from_numpy(np.ascontiguousarray(self.nmod, dtype=np.int32)).to(dev)
        self._asl_d, self._arr_d = t(a_sl), t(a_rr)
        self._wsl_d, self._wrr_d = t(self.w_sl), t(self.w_rr)
        self._ones_d = t(np.ones((max(K, 1), lmax + 1)))
        self._zeros_d = t(np.zeros((max(K, 1), lmax + 1)))
        nc = sht.ncoef
        # degree of every packed coefficient (order-major triangle: m = 0..L, l = m..L)
        self._l_of = torch.from_numpy(np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])).to(dev)
        self.R = torch.zeros((max(K, 1), nc), dtype=torch.complex128, device=dev)
        self._p_load = torch.zeros(nc, dtype=torch.complex128, device=dev)
        self._visc = torch.zeros(nc, dtype=torch.complex128, device=dev)
        self._rad = torch.zeros(nc, dtype=torch.complex128, device=dev)
        self._out1 = torch.zeros(nc, dtype=torch.complex128, device=dev)
        self._out2 = torch.zeros(nc, dtype=torch.complex128, device=dev)
        self.r_il = np.zeros((K, 3, 3))
        self.r_m = np.zeros((K, 3))
        self.r_lambda = np.zeros((K, 3), dtype=np.complex128)
        self.p_il = np.zeros((3, 3))
        self.p_m = np.zeros(3)
        self.p_lambda = np.zeros(3, dtype=np.complex128)
        self.dt_years = None
        self.set_dt(dt_years)

    def _t(self, a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.dev)

    def set_dt(self, dt_years):
        """Decay factors and the ramp coefficients for steps of dt_years (solver.py:331-367);
        recomputed only when the step length changes."""
        dt_years = float(dt_years)
        if dt_years == self.dt_years:
            return
        d, o = decay_factors(self._s, self.nmod, dt_years)
        self.decay, self.omd = d, o
        self._dec_d = self._t(d)
        self._omd_d = self._t(o)
        self.dt_years = dt_years
        self._set_ramp_coeffs(dt_years)

    def _set_ramp_coeffs(self, dt_years):
        """phi, 1 - phi per mode and the per-degree sums c_sl, c_rr, ws, wsphi; the degree-2
        c_il, c_m, c_tsl, c_trr (solver.py:348-367), summed over k in order."""
        phi, omp = ramp_factor(self._s * (dt_years / 1000.0))
        exists = self._s > 0
        self.phi = np.where(exists, phi, 1.0)
        self.omphi = np.where(exists, omp, 0.0)
        self.c_sl = (self.a_sl * self.omphi).sum(0)
        self.c_rr = (self.a_rr * self.omphi).sum(0)
        self.ws_sl = self.w_sl.sum(0)
        self.ws_rr = self.w_rr.sum(0)
        self.wsphi_sl = (self.w_sl * self.phi).sum(0)
        self.wsphi_rr = (self.w_rr * self.phi).sum(0)
        n2 = self.nmod2
        c2 = self.omphi[:n2, 2] if self.lmax >= 2 else np.zeros(0)
        self.c_il = float(c2 @ self.a_il[:n2])
        self.c_m = float(c2 @ self.a_m[:n2])
        self.c_tsl = float(c2 @ self.a_tsl[:n2])
        self.c_trr = float(c2 @ self.a_trr[:n2])
        self._omphi_d = self._t(self.omphi)
        per_coef = lambda v: self._t(v)[self._l_of]  # noqa: E731
        self._ws_sl_c, self._ws_rr_c = per_coef(self.ws_sl), per_coef(self.ws_rr)
        self._wsphi_sl_c, self._wsphi_rr_c = per_coef(self.wsphi_sl), per_coef(self.wsphi_rr)

    @property
    def p_load(self):
        """Committed load (dense [l, m] view, a copy)."""
        return self._sht.to_dense(self._p_load)

    def _pass(self, decay, omd, w1, w2, P, out1, out2):
        """One pass of the modal kernel: R_k <- decay R_k + omd P, out1 = sum_k w1 R_k, out2 = sum_k w2 R_k."""
        _lib.call("slx_modal_advance_read", self._sht.handle, self.kmax, self._nmod_d.data_ptr(), decay.data_ptr(),
                  omd.data_ptr(), w1.data_ptr(), w2.data_ptr(), self.R.data_ptr(), P.data_ptr(), out1.data_ptr(),
                  out2.data_ptr(), D.stream_ptr())

    def _packed(self, x):
        x = D.as_device(x, torch.complex128, self.dev)
        return self._sht.to_packed(x) if x.dim() == 2 else x

    def advance_and_read_packed(self):
        self._pass(self._dec_d, self._omd_d, self._asl_d, self._arr_d, self._p_load, self._visc, self._rad)
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

    def horizontal_weights(self, love):
        """a_hl[k, l] = l_k / s_k and the degree-2 tidal a_thl[k] (solver.py:424-434)."""
        self.a_hl, self.a_thl = horizontal_weights(love, self.lmax)
        self._ahl_d = self._t(self.a_hl)

    def read_horizontal_packed(self):
        """sum_k a_hl R_k (packed, device) and the tidal sum_k a_thl r_lambda_k; state unchanged."""
        if self.a_hl is None:
            raise RuntimeError("horizontal_weights(love) must be called before read_horizontal")
        self._pass(self._ones_d, self._zeros_d, self._ahl_d, self._zeros_d, self._p_load, self._out1, self._out2)
        thl = np.zeros(3, dtype=np.complex128)
        if self.lmax >= 2:
            for k in range(self.nmod2):
                thl += self.a_thl[k] * self.r_lambda[k]
        return self._out1.clone(), thl

    def read_horizontal(self):
        """sum_k a_hl R_k over the current states (dense [l, m]) and the degree-2 tidal sum
        (solver.py:436-452).  Valid between this step's advance_and_read and the next one."""
        hz, thl = self.read_horizontal_packed()
        return self._sht.to_dense(hz), thl

    def _rate_sums(self):
        self._pass(self._ones_d, self._zeros_d, self._wsl_d, self._wrr_d, self._p_load, self._out1, self._out2)
        return self._out1, self._out2

    def read_rate_packed(self, current_load):
        """rs = ws P - sum_k a s R_k + wsphi dL for sea level and radial displacement (per year),
        packed; state unchanged (solver.py:454-474)."""
        cl = self._packed(current_load)
        q_sl, q_rr = self._rate_sums()
        rs = self._ws_sl_c * self._p_load - q_sl + self._wsphi_sl_c * cl
        rr = self._ws_rr_c * self._p_load - q_rr + self._wsphi_rr_c * cl
        return rs, rr

    def read_rate(self, current_load):
        """Analytic left time derivative of the viscous sea-level and radial terms at the end of
        the step (dense [l, m]); call between advance_and_read and commit (solver.py:464-474)."""
        rs, rr = self.read_rate_packed(current_load)
        return self._sht.to_dense(rs), self._sht.to_dense(rr)

    def read_rate_rotation(self, dil, dm, dlam):
        """Degree-2 rotational states made ramp-consistent at the end of the step, r_k + (1 - phi_k) d(.),
        and s_k / 1000 (k < nmod2; solver.py:476-484).  State unchanged."""
        n2 = self.nmod2
        c = self.omphi[:n2, 2]
        r_il = self.r_il[:n2] + c[:, None, None] * np.asarray(dil)[None]
        r_m = self.r_m[:n2] + c[:, None] * np.asarray(dm)[None]
        r_la = self.r_lambda[:n2] + c[:, None] * np.asarray(dlam)[None]
        return r_il, r_m, r_la, self._s[:n2, 2] / 1000.0

    def commit(self, current_load, dil=None, dm=None, dlambda=None):
        """P += current_load (dense [l, m] tensor/array, or packed); rotational states likewise.
        With ``ramp`` the increment first adds (1 - phi_k) of itself to every state
        (solver.py:486-506)."""
        cl = self._packed(current_load)
        if self.ramp:
            self._pass(self._ones_d, self._omphi_d, self._zeros_d, self._zeros_d, cl, self._out1, self._out2)
            if self.lmax >= 2:
                n2 = self.nmod2
                c = self.omphi[:n2, 2]
                if dil is not None:
                    self.r_il[:n2] += c[:, None, None] * np.asarray(dil)[None]
                if dm is not None:
                    self.r_m[:n2] += c[:, None] * np.asarray(dm)[None]
                if dlambda is not None:
                    self.r_lambda[:n2] += c[:, None] * np.asarray(dlambda)[None]
        self._p_load += cl
        if dil is not None:
            self.p_il = self.p_il + dil
        if dm is not None:
            self.p_m = self.p_m + dm
        if dlambda is not None:
            self.p_lambda = self.p_lambda + dlambda
