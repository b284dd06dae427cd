"""Oracle recursive modal state.  Test infrastructure only (see ``oracle/__init__.py``).

Restates ``ExactMemory`` (``python/slevolve/solver.py:244-506``, default step placement, no ramp):

* per mode k and degree l with s_kl > 0 (mode exists iff k < nmod[l], solver.py:266-276):
  a_sl = (r' - r)/s, a_rr = r/s; at degree 2 also a_il = r'/s, a_m = r'T/s,
  a_tsl = (r'T - rT)/s, a_trr = rT/s;
* decay d_kl = exp(-s_kl dt / 1000) (dt in years, s in 1/kyr; solver.py:331-346);
* advance: R_k <- d R_k + (1 - d) P, visc = sum_k a_sl R_k, rad = sum_k a_rr R_k, and the same
  recursion for the degree-2 rotational states r_il (3x3), r_m (3), r_lambda (3, complex)
  (solver.py:386-422);
* commit: P += L, p_il += dil, p_m += dm, p_lambda += dlambda (solver.py:486-506).
"""
from __future__ import annotations

import numpy as np


class ExactMemory:
    def __init__(self, love, lmax: int, dt_years: float):
        nl = lmax + 1
        K = love.max_modes
        self.lmax, self.kmax = lmax, K
        self.nmod = np.array([int(love.nmod[l]) if l > 0 else 0 for l in range(nl)])
        self.s = np.zeros((K, nl))
        self.a_sl = np.zeros((K, nl))
        self.a_rr = np.zeros((K, nl))
        for l in range(1, nl):
            n = self.nmod[l]
            s = love.s[:n, l]
            self.s[:n, l] = s
            self.a_sl[:n, l] = (love.rprime[:n, l] - love.r[:n, l]) / s
            self.a_rr[:n, l] = love.r[:n, l] / s
        self.nmod2 = int(love.nmod[2]) if lmax >= 2 else 0
        n2 = self.nmod2
        self.a_il = np.zeros(K)
        self.a_m = np.zeros(K)
        self.a_tsl = np.zeros(K)
        self.a_trr = np.zeros(K)
        if n2:
            s2 = love.s[:n2, 2]
            self.a_il[:n2] = love.rprime[:n2, 2] / s2
            self.a_m[:n2] = love.rprimeT[:n2, 2] / s2
            self.a_tsl[:n2] = (love.rprimeT[:n2, 2] - love.rT[:n2, 2]) / s2
            self.a_trr[:n2] = love.rT[:n2, 2] / s2
        self.decay = np.zeros((K, nl))
        self.dt_years = None
        self.set_dt(dt_years)
        self.R = np.zeros((K, nl, nl), dtype=np.complex128)
        self.p_load = np.zeros((nl, nl), dtype=np.complex128)
        self.p_ice = np.zeros((nl, nl), dtype=np.complex128)
        self.p_ocean = np.zeros((nl, nl), dtype=np.complex128)
        self.prev_ice = np.zeros((nl, nl), dtype=np.complex128)
        self.r_il = np.zeros((K, 3, 3))
        self.r_m = np.zeros((K, 3))
        self.r_lambda = np.zeros((K, 3), dtype=np.complex128)
        self.p_il = np.zeros((3, 3))
        self.p_m = np.zeros(3)
        self.p_lambda = np.zeros(3, dtype=np.complex128)

    def set_dt(self, dt_years):
        dt_years = float(dt_years)
        if dt_years == self.dt_years:
            return
        d = np.zeros_like(self.decay)
        for l in range(1, self.lmax + 1):
            n = self.nmod[l]
            d[:n, l] = np.exp(-self.s[:n, l] * dt_years / 1000.0)
        self.decay = d
        self.dt_years = dt_years

    def advance_and_read(self):
        nl = self.lmax + 1
        visc = np.zeros((nl, nl), dtype=np.complex128)
        rad = np.zeros((nl, nl), dtype=np.complex128)
        for k in range(self.kmax):
            d = self.decay[k][:, None]
            self.R[k] = self.R[k] * d + self.p_load * (1.0 - d)
            visc += self.R[k] * self.a_sl[k][:, None]
            rad += self.R[k] * self.a_rr[k][:, None]
        sum_il = np.zeros((3, 3))
        sum_m = np.zeros(3)
        sum_tsl = np.zeros(3, dtype=np.complex128)
        sum_trr = np.zeros(3, dtype=np.complex128)
        for k in range(self.nmod2):
            d = self.decay[k, 2]
            self.r_il[k] = d * self.r_il[k] + (1.0 - d) * self.p_il
            self.r_m[k] = d * self.r_m[k] + (1.0 - d) * self.p_m
            self.r_lambda[k] = d * self.r_lambda[k] + (1.0 - d) * self.p_lambda
            sum_il += self.a_il[k] * self.r_il[k]
            sum_m += self.a_m[k] * self.r_m[k]
            sum_tsl += self.a_tsl[k] * self.r_lambda[k]
            sum_trr += self.a_trr[k] * self.r_lambda[k]
        return visc, rad, sum_il, sum_m, sum_tsl, sum_trr

    def commit(self, load, dil=None, dm=None, dlambda=None):
        self.p_load += load
        if dil is not None:
            self.p_il += dil
        if dm is not None:
            self.p_m += dm
        if dlambda is not None:
            self.p_lambda += dlambda
