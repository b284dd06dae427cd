"""Oracle recursive modal state.  Test infrastructure only (see ``oracle/__init__.py``).

Restates ``ExactMemory`` (``python/slevolve/solver.py:229-506``):

* per mode k and degree l with s_kl > 0 (mode exists iff k < nmod[l], solver.py:266-276):
  a_sl = (r' - r)/s, a_rr = r/s; at degree 2 also a_il = r'/s, a_m = r'T/s,
  a_tsl = (r'T - rT)/s, a_trr = rT/s;
* decay d_kl = exp(-s_kl dt / 1000) (dt in years, s in 1/kyr; solver.py:331-346);
* advance: R_k <- d R_k + (1 - d) P, visc = sum_k a_sl R_k, rad = sum_k a_rr R_k, and the same
  recursion for the degree-2 rotational states r_il (3x3), r_m (3), r_lambda (3, complex)
  (solver.py:386-422);
* commit: P += L, p_il += dil, p_m += dm, p_lambda += dlambda (solver.py:486-506);
* second-order placement (``ramp``): phi(x) = (1 - e^-x)/x with its series below |x| = 1e-4
  (solver.py:229-241), per mode 1 - phi(s dt) and the per-degree sums c_sl, c_rr, ws, wsphi,
  the degree-2 scalars c_il, c_m, c_tsl, c_trr (solver.py:348-367); the ramp commit
  R_k += (1 - phi_k) L and r_k += (1 - phi_k) d(.) before P += L (solver.py:486-498);
* analytic rate read-out (solver.py:454-484) and the horizontal read-out weights and sums
  (solver.py:424-452).
"""
from __future__ import annotations

import numpy as np


def ramp_factor(x):
    """phi(x) = (1 - e^-x)/x and 1 - phi(x); third-order series where |x| < 1e-4 (solver.py:229-241)."""
    x = np.asarray(x, dtype=np.float64)
    small = np.abs(x) < 1.0e-4
    xs = np.where(small, 1.0, x)
    phi = np.where(small, 1.0 - x / 2.0 + x * x / 6.0 - x**3 / 24.0, -np.expm1(-xs) / xs)
    omp = np.where(small, x / 2.0 - x * x / 6.0 + x**3 / 24.0, (xs + np.expm1(-xs)) / xs)
    return phi, omp


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
        self.ramp = False
        self.phi = np.ones((K, nl))
        self.omphi = np.zeros((K, nl))
        self.w_sl = self.a_sl * self.s / 1000.0  # rate weights a_k s_k per year
        self.w_rr = self.a_rr * self.s / 1000.0
        self.a_hl = None
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
        self._set_ramp_coeffs(dt_years)

    def _set_ramp_coeffs(self, dt_years):
        """1 - phi(s dt) per mode and its sums (solver.py:348-367)."""
        phi, omp = ramp_factor(self.s * (dt_years / 1000.0))
        exists = self.s > 0
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

    # ------------------------------------------------------------------ read-outs
    def horizontal_weights(self, love):
        """a_hl = l_k / s_k per mode and degree, a_thl = tidal l_k / s_k at degree 2 (solver.py:424-434)."""
        nl = self.lmax + 1
        self.a_hl = np.zeros((self.kmax, nl))
        self.a_thl = np.zeros(self.kmax)
        for l in range(1, nl):
            n = self.nmod[l]
            self.a_hl[:n, l] = love.rl[:n, l] / love.s[:n, l]
            if l == 2:
                self.a_thl[:n] = love.rlT[:n, l] / love.s[:n, l]

    def read_horizontal(self):
        """sum_k a_hl R_k and the degree-2 sum_k a_thl r_lambda,k over the current states (solver.py:436-452)."""
        nl = self.lmax + 1
        hz = np.zeros((nl, nl), dtype=np.complex128)
        for k in range(self.kmax):
            hz += self.a_hl[k][:, None] * self.R[k]
        thl = np.zeros(3, dtype=np.complex128)
        for k in range(self.nmod2):
            thl += self.a_thl[k] * self.r_lambda[k]
        return hz, thl

    def read_rate(self, load):
        """Left time derivative (per year) of the viscous sea-level and radial sums at the end of the
        step, the increment ``load`` accruing linearly over it:
        ws P - sum_k a_k s_k R_k + wsphi L (solver.py:454-474).  State unchanged."""
        nl = self.lmax + 1
        q_sl = np.zeros((nl, nl), dtype=np.complex128)
        q_rr = np.zeros((nl, nl), dtype=np.complex128)
        for k in range(self.kmax):
            q_sl += self.w_sl[k][:, None] * self.R[k]
            q_rr += self.w_rr[k][:, None] * self.R[k]
        rs = self.ws_sl[:, None] * self.p_load - q_sl + self.wsphi_sl[:, None] * load
        rr = self.ws_rr[:, None] * self.p_load - q_rr + self.wsphi_rr[:, None] * load
        return rs, rr

    def read_rate_rotation(self, dil, dm, dlam):
        """Degree-2 states r_k + (1 - phi_k) d(.) and s_k per year (solver.py:476-484)."""
        n2 = self.nmod2
        c = self.omphi[:n2, 2]
        return (self.r_il[:n2] + c[:, None, None] * dil[None], self.r_m[:n2] + c[:, None] * dm[None],
                self.r_lambda[:n2] + c[:, None] * dlam[None], self.s[:n2, 2] / 1000.0)

    def commit(self, load, dil=None, dm=None, dlambda=None):
        if self.ramp:
            # the increment already carries 1 - phi_k of its response in every mode (solver.py:486-498)
            for k in range(self.kmax):
                self.R[k] += self.omphi[k][:, None] * load
            n2 = self.nmod2
            if n2:
                c = self.omphi[:n2, 2]
                if dil is not None:
                    self.r_il[:n2] += c[:, None, None] * dil[None]
                if dm is not None:
                    self.r_m[:n2] += c[:, None] * dm[None]
                if dlambda is not None:
                    self.r_lambda[:n2] += c[:, None] * dlambda[None]
        self.p_load += load
        if dil is not None:
            self.p_il += dil
        if dm is not None:
            self.p_m += dm
        if dlambda is not None:
            self.p_lambda += dlambda
