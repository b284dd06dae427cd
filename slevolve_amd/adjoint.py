"""Adjoint of the sea-level time step with the ocean and grounding masks frozen (no counterpart in
the reference, whose solver has no derivatives).

Model.  With the ocean function C*_n = C_n beta_n and the grounding indicator keep_n (the marine
check of solver.py:655-660) held at the converged values of a forward run, one step of
``SeaLevelSolver.step`` (solver.py:814-1144, default step placement, rotation on) is

    R_k <- d_k R_k + (1 - d_k) P;  visc, rad = sum_k a_sl,k R_k, sum_k a_rr,k R_k       (modal)
    I* = ice_n keep_n;  I_lm = A(I*);  dI = I_lm - prevI;  DI = pI + dI                    (ice)
    x = Phi(x):  DS = pO + x;  (il, m, Lambda, rot) = rotation(DS_2m, DI_2m, modal sums)
                 CS = T E (rho_i DI + rho_w DS) + T visc + rot_sl;  dsl = S(CS)
                 rt = A(dsl C* - tinit (C* - C*0));  c = (-(rho_i/rho_w) DI_00 - rt_00) / A(C*)_00
                 Phi(x) = -pO + rt + A(C*) c                                            (eq. B18 loop)
    deltasl = dsl + c K00;  CSr = T h (rho_i DI + rho_w DS) + T rad + rot_rr;  G = deltasl + S(CSr)
    P += rho_i dI + rho_w x (l > 0);  pI = DI;  pO = DS;  prevI = I_lm;  (il, m, Lambda) kept

with A, S the transforms.  It is affine in the state and the ice except for the centrifugal
potential, quadratic in the rotation vector m (solver.py:955-959), so the reverse sweep needs no
forward state but the per-step m (for that Jacobian) and the masks: the modal states R_k are
never stored or recomputed for the backward (the adjoint recursion runs on their cotangents).
``forward`` still returns the full state every ``checkpoint`` steps (default 20), from which
``FrozenStepModel.resume`` restarts a run, e.g. to re-evaluate the misfit of a segment.

Reverse sweep of a step (the transposes of every statement above, in reverse order): the commit,
the output cotangents, the adjoint fixed point mu = Phi_x^T mu + h_x^T o_bar + x_bar (solved by the
same iteration as the forward one, whose contraction it shares), then the ice analysis
(A^T, slx_sht_analysis_adjoint), the reverse modal recursion (slx_modal_reverse) and the degree-2
rotational states.  A^T and S^T (slx_sht_synthesis_adjoint) are HIP kernels; the per-coefficient
algebra runs as device tensor operations and the rotation (a few scalars per step) on the host.

Misfit (the adjoint configuration): J = sum_p deltasl_N(p)^2 [m^2] + (dJ2/dt / 1e-11 yr^-1)^2 at
the last step N, with J2_n = -sqrt(5) Re G_20 / a (G_20 the (2, 0) coefficient of the geoid
change G = deltasl + R in the transform's convention) and dJ2/dt = (J2_N - J2_{N-1}) / dt.

``IceHistoryMisfit`` wraps model, misfit and reverse sweep as a torch.autograd.Function: the
gradient of J with respect to the ice-thickness history (N + 1 grids, FP32) in one backward call.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _device as D
from . import _lib
from .memory import decay_factors, love_weights

PI_LEGACY = 3.1415926535898
SQRT2_LEGACY = 1.41421356
J2_RATE_SCALE = 1e-11  # 1/yr


class DeviceOps:
    """The transforms, their transposes and the modal advance / reverse on the device (C-ABI),
    for a batch of B fields: grids [B, nlat, nlon] float64, spectra [B, ncoef] complex128."""

    def __init__(self, sht):
        self.sht, self.h = sht, sht.handle
        self.nc, self.nlat, self.nlon, self.L = sht.ncoef, sht.nlat, sht.nlon, sht.lmax
        self.dev = sht.dev
        self.lib = _lib.load()

    def _each(self, fn, src, dst):
        for b0 in range(0, src.shape[0], 3):
            b1 = min(src.shape[0], b0 + 3)
            _lib.check(fn(self.h, C.c_void_p(src[b0].data_ptr()), b1 - b0, C.c_void_p(dst[b0].data_ptr()),
                          C.c_void_p(D.stream_ptr())))
        return dst

    def A(self, g):
        return self._each(self.lib.slx_sht_analysis, g.contiguous(),
                          torch.empty((g.shape[0], self.nc), dtype=torch.complex128, device=self.dev))

    def S(self, c):
        return self._each(self.lib.slx_sht_synthesis, c.contiguous(),
                          torch.empty((c.shape[0], self.nlat, self.nlon), dtype=torch.float64, device=self.dev))

    def AT(self, c):
        return self._each(self.lib.slx_sht_analysis_adjoint, c.contiguous(),
                          torch.empty((c.shape[0], self.nlat, self.nlon), dtype=torch.float64, device=self.dev))

    def ST(self, g):
        return self._each(self.lib.slx_sht_synthesis_adjoint, g.contiguous(),
                          torch.empty((g.shape[0], self.nc), dtype=torch.complex128, device=self.dev))

    def modal(self, tabs, R, P):
        """R [B, K, nc] advanced (in place); returns (R, visc, rad [B, nc])."""
        B = P.shape[0]
        visc = torch.empty_like(P)
        rad = torch.empty_like(P)
        for b in range(B):
            _lib.check(self.lib.slx_modal_advance_read(
                self.h, tabs["K"], tabs["nmod"].data_ptr(), tabs["dec"].data_ptr(), tabs["omd"].data_ptr(),
                tabs["asl"].data_ptr(), tabs["arr"].data_ptr(), R[b].data_ptr(), P[b].data_ptr(), visc[b].data_ptr(),
                rad[b].data_ptr(), D.stream_ptr()), "slx_modal_advance_read")
        return R, visc, rad

    def modal_reverse(self, tabs, Rbar, vbar, rbar, Pbar):
        for b in range(Pbar.shape[0]):
            _lib.check(self.lib.slx_modal_reverse(
                self.h, tabs["K"], tabs["nmod"].data_ptr(), tabs["dec"].data_ptr(), tabs["omd"].data_ptr(),
                tabs["asl"].data_ptr(), tabs["arr"].data_ptr(), Rbar[b].data_ptr(), vbar[b].contiguous().data_ptr(),
                rbar[b].contiguous().data_ptr(), Pbar[b].data_ptr(), D.stream_ptr()), "slx_modal_reverse")


class FrozenStepModel:
    """The frozen-mask step model of a solver's configuration (see the module docstring).

    ``keep`` [N+1, nlat, nlon] and ``cstar`` [N+1, nlat, nlon] (0/1) are the masks of steps 0..N
    (index 0: the initial state's), ``tinit`` and ``cstar0`` the initial topography and ocean
    function; ``love``, ``planet``, ``lmax`` and ``dt`` as the solver's.  ``ops`` provides the
    transforms and the modal advance on ``ops.dev`` (DeviceOps: the HIP kernels); the model's
    own algebra is written with tensor operations only, so a differentiable ``ops`` makes the
    whole forward differentiable by other means (the tests build an independent CPU one)."""

    def __init__(self, ops, love, planet, lmax, dt, tinit, cstar0, keep, cstar, tol=1e-13, max_iter=200):
        self.ops = ops
        self.dev = dev = ops.dev
        self.L, self.nl, self.nc = lmax, lmax + 1, ops.nc
        self.dt, self.tol, self.max_iter = float(dt), float(tol), int(max_iter)
        p = planet
        self.ri, self.rw, self.g, self.radius = p.rhoi, p.rhow, p.gacc, p.radius
        nmod, s, a_sl, a_rr, a2, nmod2 = love_weights(love, lmax)
        dec, omd = decay_factors(s, nmod, self.dt)
        K = s.shape[0]
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self.tabs = {"K": K, "nmod": t(nmod.astype(np.int32)), "dec": t(dec), "omd": t(omd), "asl": t(a_sl),
                     "arr": t(a_rr)}
        self.K, self.nmod2 = K, nmod2
        self.d2 = torch.from_numpy(dec[:nmod2, 2].copy()) if lmax >= 2 else torch.zeros(0, dtype=torch.float64)
        self.a2 = torch.from_numpy(a2[:, :nmod2].copy())  # [4, nmod2]: a_il, a_m, a_tsl, a_trr
        ell = np.arange(self.nl, dtype=np.float64)
        ttl = np.zeros(self.nl)
        ttl[1:] = (4.0 * PI_LEGACY * p.radius ** 3) / (p.mass * (2 * ell[1:] + 1.0))
        ekhl = np.zeros(self.nl)
        ekhl[1:] = 1 + love.ke[1:self.nl] - love.he[1:self.nl]
        he = np.concatenate(([0.0], love.he[1:self.nl]))
        l_of = np.concatenate([np.arange(m, self.nl) for m in range(self.nl)])
        m_of = np.concatenate([np.full(self.nl - m, m) for m in range(self.nl)])
        self.w1 = t((ttl * ekhl)[l_of])  # T E
        self.w2 = t(ttl[l_of])           # T
        self.wh = t((ttl * he)[l_of])    # T h
        self.lpos = t((l_of > 0).astype(np.float64))
        self.i20, self.i21, self.i22 = 2, self.nl + 1, 2 * self.nl - 1  # packed (2, 0), (2, 1), (2, 2)
        assert l_of[self.i21] == 2 and m_of[self.i21] == 1 and l_of[self.i22] == 2 and m_of[self.i22] == 2
        kTE2, hTE2, kE2 = float(love.kTE[2]), float(love.hTE[2]), float(love.ke[2])
        r4 = p.radius ** 4
        self.cI33 = -8 * PI_LEGACY * r4 / (3 * math.sqrt(5.0))
        self.cI13 = -8 * PI_LEGACY * r4 / math.sqrt(30.0)
        self.cI23 = 8 * PI_LEGACY * r4 / math.sqrt(30.0)
        self.alpha = 1 / (1 - kTE2 / p.kf)
        self.inv_CA, self.one_kE2, self.inv_kf, self.inv_C = 1 / (p.moiC - p.moiA), 1 + kE2, 1 / p.kf, 1.0 / p.moiC
        r2w2 = p.radius ** 2 * p.omega ** 2
        self.K20, self.K21, self.K22 = r2w2 / (6.0 * math.sqrt(5.0)), r2w2 / math.sqrt(30.0), r2w2 / math.sqrt(5.0 * 24.0)
        self.ekhTE, self.hTE2 = 1 + kTE2 - hTE2, hTE2
        self.K00 = SQRT2_LEGACY * math.sqrt(0.5)
        self.tinit = tinit.to(dev, torch.float64)
        self.keep = keep.to(dev, torch.float64)
        self.cstar = cstar.to(dev, torch.float64)
        self.cstar0 = cstar0.to(dev, torch.float64)
        self.tO = self.tinit[None] * (self.cstar - self.cstar0[None])
        self.cs = ops.A(self.cstar)   # [N+1, nc]: A(C*_n)
        self.cs00 = self.cs[:, 0].real.clone()
        self.m_hist = {}

    def _converged(self, d, prev):
        """Relative update d at or below the tolerance, or stalled at the round-off floor (within
        100 x the tolerance and no longer shrinking)."""
        return d <= self.tol or (d <= 100.0 * self.tol and d >= 0.5 * prev)

    # ------------------------------------------------------------------------- pieces
    def _rotation(self, DS2, DI2, sums):
        """DS2, DI2: [B, 3] complex (host) degree-2 coefficients m = 0, 1, 2; sums: host tensors.
        Returns il [B, 9], m [B, 3], lam [B, 6], rot_sl [B, 6], rot_rr [B, 6] (6 = re, im of m = 0..2)."""
        rw, ri = self.rw, self.ri
        z20r = rw * DS2[:, 0].real + ri * DI2[:, 0].real
        z21 = rw * DS2[:, 1] + ri * DI2[:, 1]
        B = DS2.shape[0]
        il = torch.zeros((B, 9), dtype=torch.float64)
        il[:, 8] = self.cI33 * z20r
        il[:, 2] = self.cI13 * z21.real
        il[:, 5] = self.cI23 * z21.imag
        si, sm = sums["il"], sums["m"]
        m1 = self.alpha * (self.inv_CA * (self.one_kE2 * il[:, 2] + si[:, 2]) + self.inv_kf * sm[:, 0])
        m2 = self.alpha * (self.inv_CA * (self.one_kE2 * il[:, 5] + si[:, 5]) + self.inv_kf * sm[:, 1])
        m3 = -self.inv_C * (self.one_kE2 * il[:, 8] + si[:, 8])
        lam = torch.stack([self.K20 * (m1 * m1 + m2 * m2 - 2.0 * m3 * m3 - 4.0 * m3), torch.zeros_like(m1),
                           -self.K21 * (m1 * (1.0 + m3)), -self.K21 * (-m2 * (1.0 + m3)),
                           self.K22 * (m2 * m2 - m1 * m1), self.K22 * (2.0 * m1 * m2)], dim=1)
        rot_sl = (self.ekhTE * lam + sums["tsl"]) / self.g
        rot_rr = (self.hTE2 * lam + sums["trr"]) / self.g
        return il, torch.stack([m1, m2, m3], dim=1), lam, rot_sl, rot_rr

    def _deg2(self, a):
        """[B, nc] complex (device) -> [B, 3] complex (host): coefficients (2, 0), (2, 1), (2, 2)."""
        return a[:, [self.i20, self.i21, self.i22]].cpu()

    def _add_deg2(self, c, six):
        """c[:, (2, m)] += six[:, 2m] + i six[:, 2m + 1] (six on the host)."""
        z = torch.complex(six[:, 0::2], six[:, 1::2]).to(self.dev)
        c[:, [self.i20, self.i21, self.i22]] += z
        return c

    def initial_state(self, ice0):
        """The state before step 1 for B initial ice grids [B, nlat, nlon]."""
        B = ice0.shape[0]
        z = lambda: torch.zeros((B, self.nc), dtype=torch.complex128, device=self.dev)  # noqa: E731
        st = {"P": z(), "pI": z(), "pO": z(), "prevI": self.ops.A(ice0 * self.keep[0]),
              "R": torch.zeros((B, max(self.K, 1), self.nc), dtype=torch.complex128, device=self.dev),
              "r2": torch.zeros((B, self.nmod2, 18), dtype=torch.float64),
              "pil": torch.zeros((B, 9), dtype=torch.float64), "pm": torch.zeros((B, 3), dtype=torch.float64),
              "plam": torch.zeros((B, 6), dtype=torch.float64), "x": z()}
        return st

    def _modal_sums(self, st):
        pv = torch.cat([st["pil"], st["pm"], st["plam"]], dim=1)  # [B, 18]
        d = self.d2[None, :, None]
        st["r2"] = d * st["r2"] + (1.0 - d) * pv[:, None, :]
        a = self.a2
        r = st["r2"]
        return {"il": torch.einsum("k,bkq->bq", a[0], r[:, :, :9]), "m": torch.einsum("k,bkq->bq", a[1], r[:, :, 9:12]),
                "tsl": torch.einsum("k,bkq->bq", a[2], r[:, :, 12:]), "trr": torch.einsum("k,bkq->bq", a[3], r[:, :, 12:])}

    def step(self, st, ice, n, record=False):
        """Advance the batch state st by step n (1-based) with ice [B, nlat, nlon]; returns the
        outputs {deltasl [B, nlat, nlon], G20 [B]} (deltasl only when record)."""
        ops = self.ops
        st["R"], visc, rad = ops.modal(self.tabs, st["R"], st["P"])
        sums = self._modal_sums(st)
        Ilm = ops.A(ice * self.keep[n])
        dI = Ilm - st["prevI"]
        DI = st["pI"] + dI
        DI2 = self._deg2(DI)
        base = self.w1 * (self.ri * DI) + self.w2 * visc
        cs, cs00, cstar, tO = self.cs[n], self.cs00[n], self.cstar[n], self.tO[n]
        x = st["x"].clone()
        prev = math.inf
        for it in range(self.max_iter):
            DS = st["pO"] + x
            il, m, lam, rot_sl, rot_rr = self._rotation(self._deg2(DS), DI2, sums)
            CS = self._add_deg2(base + (self.rw * self.w1) * DS, rot_sl)
            dsl = ops.S(CS)
            rt = ops.A(dsl * cstar - tO)
            cons = (-(self.ri / self.rw) * DI[:, 0].real - rt[:, 0].real) / cs00
            xn = -st["pO"] + rt + cs[None] * cons[:, None]
            dx = float(((xn - x).abs().amax(dim=1) / xn.abs().amax(dim=1).clamp_min(1e-300)).max())
            x = xn
            if self._converged(dx, prev):
                break
            prev = dx
        else:
            raise RuntimeError(f"frozen-mask step {n}: fixed point not converged ({dx:.2e})")
        DSx = st["pO"] + x  # committed state (at the fixed point)
        CSr = self._add_deg2(self.wh * (self.ri * DI + self.rw * (st["pO"] + x)) + self.w2 * rad, rot_rr)
        out = {"G20": (CS[:, self.i20] + CSr[:, self.i20]).real.clone(), "m": m}
        if record:
            out["deltasl"] = dsl + (cons * self.K00)[:, None, None]
        st["P"] = st["P"] + self.lpos * (self.ri * dI + self.rw * x)
        st["pI"], st["pO"], st["prevI"], st["x"] = DI, DSx, Ilm, x
        st["pil"], st["pm"], st["plam"] = il, m, lam
        return out

    # ------------------------------------------------------------------------- runs
    def forward(self, ice, nsteps=None, keep_m=False, checkpoint=20):
        """Run steps 1..N for a batch of ice histories ice [B, N+1, nlat, nlon] (or [N+1, ...]);
        returns (deltasl_N [B, nlat, nlon], G20 [B, N+1] with column n = step n, checkpoints)."""
        if ice.dim() == 3:
            ice = ice[None]
        ice = ice.to(self.dev, torch.float64)
        N = ice.shape[1] - 1 if nsteps is None else int(nsteps)
        st = self.initial_state(ice[:, 0])
        G20 = torch.zeros((ice.shape[0], N + 1), dtype=torch.float64)
        ckpt = {}
        for n in range(1, N + 1):
            if checkpoint and (n - 1) % checkpoint == 0:
                ckpt[n - 1] = {k: (v.clone() if torch.is_tensor(v) else v) for k, v in st.items()}
            out = self.step(st, ice[:, n], n, record=(n == N))
            G20[:, n] = out["G20"].cpu()
            if keep_m:
                self.m_hist[n] = out["m"].clone()
        return out["deltasl"], G20, ckpt

    def resume(self, state, n0, ice, n1):
        """Steps n0+1..n1 from a checkpointed state (as returned by forward for step n0) with the
        ice history ice [B, >= n1+1, nlat, nlon]; returns (deltasl_{n1}, G20 of steps n0+1..n1)."""
        st = {k: (v.clone() if torch.is_tensor(v) else v) for k, v in state.items()}
        ice = ice.to(self.dev, torch.float64)
        G20 = []
        for n in range(n0 + 1, n1 + 1):
            out = self.step(st, ice[:, n], n, record=(n == n1))
            G20.append(out["G20"].cpu())
        return out["deltasl"], torch.stack(G20, dim=1)

    # ------------------------------------------------------------------------- misfit
    def misfit(self, deltasl_N, G20, points):
        """J per batch member: sum over the points (flat grid indices) of deltasl_N^2 plus the
        squared J2 rate in units of 1e-11 / yr."""
        v = deltasl_N.reshape(deltasl_N.shape[0], -1)[:, points]
        j2 = -math.sqrt(5.0) * G20 / self.radius
        rate = (j2[:, -1] - j2[:, -2]) / self.dt / J2_RATE_SCALE
        return (v * v).sum(dim=1).cpu() + rate * rate

    # ------------------------------------------------------------------------- reverse
    def _back_phi(self, n, xbar, o):
        """Cotangents through one evaluation of Phi at step n: xbar of its output and o of the
        intermediates -> (cotangent of its input x, cotangents of its other inputs)."""
        ops = self.ops
        cs, cs00 = self.cs[n], float(self.cs00[n])
        pOb = -xbar
        rtb = xbar.clone()
        consb = o["cons"] + (xbar.conj() * cs[None]).real.sum(dim=1)
        rtb[:, 0] += (-consb / cs00).to(rtb.dtype)
        DIb = torch.zeros_like(xbar)
        DIb[:, 0] += (-(self.ri / self.rw) * consb / cs00).to(DIb.dtype)
        dslb = ops.AT(rtb) * self.cstar[n]
        if o["dsl"] is not None:
            dslb = dslb + o["dsl"]
        CSb = ops.ST(dslb)
        if o["CS"] is not None:
            CSb = CSb + o["CS"]
        DSb = o["DS"] + (self.rw * self.w1) * CSb
        DIb = DIb + (self.ri * self.w1) * CSb
        viscb = self.w2 * CSb
        c2 = self._deg2(CSb)
        slb = torch.stack([c2[:, 0].real, c2[:, 0].imag, c2[:, 1].real, c2[:, 1].imag, c2[:, 2].real, c2[:, 2].imag], 1)
        lamb = o["lam"] + (self.ekhTE / self.g) * slb + (self.hTE2 / self.g) * o["rr"]
        m = self.m_hist[n]
        m1, m2, m3 = m[:, 0], m[:, 1], m[:, 2]
        K20, K21, K22 = self.K20, self.K21, self.K22
        m1b = (o["m"][:, 0] + lamb[:, 0] * 2 * K20 * m1 + lamb[:, 2] * (-K21 * (1 + m3)) + lamb[:, 4] * (-2 * K22 * m1)
               + lamb[:, 5] * (2 * K22 * m2))
        m2b = (o["m"][:, 1] + lamb[:, 0] * 2 * K20 * m2 + lamb[:, 3] * (K21 * (1 + m3)) + lamb[:, 4] * (2 * K22 * m2)
               + lamb[:, 5] * (2 * K22 * m1))
        m3b = o["m"][:, 2] + lamb[:, 0] * K20 * (-4 * m3 - 4) + lamb[:, 2] * (-K21 * m1) + lamb[:, 3] * (K21 * m2)
        ilb = o["il"].clone()
        a_ca = self.alpha * self.inv_CA
        ilb[:, 2] += a_ca * self.one_kE2 * m1b
        ilb[:, 5] += a_ca * self.one_kE2 * m2b
        ilb[:, 8] += -self.inv_C * self.one_kE2 * m3b
        B = xbar.shape[0]
        s_il = torch.zeros((B, 9), dtype=torch.float64)
        s_il[:, 2], s_il[:, 5], s_il[:, 8] = a_ca * m1b, a_ca * m2b, -self.inv_C * m3b
        s_m = torch.zeros((B, 3), dtype=torch.float64)
        s_m[:, 0], s_m[:, 1] = self.alpha * self.inv_kf * m1b, self.alpha * self.inv_kf * m2b
        z20 = self.cI33 * ilb[:, 8]
        z21 = torch.complex(self.cI13 * ilb[:, 2], self.cI23 * ilb[:, 5])
        zs = torch.stack([torch.complex(z20, torch.zeros_like(z20)), z21, torch.zeros_like(z21)], 1).to(self.dev)
        idx = [self.i20, self.i21, self.i22]
        DSb[:, idx] += self.rw * zs
        DIb[:, idx] += self.ri * zs
        th = {"pO": pOb + DSb, "DI": DIb, "visc": viscb, "il": s_il, "m": s_m, "tsl": slb / self.g,
              "trr": o["rr"] / self.g}
        return DSb, th

    def gradient(self, ice, points):
        """J and dJ/d(ice history) [N+1, nlat, nlon] for one ice history [N+1, nlat, nlon]."""
        ice = ice.to(self.dev, torch.float64)
        N = ice.shape[0] - 1
        self.m_hist = {}
        dslN, G20, _ = self.forward(ice, keep_m=True)
        J = self.misfit(dslN, G20, points)
        # output cotangents
        dslb = torch.zeros_like(dslN).reshape(1, -1)
        dslb[:, points] = 2.0 * dslN.reshape(1, -1)[:, points]
        dslb = dslb.reshape(dslN.shape)
        j2 = -math.sqrt(5.0) * G20[0] / self.radius
        rate = (j2[N] - j2[N - 1]) / self.dt / J2_RATE_SCALE
        gN = float(2.0 * rate / J2_RATE_SCALE / self.dt * (-math.sqrt(5.0) / self.radius))
        G20b = {N: gN, N - 1: -gN}
        z = lambda: torch.zeros((1, self.nc), dtype=torch.complex128, device=self.dev)  # noqa: E731
        a = {"P": z(), "pI": z(), "pO": z(), "prevI": z(),
             "R": torch.zeros((1, max(self.K, 1), self.nc), dtype=torch.complex128, device=self.dev),
             "r2": torch.zeros((1, self.nmod2, 18), dtype=torch.float64), "pil": torch.zeros((1, 9), dtype=torch.float64),
             "pm": torch.zeros((1, 3), dtype=torch.float64), "plam": torch.zeros((1, 6), dtype=torch.float64)}
        grad = torch.zeros((N + 1,) + tuple(ice.shape[1:]), dtype=torch.float64, device=self.dev)
        for n in range(N, 0, -1):
            grad[n] = self._reverse_step(n, a, dslb if n == N else None, G20b.get(n, 0.0))[0]
        grad[0] = (self.ops.AT(a["prevI"]) * self.keep[0])[0]
        return J, grad

    def _reverse_step(self, n, a, dslb, g20b):
        ops = self.ops
        curb = self.lpos * a["P"]
        dIb = self.ri * curb
        xdir = self.rw * curb
        B = a["P"].shape[0]
        zero6 = torch.zeros((B, 6), dtype=torch.float64)
        o = {"dsl": dslb, "cons": torch.zeros(B, dtype=torch.float64, device=self.dev), "CS": None,
             "DS": a["pO"].clone(), "il": a["pil"].clone(), "m": a["pm"].clone(), "lam": a["plam"].clone(),
             "rr": zero6.clone()}
        if dslb is not None:
            o["cons"] = self.K00 * dslb.reshape(B, -1).sum(dim=1)
        DIb = a["pI"].clone()
        radb = torch.zeros_like(a["P"])
        if g20b:
            e = torch.zeros_like(a["P"])
            e[:, self.i20] = g20b
            o["CS"] = e
            # the radial coefficients CSr (the G20 output): h-outputs of the fixed point
            o["DS"] = o["DS"] + (self.rw * self.wh) * e
            DIb = DIb + (self.ri * self.wh) * e
            radb = self.w2 * e
            o["rr"][:, 0] += g20b
        # adjoint fixed point mu = Phi_x^T mu + h_x^T o + xdir
        mu = xdir.clone()
        prev = math.inf
        for it in range(self.max_iter):
            xin, _ = self._back_phi(n, mu, o)
            mun = xin + xdir
            d = float((mun - mu).abs().amax() / mun.abs().amax().clamp_min(1e-300))
            mu = mun
            if self._converged(d, prev):
                break
            prev = d
        else:
            raise RuntimeError(f"adjoint fixed point of step {n} not converged")
        _, th = self._back_phi(n, mu, o)
        DIb = DIb + th["DI"]
        pOb_prev = th["pO"]
        # ice analysis
        dIt = dIb + DIb
        Ilmb = a["prevI"] + dIt
        iceb = ops.AT(Ilmb) * self.keep[n]
        # modal states
        Pb_prev = a["P"].clone()
        ops.modal_reverse(self.tabs, a["R"], th["visc"], radb, Pb_prev)
        # degree-2 rotational states: r2_new = d r2_old + (1 - d) pv
        an = self.a2
        r2b = a["r2"].clone()
        r2b[:, :, :9] += an[0][None, :, None] * th["il"][:, None, :]
        r2b[:, :, 9:12] += an[1][None, :, None] * th["m"][:, None, :]
        r2b[:, :, 12:] += an[2][None, :, None] * th["tsl"][:, None, :] + an[3][None, :, None] * th["trr"][:, None, :]
        d = self.d2[None, :, None]
        pvb = ((1.0 - d) * r2b).sum(dim=1)
        a.update({"P": Pb_prev, "pI": DIb, "pO": pOb_prev, "prevI": -dIt, "r2": d * r2b, "pil": pvb[:, :9],
                  "pm": pvb[:, 9:12], "plam": pvb[:, 12:]})
        return iceb


def frozen_model_from_run(solver, ice, nsteps, ops=None, record_outputs=False):
    """Run the device solver (already constructed and initialised with ice[0]) for nsteps steps of
    the ice history ice [>= nsteps+1, nlat, nlon] (device float64) and build the frozen-mask model
    of that trajectory: keep_n = the grounding indicator of the converged step n (I*_n > 0 where
    there is ice, land elsewhere), C*_n = ocean_n beta_n.  Returns (model, outputs) where outputs
    holds the solver's tgrid of the last step and G20 per step when record_outputs."""
    from .solver import _copy_dev

    dev = solver.dev
    nlat, nlon = solver.nlat, solver.nlon
    keep = torch.zeros((nsteps + 1, nlat, nlon), dtype=torch.float64, device=dev)
    cstar = torch.zeros_like(keep)
    ist = torch.empty((nlat, nlon), dtype=torch.float64, device=dev)
    _copy_dev(ist, solver._fields.icestar, ist.numel() * 8)
    keep[0] = torch.where(ice[0] > 0, (ist > 0).double(), (solver.tinit > 0).double())
    cstar[0] = solver.cxy0 * solver.beta0
    outs = {"tgrid": [], "mm": []}
    for n in range(1, nsteps + 1):
        r = solver.step(ice[n])
        _copy_dev(ist, solver._fields.icestar, ist.numel() * 8)
        keep[n] = torch.where(ice[n] > 0, (ist > 0).double(), (r["tgrid"] > 0).double())
        cstar[n] = r["ocean"] * r["beta"]
        if record_outputs:
            outs["tgrid"].append(r["tgrid"].clone())
            outs["mm"].append(np.array(r["mm"]))
    model = FrozenStepModel(ops or DeviceOps(solver.sht), solver.love, solver.p, solver.norder, solver.dt,
                            solver.tinit, solver.cxy0 * solver.beta0, keep, cstar)
    return model, outs


class IceHistoryMisfit(torch.autograd.Function):
    """J(ice history) of a frozen-mask model and misfit points; backward by the hand-written
    reverse sweep.  ice: [N+1, nlat, nlon] float32 or float64 (device); the gradient has its dtype."""

    @staticmethod
    def forward(ctx, ice, model, points):
        J, grad = model.gradient(ice.detach(), points)
        ctx.save_for_backward(grad.to(ice.dtype))
        return J[0].to(ice.device, ice.dtype)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return gJ * grad, None, None


def misfit_points(model, n_points=200, seed=None):
    """n_points seeded grid cells (flat indices) that are ocean (C* = 1) at every step."""
    ocean = (model.cstar > 0.5).all(dim=0).reshape(-1).nonzero().reshape(-1).cpu().numpy()
    rng = np.random.default_rng(seed if seed is not None else np.random.SeedSequence(20261203 + 5).spawn(5)[4])
    return torch.from_numpy(np.sort(rng.choice(ocean, size=min(n_points, ocean.size), replace=False))).to(model.dev)


__all__ = ["DeviceOps", "FrozenStepModel", "IceHistoryMisfit", "frozen_model_from_run", "misfit_points"]
