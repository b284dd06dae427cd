"""Sea-level solver on the GPU with the interface of the reference's ``SeaLevelSolver``
(python/slevolve/solver.py:499-1144) and ``OptSeaLevelSolver`` (python/slevolve/solver_opt.py:267-979).

The whole time step runs on the device (libslx.so): modal advance, grounded-ice projection, the
inner fixed point with its ocean-function / grounded-ice re-guesses and FP64 convergence test,
the radial displacement, the commit and the degree-0 diagnostics.  The host enqueues one step,
reads one small status block back, and returns the reference's result dictionary; the grids in
it are device tensors that the solver overwrites at the next step (the reference also returns
its reused buffers).

Supported options: tpw, checkmarine, calcRG, epsilon1, dt per step, inner "off" | "extrap" |
"anderson" (Anderson mixing with the frozen ocean mask, OptSeaLevelSolver), the
second-order load placement (ramp), the analytic rates (step(rates=True)), the horizontal read-out
(solver.horizontal = True, result key "horiz"), the cross-pass reuse modes of OptSeaLevelSolver
("warm", "warm+corr", "warm+corr2", "warm+ice"; per-step converged states kept on the device), its
raw-mode spectra (opt "raw" with a raw-mode SHT: dense spectral outputs in V = (-1)^m U) and the
per-stage device times of step(profile=True).  The axial-rotation variants are not provided and
raise.

``precision="mixed"`` (no counterpart in the reference) runs the early inner sweeps of every step
with the FP32-product Legendre contractions; a sweep takes the FP64 contractions once the previous
one changed no ocean-function cell and no grounded-ice flag, or once the convergence ratio is at or
below ``eps_switch`` (default 1000 epsilon1), and so do all later sweeps of the step; the step
converges on an FP64 sweep only.  Everything else stays FP64.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from . import _device as D
from . import _lib
from .memory import decay_factors, horizontal_weights, love_weights
from .sht import SHT

PI = 3.1415926535898
EPS0 = float(np.finfo(np.float64).eps)
GR_MAXACC = 3
ALL_OPTS = ("raw", "diag", "dsl", "numba", "sums")
INNER_MODES = ("off", "extrap", "anderson")
# cross-pass reuse (solver_opt.py:272-286).  "warm+ice" starts as "warm": the reference reuses the
# stored ice analysis only where the grounded support is unchanged, i.e. where it equals the
# recomputed one; here the step-start ice pass is incremental anyway (it transforms the change of
# I* only), so it is always recomputed (reuse_stats["ice_recomputed"]).
REUSE_MODES = ("off", "warm", "warm+ice", "warm+corr", "warm+corr2")


class Planet:
    """Earth constants (solver.py:98-110)."""

    def __init__(self):
        self.radius = 6.371e6
        self.mass = 5.976e24
        self.rhoi = 920.0
        self.rhow = 1000.0
        self.gacc = 9.80665
        self.omega = 7.292e-5
        self.moiA = 0.3296145 * self.mass * self.radius**2
        self.moiC = 0.3307007 * self.mass * self.radius**2
        self.kf = 0.9342 + 0.008


def parse_opts(spec):
    """``"all"`` / ``""`` / ``"raw,dsl"`` -> set of option names (solver_opt.py:168-185)."""
    if spec is None:
        return set()
    spec = spec.strip()
    if spec in ("", "none", "0"):
        return set()
    if spec == "all":
        return set(ALL_OPTS)
    out = set()
    for tok in (t.strip() for t in spec.split(",")):
        if tok and tok not in ALL_OPTS:
            raise ValueError(f"unknown --opt {tok!r} (known: {','.join(ALL_OPTS)})")
        if tok:
            out.add(tok)
    return out


class _StateView:
    """Dense [l, m] views of the device-resident committed state (solver.state in the reference)."""

    def __init__(self, slv):
        self._slv = slv

    def _dense(self, name):
        f = self._slv._fields
        n = self._slv.sht.ncoef
        t = torch.empty(n, dtype=torch.complex128, device=self._slv.dev)
        _copy_dev(t, getattr(f, name), n * 16)
        return self._slv.sht.to_dense(t)

    p_load = property(lambda self: self._dense("p_load"))
    p_ice = property(lambda self: self._dense("p_ice"))
    p_ocean = property(lambda self: self._dense("p_ocean"))
    prev_ice = property(lambda self: self._dense("prev_ice"))


def _copy_dev(dst: torch.Tensor, src_ptr: int, nbytes: int):
    """Device-to-device copy from a library-owned buffer into a torch tensor (current stream)."""
    lib = _HIPRT.get()
    rc = lib.hipMemcpyAsync(C.c_void_p(dst.data_ptr()), C.c_void_p(src_ptr), C.c_size_t(nbytes), 3,
                            C.c_void_p(D.stream_ptr()))
    if rc != 0:
        raise RuntimeError(f"hipMemcpyAsync failed ({rc})")


class _HIPRT:
    _lib = None

    @classmethod
    def get(cls):
        if cls._lib is None:
            cls._lib = C.CDLL("libamdhip64.so")
            cls._lib.hipMemcpyAsync.restype = C.c_int
            cls._lib.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        return cls._lib


def _status_record(st):
    lam = np.ctypeslib.as_array(st.lam).reshape(3, 3, 2)
    return {"iter": st.step, "numiter": st.numiter, "ocean_area": st.ocean_area, "oceanBeta_area": st.oceanBeta_area,
            "gmslc_ocn": st.gmslc_ocn, "gmslc_ocnBeta": st.gmslc_ocnBeta, "ice_volume": st.ice_volume,
            "il": np.ctypeslib.as_array(st.il).reshape(3, 3).copy(), "mm": np.ctypeslib.as_array(st.mm).copy(),
            "lambda": lam[..., 0] + 1j * lam[..., 1], "reguess": st.reguess, "ice_analyses": st.ice_analyses,
            "sweeps32": st.sweeps32}


class InnerLoopError(RuntimeError):
    """An inner fixed point did not converge within the iteration limit (solver.py:1049-1050).
    ``records`` holds the scalar records of the steps a ``run`` finished before the failing one,
    ``status`` the failing step's (not committed, as in the reference)."""

    def __init__(self, msg, records=(), status=None):
        super().__init__(msg)
        self.records = list(records)
        self.status = status


class SeaLevelSolver:
    """Exact-memory sea-level solver; constructor and ``step`` as in solver.py:521-1144."""

    def __init__(self, sht, love, topo0, ice0, *, norder=512, dt=1, tpw=True, checkmarine=True, calcRG=True,
                 epsilon1=1.0e-5, planet=None, mem_threads=1, ramp=False, axial_rotation="off",
                 axial_tau_c=None, inner="off", _anderson=(3, 2, 3, False), precision="fp64", eps_switch=None,
                 _max_inner=9999):
        if axial_rotation != "off":
            raise NotImplementedError("axial_rotation other than 'off' is not provided")
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
        if precision == "mixed" and inner == "anderson":
            raise ValueError("precision='mixed' needs inner 'extrap' or 'off'")
        if inner not in INNER_MODES:
            raise ValueError(f"inner must be one of {INNER_MODES}, got {inner!r}")
        if not isinstance(sht, SHT):
            raise TypeError("sht must be a slevolve_amd.SHT (the GPU transform)")
        if sht.lmax != norder:
            raise ValueError("norder must equal the transform's lmax")
        if sht.max_fields < 3:
            raise ValueError("the solver needs an SHT with max_fields >= 3")
        if sht.raw and not getattr(self, "_raw_ok", False):
            raise ValueError("a raw-mode SHT needs OptSeaLevelSolver(opt={'raw', ...}) (solver_opt.py:299-303)")
        self.sht, self.love = sht, love
        self.norder, self.nl = int(norder), int(norder) + 1
        self.dt = float(dt)
        self.tpw, self.checkmarine, self.calcRG = bool(tpw), bool(checkmarine), bool(calcRG)
        self.epsilon1 = float(epsilon1)
        self.precision = precision
        self.eps_switch = float(1000.0 * self.epsilon1 if eps_switch is None else eps_switch)
        if precision == "mixed" and not self.eps_switch > self.epsilon1:
            raise ValueError("eps_switch must exceed epsilon1")
        self.inner = inner
        self.horizontal = False  # set True for the horizontal read-out (result key "horiz")
        self.ramp = bool(ramp)
        self.p = planet or Planet()
        self.nlat, self.nlon = sht.nlat, sht.nlon
        self.dev = dev = sht.dev
        p = self.p
        nl = self.nl

        nmod, s, a_sl, a_rr, a2, nmod2 = love_weights(love, norder)
        self._nmod, self._s = nmod, s
        K = s.shape[0]
        ell = np.arange(nl, dtype=np.float64)
        ttl = np.zeros(nl)
        ttl[1:] = (4.0 * PI * (p.radius**3)) / (p.mass * (2 * ell[1:] + 1.0))
        ekhl = np.zeros(nl)
        ekhl[1:] = 1 + love.ke[1:nl] - love.he[1:nl]
        he = np.concatenate(([0.0], love.he[1:nl]))
        le = np.concatenate(([0.0], love.le[1:nl]))
        a_hl, a_thl = horizontal_weights(love, norder)
        self._host = dict(nmod=np.ascontiguousarray(nmod, dtype=np.int32), s=np.ascontiguousarray(s),
                          a_sl=np.ascontiguousarray(a_sl), a_rr=np.ascontiguousarray(a_rr),
                          a2=np.ascontiguousarray(a2), ttl=ttl, ttlekhl=ttl * ekhl, ttlhe=ttl * he,
                          a_hl=np.ascontiguousarray(a_hl), a_thl=np.ascontiguousarray(a_thl), ttlle=ttl * le)
        h = self._host
        dp = lambda a: a.ctypes.data_as(_lib.P_double)  # noqa: E731
        prm = _lib.SlxSolverParams(
            lmax=self.norder, kmax=K, nmod2=nmod2, tpw=int(self.tpw), checkmarine=int(self.checkmarine),
            calcRG=int(self.calcRG), extrap=int(inner != "off"), epsilon1=self.epsilon1,
            radius=p.radius, mass=p.mass, rhoi=p.rhoi, rhow=p.rhow, gacc=p.gacc, omega=p.omega, moiA=p.moiA,
            moiC=p.moiC, kf=p.kf, kTE2=float(love.kTE[2]) if norder >= 2 else 0.0,
            hTE2=float(love.hTE[2]) if norder >= 2 else 0.0, kE2=float(love.ke[2]) if norder >= 2 else 0.0,
            nmod=h["nmod"].ctypes.data_as(_lib.P_int), s=dp(h["s"]), a_sl=dp(h["a_sl"]), a_rr=dp(h["a_rr"]),
            a_il=dp(h["a2"][0]), a_m=dp(h["a2"][1]), a_tsl=dp(h["a2"][2]), a_trr=dp(h["a2"][3]),
            ttl=dp(h["ttl"]), ttlekhl=dp(h["ttlekhl"]), ttlhe=dp(h["ttlhe"]), ramp=int(self.ramp),
            a_hl=dp(h["a_hl"]), a_thl=dp(h["a_thl"]), ttlle=dp(h["ttlle"]),
            lTE2=float(love.lTE[2]) if norder >= 2 else 0.0, anderson=int(inner == "anderson"),
            anderson_depth=int(_anderson[0]), anderson_freeze=int(_anderson[1]), max_mask_refresh=int(_anderson[2]),
            anderson_keep=int(bool(_anderson[3])), precision=_lib.PRECISIONS[precision],
            eps_switch=self.eps_switch, max_inner=int(_max_inner))
        g = lambda: torch.zeros((self.nlat, self.nlon), dtype=torch.float64, device=dev)  # noqa: E731
        self.topoxy, self.cxy, self.beta, self.rrxy, self.Gxy = g(), g(), g(), g(), g()
        outs = _lib.SlxOutputs(tgrid=self.topoxy.data_ptr(), ocean=self.cxy.data_ptr(), beta=self.beta.data_ptr(),
                               R=self.rrxy.data_ptr(), G=self.Gxy.data_ptr())
        handle = C.c_void_p()
        _lib.check(_lib.load().slx_solver_create(sht.handle, C.byref(prm), C.byref(outs), C.byref(handle)),
                   "slx_solver_create")
        self._s_handle = handle
        f = _lib.SlxFields()
        _lib.call("slx_solver_fields", handle, C.byref(f))
        self._fields = f
        topo = D.as_device(topo0, torch.float64, dev, (self.nlat, self.nlon))
        ice = D.as_device(ice0, torch.float64, dev, (self.nlat, self.nlon))
        _lib.call("slx_solver_init", handle, topo.data_ptr(), ice.data_ptr(), D.stream_ptr())
        _lib.call("slx_solver_set_dt", handle, self.dt)
        st = self._status()
        self.ocean_area0, self.oceanBeta_area0, self.ice_volume0 = st.ocean_area0, st.oceanBeta_area0, st.ice_volume0
        self.tinit = torch.empty_like(self.topoxy)
        _copy_dev(self.tinit, f.tinit, self.tinit.numel() * 8)
        self.tinit_0 = self.tinit
        self.cxy0 = torch.empty_like(self.topoxy)
        _copy_dev(self.cxy0, f.cxy0, self.cxy0.numel() * 8)
        self.beta0 = torch.empty_like(self.topoxy)
        _copy_dev(self.beta0, f.beta0, self.beta0.numel() * 8)
        self.iter = 0
        self.gr_counts = {"reguess": 0, "ice_analyses": 0}
        self._dt_hist = [self.dt, self.dt]  # lengths of steps n-1, n-2
        self._ice_buf = torch.empty((self.nlat, self.nlon), dtype=torch.float64, device=dev)
        self._dS = torch.empty((self.nl, self.nl), dtype=torch.complex128, device=dev)
        self._spec_tmp = torch.empty(sht.ncoef, dtype=torch.complex128, device=dev)
        self._guess_iters = 4
        self._recent = []  # inner-iteration counts of the last four steps, newest first
        self._reuse, self._reuse_store = "off", {}
        self.reuse_stats = {"warm_steps": 0}
        self.prof = {}
        self.state = _StateView(self)
        self.il = np.zeros((3, 3))
        self.mm = np.zeros(3)
        self.lam = np.zeros((3, 3), dtype=np.complex128)
        self._ro_bufs = None

    # ---------------------------------------------------------------------------- helpers
    def _readouts(self, rates, horizontal):
        """Device buffers of the read-outs (allocated once) and the slx_readouts block."""
        if self._ro_bufs is None:
            dev, n = self.dev, self.sht.ncoef
            g = lambda: torch.zeros((self.nlat, self.nlon), dtype=torch.float64, device=dev)  # noqa: E731
            c = lambda: torch.zeros(n, dtype=torch.complex128, device=dev)  # noqa: E731
            self._ro_bufs = {"dSLdt": g(), "dRdt": g(), "dSLdt_lm": c(), "dRdt_lm": c(),
                             "rot": torch.zeros(30, dtype=torch.float64, device=dev), "horiz": c()}
        b = self._ro_bufs
        ro = _lib.SlxReadouts()
        if rates:
            ro.dSLdt, ro.dRdt = b["dSLdt"].data_ptr(), b["dRdt"].data_ptr()
            ro.dSLdt_lm, ro.dRdt_lm = b["dSLdt_lm"].data_ptr(), b["dRdt_lm"].data_ptr()
            ro.rot_rates = b["rot"].data_ptr()
        if horizontal:
            ro.horiz_lm = b["horiz"].data_ptr()
        return ro
    def _status(self):
        st = _lib.SlxStatus()
        _lib.call("slx_solver_status", self._s_handle, C.byref(st), D.stream_ptr())
        return st

    def set_step_dt(self, dt):
        if dt is not None:
            self.dt = float(dt)
        _lib.call("slx_solver_set_dt", self._s_handle, self.dt)

    def __del__(self):
        h = getattr(self, "_s_handle", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.dev)
                _lib.load().slx_solver_destroy(h)
            except Exception:
                pass
            self._s_handle = None

    # ------------------------------------------------------------------------------- step
    def step(self, icexy_cur, profile=False, dt=None, rates=False):
        """Advance one time step (solver.py:814-1144); returns the reference's result dict.
        rates: also the analytic rates dSL/dt, dR/dt at the end of the step (grids in m/yr, dense
        spectra, and with rotation il-dot, m-dot, Lambda-dot; result key "rates")."""
        t_all = time.perf_counter()
        horizontal = bool(self.horizontal)
        ro = self._readouts(rates, horizontal) if (rates or horizontal) else None
        self.set_step_dt(dt)
        self.iter += 1
        if isinstance(icexy_cur, torch.Tensor) and icexy_cur.device == self.dev and \
                icexy_cur.dtype == torch.float64 and icexy_cur.is_contiguous():
            ice = icexy_cur
        else:
            ice = self._ice_buf
            ice.copy_(D.as_device(icexy_cur, torch.float64, self.dev, (self.nlat, self.nlon)))
        h = self._s_handle
        sp = D.stream_ptr()
        if profile:
            _lib.call("slx_solver_profile", h, 1)
        rs = self._reuse_store.get(self.iter) if self._reuse != "off" else None
        if rs is not None:  # warm start from the previous outer pass (solver_opt.py:400-402,515-528)
            guess = self._warm_guess(rs)
            _lib.call("slx_step_begin_warm", h, ice.data_ptr(), self.dt, self._dt_hist[0], self._dt_hist[1],
                      guess.data_ptr(), rs["cxy"].data_ptr(), sp)
            self.reuse_stats["warm_steps"] += 1
            if self._reuse == "warm+ice":
                self.reuse_stats["ice_recomputed"] += 1
        else:
            _lib.call("slx_step_begin", h, ice.data_ptr(), self.dt, self._dt_hist[0], self._dt_hist[1], sp)
        n = max([self._guess_iters] + self._recent)  # first batch: the largest count of the last four steps
        total = 0
        while True:
            _lib.call("slx_step_iterate", h, n, sp)
            total += n
            if ro is None:
                _lib.call("slx_step_finish", h, sp)
            else:
                _lib.call("slx_step_finish_readouts", h, C.byref(ro), sp)
            st = self._status()
            if st.failed:  # the step is not committed (its finish kernels skip a failed step)
                if profile:
                    _lib.call("slx_solver_profile", h, 0)
                raise InnerLoopError(f"inner loop failed to converge after {st.numiter} iterations, xi={st.xi}",
                                     status=_status_record(st))
            if st.finished:
                break
            n = 1 if total < 16 else max(1, total // 2)
        self._guess_iters = max(1, st.numiter)
        self._recent = ([st.numiter] + self._recent)[:4]
        self._last_status = st
        self._dt_hist = [self.dt, self._dt_hist[0]]
        self.gr_counts = {"reguess": st.reguess, "ice_analyses": st.ice_analyses}
        self.il = np.array(st.il).reshape(3, 3)
        self.mm = np.array(st.mm)
        lam = np.array(st.lam).reshape(3, 3, 2)
        self.lam = lam[..., 0] + 1j * lam[..., 1]
        _copy_dev(self._spec_tmp, self._fields.dS_converged, self.sht.ncoef * 16)
        self.sht.to_dense(self._spec_tmp, self._dS)
        if self._reuse != "off":  # store this step's converged state (solver_opt.py:932-941)
            ent = self._reuse_store.get(self.iter)
            if ent is None:
                ent = self._reuse_store[self.iter] = {
                    "dS": torch.empty(self.sht.ncoef, dtype=torch.complex128, device=self.dev),
                    "cxy": torch.empty_like(self.cxy)}
            elif self._reuse in ("warm+corr", "warm+corr2"):  # keep the previous pass's increment
                if "dS_old" not in ent:
                    ent["dS_old"] = torch.empty_like(ent["dS"])
                ent["dS_old"].copy_(ent["dS"])
            _copy_dev(ent["dS"], self._fields.load_inc, self.sht.ncoef * 16)
            ent["cxy"].copy_(self.cxy)
        rate_out = hz = None
        if rates:
            b = self._ro_bufs
            rate_out = {"dSLdt": b["dSLdt"], "dRdt": b["dRdt"], "dSLdt_lm": self.sht.to_dense(b["dSLdt_lm"]),
                        "dRdt_lm": self.sht.to_dense(b["dRdt_lm"])}
            if self.tpw and self.norder >= 2:
                r = b["rot"].cpu().numpy()
                rate_out.update({"ildot": r[:9].reshape(3, 3).copy(), "mdot": r[9:12].copy(),
                                 "lambdadot": r[12:30:2].reshape(3, 3) + 1j * r[13:30:2].reshape(3, 3)})
        if horizontal:
            hz = self.sht.to_dense(self._ro_bufs["horiz"])
        wall = time.perf_counter() - t_all
        if profile:
            self.prof = self._stage_times()
            self.prof["total"] = wall
        return {
            "iter": self.iter,
            "numiter": st.numiter,
            "tgrid": self.topoxy,
            "ocean": self.cxy,
            "beta": self.beta,
            "R": self.rrxy if self.calcRG else None,
            "G": self.Gxy if self.calcRG else None,
            "horiz": hz,
            "dS_converged": self._dS,
            "ocean_area": st.ocean_area,
            "oceanBeta_area": st.oceanBeta_area,
            "gmslc_ocn": st.gmslc_ocn,
            "gmslc_ocnBeta": st.gmslc_ocnBeta,
            "ice_volume": st.ice_volume,
            "il": self.il.copy(),
            "mm": self.mm.copy(),
            "lambda": self.lam.copy(),
            "rates": rate_out,
            "wall": wall,
            "prof": self.prof if profile else None,
        }

    def _warm_guess(self, rs):
        """Start increment of a warm step (solver_opt.py:515-528): the stored converged increment of
        the previous pass, plus for "warm+corr" the previous step's pass-to-pass change
        D_{n-1} = dS_{n-1}^(k) - dS_{n-1}^(k-1) and for "warm+corr2" 2 D_{n-1} - D_{n-2} (D_{n-1}
        alone while step n-2 has no previous-pass value); packed device arrays, elementwise."""
        if self._reuse not in ("warm+corr", "warm+corr2"):
            return rs["dS"]
        p1 = self._reuse_store.get(self.iter - 1)
        if p1 is None or "dS_old" not in p1:
            return rs["dS"]
        g = getattr(self, "_guess_buf", None)
        if g is None:
            g = self._guess_buf = torch.empty_like(rs["dS"])
        g.copy_(rs["dS"])
        d1 = p1["dS"] - p1["dS_old"]
        p2 = self._reuse_store.get(self.iter - 2) if self._reuse == "warm+corr2" else None
        if p2 is not None and "dS_old" in p2:
            g += 2.0 * d1 - (p2["dS"] - p2["dS_old"])
        else:
            g += d1
        self.reuse_stats["corr_steps"] = self.reuse_stats.get("corr_steps", 0) + 1
        return g

    def _stage_times(self):
        """Device time per stage of the step just profiled, in seconds, under the reference's keys
        (solver.py:829-831): memory, sht, spec, grid, rot; profiling is switched off again."""
        ms = (C.c_double * 5)()
        try:
            _lib.call("slx_solver_stage_times", self._s_handle, ms)
        finally:
            _lib.call("slx_solver_profile", self._s_handle, 0)
        return {k: ms[i] / 1e3 for i, k in enumerate(("memory", "sht", "spec", "grid", "rot"))}

    def run(self, ice, k0=1, k1=None, dts=None):
        """Steps k0..k1-1 with the ice grids ice[k] (a device float64 tensor [n][nlat][nlon], e.g.
        ``IceStack.grids``) through the native step loop (slx_run_steps: the per-step loop of
        run_ais.py:389-425 without per-step host outputs).  dts: the step lengths (default: the current
        dt for every step).  Returns one record of the step's scalar outputs per step; the solver's
        grids (tgrid, ocean, beta, R, G) hold the last step's."""
        if getattr(self, "_reuse", "off") != "off":
            raise NotImplementedError("run() does not keep the warm-start store; use step()")
        if not (isinstance(ice, torch.Tensor) and ice.is_cuda and ice.dtype == torch.float64 and ice.is_contiguous()
                and ice.dim() == 3 and tuple(ice.shape[1:]) == (self.nlat, self.nlon)):
            raise ValueError("ice must be a contiguous float64 device tensor [n][nlat][nlon]")
        k1 = ice.shape[0] if k1 is None else int(k1)
        n = k1 - k0
        if n <= 0:
            return []
        if k0 < 0 or k1 > ice.shape[0]:
            raise IndexError("steps outside the ice stack")
        d = np.full(n, self.dt) if dts is None else np.ascontiguousarray(dts, dtype=np.float64)
        if d.shape != (n,):
            raise ValueError(f"dts must have {n} entries")
        prev = np.array(self._dt_hist, dtype=np.float64)
        guess = C.c_int(self._guess_iters)
        out = (_lib.SlxStatus * n)()
        stride = self.nlat * self.nlon
        rc = _lib.load().slx_run_steps(self._s_handle, C.c_void_p(ice.data_ptr() + k0 * stride * 8), stride, n,
                                       d.ctypes.data_as(_lib.P_double), prev.ctypes.data_as(_lib.P_double),
                                       C.byref(guess), out, D.stream_ptr())
        if rc == -34:  # step k0 + j failed: steps before it finished and were committed
            j = next(i for i in range(n) if out[i].failed)
            if j > 0:
                self._advance_bookkeeping(out, d, j)
            self.iter += 1  # the failed step counts as started, as in the reference's step()
            raise InnerLoopError(_lib.load().slx_last_error().decode(errors="replace"),
                                 records=[_status_record(x) for x in out[:j]], status=_status_record(out[j]))
        _lib.check(rc, "slx_run_steps")
        return self._advance_bookkeeping(out, d, n, guess=guess.value)

    def _advance_bookkeeping(self, out, d, n, guess=None):
        """Host state after n finished steps of the native loop; returns their records."""
        outs = list(out)[:n]
        self.iter += n
        self._guess_iters = max(1, guess if guess is not None else outs[-1].numiter)
        self._recent = ([x.numiter for x in outs][::-1] + self._recent)[:4]
        self.dt = float(d[n - 1])
        self._dt_hist = [float(d[n - 1]), float(d[n - 2]) if n >= 2 else self._dt_hist[0]]
        st = outs[-1]
        self._last_status = st
        self.gr_counts = {"reguess": st.reguess, "ice_analyses": st.ice_analyses}
        rec = [_status_record(x) for x in outs]
        self.il, self.mm, self.lam = rec[-1]["il"], rec[-1]["mm"], rec[-1]["lambda"]
        return rec


class OptSeaLevelSolver(SeaLevelSolver):
    """``OptSeaLevelSolver`` signature (solver_opt.py:268-296).  The optimisation switches ``opt``
    (a set, or the --opt string) are validated as the reference does: unknown names raise, and
    "raw" must come with a raw-mode SHT, whose V = (-1)^m U convention the dense spectral outputs
    then use (the device path itself always runs the fused forms in the U convention); ``inner``
    may be "off", "extrap" or "anderson" (with anderson_depth 1..5, anderson_freeze,
    max_mask_refresh, anderson_keep; ``inner_stats`` counts the mixing outcomes over the run)."""

    def __init__(self, *args, opt=(), inner="off", anderson_depth=3, max_mask_refresh=3, anderson_freeze=2,
                 anderson_keep=False, reuse="off", reuse_store=None, **kw):
        if inner not in INNER_MODES:
            raise ValueError(f"inner must be one of {INNER_MODES}, got {inner!r}")
        if reuse not in REUSE_MODES:
            raise ValueError(f"reuse must be off|warm|warm+ice|warm+corr|warm+corr2, got {reuse!r}")
        self.opt = parse_opts(opt) if isinstance(opt, str) else set(opt)
        unknown = self.opt - set(ALL_OPTS)
        if unknown:
            raise ValueError(f"unknown opt {sorted(unknown)} (known: {','.join(ALL_OPTS)})")
        sht = args[0] if args else kw.get("sht")
        if ("raw" in self.opt) != bool(getattr(sht, "raw", False)):
            raise ValueError("--opt raw must be used with a raw-mode SHT backend")  # solver_opt.py:299-301
        self._raw_ok = "raw" in self.opt
        self.anderson_depth, self.anderson_freeze = int(anderson_depth), int(anderson_freeze)
        self.max_mask_refresh, self.anderson_keep = int(max_mask_refresh), bool(anderson_keep)
        super().__init__(*args, inner=inner, _anderson=(anderson_depth, anderson_freeze, max_mask_refresh,
                                                        anderson_keep), **kw)
        # cross-pass reuse (solver_opt.py:272-286): the caller-owned store keeps, per step, the
        # converged load increment (packed, device) and ocean function (device grid)
        self._reuse = reuse
        self._reuse_store = reuse_store if reuse_store is not None else {}
        self.reuse_store = self._reuse_store
        self.reuse_stats = {"warm_steps": 0, "ice_reused": 0, "ice_recomputed": 0}

    def step(self, icexy_cur, profile=False, dt=None, rates=False):
        r = super().step(icexy_cur, profile=profile, dt=dt, rates=rates)
        r["inner_stats"] = dict(self.inner_stats)
        return r

    @property
    def inner_stats(self):
        st = getattr(self, "_last_status", None)
        if st is None:
            return {"mixed": 0, "plain": 0, "fallback": 0, "singular": 0, "mask_refresh": 0, "frozen_iters": 0}
        return {"mixed": st.aa_mixed, "plain": st.aa_plain, "fallback": st.aa_fallback, "singular": st.aa_singular,
                "mask_refresh": st.aa_mask_refresh, "frozen_iters": st.aa_frozen_iters}


__all__ = ["SeaLevelSolver", "OptSeaLevelSolver", "InnerLoopError", "Planet", "PI", "EPS0", "GR_MAXACC",
           "parse_opts"]
