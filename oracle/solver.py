"""Oracle FP64 sea-level time step.  Test infrastructure only (see ``oracle/__init__.py``).

Restates the reference's exact-memory solver:

* ``SeaLevelSolver.__init__`` step-0 set-up (``python/slevolve/solver.py:521-605``);
* ``SeaLevelSolver.step`` with ``axial_rotation`` off (solver.py:814-1144): modal advance,
  grounded-ice projection on the previous converged topography, eustatic first guess, the inner
  fixed point (rotation, eq. B18, synthesis, rO, analysis, conservation constant, ocean-load
  update, convergence ratio, ocean-function and grounded-ice re-guess), radial displacement,
  commit and the degree-0 diagnostics;
* the optional second-order load placement ``ramp`` (solver.py:904-988,1065-1068), the horizontal
  read-out ``horizontal`` (solver.py:687-720) and the analytic rates ``step(rates=True)``
  (solver.py:722-806);
* ``inner="extrap"``: the extrapolated inner start of ``OptSeaLevelSolver``
  (``solver_opt.py:148-165,504-527,817-822``);
* ``inner="anderson"``: the same start, then Anderson mixing on the load increment with the ocean
  mask frozen after ``anderson_freeze`` plain iterations, and at convergence the mask re-checked
  (at most ``max_mask_refresh`` refreshes) (``solver_opt.py:532-535,705-775``).

Constants are those of ``solver.py:31-39,98-110``.
"""
from __future__ import annotations

import math

import numpy as np

from .anderson import Anderson
from .memory import ExactMemory

PI = 3.1415926535898
EPS0 = float(np.finfo(np.float64).eps)
GR_MAXACC = 3
MAX_INNER = 9999


class Planet:
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


def extrap_guess(h1, h2, dt0, dt1, dt2):
    """dS0_j = dt_j [Sdot_{j-1} + (Sdot_{j-1} - Sdot_{j-2}) (dt_j + dt_{j-1})/(dt_{j-1} + dt_{j-2})],
    Sdot_i = dS_i/dt_i; equal steps give exactly 2 h1 - h2 (solver_opt.py:148-165)."""
    if dt0 == dt1 == dt2:
        return h1 * 2.0 - h2
    w = (dt0 + dt1) / (dt1 + dt2)
    return h1 * (dt0 / dt1 * (1.0 + w)) - (dt0 / dt2 * w) * h2


def convergence_ratio(sa, so):
    """xi of solver.py:1025-1030 with its EPS0 guards."""
    if abs(sa - so) < EPS0 and abs(so) < EPS0:
        return 0.0
    if abs(so) < EPS0:
        return abs((sa - so) / (EPS0 * 10))
    return abs((sa - so) / so)


REUSE_MODES = ("off", "warm", "warm+ice", "warm+corr", "warm+corr2")


def reuse_guess(store, it, mode):
    """Warm-start load increment of step it (solver_opt.py:515-528), or None without a stored entry:
    the stored converged increment of the previous pass, dS_n^(k-1), plus for "warm+corr" the
    previous step's pass-to-pass change D_{n-1} = dS_{n-1}^(k) - dS_{n-1}^(k-1) and for
    "warm+corr2" 2 D_{n-1} - D_{n-2} (D_{n-1} alone when step n-2 has no previous-pass value).
    Returns (guess, corrected)."""
    rs = store.get(it) if mode != "off" else None
    if rs is None:
        return None, False
    g = rs["dS"].copy()
    if mode in ("warm+corr", "warm+corr2"):
        p1 = store.get(it - 1)
        if p1 is not None and "dS_old" in p1:
            d1 = p1["dS"] - p1["dS_old"]
            p2 = store.get(it - 2) if mode == "warm+corr2" else None
            if p2 is not None and "dS_old" in p2:
                g += 2.0 * d1 - (p2["dS"] - p2["dS_old"])
            else:
                g += d1
            return g, True
    return g, False


def reuse_store_step(store, it, mode, dS, cxy):
    """Keeps step it's converged increment and ocean function (solver_opt.py:932-941); the corr
    modes keep the previous pass's increment as dS_old."""
    ent = store.get(it)
    if ent is None:
        store[it] = {"dS": dS.copy(), "cxy": cxy != 0.0}
        return
    if mode in ("warm+corr", "warm+corr2"):
        ent["dS_old"] = ent["dS"].copy()
    ent["dS"] = dS.copy()
    ent["cxy"] = cxy != 0.0


class SeaLevelSolver:
    """The step of solver.py:521-1144 with the inner starts of solver_opt.py ("off", "extrap") and
    the cross-pass reuse of solver_opt.py:272-286: with reuse="warm" and a caller-owned
    reuse_store dict (keyed by step number), step n starts from the ocean function and the
    converged load increment that step n stored in the previous outer pass
    (solver_opt.py:400-402,517-528) and stores its own at the end (solver_opt.py:932-941);
    "warm+corr" / "warm+corr2" add the previous steps' pass-to-pass changes to that start
    (reuse_guess); "warm+ice" starts as "warm" (its stored ice analysis equals the recomputed one
    whenever the reference would reuse it: same grounded support, same ice)."""

    def __init__(self, sht, love, topo0, ice0, *, norder, dt=1.0, tpw=True, checkmarine=True,
                 calcRG=True, epsilon1=1.0e-5, inner="off", planet=None, reuse="off", reuse_store=None,
                 ramp=False, anderson_depth=3, max_mask_refresh=3, anderson_freeze=2, anderson_keep=False):
        if inner not in ("off", "extrap", "anderson"):
            raise ValueError("inner must be 'off', 'extrap' or 'anderson'")
        self.aa = Anderson(anderson_depth) if inner == "anderson" else None
        self.max_mask_refresh = int(max_mask_refresh)
        self.anderson_freeze = int(anderson_freeze)
        self.anderson_keep = bool(anderson_keep)
        self.inner_stats = {"mixed": 0, "plain": 0, "fallback": 0, "singular": 0, "mask_refresh": 0,
                            "frozen_iters": 0}
        if reuse not in REUSE_MODES:
            raise ValueError(f"reuse must be one of {REUSE_MODES}")
        self.reuse = reuse
        self.reuse_store = reuse_store if reuse_store is not None else {}
        self.sht, self.love = sht, love
        self.L = norder
        self.nl = norder + 1
        self.dt = float(dt)
        self.tpw, self.checkmarine, self.calcRG = tpw, checkmarine, calcRG
        self.epsilon1 = epsilon1
        self.inner = inner
        self.p = planet or Planet()
        p = self.p
        self.tinit = np.array(topo0, dtype=np.float64)
        self.cxy0 = (self.tinit < 0).astype(np.float64)
        g0 = self._grounded(ice0, self.tinit) if checkmarine else np.asarray(ice0, np.float64)
        self.beta0 = (g0 < EPS0).astype(np.float64)
        self.cstar0 = self.cxy0 * self.beta0
        self.cxy = self.cxy0.copy()
        self.topo_prev = self.tinit.copy()
        area = 4 * PI * p.radius**2
        self.ocean_area0 = sht.analysis(self.cxy0)[0, 0].real * area
        self.oceanBeta_area0 = sht.analysis(self.cstar0)[0, 0].real * area
        self.ice_volume0 = sht.analysis(g0)[0, 0].real * area
        self.state = ExactMemory(love, norder, self.dt)
        self.ramp = bool(ramp)
        self.state.ramp = self.ramp
        self.horizontal = False  # set True for the horizontal read-out (result key "horiz")
        nl = self.nl
        ell = np.arange(nl, dtype=np.float64)
        ttl = np.zeros(nl)
        ttl[1:] = (4.0 * PI * p.radius**3) / (p.mass * (2 * ell[1:] + 1.0))
        ekhl = np.zeros(nl)
        ekhl[1:] = 1 + love.ke[1:nl] - love.he[1:nl]
        he = np.zeros(nl)
        he[1:] = love.he[1:nl]
        le = np.zeros(nl)
        le[1:] = love.le[1:nl]
        self.ttl = ttl[:, None]
        self.ttlekhl = (ttl * ekhl)[:, None]
        self.ttlhe = (ttl * he)[:, None]
        self.ttlle = (ttl * le)[:, None]
        self.il = np.zeros((3, 3))
        self.mm = np.zeros(3)
        self.lam = np.zeros((3, 3), dtype=np.complex128)
        self.beta = np.zeros_like(self.tinit)
        self.icestarxy = np.zeros_like(self.tinit)
        self.ice_prev = np.array(ice0, dtype=np.float64)
        self.iter = 0
        self.hist = []  # converged load_inc of the last two steps (newest first), with dt
        self.gr_counts = {"reguess": 0, "ice_analyses": 0}

    def _grounded(self, ice, topo):
        """I* = ice where (topo > 0) or ice > |topo| rho_w / rho_i, else 0 (solver.py:655-660)."""
        p = self.p
        ice = np.asarray(ice, dtype=np.float64)
        keep = (topo > 0) | (ice > np.abs(topo) * p.rhow / p.rhoi)
        return ice * keep

    def set_step_dt(self, dt):
        if dt is not None:
            self.dt = float(dt)
            self.state.set_dt(self.dt)

    # ------------------------------------------------------------------------------------
    def horizontal_lm(self, dI, dS_prev, load_inc, dice, rot_tot):
        """Psi_lm = T_l [l^E_l A_lm + sum_k (l_k/s_k) R_k,lm] (+ the ramp's same-step term and the
        degree-2 rotation with the tidal l numbers), U convention (solver.py:687-720)."""
        p, st = self.p, self.state
        if st.a_hl is None:
            st.horizontal_weights(self.love)
        visc_h, thl = st.read_horizontal()
        A = p.rhoi * dI + p.rhow * dS_prev
        A += p.rhow * load_inc
        hz = self.ttlle * A + self.ttl * visc_h
        if self.ramp:
            c_hl = (st.a_hl * st.omphi).sum(0)
            hz += (self.ttl[:, 0] * c_hl)[:, None] * (p.rhoi * dice + p.rhow * load_inc)
        hz[0, :] = 0.0
        if self.tpw and rot_tot is not None and self.L >= 2:
            if self.ramp:
                c_thl = float(st.omphi[:st.nmod2, 2] @ st.a_thl[:st.nmod2])
                thl = thl + c_thl * (rot_tot - st.p_lambda)
            hz[2, 0:3] += (self.love.lTE[2] * rot_tot + thl) / p.gacc
        return hz

    def _rotation_rate(self, old_il, old_m, lam_prev2):
        """il-dot, m-dot (Liouville equations differentiated, the ramp's current-increment term as
        s_k (m - r_m,k)), Lambda-dot by the chain rule, and the rotational sea-level / radial rates
        at degree 2 (solver.py:722-767).  U convention."""
        p, st = self.p, self.state
        h = float(self.dt)
        dil = self.il - old_il
        dm = self.mm - old_m
        tot = self.lam[2, 0:3].copy()
        r_il, r_m, r_la, sk = st.read_rate_rotation(dil, dm, tot - lam_prev2)
        n2 = st.nmod2
        kTE2, hTE2, kE2 = self.love.kTE[2], self.love.hTE[2], self.love.ke[2]
        ild = dil / h
        sild = np.zeros((3, 3))
        smd = np.zeros(3)
        tsld = np.zeros(3, dtype=np.complex128)
        trrd = np.zeros(3, dtype=np.complex128)
        for k in range(n2):
            sild += (st.a_il[k] * sk[k]) * (self.il - r_il[k])
            smd += (st.a_m[k] * sk[k]) * (self.mm - r_m[k])
            tsld += (st.a_tsl[k] * sk[k]) * (tot - r_la[k])
            trrd += (st.a_trr[k] * sk[k]) * (tot - r_la[k])
        alpha = 1 / (1 - kTE2 / p.kf)
        md = np.zeros(3)
        for i in (0, 1):
            md[i] = alpha * ((1 / (p.moiC - p.moiA)) * ((1 + kE2) * ild[i, 2] + sild[i, 2]) + smd[i] / p.kf)
        md[2] = -(1.0 / p.moiC) * ((1 + kE2) * ild[2, 2] + sild[2, 2])
        m1, m2, m3 = self.mm
        d1, d2, d3 = md
        r2w2 = p.radius**2 * p.omega**2
        lad = np.zeros((3, 3), dtype=np.complex128)
        lad[0, 0] = (r2w2 / 3.0) * (2 * m1 * d1 + 2 * m2 * d2 + 2 * m3 * d3 + 2.0 * d3)
        lad[2, 0] = (r2w2 / (6.0 * math.sqrt(5.0))) * (2 * m1 * d1 + 2 * m2 * d2 - 4 * m3 * d3 - 4 * d3)
        lad[2, 1] = -(r2w2 / math.sqrt(30.0)) * ((d1 - 1j * d2) * (1.0 + m3) + (m1 - 1j * m2) * d3)
        lad[2, 2] = (r2w2 / math.sqrt(5.0 * 24.0)) * (2 * m2 * d2 - 2 * m1 * d1 + 2.0 * 1j * (d1 * m2 + m1 * d2))
        rate_sl = ((1 + kTE2 - hTE2) * lad[2, 0:3] + tsld) / p.gacc
        rate_rr = (hTE2 * lad[2, 0:3] + trrd) / p.gacc
        return rate_sl, rate_rr, {"ildot": ild, "mdot": md, "lambdadot": lad}

    def _rate_spectra(self, cur, old_il, old_m, lam_prev2):
        """dSL/dt (degrees >= 1) and dR/dt spectra at the end of the step, before the commit
        (solver.py:769-790)."""
        h = float(self.dt)
        rv_sl, rv_rr = self.state.read_rate(cur)
        dsl = self.ttlekhl * cur / h + self.ttl * rv_sl
        drr = self.ttlhe * cur / h + self.ttl * rv_rr
        dsl[0, :] = 0.0
        drr[0, :] = 0.0
        rot = None
        if self.tpw and self.L >= 2:
            rs, rr, rot = self._rotation_rate(old_il, old_m, lam_prev2)
            dsl[2, 0:3] += rs
            drr[2, 0:3] += rr
        return dsl, drr, rot

    def _rate_grids(self, dsl, drr, dice00, rot):
        """Rate grids; degree 0 of dSL/dt from the conservation constant on the end-of-step ocean
        function C* = C beta (solver.py:792-806)."""
        p, sht = self.p, self.sht
        g = sht.synthesis(dsl)
        cst = self.cxy * self.beta
        a = sht.analysis(g * cst)[0, 0].real
        c00 = sht.analysis(cst)[0, 0].real
        dsl[0, 0] = (-(p.rhoi / p.rhow) * dice00.real / float(self.dt) - a) / c00
        out = {"dSLdt": sht.synthesis(dsl), "dRdt": sht.synthesis(drr), "dSLdt_lm": dsl, "dRdt_lm": drr}
        if rot is not None:
            out.update(rot)
        return out

    def step(self, ice_cur, dt=None, rates=False):
        self.set_step_dt(dt)
        p, sht, st = self.p, self.sht, self.state
        self.iter += 1
        ice_cur = np.asarray(ice_cur, dtype=np.float64)
        visc, rad, sum_il, sum_m, sum_tsl, sum_trr = st.advance_and_read()
        dS_prev = st.p_ocean.copy()
        old_il, old_m, lam_prev2 = st.p_il.copy(), st.p_m.copy(), st.p_lambda.copy()
        rs = self.reuse_store.get(self.iter) if self.reuse != "off" else None
        if rs is not None:
            self.cxy = rs["cxy"].astype(np.float64)  # converged mask of this step, previous pass

        # ---- ice: grounded projection on the previous converged topography
        if self.checkmarine:
            self.icestarxy = self._grounded(ice_cur, self.topo_prev)
        else:
            self.icestarxy = ice_cur.copy()
        icestar_lm = sht.analysis(self.icestarxy)
        if self.iter == 1:
            src = self._grounded(self.ice_prev, self.tinit) if self.checkmarine else self.ice_prev
            oldice = sht.analysis(src)
        else:
            oldice = st.prev_ice.copy()
        dice = icestar_lm - oldice
        dI = st.p_ice + dice

        # ---- ocean function on the last converged mask
        self.beta = (self.icestarxy < EPS0).astype(np.float64)
        cstar = self.cxy * self.beta
        cstar_lm = sht.analysis(cstar)
        tO = self.tinit * (cstar - self.cstar0)
        tO_lm = sht.analysis(tO)

        # ---- first guess
        load_inc = (cstar_lm / cstar_lm[0, 0]) * ((-p.rhoi / p.rhow) * dice[0, 0])
        dS_cur = np.zeros_like(load_inc)
        if self.inner != "off" and len(self.hist) >= 2:
            (h1, dt1), (h2, dt2) = self.hist[0], self.hist[1]
            load_inc = extrap_guess(h1, h2, self.dt, dt1, dt2)
            dS_cur = dS_prev + load_inc
        if rs is not None:  # warm start: the stored converged increment replaces the guess
            load_inc, _ = reuse_guess(self.reuse_store, self.iter, self.reuse)
            dS_cur = dS_prev + load_inc

        kTE2, hTE2, kE2 = self.love.kTE[2], self.love.hTE[2], self.love.ke[2]
        r2w2 = p.radius**2 * p.omega**2
        dil = np.zeros((3, 3))
        dm = np.zeros(3)
        dlam = np.zeros(3, dtype=np.complex128)
        dsl_rot = np.zeros(3, dtype=np.complex128)
        rr_rot = np.zeros(3, dtype=np.complex128)
        n_acc = 0
        numiter = 1
        frozen = refreshing = False
        n_refresh = 0
        if self.aa is not None:
            self.aa.reset()
        while True:
            if self.tpw:
                # products of inertia from the degree-2 load (solver.py:912-921)
                z20 = p.rhow * dS_cur[2, 0] + p.rhoi * dI[2, 0]
                z21 = p.rhow * dS_cur[2, 1] + p.rhoi * dI[2, 1]
                c4 = 8 * PI * p.radius**4
                self.il[2, 2] = (-c4 / (3 * math.sqrt(5.0))) * z20.real
                self.il[0, 2] = (-c4 / math.sqrt(30.0)) * z21.real
                self.il[1, 2] = (c4 / math.sqrt(30.0)) * (-1j * z21).real
                dil = self.il - old_il
                if self.ramp:
                    # current increments carry 1 - phi of their viscous response: implicit in m
                    # (solver.py:922-933)
                    sil_r = sum_il + st.c_il * dil
                    al = 1 / (1 - kTE2 / p.kf)
                    for i in (0, 1):
                        self.mm[i] = al * (
                            (1 / (p.moiC - p.moiA)) * ((1 + kE2) * self.il[i, 2] + sil_r[i, 2])
                            + (1 / p.kf) * (sum_m[i] - st.c_m * old_m[i])) / (1 - al * st.c_m / p.kf)
                    self.mm[2] = -(1.0 / p.moiC) * ((1 + kE2) * self.il[2, 2] + sil_r[2, 2])
                else:
                    # Liouville equations (solver.py:940-946)
                    for i in (0, 1):
                        self.mm[i] = (1 / (1 - kTE2 / p.kf)) * (
                            (1 / (p.moiC - p.moiA)) * ((1 + kE2) * self.il[i, 2] + sum_il[i, 2])
                            + (1 / p.kf) * sum_m[i])
                    self.mm[2] = -(1.0 / p.moiC) * ((1 + kE2) * self.il[2, 2] + sum_il[2, 2])
                dm = self.mm - old_m
                m1, m2, m3 = self.mm
                self.lam[0, 0] = (r2w2 / 3.0) * ((m1**2 + m2**2 + m3**2) + 2.0 * m3)
                self.lam[2, 0] = (r2w2 / (6.0 * math.sqrt(5.0))) * (m1**2 + m2**2 - 2.0 * m3**2 - 4.0 * m3)
                self.lam[2, 1] = -(r2w2 / math.sqrt(30.0)) * ((m1 - 1j * m2) * (1.0 + m3))
                self.lam[2, 2] = (r2w2 / math.sqrt(5.0 * 24.0)) * (m2**2 - m1**2 + 2.0 * 1j * m1 * m2)
                dlam = self.lam[2, 0:3] - lam_prev2
                tot = lam_prev2 + dlam
                tsl_v, trr_v = sum_tsl, sum_trr
                if self.ramp:
                    tsl_v = sum_tsl + st.c_tsl * dlam
                    trr_v = sum_trr + st.c_trr * dlam
                dsl_rot = ((1 + kTE2 - hTE2) * tot + tsl_v) / p.gacc
                rr_rot = (hTE2 * tot + trr_v) / p.gacc if self.calcRG else np.zeros(3, complex)

            # ---- eq. B18 and the grid round trip
            A = (p.rhoi * dI + p.rhow * dS_prev) + p.rhow * load_inc
            sl_inc = A * self.ttlekhl + visc * self.ttl
            if self.ramp:  # same-step viscous weight of the current increment (solver.py:986-989)
                sl_inc += (self.ttl[:, 0] * st.c_sl)[:, None] * (p.rhoi * dice + p.rhow * load_inc)
            if self.tpw:
                sl_inc[2, 0:3] += dsl_rot
            dslxy = sht.synthesis(sl_inc)
            rO = dslxy * cstar
            rO_lm = sht.analysis(rO)
            conserv = ((1 / cstar_lm[0, 0])
                       * (-(p.rhoi / p.rhow) * dI[0, 0] - rO_lm[0, 0] + tO_lm[0, 0])).real
            old = load_inc
            load_inc = ((-dS_prev + rO_lm) + cstar_lm * conserv) - tO_lm
            dS_cur = dS_prev + load_inc
            sl_tot = sl_inc.copy()
            sl_tot[0, 0] += conserv
            deltasl = sht.synthesis(sl_tot)

            xi = convergence_ratio(np.abs(load_inc).sum(), np.abs(old).sum())
            if self.aa is not None:
                # Anderson mixing with the ocean mask frozen (solver_opt.py:705-775)
                if xi <= self.epsilon1:
                    # converged on the frozen mask: accept if this iterate's own mask equals it
                    same = np.array_equal(self.tinit - deltasl < 0.0, self.cxy != 0.0)
                    if same or n_refresh >= self.max_mask_refresh:
                        break
                    n_refresh += 1
                    self.inner_stats["mask_refresh"] += 1
                    refreshing = True
                elif frozen:
                    self.aa.push(old, load_inc)
                    load_inc, kind = self.aa.mix()
                    self.inner_stats[kind] += 1
                    self.inner_stats["frozen_iters"] += 1
                    dS_cur = dS_prev + load_inc
                    if numiter == MAX_INNER:
                        raise RuntimeError(f"inner loop failed to converge, xi={xi}")
                    numiter += 1
                    continue
                # plain step (load_inc = G(x)) with the mask update, then freeze
                if refreshing:
                    if not self.anderson_keep:
                        self.aa.reset()
                    refreshing = False
                    frozen = True
                elif numiter >= self.anderson_freeze:
                    self.aa.reset()
                    self.aa.push(old, load_inc)
                    frozen = True
            if xi > self.epsilon1 or self.aa is not None:
                topo = self.tinit - deltasl
                self.cxy = (topo < 0.0).astype(np.float64)
                if self.checkmarine and n_acc < GR_MAXACC:
                    # grounded-ice re-guess on the topography iterate (solver.py:661-684)
                    cand = self._grounded(ice_cur, topo)
                    bnew = cand < EPS0
                    self.gr_counts["reguess"] += 1
                    if np.count_nonzero(bnew != (self.beta != 0.0)):
                        self.icestarxy = cand
                        self.beta = bnew.astype(np.float64)
                        icestar_lm = sht.analysis(self.icestarxy)
                        dice = icestar_lm - oldice
                        dI = st.p_ice + dice
                        self.gr_counts["ice_analyses"] += 1
                        n_acc += 1
                cstar = self.cxy * self.beta
                cstar_lm = sht.analysis(cstar)
                tO = self.tinit * (cstar - self.cstar0)
                tO_lm = sht.analysis(tO)
            else:
                break
            if numiter == MAX_INNER:
                raise RuntimeError(f"inner loop failed to converge, xi={xi}")
            numiter += 1

        if self.inner != "off":
            self.hist = [(load_inc.copy(), self.dt)] + self.hist[:1]

        rrxy = None
        if self.calcRG:
            A = (p.rhoi * dI + p.rhow * dS_prev) + p.rhow * load_inc
            rr_lm = A * self.ttlhe + rad * self.ttl
            if self.ramp:
                rr_lm += (self.ttl[:, 0] * st.c_rr)[:, None] * (p.rhoi * dice + p.rhow * load_inc)
            if self.tpw:
                rr_lm[2, 0:3] += rr_rot
            rrxy = sht.synthesis(rr_lm)

        hz = None
        if self.horizontal:
            hz = self.horizontal_lm(dI, dS_prev, load_inc, dice, lam_prev2 + dlam if self.tpw else None)

        # ---- commit (row l = 0 of the load excluded, solver.py:1080-1095)
        cur = p.rhoi * dice + p.rhow * load_inc
        cur[0, :] = 0.0
        rate_sp = None
        if rates:
            rate_sp = self._rate_spectra(cur, old_il, old_m, lam_prev2)
            dice00 = complex(dice[0, 0])
        if self.tpw:
            st.commit(cur, dil, dm, dlam)
        else:
            st.commit(cur)
        st.p_ice = dI.copy()
        st.p_ocean = dS_cur.copy()
        st.prev_ice = icestar_lm.copy()

        topo = self.tinit - deltasl
        self.cxy = (topo < 0.0).astype(np.float64)
        self.topo_prev = topo.copy()
        if self.reuse != "off":
            reuse_store_step(self.reuse_store, self.iter, self.reuse, load_inc, self.cxy)
        area = 4 * PI * p.radius**2
        slc = deltasl * self.cxy
        gmslc_ocn = (sht.analysis(slc)[0, 0] / sht.analysis(self.cxy)[0, 0]).real
        clm00 = sht.analysis(self.cxy)[0, 0].real
        cst = self.cxy * self.beta
        cstar00 = sht.analysis(cst)[0, 0].real
        gmslc_beta = (sht.analysis(slc * self.beta)[0, 0].real) / cstar00
        self.ice_prev = ice_cur
        rate_out = self._rate_grids(*rate_sp[:2], dice00, rate_sp[2]) if rates else None
        return {
            "iter": self.iter,
            "numiter": numiter,
            "tgrid": topo,
            "ocean": self.cxy.copy(),
            "beta": self.beta.copy(),
            "R": rrxy,
            "G": (deltasl + rrxy) if rrxy is not None else None,
            "horiz": hz,
            "rates": rate_out,
            "dS_converged": dS_cur,
            "ocean_area": clm00 * area,
            "oceanBeta_area": cstar00 * area,
            "gmslc_ocn": gmslc_ocn,
            "gmslc_ocnBeta": gmslc_beta,
            "ice_volume": icestar_lm[0, 0].real * area,
            "il": self.il.copy(),
            "mm": self.mm.copy(),
            "lambda": self.lam.copy(),
        }
