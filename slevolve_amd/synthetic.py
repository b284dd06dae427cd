"""Synthetic inputs: the reference's synthetic Earth and the benchmark workloads.

Input generation only -- nothing here is on the solver path.  Everything is seeded and
deterministic, so the CPU oracle and the GPU solver see identical arrays.

* ``synthetic_earth(nglv, topo_seed)``: the synthetic Earth of the reference
  (``tools/make_synthetic_case.py:75-190``): the 512 x 1024 topography seed sampled to the grid,
  two Vialov ice sheets (south pole, R 17 deg, H0 3000 m; 56 N 132 E, R 12 deg, H0 2500 m) melting
  over 8 steps, ice only where the bed is above -500 m, everything rounded to whole millimetres.
* ``small_case(gaulats, nlon)``: the idealised polar cap of ``tests/test_small_case.py:92-99``.
* ``love_file_text(...)``: a Love-number file in the reference layout
  (``python/slevolve/tools/write_lovenum.py:9-33``).
* ``workload(name)``: the benchmark configurations "lmax63" (64 x 128, 100 x 1 yr, 6 modes),
  "lmax512" (512 x 1024, 550 x 1 yr, 30 modes, polar cap + retreating marine sector) and
  "lmax1023" (1024 x 2048, 240 x 0.5 kyr, 30 modes, three caps over a glacial cycle) and
  "lmax256" (256 x 512, 300 x 1 yr, 30 modes, the lmax512 ice geometry; the adjoint case).
"""
from __future__ import annotations

import math

import numpy as np

D2R = np.pi / 180.0
POLAR = [3000.0, 2625.0, 2250.0, 1875.0, 1500.0, 1125.0, 750.0, 375.0, 0.0]
MIDLAT = [2500.0, 2400.0, 2150.0, 1800.0, 1350.0, 900.0, 500.0, 200.0, 0.0]
SMALL_H0 = [0.0, 300.0, 700.0, 1000.0, 1000.0, 800.0, 400.0, 150.0, 0.0]


def round_mm(f):
    return np.round(f * 1000.0) / 1000.0


def gl_lat_lon(nglv):
    """Latitudes (deg, north -> south) and longitudes (deg) of the nglv x 2 nglv GL grid."""
    x, _ = np.polynomial.legendre.leggauss(nglv)
    return np.arcsin(x[::-1]) / D2R, np.arange(2 * nglv) * 360.0 / (2 * nglv)


def angular_distance(LAT, LON, lat0, lon0):
    p, p0, dl = LAT * D2R, lat0 * D2R, (LON - lon0) * D2R
    c = np.sin(p) * np.sin(p0) + np.cos(p) * np.cos(p0) * np.cos(dl)
    return np.arccos(np.clip(c, -1.0, 1.0)) / D2R


def vialov(d, R):
    return np.clip(1.0 - (np.minimum(d, R) / R) ** (4.0 / 3.0), 0.0, None) ** 0.375


def sample_seed(seed, nglv):
    """Nearest-point sample of the 512 x 1024 seed on the nglv x 2 nglv grid."""
    ns = seed.shape[0]
    if nglv == ns:
        return np.array(seed, dtype=np.float64)
    lat_s, _ = gl_lat_lon(ns)
    lat, _ = gl_lat_lon(nglv)
    rows = np.array([int(np.argmin(np.abs(lat_s - la))) for la in lat])
    cols = (2 * np.arange(2 * nglv) * ns + nglv) // (2 * nglv) % (2 * ns)
    return np.ascontiguousarray(seed[rows][:, cols], dtype=np.float64)


def synthetic_earth(nglv, topo_seed):
    """(lat, lon, topo, [ice_0 .. ice_8]) of the reference's synthetic Earth."""
    lat, lon = gl_lat_lon(nglv)
    LON, LAT = np.meshgrid(lon, lat)
    topo = sample_seed(topo_seed, nglv)
    polar = vialov(90.0 + LAT, 17.0)
    mid = vialov(angular_distance(LAT, LON, 56.0, 132.0), 12.0)
    grounded = topo > -500.0
    ice = [round_mm((hp * polar + hm * mid) * grounded) for hp, hm in zip(POLAR, MIDLAT)]
    return lat, lon, topo, ice


def small_case(gaulats, nlon):
    """Sloping polar continent, -4000 m ocean, parabolic 14-degree cap grown and melted."""
    colat = np.degrees(np.arccos(gaulats))[:, None] * np.ones((1, nlon))
    topo = np.where(colat <= 20.0, 300.0 - 1100.0 * colat / 20.0, -800.0)
    topo = np.where(colat > 25.0, -4000.0,
                    np.where(colat > 20.0, -800.0 - 640.0 * (colat - 20.0), topo))
    shape = np.sqrt(np.clip(1.0 - (colat / 14.0) ** 2, 0.0, None))
    return topo, [h * shape for h in SMALL_H0]


# ------------------------------------------------------------------------ Love numbers
def _fmt(v):
    return " ".join(f"{x:.17e}" for x in np.atleast_1d(v))


def love_file_text(s, elastic, amp, tidal_elastic, tidal_amp, header="synthetic"):
    """Love-number file text in the reference layout.

    ``s[l]`` (K_l,) inverse relaxation times (1/kyr); ``elastic[l] = (h, l, k)``;
    ``amp[l]`` (3, K_l) modal amplitudes a_k of h, l, k (response h(t) = h^E + sum a (1 - e^{-s t}));
    the file stores residues r = a s, and l and k (elastic, fluid, residues) times the degree.
    ``tidal_*`` likewise for l >= 2.  Index 0 of the lists is unused.
    """
    L = len(s) - 1
    out = [header]
    for l in range(1, L + 1):
        K = len(s[l])
        out.append(f"{l} {K}")
        out.append(_fmt(s[l]))
        for el, a in ((elastic[l], amp[l]),) + (((tidal_elastic[l], tidal_amp[l]),) if l >= 2 else ()):
            h, ll, k = el
            fl = (h + a[0].sum(), ll + a[1].sum(), k + a[2].sum())
            out.append(_fmt([h, l * ll, 0.0, 0.0, l * k]))
            out.append(_fmt([fl[0], l * fl[1], 0.0, 0.0, l * fl[2]]))
            out.append(_fmt(a[0] * s[l]))
            out.append(_fmt(l * a[1] * s[l]))
            out.append(_fmt(l * a[2] * s[l]))
    return "\n".join(out) + "\n"


def synthetic_love_text(lmax, nmodes, seq_love, seq_tidal):
    """Seeded synthetic Earth response: h^E = -1/(1+l/20), k^E = -0.3/(1+l/10), l^E = 0.05/(1+l/30);
    nmodes modes per degree with s = 1000 * 10^U(-5,-1) per kyr (sorted) and amplitudes
    U(0,1) * elastic / nmodes, for loading and (l >= 2) tidal numbers."""
    rl = np.random.default_rng(seq_love)
    rt = np.random.default_rng(seq_tidal)
    s, el, amp, tel, tamp = [None], [None], [None], [None], [None]
    for l in range(1, lmax + 1):
        e = np.array([-1.0 / (1 + l / 20.0), 0.05 / (1 + l / 30.0), -0.3 / (1 + l / 10.0)])
        s.append(np.sort(1000.0 * 10.0 ** rl.uniform(-5.0, -1.0, nmodes)))
        el.append(e)
        amp.append(rl.uniform(0.0, 1.0, (3, nmodes)) * e[:, None] / nmodes)
        tel.append(e)
        tamp.append(rt.uniform(0.0, 1.0, (3, nmodes)) * e[:, None] / nmodes)
    return love_file_text(s, el, amp, tel, tamp, header=f"synthetic lmax {lmax} modes {nmodes}")


# ------------------------------------------------------------------- random topography
def random_sh_field(nlat, nlon, lcut, rng, power=1.5):
    """Sum_{l<=lcut, m<=l} c_lm P~_lm(mu) e^{i m phi} with |c_lm| ~ (l+1)^-power, Gaussian phases,
    evaluated on the nlat x 2 nlat GL grid (rows north -> south).  Input generation only."""
    mu, _ = np.polynomial.legendre.leggauss(nlat)
    mu = mu[::-1]
    amp = (np.arange(lcut + 1) + 1.0) ** (-power)
    F = np.zeros((nlat, nlon // 2 + 1), dtype=np.complex128)
    s2 = 1.0 - mu * mu
    pmm = np.full(nlat, math.sqrt(0.5))
    for m in range(lcut + 1):
        if m:
            pmm = pmm * np.sqrt(s2 * (2 * m + 1) / (2 * m))
        c = (rng.standard_normal(lcut + 1 - m) + 1j * rng.standard_normal(lcut + 1 - m)) / math.sqrt(2.0)
        if m == 0:
            c = c.real * math.sqrt(2.0)
        c = c * amp[m:]
        p0, acc = pmm.copy(), c[0] * pmm
        if lcut > m:
            p1 = math.sqrt(2 * m + 3) * mu * p0
            acc = acc + c[1] * p1
            for n in range(m + 2, lcut + 1):
                a = math.sqrt((4 * n * n - 1) / (n * n - m * m))
                b = math.sqrt(((n - 1) ** 2 - m * m) / (4 * (n - 1) ** 2 - 1))
                p0, p1 = p1, a * (mu * p1 - b * p0)
                acc = acc + c[n - m] * p1
        F[:, m] = acc
    return np.fft.irfft(F * nlon, n=nlon, axis=1)


def scaled_topography(nlat, lcut, rng, mean=-1500.0, std=2500.0):
    nlon = 2 * nlat
    g = random_sh_field(nlat, nlon, lcut, rng)
    _, w = np.polynomial.legendre.leggauss(nlat)
    W = w[::-1][:, None] * np.ones((1, nlon))
    m = (W * g).sum() / W.sum()
    sd = math.sqrt((W * (g - m) ** 2).sum() / W.sum())
    return round_mm(mean + std * (g - m) / sd)


# ------------------------------------------------------------------------- workloads
class Workload:
    """A seeded benchmark configuration: grid, truncation, steps, Love text, topography, ice(k)."""

    def __init__(self, name, nlat, lmax, nsteps, dt, nmodes, love_text, topo, ice_fn, note):
        self.name, self.nlat, self.nlon, self.lmax = name, nlat, 2 * nlat, lmax
        self.nsteps, self.dt, self.nmodes = nsteps, dt, nmodes
        self.love_text, self.topo, self._ice_fn, self.note = love_text, topo, ice_fn, note

    def ice(self, k):
        return self._ice_fn(k)


def _caps(LAT, LON, caps):
    """[(lat0, lon0, radius_deg, H0)] -> list of (shape, H0), shape = sqrt(1 - (d/R)^2)."""
    out = []
    for la, lo, R, H in caps:
        d = angular_distance(LAT, LON, la, lo)
        out.append((np.sqrt(np.clip(1.0 - (d / R) ** 2, 0.0, None)), H))
    return out


def workload(name):
    if name == "lmax63":
        cfg, nlat, lmax, nsteps, dt, K, lt = 1, 64, 63, 100, 1.0, 6, 32
    elif name == "lmax512":
        cfg, nlat, lmax, nsteps, dt, K, lt = 2, 512, 512, 550, 1.0, 30, 256
    elif name == "lmax1023":
        cfg, nlat, lmax, nsteps, dt, K, lt = 3, 1024, 1023, 240, 500.0, 30, 256
    elif name == "lmax256":  # the adjoint configuration: 256 x 512, lmax 256, 300 x 1 yr
        cfg, nlat, lmax, nsteps, dt, K, lt = 5, 256, 256, 300, 1.0, 30, 128
    else:
        raise ValueError(f"unknown workload {name!r}")
    seqs = np.random.SeedSequence(20261203 + cfg).spawn(5)
    love = synthetic_love_text(lmax, K, seqs[0], seqs[1])
    rng_topo = np.random.default_rng(seqs[2])
    rng_ice = np.random.default_rng(seqs[3])
    topo = scaled_topography(nlat, lt, rng_topo)
    lat, lon = gl_lat_lon(nlat)
    LON, LAT = np.meshgrid(lon, lat)
    if name == "lmax63":
        caps = _caps(LAT, LON, [(-80.0, 0.0, 12.0, 3000.0), (70.0, -60.0, 15.0, 2500.0)])

        def ice(k):
            f = 1.0 - k / nsteps
            return round_mm(sum(sh * H * f for sh, H in caps))
        note = "two caps (-80,0) R12 H3000 and (70,-60) R15 H2500, linear melt over 100 steps"
    elif name in ("lmax512", "lmax256"):
        # polar cap (R 20 deg, 3500 m) and a marine sector (lon 150-210, colat 20-32 deg) whose bed
        # is -500 +- 300 m; the sector's ice front retreats 2 km/yr towards the cap
        colat = 90.0 + LAT
        cap = np.sqrt(np.clip(1.0 - (colat / 20.0) ** 2, 0.0, None)) * 3500.0
        sector = (LON >= 150.0) & (LON <= 210.0) & (colat >= 20.0) & (colat <= 32.0)
        topo = np.where(sector, round_mm(-500.0 + 300.0 * rng_ice.uniform(-1.0, 1.0, topo.shape)), topo)
        km_per_deg = 6371.0 * D2R

        def ice(k):
            front = 32.0 - 2.0 * k * dt / km_per_deg
            x = np.clip((front - colat) / (front - 20.0), 0.0, 1.0)
            marine = np.where(sector & (colat <= front), 1500.0 * np.sqrt(x), 0.0)
            return round_mm(np.maximum(cap, marine))
        note = ("polar cap R20 H3500 + marine sector lon 150-210 colat 20-32, bed -500+-300 m, "
                "front retreating 2 km/yr")
        if name == "lmax256":
            note += "; adjoint misfit: 200 seeded ocean points and the J2 rate (seed stream 4)"
    else:
        caps = _caps(LAT, LON, [(-90.0, 0.0, 20.0, 3500.0), (60.0, -95.0, 12.0, 3000.0),
                                (65.0, 20.0, 8.0, 2000.0)])

        def ice(k):
            f = math.sin(math.pi * k / nsteps)
            return round_mm(sum(sh * H * f for sh, H in caps))
        note = "three caps grown and melted as sin(pi t/T), peaks 3.5, 3.0, 2.0 km"
    return Workload(name, nlat, lmax, nsteps, dt, K, love, topo, ice, note)
