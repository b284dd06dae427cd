"""Oracle Gauss-Legendre spherical-harmonic transform in the legacy ``spharmt`` conventions.

Test infrastructure only (see ``oracle/__init__.py``).  Restates ``python/slevolve/sht.py``:

* grid ``g[j, i]``: GL rows north -> south, longitude fastest (sht.py:8-12);
* orthonormal P~_lm without Condon-Shortley phase, by the ``LEGEND`` recursion (sht.py:105-144);
* analysis   ``U[l, m] = sum_j P~_lm(mu_j) w_j rfft(g)[j, m] / nlon / 1.41421356`` (sht.py:274-302);
* synthesis  ``g = irfft(nlon * sum_l 1.41421356 U[l, m] P~_lm, n=nlon)`` (sht.py:305-343);
  ``irfft`` drops Im of the m = 0 and Nyquist bins.

The implementation is the plain per-order form (one small matrix product per m, hemispheres
folded by the parity of l - m), not the chunked BLAS layout of the reference.
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft

SQRT2_LEGACY = 1.41421356


def gauss_legendre(nlat: int):
    """Nodes mu_j (north -> south, mu_0 > 0) and weights of the nlat-point GL rule.

    Newton iteration on P_nlat from the Tricomi start cos(pi (j - 1/4) / (nlat + 1/2)),
    at most 10 sweeps, |dz| < 1e-15 (``gaulw``, sht.py:65-101).
    """
    pi = 4.0 * np.arctan(1.0)
    z = np.cos(pi * (np.arange(1, nlat + 1, dtype=np.float64) - 0.25) / (nlat + 0.5))
    nodes = np.zeros(nlat)
    deriv = np.zeros(nlat)
    done = np.zeros(nlat, dtype=bool)
    for _ in range(10):
        pn, pnm1 = np.ones(nlat), np.zeros(nlat)
        for n in range(1, nlat + 1):
            pn, pnm1 = ((2.0 * n - 1.0) * z * pn - (n - 1.0) * pnm1) / n, pn
        dp = nlat * (z * pn - pnm1) / (z * z - 1.0)
        znew = z - pn / dp
        conv = np.abs(znew - z) < 1e-15
        first = conv & ~done
        nodes[first] = znew[first]
        deriv[first] = dp[first]
        done |= conv
        z = znew
        if done.all():
            break
    if not done.all():
        raise RuntimeError("Gauss-Legendre nodes did not converge")
    w = 2.0 / ((1.0 - nodes * nodes) * deriv * deriv)
    return nodes, w


SCALE_K = 600  # one scale unit: 2^600
SCALE_T = 300  # a scaled value is renormalised once it exceeds 2^300


def sectoral_starts(x: np.ndarray, lmax: int):
    """Extended-range sectoral start values of ``legendre_block``: for m = 0..lmax the arrays
    (p0[m], sc[m]) with P~_mm(x_j) = p0 2^(-600 sc) (sc = 0: the value itself)."""
    x = np.asarray(x, dtype=np.float64)
    s2 = 1.0 - x * x
    prod = np.ones_like(x)
    E = np.zeros(x.size, dtype=np.int64)
    a, b = 1.0, 0.0
    p0s, scs = [], []
    for m in range(lmax + 1):
        if m:
            a += 2.0
            b += 2.0
            prod = prod * s2 * a / b
            low = prod < 2.0**-500
            prod[low] *= 2.0**500
            E[low] -= 500
        base = np.sqrt(0.5 * prod)
        e2 = E // 2
        lg = np.frexp(base)[1] - 1 + e2  # floor(log2 |P~_mm|)
        sc = np.maximum(0, np.ceil((-SCALE_T - lg) / SCALE_K)).astype(np.int64)
        p0s.append(np.ldexp(base, e2 + SCALE_K * sc))
        scs.append(sc)
    return p0s, scs


def legendre_block(x: np.ndarray, lmax: int):
    """List over m = 0..lmax of arrays ``P[m][l - m, j] = P~_lm(x_j)``.

    Sectoral start P~_mm = sqrt(0.5 prod_m), prod_m = prod_{i<=m} (1 - x^2)(2i+1)/(2i);
    P~_{m+1,m} = sqrt(2m+3) x P~_mm; three-term step
    P~_n = (x P~_{n-1} - eps_{n-1} P~_{n-2}) / eps_n, eps_n = sqrt((n^2 - m^2)/(4 n^2 - 1))
    (sht.py:105-144).

    Extended-range arithmetic: prod is kept as mantissa * 2^E (exact power-of-two rescaling),
    and a start value below 2^-300 is carried as P~ 2^(600 s) with an integer s > 0 that the
    recurrence renormalises when the scaled value passes 2^300; values with s > 0 (true value
    < 2^-300) are set to 0.  Wherever the reference's prod stays a normal number the results
    are identical to sht.py; where it underflows (nlat >~ 700 at high degree) the reference's
    start values lose precision and its high-degree functions are wrong, these are not.
    """
    x = np.asarray(x, dtype=np.float64)
    p0s, scs = sectoral_starts(x, lmax)
    out = []
    for m in range(lmax + 1):
        nl = lmax + 1 - m
        P = np.zeros((nl, x.size))
        p0, sc = p0s[m], scs[m].copy()
        P[0] = np.where(sc == 0, p0, 0.0)
        if nl >= 2:
            p1 = np.sqrt(float(2 * m + 3)) * x * p0
            P[1] = np.where(sc == 0, p1, 0.0)
            eps_prev = np.sqrt(float((m + 1) ** 2 - m * m) / float(4 * (m + 1) ** 2 - 1))
            for i in range(2, nl):
                n = m + i
                eps = np.sqrt(float(n * n - m * m) / float(4 * n * n - 1))
                p2 = (x * p1 - eps_prev * p0) / eps
                big = (sc > 0) & (np.abs(p2) >= 2.0**SCALE_T)
                if big.any():
                    p2[big] *= 2.0**-SCALE_K
                    p1[big] *= 2.0**-SCALE_K
                    sc[big] -= 1
                P[i] = np.where(sc == 0, p2, 0.0)
                p0, p1, eps_prev = p1, p2, eps
        out.append(P)
    return out


class OracleSHT:
    """FP64 GL transform, dense ``[l, m]`` coefficients (zero above the diagonal)."""

    def __init__(self, nlat: int, nlon: int, lmax: int):
        if nlat % 2:
            raise ValueError("nlat must be even")
        if lmax > nlon // 2:
            raise ValueError("lmax must be <= nlon/2")
        self.nlat, self.nlon, self.lmax = int(nlat), int(nlon), int(lmax)
        self.nh = self.nlat // 2
        self.gaulats, self.weights = gauss_legendre(self.nlat)
        self.P = legendre_block(self.gaulats[: self.nh], self.lmax)
        self._wn = self.weights[: self.nh] / self.nlon
        self._ws = self.weights[::-1][: self.nh] / self.nlon  # row nlat-1-j

    def analysis(self, grid: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        nh, L = self.nh, self.lmax
        F = sfft.rfft(np.asarray(grid, dtype=np.float64), axis=1)[:, : L + 1]
        north = F[:nh] * self._wn[:, None]
        south = F[::-1][:nh] * self._ws[:, None]
        S, D = north + south, north - south
        U = np.zeros((L + 1, L + 1), dtype=np.complex128) if out is None else out
        U[...] = 0.0
        for m in range(L + 1):
            P = self.P[m]
            U[m::2, m] = (P[0::2] @ S[:, m]) / SQRT2_LEGACY
            U[m + 1 :: 2, m] = (P[1::2] @ D[:, m]) / SQRT2_LEGACY
        return U

    def synthesis(self, coeffs: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        nh, L = self.nh, self.lmax
        c = np.asarray(coeffs) * (SQRT2_LEGACY * self.nlon)
        F = np.zeros((self.nlat, self.nlon // 2 + 1), dtype=np.complex128)
        for m in range(L + 1):
            P = self.P[m]
            E = P[0::2].T @ c[m::2, m]
            O = P[1::2].T @ c[m + 1 :: 2, m]
            F[:nh, m] = E + O
            F[::-1][:nh, m] = E - O
        g = sfft.irfft(F, n=self.nlon, axis=1)
        if out is not None:
            out[...] = g
            return out
        return g

    def mean00_weights(self) -> np.ndarray:
        """Row weights with ``analysis(g)[0, 0] == w @ g.sum(axis=1)`` (P~_00 = 1/sqrt 2)."""
        return self.weights / self.nlon / np.sqrt(2.0) / SQRT2_LEGACY
