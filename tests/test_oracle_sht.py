"""Pins of the oracle transform (oracle/sht.py) by closed forms, a library and quadrature."""
import numpy as np
import pytest
from scipy.special import lpmv, factorial

from oracle.sht import OracleSHT, SQRT2_LEGACY, gauss_legendre, legendre_block


def test_nodes_match_numpy_leggauss():
    for n in (8, 64, 512):
        x, w = gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.abs(x - xr[::-1]).max() < 1e-15
        assert np.abs(w - wr[::-1]).max() < 1e-14 * max(1, n // 64)
        assert x[0] > 0.96 and abs(w.sum() - 2.0) < 1e-13


def test_legendre_closed_forms():
    x = np.linspace(-0.99, 0.99, 17)
    P = legendre_block(x, 4)
    s = np.sqrt(1 - x * x)
    assert np.allclose(P[0][0], 1 / np.sqrt(2), rtol=0, atol=2e-16)
    assert np.allclose(P[1][0], np.sqrt(3) / 2 * s, atol=1e-15)
    assert np.allclose(P[0][1], np.sqrt(1.5) * x, atol=1e-15)
    assert np.allclose(P[0][2], np.sqrt(5 / 8) * (3 * x * x - 1), atol=1e-15)
    assert np.allclose(P[2][0], np.sqrt(15) / 4 * s * s, atol=1e-15)


def test_legendre_matches_scipy_without_condon_shortley():
    x = np.linspace(-0.95, 0.95, 11)
    L = 20
    P = legendre_block(x, L)
    for m in range(0, L + 1, 3):
        for l in range(m, L + 1):
            norm = np.sqrt((2 * l + 1) / 2 * factorial(l - m) / factorial(l + m))
            ref = (-1) ** m * lpmv(m, l, x) * norm  # scipy includes the CS phase
            assert np.allclose(P[m][l - m], ref, rtol=1e-11, atol=1e-13), (l, m)


def test_orthonormal_by_quadrature():
    x, w = gauss_legendre(64)
    P = legendre_block(x, 63)
    for m in (0, 1, 7, 30, 63):
        G = (P[m] * w) @ P[m].T
        assert np.abs(G - np.eye(G.shape[0])).max() < 1e-12, m


def test_constant_field_and_single_mode():
    sht = OracleSHT(64, 128, 31)
    U = sht.analysis(np.full((64, 128), 2.5))
    assert abs(U[0, 0] - 2.5 * np.sqrt(2) / SQRT2_LEGACY) < 1e-14
    U[0, 0] = 0
    assert np.abs(U).max() < 1e-13
    # one mode P~_lm e^{i m phi}: analysis returns 1/(2 pi sqrt2_legacy ...) times the unit
    # coefficient in the legacy scaling, i.e. synthesis(e_lm) is analysed back to e_lm
    for l, m in ((5, 3), (31, 31), (12, 0)):
        c = np.zeros((32, 32), complex)
        c[l, m] = 1.0 + (0.5j if m else 0)
        g = sht.synthesis(c)
        phi = np.arange(128) * 2 * np.pi / 128
        Plm = legendre_block(sht.gaulats, 31)[m][l - m]
        expect = SQRT2_LEGACY * Plm[:, None] * (c[l, m] * np.exp(1j * m * phi)[None, :]).real * (2 if m else 1)
        assert np.abs(g - expect).max() < 1e-12
        assert np.abs(sht.analysis(g) - c).max() < 1e-13


@pytest.mark.parametrize("lmax", [63, 511])
def test_round_trip_512x1024(lmax):
    sht = OracleSHT(512, 1024, lmax)
    rng = np.random.default_rng(3)
    c = np.tril(rng.standard_normal((lmax + 1,) * 2) + 1j * rng.standard_normal((lmax + 1,) * 2))
    c[:, 0] = c[:, 0].real
    back = sht.analysis(sht.synthesis(c))
    assert np.abs(back - c).max() / np.abs(c).max() < 1e-12


def test_mean00_weights():
    sht = OracleSHT(64, 128, 63)
    g = np.random.default_rng(1).standard_normal((64, 128))
    assert abs(sht.mean00_weights() @ g.sum(axis=1) - sht.analysis(g)[0, 0].real) < 1e-15


@pytest.mark.parametrize("m", [0, 1, 5, 20])
def test_aliased_top_degree_row(m):
    """lmax = nlat (the lmax-512 workload's shape): the GL rule of nlat points does not integrate
    P~_{nlat,m}^2 exactly, so the analysis of the pure mode P~_{nlat,m}(mu) cos(m phi) is not 1 in its
    own coefficient but sum_j w_j P~^2 / 2 (m > 0; m = 0: sum_j w_j P~^2) / 1.41421356, with P~ from
    scipy (an implementation independent of the oracle's recursion), and the aliased mass shows in
    the degree nlat - 2 of the same order as sum_j w_j P~_{nlat} P~_{nlat-2}."""
    from scipy.special import gammaln

    nlat, nlon = 64, 128
    L = nlat
    sht = OracleSHT(nlat, nlon, L)
    x, w = sht.gaulats, sht.weights

    def pt(l):
        lognorm = 0.5 * (np.log((2 * l + 1) / 2) + gammaln(l - m + 1) - gammaln(l + m + 1))
        return (-1) ** m * lpmv(m, l, x) * np.exp(lognorm)

    p_top, p_low = pt(L), pt(L - 2)
    phi = 2 * np.pi * np.arange(nlon) / nlon
    g = p_top[:, None] * np.cos(m * phi)[None, :]
    U = sht.analysis(g)
    fac = 1.0 if m == 0 else 0.5
    top = fac * float((w * p_top * p_top).sum()) / SQRT2_LEGACY
    low = fac * float((w * p_top * p_low).sum()) / SQRT2_LEGACY
    assert abs(top - fac / SQRT2_LEGACY) > 1e-4  # the row is aliased, not exact
    assert abs(U[L, m].real - top) < 1e-12 and abs(U[L, m].imag) < 1e-12
    assert abs(U[L - 2, m].real - low) < 1e-12
