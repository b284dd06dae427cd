"""Pins of the oracle's Anderson mixing (oracle/anderson.py) and of the frozen-mask inner loop
(oracle/solver.py, inner="anderson"; python/slevolve/solver_opt.py:67-145,705-775).  The reference
ships no test for this experimental option, so the mixer is pinned by what the mathematics fixes:
the mixed point is g_0 - dG gamma with gamma the least-squares solution of dF gamma = f_0 (checked
against a QR solve of the tall system), full-memory type-II Anderson on an affine map of R^n finds
the fixed point after n + 1 evaluations (its equivalence with GMRES), and the branch rules; the
step is pinned by agreement with the extrapolated start within the stopping test."""
import numpy as np
import pytest

from oracle.anderson import Anderson, real_dot
from oracle.love import LoveNumbers
from oracle.sht import OracleSHT
from oracle.solver import SeaLevelSolver
from slevolve_amd import synthetic as syn


def test_mix_is_the_least_squares_step():
    rng = np.random.default_rng(0)
    aa = Anderson(3)
    xs = [rng.normal(size=(5, 5)) + 1j * rng.normal(size=(5, 5)) for _ in range(4)]
    gs = [x * 0.3 + rng.normal(size=(5, 5)) * 0.1 * (4 - i) for i, x in enumerate(xs)]
    for x, g in zip(xs, gs):
        aa.push(x, g)
    out, kind = aa.mix()
    assert kind == "mixed"
    f = [g - x for x, g in zip(xs, gs)][::-1]  # newest first
    g = gs[::-1]
    real = lambda z: np.concatenate([z.real.ravel(), z.imag.ravel()])  # noqa: E731
    dF = np.stack([real(f[j] - f[j + 1]) for j in range(3)], 1)
    q, r = np.linalg.qr(dF)
    gam = np.linalg.solve(r, q.T @ real(f[0]))
    ref = g[0] - sum(gam[j] * (g[j] - g[j + 1]) for j in range(3))
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()
    assert abs(real_dot(f[0], f[0]) - np.sum(np.abs(f[0]) ** 2)) < 1e-12 * np.sum(np.abs(f[0]) ** 2)


def test_full_memory_solves_an_affine_map_in_n_plus_one_evaluations():
    """G(x) = M x + b on C^2 (R^4, M real-linear on the real and imaginary parts), spectral radius
    0.95: plain iteration is slow, Anderson with depth 4 reaches the fixed point at the 5th mix."""
    rng = np.random.default_rng(1)
    Q, _ = np.linalg.qr(rng.normal(size=(4, 4)))
    M = Q @ np.diag([0.95, -0.9, 0.5, 0.2]) @ Q.T
    b = rng.normal(size=4)
    pack = lambda v: v[:2] + 1j * v[2:]  # noqa: E731
    unpack = lambda z: np.concatenate([z.real, z.imag])  # noqa: E731
    G = lambda z: pack(M @ unpack(z) + b)  # noqa: E731
    xstar = pack(np.linalg.solve(np.eye(4) - M, b))
    aa = Anderson(4)
    x = np.zeros(2, complex)
    kinds = []
    for _ in range(5):
        aa.push(x, G(x))
        x, kind = aa.mix()
        kinds.append(kind)
    assert kinds[0] == "plain" and "singular" not in kinds
    assert np.abs(x - xstar).max() < 1e-10 * np.abs(xstar).max()
    y = np.zeros(2, complex)
    for _ in range(5):
        y = G(y)
    assert np.abs(y - xstar).max() > 1e-2 * np.abs(xstar).max()


def test_branch_rules():
    aa = Anderson(3)
    x = np.ones(4, complex)
    aa.push(x, 0.5 * x)
    out, kind = aa.mix()
    assert kind == "plain" and np.array_equal(out, 0.5 * x)
    aa.push(0.5 * x, 0.5 * x + 3.0)  # residual norm grew: history restarts from the newest pair
    out, kind = aa.mix()
    assert kind == "fallback" and np.array_equal(out, 0.5 * x + 3.0) and len(aa.pairs) == 1
    aa2 = Anderson(3)
    aa2.push(x, 0.5 * x)
    aa2.push(0.5 * x, np.full(4, np.nan + 0j))  # a NaN residual norm does not compare as grown
    assert aa2.mix()[1] == "singular"
    for _ in range(6):
        aa.push(x, x)
    assert len(aa.pairs) == 4  # depth + 1 pairs kept


@pytest.fixture(scope="module")
def small(prem_text):
    sht = OracleSHT(128, 256, 63)
    love = LoveNumbers(prem_text, 63, text=True)
    topo, ice = syn.small_case(sht.gaulats, 256)
    return sht, love, topo, ice


def test_anderson_step_reaches_the_same_fixed_point(small):
    sht, love, topo, ice = small
    runs = {}
    for inner in ("extrap", "anderson"):
        slv = SeaLevelSolver(sht, love, topo, ice[0], norder=63, dt=1000.0, epsilon1=1e-10, inner=inner)
        runs[inner] = [slv.step(ice[k]) for k in range(1, 9)]
        if inner == "anderson":
            stats = slv.inner_stats
    assert stats["mixed"] > 0 and stats["frozen_iters"] > 0
    for a, b in zip(runs["extrap"], runs["anderson"]):
        assert np.array_equal(a["ocean"], b["ocean"])
        assert np.abs(a["tgrid"] - b["tgrid"]).max() < 1e-5
        assert np.abs(a["R"] - b["R"]).max() < 1e-5


def test_anderson_saves_iterations_on_the_lmax63_workload():
    w = syn.workload("lmax63")
    sht = OracleSHT(w.nlat, w.nlon, w.lmax)
    love = LoveNumbers(w.love_text, w.lmax, text=True)
    tot = {}
    for inner in ("extrap", "anderson"):
        slv = SeaLevelSolver(sht, love, w.topo, w.ice(0), norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner=inner)
        tot[inner] = sum(slv.step(w.ice(k))["numiter"] for k in range(1, 31))
    assert tot == {"extrap": 206, "anderson": 176}
