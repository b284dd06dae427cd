"""Cross-pass reuse modes of the oracle (solver_opt.py:272-286,515-528,932-941), CPU.

The corrected warm starts are initial guesses only; they are pinned by what they must reproduce
exactly: when the pass-to-pass change D_j = dS_j^(k) - dS_j^(k-1) of the converged increments is the
same for every step j, "warm+corr" starts step n at its new converged value; when D_j is linear in
j, "warm+corr2" does (2 D_{n-1} - D_{n-2} = D_n).  A wrong sign, step index or pass fails these.
Whole outer passes converge to the same fixed point as "warm" within the stopping test."""
import numpy as np
import pytest

from oracle.solver import reuse_guess, reuse_store_step


def _stores(D):
    """Stores after pass k-1 (states dS_j^(k-1) = A_j) and during pass k up to step n-1
    (dS_j^(k) = A_j + D_j for j < n); D: callable j -> array."""
    rng = np.random.default_rng(3)
    A = {j: rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6)) for j in range(1, 8)}
    store = {}
    cxy = np.ones((4, 4))
    for j in range(1, 8):  # pass k-1
        reuse_store_step(store, j, "warm+corr2", A[j], cxy)
    return A, store, cxy


@pytest.mark.parametrize("mode", ["warm+corr", "warm+corr2"])
def test_corrected_start_is_exact_for_its_drift(mode):
    rng = np.random.default_rng(4)
    a = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    b = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    D = (lambda j: a) if mode == "warm+corr" else (lambda j: a + j * b)
    A, store, cxy = _stores(D)
    n = 5
    for j in range(1, n):  # pass k, steps before n
        reuse_store_step(store, j, mode, A[j] + D(j), cxy)
    g, corrected = reuse_guess(store, n, mode)
    assert corrected
    assert np.allclose(g, A[n] + D(n), rtol=0, atol=1e-12)
    # "warm" starts at the previous pass's value; "warm+corr" misses a drift linear in j by b
    g0, c0 = reuse_guess(store, n, "warm")
    assert not c0 and np.array_equal(g0, A[n])
    if mode == "warm+corr2":
        g1, _ = reuse_guess(store, n, "warm+corr")
        assert np.allclose(g1 - (A[n] + D(n)), -b, atol=1e-12)


def test_first_steps_and_first_pass():
    """Without a previous-pass value at step n-1 the start is the stored value; warm+corr2 falls
    back to D_{n-1} alone at step 2 (no step 0)."""
    A, store, cxy = _stores(lambda j: 0)
    g, c = reuse_guess(store, 1, "warm+corr")
    assert not c and np.array_equal(g, A[1])
    reuse_store_step(store, 1, "warm+corr2", A[1] + 1.0, cxy)
    g, c = reuse_guess(store, 2, "warm+corr2")
    assert c and np.allclose(g, A[2] + 1.0)
    assert reuse_guess({}, 3, "warm+corr")[0] is None


def test_outer_passes_same_fixed_point(prem_text):
    """Three Picard passes of the lmax-63 workload: every reuse mode ends within the stopping
    tolerance of "off"."""
    from oracle.love import LoveNumbers as OLove
    from oracle.outer import outer_iteration
    from oracle.sht import OracleSHT
    from slevolve_amd import synthetic as syn

    w = syn.workload("lmax63")
    ice = [w.ice(k) for k in range(7)]
    sht, love = OracleSHT(w.nlat, w.nlon, w.lmax), OLove(w.love_text, w.lmax, text=True)
    kw = dict(lmax=w.lmax, dt=w.dt, iters=3, outer="picard", epsilon1=1e-10, inner="extrap")
    _, _, ref = outer_iteration(sht, love, w.topo, ice, reuse="off", **kw)
    its = {}
    for mode in ("warm", "warm+ice", "warm+corr", "warm+corr2"):
        rec, _, pred = outer_iteration(sht, love, w.topo, ice, reuse=mode, **kw)
        its[mode] = [r["numiter"] for r in rec]
        assert np.abs(pred[-1] - ref[-1]).max() < 1e-4, mode
    assert its["warm+ice"] == its["warm"]
    print(its)
