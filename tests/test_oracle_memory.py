"""Pins of the oracle modal state (oracle/memory.py) and extrapolated start by brute force."""
import numpy as np

from oracle.memory import ExactMemory
from oracle.solver import extrap_guess

NPAM = 500


class RandomSpectrum:
    """Random decay rates and residues, shaped like a LoveNumbers object."""

    def __init__(self, lmax, kmax, rng):
        nl = lmax + 1
        self.nmod = np.zeros(nl, dtype=np.int64)
        self.s = np.zeros((NPAM, nl))
        self.r, self.rprime, self.rT, self.rprimeT = (np.zeros((NPAM, nl)) for _ in range(4))
        for l in range(1, nl):
            n = kmax if l == 2 else int(rng.integers(1, kmax + 1))
            self.nmod[l] = n
            self.s[:n, l] = 10.0 ** rng.uniform(-2, 2, n)
            for a in (self.r, self.rprime, self.rT, self.rprimeT):
                a[:n, l] = rng.normal(size=n)
        self.max_modes = int(self.nmod.max())


def test_recursion_equals_history_sum_nonuniform_steps():
    rng = np.random.default_rng(12345)
    lmax, kmax, nsteps = 12, 8, 60
    love = RandomSpectrum(lmax, kmax, rng)
    st = ExactMemory(love, lmax, 1.0)
    dts = rng.uniform(1.0, 2000.0, nsteps)
    t = np.concatenate(([0.0], np.cumsum(dts)))
    nl = lmax + 1
    hist = []
    worst = 0.0
    for n in range(1, nsteps + 1):
        st.set_dt(dts[n - 1])
        # decay factors are exactly exp(-s dt / 1000)
        for l in range(1, nl):
            k = st.nmod[l]
            assert np.array_equal(st.decay[:k, l], np.exp(-st.s[:k, l] * dts[n - 1] / 1000.0))
        visc, rad, sil, sm, stsl, strr = st.advance_and_read()
        ev, er = np.zeros((nl, nl), complex), np.zeros((nl, nl), complex)
        esil, esm, estsl = np.zeros((3, 3)), np.zeros(3), np.zeros(3, complex)
        for j, (L, il, m, lam) in enumerate(hist, start=1):
            age = t[n] - t[j]  # increment j committed at the end of step j
            f = np.where(st.s > 0, 1.0 - np.exp(-st.s * age / 1000.0), 0.0)
            ev += (st.a_sl * f).sum(0)[:, None] * L
            er += (st.a_rr * f).sum(0)[:, None] * L
            f2 = f[: st.nmod2, 2]
            esil += (st.a_il[: st.nmod2] * f2).sum() * il
            esm += (st.a_m[: st.nmod2] * f2).sum() * m
            estsl += (st.a_tsl[: st.nmod2] * f2).sum() * lam
        for a, b in ((visc, ev), (rad, er), (sil, esil), (sm, esm), (stsl, estsl)):
            scale = max(np.abs(b).max(), 1e-300)
            worst = max(worst, np.abs(a - b).max() / scale)
        L = np.tril(rng.standard_normal((nl, nl)) + 1j * rng.standard_normal((nl, nl)))
        L[0] = 0
        il, m, lam = rng.standard_normal((3, 3)), rng.standard_normal(3), rng.standard_normal(3) + 0j
        st.commit(L, il, m, lam)
        hist.append((L, il, m, lam))
    assert worst < 1e-12, worst


def test_extrap_closed_form_and_uniform_exact():
    rng = np.random.default_rng(7)
    h1, h2 = rng.standard_normal(50) + 1j * rng.standard_normal(50), rng.standard_normal(50) + 0j
    assert np.array_equal(extrap_guess(h1, h2, 1.0, 1.0, 1.0), 2.0 * h1 - h2)
    for dt0, dt1, dt2 in ((2.0, 1.0, 1.0), (5.0, 2.0, 1.0), (37.5, 200.0, 100.0)):
        s1, s2 = h1 / dt1, h2 / dt2
        ref = dt0 * (s1 + (s1 - s2) * (dt0 + dt1) / (dt1 + dt2))
        got = extrap_guess(h1, h2, dt0, dt1, dt2)
        assert np.abs(got - ref).max() <= 4e-16 * np.abs(ref).max() * 8
