"""Whole-step pins of the oracle (oracle/solver.py) by the reference's known answers."""
import numpy as np
import pytest

from oracle.love import LoveNumbers
from oracle.memory import ExactMemory
from oracle.sht import OracleSHT
from oracle.solver import SeaLevelSolver
from slevolve_amd import synthetic as syn


class HistoryMemory(ExactMemory):
    """Viscous terms as the explicit sum over committed increments (no recursion)."""

    def __init__(self, love, lmax, dt):
        super().__init__(love, lmax, dt)
        self.hist, self.n = [], 0

    def advance_and_read(self):
        self.n += 1
        nl = self.lmax + 1
        out = [np.zeros((nl, nl), complex), np.zeros((nl, nl), complex), np.zeros((3, 3)),
               np.zeros(3), np.zeros(3, complex), np.zeros(3, complex)]
        k2 = self.nmod2
        for j, (L, il, m, lam) in enumerate(self.hist, start=1):
            f = np.where(self.s > 0, 1.0 - np.exp(-self.s * (self.n - j) * self.dt_years / 1000.0), 0.0)
            out[0] += (self.a_sl * f).sum(0)[:, None] * L
            out[1] += (self.a_rr * f).sum(0)[:, None] * L
            f2 = f[:k2, 2]
            out[2] += (self.a_il[:k2] * f2).sum() * il
            out[3] += (self.a_m[:k2] * f2).sum() * m
            out[4] += (self.a_tsl[:k2] * f2).sum() * lam
            out[5] += (self.a_trr[:k2] * f2).sum() * lam
        return tuple(out)

    def commit(self, L, dil=None, dm=None, dlam=None):
        self.hist.append((L.copy(), dil.copy(), dm.copy(), dlam.copy()))
        super().commit(L, dil, dm, dlam)


def test_quickstart_known_answers(prem_text, topo_seed):
    """examples/quickstart.py printed values (docs/src/quickstart.md:76-82)."""
    sht = OracleSHT(128, 256, 63)
    love = LoveNumbers(prem_text, 63, text=True)
    lat, lon, topo, ice = syn.synthetic_earth(128, topo_seed)
    res = {}
    for cm in (True, False):
        slv = SeaLevelSolver(sht, love, topo, ice[0], norder=63, dt=1000.0, tpw=True,
                             checkmarine=cm, calcRG=True, epsilon1=1e-5)
        out = [slv.step(ice[k]) for k in range(1, 9)]
        res[cm] = ([r["numiter"] for r in out], -(out[-1]["tgrid"] - topo))
    assert res[True][0] == [6, 4, 6, 7, 7, 7, 7, 7]
    assert res[False][0] == [6, 4, 6, 7, 7, 7, 7, 7]
    sites = {(-82.0, 270.0): (-481.274, -481.282), (-55.0, 0.0): (65.413, 65.428),
             (10.0, 180.0): (78.067, 78.065)}
    for (la, lo), (on, off) in sites.items():
        i, j = int(np.argmin(np.abs(lat - la))), int(np.argmin(np.abs(lon - lo)))
        assert round(res[True][1][i, j], 3) == on
        assert round(res[False][1][i, j], 3) == off
    w = np.polynomial.legendre.leggauss(128)[1][:, None] * np.ones((1, 256))
    ocean = topo < 0
    assert round((w * res[True][1])[ocean].sum() / w[ocean].sum(), 3) == 76.463


def test_small_case_floating_ice_and_history(prem_text):
    """tests/test_floating_ice.py numiter [4,4,5,8,6,6,5,6]; recursion vs explicit history
    (tests/test_small_case.py (a) vs (c), < 1e-9 m)."""
    sht = OracleSHT(128, 256, 63)
    love = LoveNumbers(prem_text, 63, text=True)
    topo, ice = syn.small_case(sht.gaulats, 256)
    a = SeaLevelSolver(sht, love, topo, ice[0], norder=63, dt=1000.0, epsilon1=1e-5)
    c = SeaLevelSolver(sht, love, topo, ice[0], norder=63, dt=1000.0, epsilon1=1e-5)
    c.state = HistoryMemory(love, 63, 1000.0)
    ra = [a.step(ice[k]) for k in range(1, 9)]
    rc = [c.step(ice[k]) for k in range(1, 9)]
    assert [r["numiter"] for r in ra] == [4, 4, 5, 8, 6, 6, 5, 6]
    assert a.gr_counts == {"reguess": 36, "ice_analyses": 1}
    for x, y in zip(ra, rc):
        assert np.abs(x["tgrid"] - y["tgrid"]).max() < 1e-9
        assert np.abs(x["R"] - y["R"]).max() < 1e-9


@pytest.mark.slow
def test_reference_fixture_tgrid8(prem_text, topo_seed, tgrid8_change):
    """reference/synthetic/tgrid8.npy: 512 x 1024, lmax 511, 8 x 1 yr, extrapolated start,
    rotation and marine check on, epsilon1 1e-5; 1e-9 m and numiter 5,5,6,6,5,5,5,5
    (tests/test_fortran_synthetic.sh:91-96, docs/src/verify.md:353)."""
    sht = OracleSHT(512, 1024, 511)
    love = LoveNumbers(prem_text, 511, text=True)
    lat, lon, topo, ice = syn.synthetic_earth(512, topo_seed)
    slv = SeaLevelSolver(sht, love, topo, ice[0], norder=511, dt=1.0, epsilon1=1e-5, inner="extrap")
    out = [slv.step(ice[k]) for k in range(1, 9)]
    assert [r["numiter"] for r in out] == [5, 5, 6, 6, 5, 5, 5, 5]
    assert np.abs(out[-1]["tgrid"] - (topo_seed + tgrid8_change)).max() < 1e-9


def test_love_writer_reader_round_trip():
    seqs = np.random.SeedSequence(99).spawn(2)
    txt = syn.synthetic_love_text(12, 4, seqs[0], seqs[1])
    love = LoveNumbers(txt, 12, text=True)
    rl = np.random.default_rng(seqs[0])
    for l in range(1, 13):
        e = np.array([-1.0 / (1 + l / 20.0), 0.05 / (1 + l / 30.0), -0.3 / (1 + l / 10.0)])
        s = np.sort(1000.0 * 10.0 ** rl.uniform(-5.0, -1.0, 4))
        a = rl.uniform(0.0, 1.0, (3, 4)) * e[:, None] / 4
        assert love.nmod[l] == 4
        assert np.allclose(love.s[:4, l], s, rtol=1e-15)
        assert np.isclose(love.he[l], e[0], rtol=1e-15) and np.isclose(love.ke[l], e[2], rtol=1e-15)
        assert np.allclose(love.r[:4, l], a[0] * s, rtol=1e-14)
        assert np.allclose(love.rprime[:4, l], a[2] * s, rtol=1e-14)
