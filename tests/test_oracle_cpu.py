"""The CPU baseline's C kernels (oracle/cpu.py, oracle/cpu_kernels.c) against the oracle: transforms
to rounding at several grids (incl. lmax = nlon/2 and the extended-range start values at nlat
1024), the dense modal advance bit for bit, and whole steps with identical iteration counts."""
import os

import numpy as np
import pytest

from oracle import cpu as ocpu
from oracle.sht import OracleSHT


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(ocpu.LIB):
        ocpu.build()


@pytest.mark.parametrize("nlat,L", [(64, 63), (64, 64), (128, 63), (1024, 1023)])
def test_transforms_match_oracle(nlat, L):
    rng = np.random.default_rng(nlat + L)
    a, b = ocpu.FastCpuSHT(nlat, 2 * nlat, L), OracleSHT(nlat, 2 * nlat, L)
    g = rng.standard_normal((nlat, 2 * nlat))
    assert np.abs(a.analysis(g) - b.analysis(g)).max() < 1e-15 * max(1, L / 64)
    c = np.tril(rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1)))
    c[:, 0] = c[:, 0].real
    gb = b.synthesis(c)
    assert np.abs(a.synthesis(c) - gb).max() < 1e-13 * np.abs(gb).max() * max(1, L / 64)


def test_modal_advance_bit_exact(prem_text):
    from oracle.love import LoveNumbers
    from oracle.memory import ExactMemory

    L = 63
    love = LoveNumbers(prem_text, L, text=True)
    a, b = ExactMemory(love, L, 500.0), ExactMemory(love, L, 500.0)
    a.__class__ = ocpu.FastExactMemory
    rng = np.random.default_rng(2)
    for _ in range(3):
        ra, rb = a.advance_and_read(), b.advance_and_read()
        for x, y in zip(ra, rb):
            assert np.array_equal(x, y)
        load = np.tril(rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1)))
        a.commit(load)
        b.commit(load)


def test_steps_match_oracle():
    from oracle.love import LoveNumbers
    from oracle.solver import SeaLevelSolver
    from slevolve_amd import synthetic as syn

    w = syn.workload("lmax63")
    kw = dict(norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner="extrap")
    a = ocpu.FastCpuSolver(ocpu.FastCpuSHT(w.nlat, w.nlon, w.lmax), LoveNumbers(w.love_text, w.lmax, text=True),
                           w.topo, w.ice(0), **kw)
    b = SeaLevelSolver(OracleSHT(w.nlat, w.nlon, w.lmax), LoveNumbers(w.love_text, w.lmax, text=True), w.topo,
                       w.ice(0), **kw)
    for k in range(1, 5):
        ra, rb = a.step(w.ice(k)), b.step(w.ice(k))
        assert ra["numiter"] == rb["numiter"]
        assert np.abs(ra["tgrid"] - rb["tgrid"]).max() < 1e-9
