"""A pin of the horizontal read-out's composition (solver.py:687-720): with a Love-number file
whose l numbers are its h numbers (elastic, modal residues and tidal degree-2 alike), the
horizontal potential Psi_lm = T l^E A + T sum_k (l_k/s_k) R_k (+ degree-2 rotation with the tidal
l numbers) must equal the radial displacement's spectrum T h^E A + T sum_k (r_k/s_k) R_k (+ the
rotation with the tidal h numbers), which the reference's fixtures pin.  A wrong weight in Psi (a
missing T, the sea-level residues instead of the horizontal ones, the rotation without its tidal
numbers or with the loading ones, another load) fails it.  The oracle is checked on the CPU; the
device path in the gpu test."""
import numpy as np
import pytest

from slevolve_amd import synthetic as syn


def l_equals_h_love_text(lmax, nmodes, seed):
    rng = np.random.default_rng(seed)
    s, el, amp, tel, tamp = [None], [None], [None], [None], [None]
    for l in range(1, lmax + 1):
        h = -1.0 / (1 + l / 20.0)
        k = -0.3 / (1 + l / 10.0)
        e = np.array([h, h, k])  # (h, l, k) with l = h
        s.append(np.sort(1000.0 * 10.0 ** rng.uniform(-4.0, -1.0, nmodes)))
        a = rng.uniform(0.0, 1.0, (3, nmodes)) * np.array([h, h, k])[:, None] / nmodes
        a[1] = a[0]
        el.append(e)
        amp.append(a)
        t = rng.uniform(0.0, 1.0, (3, nmodes)) * np.array([h, h, k])[:, None] / nmodes
        t[1] = t[0]
        tel.append(e.copy())
        tamp.append(t)
    return syn.love_file_text(s, el, amp, tel, tamp, header="l = h")


def test_oracle_horizontal_equals_radial_when_l_is_h(topo_seed):
    from oracle.love import LoveNumbers
    from oracle.sht import OracleSHT
    from oracle.solver import SeaLevelSolver

    L, nlat = 31, 64
    lat, lon, topo, ice = syn.synthetic_earth(nlat, topo_seed)
    text = l_equals_h_love_text(L, 4, 11)
    sht = OracleSHT(nlat, 2 * nlat, L)
    slv = SeaLevelSolver(sht, LoveNumbers(text, L, text=True), topo, ice[0], norder=L, dt=1000.0, epsilon1=1e-8)
    slv.horizontal = True
    for k in range(1, 5):
        r = slv.step(ice[k])
        R_lm = sht.analysis(r["R"])
        err = np.abs(r["horiz"] - R_lm).max() / np.abs(R_lm).max()
        assert err < 1e-11, (k, err)
    # and a deliberately wrong Love file (l residues halved) is caught
    bad = LoveNumbers(text, L, text=True)
    bad.rl = bad.rl * 0.5
    slv2 = SeaLevelSolver(sht, bad, topo, ice[0], norder=L, dt=1000.0, epsilon1=1e-8)
    slv2.horizontal = True
    for k in range(1, 3):
        r = slv2.step(ice[k])
    R_lm = sht.analysis(r["R"])
    assert np.abs(r["horiz"] - R_lm).max() / np.abs(R_lm).max() > 1e-6


@pytest.mark.gpu
def test_device_horizontal_equals_radial_when_l_is_h(topo_seed):
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    L, nlat = 63, 128
    lat, lon, topo, ice = syn.synthetic_earth(nlat, topo_seed)
    text = l_equals_h_love_text(L, 6, 12)
    sht = SHT(nlat, 2 * nlat, L)
    slv = OptSeaLevelSolver(sht, LoveNumbers(text, L, text=True), topo, ice[0], norder=L, dt=1000.0, epsilon1=1e-8,
                            inner="extrap")
    slv.horizontal = True
    for k in range(1, 6):
        r = slv.step(ice[k])
        R_lm = sht.analysis(r["R"]).cpu().numpy()
        hz = r["horiz"].cpu().numpy()
        err = np.abs(hz - R_lm).max() / np.abs(R_lm).max()
        print(f"[horizontal pin] step {k}: max |Psi - R_lm| / max |R_lm| = {err:.2e}", flush=True)
        assert err < 1e-11, (k, err)
