"""Long-horizon parity of the timed path (SeaLevelSolver.run: the native step loop with the
incremental grounded-ice analyses carried across steps) against the oracle's stored long runs
(tests/golden/make_oracle_long.py): the whole 550-step lmax-512 workload and 40 steps of the
lmax-1023 workload.  Iteration counts per step must be identical; fields at the checkpoints agree
to 1e-9 m on the stored subsample, and full-grid sums to the same 1e-9 m per point."""
import os

import numpy as np
import pytest
import torch

from slevolve_amd import synthetic as syn

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name", ["lmax512", "lmax1023"])
def test_native_loop_long_run_vs_oracle(name):
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    f = dict(np.load(os.path.join(GOLDEN, f"oracle_{name}_long.npz")))
    assert bool(f["complete"]), "the fixture run did not finish"
    w = syn.workload(name)
    N = int(f["numiter"].size)
    ice = torch.empty((N + 1, w.nlat, w.nlon), dtype=torch.float64, device="cuda")
    for k in range(N + 1):
        ice[k].copy_(torch.from_numpy(w.ice(k)))
    slv = OptSeaLevelSolver(SHT(w.nlat, w.nlon, w.lmax), LoveNumbers(w.love_text, w.lmax, text=True), w.topo, ice[0],
                            norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner="extrap")
    s = int(f["sub"])
    npts = w.nlat * w.nlon
    k0, numiter, worst = 1, [], {"tgrid": 0.0, "R": 0.0}
    for i, cp in enumerate(int(c) for c in f["check"]):
        rec = slv.run(ice, k0, cp + 1)
        numiter += [r["numiter"] for r in rec]
        for fld, g in (("tgrid", slv.topoxy), ("R", slv.rrxy)):
            a = g.cpu().numpy()
            err = float(np.abs(a[::s, ::s] - f[fld + "_sub"][i]).max())
            worst[fld] = max(worst[fld], err)
            assert err < 1e-9, (name, cp, fld, err)
            assert abs(a.sum() - f[fld + "_sum"][i]) < 1e-9 * npts + 1e-12 * abs(f[fld + "_sum"][i])
        assert np.allclose(rec[-1]["mm"], f["mm"][cp - 1], rtol=1e-9, atol=1e-20)
        print(f"[long {name}] step {cp}: max |dtgrid| {worst['tgrid']:.2e} m, |dR| {worst['R']:.2e} m", flush=True)
        k0 = cp + 1
    mism = [k + 1 for k in range(N) if numiter[k] != int(f["numiter"][k])]
    print(f"[long {name}] {N} steps, numiter mismatches: {mism}", flush=True)
    assert not mism
