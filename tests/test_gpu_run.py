"""The native step loop (SeaLevelSolver.run -> slx_run_steps, which enqueues each inner iteration
on the device's posted decision of the previous one) against the Python step() loop (batches of
predicated iterations): the same steps bit for bit (iteration counts, grids, every step's
scalars), uniform and variable step lengths, extrapolated and Anderson inner loops, and both loops
can continue each other's run."""
import numpy as np
import pytest

from slevolve_amd import synthetic as syn

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("inner", ["extrap", "anderson"])
@pytest.mark.parametrize("variable", [False, True])
def test_native_loop_equals_step_loop(prem_text, topo_seed, variable, inner):
    import torch

    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    lat, lon, topo, ice = syn.synthetic_earth(128, topo_seed)
    ice_dev = torch.from_numpy(np.stack(ice)).cuda()
    dts = [1000.0, 500.0, 2000.0, 250.0, 1000.0, 750.0, 1500.0, 500.0] if variable else [1000.0] * 8
    mk = lambda: OptSeaLevelSolver(SHT(128, 256, 63), LoveNumbers(prem_text, 63, text=True), topo, ice[0],  # noqa: E731
                                   norder=63, dt=dts[0], epsilon1=1e-5, inner=inner)
    a, b = mk(), mk()
    ra = [a.step(ice_dev[k], dt=dts[k - 1]) for k in range(1, 7)]
    na = [r["numiter"] for r in ra]
    last_a = {k: ra[-1][k].cpu().numpy().copy() for k in ("tgrid", "R", "G", "ocean", "beta")}
    rb = b.run(ice_dev, 1, 7, dts=dts[:6])
    assert [r["numiter"] for r in rb] == na
    for key, v in last_a.items():
        assert np.array_equal(getattr(b, {"tgrid": "topoxy", "R": "rrxy", "G": "Gxy", "ocean": "cxy",
                                          "beta": "beta"}[key]).cpu().numpy(), v), key
    for k in range(6):
        for key in ("gmslc_ocn", "gmslc_ocnBeta", "ocean_area", "ice_volume"):
            assert rb[k][key] == ra[k][key], (k, key)
        assert np.array_equal(rb[k]["mm"], ra[k]["mm"]), k
    # continue: a with run(), b with step(); still identical
    ra2 = a.run(ice_dev, 7, 9, dts=dts[6:8])
    rb2 = [b.step(ice_dev[k], dt=dts[k - 1]) for k in (7, 8)]
    assert [r["numiter"] for r in ra2] == [r["numiter"] for r in rb2]
    assert np.array_equal(a.topoxy.cpu().numpy(), rb2[-1]["tgrid"].cpu().numpy())
    assert a.iter == b.iter == 8
