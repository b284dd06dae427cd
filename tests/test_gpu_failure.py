"""Failure handling of the inner loop, forced from the host: an iteration limit (the reference aborts
at 9999, solver.py:1049-1050) small enough that a later step of the quickstart case (numiter
6, 4, 6, 7, ...) cannot converge.  Every kernel exits normally; the host gives up at the limit.
The failing step is not committed, so the solver's committed state stays that of the last
finished step, as when the reference raises."""
import numpy as np
import pytest

from slevolve_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _solver(prem_text, topo, ice0, max_inner):
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    return OptSeaLevelSolver(SHT(128, 256, 63), LoveNumbers(prem_text, 63, text=True), topo, ice0, norder=63,
                             dt=1000.0, epsilon1=1e-5, _max_inner=max_inner)


def test_step_raises_and_run_reports_finished_steps(prem_text, topo_seed):
    import torch

    from slevolve_amd import InnerLoopError

    lat, lon, topo, ice = syn.synthetic_earth(128, topo_seed)
    # step() path: steps 1-3 need 6, 4, 6 iterations; step 4 needs 7 > 6
    a = _solver(prem_text, topo, ice[0], 6)
    for k in range(1, 4):
        print(f"[failure] step() {k}", flush=True)
        assert a.step(ice[k])["numiter"] == [6, 4, 6][k - 1]
    p_before = a.state.p_load.clone()
    with pytest.raises(InnerLoopError) as ei:
        a.step(ice[4])
    assert ei.value.status["numiter"] == 6
    assert torch.equal(a.state.p_load, p_before)  # not committed
    # run() path: the same failure inside one native loop
    stack = torch.from_numpy(np.ascontiguousarray(ice[:9])).cuda()
    b = _solver(prem_text, topo, ice[0], 6)
    print("[failure] run()", flush=True)
    with pytest.raises(InnerLoopError) as er:
        b.run(stack, 1, 9)
    assert [r["numiter"] for r in er.value.records] == [6, 4, 6]
    assert er.value.status["numiter"] == 6
    assert b.iter == 4 and b._last_status.step == 3
    assert torch.equal(b.state.p_load, p_before)  # the same committed state as the step() path
    # with the reference's limit the same run finishes
    c = _solver(prem_text, topo, ice[0], 9999)
    assert [r["numiter"] for r in c.run(stack, 1, 9)] == [6, 4, 6, 7, 7, 7, 7, 7]
