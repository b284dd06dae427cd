"""The grounded re-guess analysed incrementally (the ring pass transforms only the rings whose I*
changed and a sparse contraction adds the change's analysis to I*_lm) against the full re-analysis
(SLX_ICE_REANALYSIS=full, read when a solver is created): the same iteration counts and re-guess
bookkeeping, grids within 1e-9 m.  On the lmax 1023 workload nearly every inner iteration accepts a
re-guess, which changes I* at some tens of points in some tens of rings."""
import numpy as np
import pytest

from slevolve_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def run(w, nsteps):
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    slv = OptSeaLevelSolver(SHT(w.nlat, w.nlon, w.lmax), LoveNumbers(w.love_text, w.lmax, text=True), w.topo, w.ice(0),
                            norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner="extrap")
    out = []
    for k in range(1, nsteps + 1):
        r = slv.step(w.ice(k))
        out.append({"numiter": r["numiter"], "tgrid": r["tgrid"].cpu().numpy().copy(), "R": r["R"].cpu().numpy().copy(),
                    "G": r["G"].cpu().numpy().copy(), "gr": dict(slv.gr_counts)})
    return out


@pytest.mark.parametrize("name,nsteps", [("lmax1023", 4), ("lmax512", 6)])
def test_incremental_reguess_matches_full_reanalysis(monkeypatch, name, nsteps):
    w = syn.workload(name)
    print(f"[reguess] {name} incremental", flush=True)
    monkeypatch.delenv("SLX_ICE_REANALYSIS", raising=False)
    inc = run(w, nsteps)
    print(f"[reguess] {name} full", flush=True)
    monkeypatch.setenv("SLX_ICE_REANALYSIS", "full")
    full = run(w, nsteps)
    for k, (a, b) in enumerate(zip(inc, full), start=1):
        assert a["numiter"] == b["numiter"], (k, a["numiter"], b["numiter"])
        assert a["gr"] == b["gr"], (k, a["gr"], b["gr"])
        e = {q: float(np.abs(a[q] - b[q]).max()) for q in ("tgrid", "R", "G")}
        print(f"[reguess] {name} step {k} numiter {a['numiter']} re-guesses {a['gr']} max diff {e}", flush=True)
        assert max(e.values()) < 1e-9, (k, e)
    assert inc[-1]["gr"]["ice_analyses"] > 0  # the incremental path ran
