"""The forward-run driver (slevolve_amd.forward, the command line of run_ais.py) on a small case:
its files equal the solver's own step results, run_log.json carries the reference's keys, the
ice cache with its sidecar is built and then reused, and the native-loop mode agrees."""
import gzip
import json
import os

import numpy as np
import pytest

from slevolve_amd import synthetic as syn

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _inputs(tmp_path, topo_seed):
    from slevolve_amd import npyio

    lat, lon, topo, ice = syn.synthetic_earth(64, topo_seed)
    icedir = tmp_path / "ice"
    icedir.mkdir()
    for k in range(5):  # mixed encodings: text for even k, .npy for odd k
        if k % 2:
            npyio.save_grid(str(icedir / f"ice{k}.npy"), ice[k])
        else:
            npyio.write_text_grid(str(icedir / f"ice{k}"), ice[k], "row-e.18", ncol=128)
    np.save(tmp_path / "topo.npy", topo)
    with gzip.open(os.path.join(GOLDEN, "prem_512.l96C.ump3.lm1.gz"), "rt") as fh:
        (tmp_path / "earth.txt").write_text(fh.read())
    return topo, ice, [f"--icedir={icedir}", "--icepfx=ice", f"--topo={tmp_path / 'topo.npy'}",
                       f"--earth={tmp_path / 'earth.txt'}", "--nglv=64", "--lmax=31", "--dt=1000",
                       "--steps=4", "--inner=extrap"]


def test_forward_driver_files_and_log(tmp_path, topo_seed, prem_text):
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver
    from slevolve_amd.forward import main

    topo, ice, base = _inputs(tmp_path, topo_seed)
    out = tmp_path / "run"
    print("[forward] step-by-step run", flush=True)
    assert main(base + [f"--out={out}", "--save-fields=tgrid,ocean,beta,R,G,dS", f"--cache-dir={tmp_path / 'cache'}"]) == 0
    slv = OptSeaLevelSolver(SHT(64, 128, 31), LoveNumbers(prem_text, 31, text=True), topo, ice[0], norder=31,
                            dt=1000.0, inner="extrap")
    nit = []
    for k in range(1, 5):
        r = slv.step(ice[k])
        nit.append(r["numiter"])
        for f in ("tgrid", "R", "G", "ocean", "beta"):
            assert np.array_equal(np.load(out / f"{f}{k}.npy"), r[f].cpu().numpy()), (f, k)
        assert np.array_equal(np.load(out / f"dS_converged{k}.npy"), r["dS_converged"].cpu().numpy())
    assert list(np.load(out / "numiter.npy")) == nit
    assert np.array_equal(np.load(out / "tgrid0.npy"), topo)
    log = json.loads((out / "run_log.json").read_text())
    for key in ("argv", "steps", "norder", "nglv", "epsilon1", "numiter_total", "numiter_mean", "scalars",
                "total_solver_seconds", "per_step_median", "precision", "device"):
        assert key in log, key
    assert len(log["scalars"]) == 5 and log["numiter_total"] == sum(nit)
    cache = [f for f in os.listdir(tmp_path / "cache") if f.endswith(".npy")]
    assert cache and os.path.exists(tmp_path / "cache" / (cache[0] + ".json"))
    # the native loop: same iteration counts and final grid; the cache is reused
    out2 = tmp_path / "run2"
    print("[forward] native run", flush=True)
    assert main(base + [f"--out={out2}", "--native", "--save-fields=tgrid", f"--cache-dir={tmp_path / 'cache'}"]) == 0
    assert list(np.load(out2 / "numiter.npy")) == nit
    assert np.array_equal(np.load(out2 / "tgrid4.npy"), np.load(out / "tgrid4.npy"))
    assert json.loads((out2 / "run_log.json").read_text())["ice_cache_seconds"] == 0.0
    # mixed precision through the driver, within the metric's 1e-6 m of the FP64 run
    out3 = tmp_path / "run3"
    assert main(base + [f"--out={out3}", "--precision=mixed", "--save-fields=tgrid", f"--cache-dir={tmp_path / 'cache'}"]) == 0
    assert np.abs(np.load(out3 / "tgrid4.npy") - np.load(out / "tgrid4.npy")).max() <= 1e-6
