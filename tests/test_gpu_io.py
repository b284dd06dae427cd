"""Step I/O on the device (slevolve_amd/snapshots.py): the asynchronous snapshot writer produces the
same files, byte for byte, as writing each step's outputs synchronously; the ice history loaded
into HBM (from the ice directory or through the manifest-checked cache) equals the files."""
import os

import numpy as np
import pytest

from slevolve_amd import npyio
from slevolve_amd import synthetic as syn

pytestmark = pytest.mark.gpu

FIELDS = ("tgrid", "ocean", "beta", "R", "G", "dS", "dSLdt", "dRdt")


@pytest.mark.parametrize("io_compat", ["fortran", "legacy"])
def test_async_snapshots_equal_synchronous_files(tmp_path, prem_text, topo_seed, io_compat):
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver
    from slevolve_amd.snapshots import SnapshotWriter

    lat, lon, topo, ice = syn.synthetic_earth(64, topo_seed)
    slv = OptSeaLevelSolver(SHT(64, 128, 63), LoveNumbers(prem_text, 63, text=True), topo, ice[0], norder=63,
                            dt=500.0, epsilon1=1e-7, ramp=True)
    slv.horizontal = True
    sync_dir, async_dir = tmp_path / "sync", tmp_path / "async"
    sync_dir.mkdir()
    ref = SnapshotWriterReference(str(sync_dir), io_compat)
    with SnapshotWriter(str(async_dir), FIELDS, io_compat=io_compat, depth=2) as w:
        for k in range(1, 7):
            print(f"[snapshots {io_compat}] step {k}", flush=True)
            r = slv.step(ice[k], rates=True)
            w.write(k, r)
            ref.write(k, r)
    names = sorted(os.listdir(sync_dir))
    assert len(names) == 6 * 10 and names == sorted(os.listdir(async_dir))
    for n in names:
        assert open(sync_dir / n, "rb").read() == open(async_dir / n, "rb").read(), n


class SnapshotWriterReference:
    """The run_ais.py:400-425 output block, synchronously, through npyio."""

    def __init__(self, out, io_compat):
        self.out, self.io_compat = out, io_compat

    def write(self, k, r):
        p = lambda name: os.path.join(self.out, name)  # noqa: E731
        h = lambda t: t.cpu().numpy()  # noqa: E731
        npyio.save_grid(p(f"tgrid{k}.npy"), h(r["tgrid"]))
        npyio.save_mask(p(f"ocean{k}.npy"), h(r["ocean"]), self.io_compat)
        npyio.save_mask(p(f"beta{k}.npy"), h(r["beta"]), self.io_compat)
        npyio.save_grid(p(f"R{k}.npy"), h(r["R"]))
        npyio.save_grid(p(f"G{k}.npy"), h(r["G"]))
        if self.io_compat == "fortran":
            npyio.save_spectral(p(f"dS_converged{k}.npy"), h(r["dS_converged"]))
        else:
            np.save(p(f"dS_converged{k}.npy"), h(r["dS_converged"]))
        rates = {q: (h(v) if hasattr(v, "cpu") else v) for q, v in r["rates"].items()}
        npyio.save_rates(self.out, f"{k}", rates, FIELDS)
        npyio.save_spectral(p(f"horiz{k}.npy"), h(r["horiz"]))


def test_ice_stack_in_hbm(tmp_path):
    import torch

    from slevolve_amd.snapshots import IceStack

    nlat, nlon, n = 16, 32, 5
    d = tmp_path / "ice"
    d.mkdir()
    rng = np.random.default_rng(7)
    src = [np.round(rng.random((nlat, nlon)) * 3000.0, 3) for _ in range(n + 1)]
    for k, g in enumerate(src):
        if k % 3 == 0:
            np.save(d / f"ice{k}.npy", g)
        elif k % 3 == 1:
            g.astype("<f8").tofile(d / f"ice{k}.bin")
        else:
            npyio.write_text_grid(str(d / f"ice{k}"), g, "row-e.18", ncol=nlon)
    direct = IceStack(str(d), "ice", n, nlat, nlon, chunk=2)
    cached = IceStack(str(d), "ice", n, nlat, nlon, cache_dir=str(tmp_path / "cache"), chunk=4)
    again = IceStack(str(d), "ice", n, nlat, nlon, cache_dir=str(tmp_path / "cache"))
    assert len(direct) == n + 1 and again.cache_seconds == 0.0 and cached.cache_seconds > 0.0
    for k in range(n + 1):
        for s in (direct, cached, again):
            assert s[k].is_cuda and s[k].dtype == torch.float64
            assert np.array_equal(s[k].cpu().numpy(), src[k]), k
