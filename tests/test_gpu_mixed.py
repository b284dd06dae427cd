"""Mixed precision: the FP32-product Legendre contractions (FP64 recurrence, FP32 matrix-core
products) against the oracle, and the mixed-precision solver (FP32-product sweeps until the masks
stop changing or the convergence ratio reaches eps_switch = 1000 epsilon1, then FP64 sweeps to
convergence) against the FP64 solver, field by field and iteration count by iteration count, and
against the oracle's lmax-1023 fixture.

Tolerances: an FP32 product of a value rounded to 2^-24 relative and a sum of at most 512 ring
pairs (analysis) or 1024 degrees (synthesis) in FP32 are good to a few 1e-7 of the largest term;
the transforms are held to 2e-6 of the largest output.  The solver's bar is the metric's:
max |dSL| <= 1e-6 m against the FP64 path (dSL = tinit - tgrid, so dSL differences are tgrid
differences), reported separately for R and the rotation vector m.
"""
import os

import numpy as np
import pytest
import torch

from oracle.sht import OracleSHT
from slevolve_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def rand_coeffs(L, rng):
    c = np.tril(rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1)))
    c[:, 0] = c[:, 0].real
    return c


@pytest.mark.parametrize("nlat,L", [(64, 63), (64, 31), (128, 63), (256, 255), (512, 512), (1024, 1023)])
def test_fp32_product_transforms_vs_oracle(nlat, L):
    from slevolve_amd import SHT

    print(f"[x32 sht] nlat {nlat} lmax {L}", flush=True)
    nlon = 2 * nlat
    rng = np.random.default_rng(nlat + 3 * L)
    gpu, ref = SHT(nlat, nlon, L, precision="mixed"), OracleSHT(nlat, nlon, L)
    c = rand_coeffs(L, rng)
    g_ref = ref.synthesis(c)
    g_gpu = gpu.synthesis(torch.from_numpy(c).cuda()).cpu().numpy()
    es = float(np.abs(g_gpu - g_ref).max() / np.abs(g_ref).max())
    fields = rng.standard_normal((2, nlat, nlon))
    out = torch.zeros((2, gpu.ncoef), dtype=torch.complex128, device="cuda")
    gpu.analysis_packed(torch.from_numpy(fields).cuda(), out)
    ea = 0.0
    for f in range(2):
        U_ref = ref.analysis(fields[f])
        U_gpu = gpu.to_dense(out[f]).cpu().numpy()
        ea = max(ea, float(np.abs(U_gpu - U_ref).max() / np.abs(U_ref).max()))
        assert np.abs(np.triu(U_gpu, 1)).max() == 0.0
    one = torch.zeros((1, gpu.ncoef), dtype=torch.complex128, device="cuda")
    gpu.analysis_packed(torch.from_numpy(fields[1:2].copy()).cuda(), one)
    e1 = float((one[0] - out[1]).abs().max().item() / out[1].abs().max().item())
    print(f"[x32 sht] nlat {nlat} lmax {L}: synthesis {es:.2e}, analysis {ea:.2e}, 1 vs 2 fields {e1:.2e}",
          flush=True)
    assert es < 2e-6 and ea < 2e-6 and e1 < 2e-6, (es, ea, e1)


def _pair(w, nsteps, **kw):
    """FP64 and mixed solvers on workload w side by side for nsteps steps; per-step max errors."""
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    def make(prec):
        return OptSeaLevelSolver(SHT(w.nlat, w.nlon, w.lmax), LoveNumbers(w.love_text, w.lmax, text=True), w.topo,
                                 w.ice(0), norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner="extrap", precision=prec,
                                 **kw)

    a, b = make("fp64"), make("mixed")
    rows = []
    for k in range(1, nsteps + 1):
        ra, rb = a.step(w.ice(k)), b.step(w.ice(k))
        e = {"step": k, "numiter": (ra["numiter"], rb["numiter"]), "sweeps32": b._last_status.sweeps32,
             "dSL": float((ra["tgrid"] - rb["tgrid"]).abs().max().item()),
             "dR": float((ra["R"] - rb["R"]).abs().max().item()),
             "dm": float(np.abs(ra["mm"] - rb["mm"]).max()), "m": float(np.abs(ra["mm"]).max())}
        rows.append(e)
        print(f"[mixed {w.name}] {e}", flush=True)
    return a, b, rows, rb


@pytest.mark.parametrize("name,nsteps", [("lmax63", 12), ("lmax512", 10), ("lmax1023", 10)])
def test_mixed_solver_vs_fp64(name, nsteps):
    w = syn.workload(name)
    a, b, rows, _ = _pair(w, nsteps)
    dsl = max(r["dSL"] for r in rows)
    drr = max(r["dR"] for r in rows)
    dm = max(r["dm"] / max(r["m"], 1e-30) for r in rows)
    print(f"[mixed {name}] max |dSL| {dsl:.2e} m, max |dR| {drr:.2e} m, max |dm|/|m| {dm:.2e}", flush=True)
    assert sum(r["sweeps32"] for r in rows) > 0  # the reduced sweeps ran
    assert all(r["numiter"][0] == r["numiter"][1] for r in rows)
    assert dsl <= 1e-6 and drr <= 1e-6 and dm <= 1e-6


def test_mixed_lmax1023_vs_oracle_fixture():
    """The mixed path on the 1024 x 2048, lmax 1023 workload against the oracle's stored outputs
    (tests/golden/oracle_lmax1023.npz), to the metric's 1e-6 m."""
    from slevolve_amd import SHT, LoveNumbers, OptSeaLevelSolver

    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_lmax1023.npz"))
    w = syn.workload("lmax1023")
    slv = OptSeaLevelSolver(SHT(w.nlat, w.nlon, w.lmax), LoveNumbers(w.love_text, w.lmax, text=True), w.topo,
                            w.ice(0), norder=w.lmax, dt=w.dt, epsilon1=1e-5, inner="extrap", precision="mixed")
    s = int(f["sub"])
    for k in range(f["numiter"].size):
        r = slv.step(w.ice(k + 1))
        errs = {}
        for name in ("tgrid", "R"):
            g = r[name].cpu().numpy()
            errs[name] = float(np.abs(g[::s, ::s] - f[name + "_sub"][k]).max())
        print(f"[mixed lmax1023 vs oracle] step {k + 1} numiter {r['numiter']} (oracle {int(f['numiter'][k])}) "
              f"errors {errs}", flush=True)
        assert errs["tgrid"] <= 1e-6 and errs["R"] <= 1e-6, errs
        assert np.allclose(r["mm"], f["mm"][k], rtol=1e-6, atol=1e-20)
