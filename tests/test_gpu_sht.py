"""GPU transform and modal-state parity against the oracle (FP64)."""
import numpy as np
import pytest
import torch

from oracle.memory import ExactMemory as OMemory
from oracle.sht import OracleSHT

pytestmark = pytest.mark.gpu


def rand_coeffs(L, rng):
    c = np.tril(rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1)))
    c[:, 0] = c[:, 0].real
    return c


@pytest.mark.parametrize("variant", ["mfma2", "mfma4", "mw1", "mw2", "mw1x8", "mw2x8", "scalar"])
@pytest.mark.parametrize("nlat,L", [(64, 63), (128, 63), (64, 31), (64, 62), (16, 4), (512, 511), (512, 512), (1024, 1023)])
def test_synthesis_and_analysis_match_oracle(nlat, L, variant, monkeypatch):
    """Every synthesis kernel (matrix cores with the waves over segment groups or over rings, and
    scalar-fed VALU; SLX_SYN_VARIANT is read when a transform handle is created) and the analysis
    against the oracle."""
    from slevolve_amd import SHT

    monkeypatch.setenv("SLX_SYN_VARIANT", variant)
    print(f"[sht] nlat {nlat} lmax {L} synthesis {variant}", flush=True)
    nlon = 2 * nlat
    rng = np.random.default_rng(nlat + L)
    gpu, ref = SHT(nlat, nlon, L), OracleSHT(nlat, nlon, L)
    assert np.abs(gpu.gaulats - ref.gaulats).max() < 1e-15
    c = rand_coeffs(L, rng)
    g_ref = ref.synthesis(c)
    g_gpu = gpu.synthesis(torch.from_numpy(c).cuda()).cpu().numpy()
    scale = np.abs(g_ref).max()
    assert np.abs(g_gpu - g_ref).max() < 1e-13 * scale * max(1, L / 64)
    field = rng.standard_normal((nlat, nlon))
    U_ref = ref.analysis(field)
    U_gpu = gpu.analysis(torch.from_numpy(field).cuda()).cpu().numpy()
    assert np.abs(U_gpu - U_ref).max() < 1e-14 * max(1, L / 64)
    assert np.abs(np.triu(U_gpu, 1)).max() == 0.0
    # numpy in -> numpy out, same numbers
    assert np.array_equal(gpu.analysis(field), U_gpu)
    if L < nlat:  # exact round trip needs lmax <= nlat - 1 (aliased top degree otherwise)
        back = gpu.analysis(gpu.synthesis(torch.from_numpy(c).cuda())).cpu().numpy()
        assert np.abs(back - c).max() / np.abs(c).max() < 1e-12 * max(1.0, L / 511)


@pytest.mark.parametrize("variant", ["mf2", "mf4", "ck2", "ck4", "qw2", "qw4", "qw4:48:257", "qw2:1048576:300"])
@pytest.mark.parametrize("nlat,L", [(64, 63), (512, 512), (1024, 1023)])
def test_analysis_variants_match_oracle(nlat, L, variant, monkeypatch):
    """Every matrix-core analysis kernel (one group per block, persistent with chunked
    accumulators, or walking long chains, 2 or 4 rings per lane; SLX_ANA_VARIANT is read when a
    handle is created, and a shape that does not fit falls back to 4 rings) against the oracle, 1
    and 2 fields.  "qw4:T:B" sets the walk kernel's chain length and block count (SLX_QW_T,
    SLX_QW_BLOCKS) so that blocks run several units one after another."""
    from slevolve_amd import SHT

    name, *walk = variant.split(":")
    monkeypatch.setenv("SLX_ANA_VARIANT", name)
    if walk:
        monkeypatch.setenv("SLX_QW_T", walk[0])
        monkeypatch.setenv("SLX_QW_BLOCKS", walk[1])
    print(f"[sht] nlat {nlat} lmax {L} analysis {variant}", flush=True)
    nlon = 2 * nlat
    rng = np.random.default_rng(nlat + L + 7)
    gpu, ref = SHT(nlat, nlon, L), OracleSHT(nlat, nlon, L)
    fields = rng.standard_normal((2, nlat, nlon))
    out = torch.zeros((2, gpu.ncoef), dtype=torch.complex128, device="cuda")
    gpu.analysis_packed(torch.from_numpy(fields).cuda(), out)
    for f in range(2):
        U_ref = ref.analysis(fields[f])
        U_gpu = gpu.to_dense(out[f]).cpu().numpy()
        assert np.abs(U_gpu - U_ref).max() < 1e-14 * max(1, L / 64)
    one = torch.zeros((1, gpu.ncoef), dtype=torch.complex128, device="cuda")
    gpu.analysis_packed(torch.from_numpy(fields[1:2].copy()).cuda(), one)
    assert torch.equal(one[0], out[1])


def test_multi_field_analysis_matches_single():
    from slevolve_amd import SHT

    nlat, L = 128, 127
    gpu = SHT(nlat, 2 * nlat, L)
    rng = np.random.default_rng(5)
    g = torch.from_numpy(rng.standard_normal((3, nlat, 2 * nlat))).cuda()
    out3 = torch.zeros((3, gpu.ncoef), dtype=torch.complex128, device="cuda")
    gpu.analysis_packed(g, out3)
    for f in range(3):
        one = torch.zeros((1, gpu.ncoef), dtype=torch.complex128, device="cuda")
        gpu.analysis_packed(g[f:f + 1].contiguous(), one)
        assert torch.equal(one[0], out3[f])


def test_modal_state_bit_exact(prem_text):
    from oracle.love import LoveNumbers as OLove
    from slevolve_amd import ExactMemory, LoveNumbers

    L = 63
    rng = np.random.default_rng(11)
    a = ExactMemory(LoveNumbers(prem_text, L, text=True), L, 1000.0, True)
    b = OMemory(OLove(prem_text, L, text=True), L, 1000.0)
    for step in range(6):
        if step == 3:
            a.set_dt(250.0)
            b.set_dt(250.0)
        va, ra, sila, sma, tsla, trra = a.advance_and_read()
        vb, rb, silb, smb, tslb, trrb = b.advance_and_read()
        assert np.array_equal(va.cpu().numpy(), vb) and np.array_equal(ra.cpu().numpy(), rb)
        assert np.array_equal(sila, silb) and np.array_equal(tsla, tslb)
        load = np.tril(rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1)))
        load[0] = 0
        dil, dm, dl = rng.standard_normal((3, 3)), rng.standard_normal(3), rng.standard_normal(3) + 0j
        a.commit(load, dil, dm, dl)
        b.commit(load, dil, dm, dl)


def test_modal_state_ramp_and_readouts(prem_text):
    """The second-order placement (ramp commit R_k += (1 - phi_k) L before P += L) and the
    horizontal and rate read-outs of the device ExactMemory against the oracle's, over steps of
    two lengths: states and horizontal sums bit for bit (the kernel pass with unit decay rewrites
    the states unchanged and sums k in order), the rate read-out to 1e-14 of its largest value."""
    from oracle.love import LoveNumbers as OLove
    from slevolve_amd import ExactMemory, LoveNumbers

    L = 63
    rng = np.random.default_rng(12)
    love_a, love_b = LoveNumbers(prem_text, L, text=True), OLove(prem_text, L, text=True)
    a = ExactMemory(love_a, L, 500.0, True)
    b = OMemory(love_b, L, 500.0)
    a.ramp = b.ramp = True
    a.horizontal_weights(love_a)
    b.horizontal_weights(love_b)
    assert np.array_equal(a.omphi, b.omphi) and np.array_equal(a.wsphi_sl, b.wsphi_sl)
    assert (a.c_il, a.c_m, a.c_tsl, a.c_trr) == (b.c_il, b.c_m, b.c_tsl, b.c_trr)
    for step in range(6):
        if step == 3:
            a.set_dt(100.0)
            b.set_dt(100.0)
        va, ra, *rot_a = a.advance_and_read()
        vb, rb, *rot_b = b.advance_and_read()
        assert np.array_equal(va.cpu().numpy(), vb) and np.array_equal(ra.cpu().numpy(), rb)
        assert all(np.array_equal(x, y) for x, y in zip(rot_a, rot_b))
        hza, thla = a.read_horizontal()
        hzb, thlb = b.read_horizontal()
        assert np.array_equal(hza.cpu().numpy(), hzb) and np.array_equal(thla, thlb)
        load = np.tril(rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1)))
        load[0] = 0
        load[:, 0] = load[:, 0].real
        dil, dm, dl = rng.standard_normal((3, 3)), rng.standard_normal(3), rng.standard_normal(3) + 0j
        rsa, rra = a.read_rate(load)
        rsb, rrb = b.read_rate(load)
        for x, y in ((rsa, rsb), (rra, rrb)):
            assert np.abs(x.cpu().numpy() - y).max() <= 1e-14 * np.abs(y).max()
        for x, y in zip(a.read_rate_rotation(dil, dm, dl), b.read_rate_rotation(dil, dm, dl)):
            assert np.array_equal(x, y)
        a.commit(load, dil, dm, dl)
        b.commit(load, dil, dm, dl)
        # states of existing modes (k < nmod[l]; the oracle's loop also carries the weightless others)
        Ra = torch.stack([a._sht.to_dense(a.R[k]) for k in range(a.kmax)]).cpu().numpy()
        live = np.arange(a.kmax)[:, None] < a.nmod[None, :]
        assert np.array_equal(Ra[live], b.R[live]) and np.array_equal(a.p_load.cpu().numpy(), b.p_load)
        assert np.array_equal(a.r_lambda, b.r_lambda) and np.array_equal(a.r_il, b.r_il)


@pytest.mark.parametrize("nlat,nlon,L", [(64, 4096, 63), (16, 8, 3)])
def test_ring_passes_at_the_row_length_limits(nlat, nlon, L):
    """The ring passes run nlon / 4 threads per ring pair up to nlon = 2048 and nlon / 8 at 4096;
    the widest and the narrowest rows against the oracle."""
    from slevolve_amd import SHT

    print(f"[sht] ring passes nlat {nlat} nlon {nlon} lmax {L}", flush=True)
    rng = np.random.default_rng(nlon + L)
    gpu, ref = SHT(nlat, nlon, L), OracleSHT(nlat, nlon, L)
    c = rand_coeffs(L, rng)
    g_ref = ref.synthesis(c)
    g_gpu = gpu.synthesis(torch.from_numpy(c).cuda()).cpu().numpy()
    assert np.abs(g_gpu - g_ref).max() < 1e-13 * np.abs(g_ref).max()
    field = rng.standard_normal((nlat, nlon))
    assert np.abs(gpu.analysis(torch.from_numpy(field).cuda()).cpu().numpy() - ref.analysis(field)).max() < 1e-14


def test_kernel_choice_is_a_rule_per_shape(monkeypatch):
    """Without SLX_ANA_VARIANT / SLX_SYN_VARIANT every handle of a shape takes the same kernels
    (a fixed rule per shape), so two handles give bit-identical transforms."""
    from slevolve_amd import SHT

    monkeypatch.delenv("SLX_ANA_VARIANT", raising=False)
    monkeypatch.delenv("SLX_SYN_VARIANT", raising=False)
    a, b = SHT(128, 256, 127), SHT(128, 256, 127)
    assert a.kernels == b.kernels == {"analysis": "k_leg_ana_ck", "analysis_rings_per_lane": 2,
                                      "synthesis": "k_leg_syn_mf", "synthesis_rings_per_lane": 2}
    rng = np.random.default_rng(3)
    field = torch.from_numpy(rng.standard_normal((128, 256))).cuda()
    assert torch.equal(a.analysis(field), b.analysis(field))
    c = torch.from_numpy(rand_coeffs(127, rng)).cuda()
    assert torch.equal(a.synthesis(c), b.synthesis(c))
    # the benchmark grids: the walk analysis, and the synthesis with the waves over rings from 256
    # ring pairs on
    assert SHT(512, 1024, 512).kernels == {"analysis": "k_leg_ana_qw", "analysis_rings_per_lane": 4,
                                           "synthesis": "k_leg_syn_mw", "synthesis_rings_per_lane": 2}
    assert SHT(1024, 2048, 1023).kernels == {"analysis": "k_leg_ana_qw", "analysis_rings_per_lane": 4,
                                             "synthesis": "k_leg_syn_mw", "synthesis_rings_per_lane": 2}
