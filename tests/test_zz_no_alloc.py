"""slx_sht_analysis_noalloc (the allocation-free analysis) through the C-ABI against
slx_sht_analysis, bit for bit: two calls in a row on different handles, the second with a
smaller input, plus the size limit and the FP32-product precision."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _both(lib, sht, grids):
    """(slx_sht_analysis, slx_sht_analysis_noalloc) outputs for the fields in grids."""
    from slevolve_amd import _device as D
    from slevolve_amd import _lib

    nf = grids.shape[0]
    ref = torch.zeros((nf, sht.ncoef), dtype=torch.complex128, device="cuda")
    out = torch.full((nf, sht.ncoef), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    _lib.check(lib.slx_sht_analysis(sht.handle, grids.data_ptr(), nf, ref.data_ptr(), D.stream_ptr()))
    _lib.check(lib.slx_sht_analysis_noalloc(sht.handle, grids.data_ptr(), nf, out.data_ptr(), D.stream_ptr()))
    torch.cuda.synchronize()
    return ref, out


def test_noalloc_analysis_matches_the_allocating_entry_point():
    from slevolve_amd import SHT, _lib

    lib = _lib.load()
    rng = np.random.default_rng(11)
    # first call: three fields at lmax 512 on 512 x 1024 (12.6 MB of the 24 MiB scratch)
    big = SHT(512, 1024, 512)
    g = torch.from_numpy(rng.standard_normal((3, 512, 1024))).cuda()
    print("[noalloc] 3 fields, 512 x 1024, lmax 512", flush=True)
    ref, out = _both(lib, big, g)
    assert torch.equal(ref, out)
    # second call, smaller input: one field at lmax 63 on 128 x 256 (the array keeps the first call's
    # contents beyond what this call writes)
    small = SHT(128, 256, 63)
    g2 = torch.from_numpy(rng.standard_normal((1, 128, 256))).cuda()
    print("[noalloc] 1 field, 128 x 256, lmax 63", flush=True)
    ref2, out2 = _both(lib, small, g2)
    assert torch.equal(ref2, out2)
    # FP32-product handle
    mixed = SHT(128, 256, 63, precision="mixed")
    ref3, out3 = _both(lib, mixed, g2)
    assert torch.equal(ref3, out3)
    # beyond the scratch (rejected before any launch): two fields at lmax 1023 on 1024 x 2048 need
    # 2 x 2 x 1024 x 512 x 16 B = 32 MiB; one field (16 MiB) fits
    from slevolve_amd import _device as D

    huge = SHT(1024, 2048, 1023)
    g3 = torch.zeros((2, 1024, 2048), dtype=torch.float64, device="cuda")
    out4 = torch.zeros((2, huge.ncoef), dtype=torch.complex128, device="cuda")
    rc = lib.slx_sht_analysis_noalloc(huge.handle, g3.data_ptr(), 2, out4.data_ptr(), D.stream_ptr())
    assert rc == -22 and b"static scratch" in lib.slx_last_error()
    g4 = torch.from_numpy(rng.standard_normal((1, 1024, 2048))).cuda()
    print("[noalloc] 1 field, 1024 x 2048, lmax 1023", flush=True)
    ref5, out5 = _both(lib, huge, g4)
    assert torch.equal(ref5, out5)
