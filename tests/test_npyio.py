"""File conventions (slevolve_amd/npyio.py) against fortran/IO_FORMATS.md and python/slevolve/npyio.py:
.npy headers and payload order, masks, spectral Fortran order, f4 losslessness, text layouts
(byte-for-byte text -> npy -> text), reader resolution order, times file, ice-cache manifest,
the TPWrate record and the NetCDF export."""
import os

import numpy as np
import pytest

from slevolve_amd import npyio

HEADER_512 = b"\x93NUMPY\x01\x00v\x00{'descr': '<f8', 'fortran_order': False, 'shape': (512, 1024), }"


def payload(path):
    raw = open(path, "rb").read()
    hlen = int.from_bytes(raw[8:10], "little")
    return raw[: 10 + hlen], raw[10 + hlen:]


def test_grid_header_and_payload(tmp_path):
    a = np.random.default_rng(0).normal(size=(512, 1024))
    p = str(tmp_path / "tgrid1.npy")
    assert npyio.save_grid(p, a) == "f8"
    head, body = payload(p)
    assert head.startswith(HEADER_512) and head.endswith(b"\n") and len(head) == 128
    assert body == a.astype("<f8").tobytes()  # stream order = the headerless .bin cache


def test_f4_only_when_lossless(tmp_path):
    exact = np.arange(512 * 1024, dtype=np.float32).reshape(512, 1024).astype(np.float64) / 8
    p = str(tmp_path / "ice0.npy")
    assert npyio.f4_exact(exact) and npyio.save_grid(p, exact, "f4") == "f4"
    assert np.load(p).dtype == np.dtype("<f4") and np.array_equal(npyio.read_grid(p), exact)
    lossy = exact + 1e-9
    assert not npyio.f4_exact(lossy) and npyio.save_grid(p, lossy, "f4") == "f8"
    assert np.array_equal(npyio.read_grid(p), lossy)


def test_spectral_fortran_order_and_masks(tmp_path):
    L = 7
    z = np.tril(np.random.default_rng(1).normal(size=(L + 1, L + 1)) + 1j)
    p = str(tmp_path / "dS_converged3.npy")
    npyio.save_spectral(p, z)
    head, body = payload(p)
    assert b"'fortran_order': True" in head and b"'<c16'" in head
    assert body == z.astype("<c16").ravel(order="F").tobytes()  # Fortran dS(0:N, 0:N) storage
    assert np.array_equal(np.load(p), z)
    m = (np.random.default_rng(2).random((512, 1024)) > 0.5).astype(np.float64)
    npyio.save_mask(str(tmp_path / "ocean1.npy"), m)
    npyio.save_mask(str(tmp_path / "beta1.npy"), m, io_compat="legacy")
    assert np.load(str(tmp_path / "ocean1.npy")).dtype == np.dtype("<f8")
    b = np.load(str(tmp_path / "beta1.npy"))
    assert b.dtype == np.uint8 and np.array_equal(b, m.astype(np.uint8))


@pytest.mark.parametrize("fmt", sorted(npyio.TEXT_FORMATS))
def test_text_layouts_round_trip_byte_for_byte(tmp_path, fmt):
    nlat, nlon = 8, 1024
    rng = np.random.default_rng(3)
    a = rng.normal(scale=1e3, size=(nlat, nlon)) * (rng.random((nlat, nlon)) > 0.2)
    t1 = str(tmp_path / "grid")
    npyio.write_text_grid(t1, a, fmt)
    f, sep, per = npyio.TEXT_FORMATS[fmt]
    lines = open(t1).read().splitlines()
    assert len(lines) == nlat * nlon // per
    assert lines[0] == sep.join(f % v for v in a.ravel()[:per])
    # parse -> .npy -> text reproduces the file; the values are the formatted decimals
    (npy, _), = npyio.text2npy([t1], nlat, nlon)
    g = np.load(npy)
    assert np.array_equal(g, np.array([float(f % v) for v in a.ravel()]).reshape(nlat, nlon))
    (tmp_path / "back").mkdir()
    out = npyio.npy2text([npy], fmt, outdir=str(tmp_path / "back"))
    assert open(out[0], "rb").read() == open(t1, "rb").read()
    if fmt == "row-e.18":
        assert np.array_equal(g, a)  # 19 significant digits round-trip a double exactly


def test_text_layout_partial_last_line(tmp_path):
    a = np.arange(10.0)
    p = str(tmp_path / "t")
    npyio.write_text_grid(p, a, "row-e.18", ncol=4)
    assert [len(x.split()) for x in open(p).read().splitlines()] == [4, 4, 2]
    assert np.array_equal(npyio.read_text_grid(p, 2, 5).ravel(), a)


def test_reader_resolution_order(tmp_path):
    nlat, nlon = 4, 8
    a = np.arange(32.0).reshape(nlat, nlon)
    d = str(tmp_path)
    npyio.write_text_grid(os.path.join(d, "ice0"), a + 2)
    assert np.array_equal(npyio.find_grid(d, "ice0", nlat, nlon), a + 2)
    (a + 1).astype("<f8").tofile(os.path.join(d, "ice0.bin"))
    assert np.array_equal(npyio.find_grid(d, "ice0", nlat, nlon), a + 1)
    np.save(os.path.join(d, "ice0.npy"), a.ravel())  # flat: taken in stream order
    assert np.array_equal(npyio.find_grid(d, "ice0", nlat, nlon), a)
    assert npyio.grid_source(d, "ice0").endswith(".npy") and npyio.grid_source(d, "nope") is None
    np.zeros(5).tofile(os.path.join(d, "bad.bin"))
    with pytest.raises(ValueError):
        npyio.read_grid(os.path.join(d, "bad.bin"), nlat, nlon)
    assert np.array_equal(npyio.IceDir(d, "ice", nlat, nlon)[0], a)


def test_times_file(tmp_path):
    p = str(tmp_path / "times")
    np.savetxt(p, [-20000.0, -19500.0, -19000.0, -18000.0])
    t = npyio.read_times(p)
    assert np.array_equal(np.diff(t), [500.0, 500.0, 1000.0])
    assert npyio.read_times(p, 2).size == 3
    with pytest.raises(ValueError):
        npyio.read_times(p, 5)
    np.savetxt(p, [0.0, 1.0, 1.0])
    with pytest.raises(ValueError):
        npyio.read_times(p)
    np.savetxt(p, [0.0])
    with pytest.raises(ValueError):
        npyio.read_times(p)


def test_ice_cache_manifest(tmp_path):
    nlat, nlon = 4, 8
    d = tmp_path / "ice"
    d.mkdir()
    for k in range(4):
        if k % 2:
            np.save(d / f"ice{k}.npy", np.full((nlat, nlon), float(k)))
        else:
            npyio.write_text_grid(str(d / f"ice{k}"), np.full((nlat, nlon), float(k)))
    cache = str(tmp_path / "cache")
    path, secs = npyio.build_ice_cache(str(d), "ice", 3, nlat, nlon, cache)
    stack = np.load(path)
    assert stack.shape == (4, nlat, nlon) and all((stack[k] == k).all() for k in range(4))
    assert npyio.build_ice_cache(str(d), "ice", 3, nlat, nlon, cache)[1] == 0.0  # reused
    assert npyio.build_ice_cache(str(d), "ice", 2, nlat, nlon, cache) == (path, 0.0)  # longer cache accepted
    man = npyio.ice_manifest(str(d), "ice", 4)
    assert [f[0] for f in man["files"]] == ["ice0", "ice1.npy", "ice2", "ice3.npy"]
    assert npyio.cache_matches(path, man) == (True, "")
    assert npyio.cache_matches(path + "x", man) == (False, "cache file missing")
    assert npyio.cache_matches(path, dict(man, icedir="/elsewhere"))[1] == "built from another ice directory"
    assert npyio.cache_matches(path, dict(man, prefix="other"))[1] == "built for another file prefix"
    assert npyio.cache_matches(path, npyio.ice_manifest(str(d), "ice", 3))[0]
    np.save(d / "ice4.npy", np.zeros((nlat, nlon)))
    assert npyio.cache_matches(path, npyio.ice_manifest(str(d), "ice", 5))[1] == "holds fewer ice files than needed"
    st = os.stat(d / "ice1.npy")
    os.utime(d / "ice1.npy", ns=(st.st_atime_ns, st.st_mtime_ns + 10**9))
    assert npyio.cache_matches(path, npyio.ice_manifest(str(d), "ice", 4))[1] == "an ice file's size or modification time differs"
    path2, secs2 = npyio.build_ice_cache(str(d), "ice", 3, nlat, nlon, cache, log=lambda *a: None)
    assert path2 == path and secs2 > 0.0
    os.remove(npyio.manifest_path(path))
    assert npyio.cache_matches(path, man)[1] == "sidecar missing"


def test_tpw_rate_record(tmp_path):
    rng = np.random.default_rng(4)
    rates = {"dSLdt": rng.normal(size=(4, 8)), "dRdt": rng.normal(size=(4, 8)), "ildot": rng.normal(size=(3, 3)),
             "mdot": rng.normal(size=3), "lambdadot": rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))}
    npyio.save_rates(str(tmp_path), "7", rates, fields=("dSLdt",))
    assert os.path.exists(tmp_path / "dSLdt7.npy") and not os.path.exists(tmp_path / "dRdt7.npy")
    lines = open(tmp_path / "TPWrate7.txt").read().splitlines()
    assert [len(x) for x in lines] == [9 * 19, 3 * 19, 18 * 19]
    vals = [np.array([float(x[i:i + 19]) for i in range(0, len(x), 19)]) for x in lines]
    assert np.allclose(vals[0], rates["ildot"].ravel(order="F"), rtol=1e-8)
    assert np.allclose(vals[1], rates["mdot"], rtol=1e-8)
    lam = rates["lambdadot"].ravel(order="F")
    assert np.allclose(vals[2][0::2], lam.real, rtol=1e-8) and np.allclose(vals[2][1::2], lam.imag, rtol=1e-8)


def test_netcdf_export(tmp_path):
    from scipy.io import netcdf_file

    nlat, nlon = 16, 32
    run = tmp_path / "run"
    run.mkdir()
    rng = np.random.default_rng(5)
    grids = {}
    for k in (0, 1, 2):
        grids[k] = rng.normal(size=(nlat, nlon))
        npyio.save_grid(str(run / f"tgrid{k}.npy"), grids[k])
        npyio.save_mask(str(run / f"ocean{k}.npy"), grids[k] < 0)
    out = npyio.npy2nc(str(run), str(tmp_path / "run.nc"), fields=("tgrid", "ocean"), dt=500.0, earth="prem")
    with netcdf_file(out, "r", mmap=False) as nc:
        assert nc.dimensions["time"] == 3 and nc.dimensions["lat"] == nlat and nc.dimensions["lon"] == nlon
        assert np.array_equal(nc.variables["time"][:], [0.0, 500.0, 1000.0])
        lat = nc.variables["lat"][:]
        x, _ = np.polynomial.legendre.leggauss(nlat)
        assert np.all(np.diff(lat) < 0) and np.allclose(np.sin(np.radians(lat)), x[::-1], atol=1e-14)
        assert np.array_equal(nc.variables["lon"][:], np.arange(nlon) * 360.0 / nlon)
        for k in range(3):
            assert np.array_equal(nc.variables["tgrid"][k], grids[k])
            assert np.array_equal(nc.variables["ocean"][k], (grids[k] < 0).astype(float))
        assert nc.variables["tgrid"].units == b"m" and nc.earth_model == b"prem"
    np.savetxt(str(run / "times"), [-100.0, -50.0, -25.0])
    out = npyio.npy2nc(str(run), str(tmp_path / "run2.nc"), lon_convention="sampled")
    with netcdf_file(out, "r", mmap=False) as nc:
        assert np.array_equal(nc.variables["time"][:], [-100.0, -50.0, -25.0])
        assert nc.variables["lon"][0] == 360.0 / nlon
