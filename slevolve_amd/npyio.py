"""File conventions of the reference's solvers and their converters.

Formats (python/slevolve/npyio.py:1-14; fortran/IO_FORMATS.md sections 1-5):

* grids (ice, topography, tgrid/ocean/beta/R/G snapshots): ``.npy`` '<f8', shape (nlat, nlon),
  C order, row j = Gauss-Legendre latitude north -> south, column i = longitude; the payload is
  the stream order of the legacy text files and of the headerless little-endian '.bin' caches.
  Ice inputs may be '<f4' when that is lossless.
* masks (ocean, beta): '<f8' 0/1 with io_compat "fortran" (the default), uint8 with "legacy".
* spectral fields (dS_converged, horiz): '<c16' (N+1, N+1), a[l, m], written fortran_order=True
  so the payload is the Fortran dS(0:N, 0:N) storage order.
* text grids: whitespace-separated values in stream order; four writer layouts.
* times file: one time (years) per ice file, strictly increasing (npyio.py:180-193).
* ice cache: one (n+1, nlat, nlon) '<f8' stack with a JSON manifest of its sources
  (npyio.py:130-177, run_ais.py:180-227).
* rate outputs: dSLdt<N>.npy, dRdt<N>.npy and the TPWrate<N>.txt rotation record (npyio.py:196-209).
* NetCDF export of a run directory (python/slevolve/tools/npy2nc.py), written as NetCDF-3 with
  64-bit offsets through scipy (the netCDF4 library is not a dependency here).

Readers return float64 C-contiguous host arrays.  Device tensors given to writers are copied to
the host first; the asynchronous device-side snapshot path is ``slevolve_amd.snapshots``.
"""
from __future__ import annotations

import datetime
import json
import math
import os
import re

import numpy as np

NLAT, NLON = 512, 1024

# snapshot kind -> (dtype, fortran_order) of the file
ENCODINGS = {
    "grid": ("<f8", False),
    "mask": ("<f8", False),
    "mask_legacy": ("|u1", False),
    "spectral": ("<c16", True),
}

# text layout -> (printf format of one value, separator, values per line; 1 = one per line)
TEXT_FORMATS = {
    "es16.9": ("%16.9E", "", 1),
    "row-e16.7": ("%16.7e", "", 1024),
    "row-e.18": ("%.18e", " ", 1024),
    "row-E18.10": ("%18.10E", " ", 1024),
}

# CF units and long names of the exported snapshot fields
FIELD_META = {
    "tgrid": ("m", "topography (bedrock + grounded ice), relative to sea level"),
    "ocean": ("1", "ocean function (1 = ocean)"),
    "beta": ("1", "grounded-ice mask function beta (1 = ice free)"),
    "R": ("m", "radial displacement of the solid surface"),
    "G": ("m", "geoid (sea-surface) perturbation"),
}


def _host(a):
    """numpy view of a host array or a device tensor (copied to the host)."""
    if hasattr(a, "detach") and hasattr(a, "cpu"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


# ------------------------------------------------------------------------------ writers
def encode(kind, a):
    """The array exactly as a snapshot of ``kind`` is stored (dtype and memory order)."""
    dtype, forder = ENCODINGS[kind]
    a = _host(a)
    if kind == "mask_legacy":
        return np.ascontiguousarray(a).astype(dtype)
    out = np.asarray(a, dtype=dtype)
    return np.asfortranarray(out) if forder else np.ascontiguousarray(out)


def save(path, kind, a):
    np.save(path, encode(kind, a))


def f4_exact(a):
    """True if every value survives a round trip through float32."""
    a = np.asarray(a, dtype=np.float64)
    with np.errstate(over="ignore", invalid="ignore"):
        back = a.astype(np.float32).astype(np.float64)
    return bool(np.array_equal(back, a))


def save_grid(path, a, dtype="f8"):
    """Grid as '<f8' ('<f4' if asked and lossless); returns the dtype written."""
    g = np.ascontiguousarray(_host(a), dtype=np.float64)
    if dtype == "f4" and f4_exact(g):
        np.save(path, g.astype("<f4"))
        return "f4"
    np.save(path, g.astype("<f8", copy=False))
    return "f8"


def save_mask(path, m, io_compat="fortran"):
    save(path, "mask" if io_compat == "fortran" else "mask_legacy", m)


def save_spectral(path, z):
    save(path, "spectral", z)


def write_text_grid(path, a, fmt="es16.9", ncol=None):
    """Grid in one of the legacy text layouts, values in C-flat (stream) order."""
    f, sep, per = TEXT_FORMATS[fmt]
    if ncol is not None and per != 1:
        per = int(ncol)
    flat = np.ascontiguousarray(_host(a), dtype=np.float64).ravel()
    nfull = (flat.size // per) * per
    with open(path, "w") as fh:
        if nfull:
            np.savetxt(fh, flat[:nfull].reshape(-1, per), fmt=f, delimiter=sep)
        if nfull < flat.size:
            fh.write(sep.join(f % v for v in flat[nfull:].tolist()) + "\n")


def save_rates(out, tag, rates, fields=("dSLdt", "dRdt")):
    """dSLdt<tag>.npy, dRdt<tag>.npy (m/yr grids) and, with rotation, TPWrate<tag>.txt: three lines
    of %19.8E values: il-dot (9, column-major), m-dot (3), Lambda-dot (3 x 3 complex, column-major,
    re and im interleaved)."""
    for f in ("dSLdt", "dRdt"):
        if f in fields:
            save(os.path.join(out, f"{f}{tag}.npy"), "grid", rates[f])
    if "mdot" in rates:
        il = np.asarray(rates["ildot"], dtype=np.float64).T.ravel()
        lam = np.asarray(rates["lambdadot"], dtype=np.complex128).T.ravel()
        lv = np.empty(2 * lam.size)
        lv[0::2], lv[1::2] = lam.real, lam.imag
        line = lambda v: "".join("%19.8E" % x for x in np.asarray(v, dtype=np.float64))  # noqa: E731
        with open(os.path.join(out, f"TPWrate{tag}.txt"), "w") as fh:
            fh.write("\n".join((line(il), line(rates["mdot"]), line(lv))) + "\n")


# ------------------------------------------------------------------------------ readers
def read_text_grid(path, nlat=NLAT, nlon=NLON):
    """Whitespace-separated text grid (any values per line) -> (nlat, nlon), stream order."""
    a = np.fromfile(path, sep=" ")
    if a.size != nlat * nlon:
        raise ValueError(f"{path}: got {a.size} values, expected {nlat * nlon}")
    return a.reshape(nlat, nlon)


def load_npy(path, mmap=False):
    return np.load(path, mmap_mode="r" if mmap else None, allow_pickle=False)


def read_grid(path, nlat=NLAT, nlon=NLON):
    """A grid from '.npy' (any real dtype, either order; a flat array of nlat*nlon values is taken
    in stream order), a headerless little-endian f8 '.bin', or text.  A path without either
    extension is looked up as <path>.npy, then <path>.bin, then the text file itself."""
    if path.endswith(".npy"):
        a = np.load(path, allow_pickle=False)
        if a.shape != (nlat, nlon):
            if a.size != nlat * nlon:
                raise ValueError(f"{path}: shape {a.shape}, expected {(nlat, nlon)}")
            a = a.reshape(nlat, nlon)
        return np.ascontiguousarray(a, dtype=np.float64)
    if path.endswith(".bin"):
        nbytes = os.path.getsize(path)
        if nbytes != 8 * nlat * nlon:
            raise ValueError(f"{path}: {nbytes} bytes, expected {8 * nlat * nlon} (LE f8)")
        return np.fromfile(path, dtype="<f8").reshape(nlat, nlon)
    src = grid_source(os.path.dirname(path) or ".", os.path.basename(path))
    if src is not None and src != path:
        return read_grid(src, nlat, nlon)
    return read_text_grid(path, nlat, nlon)


def grid_source(dirname, name):
    """The file a grid <dirname>/<name> is read from (.npy, then .bin, then text), or None."""
    for ext in (".npy", ".bin", ""):
        p = os.path.join(dirname, name + ext)
        if os.path.isfile(p):
            return p
    return None


def find_grid(dirname, name, nlat=NLAT, nlon=NLON):
    return read_grid(os.path.join(dirname, name), nlat, nlon)


class IceDir:
    """ice[k] read on demand from <dir>/<pfx><k> (.npy, .bin or text)."""

    def __init__(self, d, pfx, nlat=NLAT, nlon=NLON):
        self.d, self.pfx, self.nlat, self.nlon = d, pfx, nlat, nlon

    def __getitem__(self, k):
        return find_grid(self.d, f"{self.pfx}{k}", self.nlat, self.nlon)


def read_times(path, nsteps=None):
    """Times t_0..t_N (years) of the ice files, strictly increasing; with nsteps, the first
    nsteps + 1 (an error if fewer).  Step j lasts t_j - t_{j-1}."""
    t = np.loadtxt(path, dtype=np.float64, ndmin=1)
    if t.ndim != 1 or t.size < 2:
        raise ValueError(f"{path}: need one time per line, at least 2 values")
    if nsteps is not None:
        if t.size < nsteps + 1:
            raise ValueError(f"{path}: {t.size} times, need {nsteps + 1} (files 0..{nsteps})")
        t = t[: nsteps + 1]
    if not (np.isfinite(t).all() and (np.diff(t) > 0).all()):
        raise ValueError(f"{path}: times must be finite and strictly increasing")
    return t


# --------------------------------------------------------------------------- ice cache
def ice_manifest(dirname, pfx, nfiles):
    """Identity of an ice cache's sources: directory, prefix, count and, per file k < nfiles, the
    [name, size, mtime_ns] of the file a reader would use."""
    files = []
    for k in range(nfiles):
        src = grid_source(dirname, f"{pfx}{k}")
        if src is None:
            raise FileNotFoundError(os.path.join(dirname, f"{pfx}{k}"))
        st = os.stat(src)
        files.append([os.path.basename(src), int(st.st_size), int(st.st_mtime_ns)])
    return {"icedir": os.path.abspath(dirname), "prefix": pfx, "nfiles": int(nfiles), "files": files}


def manifest_path(cache_path):
    return cache_path + ".json"


def write_manifest(cache_path, man):
    with open(manifest_path(cache_path), "w") as fh:
        json.dump(man, fh, indent=0)


def cache_matches(cache_path, man):
    """(ok, reason): whether the cache was built from the files of ``man``.  A cache of more files
    matches when its leading entries do."""
    if not os.path.exists(cache_path):
        return False, "no cache"
    mp = manifest_path(cache_path)
    if not os.path.exists(mp):
        return False, "no manifest"
    try:
        with open(mp) as fh:
            old = json.load(fh)
    except (OSError, ValueError):
        return False, "unreadable manifest"
    checks = (
        (old.get("icedir") == man["icedir"], "ice directory changed"),
        (old.get("prefix") == man["prefix"], "prefix changed"),
        (int(old.get("nfiles", -1)) >= man["nfiles"], "file count changed"),
        ([list(e) for e in old.get("files", [])[: man["nfiles"]]] == man["files"], "ice file size or mtime changed"),
    )
    for ok, why in checks:
        if not ok:
            return False, why
    return True, ""


def build_ice_cache(icedir, pfx, nmax, nlat=NLAT, nlon=NLON, cache_dir="cache", rebuild=False, log=print):
    """ice_<pfx>0_<n>.npy in cache_dir holding files 0..nmax: an existing cache of n >= nmax files
    whose manifest matches is reused; otherwise one is written (to a temporary name, then renamed)
    with its manifest.  Returns (path, seconds spent building; 0 when reused)."""
    import time

    os.makedirs(cache_dir, exist_ok=True)
    stem = f"ice_{pfx}0_"
    path = os.path.join(cache_dir, f"{stem}{nmax}.npy")
    man = ice_manifest(icedir, pfx, nmax + 1)
    if not rebuild:
        reason, found = "no cache", False
        for f in sorted(os.listdir(cache_dir)):
            m = re.fullmatch(re.escape(stem) + r"(\d+)\.npy", f)
            if not m or int(m.group(1)) < nmax:
                continue
            found = True
            ok, why = cache_matches(os.path.join(cache_dir, f), man)
            if ok:
                return os.path.join(cache_dir, f), 0.0
            reason = why
        if found:
            log(f"ice cache out of date ({reason}): rebuilding {path}")
    t0 = time.perf_counter()
    tmp = path + ".tmp"
    stack = np.lib.format.open_memmap(tmp, mode="w+", dtype=np.float64, shape=(nmax + 1, nlat, nlon))
    for k in range(nmax + 1):
        stack[k] = find_grid(icedir, f"{pfx}{k}", nlat, nlon)
    stack.flush()
    del stack
    os.replace(tmp, path)
    write_manifest(path, man)
    return path, time.perf_counter() - t0


# ------------------------------------------------------------------------- converters
def text2npy(files, nlat=NLAT, nlon=NLON, f4=False, outdir=None):
    """Text grids -> <outdir or the input's directory>/<name>.npy; returns [(path, dtype)]."""
    done = []
    for f in files:
        out = os.path.join(outdir or os.path.dirname(f), os.path.basename(f) + ".npy")
        done.append((out, save_grid(out, read_text_grid(f, nlat, nlon), "f4" if f4 else "f8")))
    return done


def npy2text(files, fmt="es16.9", ncol=None, outdir=None):
    """'.npy' grids -> text files named without the '.npy' suffix; returns the paths."""
    done = []
    for f in files:
        base = os.path.basename(f)
        out = os.path.join(outdir or os.path.dirname(f), base[:-4] if base.endswith(".npy") else base)
        write_text_grid(out, np.load(f, allow_pickle=False), fmt, ncol)
        done.append(out)
    return done


def snapshot_steps(rundir, field):
    """Sorted step numbers N of <rundir>/<field><N>.npy."""
    pat = re.compile(re.escape(field) + r"(\d+)\.npy")
    return sorted(int(m.group(1)) for f in os.listdir(rundir) if (m := pat.fullmatch(f)))


def gl_lat_deg(nlat):
    """Gauss-Legendre latitudes, degrees north, rows north -> south."""
    x, _ = np.polynomial.legendre.leggauss(nlat)
    return np.degrees(np.arcsin(x[::-1]))


def npy2nc(rundir, out, fields=("tgrid",), times=None, dt=1.0, lon_convention="solver", earth=None,
           solver=None, attrs=None, planet=None):
    """Per-step snapshots <field><N>.npy of a run directory -> one NetCDF file with dimensions
    (time, lat, lon) and the coordinates, units and global attributes of npy2nc.py: lat from the
    Gauss-Legendre nodes, lon = (i + off) 360/nlon (off 1 for lon_convention "sampled"), time from
    the times file (``times`` or <rundir>/times, row = step) or step * dt."""
    from scipy.io import netcdf_file

    fields = list(fields)
    steps = snapshot_steps(rundir, fields[0])
    if not steps:
        raise FileNotFoundError(f"no {fields[0]}<N>.npy in {rundir}")
    for f in fields[1:]:
        if snapshot_steps(rundir, f) != steps:
            raise ValueError(f"field {f} has other steps than {fields[0]}")
    nlat, nlon = np.load(os.path.join(rundir, f"{fields[0]}{steps[0]}.npy"), mmap_mode="r").shape
    tfile = times or (os.path.join(rundir, "times") if os.path.exists(os.path.join(rundir, "times")) else None)
    if tfile:
        tval = np.loadtxt(tfile, ndmin=1)[steps]
        tsrc = f"file {os.path.abspath(tfile)} (row = step)"
    else:
        tval = np.asarray(steps, dtype=np.float64) * dt
        tsrc = f"step * {dt} yr"
    if planet is None:
        from .solver import Planet

        planet = Planet()
    nc = netcdf_file(out, "w", version=2)
    try:
        nc.createDimension("time", len(steps))
        nc.createDimension("lat", nlat)
        nc.createDimension("lon", nlon)
        v = nc.createVariable("time", "f8", ("time",))
        v[:] = tval
        v.units, v.long_name, v.source = "years since start of simulation", "model time", tsrc
        v = nc.createVariable("step", "i4", ("time",))
        v[:] = np.asarray(steps, dtype=np.int32)
        v.long_name = "solver time-step index (file suffix N)"
        v = nc.createVariable("lat", "f8", ("lat",))
        v[:] = gl_lat_deg(nlat)
        v.units, v.long_name = "degrees_north", "Gauss-Legendre latitude (row j of the grid, north to south)"
        off = 0 if lon_convention == "solver" else 1
        v = nc.createVariable("lon", "f8", ("lon",))
        v[:] = (np.arange(nlon) + off) * 360.0 / nlon
        v.units, v.long_name = "degrees_east", "longitude (column i of the grid)"
        v.convention = ("i*360/nlon (solver SHT grid)" if off == 0
                        else "(i+1)*360/nlon (one-cell-shifted ice-sampling convention)")
        for f in fields:
            var = nc.createVariable(f, "f8", ("time", "lat", "lon"))
            var.units, var.long_name = FIELD_META.get(f, ("1", f))
            for k, s in enumerate(steps):
                var[k] = np.load(os.path.join(rundir, f"{f}{s}.npy"), allow_pickle=False).astype(np.float64)
        g = {
            "title": "sea-level model snapshots", "Conventions": "CF-1.8",
            "solver": solver or "slevolve_amd", "earth_model": earth or "unknown",
            "grid": f"Gauss-Legendre {nlat} x {nlon}",
            "planet_radius_m": planet.radius, "planet_mass_kg": planet.mass,
            "rho_ice_kg_m3": planet.rhoi, "rho_water_kg_m3": planet.rhow, "g_m_s2": planet.gacc,
            "omega_rad_s": planet.omega, "moi_A_kg_m2": planet.moiA, "moi_C_kg_m2": planet.moiC,
            "fluid_love_kf": planet.kf, "source_directory": os.path.abspath(rundir),
            "history": f"{datetime.datetime.now().isoformat(timespec='seconds')} npy2nc {os.path.abspath(rundir)}",
        }
        g.update(attrs or {})
        for k, val in g.items():
            setattr(nc, k, val if not isinstance(val, float) or math.isfinite(val) else str(val))
    finally:
        nc.close()
    return out
