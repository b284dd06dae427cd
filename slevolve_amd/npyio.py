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
    """True when float32 holds every value of ``a`` exactly, so that an '<f4' file loses nothing
    (values outside the float32 range, and NaN, do not qualify)."""
    x = np.asarray(a, dtype=np.float64)
    with np.errstate(over="ignore", invalid="ignore"):
        return bool((x.astype(np.float32) == x).all())


def save_grid(path, a, dtype="f8"):
    """Write a grid as '<f8', or as '<f4' when asked for and exact; returns the dtype written."""
    g = np.ascontiguousarray(_host(a), dtype=np.float64)
    kind = "f4" if dtype == "f4" and f4_exact(g) else "f8"
    np.save(path, g.astype("<" + kind, copy=False))
    return kind


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
class GridFileError(ValueError):
    """A grid or times file whose content does not fit what it is read for."""


def _check_count(path, count, nlat, nlon, unit):
    if count != nlat * nlon:
        raise GridFileError(f"{path} holds {count} {unit}; a {nlat} x {nlon} grid needs {nlat * nlon}")


def read_text_grid(path, nlat=NLAT, nlon=NLON):
    """A text grid: whitespace-separated values in stream order, any number per line."""
    with open(path) as fh:
        vals = np.array(fh.read().split(), dtype=np.float64)
    _check_count(path, vals.size, nlat, nlon, "values")
    return vals.reshape(nlat, nlon)


def load_npy(path, mmap=False):
    return np.load(path, mmap_mode="r" if mmap else None, allow_pickle=False)


def _read_npy_grid(path, nlat, nlon):
    arr = np.load(path, allow_pickle=False)
    _check_count(path, arr.size, nlat, nlon, "values")
    if arr.shape != (nlat, nlon):  # a flat or otherwise shaped payload, taken in stream order
        arr = arr.reshape(nlat, nlon)
    return np.ascontiguousarray(arr, dtype=np.float64)


def _read_bin_grid(path, nlat, nlon):
    nbytes = os.path.getsize(path)
    if nbytes % 8:
        raise GridFileError(f"{path}: {nbytes} bytes is not a whole number of little-endian float64 values")
    _check_count(path, nbytes // 8, nlat, nlon, "float64 values")
    return np.fromfile(path, dtype="<f8").reshape(nlat, nlon)


# the encodings a grid name may have on disk, in the order a reader prefers them
_GRID_ENCODINGS = ((".npy", _read_npy_grid), (".bin", _read_bin_grid), ("", read_text_grid))


def grid_source(dirname, name):
    """Path of the file grid ``name`` in ``dirname`` is read from: <name>.npy, else <name>.bin, else
    the text file <name>; None when there is none of them."""
    base = os.path.join(dirname, name)
    return next((base + suf for suf, _ in _GRID_ENCODINGS if os.path.isfile(base + suf)), None)


def read_grid(path, nlat=NLAT, nlon=NLON):
    """A float64 (nlat, nlon) grid from '.npy' (any real dtype and memory order), a headerless
    little-endian '.bin' of float64 values, or text.  A path with neither suffix names a grid whose
    preferred encoding is read (see grid_source)."""
    for suf, reader in _GRID_ENCODINGS[:2]:
        if path.endswith(suf):
            return reader(path, nlat, nlon)
    src = grid_source(os.path.dirname(path) or ".", os.path.basename(path)) or path
    reader = next(r for suf, r in _GRID_ENCODINGS if src.endswith(suf))
    return reader(src, nlat, nlon)


def find_grid(dirname, name, nlat=NLAT, nlon=NLON):
    return read_grid(os.path.join(dirname, name), nlat, nlon)


class IceDir:
    """The ice history of a directory as a lazily read sequence: ``ice[k]`` is grid <pfx><k>."""

    def __init__(self, d, pfx, nlat=NLAT, nlon=NLON):
        self.directory, self.prefix, self.shape = d, pfx, (nlat, nlon)

    def __getitem__(self, k):
        return find_grid(self.directory, f"{self.prefix}{int(k)}", *self.shape)


def read_times(path, nsteps=None):
    """The times t_0, t_1, ... (years) of the ice files, one per line (the first number of a line;
    blank lines are skipped).  They must be finite and strictly increasing, at least two of them;
    with ``nsteps`` only t_0..t_nsteps are returned and fewer is an error.  Step j lasts
    t_j - t_{j-1}."""
    with open(path) as fh:
        t = np.array([float(line.split()[0]) for line in fh if line.strip()], dtype=np.float64)
    if t.size < 2:
        raise GridFileError(f"{path}: a times file lists at least two times, found {t.size}")
    if nsteps is not None:
        if t.size <= nsteps:
            raise GridFileError(f"{path}: {nsteps} steps need {nsteps + 1} times, the file lists {t.size}")
        t = t[: nsteps + 1]
    if not np.isfinite(t).all():
        raise GridFileError(f"{path}: non-finite time")
    if (np.diff(t) <= 0).any():
        raise GridFileError(f"{path}: times do not increase strictly")
    return t


# --------------------------------------------------------------------------- ice cache
# A cache stacks ice files 0..n-1 as one (n, nlat, nlon) '<f8' array; its JSON sidecar <cache>.json
# records where the stack came from: {"icedir": absolute directory, "prefix": file prefix,
# "nfiles": n, "files": [[file name, size in bytes, mtime in ns], ...] of the file a reader takes
# for each k} (the reference's schema, so that caches are shared with it).
def _file_identity(path):
    st = os.stat(path)
    return [os.path.basename(path), int(st.st_size), int(st.st_mtime_ns)]


def ice_manifest(dirname, pfx, nfiles):
    """The sidecar record of a cache built now from ice files 0..nfiles-1 of ``dirname``."""
    ident = []
    for k in range(nfiles):
        src = grid_source(dirname, f"{pfx}{k}")
        if src is None:
            raise FileNotFoundError(f"ice file {pfx}{k} (.npy, .bin or text) not found in {dirname}")
        ident.append(_file_identity(src))
    return {"icedir": os.path.abspath(dirname), "prefix": pfx, "nfiles": int(nfiles), "files": ident}


def manifest_path(cache_path):
    return f"{cache_path}.json"


def write_manifest(cache_path, man):
    with open(manifest_path(cache_path), "w") as fh:
        fh.write(json.dumps(man, indent=1))


def _stale(cache_path, man):
    """Why the cache at ``cache_path`` cannot stand for the sources ``man`` describes, or None.  A
    cache of more files serves when its leading files are the same."""
    if not os.path.isfile(cache_path):
        return "cache file missing"
    try:
        with open(manifest_path(cache_path)) as fh:
            rec = json.load(fh)
    except FileNotFoundError:
        return "sidecar missing"
    except (OSError, ValueError):
        return "sidecar not readable as JSON"
    if rec.get("icedir") != man["icedir"]:
        return "built from another ice directory"
    if rec.get("prefix") != man["prefix"]:
        return "built for another file prefix"
    if int(rec.get("nfiles", -1)) < man["nfiles"]:
        return "holds fewer ice files than needed"
    have = [list(f) for f in rec.get("files", ())][: man["nfiles"]]
    if have != man["files"]:
        return "an ice file's size or modification time differs"
    return None


def cache_matches(cache_path, man):
    """(True, "") when the cache can stand for the sources ``man`` describes, else (False, why)."""
    why = _stale(cache_path, man)
    return (why is None, why or "")


def _stacks(cache_dir, stem):
    """[(n, path)] of the cached stacks <stem><n>.npy in cache_dir, largest n first."""
    found = []
    for f in os.listdir(cache_dir):
        if f.startswith(stem) and f.endswith(".npy") and f[len(stem):-4].isdigit():
            found.append((int(f[len(stem):-4]), os.path.join(cache_dir, f)))
    return sorted(found, reverse=True)


def build_ice_cache(icedir, pfx, nmax, nlat=NLAT, nlon=NLON, cache_dir="cache", rebuild=False, log=print):
    """The stack of ice files 0..nmax as cache_dir/ice_<pfx>0_<n>.npy (the reference's naming): a
    cached stack of at least nmax + 1 files whose sidecar matches the current files is reused;
    otherwise the stack is written to a temporary file that replaces ice_<pfx>0_<nmax>.npy once
    complete, with its sidecar.  Returns (path, seconds spent writing; 0.0 when reused)."""
    import tempfile
    import time

    os.makedirs(cache_dir, exist_ok=True)
    stem = f"ice_{pfx}0_"
    man = ice_manifest(icedir, pfx, nmax + 1)
    if not rebuild:
        reasons = []
        for n, cand in _stacks(cache_dir, stem):
            if n < nmax:
                break
            why = _stale(cand, man)
            if why is None:
                return cand, 0.0
            reasons.append(f"{os.path.basename(cand)}: {why}")
        if reasons:
            log(f"[slevolve_amd] cached ice stacks not usable ({'; '.join(reasons)}); writing a new one")
    t0 = time.perf_counter()
    target = os.path.join(cache_dir, f"{stem}{nmax}.npy")
    fd, part = tempfile.mkstemp(suffix=".part", dir=cache_dir)
    os.close(fd)
    try:
        out = np.lib.format.open_memmap(part, mode="w+", dtype="<f8", shape=(nmax + 1, nlat, nlon))
        src = IceDir(icedir, pfx, nlat, nlon)
        for k in range(nmax + 1):
            out[k] = src[k]
        out.flush()
        del out
        os.replace(part, target)
    except BaseException:
        if os.path.exists(part):
            os.remove(part)
        raise
    write_manifest(target, man)
    return target, time.perf_counter() - t0


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
