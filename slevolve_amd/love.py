"""Love-number file reader (host), the layout of the reference's LoveNumbers
(python/slevolve/solver.py:114-225; file layout python/slevolve/tools/write_lovenum.py:9-33)."""
from __future__ import annotations

import numpy as np

NPAM = 500


class _Records:
    """Fortran list-directed reads: every read starts a new record; surplus items are dropped."""

    def __init__(self, text):
        self.lines = text.split("\n")
        self.pos = 0

    def skip(self):
        self.pos += 1

    def take(self, n):
        out = []
        while len(out) < n:
            if self.pos >= len(self.lines):
                raise ValueError("Love-number file ended early")
            out.extend(self.lines[self.pos].split())
            self.pos += 1
        return out[:n]

    def floats(self, n):
        return np.array([float(v) for v in self.take(n)], dtype=np.float64)


class LoveNumbers:
    """Degrees 1..norder of a Love-number file.  ``path`` may be a file name or, with
    ``text=True``, the file's contents."""

    def __init__(self, path, norder, text=False):
        if text:
            body = path
        else:
            with open(path) as fh:
                body = fh.read()
        nl = norder + 1
        rd = _Records(body)
        rd.skip()
        self.norder = norder
        self.nmod = np.zeros(nl, dtype=np.int64)
        self.s = np.zeros((NPAM, nl))
        el, fl, tel, tfl = (np.zeros((3, nl)) for _ in range(4))
        res = {k: np.zeros((NPAM, nl)) for k in ("h", "l", "k", "th", "tl", "tk")}
        for _ in range(norder):
            l, K = (int(v) for v in rd.take(2))
            if not 1 <= l <= norder or K > NPAM:
                raise ValueError(f"Love-number file: bad degree/mode count {l} {K}")
            self.nmod[l] = K
            self.s[:K, l] = rd.floats(K)
            v = rd.floats(5)
            el[:, l] = v[0], v[1], v[4]
            v = rd.floats(5)
            fl[:, l] = v[0], v[1], v[4]
            res["h"][:K, l] = rd.floats(K)
            res["l"][:K, l] = rd.floats(K)
            res["k"][:K, l] = rd.floats(K)
            if l >= 2:
                v = rd.floats(5)
                tel[:, l] = v[0], v[1], v[4]
                v = rd.floats(5)
                tfl[:, l] = v[0], v[1], v[4]
                res["th"][:K, l] = rd.floats(K)
                res["tl"][:K, l] = rd.floats(K)
                res["tk"][:K, l] = rd.floats(K)
        deg = np.arange(1, nl, dtype=np.float64)
        for key in ("l", "k", "tl", "tk"):
            res[key][:, 1:] /= deg
        z = lambda: np.zeros(nl)  # noqa: E731
        self.he, self.ke, self.le = z(), z(), z()
        self.hTE, self.kTE, self.lTE = z(), z(), z()
        self.h_fluid, self.l_fluid, self.k_fluid = z(), z(), z()
        self.hT_fluid, self.lT_fluid, self.kT_fluid = z(), z(), z()
        self.he[1:], self.le[1:], self.ke[1:] = el[0, 1:], el[1, 1:] / deg, el[2, 1:] / deg
        self.hTE[1:], self.lTE[1:], self.kTE[1:] = tel[0, 1:], tel[1, 1:] / deg, tel[2, 1:] / deg
        self.h_fluid[1:], self.l_fluid[1:], self.k_fluid[1:] = fl[0, 1:], fl[1, 1:] / deg, fl[2, 1:] / deg
        self.hT_fluid[1:], self.lT_fluid[1:], self.kT_fluid[1:] = tfl[0, 1:], tfl[1, 1:] / deg, tfl[2, 1:] / deg
        self.r, self.rl, self.rprime = res["h"], res["l"], res["k"]
        self.rT, self.rlT, self.rprimeT = res["th"], res["tl"], res["tk"]
        self.max_modes = int(self.nmod.max())
