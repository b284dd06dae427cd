"""Oracle Love-number reader.  Test infrastructure only (see ``oracle/__init__.py``).

Restates ``LoveNumbers`` (``python/slevolve/solver.py:114-225``): Fortran list-directed reads
(every READ starts a new record, surplus items of the last record are dropped), per degree
l = 1..norder the groups ``l K`` / ``s_1..s_K`` / elastic line / fluid line / h, l, k residues
and, for l >= 2, the tidal elastic / fluid lines and residues.  l- and k-type numbers are stored
times the degree in the file and divided by l here.  The tangential (l-type) numbers feed only the
horizontal read-out.
"""
from __future__ import annotations

import numpy as np

NPAM = 500


class _Records:
    def __init__(self, text: str):
        self.lines = text.split("\n")
        self.i = 0

    def skip(self):
        self.i += 1

    def take(self, n):
        vals = []
        while len(vals) < n:
            vals += self.lines[self.i].split()
            self.i += 1
        return vals[:n]

    def floats(self, n):
        return np.array([float(v) for v in self.take(n)])


class LoveNumbers:
    def __init__(self, path_or_text, norder: int, text: bool = False):
        if text:
            body = path_or_text
        else:
            with open(path_or_text) as fh:
                body = fh.read()
        rd = _Records(body)
        rd.skip()
        nl = norder + 1
        self.norder = norder
        self.nmod = np.zeros(nl, dtype=np.int64)
        self.s = np.zeros((NPAM, nl))
        el, tel = np.zeros((3, nl)), np.zeros((3, nl))
        res = {k: np.zeros((NPAM, nl)) for k in ("h", "l", "k", "th", "tl", "tk")}
        for _ in range(norder):
            hdr = rd.take(2)
            l, K = int(hdr[0]), int(hdr[1])
            self.nmod[l] = K
            self.s[:K, l] = rd.floats(K)
            v = rd.floats(5)
            el[:, l] = v[0], v[1], v[4]
            rd.floats(5)  # fluid limit, unused by the step
            for key in ("h", "l", "k"):
                res[key][:K, l] = rd.floats(K)
            if l < 2:
                continue
            v = rd.floats(5)
            tel[:, l] = v[0], v[1], v[4]
            rd.floats(5)
            for key in ("th", "tl", "tk"):
                res[key][:K, l] = rd.floats(K)
        deg = np.arange(1, nl, dtype=np.float64)
        for key in ("l", "k", "tl", "tk"):
            res[key][:, 1:] /= deg
        self.he = np.zeros(nl)
        self.ke = np.zeros(nl)
        self.le = np.zeros(nl)
        self.hTE = np.zeros(nl)
        self.kTE = np.zeros(nl)
        self.lTE = np.zeros(nl)
        self.he[1:] = el[0, 1:]
        self.ke[1:] = el[2, 1:] / deg
        self.le[1:] = el[1, 1:] / deg  # tangential numbers (solver.py:202-208)
        self.hTE[1:] = tel[0, 1:]
        self.kTE[1:] = tel[2, 1:] / deg
        self.lTE[1:] = tel[1, 1:] / deg
        self.r = res["h"]
        self.rprime = res["k"]
        self.rl = res["l"]
        self.rT = res["th"]
        self.rprimeT = res["tk"]
        self.rlT = res["tl"]
        self.max_modes = int(self.nmod.max())
