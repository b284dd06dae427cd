"""Oracle Anderson mixing.  Test infrastructure only (see ``oracle/__init__.py``).

Restates ``_Anderson`` (``python/slevolve/solver_opt.py:67-145``): type-II Anderson mixing
(Walker & Ni 2011) on a complex spectral array, with the real inner product over its real and
imaginary parts.  The last depth + 1 pairs (input x_i, output g_i = G(x_i)) are kept with
residuals f_i = g_i - x_i.  ``mix`` forms dF_j = f_j - f_{j+1}, dG_j = g_j - g_{j+1} (newest
first), solves min || f_0 - dF gamma || through the normal equations with a least-squares solve
(rcond 1e-12) and returns x_next = g_0 - sum_j gamma_j dG_j.  With fewer than two pairs the plain
step g_0 is returned ("plain"); if the newest residual norm grew, the history restarts from the
newest pair and the plain step is returned ("fallback"); a non-finite gamma gives "singular".
"""
from __future__ import annotations

import math

import numpy as np


def real_dot(a, b):
    return float(np.dot(np.ascontiguousarray(a).view(np.float64).ravel(),
                        np.ascontiguousarray(b).view(np.float64).ravel()))


class Anderson:
    def __init__(self, depth):
        self.depth = int(depth)
        self.reset()

    def reset(self):
        self.pairs = []  # (x, g, f, |f|), oldest first

    def push(self, x, g):
        f = g - x
        self.pairs.append((x.copy(), g.copy(), f, math.sqrt(real_dot(f, f))))
        del self.pairs[: -(self.depth + 1)]

    def mix(self):
        """(next input, kind) with kind in mixed | plain | fallback | singular."""
        x0, g0, f0, n0 = self.pairs[-1]
        if len(self.pairs) >= 2 and n0 > self.pairs[-2][3]:
            self.reset()
            self.push(x0, g0)
            return g0.copy(), "fallback"
        if len(self.pairs) < 2:
            return g0.copy(), "plain"
        newest_first = self.pairs[::-1]
        m = len(newest_first) - 1
        dF = [newest_first[j][2] - newest_first[j + 1][2] for j in range(m)]
        dG = [newest_first[j][1] - newest_first[j + 1][1] for j in range(m)]
        A = np.empty((m, m))
        rhs = np.empty(m)
        for i in range(m):
            rhs[i] = real_dot(dF[i], f0)
            for j in range(i, m):
                A[i, j] = A[j, i] = real_dot(dF[i], dF[j])
        try:
            gam = np.linalg.lstsq(A, rhs, rcond=1e-12)[0]
        except np.linalg.LinAlgError:
            return g0.copy(), "singular"
        if not np.all(np.isfinite(gam)):
            return g0.copy(), "singular"
        out = g0.copy()
        for j in range(m):
            out -= gam[j] * dG[j]
        return out, "mixed"
