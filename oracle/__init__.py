"""CPU FP64 oracle of the SLEvolve sea-level time step -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy/SciPy FP64, the algorithm of the reference's Python
solver (``python/slevolve/sht.py``, ``solver.py``, ``solver_opt.py``) so that the HIP path of
``slevolve_amd`` can be checked against it on identical inputs.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it, and
only as the checker / CPU baseline: the product path never calls into it.

Pins (what checks each function; see tests/test_oracle_*.py):

* ``sht.gauss_legendre``  -- ``numpy.polynomial.legendre.leggauss`` nodes and weights.
* ``sht.legendre_block``  -- closed forms P~_00 = 1/sqrt(2), P~_11 = sqrt(3)/2 sin(theta),
  P~_20, P~_22; orthonormality by Gauss quadrature (all l, m to lmax 63); brute-force
  comparison with ``scipy.special.lpmv`` scaled to the orthonormal, no-Condon-Shortley form.
* ``sht.OracleSHT``       -- analysis of a constant field c gives U_00 = c sqrt(2)/1.41421356;
  analysis -> synthesis round trip < 1e-12 (band-limited fields, lmax 63 and 511 on 512 x 1024);
  a single e^{i m phi} P~_lm mode is analysed to the unit coefficient.
* ``memory.ExactMemory``  -- explicit history sum over committed increments (non-uniform
  steps, relative error < 1e-12); decay factors equal exp(-s dt / 1000).
* ``solver.extrap_guess`` -- closed form of the non-uniform extrapolation and exact 2h1 - h2.
* ``solver.SeaLevelSolver`` -- the reference's golden fixture ``reference/synthetic/tgrid8.npy``
  (512 x 1024, lmax 511, 8 steps, extrapolated start) to 1e-9 m with numiter 5,5,6,6,5,5,5,5;
  the quickstart known answers (128 x 256, lmax 63, numiter [6,4,6,7,7,7,7,7], printed RSL
  values); the explicit-history variant of the whole step to 1e-9 m.
* ``love.LoveNumbers`` -- write -> read round trip of the file layout, and the PREM file's
  values enter every whole-step pin above.
"""
