/* Multi-threaded CPU kernels of the oracle's optimised CPU baseline (test infrastructure only:
 * timed beside the GPU path, never part of the product).  Restates the same algorithms as
 * oracle/sht.py and oracle/memory.py:
 *
 *   analysis   U[l, m] = sum_j P~_lm(mu_j) (w_j F_j,m / nlon) / 1.41421356, hemispheres folded
 *              into S = north + south (l - m even) and D = north - south (l - m odd)
 *              (python/slevolve/sht.py:274-302);
 *   synthesis  F_j,m = sum_l 1.41421356 nlon U[l, m] P~_lm(mu_j), north = E + O, south = E - O
 *              (python/slevolve/sht.py:305-343);
 *   P~_lm      regenerated per (order, ring block) by the three-term recurrence
 *              P~_n = (x P~_{n-1} - eps_{n-1} P~_{n-2}) / eps_n, eps_n = sqrt((n^2 - m^2)/(4 n^2 - 1))
 *              (sht.py:105-144), with the oracle's extended-range start values (scale count s:
 *              value P~ 2^(600 s), renormalised past 2^300, contributing nothing while s > 0);
 *   modal      R_k <- R_k d + P (1 - d), visc += R_k a_sl, rad += R_k a_rr over the dense lower
 *              triangle (python/slevolve/solver.py:386-422).
 *
 * Orders are dealt to OpenMP threads in heavy/light pairs (m, L - m); rings are processed in
 * blocks of RB so that the recurrence vectorises across rings.  The per-ring recurrence is the
 * oracle's; sums over rings run in a different order than NumPy's products, so results agree
 * with the oracle to rounding, not bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define RB 8
#define SCALE_K 600
#define SCALE_T 300

static inline int order_of(int idx, int L) { return (idx & 1) ? L - idx / 2 : idx / 2; }

/* eps_n for n = m .. L into e[n - m] (e[0] unused) and their reciprocals */
static void eps_table(int m, int L, double* e, double* ie) {
  for (int n = m + 1; n <= L; ++n) {
    const double v = sqrt((double)((long)n * n - (long)m * m) / (double)(4L * n * n - 1));
    e[n - m] = v;
    ie[n - m] = 1.0 / v;
  }
}

/* Walks the recurrence of order m for the rings j0 .. j0 + nr - 1 and calls BODY(i, p) for every
 * degree index i = l - m with the values p[r] (0 where the ring's scale count is still > 0). */
#define WALK_ORDER(BODY)                                                                          \
  do {                                                                                            \
    double p0[RB], p1[RB], p2[RB], xr[RB], v[RB];                                                 \
    int sc[RB], any = 0;                                                                          \
    for (int r = 0; r < RB; ++r) {                                                                \
      const int j = j0 + (r < nr ? r : 0);                                                        \
      xr[r] = x[j];                                                                               \
      p0[r] = r < nr ? pmm[(long)m * nh + j] : 0.0;                                               \
      sc[r] = r < nr ? pms[(long)m * nh + j] : 1;                                                 \
      any |= sc[r] > 0;                                                                           \
    }                                                                                             \
    for (int r = 0; r < RB; ++r) v[r] = sc[r] == 0 ? p0[r] : 0.0;                                 \
    BODY(0, v);                                                                                   \
    if (nl < 2) break;                                                                            \
    const double s23 = sqrt((double)(2 * m + 3));                                                 \
    for (int r = 0; r < RB; ++r) {                                                                \
      p1[r] = s23 * xr[r] * p0[r];                                                                \
      v[r] = sc[r] == 0 ? p1[r] : 0.0;                                                            \
    }                                                                                             \
    BODY(1, v);                                                                                   \
    for (int i = 2; i < nl; ++i) {                                                                \
      const double ep = e[i - 1], ie_ = ie[i];                                                    \
      for (int r = 0; r < RB; ++r) p2[r] = (xr[r] * p1[r] - ep * p0[r]) * ie_;                    \
      if (any) {                                                                                  \
        any = 0;                                                                                  \
        for (int r = 0; r < RB; ++r) {                                                            \
          if (sc[r] > 0 && fabs(p2[r]) >= 0x1p300) {                                              \
            p2[r] = ldexp(p2[r], -SCALE_K);                                                       \
            p1[r] = ldexp(p1[r], -SCALE_K);                                                       \
            sc[r] -= 1;                                                                           \
          }                                                                                       \
          any |= sc[r] > 0;                                                                       \
          v[r] = sc[r] == 0 ? p2[r] : 0.0;                                                        \
        }                                                                                         \
      } else {                                                                                    \
        for (int r = 0; r < RB; ++r) v[r] = p2[r];                                                \
      }                                                                                           \
      BODY(i, v);                                                                                 \
      for (int r = 0; r < RB; ++r) {                                                              \
        p0[r] = p1[r];                                                                            \
        p1[r] = p2[r];                                                                            \
      }                                                                                           \
    }                                                                                             \
  } while (0)

/* S, D: [nm][nh] complex (interleaved), weights applied; U: dense [L+1][L+1] complex, (l, m) at
 * l (L+1) + m; written for l >= m. */
void cpu_legendre_analysis(int L, int nh, const double* x, const double* pmm, const int* pms, const double* S,
                           const double* D, double* U, double scale) {
#pragma omp parallel
  {
    double e[L + 2], ie[L + 2];
    /* per-ring partial sums [l - m][r] (re, im): the sum over a ring block's lanes is taken once per
       order, after all blocks, so the degree loop has no cross-lane reduction */
    double* acc = (double*)malloc(sizeof(double) * 2 * RB * (L + 1));
#pragma omp for schedule(dynamic, 1)
    for (int idx = 0; idx <= L; ++idx) {
      const int m = order_of(idx, L), nl = L + 1 - m;
      eps_table(m, L, e, ie);
      memset(acc, 0, sizeof(double) * 2 * RB * nl);
      for (int j0 = 0; j0 < nh; j0 += RB) {
        const int nr = nh - j0 < RB ? nh - j0 : RB;
        double sr[RB], si[RB], dr[RB], di[RB];
        for (int r = 0; r < RB; ++r) {
          const long q = 2 * ((long)m * nh + j0 + (r < nr ? r : 0));
          const double keep = r < nr ? 1.0 : 0.0;
          sr[r] = keep * S[q];
          si[r] = keep * S[q + 1];
          dr[r] = keep * D[q];
          di[r] = keep * D[q + 1];
        }
#define ANA_BODY(i, v)                                              \
  do {                                                              \
    double* ar = acc + 2 * RB * (i);                                \
    if (((i) & 1) == 0) {                                           \
      for (int r = 0; r < RB; ++r) {                                \
        ar[r] += v[r] * sr[r];                                      \
        ar[RB + r] += v[r] * si[r];                                 \
      }                                                             \
    } else {                                                        \
      for (int r = 0; r < RB; ++r) {                                \
        ar[r] += v[r] * dr[r];                                      \
        ar[RB + r] += v[r] * di[r];                                 \
      }                                                             \
    }                                                               \
  } while (0)
        WALK_ORDER(ANA_BODY);
#undef ANA_BODY
      }
      for (int i = 0; i < nl; ++i) {
        double re = 0.0, im = 0.0;
        for (int r = 0; r < RB; ++r) {
          re += acc[2 * RB * i + r];
          im += acc[2 * RB * i + RB + r];
        }
        const long q = 2 * ((long)(m + i) * (L + 1) + m);
        U[q] = re * scale;
        U[q + 1] = im * scale;
      }
    }
    free(acc);
  }
}

/* U: dense [L+1][L+1] complex, already multiplied by the synthesis factor; FN, FS: [nh][nlon/2+1]
 * complex rows (north rows j, south rows nlat-1-j), orders 0..L written, the rest left as given. */
void cpu_legendre_synthesis(int L, int nh, int nf, const double* x, const double* pmm, const int* pms,
                            const double* U, double* FN, double* FS) {
#pragma omp parallel
  {
    double e[L + 2], ie[L + 2], c[2 * (L + 1)];
#pragma omp for schedule(dynamic, 1)
    for (int idx = 0; idx <= L; ++idx) {
      const int m = order_of(idx, L), nl = L + 1 - m;
      eps_table(m, L, e, ie);
      for (int i = 0; i < nl; ++i) {
        const long q = 2 * ((long)(m + i) * (L + 1) + m);
        c[2 * i] = U[q];
        c[2 * i + 1] = U[q + 1];
      }
      for (int j0 = 0; j0 < nh; j0 += RB) {
        const int nr = nh - j0 < RB ? nh - j0 : RB;
        double er[RB] = {0}, ei[RB] = {0}, or_[RB] = {0}, oi[RB] = {0};
#define SYN_BODY(i, v)                                \
  do {                                                \
    const double cr = c[2 * (i)], ci = c[2 * (i) + 1]; \
    if (((i) & 1) == 0) {                             \
      for (int r = 0; r < RB; ++r) {                  \
        er[r] += v[r] * cr;                           \
        ei[r] += v[r] * ci;                           \
      }                                               \
    } else {                                          \
      for (int r = 0; r < RB; ++r) {                  \
        or_[r] += v[r] * cr;                          \
        oi[r] += v[r] * ci;                           \
      }                                               \
    }                                                 \
  } while (0)
        WALK_ORDER(SYN_BODY);
#undef SYN_BODY
        for (int r = 0; r < nr; ++r) {
          const long qn = 2 * ((long)(j0 + r) * nf + m);
          FN[qn] = er[r] + or_[r];
          FN[qn + 1] = ei[r] + oi[r];
          FS[qn] = er[r] - or_[r];
          FS[qn + 1] = ei[r] - oi[r];
        }
      }
    }
  }
}

/* Dense modal advance-and-read: R [K][nl][nl] complex, P [nl][nl] complex, decay/a_sl/a_rr [K][nl];
 * lower triangle m <= l only.  Same operations per element as oracle/memory.py. */
void cpu_modal_advance(int K, int nl, double* R, const double* P, const double* decay, const double* a_sl,
                       const double* a_rr, double* visc, double* rad) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int l = 0; l < nl; ++l) {
    const long row = (long)l * nl;
    for (int m = 0; m <= l; ++m) {
      visc[2 * (row + m)] = visc[2 * (row + m) + 1] = 0.0;
      rad[2 * (row + m)] = rad[2 * (row + m) + 1] = 0.0;
    }
    for (int k = 0; k < K; ++k) {
      const double d = decay[(long)k * nl + l], o = 1.0 - d, s = a_sl[(long)k * nl + l], q = a_rr[(long)k * nl + l];
      double* Rk = R + 2 * ((long)k * nl * nl + row);
      for (int m = 0; m <= l; ++m) {
        const double xr = Rk[2 * m] * d + P[2 * (row + m)] * o;
        const double xi = Rk[2 * m + 1] * d + P[2 * (row + m) + 1] * o;
        Rk[2 * m] = xr;
        Rk[2 * m + 1] = xi;
        visc[2 * (row + m)] += xr * s;
        visc[2 * (row + m) + 1] += xi * s;
        rad[2 * (row + m)] += xr * q;
        rad[2 * (row + m) + 1] += xi * q;
      }
    }
  }
}

int cpu_threads(void);
#ifdef _OPENMP
#include <omp.h>
int cpu_threads(void) { return omp_get_max_threads(); }
#else
int cpu_threads(void) { return 1; }
#endif
