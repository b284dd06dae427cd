// Device building blocks of the Gauss-Legendre transform (FP64), templated on how a kernel
// reads its input and writes its output so that the solver can fuse its grid and spectral
// algebra into the transform passes.
//
// Analysis  = ring-pair FFT + hemisphere fold (k_fft_ana)  ->  per-order contraction (k_leg_ana)
// Synthesis = per-order contraction (k_leg_syn)            ->  ring-pair inverse FFT (k_fft_syn)
//
// Conventions (python/slevolve/sht.py:19-25): U[l,m] = sum_j P~_lm(mu_j) w_j rfft(g)[j,m] / nlon
// / 1.41421356; g = irfft(nlon sum_l 1.41421356 U P~).  P~ orthonormal, no Condon-Shortley phase,
// regenerated on the fly by the three-term recurrence of sht.py:105-144 in FP64.
#pragma once
#include "slx_internal.h"

namespace slx {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// -------------------------------------------------------------------------------------------
// Complex FFT of length N = 2^log2N held in LDS (a), Stockham autosort with ping-pong buffer b:
// one radix-2 stage first when log2N is odd, then radix-4 stages.  Forward uses
// exp(-2 pi i k n / N), inverse exp(+...), both unnormalised.  Returns the buffer holding the
// result.  Every thread of the block must call it.
template <int NT>
__device__ double2* fft_lds(double2* a, double2* b, const double2* __restrict__ tw, int N, int log2N,
                            bool inverse) {
  const int tid = threadIdx.x;
  int Ns = 1;
  if (log2N & 1) {
    for (int j = tid; j < N / 2; j += NT) {
      double2 v0 = a[j], v1 = a[j + N / 2];
      b[2 * j] = cadd(v0, v1);
      b[2 * j + 1] = csub(v0, v1);
    }
    __syncthreads();
    double2* t = a; a = b; b = t;
    Ns = 2;
  }
  const int q = N / 4;
  for (; Ns < N; Ns *= 4) {
    const int tstride = N / (4 * Ns);
    for (int j = tid; j < q; j += NT) {
      const int k = j & (Ns - 1);
      double2 v0 = a[j], v1 = a[j + q], v2 = a[j + 2 * q], v3 = a[j + 3 * q];
      if (Ns > 1) {
        const int st = k * tstride;
        if (!inverse) {
          v1 = cmul(v1, tw[st]);
          v2 = cmul(v2, tw[2 * st]);
          v3 = cmul(v3, tw[3 * st]);
        } else {
          v1 = cmulc(v1, tw[st]);
          v2 = cmulc(v2, tw[2 * st]);
          v3 = cmulc(v3, tw[3 * st]);
        }
      }
      double2 s02 = cadd(v0, v2), d02 = csub(v0, v2), s13 = cadd(v1, v3), d13 = csub(v1, v3);
      // -i * d13 = (d13.y, -d13.x)
      double2 y1, y3;
      if (!inverse) {
        y1 = make_double2(d02.x + d13.y, d02.y - d13.x);
        y3 = make_double2(d02.x - d13.y, d02.y + d13.x);
      } else {
        y1 = make_double2(d02.x - d13.y, d02.y + d13.x);
        y3 = make_double2(d02.x + d13.y, d02.y - d13.x);
      }
      const int o = (j - k) * 4 + k;
      b[o] = cadd(s02, s13);
      b[o + Ns] = y1;
      b[o + 2 * Ns] = csub(s02, s13);
      b[o + 3 * Ns] = y3;
    }
    __syncthreads();
    double2* t = a; a = b; b = t;
  }
  return a;
}

// Predication of a launch on device flags: the kernel runs only if (*f0 != 0) == w0 and
// (*f1 != 0) == w1 for the flags given (lets the host enqueue work ahead of a device-side test,
// e.g. inner iterations ahead of the convergence flag).
struct Pred {
  const int* f0 = nullptr;
  int w0 = 0;
  const int* f1 = nullptr;
  int w1 = 0;
  __device__ __forceinline__ bool active() const {
    if (f0 && ((*(volatile const int*)f0 != 0) != (w0 != 0))) return false;
    if (f1 && ((*(volatile const int*)f1 != 0) != (w1 != 0))) return false;
    return true;
  }
};

// Device view of a transform handle.
struct ShtDev {
  int nlat, nlon, lmax, nh, nm, log2n;
  const double* x;
  const double* wn;
  const double* ws;
  const double* pmm;
  const int* pms;
  const double2* rec;
  const double* s2m3;
  const int* moff;
  const int* jb;
  const double2* tw;
};

// -------------------------------------------------------------------------------------------
// Analysis front end: one block per (northern ring j, field f).  The two rows j and nlat-1-j are
// packed as the real and imaginary parts of one complex sequence, transformed once, and split
// into the two real spectra; the fold with the quadrature weights gives
//   S[m] = Fn w_n + Fs w_s   (pairs with even l - m),   D[m] = Fn w_n - Fs w_s   (odd l - m),
// written order-major: WS[f][m][j], WD[f][m][j].
// Loader: double operator()(int f, long idx) const -> grid value of field f at flat index idx.

template <int NT, class Loader>
__global__ void __launch_bounds__(NT) k_fft_ana(ShtDev d, Loader ld, double2* __restrict__ WS,
                                                double2* __restrict__ WD, Pred pr) {
  if (!pr.active()) return;
  extern __shared__ double2 smem[];
  const int N = d.nlon;
  double2* a = smem;
  double2* b = smem + N;
  const int j = blockIdx.x, f = blockIdx.y, js = d.nlat - 1 - j;
  const long rn = (long)j * N, rs = (long)js * N;
  for (int i = threadIdx.x; i < N; i += NT) a[i] = make_double2(ld(f, rn + i), ld(f, rs + i));
  __syncthreads();
  double2* Z = fft_lds<NT>(a, b, d.tw, N, d.log2n, false);
  const double wn = d.wn[j], ws = d.ws[j];
  const long base = (long)f * d.nm * d.nh;
  for (int m = threadIdx.x; m <= d.lmax; m += NT) {
    double2 zm = Z[m], zc = Z[(N - m) & (N - 1)];
    // Fn = (Z_m + conj Z_{N-m}) / 2 ; Fs = (Z_m - conj Z_{N-m}) / (2i)
    double2 fn = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y - zc.y));
    double2 fs = make_double2(0.5 * (zm.y + zc.y), -0.5 * (zm.x - zc.x));
    double2 tn = make_double2(fn.x * wn, fn.y * wn), ts = make_double2(fs.x * ws, fs.y * ws);
    WS[base + (long)m * d.nh + j] = cadd(tn, ts);
    WD[base + (long)m * d.nh + j] = csub(tn, ts);
  }
}

// -------------------------------------------------------------------------------------------
// Per-order analysis contraction.  One block of NW waves per order m; lane owns R rings
// (j = lane + 64 (wave + NW r)); the FP64 recurrence runs per ring along l; each chunk of
// CH = 32/NF degrees produces 64 partial sums per lane (degree x field x re/im) that are summed
// over the 64 lanes by a reduce-scatter (lane i ends with value i), then over the waves in LDS
// in a fixed order.  out[f] is packed; the result is divided by 1.41421356.
template <int NF>
struct OutPtrs {
  double2* p[NF];
};

template <int NF, int R, int NW>
__global__ void __launch_bounds__(64 * NW) k_leg_ana(ShtDev d, const double2* __restrict__ WS,
                                                     const double2* __restrict__ WD, OutPtrs<NF> out,
                                                     Pred pr) {
  if (!pr.active()) return;
  constexpr int CH = 32 / NF;
  __shared__ double red[2][NW][64];
  const int m = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L = d.lmax, nh = d.nh;
  const long fstride = (long)d.nm * nh;
  double xr[R], pa[R], pb[R];
  int sc[R];
  double2 S[R][NF], D[R][NF];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = lane + 64 * (wave + NW * r);
    const bool ok = j < nh;
    xr[r] = ok ? d.x[j] : 0.0;
    pa[r] = 0.0;
    pb[r] = ok ? d.pmm[(long)m * nh + j] : 0.0;  // P_{m}; pa is P_{m-1} := 0
    sc[r] = ok ? d.pms[(long)m * nh + j] : 0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      S[r][f] = ok ? WS[f * fstride + (long)m * nh + j] : make_double2(0.0, 0.0);
      D[r][f] = ok ? WD[f * fstride + (long)m * nh + j] : make_double2(0.0, 0.0);
    }
  }
  const double s23 = d.s2m3[m];
  const int mo = d.moff[m];
  int buf = 0;
  for (int l0 = m; l0 <= L; l0 += CH) {
    double acc[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] = 0.0;
    double mk[R];  // 1 where the ring's values are in range this chunk, 0 while still scaled
#pragma unroll
    for (int r = 0; r < R; ++r) mk[r] = sc[r] == 0 ? 1.0 : 0.0;
#pragma unroll
    for (int il = 0; il < CH; ++il) {
      const int l = l0 + il;
      double c0 = 0.0, c1 = 0.0;  // (eps_{l-1}, 1/eps_l)
      const bool first = (l == m), second = (l == m + 1);
      if (l > m + 1 && l <= L) {
        double2 c = d.rec[mo + (l - m)];
        c0 = c.x;
        c1 = c.y;
      }
      const bool even = ((l - m) & 1) == 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double p;
        if (first) {
          p = pb[r];
        } else if (second) {
          p = s23 * xr[r] * pb[r];
          pa[r] = pb[r];
          pb[r] = p;
        } else {
          p = (xr[r] * pb[r] - c0 * pa[r]) * c1;
          pa[r] = pb[r];
          pb[r] = p;
        }
        p *= mk[r];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          double2 v = even ? S[r][f] : D[r][f];
          acc[(il * NF + f) * 2 + 0] += p * v.x;
          acc[(il * NF + f) * 2 + 1] += p * v.y;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (sc[r] > 0 && fabs(pb[r]) >= 0x1p300) {
        pa[r] *= 0x1p-600;
        pb[r] *= 0x1p-600;
        sc[r] -= 1;
      }
    }
    // reduce-scatter over the 64 lanes: after the step with offset o, lane keeps the half of
    // the remaining values selected by its bit o; lane i ends with the sum of value i
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const bool up = (lane & o) != 0;
#pragma unroll
      for (int i = 0; i < o; ++i) {
        double send = up ? acc[i] : acc[i + o];
        double keep = up ? acc[i + o] : acc[i];
        acc[i] = keep + __shfl_xor(send, o, 64);
      }
    }
    red[buf][wave][lane] = acc[0];
    __syncthreads();
    if (wave == 0) {
      double v = red[buf][0][lane];
#pragma unroll
      for (int w = 1; w < NW; ++w) v += red[buf][w][lane];
      const int il = lane / (2 * NF), f = (lane / 2) % NF, c = lane & 1;
      const int l = l0 + il;
      if (l <= L) reinterpret_cast<double*>(out.p[f])[2 * (long)(mo + l - m) + c] = v / SLX_SQRT2_LEGACY;
    }
    buf ^= 1;
  }
}

// -------------------------------------------------------------------------------------------
// Per-order synthesis contraction.  Block (order m, ring block); one thread per northern ring;
// the order's coefficients are staged in LDS through the Coef loader (which may assemble them
// from several spectral arrays); E = sum over even l-m, O = sum over odd l-m; outputs
// FN = 1.41421356 (E + O) (north), FS = 1.41421356 (E - O) (south), order-major [f][m][j].
// Coef: double2 operator()(int f, int idx, int l, int m) const, idx = packed index.
template <int NF, int NT, class Coef>
__global__ void __launch_bounds__(NT) k_leg_syn(ShtDev d, Coef cf, double2* __restrict__ FN,
                                                double2* __restrict__ FS, Pred pr) {
  if (!pr.active()) return;
  extern __shared__ double2 cl[];  // [NF][L+1]
  const int m = blockIdx.x;
  const int L = d.lmax, nl = L + 1 - m;
  const int mo = d.moff[m];
  for (int i = threadIdx.x; i < nl; i += NT) {
#pragma unroll
    for (int f = 0; f < NF; ++f) cl[f * (L + 1) + i] = cf(f, mo + i, m + i, m);
  }
  __syncthreads();
  const int j = blockIdx.y * NT + threadIdx.x;
  if (j >= d.nh) return;
  const double x = d.x[j];
  double pa = 0.0, pb = d.pmm[(long)m * d.nh + j];
  int sc = d.pms[(long)m * d.nh + j];
  double mk = sc == 0 ? 1.0 : 0.0;
  double2 E[NF], O[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double2 c = cl[f * (L + 1)];
    E[f] = make_double2(pb * mk * c.x, pb * mk * c.y);
    O[f] = make_double2(0.0, 0.0);
  }
  if (nl > 1) {
    double p = d.s2m3[m] * x * pb;
    pa = pb;
    pb = p;
    p *= mk;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double2 c = cl[f * (L + 1) + 1];
      O[f].x += p * c.x;
      O[f].y += p * c.y;
    }
  }
  int i = 2;
  for (; i + 1 < nl; i += 2) {
    double2 c2 = d.rec[mo + i], c3 = d.rec[mo + i + 1];
    double p2 = (x * pb - c2.x * pa) * c2.y;
    double p3 = (x * p2 - c3.x * pb) * c3.y;
    pa = p2;
    pb = p3;
    p2 *= mk;
    p3 *= mk;
    if (sc > 0 && fabs(pb) >= 0x1p300) {
      pa *= 0x1p-600;
      pb *= 0x1p-600;
      sc -= 1;
      mk = sc == 0 ? 1.0 : 0.0;
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double2 ce = cl[f * (L + 1) + i], co = cl[f * (L + 1) + i + 1];
      E[f].x += p2 * ce.x;
      E[f].y += p2 * ce.y;
      O[f].x += p3 * co.x;
      O[f].y += p3 * co.y;
    }
  }
  if (i < nl) {
    double2 c2 = d.rec[mo + i];
    double p2 = (x * pb - c2.x * pa) * c2.y * mk;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double2 ce = cl[f * (L + 1) + i];
      E[f].x += p2 * ce.x;
      E[f].y += p2 * ce.y;
    }
  }
  const long fstride = (long)d.nm * d.nh;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const long o = f * fstride + (long)m * d.nh + j;
    FN[o] = make_double2(SLX_SQRT2_LEGACY * (E[f].x + O[f].x), SLX_SQRT2_LEGACY * (E[f].y + O[f].y));
    FS[o] = make_double2(SLX_SQRT2_LEGACY * (E[f].x - O[f].x), SLX_SQRT2_LEGACY * (E[f].y - O[f].y));
  }
}

// -------------------------------------------------------------------------------------------
// Synthesis back end: one block per (northern ring j, field f).  Builds the Hermitian spectra of
// rows j and nlat-1-j (Im of m = 0 and of the Nyquist bin dropped, as irfft does), packs them as
// Z = Xn + i Xs, inverse-transforms once and hands (row, column, value) to the Store functor:
// void operator()(int f, long idx, double v) const.
template <int NT, class Store>
__global__ void __launch_bounds__(NT) k_fft_syn(ShtDev d, const double2* __restrict__ FN,
                                                const double2* __restrict__ FS, Store st,
                                                Pred pr) {
  if (!pr.active()) return;
  extern __shared__ double2 smem[];
  const int N = d.nlon, L = d.lmax;
  double2* a = smem;
  double2* b = smem + N;
  const int j = blockIdx.x, f = blockIdx.y, js = d.nlat - 1 - j;
  const long base = (long)f * d.nm * d.nh;
  for (int i = threadIdx.x; i < N; i += NT) a[i] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int m = threadIdx.x; m <= L; m += NT) {
    double2 fn = FN[base + (long)m * d.nh + j], fs = FS[base + (long)m * d.nh + j];
    if (m == 0 || 2 * m == N) {
      a[m] = make_double2(fn.x, fs.x);
    } else {
      a[m] = make_double2(fn.x - fs.y, fn.y + fs.x);          // Fn + i Fs
      a[N - m] = make_double2(fn.x + fs.y, -fn.y + fs.x);     // conj(Fn) + i conj(Fs)
    }
  }
  __syncthreads();
  double2* z = fft_lds<NT>(a, b, d.tw, N, d.log2n, true);
  const long rn = (long)j * N, rs = (long)js * N;
  for (int i = threadIdx.x; i < N; i += NT) {
    double2 v = z[i];
    st(f, rn + i, v.x);
    st(f, rs + i, v.y);
  }
}

}  // namespace slx
