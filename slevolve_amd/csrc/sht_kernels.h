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
#include <type_traits>
#include <utility>
#include "slx_internal.h"

namespace slx {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// complex products with their fused multiply-adds written out, so that every kernel rounds them
// the same way whatever the compiler's contraction choices
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -(a.x * b.y)));
}

// -------------------------------------------------------------------------------------------
// Ring passes run one block of nt = N / P threads per ring pair (N = nlon): thread t owns points
// t + k nt (k < P) of both rows.  The passes are latency-bound (a few hundred blocks, a short chain
// of dependent phases each), so every global load of a pass is issued into registers before the
// first LDS phase and the FFT loops have exact trip counts.  P = 4 up to N = 2048 (nt <= 512), 8
// at N = 4096 (two rows and their ping-pong buffer fill the LDS at that size).
constexpr int RING_MAXNT = 512;

// FFT twiddles.  stw (ShtDev::stw) holds, for every radix-4 stage with Ns > 1 in stage order,
// W1[0..Ns), W2[0..Ns), W3[0..Ns) consecutively (fewer than N entries); a butterfly j = tid + r nt
// of a stage uses W1[k], W2[k], W3[k] with k = j & (Ns - 1), at offset off of its stage.
//
// TwReg: the thread's twiddles of every stage held in registers.  A ring pass with 4 points per
// thread loads them once, together with its other global reads, and every FFT of the pass
// (forward and inverse) uses them, so no stage waits on a memory round trip.
template <int P>
struct TwReg {
  static constexpr int MAXST = P == 4 ? 5 : 6;  // radix-4 stages of N <= 4 P RING_MAXNT
  static constexpr int NB = P / 4;
  double2 w[MAXST][3][NB];
  __device__ __forceinline__ void load(const double2* __restrict__ stw, int N, int log2N) {
    const int tid = threadIdx.x, nt = N / P;
    int Ns = (log2N & 1) ? 2 : 1;
    const double2* w0 = stw;
#pragma unroll
    for (int st = 0; st < MAXST; ++st) {
      const bool on = Ns > 1 && Ns < N;
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int k = (tid + r * nt) & (Ns - 1);
        w[st][0][r] = on ? w0[k] : make_double2(1.0, 0.0);
        w[st][1][r] = on ? w0[Ns + k] : make_double2(1.0, 0.0);
        w[st][2][r] = on ? w0[2 * Ns + k] : make_double2(1.0, 0.0);
      }
      if (on) w0 += 3 * Ns;
      if (Ns < N) Ns *= 4;
    }
  }
  __device__ __forceinline__ void get(int st, int r, int, int, int, double2& w1, double2& w2, double2& w3) const {
    w1 = w[st][0][r];
    w2 = w[st][1][r];
    w3 = w[st][2][r];
  }
};
// TwMem: the twiddles read from stw at each stage (8 points per thread, where a register copy of
// every stage would take 144 registers).
struct TwMem {
  const double2* stw;
  __device__ __forceinline__ void load(const double2* __restrict__ t, int, int) { stw = t; }
  __device__ __forceinline__ void get(int, int, int off, int Ns, int k, double2& w1, double2& w2,
                                      double2& w3) const {
    w1 = stw[off + k];
    w2 = stw[off + Ns + k];
    w3 = stw[off + 2 * Ns + k];
  }
};
// the twiddle holder of a ring pass with P points per thread (REG = false: per-stage reads)
template <int P, bool REG = true>
using TwOf = typename std::conditional<P == 4 && REG, TwReg<4>, TwMem>::type;

// Complex FFT of length N = 2^log2N held in LDS (a), Stockham autosort with ping-pong buffer b:
// one radix-2 stage first when log2N is odd, then radix-4 stages.  Forward uses
// exp(-2 pi i k n / N), inverse exp(+...), both unnormalised; tw supplies each stage's twiddles.
// Returns the buffer holding the result.  Every thread of the block (nt = N / P of them) must
// call it.
template <int P, class TW>
__device__ double2* fft_lds(double2* a, double2* b, const TW& tw, int N, int log2N, bool inverse) {
  const int tid = threadIdx.x, nt = N / P;
  int Ns = 1;
  if (log2N & 1) {
#pragma unroll
    for (int r = 0; r < P / 2; ++r) {
      const int j = tid + r * nt;
      double2 v0 = a[j], v1 = a[j + N / 2];
      b[2 * j] = cadd(v0, v1);
      b[2 * j + 1] = csub(v0, v1);
    }
    __syncthreads();
    double2* t = a; a = b; b = t;
    Ns = 2;
  }
  const int q = N / 4;
  int off = 0;  // this stage's W1 in stw
#pragma unroll
  for (int st = 0; st < 6; ++st) {  // N <= 4096
    if (Ns >= N) break;
#pragma unroll
    for (int r = 0; r < P / 4; ++r) {
      const int j = tid + r * nt;
      const int k = j & (Ns - 1);
      double2 v0 = a[j], v1 = a[j + q], v2 = a[j + 2 * q], v3 = a[j + 3 * q];
      if (Ns > 1) {
        double2 w1, w2, w3;
        tw.get(st, r, off, Ns, k, w1, w2, w3);
        if (!inverse) {
          v1 = cmul(v1, w1);
          v2 = cmul(v2, w2);
          v3 = cmul(v3, w3);
        } else {
          v1 = cmulc(v1, w1);
          v2 = cmulc(v2, w2);
          v3 = cmulc(v3, w3);
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
    if (Ns > 1) off += 3 * Ns;
    __syncthreads();
    double2* t = a; a = b; b = t;
    Ns *= 4;
  }
  return a;
}

// Fourier coefficients of a ring pair, read into registers: m = tid + k nt for k < MP covers
// 0..L (L <= N/2 < MP nt); entries past L are zero.
template <int P>
struct RingCoef {
  static constexpr int MP = P / 2 + 1;
  double2 fn[MP], fs[MP];
  __device__ __forceinline__ void load(const double2* __restrict__ FN, const double2* __restrict__ FS, long base,
                                       int nh, int j, int L, int nt) {
#pragma unroll
    for (int k = 0; k < MP; ++k) {
      const int m = threadIdx.x + k * nt;
      fn[k] = fs[k] = make_double2(0.0, 0.0);
      if (m <= L) {
        fn[k] = FN[base + (long)m * nh + j];
        fs[k] = FS[base + (long)m * nh + j];
      }
    }
  }
  // Hermitian spectra of rows j (Fn) and nlat-1-j (Fs) packed as Z = Xn + i Xs, Im of m = 0 and
  // of the Nyquist bin dropped as irfft does: Z[m] = Fn + i Fs, Z[N-m] = conj(Fn) + i conj(Fs),
  // zero for L < i < N - L.  Every index is written exactly once (no barrier between the parts).
  __device__ __forceinline__ void scatter(double2* a, int N, int L, int nt) const {
    for (int i = L + 1 + threadIdx.x; i < N - L; i += nt) a[i] = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < MP; ++k) {
      const int m = threadIdx.x + k * nt;
      if (m > L) continue;
      const double2 f = fn[k], g = fs[k];
      if (m == 0 || 2 * m == N) {
        a[m] = make_double2(f.x, g.x);
      } else {
        a[m] = make_double2(f.x - g.y, f.y + g.x);          // Fn + i Fs
        a[N - m] = make_double2(f.x + g.y, -f.y + g.x);     // conj(Fn) + i conj(Fs)
      }
    }
  }
};

// Ring of a ring-pass block, XCD-aware.  The hardware deals block b to XCD b % 8; block b takes
// ring start(b % 8) + b / 8, so that every XCD handles one contiguous run of n / 8 (+1) rings.  The
// order-major arrays [m][j] hold eight neighbouring rings in one 128-byte line, which is then
// written (analysis) or read (synthesis) through one XCD's L2 instead of eight.
__device__ __forceinline__ int xcd_ring(int b, int n) {
  const int x = b & 7, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
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
    if (!f0 && !f1) return true;
    // both flags are loaded, from valid addresses, before either is tested (one memory round
    // trip, not two); the flags were written by earlier kernels of the stream
    const int* p0 = f0 ? f0 : f1;
    const int* p1 = f1 ? f1 : f0;
    const int v0 = __hip_atomic_load(p0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int v1 = __hip_atomic_load(p1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (!f0 || (v0 != 0) == (w0 != 0)) && (!f1 || (v1 != 0) == (w1 != 0));
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
  const double2* stw;  // per-stage FFT twiddles (fft_lds)
  const int* l_of;  // degree of each packed index
  const int* m_of;  // order of each packed index
  const double* qc;  // scaled recurrence coefficients (matrix-core kernels)
  const double* qg;  // P~_l = qg_l Q_l
};

// -------------------------------------------------------------------------------------------
// Analysis front end: one block per (northern ring j, field f).  The two rows j and nlat-1-j are
// packed as the real and imaginary parts of one complex sequence, transformed once, and split
// into the two real spectra; the fold with the quadrature weights gives
//   S[m] = Fn w_n + Fs w_s   (pairs with even l - m),   D[m] = Fn w_n - Fs w_s   (odd l - m),
// written order-major: WS[f][m][j], WD[f][m][j].
// Loader: double operator()(int f, long idx) const -> grid value of field f at flat index idx
// (reads only, so that the pass can issue all of them at once).

// A loader may carry a device-side gate() (evaluated by every thread; the pass returns when it
// is false) and a commit(f, idx, v) that stores what it derives from the values it supplied
// (called after all loads are in flight); loaders without them always run and store nothing.
template <class L, class = void>
struct has_gate : std::false_type {};
template <class L>
struct has_gate<L, std::void_t<decltype(std::declval<const L&>().gate())>> : std::true_type {};
template <class L>
__device__ __forceinline__ bool loader_gate(const L& l) {
  if constexpr (has_gate<L>::value)
    return l.gate();
  else
    return true;
}
template <class L, class = void>
struct has_commit : std::false_type {};
template <class L>
struct has_commit<L, std::void_t<decltype(std::declval<const L&>().commit(0, 0L, 0.0))>> : std::true_type {};

// Split of the packed transform Z into the rows' spectra and the fold with the quadrature weights
// (shared by every analysis ring pass): Fn = (Z_m + conj Z_{N-m}) / 2, Fs = (Z_m - conj Z_{N-m}) / (2i).
__device__ __forceinline__ void ring_fold(const double2* Z, double2* __restrict__ WS, double2* __restrict__ WD,
                                          long base, int nh, int j, int L, int N, double wn, double ws, int nt) {
  for (int m = threadIdx.x; m <= L; m += nt) {
    const double2 zm = Z[m], zc = Z[(N - m) & (N - 1)];
    const double2 fn = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y - zc.y));
    const double2 fs = make_double2(0.5 * (zm.y + zc.y), -0.5 * (zm.x - zc.x));
    const double2 tn = make_double2(fn.x * wn, fn.y * wn), ts = make_double2(fs.x * ws, fs.y * ws);
    WS[base + (long)m * nh + j] = cadd(tn, ts);
    WD[base + (long)m * nh + j] = csub(tn, ts);
  }
}

template <int P, class Loader>
__global__ void __launch_bounds__(RING_MAXNT) k_fft_ana(ShtDev d, Loader ld, double2* __restrict__ WS,
                                                        double2* __restrict__ WD, Pred pr) {
  if (!pr.active() || !loader_gate(ld)) return;
  extern __shared__ double2 smem[];
  const int N = d.nlon, nt = N / P, tid = threadIdx.x;
  double2* a = smem;
  double2* b = smem + N;
  const int j = xcd_ring(blockIdx.x, d.nh), f = blockIdx.y, js = d.nlat - 1 - j;
  const long rn = (long)j * N, rs = (long)js * N;
  TwOf<P> tw;
  tw.load(d.stw, N, d.log2n);
  double vn[P], vs[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    vn[k] = ld(f, rn + tid + k * nt);
    vs[k] = ld(f, rs + tid + k * nt);
  }
#pragma unroll
  for (int k = 0; k < P; ++k) a[tid + k * nt] = make_double2(vn[k], vs[k]);
  if constexpr (has_commit<Loader>::value) {
#pragma unroll
    for (int k = 0; k < P; ++k) {
      ld.commit(f, rn + tid + k * nt, vn[k]);
      ld.commit(f, rs + tid + k * nt, vs[k]);
    }
  }
  __syncthreads();
  double2* Z = fft_lds<P>(a, b, tw, N, d.log2n, false);
  ring_fold(Z, WS, WD, (long)f * d.nm * d.nh, d.nh, j, d.lmax, N, d.wn[j], d.ws[j], nt);
}

// -------------------------------------------------------------------------------------------
// The degree recurrence, shared by every kernel that walks it so that all of them produce
// bit-identical P~_lm: explicit FMA, renormalisation test every 8 degrees (after l with
// (l - m) % 8 == 7) in extended range (SLX_SCALE_K / SLX_SCALE_T).
struct LegState {
  double pa, pb;  // P~_{l-2}, P~_{l-1} (scaled by 2^(SLX_SCALE_K sc))
  int sc;
};

// advance to degree l (> m+1) with (eps_{l-1}, 1/eps_l) = c
__device__ __forceinline__ double leg_next(double x, double2 c, LegState& s) {
  const double p = fma(x, s.pb, -(c.x * s.pa)) * c.y;
  s.pa = s.pb;
  s.pb = p;
  return p;
}
__device__ __forceinline__ void leg_renorm(LegState& s) {
  const bool big = (s.sc > 0) && (fabs(s.pb) >= 0x1p300);
  const double f = big ? 0x1p-600 : 1.0;
  s.pa *= f;
  s.pb *= f;
  s.sc -= big ? 1 : 0;
}

// Segments of the degree range: order m, first degree l0 (segments start every S degrees from
// l = m, S a multiple of 8); the recurrence state before l0 for every northern ring (seed) is
// produced once on the device by k_leg_seed with the same arithmetic.
struct SegDev {
  int nseg, S;
  int Ssyn;  // degrees per synthesis chain (a multiple of S); chain s of order m starts at seed
             // first[m] + s Ssyn / S
  const int* seg_m;
  const int* seg_l0;
  const double* sa;  // [nseg][nh] P~_{l0-2}
  const double* sb;  // [nseg][nh] P~_{l0-1}
  const int* ss;     // [nseg][nh] scale count
  const double2* sab;  // [nseg][nh] start states of the matrix-core kernels (k_seed_unscale)
  const int* first;  // [nm] index of the first segment of order m
};

// One thread per (m, ring): walks l = m .. L and records the state at every segment start.
static __global__ void k_leg_seed(ShtDev d, SegDev g, double* sa, double* sb, int* ss) {
  const int m = blockIdx.x;
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= d.nh) return;
  const int L = d.lmax, mo = d.moff[m];
  const double x = d.x[j];
  LegState s{0.0, d.pmm[(long)m * d.nh + j], d.pms[(long)m * d.nh + j]};
  int seg = g.first[m] + 1;
  for (int l = m + 1; l <= L; ++l) {
    if (l == m + 1) {
      const double p = d.s2m3[m] * x * s.pb;
      s.pa = s.pb;
      s.pb = p;
    } else {
      leg_next(x, d.rec[mo + (l - m)], s);
    }
    if (((l - m) & 7) == 7) leg_renorm(s);
    if (((l + 1 - m) % g.S) == 0 && l + 1 <= L) {
      sa[(long)seg * d.nh + j] = s.pa;
      sb[(long)seg * d.nh + j] = s.pb;
      ss[(long)seg * d.nh + j] = s.sc;
      ++seg;
    }
  }
}

// -------------------------------------------------------------------------------------------
// Analysis contraction over one segment of degrees of one order per block (NW waves; lane owns
// R rings j = lane + 64 (wave + NW r)).  Each chunk of CH degrees (a multiple of 8, so parity and
// renormalisation points are compile-time) gives 64 partial sums per lane (degree x field x
// re/im); a reduce-scatter over the 64 lanes leaves lane i with the total of value i (the two
// upper levels with the gfx950 half-wave / row swaps, the four lower ones with xor shuffles),
// then a fixed-order sum over the waves in LDS.  out[f] is packed; results are / 1.41421356.
template <int NF>
struct OutPtrs {
  double2* p[NF];
};

// acc[i] + acc[i + 32] over lane halves: lanes 0..31 end with value i, 32..63 with value i + 32
__device__ __forceinline__ double swap_add32(double a, double b) {
  const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  const double a2 = __longlong_as_double(((unsigned long long)hi[0] << 32) | lo[0]);
  const double b2 = __longlong_as_double(((unsigned long long)hi[1] << 32) | lo[1]);
  return a2 + b2;
}
// the same across 16-lane rows (bit 4 of the lane)
__device__ __forceinline__ double swap_add16(double a, double b) {
  const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  const double a2 = __longlong_as_double(((unsigned long long)hi[0] << 32) | lo[0]);
  const double b2 = __longlong_as_double(((unsigned long long)hi[1] << 32) | lo[1]);
  return a2 + b2;
}

// one lower level of the reduce-scatter: pairs (i, i + O) across lanes differing in bit O
template <int O>
__device__ __forceinline__ void xor_level(double (&acc)[64], int lane) {
  const bool up = (lane & O) != 0;
#pragma unroll
  for (int i = 0; i < O; ++i) {
    const double send = up ? acc[i] : acc[i + O];
    const double keep = up ? acc[i + O] : acc[i];
    acc[i] = keep + __shfl_xor(send, O, 64);
  }
}

// sum of acc[v] over the wave for v = lane (reduce-scatter, fixed pairing order)
__device__ __forceinline__ double reduce_scatter64(double (&acc)[64], int lane) {
#pragma unroll
  for (int i = 0; i < 32; ++i) acc[i] = swap_add32(acc[i], acc[i + 32]);
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = swap_add16(acc[i], acc[i + 16]);
  xor_level<8>(acc, lane);
  xor_level<4>(acc, lane);
  xor_level<2>(acc, lane);
  xor_level<1>(acc, lane);
  return acc[0];
}

template <int NF, int R, int NW>
__global__ void __launch_bounds__(64 * NW) k_leg_ana(ShtDev d, SegDev g, const double2* __restrict__ WS,
                                                     const double2* __restrict__ WD, OutPtrs<NF> out,
                                                     Pred pr) {
  if (!pr.active()) return;
  constexpr int CH = (NF == 1) ? 32 : (NF == 2 ? 16 : 8);
  __shared__ double red[2][NW][64];
  const int seg = blockIdx.x;
  const int m = g.seg_m[seg], l0 = g.seg_l0[seg];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int L = d.lmax, nh = d.nh;
  const int lend = min(l0 + g.S, L + 1);
  const long fstride = (long)d.nm * nh;
  double xr[R], mk[R];
  LegState st[R];
  double2 S[R][NF], D[R][NF];
  const bool head = (l0 == m);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = lane + 64 * (wave + NW * r);
    const bool ok = j < nh;
    const int jj = ok ? j : 0;
    xr[r] = ok ? d.x[jj] : 0.0;
    const double a0 = head ? 0.0 : g.sa[(long)seg * nh + jj];
    const double b0 = head ? d.pmm[(long)m * nh + jj] : g.sb[(long)seg * nh + jj];
    const int s0 = head ? d.pms[(long)m * nh + jj] : g.ss[(long)seg * nh + jj];
    st[r] = LegState{ok ? a0 : 0.0, ok ? b0 : 0.0, ok ? s0 : 0};
    mk[r] = st[r].sc == 0 ? 1.0 : 0.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      S[r][f] = ok ? WS[f * fstride + (long)m * nh + jj] : make_double2(0.0, 0.0);
      D[r][f] = ok ? WD[f * fstride + (long)m * nh + jj] : make_double2(0.0, 0.0);
    }
  }
  const double s23 = d.s2m3[m];
  const int mo = d.moff[m];
  int buf = 0;
  for (int lc = l0; lc < lend; lc += CH) {
    const bool hd = (lc == m);
    double2 cc[CH];
#pragma unroll
    for (int il = 0; il < CH; ++il) {
      const int off = min(max(lc + il - m, 0), L - m);
      const double2 c = d.rec[mo + off];
      const bool use = (lc + il >= m + 2) && (lc + il <= L);
      cc[il] = use ? c : make_double2(0.0, 0.0);
    }
    double acc[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] = 0.0;
#pragma unroll
    for (int il = 0; il < CH; ++il) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double p;
        if (il == 0) {  // head: P_m itself, the state does not move
          const double pa0 = st[r].pa, pb0 = st[r].pb;
          const double pn = leg_next(xr[r], cc[il], st[r]);
          p = hd ? pb0 : pn;
          st[r].pa = hd ? pa0 : st[r].pa;
          st[r].pb = hd ? pb0 : st[r].pb;
        } else if (il == 1) {  // head: P_{m+1} = sqrt(2m+3) x P_m
          const double ps = s23 * xr[r] * st[r].pb;
          const double pn = leg_next(xr[r], cc[il], st[r]);
          p = hd ? ps : pn;
          st[r].pb = p;
        } else {
          p = leg_next(xr[r], cc[il], st[r]);
        }
        p *= mk[r];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          const double2 v = (il & 1) == 0 ? S[r][f] : D[r][f];
          acc[(il * NF + f) * 2 + 0] += p * v.x;
          acc[(il * NF + f) * 2 + 1] += p * v.y;
        }
        if ((il & 7) == 7) {
          leg_renorm(st[r]);
          mk[r] = st[r].sc == 0 ? 1.0 : 0.0;
        }
      }
    }
    double v = reduce_scatter64(acc, lane);
    if (NW > 1) {
      red[buf][wave][lane] = v;
      __syncthreads();
      v = red[buf][0][lane];
#pragma unroll
      for (int w = 1; w < NW; ++w) v += red[buf][w][lane];
      buf ^= 1;
    }
    if (wave == 0 && lane < 2 * NF * CH) {
      const int il = lane / (2 * NF), f = (lane / 2) % NF, c2 = lane & 1;
      const int l = lc + il;
      if (l < lend) reinterpret_cast<double*>(out.p[f])[2 * (long)(mo + l - m) + c2] = v / SLX_SQRT2_LEGACY;
    }
  }
}

// -------------------------------------------------------------------------------------------
// Synthesis back end: one block per (northern ring j, field f).  Builds the Hermitian spectra of
// rows j and nlat-1-j (RingCoef::scatter), inverse-transforms them once and hands (row, column,
// value) to the Store functor: void operator()(int f, long idx, double v) const.
template <int P, class Store>
__global__ void __launch_bounds__(RING_MAXNT) k_fft_syn(ShtDev d, const double2* __restrict__ FN,
                                                        const double2* __restrict__ FS, Store st, Pred pr) {
  if (!pr.active()) return;
  extern __shared__ double2 smem[];
  const int N = d.nlon, L = d.lmax, nt = N / P, tid = threadIdx.x;
  double2* a = smem;
  double2* b = smem + N;
  const int j = xcd_ring(blockIdx.x, d.nh), f = blockIdx.y, js = d.nlat - 1 - j;
  TwOf<P> tw;
  tw.load(d.stw, N, d.log2n);
  RingCoef<P> rc;
  rc.load(FN, FS, (long)f * d.nm * d.nh, d.nh, j, L, nt);
  rc.scatter(a, N, L, nt);
  __syncthreads();
  double2* z = fft_lds<P>(a, b, tw, N, d.log2n, true);
  const long rn = (long)j * N, rs = (long)js * N;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int i = tid + k * nt;
    const double2 v = z[i];
    st(f, rn + i, v.x);
    st(f, rs + i, v.y);
  }
}

// -------------------------------------------------------------------------------------------
// Analysis contraction on the matrix cores: v_mfma_f64_4x4x4_4b_f64 (four independent 4x4x4
// blocks per wave, FP64 in and out).  Lane roles of that instruction (established on the device
// with a one-hot probe): block b = (lane >> 2) & 3; the A operand A[i][k] comes from lane
// 16 k + 4 b + i, B[k][j] from lane 16 k + 4 b + j, and D[i][j] is held by lane 16 i + 4 b + j.
//   rows i    = four consecutive degree segments ("slots") of one order m, S degrees apart;
//   K index k = four northern rings;  columns j = field x (re, im).
// A lane walks the recurrence of ring k through segment i, so at degree step t the four rows hold
// the degrees l0_i + t, whose l - m all have the parity of t (S is a multiple of 8): the folded
// Fourier value of the lane's ring for that parity (S or D) is the B operand, and the matrix core
// does the sum over rings.  Each lane owns R rings (R MFMAs per degree step into the same
// accumulator); the S accumulators of a segment group are reduced over the 4 MFMA blocks (lane
// bits 2 and 3) by a two-level reduce-scatter and over the waves in LDS, in a fixed order.
// Block = one group of 4 segments of one order; consecutive groups (same order, same Fourier
// inputs) are given to the same XCD (the hardware deals block b to XCD b % 8).
struct GrpDev {
  int ngrp;
  const int* m;   // order of the group
  const int* s0;  // first segment (index within the order) of the group
  const int4* desc = nullptr;  // (m, s0, seed row of segment s0, packed offset of order m)
};

constexpr int MF_S = 32;  // degrees per segment for the matrix-core analysis
// LDS row stride (doubles) of a segment's staged recurrence tables.  The lanes of one instruction
// read the same degree of 4 segments at once; with a stride of 32 doubles those 4 addresses fall in
// the same bank, with 34 they start 4 banks apart (conflict-free, and 16-byte aligned for b128 reads).
constexpr int MF_QS = MF_S + 2;
// LDS stride (double2) of a segment's staged synthesis coefficients: 33 entries put the 4 segments'
// same-degree entries 4 banks apart.
constexpr int MF_CS = MF_S + 1;

// acc[v] summed over the lanes that differ in bits 2 and 3 (the 4 MFMA blocks), reduce-scatter:
// afterwards acc[v] (v < 8) holds the sum of value 16 ((lane >> 3) & 1) + 8 ((lane >> 2) & 1) + v.
__device__ __forceinline__ void mf_block_reduce(double (&acc)[32], int lane) {
  {
    const bool up = (lane & 8) != 0;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const double send = up ? acc[v] : acc[v + 16];
      const double keep = up ? acc[v + 16] : acc[v];
      acc[v] = keep + __shfl_xor(send, 8, 64);
    }
  }
  {
    const bool up = (lane & 4) != 0;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const double send = up ? acc[v] : acc[v + 8];
      const double keep = up ? acc[v + 8] : acc[v];
      acc[v] = keep + __shfl_xor(send, 4, 64);
    }
  }
}

// one step of the scaled recurrence: Q_l = x Q_{l-1} - c Q_{l-2}
__device__ __forceinline__ double q_next(double x, double c, LegState& s) {
  const double q = fma(x, s.pb, -(c * s.pa));
  s.pa = s.pb;
  s.pb = q;
  return q;
}
// Start state of a segment for the matrix-core kernels, which carry no scale counts: a lane whose
// scale count is still > 0 at the next segment start (sc_next) stays below 2^-300 through the
// segment and starts from zero (its terms are below 2^-300 relative and are dropped, as the
// masked recurrence drops them); a lane whose scale count reaches 0 inside the segment started it
// with s = 1 (the functions grow by far less than 2^600 over 32 degrees), so its true value
// P~ 2^-600 lies in [2^-900, 2^-300), a normal double: it runs unscaled, keeping those terms.
__device__ __forceinline__ void mf_unscale(double& a, double& b, int sc, int sc_next) {
  if (sc <= 0) return;
  if (sc_next > 0) {
    a = b = 0.0;
  } else {
    a = ldexp(a, -SLX_SCALE_K * sc);
    b = ldexp(b, -SLX_SCALE_K * sc);
  }
}

// The start state (P~_{l0-2}, P~_{l0-1}) of every segment as the matrix-core kernels take it, one
// double2 per (segment, ring): the sectoral pair (0, P~_m) for the first segment of an order, the
// recorded seed otherwise, each through mf_unscale.  Built once per handle, so that a kernel lane
// starts its segment with one 16-byte load.
static __global__ void k_seed_unscale(ShtDev d, SegDev g, double2* sab) {
  const int nh = d.nh;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)g.nseg * nh) return;
  const int seg = (int)(i / nh), j = (int)(i - (long)seg * nh);
  const int m = g.seg_m[seg], segi = seg - g.first[m];
  const int nsm = (d.lmax - m + g.S) / g.S;
  double a0 = 0.0, b0;
  int c0;
  if (segi == 0) {
    b0 = d.pmm[(long)m * nh + j];
    c0 = d.pms[(long)m * nh + j];
  } else {
    a0 = g.sa[i];
    b0 = g.sb[i];
    c0 = g.ss[i];
  }
  mf_unscale(a0, b0, c0, segi + 1 < nsm ? g.ss[i + nh] : 0);
  sab[i] = make_double2(a0, b0);
}

// Degree steps t of one segment group for R rings of this lane (slot = its segment row).
// HEADG: the group holds the first segment of its order (lanes with head = true start with
// P~_m, P~_{m+1}).  MFMAs of two consecutive steps are interleaved so that consecutive matrix
// instructions update different accumulators.
template <int R, bool HEADG>
__device__ __forceinline__ void ana_mf_steps(double (&acc)[MF_S], LegState (&st)[R], const double (&xr)[R],
                                             const double (&wv)[2][R], const double* qcs, int slot, bool head,
                                             double s23) {
  constexpr int S = MF_S;
  double cc[8];
#pragma unroll
  for (int t = 0; t < S; t += 2) {
    if ((t & 7) == 0) {
#pragma unroll
      for (int u = 0; u < 8; ++u) cc[u] = qcs[slot * MF_QS + t + u];
    }
    double pv[2][R];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int tt = t + h;
      const double c = cc[tt & 7];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double p;
        if (HEADG && tt == 0) {  // head: P_m itself, the state does not move
          const double pa0 = st[r].pa, pb0 = st[r].pb;
          const double pn = q_next(xr[r], c, st[r]);
          p = head ? pb0 : pn;
          st[r].pa = head ? pa0 : st[r].pa;
          st[r].pb = head ? pb0 : st[r].pb;
        } else if (HEADG && tt == 1) {  // head: P_{m+1} = sqrt(2m+3) x P_m
          const double ps = s23 * xr[r] * st[r].pb;
          const double pn = q_next(xr[r], c, st[r]);
          p = head ? ps : pn;
          st[r].pb = p;
        } else {
          p = q_next(xr[r], c, st[r]);
        }
        pv[h][r] = p;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      acc[t] = __builtin_amdgcn_mfma_f64_4x4x4f64(pv[0][r], wv[0][r], acc[t], 0, 0, 0);
      acc[t + 1] = __builtin_amdgcn_mfma_f64_4x4x4f64(pv[1][r], wv[1][r], acc[t + 1], 0, 0, 0);
    }
  }
}

template <int NF, int R>
__global__ void __launch_bounds__(512) k_leg_ana_mf(ShtDev d, SegDev g, GrpDev gr, const double2* __restrict__ WS,
                                                    const double2* __restrict__ WD, OutPtrs<NF> out, Pred pr) {
  static_assert(NF == 1 || NF == 2, "one or two fields per launch");
  if (!pr.active()) return;
  constexpr int S = MF_S;
  extern __shared__ double red[];  // [NW][8][64] wave partials
  __shared__ double qcs[4 * MF_QS];  // scaled recurrence coefficients of the 4 slots
  __shared__ double qgs[4 * MF_QS];  // and the factors P~ = qg Q
  const int per = (gr.ngrp + 7) >> 3;
  const int gi = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (gi >= gr.ngrp) return;
  const int4 ds = gr.desc[gi];  // (m, s0, seed row of segment s0, packed offset of order m)
  const int m = ds.x, s0 = ds.y;
  const int L = d.lmax, nh = d.nh, mo = ds.w;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, NW = blockDim.x >> 6;
  const int nsm = (L - m + S) / S;  // segments of order m
  for (int i = threadIdx.x; i < 4 * S; i += blockDim.x) {
    const int off = (s0 + i / S) * S + i % S;  // l - m
    const bool use = off <= L - m;
    qcs[(i / S) * MF_QS + i % S] = use ? d.qc[mo + off] : 0.0;
    qgs[(i / S) * MF_QS + i % S] = use ? d.qg[mo + off] : 0.0;
  }
  __syncthreads();
  const int blk = (lane >> 2) & 3, slot = lane & 3, kk = lane >> 4;
  const int segi = s0 + slot;
  const bool sval = segi < nsm;
  const bool head = sval && segi == 0;
  const long sx = (long)(ds.z + (sval ? slot : 0)) * nh;
  const long fstride = (long)d.nm * nh;
  const int col = slot;  // B column of this lane (same lane bits as its A row)
  const bool colok = (col >> 1) < NF;
  const int fb = colok ? (col >> 1) : 0;
  double xr[R], wv[2][R];
  LegState st[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = (wave * R + r) * 16 + 4 * blk + kk;  // ring of this lane (A and B roles alike)
    const bool ok = j < nh;
    const int jj = ok ? j : 0;
    xr[r] = d.x[jj];
    const double2 ab = (ok && sval) ? g.sab[sx + jj] : make_double2(0.0, 0.0);
    st[r] = LegState{ab.x, ab.y, 0};
    const double2 vs = WS[fb * fstride + (long)m * nh + jj], vd = WD[fb * fstride + (long)m * nh + jj];
    const bool use = ok && colok;
    wv[0][r] = use ? ((col & 1) ? vs.y : vs.x) : 0.0;
    wv[1][r] = use ? ((col & 1) ? vd.y : vd.x) : 0.0;
  }
  const double s23 = d.s2m3[m];
  double acc[S];
#pragma unroll
  for (int t = 0; t < S; ++t) acc[t] = 0.0;
  if (s0 == 0)
    ana_mf_steps<R, true>(acc, st, xr, wv, qcs, slot, head, s23);
  else
    ana_mf_steps<R, false>(acc, st, xr, wv, qcs, slot, head, s23);
  mf_block_reduce(acc, lane);
  auto emit = [&](int v, int ln, double val) {
    const int t = 16 * ((ln >> 3) & 1) + 8 * ((ln >> 2) & 1) + v;
    const int row = ln >> 4, cj = ln & 3;
    const int off = (s0 + row) * S + t;
    if ((s0 + row) < nsm && off <= L - m && (cj >> 1) < NF)
      reinterpret_cast<double*>(out.p[cj >> 1])[2 * (long)(mo + off) + (cj & 1)] =
          (val * qgs[row * MF_QS + t]) / SLX_SQRT2_LEGACY;
  };
  if (NW == 1) {
#pragma unroll
    for (int v = 0; v < 8; ++v) emit(v, lane, acc[v]);
    return;
  }
#pragma unroll
  for (int v = 0; v < 8; ++v) red[(wave * 8 + v) * 64 + lane] = acc[v];
  __syncthreads();
  for (int e = threadIdx.x; e < 512; e += blockDim.x) {
    const int v = e >> 6, ln = e & 63;
    double sum = red[v * 64 + ln];
    for (int w = 1; w < NW; ++w) sum += red[(w * 8 + v) * 64 + ln];
    emit(v, ln, sum);
  }
}

// Chunked-accumulator form of ana_mf_steps: 8 degree steps per chunk, reduced over the 4 MFMA
// blocks after each chunk and written to this wave's LDS partials (8 x 64 doubles); afterwards the
// lane of partial v holds t = 8 (v >> 1) + 4 ((lane >> 3) & 1) + 2 ((lane >> 2) & 1) + (v & 1).
// The MFMA accumulation of every t and the pairing of the reduction (lanes xor 8, then xor 4) are
// those of ana_mf_steps + mf_block_reduce, so the results are identical; 16 accumulator registers
// instead of 64.
template <int R, bool HEADG>
__device__ __forceinline__ void ana_ck_steps(LegState (&st)[R], const double (&xr)[R], const double (&wv)[2][R],
                                             const double* qcs, int slot, bool head, double s23, double* redw,
                                             int lane) {
  constexpr int S = MF_S;
#pragma unroll
  for (int c = 0; c < S / 8; ++c) {
    double acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = 0.0;
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
      double pv[2][R];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int tt = 8 * c + u + h;
        const double cq = qcs[slot * MF_QS + tt];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          double p;
          if (HEADG && tt == 0) {  // head: P_m itself, the state does not move
            const double pa0 = st[r].pa, pb0 = st[r].pb;
            const double pn = q_next(xr[r], cq, st[r]);
            p = head ? pb0 : pn;
            st[r].pa = head ? pa0 : st[r].pa;
            st[r].pb = head ? pb0 : st[r].pb;
          } else if (HEADG && tt == 1) {  // head: P_{m+1} = sqrt(2m+3) x P_m
            const double ps = s23 * xr[r] * st[r].pb;
            const double pn = q_next(xr[r], cq, st[r]);
            p = head ? ps : pn;
            st[r].pb = p;
          } else {
            p = q_next(xr[r], cq, st[r]);
          }
          pv[h][r] = p;
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        acc[u] = __builtin_amdgcn_mfma_f64_4x4x4f64(pv[0][r], wv[0][r], acc[u], 0, 0, 0);
        acc[u + 1] = __builtin_amdgcn_mfma_f64_4x4x4f64(pv[1][r], wv[1][r], acc[u + 1], 0, 0, 0);
      }
    }
    {
      const bool up = (lane & 8) != 0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double send = up ? acc[v] : acc[v + 4];
        const double keep = up ? acc[v + 4] : acc[v];
        acc[v] = keep + __shfl_xor(send, 8, 64);
      }
    }
    {
      const bool up = (lane & 4) != 0;
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const double send = up ? acc[v] : acc[v + 2];
        const double keep = up ? acc[v + 2] : acc[v];
        acc[v] = keep + __shfl_xor(send, 4, 64);
      }
    }
    redw[(2 * c) * 64 + lane] = acc[0];
    redw[(2 * c + 1) * 64 + lane] = acc[1];
  }
}

// Persistent form of k_leg_ana_mf: the grid holds as many blocks as can be resident (5 waves per
// SIMD with the chunked accumulators), each looping over the groups of its XCD's share in order,
// so no late round of a few leftover groups follows the resident ones.  A group's order, first
// segment, seed row and packed offset come from one descriptor load, so its seeds, inputs and
// tables are requested after a single dependent load.  Same arithmetic as k_leg_ana_mf.
template <int NF, int R>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_leg_ana_ck(ShtDev d, SegDev g, GrpDev gr, const double2* __restrict__ WS, const double2* __restrict__ WD,
             OutPtrs<NF> out, Pred pr) {
  static_assert(NF == 1 || NF == 2, "one or two fields per launch");
  if (!pr.active()) return;
  constexpr int S = MF_S;
  extern __shared__ double red[];  // [NW][8][64] wave partials
  __shared__ double qcs[4 * MF_QS];  // scaled recurrence coefficients of the 4 slots
  __shared__ double qgs[4 * MF_QS];  // and the factors P~ = qg Q
  const int per = (gr.ngrp + 7) >> 3;
  const int x = blockIdx.x & 7, nbx = gridDim.x >> 3;
  const int gend = min(gr.ngrp, (x + 1) * per);
  const int L = d.lmax, nh = d.nh;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, NW = blockDim.x >> 6;
  const int blk = (lane >> 2) & 3, slot = lane & 3, kk = lane >> 4;
  const long fstride = (long)d.nm * nh;
  const int col = slot;  // B column of this lane (same lane bits as its A row)
  const bool colok = (col >> 1) < NF;
  const int fb = colok ? (col >> 1) : 0;
  for (int gi = x * per + (blockIdx.x >> 3); gi < gend; gi += nbx) {
    const int4 ds = gr.desc[gi];
    const int m = ds.x, s0 = ds.y, mo = ds.w;
    const int nsm = (L - m + S) / S;  // segments of order m
    for (int i = threadIdx.x; i < 4 * S; i += blockDim.x) {
      const int off = (s0 + i / S) * S + i % S;  // l - m
      const bool use = off <= L - m;
      qcs[(i / S) * MF_QS + i % S] = use ? d.qc[mo + off] : 0.0;
      qgs[(i / S) * MF_QS + i % S] = use ? d.qg[mo + off] : 0.0;
    }
    const int segi = s0 + slot;
    const bool sval = segi < nsm;
    const bool head = sval && segi == 0;
    const long sx = (long)(ds.z + (sval ? slot : 0)) * nh;
    double xr[R], wv[2][R];
    LegState st[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = (wave * R + r) * 16 + 4 * blk + kk;  // ring of this lane (A and B roles alike)
      const bool ok = j < nh;
      const int jj = ok ? j : 0;
      xr[r] = d.x[jj];
      const double2 ab = (ok && sval) ? g.sab[sx + jj] : make_double2(0.0, 0.0);
      st[r] = LegState{ab.x, ab.y, 0};
      const double2 vs = WS[fb * fstride + (long)m * nh + jj], vd = WD[fb * fstride + (long)m * nh + jj];
      const bool use = ok && colok;
      wv[0][r] = use ? ((col & 1) ? vs.y : vs.x) : 0.0;
      wv[1][r] = use ? ((col & 1) ? vd.y : vd.x) : 0.0;
    }
    const double s23 = d.s2m3[m];
    __syncthreads();
    if (s0 == 0)
      ana_ck_steps<R, true>(st, xr, wv, qcs, slot, head, s23, red + wave * 8 * 64, lane);
    else
      ana_ck_steps<R, false>(st, xr, wv, qcs, slot, head, s23, red + wave * 8 * 64, lane);
    __syncthreads();
    for (int e = threadIdx.x; e < 512; e += blockDim.x) {
      const int v = e >> 6, ln = e & 63;
      double sum = red[v * 64 + ln];
      for (int w = 1; w < NW; ++w) sum += red[(w * 8 + v) * 64 + ln];
      const int t = 8 * (v >> 1) + 4 * ((ln >> 3) & 1) + 2 * ((ln >> 2) & 1) + (v & 1);
      const int row = ln >> 4, cj = ln & 3;
      const int off = (s0 + row) * S + t;
      if ((s0 + row) < nsm && off <= L - m && (cj >> 1) < NF)
        reinterpret_cast<double*>(out.p[cj >> 1])[2 * (long)(mo + off) + (cj & 1)] =
            (sum * qgs[row * MF_QS + t]) / SLX_SQRT2_LEGACY;
    }
    __syncthreads();  // the tables and partials are refilled by the next group
  }
}

// Ring chunk ch (16 R listed rings) of the sparse analysis: ring values, start states of this
// lane's segment (zero past the order's last segment), folded inputs of field 0.
template <int R>
__device__ __forceinline__ void sp_chunk_loads(const ShtDev& d, const SegDev& g, const double2* __restrict__ WS,
                                               const double2* __restrict__ WD, const int* list, int nact, int ch,
                                               long sx, bool sval, int m, int lane, double (&xr)[R],
                                               LegState (&st)[R], double (&wv)[2][R]) {
  const int blk = (lane >> 2) & 3, slot = lane & 3, kk = lane >> 4;
  const int nh = d.nh;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int jl = ch * 16 * R + r * 16 + 4 * blk + kk;
    const bool ok = jl < nact;
    const int jj = ok ? list[jl] : 0;
    xr[r] = d.x[jj];
    const double2 ab = (ok && sval) ? g.sab[sx + jj] : make_double2(0.0, 0.0);
    st[r] = LegState{ab.x, ab.y, 0};
    const double2 vs = WS[(long)m * nh + jj], vd = WD[(long)m * nh + jj];
    const bool use = ok && slot < 2;
    wv[0][r] = use ? ((slot & 1) ? vs.y : vs.x) : 0.0;
    wv[1][r] = use ? ((slot & 1) ? vd.y : vd.x) : 0.0;
  }
}

// Incremental analysis of one field over a sparse set of rings: out += the analysis of the folded
// spectra in WS/WD field 0 restricted to the rings whose ring_tok equals token (a grounded re-guess
// changes I* at a few grid points; its ring pass writes those rings' spectra of the change).  One
// wave per block, persistent over the groups in the XCD-aware order of k_leg_ana_ck; the listed
// rings (ascending) are walked 16 R at a time, their partial sums added in that order, and each
// coefficient is updated by the one lane that emits it.  Same degree steps as k_leg_ana_ck.
template <int R>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_leg_ana_sp(ShtDev d, SegDev g, GrpDev gr, const double2* __restrict__ WS, const double2* __restrict__ WD,
             double2* __restrict__ out, const int* __restrict__ ring_tok, int token, Pred pr) {
  if (!pr.active()) return;
  constexpr int S = MF_S;
  __shared__ double redw[8 * 64];  // the wave's partials of one ring chunk
  __shared__ double qcs[4 * MF_QS];
  __shared__ double qgs[4 * MF_QS];
  __shared__ int list[513];  // nh <= 512 on the matrix-core path, + a slot for unlisted lanes
  const int L = d.lmax, nh = d.nh, lane = threadIdx.x;
  // every ring's token and this wave's first group descriptor are loaded first (one round trip),
  // then the tagged rings are listed
  const int per = (gr.ngrp + 7) >> 3;
  const int x = blockIdx.x & 7, nbx = gridDim.x >> 3;
  const int gend = min(gr.ngrp, (x + 1) * per);
  int gi = x * per + (blockIdx.x >> 3);
  if (gi >= gend) return;
  int4 ds = gr.desc[gi];
  int tk[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) tk[u] = ring_tok[min(u * 64 + lane, nh - 1)];
  int nact = 0;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const bool on = u * 64 + lane < nh && tk[u] == token;
    const unsigned long long bal = __ballot(on);
    list[on ? nact + __popcll(bal & ((1ull << lane) - 1ull)) : 512] = u * 64 + lane;  // no branch
    nact += __popcll(bal);
  }
  // (the descriptor takes part in the test, m <= L always, so that its load is not sunk below it)
  if (nact == 0 || ds.x > L) return;
  __syncthreads();
  const int slot = lane & 3;
  const int nchunk = (nact + 16 * R - 1) / (16 * R);
  const int row = lane >> 4, cj = lane & 3;
  int m = ds.x, s0 = ds.y, sz = ds.z, mo = ds.w;  // carried as scalars (an int4 went to the stack)
  for (; gi < gend; gi += nbx) {
    const int nsm = (L - m + S) / S;
    // everything this group reads is requested before the first wait: the tables, the first ring
    // chunk's values, the coefficients it adds to (no other block writes them) and the next
    // group's descriptor
    static_assert(4 * S == 128, "two table entries per lane");
    const int offa = s0 * S + lane, offb = offa + 64;  // entries lane and lane + 64 of [4][S]
    const double tqa = offa <= L - m ? d.qc[mo + offa] : 0.0, tga = offa <= L - m ? d.qg[mo + offa] : 0.0;
    const double tqb = offb <= L - m ? d.qc[mo + offb] : 0.0, tgb = offb <= L - m ? d.qg[mo + offb] : 0.0;
    const int segi = s0 + slot;
    const bool sval = segi < nsm;
    const bool head = sval && segi == 0;
    const long sx = (long)(sz + (sval ? slot : 0)) * nh;
    double xr[R], wv[2][R];
    LegState st[R];
    sp_chunk_loads<R>(d, g, WS, WD, list, nact, 0, sx, sval, m, lane, xr, st, wv);
    // coefficient of emitted value v of this lane: l - m = (s0 + row) S + t_v, written when it exists
    const int t0l = 4 * ((lane >> 3) & 1) + 2 * ((lane >> 2) & 1);
    const int offl = (s0 + row) * S + t0l;
    const bool rowok = (s0 + row) < nsm && cj < 2;
    double* const ob = reinterpret_cast<double*>(out) + 2 * (long)(mo + offl) + cj;
    double ov[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const int dt = 8 * (v >> 1) + (v & 1);
      ov[v] = (rowok && offl + dt <= L - m) ? ob[2 * dt] : 0.0;
    }
    const int4 dn = gr.desc[gi + nbx < gend ? gi + nbx : gi];
    const double s23 = d.s2m3[m];
    qcs[(lane / S) * MF_QS + lane % S] = tqa;
    qgs[(lane / S) * MF_QS + lane % S] = tga;
    qcs[(lane / S + 2) * MF_QS + lane % S] = tqb;
    qgs[(lane / S + 2) * MF_QS + lane % S] = tgb;
    __syncthreads();
    double accv[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) accv[v] = 0.0;
    for (int ch = 0; ch < nchunk; ++ch) {
      if (ch > 0) sp_chunk_loads<R>(d, g, WS, WD, list, nact, ch, sx, sval, m, lane, xr, st, wv);
      if (s0 == 0)
        ana_ck_steps<R, true>(st, xr, wv, qcs, slot, head, s23, redw, lane);
      else
        ana_ck_steps<R, false>(st, xr, wv, qcs, slot, head, s23, redw, lane);
#pragma unroll
      for (int v = 0; v < 8; ++v) accv[v] += redw[v * 64 + lane];  // this lane's own slots
    }
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const int dt = 8 * (v >> 1) + (v & 1);
      if (rowok && offl + dt <= L - m) ob[2 * dt] = ov[v] + (accv[v] * qgs[row * MF_QS + t0l + dt]) / SLX_SQRT2_LEGACY;
    }
    __syncthreads();  // the tables are refilled by the next group
    m = dn.x;
    s0 = dn.y;
    sz = dn.z;
    mo = dn.w;
  }
}

// -------------------------------------------------------------------------------------------
// Synthesis contraction with wave-uniform degree streams.  Block = (order pair m1 = p, m2 = L - p)
// x 64 northern rings, 4 waves.  The chains (Ssyn-degree segments, seeded on the device) of the
// two orders are dealt to the waves round-robin, two at a time per wave; every lane owns one ring.
// All lanes of a wave step through the same degree, so the recurrence coefficients and the
// spectral coefficients are uniform and come through scalar loads; past a chain's end both read
// as zero.  E (even l - m) and O (odd) are summed per chain, per wave, then over the 4 waves in
// LDS in a fixed order; FN = 1.41421356 (E + O) (north), FS = 1.41421356 (E - O) (south),
// written [f][m][j].  Coefficients: cf[f cstride + packed index].
template <int NF, int CQ, bool FIRST>
__device__ __forceinline__ void syn_step8(LegState (&st)[CQ], double (&mk)[CQ], const bool (&head)[CQ],
                                          const int (&gi)[CQ], const int (&cnt)[CQ], int t0, double x,
                                          const double (&s23)[CQ], const double2* __restrict__ rec,
                                          const double2* __restrict__ cf, long cstride, double2 (&Ec)[CQ][NF],
                                          double2 (&Oc)[CQ][NF]) {
#pragma unroll
  for (int il = 0; il < 8; ++il) {
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int t = t0 + il;
      const bool v = t < cnt[q];
      const int idx = v ? gi[q] + t : 0;
      const double2 c = v ? rec[idx] : make_double2(0.0, 0.0);
      double p;
      if (FIRST && il == 0) {
        const double pa0 = st[q].pa, pb0 = st[q].pb;
        const double pn = leg_next(x, c, st[q]);
        p = head[q] ? pb0 : pn;
        st[q].pa = head[q] ? pa0 : st[q].pa;
        st[q].pb = head[q] ? pb0 : st[q].pb;
      } else if (FIRST && il == 1) {
        const double ps = s23[q] * x * st[q].pb;
        const double pn = leg_next(x, c, st[q]);
        p = head[q] ? ps : pn;
        st[q].pb = p;
      } else {
        p = leg_next(x, c, st[q]);
      }
      p *= mk[q];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double2 cv = v ? cf[f * cstride + idx] : make_double2(0.0, 0.0);
        if ((il & 1) == 0) {
          Ec[q][f].x += p * cv.x;
          Ec[q][f].y += p * cv.y;
        } else {
          Oc[q][f].x += p * cv.x;
          Oc[q][f].y += p * cv.y;
        }
      }
      if (il == 7) {
        leg_renorm(st[q]);
        mk[q] = st[q].sc == 0 ? 1.0 : 0.0;
      }
    }
  }
}

template <int NF, int CQ>
__global__ void __launch_bounds__(256) k_leg_syn_s(ShtDev d, SegDev g, const double2* __restrict__ rec,
                                                   const double2* __restrict__ cf, long cstride,
                                                   double2* __restrict__ FN, double2* __restrict__ FS, Pred pr) {
  if (!pr.active()) return;
  __shared__ double2 part[4][2][NF][2][64];
  const int L = d.lmax, nl = L + 1, S = g.Ssyn, ratio = g.Ssyn / g.S, nh = d.nh;
  const int p = blockIdx.x;
  const int mord[2] = {p, L - p};
  const int nord = (L - p == p) ? 1 : 2;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int j = blockIdx.y * 64 + lane;
  const bool ok = j < nh;
  const int jj = ok ? j : 0;
  const double x = d.x[jj];
  const double s23[2] = {d.s2m3[mord[0]], d.s2m3[mord[1]]};
  const int ns0 = (nl - mord[0] + S - 1) / S;
  const int ns1 = nord == 2 ? (nl - mord[1] + S - 1) / S : 0;
  const int C = ns0 + ns1;
  double2 E[2][NF], O[2][NF];
#pragma unroll
  for (int o = 0; o < 2; ++o)
#pragma unroll
    for (int f = 0; f < NF; ++f) E[o][f] = O[o][f] = make_double2(0.0, 0.0);
  for (int c0 = CQ * wave; c0 < C; c0 += 4 * CQ) {
    LegState st[CQ];
    double mk[CQ], s23c[CQ];
    bool head[CQ];
    int gi[CQ], cnt[CQ], oo[CQ];
    double2 Ec[CQ][NF], Oc[CQ][NF];
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int c = c0 + q;
      const bool act = c < C;
      const int o = (c < ns0) ? 0 : 1;
      const int s = (c < ns0) ? c : c - ns0;
      const int m = mord[o];
      oo[q] = o;
      s23c[q] = s23[o];
      cnt[q] = act ? min(S, nl - m - s * S) : 0;
      gi[q] = d.moff[m] + s * S;
      head[q] = act && s == 0;
      const long ix = (long)m * nh + jj;
      const long sx = (long)(g.first[m] + s * ratio) * nh + jj;
      if (act && s == 0) {
        st[q] = LegState{0.0, d.pmm[ix], d.pms[ix]};
      } else if (act) {
        st[q] = LegState{g.sa[sx], g.sb[sx], g.ss[sx]};
      } else {
        st[q] = LegState{0.0, 0.0, 0};
      }
      if (!ok) st[q] = LegState{0.0, 0.0, 0};
      mk[q] = st[q].sc == 0 ? 1.0 : 0.0;
#pragma unroll
      for (int f = 0; f < NF; ++f) Ec[q][f] = Oc[q][f] = make_double2(0.0, 0.0);
    }
    int tmax = 0;
#pragma unroll
    for (int q = 0; q < CQ; ++q) tmax = max(tmax, cnt[q]);
    syn_step8<NF, CQ, true>(st, mk, head, gi, cnt, 0, x, s23c, rec, cf, cstride, Ec, Oc);
    for (int t0 = 8; t0 < tmax; t0 += 8)
      syn_step8<NF, CQ, false>(st, mk, head, gi, cnt, t0, x, s23c, rec, cf, cstride, Ec, Oc);
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        if (oo[q] == 0) {
          E[0][f] = cadd(E[0][f], Ec[q][f]);
          O[0][f] = cadd(O[0][f], Oc[q][f]);
        } else {
          E[1][f] = cadd(E[1][f], Ec[q][f]);
          O[1][f] = cadd(O[1][f], Oc[q][f]);
        }
      }
    }
  }
#pragma unroll
  for (int o = 0; o < 2; ++o)
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      part[wave][o][f][0][lane] = E[o][f];
      part[wave][o][f][1][lane] = O[o][f];
    }
  __syncthreads();
  if (wave >= nord || !ok) return;
  const int o = wave, m = mord[o];
  const long fstride = (long)d.nm * nh;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double2 e = part[0][o][f][0][lane], q = part[0][o][f][1][lane];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      e = cadd(e, part[w][o][f][0][lane]);
      q = cadd(q, part[w][o][f][1][lane]);
    }
    const long idx = f * fstride + (long)m * nh + j;
    FN[idx] = make_double2(SLX_SQRT2_LEGACY * (e.x + q.x), SLX_SQRT2_LEGACY * (e.y + q.y));
    FS[idx] = make_double2(SLX_SQRT2_LEGACY * (e.x - q.x), SLX_SQRT2_LEGACY * (e.y - q.y));
  }
}

// Spectral coefficients assembled by a loader (Coef: prologue(double*), run by the blocks for
// which prologue_needed(first index, end index) holds, then double2 operator()(f, idx, l, m, pro))
// into a plain packed array out[f ostride + idx].
template <int NF, class Coef>
__global__ void __launch_bounds__(256) k_coef_fill(ShtDev d, Coef cf, double2* __restrict__ out, long ostride,
                                                   int ncoef, Pred pr) {
  if (!pr.active()) return;
  __shared__ double pro[32];
  if (cf.prologue_needed(blockIdx.x * 256, blockIdx.x * 256 + 256)) {  // block-uniform
    cf.prologue(pro);
    __syncthreads();
  }
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= ncoef) return;
  const int l = d.l_of[idx], m = d.m_of[idx];
#pragma unroll
  for (int f = 0; f < NF; ++f) out[f * ostride + idx] = cf(f, idx, l, m, pro);
}


// -------------------------------------------------------------------------------------------
// Synthesis contraction on the matrix cores (v_mfma_f64_4x4x4_4b_f64, lane roles as for the
// analysis above): rows i = four northern rings, K index k = four degree slots (32-degree
// segments 4g .. 4g + 3 of one order), columns j = field x (re, im).  A lane walks the
// recurrence of its ring through its slot; at step t the K index runs over the degrees
// l0_k + t, all of the parity of t, so even and odd steps accumulate into separate E and O
// outputs.  Block = (order pair of syn_pair) x 16 R rings (R rings per lane, 4 waves; a 1-D grid dealt
// so that the ring chunks of one order pair share an XCD); the segment groups of both orders are
// dealt to the waves, whose E/O partials are summed in LDS in a fixed
// order.  Recurrence and spectral coefficients of both orders are staged in LDS, one order after
// the other in groups of 4 S degrees, zero past each order's last degree.
// Outputs FN = 1.41421356 (E + O), FS = 1.41421356 (E - O), [f][m][j].
// Degree steps of one segment group (synthesis).  Parities of t go to separate accumulators
// aE (even l - m) and aO (odd); consecutive MFMAs alternate between them.
template <int R, bool HEADG>
__device__ __forceinline__ void syn_mf_steps(double (&aE)[R], double (&aO)[R], LegState (&st)[R],
                                             const double (&xr)[R], const double* qco, const double* clo,
                                             bool head, double s23) {
  constexpr int S = MF_S;
  double cc[8], bb[8];
#pragma unroll
  for (int t = 0; t < S; t += 2) {
    if ((t & 7) == 0) {  // staged rows are zero past the order's last degree
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        cc[u] = qco[t + u];
        bb[u] = clo[2 * (t + u)];  // a lane of an unused B column reads the zero row
      }
    }
    double pv[2][R];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int tt = t + h;
      const double c = cc[tt & 7];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double p;
        if (HEADG && tt == 0) {
          const double pa0 = st[r].pa, pb0 = st[r].pb;
          const double pn = q_next(xr[r], c, st[r]);
          p = head ? pb0 : pn;
          st[r].pa = head ? pa0 : st[r].pa;
          st[r].pb = head ? pb0 : st[r].pb;
        } else if (HEADG && tt == 1) {
          const double ps = s23 * xr[r] * st[r].pb;
          const double pn = q_next(xr[r], c, st[r]);
          p = head ? ps : pn;
          st[r].pb = p;
        } else {
          p = q_next(xr[r], c, st[r]);
        }
        pv[h][r] = p;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      aE[r] = __builtin_amdgcn_mfma_f64_4x4x4f64(pv[0][r], bb[t & 7], aE[r], 0, 0, 0);
      aO[r] = __builtin_amdgcn_mfma_f64_4x4x4f64(pv[1][r], bb[(t + 1) & 7], aO[r], 0, 0, 0);
    }
  }
}

// Order pairs of the synthesis blocks, p < (L + 2) / 2: (p, L - p) for odd L; for even L, (p, L - 1 - p)
// for p < L / 2 and order L alone (one degree) for p = L / 2, so that every pair but that one spans
// L + 2 or L + 3 degrees (pairing p with L - p would leave order L / 2, L / 2 + 1 degrees, alone in a
// block of half the others' work).  Returns the number of orders.
__device__ __forceinline__ int syn_pair(int L, int p, int (&mord)[2]) {
  if (L & 1) {
    mord[0] = p;
    mord[1] = L - p;
    return 2;
  }
  if (p == L / 2) {
    mord[0] = mord[1] = L;
    return 1;
  }
  mord[0] = p;
  mord[1] = L - 1 - p;
  return 2;
}

// Staging of an order pair for the synthesis contractions: the scaled-recurrence coefficients qc
// and the spectral coefficients times qg of order mord[0]'s groups of 4 S degrees, then order
// mord[1]'s from entry base1 on, zero past each order's last degree.  The coefficients come from
// the loader cf: CoefPlain reads an assembled array; an assembling loader (eq. B18, the adjoint's
// scaled coefficients) forms them here, after its block prologue where its index ranges need one
// (block-uniform), so that no separate fill kernel runs.  Four entries per thread are loaded
// before any is stored, so that their loads are in flight together.
template <int NF, int NT, class Coef>
__device__ __forceinline__ void syn_stage(const ShtDev& d, const Coef& cf, const int (&mord)[2], int nord, int np2,
                                          int base1, int nsg, double* __restrict__ qcl, double2* __restrict__ cl) {
  constexpr int S = MF_S, U = 4;
  __shared__ double pro[32];
  const int nl = d.lmax + 1;
  const int mo0 = d.moff[mord[0]], mo1 = d.moff[mord[1]];
  const int cnt0 = nl - mord[0], cnt1 = nl - mord[1];
  if (cf.prologue_needed(mo0, mo0 + cnt0) || (nord == 2 && cf.prologue_needed(mo1, mo1 + cnt1))) {
    cf.prologue(pro);
    __syncthreads();
  }
  constexpr int nt = NT;  // threads of the block
  for (int i0 = threadIdx.x; i0 < np2; i0 += nt * U) {
    double qv[U], gv[U];
    double2 cv[U][NF];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + nt * u;
      const bool second = i >= base1;
      const int k = second ? i - base1 : i;
      const bool v = i < np2 && k < (second ? cnt1 : cnt0);
      const int m = second ? mord[1] : mord[0];
      const int src = (second ? mo1 : mo0) + k;
      qv[u] = v ? d.qc[src] : 0.0;
      gv[u] = v ? d.qg[src] : 0.0;
#pragma unroll
      for (int f = 0; f < NF; ++f) cv[u][f] = v ? cf(f, src, m + k, m, pro) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + nt * u;
      if (i < np2) {
        const int sgi = i / S, tq = i % S;
        qcl[sgi * MF_QS + tq] = qv[u];
#pragma unroll
        for (int f = 0; f < NF; ++f) cl[(f * nsg + sgi) * MF_CS + tq] = make_double2(cv[u][f].x * gv[u], cv[u][f].y * gv[u]);
      }
    }
  }
}

// One field with 2 rings per lane is compiled for 5 waves per SIMD (91 VGPRs, no spills; the LDS of
// 5 blocks fits a CU): 11 % faster than at the 4 waves its natural 98 registers allow.
template <int NF, int R, class Coef>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((NF == 1 && R == 2) ? 5 : 1, 8)))
k_leg_syn_mf(ShtDev d, SegDev g, Coef cf, double2* __restrict__ FN, double2* __restrict__ FS, Pred pr) {
  static_assert(NF == 1 || NF == 2, "one or two fields per launch");
  if (!pr.active()) return;
  constexpr int S = MF_S;
  extern __shared__ double2 lsy[];
  const int L = d.lmax, nl = L + 1, nh = d.nh;
  // XCD-aware block order: block b runs on XCD b % 8; the ny ring chunks of one order pair go to
  // the same XCD, so that its L2 fetches the pair's staged coefficients once
  const int ny = (nh + 16 * R - 1) / (16 * R), bt = blockIdx.x >> 3;
  const int yb = bt % ny, p = (blockIdx.x & 7) + 8 * (bt / ny);
  if (p >= (L + 2) / 2) return;
  int mord[2];
  const int nord = syn_pair(L, p, mord);
  const int ng0 = ((nl - mord[0] + S - 1) / S + 3) / 4;
  const int ng1 = nord == 2 ? ((nl - mord[1] + S - 1) / S + 3) / 4 : 0;
  // staging, packed: order mord[0]'s ng0 groups of 4 S degrees, then order mord[1]'s ng1 groups,
  // zero past each order's last degree; np2 = 4 S (ng0 + ng1) <= 4 S (ceil((L + 3) / (4 S)) + 1)
  const int np2 = 4 * S * (ng0 + ng1), base1 = 4 * S * ng0;
  const int nsg = np2 / S;  // staged segments, each padded to MF_QS / MF_CS entries
  double* qcl = reinterpret_cast<double*>(lsy);  // [nsg][MF_QS] scaled recurrence coefficients
  double2* cl = lsy + nsg * (MF_QS / 2);         // [NF][nsg][MF_CS] coefficients times qg
  double2* zr = cl + NF * nsg * MF_CS;            // [MF_CS] zeros: the B operand of unused columns
  double* part = reinterpret_cast<double*>(lsy);  // [4][2][2][R][64], reuses the staging area
  if (threadIdx.x < MF_CS) zr[threadIdx.x] = make_double2(0.0, 0.0);
  syn_stage<NF, 256>(d, cf, mord, nord, np2, base1, nsg, qcl, cl);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int blk = (lane >> 2) & 3, ri = lane & 3, slot = lane >> 4;
  const int col = lane & 3, fb = col >> 1, comp = col & 1;
  const bool colok = fb < NF;
  int jr[R];
  bool okr[R];
  double xr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    jr[r] = yb * 16 * R + r * 16 + 4 * blk + ri;
    okr[r] = jr[r] < nh;
    xr[r] = d.x[okr[r] ? jr[r] : 0];
  }
  double EO[2][2][R];  // [order][E, O][ring]
#pragma unroll
  for (int o = 0; o < 2; ++o)
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int r = 0; r < R; ++r) EO[o][e][r] = 0.0;
  const int first0 = g.first[mord[0]], first1 = g.first[mord[1]];
  const double s23o0 = d.s2m3[mord[0]], s23o1 = d.s2m3[mord[1]];
  for (int it = wave; it < ng0 + ng1; it += 4) {
    const int o = it < ng0 ? 0 : 1;
    const int gq = it < ng0 ? it : it - ng0;
    const int m = mord[o];
    const int nsm = (nl - m + S - 1) / S;
    const int sg = 4 * gq + slot;
    const bool sval = sg < nsm;
    const bool head = sval && sg == 0;
    const long sx = (long)((o ? first1 : first0) + (sval ? sg : 0)) * nh;
    LegState st[R];
    double aE[R], aO[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int jj = okr[r] ? jr[r] : 0;
      const double2 ab = (okr[r] && sval) ? g.sab[sx + jj] : make_double2(0.0, 0.0);
      st[r] = LegState{ab.x, ab.y, 0};
      aE[r] = aO[r] = 0.0;
    }
    const double s23 = o ? s23o1 : s23o0;
    const int ob = (o ? base1 : 0) / S + sg;  // staged segment
    const double* qco = qcl + ob * MF_QS;
    const double* clo = colok ? reinterpret_cast<const double*>(cl + (fb * nsg + ob) * MF_CS) + comp
                               : reinterpret_cast<const double*>(zr);
    if (gq == 0)
      syn_mf_steps<R, true>(aE, aO, st, xr, qco, clo, head, s23);
    else
      syn_mf_steps<R, false>(aE, aO, st, xr, qco, clo, head, s23);
    if (o == 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        EO[0][0][r] += aE[r];
        EO[0][1][r] += aO[r];
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        EO[1][0][r] += aE[r];
        EO[1][1][r] += aO[r];
      }
    }
  }
  __syncthreads();  // all waves are done with the staged coefficients
#pragma unroll
  for (int o = 0; o < 2; ++o)
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int r = 0; r < R; ++r) part[(((wave * 2 + o) * 2 + e) * R + r) * 64 + lane] = EO[o][e][r];
  __syncthreads();
  if (wave >= nord) return;
  const int o = wave, m = mord[o];
  if (!colok) return;
  const long fstride = (long)d.nm * nh;
  const int rowr = lane >> 4;  // ring row of the D element this lane holds
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double e = part[(((0 * 2 + o) * 2 + 0) * R + r) * 64 + lane];
    double q = part[(((0 * 2 + o) * 2 + 1) * R + r) * 64 + lane];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      e += part[(((w * 2 + o) * 2 + 0) * R + r) * 64 + lane];
      q += part[(((w * 2 + o) * 2 + 1) * R + r) * 64 + lane];
    }
    const int j = yb * 16 * R + r * 16 + 4 * blk + rowr;
    if (j < nh) {
      const long idx = fb * fstride + (long)m * nh + j;
      reinterpret_cast<double*>(FN)[2 * idx + comp] = SLX_SQRT2_LEGACY * (e + q);
      reinterpret_cast<double*>(FS)[2 * idx + comp] = SLX_SQRT2_LEGACY * (e - q);
    }
  }
}


// Synthesis contraction with the waves over rings: block = (order pair, chunk of 16 NWV R rings), wave
// w owns rings 16 R w .. 16 R w + 16 R - 1 of the chunk and walks every segment group of the pair
// (the same degree steps and staged coefficients as k_leg_syn_mf), so the waves carry equal work
// whatever the pair's group count, and each writes its own rings' FN, FS with no cross-wave sum.
template <int NF, int R, int NWV, class Coef>
__global__ void __launch_bounds__(64 * NWV) __attribute__((amdgpu_waves_per_eu(R <= 2 ? 4 : 1, 8)))
k_leg_syn_mw(ShtDev d, SegDev g, Coef cf, double2* __restrict__ FN, double2* __restrict__ FS, Pred pr) {
  static_assert(NF == 1 || NF == 2, "one or two fields per launch");
  if (!pr.active()) return;
  constexpr int S = MF_S;
  extern __shared__ double2 lsy[];
  const int L = d.lmax, nl = L + 1, nh = d.nh;
  // XCD-aware block order: block b runs on XCD b % 8; the ny ring chunks of one order pair go to
  // the same XCD, so that its L2 fetches the pair's staged coefficients once
  const int ny = (nh + 16 * NWV * R - 1) / (16 * NWV * R), bt = blockIdx.x >> 3;
  const int yb = bt % ny, p = (blockIdx.x & 7) + 8 * (bt / ny);
  if (p >= (L + 2) / 2) return;
  int mord[2];
  const int nord = syn_pair(L, p, mord);
  const int ng0 = ((nl - mord[0] + S - 1) / S + 3) / 4;
  const int ng1 = nord == 2 ? ((nl - mord[1] + S - 1) / S + 3) / 4 : 0;
  const int np2 = 4 * S * (ng0 + ng1), base1 = 4 * S * ng0;
  const int nsg = np2 / S;
  double* qcl = reinterpret_cast<double*>(lsy);  // [nsg][MF_QS] scaled recurrence coefficients
  double2* cl = lsy + nsg * (MF_QS / 2);         // [NF][nsg][MF_CS] coefficients times qg
  double2* zr = cl + NF * nsg * MF_CS;            // [MF_CS] zeros: the B operand of unused columns
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int blk = (lane >> 2) & 3, ri = lane & 3, slot = lane >> 4;
  const int col = lane & 3, fb = col >> 1, comp = col & 1;
  const bool colok = fb < NF;
  const int jw = yb * 16 * NWV * R + wave * 16 * R;  // this wave's first ring
  int jr[R];
  bool okr[R];
  double xr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    jr[r] = jw + r * 16 + 4 * blk + ri;
    okr[r] = jr[r] < nh;
    xr[r] = d.x[okr[r] ? jr[r] : 0];
  }
  const int first0 = g.first[mord[0]], first1 = g.first[mord[1]];
  // start states of segment group gq of order o (zero for lanes past the last ring or the order's
  // last segment); the next group's are loaded while the current group's degree steps run, the
  // first group's while the coefficients are staged
  const int nsm0 = (nl - mord[0] + S - 1) / S, nsm1 = (nl - mord[1] + S - 1) / S;
  const double2* __restrict__ sab = g.sab;
  auto seeds = [=](int o, int gq, double2 (&ab)[R]) {
    const int sg = 4 * gq + slot;
    const bool sval = sg < (o ? nsm1 : nsm0);
    const long sx = (long)((o ? first1 : first0) + (sval ? sg : 0)) * nh;
#pragma unroll
    for (int r = 0; r < R; ++r) ab[r] = (okr[r] && sval) ? sab[sx + (okr[r] ? jr[r] : 0)] : make_double2(0.0, 0.0);
  };
  double2 nab[R];
  seeds(0, 0, nab);
  if (threadIdx.x < MF_CS) zr[threadIdx.x] = make_double2(0.0, 0.0);
  syn_stage<NF, 64 * NWV>(d, cf, mord, nord, np2, base1, nsg, qcl, cl);
  __syncthreads();
  if (jw >= nh) return;  // a wave past the last ring (no barrier follows)
  const double s23o0 = d.s2m3[mord[0]], s23o1 = d.s2m3[mord[1]];
  const long fstride = (long)d.nm * nh;
  const int rowr = lane >> 4;  // ring row of the D element this lane holds
  for (int o = 0; o < nord; ++o) {
    const int m = mord[o], ng = o ? ng1 : ng0;
    const int nsm = (nl - m + S - 1) / S;
    const double s23 = o ? s23o1 : s23o0;
    double E[R], O[R];
#pragma unroll
    for (int r = 0; r < R; ++r) E[r] = O[r] = 0.0;
    for (int gq = 0; gq < ng; ++gq) {
      const int sg = 4 * gq + slot;
      const bool sval = sg < nsm;
      const bool head = sval && sg == 0;
      LegState st[R];
      double aE[R], aO[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        st[r] = LegState{nab[r].x, nab[r].y, 0};
        aE[r] = aO[r] = 0.0;
      }
      if (gq + 1 < ng)
        seeds(o, gq + 1, nab);
      else if (o + 1 < nord)
        seeds(o + 1, 0, nab);
      const int ob = (o ? base1 : 0) / S + sg;  // staged segment
      const double* qco = qcl + ob * MF_QS;
      const double* clo = colok ? reinterpret_cast<const double*>(cl + (fb * nsg + ob) * MF_CS) + comp
                                 : reinterpret_cast<const double*>(zr);
      if (gq == 0)
        syn_mf_steps<R, true>(aE, aO, st, xr, qco, clo, head, s23);
      else
        syn_mf_steps<R, false>(aE, aO, st, xr, qco, clo, head, s23);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        E[r] += aE[r];
        O[r] += aO[r];
      }
    }
    if (colok) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = jw + r * 16 + 4 * blk + rowr;
        if (j < nh) {
          const long idx = fb * fstride + (long)m * nh + j;
          reinterpret_cast<double*>(FN)[2 * idx + comp] = SLX_SQRT2_LEGACY * (E[r] + O[r]);
          reinterpret_cast<double*>(FS)[2 * idx + comp] = SLX_SQRT2_LEGACY * (E[r] - O[r]);
        }
      }
    }
  }
}

}  // namespace slx
