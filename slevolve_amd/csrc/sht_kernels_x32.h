// Reduced-precision Legendre contractions: the FP64 degree recurrence of sht_kernels.h (same
// seeds, same scaled recurrence, same head handling) feeds FP32 matrix-core products
// (v_mfma_f32_4x4x1_16b_f32) accumulated in FP32.  Used for the early sweeps of the inner fixed
// point (the iteration refines them in its final FP64 sweeps) and by a transform handle created
// with precision 1.  The folded Fourier inputs and the outputs stay FP64 in HBM; the products'
// B operands carry an exact power-of-two scale (2^64 for the analysis inputs, 2^-64 for the
// synthesis coefficients times qg, whose largest factor is 2^108.5 at lmax 1023), so that the
// products of the scaled recurrence values (down to ~1e-31 at the head of high orders) stay in the
// normal FP32 range.
//
// Lane roles of v_mfma_f32_4x4x1_16b_f32 (16 independent 4x4x1 blocks; established with an
// in-bounds one-hot probe): block b = lane >> 2; A[i] of block b from lane 4b + i, B[j] of block b
// from lane 4b + j, D[i][j] of block b in register i of lane 4b + j.
#pragma once
#include "sht_kernels.h"

namespace slx {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr double X32_ANA_SCALE = 0x1p64;   // analysis B operands (folded Fourier inputs) x 2^64
constexpr double X32_ANA_UNSCALE = 0x1p-64;
constexpr double X32_SYN_SCALE = 0x1p-64;  // synthesis B operands (coefficients x qg) x 2^-64
constexpr double X32_SYN_UNSCALE = 0x1p64;

__device__ __forceinline__ f32x4 mfma_x32(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}
// a over lane halves (bit 5): lanes 0..31 end with sum(a), lanes 32..63 with sum(b)
__device__ __forceinline__ float swap_add32f(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// the same across 16-lane rows (bit 4)
__device__ __forceinline__ float swap_add16f(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// One degree step of the recurrence for a lane that may hold the head of its order (first two
// steps of the order's first segment): the arithmetic of ana_ck_steps / syn_mf_steps with the head
// decided per lane.
__device__ __forceinline__ double x32_step(double x, double c, LegState& s, int tt, bool head, double s23) {
  if (tt == 0) {
    const double pa0 = s.pa, pb0 = s.pb;
    const double pn = q_next(x, c, s);
    if (head) {
      s.pa = pa0;
      s.pb = pb0;
      return pb0;
    }
    return pn;
  }
  if (tt == 1) {
    const double ps = s23 * x * s.pb;
    const double pn = q_next(x, c, s);
    const double p = head ? ps : pn;
    s.pb = p;
    return p;
  }
  return q_next(x, c, s);
}

// -------------------------------------------------------------------------------------------
// Analysis.  A block takes a super-group of four consecutive segment groups (gi = 4 sg + q, each
// 4 segments of one order, as the FP64 kernels' groups) and all rings.  Lane (rg, q, i) =
// (lane >> 4, (lane >> 2) & 3, lane & 3): MFMA block b = 4 rg + q; its A row i is segment i of group
// q (the lane walks that segment's recurrence), its B column i is field i >> 1, (re, im) = i & 1 of
// the folded input of group q's order; block (rg, q) takes ring 4 k + rg of ring quad k, so each
// MFMA covers 16 segments x 4 rings and the matrix core sums nothing.  Wave w walks the R ring quads
// k = r NW + w (R chains per lane); every degree step t accumulates the R products in its own
// f32x4 (row = segment, lane = (rg, q, column)).  After each chunk of 8 steps the four ring groups
// (lane bits 4, 5) are summed with two half/row swaps into the wave's LDS partials; at the end the
// waves are summed in LDS in a fixed order and each result is scaled by qg / 1.41421356 / 2^64 in
// FP64.
constexpr int X32_NW = 8;  // waves per analysis block
template <int NF, int R>
__global__ void __launch_bounds__(64 * X32_NW) k_leg_ana_x32(ShtDev d, SegDev g, GrpDev gr,
                                                            const double2* __restrict__ WS,
                                                            const double2* __restrict__ WD, OutPtrs<NF> out,
                                                            Pred pr) {
  static_assert(NF == 1 || NF == 2, "one or two fields per launch");
  if (!pr.active()) return;
  constexpr int S = MF_S, NW = X32_NW;
  __shared__ double qcs[16 * S];  // scaled recurrence coefficients of the 16 segments
  __shared__ double qgs[16 * S];  // and the factors P~ = qg Q
  __shared__ float red[NW][S][64];
  __shared__ int4 dsc[4];
  const int nsg = (gr.ngrp + 3) >> 2;
  const int per = (nsg + 7) >> 3;
  const int sgi = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (sgi >= nsg) return;
  const int L = d.lmax, nh = d.nh, tid = threadIdx.x;
  if (tid < 4) {
    const int gi = 4 * sgi + tid;
    dsc[tid] = gi < gr.ngrp ? gr.desc[gi] : make_int4(-1, 0, 0, 0);
  }
  __syncthreads();
  for (int e = tid; e < 16 * S; e += 64 * NW) {
    const int q = e / (4 * S), sl = (e / S) & 3, t = e % S;
    const int4 ds = dsc[q];
    const int off = (ds.y + sl) * S + t;
    const bool use = ds.x >= 0 && off <= L - ds.x;
    qcs[e] = use ? d.qc[ds.w + off] : 0.0;
    qgs[e] = use ? d.qg[ds.w + off] : 0.0;
  }
  const int lane = tid & 63, wave = tid >> 6;
  const int rg = lane >> 4, q = (lane >> 2) & 3, i = lane & 3;
  const int4 ds = dsc[q];
  const bool gval = ds.x >= 0;
  const int m = gval ? ds.x : 0, s0 = ds.y;
  const int nsm = gval ? (L - m + S) / S : 0;
  const int segi = s0 + i;
  const bool sval = segi < nsm;
  const bool head = sval && segi == 0;
  const long sx = (long)(ds.z + (sval ? i : 0)) * nh;
  const long fstride = (long)d.nm * nh;
  const bool colok = gval && (i >> 1) < NF;
  const long wbase = (colok ? (i >> 1) : 0) * fstride + (long)m * nh;
  const double s23 = d.s2m3[m];
  double xr[R];
  float ws[R], wd[R];
  LegState st[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = 4 * (r * NW + wave) + rg;
    const bool ok = j < nh;
    const int jj = ok ? j : 0;
    xr[r] = d.x[jj];
    const double2 ab = (ok && sval) ? g.sab[sx + jj] : make_double2(0.0, 0.0);
    st[r] = LegState{ab.x, ab.y, 0};
    const double2 vs = WS[wbase + jj], vd = WD[wbase + jj];
    const bool use = ok && colok;
    ws[r] = use ? (float)(((i & 1) ? vs.y : vs.x) * X32_ANA_SCALE) : 0.f;
    wd[r] = use ? (float)(((i & 1) ? vd.y : vd.x) * X32_ANA_SCALE) : 0.f;
  }
  const double* qcl = qcs + (q * 4 + i) * S;
  __syncthreads();
#pragma unroll
  for (int c = 0; c < S / 8; ++c) {
    f32x4 acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
      const int t = 8 * c + u;
      const double c0 = qcl[t], c1 = qcl[t + 1];
      float pv[2][R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (t == 0) {
          pv[0][r] = (float)x32_step(xr[r], c0, st[r], 0, head, s23);
          pv[1][r] = (float)x32_step(xr[r], c1, st[r], 1, head, s23);
        } else {
          pv[0][r] = (float)q_next(xr[r], c0, st[r]);
          pv[1][r] = (float)q_next(xr[r], c1, st[r]);
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        acc[u] = mfma_x32(pv[0][r], ws[r], acc[u]);
        acc[u + 1] = mfma_x32(pv[1][r], wd[r], acc[u + 1]);
      }
    }
    // sum over the ring groups (lane bits 5, 4): lane (b5, b4, q, i) keeps segment row b4 + 2 b5
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const float u0 = swap_add32f(acc[u][0], acc[u][2]);
      const float u1 = swap_add32f(acc[u][1], acc[u][3]);
      red[wave][8 * c + u][lane] = swap_add16f(u0, u1);
    }
  }
  __syncthreads();
  for (int e = tid; e < S * 64; e += 64 * NW) {
    const int t = e >> 6, ln = e & 63;
    float sum = red[0][t][ln];
#pragma unroll
    for (int w = 1; w < NW; ++w) sum += red[w][t][ln];
    const int row = ((ln >> 4) & 1) + 2 * (ln >> 5), qq = (ln >> 2) & 3, cj = ln & 3;
    const int4 dq = dsc[qq];
    if (dq.x < 0 || (cj >> 1) >= NF) continue;
    const int mq = dq.x, nsq = (L - mq + S) / S;
    const int off = (dq.y + row) * S + t;
    if ((dq.y + row) < nsq && off <= L - mq)
      reinterpret_cast<double*>(out.p[cj >> 1])[2 * (long)(dq.w + off) + (cj & 1)] =
          ((double)sum * X32_ANA_UNSCALE * qgs[(qq * 4 + row) * S + t]) / SLX_SQRT2_LEGACY;
  }
}

// -------------------------------------------------------------------------------------------
// Synthesis.  Blocks, staging and the dealing of segment groups to the 4 waves as k_leg_syn_mf
// (order pair p, L - p; ring chunk of 16 R rings), with the coefficients x qg staged as FP32 (x 2^-64).
// Lane (rg, k, i) = (lane >> 4, (lane >> 2) & 3, lane & 3): MFMA block b = 4 rg + k; A row i is
// ring 16 r + 4 rg + i of the chunk walked through slot k (segment 4 g + k of the group), B column
// i is field i >> 1, (re, im) = i & 1 of that slot's staged coefficient; even and odd steps go to
// separate accumulators, which run over all groups of an order.  At the end the four slots (lane
// bits 2, 3) are summed with xor shuffles, the waves in LDS in a fixed order, and
// FN = 1.41421356 (E + O) 2^64, FS = 1.41421356 (E - O) 2^64 are written in FP64.
template <int NF, int R>
__global__ void __launch_bounds__(256) k_leg_syn_x32(ShtDev d, SegDev g, const double2* __restrict__ cf, long cstride,
                                                     double2* __restrict__ FN, double2* __restrict__ FS, Pred pr) {
  static_assert(NF == 1 || NF == 2, "one or two fields per launch");
  if (!pr.active()) return;
  constexpr int S = MF_S;
  extern __shared__ double2 lsy[];
  const int L = d.lmax, nl = L + 1, nh = d.nh;
  const int ny = (nh + 16 * R - 1) / (16 * R), bt = blockIdx.x >> 3;
  const int yb = bt % ny, p = (blockIdx.x & 7) + 8 * (bt / ny);
  if (p >= (L + 2) / 2) return;
  const int mord[2] = {p, L - p};
  const int nord = (L - p == p) ? 1 : 2;
  const int ng0 = ((nl - mord[0] + S - 1) / S + 3) / 4;
  const int ng1 = nord == 2 ? ((nl - mord[1] + S - 1) / S + 3) / 4 : 0;
  const int np2 = 4 * S * (ng0 + ng1), base1 = 4 * S * ng0;
  double* qcl = reinterpret_cast<double*>(lsy);           // [np2] scaled recurrence coefficients
  float2* cl = reinterpret_cast<float2*>(lsy + np2 / 2);  // [NF][np2] coefficients x qg x 2^-64
  float* part = reinterpret_cast<float*>(lsy);            // [4][2][2][R][64], reuses the staging area
  {
    const int mo0 = d.moff[mord[0]], mo1 = d.moff[mord[1]];
    const int cnt0 = nl - mord[0], cnt1 = nl - mord[1];
    for (int e = threadIdx.x; e < np2; e += 256) {
      const bool second = e >= base1;
      const int k = second ? e - base1 : e;
      const bool v = k < (second ? cnt1 : cnt0);
      const int src = (second ? mo1 : mo0) + k;
      qcl[e] = v ? d.qc[src] : 0.0;
      const double gq = v ? d.qg[src] * X32_SYN_SCALE : 0.0;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double2 c = v ? cf[f * cstride + src] : make_double2(0.0, 0.0);
        cl[f * np2 + e] = make_float2((float)(c.x * gq), (float)(c.y * gq));
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int rg = lane >> 4, slot = (lane >> 2) & 3, i = lane & 3;
  const int fb = i >> 1, comp = i & 1;
  const bool colok = fb < NF;
  int jr[R];
  bool okr[R];
  double xr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    jr[r] = yb * 16 * R + r * 16 + 4 * rg + i;
    okr[r] = jr[r] < nh;
    xr[r] = d.x[okr[r] ? jr[r] : 0];
  }
  f32x4 acc[2][2][R];  // [order][E, O][ring set]
#pragma unroll
  for (int o = 0; o < 2; ++o)
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int r = 0; r < R; ++r) acc[o][e][r] = f32x4{0.f, 0.f, 0.f, 0.f};
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
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int jj = okr[r] ? jr[r] : 0;
      const double2 ab = (okr[r] && sval) ? g.sab[sx + jj] : make_double2(0.0, 0.0);
      st[r] = LegState{ab.x, ab.y, 0};
    }
    const double s23 = o ? s23o1 : s23o0;
    const int ob = (o ? base1 : 0) + sg * S;
    const double* qco = qcl + ob;
    const float* clo = reinterpret_cast<const float*>(cl + (colok ? fb : 0) * np2 + ob) + comp;
    f32x4 aE[R], aO[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      aE[r] = acc[o][0][r];
      aO[r] = acc[o][1][r];
    }
#pragma unroll
    for (int t = 0; t < S; t += 2) {
      const double c0 = qco[t], c1 = qco[t + 1];
      const float b0 = colok ? clo[2 * t] : 0.f, b1 = colok ? clo[2 * (t + 1)] : 0.f;
      float pv[2][R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (t == 0) {
          pv[0][r] = (float)x32_step(xr[r], c0, st[r], 0, head, s23);
          pv[1][r] = (float)x32_step(xr[r], c1, st[r], 1, head, s23);
        } else {
          pv[0][r] = (float)q_next(xr[r], c0, st[r]);
          pv[1][r] = (float)q_next(xr[r], c1, st[r]);
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        aE[r] = mfma_x32(pv[0][r], b0, aE[r]);
        aO[r] = mfma_x32(pv[1][r], b1, aO[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      acc[o][0][r] = aE[r];
      acc[o][1][r] = aO[r];
    }
  }
  __syncthreads();  // all waves are done with the staged coefficients
  // sum over the four slots (lane bits 3, 2): afterwards lane (rg, b3, b2, i) holds ring row
  // b2 + 2 b3 of its ring group, column i
  const bool up8 = (lane & 8) != 0, up4 = (lane & 4) != 0;
#pragma unroll
  for (int o = 0; o < 2; ++o)
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const f32x4 v = acc[o][e][r];
        const float s0 = up8 ? v[0] : v[2], k0 = up8 ? v[2] : v[0];
        const float s1 = up8 ? v[1] : v[3], k1 = up8 ? v[3] : v[1];
        const float h0 = k0 + __shfl_xor(s0, 8, 64), h1 = k1 + __shfl_xor(s1, 8, 64);
        const float sd = up4 ? h0 : h1, kp = up4 ? h1 : h0;
        part[(((wave * 2 + o) * 2 + e) * R + r) * 64 + lane] = kp + __shfl_xor(sd, 4, 64);
      }
  __syncthreads();
  if (wave >= nord) return;
  const int o = wave, m = mord[o];
  if (!colok) return;
  const long fstride = (long)d.nm * nh;
  const int rowr = ((lane >> 2) & 1) + 2 * ((lane >> 3) & 1);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    float e = part[(((0 * 2 + o) * 2 + 0) * R + r) * 64 + lane];
    float q = part[(((0 * 2 + o) * 2 + 1) * R + r) * 64 + lane];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      e += part[(((w * 2 + o) * 2 + 0) * R + r) * 64 + lane];
      q += part[(((w * 2 + o) * 2 + 1) * R + r) * 64 + lane];
    }
    const int j = yb * 16 * R + r * 16 + 4 * rg + rowr;
    if (j < nh) {
      const long idx = fb * fstride + (long)m * nh + j;
      const double ed = (double)e * X32_SYN_UNSCALE, od = (double)q * X32_SYN_UNSCALE;
      reinterpret_cast<double*>(FN)[2 * idx + comp] = SLX_SQRT2_LEGACY * (ed + od);
      reinterpret_cast<double*>(FS)[2 * idx + comp] = SLX_SQRT2_LEGACY * (ed - od);
    }
  }
}

}  // namespace slx
