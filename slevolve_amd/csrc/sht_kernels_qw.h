// Long-chain ("walk") analysis contraction on the matrix cores (FP64).
//
// The segmented kernels (k_leg_ana_mf / k_leg_ana_ck) give every 32-degree segment of an order its
// own start state, one double2 per (segment, ring) loaded from HBM per launch (17.9 MB of the
// 32.6 MB a 2-field launch moves at lmax 512), and pad every order to whole groups of 4 segments
// (about a fifth of their matrix steps are padding).  Here an order m is cut into 4 W chains of
// equal length Q (W walks of 4 chains; Q = ceil((L - m + 1) / 4W) rounded up to even, so that
// l - m has the parity of t in every chain and one folded input, S or D, serves an instruction):
// chain c covers l = m + c Q + t, t < Q.  The four rows of an MFMA block are the four chains of
// one walk ("unit"); a chain carries its recurrence state in registers from start to end, so a
// unit needs 4 start states per ring, and the padding is at most 7 steps per chain.  W is chosen
// per order so that chains are at most a target length (longer chains: fewer start-state bytes;
// shorter: more units to spread over the CUs).
//
// Scaled recurrence as in the segmented kernels (Q_l = x Q_{l-1} - qc_l Q_{l-2}, P~_l = qg_l Q_l),
// restarted every 32 steps of a chain (t = 32k) so that qg stays below ~2^(32 log2(1/eps)): at a
// restart the running state is converted to the unscaled one by the last two factors qg_{t-2},
// qg_{t-1} of the sub-segment that ends, and the first step uses qc = eps_{l-1}.  Chain 0 begins
// with the order's head (P~_m, then P~_{m+1} = sqrt(2m+3) x P~_m) and restarts at t = 2.
//
// Extended range: a lane whose Legendre values are below 2^-300 (scale count > 0 in the seeding
// walk) holds a zero state, as in the segmented kernels, with the same rule per 32-step
// sub-segment (zero while the scale count at the sub-segment's end is > 0).  A lane that is zero
// and reaches normal range inside sub-segment k takes its unscaled state at the sub-segment's
// start from a reload table; a per-unit bit mask (4 bits per boundary k) says whether any ring of
// a chain reloads there, so the reload is a wave-uniform branch, and only the lanes inside the
// band of rings that reload there (rings are ordered from the pole, so the band is contiguous)
// read the table.
//
// Work split: block = up to 3 units (dealt longest first to the least loaded block), NW waves over
// the rings (R rings per lane, as in the segmented kernels: lane bits 2-3 pick the MFMA block,
// bits 4-5 the ring within it, bits 0-1 the chain / B column).  Per chunk of 8 steps the 4 MFMA
// blocks are reduced by the same two xor levels as k_leg_ana_ck and the wave's partials go to LDS;
// every CHB chunks the block sums the waves in a fixed order and writes the outputs (two partial
// buffers, one barrier per CHB chunks).
#pragma once
#include "sht_kernels.h"

namespace slx {

struct QwDev {
  const int4* blk;       // [nblk] the units of a block (-1 padded)
  const int4* ua;        // [nunit] (m, Q, Qpad, table offset)
  const int4* ub;        // [nunit] (sub-segments per chain nk, reload row offset, reload mask, first chain)
  const double* qc;      // [table] per (unit, chain, t < Qpad)
  const double* qg;
  const double2* seed;   // [nunit][4][nh] unscaled chain start states
  const double2* rl;     // reload states, [(roff + chain (nk - 1) + k - 1) nh + ring]
  const int2* band;      // [nunit][4][8] rings [j0, j1) that reload at boundary k (entry k - 1)
};

// 64-bit DPP moves within 16-lane rows.  dpp_merge<CA, CB, BANKS>(a, b): lanes in the 4-lane
// banks of BANKS take a moved by CA, the other lanes b moved by CB.  bank_merge<BANKS>(a, b): a in
// BANKS, b elsewhere (identity quad permutation).
template <int CTRL, int BANKS>
__device__ __forceinline__ int dpp32(int old, int src) {
  return __builtin_amdgcn_update_dpp(old, src, CTRL, 0xF, BANKS, false);
}
template <int CA, int CB, int BANKS>
__device__ __forceinline__ double dpp_merge(double a, double b) {
  const int2 ia = __builtin_bit_cast(int2, a), ib = __builtin_bit_cast(int2, b);
  int2 r;
  // the first move leaves the lanes of BANKS as they were (undefined: no register is cleared for
  // them), the second writes exactly those lanes
  r.x = __builtin_amdgcn_mov_dpp(ib.x, CB, 0xF, 0xF & ~BANKS, false);
  r.y = __builtin_amdgcn_mov_dpp(ib.y, CB, 0xF, 0xF & ~BANKS, false);
  r.x = dpp32<CA, BANKS>(r.x, ia.x);
  r.y = dpp32<CA, BANKS>(r.y, ia.y);
  return __builtin_bit_cast(double, r);
}
template <int BANKS>
__device__ __forceinline__ double bank_merge(double a, double b) {
  const int2 ia = __builtin_bit_cast(int2, a), ib = __builtin_bit_cast(int2, b);
  int2 r;
  r.x = dpp32<0xE4, BANKS>(ib.x, ia.x);
  r.y = dpp32<0xE4, BANKS>(ib.y, ia.y);
  return __builtin_bit_cast(double, r);
}

template <int R, bool HEAD>
__device__ __forceinline__ void qw_chunk(LegState (&st)[R], const double (&xr)[R], const double (&wv)[2][R],
                                         const double* qcs, int t0, bool head, double s23, double* redw,
                                         int lane) {
  double acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = 0.0;
#pragma unroll
  for (int u = 0; u < 8; u += 2) {
    double pv[2][R];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int tt = u + h;
      const double cq = qcs[t0 + tt];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double p;
        if (HEAD && tt == 0) {  // head: P_m itself, the state does not move
          const double pa0 = st[r].pa, pb0 = st[r].pb;
          const double pn = q_next(xr[r], cq, st[r]);
          p = head ? pb0 : pn;
          st[r].pa = head ? pa0 : st[r].pa;
          st[r].pb = head ? pb0 : st[r].pb;
        } else if (HEAD && tt == 1) {  // head: P_{m+1} = sqrt(2m+3) x P_m
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
  // reduce-scatter over the 4 MFMA blocks (lane bits 3, 2) with row rotations of 16-lane rows:
  // level 1 pairs lane i with i ^ 8 (row_ror 8), level 2 with i ^ 4 (row_ror 12 for lanes with
  // bit 2 clear, row_ror 4 for the others); bank masks (4-lane banks of a row) merge the two halves
  // of a level, so the lanes keep the same sums in the same order as the xor shuffles of
  // k_leg_ana_ck.
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double recv = dpp_merge<0x128, 0x128, 0x3>(acc[v], acc[v + 4]);  // partner's acc[v] / acc[v + 4]
    const double keep = bank_merge<0x3>(acc[v], acc[v + 4]);
    acc[v] = keep + recv;
  }
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    const double recv = dpp_merge<0x12C, 0x124, 0x5>(acc[v], acc[v + 2]);
    const double keep = bank_merge<0x5>(acc[v], acc[v + 2]);
    acc[v] = keep + recv;
  }
  redw[lane] = acc[0];
  redw[64 + lane] = acc[1];
}

// Dynamic LDS: partials [2][NW][2 CHB][64] doubles, then a unit's tables qc, qg [4][QS] and its
// reload bands [4][8] int2; CHB chunks of 8 steps between two cross-wave reductions.
template <int NF, int R, int CHB>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_leg_ana_qw(ShtDev d, QwDev q, const double2* __restrict__ WS, const double2* __restrict__ WD, OutPtrs<NF> out,
             Pred pr) {
  static_assert(NF == 1 || NF == 2, "one or two fields per launch");
  if (!pr.active()) return;
  extern __shared__ double qw_smem[];
  const int L = d.lmax, nh = d.nh;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, NW = blockDim.x >> 6;
  const int blk = (lane >> 2) & 3, slot = lane & 3, kk = lane >> 4;
  const long fstride = (long)d.nm * nh;
  const int col = slot;  // B column of this lane (same lane bits as its A row)
  const bool colok = (col >> 1) < NF;
  const int fb = colok ? (col >> 1) : 0;
  double* red = qw_smem;                             // [2][NW][2 CHB][64]
  double* tabs = qw_smem + 2 * NW * 2 * CHB * 64;   // qc [4][QS], then qg [4][QS]
  const int wstride = 2 * CHB * 64;                  // one wave's partials in one buffer
  const int4 bo = q.blk[blockIdx.x];
  int jr[R];
  double xr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = (wave * R + r) * 16 + 4 * blk + kk;  // ring of this lane (A and B roles alike)
    jr[r] = j < nh ? j : -1;
    xr[r] = d.x[j < nh ? j : 0];
  }
  int buf = 0;
  for (int o = 0; o < 3; ++o) {
    const int u = o == 0 ? bo.x : o == 1 ? bo.y : bo.z;
    if (u < 0) break;
    const int4 ua = q.ua[u], ub = q.ub[u];
    const int m = ua.x, Q = ua.y, Qp = ua.z, toff = ua.w;
    const int nk = ub.x, roff = ub.y, rmask = ub.z, c0 = ub.w;
    const int QS = Qp + 2;
    const int mo = d.moff[m];
    int2* bands = reinterpret_cast<int2*>(tabs + 8 * QS);  // [4][8]
    __syncthreads();  // the previous unit's outputs have been written from the tables
    for (int i = threadIdx.x; i < 4 * Qp; i += blockDim.x) {
      const int c = i / Qp, t = i - c * Qp;
      tabs[c * QS + t] = q.qc[toff + i];
      tabs[4 * QS + c * QS + t] = q.qg[toff + i];
    }
    if (rmask != 0 && threadIdx.x < 32) bands[threadIdx.x] = q.band[(long)u * 32 + threadIdx.x];
    double wv[2][R];
    LegState st[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool ok = jr[r] >= 0;
      const int jj = ok ? jr[r] : 0;
      const double2 ab = ok ? q.seed[((long)u * 4 + slot) * nh + jj] : make_double2(0.0, 0.0);
      st[r] = LegState{ab.x, ab.y, 0};
      const double2 vs = WS[fb * fstride + (long)m * nh + jj], vd = WD[fb * fstride + (long)m * nh + jj];
      const bool use = ok && colok;
      wv[0][r] = use ? ((col & 1) ? vs.y : vs.x) : 0.0;
      wv[1][r] = use ? ((col & 1) ? vd.y : vd.x) : 0.0;
    }
    const double s23 = d.s2m3[m];
    const bool head = c0 == 0 && slot == 0;
    __syncthreads();
    const double* qcs = tabs + slot * QS;
    const double* qgs = tabs + 4 * QS + slot * QS;
    const int nch = Qp >> 3;
    for (int c = 0; c < nch; ++c) {
      const int t0 = 8 * c;
      if ((t0 & 31) == 0 && t0 > 0) {  // restart of the scaled recurrence at t = 32 k
        const int k = t0 >> 5;
        const double ga = qgs[t0 - 2], gb = qgs[t0 - 1];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          st[r].pa *= ga;
          st[r].pb *= gb;
        }
        const int fl = (rmask >> (4 * (k - 1))) & 15;
        if (fl != 0) {
          if ((fl >> slot) & 1) {
            const double2* rl = q.rl + ((long)roff + slot * (nk - 1) + k - 1) * nh;
            const int2 bd = bands[slot * 8 + k - 1];
#pragma unroll
            for (int r = 0; r < R; ++r) {
              if (jr[r] >= bd.x && jr[r] < bd.y && st[r].pa == 0.0 && st[r].pb == 0.0) {
                const double2 v = rl[jr[r]];
                st[r].pa = v.x;
                st[r].pb = v.y;
              }
            }
          }
        }
      }
      const int cw = c % CHB;
      double* redw = red + (buf * NW + wave) * wstride + cw * 128;
      if (c == 0 && c0 == 0)
        qw_chunk<R, true>(st, xr, wv, qcs, t0, head, s23, redw, lane);
      else
        qw_chunk<R, false>(st, xr, wv, qcs, t0, head, s23, redw, lane);
      if (cw == CHB - 1 || c == nch - 1) {
        __syncthreads();
        const int nv = (cw + 1) * 128;  // partial values of this window
        const int cs = c - cw;
        const double* rb = red + buf * NW * wstride;
        for (int e = threadIdx.x; e < nv; e += blockDim.x) {
          const int v = e >> 6, ln = e & 63;
          double sum = rb[v * 64 + ln];
          for (int w = 1; w < NW; ++w) sum += rb[w * wstride + v * 64 + ln];
          const int t = 8 * (cs + (v >> 1)) + 4 * ((ln >> 3) & 1) + 2 * ((ln >> 2) & 1) + (v & 1);
          const int row = ln >> 4, cj = ln & 3;
          const int off = (c0 + row) * Q + t;  // l - m
          if (t < Q && off <= L - m && (cj >> 1) < NF)
            reinterpret_cast<double*>(out.p[cj >> 1])[2 * (long)(mo + off) + (cj & 1)] =
                (sum * tabs[4 * QS + row * QS + t]) / SLX_SQRT2_LEGACY;
        }
        buf ^= 1;
      }
    }
  }
}

// Per order, for the seeding kernels: (Q, chains 4 W, sub-segments per chain nk, first point),
// and the order's first unit.
struct QwOrd {
  const int4* ord;
  const int* u0;
};

// Seeding walk of the chains: one thread per (order, ring) runs the extended-range recurrence of
// k_leg_seed (same arithmetic) and records the raw state (P~_{l0-2}, P~_{l0-1}, scale count) at
// every boundary l0 = m + c Q + 32 k of the order (point c nk + k; the head point (0, 0) holds
// (0, P~_mm)).  Points past L are not written.
static __global__ void k_qw_seed_walk(ShtDev d, QwOrd qo, double2* __restrict__ ab, int* __restrict__ sc) {
  const int m = blockIdx.x;
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= d.nh) return;
  const int L = d.lmax, mo = d.moff[m], nh = d.nh;
  const int4 od = qo.ord[m];
  const int Q = od.x, nc = od.y, nk = od.z;
  const long pb = od.w;
  const double x = d.x[j];
  LegState s{0.0, d.pmm[(long)m * nh + j], d.pms[(long)m * nh + j]};
  ab[pb * nh + j] = make_double2(s.pa, s.pb);
  sc[pb * nh + j] = s.sc;
  for (int l = m + 1; l <= L; ++l) {
    if (l == m + 1) {
      const double p = d.s2m3[m] * x * s.pb;
      s.pa = s.pb;
      s.pb = p;
    } else {
      leg_next(x, d.rec[mo + (l - m)], s);
    }
    if (((l - m) & 7) == 7) leg_renorm(s);
    const int n = l + 1 - m;  // the state now starts degree l + 1 = m + n
    if (l + 1 <= L) {
      const int c = n / Q, t = n - c * Q;
      if (c < nc && (t & 31) == 0) {
        const long p = pb + (long)c * nk + (t >> 5);
        ab[p * nh + j] = make_double2(s.pa, s.pb);
        sc[p * nh + j] = s.sc;
      }
    }
  }
}

// Start and reload states from the raw walk: point (c, k) of order m exists when its degree is
// <= L; the state is unscaled by the rule of mf_unscale with the scale count at the next existing
// point (0 after the order's last one).  Sub-segment 0 of each chain gives the chain's start
// state; sub-segments k >= 1 give a reload state only for a lane that is zero before it (scale
// count > 0) and reaches normal range inside it; any such ring sets bit 4 (k - 1) + c % 4 of its
// unit's word in umask and widens the chain's band [j0, j1) of reloading rings at k.
static __global__ void k_qw_seed_unscale(ShtDev d, QwDev q, QwOrd qo, const double2* __restrict__ ab,
                                         const int* __restrict__ sc, double2* __restrict__ seed,
                                         double2* __restrict__ rl, int* __restrict__ umask, int2* __restrict__ band) {
  const int m = blockIdx.x;
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= d.nh) return;
  const int L = d.lmax, nh = d.nh;
  const int4 od = qo.ord[m];
  const int Q = od.x, nc = od.y, nk = od.z;
  const long pb = od.w;
  const int u0 = qo.u0[m];
  auto exists = [&](int c, int k) { return c < nc && 32 * k < Q && m + c * Q + 32 * k <= L; };
  for (int c = 0; c < nc; ++c) {
    const int u = u0 + c / 4, i = c % 4;
    const int roff = q.ub[u].y;
    for (int k = 0; k < nk; ++k) {
      double2 v = make_double2(0.0, 0.0);
      bool reload = false;
      if (exists(c, k)) {
        const long p = pb + (long)c * nk + k;
        int cn = c, kn = k + 1;
        if (32 * kn >= Q) {
          cn = c + 1;
          kn = 0;
        }
        const int sn = exists(cn, kn) ? sc[(pb + (long)cn * nk + kn) * nh + j] : 0;
        const int s0 = sc[p * nh + j];
        double a = ab[p * nh + j].x, b = ab[p * nh + j].y;
        mf_unscale(a, b, s0, sn);
        v = make_double2(a, b);
        reload = k > 0 && s0 > 0 && sn <= 0;
      }
      if (k == 0) {
        seed[((long)u * 4 + i) * nh + j] = v;
      } else {
        rl[((long)roff + i * (nk - 1) + k - 1) * nh + j] = reload ? v : make_double2(0.0, 0.0);
        if (reload) {
          atomicOr(&umask[u], 1 << (4 * (k - 1) + i));
          atomicMin(&band[((long)u * 4 + i) * 8 + k - 1].x, j);
          atomicMax(&band[((long)u * 4 + i) * 8 + k - 1].y, j + 1);
        }
      }
    }
  }
}

}  // namespace slx
