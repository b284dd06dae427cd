// Host-side launch helpers for the transform kernels (shared by the SHT and solver units).
#pragma once
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "sht_kernels.h"
#include "sht_kernels_qw.h"
#include "sht_kernels_x32.h"

namespace slx {

inline ShtDev dev_of(const slx_sht* h) {
  ShtDev d;
  d.nlat = h->nlat;
  d.nlon = h->nlon;
  d.lmax = h->lmax;
  d.nh = h->nh;
  d.nm = h->nm;
  d.log2n = h->log2n;
  d.x = h->x;
  d.wn = h->wn;
  d.ws = h->ws;
  d.pmm = h->pmm;
  d.pms = h->pms;
  d.rec = h->rec;
  d.s2m3 = h->s2m3;
  d.moff = h->moff;
  d.jb = h->jb;
  d.stw = h->stw;
  d.l_of = h->l_of;
  d.m_of = h->m_of;
  d.qc = h->qc;
  d.qg = h->qg;
  return d;
}

struct LoadPlain {
  const double* g;
  long fstride;
  __device__ double operator()(int f, long i) const { return g[f * fstride + i]; }
};
struct CoefPlain {
  const double2* c;
  long fstride;
  __device__ bool prologue_needed(int, int) const { return false; }
  __device__ void prologue(double*) const {}
  __device__ double2 operator()(int f, int idx, int, int, const double*) const { return c[f * fstride + idx]; }
};
struct StorePlain {
  double* g;
  long fstride;
  __device__ void operator()(int f, long i, double v) const { g[f * fstride + i] = v; }
};

// Launch helpers shared with the solver ------------------------------------------------------
inline SegDev seg_of(const slx_sht* h) {
  SegDev g;
  g.nseg = h->nseg;
  g.S = h->seg_len;
  g.Ssyn = h->syn_seg;
  g.seg_m = h->seg_m;
  g.seg_l0 = h->seg_l0;
  g.sa = h->seed_a;
  g.sb = h->seed_b;
  g.ss = h->seed_s;
  g.sab = h->seed_ab;
  g.first = h->seg_first;
  return g;
}

// Dynamic LDS above the default 64 KB is opted into per kernel, raised monotonically under a
// lock (host threads driving different handles may launch the same kernel with different sizes).
inline std::mutex& lds_mutex() {
  static std::mutex mu;
  return mu;
}
template <auto KERN>
int allow_lds(size_t lds) {
  static size_t allowed = 65536;
  std::lock_guard<std::mutex> lock(lds_mutex());
  if (lds > allowed) {
    SLX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(KERN), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    allowed = lds;
  }
  return 0;
}

inline int ana_rings_per_lane(const slx_sht* h) {
  const int r = (h->nh + 64 * h->ana_nw - 1) / (64 * h->ana_nw);
  return r <= 1 ? 1 : r <= 2 ? 2 : 4;
}

template <int NF, int R>
int launch_leg_ana_mf(const slx_sht* h, const double2* WS, const double2* WD, OutPtrs<NF> out, Pred pr,
                      hipStream_t s) {
  const int nw = (h->nh + 16 * R - 1) / (16 * R);
  GrpDev gr{h->ngrp, h->grp_m, h->grp_s0, h->grp_desc};
  dim3 grid(8 * ((h->ngrp + 7) / 8));
  const size_t lds = (nw > 1) ? (size_t)nw * 8 * 64 * sizeof(double) : 0;
  k_leg_ana_mf<NF, R><<<grid, 64 * nw, lds, s>>>(dev_of(h), seg_of(h), gr, WS, WD, out, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}

// Persistent chunked-accumulator analysis: one block per resident slot (as many as the device
// holds at this block shape, a multiple of 8 so that every XCD gets the same number), at most one
// per group.
template <int NF, int R>
int launch_leg_ana_ck(const slx_sht* h, const double2* WS, const double2* WD, OutPtrs<NF> out, Pred pr,
                      hipStream_t s) {
  const int nw = (h->nh + 16 * R - 1) / (16 * R);
  const size_t lds = (size_t)nw * 8 * 64 * sizeof(double);
  static int cached_nw = -1, cached_grid = 0;
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (nw != cached_nw) {
    int dev = 0, ncu = 0, per_cu = 0;
    SLX_HIP(hipGetDevice(&dev));
    SLX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    SLX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_leg_ana_ck<NF, R>),
                                                         64 * nw, lds));
    cached_grid = 8 * std::max(1, (ncu * std::max(per_cu, 1)) / 8);
    cached_nw = nw;
  }
  const int grid = std::min(cached_grid, 8 * ((h->ngrp + 7) / 8));
  GrpDev gr{h->ngrp, h->grp_m, h->grp_s0, h->grp_desc};
  k_leg_ana_ck<NF, R><<<grid, 64 * nw, lds, s>>>(dev_of(h), seg_of(h), gr, WS, WD, out, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}

inline QwDev qw_of(const slx_walk& w, const int4* blk) {
  return QwDev{blk, w.ua, w.ub, w.qc, w.qg, w.seed, w.rl, w.band};
}
// Walk analysis: one block per entry of the unit-block list, 16 R rings per wave; CHB = 4
// chunks between reductions when two blocks still fit a CU's LDS, else 2.
inline size_t qw_lds(const slx_sht* h, int nw, int chb) {
  return (size_t)(2 * nw * 2 * chb * 64 + 2 * 4 * h->wa->qsmax) * sizeof(double) + 32 * sizeof(int2);
}
template <int NF, int R, int CHB>
int launch_leg_ana_qw_c(const slx_sht* h, const double2* WS, const double2* WD, OutPtrs<NF> out, Pred pr,
                        hipStream_t s) {
  const int nw = (h->nh + 16 * R - 1) / (16 * R);
  const size_t lds = qw_lds(h, nw, CHB);
  if (int rc = allow_lds<&k_leg_ana_qw<NF, R, CHB>>(lds)) return rc;
  k_leg_ana_qw<NF, R, CHB><<<h->qw_nblk, 64 * nw, lds, s>>>(dev_of(h), qw_of(*h->wa, h->qw_blk), WS, WD, out, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}
template <int NF, int R>
int launch_leg_ana_qw(const slx_sht* h, const double2* WS, const double2* WD, OutPtrs<NF> out, Pred pr,
                      hipStream_t s) {
  const int nw = (h->nh + 16 * R - 1) / (16 * R);
  if (qw_lds(h, nw, 4) <= 80 * 1024) return launch_leg_ana_qw_c<NF, R, 4>(h, WS, WD, out, pr, s);
  return launch_leg_ana_qw_c<NF, R, 2>(h, WS, WD, out, pr, s);
}

template <int NF>
int launch_leg_ana(const slx_sht* h, const double2* WS, const double2* WD, OutPtrs<NF> out, Pred pr,
                   hipStream_t s) {
  if (h->ana_mfma) {
    if constexpr (NF == 3) {  // two launches: fields 0, 1 and field 2
      const long fs = (long)h->nm * h->nh;
      int rc = launch_leg_ana<2>(h, WS, WD, OutPtrs<2>{{out.p[0], out.p[1]}}, pr, s);
      if (rc) return rc;
      return launch_leg_ana<1>(h, WS + 2 * fs, WD + 2 * fs, OutPtrs<1>{{out.p[2]}}, pr, s);
    } else if (h->ana_qw) {
      return h->ana_r == 4 ? launch_leg_ana_qw<NF, 4>(h, WS, WD, out, pr, s)
                           : launch_leg_ana_qw<NF, 2>(h, WS, WD, out, pr, s);
    } else if (h->ana_ck) {
      return h->ana_r == 4 ? launch_leg_ana_ck<NF, 4>(h, WS, WD, out, pr, s)
                           : launch_leg_ana_ck<NF, 2>(h, WS, WD, out, pr, s);
    } else {
      return h->ana_r == 4 ? launch_leg_ana_mf<NF, 4>(h, WS, WD, out, pr, s)
                           : launch_leg_ana_mf<NF, 2>(h, WS, WD, out, pr, s);
    }
  }
  const ShtDev d = dev_of(h);
  const SegDev g = seg_of(h);
  const int R = ana_rings_per_lane(h);
  dim3 grid(h->nseg);
  if (h->ana_nw == 1) {
    switch (R) {
      case 1: k_leg_ana<NF, 1, 1><<<grid, 64, 0, s>>>(d, g, WS, WD, out, pr); break;
      case 2: k_leg_ana<NF, 2, 1><<<grid, 64, 0, s>>>(d, g, WS, WD, out, pr); break;
      default: k_leg_ana<NF, 4, 1><<<grid, 64, 0, s>>>(d, g, WS, WD, out, pr); break;
    }
  } else {
    switch (R) {
      case 1: k_leg_ana<NF, 1, 2><<<grid, 128, 0, s>>>(d, g, WS, WD, out, pr); break;
      case 2: k_leg_ana<NF, 2, 2><<<grid, 128, 0, s>>>(d, g, WS, WD, out, pr); break;
      default: k_leg_ana<NF, 4, 2><<<grid, 128, 0, s>>>(d, g, WS, WD, out, pr); break;
    }
  }
  SLX_LAUNCH_CHECK();
  return 0;
}
}  // namespace slx

namespace slx {
// Ring passes: points per thread (4 up to nlon = 2048, else 8), LDS = two rows.
inline int ring_points(const slx_sht* h) { return h->nlon <= 4 * RING_MAXNT ? 4 : 8; }
inline size_t ring_lds(const slx_sht* h) { return 2 * (size_t)h->nlon * sizeof(double2); }

// (WS, WD: the folded-spectra work area, the handle's unless given)
template <int P, class Loader>
int launch_fft_ana_p(const slx_sht* h, Loader ld, int nf, Pred pr, hipStream_t s, double2* WS, double2* WD) {
  const size_t lds = ring_lds(h);
  if (int rc = allow_lds<&k_fft_ana<P, Loader>>(lds)) return rc;
  k_fft_ana<P, Loader><<<dim3(h->nh, nf), h->nlon / P, lds, s>>>(dev_of(h), ld, WS, WD, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}
template <class Loader>
int launch_fft_ana(const slx_sht* h, Loader ld, int nf, Pred pr, hipStream_t s, double2* WS = nullptr,
                   double2* WD = nullptr) {
  if (!WS) WS = h->WS;
  if (!WD) WD = h->WD;
  return ring_points(h) == 4 ? launch_fft_ana_p<4>(h, ld, nf, pr, s, WS, WD)
                             : launch_fft_ana_p<8>(h, ld, nf, pr, s, WS, WD);
}
template <int P, class Store>
int launch_fft_syn_p(const slx_sht* h, Store st, int nf, Pred pr, hipStream_t s) {
  const size_t lds = ring_lds(h);
  if (int rc = allow_lds<&k_fft_syn<P, Store>>(lds)) return rc;
  k_fft_syn<P, Store><<<dim3(h->nh, nf), h->nlon / P, lds, s>>>(dev_of(h), h->FN, h->FSo, st, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}
template <class Store>
int launch_fft_syn(const slx_sht* h, Store st, int nf, Pred pr, hipStream_t s) {
  return ring_points(h) == 4 ? launch_fft_syn_p<4>(h, st, nf, pr, s) : launch_fft_syn_p<8>(h, st, nf, pr, s);
}
// Synthesis with the waves over rings: 8 XCD lanes x chunks of 16 NWV R rings x ceil(order pairs /
// 8) blocks of NWV waves; LDS = the pair's staged tables and coefficients.
template <int NF, int R, int NWV, class Coef>
int launch_leg_syn_mw(const slx_sht* h, Coef cf, long off, Pred pr, hipStream_t s) {
  const int ny = (h->nh + 16 * NWV * R - 1) / (16 * NWV * R), np = (h->lmax + 2) / 2;
  const dim3 grid(8 * ny * ((np + 7) / 8));
  const int np2 = 4 * MF_S * ((h->lmax + 3 + 4 * MF_S - 1) / (4 * MF_S) + 1);
  const size_t lds = (size_t)(np2 / MF_S) * (MF_QS * sizeof(double) + NF * MF_CS * sizeof(double2)) +
                     MF_CS * sizeof(double2);  // + the zero row
  if (int rc = allow_lds<&k_leg_syn_mw<NF, R, NWV, Coef>>(lds)) return rc;
  k_leg_syn_mw<NF, R, NWV, Coef><<<grid, 64 * NWV, lds, s>>>(dev_of(h), seg_of(h), cf, h->FN + off, h->FSo + off,
                                                             pr);
  SLX_LAUNCH_CHECK();
  return 0;
}
template <int NF, int R, class Coef>
int launch_leg_syn_mf(const slx_sht* h, Coef cf, long off, Pred pr, hipStream_t s) {
  // 1-D grid: 8 XCD lanes x ring chunks x ceil(order pairs / 8) (blocks past the last pair exit)
  const int ny = (h->nh + 16 * R - 1) / (16 * R), np = (h->lmax + 2) / 2;
  const dim3 grid(8 * ny * ((np + 7) / 8));
  // staged entries of an order pair: ceil(n0 / 4S) + ceil(n1 / 4S) groups of 4S degrees with
  // n0 + n1 <= L + 3 (syn_pair), at most ceil((L + 3) / 4S) + 1 groups
  const int np2 = 4 * MF_S * ((h->lmax + 3 + 4 * MF_S - 1) / (4 * MF_S) + 1);
  const size_t lds = std::max((size_t)(np2 / MF_S) * (MF_QS * sizeof(double) + NF * MF_CS * sizeof(double2)) +
                                  MF_CS * sizeof(double2),  // + the zero row
                              (size_t)4 * 2 * 2 * R * 64 * sizeof(double));
  if (int rc = allow_lds<&k_leg_syn_mf<NF, R, Coef>>(lds)) return rc;
  k_leg_syn_mf<NF, R, Coef><<<grid, 256, lds, s>>>(dev_of(h), seg_of(h), cf, h->FN + off, h->FSo + off, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}

// Synthesis contraction of NF fields (FN/FS fields field .. field + NF - 1).  The matrix-core
// kernels take the coefficient loader itself and form the coefficients while they stage them (a
// CoefPlain source is read in place); the scalar-fed kernel reads an array, so any other loader is
// first assembled into the handle's CS array for it.
template <int NF, class Coef>
int launch_leg_syn(const slx_sht* h, Coef cf, Pred pr, hipStream_t s, int field = 0) {
  const long off = (long)field * h->nm * h->nh;
  if constexpr (NF <= 2) {
    if (h->syn_mfma == 2) {
      if (h->syn_nw == 8)
        return h->syn_r == 1 ? launch_leg_syn_mw<NF, 1, 8>(h, cf, off, pr, s)
                             : launch_leg_syn_mw<NF, 2, 8>(h, cf, off, pr, s);
      return h->syn_r == 1 ? launch_leg_syn_mw<NF, 1, 4>(h, cf, off, pr, s)
                           : launch_leg_syn_mw<NF, 2, 4>(h, cf, off, pr, s);
    }
    if (h->syn_mfma) return h->syn_r == 2 ? launch_leg_syn_mf<NF, 2>(h, cf, off, pr, s)
                                          : launch_leg_syn_mf<NF, 4>(h, cf, off, pr, s);
  }
  const double2* src;
  long cstride;
  if constexpr (std::is_same<Coef, CoefPlain>::value) {
    src = cf.c;
    cstride = cf.fstride;
  } else {
    k_coef_fill<NF, Coef><<<(h->ncoef + 255) / 256, 256, 0, s>>>(dev_of(h), cf, h->CS, h->ncoef, h->ncoef, pr);
    SLX_LAUNCH_CHECK();
    slx_mark(h, SLX_STAGE_SPEC, s);
    src = h->CS;
    cstride = h->ncoef;
  }
  dim3 grid((h->lmax + 2) / 2, (h->nh + 63) / 64);
  k_leg_syn_s<NF, 1><<<grid, 256, 0, s>>>(dev_of(h), seg_of(h), h->rec, src, cstride, h->FN + off, h->FSo + off, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}

// Reduced-precision contractions (FP64 recurrence, FP32 matrix-core products).  Analysis: one block
// of X32_NW waves per super-group of 4 segment groups (dealt XCD-aware as the FP64 kernels'
// groups), R ring chains per lane with 4 R X32_NW >= nh.
constexpr int X32_SYN_R = 2;  // rings per lane of the synthesis
template <int NF>
int launch_leg_ana_x32(const slx_sht* h, const double2* WS, const double2* WD, OutPtrs<NF> out, Pred pr,
                       hipStream_t s) {
  const int nsg = (h->ngrp + 3) / 4;
  GrpDev gr{h->ngrp, h->grp_m, h->grp_s0, h->grp_desc};
  const int rneed = (h->nh + 4 * X32_NW - 1) / (4 * X32_NW);
  const dim3 grid(8 * ((nsg + 7) / 8)), blk(64 * X32_NW);
  const ShtDev d = dev_of(h);
  const SegDev g = seg_of(h);
  if (rneed <= 1)
    k_leg_ana_x32<NF, 1><<<grid, blk, 0, s>>>(d, g, gr, WS, WD, out, pr);
  else if (rneed <= 2)
    k_leg_ana_x32<NF, 2><<<grid, blk, 0, s>>>(d, g, gr, WS, WD, out, pr);
  else if (rneed <= 4)
    k_leg_ana_x32<NF, 4><<<grid, blk, 0, s>>>(d, g, gr, WS, WD, out, pr);
  else if (rneed <= 8)
    k_leg_ana_x32<NF, 8><<<grid, blk, 0, s>>>(d, g, gr, WS, WD, out, pr);
  else
    k_leg_ana_x32<NF, 16><<<grid, blk, 0, s>>>(d, g, gr, WS, WD, out, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}
template <int NF>
int launch_leg_syn_x32(const slx_sht* h, const double2* src, long cstride, long off, Pred pr, hipStream_t s) {
  constexpr int R = X32_SYN_R;
  const int ny = (h->nh + 16 * R - 1) / (16 * R), np = (h->lmax + 2) / 2;
  const dim3 grid(8 * ny * ((np + 7) / 8));
  const int np2 = 4 * MF_S * ((h->lmax + 2 + 4 * MF_S - 1) / (4 * MF_S) + 1);
  const size_t lds = std::max((size_t)np2 * (sizeof(double) + NF * sizeof(float2)),
                              (size_t)4 * 2 * 2 * R * 64 * sizeof(float));
  if (int rc = allow_lds<&k_leg_syn_x32<NF, R>>(lds)) return rc;
  k_leg_syn_x32<NF, R><<<grid, 256, lds, s>>>(dev_of(h), seg_of(h), src, cstride, h->FN + off, h->FSo + off, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}
// FP32-product analysis of NF fields from the folded spectra in WS/WD (matrix-core path only)
template <int NF>
int launch_leg_ana32(const slx_sht* h, const double2* WS, const double2* WD, OutPtrs<NF> out, Pred pr, hipStream_t s) {
  if constexpr (NF == 3) {
    const long fs = (long)h->nm * h->nh;
    int rc = launch_leg_ana_x32<2>(h, WS, WD, OutPtrs<2>{{out.p[0], out.p[1]}}, pr, s);
    if (rc) return rc;
    return launch_leg_ana_x32<1>(h, WS + 2 * fs, WD + 2 * fs, OutPtrs<1>{{out.p[2]}}, pr, s);
  } else {
    return launch_leg_ana_x32<NF>(h, WS, WD, out, pr, s);
  }
}
// FP32-product synthesis contraction of NF fields (coefficient loaders as launch_leg_syn)
template <int NF, class Coef>
int launch_leg_syn32(const slx_sht* h, Coef cf, Pred pr, hipStream_t s, int field = 0) {
  const double2* src;
  long cstride;
  if constexpr (std::is_same<Coef, CoefPlain>::value) {
    src = cf.c;
    cstride = cf.fstride;
  } else {
    k_coef_fill<NF, Coef><<<(h->ncoef + 255) / 256, 256, 0, s>>>(dev_of(h), cf, h->CS, h->ncoef, h->ncoef, pr);
    SLX_LAUNCH_CHECK();
    slx_mark(h, SLX_STAGE_SPEC, s);
    src = h->CS;
    cstride = h->ncoef;
  }
  const long off = (long)field * h->nm * h->nh;
  if constexpr (NF <= 2) {
    return launch_leg_syn_x32<NF>(h, src, cstride, off, pr, s);
  } else {
    int rc = launch_leg_syn_x32<2>(h, src, cstride, off, pr, s);
    if (rc) return rc;
    return launch_leg_syn_x32<1>(h, src + 2 * cstride, cstride, off + 2L * h->nm * h->nh, pr, s);
  }
}
}  // namespace slx
