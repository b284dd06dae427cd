// Internal definitions shared by the libslx translation units (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#define SLX_SQRT2_LEGACY 1.41421356
// Extended-range Legendre values: a value below 2^-SLX_SCALE_T is carried as P~ 2^(SLX_SCALE_K s)
// with an integer scale count s > 0 and renormalised once the scaled value passes 2^SLX_SCALE_T;
// while s > 0 it contributes nothing (its true magnitude is then below about 2^-180, the bound
// including the growth between two renormalisation checks).
#define SLX_SCALE_K 600
#define SLX_SCALE_T 300

// ------------------------------------------------------------------ error plumbing
void slx_set_error(const std::string& msg);
#define SLX_HIP(call)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      slx_set_error(std::string(#call) + ": " + hipGetErrorString(e_));                      \
      return (int)e_ == 0 ? -1 : (int)e_;                                                    \
    }                                                                                        \
  } while (0)
#define SLX_CHECK(cond, msg)                                                                 \
  do {                                                                                       \
    if (!(cond)) {                                                                           \
      slx_set_error(msg);                                                                    \
      return -22;                                                                            \
    }                                                                                        \
  } while (0)
#define SLX_LAUNCH_CHECK()                                                                   \
  do {                                                                                       \
    hipError_t e_ = hipGetLastError();                                                       \
    if (e_ != hipSuccess) {                                                                  \
      slx_set_error(std::string("kernel launch: ") + hipGetErrorString(e_));                 \
      return (int)e_;                                                                        \
    }                                                                                        \
  } while (0)

// ------------------------------------------------------------------ the transform handle
// Spectral arrays on the device are "packed": order-major triangles, coefficient (l, m) at
// moff[m] + (l - m), moff[m] = m (L+1) - m (m-1)/2, complex as double2.
struct slx_sht {
  int nlat = 0, nlon = 0, lmax = 0, nh = 0, nm = 0, ncoef = 0, log2n = 0, max_fields = 0;
  int precision = 0;      // SLX_PREC_FP64, or SLX_PREC_FP32: the transform entry points take the
                          // FP32-product contractions
  double* x = nullptr;    // [nh] mu_j of the northern rings (north -> equator)
  double* wn = nullptr;   // [nh] w_j / nlon
  double* ws = nullptr;   // [nh] w_{nlat-1-j} / nlon
  double* pmm = nullptr;  // [nm][nh] sectoral P~_mm(mu_j), scaled by 2^(SLX_SCALE_K pms)
  int* pms = nullptr;     // [nm][nh] scale counts of pmm
  double2* rec = nullptr; // [ncoef] (eps_{l-1,m}, 1/eps_{l,m}) at (l, m), l >= m+2
  double* s2m3 = nullptr; // [nm] sqrt(2m+3)
  int* moff = nullptr;    // [nm+1]
  int* l_of = nullptr;    // [ncoef] degree of each packed index
  int* jb = nullptr;      // [nm] first ring (from the pole) with non-negligible P~_lm, multiple of 64
  double2* stw = nullptr;  // per-stage FFT twiddles (fft_lds), < nlon entries
  int* m_of = nullptr;    // [ncoef] order of each packed index
  // scaled recurrence of the matrix-core kernels, restarted at every 32-degree segment start
  // ls = max(l0, m + 2) from the unscaled state (P~_{ls-2}, P~_{ls-1}):
  //   Q_l = x Q_{l-1} - qc_l Q_{l-2},  P~_l = qg_l Q_l,
  //   qc_ls = eps_{ls-1}, qc_l = eps_{l-1}^2 (l > ls), qg_l = prod_{k = ls..l} 1 / eps_k
  double* qc = nullptr;   // [ncoef]
  double* qg = nullptr;   // [ncoef] (1 for l < m + 2)
  // degree segments of the contractions (order m, first degree l0) and recurrence seeds
  int nseg = 0, seg_len = 0, ana_nw = 1;
  int syn_seg = 0;        // degrees per synthesis chain (a multiple of seg_len)
  int *seg_m = nullptr, *seg_l0 = nullptr, *seg_first = nullptr;
  double *seed_a = nullptr, *seed_b = nullptr;
  int* seed_s = nullptr;
  double2* seed_ab = nullptr;  // [nseg][nh] start states unscaled for the matrix-core kernels
  // analysis on the matrix cores: groups of 4 consecutive segments of one order
  int ana_mfma = 0, ana_r = 2, ngrp = 0;
  int ana_ck = 0;         // chunked-accumulator persistent analysis kernel (else one group per block)
  int4* grp_desc = nullptr;  // [ngrp] (m, first segment s0, seed row first[m] + s0, moff[m])
  // walk analysis (k_leg_ana_qw): orders cut into chains of at most a target length, 4 chains per
  // unit, units packed onto blocks
  int ana_qw = 0;         // the analysis runs the walk kernel (ana_r rings per lane)
  struct slx_walk* wa = nullptr;  // its layout
  int qw_nblk = 0;
  int4* qw_blk = nullptr;   // [qw_nblk] units of a block, -1 padded
  int syn_mfma = 0;       // synthesis: 0 wave-uniform scalar-fed VALU, 1 matrix cores with the waves over
                          // segment groups (k_leg_syn_mf), 2 with the waves over rings (k_leg_syn_mw)
  int syn_r = 4;          // rings per lane of the matrix-core synthesis (1, 2 or 4)
  int syn_nw = 4;         // waves per block of the ring-split synthesis (4 or 8)
  int *grp_m = nullptr, *grp_s0 = nullptr;
  double2* CS = nullptr;  // [max_fields][ncoef] synthesis coefficients assembled by a loader
  double2* WS = nullptr;  // [max_fields][nm][nh] analysis: folded even parts
  double2* WD = nullptr;  // [max_fields][nm][nh] analysis: folded odd parts
  double2* FN = nullptr;  // [max_fields][nm][nh] synthesis: northern Fourier coefficients
  double2* FSo = nullptr; // [max_fields][nm][nh] synthesis: southern Fourier coefficients
  std::vector<double> h_x, h_w;  // host copies: all nlat nodes / weights (north -> south)
  std::vector<int> h_moff;
  struct slx_stage_marks* marks = nullptr;  // stage timing of a profiled solver step (else null)
};

// Walk layout of the chain analysis: order m cut into 4 W chains of Q steps (Q even, W the least
// with Q <= T); unit = 4 chains of one order with its scaled-recurrence tables (restarted every 32
// steps), start states, reload rows and ring bands for the extended range.
struct slx_walk {
  int T = 0, nunit = 0, qsmax = 0;
  int4* ua = nullptr;     // [nunit] (m, Q, Qpad, table offset)
  int4* ub = nullptr;     // [nunit] (sub-segments per chain nk, reload row offset, reload mask, first chain)
  double *qc = nullptr, *qg = nullptr;  // [table] per (unit, chain, t < Qpad)
  double2 *seed = nullptr, *rl = nullptr;
  int2* band = nullptr;   // [nunit][4][8] rings [j0, j1) reloading at boundary k (entry k - 1)
  std::vector<int4> h_ua;
};

// ------------------------------------------------------------------ stage timing
// Stages of a profiled step (the keys of the reference's step(profile=True), solver.py:829-831):
// events recorded on the stream after each group of launches; the time between two consecutive
// events belongs to the later event's stage (stage -1: not attributed, e.g. the host's gap before
// a step starts).
enum { SLX_STAGE_MEMORY = 0, SLX_STAGE_SHT, SLX_STAGE_SPEC, SLX_STAGE_GRID, SLX_STAGE_ROT, SLX_NSTAGE };
struct slx_stage_marks {
  std::vector<hipEvent_t> ev;
  std::vector<int> stage;
  int n = 0;
  hipError_t err = hipSuccess;
};
static inline void slx_mark(const slx_sht* h, int stage, hipStream_t s) {
  slx_stage_marks* m = h->marks;
  if (!m || m->err != hipSuccess) return;
  if (m->n >= 16384) {  // read out with slx_solver_stage_times at least every few steps
    m->err = hipErrorOutOfMemory;
    return;
  }
  if (m->n == (int)m->ev.size()) {
    hipEvent_t e;
    m->err = hipEventCreate(&e);
    if (m->err != hipSuccess) return;
    m->ev.push_back(e);
    m->stage.push_back(-1);
  }
  m->err = hipEventRecord(m->ev[m->n], s);
  if (m->err == hipSuccess) m->stage[m->n++] = stage;
}

static inline int slx_moff(int m, int L) { return m * (L + 1) - m * (m - 1) / 2; }
