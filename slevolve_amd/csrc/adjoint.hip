// Transposed transforms and the reverse modal recursion: the linear building blocks of the
// adjoint of the time step (slevolve_amd/adjoint.py).  Both transposes are taken with respect to
// the real inner products of the arrays as stored (a grid's values; a packed spectrum's real and
// imaginary parts), so <A g, U> = <g, A^T U> and <S U, g> = <U, S^T g> hold to round-off.
//
//   A (analysis):  U_lm = sum_j P~_lm(mu_j) w_j F_jm / nlon / 1.41421356,  F_j = rfft(row j)
//   S (synthesis): g_j  = irfft(nlon 1.41421356 sum_l U_lm P~_lm(mu_j))   (Im of m = 0 and of the
//                  Nyquist order ignored)
// The transpose of rfft is nlon irfft(c F) and that of irfft is d rfft / nlon, with c_m = 1 at
// m = 0 and at the Nyquist order (2 m = nlon), 1/2 otherwise, and d_m = 1 / c_m (and Im of those
// orders dropped), which gives
//   A^T U = w_j / (nlon 1.41421356^2) S(c U),       S^T g = d_m 1.41421356^2 nlon A(g / w_j),
// evaluated here with the FP64 transform kernels and the scalings applied as the coefficients and
// grid values are read or stored.
#include "sht_launch.h"
#include "slx.h"

namespace slx {
namespace {

constexpr double SQ2L2 = SLX_SQRT2_LEGACY * SLX_SQRT2_LEGACY;

// c_m U (coefficients of S for A^T)
struct CoefAnaT {
  const double2* c;
  long fstride;
  int nlon;
  __device__ bool prologue_needed(int, int) const { return false; }
  __device__ void prologue(double*) const {}
  __device__ double2 operator()(int f, int idx, int, int m, const double*) const {
    const double2 v = c[f * fstride + idx];
    const double s = (m == 0 || 2 * m == nlon) ? 1.0 : 0.5;
    return make_double2(v.x * s, v.y * s);
  }
};
// grid row r times w_r / (nlon 1.41421356^2) (wn, ws hold w / nlon of the ring pair's rows)
struct StoreAnaT {
  double* g;
  long fstride;
  const double* wn;
  const double* ws;
  int nlat, nlon;
  __device__ void operator()(int f, long i, double v) const {
    const int r = (int)(i / nlon);
    const double w = r < nlat / 2 ? wn[r] : ws[nlat - 1 - r];
    g[f * fstride + i] = v * (w / SQ2L2);
  }
};
// grid value times 1.41421356^2 nlon / w_r (input of A for S^T)
struct LoadSynT {
  const double* g;
  long fstride;
  const double* wn;
  const double* ws;
  int nlat, nlon;
  __device__ double operator()(int f, long i) const {
    const int r = (int)(i / nlon);
    const double w = r < nlat / 2 ? wn[r] : ws[nlat - 1 - r];
    return g[f * fstride + i] * (SQ2L2 / w);
  }
};

// d_m and the dropped imaginary parts of S^T
__global__ void k_syn_t_orders(double2* __restrict__ c, long n, int ncoef, const int* __restrict__ m_of, int nlon) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = m_of[i % ncoef];
  double2 v = c[i];
  if (m == 0 || 2 * m == nlon) {
    v.y = 0.0;
  } else {
    v.x *= 2.0;
    v.y *= 2.0;
  }
  c[i] = v;
}

// Reverse of one modal advance (python/slevolve/solver.py:386-422 run backwards): with the
// cotangents visc_bar, rad_bar of the read-out and Rbar of the advanced states,
//   Rbar_k <- decay_k (Rbar_k + a_sl_k visc_bar + a_rr_k rad_bar),
//   Pbar   += sum_k omd_k (Rbar_k + a_sl_k visc_bar + a_rr_k rad_bar)   (k < nmod[l], in k order).
__global__ void __launch_bounds__(256) k_modal_reverse(int ncoef, int nl, int kmax, const int* __restrict__ l_of,
                                                       const int* __restrict__ nmod, const double* __restrict__ decay,
                                                       const double* __restrict__ omd, const double* __restrict__ a_sl,
                                                       const double* __restrict__ a_rr, double2* __restrict__ Rbar,
                                                       const double2* __restrict__ vbar,
                                                       const double2* __restrict__ rbar, double2* __restrict__ Pbar) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncoef) return;
  const int l = l_of[i];
  const int nk = nmod[l] < kmax ? nmod[l] : kmax;
  const double2 v = vbar[i], q = rbar[i];
  double2 p = Pbar[i];
  for (int k = 0; k < nk; ++k) {
    const long o = (long)k * nl + l;
    const double s = a_sl[o], a = a_rr[o];
    const double2 r = Rbar[(long)k * ncoef + i];
    const double2 rn = make_double2(r.x + s * v.x + a * q.x, r.y + s * v.y + a * q.y);
    p.x += omd[o] * rn.x;
    p.y += omd[o] * rn.y;
    Rbar[(long)k * ncoef + i] = make_double2(decay[o] * rn.x, decay[o] * rn.y);
  }
  Pbar[i] = p;
}

}  // namespace
}  // namespace slx

using namespace slx;

extern "C" int slx_sht_analysis_adjoint(slx_sht* h, const void* coef_v, int nf, double* grid, void* stream) {
  const double2* coef = static_cast<const double2*>(coef_v);
  SLX_CHECK(h && coef && grid, "slx_sht_analysis_adjoint: null argument");
  SLX_CHECK(nf >= 1 && nf <= h->max_fields, "slx_sht_analysis_adjoint: nf out of range");
  hipStream_t s = (hipStream_t)stream;
  for (int f = 0; f < nf; ++f) {
    const int rc = launch_leg_syn<1>(h, CoefAnaT{coef + (long)f * h->ncoef, 0, h->nlon}, Pred{}, s, f);
    if (rc) return rc;
  }
  const ShtDev d = dev_of(h);
  return launch_fft_syn(h, StoreAnaT{grid, (long)h->nlat * h->nlon, d.wn, d.ws, h->nlat, h->nlon}, nf, Pred{}, s);
}

extern "C" int slx_sht_synthesis_adjoint(slx_sht* h, const double* grid, int nf, void* coef_v, void* stream) {
  double2* coef = static_cast<double2*>(coef_v);
  SLX_CHECK(h && coef && grid, "slx_sht_synthesis_adjoint: null argument");
  SLX_CHECK(nf >= 1 && nf <= h->max_fields, "slx_sht_synthesis_adjoint: nf out of range");
  hipStream_t s = (hipStream_t)stream;
  const ShtDev d = dev_of(h);
  int rc = launch_fft_ana(h, LoadSynT{grid, (long)h->nlat * h->nlon, d.wn, d.ws, h->nlat, h->nlon}, nf, Pred{}, s);
  if (rc) return rc;
  const long n = h->ncoef;
  switch (nf) {
    case 1: rc = launch_leg_ana<1>(h, h->WS, h->WD, OutPtrs<1>{{coef}}, Pred{}, s); break;
    case 2: rc = launch_leg_ana<2>(h, h->WS, h->WD, OutPtrs<2>{{coef, coef + n}}, Pred{}, s); break;
    default: rc = launch_leg_ana<3>(h, h->WS, h->WD, OutPtrs<3>{{coef, coef + n, coef + 2 * n}}, Pred{}, s); break;
  }
  if (rc) return rc;
  const long tot = (long)nf * n;
  k_syn_t_orders<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(coef, tot, h->ncoef, h->m_of, h->nlon);
  SLX_LAUNCH_CHECK();
  return 0;
}

extern "C" int slx_modal_reverse(const slx_sht* h, int kmax, const int* nmod, const double* decay, const double* omd,
                                 const double* a_sl, const double* a_rr, void* Rbar, const void* visc_bar,
                                 const void* rad_bar, void* Pbar, void* stream) {
  SLX_CHECK(h && nmod && decay && omd && a_sl && a_rr && Rbar && visc_bar && rad_bar && Pbar,
            "slx_modal_reverse: null argument");
  SLX_CHECK(kmax >= 0, "kmax must be >= 0");
  const int n = h->ncoef;
  k_modal_reverse<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(
      n, h->lmax + 1, kmax, h->l_of, nmod, decay, omd, a_sl, a_rr, static_cast<double2*>(Rbar),
      static_cast<const double2*>(visc_bar), static_cast<const double2*>(rad_bar), static_cast<double2*>(Pbar));
  SLX_LAUNCH_CHECK();
  return 0;
}
