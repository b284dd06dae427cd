// Recursive viscoelastic modal state: fused decay-and-accumulate (FP64, HBM-bound).
#include "sht_launch.h"
#include "slx.h"

namespace slx {

// R_k <- d R_k + (1 - d) P ; visc = sum_k a_sl R_k ; rad = sum_k a_rr R_k, per packed (l, m),
// modes k < nmod[l] in order (python/slevolve/kernels.py:52-67).  Contraction is off so every
// product and sum rounds as the NumPy statements do.
__global__ void __launch_bounds__(256) k_modal_advance(int ncoef, int nl, int kmax, const int* __restrict__ l_of,
                                                       const int* __restrict__ nmod, const double* __restrict__ decay,
                                                       const double* __restrict__ omd, const double* __restrict__ a_sl,
                                                       const double* __restrict__ a_rr, double2* __restrict__ R,
                                                       const double2* __restrict__ P, double2* __restrict__ visc,
                                                       double2* __restrict__ rad) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncoef) return;
  const int l = l_of[i];
  const int nk = nmod[l] < kmax ? nmod[l] : kmax;
  const double2 p = P[i];
  double vr = 0.0, vi = 0.0, rr = 0.0, ri = 0.0;
  for (int k = 0; k < nk; ++k) {
    const double d = decay[k * nl + l], o = omd[k * nl + l], s = a_sl[k * nl + l], q = a_rr[k * nl + l];
    double2 r = R[(long)k * ncoef + i];
    double xr = r.x * d + p.x * o;
    double xi = r.y * d + p.y * o;
    R[(long)k * ncoef + i] = make_double2(xr, xi);
    vr += xr * s;
    vi += xi * s;
    rr += xr * q;
    ri += xi * q;
  }
  visc[i] = make_double2(vr, vi);
  rad[i] = make_double2(rr, ri);
}

int launch_modal(const slx_sht* h, int kmax, const int* nmod, const double* decay, const double* omd,
                 const double* a_sl, const double* a_rr, double2* R, const double2* P, double2* visc,
                 double2* rad, hipStream_t s) {
  const int n = h->ncoef;
  k_modal_advance<<<(n + 255) / 256, 256, 0, s>>>(n, h->lmax + 1, kmax, h->l_of, nmod, decay, omd, a_sl, a_rr,
                                                  R, P, visc, rad);
  SLX_LAUNCH_CHECK();
  return 0;
}

}  // namespace slx

extern "C" int slx_modal_advance_read(const slx_sht* h, int kmax, const int* nmod, const double* decay,
                                      const double* omd, const double* a_sl, const double* a_rr, void* R,
                                      const void* P, void* visc, void* rad, void* stream) {
  SLX_CHECK(h && nmod && decay && omd && a_sl && a_rr && R && P && visc && rad, "slx_modal_advance_read: null argument");
  SLX_CHECK(kmax >= 0, "kmax must be >= 0");
  return slx::launch_modal(h, kmax, nmod, decay, omd, a_sl, a_rr, static_cast<double2*>(R),
                           static_cast<const double2*>(P), static_cast<double2*>(visc), static_cast<double2*>(rad),
                           (hipStream_t)stream);
}
