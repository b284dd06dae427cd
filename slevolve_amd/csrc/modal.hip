// Recursive viscoelastic modal state: fused decay-and-accumulate (FP64, HBM-bound).
#include "sht_launch.h"
#include "slx.h"

namespace slx {

// R_k <- d R_k + (1 - d) P ; visc = sum_k a_sl R_k ; rad = sum_k a_rr R_k, per packed (l, m),
// modes k < nmod[l] in order (python/slevolve/kernels.py:52-67).  Contraction is off so every
// product and sum rounds as the NumPy statements do.  NT: the states are read and written
// non-temporally (states too large to stay in the Infinity Cache from one step to the next).
typedef double state2 __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ double2 state_load(const double2* p) {
  if constexpr (NT) {
    const state2 v = __builtin_nontemporal_load(reinterpret_cast<const state2*>(p));
    return make_double2(v.x, v.y);
  } else {
    return *p;
  }
}
template <bool NT>
__device__ __forceinline__ void state_store(double2* p, double x, double y) {
  if constexpr (NT) {
    state2 v;
    v.x = x;
    v.y = y;
    __builtin_nontemporal_store(v, reinterpret_cast<state2*>(p));
  } else {
    *p = make_double2(x, y);
  }
}
template <bool NT>
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
    const double2 r = state_load<NT>(R + (long)k * ncoef + i);
    double xr = r.x * d + p.x * o;
    double xi = r.y * d + p.y * o;
    state_store<NT>(R + (long)k * ncoef + i, xr, xi);
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
  // non-temporal state access once the states fill half the 256 MB Infinity Cache (the step streams
  // its other arrays through it between two advances): 123 -> 113 us per launch at lmax 1023 (252 MB
  // of states); at lmax 512 (63 MB) the states stay cached and the temporal form is faster (24 us
  // against 35 us)
  const bool nt = (double)n * kmax * sizeof(double2) >= 128.0 * (1 << 20);
  if (nt)
    k_modal_advance<true><<<(n + 255) / 256, 256, 0, s>>>(n, h->lmax + 1, kmax, h->l_of, nmod, decay, omd, a_sl,
                                                          a_rr, R, P, visc, rad);
  else
    k_modal_advance<false><<<(n + 255) / 256, 256, 0, s>>>(n, h->lmax + 1, kmax, h->l_of, nmod, decay, omd, a_sl,
                                                           a_rr, R, P, visc, rad);
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
