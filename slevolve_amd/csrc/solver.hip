// Device-resident sea-level time step (FP64).  Restates python/slevolve/solver.py:521-1144 and
// the extrapolated start of solver_opt.py:148-165,504-527 with the grid and spectral algebra fused
// into the transform passes:
//   * rO - tO is analysed as one field (its analysis is all the update uses, solver.py:1002-1015);
//   * the ocean function C* is analysed in the same sweep as rO - tO of the next iteration;
//   * deltasl = dslxy + conserv K00 (one synthesis less per iteration, solver_opt.py:28-32);
//   * eq. B18 is assembled while the synthesis stages its coefficients.
// The inner iterations are launched ahead and predicated on a device convergence flag.
#include <cmath>
#include <cstring>

#include "sht_launch.h"
#include "slx.h"

namespace slx {
int launch_modal(const slx_sht* h, int kmax, const int* nmod, const double* decay, const double* omd,
                 const double* a_sl, const double* a_rr, double2* R, const double2* P, double2* visc,
                 double2* rad, hipStream_t s);
}

using namespace slx;

static constexpr double EPS0 = 2.220446049250313e-16;
static constexpr int GR_MAXACC = 3;
static constexpr int MAX_INNER = 9999;
static constexpr double PI_LEGACY = 3.1415926535898;
static constexpr int PSLOT = 1024;  // block-partial slots per reduction (3 of them: 2 iteration parities + guess)

// Scalar state of a run.  The status block comes first so it can be copied out as is.
struct DevState {
  slx_status st;
  int ice_changed, counter;
  int nflip[2], nacc2[2];  // grounded-ice flip count / accepted re-guesses, by iteration parity
  double sa, so;
  double p_il[9], p_m[3], p_lam[6];
  double dil[9], dm[3], dlam[6];
  double sum_il[9], sum_m[3], sum_tsl[6], sum_trr[6];
  double dsl_rot[6], rr_rot[6];
};

// Constants evaluated on the host once, in the reference's operation order.
struct Consts {
  double rhoi, rhow, gacc, K00, eps1, area;  // area = 4 PI a^2
  double cI33, cI13, cI23;                   // inertia factors (solver.py:915-920)
  double alpha, inv_CA, one_kE2, inv_kf, inv_C;
  double K0, K20, K21, K22;                  // Lambda factors (solver.py:955-959)
  double ekhTE, hTE2;
  int L, nl, ncoef, npts, nlat, nlon, kmax, nmod2, tpw, checkmarine, calcRG;
  int idx20, idx21, idx22;
};

struct slx_solver {
  slx_sht* h = nullptr;
  Consts c;
  int extrap = 0;
  int nhist = 0;         // converged increments stored for the extrapolated start
  int step = 0;
  int iters_enqueued = 0;
  // device arrays --------------------------------------------------------------------------
  DevState* st = nullptr;
  double* partial = nullptr;  // block partial sums
  double* rows = nullptr;     // [4][nlat] row sums
  double* w00 = nullptr;      // [nlat] degree-0 quadrature weights
  int* nmod = nullptr;
  double *s = nullptr, *decay = nullptr, *omd = nullptr, *a_sl = nullptr, *a_rr = nullptr;
  double *a2 = nullptr;       // [4][kmax] a_il, a_m, a_tsl, a_trr
  double *r2 = nullptr;       // [kmax][18] degree-2 states r_il(9), r_m(3), r_lambda(6)
  double *ttl = nullptr, *ttlekhl = nullptr, *ttlhe = nullptr;
  // grids
  double *tinit, *cxy0, *beta0, *cstar0, *cxy, *beta, *cstar, *tO, *icestar, *dsl, *rr, *G, *topo, *topo_prev,
      *ice_prev, *tmp;
  // spectra (packed)
  double2 *R, *p_load, *p_ice, *p_ocean, *prev_ice, *icestar_lm, *oldice_lm, *dice, *dI, *cstar_lm, *rt_lm,
      *load_inc, *dS_cur, *visc, *rad, *h1, *h2, *scratch_lm;
  std::vector<double> h_s;  // host copy of s [kmax][nl]
  slx_status* pinned = nullptr;
  double dt = -1.0;
  const double* ice_cur = nullptr;  // ice grid of the step in progress (device)
  int closed = 1;                   // host bookkeeping of the last step done
};

// ================================================================================== kernels
namespace {

struct Pts {  // all device pointers a kernel may need, passed by value
  DevState* st;
  double* partial;
  const double *tinit, *cxy0, *beta0, *cstar0;
  double *cxy, *beta, *cstar, *tO, *icestar, *dsl, *rr, *G, *topo, *topo_prev, *ice_prev, *tmp, *rows;
  const double* w00;
  const double* ice_cur;
  double2 *p_load, *p_ice, *p_ocean, *prev_ice, *icestar_lm, *oldice, *dice, *dI, *cstar_lm, *rt_lm, *load_inc,
      *dS_cur, *visc, *rad, *h1, *h2;
  const double *ttl, *ttlekhl, *ttlhe, *decay, *a2;
  double* r2;
  const int* l_of;
};

__device__ __forceinline__ double keep_factor(double ice, double topo, double rhow, double rhoi) {
  // (topo > 0) | (ice > |topo| rho_w / rho_i), as a 0/1 multiplier (solver.py:655-660)
  return (topo > 0.0 || ice > fabs(topo) * rhow / rhoi) ? 1.0 : 0.0;
}

__device__ __forceinline__ double cabs_d(double2 z) { return hypot(z.x, z.y); }

// Block sum of v (blockDim 256) -> lane 0; fixed order.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  return sh[0];
}

// rotation block (solver.py:910-975): products of inertia from the degree-2 load, Liouville
// equations, centrifugal potential Lambda and the rotational sea-level / radial terms, from the
// (2,0), (2,1) coefficients of the ocean-load increment dS_cur and the ice increment dI.
struct RotOut {
  double il[9], mm[3], lam[18], dil[9], dm[3], dlam[6], dsl[6], rr[6];
};

__device__ void rotation_eval(const DevState* s, const Consts& c, double2 dS20, double2 dS21, double2 dI20,
                              double2 dI21, RotOut& o) {
  for (int q = 0; q < 9; ++q) o.il[q] = s->st.il[q];
  for (int q = 0; q < 3; ++q) o.mm[q] = s->st.mm[q];
  for (int q = 0; q < 18; ++q) o.lam[q] = s->st.lam[q];
  const double z20r = c.rhow * dS20.x + c.rhoi * dI20.x;
  const double z21r = c.rhow * dS21.x + c.rhoi * dI21.x;
  const double z21i = c.rhow * dS21.y + c.rhoi * dI21.y;
  o.il[8] = c.cI33 * z20r;
  o.il[2] = c.cI13 * z21r;
  o.il[5] = c.cI23 * z21i;
  for (int q = 0; q < 9; ++q) o.dil[q] = o.il[q] - s->p_il[q];
  for (int i = 0; i < 2; ++i)
    o.mm[i] = c.alpha * (c.inv_CA * (c.one_kE2 * o.il[3 * i + 2] + s->sum_il[3 * i + 2]) + c.inv_kf * s->sum_m[i]);
  o.mm[2] = -c.inv_C * (c.one_kE2 * o.il[8] + s->sum_il[8]);
  for (int q = 0; q < 3; ++q) o.dm[q] = o.mm[q] - s->p_m[q];
  const double m1 = o.mm[0], m2 = o.mm[1], m3 = o.mm[2];
  o.lam[0] = c.K0 * ((m1 * m1 + m2 * m2 + m3 * m3) + 2.0 * m3);
  o.lam[1] = 0.0;
  o.lam[12] = c.K20 * (m1 * m1 + m2 * m2 - 2.0 * m3 * m3 - 4.0 * m3);
  o.lam[13] = 0.0;
  o.lam[14] = -c.K21 * (m1 * (1.0 + m3));
  o.lam[15] = -c.K21 * (-m2 * (1.0 + m3));
  o.lam[16] = c.K22 * (m2 * m2 - m1 * m1);
  o.lam[17] = c.K22 * (2.0 * m1 * m2);
  for (int q = 0; q < 6; ++q) o.dlam[q] = o.lam[12 + q] - s->p_lam[q];
  const double ig = 1.0 / c.gacc;
  for (int q = 0; q < 6; ++q) {
    const double tot = s->p_lam[q] + o.dlam[q];
    o.dsl[q] = (c.ekhTE * tot + s->sum_tsl[q]) * ig;
    o.rr[q] = c.calcRG ? (c.hTE2 * tot + s->sum_trr[q]) * ig : 0.0;
  }
}

__device__ void rotation_store(DevState* s, const RotOut& o) {
  for (int q = 0; q < 9; ++q) {
    s->st.il[q] = o.il[q];
    s->dil[q] = o.dil[q];
  }
  for (int q = 0; q < 3; ++q) {
    s->st.mm[q] = o.mm[q];
    s->dm[q] = o.dm[q];
  }
  for (int q = 0; q < 18; ++q) s->st.lam[q] = o.lam[q];
  for (int q = 0; q < 6; ++q) {
    s->dlam[q] = o.dlam[q];
    s->dsl_rot[q] = o.dsl[q];
    s->rr_rot[q] = o.rr[q];
  }
}

// ---------------------------------------------------------------------- set-up (step 0)
__global__ void k_init_grid(Pts p, Consts c, const double* topo0, const double* ice0) {
  const long n = c.npts;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double t = topo0[i], ice = ice0[i];
    double cx = t < 0.0 ? 1.0 : 0.0;
    double g0 = c.checkmarine ? ice * keep_factor(ice, t, c.rhow, c.rhoi) : ice;
    double b0 = g0 < EPS0 ? 1.0 : 0.0;
    const_cast<double*>(p.tinit)[i] = t;
    const_cast<double*>(p.cxy0)[i] = cx;
    const_cast<double*>(p.beta0)[i] = b0;
    const_cast<double*>(p.cstar0)[i] = cx * b0;
    p.cxy[i] = cx;
    p.topo_prev[i] = t;
    p.topo[i] = t;
    p.ice_prev[i] = ice;
    p.tmp[i] = g0;
    p.beta[i] = b0;
    p.icestar[i] = g0;
  }
}

// ---------------------------------------------------------------------- step begin
// degree-2 rotational states (solver.py:407-421): block q < 18 advances state component q
// (r_il 9, r_m 3, r_lambda 6 reals) for every mode, one lane per mode, and sums the weighted
// states over the modes in a fixed order; block 18 does the per-step resets.
__global__ void __launch_bounds__(64) k_begin(Pts p, Consts c) {
  DevState* s = p.st;
  const int q = blockIdx.x, lane = threadIdx.x;
  if (q < 18) {
    const int row = q < 9 ? 0 : (q < 12 ? 1 : 2);
    const double pv = q < 9 ? s->p_il[q] : (q < 12 ? s->p_m[q - 9] : s->p_lam[q - 12]);
    double sum1 = 0.0, sum2 = 0.0;
    if (c.L >= 2) {
      for (int k = lane; k < c.nmod2; k += 64) {
        const double d = p.decay[k * c.nl + 2], o = 1.0 - d;
        double* r = p.r2 + 18 * k + q;
        const double rv = d * *r + o * pv;
        *r = rv;
        sum1 += p.a2[row * c.kmax + k] * rv;
        if (row == 2) sum2 += p.a2[3 * c.kmax + k] * rv;
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      sum1 += __shfl_xor(sum1, off, 64);
      sum2 += __shfl_xor(sum2, off, 64);
    }
    if (lane == 0) {
      if (q < 9) {
        s->sum_il[q] = sum1;
        s->dil[q] = 0.0;
      } else if (q < 12) {
        s->sum_m[q - 9] = sum1;
        s->dm[q - 9] = 0.0;
      } else {
        s->sum_tsl[q - 12] = sum1;
        s->sum_trr[q - 12] = sum2;
        s->dsl_rot[q - 12] = 0.0;
        s->rr_rot[q - 12] = 0.0;
        s->dlam[q - 12] = 0.0;
      }
    }
  } else if (lane == 0) {
    s->st.numiter = 0;
    s->st.done = 0;
    s->st.failed = 0;
    s->st.nacc = 0;
    s->st.finished = 0;
    s->st.step += 1;
    s->ice_changed = 0;
    s->nflip[0] = s->nflip[1] = 0;
    s->nacc2[0] = s->nacc2[1] = 0;
    s->counter = 0;
  }
}

// grounded ice on the previous converged topography, beta, C*, tO (solver.py:853-887)
__global__ void k_ice_grid(Pts p, Consts c) {
  const long n = c.npts;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double ice = p.ice_cur[i];
    double ist = c.checkmarine ? ice * keep_factor(ice, p.topo_prev[i], c.rhow, c.rhoi) : ice;
    double b = ist < EPS0 ? 1.0 : 0.0;
    double cs = p.cxy[i] * b;
    p.icestar[i] = ist;
    p.beta[i] = b;
    p.cstar[i] = cs;
    p.tO[i] = p.tinit[i] * (cs - p.cstar0[i]);
  }
}

// ice increments and the first guess of the ocean-load increment (solver.py:857-892;
// solver_opt.py:504-527); so = sum |load_inc|
struct GuessArgs {
  int extrap_ready, uniform;
  double c1, c2;
  const double2* warm;  // warm start: replaces the guess (deltaS_cur = p_ocean + warm)
};
__global__ void __launch_bounds__(256) k_guess(Pts p, Consts c, GuessArgs g) {
  __shared__ double sh[256];
  const int n = c.ncoef;
  const double2 c00 = p.cstar_lm[0];
  const double2 d00 = make_double2(p.icestar_lm[0].x - p.oldice[0].x, p.icestar_lm[0].y - p.oldice[0].y);
  const double scl = 1.0 / c00.x;
  const double fr = -c.rhoi / c.rhow;
  const double2 e = make_double2(fr * d00.x, fr * d00.y);
  double a = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double2 old = p.oldice[i], ist = p.icestar_lm[i];
    const double2 d = make_double2(ist.x - old.x, ist.y - old.y);
    const double2 pi = p.p_ice[i];
    p.dice[i] = d;
    p.dI[i] = make_double2(pi.x + d.x, pi.y + d.y);
    double2 ld;
    if (g.extrap_ready) {
      const double2 a1 = p.h1[i], a2 = p.h2[i];
      if (g.uniform)
        ld = make_double2(a1.x * 2.0 - a2.x, a1.y * 2.0 - a2.y);
      else
        ld = make_double2(a1.x * g.c1 - g.c2 * a2.x, a1.y * g.c1 - g.c2 * a2.y);
      const double2 sp = p.p_ocean[i];
      p.dS_cur[i] = make_double2(sp.x + ld.x, sp.y + ld.y);
    } else {
      // (cstar_lm / cstar_lm[0,0]) * ((-rho_i/rho_w) dice[0,0]); cstar_lm[0,0] and dice[0,0] are real
      const double2 q = make_double2(p.cstar_lm[i].x * scl, p.cstar_lm[i].y * scl);
      ld = make_double2(q.x * e.x - q.y * e.y, q.x * e.y + q.y * e.x);
      p.dS_cur[i] = make_double2(0.0, 0.0);
    }
    if (g.warm) {
      ld = g.warm[i];
      const double2 sp = p.p_ocean[i];
      p.dS_cur[i] = make_double2(sp.x + ld.x, sp.y + ld.y);
    }
    p.load_inc[i] = ld;
    a += cabs_d(ld);
  }
  const double bs = block_sum(a, sh);
  if (threadIdx.x == 0) p.partial[2 * PSLOT + blockIdx.x] = bs;  // sum |load_inc| of the guess
}

// ---------------------------------------------------------------------- inner iteration
// eq. B18 coefficient, assembled while the synthesis stages its coefficients:
// w1[l] ((rho_i dI + rho_w dS_prev) + rho_w load_inc) + w2[l] resp (+ rotation at l = 2, m <= 2).
// With eval set (sea-level synthesis of an inner iteration) every block evaluates the rotation
// block in its prologue from dS_cur and dI (identical arithmetic everywhere) and block 0
// publishes it; otherwise (radial synthesis after the loop) the stored rr terms are used.
struct CoefB18 {
  const double2 *dI, *dSp, *ld, *resp, *dS;
  const double *w1, *w2;
  DevState* st;
  Consts c;
  int eval;
  // only the block that stores the rotation (block 0) and those holding the degree-2, m <= 2
  // coefficients need it
  __device__ bool prologue_needed(int lo, int hi) const {
    auto in = [&](int i) { return i >= lo && i < hi; };
    return lo == 0 || in(c.idx20) || in(c.idx21) || in(c.idx22);
  }
  __device__ void prologue(double* pro) const {
    if (threadIdx.x != 0) return;
    if (eval) {
      RotOut o;
      if (c.tpw) {
        rotation_eval(st, c, dS[c.idx20], dS[c.idx21], dI[c.idx20], dI[c.idx21], o);
        if (blockIdx.x == 0 && blockIdx.y == 0) rotation_store(st, o);
      } else {
        for (int q = 0; q < 6; ++q) o.dsl[q] = 0.0;
      }
      for (int q = 0; q < 6; ++q) pro[q] = o.dsl[q];
    } else {
      for (int q = 0; q < 6; ++q) pro[q] = c.tpw ? st->rr_rot[q] : 0.0;
    }
  }
  __device__ double2 operator()(int, int idx, int l, int m, const double* pro) const {
    const double2 a = dI[idx], b = dSp[idx], q = ld[idx], v = resp[idx];
    const double ar = (a.x * c.rhoi + b.x * c.rhow) + q.x * c.rhow;
    const double ai = (a.y * c.rhoi + b.y * c.rhow) + q.y * c.rhow;
    const double u = w1[l], t = w2[l];
    double2 r = make_double2(ar * u + v.x * t, ai * u + v.y * t);
    if (l == 2 && m <= 2) {
      r.x += pro[2 * m];
      r.y += pro[2 * m + 1];
    }
    return r;
  }
};

// analysis input of the inner sweep: field 0 = dslxy C* - tO (rO - tO), field 1 = C*
struct LoadRT {
  const double *dsl, *cstar, *tO;
  __device__ double operator()(int f, long i) const { return f == 0 ? dsl[i] * cstar[i] - tO[i] : cstar[i]; }
};
template <int NF>
struct LoadPtrs {
  const double* g[NF];
  __device__ double operator()(int f, long i) const { return g[f][i]; }
};

// Ring-pair pass of an inner iteration: the synthesised Fourier coefficients of eq. B18 are
// inverse-transformed (one complex FFT for rows j and nlat-1-j), written as dslxy, multiplied by
// C* and reduced by tO in LDS (rO - tO, solver.py:993-996), forward-transformed and folded into
// analysis field 0; with nf == 2 the ocean function C* of the same rows is transformed into
// analysis field 1 as well.  Replaces separate synthesis and analysis ring passes.
template <int NT>
__global__ void __launch_bounds__(NT) k_fft_synana(ShtDev d, const double2* __restrict__ FN,
                                                   const double2* __restrict__ FS, double* __restrict__ dsl,
                                                   const double* __restrict__ cstar, const double* __restrict__ tO,
                                                   double2* __restrict__ WS, double2* __restrict__ WD, int nf,
                                                   Pred pr) {
  if (!pr.active()) return;
  extern __shared__ double2 smem[];
  const int N = d.nlon, L = d.lmax;
  double2* a = smem;
  double2* b = smem + N;
  double2* tws = smem + 2 * N;  // quarter-period twiddles, staged once per block
  const int j = blockIdx.x, js = d.nlat - 1 - j;
  const long rn = (long)j * N, rs = (long)js * N;
  for (int i = threadIdx.x; i < N; i += NT) a[i] = make_double2(0.0, 0.0);
  stage_tw4<NT>(tws, d.tw, N);
  __syncthreads();
  for (int m = threadIdx.x; m <= L; m += NT) {
    const double2 fn = FN[(long)m * d.nh + j], fs = FS[(long)m * d.nh + j];
    if (m == 0 || 2 * m == N) {
      a[m] = make_double2(fn.x, fs.x);
    } else {
      a[m] = make_double2(fn.x - fs.y, fn.y + fs.x);
      a[N - m] = make_double2(fn.x + fs.y, -fn.y + fs.x);
    }
  }
  __syncthreads();
  double2* z = fft_lds<NT>(a, b, tws, N, d.log2n, true);
  double2* w = (z == a) ? b : a;
  for (int i = threadIdx.x; i < N; i += NT) {
    const double2 v = z[i];
    dsl[rn + i] = v.x;
    dsl[rs + i] = v.y;
    z[i] = make_double2(v.x * cstar[rn + i] - tO[rn + i], v.y * cstar[rs + i] - tO[rs + i]);
  }
  __syncthreads();
  const double wn = d.wn[j], ws = d.ws[j];
  const long fstride = (long)d.nm * d.nh;
  for (int f = 0; f < nf; ++f) {
    if (f == 1) {
      for (int i = threadIdx.x; i < N; i += NT) z[i] = make_double2(cstar[rn + i], cstar[rs + i]);
      __syncthreads();
    }
    double2* Z = fft_lds<NT>(z, w, tws, N, d.log2n, false);
    for (int m = threadIdx.x; m <= L; m += NT) {
      const double2 zm = Z[m], zc = Z[(N - m) & (N - 1)];
      const double2 fn = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y - zc.y));
      const double2 fs = make_double2(0.5 * (zm.y + zc.y), -0.5 * (zm.x - zc.x));
      const double2 tn = make_double2(fn.x * wn, fn.y * wn), ts = make_double2(fs.x * ws, fs.y * ws);
      WS[f * fstride + (long)m * d.nh + j] = cadd(tn, ts);
      WD[f * fstride + (long)m * d.nh + j] = csub(tn, ts);
    }
    __syncthreads();
    if (Z != z) {  // keep z as the buffer the next field is loaded into
      double2* t = z;
      z = w;
      w = t;
    }
  }
}

// conservation constant and ocean-load update (solver.py:1002-1015); per-block partial sums of
// |load_inc| for the convergence ratio, which the refresh kernel forms
__device__ __forceinline__ double conserv_of(const Pts& p, const Consts& c) {
  return (1.0 / p.cstar_lm[0].x) * (-(c.rhoi / c.rhow) * p.dI[0].x - p.rt_lm[0].x);
}

__global__ void __launch_bounds__(256) k_update(Pts p, Consts c, int it) {
  __shared__ double sh[256];
  DevState* s = p.st;
  if (*(volatile int*)&s->st.done) return;
  const int n = c.ncoef;
  const double conserv = conserv_of(p, c);
  double a = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double2 sp = p.p_ocean[i], rt = p.rt_lm[i], cs = p.cstar_lm[i];
    const double2 nw = make_double2((-sp.x + rt.x) + cs.x * conserv, (-sp.y + rt.y) + cs.y * conserv);
    p.load_inc[i] = nw;
    p.dS_cur[i] = make_double2(sp.x + nw.x, sp.y + nw.y);
    a += cabs_d(nw);
  }
  const double bs = block_sum(a, sh);
  if (threadIdx.x == 0) p.partial[(it & 1) * PSLOT + blockIdx.x] = bs;
}

// topography / ocean-function re-guess and the grounded-ice flip count (solver.py:1032-1040,661-684).
// it = iteration index within the step: the counters of this iteration live in slot it & 1.
// Every block first forms xi = |sum|new| - sum|old|| / sum|old| (solver.py:1023-1030) from the
// update's block partials in a fixed order (sum|old| = the previous iteration's partials, or the
// guess's), so that all blocks take the same decision; block 0 publishes it.
__global__ void __launch_bounds__(256) k_refresh(Pts p, Consts c, int it, int nbu, int nbg) {
  __shared__ double sh[256];
  __shared__ int shf;
  DevState* s = p.st;
  if (*(volatile int*)&s->st.done) return;
  double v = 0.0;
  for (int b = threadIdx.x; b < nbu; b += blockDim.x) v += p.partial[(it & 1) * PSLOT + b];
  const double sa = block_sum(v, sh);
  __syncthreads();
  v = 0.0;
  if (it == 0) {
    for (int b = threadIdx.x; b < nbg; b += blockDim.x) v += p.partial[2 * PSLOT + b];
  } else {
    for (int b = threadIdx.x; b < nbu; b += blockDim.x) v += p.partial[((it - 1) & 1) * PSLOT + b];
  }
  const double so = block_sum(v, sh);
  double xi;
  if (fabs(sa - so) < EPS0 && fabs(so) < EPS0)
    xi = 0.0;
  else if (fabs(so) < EPS0)
    xi = fabs((sa - so) / (EPS0 * 10));
  else
    xi = fabs((sa - so) / so);
  const double conserv = conserv_of(p, c);
  const bool conv = xi <= c.eps1;
  const bool fail = !conv && it + 1 >= MAX_INNER;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s->st.numiter = it + 1;
    s->st.xi = xi;
    s->st.conserv = conserv;
    s->sa = sa;
    s->so = so;
    if (conv || fail) s->st.done = 1;
    if (fail) s->st.failed = 1;
  }
  if (conv || fail) return;
  if (threadIdx.x == 0) shf = 0;
  __syncthreads();
  const long n = c.npts;
  const double cadd = conserv * c.K00;
  const bool regs = c.checkmarine && s->nacc2[it & 1] < GR_MAXACC;
  int flips = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double t = p.tinit[i] - (p.dsl[i] + cadd);
    p.topo[i] = t;
    p.cxy[i] = t < 0.0 ? 1.0 : 0.0;
    if (regs) {
      const double ice = p.ice_cur[i];
      const double cand = ice * keep_factor(ice, t, c.rhow, c.rhoi);
      const bool bnew = cand < EPS0;
      if (bnew != (p.beta[i] != 0.0)) flips += 1;
    }
  }
  if (flips) atomicAdd(&shf, flips);
  __syncthreads();
  if (threadIdx.x == 0 && shf) atomicAdd(&s->nflip[it & 1], shf);
}

// apply the re-guess when a grounded flag flipped (at most GR_MAXACC times per step); C*, tO
__global__ void k_refresh2(Pts p, Consts c, int it) {
  DevState* s = p.st;
  if (*(volatile int*)&s->st.done) return;
  const int slot = it & 1;
  const bool regs = c.checkmarine && s->nacc2[slot] < GR_MAXACC;
  const bool changed = regs && s->nflip[slot] > 0;
  const long n = c.npts;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    if (changed) {
      const double ice = p.ice_cur[i];
      const double cand = ice * keep_factor(ice, p.topo[i], c.rhow, c.rhoi);
      p.icestar[i] = cand;
      p.beta[i] = cand < EPS0 ? 1.0 : 0.0;
    }
    const double cs = p.cxy[i] * p.beta[i];
    p.cstar[i] = cs;
    p.tO[i] = p.tinit[i] * (cs - p.cstar0[i]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s->ice_changed = changed ? 1 : 0;
    s->nacc2[slot ^ 1] = s->nacc2[slot] + (changed ? 1 : 0);
    s->nflip[slot ^ 1] = 0;
    if (regs) s->st.reguess += 1;
    if (changed) {
      s->st.ice_analyses += 1;
      s->st.nacc += 1;
    }
  }
}

__global__ void k_dice(Pts p, Consts c) {
  DevState* s = p.st;
  if (*(volatile int*)&s->st.done || !*(volatile int*)&s->ice_changed) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.ncoef) return;
  const double2 ist = p.icestar_lm[i], old = p.oldice[i], pi = p.p_ice[i];
  const double2 d = make_double2(ist.x - old.x, ist.y - old.y);
  p.dice[i] = d;
  p.dI[i] = make_double2(pi.x + d.x, pi.y + d.y);
}

// ---------------------------------------------------------------------- step finish
// commit (solver.py:1080-1095): load increment without degree 0 into P; ice, ocean, previous ice
__device__ __forceinline__ bool finish_active(const DevState* s) {
  return *(volatile const int*)&s->st.done && !*(volatile const int*)&s->st.finished;
}

__global__ void k_commit(Pts p, Consts c, int extrap) {
  if (!finish_active(p.st)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.ncoef) return;
  const int l = p.l_of[i];
  const double2 d = p.dice[i], q = p.load_inc[i];
  double2 cur = make_double2(d.x * c.rhoi + q.x * c.rhow, d.y * c.rhoi + q.y * c.rhow);
  if (l == 0) cur = make_double2(0.0, 0.0);
  const double2 pl = p.p_load[i];
  p.p_load[i] = make_double2(pl.x + cur.x, pl.y + cur.y);
  p.p_ice[i] = p.dI[i];
  p.p_ocean[i] = p.dS_cur[i];
  p.prev_ice[i] = p.icestar_lm[i];
  if (extrap) p.h2[i] = q;  // becomes the newest history entry (the host swaps h1/h2)
}

// final topography, ocean function, G and the row sums of the degree-0 diagnostics (one block per row)
__global__ void __launch_bounds__(256) k_post_grid(Pts p, Consts c) {
  if (!finish_active(p.st)) return;
  __shared__ double sh[256];
  const int j = blockIdx.x;
  const double cadd = p.st->st.conserv * c.K00;
  double a = 0.0, b = 0.0, e = 0.0, f = 0.0;
  for (int i = threadIdx.x; i < c.nlon; i += blockDim.x) {
    const long k = (long)j * c.nlon + i;
    const double dsl = p.dsl[k] + cadd;
    const double t = p.tinit[k] - dsl;
    const double o = t < 0.0 ? 1.0 : 0.0;
    p.topo[k] = t;
    p.topo_prev[k] = t;
    p.cxy[k] = o;
    if (c.calcRG) p.G[k] = dsl + p.rr[k];
    const double sl = dsl * o, bt = p.beta[k];
    a += sl;
    b += o;
    e += sl * bt;
    f += o * bt;
  }
  double r;
  r = block_sum(a, sh);
  if (threadIdx.x == 0) p.rows[j] = r;
  __syncthreads();
  r = block_sum(b, sh);
  if (threadIdx.x == 0) p.rows[c.nlat + j] = r;
  __syncthreads();
  r = block_sum(e, sh);
  if (threadIdx.x == 0) p.rows[2 * c.nlat + j] = r;
  __syncthreads();
  r = block_sum(f, sh);
  if (threadIdx.x == 0) p.rows[3 * c.nlat + j] = r;
}

// degree-0 diagnostics and the rotational commit (solver.py:1060-1110; solver_opt.py:525-547)
__global__ void __launch_bounds__(256) k_finish(Pts p, Consts c) {
  __shared__ double sh[256];
  DevState* s = p.st;
  if (!finish_active(s)) return;
  double q[4];
  for (int k = 0; k < 4; ++k) {
    double v = 0.0;
    for (int j = threadIdx.x; j < c.nlat; j += blockDim.x) v += p.w00[j] * p.rows[k * c.nlat + j];
    q[k] = block_sum(v, sh);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  s->st.gmslc_ocn = q[0] / q[1];
  s->st.gmslc_ocnBeta = q[2] / q[3];
  s->st.ocean_area = q[1] * c.area;
  s->st.oceanBeta_area = q[3] * c.area;
  s->st.ice_volume = p.icestar_lm[0].x * c.area;
  if (c.tpw) {
    for (int k = 0; k < 9; ++k) s->p_il[k] += s->dil[k];
    for (int k = 0; k < 3; ++k) s->p_m[k] += s->dm[k];
    for (int k = 0; k < 6; ++k) s->p_lam[k] += s->dlam[k];
  }
  s->st.finished = 1;
}

__global__ void k_init_areas(Pts p, Consts c, const double2* c0, const double2* cs0, const double2* i0) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  DevState* s = p.st;
  s->st.ocean_area0 = c0[0].x * c.area;
  s->st.oceanBeta_area0 = cs0[0].x * c.area;
  s->st.ice_volume0 = i0[0].x * c.area;
}

}  // namespace

// ============================================================================== host side
static Pts make_pts(slx_solver* s, const double* ice_cur, const double2* oldice) {
  Pts p;
  p.st = s->st;
  p.partial = s->partial;
  p.tinit = s->tinit;
  p.cxy0 = s->cxy0;
  p.beta0 = s->beta0;
  p.cstar0 = s->cstar0;
  p.cxy = s->cxy;
  p.beta = s->beta;
  p.cstar = s->cstar;
  p.tO = s->tO;
  p.icestar = s->icestar;
  p.dsl = s->dsl;
  p.rr = s->rr;
  p.G = s->G;
  p.topo = s->topo;
  p.topo_prev = s->topo_prev;
  p.ice_prev = s->ice_prev;
  p.tmp = s->tmp;
  p.rows = s->rows;
  p.w00 = s->w00;
  p.ice_cur = ice_cur;
  p.p_load = s->p_load;
  p.p_ice = s->p_ice;
  p.p_ocean = s->p_ocean;
  p.prev_ice = s->prev_ice;
  p.icestar_lm = s->icestar_lm;
  p.oldice = const_cast<double2*>(oldice);
  p.dice = s->dice;
  p.dI = s->dI;
  p.cstar_lm = s->cstar_lm;
  p.rt_lm = s->rt_lm;
  p.load_inc = s->load_inc;
  p.dS_cur = s->dS_cur;
  p.visc = s->visc;
  p.rad = s->rad;
  p.h1 = s->h1;
  p.h2 = s->h2;
  p.ttl = s->ttl;
  p.ttlekhl = s->ttlekhl;
  p.ttlhe = s->ttlhe;
  p.decay = s->decay;
  p.a2 = s->a2;
  p.r2 = s->r2;
  p.l_of = s->h->l_of;
  return p;
}

static int grid_blocks(long n) {
  long b = (n + 255) / 256;
  return (int)(b > 2048 ? 2048 : b);
}
// blocks of the spectral kernels that post block partials (<= PSLOT)
static int red_blocks(long n) {
  long b = (n + 511) / 512;
  return (int)(b > 512 ? 512 : (b < 1 ? 1 : b));
}

extern "C" int slx_solver_create(slx_sht* h, const slx_solver_params* p, const slx_outputs* o, slx_solver** out) {
  SLX_CHECK(h && p && o && out, "slx_solver_create: null argument");
  SLX_CHECK(o->tgrid && o->ocean && o->beta && o->R && o->G, "slx_solver_create: all output grids are required");
  SLX_CHECK(p->lmax == h->lmax, "solver lmax must equal the transform's lmax");
  SLX_CHECK(h->max_fields >= 3, "the solver needs a transform handle with max_fields >= 3");
  SLX_CHECK(p->kmax >= 0 && p->nmod2 >= 0 && p->nmod2 <= p->kmax, "bad mode counts");
  SLX_CHECK(!p->tpw || p->lmax >= 2, "rotational feedback needs lmax >= 2");
  auto* s = new slx_solver();
  s->h = h;
  const int L = p->lmax, nl = L + 1, K = p->kmax;
  Consts& c = s->c;
  c.rhoi = p->rhoi;
  c.rhow = p->rhow;
  c.gacc = p->gacc;
  c.K00 = SLX_SQRT2_LEGACY * std::sqrt(0.5);
  c.eps1 = p->epsilon1;
  c.area = 4 * PI_LEGACY * p->radius * p->radius;
  const double r4 = p->radius * p->radius * p->radius * p->radius;
  c.cI33 = (-8 * PI_LEGACY * r4 / (3 * std::sqrt(5.0)));
  c.cI13 = (-8 * PI_LEGACY * r4 / (std::sqrt(30.0)));
  c.cI23 = (8 * PI_LEGACY * r4 / (std::sqrt(30.0)));
  c.alpha = 1 / (1 - p->kTE2 / p->kf);
  c.inv_CA = 1 / (p->moiC - p->moiA);
  c.one_kE2 = 1 + p->kE2;
  c.inv_kf = 1 / p->kf;
  c.inv_C = 1.0 / p->moiC;
  const double r2w2 = p->radius * p->radius * p->omega * p->omega;
  c.K0 = r2w2 / 3.0;
  c.K20 = r2w2 / (6.0 * std::sqrt(5.0));
  c.K21 = r2w2 / std::sqrt(30.0);
  c.K22 = r2w2 / std::sqrt(5.0 * 24.0);
  c.ekhTE = 1 + p->kTE2 - p->hTE2;
  c.hTE2 = p->hTE2;
  c.L = L;
  c.nl = nl;
  c.ncoef = h->ncoef;
  c.nlat = h->nlat;
  c.nlon = h->nlon;
  c.npts = h->nlat * h->nlon;
  c.kmax = K;
  c.nmod2 = p->nmod2;
  c.tpw = p->tpw;
  c.checkmarine = p->checkmarine;
  c.calcRG = p->calcRG;
  c.idx20 = 2;
  c.idx21 = h->h_moff[1] + 1;
  c.idx22 = h->h_moff[2];
  s->extrap = p->extrap;
  const size_t G = (size_t)c.npts * sizeof(double), S = (size_t)h->ncoef * sizeof(double2);
  const size_t nblk_max = 3 * PSLOT;
  SLX_HIP(hipMalloc(&s->st, sizeof(DevState)));
  SLX_HIP(hipMemset(s->st, 0, sizeof(DevState)));
  SLX_HIP(hipMalloc(&s->partial, nblk_max * sizeof(double)));
  SLX_HIP(hipMalloc(&s->rows, 4 * (size_t)c.nlat * sizeof(double)));
  SLX_HIP(hipMalloc(&s->w00, (size_t)c.nlat * sizeof(double)));
  std::vector<double> w00(c.nlat);
  for (int j = 0; j < c.nlat; ++j) w00[j] = h->h_w[j] / c.nlon * std::sqrt(0.5) / SLX_SQRT2_LEGACY;
  SLX_HIP(hipMemcpy(s->w00, w00.data(), w00.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<int> nmod(nl);
  for (int l = 0; l < nl; ++l) nmod[l] = p->nmod[l];
  SLX_HIP(hipMalloc(&s->nmod, nl * sizeof(int)));
  SLX_HIP(hipMemcpy(s->nmod, nmod.data(), nl * sizeof(int), hipMemcpyHostToDevice));
  const size_t KL = (size_t)std::max(K, 1) * nl * sizeof(double);
  s->h_s.assign(p->s, p->s + (size_t)K * nl);
  for (double** q : {&s->s, &s->decay, &s->omd, &s->a_sl, &s->a_rr}) SLX_HIP(hipMalloc(q, KL));
  SLX_HIP(hipMemset(s->decay, 0, KL));
  SLX_HIP(hipMemset(s->omd, 0, KL));
  if (K > 0) {
    SLX_HIP(hipMemcpy(s->s, p->s, (size_t)K * nl * sizeof(double), hipMemcpyHostToDevice));
    SLX_HIP(hipMemcpy(s->a_sl, p->a_sl, (size_t)K * nl * sizeof(double), hipMemcpyHostToDevice));
    SLX_HIP(hipMemcpy(s->a_rr, p->a_rr, (size_t)K * nl * sizeof(double), hipMemcpyHostToDevice));
  }
  std::vector<double> a2(4 * (size_t)std::max(K, 1), 0.0);
  for (int k = 0; k < K; ++k) {
    a2[k] = p->a_il[k];
    a2[K + k] = p->a_m[k];
    a2[2 * K + k] = p->a_tsl[k];
    a2[3 * K + k] = p->a_trr[k];
  }
  SLX_HIP(hipMalloc(&s->a2, a2.size() * sizeof(double)));
  SLX_HIP(hipMemcpy(s->a2, a2.data(), a2.size() * sizeof(double), hipMemcpyHostToDevice));
  SLX_HIP(hipMalloc(&s->r2, 18 * (size_t)std::max(K, 1) * sizeof(double)));
  SLX_HIP(hipMemset(s->r2, 0, 18 * (size_t)std::max(K, 1) * sizeof(double)));
  for (auto pr : {std::make_pair(&s->ttl, p->ttl), std::make_pair(&s->ttlekhl, p->ttlekhl),
                  std::make_pair(&s->ttlhe, p->ttlhe)}) {
    SLX_HIP(hipMalloc(pr.first, nl * sizeof(double)));
    SLX_HIP(hipMemcpy(*pr.first, pr.second, nl * sizeof(double), hipMemcpyHostToDevice));
  }
  s->topo = o->tgrid;
  s->cxy = o->ocean;
  s->beta = o->beta;
  s->rr = o->R;
  s->G = o->G;
  for (double** q : {&s->tinit, &s->cxy0, &s->beta0, &s->cstar0, &s->cstar, &s->tO, &s->icestar, &s->dsl,
                     &s->topo_prev, &s->ice_prev, &s->tmp}) {
    SLX_HIP(hipMalloc(q, G));
    SLX_HIP(hipMemset(*q, 0, G));
  }
  for (double2** q : {&s->p_load, &s->p_ice, &s->p_ocean, &s->prev_ice, &s->icestar_lm, &s->oldice_lm, &s->dice,
                      &s->dI, &s->cstar_lm, &s->rt_lm, &s->load_inc, &s->dS_cur, &s->visc, &s->rad, &s->h1, &s->h2}) {
    SLX_HIP(hipMalloc(q, S));
    SLX_HIP(hipMemset(*q, 0, S));
  }
  SLX_HIP(hipMalloc(&s->scratch_lm, 3 * S));
  SLX_HIP(hipMalloc(&s->R, (size_t)std::max(K, 1) * S));
  SLX_HIP(hipMemset(s->R, 0, (size_t)std::max(K, 1) * S));
  SLX_HIP(hipHostMalloc(&s->pinned, sizeof(slx_status), hipHostMallocDefault));
  *out = s;
  return 0;
}

extern "C" int slx_solver_destroy(slx_solver* s) {
  if (!s) return 0;
  void* ptrs[] = {s->st, s->partial, s->rows, s->w00, s->nmod, s->s, s->decay, s->omd, s->a_sl, s->a_rr, s->a2,
                  s->r2, s->ttl, s->ttlekhl, s->ttlhe, s->tinit, s->cxy0, s->beta0, s->cstar0,
                  s->cstar, s->tO, s->icestar, s->dsl, s->topo_prev, s->ice_prev, s->tmp,
                  s->p_load, s->p_ice, s->p_ocean, s->prev_ice, s->icestar_lm, s->oldice_lm, s->dice, s->dI,
                  s->cstar_lm, s->rt_lm, s->load_inc, s->dS_cur, s->visc, s->rad, s->h1, s->h2, s->scratch_lm,
                  s->R};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (s->pinned) (void)hipHostFree(s->pinned);
  delete s;
  return 0;
}

extern "C" int slx_solver_set_dt(slx_solver* s, double dt) {
  SLX_CHECK(s != nullptr, "null solver");
  if (dt == s->dt) return 0;
  const int K = s->c.kmax, nl = s->c.nl;
  std::vector<int> nmod(nl);
  SLX_HIP(hipMemcpy(nmod.data(), s->nmod, nl * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<double> d((size_t)K * nl, 0.0), o((size_t)K * nl, 0.0);
  for (int l = 1; l < nl; ++l)
    for (int k = 0; k < nmod[l] && k < K; ++k) d[(size_t)k * nl + l] = std::exp(-s->h_s[(size_t)k * nl + l] * dt / 1000.0);
  for (size_t i = 0; i < d.size(); ++i) o[i] = 1.0 - d[i];
  if (K > 0) {
    SLX_HIP(hipMemcpy(s->decay, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice));
    SLX_HIP(hipMemcpy(s->omd, o.data(), o.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  s->dt = dt;
  return 0;
}

extern "C" int slx_solver_init(slx_solver* s, const double* topo0, const double* ice0, void* stream) {
  SLX_CHECK(s && topo0 && ice0, "slx_solver_init: null argument");
  hipStream_t st = (hipStream_t)stream;
  Pts p = make_pts(s, nullptr, s->oldice_lm);
  k_init_grid<<<grid_blocks(s->c.npts), 256, 0, st>>>(p, s->c, topo0, ice0);
  SLX_LAUNCH_CHECK();
  // degree-0 areas of C0, C*0 and I*0 (solver.py:571-580); the I*0 spectrum is the "previous ice"
  // of step 1 (solver.py:859-862)
  int rc = launch_fft_ana<256>(s->h, LoadPtrs<3>{{s->cxy0, s->cstar0, s->tmp}}, 3, Pred{}, st);
  if (rc) return rc;
  rc = launch_leg_ana<3>(s->h, s->h->WS, s->h->WD,
                         OutPtrs<3>{{s->scratch_lm, s->scratch_lm + s->c.ncoef, s->oldice_lm}}, Pred{}, st);
  if (rc) return rc;
  k_init_areas<<<1, 1, 0, st>>>(p, s->c, s->scratch_lm, s->scratch_lm + s->c.ncoef, s->oldice_lm);
  SLX_LAUNCH_CHECK();
  s->step = 0;
  s->nhist = 0;
  return 0;
}

extern "C" int slx_step_begin(slx_solver* s, const double* ice_cur, double dt0, double dt1, double dt2, void* stream) {
  return slx_step_begin_warm(s, ice_cur, dt0, dt1, dt2, nullptr, nullptr, stream);
}

extern "C" int slx_step_begin_warm(slx_solver* s, const double* ice_cur, double dt0, double dt1, double dt2,
                                   const void* warm_dS, const double* warm_ocean, void* stream) {
  SLX_CHECK(s && ice_cur, "slx_step_begin: null argument");
  SLX_CHECK((warm_dS == nullptr) == (warm_ocean == nullptr),
            "slx_step_begin_warm: give both warm_dS and warm_ocean, or neither");
  SLX_CHECK(s->dt > 0.0, "slx_solver_set_dt must be called before the first step");
  hipStream_t st = (hipStream_t)stream;
  SLX_CHECK(s->closed, "slx_step_begin: the previous step has not finished (check slx_solver_status)");
  s->step += 1;
  s->closed = 0;
  s->iters_enqueued = 0;
  s->ice_cur = ice_cur;
  const double2* oldice = (s->step == 1) ? s->oldice_lm : s->prev_ice;
  Pts p = make_pts(s, ice_cur, oldice);
  const Consts& c = s->c;
  k_begin<<<19, 64, 0, st>>>(p, c);
  SLX_LAUNCH_CHECK();
  int rc = launch_modal(s->h, c.kmax, s->nmod, s->decay, s->omd, s->a_sl, s->a_rr, s->R, s->p_load, s->visc,
                        s->rad, st);
  if (rc) return rc;
  if (warm_ocean)
    SLX_HIP(hipMemcpyAsync(s->cxy, warm_ocean, (size_t)c.npts * sizeof(double), hipMemcpyDeviceToDevice, st));
  k_ice_grid<<<grid_blocks(c.npts), 256, 0, st>>>(p, c);
  SLX_LAUNCH_CHECK();
  rc = launch_fft_ana<256>(s->h, LoadPtrs<2>{{s->icestar, s->cstar}}, 2, Pred{}, st);
  if (rc) return rc;
  rc = launch_leg_ana<2>(s->h, s->h->WS, s->h->WD, OutPtrs<2>{{s->icestar_lm, s->cstar_lm}}, Pred{}, st);
  if (rc) return rc;
  GuessArgs g;
  g.extrap_ready = (s->extrap && s->nhist >= 2) ? 1 : 0;
  g.uniform = (dt0 == dt1 && dt1 == dt2) ? 1 : 0;
  const double w = (dt0 + dt1) / (dt1 + dt2);
  g.c1 = dt0 / dt1 * (1.0 + w);
  g.c2 = dt0 / dt2 * w;
  g.warm = static_cast<const double2*>(warm_dS);
  k_guess<<<red_blocks(c.ncoef), 256, 0, st>>>(p, c, g);
  SLX_LAUNCH_CHECK();
  return 0;
}

extern "C" int slx_step_iterate(slx_solver* s, int n, void* stream) {
  SLX_CHECK(s != nullptr, "null solver");
  hipStream_t st = (hipStream_t)stream;
  const Consts& c = s->c;
  const double2* oldice = (s->step == 1) ? s->oldice_lm : s->prev_ice;
  Pts p = make_pts(s, s->ice_cur, oldice);
  const int* done = &s->st->st.done;
  const int* changed = &s->st->ice_changed;
  const int nsb = (c.ncoef + 255) / 256;
  for (int it = 0; it < n; ++it) {
    const bool first = (s->iters_enqueued == 0);
    const int it_step = s->iters_enqueued;
    s->iters_enqueued += 1;
    CoefB18 cb{s->dI, s->p_ocean, s->load_inc, s->visc, s->dS_cur, s->ttlekhl, s->ttl, s->st, c, 1};
    int rc = launch_leg_syn<1>(s->h, cb, Pred{done, 0}, st);
    if (rc) return rc;
    {
      const int nf = first ? 1 : 2;
      const size_t lds = (2 * (size_t)c.nlon + c.nlon / 4) * sizeof(double2);
      static size_t lds_allowed = 65536;
      if (lds > lds_allowed) {
        SLX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_synana<256>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        lds_allowed = lds;
      }
      k_fft_synana<256><<<s->h->nh, 256, lds, st>>>(dev_of(s->h), s->h->FN, s->h->FSo, s->dsl, s->cstar, s->tO,
                                                     s->h->WS, s->h->WD, nf, Pred{done, 0});
      SLX_LAUNCH_CHECK();
    }
    if (first)
      rc = launch_leg_ana<1>(s->h, s->h->WS, s->h->WD, OutPtrs<1>{{s->rt_lm}}, Pred{done, 0}, st);
    else
      rc = launch_leg_ana<2>(s->h, s->h->WS, s->h->WD, OutPtrs<2>{{s->rt_lm, s->cstar_lm}}, Pred{done, 0}, st);
    if (rc) return rc;
    k_update<<<red_blocks(c.ncoef), 256, 0, st>>>(p, c, it_step);
    SLX_LAUNCH_CHECK();
    k_refresh<<<grid_blocks(c.npts), 256, 0, st>>>(p, c, it_step, red_blocks(c.ncoef), red_blocks(c.ncoef));
    SLX_LAUNCH_CHECK();
    k_refresh2<<<grid_blocks(c.npts), 256, 0, st>>>(p, c, it_step);
    SLX_LAUNCH_CHECK();
    if (c.checkmarine) {
      rc = launch_fft_ana<256>(s->h, LoadPtrs<1>{{s->icestar}}, 1, Pred{done, 0, changed, 1}, st);
      if (!rc) rc = launch_leg_ana<1>(s->h, s->h->WS, s->h->WD, OutPtrs<1>{{s->icestar_lm}}, Pred{done, 0, changed, 1}, st);
      if (rc) return rc;
      k_dice<<<nsb, 256, 0, st>>>(p, c);
      SLX_LAUNCH_CHECK();
    }
  }
  return 0;
}

extern "C" int slx_step_finish(slx_solver* s, void* stream) {
  SLX_CHECK(s != nullptr, "null solver");
  hipStream_t st = (hipStream_t)stream;
  const Consts& c = s->c;
  const double2* oldice = (s->step == 1) ? s->oldice_lm : s->prev_ice;
  Pts p = make_pts(s, s->ice_cur, oldice);
  if (c.calcRG) {
    CoefB18 cb{s->dI, s->p_ocean, s->load_inc, s->rad, s->dS_cur, s->ttlhe, s->ttl, s->st, c, 0};
    const Pred fin{&s->st->st.done, 1, &s->st->st.finished, 0};
    int rc = launch_leg_syn<1>(s->h, cb, fin, st);
    if (rc) return rc;
    rc = launch_fft_syn<256>(s->h, StorePlain{s->rr, 0}, 1, fin, st);
    if (rc) return rc;
  }
  k_commit<<<(c.ncoef + 255) / 256, 256, 0, st>>>(p, c, s->extrap);
  SLX_LAUNCH_CHECK();
  k_post_grid<<<c.nlat, 256, 0, st>>>(p, c);
  SLX_LAUNCH_CHECK();
  k_finish<<<1, 256, 0, st>>>(p, c);
  SLX_LAUNCH_CHECK();
  return 0;
}

// Host bookkeeping once the step's finish has run on the device (status.finished): the newest
// converged increment (written to h2 by the commit) becomes h1.
static void step_close(slx_solver* s) {
  if (s->extrap) {
    std::swap(s->h1, s->h2);
    s->nhist = std::min(2, s->nhist + 1);
  }
  s->closed = 1;
}

extern "C" int slx_solver_status(slx_solver* s, slx_status* out, void* stream) {
  SLX_CHECK(s && out, "slx_solver_status: null argument");
  hipStream_t st = (hipStream_t)stream;
  SLX_HIP(hipMemcpyAsync(s->pinned, &s->st->st, sizeof(slx_status), hipMemcpyDeviceToHost, st));
  SLX_HIP(hipStreamSynchronize(st));
  std::memcpy(out, s->pinned, sizeof(slx_status));
  if (out->finished && !s->closed && s->step > 0) step_close(s);
  return 0;
}

extern "C" int slx_solver_fields(slx_solver* s, slx_fields* o) {
  SLX_CHECK(s && o, "slx_solver_fields: null argument");
  o->icestar = s->icestar;
  o->tinit = s->tinit;
  o->cstar0 = s->cstar0;
  o->cxy0 = s->cxy0;
  o->beta0 = s->beta0;
  o->dS_converged = s->dS_cur;
  o->p_load = s->p_load;
  o->p_ice = s->p_ice;
  o->p_ocean = s->p_ocean;
  o->prev_ice = s->prev_ice;
  o->load_inc = s->load_inc;
  o->modal_R = s->R;
  return 0;
}
