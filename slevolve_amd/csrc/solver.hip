// Device-resident sea-level time step (FP64; optionally FP32-product contractions in the early
// inner sweeps).  Restates python/slevolve/solver.py:521-1144 (step
// or second-order load placement; the optional rate and horizontal read-outs of solver.py:687-806)
// and the extrapolated start and Anderson mixing of solver_opt.py:67-165,504-535,705-775 with the
// grid and spectral algebra fused into the transform passes:
//   * rO - tO is analysed as one field (its analysis is all the update uses, solver.py:1002-1015);
//   * the ocean function C* is analysed in the same sweep as rO - tO of the next iteration;
//   * deltasl = dslxy + conserv K00 (one synthesis less per iteration, solver_opt.py:28-32);
//   * eq. B18 is formed by the synthesis contraction as it stages each order pair's coefficients
//     (the CoefB18 loader, which also evaluates the rotation; the mixed policy assembles them once
//     with k_coef_fill for whichever of its two contraction forms runs);
//   * with the marine check, the load update and the refresh of topography and ocean function
//     share one launch and the re-guess ring pass takes the convergence decision.
// The inner iterations are launched ahead and predicated on a device convergence flag.
#include <cmath>
#include <cstring>
#include <string>

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
static constexpr int MAX_INNER = 9999;  // the reference's abort count (solver.py:1049-1050)
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
  double rate_sl[6], rate_rr[6], hz_rot[6], rate00;  // read-outs of the finishing step
  double dice00;                                     // degree-0 ice increment of the finishing step
  // inner = "anderson", per step: mask frozen / being refreshed, refreshes done, pairs stored,
  // newest slot, accepted grounded re-guesses, C* analysed by the next iteration, this
  // iteration's re-guess accepted; the iteration index in which each branch runs (tokens)
  int aa_frozen, aa_refreshing, aa_nrefresh, aa_n, aa_head, aa_nacc, aa_cst, aa_changed;
  int aa_tok_mask, aa_tok_accept, aa_tok_push, aa_tok_mix;
  int aa_mdiff[2];   // ocean-mask cells that would change, by iteration parity
  int aa_kind;       // last mix: 0 mixed, 1 plain, 2 fallback, 3 singular
  double aa_fn[6];   // residual norm of each history slot
  double aa_gam[5];
  // mixed-precision sweeps: 1 = iteration of that parity takes FP64 contractions, 0 = FP32 products;
  // ocean-function cells the refresh of the iteration of that parity changed
  int sweep64[2];
  int mchg[2];
};

// Constants evaluated on the host once, in the reference's operation order.
struct Consts {
  double rhoi, rhow, gacc, K00, eps1, area;  // area = 4 PI a^2
  double cI33, cI13, cI23;                   // inertia factors (solver.py:915-920)
  double alpha, inv_CA, one_kE2, inv_kf, inv_C;
  double K0, K20, K21, K22;                  // Lambda factors (solver.py:955-959)
  double ekhTE, hTE2, lTE2;
  int L, nl, ncoef, npts, nlat, nlon, kmax, nmod2, tpw, checkmarine, calcRG;
  int idx20, idx21, idx22;
  int ramp;                                  // second-order load placement
  int aa, aa_depth, aa_freeze, aa_maxref, aa_keep;  // inner = "anderson"
  double dt;                                 // length of the current step (years)
  double c_il, c_m, c_tsl, c_trr, c_thl;     // sum_k (1 - phi_k) a_k at degree 2 (solver.py:362-367)
  int mixed;                                 // FP32-product sweeps while xi > eps_sw
  double eps_sw;
  int max_inner;                             // inner iterations before a step fails (<= MAX_INNER)
};

struct slx_solver {
  slx_sht* h = nullptr;
  slx_stage_marks marks;  // stage timing while profiling (slx_solver_profile)
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
  // second-order placement and read-outs: 1 - phi(s dt) [kmax][nl]; rate weights a s / 1000
  // [kmax][nl]; horizontal weights a_hl [kmax][nl], a_thl [kmax]; T_l l^E_l [nl]; per-degree
  // [7][nl]: T c_sl, T c_rr, ws_sl, ws_rr, wsphi_sl, wsphi_rr, T c_hl
  double *omphi = nullptr, *w_sl = nullptr, *w_rr = nullptr, *a_hl = nullptr, *a_thl = nullptr, *ttlle = nullptr,
         *deg = nullptr;
  int has_horiz = 0;
  // inner = "anderson": history X, G, F [depth+1][ncoef], this iteration's output G(x) [ncoef],
  // block partials of the Gram sums [32][PSLOT]
  double2 *aaX = nullptr, *aaG = nullptr, *aaF = nullptr, *gnew = nullptr;
  double* aa_part = nullptr;
  std::vector<double> h_a_sl, h_a_rr, h_a2, h_a_hl, h_a_thl, h_ttl;  // host copies for the ramp sums
  std::vector<int> h_nmod;
  // grids
  double *tinit, *cxy0, *beta0, *cstar0, *cxy, *beta, *cstar, *icestar, *dsl, *rr, *G, *topo, *topo_prev,
      *ice_prev, *tmp;
  // spectra (packed)
  double2 *R, *p_load, *p_ice, *p_ocean, *prev_ice, *icestar_lm, *oldice_lm, *cstar_lm, *rt_lm,
      *load_inc, *dS_cur, *visc, *rad, *h1, *h2, *scratch_lm;
  std::vector<double> h_s;  // host copy of s [kmax][nl]
  slx_status* pinned = nullptr;
  double dt = -1.0;
  const double* ice_cur = nullptr;  // ice grid of the step in progress (device)
  int closed = 1;                   // host bookkeeping of the last step done
  // native step loop: per-iteration decision flags [MAX_INNER] (coherent pinned host memory, host
  // and device views), whether the kernels post them, and the device view of the status slot the
  // finishing step is copied to
  int* hflag = nullptr;
  int* hflag_dev = nullptr;
  int pipelined = 0;
  // the stream of the latest enqueued work (the step tables that slx_solver_set_dt rewrites are
  // read by it until it drains)
  hipStream_t last_stream = nullptr;
  int used = 0;
  // token of the step in progress in the posted decisions: a per-handle count of started steps that
  // no re-initialisation resets, so that a decision posted before slx_solver_init never matches
  // a later step
  unsigned epoch = 0;
  // per-step status slots of the native step loop (coherent pinned host memory, grown on demand)
  slx_status* run_stat = nullptr;
  int run_stat_cap = 0;
  // incremental grounded re-guess (matrix-core path): per-ring token of the last re-analysis that
  // changed the ring [nh], the next token, and whether the incremental form is used
  int* ring_tok = nullptr;
  int next_tok = 0;
  int ice_incr = 0;
  // the step started with the incremental ice pass: the first inner iteration analyses C* too
  int cstar_pending = 0;
  slx_status* hstat_dev = nullptr;
};

// ================================================================================== kernels
namespace {

struct Pts {  // all device pointers a kernel may need, passed by value
  DevState* st;
  double* partial;
  const double *tinit, *cxy0, *beta0, *cstar0;
  double *cxy, *beta, *cstar, *icestar, *dsl, *rr, *G, *topo, *topo_prev, *ice_prev, *tmp, *rows;
  const double* w00;
  const double* ice_cur;
  double2 *p_load, *p_ice, *p_ocean, *prev_ice, *icestar_lm, *oldice, *cstar_lm, *rt_lm, *load_inc,
      *dS_cur, *visc, *rad, *h1, *h2;
  const double *ttl, *ttlekhl, *ttlhe, *decay, *a2;
  double* r2;
  const int* l_of;
  const int* m_of;
  const int* nmod;
  double2* R;
  const double *s, *omphi, *w_sl, *w_rr, *a_hl, *a_thl, *ttlle, *deg;
  double2 *aaX, *aaG, *aaF, *gnew;
  double* aa_part;
  // native step loop: decision flag of each inner iteration (token << 2 | 1 continue, 2 converged,
  // 3 failed) and the finishing step's status, both in coherent pinned host memory (null otherwise)
  int* hflag;
  int htoken;
  slx_status* hstat;
};

// Posts an inner iteration's decision to the host (system-scope store into pinned memory).
__device__ __forceinline__ void post_decision(const Pts& p, int it, bool done, bool failed) {
  if (p.hflag)
    __hip_atomic_store(p.hflag + it, (p.htoken << 2) | (failed ? 3 : done ? 2 : 1), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ double keep_factor(double ice, double topo, double rhow, double rhoi) {
  // (topo > 0) | (ice > |topo| rho_w / rho_i), as a 0/1 multiplier (solver.py:655-660)
  return (topo > 0.0 || ice > fabs(topo) * rhow / rhoi) ? 1.0 : 0.0;
}

__device__ __forceinline__ double cabs_d(double2 z) { return hypot(z.x, z.y); }

// ice increments of the step (solver.py:868-869): dicestar = icestar_lm - oldice, deltaicestar =
// p_ice + dicestar, formed where they are used rather than stored
__device__ __forceinline__ double2 ice_inc(const double2* ist, const double2* old, long i) {
  const double2 a = ist[i], b = old[i];
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 ice_tot(const double2* pice, long i, double2 d) {
  const double2 a = pice[i];
  return make_double2(a.x + d.x, a.y + d.y);
}

// NV block sums at once (blockDim 256, sh holds NV x 256 doubles) -> every thread.  Fixed pairwise
// tree: at stride s = 128, 64, ..., 1 element t < s takes element t + s; strides 128 and 64 through
// LDS, the rest inside wave 0 with lane shuffles (lane t adds lane t + s's partial: the same pairs
// in the same order, without a barrier per level).
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double* sh) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < NV; ++q) sh[q * 256 + t] = v[q];
  __syncthreads();
  if (t < 128) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q * 256 + t] += sh[q * 256 + t + 128];
  }
  __syncthreads();
  if (t < 64) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double x = sh[q * 256 + t] + sh[q * 256 + t + 64];
#pragma unroll
      for (int s = 32; s > 0; s >>= 1) x += __shfl_down(x, s, 64);
      if (t == 0) sh[q * 256] = x;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = sh[q * 256];
}

// Block sum of v (blockDim 256) -> every thread; the tree of block_sums.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  double a[1] = {v};
  block_sums<1>(a, sh);
  return a[0];
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
  if (c.ramp) {
    // the current increments carry 1 - phi of their viscous response: implicit in m
    // (solver.py:922-933)
    double silr[9];
    for (int q = 0; q < 9; ++q) silr[q] = s->sum_il[q] + c.c_il * o.dil[q];
    const double den = 1.0 - c.alpha * c.c_m * c.inv_kf;
    for (int i = 0; i < 2; ++i)
      o.mm[i] = c.alpha *
                (c.inv_CA * (c.one_kE2 * o.il[3 * i + 2] + silr[3 * i + 2]) + c.inv_kf * (s->sum_m[i] - c.c_m * s->p_m[i])) /
                den;
    o.mm[2] = -c.inv_C * (c.one_kE2 * o.il[8] + silr[8]);
  } else {
    for (int i = 0; i < 2; ++i)
      o.mm[i] = c.alpha * (c.inv_CA * (c.one_kE2 * o.il[3 * i + 2] + s->sum_il[3 * i + 2]) + c.inv_kf * s->sum_m[i]);
    o.mm[2] = -c.inv_C * (c.one_kE2 * o.il[8] + s->sum_il[8]);
  }
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
    double tsl = s->sum_tsl[q], trr = s->sum_trr[q];
    if (c.ramp) {
      tsl += c.c_tsl * o.dlam[q];
      trr += c.c_trr * o.dlam[q];
    }
    o.dsl[q] = (c.ekhTE * tot + tsl) * ig;
    o.rr[q] = c.calcRG ? (c.hTE2 * tot + trr) * ig : 0.0;
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
    s->aa_frozen = s->aa_refreshing = s->aa_nrefresh = s->aa_n = s->aa_nacc = s->aa_cst = s->aa_changed = 0;
    s->aa_head = 0;
    s->aa_tok_mask = s->aa_tok_accept = s->aa_tok_push = s->aa_tok_mix = -1;
    s->aa_mdiff[0] = s->aa_mdiff[1] = 0;
    s->sweep64[0] = c.mixed ? 0 : 1;
    s->sweep64[1] = 1;
    s->mchg[0] = s->mchg[1] = 0;
    s->st.sweeps32 = 0;
  }
}

// Grounded ice on the previous converged topography, beta, C* (solver.py:853-887), read by the
// 2-field ring pass of slx_step_begin: field 0 = I* (the ice grid, marine-checked on the previous
// step's topography), field 1 = C* = C beta with beta = (I* < eps); the field-0 blocks store I*
// and beta, the field-1 blocks C*.
struct LoadIceGrid {
  const double *ice, *topo_prev, *cxy;
  double *icestar, *beta, *cstar;
  double rhow, rhoi;
  int checkmarine;
  __device__ double operator()(int f, long i) const {
    const double v = ice[i];
    const double ist = ist_of(v, checkmarine ? topo_prev[i] : 0.0);
    return f == 0 ? ist : cxy[i] * (ist < EPS0 ? 1.0 : 0.0);
  }
  // I* of a point from its ice and previous topography (the topography is not read unchecked)
  __device__ double ist_of(double v, double t) const {
    return checkmarine ? v * keep_factor(v, t, rhow, rhoi) : v;
  }
  __device__ void commit(int f, long i, double val) const {
    if (f == 0) {
      icestar[i] = val;
      beta[i] = val < EPS0 ? 1.0 : 0.0;
    } else {
      cstar[i] = val;
    }
  }
};

// first guess of the ocean-load increment (solver.py:888-892; solver_opt.py:504-527);
// so = sum |load_inc|
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
// w1[l] ((rho_i dI + rho_w dS_prev) + rho_w load_inc) + w2[l] resp
// (+ w3[l] (rho_i dice + rho_w load_inc) with the ramp, solver.py:986-989,1065-1068)
// (+ rotation at l = 2, m <= 2).
// With eval set (sea-level synthesis of an inner iteration) every block evaluates the rotation
// block in its prologue from dS_cur and dI (identical arithmetic everywhere) and block 0
// publishes it; otherwise (radial synthesis after the loop) the stored rr terms are used.
struct CoefB18 {
  const double2 *ist, *old, *pice;  // dicestar = ist - old, deltaicestar = pice + dicestar
  const double2 *dSp, *ld, *resp, *dS;
  const double *w1, *w2;
  const double* w3;
  DevState* st;
  Consts c;
  int eval;
  int it = -1;  // inner iteration whose sweep this assembles (mixed precision: its block 0 decides)
  // only the block that stores the rotation (block 0) and those holding the degree-2, m <= 2
  // coefficients need it
  __device__ bool prologue_needed(int lo, int hi) const {
    auto in = [&](int i) { return i >= lo && i < hi; };
    return lo == 0 || in(c.idx20) || in(c.idx21) || in(c.idx22);
  }
  __device__ void prologue(double* pro) const {
    if (threadIdx.x != 0) return;
    if (c.mixed && it > 0 && blockIdx.x == 0 && blockIdx.y == 0) {
      // mixed precision: the sweep after one that changed neither the ocean function nor a
      // grounded-ice flag (the masks are final) and every later one take FP64 contractions
      const int prev = (it - 1) & 1;
      if (!st->sweep64[it & 1] && st->mchg[prev] == 0 && st->nflip[prev] == 0) st->sweep64[it & 1] = 1;
    }
    if (eval) {
      RotOut o;
      if (c.tpw) {
        rotation_eval(st, c, dS[c.idx20], dS[c.idx21], ice_tot(pice, c.idx20, ice_inc(ist, old, c.idx20)),
                      ice_tot(pice, c.idx21, ice_inc(ist, old, c.idx21)), o);
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
    const double2 d = ice_inc(ist, old, idx);
    const double2 a = ice_tot(pice, idx, d), b = dSp[idx], q = ld[idx], v = resp[idx];
    const double ar = (a.x * c.rhoi + b.x * c.rhow) + q.x * c.rhow;
    const double ai = (a.y * c.rhoi + b.y * c.rhow) + q.y * c.rhow;
    const double u = w1[l], t = w2[l];
    double2 r = make_double2(ar * u + v.x * t, ai * u + v.y * t);
    if (c.ramp) {
      const double w = w3[l];
      r.x += w * (c.rhoi * d.x + c.rhow * q.x);
      r.y += w * (c.rhoi * d.y + c.rhow * q.y);
    }
    if (l == 2 && m <= 2) {
      r.x += pro[2 * m];
      r.y += pro[2 * m + 1];
    }
    return r;
  }
};

template <int NF>
struct LoadPtrs {
  const double* g[NF];
  __device__ double operator()(int f, long i) const { return g[f][i]; }
};

// Ring-pair pass of an inner iteration: the synthesised Fourier coefficients of eq. B18 are
// inverse-transformed (one complex FFT for rows j and nlat-1-j), written as dslxy, multiplied by
// C* and reduced by tO (both formed from C, beta, tinit, C*0 as they are read) in LDS (rO - tO,
// solver.py:993-996), forward-transformed and folded into
// analysis field 0; with nf == 2 (or a set cst_flag) the ocean function C* of the same rows is
// transformed into analysis field 1 as well, after it or (up to nlon 1024) by a second block per
// ring pair (blockIdx.y = 1) that runs beside the rO - tO path.  Replaces separate synthesis and
// analysis ring passes.
struct OceanFn {  // C* = C beta and tO = tinit (C* - C*0) of a grid point (solver.py:881-884)
  const double *cxy, *beta, *tinit, *cstar0;
  __device__ __forceinline__ double cstar(long i) const { return cxy[i] * beta[i]; }
  __device__ __forceinline__ double tO(long i, double cs) const { return tinit[i] * (cs - cstar0[i]); }
};

// REG: the FFT twiddles in registers (TwReg), where one block per CU is what the shape has anyway
// (nlon <= 1024); at nlon 2048 their registers would halve the resident blocks (two blocks of 8
// waves per CU need <= 128 registers), so that shape reads them per stage.
template <int P, bool REG>
__global__ void __launch_bounds__(RING_MAXNT) k_fft_synana(ShtDev d, const double2* __restrict__ FN,
                                                           const double2* __restrict__ FS, double* __restrict__ dsl,
                                                           OceanFn oc, double2* __restrict__ WS,
                                                           double2* __restrict__ WD, int nf, const int* cst_flag,
                                                           Pred pr) {
  if (!pr.active()) return;
  if (cst_flag) nf = *(volatile const int*)cst_flag ? 2 : 1;  // C* analysed only after a mask update
  extern __shared__ double2 smem[];
  const int N = d.nlon, L = d.lmax, nt = N / P, tid = threadIdx.x;
  double2* a = smem;
  double2* b = smem + N;
  const int j = xcd_ring(blockIdx.x, d.nh), js = d.nlat - 1 - j;
  const long rn = (long)j * N, rs = (long)js * N;
  const double wn = d.wn[j], ws = d.ws[j];
  const long fstride = (long)d.nm * d.nh;
  // every global read of the pass first: the FFT twiddles, the ring pair's Fourier coefficients,
  // then C* and tO of the thread's points of both rows
  TwOf<P, REG> tw;
  tw.load(d.stw, N, d.log2n);
  if (gridDim.y == 2) {  // C* has blocks of its own (blockIdx.y = 1), concurrent with the rO - tO path
    if (blockIdx.y == 1) {
      if (nf < 2) return;
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const long i = tid + k * nt;
        a[i] = make_double2(oc.cstar(rn + i), oc.cstar(rs + i));
      }
      __syncthreads();
      double2* Z = fft_lds<P>(a, b, tw, N, d.log2n, false);
      ring_fold(Z, WS, WD, fstride, d.nh, j, L, N, wn, ws, nt);
      return;
    }
    nf = 1;
  }
  RingCoef<P> rc;
  rc.load(FN, FS, 0, d.nh, j, L, nt);
  double cn[P], cs[P], tn[P], ts[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const long i = tid + k * nt;
    cn[k] = oc.cstar(rn + i);
    cs[k] = oc.cstar(rs + i);
    tn[k] = oc.tO(rn + i, cn[k]);
    ts[k] = oc.tO(rs + i, cs[k]);
  }
  rc.scatter(a, N, L, nt);
  __syncthreads();
  double2* z = fft_lds<P>(a, b, tw, N, d.log2n, true);
  double2* w = (z == a) ? b : a;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int i = tid + k * nt;
    const double2 v = z[i];
    dsl[rn + i] = v.x;
    dsl[rs + i] = v.y;
    z[i] = make_double2(v.x * cn[k] - tn[k], v.y * cs[k] - ts[k]);
  }
  __syncthreads();
  for (int f = 0; f < nf; ++f) {
    if (f == 1) {
#pragma unroll
      for (int k = 0; k < P; ++k) z[tid + k * nt] = make_double2(cn[k], cs[k]);
      __syncthreads();
    }
    double2* Z = fft_lds<P>(z, w, tw, N, d.log2n, false);
    ring_fold(Z, WS, WD, f * fstride, d.nh, j, L, N, wn, ws, nt);
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
  const double dI00 = ice_tot(p.p_ice, 0, ice_inc(p.icestar_lm, p.oldice, 0)).x;
  return (1.0 / p.cstar_lm[0].x) * (-(c.rhoi / c.rhow) * dI00 - p.rt_lm[0].x);
}

// out = load_inc, or with Anderson mixing the separate output buffer G(x) (the input x is kept for
// the mixing history)
__global__ void __launch_bounds__(256) k_update(Pts p, Consts c, int it, double2* out) {
  __shared__ double sh[256];
  DevState* s = p.st;
  if (*(volatile int*)&s->st.done) return;
  const int n = c.ncoef;
  const double conserv = conserv_of(p, c);
  double a = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double2 sp = p.p_ocean[i], rt = p.rt_lm[i], cs = p.cstar_lm[i];
    const double2 nw = make_double2((-sp.x + rt.x) + cs.x * conserv, (-sp.y + rt.y) + cs.y * conserv);
    out[i] = nw;
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
// sh: 2 x 256 doubles; both sums are reduced together, each by the tree of block_sum
// Any block of 64 or more threads (a multiple of 64) forms the same sums in the same order: 256
// virtual lanes each add every 256th partial, then the tree of block_sums.
__device__ double convergence_xi(const Pts& p, int it, int nbu, int nbg, double* sh, double& sa, double& so) {
  const int t = threadIdx.x, nt = blockDim.x;
  // a virtual lane's partials (at most PSLOT / 256 of each sum) are loaded together, without a
  // branch (every index < PSLOT lies in its slot), then added in index order with the absent ones
  // as +0.0, which leaves the sums of non-negative partials unchanged
  constexpr int NQ = PSLOT / 256;
  const double* pa = p.partial + (it & 1) * PSLOT;
  const double* po = p.partial + (it == 0 ? 2 : ((it - 1) & 1)) * PSLOT;
  const int nbo = it == 0 ? nbg : nbu;
  for (int vt = t; vt < 256; vt += nt) {
    double av[NQ], ov[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      av[q] = pa[vt + 256 * q];
      ov[q] = po[vt + 256 * q];
    }
    double a = 0.0, o = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      a += vt + 256 * q < nbu ? av[q] : 0.0;
      o += vt + 256 * q < nbo ? ov[q] : 0.0;
    }
    sh[vt] = a;
    sh[256 + vt] = o;
  }
  __syncthreads();
  for (int vt = t; vt < 128; vt += nt) {
    sh[vt] += sh[vt + 128];
    sh[256 + vt] += sh[256 + vt + 128];
  }
  __syncthreads();
  if (t < 64) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      double x = sh[q * 256 + t] + sh[q * 256 + t + 64];
#pragma unroll
      for (int s = 32; s > 0; s >>= 1) x += __shfl_down(x, s, 64);
      if (t == 0) sh[q * 256] = x;
    }
  }
  __syncthreads();
  double v[2] = {sh[0], sh[256]};
  sa = v[0];
  so = v[1];
  if (fabs(sa - so) < EPS0 && fabs(so) < EPS0) return 0.0;
  if (fabs(so) < EPS0) return fabs((sa - so) / (EPS0 * 10));
  return fabs((sa - so) / so);
}

// The iteration's decision from its convergence ratio (one thread): posted to the host, the next
// sweep's precision, the status block; returns whether the step ends here.
__device__ bool refresh_decide(const Pts& p, const Consts& c, int it, double xi, double sa, double so, double conserv,
                               bool post) {
  DevState* s = p.st;
  // a step converges on an FP64 sweep only; FP32-product sweeps run while xi > eps_sw
  const int fp64 = *(volatile const int*)&s->sweep64[it & 1];
  const bool conv = xi <= c.eps1 && fp64;
  const bool fail = !conv && it + 1 >= c.max_inner;
  if (post) {
    post_decision(p, it, conv || fail, fail);
    s->sweep64[(it + 1) & 1] = (fp64 || xi <= c.eps_sw) ? 1 : 0;
    if (!fp64) s->st.sweeps32 += 1;
    s->st.numiter = it + 1;
    s->st.xi = xi;
    s->st.conserv = conserv;
    s->sa = sa;
    s->so = so;
    if (conv || fail) s->st.done = 1;
    if (fail) s->st.failed = 1;
  }
  return conv || fail;
}

// Topography, ocean function and grounded-flip count of the refresh over the grid points of blocks
// b0 .. b0 + nb - 1 (block b = blockIdx.x - first block of the grid part).
__device__ void refresh_grid(const Pts& p, const Consts& c, int it, double conserv, int b, int nb, int* shf,
                             int* shm) {
  DevState* s = p.st;
  if (threadIdx.x == 0) *shf = *shm = 0;
  __syncthreads();
  const long n = c.npts;
  const double cadd = conserv * c.K00;
  const bool regs = c.checkmarine && s->nacc2[it & 1] < GR_MAXACC;
  int flips = 0, mflips = 0;
  // one grid point: topography and ocean function from the sea level, flip counts
  auto point = [&](double ti, double ds, double cx, double ice, double be, double& t, double& cn) {
    t = ti - (ds + cadd);
    cn = t < 0.0 ? 1.0 : 0.0;
    if (c.mixed && cn != cx) mflips += 1;
    if (regs) {
      const double cand = ice * keep_factor(ice, t, c.rhow, c.rhoi);
      const bool bnew = cand < EPS0;
      if (bnew != (be != 0.0)) flips += 1;
    }
  };
  const uintptr_t al = reinterpret_cast<uintptr_t>(p.tinit) | reinterpret_cast<uintptr_t>(p.dsl) |
                       reinterpret_cast<uintptr_t>(p.topo) | reinterpret_cast<uintptr_t>(p.cxy) |
                       reinterpret_cast<uintptr_t>(p.ice_cur) | reinterpret_cast<uintptr_t>(p.beta);
  if ((n & 1) == 0 && (al & 15) == 0) {  // two points per thread, 16-byte accesses
    const long n2 = n >> 1;
    for (long i = b * (long)blockDim.x + threadIdx.x; i < n2; i += (long)nb * blockDim.x) {
      const double2 ti = reinterpret_cast<const double2*>(p.tinit)[i];
      const double2 ds = reinterpret_cast<const double2*>(p.dsl)[i];
      const double2 cx = c.mixed ? reinterpret_cast<const double2*>(p.cxy)[i] : make_double2(0.0, 0.0);
      const double2 ice = regs ? reinterpret_cast<const double2*>(p.ice_cur)[i] : make_double2(0.0, 0.0);
      const double2 be = regs ? reinterpret_cast<const double2*>(p.beta)[i] : make_double2(0.0, 0.0);
      double2 t, cn;
      point(ti.x, ds.x, cx.x, ice.x, be.x, t.x, cn.x);
      point(ti.y, ds.y, cx.y, ice.y, be.y, t.y, cn.y);
      reinterpret_cast<double2*>(p.topo)[i] = t;
      reinterpret_cast<double2*>(p.cxy)[i] = cn;
    }
  } else {
    for (long i = b * (long)blockDim.x + threadIdx.x; i < n; i += (long)nb * blockDim.x) {
      double t, cn;
      point(p.tinit[i], p.dsl[i], c.mixed ? p.cxy[i] : 0.0, regs ? p.ice_cur[i] : 0.0, regs ? p.beta[i] : 0.0, t, cn);
      p.topo[i] = t;
      p.cxy[i] = cn;
    }
  }
  if (flips) atomicAdd(shf, flips);
  if (mflips) atomicAdd(shm, mflips);
  __syncthreads();
  if (threadIdx.x == 0 && *shf) atomicAdd(&s->nflip[it & 1], *shf);
  if (threadIdx.x == 0 && *shm) atomicAdd(&s->mchg[it & 1], *shm);
}

__global__ void __launch_bounds__(256) k_refresh(Pts p, Consts c, int it, int nbu, int nbg) {
  __shared__ double sh[2 * 256];
  __shared__ int shf, shm;
  DevState* s = p.st;
  if (*(volatile int*)&s->st.done) return;
  double sa, so;
  const double xi = convergence_xi(p, it, nbu, nbg, sh, sa, so);
  const double conserv = conserv_of(p, c);
  if (refresh_decide(p, c, it, xi, sa, so, conserv, blockIdx.x == 0 && threadIdx.x == 0)) return;
  if (c.mixed && blockIdx.x == 0 && threadIdx.x == 0) s->mchg[(it + 1) & 1] = 0;
  refresh_grid(p, c, it, conserv, blockIdx.x, gridDim.x, &shf, &shm);
}

// The ocean-load update and the refresh in one launch (solver.py:1002-1040), for the loop whose
// grounded re-guess decides the iteration (k_reguess_ring with decide set).  Blocks 0 .. nbu - 1
// update the coefficients and post their partial sums of |load_inc| (as k_update); the other
// blocks refresh the grids with this iteration's sea level and conservation constant without
// waiting for the convergence test: when the step converges here, the post-loop pass (k_post_grid)
// rewrites topo and the ocean function from the same sea level and constant, as the reference
// recomputes both after its loop (solver.py:1053-1058), so this refresh leaves no trace; its flip
// counts only feed the re-guess, which does not run once the step is done.
__global__ void __launch_bounds__(256) k_update_refresh(Pts p, Consts c, int it, int nbu, double2* out) {
  __shared__ double sh[256];
  __shared__ int shf, shm;
  DevState* s = p.st;
  if (*(volatile int*)&s->st.done) return;
  const double conserv = conserv_of(p, c);
  if (blockIdx.x >= nbu) {
    if (c.mixed && blockIdx.x == nbu && threadIdx.x == 0) s->mchg[(it + 1) & 1] = 0;
    refresh_grid(p, c, it, conserv, blockIdx.x - nbu, gridDim.x - nbu, &shf, &shm);
    return;
  }
  const int n = c.ncoef;
  double a = 0.0;
  // two strides per pass, both coefficients' reads before either's stores (which may alias them);
  // |load_inc| is added in the same order as one stride at a time
  const int st = nbu * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += 2 * st) {
    const int i2 = i + st;
    const bool two = i2 < n;
    const double2 sp = p.p_ocean[i], rt = p.rt_lm[i], cs = p.cstar_lm[i];
    const double2 sp2 = p.p_ocean[two ? i2 : i], rt2 = p.rt_lm[two ? i2 : i], cs2 = p.cstar_lm[two ? i2 : i];
    const double2 nw = make_double2((-sp.x + rt.x) + cs.x * conserv, (-sp.y + rt.y) + cs.y * conserv);
    const double2 nw2 = make_double2((-sp2.x + rt2.x) + cs2.x * conserv, (-sp2.y + rt2.y) + cs2.y * conserv);
    out[i] = nw;
    p.dS_cur[i] = make_double2(sp.x + nw.x, sp.y + nw.y);
    a += cabs_d(nw);
    if (two) {
      out[i2] = nw2;
      p.dS_cur[i2] = make_double2(sp2.x + nw2.x, sp2.y + nw2.y);
      a += cabs_d(nw2);
    }
  }
  const double bs = block_sum(a, sh);
  if (threadIdx.x == 0) p.partial[(it & 1) * PSLOT + blockIdx.x] = bs;
}

// The grounded-ice re-guess of an inner iteration, applied by the ice ring pass
// (solver.py:661-684): gate() accepts it when a grounded flag flipped on the topography iterate
// (at most GR_MAXACC times per step) and block 0 records the bookkeeping; the loader reads
// I* = ice where (topo > 0) or ice > |topo| rho_w / rho_i, else 0, and its commit stores I* and
// beta = (I* < eps).
struct LoadReguess {
  const double *ice, *topo;
  double *icestar, *beta;
  double rhow, rhoi;
  DevState* st;
  int it, checkmarine;
  // the gate's counters (nacc2, nflip of this iteration's slot), loaded together
  __device__ int2 counters() const {
    const int slot = it & 1;
    return make_int2(st->nacc2[slot], st->nflip[slot]);
  }
  __device__ bool gate() const { return gate(counters()); }
  __device__ bool gate(int2 cnt) const {
    const int slot = it & 1;
    const int nacc = cnt.x, nflip = cnt.y;
    const bool regs = checkmarine && nacc < GR_MAXACC;
    const bool changed = regs && nflip > 0;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
      st->ice_changed = changed ? 1 : 0;
      st->nacc2[slot ^ 1] = st->nacc2[slot] + (changed ? 1 : 0);
      st->nflip[slot ^ 1] = 0;
      if (regs) st->st.reguess += 1;
      if (changed) {
        st->st.ice_analyses += 1;
        st->st.nacc += 1;
      }
    }
    return changed;
  }
  __device__ double operator()(int, long i) const { return of(ice[i], topo[i]); }
  __device__ double of(double v, double t) const { return v * keep_factor(v, t, rhow, rhoi); }
  __device__ void commit(int, long i, double cand) const {
    icestar[i] = cand;
    beta[i] = cand < EPS0 ? 1.0 : 0.0;
  }
};

// Grounded re-guess ring pass, incremental form: when gate() accepts the re-guess (as for
// LoadReguess), every block stores I* and beta of its ring pair where they change (as LoadReguess
// does everywhere) and forms the change dI* = I*new - I*old; a ring pair whose rows did not change ends there, the
// others transform the change, fold it into WS/WD field 0 (their own columns) and tag the ring
// with token for k_leg_ana_sp, which adds the analysis of the change to I*_lm.  The gate's counters
// are loaded with the convergence test's partials (the ring pair's reads wait for the gate: about
// half the iterations flip no grounded flag and read nothing more).
template <int P>
__global__ void __launch_bounds__(RING_MAXNT) k_reguess_ring(ShtDev d, LoadReguess lr, double2* __restrict__ WS,
                                                             double2* __restrict__ WD, int* __restrict__ ring_tok,
                                                             int token, Pred pr, Pts p, Consts c, int nbu,
                                                             int decide) {
  if (!pr.active()) return;
  const int2 gc = lr.counters();
  if (decide) {  // the iteration's convergence test (k_update_refresh left it to this pass)
    __shared__ double shx[2 * 256];
    double sa, so;
    const double xi = convergence_xi(p, lr.it, nbu, nbu, shx, sa, so);
    if (refresh_decide(p, c, lr.it, xi, sa, so, conserv_of(p, c), blockIdx.x == 0 && threadIdx.x == 0)) return;
  }
  if (!lr.gate(gc)) return;
  extern __shared__ double2 smem[];
  const int N = d.nlon, nt = N / P, tid = threadIdx.x;
  double2* a = smem;
  double2* b = smem + N;
  const int j = xcd_ring(blockIdx.x, d.nh), js = d.nlat - 1 - j;
  const long rn = (long)j * N, rs = (long)js * N;
  // every point's reads before any store (the stores could alias them, so interleaved they would
  // take a round trip per point)
  double on[P], os[P], vn[P], vs[P], tn[P], ts[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const long in = rn + tid + k * nt, is = rs + tid + k * nt;
    on[k] = lr.icestar[in];
    os[k] = lr.icestar[is];
    vn[k] = lr.ice[in];
    vs[k] = lr.ice[is];
    tn[k] = lr.topo[in];
    ts[k] = lr.topo[is];
  }
  double dn[P], dv[P];
  int any = 0;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const long in = rn + tid + k * nt, is = rs + tid + k * nt;
    const double cn = lr.of(vn[k], tn[k]), cs = lr.of(vs[k], ts[k]);
    // I* and beta = (I* < eps) are always stored together, so only changed points need storing
    if (cn != on[k]) lr.commit(0, in, cn);
    if (cs != os[k]) lr.commit(0, is, cs);
    dn[k] = cn - on[k];
    dv[k] = cs - os[k];
    any |= (dn[k] != 0.0 || dv[k] != 0.0) ? 1 : 0;
  }
  if (!__syncthreads_or(any)) return;
  TwOf<P> tw;  // only the blocks whose rows changed transform
  tw.load(d.stw, N, d.log2n);
#pragma unroll
  for (int k = 0; k < P; ++k) a[tid + k * nt] = make_double2(dn[k], dv[k]);
  __syncthreads();
  double2* Z = fft_lds<P>(a, b, tw, N, d.log2n, false);
  ring_fold(Z, WS, WD, 0, d.nh, j, d.lmax, N, d.wn[j], d.ws[j], nt);
  if (tid == 0) ring_tok[j] = token;
}

// Step-start ice pass, incremental form (used with the extrapolated start, whose first guess does
// not read C*_lm; the first inner iteration then analyses C* with rO - tO): every block forms I* of
// its ring pair from the step's ice on the previous converged topography (solver.py:853-857) as
// LoadIceGrid does, stores I* and beta where they change and C* = C beta everywhere, and forms
// dI* = I*new - I*old; a ring pair whose rows did not change ends there, the others transform the
// change, fold it into WS/WD field 0 and tag the ring with token for k_leg_ana_sp, which adds the
// analysis of the change to I*_lm.
template <int P>
__global__ void __launch_bounds__(RING_MAXNT) k_begin_ring(ShtDev d, LoadIceGrid lg, double2* __restrict__ WS,
                                                           double2* __restrict__ WD, int* __restrict__ ring_tok,
                                                           int token) {
  extern __shared__ double2 smem[];
  const int N = d.nlon, nt = N / P, tid = threadIdx.x;
  double2* a = smem;
  double2* b = smem + N;
  const int j = xcd_ring(blockIdx.x, d.nh), js = d.nlat - 1 - j;
  const long rn = (long)j * N, rs = (long)js * N;
  // every point's reads before any store (the stores could alias them, so interleaved they would
  // take a round trip per point)
  double on[P], os[P], vn[P], vs[P], tn[P], ts[P], xn[P], xs[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const long in = rn + tid + k * nt, is = rs + tid + k * nt;
    on[k] = lg.icestar[in];
    os[k] = lg.icestar[is];
    vn[k] = lg.ice[in];
    vs[k] = lg.ice[is];
    tn[k] = lg.checkmarine ? lg.topo_prev[in] : 0.0;
    ts[k] = lg.checkmarine ? lg.topo_prev[is] : 0.0;
    xn[k] = lg.cxy[in];
    xs[k] = lg.cxy[is];
  }
  double dn[P], dv[P];
  int any = 0;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const long in = rn + tid + k * nt, is = rs + tid + k * nt;
    const double cn = lg.ist_of(vn[k], tn[k]), cs = lg.ist_of(vs[k], ts[k]);
    if (cn != on[k]) lg.commit(0, in, cn);
    if (cs != os[k]) lg.commit(0, is, cs);
    lg.cstar[in] = xn[k] * (cn < EPS0 ? 1.0 : 0.0);
    lg.cstar[is] = xs[k] * (cs < EPS0 ? 1.0 : 0.0);
    dn[k] = cn - on[k];
    dv[k] = cs - os[k];
    any |= (dn[k] != 0.0 || dv[k] != 0.0) ? 1 : 0;
  }
  if (!__syncthreads_or(any)) return;
  TwOf<P> tw;  // only the blocks whose rows changed transform
  tw.load(d.stw, N, d.log2n);
#pragma unroll
  for (int k = 0; k < P; ++k) a[tid + k * nt] = make_double2(dn[k], dv[k]);
  __syncthreads();
  double2* Z = fft_lds<P>(a, b, tw, N, d.log2n, false);
  ring_fold(Z, WS, WD, 0, d.nh, j, d.lmax, N, d.wn[j], d.ws[j], nt);
  if (tid == 0) ring_tok[j] = token;
}

// ---------------------------------------------------------------------- Anderson mixing
// Inner iteration with inner = "anderson" (solver_opt.py:705-775): xi as above in every block;
// when the iteration has converged, or the mask is not frozen, count the ocean-mask cells the
// topography iterate would change and the grounded flips on it (no grid is written here).
__global__ void __launch_bounds__(256) k_refresh_aa(Pts p, Consts c, int it, int nbu, int nbg) {
  __shared__ double sh[2 * 256];
  __shared__ int shm, shf;
  DevState* s = p.st;
  if (*(volatile int*)&s->st.done) return;
  double sa, so;
  const double xi = convergence_xi(p, it, nbu, nbg, sh, sa, so);
  const double conserv = conserv_of(p, c);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s->st.numiter = it + 1;
    s->st.xi = xi;
    s->st.conserv = conserv;
    s->sa = sa;
    s->so = so;
  }
  const bool conv = xi <= c.eps1;
  if (!conv && s->aa_frozen) return;
  if (threadIdx.x == 0) shm = shf = 0;
  __syncthreads();
  const double cadd = conserv * c.K00;
  const bool regs = c.checkmarine && s->aa_nacc < GR_MAXACC;
  int md = 0, flips = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < c.npts; i += (long)gridDim.x * blockDim.x) {
    const double t = p.tinit[i] - (p.dsl[i] + cadd);
    if ((t < 0.0) != (p.cxy[i] != 0.0)) md += 1;
    if (regs) {
      const double ice = p.ice_cur[i];
      if ((ice * keep_factor(ice, t, c.rhow, c.rhoi) < EPS0) != (p.beta[i] != 0.0)) flips += 1;
    }
  }
  if (md) atomicAdd(&shm, md);
  if (flips) atomicAdd(&shf, flips);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (shm) atomicAdd(&s->aa_mdiff[it & 1], shm);
    if (shf) atomicAdd(&s->nflip[it & 1], shf);
  }
}

// The branch of the iteration (one lane): converged with an unchanged mask (or refreshes spent)
// -> done; converged otherwise -> refresh the mask; frozen -> mix; else a plain step with the mask
// update, freezing the mask from iteration anderson_freeze on (the history restarts with this
// step's pair).  Sets the tokens of the kernels that follow and the grounded re-guess bookkeeping.
__global__ void k_decide_aa(Pts p, Consts c, int it) {
  DevState* s = p.st;
  if (threadIdx.x != 0 || blockIdx.x != 0 || *(volatile int*)&s->st.done) return;
  const int slot = it & 1;
  const bool conv = s->st.xi <= c.eps1;
  bool mask = false, mix = false, push = false, accept = false;
  if (conv) {
    if (s->aa_mdiff[slot] == 0 || s->aa_nrefresh >= c.aa_maxref) {
      s->st.done = 1;
      accept = true;
    } else {
      s->aa_nrefresh += 1;
      s->st.aa_mask_refresh += 1;
      s->aa_refreshing = 1;
    }
  } else if (s->aa_frozen) {
    mix = push = true;
    s->st.aa_frozen_iters += 1;
  }
  if (!s->st.done && !mix) {
    if (s->aa_refreshing) {
      if (!c.aa_keep) s->aa_n = 0;
      s->aa_refreshing = 0;
      s->aa_frozen = 1;
    } else if (it + 1 >= c.aa_freeze) {
      s->aa_n = 0;
      push = true;
      s->aa_frozen = 1;
    }
    mask = accept = true;
  }
  if (!s->st.done && it + 1 >= c.max_inner) {
    s->st.failed = 1;
    s->st.done = 1;
    mask = mix = push = accept = false;
  }
  post_decision(p, it, s->st.done != 0, s->st.failed != 0);
  if (push) {
    s->aa_head = (s->aa_head + 1) % (c.aa_depth + 1);
    s->aa_n = min(s->aa_n + 1, c.aa_depth + 1);
  }
  bool changed = false;
  if (mask) {
    const bool regs = c.checkmarine && s->aa_nacc < GR_MAXACC;
    changed = regs && s->nflip[slot] > 0;
    if (regs) s->st.reguess += 1;
    if (changed) {
      s->aa_nacc += 1;
      s->st.ice_analyses += 1;
      s->st.nacc += 1;
    }
  }
  s->aa_changed = changed ? 1 : 0;
  s->ice_changed = changed ? 1 : 0;
  s->aa_cst = mask ? 1 : 0;
  s->aa_tok_mask = mask ? it : -1;
  s->aa_tok_accept = accept ? it : -1;
  s->aa_tok_push = push ? it : -1;
  s->aa_tok_mix = mix ? it : -1;
  s->aa_mdiff[slot ^ 1] = 0;
  s->nflip[slot ^ 1] = 0;
}

// mask update of a plain / refresh step: topography, ocean function, grounded re-guess (C* and tO
// are formed from them by the next ring pass)
__global__ void k_refresh2_aa(Pts p, Consts c, int it) {
  DevState* s = p.st;
  if (*(volatile int*)&s->aa_tok_mask != it) return;
  const bool changed = s->aa_changed;
  const double cadd = s->st.conserv * c.K00;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < c.npts; i += (long)gridDim.x * blockDim.x) {
    const double t = p.tinit[i] - (p.dsl[i] + cadd);
    const double cx = t < 0.0 ? 1.0 : 0.0;
    p.topo[i] = t;
    p.cxy[i] = cx;
    if (changed) {
      const double ice = p.ice_cur[i];
      const double cand = ice * keep_factor(ice, t, c.rhow, c.rhoi);
      p.icestar[i] = cand;
      p.beta[i] = cand < EPS0 ? 1.0 : 0.0;
    }
  }
}

// Store the pair (x = load_inc, g = G(x)) in the newest history slot with f = g - x, and form the
// block partials of f.f, dF_j . f and dF_j . dF_k (dF_j = f_j - f_{j+1}, newest first; real inner
// product over real and imaginary parts).  Partial q of block b at aa_part[q * PSLOT + b]:
// q = 0 f.f, 1..m dF.f, then the upper triangle of dF.dF row by row.
__global__ void __launch_bounds__(256) k_aa_push(Pts p, Consts c, int it) {
  __shared__ double sh[256];
  DevState* s = p.st;
  if (*(volatile int*)&s->aa_tok_push != it) return;
  const int D1 = c.aa_depth + 1, h = s->aa_head, n = s->aa_n, m = n - 1;
  const long nc = c.ncoef;
  int sl[6];
  for (int j = 0; j < 6; ++j) sl[j] = ((h - j) % D1 + D1) % D1;
  double acc[21];
  for (int q = 0; q < 21; ++q) acc[q] = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nc; i += (long)gridDim.x * blockDim.x) {
    const double2 x = p.load_inc[i], g = p.gnew[i];
    const double2 f = make_double2(g.x - x.x, g.y - x.y);
    p.aaX[h * nc + i] = x;
    p.aaG[h * nc + i] = g;
    p.aaF[h * nc + i] = f;
    acc[0] += f.x * f.x + f.y * f.y;
    double2 df[5];
    double2 fa = f;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      if (j < m) {
        const double2 fb = p.aaF[sl[j + 1] * nc + i];
        df[j] = make_double2(fa.x - fb.x, fa.y - fb.y);
        fa = fb;
        acc[1 + j] += df[j].x * f.x + df[j].y * f.y;
      }
    }
    int q = 6;
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
      for (int k = j; k < 5; ++k, ++q)
        if (k < m) acc[q] += df[j].x * df[k].x + df[j].y * df[k].y;
  }
  for (int q = 0; q < 21; ++q) {
    const double v = block_sum(acc[q], sh);
    if (threadIdx.x == 0) p.aa_part[q * PSLOT + blockIdx.x] = v;
    __syncthreads();
  }
}

// Least squares of the mixing step (one block): the partials reduced in a fixed order; the
// residual norm of the new pair; with a mix, the fallback / plain rules and
// gamma = pinv(A) rhs of the normal equations A = dF.dF, rhs = dF.f (cyclic Jacobi
// eigendecomposition of the symmetric A, eigenvalues of magnitude <= 1e-12 of the largest
// dropped, as a least-squares solve with rcond 1e-12; a non-finite gamma is "singular").
__global__ void __launch_bounds__(256) k_aa_solve(Pts p, Consts c, int it, int nb) {
  __shared__ double sh[256];
  __shared__ double tot[21];
  DevState* s = p.st;
  if (*(volatile int*)&s->aa_tok_push != it) return;
  for (int q = 0; q < 21; ++q) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += p.aa_part[q * PSLOT + b];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) tot[q] = v;
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const int D1 = c.aa_depth + 1, h = s->aa_head, n = s->aa_n, m = n - 1;
  const double fn = sqrt(tot[0]);
  const double fprev = s->aa_fn[(h - 1 + D1) % D1];
  s->aa_fn[h] = fn;
  if (s->aa_tok_mix != it) return;
  int kind;
  if (n >= 2 && fn > fprev) {
    kind = 2;
    s->aa_n = 1;  // the history restarts from the newest pair
  } else if (n < 2) {
    kind = 1;
  } else {
    double a[5][5], v[5][5], rhs[5];
    int q = 6;
    for (int j = 0; j < 5; ++j)
      for (int k = j; k < 5; ++k, ++q)
        if (k < m) a[j][k] = a[k][j] = tot[q];
    for (int j = 0; j < m; ++j) {
      rhs[j] = tot[1 + j];
      for (int k = 0; k < m; ++k) v[j][k] = j == k ? 1.0 : 0.0;
    }
    for (int sweep = 0; sweep < 64; ++sweep) {
      double off = 0.0;
      for (int j = 0; j < m; ++j)
        for (int k = j + 1; k < m; ++k) off += a[j][k] * a[j][k];
      if (!(off > 0.0)) break;
      for (int j = 0; j < m; ++j)
        for (int k = j + 1; k < m; ++k) {
          if (a[j][k] == 0.0) continue;
          const double th = (a[k][k] - a[j][j]) / (2.0 * a[j][k]);
          const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
          const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
          for (int r = 0; r < m; ++r) {  // columns j, k
            const double arj = a[r][j], ark = a[r][k];
            a[r][j] = cs * arj - sn * ark;
            a[r][k] = sn * arj + cs * ark;
          }
          for (int r = 0; r < m; ++r) {  // rows j, k
            const double ajr = a[j][r], akr = a[k][r];
            a[j][r] = cs * ajr - sn * akr;
            a[k][r] = sn * ajr + cs * akr;
          }
          for (int r = 0; r < m; ++r) {
            const double vrj = v[r][j], vrk = v[r][k];
            v[r][j] = cs * vrj - sn * vrk;
            v[r][k] = sn * vrj + cs * vrk;
          }
        }
    }
    double lmax = 0.0;
    for (int j = 0; j < m; ++j) lmax = fmax(lmax, fabs(a[j][j]));
    double gam[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = 0; j < m; ++j) {
      const double lam = a[j][j];
      if (!(fabs(lam) > 1e-12 * lmax)) continue;
      double proj = 0.0;
      for (int r = 0; r < m; ++r) proj += v[r][j] * rhs[r];
      for (int r = 0; r < m; ++r) gam[r] += v[r][j] * (proj / lam);
    }
    bool finite = true;
    for (int j = 0; j < m; ++j) finite = finite && isfinite(gam[j]);
    for (int j = 0; j < m; ++j) finite = finite && isfinite(a[j][j]);
    kind = finite ? 0 : 3;
    for (int j = 0; j < 5; ++j) s->aa_gam[j] = finite && j < m ? gam[j] : 0.0;
  }
  s->aa_kind = kind;
  if (kind == 0) s->st.aa_mixed += 1;
  if (kind == 1) s->st.aa_plain += 1;
  if (kind == 2) s->st.aa_fallback += 1;
  if (kind == 3) s->st.aa_singular += 1;
}

// next input x = g_0 - sum_j gamma_j (g_j - g_{j+1}) (g_0 alone unless "mixed"); deltaS_cur; the
// block partials of |x| for the next convergence ratio
__global__ void __launch_bounds__(256) k_aa_mix(Pts p, Consts c, int it) {
#pragma clang fp contract(off)
  __shared__ double sh[256];
  DevState* s = p.st;
  if (*(volatile int*)&s->aa_tok_mix != it) return;
  const int D1 = c.aa_depth + 1, h = s->aa_head, m = s->aa_n - 1, kind = s->aa_kind;
  const long nc = c.ncoef;
  int sl[6];
  for (int j = 0; j < 6; ++j) sl[j] = ((h - j) % D1 + D1) % D1;
  double gam[5];
  for (int j = 0; j < 5; ++j) gam[j] = s->aa_gam[j];
  double a = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nc; i += (long)gridDim.x * blockDim.x) {
    double2 g = p.aaG[h * nc + i];
    double2 out = g;
    if (kind == 0) {
      for (int j = 0; j < m; ++j) {
        const double2 gb = p.aaG[sl[j + 1] * nc + i];
        const double2 dg = make_double2(g.x - gb.x, g.y - gb.y);
        out.x -= gam[j] * dg.x;
        out.y -= gam[j] * dg.y;
        g = gb;
      }
    }
    p.load_inc[i] = out;
    const double2 sp = p.p_ocean[i];
    p.dS_cur[i] = make_double2(sp.x + out.x, sp.y + out.y);
    a += cabs_d(out);
  }
  const double bs = block_sum(a, sh);
  if (threadIdx.x == 0) p.partial[(it & 1) * PSLOT + blockIdx.x] = bs;
}

// a plain or final step takes the output as the next input (deltaS_cur and |x| are already those)
__global__ void k_aa_plain(Pts p, Consts c, int it) {
  if (*(volatile int*)&p.st->aa_tok_accept != it) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < c.ncoef; i += (long)gridDim.x * blockDim.x)
    p.load_inc[i] = p.gnew[i];
}

// ---------------------------------------------------------------------- step finish
// commit (solver.py:1080-1095): load increment without degree 0 into P; ice, ocean, previous ice
// (a step whose inner loop failed is not committed: the reference raises before its commit,
// solver.py:1049-1050)
__device__ __forceinline__ bool finish_active(const DevState* s) {
  return *(volatile const int*)&s->st.done && !*(volatile const int*)&s->st.finished &&
         !*(volatile const int*)&s->st.failed;
}

__global__ void k_commit(Pts p, Consts c, int extrap) {
  if (!finish_active(p.st)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.ncoef) return;
  const int l = p.l_of[i];
  const double2 d = ice_inc(p.icestar_lm, p.oldice, i), q = p.load_inc[i];
  const double2 dI = ice_tot(p.p_ice, i, d);
  double2 cur = make_double2(d.x * c.rhoi + q.x * c.rhow, d.y * c.rhoi + q.y * c.rhow);
  if (l == 0) cur = make_double2(0.0, 0.0);
  if (c.ramp) {
    // the increment already carries 1 - phi_k of its response in every mode (solver.py:486-498)
#pragma clang fp contract(off)
    const int nk = p.nmod[l] < c.kmax ? p.nmod[l] : c.kmax;
    for (int k = 0; k < nk; ++k) {
      const double o = p.omphi[k * c.nl + l];
      const double2 r = p.R[(long)k * c.ncoef + i];
      p.R[(long)k * c.ncoef + i] = make_double2(r.x + o * cur.x, r.y + o * cur.y);
    }
  }
  const double2 pl = p.p_load[i];
  p.p_load[i] = make_double2(pl.x + cur.x, pl.y + cur.y);
  p.p_ice[i] = dI;
  p.p_ocean[i] = p.dS_cur[i];
  p.prev_ice[i] = p.icestar_lm[i];
  if (extrap) p.h2[i] = q;  // becomes the newest history entry (the host swaps h1/h2)
}

// final topography, ocean function, G and the row sums of the degree-0 diagnostics (one block per
// row); with g0 (the dSL/dt grid without its degree 0) also the row sums of g0 C* on the final C*
__global__ void __launch_bounds__(256) k_post_grid(Pts p, Consts c, const double* g0) {
  if (!finish_active(p.st)) return;
  __shared__ double sh[5 * 256];
  const int j = blockIdx.x;
  const double cadd = p.st->st.conserv * c.K00;
  double a = 0.0, b = 0.0, e = 0.0, f = 0.0, r5 = 0.0;
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
    if (g0) r5 += g0[k] * (o * bt);
  }
  double r[5] = {a, b, e, f, r5};
  block_sums<5>(r, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) p.rows[q * c.nlat + j] = r[q];
    if (g0) p.rows[4 * c.nlat + j] = r[4];
  }
}

// degree-0 diagnostics and the rotational commit (solver.py:1060-1110; solver_opt.py:525-547);
// with the ramp the degree-2 states take their share of the increments (solver.py:490-498).
// With rate read-outs (rates != 0) the degree-0 coefficient of dSL/dt from the conservation
// constant on the end-of-step C* (solver.py:792-806) goes to rate00 and dsl_lm[0]; close = 0
// leaves status.finished to k_close.
__global__ void __launch_bounds__(256) k_finish(Pts p, Consts c, int rates, double2* dsl_lm, int close) {
  __shared__ double sh[256];
  DevState* s = p.st;
  if (!finish_active(s)) return;
  __shared__ double sh5[5 * 256];
  double q[5];
  const int nq = rates ? 5 : 4;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    double v = 0.0;
    if (k < nq)
      for (int j = threadIdx.x; j < c.nlat; j += blockDim.x) v += p.w00[j] * p.rows[k * c.nlat + j];
    q[k] = v;
  }
  block_sums<5>(q, sh5);
  __syncthreads();
  if (c.ramp && c.tpw) {
    for (int idx = threadIdx.x; idx < 18 * c.nmod2; idx += blockDim.x) {
      const int k = idx / 18, e = idx % 18;
      const double d = e < 9 ? s->dil[e] : (e < 12 ? s->dm[e - 9] : s->dlam[e - 12]);
      p.r2[idx] += p.omphi[k * c.nl + 2] * d;
    }
  }
  if (threadIdx.x == 0) {
    if (rates) {
      const double d00 = ((-(c.rhoi / c.rhow) * s->dice00) / c.dt - q[4]) / q[3];
      s->rate00 = d00;
      dsl_lm[0] = make_double2(d00, 0.0);
    }
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
    if (close) s->st.finished = 1;
  }
  if (close && p.hstat) {  // the final status into the host slot, one word per lane
    __syncthreads();
    constexpr int nw = (int)(sizeof(slx_status) / sizeof(int));
    static_assert(nw <= 256, "status larger than a block");
    if (threadIdx.x < nw)
      __hip_atomic_store(reinterpret_cast<int*>(p.hstat) + threadIdx.x,
                         reinterpret_cast<const int*>(&s->st)[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---------------------------------------------------------------------- read-outs
// Rotational read-outs of the finishing step, before its commit (one lane):
// * rates (solver.py:722-767): il-dot, m-dot from the differentiated Liouville equations with the
//   ramp-consistent states r_k + (1 - phi_k) d(.), Lambda-dot by the chain rule, and the degree-2
//   sea-level / radial rate terms (rate_sl, rate_rr);
// * horizontal (solver.py:708-716): the degree-2 rotational term with the tidal l numbers (hz_rot).
__global__ void k_rot_readout(Pts p, Consts c, int rates, int hor, double* rot_rates) {
  DevState* s = p.st;
  if (!finish_active(s) || threadIdx.x != 0 || blockIdx.x != 0) return;
  const int n2 = c.nmod2, nl = c.nl, K = c.kmax;
  const bool rot = c.tpw && c.L >= 2;
  if (rates) {
    s->dice00 = ice_inc(p.icestar_lm, p.oldice, 0).x;  // before the commit replaces the previous ice
    double ild[9] = {}, md[3] = {}, lad[18] = {};
    for (int q = 0; q < 6; ++q) s->rate_sl[q] = s->rate_rr[q] = 0.0;
    if (rot) {
      const double h = c.dt;
      const double* il = s->st.il;
      const double* mm = s->st.mm;
      const double* tot = s->st.lam + 12;
      double dil[9], dm[3], dla[6], sild[9] = {}, smd[3] = {}, tsld[6] = {}, trrd[6] = {};
      for (int q = 0; q < 9; ++q) dil[q] = il[q] - s->p_il[q];
      for (int q = 0; q < 3; ++q) dm[q] = mm[q] - s->p_m[q];
      for (int q = 0; q < 6; ++q) dla[q] = tot[q] - s->p_lam[q];
      for (int q = 0; q < 9; ++q) ild[q] = dil[q] / h;
      for (int k = 0; k < n2; ++k) {
        const double o = p.omphi[k * nl + 2], sk = p.s[k * nl + 2] / 1000.0;
        const double* r = p.r2 + 18 * k;
        const double wil = p.a2[k] * sk, wm = p.a2[K + k] * sk, wt = p.a2[2 * K + k] * sk,
                     wr = p.a2[3 * K + k] * sk;
        for (int q = 0; q < 9; ++q) sild[q] += wil * (il[q] - (r[q] + o * dil[q]));
        for (int q = 0; q < 3; ++q) smd[q] += wm * (mm[q] - (r[9 + q] + o * dm[q]));
        for (int q = 0; q < 6; ++q) {
          const double rl = r[12 + q] + o * dla[q];
          tsld[q] += wt * (tot[q] - rl);
          trrd[q] += wr * (tot[q] - rl);
        }
      }
      for (int i = 0; i < 2; ++i)
        md[i] = c.alpha * (c.inv_CA * (c.one_kE2 * ild[3 * i + 2] + sild[3 * i + 2]) + smd[i] * c.inv_kf);
      md[2] = -c.inv_C * (c.one_kE2 * ild[8] + sild[8]);
      const double m1 = mm[0], m2 = mm[1], m3 = mm[2], d1 = md[0], d2 = md[1], d3 = md[2];
      lad[0] = c.K0 * (2 * m1 * d1 + 2 * m2 * d2 + 2 * m3 * d3 + 2.0 * d3);
      lad[12] = c.K20 * (2 * m1 * d1 + 2 * m2 * d2 - 4 * m3 * d3 - 4 * d3);
      // order 1 in the U basis: -(r^2 w^2 / sqrt 30) ((d1 - i d2)(1 + m3) + (m1 - i m2) d3)
      lad[14] = -c.K21 * (d1 * (1.0 + m3) + m1 * d3);
      lad[15] = -c.K21 * (-d2 * (1.0 + m3) - m2 * d3);
      lad[16] = c.K22 * (2 * m2 * d2 - 2 * m1 * d1);
      lad[17] = c.K22 * (2.0 * (d1 * m2 + m1 * d2));
      const double ig = 1.0 / c.gacc;
      for (int q = 0; q < 6; ++q) {
        s->rate_sl[q] = (c.ekhTE * lad[12 + q] + tsld[q]) * ig;
        s->rate_rr[q] = (c.hTE2 * lad[12 + q] + trrd[q]) * ig;
      }
    }
    for (int q = 0; q < 9; ++q) rot_rates[q] = ild[q];
    for (int q = 0; q < 3; ++q) rot_rates[9 + q] = md[q];
    for (int q = 0; q < 18; ++q) rot_rates[12 + q] = lad[q];
  }
  if (hor) {
    for (int q = 0; q < 6; ++q) s->hz_rot[q] = 0.0;
    if (rot) {
      for (int q = 0; q < 6; ++q) {
        double thl = 0.0;
        for (int k = 0; k < n2; ++k) thl += p.a_thl[k] * p.r2[18 * k + 12 + q];
        const double tot = s->p_lam[q] + s->dlam[q];
        if (c.ramp) thl = thl + c.c_thl * (tot - s->p_lam[q]);
        s->hz_rot[q] = (c.lTE2 * tot + thl) / c.gacc;
      }
    }
  }
}

// Spectral read-outs of the finishing step, before its commit, one packed (l, m) per lane:
// * rates (solver.py:454-474,769-790): cur = rho_i dice + rho_w load_inc (row l = 0 zero),
//   rs = ws P - sum_k w_sl R_k + wsphi cur, dSL/dt = T E cur / dt + T rs (row l = 0 zero) plus
//   the rotational term; dR/dt likewise with h and w_rr;
// * horizontal (solver.py:687-720): Psi = T l^E A + T sum_k a_hl R_k
//   (+ T c_hl (rho_i dice + rho_w load_inc) with the ramp), row l = 0 zero, plus hz_rot.
// The modal sums run over k in order, each product rounded before its sum.
struct ReadoutSpec {
  double2 *dsl, *drr, *dsl_user, *drr_user, *hz;
};
__global__ void __launch_bounds__(256) k_readout_spec(Pts p, Consts c, ReadoutSpec a) {
#pragma clang fp contract(off)
  if (!finish_active(p.st)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.ncoef) return;
  const DevState* s = p.st;
  const int l = p.l_of[i], m = p.m_of[i], nl = c.nl;
  const int nk = p.nmod[l] < c.kmax ? p.nmod[l] : c.kmax;
  const double2 d = ice_inc(p.icestar_lm, p.oldice, i), q = p.load_inc[i];
  const double2 cur = l == 0 ? make_double2(0.0, 0.0)
                             : make_double2(d.x * c.rhoi + q.x * c.rhow, d.y * c.rhoi + q.y * c.rhow);
  const bool rates = a.dsl != nullptr, hor = a.hz != nullptr;
  double sr = 0.0, si = 0.0, tr = 0.0, ti = 0.0, hr = 0.0, hi = 0.0;
  for (int k = 0; k < nk; ++k) {
    const double2 r = p.R[(long)k * c.ncoef + i];
    if (rates) {
      const double u = p.w_sl[k * nl + l], v = p.w_rr[k * nl + l];
      sr += u * r.x;
      si += u * r.y;
      tr += v * r.x;
      ti += v * r.y;
    }
    if (hor) {
      const double w = p.a_hl[k * nl + l];
      hr += w * r.x;
      hi += w * r.y;
    }
  }
  const bool rot = c.tpw && l == 2 && m <= 2;
  if (rates) {
    const double2 P = p.p_load[i];
    const double* g = p.deg;
    const double ws = g[2 * nl + l], wr = g[3 * nl + l], wps = g[4 * nl + l], wpr = g[5 * nl + l];
    const double rsx = (ws * P.x - sr) + wps * cur.x, rsy = (ws * P.y - si) + wps * cur.y;
    const double rrx = (wr * P.x - tr) + wpr * cur.x, rry = (wr * P.y - ti) + wpr * cur.y;
    const double t = p.ttl[l], e = p.ttlekhl[l], hh = p.ttlhe[l];
    double2 ds = make_double2(e * cur.x / c.dt + t * rsx, e * cur.y / c.dt + t * rsy);
    double2 dr = make_double2(hh * cur.x / c.dt + t * rrx, hh * cur.y / c.dt + t * rry);
    if (l == 0) ds = dr = make_double2(0.0, 0.0);
    if (rot) {
      ds.x += s->rate_sl[2 * m];
      ds.y += s->rate_sl[2 * m + 1];
      dr.x += s->rate_rr[2 * m];
      dr.y += s->rate_rr[2 * m + 1];
    }
    a.dsl[i] = ds;
    a.drr[i] = dr;
    a.dsl_user[i] = ds;
    a.drr_user[i] = dr;
  }
  if (hor) {
    const double2 I = ice_tot(p.p_ice, i, d), Sp = p.p_ocean[i];
    const double Ax = (c.rhoi * I.x + c.rhow * Sp.x) + c.rhow * q.x;
    const double Ay = (c.rhoi * I.y + c.rhow * Sp.y) + c.rhow * q.y;
    const double t = p.ttl[l], e = p.ttlle[l];
    double2 h = make_double2(e * Ax + t * hr, e * Ay + t * hi);
    if (c.ramp) {
      const double w = p.deg[6 * nl + l];
      h.x += w * (c.rhoi * d.x + c.rhow * q.x);
      h.y += w * (c.rhoi * d.y + c.rhow * q.y);
    }
    if (l == 0) h = make_double2(0.0, 0.0);
    if (rot) {
      h.x += s->hz_rot[2 * m];
      h.y += s->hz_rot[2 * m + 1];
    }
    a.hz[i] = h;
  }
}

// dSL/dt grid: add the degree-0 term rate00 (synthesised as the constant rate00 K00)
__global__ void k_rate_grid(Pts p, Consts c, double* g) {
  if (!finish_active(p.st)) return;
  const double add = p.st->rate00 * c.K00;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < c.npts; i += (long)gridDim.x * blockDim.x)
    g[i] += add;
}

__global__ void k_close(Pts p) {
  if (!finish_active(p.st) || threadIdx.x != 0 || blockIdx.x != 0) return;
  p.st->st.finished = 1;
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
  p.hflag = s->pipelined ? s->hflag_dev : nullptr;
  p.htoken = (int)(s->epoch & 0x1fffffffu);
  p.hstat = s->hstat_dev;
  p.st = s->st;
  p.partial = s->partial;
  p.tinit = s->tinit;
  p.cxy0 = s->cxy0;
  p.beta0 = s->beta0;
  p.cstar0 = s->cstar0;
  p.cxy = s->cxy;
  p.beta = s->beta;
  p.cstar = s->cstar;
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
  p.m_of = s->h->m_of;
  p.nmod = s->nmod;
  p.R = s->R;
  p.s = s->s;
  p.omphi = s->omphi;
  p.w_sl = s->w_sl;
  p.w_rr = s->w_rr;
  p.a_hl = s->a_hl;
  p.a_thl = s->a_thl;
  p.ttlle = s->ttlle;
  p.deg = s->deg;
  p.aaX = s->aaX;
  p.aaG = s->aaG;
  p.aaF = s->aaF;
  p.gnew = s->gnew;
  p.aa_part = s->aa_part;
  return p;
}

static int grid_blocks(long n) {
  long b = (n + 255) / 256;
  return (int)(b > 2048 ? 2048 : b);
}
// blocks of the grid part of the refresh (two points per thread)
static int refresh_blocks(long n) { return grid_blocks((n + 1) / 2); }
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
  c.lTE2 = p->lTE2;
  c.ramp = p->ramp ? 1 : 0;
  c.dt = -1.0;
  c.c_il = c.c_m = c.c_tsl = c.c_trr = c.c_thl = 0.0;
  c.aa = p->anderson ? 1 : 0;
  c.aa_depth = p->anderson_depth;
  c.aa_freeze = p->anderson_freeze;
  c.aa_maxref = p->max_mask_refresh;
  c.aa_keep = p->anderson_keep ? 1 : 0;
  SLX_CHECK(p->precision == SLX_PREC_FP64 || p->precision == SLX_PREC_FP32, "precision must be 0 or 1");
  c.mixed = p->precision == SLX_PREC_FP32 ? 1 : 0;
  c.eps_sw = p->eps_switch;
  c.max_inner = p->max_inner > 0 ? std::min(p->max_inner, MAX_INNER) : MAX_INNER;
  SLX_CHECK(!c.mixed || (!c.aa && p->eps_switch > p->epsilon1 && h->ana_mfma),
            "mixed-precision sweeps need inner extrap or off and eps_switch > epsilon1");
  SLX_CHECK(!c.aa || (p->extrap && p->anderson_depth >= 1 && p->anderson_depth <= 5 && p->anderson_freeze >= 1 &&
                      p->max_mask_refresh >= 0),
            "anderson needs extrap = 1, depth 1..5, freeze >= 1, max_mask_refresh >= 0");
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
  SLX_HIP(hipMalloc(&s->rows, 5 * (size_t)c.nlat * sizeof(double)));
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
  // second-order placement and read-outs
  s->h_nmod = nmod;
  s->h_a_sl.assign(p->a_sl, p->a_sl + (size_t)K * nl);
  s->h_a_rr.assign(p->a_rr, p->a_rr + (size_t)K * nl);
  s->h_a2 = a2;
  s->h_ttl.assign(p->ttl, p->ttl + nl);
  s->has_horiz = p->a_hl != nullptr;
  SLX_CHECK(!s->has_horiz || (p->a_thl && p->ttlle), "slx_solver_create: a_hl needs a_thl and ttlle");
  s->h_a_hl.assign((size_t)K * nl, 0.0);
  s->h_a_thl.assign((size_t)std::max(K, 1), 0.0);
  std::vector<double> ttlle(nl, 0.0);
  if (s->has_horiz) {
    std::copy(p->a_hl, p->a_hl + (size_t)K * nl, s->h_a_hl.begin());
    std::copy(p->a_thl, p->a_thl + K, s->h_a_thl.begin());
    std::copy(p->ttlle, p->ttlle + nl, ttlle.begin());
  }
  std::vector<double> wsl((size_t)K * nl), wrr((size_t)K * nl);
  for (size_t i = 0; i < wsl.size(); ++i) {  // rate weights a_k s_k per year (solver.py:295-296)
    wsl[i] = s->h_a_sl[i] * s->h_s[i] / 1000.0;
    wrr[i] = s->h_a_rr[i] * s->h_s[i] / 1000.0;
  }
  for (double** q : {&s->omphi, &s->w_sl, &s->w_rr, &s->a_hl}) {
    SLX_HIP(hipMalloc(q, KL));
    SLX_HIP(hipMemset(*q, 0, KL));
  }
  SLX_HIP(hipMalloc(&s->a_thl, (size_t)std::max(K, 1) * sizeof(double)));
  SLX_HIP(hipMalloc(&s->ttlle, nl * sizeof(double)));
  SLX_HIP(hipMalloc(&s->deg, 7 * (size_t)nl * sizeof(double)));
  SLX_HIP(hipMemset(s->deg, 0, 7 * (size_t)nl * sizeof(double)));
  if (K > 0) {
    SLX_HIP(hipMemcpy(s->w_sl, wsl.data(), wsl.size() * sizeof(double), hipMemcpyHostToDevice));
    SLX_HIP(hipMemcpy(s->w_rr, wrr.data(), wrr.size() * sizeof(double), hipMemcpyHostToDevice));
    SLX_HIP(hipMemcpy(s->a_hl, s->h_a_hl.data(), s->h_a_hl.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  SLX_HIP(hipMemcpy(s->a_thl, s->h_a_thl.data(), s->h_a_thl.size() * sizeof(double), hipMemcpyHostToDevice));
  SLX_HIP(hipMemcpy(s->ttlle, ttlle.data(), nl * sizeof(double), hipMemcpyHostToDevice));
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
  for (double** q : {&s->tinit, &s->cxy0, &s->beta0, &s->cstar0, &s->cstar, &s->icestar, &s->dsl,
                     &s->topo_prev, &s->ice_prev, &s->tmp}) {
    SLX_HIP(hipMalloc(q, G));
    SLX_HIP(hipMemset(*q, 0, G));
  }
  for (double2** q : {&s->p_load, &s->p_ice, &s->p_ocean, &s->prev_ice, &s->icestar_lm, &s->oldice_lm, &s->cstar_lm, &s->rt_lm, &s->load_inc, &s->dS_cur, &s->visc, &s->rad, &s->h1, &s->h2}) {
    SLX_HIP(hipMalloc(q, S));
    SLX_HIP(hipMemset(*q, 0, S));
  }
  SLX_HIP(hipMalloc(&s->scratch_lm, 3 * S));
  if (c.aa) {
    const size_t H = (size_t)(c.aa_depth + 1) * S;
    for (double2** q : {&s->aaX, &s->aaG, &s->aaF}) {
      SLX_HIP(hipMalloc(q, H));
      SLX_HIP(hipMemset(*q, 0, H));
    }
    SLX_HIP(hipMalloc(&s->gnew, S));
    SLX_HIP(hipMemset(s->gnew, 0, S));
    SLX_HIP(hipMalloc(&s->aa_part, 32 * (size_t)PSLOT * sizeof(double)));
  }
  SLX_HIP(hipMalloc(&s->R, (size_t)std::max(K, 1) * S));
  SLX_HIP(hipMemset(s->R, 0, (size_t)std::max(K, 1) * S));
  SLX_HIP(hipHostMalloc(&s->pinned, sizeof(slx_status), hipHostMallocDefault));
  SLX_HIP(hipHostMalloc((void**)&s->hflag, MAX_INNER * sizeof(int), hipHostMallocCoherent));
  std::memset(s->hflag, 0, MAX_INNER * sizeof(int));
  SLX_HIP(hipHostGetDevicePointer((void**)&s->hflag_dev, s->hflag, 0));
  // the incremental re-guess needs the matrix-core tables (nh <= 512); SLX_ICE_REANALYSIS=full
  // keeps the full re-analysis
  {
    const char* e = std::getenv("SLX_ICE_REANALYSIS");
    s->ice_incr = (h->ana_mfma && !(e && std::string(e) == "full")) ? 1 : 0;
  }
  SLX_HIP(hipMalloc((void**)&s->ring_tok, (size_t)h->nh * sizeof(int)));
  SLX_HIP(hipMemset(s->ring_tok, 0, (size_t)h->nh * sizeof(int)));
  // the set-up's memsets and copies went to the null stream, which does not order itself with
  // non-blocking streams: they complete before any stream may use the solver
  SLX_HIP(hipStreamSynchronize(nullptr));
  *out = s;
  return 0;
}

extern "C" int slx_solver_destroy(slx_solver* s) {
  if (!s) return 0;
  void* ptrs[] = {s->st, s->partial, s->rows, s->w00, s->nmod, s->s, s->decay, s->omd, s->a_sl, s->a_rr, s->a2,
                  s->r2, s->ttl, s->ttlekhl, s->ttlhe, s->omphi, s->w_sl, s->w_rr, s->a_hl, s->a_thl, s->ttlle,
                  s->deg, s->aaX, s->aaG, s->aaF, s->gnew, s->aa_part, s->tinit, s->cxy0, s->beta0, s->cstar0,
                  s->cstar, s->icestar, s->dsl, s->topo_prev, s->ice_prev, s->tmp,
                  s->p_load, s->p_ice, s->p_ocean, s->prev_ice, s->icestar_lm, s->oldice_lm,
                  s->cstar_lm, s->rt_lm, s->load_inc, s->dS_cur, s->visc, s->rad, s->h1, s->h2, s->scratch_lm,
                  s->R};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (s->pinned) (void)hipHostFree(s->pinned);
  if (s->hflag) (void)hipHostFree(s->hflag);
  if (s->run_stat) (void)hipHostFree(s->run_stat);
  if (s->ring_tok) (void)hipFree(s->ring_tok);
  if (s->h && s->h->marks == &s->marks) s->h->marks = nullptr;
  for (hipEvent_t e : s->marks.ev) (void)hipEventDestroy(e);
  delete s;
  return 0;
}

// phi(x) = (1 - e^-x)/x and 1 - phi(x), third-order series where |x| < 1e-4 (solver.py:229-241)
static void ramp_factor(double x, double& phi, double& omp) {
  if (std::fabs(x) < 1.0e-4) {
    phi = 1.0 - x / 2.0 + x * x / 6.0 - std::pow(x, 3.0) / 24.0;
    omp = x / 2.0 - x * x / 6.0 + std::pow(x, 3.0) / 24.0;
  } else {
    phi = -std::expm1(-x) / x;
    omp = (x + std::expm1(-x)) / x;
  }
}

extern "C" int slx_solver_set_dt(slx_solver* s, double dt) {
  SLX_CHECK(s != nullptr, "null solver");
  if (dt == s->dt) return 0;
  // enqueued steps still read the decay, ramp and degree tables rewritten below
  if (s->used) SLX_HIP(hipStreamSynchronize(s->last_stream));
  const int K = s->c.kmax, nl = s->c.nl;
  const std::vector<int>& nmod = s->h_nmod;
  std::vector<double> d((size_t)K * nl, 0.0), o((size_t)K * nl, 0.0);
  for (int l = 1; l < nl; ++l)
    for (int k = 0; k < nmod[l] && k < K; ++k) d[(size_t)k * nl + l] = std::exp(-s->h_s[(size_t)k * nl + l] * dt / 1000.0);
  for (size_t i = 0; i < d.size(); ++i) o[i] = 1.0 - d[i];
  // ramp factors and their sums over the modes, k in order (solver.py:348-367)
  std::vector<double> phi((size_t)K * nl), omphi((size_t)K * nl), deg(7 * (size_t)nl, 0.0), chl(nl, 0.0);
  for (size_t i = 0; i < phi.size(); ++i) {
    const double sv = s->h_s[i];
    double ph, om;
    ramp_factor(sv * (dt / 1000.0), ph, om);
    phi[i] = sv > 0 ? ph : 1.0;
    omphi[i] = sv > 0 ? om : 0.0;
  }
  std::vector<double> csl(nl, 0.0), crr(nl, 0.0), wssl(nl, 0.0), wsrr(nl, 0.0), wpsl(nl, 0.0), wprr(nl, 0.0);
  for (int k = 0; k < K; ++k)
    for (int l = 0; l < nl; ++l) {
      const size_t i = (size_t)k * nl + l;
      const double wsl = s->h_a_sl[i] * s->h_s[i] / 1000.0, wrr = s->h_a_rr[i] * s->h_s[i] / 1000.0;
      csl[l] += s->h_a_sl[i] * omphi[i];
      crr[l] += s->h_a_rr[i] * omphi[i];
      wssl[l] += wsl;
      wsrr[l] += wrr;
      wpsl[l] += wsl * phi[i];
      wprr[l] += wrr * phi[i];
      chl[l] += s->h_a_hl[i] * omphi[i];
    }
  for (int l = 0; l < nl; ++l) {
    const double t = s->h_ttl[l];
    deg[l] = t * csl[l];
    deg[nl + l] = t * crr[l];
    deg[2 * nl + l] = wssl[l];
    deg[3 * nl + l] = wsrr[l];
    deg[4 * nl + l] = wpsl[l];
    deg[5 * nl + l] = wprr[l];
    deg[6 * nl + l] = t * chl[l];
  }
  Consts& c = s->c;
  c.c_il = c.c_m = c.c_tsl = c.c_trr = c.c_thl = 0.0;
  if (c.L >= 2)
    for (int k = 0; k < c.nmod2; ++k) {
      const double w = omphi[(size_t)k * nl + 2];
      c.c_il += w * s->h_a2[k];
      c.c_m += w * s->h_a2[K + k];
      c.c_tsl += w * s->h_a2[2 * K + k];
      c.c_trr += w * s->h_a2[3 * K + k];
      c.c_thl += w * s->h_a_thl[k];
    }
  c.dt = dt;
  if (K > 0) {
    SLX_HIP(hipMemcpy(s->decay, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice));
    SLX_HIP(hipMemcpy(s->omd, o.data(), o.size() * sizeof(double), hipMemcpyHostToDevice));
    SLX_HIP(hipMemcpy(s->omphi, omphi.data(), omphi.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  SLX_HIP(hipMemcpy(s->deg, deg.data(), deg.size() * sizeof(double), hipMemcpyHostToDevice));
  SLX_HIP(hipStreamSynchronize(nullptr));  // the copies land before work on any stream reads them
  s->dt = dt;
  return 0;
}

extern "C" int slx_solver_init(slx_solver* s, const double* topo0, const double* ice0, void* stream) {
  SLX_CHECK(s && topo0 && ice0, "slx_solver_init: null argument");
  hipStream_t st = (hipStream_t)stream;
  s->last_stream = st;
  s->used = 1;
  Pts p = make_pts(s, nullptr, s->oldice_lm);
  k_init_grid<<<grid_blocks(s->c.npts), 256, 0, st>>>(p, s->c, topo0, ice0);
  SLX_LAUNCH_CHECK();
  // degree-0 areas of C0, C*0 and I*0 (solver.py:571-580); the I*0 spectrum is the "previous ice"
  // of step 1 (solver.py:859-862)
  int rc = launch_fft_ana(s->h, LoadPtrs<3>{{s->cxy0, s->cstar0, s->tmp}}, 3, Pred{}, st);
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

// sparse contraction of the ring pairs tagged with token (the change of I*, WS/WD field 0), added
// to I*_lm: a persistent grid of one-wave blocks, sized once to what is resident
static int launch_ice_sparse(slx_solver* s, int token, Pred pr, hipStream_t st) {
  const slx_sht* h = s->h;
  static int cap = 0;
  if (cap == 0) {
    int dev = 0, ncu = 0, per_cu = 0;
    SLX_HIP(hipGetDevice(&dev));
    SLX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    SLX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_leg_ana_sp<2>), 64, 0));
    cap = 8 * std::max(1, ncu * std::max(per_cu, 1) / 8);
  }
  const int grid = std::min(cap, 8 * ((h->ngrp + 7) / 8));
  GrpDev gr{h->ngrp, h->grp_m, h->grp_s0, h->grp_desc};
  k_leg_ana_sp<2><<<grid, 64, 0, st>>>(dev_of(h), seg_of(h), gr, h->WS, h->WD, s->icestar_lm, s->ring_tok, token, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}

template <int P>
static int launch_begin_ring_p(slx_solver* s, const LoadIceGrid& lg, int token, hipStream_t st) {
  const slx_sht* h = s->h;
  const size_t lds = ring_lds(h);
  if (int rc = allow_lds<&k_begin_ring<P>>(lds)) return rc;
  k_begin_ring<P><<<h->nh, h->nlon / P, lds, st>>>(dev_of(h), lg, h->WS, h->WD, s->ring_tok, token);
  SLX_LAUNCH_CHECK();
  return 0;
}

extern "C" int slx_step_begin_warm(slx_solver* s, const double* ice_cur, double dt0, double dt1, double dt2,
                                   const void* warm_dS, const double* warm_ocean, void* stream) {
  SLX_CHECK(s && ice_cur, "slx_step_begin: null argument");
  SLX_CHECK((warm_dS == nullptr) == (warm_ocean == nullptr),
            "slx_step_begin_warm: give both warm_dS and warm_ocean, or neither");
  SLX_CHECK(s->dt > 0.0, "slx_solver_set_dt must be called before the first step");
  hipStream_t st = (hipStream_t)stream;
  SLX_CHECK(s->closed, "slx_step_begin: the previous step has not finished (check slx_solver_status)");
  s->last_stream = st;
  s->used = 1;
  s->step += 1;
  s->epoch += 1;
  s->closed = 0;
  s->iters_enqueued = 0;
  s->ice_cur = ice_cur;
  const double2* oldice = (s->step == 1) ? s->oldice_lm : s->prev_ice;
  Pts p = make_pts(s, ice_cur, oldice);
  const Consts& c = s->c;
  slx_mark(s->h, -1, st);
  k_begin<<<19, 64, 0, st>>>(p, c);
  SLX_LAUNCH_CHECK();
  slx_mark(s->h, SLX_STAGE_ROT, st);
  int rc = launch_modal(s->h, c.kmax, s->nmod, s->decay, s->omd, s->a_sl, s->a_rr, s->R, s->p_load, s->visc,
                        s->rad, st);
  if (rc) return rc;
  slx_mark(s->h, SLX_STAGE_MEMORY, st);
  if (warm_ocean) {
    SLX_HIP(hipMemcpyAsync(s->cxy, warm_ocean, (size_t)c.npts * sizeof(double), hipMemcpyDeviceToDevice, st));
    slx_mark(s->h, SLX_STAGE_GRID, st);
  }
  const int extrap_ready = (s->extrap && s->nhist >= 2) ? 1 : 0;
  const LoadIceGrid lg{ice_cur, s->topo_prev, s->cxy, s->icestar, s->beta, s->cstar, c.rhow, c.rhoi, c.checkmarine};
  // the eustatic first guess reads all of C*_lm, so only the extrapolated start defers C*
  s->cstar_pending = (s->ice_incr && extrap_ready && !c.aa) ? 1 : 0;
  if (s->cstar_pending) {
    const int token = ++s->next_tok;
    rc = ring_points(s->h) == 4 ? launch_begin_ring_p<4>(s, lg, token, st) : launch_begin_ring_p<8>(s, lg, token, st);
    if (!rc) rc = launch_ice_sparse(s, token, Pred{}, st);
    if (rc) return rc;
  } else {
    rc = launch_fft_ana(s->h, lg, 2, Pred{}, st);
    if (rc) return rc;
    rc = launch_leg_ana<2>(s->h, s->h->WS, s->h->WD, OutPtrs<2>{{s->icestar_lm, s->cstar_lm}}, Pred{}, st);
    if (rc) return rc;
  }
  slx_mark(s->h, SLX_STAGE_SHT, st);
  GuessArgs g;
  g.extrap_ready = extrap_ready;
  g.uniform = (dt0 == dt1 && dt1 == dt2) ? 1 : 0;
  const double w = (dt0 + dt1) / (dt1 + dt2);
  g.c1 = dt0 / dt1 * (1.0 + w);
  g.c2 = dt0 / dt2 * w;
  g.warm = static_cast<const double2*>(warm_dS);
  k_guess<<<red_blocks(c.ncoef), 256, 0, st>>>(p, c, g);
  SLX_LAUNCH_CHECK();
  slx_mark(s->h, SLX_STAGE_SPEC, st);
  return 0;
}

// n inner iterations with Anderson mixing (solver_opt.py:532-535,705-775): the same ring and
// Legendre passes (C* analysed only after a mask update), then the decision lane and the
// kernels of the chosen branch, each gated by its iteration token.
// incremental grounded re-guess: the ring pass of the change, then the sparse contraction (a
// persistent grid of one-wave blocks, sized once to what is resident)
template <int P>
static int launch_reguess_ring_p(slx_solver* s, const LoadReguess& ld, int token, Pred pr, const Pts& p, int decide,
                                 hipStream_t st) {
  const slx_sht* h = s->h;
  const size_t lds = ring_lds(h);
  if (int rc = allow_lds<&k_reguess_ring<P>>(lds)) return rc;
  k_reguess_ring<P><<<h->nh, h->nlon / P, lds, st>>>(dev_of(h), ld, h->WS, h->WD, s->ring_tok, token, pr, p, s->c,
                                                     red_blocks(s->c.ncoef), decide);
  SLX_LAUNCH_CHECK();
  return 0;
}
// decide: the ring pass first takes the iteration's convergence decision (after k_update_refresh)
static int launch_reguess_incremental(slx_solver* s, const LoadReguess& ld, Pred pr_ring, Pred pr_ana, const Pts& p,
                                      int decide, hipStream_t st) {
  const slx_sht* h = s->h;
  const int token = ++s->next_tok;
  int rc = ring_points(h) == 4 ? launch_reguess_ring_p<4>(s, ld, token, pr_ring, p, decide, st)
                               : launch_reguess_ring_p<8>(s, ld, token, pr_ring, p, decide, st);
  if (rc) return rc;
  slx_mark(h, SLX_STAGE_GRID, st);
  return launch_ice_sparse(s, token, pr_ana, st);
}

// ring pass of an inner iteration (k_fft_synana) on the solver's grids
template <int P, bool REG>
static int launch_fft_synana_p(slx_solver* s, int nf, const int* cst, Pred pr, hipStream_t st) {
  const slx_sht* h = s->h;
  const size_t lds = ring_lds(h);
  if (int rc = allow_lds<&k_fft_synana<P, REG>>(lds)) return rc;
  // C* transformed by blocks of its own where a ring pair per CU leaves the CUs idle (1 872 -> 1 885
  // steps/s at lmax 512; at nlon 2048 two blocks per CU are resident already: 585 -> 577 at lmax 1023)
  const int ny = ((nf == 2 || cst != nullptr) && h->nlon <= 1024) ? 2 : 1;
  k_fft_synana<P, REG><<<dim3(h->nh, ny), h->nlon / P, lds, st>>>(dev_of(h), h->FN, h->FSo, s->dsl,
                                                         OceanFn{s->cxy, s->beta, s->tinit, s->cstar0}, h->WS, h->WD,
                                                         nf, cst, pr);
  SLX_LAUNCH_CHECK();
  return 0;
}
static int launch_fft_synana(slx_solver* s, int nf, const int* cst, Pred pr, hipStream_t st) {
  if (ring_points(s->h) != 4) return launch_fft_synana_p<8, false>(s, nf, cst, pr, st);
  return s->h->nlon <= 1024 ? launch_fft_synana_p<4, true>(s, nf, cst, pr, st)
                            : launch_fft_synana_p<4, false>(s, nf, cst, pr, st);
}

static int iterate_anderson(slx_solver* s, int n, hipStream_t st) {
  const Consts& c = s->c;
  const double2* oldice = (s->step == 1) ? s->oldice_lm : s->prev_ice;
  Pts p = make_pts(s, s->ice_cur, oldice);
  const int* done = &s->st->st.done;
  const int* changed = &s->st->ice_changed;
  const int* cst = &s->st->aa_cst;
  const int nrb = red_blocks(c.ncoef);
  for (int it = 0; it < n; ++it) {
    const int it_step = s->iters_enqueued;
    s->iters_enqueued += 1;
    CoefB18 cb{s->icestar_lm, oldice, s->p_ice, s->p_ocean, s->load_inc, s->visc, s->dS_cur, s->ttlekhl, s->ttl, s->deg,
               s->st, c, 1};
    int rc = launch_leg_syn<1>(s->h, cb, Pred{done, 0}, st);
    if (rc) return rc;
    rc = launch_fft_synana(s, 1, cst, Pred{done, 0}, st);
    if (rc) return rc;
    rc = launch_leg_ana<2>(s->h, s->h->WS, s->h->WD, OutPtrs<2>{{s->rt_lm, s->cstar_lm}}, Pred{done, 0, cst, 1}, st);
    if (!rc) rc = launch_leg_ana<1>(s->h, s->h->WS, s->h->WD, OutPtrs<1>{{s->rt_lm}}, Pred{done, 0, cst, 0}, st);
    if (rc) return rc;
    slx_mark(s->h, SLX_STAGE_SHT, st);
    k_update<<<nrb, 256, 0, st>>>(p, c, it_step, s->gnew);
    SLX_LAUNCH_CHECK();
    slx_mark(s->h, SLX_STAGE_SPEC, st);
    k_refresh_aa<<<grid_blocks(c.npts), 256, 0, st>>>(p, c, it_step, nrb, nrb);
    SLX_LAUNCH_CHECK();
    k_decide_aa<<<1, 64, 0, st>>>(p, c, it_step);
    SLX_LAUNCH_CHECK();
    k_refresh2_aa<<<grid_blocks(c.npts), 256, 0, st>>>(p, c, it_step);
    SLX_LAUNCH_CHECK();
    slx_mark(s->h, SLX_STAGE_GRID, st);
    if (c.checkmarine) {
      rc = launch_fft_ana(s->h, LoadPtrs<1>{{s->icestar}}, 1, Pred{done, 0, changed, 1}, st);
      if (!rc) rc = launch_leg_ana<1>(s->h, s->h->WS, s->h->WD, OutPtrs<1>{{s->icestar_lm}}, Pred{done, 0, changed, 1}, st);
      if (rc) return rc;
      slx_mark(s->h, SLX_STAGE_SHT, st);
    }
    k_aa_push<<<nrb, 256, 0, st>>>(p, c, it_step);
    SLX_LAUNCH_CHECK();
    k_aa_solve<<<1, 256, 0, st>>>(p, c, it_step, nrb);
    SLX_LAUNCH_CHECK();
    k_aa_mix<<<nrb, 256, 0, st>>>(p, c, it_step);
    SLX_LAUNCH_CHECK();
    k_aa_plain<<<grid_blocks(c.ncoef), 256, 0, st>>>(p, c, it_step);
    SLX_LAUNCH_CHECK();
    slx_mark(s->h, SLX_STAGE_SPEC, st);
  }
  return 0;
}

// An inner iteration (extrapolated or plain start) in two parts: the head (B18 coefficients and the
// synthesis contraction) and the tail (the fused ring pass through the grounded re-guess).  The
// native step loop enqueues the next iteration's head before the current one's decision is known;
// both parts are predicated on the device's done flag.
static int iter_head(slx_solver* s, hipStream_t st) {
  const double2* oldice = (s->step == 1) ? s->oldice_lm : s->prev_ice;
  CoefB18 cb{s->icestar_lm, oldice, s->p_ice, s->p_ocean, s->load_inc, s->visc, s->dS_cur, s->ttlekhl, s->ttl, s->deg,
             s->st, s->c, 1};
  const int* done = &s->st->st.done;
  if (!s->c.mixed) {
    const int rc = launch_leg_syn<1>(s->h, cb, Pred{done, 0}, st);
    slx_mark(s->h, SLX_STAGE_SHT, st);
    return rc;
  }
  cb.it = s->iters_enqueued;
  // mixed sweeps: the coefficients once, then the FP64 or the FP32-product contraction as the
  // device decided for this iteration (head of iteration iters_enqueued)
  const slx_sht* h = s->h;
  k_coef_fill<1, CoefB18><<<(h->ncoef + 255) / 256, 256, 0, st>>>(dev_of(h), cb, h->CS, h->ncoef, h->ncoef,
                                                                   Pred{done, 0});
  SLX_LAUNCH_CHECK();
  slx_mark(h, SLX_STAGE_SPEC, st);
  const int* sw = &s->st->sweep64[s->iters_enqueued & 1];
  int rc = launch_leg_syn<1>(h, CoefPlain{h->CS, h->ncoef}, Pred{done, 0, sw, 1}, st);
  if (!rc) rc = launch_leg_syn32<1>(h, CoefPlain{h->CS, h->ncoef}, Pred{done, 0, sw, 0}, st);
  slx_mark(h, SLX_STAGE_SHT, st);
  return rc;
}

static int iter_tail(slx_solver* s, hipStream_t st) {
  const Consts& c = s->c;
  const double2* oldice = (s->step == 1) ? s->oldice_lm : s->prev_ice;
  Pts p = make_pts(s, s->ice_cur, oldice);
  const int* done = &s->st->st.done;
  const int* changed = &s->st->ice_changed;
  const bool first = (s->iters_enqueued == 0);
  const int it_step = s->iters_enqueued;
  s->iters_enqueued += 1;
  // C* was analysed at the step's start unless that pass was incremental
  const bool rt_only = first && !s->cstar_pending;
  int rc = launch_fft_synana(s, rt_only ? 1 : 2, nullptr, Pred{done, 0}, st);
  if (rc) return rc;
  if (!c.mixed) {
    if (rt_only)
      rc = launch_leg_ana<1>(s->h, s->h->WS, s->h->WD, OutPtrs<1>{{s->rt_lm}}, Pred{done, 0}, st);
    else
      rc = launch_leg_ana<2>(s->h, s->h->WS, s->h->WD, OutPtrs<2>{{s->rt_lm, s->cstar_lm}}, Pred{done, 0}, st);
  } else {  // the FP64 or the FP32-product contraction, as decided for this iteration
    const int* sw = &s->st->sweep64[it_step & 1];
    const Pred p64{done, 0, sw, 1}, p32{done, 0, sw, 0};
    if (rt_only) {
      rc = launch_leg_ana<1>(s->h, s->h->WS, s->h->WD, OutPtrs<1>{{s->rt_lm}}, p64, st);
      if (!rc) rc = launch_leg_ana32<1>(s->h, s->h->WS, s->h->WD, OutPtrs<1>{{s->rt_lm}}, p32, st);
    } else {
      rc = launch_leg_ana<2>(s->h, s->h->WS, s->h->WD, OutPtrs<2>{{s->rt_lm, s->cstar_lm}}, p64, st);
      if (!rc) rc = launch_leg_ana32<2>(s->h, s->h->WS, s->h->WD, OutPtrs<2>{{s->rt_lm, s->cstar_lm}}, p32, st);
    }
  }
  if (rc) return rc;
  slx_mark(s->h, SLX_STAGE_SHT, st);
  // With the incremental re-guess (and no stage timing, which keeps the launches apart) the update
  // and the refresh share one launch and the re-guess ring pass takes the convergence decision.
  const bool fused = c.checkmarine && s->ice_incr && s->h->marks != &s->marks && s->h->nlon / ring_points(s->h) >= 64;
  if (fused) {
    const int nbu = red_blocks(c.ncoef);
    k_update_refresh<<<nbu + refresh_blocks(c.npts), 256, 0, st>>>(p, c, it_step, nbu, s->load_inc);
  } else {
    k_update<<<red_blocks(c.ncoef), 256, 0, st>>>(p, c, it_step, s->load_inc);
    SLX_LAUNCH_CHECK();
    slx_mark(s->h, SLX_STAGE_SPEC, st);
    k_refresh<<<refresh_blocks(c.npts), 256, 0, st>>>(p, c, it_step, red_blocks(c.ncoef), red_blocks(c.ncoef));
  }
  SLX_LAUNCH_CHECK();
  slx_mark(s->h, SLX_STAGE_GRID, st);
  if (c.checkmarine) {
    const LoadReguess ld{s->ice_cur, s->topo, s->icestar, s->beta, c.rhow, c.rhoi, s->st, it_step, c.checkmarine};
    if (s->ice_incr) {  // the re-guess changes I* at a few points: analyse the change only
      rc = launch_reguess_incremental(s, ld, Pred{done, 0}, Pred{done, 0, changed, 1}, p, fused ? 1 : 0, st);
    } else {
      rc = launch_fft_ana(s->h, ld, 1, Pred{done, 0}, st);
      slx_mark(s->h, SLX_STAGE_GRID, st);
      if (!rc)
        rc = launch_leg_ana<1>(s->h, s->h->WS, s->h->WD, OutPtrs<1>{{s->icestar_lm}}, Pred{done, 0, changed, 1}, st);
    }
    if (rc) return rc;
    slx_mark(s->h, SLX_STAGE_SHT, st);
  }
  return 0;
}

extern "C" int slx_step_iterate(slx_solver* s, int n, void* stream) {
  SLX_CHECK(s != nullptr, "null solver");
  hipStream_t st = (hipStream_t)stream;
  if (s->c.aa) return iterate_anderson(s, n, st);
  for (int it = 0; it < n; ++it) {
    int rc = iter_head(s, st);
    if (!rc) rc = iter_tail(s, st);
    if (rc) return rc;
  }
  return 0;
}

namespace {
struct StorePair {  // synthesis output of two fields into two separate grids
  double *g0, *g1;
  __device__ void operator()(int f, long i, double v) const { (f == 0 ? g0 : g1)[i] = v; }
};
}  // namespace

extern "C" int slx_step_finish(slx_solver* s, void* stream) { return slx_step_finish_readouts(s, nullptr, stream); }

extern "C" int slx_step_finish_readouts(slx_solver* s, const slx_readouts* ro, void* stream) {
  SLX_CHECK(s != nullptr, "null solver");
  hipStream_t st = (hipStream_t)stream;
  const Consts& c = s->c;
  const double2* oldice = (s->step == 1) ? s->oldice_lm : s->prev_ice;
  Pts p = make_pts(s, s->ice_cur, oldice);
  const Pred fin{&s->st->st.done, 1, &s->st->st.finished, 0};
  bool rates = false, hor = false;
  if (ro) {
    const int given = (ro->dSLdt != nullptr) + (ro->dRdt != nullptr) + (ro->dSLdt_lm != nullptr) +
                      (ro->dRdt_lm != nullptr) + (ro->rot_rates != nullptr);
    SLX_CHECK(given == 0 || given == 5,
              "slx_step_finish_readouts: give all of dSLdt, dRdt, dSLdt_lm, dRdt_lm, rot_rates, or none");
    rates = given == 5;
    hor = ro->horiz_lm != nullptr;
    SLX_CHECK(!hor || s->has_horiz, "slx_step_finish_readouts: horizontal read-out needs a_hl at solver creation");
  }
  if (c.calcRG) {
    CoefB18 cb{s->icestar_lm, oldice, s->p_ice, s->p_ocean, s->load_inc, s->rad, s->dS_cur, s->ttlhe, s->ttl,
               s->deg + c.nl, s->st, c, 0};
    int rc = launch_leg_syn<1>(s->h, cb, fin, st);
    if (rc) return rc;
    rc = launch_fft_syn(s->h, StorePlain{s->rr, 0}, 1, fin, st);
    if (rc) return rc;
    slx_mark(s->h, SLX_STAGE_SHT, st);
  }
  double2* dsl_user = rates ? static_cast<double2*>(ro->dSLdt_lm) : nullptr;
  if (rates || hor) {
    // read-outs on the states before the commit (solver.py:1074-1093)
    k_rot_readout<<<1, 64, 0, st>>>(p, c, rates ? 1 : 0, hor ? 1 : 0, rates ? ro->rot_rates : nullptr);
    SLX_LAUNCH_CHECK();
    slx_mark(s->h, SLX_STAGE_ROT, st);
    ReadoutSpec a{};
    if (rates) {
      a.dsl = s->scratch_lm;
      a.drr = s->scratch_lm + c.ncoef;
      a.dsl_user = dsl_user;
      a.drr_user = static_cast<double2*>(ro->dRdt_lm);
    }
    if (hor) a.hz = static_cast<double2*>(ro->horiz_lm);
    k_readout_spec<<<(c.ncoef + 255) / 256, 256, 0, st>>>(p, c, a);
    SLX_LAUNCH_CHECK();
    slx_mark(s->h, SLX_STAGE_SPEC, st);
    if (rates) {  // rate grids without the degree-0 term (solver.py:797-798)
      int rc = launch_leg_syn<2>(s->h, CoefPlain{s->scratch_lm, c.ncoef}, fin, st);
      if (rc) return rc;
      rc = launch_fft_syn(s->h, StorePair{ro->dSLdt, ro->dRdt}, 2, fin, st);
      if (rc) return rc;
      slx_mark(s->h, SLX_STAGE_SHT, st);
    }
  }
  k_commit<<<(c.ncoef + 255) / 256, 256, 0, st>>>(p, c, s->extrap);
  SLX_LAUNCH_CHECK();
  slx_mark(s->h, SLX_STAGE_MEMORY, st);
  k_post_grid<<<c.nlat, 256, 0, st>>>(p, c, rates ? ro->dSLdt : nullptr);
  SLX_LAUNCH_CHECK();
  slx_mark(s->h, SLX_STAGE_GRID, st);
  k_finish<<<1, 256, 0, st>>>(p, c, rates ? 1 : 0, dsl_user, rates ? 0 : 1);
  SLX_LAUNCH_CHECK();
  slx_mark(s->h, SLX_STAGE_ROT, st);
  if (rates) {
    k_rate_grid<<<grid_blocks(c.npts), 256, 0, st>>>(p, c, ro->dSLdt);
    SLX_LAUNCH_CHECK();
    slx_mark(s->h, SLX_STAGE_GRID, st);
    k_close<<<1, 64, 0, st>>>(p);
    SLX_LAUNCH_CHECK();
    slx_mark(s->h, SLX_STAGE_ROT, st);
  }
  return 0;
}

extern "C" int slx_solver_profile(slx_solver* s, int on) {
  SLX_CHECK(s != nullptr, "null solver");
  SLX_CHECK(!on || s->h->marks == nullptr || s->h->marks == &s->marks,
            "slx_solver_profile: another solver is profiling on this transform handle");
  if (on) {
    s->marks.n = 0;
    s->marks.err = hipSuccess;
    s->h->marks = &s->marks;
  } else if (s->h->marks == &s->marks) {
    s->h->marks = nullptr;
  }
  return 0;
}

extern "C" int slx_solver_stage_times(slx_solver* s, double* ms) {
  SLX_CHECK(s && ms, "slx_solver_stage_times: null argument");
  slx_stage_marks& m = s->marks;
  for (int i = 0; i < SLX_NSTAGE; ++i) ms[i] = 0.0;
  if (m.err != hipSuccess) {
    const hipError_t e = m.err;
    m.err = hipSuccess;
    m.n = 0;
    slx_set_error(std::string("stage timing: ") + hipGetErrorString(e));
    return (int)e;
  }
  if (m.n > 0) SLX_HIP(hipEventSynchronize(m.ev[m.n - 1]));
  for (int i = 1; i < m.n; ++i) {
    if (m.stage[i] < 0) continue;
    float t = 0.0f;
    SLX_HIP(hipEventElapsedTime(&t, m.ev[i - 1], m.ev[i]));
    ms[m.stage[i]] += t;
  }
  m.n = 0;
  return 0;
}

// Host bookkeeping once the step's finish has run on the device (status.finished), or has been
// enqueued after a posted convergence decision: the newest converged increment (written to h2 by
// the commit) becomes h1.
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

// Waits (spinning) until the device posts the decision of inner iteration it of the current step;
// code = 1 continue, 2 converged, 3 failed.  The stream is polled now and then, so that work that
// ended without posting (or an error) ends the wait.
static int wait_decision(slx_solver* s, int it, hipStream_t st, int& code) {
  const volatile int* f = s->hflag + it;
  const int token = (int)(s->epoch & 0x1fffffffu);
  for (long spin = 1;; ++spin) {
    int v = *f;
    if ((v >> 2) == token && (v & 3)) {
      code = v & 3;
      return 0;
    }
    if ((spin & 4095) == 0) {
      const hipError_t q = hipStreamQuery(st);
      if (q == hipSuccess) {
        v = *f;
        if ((v >> 2) == token && (v & 3)) {
          code = v & 3;
          return 0;
        }
        slx_set_error("slx_run_steps: the stream drained without the decision of iteration " + std::to_string(it));
        return -35;
      }
      if (q != hipErrorNotReady) {
        slx_set_error(std::string("slx_run_steps: ") + hipGetErrorString(q));
        return -36;
      }
    }
    __builtin_ia32_pause();
  }
}

// Native step loop (the loop of run_ais.py:389-425 without per-step host outputs), pipelined: the
// kernel that decides an inner iteration (the refresh pass, or the Anderson branch kernel) posts
// the decision to pinned host memory as it starts, and the host enqueues the next iteration only
// on a "continue", while the rest of the current one still runs on the device; the head of the
// iteration after that (B18 coefficients and synthesis, iter_head) is enqueued ahead of the
// decision as well, so the device need not wait for the host between iterations (after
// convergence that head runs predicated off).  No iteration tail is enqueued past convergence and
// the host never waits for the device to drain between steps: on "converged" the step's finish and
// the next step's start follow at once; each finish copies the
// step's status into a pinned slot, read after the last step.  *guess is not needed by this
// schedule (kept for the interface; set to the last step's count).
extern "C" int slx_run_steps(slx_solver* s, const double* ice, long ice_stride, int nsteps, const double* dts,
                             const double* dt_prev, int* guess, slx_status* out, void* stream) {
  SLX_CHECK(s && ice && dts && dt_prev && guess && out, "slx_run_steps: null argument");
  SLX_CHECK(nsteps >= 0 && ice_stride >= s->c.npts, "slx_run_steps: bad nsteps or ice_stride");
  if (nsteps == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  struct NoMarks {  // the native loop is not stage-profiled
    slx_sht* h;
    slx_stage_marks* saved;
    ~NoMarks() { h->marks = saved; }
  } nomarks{s->h, s->h->marks};
  s->h->marks = nullptr;
  if (nsteps > s->run_stat_cap) {  // grown on demand; later calls of at most this length allocate nothing
    SLX_HIP(hipStreamSynchronize(st));
    if (s->run_stat) SLX_HIP(hipHostFree(s->run_stat));
    s->run_stat = nullptr;
    s->run_stat_cap = 0;
    SLX_HIP(hipHostMalloc((void**)&s->run_stat, (size_t)nsteps * sizeof(slx_status), hipHostMallocCoherent));
    s->run_stat_cap = nsteps;
  }
  slx_status* hst = s->run_stat;
  std::memset(hst, 0, (size_t)nsteps * sizeof(slx_status));
  slx_status* hst_dev = nullptr;
  SLX_HIP(hipHostGetDevicePointer((void**)&hst_dev, hst, 0));
  double d1 = dt_prev[0], d2 = dt_prev[1];
  int rc = 0, failed_at = -1;
  s->pipelined = 1;
  for (int k = 0; k < nsteps && !rc; ++k) {
    rc = slx_solver_set_dt(s, dts[k]);
    if (!rc) rc = slx_step_begin(s, ice + (long)k * ice_stride, dts[k], d1, d2, stream);
    int code = 0;
    if (s->c.aa) {  // Anderson mixing: one iteration enqueued per posted decision
      if (!rc) rc = slx_step_iterate(s, 1, stream);
      for (int it = 0; !rc; ++it) {
        rc = wait_decision(s, it, st, code);
        if (rc || code != 1) break;
        rc = slx_step_iterate(s, 1, stream);
      }
    } else {  // iteration 0 and the head of iteration 1; then, per continue, a tail and the next head
      if (!rc) rc = iter_head(s, st);
      if (!rc) rc = iter_tail(s, st);
      if (!rc) rc = iter_head(s, st);
      for (int it = 0; !rc; ++it) {
        rc = wait_decision(s, it, st, code);
        if (rc || code != 1) break;
        rc = iter_tail(s, st);
        if (!rc) rc = iter_head(s, st);
      }
    }
    if (rc) break;
    if (code == 3) {  // not finished, not committed; its status is read after the stream drains
      failed_at = k;
      break;
    }
    s->hstat_dev = hst_dev + k;
    rc = slx_step_finish(s, stream);
    s->hstat_dev = nullptr;
    if (!rc) step_close(s);
    d2 = d1;
    d1 = dts[k];
  }
  s->pipelined = 0;
  s->hstat_dev = nullptr;
  const hipError_t se = hipStreamSynchronize(st);
  if (!rc && se != hipSuccess) {
    slx_set_error(std::string("slx_run_steps: ") + hipGetErrorString(se));
    rc = -36;
  }
  if (!rc && failed_at >= 0) {
    for (int k = 0; k < failed_at; ++k) out[k] = hst[k];
    SLX_HIP(hipMemcpy(&out[failed_at], &s->st->st, sizeof(slx_status), hipMemcpyDeviceToHost));
    slx_set_error("slx_run_steps: inner loop failed to converge at step " + std::to_string(out[failed_at].step));
    rc = -34;
  } else if (!rc) {
    for (int k = 0; k < nsteps; ++k) {
      out[k] = hst[k];
      if (!out[k].finished) {
        slx_set_error("slx_run_steps: step " + std::to_string(k) + " did not post its status");
        rc = -37;
        break;
      }
    }
    if (!rc) *guess = std::max(1, out[nsteps - 1].numiter);
  }
  return rc;
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
