/* libslx: MI355X (gfx950) HIP kernels behind the SLEvolve sea-level time-step API.
 *
 * C-ABI only: plain pointers and sizes.  Pointers documented "device" are HIP device pointers
 * (torch tensors' data_ptr() in the Python layer); "host" pointers are ordinary host memory.
 * Every function returns 0 on success, a HIP error code (> 0) or a negative code for a bad
 * argument; slx_last_error() returns a thread-local message.  `stream` is a hipStream_t
 * (NULL = the default stream); launches are asynchronous unless stated otherwise.
 *
 * Spectral arrays on the device are packed, order-major triangles: coefficient (l, m),
 * 0 <= m <= l <= lmax, at index moff(m) + (l - m), moff(m) = m (lmax+1) - m (m-1)/2,
 * ncoef = (lmax+1)(lmax+2)/2, complex double (re, im).  Grids are row-major double
 * [nlat][nlon], Gauss-Legendre rows north -> south.  Conventions are those of the reference's
 * python/slevolve/sht.py:6-31 (legacy spharmt: orthonormal P~ without Condon-Shortley phase,
 * the truncated 1.41421356 literal).
 */
#ifndef SLX_H
#define SLX_H

#ifdef __cplusplus
extern "C" {
#endif

#define SLX_ABI_VERSION 3

/* Arithmetic of the Legendre contractions: FP64 throughout, or the FP64 degree recurrence feeding
 * FP32 matrix-core products accumulated in FP32 (inputs, outputs and every other stage FP64). */
#define SLX_PREC_FP64 0
#define SLX_PREC_FP32 1

typedef struct slx_sht slx_sht;
typedef struct slx_solver slx_solver;

const char* slx_last_error(void);
int slx_abi_version(void);

/* ------------------------------------------------------------------ transform ------------
 * Replaces the constructor of python/slevolve/sht.py:151-255 (SHT(nlat, nlon, lmax)) and of
 * sht_ducc.py:70-132 (DuccSHT), reached through make_sht (sht_ducc.py:206-217).  Builds the
 * Gauss-Legendre nodes/weights (gaulw, sht.py:65-101), the sectoral P~_mm and recurrence
 * coefficients (legend, sht.py:105-144) and FFT twiddles, and work buffers for up to
 * max_fields (1..3) fields per call.  nlat even, nlon a power of two, lmax <= nlon/2.
 * precision: SLX_PREC_FP64, or SLX_PREC_FP32 for the analysis / synthesis entry points below
 * (the reference has FP64 only; the solver chooses its sweeps' precision itself). */
int slx_sht_create(int nlat, int nlon, int lmax, int precision, int max_fields, slx_sht** out);
int slx_sht_destroy(slx_sht* h);
/* dims (host, 9 ints or NULL) = {nlat, nlon, lmax, ncoef, max_fields, analysis kernel (-1 VALU,
 * 0 matrix cores one group per block, 1 persistent chunked), its rings per lane, synthesis kernel
 * (0 scalar-fed VALU, 1 matrix cores), its rings per lane}; gaulats/weights (host, nlat doubles
 * each, or NULL) = the nodes mu_j north -> south and weights (SHT.gaulats). */
int slx_sht_info(const slx_sht* h, int* dims, double* gaulats, double* weights);
/* Replaces SHT.analysis (sht.py:274-302) / DuccSHT.analysis (sht_ducc.py:134-154) for nf fields
 * at once: grid (device) nf x [nlat][nlon] -> coef (device) nf x [ncoef] packed. */
int slx_sht_analysis(slx_sht* h, const double* grid, int nf, void* coef, void* stream);
/* The second stage of slx_sht_analysis alone: the per-order Legendre contraction of the folded
 * ring spectra left in the handle's work buffers by the last analysis of nf fields (no
 * counterpart in the reference; lets a profiler time the contraction by itself). */
int slx_sht_analysis_contraction(slx_sht* h, int nf, void* coef, void* stream);
/* slx_sht_analysis that allocates nothing and uses no work buffer of the handle: the folded ring
 * spectra go to a static device array of SLX_NOALLOC_SCRATCH_BYTES per GPU (held until the
 * process ends), which limits 2 nf (lmax+1) (nlat/2) complex doubles to that size (24 MiB: up to
 * three fields at lmax 512 on 512 x 1024, one at lmax 1023 on 1024 x 2048); larger requests return -22
 * with a message.  No counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid it).  Every call on
 * one GPU shares that GPU's array, so calls that overlap in time on one GPU (other streams or
 * host threads) race.  Same results as slx_sht_analysis, bit for bit. */
#define SLX_NOALLOC_SCRATCH_BYTES (24u << 20)
int slx_sht_analysis_noalloc(slx_sht* h, const double* grid, int nf, void* coef, void* stream);
/* Replaces SHT.synthesis (sht.py:305-343) / DuccSHT.synthesis (sht_ducc.py:156-187): coef
 * (device) nf x [ncoef] -> grid (device) nf x [nlat][nlon].  Im of the m = 0 (and Nyquist)
 * coefficients is ignored, as irfft does. */
int slx_sht_synthesis(slx_sht* h, const void* coef, int nf, double* grid, void* stream);
/* Layout conversion between packed triangles and the dense [l][m] (lmax+1)^2 complex arrays
 * the reference's Python API exchanges (zeros above the diagonal). */
int slx_sht_tri_to_dense(const slx_sht* h, const void* tri, void* dense, int nf, void* stream);
int slx_sht_dense_to_tri(const slx_sht* h, const void* dense, void* tri, int nf, void* stream);

/* ---------------------------------------------------------------- modal state ------------
 * Replaces ExactMemory.advance_and_read (python/slevolve/solver.py:386-422),
 * kernels.advance_and_read (kernels.py:28-67) and Fortran state_advance_and_read
 * (modal_state_mod.f90:286-336): for every packed (l, m) and mode k < nmod[l]
 *   R_k <- decay[k][l] R_k + omd[k][l] P ;  visc = sum_k a_sl[k][l] R_k ;  rad = sum_k a_rr[k][l] R_k
 * in k order, without FMA contraction (bit-identical to the NumPy statements).
 * R (device) [kmax][ncoef] complex; P, visc, rad (device) [ncoef] complex; nmod (device, int)
 * [lmax+1]; decay, omd, a_sl, a_rr (device) [kmax][lmax+1]. */
int slx_modal_advance_read(const slx_sht* h, int kmax, const int* nmod, const double* decay,
                           const double* omd, const double* a_sl, const double* a_rr, void* R,
                           const void* P, void* visc, void* rad, void* stream);

/* ------------------------------------------------------------------- adjoint ---------------
 * Building blocks of the hand-written adjoint of the time step (slevolve_amd/adjoint.py; the
 * reference has no adjoint).  Transposes with respect to the real inner products of the stored
 * arrays: <analysis(g), U> = <g, analysis_adjoint(U)>, <synthesis(U), g> = <U, synthesis_adjoint(g)>.
 * analysis_adjoint: coef (device) nf x [ncoef] packed -> grid (device) nf x [nlat][nlon].
 * synthesis_adjoint: grid -> coef.  Both FP64, any handle precision. */
int slx_sht_analysis_adjoint(slx_sht* h, const void* coef, int nf, double* grid, void* stream);
int slx_sht_synthesis_adjoint(slx_sht* h, const double* grid, int nf, void* coef, void* stream);
/* Reverse of slx_modal_advance_read: Rbar_k <- decay (Rbar_k + a_sl visc_bar + a_rr rad_bar) and
 * Pbar += sum_k omd (Rbar_k + a_sl visc_bar + a_rr rad_bar), per packed (l, m) and mode k < nmod[l]
 * (device arrays as for the advance). */
int slx_modal_reverse(const slx_sht* h, int kmax, const int* nmod, const double* decay, const double* omd,
                      const double* a_sl, const double* a_rr, void* Rbar, const void* visc_bar,
                      const void* rad_bar, void* Pbar, void* stream);

/* ------------------------------------------------------------------- the step ------------
 * Replaces SeaLevelSolver (solver.py:521-1144) and OptSeaLevelSolver with inner = "off",
 * "extrap" or "anderson" (solver_opt.py:369-979), step or second-order ("ramp") load
 * placement, axial rotation "off".  All state is device-resident; the inner fixed point runs as predicated
 * launches that stop on a device convergence flag. */
typedef struct slx_solver_params {
  int lmax, kmax, nmod2;
  int tpw, checkmarine, calcRG, extrap; /* booleans */
  double epsilon1;
  double radius, mass, rhoi, rhow, gacc, omega, moiA, moiC, kf; /* Planet (solver.py:98-110) */
  double kTE2, hTE2, kE2;                                      /* degree-2 Love numbers */
  const int* nmod;                                  /* host [lmax+1] modes per degree */
  const double* s;                                  /* host [kmax][lmax+1] 1/kyr, 0 = no mode */
  const double* a_sl;                               /* host [kmax][lmax+1] (r' - r)/s */
  const double* a_rr;                               /* host [kmax][lmax+1] r/s */
  const double* a_il, *a_m, *a_tsl, *a_trr;         /* host [kmax] degree-2 rotational weights */
  const double* ttl, *ttlekhl, *ttlhe;              /* host [lmax+1] T_l, T_l E_l, T_l h_l */
  int ramp;          /* second-order load placement, SeaLevelSolver(ramp=True) (solver.py:535,549-551) */
  /* horizontal read-out (solver.py:424-434,687-720); a_hl NULL = not available */
  const double* a_hl;                               /* host [kmax][lmax+1] l_k/s_k */
  const double* a_thl;                              /* host [kmax] tidal l_k/s_k at degree 2 */
  const double* ttlle;                              /* host [lmax+1] T_l l^E_l */
  double lTE2;                                      /* tidal elastic l number at degree 2 */
  /* inner = "anderson" (solver_opt.py:268-293,705-775; needs extrap = 1): Anderson mixing of
   * depth 1..5 on the ocean-load increment with the ocean mask frozen after anderson_freeze plain
   * iterations, at most max_mask_refresh mask refreshes per step, history kept across a refresh
   * if anderson_keep */
  int anderson, anderson_depth, anderson_freeze, max_mask_refresh, anderson_keep;
  /* Mixed-precision sweeps (no counterpart in the reference, whose transforms are FP64): with
   * precision = SLX_PREC_FP32 the inner iterations of a step take the FP32-product contractions
   * until one follows an iteration that changed no ocean-function cell and no grounded-ice flag,
   * or whose convergence ratio xi was <= eps_switch (> epsilon1); from then on FP64 sweeps until
   * xi <= epsilon1 after an FP64 sweep.  The step start, re-guesses, radial displacement, modal
   * state, mass integrals and the convergence norm stay FP64.  Not with anderson = 1. */
  int precision;
  double eps_switch;
  /* inner iterations after which a step fails (the reference aborts at 9999, solver.py:1049-1050);
   * <= 0 or > 9999 means 9999.  A failed step is not committed. */
  int max_inner;
} slx_solver_params;

typedef struct slx_status {
  int numiter, done, failed, nacc, reguess, ice_analyses, step;
  int finished;
  double xi, conserv;
  double il[9], mm[3], lam[18]; /* lam: 3x3 complex, row-major, (re, im) */
  double ocean_area, oceanBeta_area, gmslc_ocn, gmslc_ocnBeta, ice_volume;
  double ocean_area0, oceanBeta_area0, ice_volume0;
  /* inner = "anderson" counters over the run (OptSeaLevelSolver.inner_stats) */
  int aa_mixed, aa_plain, aa_fallback, aa_singular, aa_mask_refresh, aa_frozen_iters;
  int sweeps32;  /* inner iterations of the step that took the FP32-product contractions */
  int pad_;
} slx_status;

typedef struct slx_fields { /* device pointers owned by the solver */
  double* icestar; /* grounded ice I* */
  double* tinit, *cstar0, *cxy0, *beta0;
  void* dS_converged;  /* packed [ncoef] complex */
  void* p_load, *p_ice, *p_ocean, *prev_ice, *load_inc;
  void* modal_R;       /* packed [kmax][ncoef] complex */
} slx_fields;

typedef struct slx_outputs { /* caller-owned device grids [nlat][nlon] the solver writes in place */
  double* tgrid;   /* topography tinit - deltasl after the step ("tgrid") */
  double* ocean;   /* ocean function C, 0/1 ("ocean") */
  double* beta;    /* grounded-ice complement beta, 0/1 ("beta") */
  double* R;       /* radial displacement ("R") */
  double* G;       /* deltasl + R ("G") */
} slx_outputs;

int slx_solver_create(slx_sht* h, const slx_solver_params* p, const slx_outputs* o, slx_solver** out);
int slx_solver_destroy(slx_solver* s);
/* Step-0 set-up (solver.py:554-605): tinit, C0, marine-checked I*0, beta0, C*0 and the three
 * degree-0 areas.  topo0, ice0: device grids. */
int slx_solver_init(slx_solver* s, const double* topo0, const double* ice0, void* stream);
/* Step length (years) of the following steps: decay = exp(-s dt / 1000), the ramp factors
 * 1 - phi(s dt) and their sums (solver.py:331-367). */
int slx_solver_set_dt(slx_solver* s, double dt_years);
/* Start a step with the ice grid ice_cur (device): modal advance, grounded-ice projection,
 * ice / ocean analyses and the first guess (eustatic, or extrapolated with the step lengths
 * dt0 = this step, dt1, dt2 = the two previous ones when the solver extrapolates).  With the
 * extrapolated guess the ice analysis adds the analysis of the change of I* to the previous one
 * and C* is analysed by the first inner iteration (unless the solver was created with
 * SLX_ICE_REANALYSIS=full). */
int slx_step_begin(slx_solver* s, const double* ice_cur, double dt0, double dt1, double dt2, void* stream);
/* slx_step_begin with the "warm" start across outer passes of OptSeaLevelSolver(reuse="warm")
 * (solver_opt.py:400-402,517-528): before the ice/ocean block the ocean function becomes
 * warm_ocean (device grid, 0/1: the converged ocean function this step had in the previous
 * pass), and the first guess of the ocean-load increment becomes warm_dS (device, packed
 * [ncoef] complex: the converged increment of that pass, fields.load_inc after its finish), with
 * deltaS_cur = p_ocean + warm_dS.  Both pointers NULL = slx_step_begin. */
int slx_step_begin_warm(slx_solver* s, const double* ice_cur, double dt0, double dt1, double dt2,
                        const void* warm_dS, const double* warm_ocean, void* stream);
/* Enqueue n inner iterations; those after convergence return at once on the device. */
int slx_step_iterate(slx_solver* s, int n, void* stream);
/* Radial displacement, commit, topography / ocean function and degree-0 diagnostics.  These
 * launches run only once the inner loop has converged and only once per step, so the host may
 * enqueue them right after a batch of iterations and test status.finished afterwards. */
int slx_step_finish(slx_solver* s, void* stream);
/* Optional read-outs of a step, computed on the same state by the finish (caller-owned device
 * buffers, NULL = not wanted).  Rates (solver.py:722-806, step(rates=True)): the analytic left
 * time derivatives at the end of the step with the current increment accruing linearly over it;
 * dSLdt, dRdt [nlat][nlon] m/yr, their packed [ncoef] complex spectra (dSLdt_lm degree 0 from the
 * conservation constant on the end-of-step C*), and rot_rates[30] = il-dot (3x3), m-dot (3),
 * Lambda-dot (3x3 complex, (re, im)); give all five or none.  horiz_lm: packed [ncoef] complex
 * Psi_lm of the horizontal read-out (solver.py:687-720, result key "horiz"; needs a_hl). */
typedef struct slx_readouts {
  double* dSLdt, *dRdt;
  void* dSLdt_lm, *dRdt_lm;
  double* rot_rates;
  void* horiz_lm;
} slx_readouts;
/* slx_step_finish with read-outs (r NULL = slx_step_finish). */
int slx_step_finish_readouts(slx_solver* s, const slx_readouts* r, void* stream);
/* Synchronous: copies the status block to *out (host) after the stream's pending work. */
int slx_solver_status(slx_solver* s, slx_status* out, void* stream);
/* The step loop of run_ais.py:389-425 (SeaLevelSolver.step called per ice file) in native code:
 * steps k = 0..nsteps-1 with ice grid ice + k * ice_stride (device) and length dts[k] (host),
 * dt_prev = the lengths of the two steps before the first (host); *guess is set to the last
 * step's inner-iteration count; out[k] (host) = the status of step k after its finish.  Each inner
 * iteration is enqueued when the device has posted the decision of the previous one (to pinned host
 * memory, as the deciding kernel starts), so no iteration runs past convergence and the host never
 * waits for the stream to drain before the last step.  Outputs of intermediate steps are
 * overwritten; the solver's grids hold the last step's.  Returns -34 when step k's inner loop fails
 * to converge: out[0..k-1] hold the finished steps, out[k] the failed step's status (failed = 1);
 * step k is not committed, as the reference raises before its commit. */
int slx_run_steps(slx_solver* s, const double* ice, long ice_stride, int nsteps, const double* dts,
                  const double* dt_prev, int* guess, slx_status* out, void* stream);
int slx_solver_fields(slx_solver* s, slx_fields* out);
/* Stage timing of profiled steps, the per-stage keys of SeaLevelSolver.step(profile=True)
 * (python/slevolve/solver.py:829-831,954-956; solver_opt.py:382-384).  With on = 1, the launches
 * of slx_step_begin / _iterate / _finish on this solver's stream are followed by timing events
 * ("memory": modal advance and commit; "sht": ring passes and contractions, with the synthesis's
 * B18 coefficients, which it forms while staging them; "spec": the remaining coefficient fills,
 * guess and load update; "grid": ocean-function refresh, re-guess ring pass, post-step grids;
 * "rot": degree-2 states and the rotation vector; profiled steps launch the load update and the
 * refresh separately, which unprofiled steps with the marine check fuse, with the same results).  slx_solver_stage_times waits for the last event
 * and returns the device milliseconds per stage, ms[5] = memory, sht, spec, grid, rot, since the
 * previous call (or since profiling was switched on).  Intended for step-by-step use; the native
 * loop (slx_run_steps) is not profiled.  One solver per transform handle may profile at a time. */
int slx_solver_profile(slx_solver* s, int on);
int slx_solver_stage_times(slx_solver* s, double* ms);

#ifdef __cplusplus
}
#endif
#endif
