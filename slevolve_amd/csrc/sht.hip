// Gauss-Legendre transform handle and its C-ABI (declared in include/slx.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "sht_launch.h"
#include "slx.h"

namespace {
thread_local std::string g_err;
}

void slx_set_error(const std::string& msg) { g_err = msg; }
extern "C" const char* slx_last_error(void) { return g_err.c_str(); }
extern "C" int slx_abi_version(void) { return SLX_ABI_VERSION; }


using namespace slx;

// ------------------------------------------------------------------------------------ tables
// Gauss-Legendre nodes/weights: Newton on P_nlat from cos(pi (j - 1/4)/(nlat + 1/2)), at most 10
// sweeps, |dz| < 1e-15, node and derivative kept at first convergence (python/slevolve/sht.py:65-101).
static int gauss_legendre(int nlat, std::vector<double>& z_out, std::vector<double>& w) {
  const double pi = 4.0 * std::atan(1.0);
  z_out.assign(nlat, 0.0);
  w.assign(nlat, 0.0);
  for (int i = 0; i < nlat; ++i) {
    double z = std::cos(pi * ((double)(i + 1) - 0.25) / ((double)nlat + 0.5));
    bool done = false;
    for (int it = 0; it < 10 && !done; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int n = 1; n <= nlat; ++n) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * n - 1.0) * z * p2 - (n - 1.0) * p3) / n;
      }
      double pp = nlat * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-15) {
        done = true;
        z_out[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
      }
    }
    if (!done) return -1;
  }
  return 0;
}

// Walk layout for chain length T (see slx_walk): tables on the host, start states, reload rows,
// bands and reload masks from the device seeding walk.
static int build_walk(slx_sht* h, int T, slx_walk& w) {
  const int nm = h->nm, nh = h->nh, L = h->lmax;
#define ALLOC_COPY(ptr, vec)                                                                   \
  SLX_HIP(hipMalloc(&ptr, vec.size() * sizeof(vec[0])));                                       \
  SLX_HIP(hipMemcpy(ptr, vec.data(), vec.size() * sizeof(vec[0]), hipMemcpyHostToDevice));
  w.T = T;
  std::vector<int4> ord(nm), ua, ub;
  std::vector<double> tqc, tqg;
  std::vector<int> u0(nm, 0);
  int roff = 0, npts = 0, qsmax = 0;
  for (int m = 0; m < nm; ++m) {
    const int n = L - m + 1;
    const int W = std::max(1, (n + 4 * T - 1) / (4 * T));
    int Q = (n + 4 * W - 1) / (4 * W);
    Q += Q & 1;
    const int Qp = 8 * ((Q + 7) / 8), nk = (Q + 31) / 32;
    SLX_CHECK(nk - 1 <= 8, "walk contraction: a chain longer than 288 degrees");
    ord[m] = make_int4(Q, 4 * W, nk, npts);
    npts += 4 * W * nk;
    u0[m] = (int)ua.size();
    qsmax = std::max(qsmax, Qp + 2);
    for (int wk = 0; wk < W; ++wk) {
      ua.push_back(make_int4(m, Q, Qp, (int)tqc.size()));
      ub.push_back(make_int4(nk, roff, 0, 4 * wk));
      roff += 4 * (nk - 1);
      for (int i = 0; i < 4; ++i) {
        const int c = 4 * wk + i;
        long double g = 1.0L;
        for (int t = 0; t < Qp; ++t) {
          const int l = m + c * Q + t;
          double cq = 0.0, gg = 0.0;
          if (t < Q && l <= L) {
            if (c == 0 && t < 2) {
              gg = 1.0;  // head: P~_m, P~_{m+1} unscaled
            } else {
              const int rs = (c == 0 && t < 32) ? 2 : (t & ~31);
              if (t == rs) g = 1.0L;
              const long double lm1 = l - 1, mm = m, ll = l;
              const long double ep = std::sqrt((lm1 * lm1 - mm * mm) / (4.0L * lm1 * lm1 - 1.0L));
              const long double e = std::sqrt((ll * ll - mm * mm) / (4.0L * ll * ll - 1.0L));
              cq = (double)((t == rs) ? ep : ep * ep);
              g /= e;
              gg = (double)g;
            }
          }
          tqc.push_back(cq);
          tqg.push_back(gg);
        }
      }
    }
  }
  const int nunit = (int)ua.size();
  w.nunit = nunit;
  w.qsmax = qsmax;
  int *d_u0 = nullptr, *d_umask = nullptr;
  int4* d_ord = nullptr;
  ALLOC_COPY(w.ua, ua);
  ALLOC_COPY(w.ub, ub);
  ALLOC_COPY(w.qc, tqc);
  ALLOC_COPY(w.qg, tqg);
  ALLOC_COPY(d_ord, ord);
  ALLOC_COPY(d_u0, u0);
  SLX_HIP(hipMalloc(&w.seed, (size_t)nunit * 4 * nh * sizeof(double2)));
  SLX_HIP(hipMalloc(&w.rl, (size_t)std::max(roff, 1) * nh * sizeof(double2)));
  SLX_HIP(hipMalloc(&d_umask, (size_t)nunit * sizeof(int)));
  SLX_HIP(hipMemset(d_umask, 0, (size_t)nunit * sizeof(int)));
  {
    std::vector<int2> band((size_t)nunit * 32, make_int2(nh, 0));  // empty bands [nh, 0)
    ALLOC_COPY(w.band, band);
  }
  double2* pab = nullptr;
  int* psc = nullptr;
  SLX_HIP(hipMalloc(&pab, (size_t)npts * nh * sizeof(double2)));
  SLX_HIP(hipMalloc(&psc, (size_t)npts * nh * sizeof(int)));
  dim3 grid(nm, (nh + 63) / 64);
  const QwOrd qo{d_ord, d_u0};
  k_qw_seed_walk<<<grid, 64>>>(dev_of(h), qo, pab, psc);
  SLX_LAUNCH_CHECK();
  k_qw_seed_unscale<<<grid, 64>>>(dev_of(h), qw_of(w, nullptr), qo, pab, psc, w.seed, w.rl, d_umask, w.band);
  SLX_LAUNCH_CHECK();
  SLX_HIP(hipDeviceSynchronize());
  std::vector<int> umask(nunit);
  SLX_HIP(hipMemcpy(umask.data(), d_umask, (size_t)nunit * sizeof(int), hipMemcpyDeviceToHost));
  for (int u = 0; u < nunit; ++u) ub[u].z = umask[u];
  SLX_HIP(hipMemcpy(w.ub, ub.data(), ub.size() * sizeof(int4), hipMemcpyHostToDevice));
  w.h_ua = ua;
  SLX_HIP(hipFree(pab));
  SLX_HIP(hipFree(psc));
  SLX_HIP(hipFree(d_ord));
  SLX_HIP(hipFree(d_u0));
  SLX_HIP(hipFree(d_umask));
#undef ALLOC_COPY
  return 0;
}

static void free_walk(slx_walk* w) {
  if (!w) return;
  void* ptrs[] = {w->ua, w->ub, w->qc, w->qg, w->seed, w->rl, w->band};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete w;
}

extern "C" int slx_sht_create(int nlat, int nlon, int lmax, int precision, int max_fields, slx_sht** out) {
  SLX_CHECK(out != nullptr, "slx_sht_create: null handle pointer");
  SLX_CHECK(precision == SLX_PREC_FP64 || precision == SLX_PREC_FP32, "precision must be 0 (FP64) or 1 (FP32 products)");
  SLX_CHECK(nlat >= 2 && nlat % 2 == 0, "nlat must be even");
  SLX_CHECK(nlon >= 8 && (nlon & (nlon - 1)) == 0, "nlon must be a power of two >= 8");
  SLX_CHECK(nlon <= 4096, "nlon must be <= 4096 (a ring pass holds two rows and a work buffer in LDS)");
  SLX_CHECK(lmax >= 0 && lmax <= nlon / 2, "lmax must be <= nlon/2");
  SLX_CHECK(lmax <= nlat * 2, "lmax too large for the grid");
  SLX_CHECK(max_fields >= 1 && max_fields <= 3, "max_fields must be 1..3");
  SLX_CHECK(nlat / 2 <= 512, "nlat must be <= 1024");
  auto* h = new slx_sht();
  h->nlat = nlat;
  h->nlon = nlon;
  h->lmax = lmax;
  h->nh = nlat / 2;
  h->nm = lmax + 1;
  h->ncoef = (lmax + 1) * (lmax + 2) / 2;
  h->max_fields = max_fields;
  h->precision = precision;
  int lg = 0;
  while ((1 << lg) < nlon) ++lg;
  h->log2n = lg;
  if (gauss_legendre(nlat, h->h_x, h->h_w) != 0) {
    delete h;
    slx_set_error("Gauss-Legendre nodes did not converge");
    return -22;
  }
  const int nh = h->nh, nm = h->nm, L = lmax;
  std::vector<double> x(nh), wn(nh), ws(nh), pmm((size_t)nm * nh), s2m3(nm);
  std::vector<int> pms((size_t)nm * nh);
  for (int j = 0; j < nh; ++j) {
    x[j] = h->h_x[j];
    wn[j] = h->h_w[j] / nlon;
    ws[j] = h->h_w[nlat - 1 - j] / nlon;
  }
  // sectoral start P~_mm = sqrt(0.5 prod_m), prod_m = prod (1 - x^2) A/B (sht.py:118-126), in
  // extended range: prod = mantissa 2^E with exact power-of-two renormalisation, and a start value
  // below 2^-SLX_SCALE_T stored as P~ 2^(SLX_SCALE_K s) with its scale count s (pms).  Identical to
  // the reference wherever its prod is a normal number.
  for (int j = 0; j < nh; ++j) {
    double sinsq = 1.0 - x[j] * x[j], A = 1.0, B = 0.0, prod = 1.0;
    long E = 0;
    for (int m = 0; m < nm; ++m) {
      if (m != 0) {
        A += 2.0;
        B += 2.0;
        prod = prod * sinsq * A / B;
        if (prod < 0x1p-500) {
          prod *= 0x1p500;
          E -= 500;
        }
      }
      const double base = std::sqrt(0.5 * prod);
      const long e2 = E / 2;
      int ex = 0;
      std::frexp(base, &ex);
      const long lg = (long)(ex - 1) + e2;
      long sc = 0;
      if (-SLX_SCALE_T - lg > 0) sc = (-SLX_SCALE_T - lg + SLX_SCALE_K - 1) / SLX_SCALE_K;
      pmm[(size_t)m * nh + j] = std::ldexp(base, (int)(e2 + SLX_SCALE_K * sc));
      pms[(size_t)m * nh + j] = (int)sc;
    }
  }
  h->h_moff.resize(nm + 1);
  for (int m = 0; m <= nm; ++m) h->h_moff[m] = slx_moff(m, L);
  std::vector<double2> rec(h->ncoef);
  std::vector<int> l_of(h->ncoef), m_of(h->ncoef), jb(nm, 0);
  for (int m = 0; m < nm; ++m) {
    s2m3[m] = std::sqrt((double)(2 * m + 3));
    for (int l = m; l <= L; ++l) {
      const int idx = h->h_moff[m] + (l - m);
      l_of[idx] = l;
      m_of[idx] = m;
      if (l >= m + 2) {
        double e = std::sqrt((double)(l * l - m * m) / (double)(4 * l * l - 1));
        double ep = std::sqrt((double)((l - 1) * (l - 1) - m * m) / (double)(4 * (l - 1) * (l - 1) - 1));
        rec[idx] = make_double2(ep, 1.0 / e);
      } else {
        rec[idx] = make_double2(0.0, 0.0);
      }
    }
  }
  std::vector<double> qc(h->ncoef, 0.0), qg(h->ncoef, 1.0);
  for (int m = 0; m < nm; ++m) {
    long double g = 1.0L;
    for (int l = m + 2; l <= L; ++l) {
      const int idx = h->h_moff[m] + (l - m);
      const int ls = std::max(m + ((l - m) / MF_S) * MF_S, m + 2);
      if (l == ls) g = 1.0L;
      const long double lm1 = l - 1, mm = m, ll = l;
      const long double ep = std::sqrt((lm1 * lm1 - mm * mm) / (4.0L * lm1 * lm1 - 1.0L));
      const long double e = std::sqrt((ll * ll - mm * mm) / (4.0L * ll * ll - 1.0L));
      qc[idx] = (double)((l == ls) ? ep : ep * ep);
      g /= e;
      qg[idx] = (double)g;
    }
  }
  // FFT twiddles: the quarter period exp(-2 pi i t / nlon), t < nlon / 4, from cos/sin; the other
  // quadrants by exact rotations, (-i)^q; laid out per radix-4 stage (Ns > 1, in stage order) as
  // W1[k], W2[k], W3[k] = exp(-2 pi i r k nlon / (4 Ns) / nlon), r = 1, 2, 3, k < Ns
  std::vector<double2> tw4(nlon / 4), stw;
  const double pi2 = 8.0 * std::atan(1.0);
  for (int t = 0; t < nlon / 4; ++t) {
    double a = -pi2 * (double)t / (double)nlon;
    tw4[t] = make_double2(std::cos(a), std::sin(a));
  }
  auto twq = [&](int u) {
    const int q = u / (nlon / 4), r = u % (nlon / 4);
    const double2 w = tw4[r];
    return q == 0 ? w : q == 1 ? make_double2(w.y, -w.x) : q == 2 ? make_double2(-w.x, -w.y) : make_double2(-w.y, w.x);
  };
  for (int Ns = (lg & 1) ? 2 : 1; Ns < nlon; Ns *= 4) {
    if (Ns == 1) continue;
    for (int r = 1; r <= 3; ++r)
      for (int k = 0; k < Ns; ++k) stw.push_back(twq(r * k * (nlon / (4 * Ns))));
  }
  size_t wsz = (size_t)max_fields * nm * nh * sizeof(double2);
#define ALLOC_COPY(ptr, vec)                                                                   \
  SLX_HIP(hipMalloc(&ptr, vec.size() * sizeof(vec[0])));                                       \
  SLX_HIP(hipMemcpy(ptr, vec.data(), vec.size() * sizeof(vec[0]), hipMemcpyHostToDevice));
  ALLOC_COPY(h->x, x);
  ALLOC_COPY(h->wn, wn);
  ALLOC_COPY(h->ws, ws);
  ALLOC_COPY(h->pmm, pmm);
  ALLOC_COPY(h->pms, pms);
  ALLOC_COPY(h->rec, rec);
  ALLOC_COPY(h->s2m3, s2m3);
  ALLOC_COPY(h->moff, h->h_moff);
  ALLOC_COPY(h->l_of, l_of);
  ALLOC_COPY(h->m_of, m_of);
  ALLOC_COPY(h->qc, qc);
  ALLOC_COPY(h->qg, qg);
  ALLOC_COPY(h->jb, jb);
  ALLOC_COPY(h->stw, stw);
  // degree segments.  With nlat <= 1024 the analysis runs on the matrix cores with 32-degree
  // segments in groups of 4 per block; otherwise one block per (order, segment) with S (a multiple
  // of 8) the largest of 256, 128, 64, 32 that still gives >= 2048 waves.  Synthesis chains are
  // syn_seg = a multiple of S degrees long, about 8 chains per order pair.
  auto count = [&](int S) {
    long n = 0;
    for (int m = 0; m < nm; ++m) n += (L - m + S) / S;
    return n;
  };
  int S;
  h->ana_mfma = (nh <= 512) ? 1 : 0;
  if (h->ana_mfma) {
    S = MF_S;
  } else {
    h->ana_nw = (nh > 256) ? 2 : 1;
    S = 256;
    while (S > 32 && count(S) * h->ana_nw < 2048) S /= 2;
  }
  h->seg_len = S;
  h->syn_mfma = h->ana_mfma;
  h->syn_seg = S * std::max(1, (int)std::lround((L + 2) / (8.0 * S)));
  std::vector<int> seg_m, seg_l0, first(nm);
  for (int m = 0; m < nm; ++m) {
    first[m] = (int)seg_m.size();
    for (int l0 = m; l0 <= L; l0 += S) {
      seg_m.push_back(m);
      seg_l0.push_back(l0);
    }
  }
  h->nseg = (int)seg_m.size();
  ALLOC_COPY(h->seg_m, seg_m);
  ALLOC_COPY(h->seg_l0, seg_l0);
  ALLOC_COPY(h->seg_first, first);
  SLX_HIP(hipMalloc(&h->seed_a, (size_t)h->nseg * nh * sizeof(double)));
  SLX_HIP(hipMalloc(&h->seed_b, (size_t)h->nseg * nh * sizeof(double)));
  SLX_HIP(hipMalloc(&h->seed_s, (size_t)h->nseg * nh * sizeof(int)));
  SLX_HIP(hipMemset(h->seed_a, 0, (size_t)h->nseg * nh * sizeof(double)));
  SLX_HIP(hipMemset(h->seed_b, 0, (size_t)h->nseg * nh * sizeof(double)));
  SLX_HIP(hipMemset(h->seed_s, 0, (size_t)h->nseg * nh * sizeof(int)));
  {
    dim3 grid(nm, (nh + 63) / 64);
    k_leg_seed<<<grid, 64>>>(dev_of(h), seg_of(h), h->seed_a, h->seed_b, h->seed_s);
    SLX_LAUNCH_CHECK();
    SLX_HIP(hipDeviceSynchronize());
  }
  if (h->ana_mfma) {
    const long n = (long)h->nseg * nh;
    SLX_HIP(hipMalloc(&h->seed_ab, (size_t)n * sizeof(double2)));
    k_seed_unscale<<<(unsigned)((n + 255) / 256), 256>>>(dev_of(h), seg_of(h), h->seed_ab);
    SLX_LAUNCH_CHECK();
    SLX_HIP(hipDeviceSynchronize());
  }
  // matrix-core analysis groups: 4 consecutive segments of one order
  {
    std::vector<int> gm, gs;
    for (int m = 0; m < nm; ++m)
      for (int s0 = 0; s0 < (L - m + S) / S; s0 += 4) {
        gm.push_back(m);
        gs.push_back(s0);
      }
    h->ngrp = (int)gm.size();
    ALLOC_COPY(h->grp_m, gm);
    ALLOC_COPY(h->grp_s0, gs);
    std::vector<int4> gd(gm.size());
    for (size_t i = 0; i < gm.size(); ++i) gd[i] = make_int4(gm[i], gs[i], first[gm[i]] + gs[i], h->h_moff[gm[i]]);
    ALLOC_COPY(h->grp_desc, gd);
  }
  // walk analysis: layout with T the longest of (unbounded, 256, 128, 96, 64, 48, 32) that still
  // gives >= 768 units (3 per CU: longer chains load fewer start states, more units keep more waves
  // resident), every unit a block of its own (units are dealt longest first to the least loaded
  // block, at most 3 per block, when SLX_QW_BLOCKS asks for fewer blocks; SLX_QW_T sets T).
  if (h->ana_mfma) {
    auto units_for = [&](int T) {
      long n_u = 0;
      for (int m = 0; m < nm; ++m) n_u += std::max(1, (L - m + 1 + 4 * T - 1) / (4 * T));
      return n_u;
    };
    int T = 32;
    for (int cand : {1 << 20, 256, 128, 96, 64, 48}) {
      if (units_for(cand) >= 768) {
        T = cand;
        break;
      }
    }
    if (const char* e = std::getenv("SLX_QW_T")) T = std::max(8, std::atoi(e));
    h->wa = new slx_walk();
    if (int rc = build_walk(h, T, *h->wa)) return rc;
    const int nunit = h->wa->nunit;
    int nblk = nunit;
    if (const char* e = std::getenv("SLX_QW_BLOCKS")) nblk = std::atoi(e);
    nblk = std::max(nblk, (nunit + 2) / 3);
    {
      const std::vector<int4>& ua = h->wa->h_ua;
      std::vector<int> idx(nunit);
      for (int u = 0; u < nunit; ++u) idx[u] = u;
      std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ua[a].z > ua[b].z; });
      std::vector<long> load(nblk, 0);
      std::vector<int4> blk(nblk, make_int4(-1, -1, -1, -1));
      std::vector<int> cnt(nblk, 0);
      for (int u : idx) {
        int best = -1;
        for (int b = 0; b < nblk; ++b)
          if (cnt[b] < 3 && (best < 0 || load[b] < load[best])) best = b;
        int* slotp = cnt[best] == 0 ? &blk[best].x : cnt[best] == 1 ? &blk[best].y : &blk[best].z;
        *slotp = u;
        ++cnt[best];
        load[best] += ua[u].z;
      }
      int used = 0;  // drop empty blocks
      for (int b = 0; b < nblk; ++b)
        if (cnt[b] > 0) blk[used++] = blk[b];
      blk.resize(used);
      h->qw_nblk = used;
      ALLOC_COPY(h->qw_blk, blk);
    }
  }
#undef ALLOC_COPY
  SLX_HIP(hipMalloc(&h->WS, wsz));
  SLX_HIP(hipMalloc(&h->WD, wsz));
  SLX_HIP(hipMalloc(&h->FN, wsz));
  SLX_HIP(hipMalloc(&h->FSo, wsz));
  SLX_HIP(hipMalloc(&h->CS, (size_t)max_fields * h->ncoef * sizeof(double2)));
  SLX_HIP(hipMemset(h->WS, 0, wsz));
  SLX_HIP(hipMemset(h->WD, 0, wsz));
  // choice of the analysis contraction.  By default a fixed rule per shape, so that every handle
  // of a shape sums in the same order and results repeat bit for bit across handles and runs:
  // the walk kernel with 4 rings per lane from 256 ring pairs on (both benchmark grids), the
  // chunked-accumulator persistent kernel with 2 rings per lane below (matrix-core path); 1 wave
  // per segment where it fits, else 2 (VALU path).  SLX_ANA_VARIANT = mf2 | mf4 | ck2 | ck4 | qw2 |
  // qw4 fixes another variant;
  // SLX_ANA_VARIANT = auto times the candidates (2 fields, or 1 for a 1-field handle) on the
  // device with the handle's own shapes and keeps the fastest.
  {
    double2* tmp = nullptr;
    SLX_HIP(hipMalloc(&tmp, 2 * (size_t)h->ncoef * sizeof(double2)));
    hipEvent_t e0, e1;
    SLX_HIP(hipEventCreate(&e0));
    SLX_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int best_v = -1;
    const char* av = std::getenv("SLX_ANA_VARIANT");
    std::string awant = av ? av : "";
    if (awant.empty()) awant = h->ana_mfma ? (nh >= 256 ? "qw4" : "ck2") : "";
    const bool timed = awant == "auto";
    // v = 1, 2: one group per block; 3, 4: persistent chunked; 5, 6: quarter walk (odd v: 2 rings
    // per lane, even: 4)
    const int vmax = h->ana_mfma ? 6 : 2;
    static const char* const kind[] = {"mf", "ck", "qw"};
    for (int v = 1; v <= vmax; ++v) {
      if (h->ana_mfma) {
        const int r = (v & 1) ? 2 : 4, kv = (v - 1) / 2;
        if ((nh + 16 * r - 1) / (16 * r) > 8) continue;
        if (!timed && awant != std::string(kind[kv]) + std::to_string(r)) continue;
        if (timed && kv == 1 && r == 4) continue;  // spills at its occupancy: only when asked for
        h->ana_r = r;
        h->ana_ck = kv == 1 ? 1 : 0;
        h->ana_qw = kv == 2 ? 1 : 0;
      } else {
        if ((nh + 64 * v - 1) / (64 * v) > 4) continue;
        if (!timed && best_v > 0) continue;  // the first that fits
        h->ana_nw = v;
      }
      // as many fields as the work buffers hold (a 1-field handle has no second field to read)
      auto tune = [&]() {
        return h->max_fields >= 2 ? launch_leg_ana<2>(h, h->WS, h->WD, OutPtrs<2>{{tmp, tmp + h->ncoef}}, Pred{}, nullptr)
                                  : launch_leg_ana<1>(h, h->WS, h->WD, OutPtrs<1>{{tmp}}, Pred{}, nullptr);
      };
      int rc = tune();
      if (rc) return rc;
      SLX_HIP(hipEventRecord(e0, nullptr));
      for (int rep = 0; rep < 5; ++rep) {
        rc = tune();
        if (rc) return rc;
      }
      SLX_HIP(hipEventRecord(e1, nullptr));
      SLX_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      SLX_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_v = v;
      }
    }
    if (best_v > 0) {
      if (h->ana_mfma) {
        h->ana_r = (best_v & 1) ? 2 : 4;
        h->ana_ck = (best_v - 1) / 2 == 1 ? 1 : 0;
        h->ana_qw = (best_v - 1) / 2 == 2 ? 1 : 0;
      } else {
        h->ana_nw = best_v;
      }
    } else if (h->ana_mfma) {  // the requested variant does not fit this handle: 4 rings per lane
      h->ana_r = 4;
      h->ana_ck = 0;
      h->ana_qw = 0;
    }
    // synthesis kernel: scalar-fed VALU, or matrix cores with 2 or 4 rings per lane (the latter
    // only with the matrix-core segments).  By default a fixed rule: matrix cores with 2 rings per
    // lane, the waves of a block over rings from 256 ring pairs on (k_leg_syn_mw: every wave walks
    // all of the pair's segment groups, so 5-10 groups per pair leave no wave idle; with the ring
    // pass 26.4 -> 24.7 us at lmax 512, 111 -> 89-93 us at lmax 1023), in blocks of 8 waves from
    // 512 ring pairs on (half the blocks stage each order pair's coefficients: 573.6 -> 582.7
    // steps/s at lmax 1023; 1 818 -> 1 768 at lmax 512, kept at 4 waves there), and over the pair's
    // segment groups below (its 16 KB of partial sums leave the packed staging area the LDS limit,
    // 5 blocks per CU); scalar without the matrix-core segments;
    // SLX_SYN_VARIANT = scalar | mfma2 | mfma4 | mw1 | mw2 | mw1x8 | mw2x8 (matrix cores with the
    // waves over rings, 4 or 8 waves per block) fixes the choice, mfma times the two segmented ring
    // counts and auto all seven.
    {
      const char* v = std::getenv("SLX_SYN_VARIANT");
      const std::string want = v ? v : "";
      struct Cand { int mf, r, nw = 4; };
      std::vector<Cand> cand;
      if (want == "scalar" || !h->ana_mfma) cand = {{0, 4}};
      else if (want == "mfma2") cand = {{1, 2}};
      else if (want == "mfma4") cand = {{1, 4}};
      else if (want == "mw1") cand = {{2, 1}};
      else if (want == "mw2") cand = {{2, 2}};
      else if (want == "mw1x8") cand = {{2, 1, 8}};
      else if (want == "mw2x8") cand = {{2, 2, 8}};
      else if (want == "mfma") cand = {{1, 2}, {1, 4}};
      else if (want == "auto") cand = {{0, 4}, {1, 2}, {1, 4}, {2, 1}, {2, 2}, {2, 1, 8}, {2, 2, 8}};
      else cand = {{nh >= 256 ? 2 : 1, 2, nh >= 512 ? 8 : 4}};  // the shape rule
      float bs = 1e30f;
      Cand bc = cand[0];
      for (const Cand& c : cand) {
        if (cand.size() == 1) break;
        h->syn_mfma = c.mf;
        h->syn_r = c.r;
        h->syn_nw = c.nw;
        int rc = launch_leg_syn<1>(h, CoefPlain{tmp, 0}, Pred{}, nullptr);
        if (rc) return rc;
        SLX_HIP(hipEventRecord(e0, nullptr));
        for (int rep = 0; rep < 3; ++rep) {
          rc = launch_leg_syn<1>(h, CoefPlain{tmp, 0}, Pred{}, nullptr);
          if (rc) return rc;
        }
        SLX_HIP(hipEventRecord(e1, nullptr));
        SLX_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SLX_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < bs) {
          bs = ms;
          bc = c;
        }
      }
      h->syn_mfma = bc.mf;
      h->syn_r = bc.r;
      h->syn_nw = bc.nw;
    }
    SLX_HIP(hipEventDestroy(e0));
    SLX_HIP(hipEventDestroy(e1));
    SLX_HIP(hipFree(tmp));
  }
  // tables, seeds and work buffers were built on the null stream, which does not order itself with
  // non-blocking streams: they are complete before any stream may use the handle
  SLX_HIP(hipStreamSynchronize(nullptr));
  *out = h;
  return 0;
}

extern "C" int slx_sht_destroy(slx_sht* h) {
  if (!h) return 0;
  void* ptrs[] = {h->x, h->wn, h->ws, h->pmm, h->pms, h->rec, h->s2m3, h->moff, h->l_of, h->jb, h->stw,
                  h->WS, h->WD, h->FN, h->FSo, h->seg_m, h->seg_l0, h->seg_first, h->seed_a, h->seed_b,
                  h->seed_s, h->seed_ab, h->m_of, h->grp_m, h->grp_s0, h->CS, h->qc, h->qg, h->grp_desc,
                  h->qw_blk};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  free_walk(h->wa);
  delete h;
  return 0;
}

extern "C" int slx_sht_info(const slx_sht* h, int* dims, double* gaulats, double* weights) {
  SLX_CHECK(h != nullptr, "null handle");
  if (dims) {
    dims[0] = h->nlat;
    dims[1] = h->nlon;
    dims[2] = h->lmax;
    dims[3] = h->ncoef;
    dims[4] = h->max_fields;
    // kernels chosen at creation: analysis (0 one group per block, 1 persistent chunked, 2 quarter
    // walk; rings per lane), synthesis (0 scalar-fed VALU, 1 matrix cores with the waves over
    // segment groups, 2 matrix cores with the waves over rings; rings per lane)
    dims[5] = h->ana_mfma ? (h->ana_qw ? 2 : h->ana_ck) : -1;
    dims[6] = h->ana_mfma ? h->ana_r : 0;
    dims[7] = h->syn_mfma;
    dims[8] = h->syn_r;
  }
  if (gaulats) std::memcpy(gaulats, h->h_x.data(), h->nlat * sizeof(double));
  if (weights) std::memcpy(weights, h->h_w.data(), h->nlat * sizeof(double));
  return 0;
}


extern "C" int slx_sht_analysis_contraction(slx_sht* h, int nf, void* coef_v, void* stream) {
  double2* coef = static_cast<double2*>(coef_v);
  SLX_CHECK(h && coef, "slx_sht_analysis_contraction: null argument");
  SLX_CHECK(nf >= 1 && nf <= h->max_fields, "slx_sht_analysis_contraction: nf out of range");
  hipStream_t s = (hipStream_t)stream;
  if (h->precision == SLX_PREC_FP32) {
    switch (nf) {
      case 1: return launch_leg_ana32<1>(h, h->WS, h->WD, OutPtrs<1>{{coef}}, Pred{}, s);
      case 2: return launch_leg_ana32<2>(h, h->WS, h->WD, OutPtrs<2>{{coef, coef + h->ncoef}}, Pred{}, s);
      default:
        return launch_leg_ana32<3>(h, h->WS, h->WD, OutPtrs<3>{{coef, coef + h->ncoef, coef + 2 * (long)h->ncoef}},
                                   Pred{}, s);
    }
  }
  switch (nf) {
    case 1: return launch_leg_ana<1>(h, h->WS, h->WD, OutPtrs<1>{{coef}}, Pred{}, s);
    case 2: return launch_leg_ana<2>(h, h->WS, h->WD, OutPtrs<2>{{coef, coef + h->ncoef}}, Pred{}, s);
    default:
      return launch_leg_ana<3>(h, h->WS, h->WD, OutPtrs<3>{{coef, coef + h->ncoef, coef + 2 * (long)h->ncoef}},
                               Pred{}, s);
  }
}

extern "C" int slx_sht_analysis(slx_sht* h, const double* grid, int nf, void* coef, void* stream) {
  SLX_CHECK(h && grid && coef, "slx_sht_analysis: null argument");
  SLX_CHECK(nf >= 1 && nf <= h->max_fields, "slx_sht_analysis: nf out of range");
  const long gstride = (long)h->nlat * h->nlon;
  int rc = launch_fft_ana(h, LoadPlain{grid, gstride}, nf, Pred{}, (hipStream_t)stream);
  if (rc) return rc;
  return slx_sht_analysis_contraction(h, nf, coef, stream);
}

// ------------------------------------------------------------------ allocation-free analysis
// The folded-spectra work area of slx_sht_analysis_noalloc lives in this file-scope device array
// (one copy per GPU, held until the process ends).  Host code never passes `scratch` itself to a
// kernel: its address on each device comes from hipGetSymbolAddress, taken once per device with
// that device current and kept in scratch_ptr.
static __device__ __align__(256) double2 scratch[SLX_NOALLOC_SCRATCH_BYTES / sizeof(double2)];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static int scratch_of_current_device(double2** out) {
  int dev = 0;
  SLX_HIP(hipGetDevice(&dev));
  SLX_CHECK(dev >= 0 && dev < kMaxDevices, "slx_sht_analysis_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    SLX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    SLX_CHECK(ptr != nullptr, "slx_sht_analysis_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<double2*>(g_scratch_ptr[dev]);
  return 0;
}

extern "C" int slx_sht_analysis_noalloc(slx_sht* h, const double* grid, int nf, void* coef_v, void* stream) {
  double2* coef = static_cast<double2*>(coef_v);
  SLX_CHECK(h && grid && coef, "slx_sht_analysis_noalloc: null argument");
  SLX_CHECK(nf >= 1 && nf <= h->max_fields, "slx_sht_analysis_noalloc: nf out of range");
  const size_t need = 2 * (size_t)nf * h->nm * h->nh * sizeof(double2);
  if (need > SLX_NOALLOC_SCRATCH_BYTES) {
    slx_set_error("slx_sht_analysis_noalloc: " + std::to_string(nf) + " field(s) at lmax " + std::to_string(h->lmax) +
                  ", nlat " + std::to_string(h->nlat) + " need " + std::to_string(need) +
                  " bytes of work area, more than the " + std::to_string((size_t)SLX_NOALLOC_SCRATCH_BYTES) +
                  " of the static scratch");
    return -22;
  }
  double2* ws = nullptr;
  if (int rc = scratch_of_current_device(&ws)) return rc;
  double2* wd = ws + (size_t)nf * h->nm * h->nh;
  hipStream_t s = (hipStream_t)stream;
  int rc = launch_fft_ana(h, LoadPlain{grid, (long)h->nlat * h->nlon}, nf, Pred{}, s, ws, wd);
  if (rc) return rc;
  const long n = h->ncoef;
  if (h->precision == SLX_PREC_FP32) {
    switch (nf) {
      case 1: return launch_leg_ana32<1>(h, ws, wd, OutPtrs<1>{{coef}}, Pred{}, s);
      case 2: return launch_leg_ana32<2>(h, ws, wd, OutPtrs<2>{{coef, coef + n}}, Pred{}, s);
      default: return launch_leg_ana32<3>(h, ws, wd, OutPtrs<3>{{coef, coef + n, coef + 2 * n}}, Pred{}, s);
    }
  }
  switch (nf) {
    case 1: return launch_leg_ana<1>(h, ws, wd, OutPtrs<1>{{coef}}, Pred{}, s);
    case 2: return launch_leg_ana<2>(h, ws, wd, OutPtrs<2>{{coef, coef + n}}, Pred{}, s);
    default: return launch_leg_ana<3>(h, ws, wd, OutPtrs<3>{{coef, coef + n, coef + 2 * n}}, Pred{}, s);
  }
}

extern "C" int slx_sht_synthesis(slx_sht* h, const void* coef_v, int nf, double* grid, void* stream) {
  const double2* coef = static_cast<const double2*>(coef_v);
  SLX_CHECK(h && grid && coef, "slx_sht_synthesis: null argument");
  SLX_CHECK(nf >= 1 && nf <= h->max_fields, "slx_sht_synthesis: nf out of range");
  hipStream_t s = (hipStream_t)stream;
  for (int f = 0; f < nf; ++f) {
    const CoefPlain cp{coef + (long)f * h->ncoef, 0};
    const int rc = h->precision == SLX_PREC_FP32 ? launch_leg_syn32<1>(h, cp, Pred{}, s, f)
                                                 : launch_leg_syn<1>(h, cp, Pred{}, s, f);
    if (rc) return rc;
  }
  return launch_fft_syn(h, StorePlain{grid, (long)h->nlat * h->nlon}, nf, Pred{}, s);
}

// packed <-> dense [l][m] (L+1)x(L+1) complex, zeros above the diagonal
__global__ void k_tri_to_dense(const double2* __restrict__ tri, double2* __restrict__ dense, int L,
                               const int* __restrict__ moff) {
  const int l = blockIdx.x;
  for (int m = threadIdx.x; m <= L; m += blockDim.x)
    dense[(long)l * (L + 1) + m] = (m <= l) ? tri[moff[m] + (l - m)] : make_double2(0.0, 0.0);
}
__global__ void k_dense_to_tri(const double2* __restrict__ dense, double2* __restrict__ tri, int L,
                               const int* __restrict__ moff) {
  const int l = blockIdx.x;
  for (int m = threadIdx.x; m <= l; m += blockDim.x) tri[moff[m] + (l - m)] = dense[(long)l * (L + 1) + m];
}

extern "C" int slx_sht_tri_to_dense(const slx_sht* h, const void* tri_v, void* dense_v, int nf, void* stream) {
  const double2* tri = static_cast<const double2*>(tri_v);
  double2* dense = static_cast<double2*>(dense_v);
  SLX_CHECK(h && tri && dense && nf >= 1, "slx_sht_tri_to_dense: bad argument");
  const int L = h->lmax;
  for (int f = 0; f < nf; ++f) {
    k_tri_to_dense<<<L + 1, 256, 0, (hipStream_t)stream>>>(tri + (long)f * h->ncoef,
                                                           dense + (long)f * (L + 1) * (L + 1), L, h->moff);
    SLX_LAUNCH_CHECK();
  }
  return 0;
}
extern "C" int slx_sht_dense_to_tri(const slx_sht* h, const void* dense_v, void* tri_v, int nf, void* stream) {
  const double2* dense = static_cast<const double2*>(dense_v);
  double2* tri = static_cast<double2*>(tri_v);
  SLX_CHECK(h && tri && dense && nf >= 1, "slx_sht_dense_to_tri: bad argument");
  const int L = h->lmax;
  for (int f = 0; f < nf; ++f) {
    k_dense_to_tri<<<L + 1, 256, 0, (hipStream_t)stream>>>(dense + (long)f * (L + 1) * (L + 1),
                                                           tri + (long)f * h->ncoef, L, h->moff);
    SLX_LAUNCH_CHECK();
  }
  return 0;
}
