// Host side of the C ABI (include/mhre.h).
//
// Replaces the reference's host layer src/mcmcjob.cu / src/mcmcjob.cuh
// (CCalcModel / CMcmcFit: per-call cudaMalloc + HtoD of every array, the
// whole object copied to the device, exit(1) on errors) with device-resident
// sessions:
//   * every array of every fit lives in one device arena (one hipMalloc);
//   * the native ring+Gaussian kernels get their own packed inputs
//     (float4 visibilities, closure triplets sorted into LDS-ring windows);
//   * one iteration = one step launch per model kind + one exchange launch,
//     driven by a device iteration clock so a cycle can be captured once in a
//     hipGraph and replayed;
//   * errors are returned as codes with a message (mhre_last_error).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#define MHRE_NO_MODEL_MACROS 1
#include "ccalcmodel.h"
#include "fitdesc.h"
#include "eht_model.h"
#include "engine.h"
#include "grad.h"
#include "mhre.h"
#include "mhre_device.h"

using namespace mhre;

namespace {

thread_local std::string g_err;

void set_err(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

#define FAIL(code, ...)  \
  do {                   \
    set_err(__VA_ARGS__); \
    return code;         \
  } while (0)

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      set_err("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__);   \
      return e_ == hipErrorOutOfMemory ? MHRE_ENOMEM : MHRE_EHIP;                        \
    }                                                                                    \
  } while (0)

#ifdef MHRE_NATIVE_EHT
constexpr int kNativeEht = 1;
#else
constexpr int kNativeEht = 0;
#endif

bool env_flag(const char* name) {
  const char* v = getenv(name);
  return v && *v && strcmp(v, "0") != 0;
}

// ----------------------------------------------------------------- arena

struct Arena {
  size_t bytes = 0;
  template <class T>
  size_t reserve(size_t n) {
    const size_t off = bytes;
    bytes += ((n * sizeof(T) + 255) / 256) * 256;
    if (n == 0) bytes += 256;
    return off;
  }
};

// ------------------------------------------------------------------ fits

enum Kind { K_SGRA9 = 0, K_SGRA13 = 1, K_USER = 2, K_QUAD = 3 };

struct FitHost {
  MhreFitSpec s;
  int kind = K_USER;
  int nchain = 0;
  int nvis_real = 0, ncph = 0, cpext = 0, nvis_k = 0, ntile = 0, ring_ok = 1, cgroup = 1;
  int quad_n = 0;
  // packed host inputs of the native kernels
  std::vector<float> vis;        // 4 per visibility
  std::vector<float> rho2;       // u^2 + v^2 per visibility
  std::vector<int> clo;          // 4 per closure
  std::vector<int> clor;         // 4 per closure: ring byte offsets of the legs, datum
  std::vector<float> clow;       // 4 per closure
  std::vector<int> tinfo;        // 2 per tile: first closure, max u^2+v^2 (float bits)
  std::vector<int> clr;          // gradient: leg rounds per closure (color_closure_legs)
  int nround = 0;
  bool ring_grad = false;        // the gradient may use the single-pass ring kernel
  std::vector<float> qcoef;
  std::vector<float> qrow;       // single-coordinate update rows (fitdesc.h QROW_*), or empty
  std::vector<int> sumh;
  std::vector<uint32_t> hbits;   // packed acceptance history
  // arena offsets
  size_t o_vis, o_rho2, o_clo, o_clor, o_clow, o_tinfo, o_qcoef, o_qrow;
  size_t o_coor, o_dat, o_std2r, o_icoor, o_idat, o_pdescr, o_ivar, o_invar, o_pmint, o_pmaxt, o_beta;
  size_t o_pcur, o_chi2c, o_pstp, o_ncnt, o_nacpt, o_nhist, o_nexch, o_sumh, o_tcur, o_flag, o_ptent, o_ptentn;
  size_t o_pout, o_tout, o_chi2, o_datm, o_chi2m, o_mc;
  size_t o_pcur_b = 0, o_chi2c_b = 0, o_tcur_b = 0;   // second state set (xfused)
  bool xfused = false;           // exchange inside the next step launch (FitDesc::xfused)
  bool hwphase = false;          // FitDesc::hwphase
  bool has_datm = false;
  bool record = true;            // pout/tout/chi2 given
  FitDesc d{};
};

int validate(const MhreFitSpec& s, int idx) {
  if (s.nbeta <= 0 || s.nseq <= 0 || s.nptot <= 0 || s.nprm <= 0 || s.ndat <= 0)
    FAIL(MHRE_EINVAL, "fit %d: nbeta, nseq, nptot, nprm and ndat must be positive", idx);
  if (s.nprm > s.nptot) FAIL(MHRE_EINVAL, "fit %d: nprm (%d) > nptot (%d)", idx, s.nprm, s.nptot);
  if (s.nptot > MAXP) FAIL(MHRE_EINVAL, "fit %d: nptot %d exceeds %d", idx, s.nptot, MAXP);
  if (s.nadj <= 0) FAIL(MHRE_EINVAL, "fit %d: nadj must be positive", idx);
  if (s.nburn < 0 || s.niter < 0) FAIL(MHRE_EINVAL, "fit %d: negative nburn/niter", idx);
  // the step kernels address the acceptance history (and the quadratic-system
  // steps also pstp, n_cnt, sumh and n_acpt) with 32-bit element offsets
  if ((long long)s.nbeta * s.nseq * s.nprm * ((s.nadj + 31) / 32) >= (1ll << 31))
    FAIL(MHRE_EINVAL, "fit %d: nbeta * nseq * nprm * ceil(nadj / 32) must stay below 2^31", idx);
  if (!s.pdescr || !s.ivar || !s.invar || !s.pmint || !s.pmaxt || !s.beta || !s.pcur || !s.chi2c ||
      !s.pstp || !s.n_cnt || !s.n_acpt || !s.n_hist || !s.n_exch || !s.tcur || !s.flag || !s.ptent ||
      !s.ptentn || !s.dat || !s.std2r)
    FAIL(MHRE_EINVAL, "fit %d: a required array pointer is NULL", idx);
  if (s.niter > 0 && (!s.pout || !s.tout || !s.chi2) && (s.pout || s.tout || s.chi2))
    FAIL(MHRE_EINVAL, "fit %d: pout, tout and chi2 must be all given or all NULL", idx);
  int nv = 0;
  for (int i = 0; i < s.nptot; ++i) {
    if (s.pdescr[i] != 0) {
      if (s.invar[i] != nv) FAIL(MHRE_EINVAL, "fit %d: invar[%d] = %d, expected %d", idx, i, s.invar[i], nv);
      if (nv < s.nprm && s.ivar[nv] != i) FAIL(MHRE_EINVAL, "fit %d: ivar[%d] = %d, expected %d", idx, nv, s.ivar[nv], i);
      ++nv;
    } else if (s.invar[i] != -1) {
      FAIL(MHRE_EINVAL, "fit %d: invar[%d] must be -1 for a fixed parameter", idx, i);
    }
  }
  if (nv != s.nprm) FAIL(MHRE_EINVAL, "fit %d: %d nonzero pdescr entries but nprm = %d", idx, nv, s.nprm);
  const size_t pc = (size_t)s.nprm * s.nbeta * s.nseq;
  for (size_t q = 0; q < pc; ++q)
    if (s.n_cnt[q] < 0) FAIL(MHRE_EINVAL, "fit %d: n_cnt must be non-negative (reference: ones)", idx);
  for (size_t q = 0; q < (size_t)s.nadj * pc; ++q)
    if (s.n_hist[q] != 0 && s.n_hist[q] != 1) FAIL(MHRE_EINVAL, "fit %d: n_hist entries must be 0 or 1", idx);
  for (int b = 0; b < s.nbeta; ++b)
    if (!(s.beta[b] >= 0.0f)) FAIL(MHRE_EINVAL, "fit %d: beta[%d] must be >= 0", idx, b);
  for (int i = 0; i < s.nptot; ++i)
    if (s.pdescr[i] != 0 && !(s.pmint[i] <= s.pmaxt[i])) FAIL(MHRE_EINVAL, "fit %d: pmint[%d] > pmaxt[%d]", idx, i, i);
  return MHRE_OK;
}

// Native EHT inputs from the reference arrays (src/imgpu.py:759-831 layout):
// coor = [u0,v0,...] (+ [u1,v1,u2,v2] per closure when cpExt = 1),
// dat/std2r = [amplitudes, closures], icoor = closure triplets, idat = [nvis, nchi2m, cpExt].
// Largest |shift phase|, in revolutions, that the ring+Gaussian kernels can
// meet for this fit: the parameters' magnitudes over the prior box and the
// uploaded chain values (fixed parameters keep theirs; varied ones only take
// values inside the box), times the longest baseline of the packed
// visibilities.  Sgra9 (eht_model.h): ring shift d = ecc (1 - rq) Re, Gaussian
// shift (rq - ecc (1 - rq)) Re; Sgra13: d also times cos(eta - th), Gaussian
// shift gr Re.  A non-finite input gives a non-finite bound, and so does a
// spec without a prior (the gradient's data handle, which samples nothing).
double eht_phase_bound(const FitHost& h) {
  const MhreFitSpec& s = h.s;
  if (!s.pmint || !s.pmaxt) return INFINITY;
  const size_t N = (size_t)s.nptot, C = (size_t)s.nbeta * s.nseq;
  auto range = [&](int k, double* lo, double* hi) {
    *lo = std::min((double)s.pmint[k], (double)s.pmaxt[k]);
    *hi = std::max((double)s.pmint[k], (double)s.pmaxt[k]);
    if (!std::isfinite(*lo) || !std::isfinite(*hi)) *hi = INFINITY;
    if (s.pcur)
      for (size_t c = 0; c < C; ++c) {
        const double v = s.pcur[c * N + k];
        if (!std::isfinite(v)) *hi = INFINITY;
        *lo = std::min(*lo, v);
        *hi = std::max(*hi, v);
      }
  };
  auto amax = [&](int k) {
    double lo, hi;
    range(k, &lo, &hi);
    return std::max(std::fabs(lo), std::fabs(hi));
  };
  double rlo, rhi;
  range(2, &rlo, &rhi);
  const double one_minus_rq = std::max(std::fabs(1.0 - rlo), std::fabs(1.0 - rhi));
  const double re = amax(1) * UAS2NRAD;
  const double e = amax(3) * one_minus_rq;   // |ecc (1 - rq)|
  const double ring = e * re;
  const double gauss = (h.kind == K_SGRA9 ? amax(2) + e : amax(5)) * re;
  double rho = 0.0;
  for (int i = 0; i < h.nvis_k; ++i)
    rho = std::max(rho, std::hypot((double)h.vis[4 * i], (double)h.vis[4 * i + 1]));
  return std::max(ring, gauss) * rho;
}

int pack_eht(FitHost& h, int idx) {
  const MhreFitSpec& s = h.s;
  const int need = h.kind == K_SGRA9 ? 9 : 13;
  if (s.nptot < need) FAIL(MHRE_EINVAL, "fit %d: model needs %d parameters, nptot = %d", idx, need, s.nptot);
  if (!s.idat || s.nidat < 1 || !s.coor) FAIL(MHRE_EINVAL, "fit %d: idat/coor missing", idx);
  const int nvis = s.idat[0];
  const int cpext = s.nidat >= 3 ? s.idat[2] : 0;
  if (nvis <= 0 || nvis > s.ndat) FAIL(MHRE_EINVAL, "fit %d: idat[0] = %d not in [1, ndat]", idx, nvis);
  const int ncph = s.ndat - nvis;
  if (s.ncoor < 2 * nvis) FAIL(MHRE_EINVAL, "fit %d: ncoor %d < 2*nvis", idx, s.ncoor);
  if (ncph > 0 && cpext == 0 && (!s.icoor || s.nicoor < 3 * ncph))
    FAIL(MHRE_EINVAL, "fit %d: nicoor %d < 3*ncph", idx, s.nicoor);
  if (ncph > 0 && cpext != 0 && s.ncoor < 2 * nvis + 4 * ncph)
    FAIL(MHRE_EINVAL, "fit %d: cpExt=1 needs ncoor >= 2*nvis + 4*ncph", idx);
  h.nvis_real = nvis;
  h.ncph = ncph;
  h.cpext = cpext;
  h.nvis_k = nvis + (cpext ? 3 * ncph : 0);
  h.ntile = (h.nvis_k + TILE - 1) / TILE;
  // one extra tile of zero-weight padding: the kernel prefetches tile t + 1
  h.vis.assign((size_t)(h.ntile + 1) * TILE * 4, 0.0f);
  // amplitudes as (d sqrt(w), sqrt(w)): the weighted residual (d - |V|) sqrt(w)
  // is then one FMA per visibility
  for (int i = 0; i < nvis; ++i) {
    if (!(s.std2r[i] >= 0.0f))
      FAIL(MHRE_EINVAL, "fit %d: std2r[%d] = %g: amplitude weights must be non-negative", idx, i, (double)s.std2r[i]);
    const double sw = std::sqrt((double)s.std2r[i]);
    h.vis[4 * i + 0] = s.coor[2 * i];
    h.vis[4 * i + 1] = s.coor[2 * i + 1];
    h.vis[4 * i + 2] = (float)((double)s.dat[i] * sw);
    h.vis[4 * i + 3] = (float)sw;
  }
  std::vector<int> tri((size_t)ncph * 3);
  for (int j = 0; j < ncph; ++j) {
    if (cpext) {
      const float* c = s.coor + 2 * nvis + 4 * j;
      const float uu[3] = {c[0], c[2], -(c[0] + c[2])};
      const float vv[3] = {c[1], c[3], -(c[1] + c[3])};
      for (int m = 0; m < 3; ++m) {
        const int v = nvis + 3 * j + m;
        h.vis[4 * v + 0] = uu[m];
        h.vis[4 * v + 1] = vv[m];
        tri[3 * j + m] = v;
      }
    } else {
      for (int m = 0; m < 3; ++m) {
        const int v = s.icoor[3 * j + m];
        if (v < 0 || v >= nvis) FAIL(MHRE_EINVAL, "fit %d: icoor[%d] = %d outside [0, nvis)", idx, 3 * j + m, v);
        tri[3 * j + m] = v;
      }
    }
  }
  // Closures are formed from the LDS ring after every cgroup-th tile, as long
  // as all three indices are still in the RING-deep window at that point.
  // cgroup aims at about one wave of closures per pass.
  std::vector<int> tile_of(ncph), lo_of(ncph);
  for (int j = 0; j < ncph; ++j) {
    const int a = tri[3 * j], b = tri[3 * j + 1], c = tri[3 * j + 2];
    tile_of[j] = std::max(a, std::max(b, c)) / TILE;
    lo_of[j] = std::min(a, std::min(b, c));
  }
  auto window_ok = [&](int g) {
    for (int j = 0; j < ncph; ++j) {
      const int tp = std::min((tile_of[j] / g + 1) * g - 1, h.ntile - 1);   // tile after which it is formed
      if (lo_of[j] < (tp + 1) * TILE - RING) return false;
    }
    return true;
  };
  int g = ncph > 0 ? (int)std::lround((double)TILE * h.ntile / ncph) : 1;
  g = std::max(1, std::min(MAX_CGROUP, g));
  while (g > 1 && !window_ok(g)) --g;
  const bool ok = window_ok(g);
  h.cgroup = g;
  h.ring_ok = ok ? 1 : 0;
  std::vector<int> order(ncph);
  for (int j = 0; j < ncph; ++j) order[j] = j;
  if (ok) std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return tile_of[x] < tile_of[y]; });
  // one wave of zero padding: lanes prefetch closure c0 + lane
  h.clo.assign(((size_t)ncph + TILE) * 4, 0);
  h.clor.assign(((size_t)ncph + TILE) * 4, 0);
  h.clow.assign(((size_t)ncph + TILE) * 4, 0.0f);
  for (int q = 0; q < ncph; ++q) {
    const int j = order[q];
    h.clo[4 * q + 0] = tri[3 * j];
    h.clo[4 * q + 1] = tri[3 * j + 1];
    h.clo[4 * q + 2] = tri[3 * j + 2];
    h.clo[4 * q + 3] = nvis + j;   // datum index
    for (int m = 0; m < 3; ++m) h.clor[4 * q + m] = (tri[3 * j + m] & (RING - 1)) * (int)sizeof(float2);
    h.clor[4 * q + 3] = nvis + j;
    const double obs = s.dat[nvis + j];
    h.clow[4 * q + 0] = (float)cos(obs);
    h.clow[4 * q + 1] = (float)sin(obs);
    h.clow[4 * q + 2] = s.std2r[nvis + j];
  }
  h.tinfo.assign(2 * ((size_t)h.ntile + 2), 0);
  if (ok) {
    int q = 0;
    for (int t = 0; t <= h.ntile + 1; ++t) {
      while (q < ncph && tile_of[order[q]] < t) ++q;
      h.tinfo[2 * t] = q;
    }
  }
  h.rho2.assign((size_t)(h.ntile + 1) * TILE, 0.0f);
  for (int i = 0; i < h.nvis_k; ++i) {
    const double u = h.vis[4 * i], v = h.vis[4 * i + 1];
    h.rho2[i] = (float)(u * u + v * v);
  }
  for (int t = 0; t < h.ntile; ++t) {
    float mx = 0.0f;
    for (int i = t * TILE; i < std::min((t + 1) * TILE, h.nvis_k); ++i) mx = std::max(mx, h.rho2[i]);
    // rounded up so that se * mx bounds every lane's se * (u^2 + v^2)
    mx = std::nextafter(mx * (1.0f + 1e-6f), INFINITY);
    memcpy(&h.tinfo[2 * t + 1], &mx, 4);
  }
  // The shift phases need no reduction before the hardware sine when they
  // are bounded well inside its input domain (|h| <= 256 revolutions).
  h.hwphase = eht_phase_bound(h) <= 64.0 && !env_flag("MHRE_REDUCE_PHASES");
  return MHRE_OK;
}

// MODEL_QUADSYS: coor = [Q (neq*n*n), b (neq*n)], dat = rhs, std2r = weights.
int pack_quad(FitHost& h, int idx) {
  const MhreFitSpec& s = h.s;
  const int n = s.nptot, neq = s.ndat;
  if (n != 2 && n != 4 && n != 8) FAIL(MHRE_EINVAL, "fit %d: quadratic systems support nptot in {2,4,8}", idx);
  if (!s.coor || s.ncoor < neq * n * n + neq * n) FAIL(MHRE_EINVAL, "fit %d: coor too short for Q and b", idx);
  const int ns = n * (n + 1) / 2, qs = ns + n + 2;
  h.quad_n = n;
  h.qcoef.assign((size_t)neq * qs, 0.0f);
  for (int e = 0; e < neq; ++e) {
    const float* Q = s.coor + (size_t)e * n * n;
    const float* b = s.coor + (size_t)neq * n * n + (size_t)e * n;
    float* c = &h.qcoef[(size_t)e * qs];
    int m = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) c[m++] = (i == j) ? Q[i * n + i] : Q[i * n + j] + Q[j * n + i];
    for (int i = 0; i < n; ++i) c[ns + i] = b[i];
    c[ns + n] = s.dat[e];
    c[ns + n + 1] = s.std2r[e];
  }
  // Rows of the single-coordinate update (k_iter_quad: ladders of at most 64
  // rungs) for systems of at most QROW_MAXEQ equations with non-negative
  // weights (the rows carry sqrt(w)); MHRE_QUAD_FULL=1 keeps the evaluation
  // of every equation in full.
  h.qrow.clear();
  bool rows = s.nbeta <= 64 && neq <= QROW_MAXEQ && !env_flag("MHRE_QUAD_FULL");
  for (int e = 0; rows && e < neq; ++e) rows = s.std2r[e] >= 0.0f && std::isfinite(s.std2r[e]);
  if (rows) {
    const int ts = QROW_TS(n), E = QROW_MAXEQ;
    h.qrow.assign((size_t)QROW_FLOATS(n), 0.0f);
    for (int e = 0; e < neq; ++e) {
      const float* Q = s.coor + (size_t)e * n * n;
      const float* b = s.coor + (size_t)neq * n * n + (size_t)e * n;
      const float sw = std::sqrt(s.std2r[e]);
      for (int t = 0; t < n; ++t) {
        float* blk = &h.qrow[(size_t)t * ts];
        for (int j = 0; j < n; ++j) blk[j * E + e] = sw * (j == t ? Q[t * n + t] : Q[t * n + j] + Q[j * n + t]);
        blk[n * E + e] = sw * b[t];
        blk[(n + 1) * E + e] = sw * Q[t * n + t];
      }
      h.qrow[(size_t)n * ts + e] = sw;
    }
  }
  return MHRE_OK;
}

// ------------------------------------------------------- XCD work ranges

// Re-orders items (chunks of up to XCHUNK consecutive items of one fit) into
// NXCD contiguous ranges of nearly equal cost: chunks, in their given
// (heaviest first) order, go to the range with the least cost so far.
XcdRanges deal_xcd(std::vector<WorkItem>& items, const std::vector<double>& fit_cost) {
  struct Chunk { size_t b, e; };
  std::vector<Chunk> chunks;
  for (size_t i = 0; i < items.size();) {
    size_t j = i;
    while (j < items.size() && j - i < (size_t)XCHUNK && items[j].fit == items[i].fit) ++j;
    chunks.push_back({i, j});
    i = j;
  }
  std::vector<std::vector<size_t>> lists(NXCD);
  double load[NXCD] = {};
  for (size_t c = 0; c < chunks.size(); ++c) {
    int x = 0;
    for (int y = 1; y < NXCD; ++y)
      if (load[y] < load[x]) x = y;
    lists[x].push_back(c);
    load[x] += (double)(chunks[c].e - chunks[c].b) * fit_cost[items[chunks[c].b].fit];
  }
  std::vector<WorkItem> out;
  out.reserve(items.size());
  XcdRanges r{};
  for (int x = 0; x < NXCD; ++x) {
    r.off[x] = (int)out.size();
    for (size_t c : lists[x]) out.insert(out.end(), items.begin() + chunks[c].b, items.begin() + chunks[c].e);
  }
  r.off[NXCD] = (int)out.size();
  items.swap(out);
  return r;
}

// ------------------------------------------------- acceptance history bits

// Lanes per sequence in the exchange kernel: ladders of up to 32 rungs take
// the next power of two, so that 2 to 64 sequences share a wave.
int exchange_lanes(int nbeta) {
  if (nbeta > 32) return 64;
  int w = 1;
  while (w < nbeta) w <<= 1;
  return w;
}

int hist_words(int nadj) { return (nadj + 31) / 32; }

// Quadratic-system fits whose ladder fits one wave run step and exchange in
// one launch (k_iter_quad).
bool quad_fused(const FitHost& h) { return h.kind == K_QUAD && h.s.nbeta <= 64; }

// Ring+Gaussian fits whose ladder fits one wave exchange at the start of the
// next step launch (FitDesc::xfused) unless the caller asks for the exchange
// launch of every iteration.
bool eht_xfused(const FitHost& h, const MhreOptions& o) {
  return (h.kind == K_SGRA9 || h.kind == K_SGRA13) && h.s.nbeta <= 64 && o.exchange_launch == 0;
}

// Most concurrent sequence groups (streams) of a session.
constexpr int MAX_GROUPS = 8;

// n_hist[a][q] (0/1, q = k * nchain + chain) -> bit a % 32 of word [a / 32][q]
// Packs the 0/1 history n_hist[nadj][npc] into bits and, in the same pass,
// the running accept sums sum[q] = number of set entries of q.
// n_hist[a][k][chain] (0/1) <-> bit a % 32 of hbits[chain][a / 32][k]
// (fitdesc.h), with sum[k][chain] = the set entries of (k, chain).
void pack_hist(const int* n_hist, int nadj, int P, int C, std::vector<uint32_t>& bits, std::vector<int>& sum) {
  const int W = hist_words(nadj);
  const size_t npc = (size_t)P * C;
  bits.assign((size_t)W * npc, 0u);
  sum.assign(npc, 0);
  for (int a = 0; a < nadj; ++a) {
    const int* h = n_hist + (size_t)a * npc;
    const uint32_t bit = 1u << (a & 31);
    for (int k = 0; k < P; ++k)
      for (int c = 0; c < C; ++c) {
        const size_t q = (size_t)k * C + c;
        const int on = h[q] != 0;
        bits[((size_t)c * W + (a >> 5)) * P + k] |= on ? bit : 0u;
        sum[q] += on;
      }
  }
}

void unpack_hist(const std::vector<uint32_t>& bits, int nadj, int P, int C, int* n_hist) {
  const int W = hist_words(nadj);
  const size_t npc = (size_t)P * C;
  for (int a = 0; a < nadj; ++a)
    for (int k = 0; k < P; ++k)
      for (int c = 0; c < C; ++c)
        n_hist[(size_t)a * npc + (size_t)k * C + c] =
            (int)((bits[((size_t)c * W + (a >> 5)) * P + k] >> (a & 31)) & 1u);
}

// ------------------------------------------------------------- session

struct KindWork {
  int kind;
  int quad_n;
  bool fused = false;   // quadratic systems: step and exchange in one launch, items (fit, first sequence)
  bool rows = false;    // fused quadratic systems: every fit has single-coordinate update rows
  std::vector<WorkItem> step, eval;
  XcdRanges xr{};   // wave-per-chain kinds: step list dealt into per-XCD ranges
  WorkItem* d_step = nullptr;
  WorkItem* d_eval = nullptr;
  int qslot = 0;
};

// A group of whole sequences (of every fit) whose iterations run on their own
// stream with their own iteration clock.  Sequences are coupled only through
// the replica exchange of their own temperature ladder, so groups never share
// a chain, and the launches of different groups may run side by side: while
// one group's step launch drains its last waves, another group's waves fill
// the compute units.  Group 0 runs on the session stream.
struct Group {
  hipStream_t stream = nullptr;
  RunClock* d_clk = nullptr;
  std::vector<KindWork> kinds;   // step lists of the group's chains
  WorkItem* d_x = nullptr;       // sequences exchanged by a launch of their own every iteration
  int n_x = 0;
  XcdRanges xr_x{};
  WorkItem* d_xc = nullptr;      // xfused sequences, finished by a segment's closing launch
  XcdRanges xr_xc{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t done = nullptr;
};

struct Session {
  hipStream_t stream = nullptr;
  std::vector<FitHost> fits;
  char* arena = nullptr;
  size_t arena_bytes = 0;
  FitDesc* d_fits = nullptr;
  unsigned long long* d_nevals = nullptr;   // [nfits][2]
  int* d_err = nullptr;
  WorkItem* d_init = nullptr;
  int n_init = 0;
  std::vector<KindWork> kinds;   // evaluation lists (every chain) per model kind
  std::vector<Group> groups;
  MhreOptions opt{};
  int wpb = 4, persistent = 0, grid = 0, ncu = 1;
  int waves_per_simd = 0, lds_per_cu = 0;   // occupancy cap of the native step launches
  int graph_iters = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int iters_done = 0;
  int* d_tr_dec = nullptr;
  float* d_tr_chi2 = nullptr;
  int* d_tr_ex = nullptr;
  int tr_nit = 0, tr_it0 = 0;
  size_t tr_dec_total = 0, tr_ex_total = 0;
  std::vector<void*> extra;   // other device allocations
};

template <class T>
T* at(Session* S, size_t off) { return reinterpret_cast<T*>(S->arena + off); }

// force_generic (or MHRE_FORCE_GENERIC) sends the ring+Gaussian models
// through the compiled-in model.cuh as well.
int resolve_kind(const MhreFitSpec& s, bool force_generic) {
  const bool generic = force_generic || env_flag("MHRE_FORCE_GENERIC");
  switch (s.model) {
    case MHRE_MODEL_SGRA9: return generic ? K_USER : K_SGRA9;
    case MHRE_MODEL_SGRA13: return generic ? K_USER : K_SGRA13;
    case MHRE_MODEL_USER: return K_USER;
    case MHRE_MODEL_QUADSYS: return K_QUAD;
    default: break;
  }
  if (kNativeEht && !generic) {
    if (s.imodel == 1) return K_SGRA9;
    if (s.imodel == 2) return K_SGRA13;
  }
  return K_USER;
}

void destroy_session(Session* S) {
  if (!S) return;
  for (Group& g : S->groups)
    if (g.stream) hipStreamSynchronize(g.stream);
  if (S->stream) hipStreamSynchronize(S->stream);
  for (Group& g : S->groups) {
    if (g.gexec) hipGraphExecDestroy(g.gexec);
    if (g.graph) hipGraphDestroy(g.graph);
    if (g.done) hipEventDestroy(g.done);
    if (g.stream && g.stream != S->stream) hipStreamDestroy(g.stream);
  }
  // every device allocation except the arena is owned by S->extra
  for (void* p : S->extra) hipFree(p);
  if (S->arena) hipFree(S->arena);
  if (S->ev0) hipEventDestroy(S->ev0);
  if (S->ev1) hipEventDestroy(S->ev1);
  if (S->stream) hipStreamDestroy(S->stream);
  delete S;
}

template <class T>
int dev_alloc(Session* S, T** p, size_t n) {
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(n * sizeof(T), 16)));
  S->extra.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return MHRE_OK;
}

template <class T>
int h2d(Session* S, size_t off, const T* src, size_t n) {
  if (!src || n == 0) return MHRE_OK;
  HIPCHK(hipMemcpyAsync(S->arena + off, src, n * sizeof(T), hipMemcpyHostToDevice, S->stream));
  return MHRE_OK;
}

template <class T>
int d2h(Session* S, T* dst, size_t off, size_t n) {
  if (!dst || n == 0) return MHRE_OK;
  HIPCHK(hipMemcpyAsync(dst, S->arena + off, n * sizeof(T), hipMemcpyDeviceToHost, S->stream));
  return MHRE_OK;
}

#define CHK(x)              \
  do {                      \
    int r_ = (x);           \
    if (r_ != MHRE_OK) return r_; \
  } while (0)

// Many small arena transfers (a session of a thousand small fits has tens of
// thousands of arrays) gathered into a few large copies.  Ranges closer than
// MERGE_GAP bytes travel in one copy; ranges of at least DIRECT bytes are
// copied straight between the arena and the caller's memory.
constexpr size_t MERGE_GAP = 64 << 10;
constexpr size_t DIRECT = 4 << 20;

struct Range {
  size_t off, bytes;
  unsigned char* host;   // destination (download) or copied source (upload)
};

// Sorted, non-overlapping ranges -> runs [first, last) whose gaps are < MERGE_GAP.
static std::vector<std::pair<size_t, size_t>> merge_runs(std::vector<Range>& r) {
  std::sort(r.begin(), r.end(), [](const Range& a, const Range& b) { return a.off < b.off; });
  std::vector<std::pair<size_t, size_t>> runs;
  for (size_t i = 0; i < r.size();) {
    size_t j = i + 1, end = r[i].off + r[i].bytes;
    while (j < r.size() && r[j].off < end + MERGE_GAP) {
      end = std::max(end, r[j].off + r[j].bytes);
      ++j;
    }
    runs.push_back({i, j});
    i = j;
  }
  return runs;
}

// Uploads: each source is copied into host memory when added, so callers may
// reuse their buffers at once; flush() copies the runs and waits for them.
struct Uploads {
  std::vector<Range> r;
  std::vector<std::vector<unsigned char>> keep;
  template <class T>
  void add(size_t off, const T* src, size_t n) {
    if (!src || n == 0) return;
    keep.emplace_back((const unsigned char*)src, (const unsigned char*)src + n * sizeof(T));
    r.push_back({off, n * sizeof(T), keep.back().data()});
  }
  int flush(Session* S) {
    std::vector<std::vector<unsigned char>> stage;
    for (auto [a, b] : merge_runs(r)) {
      const size_t lo = r[a].off, hi = r[b - 1].off + r[b - 1].bytes;
      if (b - a == 1) {
        HIPCHK(hipMemcpyAsync(S->arena + lo, r[a].host, r[a].bytes, hipMemcpyHostToDevice, S->stream));
        continue;
      }
      // gap bytes are arena padding or arrays the kernels write before reading
      stage.emplace_back(hi - lo, 0);
      for (size_t i = a; i < b; ++i) std::memcpy(stage.back().data() + (r[i].off - lo), r[i].host, r[i].bytes);
      HIPCHK(hipMemcpyAsync(S->arena + lo, stage.back().data(), hi - lo, hipMemcpyHostToDevice, S->stream));
    }
    HIPCHK(hipStreamSynchronize(S->stream));   // the host copies must outlive the transfers
    r.clear();
    keep.clear();
    return MHRE_OK;
  }
};

// Downloads: flush() copies the runs into a host staging buffer (large ranges
// straight to their destination) and scatters them.
struct Downloads {
  std::vector<Range> r;
  template <class T>
  void add(T* dst, size_t off, size_t n) {
    if (!dst || n == 0) return;
    r.push_back({off, n * sizeof(T), (unsigned char*)dst});
  }
  int flush(Session* S) {
    std::vector<Range> small;
    for (const Range& x : r) {
      if (x.bytes >= DIRECT)
        HIPCHK(hipMemcpyAsync(x.host, S->arena + x.off, x.bytes, hipMemcpyDeviceToHost, S->stream));
      else
        small.push_back(x);
    }
    const auto runs = merge_runs(small);
    size_t total = 0;
    for (auto [a, b] : runs) total += small[b - 1].off + small[b - 1].bytes - small[a].off;
    std::vector<unsigned char> stage(total);
    size_t at = 0;
    for (auto [a, b] : runs) {
      const size_t lo = small[a].off, hi = small[b - 1].off + small[b - 1].bytes;
      HIPCHK(hipMemcpyAsync(stage.data() + at, S->arena + lo, hi - lo, hipMemcpyDeviceToHost, S->stream));
      at += hi - lo;
    }
    HIPCHK(hipStreamSynchronize(S->stream));
    at = 0;
    for (auto [a, b] : runs) {
      const size_t lo = small[a].off;
      for (size_t i = a; i < b; ++i) std::memcpy(small[i].host, stage.data() + at + (small[i].off - lo), small[i].bytes);
      at += small[b - 1].off + small[b - 1].bytes - lo;
    }
    r.clear();
    return MHRE_OK;
  }
};

StepLaunch make_launch(Session* S, const Group& g, const KindWork& k, int it_off) {
  StepLaunch L{};
  L.fits = S->d_fits;
  L.work = k.d_step;
  L.nwork = (int)k.step.size();
  L.xr = k.xr;
  L.clk = g.d_clk;
  L.it_off = it_off;
  L.wpb = S->wpb;
  L.persistent = S->persistent;
  L.grid = S->grid;
  L.qslot = k.qslot;
  L.waves_per_simd = S->waves_per_simd;
  L.lds_per_cu = S->lds_per_cu;
  L.tr_dec = S->d_tr_dec;
  L.tr_chi2 = S->d_tr_chi2;
  L.tr_ex = S->d_tr_ex;
  L.tr_it0 = S->tr_it0;
  L.tr_nit = S->tr_nit;
  return L;
}

// An exchange launch of group g (with the decision trace, if any).
StepLaunch exchange_launch(Session* S, const Group& g, int it_off, int close_k) {
  StepLaunch L{};
  L.fits = S->d_fits;
  L.clk = g.d_clk;
  L.it_off = it_off;
  L.close_k = close_k;
  L.tr_ex = S->d_tr_ex;
  L.tr_it0 = S->tr_it0;
  L.tr_nit = S->tr_nit;
  return L;
}

// The step launches of group g at offset i of its segment.
int launch_steps(Session* S, const Group& g, int i) {
  for (const KindWork& k : g.kinds) {
    const StepLaunch L = make_launch(S, g, k, i);
    hipError_t e = hipSuccess;
    if (k.kind == K_SGRA9) e = launch_step_eht(MODEL_SGRA9, L, g.stream);
    else if (k.kind == K_SGRA13) e = launch_step_eht(MODEL_SGRA13, L, g.stream);
    else if (k.kind == K_QUAD && k.fused) e = launch_iter_quad(k.quad_n, k.rows, L, g.stream);
    else if (k.kind == K_QUAD) e = launch_step_quad(k.quad_n, L, g.stream);
    else e = launch_step_user(L, g.stream);
    HIPCHK(e);
  }
  return MHRE_OK;
}

// The exchange launch at offset i of group g's segment, for the sequences
// that exchange in a launch of their own every iteration (none: no launch).
int launch_group_exchange(Session* S, const Group& g, int i) {
  if (g.n_x > 0) HIPCHK(launch_exchange(exchange_launch(S, g, i, 0), g.d_x, g.xr_x, g.stream));
  return MHRE_OK;
}

// Offset i of group g's segment on its stream: the step launches, then the
// exchange launch of the sequences that need one.
int launch_iteration(Session* S, const Group& g, int i) {
  CHK(launch_steps(S, g, i));
  return launch_group_exchange(S, g, i);
}

// The closing launch of group g's segment of K iterations: the xfused
// sequences' last exchange, then the clock moves on by K.
int close_segment(Session* S, const Group& g, int K) {
  HIPCHK(launch_exchange(exchange_launch(S, g, 0, K), g.d_xc, g.xr_xc, g.stream));
  return MHRE_OK;
}

// A segment of K iterations of every group, the groups' launches interleaved.
int launch_segment(Session* S, int K) {
  for (int i = 0; i < K; ++i)
    for (const Group& g : S->groups) CHK(launch_iteration(S, g, i));
  for (const Group& g : S->groups) CHK(close_segment(S, g, K));
  return MHRE_OK;
}

// The other groups' streams start after what the session stream holds so far
// (hipStreamWaitEvent on ev), and the session stream later waits for their
// work (join_groups), so one event pair on the session stream brackets all.
int fork_groups(Session* S, hipEvent_t ev) {
  for (size_t i = 1; i < S->groups.size(); ++i) HIPCHK(hipStreamWaitEvent(S->groups[i].stream, ev, 0));
  return MHRE_OK;
}
int join_groups(Session* S) {
  for (size_t i = 1; i < S->groups.size(); ++i) {
    Group& g = S->groups[i];
    HIPCHK(hipEventRecord(g.done, g.stream));
    HIPCHK(hipStreamWaitEvent(S->stream, g.done, 0));
  }
  return MHRE_OK;
}

// Iterations (in segments of BATCH) with HIP events around every launch, on
// each group's stream: average device time per iteration of a group's step
// launches and of its exchange launches (the per-iteration ones plus the
// closing launch of each segment), over groups and iterations (launches of
// different groups overlap).  All launches and events of a segment are queued
// before the host waits, so the gaps between launches are not counted.
int time_kernels(Session* S, int niter, double* step_ms, double* exch_ms) {
  (void)hipGetLastError();
  constexpr int BATCH = 32;
  const int G = (int)S->groups.size();
  std::vector<hipEvent_t> ev((size_t)3 * BATCH * G + 2 * G, nullptr);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  struct Destroy {
    std::vector<hipEvent_t>& v;
    ~Destroy() { for (auto e : v) if (e) hipEventDestroy(e); }
  } destroy{ev};
  hipEvent_t* ec = &ev[(size_t)3 * BATCH * G];   // closing launches: 2 per group
  double st = 0.0, ex = 0.0;
  HIPCHK(hipEventRecord(S->ev0, S->stream));
  CHK(fork_groups(S, S->ev0));
  for (int done = 0; done < niter;) {
    const int nb = std::min(BATCH, niter - done);
    for (int i = 0; i < nb; ++i) {
      for (int gi = 0; gi < G; ++gi) {
        const Group& g = S->groups[gi];
        hipEvent_t* e = &ev[(size_t)3 * (i * G + gi)];
        HIPCHK(hipEventRecord(e[0], g.stream));
        CHK(launch_steps(S, g, i));
        HIPCHK(hipEventRecord(e[1], g.stream));
        CHK(launch_group_exchange(S, g, i));
        HIPCHK(hipEventRecord(e[2], g.stream));
      }
    }
    for (int gi = 0; gi < G; ++gi) {
      const Group& g = S->groups[gi];
      HIPCHK(hipEventRecord(ec[2 * gi], g.stream));
      CHK(close_segment(S, g, nb));
      HIPCHK(hipEventRecord(ec[2 * gi + 1], g.stream));
    }
    for (int gi = 0; gi < G; ++gi) HIPCHK(hipEventSynchronize(ec[2 * gi + 1]));
    for (int i = 0; i < nb * G; ++i) {
      float a = 0.0f, b = 0.0f;
      HIPCHK(hipEventElapsedTime(&a, ev[(size_t)3 * i], ev[(size_t)3 * i + 1]));
      HIPCHK(hipEventElapsedTime(&b, ev[(size_t)3 * i + 1], ev[(size_t)3 * i + 2]));
      st += a;
      ex += b;
    }
    for (int gi = 0; gi < G; ++gi) {
      float c = 0.0f;
      HIPCHK(hipEventElapsedTime(&c, ec[2 * gi], ec[2 * gi + 1]));
      ex += c;
    }
    done += nb;
    S->iters_done += nb;
  }
  CHK(join_groups(S));
  HIPCHK(hipStreamSynchronize(S->stream));
  if (step_ms) *step_ms = niter > 0 ? st / ((double)niter * G) : 0.0;
  if (exch_ms) *exch_ms = niter > 0 ? ex / ((double)niter * G) : 0.0;
  return MHRE_OK;
}

void drop_graph(Session* S) {
  for (Group& g : S->groups) {
    if (g.gexec) hipGraphExecDestroy(g.gexec);
    if (g.graph) hipGraphDestroy(g.graph);
    g.gexec = nullptr;
    g.graph = nullptr;
  }
}

// Device time per iteration of niter iterations of every group (events on
// the session stream around the forked and joined group streams).
int time_iterations(Session* S, int niter, double* ms) {
  HIPCHK(hipEventRecord(S->ev0, S->stream));
  CHK(fork_groups(S, S->ev0));
  if (niter > 0) CHK(launch_segment(S, niter));
  CHK(join_groups(S));
  HIPCHK(hipEventRecord(S->ev1, S->stream));
  HIPCHK(hipEventSynchronize(S->ev1));
  float t = 0.0f;
  HIPCHK(hipEventElapsedTime(&t, S->ev0, S->ev1));
  S->iters_done += niter;
  *ms = niter > 0 ? t / niter : 0.0;
  return MHRE_OK;
}

// Picks the launch form of the wave-per-chain step kernels by timing real
// iterations (they advance the chains like any other iteration): each
// (waves per work-group, grid or persistent) candidate runs `iters`
// iterations of every group and the one with the shortest device time per
// iteration stays.  All forms give bit-identical chains.  A captured graph is
// dropped and re-captured later.
int autotune(Session* S, int iters, int* best_wpb, int* best_persistent, double* best_ms) {
  static const int cand[6][2] = {{4, 0}, {2, 0}, {1, 0}, {4, 1}, {2, 1}, {1, 1}};
  if (iters <= 0) FAIL(MHRE_EINVAL, "autotune: iters must be positive");
  double best = 1e300;
  int bw = S->wpb, bp = S->persistent;
  for (const auto& c : cand) {
    S->wpb = c[0];
    S->persistent = c[1];
    S->grid = std::max(1, S->ncu * (16 / S->wpb));
    double t = 0.0;
    CHK(time_iterations(S, iters, &t));
    if (t < best) {
      best = t;
      bw = c[0];
      bp = c[1];
    }
  }
  S->wpb = bw;
  S->persistent = bp;
  S->grid = std::max(1, S->ncu * (16 / S->wpb));
  drop_graph(S);
  if (best_wpb) *best_wpb = bw;
  if (best_persistent) *best_persistent = bp;
  if (best_ms) *best_ms = best;
  return MHRE_OK;
}

int launch_eval_all(Session* S, int outputs, int respect_flag) {
  for (auto& k : S->kinds) {
    const int n = (int)k.eval.size();
    hipError_t e = hipSuccess;
    if (k.kind == K_SGRA9) e = launch_eval_eht(MODEL_SGRA9, S->d_fits, k.d_eval, n, outputs, respect_flag, S->stream);
    else if (k.kind == K_SGRA13) e = launch_eval_eht(MODEL_SGRA13, S->d_fits, k.d_eval, n, outputs, respect_flag, S->stream);
    else if (k.kind == K_QUAD) e = launch_eval_quad(k.quad_n, S->d_fits, k.d_eval, n, S->stream);
    else e = launch_eval_user(S->d_fits, k.d_eval, n, 0, respect_flag, S->stream);
    HIPCHK(e);
  }
  return MHRE_OK;
}

int check_model_error(Session* S) {
  int err = 0;
  HIPCHK(hipMemcpyAsync(&err, S->d_err, sizeof(int), hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  if (err) FAIL(MHRE_EMODEL, "model() returned error code %d (e.g. an unknown imodel)", err);
  return MHRE_OK;
}

int create(const MhreFitSpec* specs, int nfits, const MhreOptions* opt, Session** out, int eval_only) {
  *out = nullptr;
  (void)hipGetLastError();   // a stale error of other code must not be reported as ours
  if (!specs || nfits <= 0) FAIL(MHRE_EINVAL, "no fits given");
  Session* S = new Session();
  struct Guard {
    Session*& s;
    ~Guard() { if (s) destroy_session(s); }
  } guard{S};
  if (opt) S->opt = *opt;
  S->wpb = S->opt.waves_per_block == 1 || S->opt.waves_per_block == 2 ? S->opt.waves_per_block : 4;
  S->persistent = S->opt.persistent > 0 ? 1 : 0;
  S->waves_per_simd = std::max(0, S->opt.waves_per_simd);

  const int it0 = specs[0].it0;
  S->fits.resize(nfits);
  Arena A;
  for (int i = 0; i < nfits; ++i) {
    FitHost& h = S->fits[i];
    h.s = specs[i];
    if (!eval_only) CHK(validate(h.s, i));
    if (h.s.it0 != it0) FAIL(MHRE_EINVAL, "fit %d: all fits of a session must share it0", i);
    h.kind = resolve_kind(h.s, S->opt.force_generic != 0);
    if (h.kind == K_USER && (h.s.model == MHRE_MODEL_SGRA9 || h.s.model == MHRE_MODEL_SGRA13) && !kNativeEht)
      FAIL(MHRE_EINVAL, "fit %d: the ring+Gaussian models on the model.cuh path need a library built with "
                        "examples/sgra_model_fit/model.cuh", i);
    h.nchain = h.s.nbeta * h.s.nseq;
    if (h.kind == K_SGRA9 || h.kind == K_SGRA13) CHK(pack_eht(h, i));
    if (h.kind == K_QUAD) CHK(pack_quad(h, i));
    if (h.kind == K_USER && h.s.npass <= 0) FAIL(MHRE_EINVAL, "fit %d: npass must be positive", i);
    const MhreFitSpec& s = h.s;
    const size_t C = (size_t)h.nchain, P = (size_t)s.nprm, N = (size_t)s.nptot;
    h.o_vis = A.reserve<float>(h.vis.size());
    h.o_rho2 = A.reserve<float>(h.rho2.size());
    h.o_clo = A.reserve<int>(h.clo.size());
    h.o_clor = A.reserve<int>(h.clor.size());
    h.o_clow = A.reserve<float>(h.clow.size());
    h.o_tinfo = A.reserve<int>(h.tinfo.size());
    h.o_qcoef = A.reserve<float>(h.qcoef.size());
    h.o_qrow = A.reserve<float>(h.qrow.size());
    h.o_coor = A.reserve<float>(s.coor ? s.ncoor : 0);
    h.o_dat = A.reserve<float>(s.ndat);
    h.o_std2r = A.reserve<float>(s.ndat);
    h.o_icoor = A.reserve<int>(s.icoor ? s.nicoor : 0);
    h.o_idat = A.reserve<int>(s.idat ? s.nidat : 0);
    h.o_pdescr = A.reserve<int>(N);
    h.o_ivar = A.reserve<int>(P);
    h.o_invar = A.reserve<int>(N);
    h.o_pmint = A.reserve<float>(N);
    h.o_pmaxt = A.reserve<float>(N);
    h.o_beta = A.reserve<float>(s.nbeta);
    h.o_pcur = A.reserve<float>(C * N);
    h.o_chi2c = A.reserve<float>(C);
    h.o_pstp = A.reserve<float>(P * C);
    h.o_ncnt = A.reserve<int>(P * C);
    h.o_nacpt = A.reserve<int>(P * C);
    h.o_nhist = A.reserve<uint32_t>((size_t)hist_words(s.nadj) * P * C);
    h.o_nexch = A.reserve<int>(C);
    h.o_sumh = A.reserve<int>(P * C);
    h.o_tcur = A.reserve<int>(C);
    h.xfused = eht_xfused(h, S->opt);
    h.o_pcur_b = A.reserve<float>(h.xfused ? C * N : 0);
    h.o_chi2c_b = A.reserve<float>(h.xfused ? C : 0);
    h.o_tcur_b = A.reserve<int>(h.xfused ? C : 0);
    h.o_flag = A.reserve<int>(C);
    h.o_ptent = A.reserve<float>(C);
    h.o_ptentn = A.reserve<int>(C);
    h.record = s.pout != nullptr;
    const size_t nrec = h.record ? (size_t)s.niter : 0;
    h.o_pout = A.reserve<float>(P * C * nrec);
    h.o_tout = A.reserve<int>(C * nrec);
    h.o_chi2 = A.reserve<float>((size_t)s.nseq * nrec);
    h.has_datm = h.kind == K_USER || s.datm != nullptr;
    // the model data head (C * ndat) and, for the ring+Gaussian model.cuh,
    // its phase tail (C * nvis, src/imgpu.py:770-773)
    size_t need = C * (size_t)s.ndat;
    if (h.kind == K_USER && kNativeEht && (s.imodel == 1 || s.imodel == 2) && s.idat && s.nidat >= 1)
      need += C * (size_t)std::max(s.idat[0], 0);
    const size_t ndatm = h.has_datm ? std::max<size_t>((size_t)std::max(s.ndatm, 0), need) : 0;
    if (h.kind == K_USER && ndatm > (size_t)0x7fffffff)
      FAIL(MHRE_EINVAL, "fit %d: the model.cuh contract indexes datm with int; nbeta*nseq*(ndat + tail) = %zu "
                        "exceeds 2^31 - 1", i, ndatm);
    h.o_datm = A.reserve<float>(ndatm);
    h.o_chi2m = A.reserve<float>(h.has_datm ? C * s.ndat : 0);
    h.o_mc = A.reserve<CCalcModel>(1);
    h.d.ndatm = (long long)ndatm;
  }
  // all arguments are validated and packed: now the device resources
  HIPCHK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&S->ev0));
  HIPCHK(hipEventCreate(&S->ev1));
  S->arena_bytes = A.bytes;
  HIPCHK(hipMalloc((void**)&S->arena, A.bytes));
  CHK(dev_alloc(S, &S->d_fits, nfits));
  CHK(dev_alloc(S, &S->d_nevals, 2 * (size_t)nfits));
  CHK(dev_alloc(S, &S->d_err, 1));
  HIPCHK(hipMemsetAsync(S->d_nevals, 0, 2 * sizeof(unsigned long long) * nfits, S->stream));
  HIPCHK(hipMemsetAsync(S->d_err, 0, sizeof(int), S->stream));

  // decision-trace layout: each fit's [trace_iters][nprm][nchain] decisions and
  // [trace_iters][nbeta][nseq] exchanges, concatenated in fit order
  std::vector<long long> tr_dec_off(nfits, 0), tr_ex_off(nfits, 0);
  size_t tr_dec_total = 0, tr_ex_total = 0;
  if (S->opt.trace_iters > 0) {
    for (int i = 0; i < nfits; ++i) {
      const FitHost& h = S->fits[i];
      tr_dec_off[i] = (long long)tr_dec_total;
      tr_ex_off[i] = (long long)tr_ex_total;
      tr_dec_total += (size_t)S->opt.trace_iters * h.s.nprm * h.nchain;
      tr_ex_total += (size_t)S->opt.trace_iters * h.nchain;
    }
  }

  std::vector<FitDesc> descs(nfits);
  Uploads U;
  for (int i = 0; i < nfits; ++i) {
    FitHost& h = S->fits[i];
    const MhreFitSpec& s = h.s;
    const size_t C = (size_t)h.nchain, P = (size_t)s.nprm, N = (size_t)s.nptot;
    FitDesc& d = h.d;
    const long long ndatm = d.ndatm;
    d = FitDesc{};
    d.ndatm = ndatm;
    d.model = h.kind == K_SGRA9 ? MODEL_SGRA9 : h.kind == K_SGRA13 ? MODEL_SGRA13 : h.kind == K_QUAD ? MODEL_QUADSYS : MODEL_USER;
    d.imodel = s.imodel;
    d.nbeta = s.nbeta; d.nseq = s.nseq; d.nptot = s.nptot; d.nprm = s.nprm; d.nadj = s.nadj; d.npass = s.npass;
    d.ndat = s.ndat;
    d.nvis = h.nvis_k; d.ncph = h.ncph; d.ntile = h.ntile; d.ring_ok = h.ring_ok; d.cgroup = h.cgroup;
    d.nburn = s.nburn; d.niter = s.niter; d.nchain = h.nchain;
    d.key0 = (uint32_t)(s.seed & 0xffffffffull);
    d.key1 = (uint32_t)(s.seed >> 32);
    d.adapt_target = s.adapt_target > 0.0f ? s.adapt_target : 0.30f;
    rate_threshold(d.adapt_target, &d.adapt_mid, &d.adapt_tie);
    d.quad_n = h.quad_n; d.quad_neq = s.ndat;
    d.record = h.record ? 1 : 0;
    d.vis = at<float4>(S, h.o_vis); d.rho2 = at<float>(S, h.o_rho2);
    d.clo = at<int4>(S, h.o_clo); d.clor = at<int4>(S, h.o_clor); d.clow = at<float4>(S, h.o_clow);
    d.tinfo = at<int2>(S, h.o_tinfo); d.qcoef = at<float>(S, h.o_qcoef);
    d.qrow = h.qrow.empty() ? nullptr : at<float>(S, h.o_qrow);
    d.coor = at<float>(S, h.o_coor); d.dat = at<float>(S, h.o_dat); d.std2r = at<float>(S, h.o_std2r);
    d.icoor = at<int>(S, h.o_icoor); d.idat = at<int>(S, h.o_idat);
    d.pdescr = at<int>(S, h.o_pdescr); d.ivar = at<int>(S, h.o_ivar); d.invar = at<int>(S, h.o_invar);
    d.pmint = at<float>(S, h.o_pmint); d.pmaxt = at<float>(S, h.o_pmaxt); d.beta = at<float>(S, h.o_beta);
    d.pcur = at<float>(S, h.o_pcur); d.chi2c = at<float>(S, h.o_chi2c); d.pstp = at<float>(S, h.o_pstp);
    d.n_cnt = at<int>(S, h.o_ncnt); d.n_acpt = at<int>(S, h.o_nacpt); d.hbits = at<uint32_t>(S, h.o_nhist);
    d.hwords = hist_words(s.nadj);
    d.xlanes = exchange_lanes(s.nbeta);
    d.n_exch = at<int>(S, h.o_nexch); d.sumh = at<int>(S, h.o_sumh); d.tcur = at<int>(S, h.o_tcur);
    d.flag = at<int>(S, h.o_flag); d.ptent = at<float>(S, h.o_ptent); d.ptentn = at<int>(S, h.o_ptentn);
    d.xfused = h.xfused ? 1 : 0;
    d.hwphase = h.hwphase ? 1 : 0;
    if (h.xfused) {
      d.pcur_b = at<float>(S, h.o_pcur_b); d.chi2c_b = at<float>(S, h.o_chi2c_b); d.tcur_b = at<int>(S, h.o_tcur_b);
    }
    d.pout = at<float>(S, h.o_pout); d.tout = at<int>(S, h.o_tout); d.chi2 = at<float>(S, h.o_chi2);
    d.datm = h.has_datm ? at<float>(S, h.o_datm) : nullptr;
    d.chi2m = h.has_datm ? at<float>(S, h.o_chi2m) : nullptr;
    d.mc = at<CCalcModel>(S, h.o_mc);
    d.nevals = S->d_nevals + 2 * i;
    d.errflag = S->d_err;
    d.tr_dec = tr_dec_off[i];
    d.tr_ex = tr_ex_off[i];
    descs[i] = d;

    U.add(h.o_vis, h.vis.data(), h.vis.size());
    U.add(h.o_rho2, h.rho2.data(), h.rho2.size());
    U.add(h.o_clo, h.clo.data(), h.clo.size());
    U.add(h.o_clor, h.clor.data(), h.clor.size());
    U.add(h.o_clow, h.clow.data(), h.clow.size());
    U.add(h.o_tinfo, h.tinfo.data(), h.tinfo.size());
    U.add(h.o_qcoef, h.qcoef.data(), h.qcoef.size());
    U.add(h.o_qrow, h.qrow.data(), h.qrow.size());
    U.add(h.o_coor, s.coor, s.coor ? s.ncoor : 0);
    U.add(h.o_dat, s.dat, s.ndat);
    U.add(h.o_std2r, s.std2r, s.ndat);
    U.add(h.o_icoor, s.icoor, s.icoor ? s.nicoor : 0);
    U.add(h.o_idat, s.idat, s.idat ? s.nidat : 0);
    U.add(h.o_pdescr, s.pdescr, N);
    U.add(h.o_ivar, s.ivar, P);
    U.add(h.o_invar, s.invar, N);
    U.add(h.o_pmint, s.pmint, N);
    U.add(h.o_pmaxt, s.pmaxt, N);
    U.add(h.o_beta, s.beta, s.nbeta);
    U.add(h.o_pcur, s.pcur, C * N);
    U.add(h.o_chi2c, s.chi2c, C);
    U.add(h.o_pstp, s.pstp, P * C);
    U.add(h.o_ncnt, s.n_cnt, P * C);
    U.add(h.o_nacpt, s.n_acpt, P * C);
    // the history as bits and the running accept sums over its nadj entries
    // (the device keeps sumh[k][chain] equal to the set bits of the chain's ring)
    pack_hist(s.n_hist, s.nadj, P, C, h.hbits, h.sumh);
    U.add(h.o_nhist, h.hbits.data(), h.hbits.size());
    U.add(h.o_nexch, s.n_exch, C);
    U.add(h.o_tcur, s.tcur, C);
    U.add(h.o_flag, s.flag, C);
    if (s.ptent) U.add(h.o_ptent, s.ptent, C);
    if (s.ptentn) U.add(h.o_ptentn, s.ptentn, C);
    if (h.has_datm && s.datm) U.add(h.o_datm, s.datm, (size_t)std::min<long long>(d.ndatm, s.ndatm));
    if (h.has_datm && s.chi2m) U.add(h.o_chi2m, s.chi2m, C * s.ndat);
    U.add(h.o_sumh, h.sumh.data(), h.sumh.size());
    if (h.record && it0 > s.nburn) {
      // resumed inside the sampling stage: the samples recorded before it0
      // come back unchanged with the download
      U.add(h.o_pout, s.pout, P * C * (size_t)s.niter);
      U.add(h.o_tout, s.tout, C * (size_t)s.niter);
      U.add(h.o_chi2, s.chi2, (size_t)s.nseq * s.niter);
    }
    CCalcModel mc{};
    mc.imodel = s.imodel; mc.ncoor = s.ncoor; mc.ndat = s.ndat; mc.nicoor = s.nicoor; mc.nidat = s.nidat;
    mc.ndatm = (int)std::min<long long>(d.ndatm, 0x7fffffff); mc.nptot = s.nptot; mc.nbeta = s.nbeta; mc.nseq = s.nseq;
    mc.nchi2m = (int)(C * s.ndat);
    mc.coor = const_cast<float*>(d.coor); mc.dat = const_cast<float*>(d.dat); mc.std2r = const_cast<float*>(d.std2r);
    mc.icoor = const_cast<int*>(d.icoor); mc.idat = const_cast<int*>(d.idat);
    mc.datm = d.datm; mc.chi2m = d.chi2m; mc.pcur = d.pcur; mc.flag = d.flag; mc.chi2c = d.chi2c;
    U.add(h.o_mc, &mc, 1);
  }
  CHK(U.flush(S));
  HIPCHK(hipMemcpyAsync(S->d_fits, descs.data(), sizeof(FitDesc) * nfits, hipMemcpyHostToDevice, S->stream));

  // work lists: fits grouped per kind, heaviest fits first
  std::vector<WorkItem> init;
  const int kinds_order[4] = {K_SGRA9, K_SGRA13, K_QUAD, K_USER};
  // per-chain step cost ~ data terms evaluated per chi^2
  std::vector<double> fit_cost(nfits);
  for (int i = 0; i < nfits; ++i) fit_cost[i] = (double)S->fits[i].nvis_k + S->fits[i].s.ndat + 1.0;
  std::vector<std::vector<int>> by_kind(4);
  for (int kk : kinds_order) {
    std::vector<int>& idx = by_kind[kk];
    for (int i = 0; i < nfits; ++i)
      if (S->fits[i].kind == kk) idx.push_back(i);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
      const FitHost &x = S->fits[a], &y = S->fits[b];
      return (long long)x.nvis_k + x.s.ndat > (long long)y.nvis_k + y.s.ndat;
    });
  }
  // evaluation lists: every chain, per kind (per system size for K_QUAD)
  for (int kk : kinds_order) {
    const std::vector<int>& idx = by_kind[kk];
    if (idx.empty()) continue;
    for (int n : (kk == K_QUAD ? std::vector<int>{2, 4, 8} : std::vector<int>{0})) {
      KindWork kw;
      kw.kind = kk;
      kw.quad_n = n;
      for (int i : idx) {
        if (kk == K_QUAD && S->fits[i].quad_n != n) continue;
        for (int c = 0; c < S->fits[i].nchain; ++c) kw.eval.push_back({i, c});
      }
      if (!kw.eval.empty()) S->kinds.push_back(std::move(kw));
    }
  }
  // Sequence groups: each fit's sequences are cut into G contiguous ranges
  // at multiples of its exchange wave's sequence count (64 / xlanes), so no
  // exchange wave and no chain is shared between groups.
  // At most as many groups as the largest fit has exchange waves, so every
  // group holds sequences (its step launches hand its clock to its closing
  // launch, and the download reads group 0's clock).
  int max_waves = 1;
  for (const FitHost& h : S->fits) {
    const int per = 64 / exchange_lanes(h.s.nbeta);
    max_waves = std::max(max_waves, (h.s.nseq + per - 1) / per);
  }
  // Default: two groups in the grid form (one group's waves fill the compute
  // units that the other's early-finishing waves free, and its exchange runs
  // under the other's step launch), one in the persistent form (whose grid
  // already fills the chip and would make a second group wait).
  const int want = S->opt.streams > 0 ? S->opt.streams : (S->persistent ? 1 : 2);
  const int G = std::max(1, std::min({want, MAX_GROUPS, max_waves}));
  auto seq_range = [&](int i, int g, int* s0, int* s1) {
    const int nseq = S->fits[i].s.nseq;
    const int per = 64 / exchange_lanes(S->fits[i].s.nbeta);
    const int nblk = (nseq + per - 1) / per;
    *s0 = std::min(nseq, (int)((long long)nblk * g / G) * per);
    *s1 = std::min(nseq, (int)((long long)nblk * (g + 1) / G) * per);
  };
  std::vector<double> seq_cost(nfits);
  for (int i = 0; i < nfits; ++i) {
    const FitHost& h = S->fits[i];
    seq_cost[i] = (double)h.s.nbeta * (64 / exchange_lanes(h.s.nbeta));   // per wave
    if (h.s.init_chains)
      for (int c = 0; c < h.nchain; ++c) init.push_back({i, c});
  }
  std::vector<int> xorder(nfits);
  for (int i = 0; i < nfits; ++i) xorder[i] = i;
  std::stable_sort(xorder.begin(), xorder.end(), [&](int a, int b) { return seq_cost[a] > seq_cost[b]; });
  S->groups.resize(G);
  for (int g = 0; g < G; ++g) {
    Group& grp = S->groups[g];
    int qslot = 0;
    // chains of the group's sequences, ibeta-major as in the state arrays
    auto chains = [&](int i, std::vector<WorkItem>& out) {
      int s0, s1;
      seq_range(i, g, &s0, &s1);
      const FitHost& h = S->fits[i];
      for (int b = 0; b < h.s.nbeta; ++b)
        for (int q = s0; q < s1; ++q) out.push_back({i, b * h.s.nseq + q});
    };
    for (int kk : kinds_order) {
      const std::vector<int>& idx = by_kind[kk];
      if (idx.empty()) continue;
      if (kk == K_QUAD) {
        // one kernel per compiled system size; ladders of at most 64 rungs
        // run step and exchange in one launch (one item per wave: the fit
        // and its first sequence), longer ones lane per chain
        // (fused fits with update rows and those without take kernels of
        // their own: the rows form holds every equation's residual in registers)
        for (int n : {2, 4, 8}) {
          KindWork kr, kf, kw;
          kr.kind = kf.kind = kw.kind = kk;
          kr.quad_n = kf.quad_n = kw.quad_n = n;
          kr.fused = kf.fused = true;
          kr.rows = true;
          for (int i : idx) {
            if (S->fits[i].quad_n != n) continue;
            if (quad_fused(S->fits[i])) {
              const int per = 64 / exchange_lanes(S->fits[i].s.nbeta);
              int s0, s1;
              seq_range(i, g, &s0, &s1);
              KindWork& k = S->fits[i].qrow.empty() ? kf : kr;
              for (int q = s0; q < s1; q += per) k.step.push_back({i, q});
              continue;
            }
            const size_t before = kw.step.size();
            chains(i, kw.step);
            if (kw.step.size() > before)
              while (kw.step.size() % 64) kw.step.push_back({i, -1});
          }
          if (!kr.step.empty()) grp.kinds.push_back(std::move(kr));
          if (!kf.step.empty()) grp.kinds.push_back(std::move(kf));
          if (!kw.step.empty()) grp.kinds.push_back(std::move(kw));
        }
        continue;
      }
      KindWork kw;
      kw.kind = kk;
      kw.qslot = qslot++;
      for (int i : idx) {
        const size_t before = kw.step.size();
        chains(i, kw.step);
        // lane-per-chain kernels: every wave holds chains of one fit
        if (kk == K_USER && kw.step.size() > before)
          while (kw.step.size() % 64) kw.step.push_back({i, -1});
      }
      if (kw.step.empty()) continue;
      if (kk != K_USER) kw.xr = deal_xcd(kw.step, fit_cost);
      grp.kinds.push_back(std::move(kw));
    }
    // exchange lists, heaviest fits first, dealt like the step lists; one
    // item per wave: its first sequence (64 / xlanes sequences per wave).
    // xw: sequences exchanged by a launch of their own every iteration; xc:
    // xfused sequences (their segment's closing launch finishes them).
    std::vector<WorkItem> xw, xc;
    for (int i : xorder) {
      if (quad_fused(S->fits[i])) continue;   // exchanged inside its step launch
      const int per = 64 / exchange_lanes(S->fits[i].s.nbeta);
      int s0, s1;
      seq_range(i, g, &s0, &s1);
      for (int q = s0; q < s1; q += per) (S->fits[i].xfused ? xc : xw).push_back({i, q});
    }
    grp.xr_x = deal_xcd(xw, seq_cost);
    grp.n_x = (int)xw.size();
    grp.xr_xc = deal_xcd(xc, seq_cost);
    CHK(dev_alloc(S, &grp.d_x, xw.size()));
    CHK(dev_alloc(S, &grp.d_xc, xc.size()));
    if (!xw.empty())
      HIPCHK(hipMemcpyAsync(grp.d_x, xw.data(), xw.size() * sizeof(WorkItem), hipMemcpyHostToDevice, S->stream));
    if (!xc.empty())
      HIPCHK(hipMemcpyAsync(grp.d_xc, xc.data(), xc.size() * sizeof(WorkItem), hipMemcpyHostToDevice, S->stream));
    for (auto& k : grp.kinds) {
      CHK(dev_alloc(S, &k.d_step, k.step.size()));
      HIPCHK(hipMemcpyAsync(k.d_step, k.step.data(), k.step.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                            S->stream));
    }
    CHK(dev_alloc(S, &grp.d_clk, 1));
    RunClock clk{};
    clk.it = it0;
    clk.xit = it0;
    HIPCHK(hipMemcpyAsync(grp.d_clk, &clk, sizeof(clk), hipMemcpyHostToDevice, S->stream));
    if (g == 0) {
      grp.stream = S->stream;
    } else {
      HIPCHK(hipStreamCreateWithFlags(&grp.stream, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&grp.done, hipEventDisableTiming));
    }
  }
  for (auto& k : S->kinds) {
    CHK(dev_alloc(S, &k.d_eval, k.eval.size()));
    HIPCHK(hipMemcpyAsync(k.d_eval, k.eval.data(), k.eval.size() * sizeof(WorkItem), hipMemcpyHostToDevice, S->stream));
  }
  S->n_init = (int)init.size();
  CHK(dev_alloc(S, &S->d_init, init.size()));
  if (!init.empty())
    HIPCHK(hipMemcpyAsync(S->d_init, init.data(), init.size() * sizeof(WorkItem), hipMemcpyHostToDevice, S->stream));

  // persistent grid: a few work-groups per CU
  int dev = 0, ncu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  S->ncu = ncu;
  S->grid = std::max(1, ncu * (16 / S->wpb));
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || lds <= 0)
    lds = 160 << 10;
  S->lds_per_cu = lds;

  // decision trace of iterations it0 .. it0 + trace_iters - 1
  if (S->opt.trace_iters > 0) {
    S->tr_nit = S->opt.trace_iters;
    S->tr_it0 = it0;
    S->tr_dec_total = tr_dec_total;
    S->tr_ex_total = tr_ex_total;
    CHK(dev_alloc(S, &S->d_tr_dec, tr_dec_total));
    CHK(dev_alloc(S, &S->d_tr_chi2, tr_dec_total));
    CHK(dev_alloc(S, &S->d_tr_ex, tr_ex_total));
    HIPCHK(hipMemsetAsync(S->d_tr_dec, 0xff, tr_dec_total * sizeof(int), S->stream));
    HIPCHK(hipMemsetAsync(S->d_tr_chi2, 0, tr_dec_total * sizeof(float), S->stream));
    HIPCHK(hipMemsetAsync(S->d_tr_ex, 0xff, tr_ex_total * sizeof(int), S->stream));
  }

  if (!eval_only) {
    // init_pcur + initial chi^2 (src/gpu_mcmc.cu:714-719)
    HIPCHK(launch_init(S->d_fits, S->d_init, S->n_init, S->stream));
    if (S->n_init > 0) {
      for (auto& k : S->kinds) {
        std::vector<WorkItem> ev;
        for (const WorkItem& w : k.eval)
          if (S->fits[w.fit].s.init_chains) ev.push_back(w);
        if (ev.empty()) continue;
        WorkItem* dev_ev = nullptr;
        CHK(dev_alloc(S, &dev_ev, ev.size()));
        HIPCHK(hipMemcpyAsync(dev_ev, ev.data(), ev.size() * sizeof(WorkItem), hipMemcpyHostToDevice, S->stream));
        const int n = (int)ev.size();
        hipError_t e;
        if (k.kind == K_SGRA9) e = launch_eval_eht(MODEL_SGRA9, S->d_fits, dev_ev, n, 0, 0, S->stream);
        else if (k.kind == K_SGRA13) e = launch_eval_eht(MODEL_SGRA13, S->d_fits, dev_ev, n, 0, 0, S->stream);
        else if (k.kind == K_QUAD) e = launch_eval_quad(k.quad_n, S->d_fits, dev_ev, n, S->stream);
        else e = launch_eval_user(S->d_fits, dev_ev, n, 0, 0, S->stream);
        HIPCHK(e);
        HIPCHK(hipStreamSynchronize(S->stream));
      }
    }
    HIPCHK(hipStreamSynchronize(S->stream));
    CHK(check_model_error(S));
  }
  HIPCHK(hipStreamSynchronize(S->stream));
  *out = S;
  S = nullptr;   // released from the guard
  return MHRE_OK;
}

int run(Session* S, int niter) {
  if (niter <= 0) return MHRE_OK;
  (void)hipGetLastError();
  HIPCHK(hipEventRecord(S->ev0, S->stream));
  CHK(fork_groups(S, S->ev0));
  int left = niter;
  const int G = S->opt.graph_iters;
  if (G > 0 && niter >= G) {
    // one graph per group: a segment of G iterations and its closing launch,
    // captured on the group's stream
    for (Group& grp : S->groups) {
      if (grp.gexec) continue;
      HIPCHK(hipStreamBeginCapture(grp.stream, hipStreamCaptureModeThreadLocal));
      int rc = MHRE_OK;
      for (int i = 0; i < G && rc == MHRE_OK; ++i) rc = launch_iteration(S, grp, i);
      if (rc == MHRE_OK) rc = close_segment(S, grp, G);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(grp.stream, &g);
      if (rc != MHRE_OK || ec != hipSuccess) {
        if (g) hipGraphDestroy(g);
        if (rc != MHRE_OK) return rc;
        HIPCHK(ec);
      }
      grp.graph = g;
      HIPCHK(hipGraphInstantiate(&grp.gexec, g, nullptr, nullptr, 0));
    }
    while (left >= G) {
      for (Group& grp : S->groups) HIPCHK(hipGraphLaunch(grp.gexec, grp.stream));
      left -= G;
    }
  }
  if (left > 0) CHK(launch_segment(S, left));
  CHK(join_groups(S));
  HIPCHK(hipEventRecord(S->ev1, S->stream));
  S->iters_done += niter;
  return MHRE_OK;
}

int download(Session* S, int model_data) {
  HIPCHK(hipStreamSynchronize(S->stream));
  Downloads D;
  for (FitHost& h : S->fits) {
    const MhreFitSpec& s = h.s;
    const size_t C = (size_t)h.nchain, P = (size_t)s.nprm, N = (size_t)s.nptot;
    D.add(s.pcur, h.o_pcur, C * N);
    D.add(s.chi2c, h.o_chi2c, C);
    D.add(s.pstp, h.o_pstp, P * C);
    D.add(s.n_cnt, h.o_ncnt, P * C);
    D.add(s.n_acpt, h.o_nacpt, P * C);
    h.hbits.resize((size_t)hist_words(s.nadj) * P * C);
    D.add(h.hbits.data(), h.o_nhist, h.hbits.size());
    D.add(s.n_exch, h.o_nexch, C);
    D.add(s.tcur, h.o_tcur, C);
    D.add(s.flag, h.o_flag, C);
    D.add(s.ptent, h.o_ptent, C);
    D.add(s.ptentn, h.o_ptentn, C);
    if (h.record) {
      D.add(s.pout, h.o_pout, P * C * (size_t)s.niter);
      D.add(s.tout, h.o_tout, C * (size_t)s.niter);
      D.add(s.chi2, h.o_chi2, (size_t)s.nseq * s.niter);
    }
  }
  CHK(D.flush(S));
  for (FitHost& h : S->fits)
    unpack_hist(h.hbits, h.s.nadj, h.s.nprm, h.nchain, h.s.n_hist);
  if (model_data) {
    // model data and chi^2 terms of each chain's final state; chi2c is
    // recomputed identically and already copied out above
    std::vector<float> keep;
    CHK(launch_eval_all(S, 1, 0));
    HIPCHK(hipStreamSynchronize(S->stream));
    for (FitHost& h : S->fits) {
      const MhreFitSpec& s = h.s;
      if (!h.has_datm) continue;
      const size_t C = (size_t)h.nchain;
      if (s.datm) D.add(s.datm, h.o_datm, (size_t)std::min<long long>(h.d.ndatm, s.ndatm));
      if (s.chi2m) D.add(s.chi2m, h.o_chi2m, C * s.ndat);
    }
    CHK(D.flush(S));
  }
  // resume records in rndst: seed (8 B), next iteration (4 B), chain (4 B)
  int it = 0;
  HIPCHK(hipMemcpyAsync(&it, &S->groups[0].d_clk->it, sizeof(int), hipMemcpyDeviceToHost, S->stream));
  HIPCHK(hipStreamSynchronize(S->stream));
  for (FitHost& h : S->fits) {
    if (!h.s.rndst) continue;
    for (int c = 0; c < h.nchain; ++c) {
      unsigned char* r = h.s.rndst + (size_t)c * 48;
      memset(r, 0, 48);
      memcpy(r, &h.s.seed, 8);
      memcpy(r + 8, &it, 4);
      memcpy(r + 12, &c, 4);
    }
  }
  if (S->d_tr_dec && S->opt.trace_dec) {
    HIPCHK(hipMemcpy(S->opt.trace_dec, S->d_tr_dec, S->tr_dec_total * sizeof(int), hipMemcpyDeviceToHost));
    if (S->opt.trace_chi2)
      HIPCHK(hipMemcpy(S->opt.trace_chi2, S->d_tr_chi2, S->tr_dec_total * sizeof(float), hipMemcpyDeviceToHost));
    if (S->opt.trace_ex)
      HIPCHK(hipMemcpy(S->opt.trace_ex, S->d_tr_ex, S->tr_ex_total * sizeof(int), hipMemcpyDeviceToHost));
  }
  return check_model_error(S);
}

int stats(Session* S, MhreStats* st) {
  if (!st) FAIL(MHRE_EINVAL, "stats pointer is NULL");
  HIPCHK(hipStreamSynchronize(S->stream));
  memset(st, 0, sizeof(*st));
  float ms = 0.0f;
  if (S->iters_done > 0 && hipEventElapsedTime(&ms, S->ev0, S->ev1) == hipSuccess) st->run_ms = ms;
  std::vector<unsigned long long> ev(2 * S->fits.size());
  HIPCHK(hipMemcpy(ev.data(), S->d_nevals, ev.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < S->fits.size(); ++i) {
    const FitHost& h = S->fits[i];
    st->chi2_evals += ev[2 * i];
    st->proposals += ev[2 * i + 1];
    if (h.kind == K_SGRA9 || h.kind == K_SGRA13) {
      st->vis_evals += (double)ev[2 * i] * h.nvis_k;
      st->cph_evals += (double)ev[2 * i] * h.ncph;
    } else if (h.kind == K_USER && (h.s.imodel == 1 || h.s.imodel == 2) && h.s.idat && h.s.nidat >= 1) {
      // the EHT model.cuh on the generic path: idat = [nvis, ndatm, cpExt]
      const int nv = h.s.idat[0], nc = h.s.ndat - nv;
      const int cpext = h.s.nidat >= 3 ? h.s.idat[2] : 0;
      st->vis_evals += (double)ev[2 * i] * (nv + (cpext ? 3 * nc : 0));
      st->cph_evals += (double)ev[2 * i] * nc;
    }
  }
  st->iters_done = S->iters_done;
  st->nfits = (int)S->fits.size();
  return MHRE_OK;
}

MhreFitSpec legacy_spec(float* coor, float* dat, float* std2r, int* icoor, int* idat, float* datm, float* chi2m,
                        int* pdescr, int* ivar, int* invar, float* ptotal, float* pmint, float* pmaxt, float* pout,
                        int* tout, float* chi2, float* chi2c, int* tcur, int* flag, int* n_acpt, int* n_exch,
                        int* n_cnt, int* n_hist, float* beta, float* pstp, float* ptent, int* ptentn, float* pcur,
                        unsigned char* rndst, unsigned long long seed, int imodel, int ncoor, int ndat, int nicoor,
                        int nidat, int ndatm, int nptot, int nprm, int nadj, int npass, int nbeta, int nseq,
                        int nburn, int niter) {
  MhreFitSpec s{};
  s.model = MHRE_MODEL_AUTO;
  s.imodel = imodel;
  s.coor = coor; s.dat = dat; s.std2r = std2r; s.icoor = icoor; s.idat = idat; s.datm = datm; s.chi2m = chi2m;
  s.pdescr = pdescr; s.ivar = ivar; s.invar = invar; s.ptotal = ptotal; s.pmint = pmint; s.pmaxt = pmaxt;
  s.pout = pout; s.tout = tout; s.chi2 = chi2; s.chi2c = chi2c; s.tcur = tcur; s.flag = flag;
  s.n_acpt = n_acpt; s.n_exch = n_exch; s.n_cnt = n_cnt; s.n_hist = n_hist;
  s.beta = beta; s.pstp = pstp; s.ptent = ptent; s.ptentn = ptentn; s.pcur = pcur; s.rndst = rndst;
  s.seed = seed;
  s.ncoor = ncoor; s.ndat = ndat; s.nicoor = nicoor; s.nidat = nidat; s.ndatm = ndatm;
  s.nptot = nptot; s.nprm = nprm; s.nadj = nadj; s.npass = npass; s.nbeta = nbeta; s.nseq = nseq;
  s.nburn = nburn; s.niter = niter;
  s.adapt_target = 0.30f;
  s.init_chains = 1;
  s.it0 = 0;
  return s;
}

// ----------------------------------------------- packed data for gradients

// MHRE_GRAD_RING=0 selects the per-closure gradient kernel for every fit
// (both forms compute the same gradient; the test suite compares them).
static bool ring_grad_disabled() {
  const char* e = std::getenv("MHRE_GRAD_RING");
  return e && e[0] == '0';
}

struct EhtData {
  FitHost h;
  char* dev = nullptr;
  FitDesc* d_fd = nullptr;
};

void destroy_eht_data(EhtData* E) {
  if (!E) return;
  if (E->dev) hipFree(E->dev);
  delete E;
}

// The ring gradient kernel adds each closure's adjoint to its three
// visibilities' slots in LDS.  Float atomics there are slow, so the legs of
// every 64-closure batch (the kernel's batches: cgroup tiles' closures, 64 at
// a time) are split into rounds in which no two legs name the same
// visibility; a round is then a plain read-add-write.  The rounds are a
// proper edge colouring of the bipartite multigraph lanes x visibilities,
// built with Koenig's alternating-path step, so a batch needs only as many
// rounds as its largest degree (3, or the most closures sharing one
// visibility).  More than 16 rounds leaves the data to the per-closure kernel.
static void color_closure_legs(FitHost& h) {
  constexpr int MAXR = 16;
  const int ncph = h.ncph;
  h.clr.assign((size_t)ncph + TILE, 0);
  h.nround = 0;
  h.ring_grad = false;
  if (!h.ring_ok || ncph == 0) return;
  std::vector<int> at, color, vnode;
  std::vector<int> vis_ids;
  for (int t0 = 0, c0 = 0; t0 < h.ntile; t0 += h.cgroup) {
    const int t1 = std::min(t0 + h.cgroup, h.ntile);
    const int c1 = h.tinfo[2 * t1];
    for (int b0 = c0; b0 < c1; b0 += TILE) {
      const int nb = std::min(TILE, c1 - b0);
      // right-hand nodes: the batch's distinct visibilities, after the TILE lanes
      vis_ids.clear();
      vnode.assign(3 * nb, 0);
      for (int e = 0; e < 3 * nb; ++e) {
        const int v = h.clo[4 * (b0 + e / 3) + e % 3];
        int k = (int)(std::find(vis_ids.begin(), vis_ids.end(), v) - vis_ids.begin());
        if (k == (int)vis_ids.size()) vis_ids.push_back(v);
        vnode[e] = TILE + k;
      }
      const int nnode = TILE + (int)vis_ids.size();
      at.assign((size_t)nnode * MAXR, -1);   // at[node * MAXR + c]: edge of colour c there
      color.assign(3 * nb, -1);
      auto other = [&](int e, int x) { return x == e / 3 ? vnode[e] : e / 3; };
      auto free_col = [&](int x) {
        int c = 0;
        while (c < MAXR && at[(size_t)x * MAXR + c] >= 0) ++c;
        return c;
      };
      for (int e = 0; e < 3 * nb; ++e) {
        const int u = e / 3, v = vnode[e];
        const int a = free_col(u), b = free_col(v);
        if (a == MAXR || b == MAXR) return;   // degree above MAXR
        if (at[(size_t)v * MAXR + a] >= 0) {
          // swap a <-> b along the path from v that alternates a, b; in a
          // bipartite graph it cannot reach u, so a stays free at u
          std::vector<int> path;
          int x = v, c = a;
          while (at[(size_t)x * MAXR + c] >= 0) {
            const int f = at[(size_t)x * MAXR + c];
            path.push_back(f);
            x = other(f, x);
            c = c == a ? b : a;
          }
          for (int f : path) {
            at[(size_t)(f / 3) * MAXR + color[f]] = -1;
            at[(size_t)vnode[f] * MAXR + color[f]] = -1;
          }
          for (int f : path) {
            color[f] = color[f] == a ? b : a;
            at[(size_t)(f / 3) * MAXR + color[f]] = f;
            at[(size_t)vnode[f] * MAXR + color[f]] = f;
          }
        }
        color[e] = a;
        at[(size_t)u * MAXR + a] = e;
        at[(size_t)v * MAXR + a] = e;
      }
      for (int j = 0; j < nb; ++j) {
        h.clr[b0 + j] = color[3 * j] | color[3 * j + 1] << 4 | color[3 * j + 2] << 8;
        for (int m = 0; m < 3; ++m) h.nround = std::max(h.nround, color[3 * j + m] + 1);
      }
    }
    c0 = c1;
  }
  h.ring_grad = true;
}

int eht_data_create(const float* coor, const float* dat, const float* std2r, const int* icoor, const int* idat,
                    int imodel, int ncoor, int ndat, int nicoor, int nidat, int nptot, EhtData** out) {
  *out = nullptr;
  (void)hipGetLastError();
  if (imodel != 1) FAIL(MHRE_EINVAL, "gradient kernels implement the 9-parameter model (imodel 1) only");
  if (!coor || !dat || !std2r || !idat) FAIL(MHRE_EINVAL, "coor, dat, std2r and idat are required");
  if (ndat <= 0) FAIL(MHRE_EINVAL, "ndat must be positive");
  if (nptot < 9 || nptot > MAXP) FAIL(MHRE_EINVAL, "nptot = %d not in [9, %d]", nptot, MAXP);
  EhtData* E = new EhtData();
  struct Guard {
    EhtData*& e;
    ~Guard() { if (e) destroy_eht_data(e); }
  } guard{E};
  FitHost& h = E->h;
  MhreFitSpec& s = h.s;
  s.model = MHRE_MODEL_SGRA9;
  s.imodel = imodel;
  s.coor = const_cast<float*>(coor); s.dat = const_cast<float*>(dat); s.std2r = const_cast<float*>(std2r);
  s.icoor = const_cast<int*>(icoor); s.idat = const_cast<int*>(idat);
  s.ncoor = ncoor; s.ndat = ndat; s.nicoor = nicoor; s.nidat = nidat; s.nptot = nptot;
  h.kind = K_SGRA9;
  CHK(pack_eht(h, 0));
  color_closure_legs(h);
  Arena A;
  const size_t o_vis = A.reserve<float>(h.vis.size());
  const size_t o_clo = A.reserve<int>(h.clo.size());
  const size_t o_clow = A.reserve<float>(h.clow.size());
  const size_t o_tinfo = A.reserve<int>(h.tinfo.size());
  const size_t o_clr = A.reserve<int>(h.clr.size());
  const size_t o_fd = A.reserve<FitDesc>(1);
  HIPCHK(hipMalloc((void**)&E->dev, A.bytes));
  FitDesc d{};
  d.model = MODEL_SGRA9;
  d.imodel = imodel;
  d.nptot = nptot;
  d.ndat = ndat;
  d.nvis = h.nvis_k; d.ncph = h.ncph; d.ntile = h.ntile; d.ring_ok = h.ring_ok; d.cgroup = h.cgroup;
  d.vis = reinterpret_cast<const float4*>(E->dev + o_vis);
  d.clo = reinterpret_cast<const int4*>(E->dev + o_clo);
  d.clow = reinterpret_cast<const float4*>(E->dev + o_clow);
  d.tinfo = reinterpret_cast<const int2*>(E->dev + o_tinfo);
  d.clr = reinterpret_cast<const int*>(E->dev + o_clr);
  d.nround = h.nround;
  E->d_fd = reinterpret_cast<FitDesc*>(E->dev + o_fd);
  HIPCHK(hipMemcpy(E->dev + o_vis, h.vis.data(), h.vis.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(E->dev + o_clo, h.clo.data(), h.clo.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(E->dev + o_clow, h.clow.data(), h.clow.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(E->dev + o_tinfo, h.tinfo.data(), h.tinfo.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(E->dev + o_clr, h.clr.data(), h.clr.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(E->d_fd, &d, sizeof(d), hipMemcpyHostToDevice));
  h.d = d;
  // the caller's arrays are not referenced after this point
  s.coor = nullptr; s.dat = nullptr; s.std2r = nullptr; s.icoor = nullptr; s.idat = nullptr;
  *out = E;
  E = nullptr;
  return MHRE_OK;
}

// ------------------------------------------ quadratic-system Jacobians

// Device layout of one mhre_quad_jacobian call: the host inputs staged in
// device memory and the outputs, in one block.
struct QuadJacLayout {
  size_t nq, nb, nd, nx, nr, nj, rows_sys;
  size_t oQ, ob, od, os, oX, orr, oJ, bytes;
  int nsys, neq, n, rows;
  bool has_sys;
};

int quad_jac_layout(const float* Q, const float* b, const float* dat, const int* sys, int nsys, int neq, int n,
                    const float* X, int rows, float* r, float* J, QuadJacLayout* L) {
  if (nsys <= 0 || neq <= 0 || n <= 0 || n > MAXP || rows < 0) FAIL(MHRE_EINVAL, "mhre_quad_jacobian: bad sizes");
  if (!Q || !b || !dat || !X || !r || !J) FAIL(MHRE_EINVAL, "mhre_quad_jacobian: null array");
  if (!sys && rows > nsys) FAIL(MHRE_EINVAL, "mhre_quad_jacobian: rows %d > nsys %d without a row map", rows, nsys);
  if (sys)
    for (int i = 0; i < rows; ++i)
      if (sys[i] < 0 || sys[i] >= nsys) FAIL(MHRE_EINVAL, "mhre_quad_jacobian: sys[%d] = %d out of range", i, sys[i]);
  L->nsys = nsys; L->neq = neq; L->n = n; L->rows = rows; L->has_sys = sys != nullptr;
  L->nq = (size_t)nsys * neq * n * n; L->nb = (size_t)nsys * neq * n; L->nd = (size_t)nsys * neq;
  L->nx = (size_t)rows * n; L->nr = (size_t)rows * neq; L->nj = L->nr * n;
  Arena A;
  L->oQ = A.reserve<float>(L->nq); L->ob = A.reserve<float>(L->nb); L->od = A.reserve<float>(L->nd);
  L->os = A.reserve<int>(sys ? rows : 0); L->oX = A.reserve<float>(L->nx);
  L->orr = A.reserve<float>(L->nr); L->oJ = A.reserve<float>(L->nj);
  L->bytes = A.bytes;
  return MHRE_OK;
}

// Stages the inputs into dev (L.bytes long), runs the kernel on the null
// stream and copies r and J back (synchronous).
int quad_jac_staged(char* dev, const QuadJacLayout& L, const float* Q, const float* b, const float* dat,
                    const int* sys, const float* X, float* r, float* J) {
  HIPCHK(hipMemcpy(dev + L.oQ, Q, L.nq * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dev + L.ob, b, L.nb * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dev + L.od, dat, L.nd * sizeof(float), hipMemcpyHostToDevice));
  if (sys) HIPCHK(hipMemcpy(dev + L.os, sys, (size_t)L.rows * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dev + L.oX, X, L.nx * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(launch_quad_jac(reinterpret_cast<float*>(dev + L.oQ), reinterpret_cast<float*>(dev + L.ob),
                         reinterpret_cast<float*>(dev + L.od), sys ? reinterpret_cast<int*>(dev + L.os) : nullptr,
                         reinterpret_cast<float*>(dev + L.oX), L.rows, L.neq, L.n,
                         reinterpret_cast<float*>(dev + L.orr), reinterpret_cast<float*>(dev + L.oJ), nullptr));
  HIPCHK(hipMemcpy(r, dev + L.orr, L.nr * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(J, dev + L.oJ, L.nj * sizeof(float), hipMemcpyDeviceToHost));
  return MHRE_OK;
}

// ------------------------------------------------ static scratch area

// The staging area of the entry points that allocate nothing: a file-scope
// device array, one copy on every GPU the library runs on, held until the
// process ends.  Host code never uses the array's name as an address (on the
// host it is a stand-in); it uses the device address that hipGetSymbolAddress
// returns for the current device, looked up once per device.  Calls that
// overlap on one GPU share the array and race.
constexpr size_t kScratchBytes = (size_t)32 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

int scratch_on_current_device(char** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) FAIL(MHRE_EINVAL, "device %d beyond the %d devices the scratch table holds", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return MHRE_OK;
}

int eht_check_batch(const EhtData* E, const float* P, int n, int ldp, const double* chi2) {
  if (!E) FAIL(MHRE_EINVAL, "null data handle");
  if (n < 0) FAIL(MHRE_EINVAL, "n = %d is negative", n);
  if (n > 0 && (!P || !chi2)) FAIL(MHRE_EINVAL, "params and chi2 are required");
  if (ldp < 9 || ldp > MAXP) FAIL(MHRE_EINVAL, "ldp = %d not in [9, %d]", ldp, MAXP);
  return MHRE_OK;
}

}  // namespace

// ===================================================================== C ABI

extern "C" {

const char* mhre_last_error(void) { return g_err.c_str(); }

void mhre_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]) {
  const U4 r = philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
int mhre_version(void) { return 10000; }
const char* mhre_user_model(void) { return user_model_name(); }
int mhre_native_eht(void) { return kNativeEht; }

int mhre_session_create(const MhreFitSpec* fits, int nfits, const MhreOptions* opt, void** session) {
  if (!session) FAIL(MHRE_EINVAL, "session pointer is NULL");
  Session* S = nullptr;
  const int rc = create(fits, nfits, opt, &S, 0);
  *session = S;
  return rc;
}

int mhre_session_run(void* session, int niter) {
  if (!session) FAIL(MHRE_EINVAL, "NULL session");
  return run(static_cast<Session*>(session), niter);
}

int mhre_session_time_kernels(void* session, int niter, double* step_ms, double* exch_ms) {
  if (!session) FAIL(MHRE_EINVAL, "NULL session");
  return time_kernels(static_cast<Session*>(session), niter, step_ms, exch_ms);
}

int mhre_session_autotune(void* session, int iters, int* waves_per_block, int* persistent, double* step_ms) {
  if (!session) FAIL(MHRE_EINVAL, "null session");
  (void)hipGetLastError();
  return autotune(static_cast<Session*>(session), iters, waves_per_block, persistent, step_ms);
}

int mhre_session_sync(void* session) {
  if (!session) FAIL(MHRE_EINVAL, "NULL session");
  HIPCHK(hipStreamSynchronize(static_cast<Session*>(session)->stream));
  return MHRE_OK;
}

int mhre_session_download(void* session, int model_data) {
  if (!session) FAIL(MHRE_EINVAL, "NULL session");
  return download(static_cast<Session*>(session), model_data);
}

int mhre_session_stats(void* session, MhreStats* st) {
  if (!session) FAIL(MHRE_EINVAL, "NULL session");
  return stats(static_cast<Session*>(session), st);
}

int mhre_session_fit_info(const void* session, int fit, int* info, int ninfo) {
  if (!session || !info || ninfo < 4) FAIL(MHRE_EINVAL, "mhre_session_fit_info: NULL argument or ninfo < 4");
  const Session* S = static_cast<const Session*>(session);
  if (fit < 0 || fit >= (int)S->fits.size()) FAIL(MHRE_EINVAL, "mhre_session_fit_info: no fit %d", fit);
  const FitHost& h = S->fits[fit];
  info[0] = h.kind;
  info[1] = (h.xfused || quad_fused(h)) ? 1 : 0;
  info[2] = h.hwphase ? 1 : 0;
  info[3] = (int)S->groups.size();
  if (ninfo >= 5) info[4] = h.qrow.empty() ? 0 : 1;
  return MHRE_OK;
}

int mhre_session_destroy(void* session) {
  destroy_session(static_cast<Session*>(session));
  return MHRE_OK;
}

int mhre_run_batch(const MhreFitSpec* fits, int nfits, const MhreOptions* opt, MhreStats* st) {
  Session* S = nullptr;
  int rc = create(fits, nfits, opt, &S, 0);
  if (rc != MHRE_OK) return rc;
  int total = 0;
  for (int i = 0; i < nfits; ++i) total = std::max(total, fits[i].nburn + fits[i].niter - fits[i].it0);
  rc = run(S, total);
  bool model_data = false;
  for (int i = 0; i < nfits; ++i) model_data = model_data || fits[i].datm != nullptr;
  if (rc == MHRE_OK) rc = download(S, model_data ? 1 : 0);
  if (rc == MHRE_OK && st) rc = stats(S, st);
  destroy_session(S);
  return rc;
}

int mcmc_cuda(float* coor, float* dat, float* std2r, int* icoor, int* idat, float* datm, float* chi2m,
              int* pdescr, int* ivar, int* invar, float* ptotal, float* pmint, float* pmaxt, float* pout,
              int* tout, float* chi2, float* chi2c, int* tcur, int* flag, int* n_acpt, int* n_exch, int* n_cnt,
              int* n_hist, float* beta, float* pstp, float* ptent, int* ptentn, float* pcur, unsigned char* rndst,
              unsigned long long seed, int imodel, int ncoor, int ndat, int nicoor, int nidat, int ndatm,
              int nptot, int nprm, int nadj, int npass, int nbeta, int nseq, int nburn, int niter) {
  const MhreFitSpec s = legacy_spec(coor, dat, std2r, icoor, idat, datm, chi2m, pdescr, ivar, invar, ptotal, pmint,
                                    pmaxt, pout, tout, chi2, chi2c, tcur, flag, n_acpt, n_exch, n_cnt, n_hist, beta,
                                    pstp, ptent, ptentn, pcur, rndst, seed, imodel, ncoor, ndat, nicoor, nidat, ndatm,
                                    nptot, nprm, nadj, npass, nbeta, nseq, nburn, niter);
  MhreOptions opt{};
  opt.persistent = -1;
  return mhre_run_batch(&s, 1, &opt, nullptr);
}

int calc_modchi2(float* coor, float* dat, float* std2r, int* icoor, int* idat, float* datm, float* chi2m,
                 float* pcur, float* chi2c, int* flag, int imodel, int ncoor, int ndat, int nicoor, int nidat,
                 int ndatm, int nptot, int nbeta, int nseq) {
  if (!pcur || !chi2c || !flag || !dat || !std2r || nbeta <= 0 || nseq <= 0 || ndat <= 0 || nptot <= 0)
    FAIL(MHRE_EINVAL, "calc_modchi2: invalid arguments");
  if (nptot > MAXP) FAIL(MHRE_EINVAL, "calc_modchi2: nptot %d exceeds %d", nptot, MAXP);
  const int C = nbeta * nseq;
  if (ndatm < C * ndat) FAIL(MHRE_EINVAL, "calc_modchi2: ndatm %d < nbeta*nseq*ndat", ndatm);
  std::vector<int> pdescr(nptot, 0), invar(nptot, -1), ivar(1, 0);
  std::vector<float> pmm(nptot, 0.0f), beta(nbeta, 1.0f), zf(C, 0.0f);
  std::vector<int> zi(C, 0);
  MhreFitSpec s{};
  s.model = MHRE_MODEL_AUTO; s.imodel = imodel;
  s.coor = coor; s.dat = dat; s.std2r = std2r; s.icoor = icoor; s.idat = idat; s.datm = datm; s.chi2m = chi2m;
  s.pdescr = pdescr.data(); s.ivar = ivar.data(); s.invar = invar.data(); s.pmint = pmm.data(); s.pmaxt = pmm.data();
  s.beta = beta.data(); s.pcur = pcur; s.chi2c = chi2c; s.flag = flag;
  s.pstp = zf.data(); s.n_cnt = zi.data(); s.n_acpt = zi.data(); s.n_hist = zi.data(); s.n_exch = zi.data();
  s.tcur = zi.data(); s.ptent = zf.data(); s.ptentn = zi.data();
  s.ncoor = ncoor; s.ndat = ndat; s.nicoor = nicoor; s.nidat = nidat; s.ndatm = ndatm;
  s.nptot = nptot; s.nprm = 1; s.nadj = 1; s.npass = 2; s.nbeta = nbeta; s.nseq = nseq;
  Session* S = nullptr;
  int rc = create(&s, 1, nullptr, &S, 1);
  if (rc != MHRE_OK) return rc;
  // the reference runs passes 0 and 1 (gpu_mcmc.cu:778-779)
  rc = launch_eval_all(S, 1, 1);
  const FitHost& h = S->fits[0];
  if (rc == MHRE_OK) rc = d2h(S, chi2c, h.o_chi2c, C);
  if (rc == MHRE_OK && datm) rc = d2h(S, datm, h.o_datm, (size_t)std::min<long long>(h.d.ndatm, ndatm));
  if (rc == MHRE_OK && chi2m) rc = d2h(S, chi2m, h.o_chi2m, (size_t)C * ndat);
  if (rc == MHRE_OK) rc = check_model_error(S);
  destroy_session(S);
  return rc;
}

int calc_model(float* coor, int* idat, float* datm, float* pcur, int imodel, int ncoor, int ndat, int nidat,
               int ndatm, int nptot, int nbeta, int nseq) {
  if (!pcur || !datm || !coor || nbeta <= 0 || nseq <= 0 || ndat <= 0 || nptot <= 0)
    FAIL(MHRE_EINVAL, "calc_model: invalid arguments");
  if (nptot > MAXP) FAIL(MHRE_EINVAL, "calc_model: nptot %d exceeds %d", nptot, MAXP);
  const int C = nbeta * nseq;
  if (ndatm < C * ndat) FAIL(MHRE_EINVAL, "calc_model: ndatm %d < nbeta*nseq*ndat", ndatm);
  std::vector<int> pdescr(nptot, 0), invar(nptot, -1), ivar(1, 0), flag(C, 1), zi(C, 0);
  std::vector<float> pmm(nptot, 0.0f), beta(nbeta, 1.0f), zf(C, 0.0f), dat(ndat, 0.0f), w(ndat, 0.0f);
  std::vector<int> idat_local(3, 0);
  MhreFitSpec s{};
  s.model = MHRE_MODEL_AUTO; s.imodel = imodel;
  s.coor = coor; s.dat = dat.data(); s.std2r = w.data();
  s.datm = datm; s.chi2m = nullptr;
  idat_local[0] = ndat;
  if (idat && nidat >= 1) idat_local[0] = idat[0];
  s.idat = idat ? idat : idat_local.data(); s.nidat = idat ? nidat : 3;
  s.pdescr = pdescr.data(); s.ivar = ivar.data(); s.invar = invar.data(); s.pmint = pmm.data(); s.pmaxt = pmm.data();
  s.beta = beta.data(); s.pcur = pcur; s.chi2c = zf.data(); s.flag = flag.data();
  s.pstp = zf.data(); s.n_cnt = zi.data(); s.n_acpt = zi.data(); s.n_hist = zi.data(); s.n_exch = zi.data();
  s.tcur = zi.data(); s.ptent = zf.data(); s.ptentn = zi.data();
  s.ncoor = ncoor; s.ndat = ndat; s.nicoor = 0; s.ndatm = ndatm;
  s.nptot = nptot; s.nprm = 1; s.nadj = 1; s.npass = 1; s.nbeta = nbeta; s.nseq = nseq;
  // model data only: closures are not part of calc_model (pass 0)
  std::vector<int> idat_model(s.idat, s.idat + std::max(s.nidat, 1));
  idat_model[0] = ndat;
  if (idat_model.size() >= 3) idat_model[2] = 0;
  s.idat = idat_model.data();
  Session* S = nullptr;
  int rc = create(&s, 1, nullptr, &S, 1);
  if (rc != MHRE_OK) return rc;
  const FitHost& h = S->fits[0];
  hipError_t e = hipSuccess;
  const KindWork& k = S->kinds[0];
  if (h.kind == K_SGRA9 || h.kind == K_SGRA13)
    e = launch_eval_eht(h.kind == K_SGRA9 ? MODEL_SGRA9 : MODEL_SGRA13, S->d_fits, k.d_eval, C, 1, 0, S->stream);
  else
    e = launch_eval_user(S->d_fits, k.d_eval, C, 1, 0, S->stream);
  if (e != hipSuccess) {
    set_err("calc_model launch: %s", hipGetErrorString(e));
    rc = MHRE_EHIP;
  }
  if (rc == MHRE_OK) rc = d2h(S, datm, h.o_datm, (size_t)std::min<long long>(h.d.ndatm, ndatm));
  if (rc == MHRE_OK) rc = check_model_error(S);
  destroy_session(S);
  return rc;
}

int reset_gpu(void) {
  HIPCHK(hipDeviceReset());
  return MHRE_OK;
}

int list_gpus(char* buf, int buflen) {
  if (!buf || buflen <= 0) FAIL(MHRE_EINVAL, "list_gpus: empty buffer");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  std::string out;
  for (int i = 0; i < n; ++i) {
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, i));
    char line[512];
    snprintf(line, sizeof(line), "%d: %s (%s), %d compute units, %.1f GiB, %d MHz\n", i, p.name, p.gcnArchName,
             p.multiProcessorCount, p.totalGlobalMem / 1073741824.0, p.clockRate / 1000);
    out += line;
  }
  snprintf(buf, (size_t)buflen, "%s", out.c_str());
  return n;
}

int mhre_eht_data_create(const float* coor, const float* dat, const float* std2r, const int* icoor,
                         const int* idat, int imodel, int ncoor, int ndat, int nicoor, int nidat, int nptot,
                         void** handle) {
  if (!handle) FAIL(MHRE_EINVAL, "null handle pointer");
  EhtData* E = nullptr;
  const int rc = eht_data_create(coor, dat, std2r, icoor, idat, imodel, ncoor, ndat, nicoor, nidat, nptot, &E);
  *handle = E;
  return rc;
}

int mhre_eht_chi2_grad(void* handle, const float* params, int n, int ldp, double* chi2, float* grad,
                       void* stream) {
  (void)hipGetLastError();
  const EhtData* E = static_cast<const EhtData*>(handle);
  CHK(eht_check_batch(E, params, n, ldp, chi2));
  const int ring = E->h.ring_grad && !ring_grad_disabled();
  HIPCHK(launch_eht_grad9(E->d_fd, params, n, ldp, chi2, grad, ring, static_cast<hipStream_t>(stream)));
  return MHRE_OK;
}

int mhre_eht_gn_system(void* handle, const float* params, int n, int ldp, double* chi2, double* jtr,
                       double* jtj, void* stream) {
  (void)hipGetLastError();
  const EhtData* E = static_cast<const EhtData*>(handle);
  CHK(eht_check_batch(E, params, n, ldp, chi2));
  if (n > 0 && (!jtr || !jtj)) FAIL(MHRE_EINVAL, "jtr and jtj are required");
  HIPCHK(launch_eht_gn9(E->d_fd, params, n, ldp, chi2, jtr, jtj, static_cast<hipStream_t>(stream)));
  return MHRE_OK;
}

int mhre_eht_data_destroy(void* handle) {
  destroy_eht_data(static_cast<EhtData*>(handle));
  return MHRE_OK;
}

int mhre_eht_data_info(const void* handle, int* info, int ninfo) {
  if (!handle || !info) FAIL(MHRE_EINVAL, "null handle or info");
  if (ninfo < 4) FAIL(MHRE_EINVAL, "ninfo = %d, need 4", ninfo);
  const FitHost& h = static_cast<const EhtData*>(handle)->h;
  info[0] = h.ring_grad ? 1 : 0;
  info[1] = h.nround;
  info[2] = h.cgroup;
  info[3] = h.ring_ok;
  return MHRE_OK;
}

int mhre_quad_jacobian(const float* Q, const float* b, const float* dat, const int* sys, int nsys, int neq, int n,
                       const float* X, int rows, float* r, float* J) {
  (void)hipGetLastError();
  QuadJacLayout Lq;
  CHK(quad_jac_layout(Q, b, dat, sys, nsys, neq, n, X, rows, r, J, &Lq));
  if (rows == 0) return MHRE_OK;
  char* dev = nullptr;
  HIPCHK(hipMalloc((void**)&dev, Lq.bytes));
  struct Free {
    char* p;
    ~Free() { hipFree(p); }
  } guard{dev};
  return quad_jac_staged(dev, Lq, Q, b, dat, sys, X, r, J);
}

int mhre_quad_jacobian_noalloc(const float* Q, const float* b, const float* dat, const int* sys, int nsys, int neq,
                               int n, const float* X, int rows, float* r, float* J) {
  (void)hipGetLastError();
  QuadJacLayout Lq;
  CHK(quad_jac_layout(Q, b, dat, sys, nsys, neq, n, X, rows, r, J, &Lq));
  if (rows == 0) return MHRE_OK;
  if (Lq.bytes > kScratchBytes)
    FAIL(MHRE_EINVAL, "mhre_quad_jacobian_noalloc: the call needs %zu bytes of staging, more than the %zu-byte "
                      "static scratch area (use mhre_quad_jacobian)", Lq.bytes, kScratchBytes);
  char* dev = nullptr;
  CHK(scratch_on_current_device(&dev));
  return quad_jac_staged(dev, Lq, Q, b, dat, sys, X, r, J);
}

}  // extern "C"
