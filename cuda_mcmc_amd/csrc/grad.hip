// chi^2 gradient and Gauss-Newton normal equations of the native 9-parameter
// ring+Gaussian model (examples/sgra_model_fit/model.cuh:322-462), for batches
// of parameter vectors.  No reference counterpart: the reference samples
// without derivatives.
//
// One wave64 per parameter vector; lanes stride over the data.
//   * The model of one visibility is a function of the fourteen fp32
//     invariants that Sgra9::prepare builds from the 9 parameters.  Each datum
//     back-propagates its residual's adjoint into those invariants (reverse
//     mode, fp32 per datum, fp64 accumulation per lane).
//   * The invariants' 14 x 9 Jacobian comes from a float64 forward-mode pass
//     over the same formulas: lane j carries the tangent along parameter j.
//   * gradient = J_inv^T * (sum of the invariant adjoints).
//   * k_eht_gn9 forms per datum the residual row dr/dp = J_inv^T dr/dinv and
//     accumulates J^T J (45 packed entries) and J^T r in float64, with
//     r_i = (dat_i - model_i) sqrt(w_i) (closures: the wrapped phase residual),
//     so chi^2 = |r|^2 and grad chi^2 = 2 J^T r.
// Bessel derivatives: with B(y) = J1(x)/x, C(y) = J2(x)/x^2, D(y) = J3(x)/x^3
// and y = x^2, dB/dy = -C/2 and dC/dy = -D/2 (from d/dx[x^-n J_n] = -x^-n J_n+1).
#include <hip/hip_runtime.h>
#include "mhre_device.h"
#include "eht_model.h"
#include "fitdesc.h"
#include "grad.h"

namespace mhre {

constexpr int NINV = 14;   // cth sth se rq2 kBe kCe kBi kCi kd cr cg gaU gaV kg
constexpr int NP9 = 9;
constexpr int NJTJ = NP9 * (NP9 + 1) / 2;

// ------------------------------------------------- float64 forward mode

struct Dual {
  double v, d;
};
__device__ inline Dual dc(double c) { return Dual{c, 0.0}; }
__device__ inline Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
__device__ inline Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
__device__ inline Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Dual operator*(double c, Dual a) { return Dual{c * a.v, c * a.d}; }
__device__ inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual{q, (a.d - q * b.d) / b.v};
}
__device__ inline Dual dsin(Dual a) { return Dual{sin(a.v), cos(a.v) * a.d}; }
__device__ inline Dual dcos(Dual a) { return Dual{cos(a.v), -sin(a.v) * a.d}; }

// The invariants of Sgra9::prepare (same formulas, float64 duals).
// out[]: cth sth se rq2 kBe kCe kBi kCi kd cr cg gaU gaV kg
__device__ inline void sgra9_inv_dual(const Dual* p, Dual* out) {
  const Dual Zsp = p[0], Re = p[1], rq = p[2], ecc = p[3], fade = p[4];
  const Dual gax = p[5], aq = p[6], gq = p[7], th = p[8];
  const Dual Rex = UAS2NRAD * Re;
  const Dual Rin = rq * Rex;
  const Dual d = ecc * (Rex - Rin);
  const Dual one = dc(1.0);
  const Dual den = (Rex * Rex - Rin * Rin * (one + d / Rex)) * fade + (Rex * Rex - Rin * Rin * (one - d / Rex));
  const Dual h = dc(2.0 / D_PI) / den;
  const Dual h0 = fade * h;
  out[0] = dcos(th);
  out[1] = dsin(th);
  out[2] = (2.0 * D_PI * Rex) * (2.0 * D_PI * Rex);
  out[3] = rq * rq;
  out[4] = D_PI * ((h + h0) * Rex * Rex);
  out[5] = (-2.0 * D_PI * D_PI) * ((h - h0) * Rex * Rex * Rex);
  out[6] = D_PI * ((h + h0 - d * (h - h0) / Rex) * Rin * Rin);
  out[7] = (-2.0 * D_PI * D_PI) * ((h - h0) * rq * Rin * Rin * Rin);
  out[8] = (2.0 * D_PI) * d;
  const bool nog = (p[5].v == 0.0) || (p[6].v == 0.0) || (p[7].v == 0.0);
  const Dual gsy = gax * Rex, gsx = aq * gsy;
  out[9] = nog ? Zsp : Zsp * (one - gq);
  out[10] = nog ? dc(0.0) : Zsp * gq;
  out[11] = (GCOEF * LOG2E) * (gsx * gsx);
  out[12] = (GCOEF * LOG2E) * (gsy * gsy);
  out[13] = (-2.0 * D_PI) * (Rin - d);
}

// ------------------------------------------------------ per-datum model

// B, C and D at y = x^2.  y <= 64: the quarter-interval quadratics of
// BESSEL_TAB (B, C; the fused kernels' table) and BESSEL_TABD (D), staged in
// LDS by the kernel; above, Hankel asymptotics for B and C and
// D = (4C - B)/y from J3 = 4 J2/x - J1.
struct BesselTabBCD {
  const float4* tab;    // 2 float4 per interval: B0 B1 B2 C0, C1 C2 - -
  const float4* tabd;   // 1 float4 per interval: D0 D1 D2 -
  __device__ void operator()(float y, float* B, float* C, float* D) const {
    if (y <= 64.0f) {
      const float z = 4.0f * y;
      const int j = (int)z;
      const float s = __builtin_amdgcn_fractf(z);
      const float4 a = tab[2 * j], b = tab[2 * j + 1], c = tabd[j];
      *B = fmaf(fmaf(a.z, s, a.y), s, a.x);
      *C = fmaf(fmaf(b.y, s, b.x), s, a.w);
      *D = fmaf(fmaf(c.z, s, c.y), s, c.x);
    } else {
      bessel_BC_asym(y, B, C);
      *D = (4.0f * *C - *B) * __builtin_amdgcn_rcpf(y);
    }
  }
};

// Copies BESSEL_TAB and BESSEL_TABD into the work-group's LDS tables (all
// threads take part; call before any thread of the group leaves).
__device__ inline void load_bessel_tabs(float4* tab, float4* tabd) {
  float* t = reinterpret_cast<float*>(tab);
  for (int i = (int)threadIdx.x; i < 8 * BESSEL_TAB_N; i += (int)blockDim.x) t[i] = BESSEL_TAB[i];
  float* td = reinterpret_cast<float*>(tabd);
  for (int i = (int)threadIdx.x; i < 4 * BESSEL_TAB_N; i += (int)blockDim.x) td[i] = BESSEL_TABD[i];
  __syncthreads();
}

struct Fwd9 {
  float u, v, U, W, rho2, ye, Be, Ce, De, Bi, Ci, Di, sd, cd, Fx, Fy, ga, sg, cgs, Vx, Vy;
};

// Forward model of one visibility, keeping what the backward pass needs.
// The model of Sgra9::vis (same Bessel table, sin/cos of the phases in
// revolutions) with the full-range Bessel branch; the ring term is mixed
// after the fact, as the reverse pass needs the unmixed ring.
__device__ inline Fwd9 sgra9_fwd(const Sgra9::Inv& q, float u, float v, const BesselTabBCD& bt) {
  Fwd9 f;
  f.u = u;
  f.v = v;
  f.U = fmaf(u, q.cth, v * q.sth);
  f.W = fmaf(v, q.cth, -u * q.sth);
  f.rho2 = fmaf(f.U, f.U, f.W * f.W);
  f.ye = q.r.se * f.rho2;
  const float yi = q.r.rq2 * f.ye;
  bt(f.ye, &f.Be, &f.Ce, &f.De);
  bt(yi, &f.Bi, &f.Ci, &f.Di);
  const float ree = q.r.kBe * f.Be;
  const float ime = q.r.kCe * f.U * f.Ce;
  const float rei = q.r.kBi * f.Bi;
  const float imi = q.r.kCi * f.U * f.Ci;
  sincosrev(q.r.kdr * f.U, &f.sd, &f.cd);
  f.Fx = ree - (rei * f.cd - imi * f.sd);
  f.Fy = ime - (rei * f.sd + imi * f.cd);
  if (q.r.gauss) {
    f.ga = __builtin_amdgcn_exp2f(-fmaf(q.gaU, f.U * f.U, q.gaV * (f.W * f.W)));
    sincosrev(q.kgr * f.U, &f.sg, &f.cgs);
    const float gw = q.r.cg * f.ga;
    f.Vx = fmaf(q.r.cr, f.Fx, gw * f.cgs);
    f.Vy = fmaf(q.r.cr, f.Fy, gw * f.sg);
  } else {
    f.ga = f.sg = f.cgs = 0.0f;
    f.Vx = q.r.cr * f.Fx;
    f.Vy = q.r.cr * f.Fy;
  }
  return f;
}

// Reverse pass of sgra9_fwd: adds (gx, gy) . d(Vx, Vy)/d(invariant m) to a[m].
template <class A>
__device__ inline void sgra9_bwd(const Sgra9::Inv& q, const Fwd9& f, float gx, float gy, A* a) {
  constexpr float LN2 = 0.693147180559945309f;
  a[9] += (A)(gx * f.Fx + gy * f.Fy);
  const float aFx = q.r.cr * gx, aFy = q.r.cr * gy;
  float aU = 0.0f, aW = 0.0f;
  if (q.r.gauss) {
    const float gw = q.r.cg * f.ga;
    const float agw = gx * f.cgs + gy * f.sg;
    const float aphg = gw * (gy * f.cgs - gx * f.sg);
    a[10] += (A)(agw * f.ga);
    const float aE = -LN2 * agw * gw;
    a[11] += (A)(aE * f.U * f.U);
    a[12] += (A)(aE * f.W * f.W);
    aU += 2.0f * aE * q.gaU * f.U + aphg * q.kg;
    aW += 2.0f * aE * q.gaV * f.W;
    a[13] += (A)(aphg * f.U);
  }
  const float rei = q.r.kBi * f.Bi, imi = q.r.kCi * f.U * f.Ci;
  const float arei = -aFx * f.cd - aFy * f.sd;
  const float aimi = aFx * f.sd - aFy * f.cd;
  const float asd = aFx * imi - aFy * rei;
  const float acd = -aFx * rei - aFy * imi;
  const float aphd = asd * f.cd - acd * f.sd;
  a[8] += (A)(aphd * f.U);
  aU += aphd * q.r.kd;
  a[4] += (A)(aFx * f.Be);
  const float aBe = aFx * q.r.kBe;
  a[5] += (A)(aFy * f.U * f.Ce);
  aU += aFy * q.r.kCe * f.Ce;
  const float aCe = aFy * q.r.kCe * f.U;
  a[6] += (A)(arei * f.Bi);
  const float aBi = arei * q.r.kBi;
  a[7] += (A)(aimi * f.U * f.Ci);
  aU += aimi * q.r.kCi * f.Ci;
  const float aCi = aimi * q.r.kCi * f.U;
  float aye = -0.5f * (aBe * f.Ce + aCe * f.De);
  const float ayi = -0.5f * (aBi * f.Ci + aCi * f.Di);
  a[3] += (A)(ayi * f.ye);
  aye += ayi * q.r.rq2;
  a[2] += (A)(aye * f.rho2);
  const float arho2 = aye * q.r.se;
  aU += 2.0f * f.U * arho2;
  aW += 2.0f * f.W * arho2;
  a[0] += (A)(aU * f.u + aW * f.v);
  a[1] += (A)(aU * f.v - aW * f.u);
}

// d arg(V) / d(Vx, Vy) scaled by s.
__device__ inline float2 darg(const Fwd9& f, float s) {
  const float m2 = fmaxf(fmaf(f.Vx, f.Vx, f.Vy * f.Vy), 1e-30f);
  const float k = s * __builtin_amdgcn_rcpf(m2);
  return make_float2(-f.Vy * k, f.Vx * k);
}

// Closure residual dd = wrap(obs - (pa + pb - pc)) as in closure_term.
__device__ inline float closure_resid(const Fwd9& a, const Fwd9& b, const Fwd9& e, float4 ow) {
  const float abr = a.Vx * b.Vx - a.Vy * b.Vy, abi = a.Vx * b.Vy + a.Vy * b.Vx;
  const float br = abr * e.Vx + abi * e.Vy, bi = abi * e.Vx - abr * e.Vy;
  const float xr = ow.x * br + ow.y * bi;
  const float xi = ow.y * br - ow.x * bi;
  return atan2_fast(xi, xr);
}

// ----------------------------------------------------------------- kernels

struct GradArgs {
  const FitDesc* fd;   // one fit's packed data (device)
  const float* P;      // [n][ldp] parameter vectors (device)
  int n, ldp;
  double* chi2;        // [n]
  float* grad;         // [n][ldp] or null
  double* jtr;         // [n][9]   (k_eht_gn9)
  double* jtj;         // [n][45]  (k_eht_gn9), row-major upper triangle
};

// Lane j < 9 computes column j of d(inv)/dp into jac[m * 9 + j] (LDS).
__device__ inline void inv_jacobian(float pl, double* jac) {
  const int lane = lane_id();
  Dual pd[NP9];
#pragma unroll
  for (int i = 0; i < NP9; ++i) pd[i] = Dual{(double)bcast(pl, i), i == lane ? 1.0 : 0.0};
  Dual inv[NINV];
  sgra9_inv_dual(pd, inv);
  if (lane < NP9) {
#pragma unroll
    for (int m = 0; m < NINV; ++m) jac[m * NP9 + lane] = inv[m].d;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__global__ void __launch_bounds__(256) k_eht_grad9(GradArgs g) {
  __shared__ double jac_all[4][NINV * NP9];
  __shared__ float4 btab[2 * BESSEL_TAB_N];
  __shared__ float4 btabd[BESSEL_TAB_N];
  load_bessel_tabs(btab, btabd);
  const BesselTabBCD bt{btab, btabd};
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int w = blockIdx.x * 4 + wid;
  if (w >= g.n) return;
  const int lane = lane_id();
  const FitDesc& f = *g.fd;
  const float pl = lane < NP9 ? g.P[(size_t)w * g.ldp + lane] : 0.0f;
  float p[NP9];
#pragma unroll
  for (int j = 0; j < NP9; ++j) p[j] = bcast(pl, j);
  const Sgra9::Inv q = Sgra9::prepare(p);
  double* jac = jac_all[wid];
  if (g.grad) inv_jacobian(pl, jac);

  const int nvr = f.ndat - f.ncph;
  double chi = 0.0;
  double a[NINV];
#pragma unroll
  for (int m = 0; m < NINV; ++m) a[m] = 0.0;
  for (int i = lane; i < nvr; i += 64) {
    const float4 d = gptr(f.vis)[i];
    const Fwd9 fw = sgra9_fwd(q, d.x, d.y, bt);
    const float amp = __builtin_amdgcn_sqrtf(fmaf(fw.Vx, fw.Vx, fw.Vy * fw.Vy));
    const float r = fmaf(-amp, d.w, d.z);   // (d - |V|) sqrt(w): vis holds (d sqrt(w), sqrt(w))
    chi += (double)(r * r);
    if (g.grad) {
      // d(r^2)/d|V| = -2 r sqrt(w)
      const float s = -2.0f * r * d.w * __builtin_amdgcn_rcpf(fmaxf(amp, 1e-30f));
      sgra9_bwd(q, fw, s * fw.Vx, s * fw.Vy, a);
    }
  }
  for (int c = lane; c < f.ncph; c += 64) {
    const int4 ix = gptr(f.clo)[c];
    const float4 ow = gptr(f.clow)[c];
    const GLOBAL float4* vis = gptr(f.vis);
    const float4 da = vis[ix.x], db = vis[ix.y], de = vis[ix.z];
    const Fwd9 fa = sgra9_fwd(q, da.x, da.y, bt);
    const Fwd9 fb = sgra9_fwd(q, db.x, db.y, bt);
    const Fwd9 fe = sgra9_fwd(q, de.x, de.y, bt);
    const float dd = closure_resid(fa, fb, fe, ow);
    chi += (double)(dd * dd * ow.z);
    if (g.grad) {
      // dd = obs - pa - pb + pc: d(dd^2 w)/dpa = -2 dd w
      const float s = -2.0f * dd * ow.z;
      const float2 ga = darg(fa, s), gb = darg(fb, s), ge = darg(fe, -s);
      sgra9_bwd(q, fa, ga.x, ga.y, a);
      sgra9_bwd(q, fb, gb.x, gb.y, a);
      sgra9_bwd(q, fe, ge.x, ge.y, a);
    }
  }
  chi = wave_sum(chi);
  if (lane == 0) g.chi2[w] = chi;
  if (!g.grad) return;
#pragma unroll
  for (int m = 0; m < NINV; ++m) a[m] = wave_sum(a[m]);
  if (lane < g.ldp) {
    double gj = 0.0;
    if (lane < NP9) {
#pragma unroll
      for (int m = 0; m < NINV; ++m) gj += a[m] * jac[m * NP9 + lane];
    }
    g.grad[(size_t)w * g.ldp + lane] = (float)gj;
  }
}

// Closure residual from the model visibilities alone (closure_resid).
__device__ inline float closure_resid_v(float2 a, float2 b, float2 e, float4 ow) {
  const float abr = a.x * b.x - a.y * b.y, abi = a.x * b.y + a.y * b.x;
  const float br = abr * e.x + abi * e.y, bi = abi * e.x - abr * e.y;
  const float xr = ow.x * br + ow.y * bi;
  const float xi = ow.y * br - ow.x * bi;
  return atan2_fast(xi, xr);
}

// Gradient for fits whose closures lie inside the 512-visibility ring window
// (FitDesc::ring_ok) and whose closure batches split into at most 16
// collision-free rounds (host side).  One wave per parameter point streams the 64-wide
// tiles once, forming the closures after every cgroup tiles from an LDS ring
// of model visibilities, as the step kernel does.  d(arg V)/dV is linear in
// its scale, so the closure terms of a visibility reduce to one scalar
// S_j = sum of +-(-2 dd w) over the closures that contain it; S_j is
// accumulated in a second ring, in FitDesc::nround collision-free rounds of
// plain read-add-writes (FitDesc::clr).  A visibility's reverse pass (amplitude term
// and darg(V_j, S_j) together) runs once no later closure group can reach it,
// just before its ring slot is reused, with its forward recomputed.  Per
// visibility that is two forward and one reverse pass, where k_eht_grad9 runs
// three of each for every closure.
static_assert(MAX_CGROUP <= RING / TILE, "a closure group must fit in the ring");
__global__ void __launch_bounds__(256) k_eht_grad9_ring(GradArgs g) {
  __shared__ double jac_all[4][NINV * NP9];
  __shared__ float2 ringv_all[4][RING];
  __shared__ float rings_all[4][RING];
  __shared__ float4 btab[2 * BESSEL_TAB_N];
  __shared__ float4 btabd[BESSEL_TAB_N];
  load_bessel_tabs(btab, btabd);
  const BesselTabBCD bt{btab, btabd};
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int w = blockIdx.x * 4 + wid;
  if (w >= g.n) return;
  const int lane = lane_id();
  const FitDesc& f = *g.fd;
  const float pl = lane < NP9 ? g.P[(size_t)w * g.ldp + lane] : 0.0f;
  float p[NP9];
#pragma unroll
  for (int j = 0; j < NP9; ++j) p[j] = bcast(pl, j);
  const Sgra9::Inv q = Sgra9::prepare(p);
  double* jac = jac_all[wid];
  float2* ringv = ringv_all[wid];
  float* rings = rings_all[wid];
  const bool grad = g.grad != nullptr;
  if (grad) inv_jacobian(pl, jac);

  const GLOBAL float4* vis = gptr(f.vis);
  const GLOBAL int4* clo = gptr(f.clo);
  const GLOBAL float4* clow = gptr(f.clow);
  const GLOBAL int2* tinfo = gptr(f.tinfo);
  const GLOBAL int* clr = gptr(f.clr);
  const int nvr = f.ndat - f.ncph, nvis = f.nvis, ntile = f.ntile, cg = f.cgroup, nround = f.nround;
  constexpr int RT = RING / TILE;   // tiles the ring holds
  double chi = 0.0;
  // the invariants' adjoints: float partial sums over a few reverse passes
  // (one group's, or the last ring's: at most RT tiles), then fp64 totals
  double a[NINV];
  float af[NINV];
#pragma unroll
  for (int m = 0; m < NINV; ++m) {
    a[m] = 0.0;
    af[m] = 0.0f;
  }
  auto flush = [&]() {
#pragma unroll
    for (int m = 0; m < NINV; ++m) {
      a[m] += (double)af[m];
      af[m] = 0.0f;
    }
  };
  // reverse pass of tile tb, whose closures have all been formed
  auto reverse = [&](int tb, float4 d) {
    const int j = tb * TILE + lane;
    if (j < nvis) {
      const Fwd9 fw = sgra9_fwd(q, d.x, d.y, bt);
      const float m2 = fmaxf(fmaf(fw.Vx, fw.Vx, fw.Vy * fw.Vy), 1e-30f);
      const float rs = __builtin_amdgcn_rsqf(m2);   // 1/|V|: |V| = m2 rs, 1/|V|^2 = rs^2
      float sa = 0.0f;
      if (j < nvr) sa = -2.0f * fmaf(-m2 * rs, d.w, d.z) * d.w * rs;   // -2 (d - |V|) w / |V|
      // amplitude: s (Vx, Vy); closures: darg(V, S) = S (-Vy, Vx) / |V|^2
      const float k = rings[j & (RING - 1)] * (rs * rs);
      sgra9_bwd(q, fw, fmaf(sa, fw.Vx, -k * fw.Vy), fmaf(sa, fw.Vy, k * fw.Vx), af);
    }
  };
  int c0 = 0, rnext = 0;
  float4 dnext = vis[lane];
  for (int t0 = 0; t0 < ntile; t0 += cg) {
    const int t1 = t0 + cg < ntile ? t0 + cg : ntile;
    for (int t = t0; t < t1; ++t) {
      const int i = t * TILE + lane;
      const float4 d = dnext;
      dnext = vis[i + TILE];   // padded by one zero-weight tile
      const Fwd9 fw = sgra9_fwd(q, d.x, d.y, bt);
      if (i < nvr) {
        const float r = fmaf(-__builtin_amdgcn_sqrtf(fmaf(fw.Vx, fw.Vx, fw.Vy * fw.Vy)), d.w, d.z);
        chi += (double)(r * r);
      }
      ringv[i & (RING - 1)] = make_float2(fw.Vx, fw.Vy);
      rings[i & (RING - 1)] = 0.0f;
    }
    wave_sync_lds();
    // closures whose largest visibility index lies in tiles [t0, t1)
    const int c1 = tinfo[t1].x;
    for (int cb = c0; cb < c1; cb += TILE) {
      const int c = cb + lane;
      if (c < c1) {
        const int4 ix = clo[c];
        const float4 ow = clow[c];
        const float2 va = ringv[ix.x & (RING - 1)], vb = ringv[ix.y & (RING - 1)], ve = ringv[ix.z & (RING - 1)];
        const float dd = closure_resid_v(va, vb, ve, ow);
        chi += (double)(dd * dd * ow.z);
        if (grad) {
          // dd = obs - pa - pb + pc: d(dd^2 w)/dpa = -2 dd w
          const float s = -2.0f * dd * ow.z;
          const int code = clr[c];
          const int ra = code & 15, rb = (code >> 4) & 15, re = (code >> 8) & 15;
          for (int r = 0; r < nround; ++r) {
            // at most one leg of this closure, and no other lane's leg of
            // this round, names a given visibility
            if (r == ra || r == rb || r == re) {
              const int v = r == ra ? ix.x : r == rb ? ix.y : ix.z;
              rings[v & (RING - 1)] += r == re ? -s : s;
            }
            wave_sync_lds();
          }
        }
      }
    }
    c0 = c1;
    wave_sync_lds();
    // The next group [t1, t2) forms its closures after tile t2 - 1, so they
    // reach back to tile t2 - RT at the earliest, and its tiles reuse the
    // slots of tiles t1 - RT .. t2 - 1 - RT: those are final now, and
    // computed (t2 - t1 <= cg <= RT).
    if (grad) {
      const int rend = (t1 + cg < ntile ? t1 + cg : ntile) - RT;
      if (rnext < rend) {
        float4 dr = vis[rnext * TILE + lane];
        for (; rnext < rend; ++rnext) {
          const float4 d = dr;
          if (rnext + 1 < rend) dr = vis[(rnext + 1) * TILE + lane];
          reverse(rnext, d);
        }
        flush();
      }
    }
  }
  if (grad) {
    wave_sync_lds();
    for (; rnext < ntile; ++rnext) reverse(rnext, vis[rnext * TILE + lane]);
    flush();
  }
  chi = wave_sum(chi);
  if (lane == 0) g.chi2[w] = chi;
  if (!grad) return;
#pragma unroll
  for (int m = 0; m < NINV; ++m) a[m] = wave_sum(a[m]);
  if (lane < g.ldp) {
    double gj = 0.0;
    if (lane < NP9) {
#pragma unroll
      for (int m = 0; m < NINV; ++m) gj += a[m] * jac[m * NP9 + lane];
    }
    g.grad[(size_t)w * g.ldp + lane] = (float)gj;
  }
}

// Adds the residual row of one datum: dr/dp = J_inv^T dr/dinv.
__device__ inline void gn_accumulate(const float* rinv, float r, const double* jac, double* jtr, double* jtj) {
  float row[NP9];
#pragma unroll
  for (int j = 0; j < NP9; ++j) {
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < NINV; ++m) s += (double)rinv[m] * jac[m * NP9 + j];
    row[j] = (float)s;
  }
  int k = 0;
#pragma unroll
  for (int i = 0; i < NP9; ++i) {
    jtr[i] += (double)row[i] * (double)r;
#pragma unroll
    for (int j = i; j < NP9; ++j) jtj[k++] += (double)row[i] * (double)row[j];
  }
}

__global__ void __launch_bounds__(64) k_eht_gn9(GradArgs g) {
  __shared__ double jac[NINV * NP9];
  __shared__ float4 btab[2 * BESSEL_TAB_N];
  __shared__ float4 btabd[BESSEL_TAB_N];
  load_bessel_tabs(btab, btabd);
  const BesselTabBCD bt{btab, btabd};
  const int w = blockIdx.x;
  if (w >= g.n) return;
  const int lane = lane_id();
  const FitDesc& f = *g.fd;
  const float pl = lane < NP9 ? g.P[(size_t)w * g.ldp + lane] : 0.0f;
  float p[NP9];
#pragma unroll
  for (int j = 0; j < NP9; ++j) p[j] = bcast(pl, j);
  const Sgra9::Inv q = Sgra9::prepare(p);
  inv_jacobian(pl, jac);

  const int nvr = f.ndat - f.ncph;
  double chi = 0.0;
  double jtr[NP9], jtj[NJTJ];
#pragma unroll
  for (int i = 0; i < NP9; ++i) jtr[i] = 0.0;
#pragma unroll
  for (int i = 0; i < NJTJ; ++i) jtj[i] = 0.0;
  for (int i = lane; i < nvr; i += 64) {
    const float4 d = gptr(f.vis)[i];
    const Fwd9 fw = sgra9_fwd(q, d.x, d.y, bt);
    const float amp = __builtin_amdgcn_sqrtf(fmaf(fw.Vx, fw.Vx, fw.Vy * fw.Vy));
    const float sw = d.w;   // vis holds (d sqrt(w), sqrt(w))
    const float r = fmaf(-amp, sw, d.z);
    chi += (double)r * (double)r;
    // dr/dV = -sqrt(w) V / amp
    const float s = -sw * __builtin_amdgcn_rcpf(fmaxf(amp, 1e-30f));
    float rinv[NINV];
#pragma unroll
    for (int m = 0; m < NINV; ++m) rinv[m] = 0.0f;
    sgra9_bwd(q, fw, s * fw.Vx, s * fw.Vy, rinv);
    gn_accumulate(rinv, r, jac, jtr, jtj);
  }
  for (int c = lane; c < f.ncph; c += 64) {
    const int4 ix = gptr(f.clo)[c];
    const float4 ow = gptr(f.clow)[c];
    const GLOBAL float4* vis = gptr(f.vis);
    const float4 da = vis[ix.x], db = vis[ix.y], de = vis[ix.z];
    const Fwd9 fa = sgra9_fwd(q, da.x, da.y, bt);
    const Fwd9 fb = sgra9_fwd(q, db.x, db.y, bt);
    const Fwd9 fe = sgra9_fwd(q, de.x, de.y, bt);
    const float sw = __builtin_amdgcn_sqrtf(ow.z);
    const float r = closure_resid(fa, fb, fe, ow) * sw;
    chi += (double)r * (double)r;
    // r = sqrt(w) (obs - pa - pb + pc)
    const float2 ga = darg(fa, -sw), gb = darg(fb, -sw), ge = darg(fe, sw);
    float rinv[NINV];
#pragma unroll
    for (int m = 0; m < NINV; ++m) rinv[m] = 0.0f;
    sgra9_bwd(q, fa, ga.x, ga.y, rinv);
    sgra9_bwd(q, fb, gb.x, gb.y, rinv);
    sgra9_bwd(q, fe, ge.x, ge.y, rinv);
    gn_accumulate(rinv, r, jac, jtr, jtj);
  }
  chi = wave_sum(chi);
#pragma unroll
  for (int i = 0; i < NP9; ++i) jtr[i] = wave_sum(jtr[i]);
#pragma unroll
  for (int i = 0; i < NJTJ; ++i) jtj[i] = wave_sum(jtj[i]);
  if (lane == 0) {
    g.chi2[w] = chi;
#pragma unroll
    for (int i = 0; i < NP9; ++i) g.jtr[(size_t)w * NP9 + i] = jtr[i];
#pragma unroll
    for (int i = 0; i < NJTJ; ++i) g.jtj[(size_t)w * NJTJ + i] = jtj[i];
  }
}

hipError_t launch_eht_grad9(const FitDesc* fd, const float* P, int n, int ldp, double* chi2, float* grad,
                            int ring, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  GradArgs g{fd, P, n, ldp, chi2, grad, nullptr, nullptr};
  if (ring) k_eht_grad9_ring<<<(n + 3) / 4, 256, 0, s>>>(g);
  else k_eht_grad9<<<(n + 3) / 4, 256, 0, s>>>(g);
  return hipGetLastError();
}

hipError_t launch_eht_gn9(const FitDesc* fd, const float* P, int n, int ldp, double* chi2, double* jtr,
                          double* jtj, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  GradArgs g{fd, P, n, ldp, chi2, nullptr, jtr, jtj};
  k_eht_gn9<<<n, 64, 0, s>>>(g);
  return hipGetLastError();
}

// ------------------------------------------------ quadratic systems (GN)

// Residuals r_e = x^T Q_e x + b_e . x - dat_e (float64 accumulation, rounded)
// and Jacobian J_ek = sum_j (Q_ekj + Q_ejk) x_j + b_ek of row i's system
// sys[i] (sys == nullptr: system i).  One thread per row.
__global__ void __launch_bounds__(256) k_quad_jac(const float* Q, const float* b, const float* dat, const int* sys,
                                                  const float* X, int rows, int neq, int n, float* r, float* J) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const size_t s = sys ? (size_t)sys[i] : (size_t)i;
  const float* x = X + (size_t)i * n;
  for (int e = 0; e < neq; ++e) {
    const float* q = Q + (s * neq + e) * (size_t)n * n;
    const float* be = b + (s * neq + e) * (size_t)n;
    double lhs = 0.0;
    for (int j = 0; j < n; ++j) {
      double row = 0.0;
      for (int k = 0; k < n; ++k) row += (double)q[j * n + k] * (double)x[k];
      lhs += ((double)be[j] + row) * (double)x[j];
    }
    r[(size_t)i * neq + e] = (float)(lhs - (double)dat[s * neq + e]);
    for (int k = 0; k < n; ++k) {
      float g = be[k];
      for (int j = 0; j < n; ++j) g = fmaf(q[k * n + j] + q[j * n + k], x[j], g);
      J[((size_t)i * neq + e) * n + k] = g;
    }
  }
}

hipError_t launch_quad_jac(const float* Q, const float* b, const float* dat, const int* sys, const float* X,
                           int rows, int neq, int n, float* r, float* J, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  k_quad_jac<<<(rows + 255) / 256, 256, 0, s>>>(Q, b, dat, sys, X, rows, neq, n, r, J);
  return hipGetLastError();
}

}  // namespace mhre
