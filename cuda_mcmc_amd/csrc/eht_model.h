// Native fp32 ring+Gaussian visibility models for the fused chi^2 kernels.
//
// Sgra9 restates examples/sgra_model_fit/model.cuh:322-462 (sgra_model9) and
// Sgra13 restates model.cuh:120-315 (sgra_model13).  Per-chain invariants are
// computed once per proposal (in fp64, rounded to fp32); the per-visibility
// work uses the reduced form
//     ree =  pi (h+h0) Re^2 B(ye)              ime = -2 pi^2 (h-h0) Re^3 U C(ye)
//     rei =  pi (h+h0-d(h-h0)/Re) Ri^2 B(yi)   imi = -2 pi^2 (h-h0) rq Ri^3 U C(yi)
// with ye = (2 pi Re)^2 rho^2, yi = rq^2 ye, B(y) = J1(x)/x, C(y) = J2(x)/x^2,
// which follows from J1' = (J0 - J2)/2 and J0 + J2 = 2 J1/x and removes every
// 1/rho and the separate rho -> 0 branch of the reference (B(0) = 1/2 and
// C(0) = 1/8 reproduce its limits, model.cuh:411-416).
#pragma once
#include "mhre_device.h"

namespace mhre {

constexpr double D_PI = 3.141592653589793;
constexpr double UAS2NRAD = D_PI / 180.0 / 3600.0 * 1e3;
constexpr double LOG2E = 1.4426950408889634;
// 2 (pi / (2 sqrt(2 ln 2)))^2: FWHM^2 -> Gaussian exponent (model.cuh:372)
constexpr double GCOEF = 3.5597073019331127;

struct RingInv {
  float se, rq2;           // ye = se * rho^2, yi = rq2 * ye
  float kBe, kCe, kBi, kCi;
  float kd;                // inner-disk shift phase per unit U: 2 pi d
  float kdr;               // the same in revolutions: d
  float cr, cg;            // ring and Gaussian mixing weights (Zsp folded in)
  int gauss;               // Gaussian present
  float mBe, mCe, mBi, mCi;   // cr * (kBe, kCe, kBi, kCi): the ring term already mixed
  // float bits of 64 / (se max(1, rq2)): a tile whose largest u^2 + v^2 is
  // above it needs the large-argument Bessel branch (the larger of the two
  // arguments, max(1, rq^2) se rho^2, exceeds 64; rq > 1 makes the inner one
  // the larger)
  uint32_t r2thr;
};

// The mixed ring coefficients, once cr and the k's are known (rounded to
// fp32 from the fp32 factors, as the per-visibility product would be).
__device__ inline void mix_ring(RingInv* q) {
  q->mBe = q->cr * q->kBe;
  q->mCe = q->cr * q->kCe;
  q->mBi = q->cr * q->kBi;
  q->mCi = q->cr * q->kCi;
}

// h, h0 and the ring coefficients (model.cuh:366-375) for radii in nrad,
// given the inner disk's shift as a fraction of the outer radius, e = d / Re
// (the models know it without a division: e = ecc (1 - rq), times
// cos(eta - th) for the 13-parameter model).  With r = rq,
//   h = (2/pi) / (Re^2 [(1 - r^2 (1 + e)) fade + (1 - r^2 (1 - e))]),
// the reference's expression with Ri = r Re and d = e Re factored out: one
// float64 division per proposal.
__device__ inline void ring_coefs(double Rex, double rq, double fade, double e, RingInv* q) {
  const double Rin = rq * Rex;
  const double d = e * Rex;
  const double r2 = rq * rq;
  const double h = (2.0 / D_PI) / ((Rex * Rex) * ((1.0 - r2 * (1.0 + e)) * fade + (1.0 - r2 * (1.0 - e))));
  const double h0 = fade * h;
  q->se = (float)((2.0 * D_PI * Rex) * (2.0 * D_PI * Rex));
  q->rq2 = (float)r2;
  q->kBe = (float)(D_PI * (h + h0) * Rex * Rex);
  q->kCe = (float)(-2.0 * D_PI * D_PI * (h - h0) * Rex * Rex * Rex);
  q->kBi = (float)(D_PI * (h + h0 - e * (h - h0)) * Rin * Rin);
  q->kCi = (float)(-2.0 * D_PI * D_PI * (h - h0) * rq * Rin * Rin * Rin);
  q->kd = (float)(2.0 * D_PI * d);
  q->kdr = (float)d;
  q->r2thr = __float_as_uint(64.0f / (q->se * fmaxf(1.0f, q->rq2)));
}

// Ring visibility cr (Fex - Fin) at rotated coordinate U and squared
// baseline length rho2 = u^2 + v^2 (the rotation leaves it unchanged, so
// callers pass it precomputed and V is never formed).
// any_large == false: the caller guarantees ye <= 64 on every lane, so only
// the y <= 64 evaluation bc of B, C runs (BesselPoly or BesselTab).
// bc takes its argument in units of 1/BC::SCALE (the scale is folded into se
// once per chain, outside the visibility loop).  HW: the shift phase is known
// to lie inside the hardware sine's input domain (sincos_phase).
template <bool HW, class BC>
__device__ inline float2 ring_vis(const RingInv& q, float U, float rho2, bool any_large, const BC& bc) {
  constexpr float S = BC::SCALE;
  const float ye = (q.se * S) * rho2;
  const float yi = q.rq2 * ye;
  float Be, Ce, Bi, Ci;
  if (!any_large) {
    bc(ye, &Be, &Ce);
    bc(yi, &Bi, &Ci);
  } else {
    if (ye <= 64.0f * S) bc(ye, &Be, &Ce);
    else bessel_BC_asym(ye * (1.0f / S), &Be, &Ce);
    if (yi <= 64.0f * S) bc(yi, &Bi, &Ci);
    else bessel_BC_asym(yi * (1.0f / S), &Bi, &Ci);
  }
  const float ree = q.mBe * Be;
  const float ime = q.mCe * U * Ce;
  const float rei = q.mBi * Bi;
  const float imi = q.mCi * U * Ci;
  float sd, cd;
  sincos_phase<HW>(q.kdr * U, &sd, &cd);
  float2 F;
  F.x = fmaf(imi, sd, fmaf(-rei, cd, ree));
  F.y = fmaf(-imi, cd, fmaf(-rei, sd, ime));
  return F;
}

// ------------------------------------------------------------ 9 parameters
struct Sgra9 {
  static constexpr int NP = 9;
  struct Inv {
    RingInv r;
    float cth, sth;       // rotation by th
    float gaU, gaV;       // Gaussian exponent (log2 units) per U^2, V^2
    float gaD;            // gaU - gaV: the exponent is gaV rho^2 + gaD U^2
    float kg;             // Gaussian shift phase per unit U: -2 pi (Ri - d)
    float kgr;            // the same in revolutions: -(Ri - d)
  };

  // p[i] = parameter i (Zsp, Re, rq, ecc, fade, gax, aq, gq, th).
  // The invariants in four groups by the parameters they depend on.
  __device__ static void ring(Inv& q, const float* p) {   // Re, rq, ecc, fade
    const double Re = p[1], rq = p[2], ecc = p[3], fade = p[4];
    const double Rex = Re * UAS2NRAD;
    const double e = ecc * (1.0 - rq);   // d / Re, d = ecc (Re - Ri)
    ring_coefs(Rex, rq, fade, e, &q.r);
    const double gs = (rq - e) * Rex;     // Ri - d: the Gaussian's shift
    q.kg = (float)(-2.0 * D_PI * gs);
    q.kgr = (float)(-gs);
  }
  // The float64 trigonometric invariants (cos th, sin th).  The step kernel
  // keeps them from one proposal to the next and recomputes them only when
  // the proposal moves a parameter they depend on (trig_uses).
  struct Trig {
    float cth, sth;
  };
  __device__ static bool trig_uses(int t) { return t == 8; }
  __device__ static Trig trig(const float* p) {
    return Trig{(float)cos((double)p[8]), (float)sin((double)p[8])};
  }
  __device__ static void gaussian(Inv& q, const float* p) {   // gax, aq (and Re)
    const double Rex = (double)p[1] * UAS2NRAD;
    const double gsy = (double)p[5] * Rex, gsx = (double)p[6] * gsy;
    q.gaU = (float)(GCOEF * LOG2E * gsx * gsx);
    q.gaV = (float)(GCOEF * LOG2E * gsy * gsy);
    q.gaD = (float)(GCOEF * LOG2E * (gsx * gsx - gsy * gsy));
  }
  __device__ static void mixing(Inv& q, const float* p) {   // Zsp, gq (and gax, aq: Gaussian present)
    const double Zsp = p[0], gq = p[7];
    const bool nog = (p[5] == 0.0f) || (p[6] == 0.0f) || (p[7] == 0.0f);
    q.r.gauss = nog ? 0 : 1;
    q.r.cr = (float)(nog ? Zsp : Zsp * (1.0 - gq));
    q.r.cg = (float)(nog ? 0.0 : Zsp * gq);
    mix_ring(&q.r);
  }

  // The ring group (Re, rq, ecc, fade): the costly part of prepare (float64,
  // one division).  The step kernel keeps the current state's copy and
  // recomputes it only for proposals that move one of these parameters.
  struct RingPart {
    RingInv r;
    float kg, kgr;
  };
  __device__ static bool ring_uses(int t) { return t >= 1 && t <= 4; }
  __device__ static RingPart ring_part(const float* p, const Trig&) {
    Inv q;
    ring(q, p);
    return RingPart{q.r, q.kg, q.kgr};
  }
  __device__ static Inv assemble(const float* p, const Trig& tg, const RingPart& rp) {
    Inv q;
    q.r = rp.r;
    q.kg = rp.kg;
    q.kgr = rp.kgr;
    q.cth = tg.cth;
    q.sth = tg.sth;
    gaussian(q, p);
    mixing(q, p);
    return q;
  }
  __device__ static Inv prepare(const float* p, const Trig& tg) { return assemble(p, tg, ring_part(p, tg)); }
  __device__ static Inv prepare(const float* p) { return prepare(p, trig(p)); }

  // Visibility at (u, v) with rho2 = u^2 + v^2.  The Gaussian exponent
  // gaU U^2 + gaV V^2 is formed as gaV rho2 + (gaU - gaV) U^2.  Without a
  // Gaussian (cg = 0) its term is added as zero, so the code has no branch.
  // HW: both shift phases lie inside the hardware sine's input domain.
  template <class BC = BesselPoly, bool HW = false>
  __device__ static float2 vis(const Inv& q, float u, float v, float rho2, bool any_large, const BC& bc = BC()) {
    const float U = fmaf(u, q.cth, v * q.sth);
    float2 out = ring_vis<HW>(q.r, U, rho2, any_large, bc);
    const float ga = __builtin_amdgcn_exp2f(-fmaf(q.gaD, U * U, q.gaV * rho2));
    float sg, cgs;
    sincos_phase<HW>(q.kgr * U, &sg, &cgs);
    const float gw = q.r.cg * ga;
    out.x = fmaf(gw, cgs, out.x);
    out.y = fmaf(gw, sg, out.y);
    return out;
  }
  template <class BC = BesselPoly>
  __device__ static float2 vis(const Inv& q, float u, float v, bool any_large, const BC& bc = BC()) {
    return vis(q, u, v, fmaf(u, u, v * v), any_large, bc);
  }
};

// ----------------------------------------------------------- 13 parameters
struct Sgra13 {
  static constexpr int NP = 13;
  struct Inv {
    RingInv r;
    float cth, sth;       // ring rotation by th
    float cb, sb;         // Gaussian rotation by beta
    float gaU, gaV;
    float gaD;            // gaU - gaV (exponent gaV rho^2 + gaD Ub^2)
    float kgu, kgv;       // Gaussian phase per unit (u, v) in revolutions: -gr Re (cos alpha, sin alpha)
    int gonly;            // gq == 1: Gaussian only
  };

  // float64 trigonometric invariants of alpha, beta, eta - th and th
  // (parameters 9..12), kept by the step kernel across proposals that move
  // none of them
  struct Trig {
    double ceth, ca, sa;
    float cth, sth, cb, sb;
  };
  __device__ static bool trig_uses(int t) { return t >= 9; }
  __device__ static Trig trig(const float* p);
  // The ring group (Re, rq, ecc, fade and, through cos(eta - th), eta and
  // th): see Sgra9::RingPart.
  struct RingPart {
    RingInv r;
  };
  __device__ static bool ring_uses(int t) { return (t >= 1 && t <= 4) || t == 11 || t == 12; }
  __device__ static RingPart ring_part(const float* p, const Trig& tg);
  __device__ static Inv assemble(const float* p, const Trig& tg, const RingPart& rp);
  __device__ static Inv prepare(const float* p, const Trig& tg) { return assemble(p, tg, ring_part(p, tg)); }
  __device__ static Inv prepare(const float* p) { return prepare(p, trig(p)); }

  // Visibility at (u, v) with rho2 = u^2 + v^2 (see Sgra9::vis).  The ring
  // is skipped for a Gaussian-only source, as in the reference.
  template <class BC = BesselPoly, bool HW = false>
  __device__ static float2 vis(const Inv& q, float u, float v, float rho2, bool any_large, const BC& bc = BC()) {
    float2 out = make_float2(0.0f, 0.0f);
    if (!q.gonly) out = ring_vis<HW>(q.r, fmaf(u, q.cth, v * q.sth), rho2, any_large, bc);
    const float Ub = fmaf(u, q.cb, v * q.sb);
    const float ga = __builtin_amdgcn_exp2f(-fmaf(q.gaD, Ub * Ub, q.gaV * rho2));
    float sg, cgs;
    sincos_phase<HW>(fmaf(q.kgu, u, q.kgv * v), &sg, &cgs);
    const float gw = q.r.cg * ga;
    out.x = fmaf(gw, cgs, out.x);
    out.y = fmaf(gw, sg, out.y);
    return out;
  }
  template <class BC = BesselPoly>
  __device__ static float2 vis(const Inv& q, float u, float v, bool any_large, const BC& bc = BC()) {
    return vis(q, u, v, fmaf(u, u, v * v), any_large, bc);
  }
};

// p: Zsp, Re, rq, ecc, fade, gr, gax, aq, gq, alpha, beta, eta, th
__device__ inline Sgra13::Trig Sgra13::trig(const float* p) {
  const double alpha = p[9], beta = p[10], eta = p[11], th = p[12];
  Trig tg;
  double sn, cs;
  tg.ceth = cos(eta - th);
  sincos(alpha, &tg.sa, &tg.ca);   // one argument reduction per angle
  sincos(th, &sn, &cs);
  tg.cth = (float)cs;
  tg.sth = (float)sn;
  sincos(beta, &sn, &cs);
  tg.cb = (float)cs;
  tg.sb = (float)sn;
  return tg;
}

__device__ inline Sgra13::RingPart Sgra13::ring_part(const float* p, const Trig& tg) {
  RingPart rp;
  const double Re = p[1], rq = p[2], ecc = p[3], fade = p[4];
  ring_coefs(Re * UAS2NRAD, rq, fade, ecc * (1.0 - rq) * tg.ceth, &rp.r);   // d / Re, d = ecc (Re - Ri) cos(eta - th)
  return rp;
}

__device__ inline Sgra13::Inv Sgra13::assemble(const float* p, const Trig& tg, const RingPart& rp) {
  Inv q;
  const double Zsp = p[0], Re = p[1];
  const double gr = p[5], gax = p[6], aq = p[7], gq = p[8];
  const double Rex = Re * UAS2NRAD;
  q.r = rp.r;
  q.cth = tg.cth;
  q.sth = tg.sth;
  q.cb = tg.cb;
  q.sb = tg.sb;
  const bool gauss = (p[6] != 0.0f) && (p[7] != 0.0f) && (p[8] != 0.0f);
  const double gsy = gax * Rex, gsx = aq * gsy;
  q.gaU = (float)(GCOEF * LOG2E * gsx * gsx);
  q.gaV = (float)(GCOEF * LOG2E * gsy * gsy);
  q.gaD = (float)(GCOEF * LOG2E * (gsx * gsx - gsy * gsy));
  q.kgu = (float)(-gr * Rex * tg.ca);
  q.kgv = (float)(-gr * Rex * tg.sa);
  q.gonly = (p[8] == 1.0f) ? 1 : 0;
  q.r.gauss = gauss ? 1 : 0;
  q.r.cr = (float)(gauss ? Zsp * (1.0 - gq) : Zsp);
  q.r.cg = (float)(gauss ? Zsp * gq : 0.0);
  if (q.gonly) { q.r.cr = 0.0f; q.r.cg = (float)(gauss ? Zsp : 0.0); }
  mix_ring(&q.r);
  return q;
}

}  // namespace mhre
