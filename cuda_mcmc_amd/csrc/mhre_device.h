// Device-side building blocks of the MH-RE engine for gfx950 (CDNA4).
//
//  * Philox4x32-10 counter RNG (Salmon et al., SC'11) and the engine's draws:
//    uniform in (0,1], integer pick with redraw, Box-Muller normal.  The
//    stream layout (counter = chain, iteration, substep, purpose) replaces the
//    reference's per-thread cuRAND XORWOW states (src/gpu_mcmc.cu:228-239).
//  * fp32 special functions used by the visibility model, written from their
//    mathematical definitions: Cody-Waite reduced sincos with least-squares kernels,
//    atan2 from a Chebyshev-fitted odd polynomial, and the Bessel-derived
//    B(y) = J1(x)/x and C(y) = J2(x)/x^2 (y = x^2) from Chebyshev-fitted
//    polynomials on x <= 8 and Hankel's asymptotic expansion beyond.
//  * wave64 helpers (lane broadcast, cross-lane sums).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "approx_coeffs.h"

namespace mhre {

constexpr uint32_t PHILOX_M0 = 0xD2511F53u;
constexpr uint32_t PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u;
constexpr uint32_t PHILOX_W1 = 0xBB67AE85u;

constexpr uint32_t RNG_EXCH = 0x40000000u;
constexpr uint32_t RNG_INIT = 0x80000000u;
constexpr uint32_t RNG_INIT_ITER = 0xFFFFFFFFu;
constexpr int RNG_MAX_RETRY = 8;

struct U4 { uint32_t x, y, z, w; };

__host__ __device__ inline U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
    const uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    const uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  return U4{c0, c1, c2, c3};
}

// Uniform in (0, 1]: ((w >> 8) + 1) * 2^-24, exact in fp32.
__host__ __device__ inline float u01(uint32_t w) { return (float)((w >> 8) + 1u) * 5.9604644775390625e-8f; }

// floor(n * u) with u = (w + 1) * 2^-32; equals n only when w == 2^32 - 1.
__host__ __device__ inline int pick_index(uint32_t w, int n) {
  return (int)(((uint64_t)(uint32_t)n * ((uint64_t)w + 1ull)) >> 32);
}

__device__ inline float box_muller(uint32_t w1, uint32_t w2) {
  const float u1 = u01(w1), u2 = u01(w2);
  return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
}

// ------------------------------------------------- step-size adaptation
// The reference's rule (src/gpu_mcmc.cu:493, :518-522): modn = n_cnt % nadj,
// rate = (float)sumh / (float)min(nadj, n_cnt), then the float pstp /= 1.01
// when rate < target, else pstp *= 1.01 (both in double).  The forms below
// give the same bits without integer or floating-point divisions.

// cnt % n for 0 <= cnt, n >= 1, rn = 1.0f / n: below 2^22 the quotient
// estimate trunc(cnt * rn) is within one of floor(cnt / n) (relative error
// below 2^-23), so one correction each way is exact.
__host__ __device__ inline int mod_small(int cnt, int n, float rn) {
  if (cnt >= (1 << 22)) return cnt % n;
  int a = cnt - (int)((float)cnt * rn) * n;
  a += a < 0 ? n : 0;
  a -= a >= n ? n : 0;
  return a;
}

// (float)sh / (float)den < T for a float target T > 0, given the midpoint
// mid = (pred(T) + T) / 2 (pred: the next float toward zero) and
// tie_below = 1 when pred(T) has an even significand: the correctly rounded
// quotient lies below T exactly when sh / den lies below mid, or on it when
// the tie goes to pred(T).  mid * den is exact in double (25 significant bits
// times den < 2^28).  den = 0: sh / 0 is -inf (below), +inf or NaN (not).
__host__ __device__ inline bool rate_below(int sh, int den, double mid, int tie_below) {
  if (den <= 0) return sh < 0;
  const double md = mid * (double)den, s = (double)sh;
  return s < md || (s == md && tie_below != 0);
}

// mid and tie_below of rate_below for the target T (a float above zero).
inline void rate_threshold(float T, double* mid, int* tie_below) {
  const float pred = nextafterf(T, 0.0f);
  uint32_t pb;
  memcpy(&pb, &pred, sizeof pb);
  *mid = 0.5 * ((double)pred + (double)T);
  *tie_below = (pb & 1u) == 0 ? 1 : 0;
}

// (float)((double)sd / 1.01): q0 = sd * RN(1 / 1.01) corrected once by its
// FMA residual, the correctly rounded double quotient (Markstein); zeros and
// infinities, whose residual is not finite or loses the sign, take q0, which
// is exact for them.  Identical to the division for every float sd
// (tests/step_adapt_check.hip checks all 2^32 bit patterns).
__host__ __device__ inline float step_shrink(float sd) {
#pragma clang fp contract(off)
  constexpr double k = 1.01, rk = 1.0 / 1.01;
  const double x = (double)sd;
  const double q0 = x * rk;
  const double q1 = fma(fma(-q0, k, x), rk, q0);
  return (float)((x == 0.0 || __builtin_isinf(x)) ? q0 : q1);
}

__host__ __device__ inline float step_grow(float sd) { return (float)((double)sd * 1.01); }

// Angle wrap into [-pi, pi] in float64 (shared bit-for-bit with the oracle).
__host__ __device__ inline float wrap_angle(float p) {
  const double two_pi = 6.283185307179586;
  const double d = (double)p;
  return (float)(d - two_pi * rint(d / two_pi));
}

// ---------------------------------------------------------------- fp32 math

__device__ inline float poly_eval(const float* c, int n, float t) {
  float acc = c[n - 1];
  for (int k = n - 2; k >= 0; --k) acc = fmaf(acc, t, c[k]);
  return acc;
}

template <int N>
__device__ inline float horner(const float (&c)[N], float t) {
  float acc = c[N - 1];
#pragma unroll
  for (int k = N - 2; k >= 0; --k) acc = fmaf(acc, t, c[k]);
  return acc;
}

// sin and cos of theta: k = rint(theta/pi), r = theta - k*pi in [-pi/2, pi/2]
// (two-constant Cody-Waite), least-squares odd / even kernels of degree 9 / 10,
// sign (-1)^k.
__device__ inline void sincos_fast(float theta, float* s, float* c) {
  const float PI_HI = 3.14159274101257324219f;      // fp32(pi)
  const float PI_LO = -8.74227800037248274685e-08f; // pi - PI_HI
  const float INV_PI = 0.318309886183790671538f;
  const float k = rintf(theta * INV_PI);
  float r = fmaf(-k, PI_HI, theta);
  r = fmaf(-k, PI_LO, r);
  const float r2 = r * r;
  float sp = horner(SIN_T, r2);
  float sn = fmaf(r * r2, sp, r);
  float cs = fmaf(r2, horner(COS_T, r2), 1.0f);
  const uint32_t flip = ((uint32_t)(int)k & 1u) << 31;
  *s = __uint_as_float(__float_as_uint(sn) ^ flip);
  *c = __uint_as_float(__float_as_uint(cs) ^ flip);
}

// sin(2 pi h) and cos(2 pi h), h in revolutions: the reduction
// r = h - rint(h) into [-1/2, 1/2], exact in fp32 (h - floor(h) is not: it
// rounds 1 + h for small negative h), then the hardware sine and cosine,
// which take their argument in revolutions (input domain |r| <= 256).
// Measured on MI355X against float64: max abs error 1.25e-7.  Callers keep
// their phase constants in revolutions (1/(2 pi) folded in once per chain),
// so an evaluation costs four instructions.
__device__ inline void sincosrev(float h, float* s, float* c) {
  const float r = h - rintf(h);
  *s = __builtin_amdgcn_sinf(r);
  *c = __builtin_amdgcn_cosf(r);
}

// The same for |h| <= 255, inside the hardware's input domain, without the
// reduction: the instructions reduce their argument themselves, exactly
// (measured on MI355X against float64 over |h| <= 0.5, 1, 2, 4, 16, 64 and 255
// revolutions: max abs error 1.25e-7 with or without the explicit step).
// Two instructions fewer per evaluation; the host enables it only for fits
// whose phases it has bounded below that (FitDesc::hwphase).
__device__ inline void sincosrev_hw(float h, float* s, float* c) {
  *s = __builtin_amdgcn_sinf(h);
  *c = __builtin_amdgcn_cosf(h);
}

// sincosrev, or without the reduction (HW) for arguments known to lie inside
// the hardware's input domain.
template <bool HW>
__device__ inline void sincos_phase(float h, float* s, float* c) {
  if (HW) sincosrev_hw(h, s, c);
  else sincosrev(h, s, c);
}

// atan2(y, x) in (-pi, pi]: t = min/max in [0,1], atan(t) = t P(t^2), octant fix-up.
__device__ inline float atan2_fast(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  const float t = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
  float p = t * horner(ATAN_P, t * t);
  p = ay > ax ? 1.57079632679489661923f - p : p;
  p = x < 0.0f ? 3.14159265358979323846f - p : p;
  return copysignf(p, y);
}

// B(y) = J1(x)/x and C(y) = J2(x)/x^2, x = sqrt(y).
//   y <= 64: Chebyshev-fitted polynomials in t = y/32 - 1.
//   y  > 64: Hankel asymptotics J_n ~ sqrt(2/(pi x)) (P_n cos chi_n - Q_n sin chi_n),
//            chi_1 = x - 3pi/4, chi_2 = chi_1 - pi/2 (one shared sincos).
__device__ inline void bessel_BC_poly(float y, float* B, float* C) {
  const float t = fmaf(y, 0.03125f, -1.0f);
  *B = horner(BESSEL_B, t);
  *C = horner(BESSEL_C, t);
}

__device__ inline void bessel_BC_asym(float y, float* B, float* C) {
  const float x = __builtin_amdgcn_sqrtf(y);
  const float ix = __builtin_amdgcn_rcpf(x);
  const float z = 0.125f * ix;
  const float z2 = z * z;
  const float P1 = horner(HANKEL_P1, z2), Q1 = z * horner(HANKEL_Q1, z2);
  const float P2 = horner(HANKEL_P2, z2), Q2 = z * horner(HANKEL_Q2, z2);
  const float amp = __builtin_amdgcn_sqrtf(0.636619772367581343f * ix);
  float s, c;
  sincos_fast(x - 2.35619449019234492885f, &s, &c);
  const float j1 = amp * (P1 * c - Q1 * s);
  const float j2 = amp * (P2 * s + Q2 * c);
  *B = j1 * ix;
  *C = j2 * ix * ix;
}

// Full-range evaluation for one lane (used where divergence is harmless).
__device__ inline void bessel_BC(float y, float* B, float* C) {
  if (y <= 64.0f) bessel_BC_poly(y, B, C);
  else bessel_BC_asym(y, B, C);
}

// B and C for y <= 64 from polynomials (no table needed).  The argument is y
// itself (SCALE 1).
struct BesselPoly {
  static constexpr float SCALE = 1.0f;
  __device__ void operator()(float y, float* B, float* C) const { bessel_BC_poly(y, B, C); }
};

// B and C for y in [0, 64] from the piecewise quadratics of BESSEL_TAB,
// staged in LDS by the kernel (2 float4 per quarter-unit interval of y:
// B0 B1 B2 C0, C1 C2 - -).  The argument is z = SCALE * y = 4y (callers fold
// the 4 into their per-chain constants); interval j = trunc(z), s = z - j.
struct BesselTab {
  static constexpr float SCALE = 4.0f;
  const float4* tab;
  __device__ void operator()(float z, float* B, float* C) const {
    const int j = (int)z;
    const float s = __builtin_amdgcn_fractf(z);   // z - j exactly, z >= 0
    const float4 a = tab[2 * j], b = tab[2 * j + 1];
    *B = fmaf(fmaf(a.z, s, a.y), s, a.x);
    *C = fmaf(fmaf(b.y, s, b.x), s, a.w);
  }
};

// Copies BESSEL_TAB into the work-group's LDS table (all threads take part;
// call before any thread of the group leaves).
__device__ inline void load_bessel_tab(float4* tab) {
  float* t = reinterpret_cast<float*>(tab);
  for (int i = (int)threadIdx.x; i < 8 * BESSEL_TAB_N; i += (int)blockDim.x) t[i] = BESSEL_TAB[i];
  __syncthreads();
}

// ------------------------------------------------------- address spaces

// Global address space for the device pass (plain pointers in the host pass,
// which never dereferences them): loads become global_load with scalar
// bases instead of flat loads, which also count against the LDS counter.
#if defined(__HIP_DEVICE_COMPILE__)
#define GLOBAL __attribute__((address_space(1)))
#define CONSTANT __attribute__((address_space(4)))
#else
#define GLOBAL
#define CONSTANT
#endif
template <class T>
__device__ inline GLOBAL T* gptr(T* p) { return (GLOBAL T*)p; }

// A pointer to data that no kernel writes while it runs, known by the caller
// to be the same on every lane: made provably wave-uniform (scalar registers)
// in the constant address space, so loads through it with uniform offsets
// become scalar loads whose results feed the VALU from scalar registers.
template <class T>
__device__ inline const CONSTANT T* wave_const(const T* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (const CONSTANT T*)(((uint64_t)hi << 32) | lo);
}


// ------------------------------------------------------- buffer resources

// A raw buffer over [base, base + bytes) (stride 0, the gfx9 data format
// word).  A load through it adds a wave-uniform offset held in a scalar
// register to the lane's own byte offset, so stepping a wave through an array
// costs scalar adds only; a load past the end returns zero.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __amdgpu_buffer_rsrc_t BufRsrc;
__device__ inline BufRsrc make_buf(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ inline float4 buf_load_f4(BufRsrc r, uint32_t lane_off, uint32_t uni_off) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)lane_off, (int)uni_off, 0);
  return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}
__device__ inline int4 buf_load_i4(BufRsrc r, uint32_t lane_off, uint32_t uni_off) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)lane_off, (int)uni_off, 0);
  return make_int4((int)v[0], (int)v[1], (int)v[2], (int)v[3]);
}
__device__ inline float buf_load_f(BufRsrc r, uint32_t lane_off, uint32_t uni_off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)lane_off, (int)uni_off, 0));
}
#else   // the host pass never runs device code
struct BufRsrc {};
__device__ inline BufRsrc make_buf(const void*, uint32_t) { return BufRsrc{}; }
__device__ inline float4 buf_load_f4(BufRsrc, uint32_t, uint32_t) { return float4{}; }
__device__ inline int4 buf_load_i4(BufRsrc, uint32_t, uint32_t) { return int4{}; }
__device__ inline float buf_load_f(BufRsrc, uint32_t, uint32_t) { return 0.0f; }
#endif

// ---------------------------------------------------------------- wave64

__device__ inline int lane_id() { return (int)(threadIdx.x & 63u); }

__device__ inline float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ inline int bcast(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ inline double bcast(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v from another lane of the same 16-lane row, picked by the DPP control
// CTRL (both 32-bit halves moved by DPP, no LDS round trip).
template <int CTRL>
__device__ inline double dpp_double(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// double of one lane, as a wave-uniform value
__device__ inline double readlane_double(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the 64 lanes (all active), the same on every lane: inside each
// 16-lane row by DPP (lane ^ 1, lane ^ 2 by quad permutes, then the half-row
// and row mirrors, which pair quads and half-rows once every lane of a quad,
// then of a half-row, holds the same partial sum), then the four row sums.
__device__ inline double wave_sum(double v) {
  v += dpp_double<0xB1>(v);    // quad_perm [1,0,3,2]
  v += dpp_double<0x4E>(v);    // quad_perm [2,3,0,1]
  v += dpp_double<0x141>(v);   // row_half_mirror
  v += dpp_double<0x140>(v);   // row_mirror
  return (readlane_double(v, 0) + readlane_double(v, 16)) + (readlane_double(v, 32) + readlane_double(v, 48));
}

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace mhre
