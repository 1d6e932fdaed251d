// MH-RE engine kernels for gfx950.
//
// Reference: src/gpu_mcmc.cu.  The reference launches, per iteration and per
// parameter, gen_proposal, npass x calc_chi2_terms, spmp_mcmc and
// adjust_steps, then replica_exchange (:727-757) -- about 46 launches per
// burn-in iteration with every intermediate in global memory.  Here one launch
// of the step kernel runs all nprm proposals of an iteration for every chain
// with the chain state in registers, and one launch of the exchange kernel
// does the swaps and the output recording:
//
//   k_init_chains    init_pcur (:245-267)
//   k_eval_eht       calc_chi2_terms + calc_chi2 (:273-350) for the native
//                    ring+Gaussian models; also fills datm/chi2m
//   k_step_eht       gen_proposal + calc_chi2_terms + spmp_mcmc + adjust_steps
//                    (:361-526) fused; one wave64 per chain, uv data streamed
//                    through the wave, model visibilities kept in a per-wave
//                    LDS ring for the closure phases, chi^2 accumulated in fp64
//                    per lane and reduced across the wave
//   k_step_quad      the same for systems of quadratic equations, one lane per
//                    chain (wave-uniform coefficients)
//   k_exchange       replica_exchange (:533-626); one lane per temperature
//                    slot, 64 / xlanes sequences per wave
#include <hip/hip_runtime.h>
#include <type_traits>
#include "mhre_device.h"
#include "eht_model.h"
#include "fitdesc.h"
#include "engine.h"

namespace mhre {

// ------------------------------------------------------------------ helpers

__device__ inline int pick_with_retry(uint32_t w0, int n, uint32_t c0, uint32_t c1, uint32_t c2,
                                      uint32_t c3base, uint32_t k0, uint32_t k1) {
  int k = pick_index(w0, n);
  for (int r = 1; k >= n && r <= RNG_MAX_RETRY; ++r)
    k = pick_index(philox4x32_10(c0, c1, c2, c3base + (uint32_t)r, k0, k1).x, n);
  return k >= n ? n - 1 : k;
}

// Box-Muller in float64 (same formula as the oracle, so proposals round to
// the same fp32 value).
__device__ inline double normal64(uint32_t w1, uint32_t w2) {
  const double u1 = (double)u01(w1), u2 = (double)u01(w2);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// ------------------------------------------------------------- init chains

__global__ void k_init_chains(const FitDesc* fits, const WorkItem* work, int nwork) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nwork) return;
  const WorkItem w = work[g];
  const FitDesc& f = fits[w.fit];
  const int ch = w.idx;
  for (int i = 0; i < f.nprm; ++i) {
    const int k = f.ivar[i];
    const uint32_t x = philox4x32_10((uint32_t)ch, RNG_INIT_ITER, (uint32_t)i, RNG_INIT, f.key0, f.key1).x;
    const double lo = f.pmint[k], hi = f.pmaxt[k];
    double p = lo + (hi - lo) * (double)u01(x);
    if (f.pdescr[k] == 2) p = p - 6.283185307179586 * rint(p / 6.283185307179586);
    f.pcur[(size_t)ch * f.nptot + k] = (float)p;
  }
  f.flag[ch] = 1;
}

// -------------------------------------------------- native chi^2 (one wave)

// Wave-uniform view of one fit's native-kernel inputs; built once per wave
// from a uniform FitDesc address so every field lives in scalar registers.
// Element idx of a global array through a 32-bit byte offset from a
// wave-uniform base (one global_load with a scalar base, no 64-bit address
// arithmetic per lane).
template <class T>
__device__ inline T ldg(const GLOBAL T* base, int idx) {
  return *(const GLOBAL T*)((const GLOBAL char*)base + (uint32_t)idx * (uint32_t)sizeof(T));
}
#if defined(__HIP_DEVICE_COMPILE__)   // the host pass has one address space
template <class T>
__device__ inline T ldg(const CONSTANT T* base, int idx) {
  return base[idx];
}
#endif

struct EhtView {
  const GLOBAL float4* vis;
  const GLOBAL float* rho2;
  const GLOBAL int4* clo;
  const GLOBAL int4* clor;
  const GLOBAL float4* clow;
  const CONSTANT int2* tinfo;   // per-tile entries: read with scalar loads
  GLOBAL float* datm;
  GLOBAL float* chi2m;
  long long ndatm;
  int nvis, ncph, ntile, ring_ok, cgroup, ndat, nchain;
};

__device__ inline EhtView eht_view(const FitDesc& f) {
  EhtView v;
  v.vis = gptr(f.vis); v.rho2 = gptr(f.rho2); v.clo = gptr(f.clo); v.clor = gptr(f.clor);
  v.clow = gptr(f.clow); v.tinfo = wave_const(f.tinfo);
  v.datm = gptr(f.datm); v.chi2m = gptr(f.chi2m); v.ndatm = f.ndatm;
  v.nvis = f.nvis; v.ncph = f.ncph; v.ntile = f.ntile; v.ring_ok = f.ring_ok; v.cgroup = f.cgroup;
  v.ndat = f.ndat; v.nchain = f.nchain;
  return v;
}

// Orders this wave's LDS ring stores before its later ring loads (and loads
// before later stores).  LDS operations of one wave execute in order, so
// only compiler reordering must be prevented; global memory is unaffected.
__device__ inline void ring_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Closure term from three model visibilities: the residual
// wrap(obs - (pa + pb - pc)) as arg(e^{i obs} conj(a b conj(c))).
__device__ inline float closure_term(float2 a, float2 b, float2 e, float4 ow, float* model_cph) {
  const float abr = a.x * b.x - a.y * b.y, abi = a.x * b.y + a.y * b.x;
  const float br = abr * e.x + abi * e.y, bi = abi * e.x - abr * e.y;   // a b conj(e)
  const float xr = ow.x * br + ow.y * bi;
  const float xi = ow.y * br - ow.x * bi;
  const float dd = atan2_fast(xi, xr);
  if (model_cph) *model_cph = atan2_fast(bi, br);
  return dd * dd * ow.z;
}

// Model visibility in the ring at LDS byte offset off.
__device__ inline float2 ring_at(const float2* ring, int off) {
  return *(const float2*)((const char*)ring + off);
}

// chi^2 of one chain whose model invariants are q.  Lanes stream 64
// visibilities per step (the next tile's data and the pending closures' inputs
// are fetched before the model is evaluated); model visibilities go to a
// per-wave LDS ring, from which the closures are formed after every cgroup-th
// tile, so that one pass over the closures keeps most lanes busy.
// The fp32 terms are summed per lane in fp32 over blocks of 8 tiles (at most
// ~10 terms), each block sum is added to a per-lane fp64 total, and the
// totals are summed across the wave.
// OUT: also write datm/chi2m of chain ch in the reference layout.  HW: the
// fit's shift phases are known to stay inside the hardware sine's input
// domain (FitDesc::hwphase), so they are not reduced first.
template <class M, bool OUT, bool HW = false>
__device__ double eht_chi2_wave(const EhtView& v, const typename M::Inv& q, float2* ring, const BesselTab& bt,
                                int ch) {
  const int lane = lane_id();
  const bool out = OUT && v.datm != nullptr;
  const int nvr = v.ndat - v.ncph;   // visibilities with an amplitude datum
  double acc = 0.0;
  float part = 0.0f;
  // The packed arrays are padded (vis and rho2 by one tile of zero weight,
  // closures by one wave, tinfo by two entries), so every load below is in
  // bounds and the loads can stay in flight across the model evaluation.
  // Tile t of vis / rho2 and closure c of clor / clow are read through
  // buffer resources at the scalar byte offsets 1024 t, 256 t and 16 c.
  const uint32_t l16 = (uint32_t)lane * 16u, l4 = (uint32_t)lane * 4u;
  const BufRsrc vb = make_buf(v.vis, (uint32_t)(v.ntile + 1) * TILE * 16u);
  const BufRsrc rb = make_buf(v.rho2, (uint32_t)(v.ntile + 1) * TILE * 4u);
  const BufRsrc cb = make_buf(v.clor, (uint32_t)(v.ncph + TILE) * 16u);
  const BufRsrc wb = make_buf(v.clow, (uint32_t)(v.ncph + TILE) * 16u);
  float2* const ring_lane = ring + lane;                       // ring slot of tile t: + 64 (t % 8)
  int c0 = 0;           // first closure not yet formed
  int kleft = v.cgroup;
  int2 ti = ldg(v.tinfo, 1);
  float r2cur = __int_as_float(ldg(v.tinfo, 0).y);
  // tile t needs the large-argument branch when its largest u^2 + v^2 is
  // above q.r.r2thr (RingInv): a compare of the (non-negative) float bits, in
  // scalar registers
  const uint32_t r2_thr = __builtin_amdgcn_readfirstlane(q.r.r2thr);
  // tile t (ring slot J = t % 8, a constant, so the ring store's address is
  // an instruction offset) with its data (d, r2); fetches tile t + 1 into
  // (dn, r2n)
  auto tile = [&](auto J, const int t, const float4& d, const float r2, float4& dn, float& r2n) {
    constexpr int slot = decltype(J)::value;
    dn = buf_load_f4(vb, l16, (uint32_t)(t + 1) * 1024u);
    r2n = buf_load_f(rb, l4, (uint32_t)(t + 1) * 256u);
    const int2 tin = ldg(v.tinfo, t + 2);
    int4 ix = buf_load_i4(cb, l16, (uint32_t)c0 * 16u);
    float4 ow = buf_load_f4(wb, l16, (uint32_t)c0 * 16u);
    const bool any_large = __float_as_uint(r2cur) > r2_thr;
    const float2 V = M::template vis<BesselTab, HW>(q, d.x, d.y, r2, any_large, bt);
    const float amp = __builtin_amdgcn_sqrtf(fmaf(V.x, V.x, V.y * V.y));
    const float r = fmaf(-amp, d.w, d.z);   // (d - |V|) sqrt(w); zero past the amplitude data
    part = fmaf(r, r, part);
    if (out) {
      const int i = t * TILE + lane;
      if (i < nvr) {
        const size_t imd = (size_t)ch * v.ndat + i;
        v.datm[imd] = amp;
        v.chi2m[imd] = r * r;
        const long long tail = (long long)v.nchain * v.ndat + (long long)ch * nvr + i;
        if (tail < v.ndatm) v.datm[tail] = atan2_fast(V.y, V.x);
      }
    }
    {
      // Lanes past nvis write slots of indices RING back, which no pending
      // closure reads (the host checked every closure's window).  Without
      // ring-ordered closures (none, or ring_ok = 0) every tinfo entry names
      // closure 0, so no closure pass runs here.
      ring_lane[slot * TILE] = V;
      const int c1 = ti.x;   // closures whose largest index is in tiles <= t
      if (--kleft == 0 || t + 1 == v.ntile) {
        kleft = v.cgroup;
        if (c1 > c0) {
          ring_fence();
          for (int c = c0; c < c1; c += TILE) {
            if (c != c0) {
              ix = buf_load_i4(cb, l16, (uint32_t)c * 16u);
              ow = buf_load_f4(wb, l16, (uint32_t)c * 16u);
            }
            // lanes past c1 hold closures of later groups, whose legs are
            // not in the ring yet
            if (c + lane < c1) {
              float mc;
              // the legs' ring byte offsets come packed (index % RING times 8)
              const float cterm = closure_term(ring_at(ring, ix.x), ring_at(ring, ix.y), ring_at(ring, ix.z), ow,
                                               out ? &mc : nullptr);
              part += cterm;
              if (out) {
                const size_t imd = (size_t)ch * v.ndat + ix.w;
                v.datm[imd] = mc;
                v.chi2m[imd] = cterm;
              }
            }
          }
          ring_fence();
          c0 = c1;
        }
      }
    }
    r2cur = __int_as_float(ti.y);
    ti = tin;
  };
  // eight tiles (one pass round the ring) per trip: the prefetched tile
  // alternates between the a and b registers, so the prefetch costs no
  // register copies; the fp32 block sum goes into the fp64 total after each
  // trip
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  using S2 = std::integral_constant<int, 2>;
  using S3 = std::integral_constant<int, 3>;
  using S4 = std::integral_constant<int, 4>;
  using S5 = std::integral_constant<int, 5>;
  using S6 = std::integral_constant<int, 6>;
  using S7 = std::integral_constant<int, 7>;
  static_assert(RING / TILE == 8, "one trip of the tile loop goes once round the ring");
  float4 da = buf_load_f4(vb, l16, 0u), db;
  float ra = buf_load_f(rb, l4, 0u), rbv;
  for (int t = 0; t < v.ntile; t += 8) {
    tile(S0{}, t, da, ra, db, rbv);
    if (t + 1 < v.ntile) tile(S1{}, t + 1, db, rbv, da, ra);
    if (t + 2 < v.ntile) tile(S2{}, t + 2, da, ra, db, rbv);
    if (t + 3 < v.ntile) tile(S3{}, t + 3, db, rbv, da, ra);
    if (t + 4 < v.ntile) tile(S4{}, t + 4, da, ra, db, rbv);
    if (t + 5 < v.ntile) tile(S5{}, t + 5, db, rbv, da, ra);
    if (t + 6 < v.ntile) tile(S6{}, t + 6, da, ra, db, rbv);
    if (t + 7 < v.ntile) tile(S7{}, t + 7, db, rbv, da, ra);
    acc += (double)part;
    part = 0.0f;
  }
  if (v.ncph && !v.ring_ok) {
    // General closure layout: recompute the three model visibilities.
    for (int c = lane; c < v.ncph; c += TILE) {
      const int4 ix = ldg(v.clo, c);
      const float4 ow = ldg(v.clow, c);
      const float4 da = ldg(v.vis, ix.x), db = ldg(v.vis, ix.y), de = ldg(v.vis, ix.z);
      float mc;
      const float term = closure_term(M::vis(q, da.x, da.y, true, bt), M::vis(q, db.x, db.y, true, bt),
                                      M::vis(q, de.x, de.y, true, bt), ow, out ? &mc : nullptr);
      acc += (double)term;
      if (out) {
        const size_t imd = (size_t)ch * v.ndat + ix.w;
        v.datm[imd] = mc;
        v.chi2m[imd] = term;
      }
    }
  }
  return wave_sum(acc);
}

// Gather the chain's parameters (lane j holds parameter j) into a uniform
// array and build the model invariants.
template <class M>
__device__ inline void gather_params(float pl, float (&p)[M::NP]) {
#pragma unroll
  for (int j = 0; j < M::NP; ++j) p[j] = bcast(pl, j);
}
template <class M>
__device__ inline typename M::Inv prepare_from_lanes(float pl) {
  float p[M::NP];
  gather_params<M>(pl, p);
  return M::prepare(p);
}

// ------------------------------------------------------------ eval kernel

template <class M, bool OUT>
__global__ void __launch_bounds__(256) k_eval_eht(const FitDesc* fits, const WorkItem* work, int nwork,
                                                  int respect_flag) {
  __shared__ float2 ring_all[4][RING];
  __shared__ float4 btab[2 * BESSEL_TAB_N];
  load_bessel_tab(btab);
  const BesselTab bt{btab};
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int wv = blockIdx.x * 4 + wid;
  if (wv >= nwork) return;
  float2* ring = ring_all[wid];
  const WorkItem w = work[wv];
  const FitDesc& f = wave_fit(fits, w.fit);
  const int ch = __builtin_amdgcn_readfirstlane(w.idx);
  const int lane = lane_id();
  if (respect_flag && f.flag[ch] == 0) {
    double s = 0.0;
    if (f.chi2m)
      for (int i = lane; i < f.ndat; i += TILE) s += (double)f.chi2m[(size_t)ch * f.ndat + i];
    s = wave_sum(s);
    if (lane == 0) f.chi2c[ch] = (float)s;
    return;
  }
  const float pl = lane < f.nptot ? f.pcur[(size_t)ch * f.nptot + lane] : 0.0f;
  const typename M::Inv q = prepare_from_lanes<M>(pl);
  const double c2 = eht_chi2_wave<M, OUT>(eht_view(f), q, ring, bt, ch);
  if (lane == 0) f.chi2c[ch] = (float)c2;
}

// ------------------------------------------------------------ step kernel

// The replica exchange of iteration itp of sequence iseq (reference:
// gpu_mcmc.cu:548-624, the same draws and arithmetic as exchange_wave),
// computed inside one wave whose lane j holds temperature slot j (B <= 64)
// with chi^2 and labels read from c2s and tcs.  Every wave of the sequence
// computes it and keeps what its own slot sl needs: the slot whose row ends
// up in sl, and sl's chi^2 and label after the swaps.  It also does the
// exchange launch's per-slot writes for sl (n_exch) and, in slot 0's wave,
// the sequence's trace entries.  Every operand of the attempts is
// wave-uniform (lane reads; the n_exch count and the accept bits stay in
// scalar registers); labels move with the rows, so sl's label is the one its
// row started with.
struct SlotAfterSwaps {
  int src;
  float c2;
  int tc;
};
__device__ inline SlotAfterSwaps seq_exchange(const FitDesc& f, int iseq, int sl, int itp, const float* c2s,
                                              const int* tcs, const TraceBuf& tr) {
  const int lane = lane_id();
  const int B = f.nbeta, S = f.nseq;
  const bool burn = itp < f.nburn;
  const bool slot = lane < B;
  float c2 = slot ? c2s[(size_t)lane * S + iseq] : 0.0f;
  const int tc0 = slot ? tcs[(size_t)lane * S + iseq] : 0;
  int perm = lane;
  int nx = 0;
  if (B > 1) {
    // attempt a's pair pick and log u on lane a, all drawn up front
    int k_l = 0;
    double lu_l = 0.0, hb_l = 0.0;
    if (slot) {
      const U4 r = philox4x32_10((uint32_t)iseq, (uint32_t)itp, (uint32_t)lane, RNG_EXCH, f.key0, f.key1);
      k_l = pick_with_retry(r.x, B - 1, (uint32_t)iseq, (uint32_t)itp, (uint32_t)lane, RNG_EXCH, f.key0, f.key1);
      lu_l = log((double)u01(r.y));
    }
    if (lane < B - 1) hb_l = -0.5 * ((double)f.beta[lane + 1] - (double)f.beta[lane]);
    uint64_t accm = 0;   // bit a: attempt a swapped
    for (int a = 0; a < B; ++a) {
      const int k = bcast(k_l, a);
      const float ck = bcast(c2, k), ck1 = bcast(c2, k + 1);
      const double x = bcast(hb_l, k) * ((double)ck - (double)ck1);
      if (bcast(lu_l, a) <= x) {
        accm |= 1ull << a;
        const int pk = bcast(perm, k), pk1 = bcast(perm, k + 1);
        if (lane == k) { c2 = ck1; perm = pk1; }
        if (lane == k + 1) { c2 = ck; perm = pk; }
        if (k == sl) ++nx;
      }
    }
    if (tr.ex && sl == 0 && slot && itp >= tr.it0 && itp < tr.it0 + tr.nit)
      tr.ex[(size_t)f.tr_ex + ((size_t)(itp - tr.it0) * B + lane) * S + iseq] =
          k_l | ((int)((accm >> lane) & 1ull) << 16);
  }
  if (!burn && nx && lane == 0) f.n_exch[(size_t)sl * S + iseq] += nx;
  const int src = bcast(perm, sl);
  return SlotAfterSwaps{src, bcast(c2, sl), bcast(tc0, src)};
}

// One MH iteration of chain ch (all nprm single-parameter proposals) by one
// wave.  With f.xfused the launch at offset it_off > 0 of its segment first
// runs the exchange of iteration it - 1 (the state arrays alternate between
// the two sets by the parity of it_off); for it past the fit's last iteration
// only that exchange runs.
template <class M>
__device__ __attribute__((always_inline)) void step_one_chain(const FitDesc& f, int ch, int it, int it_off, float2* ring,
                               typename M::RingPart* rcache, const BesselTab& bt, const TraceBuf& tr) {
  const int lane = lane_id();
  const int C = f.nchain, P = f.nprm, N = f.nptot;
  const int ibeta = ch / f.nseq;
  const bool burn = it < f.nburn;
  const int nit = f.nburn + f.niter;

  // lane-distributed constants and per-slot state (not moved by the
  // exchange: read first, so that they arrive while the exchange runs)
  const int pdl = lane < N ? f.pdescr[lane] : 0;
  const float pminl = lane < N ? f.pmint[lane] : 0.0f;
  const float pmaxl = lane < N ? f.pmaxt[lane] : 0.0f;
  const int ivl = lane < P ? f.ivar[lane] : 0;
  float stp = lane < P ? f.pstp[(size_t)lane * C + ch] : 0.0f;
  int cnt = lane < P ? f.n_cnt[(size_t)lane * C + ch] : 0;
  int nac = lane < P ? f.n_acpt[(size_t)lane * C + ch] : 0;
  int sh = (burn && lane < P) ? f.sumh[(size_t)lane * C + ch] : 0;
  const double beta = (double)f.beta[ibeta];

  // the chain's state before the step (after the previous exchange), and
  // where the state after the step goes
  float* pdst = f.pcur;
  float* cdst = f.chi2c;
  int* tdst = nullptr;
  float pl;
  float chi2c;
  int tc = 0;
  if (f.xfused) {
    const bool odd = (it_off & 1) != 0;
    if (odd) { pdst = f.pcur_b; cdst = f.chi2c_b; }
    tdst = odd ? f.tcur_b : f.tcur;
    if (it_off > 0) {
      const int iseq = ch - ibeta * f.nseq;
      const float* ps = odd ? f.pcur : f.pcur_b;
      const SlotAfterSwaps x = seq_exchange(f, iseq, ibeta, it - 1, odd ? f.chi2c : f.chi2c_b,
                                            odd ? f.tcur : f.tcur_b, tr);
      pl = lane < N ? ps[((size_t)x.src * f.nseq + iseq) * N + lane] : 0.0f;
      chi2c = x.c2;
      tc = x.tc;
      if (it - 1 >= f.nburn && f.record) {
        // the sample of iteration it - 1 in slot ibeta (what the exchange launch records)
        const int itr = it - 1 - f.nburn;
        if (lane < N) {
          const int ip = f.invar[lane];
          if (ip >= 0) f.pout[(((size_t)f.nbeta * ip + ibeta) * f.nseq + iseq) * f.niter + itr] = pl;
        }
        if (lane == 0) {
          f.tout[(size_t)ch * f.niter + itr] = tc;
          if (ibeta == 0) f.chi2[(size_t)iseq * f.niter + itr] = chi2c;
        }
      }
      if (it >= nit) {   // the fit's last exchange: no step
        if (lane < N) pdst[(size_t)ch * N + lane] = pl;
        if (lane == 0) {
          cdst[ch] = chi2c;
          tdst[ch] = tc;
        }
        return;
      }
    } else {
      pl = lane < N ? f.pcur[(size_t)ch * N + lane] : 0.0f;
      chi2c = f.chi2c[ch];
      tc = f.tcur[ch];
    }
  } else {
    pl = lane < N ? f.pcur[(size_t)ch * N + lane] : 0.0f;
    chi2c = f.chi2c[ch];
  }

  // lane j keeps the history word holding entry cnt % nadj of parameter j
  int hw_idx = 0;
  uint32_t hword = 0u;
  if (burn && lane < P) {
    hw_idx = (cnt % f.nadj) >> 5;
    hword = f.hbits[(ch * f.hwords + hw_idx) * P + lane];
  }
  int flag = 0, tlast = 0;
  float told = 0.0f;
  unsigned nev = 0;
  const EhtView view = eht_view(f);
  // the float64 trigonometric invariants of the current state, and its ring
  // group (M::RingPart) in the wave's LDS slot rcache[rci]: a proposal that
  // moves neither recomputes only the cheap rest of the invariants.  A
  // proposal that moves the ring group writes its own to the other slot,
  // which becomes the current one if it is accepted (so no copy of it stays
  // in registers across the chi^2).
  typename M::Trig tg;
  int rci = 0;
  {
    float p[M::NP];
    gather_params<M>(pl, p);
    tg = M::trig(p);
    rcache[0] = M::ring_part(p, tg);
  }

  // This iteration's random numbers, substep s on lane s: the coordinate
  // pick, the Gaussian step and log u of the accept test (drawn for every
  // substep; counter-based, so unused draws shift nothing).
  int k_l = 0;
  double z_l = 0.0, lu_l = 0.0;
  if (lane < P) {
    const U4 r = philox4x32_10((uint32_t)ch, (uint32_t)it, (uint32_t)lane, 0u, f.key0, f.key1);
    k_l = pick_with_retry(r.x, P, (uint32_t)ch, (uint32_t)it, (uint32_t)lane, 0u, f.key0, f.key1);
    z_l = normal64(r.y, r.z);
    lu_l = log((double)u01(r.w));
  }

  for (int s = 0; s < P; ++s) {
    const int k = bcast(k_l, s);
    const int t = bcast(ivl, k);
    const float old = bcast(pl, t);
    const float sd = bcast(stp, k);
    double pd = (double)old + (double)sd * bcast(z_l, s);
    if (bcast(pdl, t) == 2) pd = pd - 6.283185307179586 * rint(pd / 6.283185307179586);
    const float pn = (float)pd;
    const bool inprior = !(pn < bcast(pminl, t) || pn > bcast(pmaxl, t));
    int acc = 0;
    double c2 = 0.0;
    if (inprior) {
      if (lane == t) pl = pn;
      float p[M::NP];
      gather_params<M>(pl, p);
      const typename M::Trig tn = M::trig_uses(t) ? M::trig(p) : tg;
      const bool ring_moved = M::ring_uses(t);
      typename M::RingPart rp;
      if (ring_moved) {
        rp = M::ring_part(p, tn);
        rcache[rci ^ 1] = rp;
      } else {
        rp = rcache[rci];
      }
      const typename M::Inv q = M::assemble(p, tn, rp);
      // one copy of the chi^2 loop per phase form (the branch is per wave)
      c2 = f.hwphase ? eht_chi2_wave<M, false, true>(view, q, ring, bt, ch)
                     : eht_chi2_wave<M, false, false>(view, q, ring, bt, ch);
      ++nev;
      const double x = -0.5 * beta * (c2 - (double)chi2c);
      if (bcast(lu_l, s) <= x) {
        acc = 1;
        chi2c = (float)c2;
        tg = tn;
        if (ring_moved) rci ^= 1;
        if (lane == k) ++nac;
      } else if (lane == t) {
        pl = old;
      }
    }
    flag = acc;
    told = old;
    tlast = t;
    if (burn && lane == k) {
      // history entry cnt % nadj: the oldest one, replaced by this outcome
      // (the division-free forms of mhre_device.h give the reference's bits)
      const int nadj = f.nadj;
      const float rnadj = 1.0f / (float)nadj;
      const uint32_t bit = 1u << (mod_small(cnt, nadj, rnadj) & 31);
      const int hold = (hword & bit) ? 1 : 0;
      hword = acc ? (hword | bit) : (hword & ~bit);
      sh += acc - hold;
      const int den = cnt < nadj ? cnt : nadj;
      stp = rate_below(sh, den, f.adapt_mid, f.adapt_tie) ? step_shrink(stp) : step_grow(stp);
      ++cnt;
      const int w2 = mod_small(cnt, nadj, rnadj) >> 5;
      if (w2 != hw_idx) {
        f.hbits[(ch * f.hwords + hw_idx) * P + lane] = hword;
        hw_idx = w2;
        hword = f.hbits[(ch * f.hwords + hw_idx) * P + lane];
      }
    }
    if (tr.dec && it >= tr.it0 && it < tr.it0 + tr.nit && lane == 0) {
      const size_t ti = (size_t)f.tr_dec + ((size_t)(it - tr.it0) * P + s) * C + ch;
      tr.dec[ti] = k | ((inprior ? 1 : 0) << 8) | (acc << 9);
      tr.chi2[ti] = inprior ? (float)c2 : 0.0f;
    }
  }

  if (lane < N) pdst[(size_t)ch * N + lane] = pl;
  if (lane < P) {
    f.pstp[(size_t)lane * C + ch] = stp;
    f.n_cnt[(size_t)lane * C + ch] = cnt;
    f.n_acpt[(size_t)lane * C + ch] = nac;
    if (burn) {
      f.sumh[(size_t)lane * C + ch] = sh;
      f.hbits[(ch * f.hwords + hw_idx) * P + lane] = hword;
    }
  }
  if (lane == 0) {
    cdst[ch] = chi2c;
    if (tdst) tdst[ch] = tc;
    f.flag[ch] = flag;
    f.ptent[ch] = told;
    f.ptentn[ch] = tlast;
    if (f.nevals) {
      atomicAdd(&f.nevals[0], (unsigned long long)nev);
      atomicAdd(&f.nevals[1], (unsigned long long)P);
    }
  }
}

// Next work item of a persistent step kernel's queue q of head set hs: one
// atomic per wave (leader lane), broadcast to the wave.
__device__ inline int grab_item(RunClock* clk, int hs, int q) {
  const int leader = __builtin_amdgcn_readfirstlane(lane_id());
  int v = 0;
  if (lane_id() == leader) v = atomicAdd(&clk->qhead[hs][q], 1);
  return __builtin_amdgcn_readfirstlane(v);
}

// The last iteration at which a launch has work for fit f: its last step,
// or, with f.xfused and a launch that is not the first of its segment, the
// exchange of that step.
__device__ inline int last_busy_it(const FitDesc& f, int it_off) {
  return f.nburn + f.niter - 1 + ((f.xfused && it_off > 0) ? 1 : 0);
}

// Grid form: work-group b serves the range of XCD b % NXCD.  Persistent form:
// each wave drains its own XCD's queue, then helps the other ranges in turn;
// every queue head only grows, so every wave leaves once all are exhausted.
// The queues use head set it_off & 1; the first thread clears the other set
// for the next launch of the segment.
template <class M, int WPB, bool PERSIST>
__global__ void __launch_bounds__(64 * WPB) k_step_eht(const FitDesc* fits, const WorkItem* work, XcdRanges xr,
                                                       RunClock* clk, TraceBuf tr, int qslot, int it_off) {
  __shared__ float2 ring_all[WPB][RING];
  __shared__ float4 btab[2 * BESSEL_TAB_N];
  __shared__ typename M::RingPart rcache_all[WPB][2];
  pass_clock(clk);
  if (PERSIST && blockIdx.x == 0 && threadIdx.x < NXCD) clk->qhead[(it_off + 1) & 1][qslot * NXCD + threadIdx.x] = 0;
  load_bessel_tab(btab);
  const BesselTab bt{btab};
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  float2* ring = ring_all[wid];
  typename M::RingPart* rcache = rcache_all[wid];
  const int it = __builtin_amdgcn_readfirstlane(load_clock(clk, it_off));
  const int xb = (int)(blockIdx.x % NXCD);
  if (!PERSIST) {
    const int wv = xr.off[xb] + (int)(blockIdx.x / NXCD) * WPB + wid;
    if (wv >= xr.off[xb + 1]) return;
    const WorkItem w = work[wv];
    const FitDesc& f = wave_fit(fits, w.fit);
    if (it > last_busy_it(f, it_off)) return;
    step_one_chain<M>(f, __builtin_amdgcn_readfirstlane(w.idx), it, it_off, ring, rcache, bt, tr);
  } else {
    const int hs = it_off & 1;
    for (int r = 0; r < NXCD; ++r) {
      const int x = (xb + r) % NXCD;
      const int base = xr.off[x], len = xr.off[x + 1] - base;
      const int q = qslot * NXCD + x;
      for (int j = grab_item(clk, hs, q); j < len; j = grab_item(clk, hs, q)) {
        const WorkItem w = work[base + j];
        const FitDesc& f = wave_fit(fits, w.fit);
        if (it <= last_busy_it(f, it_off))
          step_one_chain<M>(f, __builtin_amdgcn_readfirstlane(w.idx), it, it_off, ring, rcache, bt, tr);
      }
    }
  }
}

// --------------------------------------------------- quadratic systems (TPC)

// Per equation e the packed coefficients are qcoef[fit][e][QS]:
//   n(n+1)/2 symmetric products (i <= j, Q_ij + Q_ji off the diagonal),
//   n linear terms b_e, then rhs = dat[e] and the weight std2r[e].
template <int NQ>
struct QuadSys {
  static constexpr int NS = NQ * (NQ + 1) / 2;
  static constexpr int QS = NS + NQ + 2;
  // qc: the fit's coefficients; the step kernel passes a wave-uniform
  // constant-space pointer (scalar loads), the evaluation kernel a global one
  // Each equation as sum_i x_i (b_i + sum_{j>=i} c_ij x_j): NS + NQ FMAs
  // and no table of the NS products x_i x_j, so a lane holds little more
  // than x (the NQ inner sums are independent chains).
  template <class CP>
  __device__ static float lhs(CP c, const float (&x)[NQ]) {
    float l = 0.0f;
    int m = 0;
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      float t = c[NS + i];
#pragma unroll
      for (int j = i; j < NQ; ++j) t = fmaf(c[m++], x[j], t);
      l = fmaf(x[i], t, l);
    }
    return l;
  }
  template <class CP>
  __device__ static double chi2(CP qc, int neq, const float (&x)[NQ]) {
    double acc = 0.0;
    for (int e = 0; e < neq; ++e) {
      const CP c = qc + (size_t)e * QS;
      float l = 0.0f;
      int m = 0;
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        float t = c[NS + i];
#pragma unroll
        for (int j = i; j < NQ; ++j) t = fmaf(c[m++], x[j], t);
        l = fmaf(x[i], t, l);
      }
      const float r = c[NS + NQ] - l;
      acc += (double)(r * r * c[NS + NQ + 1]);
    }
    return acc;
  }
  // Weighted residuals sqrt(w_e) (lhs_e - rhs_e) of the first neq equations
  // (at most QROW_MAXEQ), sw = the row table's sqrt(w_e).
  template <class CP>
  __device__ static void residuals(CP qc, const CONSTANT float* sw, int neq, const float (&x)[NQ],
                                   float (&r)[QROW_MAXEQ]) {
#pragma unroll
    for (int e = 0; e < QROW_MAXEQ; ++e) {
      r[e] = 0.0f;
      if (e < neq) {
        const CP c = qc + (size_t)e * QS;
        r[e] = (lhs(c, x) - c[NS + NQ]) * sw[e];
      }
    }
  }
  // The residuals after x_t moves from x[t] to xn (d = xn - x[t]), from the
  // chain's LDS block blk = rows + t * QROW_TS (every equation side by side;
  // rows past the fit's equations are zero, as are their residuals); returns
  // their sum of squares (fp32: at most QROW_MAXEQ terms).
  __device__ static float moved(const float* blk, const float (&x)[NQ], float xn, float d,
                                const float (&r)[QROW_MAXEQ], float (&rn)[QROW_MAXEQ]) {
    static_assert(QROW_MAXEQ == 8, "two float4 per term");
    const float4* q = reinterpret_cast<const float4*>(blk);
    float g[8];
    {
      const float4 b0 = q[2 * NQ], b1 = q[2 * NQ + 1];
      g[0] = b0.x; g[1] = b0.y; g[2] = b0.z; g[3] = b0.w;
      g[4] = b1.x; g[5] = b1.y; g[6] = b1.z; g[7] = b1.w;
    }
#pragma unroll
    for (int j = 0; j <= NQ; ++j) {
      const float v = j < NQ ? x[j] : xn;
      const int k = j < NQ ? 2 * j : 2 * NQ + 2;   // term n + 1: a_tt times the new value
      const float4 c0 = q[k], c1 = q[k + 1];
      g[0] = fmaf(c0.x, v, g[0]); g[1] = fmaf(c0.y, v, g[1]);
      g[2] = fmaf(c0.z, v, g[2]); g[3] = fmaf(c0.w, v, g[3]);
      g[4] = fmaf(c1.x, v, g[4]); g[5] = fmaf(c1.y, v, g[5]);
      g[6] = fmaf(c1.z, v, g[6]); g[7] = fmaf(c1.w, v, g[7]);
    }
    float a = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      rn[e] = fmaf(d, g[e], r[e]);
      a = fmaf(rn[e], rn[e], a);
    }
    return a;
  }
};

template <int NQ>
__global__ void __launch_bounds__(256) k_eval_quad(const FitDesc* fits, const WorkItem* work, int nwork) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nwork) return;
  const WorkItem w = work[g];
  const FitDesc& f = fits[w.fit];
  float x[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) x[i] = f.pcur[(size_t)w.idx * NQ + i];
  f.chi2c[w.idx] = (float)QuadSys<NQ>::chi2(gptr(f.qcoef), f.quad_neq, x);
}

// The chain's nprm proposals of this iteration (one lane) on the state x,
// chi2c held by the caller; returns the chi^2 evaluations performed.
// CACHE: in the burn-in stage the adaptation counters n_cnt and sumh of all
// the chain's parameters are read once (one coalesced row per parameter) and
// kept in registers for the iteration, instead of one scattered
// read-modify-write per proposal (lanes pick different parameters); the
// step sizes and acceptance counts of all its parameters sit in the wave's
// LDS for the iteration (stage: [NQ][64] floats then [NQ][64] ints, column
// = lane), read and written back once.  Scattered single-word accesses per
// proposal had made the launch HBM-bound (every lane a cache line of its own).
// ROWS: the weighted residual of every equation is evaluated in full once,
// then kept in registers; a proposal, which moves one coordinate, updates
// them from its block of the fit's row table in LDS (QROW_* layout) with
// n + 3 operations per equation instead of n(n+1)/2 + n.  Its chi^2 is the
// same sum of squares to float rounding (sqrt(w) r)^2 against r^2 w.
template <int NQ, bool CACHE, bool ROWS = false>
__device__ unsigned quad_substeps(const FitDesc& f, int ch, int it, const TraceBuf& tr, float (&x)[NQ],
                                  float& chi2c, const float* rows = nullptr, float* stage = nullptr) {
  const int C = f.nchain, P = f.nprm;
  const CONSTANT float* qc = wave_const(f.qcoef);
  const int neq = __builtin_amdgcn_readfirstlane(f.quad_neq);
  GLOBAL float* pstp = gptr(f.pstp);
  GLOBAL int* n_acpt = gptr(f.n_acpt);
  GLOBAL int* n_cnt = gptr(f.n_cnt);
  GLOBAL int* sumh = gptr(f.sumh);
  GLOBAL uint32_t* hbits = gptr(f.hbits);
  const CONSTANT int* ivar = wave_const(f.ivar);
  const CONSTANT int* pdescr = wave_const(f.pdescr);
  const CONSTANT float* pmint = wave_const(f.pmint);
  const CONSTANT float* pmaxt = wave_const(f.pmaxt);
  const int ibeta = ch / f.nseq;
  const bool burn = it < f.nburn;
  const double beta = (double)gptr(f.beta)[ibeta];
  const int nadj = __builtin_amdgcn_readfirstlane(f.nadj);
  const int hwn = __builtin_amdgcn_readfirstlane(f.hwords);
  const float rnadj = 1.0f / (float)nadj;
  int cnt_c[NQ], sh_c[NQ];
  if (CACHE && burn) {
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      cnt_c[i] = i < P ? n_cnt[(size_t)i * C + ch] : 0;
      sh_c[i] = i < P ? sumh[(size_t)i * C + ch] : 0;
    }
  }
  const int ln = lane_id();
  float* sp = stage;
  int* sa = reinterpret_cast<int*>(stage + NQ * 64);
  if (CACHE) {
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      if (i < P) {
        sp[i * 64 + ln] = pstp[i * C + ch];
        sa[i * 64 + ln] = n_acpt[i * C + ch];
      }
    }
  }
  float res[QROW_MAXEQ], resn[QROW_MAXEQ];
  if (ROWS) QuadSys<NQ>::residuals(qc, wave_const(f.qrow) + NQ * QROW_TS(NQ), neq, x, res);
  // CACHE: bit k = a proposal of parameter k was accepted
  uint32_t chg = 0;
  int flag = 0, tlast = 0;
  float told = 0.0f;
  unsigned nev = 0;
  for (int s = 0; s < P; ++s) {
    const U4 r = philox4x32_10((uint32_t)ch, (uint32_t)it, (uint32_t)s, 0u, f.key0, f.key1);
    const int k = pick_with_retry(r.x, P, (uint32_t)ch, (uint32_t)it, (uint32_t)s, 0u, f.key0, f.key1);
    const int t = ivar[k];
    const int pk = k * C + ch;   // 32-bit element offsets: the host bounds nprm * nchain * hwords
    float old = 0.0f;
#pragma unroll
    for (int i = 0; i < NQ; ++i) old = (i == t) ? x[i] : old;
    const float sd = CACHE ? sp[k * 64 + ln] : pstp[pk];
    // burn-in: the acceptance-history word of this outcome's entry
    // cnt % nadj (the oldest one), read before the proposal so that its
    // latency overlaps the chi^2
    int cnt = 0, sh0 = 0, ha = 0;
    uint32_t hword = 0;
    if (burn) {
      if (CACHE) {
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
          cnt = (i == k) ? cnt_c[i] : cnt;
          sh0 = (i == k) ? sh_c[i] : sh0;
        }
      } else {
        cnt = n_cnt[pk];
        sh0 = sumh[pk];
      }
      ha = mod_small(cnt, nadj, rnadj);
      hword = hbits[(ch * hwn + (ha >> 5)) * P + k];
    }
    // The Gaussian step in fp32 with the hardware log2 and cosine (which takes
    // revolutions, so cos(2 pi u2) is one instruction; its absolute error is
    // below 2e-7): z is within about 2e-6 of the float64 Box-Muller value at
    // |z| ~ 5, so the proposal may differ from the float64 one by ~1e-6 step
    // sizes.
    const float z = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u01(r.y))) *
                    __builtin_amdgcn_cosf(u01(r.z));
    float pn = fmaf(sd, z, old);
    if (pdescr[t] == 2) pn = wrap_angle(pn);
    const bool inprior = !(pn < pmint[t] || pn > pmaxt[t]);
    int acc = 0;
    double c2 = 0.0;
    if (inprior) {
      if (ROWS) {
        c2 = (double)QuadSys<NQ>::moved(rows + t * QROW_TS(NQ), x, pn, pn - old, res, resn);
      } else {
#pragma unroll
        for (int i = 0; i < NQ; ++i) x[i] = (i == t) ? pn : x[i];
        c2 = QuadSys<NQ>::chi2(qc, neq, x);
      }
      ++nev;
      const double xx = -0.5 * beta * (c2 - (double)chi2c);
      // log u in fp32 (hardware log2): its error, ~1e-7 |log u|, is far inside the tie band
      if ((double)(0.69314718055994531f * __builtin_amdgcn_logf(u01(r.w))) <= xx) {
        acc = 1;
        chi2c = (float)c2;
        if (CACHE) {
          sa[k * 64 + ln] += 1;
          chg |= 1u << k;
        } else {
          n_acpt[pk] += 1;
        }
        if (ROWS) {
#pragma unroll
          for (int i = 0; i < NQ; ++i) x[i] = (i == t) ? pn : x[i];
#pragma unroll
          for (int e = 0; e < QROW_MAXEQ; ++e) res[e] = resn[e];
        }
      } else if (!ROWS) {
#pragma unroll
        for (int i = 0; i < NQ; ++i) x[i] = (i == t) ? old : x[i];
      }
    }
    flag = acc;
    told = old;
    tlast = t;
    if (burn) {
      const uint32_t bit = 1u << (ha & 31);
      const int hold = (hword & bit) ? 1 : 0;
      // stored only when the entry changes: the lanes of a wave touch one
      // history line each, so every lane left out saves a line
      if (acc != hold) hbits[(ch * hwn + (ha >> 5)) * P + k] = hword ^ bit;
      const int sh = sh0 + acc - hold;
      const int den = cnt < nadj ? cnt : nadj;
      const float nsd = rate_below(sh, den, f.adapt_mid, f.adapt_tie) ? step_shrink(sd) : step_grow(sd);
      if (CACHE) sp[k * 64 + ln] = nsd;
      else pstp[pk] = nsd;
      if (CACHE) {
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
          cnt_c[i] = (i == k) ? cnt + 1 : cnt_c[i];
          sh_c[i] = (i == k) ? sh : sh_c[i];
        }
      } else {
        sumh[pk] = sh;
        n_cnt[pk] = cnt + 1;
      }
    }
    if (tr.dec && it >= tr.it0 && it < tr.it0 + tr.nit) {
      const size_t ti = (size_t)f.tr_dec + ((size_t)(it - tr.it0) * P + s) * C + ch;
      gptr(tr.dec)[ti] = k | ((inprior ? 1 : 0) << 8) | (acc << 9);
      gptr(tr.chi2)[ti] = inprior ? (float)c2 : 0.0f;
    }
  }
  // write back; acceptance counts only where a proposal was accepted (a
  // lane left out touches no line)
  if (CACHE) {
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      if (i < P) {
        if (burn) {
          n_cnt[i * C + ch] = cnt_c[i];
          sumh[i * C + ch] = sh_c[i];
          pstp[i * C + ch] = sp[i * 64 + ln];
        }
        if ((chg >> i) & 1u) n_acpt[i * C + ch] = sa[i * 64 + ln];
      }
    }
  }
  gptr(f.flag)[ch] = flag;
  gptr(f.ptent)[ch] = told;
  gptr(f.ptentn)[ch] = tlast;
  return nev;
}

template <int NQ>
__device__ unsigned quad_one_chain(const FitDesc& f, int ch, int it, const TraceBuf& tr) {
  GLOBAL float* pcur = gptr(f.pcur);
  float x[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) x[i] = pcur[(size_t)ch * NQ + i];
  float chi2c = gptr(f.chi2c)[ch];
  const unsigned nev = quad_substeps<NQ, false>(f, ch, it, tr, x, chi2c);
#pragma unroll
  for (int i = 0; i < NQ; ++i) pcur[(size_t)ch * NQ + i] = x[i];
  gptr(f.chi2c)[ch] = chi2c;
  return nev;
}


// One lane per chain; all chains of a wave belong to one fit (the host pads
// every fit's chain list to whole waves), so the coefficients are uniform.
// Registers are capped so that 8 waves fit on a SIMD: the lane-per-chain
// work waits on the scalar coefficient loads and the state loads, which more
// resident waves hide.
template <int NQ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
k_step_quad(const FitDesc* fits, const WorkItem* work, int nwork, RunClock* clk, TraceBuf tr, int it_off) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (__builtin_amdgcn_readfirstlane(g) >= nwork) return;   // whole waves (the list is padded to 64)
  const WorkItem w = work[g];
  // every lane of a wave belongs to the same fit: its coefficients and sizes
  // are read through scalar loads
  const FitDesc& f = wave_fit(fits, w.fit);
  pass_clock(clk);
  const int it = __builtin_amdgcn_readfirstlane(load_clock(clk, it_off));
  if (it >= f.nburn + f.niter) return;
  const bool live = w.idx >= 0;   // padding lanes only take part in the counting
  const unsigned nev = live ? quad_one_chain<NQ>(f, w.idx, it, tr) : 0u;
  wave_count(f, nev, live ? (unsigned)f.nprm : 0u);
}

// ---------------------------------------------------------------- exchange

// Sequences iseq0 .. iseq0 + 64/W - 1 of the fit, W = f.xlanes lanes each:
// lane g*W + j holds temperature slot j of sequence iseq0 + g.  The swap
// attempts of one sequence are sequential (reference: gpu_mcmc.cu:548-624);
// the groups of a wave run theirs side by side, exchanging values with lane
// shuffles inside each group.  The state is read from (ps, cs, ts) and
// written to (pd, cd, td), which may be the same arrays; with swaps = false
// the state is only copied (no attempt, no sample, no trace).
__device__ void exchange_wave(const FitDesc& f, int iseq0, int it, bool swaps, const float* ps, const float* cs,
                              const int* ts, float* pd, float* cd, int* td, const TraceBuf& tr) {
  const int lane = lane_id();
  const int B = f.nbeta, S = f.nseq, N = f.nptot, W = f.xlanes;
  const bool burn = it < f.nburn;
  const int sl = lane & (W - 1);   // temperature slot
  const int gb = lane - sl;        // first lane of this sequence's group
  const int iseq = iseq0 + lane / W;
  const bool seq_ok = iseq < S;
  const bool slot = seq_ok && sl < B;
  const size_t me = (size_t)sl * S + iseq;
  float c2 = slot ? cs[me] : 0.0f;
  int tc = slot ? ts[me] : 0;
  int perm = sl;
  int nex = 0;
  if (swaps && B > 1) {
    // attempt a's pair pick and log u on slot a, all drawn up front
    int k_l = 0;
    double lu_l = 0.0, hb_l = 0.0;
    if (slot) {
      const U4 r = philox4x32_10((uint32_t)iseq, (uint32_t)it, (uint32_t)sl, RNG_EXCH, f.key0, f.key1);
      k_l = pick_with_retry(r.x, B - 1, (uint32_t)iseq, (uint32_t)it, (uint32_t)sl, RNG_EXCH, f.key0, f.key1);
      lu_l = log((double)u01(r.y));
    }
    // -0.5 (beta[k+1] - beta[k]) on slot k
    if (sl < B - 1) hb_l = -0.5 * ((double)f.beta[sl + 1] - (double)f.beta[sl]);
    for (int a = 0; a < B; ++a) {   // every lane takes part in every shuffle
      const int k = __shfl(k_l, gb + a);
      const float ck = __shfl(c2, gb + k), ck1 = __shfl(c2, gb + k + 1);
      const double x = __shfl(hb_l, gb + k) * ((double)ck - (double)ck1);
      const bool acc = __shfl(lu_l, gb + a) <= x;
      const int tk = __shfl(tc, gb + k), tk1 = __shfl(tc, gb + k + 1);
      const int pk = __shfl(perm, gb + k), pk1 = __shfl(perm, gb + k + 1);
      if (acc) {
        if (sl == k) { c2 = ck1; tc = tk1; perm = pk1; if (!burn) ++nex; }
        if (sl == k + 1) { c2 = ck; tc = tk; perm = pk; }
      }
      if (tr.ex && sl == 0 && seq_ok && it >= tr.it0 && it < tr.it0 + tr.nit)
        tr.ex[(size_t)f.tr_ex + ((size_t)(it - tr.it0) * B + a) * S + iseq] = k | ((acc ? 1 : 0) << 16);
    }
  }
  // move parameter rows: slot lane receives the row that started in slot perm
  float row[MAXP];
#pragma unroll
  for (int j = 0; j < MAXP; ++j)
    row[j] = (slot && j < N) ? ps[((size_t)perm * S + iseq) * N + j] : 0.0f;
  if (slot && (perm != sl || pd != ps)) {
#pragma unroll
    for (int j = 0; j < MAXP; ++j)
      if (j < N) pd[me * N + j] = row[j];
  }
  if (slot) {
    cd[me] = c2;
    td[me] = tc;
    if (!burn && nex) f.n_exch[me] += nex;
  }
  if (swaps && !burn && f.record) {
    const int itr = it - f.nburn;
    if (slot) {
#pragma unroll
      for (int j = 0; j < MAXP; ++j) {
        if (j < N) {
          const int ip = f.invar[j];
          if (ip >= 0) f.pout[(((size_t)B * ip + sl) * S + iseq) * f.niter + itr] = row[j];
        }
      }
      f.tout[me * f.niter + itr] = tc;
    }
    if (sl == 0 && seq_ok) f.chi2[(size_t)iseq * f.niter + itr] = c2;
  }
}

// nbeta > 64: the reference's serial loop on one lane (gpu_mcmc.cu:548-624).
__device__ void exchange_serial(const FitDesc& f, int iseq, int it, bool burn, const TraceBuf& tr) {
  if (lane_id() != 0) return;
  const int B = f.nbeta, S = f.nseq, N = f.nptot;
  for (int a = 0; a < B; ++a) {
    const U4 r = philox4x32_10((uint32_t)iseq, (uint32_t)it, (uint32_t)a, RNG_EXCH, f.key0, f.key1);
    const int k = pick_with_retry(r.x, B - 1, (uint32_t)iseq, (uint32_t)it, (uint32_t)a, RNG_EXCH,
                                  f.key0, f.key1);
    const size_t k0 = (size_t)k * S + iseq, k1 = (size_t)(k + 1) * S + iseq;
    const double x = -0.5 * ((double)f.beta[k + 1] - (double)f.beta[k]) *
                     ((double)f.chi2c[k0] - (double)f.chi2c[k1]);
    const bool acc = log((double)u01(r.y)) <= x;
    if (acc) {
      int ti = f.tcur[k0]; f.tcur[k0] = f.tcur[k1]; f.tcur[k1] = ti;
      for (int j = 0; j < N; ++j) {
        float tmp = f.pcur[k0 * N + j]; f.pcur[k0 * N + j] = f.pcur[k1 * N + j]; f.pcur[k1 * N + j] = tmp;
      }
      float tc = f.chi2c[k0]; f.chi2c[k0] = f.chi2c[k1]; f.chi2c[k1] = tc;
      if (!burn) f.n_exch[k0] += 1;
    }
    if (tr.ex && it >= tr.it0 && it < tr.it0 + tr.nit)
      tr.ex[(size_t)f.tr_ex + ((size_t)(it - tr.it0) * B + a) * S + iseq] = k | ((acc ? 1 : 0) << 16);
  }
  if (!burn && f.record) {
    const int itr = it - f.nburn;
    for (int ib = 0; ib < B; ++ib) {
      const size_t me = (size_t)ib * S + iseq;
      for (int j = 0; j < N; ++j) {
        const int ip = f.invar[j];
        if (ip >= 0) f.pout[(((size_t)B * ip + ib) * S + iseq) * f.niter + itr] = f.pcur[me * N + j];
      }
      f.tout[me * f.niter + itr] = f.tcur[me];
    }
    f.chi2[(size_t)iseq * f.niter + itr] = f.chi2c[iseq];
  }
}

// close_k = 0: the exchange of iteration clk.it + it_off of fits without
// xfused.  close_k = K > 0: the closing launch of a segment of K iterations
// (clk.xit is its first): for each xfused fit, the exchange of the segment's
// last iteration from the set that launch wrote, into the first set; or, for
// a fit whose last iteration came earlier in the segment, a copy of the state
// that the launch after it left in the second set.  Its first thread then
// clears the queue heads and moves the clock on.
__global__ void __launch_bounds__(64) k_exchange(const FitDesc* fits, const WorkItem* xwork, XcdRanges xr,
                                                 RunClock* clk, TraceBuf tr, int it_off, int close_k) {
  const int base = close_k > 0 ? __atomic_load_n(&clk->xit, __ATOMIC_RELAXED)
                               : __atomic_load_n(&clk->it, __ATOMIC_RELAXED);
  const int xb = (int)(blockIdx.x % NXCD);
  const int wv = xr.off[xb] + (int)(blockIdx.x / NXCD);
  if (wv < xr.off[xb + 1]) {
    const WorkItem w = xwork[wv];
    const FitDesc& f = wave_fit(fits, w.fit);
    const int lf = f.nburn + f.niter - 1;   // the fit's last iteration
    if (close_k == 0) {
      const int it = base + it_off;
      if (it <= lf) {
        if (f.nbeta <= 64) exchange_wave(f, w.idx, it, true, f.pcur, f.chi2c, f.tcur, f.pcur, f.chi2c, f.tcur, tr);
        else exchange_serial(f, w.idx, it, it < f.nburn, tr);
      }
    } else if (lf >= base) {
      const int last = base + close_k - 1;
      const int at = lf < last ? lf + 1 - base : close_k - 1;   // launch offset that wrote the state
      const bool odd = (at & 1) != 0;
      if (lf >= last || odd)
        exchange_wave(f, w.idx, last, lf >= last, odd ? f.pcur_b : f.pcur, odd ? f.chi2c_b : f.chi2c,
                      odd ? f.tcur_b : f.tcur, f.pcur, f.chi2c, f.tcur, tr);
    }
  }
  if (close_k > 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    for (int q = 0; q < NQSLOT * NXCD; ++q) {
      clk->qhead[0][q] = 0;
      clk->qhead[1][q] = 0;
    }
    __atomic_store_n(&clk->it, base + close_k, __ATOMIC_RELAXED);
  }
}

// ------------------------------------- quadratic systems: step + exchange

// One whole MH-RE iteration of quadratic-system sequences in one launch: the
// lane-per-chain proposals of k_step_quad, then the replica exchange of every
// sequence among the lanes that hold its temperature slots, with the chain
// states still in registers (parameter rows move by lane shuffles, not through
// memory).  A wave holds G = 64 / W sequences (W = xlanes, ladders of at most
// 64 rungs); lane l is temperature slot l / G of sequence iseq0 + l % G, so
// the G sequences of one slot are adjacent in the state arrays.  The random
// counters are those of k_step_quad and k_exchange (reference update rule:
// gpu_mcmc.cu:361-526 then :533-626), so the chains are the same bit for bit.
// The burn-in counters held in registers (quad_substeps<NQ, true>) take the
// register cap to 7 waves per SIMD.
// The body of k_iter_quad for the wave's G sequences from iseq0; stage is
// the wave's LDS for its chains' step sizes and acceptance counts
// (2 * NQ * 64 words), and with ROWS (fits that have update rows) qlds the
// wave's LDS for the rows (QROW_LDS floats).
template <int NQ, bool ROWS>
__device__ void iter_quad_wave(const FitDesc& f, int iseq0, int it, const TraceBuf& tr, float* qlds, float* stage) {
  const int lane = lane_id();
  const int B = f.nbeta, S = f.nseq, W = f.xlanes;
  const int lg = 6 - __builtin_ctz((unsigned)W);   // log2 G
  const int G = 1 << lg;
  const int sl = lane >> lg, g = lane & (G - 1);
  const int iseq = __builtin_amdgcn_readfirstlane(iseq0) + g;
  const bool seq_ok = iseq < S;
  const bool slot = seq_ok && sl < B;
  const int ch = sl * S + iseq;
  const bool burn = it < f.nburn;
  GLOBAL float* pcur = gptr(f.pcur);
  GLOBAL float* chi2c_a = gptr(f.chi2c);
  GLOBAL int* tcur_a = gptr(f.tcur);
  float x[NQ];
  float c2 = 0.0f;
  int tc = 0;
#pragma unroll
  for (int i = 0; i < NQ; ++i) x[i] = slot ? pcur[(size_t)ch * NQ + i] : 0.0f;
  if (slot) {
    c2 = chi2c_a[ch];
    tc = tcur_a[ch];
  }
  if (ROWS) {
    // the fit's update rows into this wave's LDS, every lane taking part
    constexpr int n4 = NQ * QROW_TS(NQ) / 4;
    const GLOBAL float4* src = (const GLOBAL float4*)gptr(f.qrow);
    float4* dst = reinterpret_cast<float4*>(qlds);
    for (int i = lane; i < n4; i += 64) dst[i] = src[i];
    ring_fence();
  }
  const unsigned nev = slot ? quad_substeps<NQ, true, ROWS>(f, ch, it, tr, x, c2, qlds, stage) : 0u;
  wave_count(f, nev, slot ? (unsigned)f.nprm : 0u);

  // exchange: as exchange_wave, with the partner of slot k at lane k * G + g.
  // Each lane follows where its starting state goes (pos: no shuffle per
  // attempt); afterwards one forward permute gives every slot the slot its
  // state started in (perm), and the labels follow from it.
  int pos = sl;
  int nex = 0;
  if (B > 1) {
    int k_l = 0;
    double lu_l = 0.0, hb_l = 0.0;
    if (slot) {
      const U4 r = philox4x32_10((uint32_t)iseq, (uint32_t)it, (uint32_t)sl, RNG_EXCH, f.key0, f.key1);
      k_l = pick_with_retry(r.x, B - 1, (uint32_t)iseq, (uint32_t)it, (uint32_t)sl, RNG_EXCH, f.key0, f.key1);
      lu_l = log((double)u01(r.y));
    }
    if (sl < B - 1) hb_l = -0.5 * ((double)f.beta[sl + 1] - (double)f.beta[sl]);
    for (int a = 0; a < B; ++a) {   // every lane takes part in every shuffle
      const int k = __shfl(k_l, (a << lg) + g);
      const int lk = (k << lg) + g, lk1 = lk + G;
      const float ck = __shfl(c2, lk), ck1 = __shfl(c2, lk1);
      const double xx = __shfl(hb_l, lk) * ((double)ck - (double)ck1);
      const bool acc = __shfl(lu_l, (a << lg) + g) <= xx;
      if (acc) {
        if (sl == k) { c2 = ck1; if (!burn) ++nex; }
        if (sl == k + 1) c2 = ck;
        pos = pos == k ? k + 1 : (pos == k + 1 ? k : pos);
      }
      if (tr.ex && sl == 0 && seq_ok && it >= tr.it0 && it < tr.it0 + tr.nit)
        tr.ex[(size_t)f.tr_ex + ((size_t)(it - tr.it0) * B + a) * S + iseq] = k | ((acc ? 1 : 0) << 16);
    }
  }
  // pos is a permutation of each sequence's slots (lanes past the ladder keep
  // their own), so the forward permute writes every lane once
  const int perm = __builtin_amdgcn_ds_permute(((pos << lg) + g) << 2, sl);
  const int src = (perm << lg) + g;
  tc = __shfl(tc, src);
  // slot lane receives the row that started in slot perm
  float row[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) row[i] = __shfl(x[i], src);
  if (slot) {
#pragma unroll
    for (int i = 0; i < NQ; ++i) pcur[(size_t)ch * NQ + i] = row[i];
    chi2c_a[ch] = c2;
    tcur_a[ch] = tc;
    if (!burn && nex) gptr(f.n_exch)[ch] += nex;
  }
  if (!burn && f.record) {
    const int itr = it - f.nburn;
    if (slot) {
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        const int ip = f.invar[i];
        if (ip >= 0) f.pout[(((size_t)B * ip + sl) * S + iseq) * f.niter + itr] = row[i];
      }
      f.tout[(size_t)ch * f.niter + itr] = tc;
    }
    if (sl == 0 && seq_ok) f.chi2[(size_t)iseq * f.niter + itr] = c2;
  }
}

// The rows form keeps every equation's residual in registers; its LDS
// (rows and stage, 27 KB per work-group at NQ = 8) already limits it to 5
// waves per SIMD, which is also its register cap.  The full form takes 16 KB
// per work-group and the register cap of 6 waves (at 7 it spills).
template <int NQ, bool ROWS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ROWS ? 5 : 6)))
k_iter_quad(const FitDesc* fits, const WorkItem* work, int nwork, RunClock* clk, TraceBuf tr, int it_off) {
  const int wv = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (wv >= nwork) return;
  pass_clock(clk);
  const WorkItem w = work[wv];
  const FitDesc& f = wave_fit(fits, w.fit);
  const int it = __builtin_amdgcn_readfirstlane(load_clock(clk, it_off));
  // 256-thread work-groups: 4 waves
  __shared__ __align__(16) float qlds[ROWS ? 4 : 1][ROWS ? QROW_LDS : 4];
  __shared__ float qstage[4][2 * NQ * 64];
  if (it < f.nburn + f.niter)
    iter_quad_wave<NQ, ROWS>(f, w.idx, it, tr, ROWS ? qlds[threadIdx.x >> 6] : nullptr, qstage[threadIdx.x >> 6]);
}

// ---------------------------------------------------------------- launchers

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

hipError_t launch_init(const FitDesc* fits, const WorkItem* work, int nwork, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  k_init_chains<<<cdiv(nwork, 256), 256, 0, s>>>(fits, work, nwork);
  return hipGetLastError();
}

hipError_t launch_eval_eht(int model, const FitDesc* fits, const WorkItem* work, int nwork, int outputs,
                           int respect_flag, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  const int g = cdiv(nwork, 4);
  if (model == MODEL_SGRA9) {
    if (outputs) k_eval_eht<Sgra9, true><<<g, 256, 0, s>>>(fits, work, nwork, respect_flag);
    else k_eval_eht<Sgra9, false><<<g, 256, 0, s>>>(fits, work, nwork, respect_flag);
  } else {
    if (outputs) k_eval_eht<Sgra13, true><<<g, 256, 0, s>>>(fits, work, nwork, respect_flag);
    else k_eval_eht<Sgra13, false><<<g, 256, 0, s>>>(fits, work, nwork, respect_flag);
  }
  return hipGetLastError();
}

// Work-groups of the grid form: NXCD x (longest range / WPB).
static inline int xcd_grid(const XcdRanges& xr, int per_group) {
  int mx = 0;
  for (int x = 0; x < NXCD; ++x) mx = xr.off[x + 1] - xr.off[x] > mx ? xr.off[x + 1] - xr.off[x] : mx;
  return NXCD * cdiv(mx, per_group);
}

// Dynamic LDS a step launch reserves (unused by the kernel) so that at most
// waves_per_simd waves of it are resident per SIMD: a work-group of WPB waves
// puts one wave on each of WPB SIMDs, so 4 * cap / WPB work-groups may share a
// compute unit, and each then takes lds_per_cu / that many bytes (in 512-byte
// allocation granules).  0 when no cap is asked for or the static LDS already
// enforces it.
static inline size_t occupancy_lds(const StepLaunch& L, int wpb, size_t static_lds) {
  if (L.waves_per_simd <= 0 || L.lds_per_cu <= 0) return 0;
  const int groups = (4 * L.waves_per_simd) / wpb;
  if (groups <= 0) return 0;
  const size_t want = ((size_t)L.lds_per_cu / (size_t)groups) / 512 * 512;
  return want > static_lds ? want - static_lds : 0;
}

template <class M, int WPB, bool PERSIST>
static void launch_step_form(const StepLaunch& L, const TraceBuf& tr, hipStream_t s) {
  constexpr size_t static_lds =
      sizeof(float2) * WPB * RING + sizeof(float4) * 2 * BESSEL_TAB_N + sizeof(typename M::RingPart) * 2 * WPB;
  const size_t dyn = occupancy_lds(L, WPB, static_lds);
  const int g = PERSIST ? L.grid : xcd_grid(L.xr, WPB);
  k_step_eht<M, WPB, PERSIST><<<g, 64 * WPB, dyn, s>>>(L.fits, L.work, L.xr, L.clk, tr, L.qslot, L.it_off);
}

template <class M>
static void launch_step_m(const StepLaunch& L, hipStream_t s) {
  const TraceBuf tr{L.tr_dec, L.tr_chi2, L.tr_ex, L.tr_it0, L.tr_nit};
  if (L.persistent) {
    switch (L.wpb) {
      case 1: launch_step_form<M, 1, true>(L, tr, s); break;
      case 2: launch_step_form<M, 2, true>(L, tr, s); break;
      default: launch_step_form<M, 4, true>(L, tr, s); break;
    }
  } else {
    switch (L.wpb) {
      case 1: launch_step_form<M, 1, false>(L, tr, s); break;
      case 2: launch_step_form<M, 2, false>(L, tr, s); break;
      default: launch_step_form<M, 4, false>(L, tr, s); break;
    }
  }
}

hipError_t launch_step_eht(int model, const StepLaunch& L, hipStream_t s) {
  if (L.nwork <= 0) return hipSuccess;
  if (model == MODEL_SGRA9) launch_step_m<Sgra9>(L, s);
  else launch_step_m<Sgra13>(L, s);
  return hipGetLastError();
}

hipError_t launch_eval_quad(int n, const FitDesc* fits, const WorkItem* work, int nwork, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  switch (n) {
    case 2: k_eval_quad<2><<<cdiv(nwork, 256), 256, 0, s>>>(fits, work, nwork); break;
    case 4: k_eval_quad<4><<<cdiv(nwork, 256), 256, 0, s>>>(fits, work, nwork); break;
    case 8: k_eval_quad<8><<<cdiv(nwork, 256), 256, 0, s>>>(fits, work, nwork); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_step_quad(int n, const StepLaunch& L, hipStream_t s) {
  if (L.nwork <= 0) return hipSuccess;
  const TraceBuf tr{L.tr_dec, L.tr_chi2, L.tr_ex, L.tr_it0, L.tr_nit};
  const int g = cdiv(L.nwork, 256);
  switch (n) {
    case 2: k_step_quad<2><<<g, 256, 0, s>>>(L.fits, L.work, L.nwork, L.clk, tr, L.it_off); break;
    case 4: k_step_quad<4><<<g, 256, 0, s>>>(L.fits, L.work, L.nwork, L.clk, tr, L.it_off); break;
    case 8: k_step_quad<8><<<g, 256, 0, s>>>(L.fits, L.work, L.nwork, L.clk, tr, L.it_off); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <bool ROWS>
static hipError_t launch_iter_quad_form(int n, const StepLaunch& L, hipStream_t s) {
  const TraceBuf tr{L.tr_dec, L.tr_chi2, L.tr_ex, L.tr_it0, L.tr_nit};
  const int g = cdiv(L.nwork, 4);   // one work item per wave, 4 waves per work-group
  switch (n) {
    case 2: k_iter_quad<2, ROWS><<<g, 256, 0, s>>>(L.fits, L.work, L.nwork, L.clk, tr, L.it_off); break;
    case 4: k_iter_quad<4, ROWS><<<g, 256, 0, s>>>(L.fits, L.work, L.nwork, L.clk, tr, L.it_off); break;
    case 8: k_iter_quad<8, ROWS><<<g, 256, 0, s>>>(L.fits, L.work, L.nwork, L.clk, tr, L.it_off); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_iter_quad(int n, bool rows, const StepLaunch& L, hipStream_t s) {
  if (L.nwork <= 0) return hipSuccess;
  return rows ? launch_iter_quad_form<true>(n, L, s) : launch_iter_quad_form<false>(n, L, s);
}

// A closing launch with no sequence to exchange still runs one work-group, to
// move the clock on.
hipError_t launch_exchange(const StepLaunch& L, const WorkItem* xwork, const XcdRanges& xr, hipStream_t s) {
  const TraceBuf tr{L.tr_dec, L.tr_chi2, L.tr_ex, L.tr_it0, L.tr_nit};
  k_exchange<<<xr.off[NXCD] > 0 ? xcd_grid(xr, 1) : 1, 64, 0, s>>>(L.fits, xwork, xr, L.clk, tr, L.it_off,
                                                                     L.close_k);
  return hipGetLastError();
}

}  // namespace mhre
