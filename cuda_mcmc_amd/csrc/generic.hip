// Generic path: the project's own model.cuh, compiled into the library as in
// the reference (src/Makefile:44-45 builds gpu_mcmc.cu with -I$(MODELDIR)).
//
// model() keeps the reference contract (ccalcmodel.h): per datum, per pass,
// results in mc->datm[imd], return 0/1/2/-1.  The kernels here fuse what the
// reference splits into calc_chi2_terms (src/gpu_mcmc.cu:273-325),
// calc_chi2 (:331-350), gen_proposal (:361-411), spmp_mcmc (:426-474) and
// adjust_steps (:480-526): one thread per chain runs every pass over every
// datum in order, so pass p+1 may read what pass p stored in datm exactly as
// in the reference, and the chain's whole iteration is one launch.
#include <hip/hip_runtime.h>
#include "mhre_device.h"
#include "fitdesc.h"
#include "engine.h"
#include "ccalcmodel.h"
#include "model.cuh"

#ifndef MHRE_MODEL_NAME
#define MHRE_MODEL_NAME "model.cuh"
#endif

namespace mhre {

// chi^2 of chain ch from the user's model (all passes), written to chi2m.
// Returns the sum over all ndat terms (the reference sums every chi2m entry
// of the chain, including entries no pass rewrote).  Sets *err on iret < 0.
__device__ inline double user_chi2(const FitDesc& f, int ch, int* err) {
  CCalcModel* mc = f.mc;
  const int ndat = f.ndat;
  const int ipt = ch * f.nptot;
  const size_t base = (size_t)ch * ndat;
  for (int ipass = 0; ipass < f.npass; ++ipass) {
    for (int id = 0; id < ndat; ++id) {
      const int imd = (int)(base + id);
      const int iret = model(mc, id, ipt, imd, ipass);
      if (iret == 0) continue;
      if (iret < 0) { *err = iret; continue; }
      float dif = f.dat[id] - f.datm[imd];
      if (iret == 2) dif = atan2f(sinf(dif), cosf(dif));
      f.chi2m[imd] = (dif * dif) * f.std2r[id];
    }
  }
  double s = 0.0;
  for (int id = 0; id < ndat; ++id) s += (double)f.chi2m[base + id];
  return s;
}

__global__ void __launch_bounds__(64) k_eval_user(const FitDesc* fits, const WorkItem* work, int nwork,
                                                  int model_only, int respect_flag) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nwork) return;
  const WorkItem w = work[g];
  const FitDesc& f = fits[w.fit];
  const int ch = w.idx;
  int err = 0;
  if (model_only) {
    // calc_model (src/gpu_mcmc.cu:633-658): pass 0 only, no chi^2
    for (int id = 0; id < f.ndat; ++id) {
      const int iret = model(f.mc, id, ch * f.nptot, ch * f.ndat + id, 0);
      if (iret < 0) err = iret;
    }
  } else {
    double s;
    if (respect_flag && f.flag[ch] == 0) {
      s = 0.0;
      for (int id = 0; id < f.ndat; ++id) s += (double)f.chi2m[(size_t)ch * f.ndat + id];
    } else {
      s = user_chi2(f, ch, &err);
    }
    f.chi2c[ch] = (float)s;
  }
  if (err && f.errflag) atomicExch(f.errflag, err);
}

// The chain's nprm proposals of this iteration (one lane); returns the
// chi^2 evaluations performed.
__device__ unsigned user_one_chain(const FitDesc& f, int ch, int it, const TraceBuf& tr) {
  const int C = f.nchain, P = f.nprm, N = f.nptot;
  const int ibeta = ch / f.nseq;
  const bool burn = it < f.nburn;
  const double beta = (double)f.beta[ibeta];
  float chi2c = f.chi2c[ch];
  int flag = 0, tlast = 0, err = 0;
  float told = 0.0f;
  unsigned nev = 0;
  for (int s = 0; s < P; ++s) {
    const U4 r = philox4x32_10((uint32_t)ch, (uint32_t)it, (uint32_t)s, 0u, f.key0, f.key1);
    int k = pick_index(r.x, P);
    for (int rr = 1; k >= P && rr <= RNG_MAX_RETRY; ++rr)
      k = pick_index(philox4x32_10((uint32_t)ch, (uint32_t)it, (uint32_t)s, (uint32_t)rr, f.key0, f.key1).x, P);
    if (k >= P) k = P - 1;
    const int t = f.ivar[k];
    const size_t pk = (size_t)k * C + ch;
    float* pt = f.pcur + (size_t)ch * N + t;
    const float old = *pt;
    const double u1 = (double)u01(r.y), u2 = (double)u01(r.z);
    double pd = (double)old + (double)f.pstp[pk] * (sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    if (f.pdescr[t] == 2) pd = pd - 6.283185307179586 * rint(pd / 6.283185307179586);
    const float pn = (float)pd;
    const bool inprior = !(pn < f.pmint[t] || pn > f.pmaxt[t]);
    int acc = 0;
    double c2 = 0.0;
    if (inprior) {
      *pt = pn;
      c2 = user_chi2(f, ch, &err);
      ++nev;
      const double x = -0.5 * beta * (c2 - (double)chi2c);
      if (log((double)u01(r.w)) <= x) {
        acc = 1;
        chi2c = (float)c2;
        f.n_acpt[pk] += 1;
      } else {
        *pt = old;
      }
    }
    flag = acc;
    told = old;
    tlast = t;
    if (tr.dec && it >= tr.it0 && it < tr.it0 + tr.nit) {
      const size_t ti = (size_t)f.tr_dec + ((size_t)(it - tr.it0) * P + s) * C + ch;
      tr.dec[ti] = k | ((inprior ? 1 : 0) << 8) | (acc << 9);
      tr.chi2[ti] = inprior ? (float)c2 : 0.0f;
    }
    if (burn) {
      const int cnt = f.n_cnt[pk];
      const int a = cnt % f.nadj;   // oldest history entry, replaced by this outcome
      uint32_t* hw = f.hbits + ((size_t)ch * f.hwords + (a >> 5)) * P + k;
      const uint32_t word = *hw, bit = 1u << (a & 31);
      const int hold = (word & bit) ? 1 : 0;
      *hw = acc ? (word | bit) : (word & ~bit);
      const int sh = f.sumh[pk] + acc - hold;
      f.sumh[pk] = sh;
      const int den = cnt < f.nadj ? cnt : f.nadj;
      const float rate = (float)sh / (float)den;
      const float st = f.pstp[pk];
      f.pstp[pk] = rate < f.adapt_target ? (float)((double)st / 1.01) : (float)((double)st * 1.01);
      f.n_cnt[pk] = cnt + 1;
    }
  }
  f.chi2c[ch] = chi2c;
  f.flag[ch] = flag;
  f.ptent[ch] = told;
  f.ptentn[ch] = tlast;
  if (err && f.errflag) atomicExch(f.errflag, err);
  return nev;
}

// One lane per chain; every fit's chain list is padded to whole waves, so a
// wave's lanes share one fit and its counters take one atomic per wave.
__global__ void __launch_bounds__(64) k_step_user(const FitDesc* fits, const WorkItem* work, int nwork,
                                                  RunClock* clk, TraceBuf tr, int it_off) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (__builtin_amdgcn_readfirstlane(g) >= nwork) return;
  const WorkItem w = work[g];
  const FitDesc& f = wave_fit(fits, w.fit);
  pass_clock(clk);
  const int it = __builtin_amdgcn_readfirstlane(load_clock(clk, it_off));
  if (it >= f.nburn + f.niter) return;
  const bool live = w.idx >= 0;
  const unsigned nev = live ? user_one_chain(f, w.idx, it, tr) : 0u;
  wave_count(f, nev, live ? (unsigned)f.nprm : 0u);
}

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

hipError_t launch_eval_user(const FitDesc* fits, const WorkItem* work, int nwork, int model_only,
                            int respect_flag, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  k_eval_user<<<cdiv(nwork, 64), 64, 0, s>>>(fits, work, nwork, model_only, respect_flag);
  return hipGetLastError();
}

hipError_t launch_step_user(const StepLaunch& L, hipStream_t s) {
  if (L.nwork <= 0) return hipSuccess;
  const TraceBuf tr{L.tr_dec, L.tr_chi2, L.tr_ex, L.tr_it0, L.tr_nit};
  k_step_user<<<cdiv(L.nwork, 64), 64, 0, s>>>(L.fits, L.work, L.nwork, L.clk, tr, L.it_off);
  return hipGetLastError();
}

const char* user_model_name() { return MHRE_MODEL_NAME; }

}  // namespace mhre
