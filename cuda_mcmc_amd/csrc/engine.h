// Host-callable launchers of the engine kernels (engine.hip, generic.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "fitdesc.h"
#include "mhre_device.h"

namespace mhre {

// The descriptor of fit `fit`, which must be the same on every lane of the
// wave.  Descriptors are written by the host before a launch and by no
// kernel, so the uniform address is taken through the constant address space
// and every field is read with a scalar load; the returned plain reference
// keeps that address space once the callers are inlined.
__device__ inline const FitDesc& wave_fit(const FitDesc* fits, int fit) {
  return *(const FitDesc*)wave_const(fits + __builtin_amdgcn_readfirstlane(fit));
}

// Step launches: the first thread of the grid hands the segment's first
// iteration on to the segment's closing exchange launch (RunClock); every
// thread reads its iteration with load_clock (clk.it plus the launch's offset
// in the segment).  Kept as two calls so that a kernel can do the hand-over
// before its set-up (which, in the fused step kernel, keeps its register
// allocation down).
__device__ inline void pass_clock(RunClock* clk) {
  if (blockIdx.x == 0 && threadIdx.x == 0) clk->xit = clk->it;
}
__device__ inline int load_clock(const RunClock* clk, int it_off) {
  return __atomic_load_n(&clk->it, __ATOMIC_RELAXED) + it_off;
}

// Adds the wave's performed evaluations and proposals to the fit's counters
// with one atomic each (every lane of the wave, padding included, calls it).
__device__ inline void wave_count(const FitDesc& f, unsigned nev, unsigned nprop) {
  unsigned a = nev, b = nprop;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    a += (unsigned)__shfl_xor((int)a, m, 64);
    b += (unsigned)__shfl_xor((int)b, m, 64);
  }
  if (f.nevals && lane_id() == 0) {
    atomicAdd(&f.nevals[0], (unsigned long long)a);
    atomicAdd(&f.nevals[1], (unsigned long long)b);
  }
}

// Optional record of the decisions of iterations it0 .. it0+nit-1.
struct TraceBuf {
  int* dec;      // [nit][nprm][nchain] : k | inprior << 8 | accept << 9
  float* chi2;   // [nit][nprm][nchain]
  int* ex;       // [nit][nbeta][nseq] : k | accept << 16
  int it0, nit;
};

struct StepLaunch {
  const FitDesc* fits;   // device
  const WorkItem* work;  // device, step work list
  int nwork;
  XcdRanges xr;          // wave-per-chain kernels: per-XCD ranges of work
  RunClock* clk;         // device
  int it_off;            // the launch's offset in its segment: iteration clk.it + it_off
  int close_k;           // exchange launch: 0 = the exchange of iteration clk.it + it_off (fits
                         // without xfused), K > 0 = closing launch of a segment of K iterations
  int wpb;               // waves per work-group (1, 2, 4)
  int persistent;        // 1: grid-stride work queue
  int grid;              // work-groups for the persistent form
  int qslot;             // RunClock::qhead slot of the persistent form
  int waves_per_simd;    // > 0: cap on resident waves per SIMD (LDS reserved per work-group)
  int lds_per_cu;        // bytes of LDS per compute unit
  int* tr_dec; float* tr_chi2; int* tr_ex; int tr_it0; int tr_nit;   // optional decision trace
};

hipError_t launch_init(const FitDesc* fits, const WorkItem* work, int nwork, hipStream_t s);
hipError_t launch_eval_eht(int model, const FitDesc* fits, const WorkItem* work, int nwork, int outputs,
                           int respect_flag, hipStream_t s);
hipError_t launch_step_eht(int model, const StepLaunch& L, hipStream_t s);
hipError_t launch_eval_quad(int n, const FitDesc* fits, const WorkItem* work, int nwork, hipStream_t s);
hipError_t launch_step_quad(int n, const StepLaunch& L, hipStream_t s);
// step + exchange of whole sequences (ladders of at most 64 rungs); work items
// are (fit, first sequence of the wave); rows: every fit of the launch has
// single-coordinate update rows (FitDesc::qrow)
hipError_t launch_iter_quad(int n, bool rows, const StepLaunch& L, hipStream_t s);
// Exchange of the sequences in xwork (items: fit, first sequence of the wave):
// per iteration for fits without xfused (L.close_k = 0); the closing launch of
// a segment (L.close_k = K) finishes the xfused fits' last exchange, leaves
// their state in the first set of arrays and moves the clock on by K.
hipError_t launch_exchange(const StepLaunch& L, const WorkItem* xwork, const XcdRanges& xr, hipStream_t s);

// generic.hip: the library's compiled-in model.cuh (reference contract)
hipError_t launch_eval_user(const FitDesc* fits, const WorkItem* work, int nwork, int model_only,
                            int respect_flag, hipStream_t s);
hipError_t launch_step_user(const StepLaunch& L, hipStream_t s);
const char* user_model_name();

}  // namespace mhre
