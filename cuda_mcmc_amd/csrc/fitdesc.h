// Device-resident description of one MH-RE fit.  A batch is an array of
// these; a plain (reference-style) run is a batch of one.
//
// All chain-state arrays use the reference's layouts (src/mcmc_interf.pyx:
// 232-263) so they can be copied to and from the caller's buffers verbatim:
//   pcur[nbeta][nseq][nptot], chi2c/flag/tcur/ptent/ptentn[nbeta][nseq],
//   pstp/n_cnt/n_acpt[nprm][nbeta][nseq],
//   n_exch[nbeta][nseq], pout[nprm][nbeta][nseq][niter],
//   tout[nbeta][nseq][niter], chi2[nseq][niter].
// Chain index ist = nseq * ibeta + iseq (src/gpu_mcmc.cu:247).
// The acceptance history n_hist[nadj][nprm][nbeta][nseq] (0/1 entries) is
// kept as bits, chain-major: hbits[nchain][hwords][nprm], entry a of
// (k, chain) is bit a % 32 of word [chain][a / 32][k], so a chain's whole
// history (hwords * nprm words, 128 bytes for 8 parameters and nadj 100)
// shares its cache lines; it is packed and unpacked on the host.
#pragma once
#include <stdint.h>

class CCalcModel;

namespace mhre {

enum ModelKind : int {
  MODEL_SGRA9 = 1,    // 9-parameter ring + Gaussian (native fused kernel)
  MODEL_SGRA13 = 2,   // 13-parameter ring + Gaussian (native fused kernel)
  MODEL_USER = 10,    // the library's compiled-in model.cuh (generic path)
  MODEL_QUADSYS = 20, // systems of quadratic equations (native, thread per chain)
};

constexpr int MAXP = 32;       // max nptot handled by the native kernels
constexpr int RING = 512;      // per-wave LDS ring of model visibilities
constexpr int MAX_CGROUP = 4;  // most tiles whose closures are formed together
constexpr int TILE = 64;       // visibilities per wave step

// Quadratic systems, single-coordinate update: moving x_t by d changes the
// left side of equation e by d (b_t + sum_j a_tj x_j + a_tt x'_t), x'_t the
// new value, where a_tj is the coefficient of x_t x_j (Q_tj + Q_jt for j != t,
// Q_tt for j = t).  Block t of the table holds, for term j = 0 .. n + 1, the
// QROW_MAXEQ values sqrt(w_e) a_tj (j < n), sqrt(w_e) b_t (j = n) and
// sqrt(w_e) a_tt (j = n + 1) of equations e = 0 .. QROW_MAXEQ - 1 (zero past
// the fit's equations), so one lane updates all equations side by side with
// two 16-byte LDS reads per term.  Blocks lie QROW_TS(n) floats apart, an odd
// multiple of 16 bytes, so the n blocks that the lanes of a wave read at once
// fall in disjoint LDS banks.  Then sqrt(w_e) per equation.
constexpr int QROW_MAXEQ = 8;
constexpr int QROW_TS(int n) { return (n + 2) * QROW_MAXEQ + 4; }
constexpr int QROW_FLOATS(int n) { return n * QROW_TS(n) + QROW_MAXEQ; }
constexpr int QROW_LDS = 8 * QROW_TS(8);   // per wave: the blocks of the largest system

struct FitDesc {
  // ---- sizes and options
  int model;          // ModelKind
  int imodel;         // value handed to the user's model() as mc->imodel
  int nbeta, nseq, nptot, nprm, nadj, npass;
  int ndat;           // number of chi^2 terms
  int nvis;           // visibilities evaluated by the native kernel (incl. cpExt pseudo-vis)
  int ncph;           // closure phases
  int ntile;          // ceil(nvis / TILE)
  int ring_ok;        // every closure lies within one ring window
  int cgroup;         // closures are formed after every cgroup-th tile
  int nburn, niter;
  int nchain;         // nbeta * nseq
  uint32_t key0, key1;
  float adapt_target; // acceptance-rate threshold of the step adaptation (reference: 0.30)
  int adapt_tie;      // rate_below (mhre_device.h): pred(adapt_target) has an even significand
  double adapt_mid;   // rate_below: midpoint between adapt_target and the next float toward 0
  int quad_n, quad_neq;  // MODEL_QUADSYS sizes
  int record;            // 1: write pout/tout/chi2 in the sampling stage
  int hwords;            // words per acceptance history: ceil(nadj / 32)
  int xlanes;            // exchange: lanes per sequence (a power of two >= nbeta, or 64),
                         // so one wave swaps 64 / xlanes sequences side by side
  int hwphase;           // ring+Gaussian step kernel: every shift phase (revolutions) is
                         // bounded by 64 over the prior box, the chains' values and the
                         // data's baselines, so the hardware sine needs no reduction
  long long ndatm;       // length of datm (head nchain*ndat, then the phase tail)
  long long tr_dec;      // decision trace: this fit's first entry in the session's
  long long tr_ex;       // dec/chi2 and exchange trace arrays (fits in session order)

  // ---- static inputs (device pointers)
  const float4* vis;       // [(ntile + 1) * TILE] (u, v, amp / sigma, 1 / sigma) for the
                           // native kernels (amp sqrt(std2r), sqrt(std2r)), zero past nvis
  const float* rho2;       // [(ntile + 1) * TILE] u^2 + v^2 of each visibility, zero past nvis
  const int4* clo;         // [ncph + TILE] vis indices (i, j, k, datum index) sorted by
                           // ring tile, zero past ncph
  const int4* clor;        // [ncph + TILE] the same legs as LDS ring byte offsets
                           // ((index % RING) * 8) and the datum index
  const float4* clow;      // [ncph + TILE] (cos obs, sin obs, 1/sigma^2, 0)
  const int2* tinfo;       // [ntile + 2] per tile t: first closure whose largest index
                           // lies past tile t - 1, and the max u^2+v^2 of tile t's
                           // visibilities (float bits)
  const int* clr;          // [ncph + TILE] gradient only: round of each closure leg
                           // (4 bits per leg), no two legs of one 64-closure batch
                           // and round naming the same visibility; null elsewhere
  int nround;              // rounds per batch (gradient only)
  const float* qcoef;      // MODEL_QUADSYS packed coefficients
  const float* qrow;       // MODEL_QUADSYS single-coordinate update rows (QROW_* layout),
                           // null where the step evaluates every equation in full
  const float* coor; const float* dat; const float* std2r;
  const int* icoor; const int* idat;
  const int* pdescr; const int* ivar; const int* invar;
  const float* pmint; const float* pmaxt; const float* beta;

  // ---- chain state
  float* pcur; float* chi2c; float* pstp;
  int* n_cnt; int* n_acpt; uint32_t* hbits; int* n_exch; int* sumh;
  int* tcur; int* flag; float* ptent; int* ptentn;
  // xfused = 1: the replica exchange of iteration it - 1 runs at the start of
  // the step launch of iteration it (every wave of a sequence computes the
  // sequence's swaps and takes its own slot's result), so no exchange launch
  // separates two step launches.  Launch i of a segment then writes the
  // swapped state (pcur, chi2c, tcur) to the first set of arrays for even i
  // and to the second (pcur_b, chi2c_b, tcur_b) for odd i, and reads launch
  // i - 1's; the segment's closing launch leaves it in the first set.
  int xfused;
  float* pcur_b; float* chi2c_b; int* tcur_b;

  // ---- outputs
  float* pout; int* tout; float* chi2;

  // ---- workspaces
  float* datm; float* chi2m;     // reference layout (generic path and eval outputs)
  CCalcModel* mc;                // device CCalcModel for the user's model()

  // ---- counters
  unsigned long long* nevals;    // [2]: performed chi^2 evaluations, proposals
  int* errflag;                  // first negative model() return code
};

// Work lists of the wave-per-chain kernels are split into NXCD contiguous
// ranges of nearly equal cost, one per XCD (each XCD has its own L2).  Work is
// dealt to the ranges in chunks of XCHUNK consecutive chains (sequences) of
// one fit, so that the 128-byte lines of the [..][chain] state arrays are
// touched by one XCD only.  Work-group b runs on XCD b % NXCD (round-robin
// dispatch) and takes its work from that XCD's range.
constexpr int NXCD = 8;
constexpr int XCHUNK = 32;
struct XcdRanges {
  int off[NXCD + 1];   // range x = [off[x], off[x + 1])
};

// Iteration clock of a sequence group.  The host launches the group's
// iterations in segments of K (one hipGraph replay, or the iterations of one
// run call): launch i of a segment (i = 0 .. K-1) runs iteration
// clk.it + i, the offset i being a kernel argument, and the segment ends with
// one exchange launch that finishes the segment's last exchange and moves the
// clock on by K.  Each kind of launch reads its own field and writes the
// other kind's, which no running work-group reads (launches on one stream
// never overlap):
//   step launch:     it = clk.it + i;  one thread stores clk.xit = clk.it
//   closing launch:  it = clk.xit + K - 1;  one thread stores clk.it = clk.xit + K
// so no launch has to count its finished work-groups.  The persistent step
// kernels' queue heads alternate between two sets by the parity of i: launch
// i clears the set that launch i + 1 uses, the closing launch clears both.
constexpr int NQSLOT = 6;
struct RunClock {
  int it;                        // first iteration of the current segment (burn-in first, then sampling)
  int xit;                       // copy of it for the closing exchange launch
  int qhead[2][NQSLOT * NXCD];   // work-queue heads of the persistent step kernels, per XCD range
};

struct WorkItem {
  int fit;
  int idx;    // chain (step) or sequence (exchange) within the fit
};

}  // namespace mhre
