/*
 * mhre.h -- C ABI of the MI355X MH-RE (parallel-tempering MCMC) engine.
 *
 * Plain pointers and sizes only.  All functions return 0 on success and a
 * negative MHRE_E* code on failure; mhre_last_error() gives the message.
 * (The reference's functions return void and call exit(1) on CUDA errors,
 * src/mcmcjob.cu:212-219.)
 *
 * Part 1 -- drop-in replacements for the reference's C ABI (src/mcmc.h) with
 * the same argument order and meanings.  Host buffers are C-contiguous arrays
 * in the reference layouts (src/mcmc_interf.pyx:232-263); results are written
 * back in place.
 *
 * Part 2 -- extensions: device-resident sessions over batches of fits (each
 * fit described by the same arrays), statistics and decision traces.
 */
#ifndef MHRE_H
#define MHRE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHRE_OK 0
#define MHRE_EINVAL (-1)     /* inconsistent sizes / arguments */
#define MHRE_EHIP (-2)       /* HIP runtime error */
#define MHRE_EMODEL (-3)     /* model() returned an error code (e.g. bad imodel) */
#define MHRE_ENOMEM (-4)     /* device allocation failed */

/* ------------------------------------------------------------ Part 1 */

/* MH-RE fit: burn-in with step adaptation, then sampling.
 * Replaces mcmc_cuda of src/mcmc.h:7-20 (src/mcmcjob.cu:475-545, run_mcmc of
 * src/gpu_mcmc.cu:682-763).  Arrays: coor[ncoor] dat[ndat] std2r[ndat]
 * icoor[nicoor] idat[nidat] datm[ndatm] chi2m[nbeta][nseq][ndat]
 * pdescr/invar/ptotal/pmint/pmaxt[nptot] ivar[nprm]
 * pout[nprm][nbeta][nseq][niter] tout[nbeta][nseq][niter] chi2[nseq][niter]
 * chi2c/tcur/flag/ptent/ptentn[nbeta][nseq] n_acpt/n_cnt/pstp[nprm][nbeta][nseq]
 * n_exch[nbeta][nseq] n_hist[nadj][nprm][nbeta][nseq] beta[nbeta]
 * pcur[nbeta][nseq][nptot] rndst[nbeta][nseq][48].
 * rndst receives each chain's resume record (seed, next iteration, chain id)
 * instead of cuRAND states.  datm/chi2m receive the model data and chi^2
 * terms of the final state of every chain. */
int mcmc_cuda(float *coor, float *dat, float *std2r,
              int *icoor, int *idat, float *datm, float *chi2m,
              int *pdescr, int *ivar, int *invar, float *ptotal,
              float *pmint, float *pmaxt,
              float *pout, int *tout, float *chi2,
              float *chi2c, int *tcur, int *flag,
              int *n_acpt, int *n_exch, int *n_cnt, int *n_hist,
              float *beta, float *pstp, float *ptent, int *ptentn,
              float *pcur, unsigned char *rndst,
              unsigned long long seed, int imodel,
              int ncoor, int ndat, int nicoor, int nidat, int ndatm,
              int nptot, int nprm, int nadj, int npass,
              int nbeta, int nseq, int nburn, int niter);

/* Batched model data and chi^2 for every [nbeta][nseq] parameter vector.
 * Replaces calc_modchi2 of src/mcmc.h:23-27 (src/mcmcjob.cu:227-254,
 * run_calcmodchi2 of src/gpu_mcmc.cu:769-783).  Chains with flag == 0 are
 * not evaluated; chi2c is the sum of each chain's chi2m row. */
int calc_modchi2(float *coor, float *dat, float *std2r,
                 int *icoor, int *idat, float *datm, float *chi2m,
                 float *pcur, float *chi2c, int *flag, int imodel,
                 int ncoor, int ndat, int nicoor, int nidat, int ndatm,
                 int nptot, int nbeta, int nseq);

/* Model data only (pass 0) for every parameter vector.
 * Replaces calc_model of src/mcmc.h:30-32 (src/mcmcjob.cu:257-298,
 * run_calcmodel of src/gpu_mcmc.cu:790-802).  coor holds interleaved
 * (u, v) pairs, the layout model() reads. */
int calc_model(float *coor, int *idat, float *datm, float *pcur,
               int imodel, int ncoor, int ndat, int nidat,
               int ndatm, int nptot, int nbeta, int nseq);

/* Release every device resource of this process on the current device.
 * Replaces reset_gpu / cudaDeviceReset of src/mcmc_interf.pyx:342-343. */
int reset_gpu(void);

/* Text listing of the visible GPUs (name, gfx target, CUs, memory), one line
 * each.  Replaces the list_gpus program of src/list_gpus.cu:13-22. */
int list_gpus(char *buf, int buflen);

/* ------------------------------------------------------------ Part 2 */

#define MHRE_MODEL_AUTO 0      /* from imodel like mcmc_cuda */
#define MHRE_MODEL_SGRA9 1     /* native 9-parameter ring + Gaussian */
#define MHRE_MODEL_SGRA13 2    /* native 13-parameter ring + Gaussian */
#define MHRE_MODEL_USER 10     /* the library's compiled-in model.cuh */
#define MHRE_MODEL_QUADSYS 20  /* sum_e (dat_e - x'Q_e x - b_e.x)^2 std2r_e;
                                  coor = [Q (ndat*nptot*nptot), b (ndat*nptot)] */

/* One fit: the arrays and sizes of mcmc_cuda plus options.  No reference
 * counterpart: the batched engine's unit of work. */
typedef struct MhreFitSpec {
  int model;                    /* MHRE_MODEL_* */
  int imodel;
  float *coor, *dat, *std2r;
  int *icoor, *idat;
  float *datm, *chi2m;          /* may be NULL: final-state model data not wanted */
  int *pdescr, *ivar, *invar;
  float *ptotal, *pmint, *pmaxt;
  float *pout; int *tout; float *chi2;  /* all NULL: sampling-stage states are
                                           not recorded (final state only) */
  float *chi2c; int *tcur; int *flag;
  int *n_acpt, *n_exch, *n_cnt, *n_hist;
  float *beta, *pstp, *ptent; int *ptentn;
  float *pcur; unsigned char *rndst;   /* rndst may be NULL */
  unsigned long long seed;
  int ncoor, ndat, nicoor, nidat, ndatm;
  int nptot, nprm, nadj, npass, nbeta, nseq, nburn, niter;
  float adapt_target;           /* step-adaptation threshold; 0 -> 0.30 */
  int init_chains;              /* 1: random start in the prior + initial chi^2
                                   (reference); 0: start from pcur and chi2c */
  int it0;                      /* first iteration index (resume) */
} MhreFitSpec;

typedef struct MhreOptions {
  int waves_per_block;          /* native step kernel: 1, 2 or 4; 0 = default */
  int persistent;               /* 1: persistent step kernel with a device work
                                   queue; 0: one wave per chain; -1 = default */
  int graph_iters;              /* >0: capture this many iterations in a
                                   hipGraph and replay it */
  int trace_iters;              /* record the decisions of iterations
                                   it0 .. it0+trace_iters-1; each array below is
                                   the concatenation, in fit order, of every
                                   fit's block of the shape given */
  int *trace_dec;               /* [trace_iters][nprm][nbeta*nseq]:
                                   k | inprior<<8 | accept<<9 */
  float *trace_chi2;            /* same shape: proposal chi^2 (0 if out of prior) */
  int *trace_ex;                /* [trace_iters][nbeta][nseq]: k | accept<<16 */
  int force_generic;            /* run imodel 1/2 through the compiled-in model.cuh */
  int waves_per_simd;           /* native step kernel: at most this many resident
                                   waves per SIMD (the launch reserves LDS so that
                                   no more work-groups fit on a compute unit);
                                   0 = as many as registers and LDS allow */
  int streams;                  /* split every fit's sequences into this many
                                   groups (1..8, at most the largest fit's
                                   exchange-wave count), each iterated on its
                                   own stream so that the groups' launches
                                   overlap; 0 = default (2 in the grid form, 1
                                   in the persistent form).  Results do not
                                   depend on it. */
  int exchange_launch;          /* 1: the ring+Gaussian fits' replica exchange
                                   runs in a launch of its own after every step
                                   launch; 0 = default: inside the next step
                                   launch (ladders of at most 64 rungs), which
                                   saves a launch per iteration.  Results do
                                   not depend on it. */
} MhreOptions;

typedef struct MhreStats {
  double run_ms;                /* device time of the last mhre_session_run */
  unsigned long long chi2_evals;/* chi^2 evaluations performed (in-prior proposals) */
  unsigned long long proposals; /* proposals made */
  double vis_evals;             /* visibility model evaluations performed */
  double cph_evals;             /* closure-phase terms evaluated */
  int iters_done;               /* iterations run so far */
  int nfits;
} MhreStats;

/* Upload a batch of fits and initialise their chains (per init_chains). */
int mhre_session_create(const MhreFitSpec *fits, int nfits, const MhreOptions *opt, void **session);
/* Run niter MH-RE iterations of every fit (asynchronous to the host). */
int mhre_session_run(void *session, int niter);
/* Run niter iterations timing every launch with HIP events on the stream of
 * its sequence group; returns the average device time of one group's step
 * launch(es) per iteration and of its exchange launches per iteration (those
 * that run every iteration, plus each segment's closing launch spread over
 * the segment's iterations), in milliseconds (the launches of different
 * groups overlap, so these do not add up to the iteration time). */
int mhre_session_time_kernels(void *session, int niter, double *step_ms, double *exch_ms);
/* Choose the step kernels' launch form (waves per work-group 1/2/4, grid or
 * persistent work queue) by timing each candidate on `iters` real
 * iterations; the 6 * iters iterations advance the chains like any others
 * (every form gives bit-identical chains).  Returns the choice and its step
 * time per iteration.  No reference counterpart. */
int mhre_session_autotune(void *session, int iters, int *waves_per_block, int *persistent,
                          double *step_ms);
/* Wait for the session's stream. */
int mhre_session_sync(void *session);
/* Copy chain state and outputs back into the fit specs' host arrays; with
 * model_data != 0 also fill datm/chi2m for the final state. */
int mhre_session_download(void *session, int model_data);
int mhre_session_stats(void *session, MhreStats *stats);
/* How fit `fit` of the session is run (no reference counterpart):
 * info[0] = step kernel (0 native 9-parameter ring+Gaussian, 1 native
 * 13-parameter, 2 the compiled-in model.cuh, 3 quadratic systems),
 * info[1] = 1 if its replica exchange runs inside a step launch (no exchange
 * launch per iteration), info[2] = 1 if its ring+Gaussian shift phases skip
 * the reduction before the hardware sine (bounded by 64 revolutions),
 * info[3] = the session's sequence groups, and with ninfo >= 5 info[4] = 1 if
 * its quadratic-system steps update each equation's residual from a row
 * table when a proposal moves one coordinate, instead of evaluating every
 * equation in full.  ninfo >= 4. */
int mhre_session_fit_info(const void *session, int fit, int *info, int ninfo);
int mhre_session_destroy(void *session);
/* create + run(nburn + niter) + download + destroy. */
int mhre_run_batch(const MhreFitSpec *fits, int nfits, const MhreOptions *opt, MhreStats *stats);

const char *mhre_last_error(void);
/* The engine's Philox4x32-10 bijection (host implementation of the device
 * generator): out = Philox(counter ctr, key key).  No reference counterpart
 * (the reference draws from cuRAND XORWOW states, src/gpu_mcmc.cu:228-239). */
void mhre_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]);
int mhre_version(void);
/* Name of the model.cuh compiled into this library build. */
const char *mhre_user_model(void);
/* 1 if imodel 1/2 use the native fused kernels in this build. */
int mhre_native_eht(void);

/* ------------------------------------------------------------ Part 3 */

/* chi^2 gradients of the 9-parameter ring+Gaussian model (imodel 1) for
 * batches of parameter vectors.  No reference counterpart (the reference
 * samples without derivatives); the chi^2 is the one of calc_modchi2
 * (src/mcmc.h:23-27) for the same data arrays.  params, chi2, grad, jtr and
 * jtj are DEVICE pointers; stream is a hipStream_t (NULL: the null stream);
 * the calls are asynchronous to the host. */

/* Pack one data set (the calc_modchi2 data arrays) on the current device. */
int mhre_eht_data_create(const float *coor, const float *dat, const float *std2r,
                         const int *icoor, const int *idat, int imodel,
                         int ncoor, int ndat, int nicoor, int nidat, int nptot,
                         void **handle);
/* chi2[n] and, if grad != NULL, grad[n][ldp] = d chi^2 / d params for the
 * parameter vectors params[n][ldp] (9 <= ldp <= 32; entries past 9 get 0).
 * Data whose closures lie within the 512-visibility ring window use the
 * single-pass ring kernel; the environment variable MHRE_GRAD_RING=0 selects
 * the per-closure kernel instead (same results to float rounding). */
int mhre_eht_chi2_grad(void *handle, const float *params, int n, int ldp,
                       double *chi2, float *grad, void *stream);
/* Gauss-Newton normal equations of the weighted residuals r (amplitudes:
 * (dat - |V|) sqrt(std2r); closures: the wrapped phase residual times
 * sqrt(std2r)), so chi^2 = |r|^2: chi2[n], jtr[n][9] = J^T r and the upper
 * triangle of J^T J row by row, jtj[n][45]. */
int mhre_eht_gn_system(void *handle, const float *params, int n, int ldp,
                       double *chi2, double *jtr, double *jtj, void *stream);
int mhre_eht_data_destroy(void *handle);
/* How the data were packed (no reference counterpart): info[0] = 1 if the
 * gradient uses the single-pass ring kernel, info[1] = its closure-adjoint
 * rounds per 64-closure batch, info[2] = tiles per closure group (cgroup),
 * info[3] = 1 if every closure lies in the ring window.  ninfo >= 4. */
int mhre_eht_data_info(const void *handle, int *info, int ninfo);

/* Residuals r[rows][neq] = x^T Q_e x + b_e . x - dat_e and Jacobian
 * J[rows][neq][n] of the quadratic systems Q[nsys][neq][n][n],
 * b[nsys][neq][n], dat[nsys][neq] at the rows X[rows][n]; row i belongs to
 * system sys[i] (sys == NULL: system i).  Host pointers, synchronous.  Drives
 * the Gauss-Newton polish of solved systems; no reference counterpart (the
 * reference's solvsys example is sampled only, examples/solvsys/model.cuh). */
int mhre_quad_jacobian(const float *Q, const float *b, const float *dat, const int *sys,
                       int nsys, int neq, int n, const float *X, int rows, float *r, float *J);

/* mhre_quad_jacobian without any device allocation, for callers that cannot
 * allocate (code under stream capture, tight polishing loops): the inputs are
 * staged in a static 32 MiB device area of the library (one per GPU, kept
 * until the process ends) instead of a per-call allocation.  Same arguments,
 * results and errors, plus MHRE_EINVAL when the call's staging (all of Q, b,
 * dat, sys, X, r and J, each rounded up to 256 bytes) exceeds 32 MiB.  Calls
 * that overlap on one GPU share the area and race.  No counterpart in the
 * reference. */
int mhre_quad_jacobian_noalloc(const float *Q, const float *b, const float *dat, const int *sys,
                               int nsys, int neq, int n, const float *X, int rows, float *r, float *J);

#ifdef __cplusplus
}
#endif

#endif /* MHRE_H */
