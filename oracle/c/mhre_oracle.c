/*
 * C / OpenMP float64 port of the oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Used as the timed CPU baseline of bench.py and cross-checked against the
 * NumPy oracle (oracle/mhre.py, oracle/models.py) by tests/test_oracle_c.py.
 * It restates, in float64 with the C library's j0/j1/jn:
 *   sgra_model9      examples/sgra_model_fit/model.cuh:322-462
 *   closures / chi^2 examples/sgra_model_fit/model.cuh:73-111, src/gpu_mcmc.cu:305-323
 *   MH-RE iteration  src/gpu_mcmc.cu:361-626 with the engine's Philox layout
 *                    (oracle/philox.py) and the reference's storage types
 *                    (float32 parameters / steps / chi^2, int32 counters).
 * Parallel over chains (MH proposals) and sequences (exchanges).
 */
#define _DEFAULT_SOURCE 1   /* POSIX j0/j1/jn declarations */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.141592653589793
#define UAS2NRAD (PI / 180.0 / 3600.0 * 1e3)
#define EXCH 0x40000000u
#define INIT 0x80000000u

typedef struct { uint32_t x, y, z, w; } u4;

static u4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1; c3 = (uint32_t)p0; c0 = n0; c2 = n2;
  }
  u4 o = {c0, c1, c2, c3};
  return o;
}

static double u01(uint32_t w) { return (double)((w >> 8) + 1u) * 5.9604644775390625e-8; }
static int pick(uint32_t w, int n) { return (int)(((uint64_t)(uint32_t)n * ((uint64_t)w + 1ull)) >> 32); }

/* 9-parameter visibility, literal restatement of model.cuh:322-462 in double */
static void sgra9(const double* p, double ul, double vl, double* vre, double* vim) {
  double Zsp = p[0], Rex = p[1] * UAS2NRAD, rq = p[2], ecc = p[3], fade = p[4];
  double gax = p[5], aq = p[6], gq = p[7], th = p[8];
  double Rin = rq * Rex, d = ecc * (Rex - Rin), gsy = gax * Rex, gsx = aq * gsy;
  double gcoef = 2.0 * pow(PI / (2.0 * sqrt(2.0 * log(2.0))), 2);
  double h = (2.0 / PI) / ((Rex * Rex - Rin * Rin * (1.0 + d / Rex)) * fade + (Rex * Rex - Rin * Rin * (1.0 - d / Rex)));
  double h0 = fade * h;
  double U = ul * cos(th) + vl * sin(th), V = -ul * sin(th) + vl * cos(th);
  double rho = sqrt(U * U + V * V), ree, ime, rei = 0, imi = 0;
  double pi4 = 4.0 * PI;
  if (rho > 1e-7) {
    double arg = 2.0 * PI * rho, aRex = arg * Rex;
    ree = .5 * (h + h0) * Rex * j1(aRex) / rho;
    ime = (h - h0) / pi4 * (PI * Rex * (j0(aRex) - jn(2, aRex)) / (rho * rho) - j1(aRex) / (rho * rho * rho)) * U;
    if (Rin != 0.0) {
      double aRin = arg * Rin;
      rei = .5 * (h + h0 - d * (h - h0) / Rex) * Rin * j1(aRin) / rho;
      imi = (h - h0) / pi4 * rq * (PI * Rin * (j0(aRin) - jn(2, aRin)) / (rho * rho) - j1(aRin) / (rho * rho * rho)) * U;
    }
  } else {
    ree = .5 * (h + h0) * PI * Rex * Rex;
    ime = -.25 * PI * PI * (h - h0) * Rex * Rex * Rex * U;
    if (Rin != 0.0) {
      rei = .5 * (h + h0 - d * (h - h0) / Rex) * PI * Rin * Rin;
      imi = -.25 * PI * PI * (h - h0) * (Rin * Rin * Rin * Rin / Rex) * U;
    }
  }
  double Frr = ree, Fri = ime;
  if (Rin != 0.0) {
    double dr = cos(2.0 * PI * d * U), di = sin(2.0 * PI * d * U);
    Frr = ree - (rei * dr - imi * di);
    Fri = ime - (rei * di + imi * dr);
  }
  if (gax == 0.0 || aq == 0.0 || gq == 0.0) {
    *vre = Zsp * Frr; *vim = Zsp * Fri; return;
  }
  double gshift = Rin - d;
  double Gamp = exp(-gcoef * ((U * gsx) * (U * gsx) + (V * gsy) * (V * gsy)));
  double Gph = -2.0 * PI * gshift * U;
  *vre = Zsp * ((1.0 - gq) * Frr + gq * Gamp * cos(Gph));
  *vim = Zsp * ((1.0 - gq) * Fri + gq * Gamp * sin(Gph));
}

static double wrap(double x) { return atan2(sin(x), cos(x)); }

/* chi^2 of one parameter vector: amplitudes + closures (cpExt = 0). pha: [nvis] scratch */
double oracle_eht_chi2(const double* p, int nvis, int ncph, const float* coor, const float* dat,
                       const float* std2r, const int* icoor, double* pha) {
  double s = 0.0;
  for (int i = 0; i < nvis; ++i) {
    double re, im;
    sgra9(p, coor[2 * i], coor[2 * i + 1], &re, &im);
    double amp = sqrt(re * re + im * im);
    pha[i] = atan2(im, re);
    double dd = (double)dat[i] - amp;
    s += dd * dd * (double)std2r[i];
  }
  for (int j = 0; j < ncph; ++j) {
    double cph = wrap(pha[icoor[3 * j]] + pha[icoor[3 * j + 1]] - pha[icoor[3 * j + 2]]);
    double dd = wrap((double)dat[nvis + j] - cph);
    s += dd * dd * (double)std2r[nvis + j];
  }
  return s;
}

void oracle_eht_chi2_batch(int n, const float* P, int nptot, int nvis, int ncph, const float* coor,
                           const float* dat, const float* std2r, const int* icoor, double* out, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    double* pha = (double*)malloc(sizeof(double) * (nvis > 0 ? nvis : 1));
    double p[16];
#pragma omp for schedule(dynamic, 4)
    for (int c = 0; c < n; ++c) {
      for (int k = 0; k < 9; ++k) p[k] = P[(size_t)c * nptot + k];
      out[c] = oracle_eht_chi2(p, nvis, ncph, coor, dat, std2r, icoor, pha);
    }
    free(pha);
  }
}

/* One MH-RE iteration of the EHT fit (oracle/mhre.py iteration()).  State
 * arrays in the reference layouts; returns the number of chi^2 evaluations. */
long long oracle_eht_iteration(int nbeta, int nseq, int nptot, int nprm, int nadj, int it, int burn, int itr,
                               uint64_t seed, float adapt_target,
                               const int* pdescr, const int* ivar, const int* invar, const float* pmint,
                               const float* pmaxt, const float* beta,
                               int nvis, int ncph, const float* coor, const float* dat, const float* std2r,
                               const int* icoor,
                               float* pcur, float* chi2c, float* pstp, int* n_cnt, int* n_acpt, int* n_hist,
                               int* n_exch, int* tcur, int* flag, float* ptent, int* ptentn,
                               float* pout, int* tout, float* chi2, int niter, int nthreads) {
  const int C = nbeta * nseq;
  const uint32_t k0 = (uint32_t)(seed & 0xffffffffu), k1 = (uint32_t)(seed >> 32);
  long long nev = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : nev)
  {
    double* pha = (double*)malloc(sizeof(double) * (nvis > 0 ? nvis : 1));
    double p[32];
#pragma omp for schedule(dynamic, 2)
    for (int ch = 0; ch < C; ++ch) {
      const int ib = ch / nseq;
      for (int s = 0; s < nprm; ++s) {
        u4 r = philox((uint32_t)ch, (uint32_t)it, (uint32_t)s, 0u, k0, k1);
        int k = pick(r.x, nprm);
        for (int rr = 1; k >= nprm && rr <= 8; ++rr) k = pick(philox((uint32_t)ch, (uint32_t)it, (uint32_t)s, (uint32_t)rr, k0, k1).x, nprm);
        if (k >= nprm) k = nprm - 1;
        const int t = ivar[k];
        const size_t pk = (size_t)k * C + ch;
        float* pt = pcur + (size_t)ch * nptot + t;
        const float old = *pt;
        ptent[ch] = old;
        ptentn[ch] = t;
        double pd = (double)old + (double)pstp[pk] * (sqrt(-2.0 * log(u01(r.y))) * cos(2.0 * PI * u01(r.z)));
        if (pdescr[t] == 2) pd = pd - 2.0 * PI * rint(pd / (2.0 * PI));
        const float pn = (float)pd;
        int acc = 0;
        if (!(pn < pmint[t] || pn > pmaxt[t])) {
          *pt = pn;
          for (int q = 0; q < nptot; ++q) p[q] = pcur[(size_t)ch * nptot + q];
          double c2 = oracle_eht_chi2(p, nvis, ncph, coor, dat, std2r, icoor, pha);
          ++nev;
          double x = -0.5 * (double)beta[ib] * (c2 - (double)chi2c[ch]);
          if (log(u01(r.w)) <= x) {
            acc = 1;
            chi2c[ch] = (float)c2;
            n_acpt[pk] += 1;
          } else {
            *pt = old;
          }
        }
        flag[ch] = acc;
        if (burn) {
          const int cnt = n_cnt[pk];
          const size_t hidx = ((size_t)nprm * (size_t)(cnt % nadj) + (size_t)k) * C + ch;
          n_hist[hidx] = acc;
          int sh = 0;
          for (int a = 0; a < nadj; ++a) sh += n_hist[((size_t)nprm * a + k) * C + ch];
          const int den = cnt < nadj ? cnt : nadj;
          const float rate = (float)sh / (float)den;
          pstp[pk] = rate < adapt_target ? (float)((double)pstp[pk] / 1.01) : (float)((double)pstp[pk] * 1.01);
          n_cnt[pk] = cnt + 1;
        }
      }
    }
    free(pha);
  }
  /* exchange (gpu_mcmc.cu:533-626), one sequence per thread */
#pragma omp parallel for schedule(static)
  for (int q = 0; q < nseq; ++q) {
    if (nbeta > 1) {
      for (int a = 0; a < nbeta; ++a) {
        u4 r = philox((uint32_t)q, (uint32_t)it, (uint32_t)a, EXCH, k0, k1);
        int k = pick(r.x, nbeta - 1);
        for (int rr = 1; k >= nbeta - 1 && rr <= 8; ++rr) k = pick(philox((uint32_t)q, (uint32_t)it, (uint32_t)a, EXCH + (uint32_t)rr, k0, k1).x, nbeta - 1);
        if (k >= nbeta - 1) k = nbeta - 2;
        const size_t a0 = (size_t)k * nseq + q, a1 = (size_t)(k + 1) * nseq + q;
        double x = -0.5 * ((double)beta[k + 1] - (double)beta[k]) * ((double)chi2c[a0] - (double)chi2c[a1]);
        if (log(u01(r.y)) <= x) {
          int ti = tcur[a0]; tcur[a0] = tcur[a1]; tcur[a1] = ti;
          float tc = chi2c[a0]; chi2c[a0] = chi2c[a1]; chi2c[a1] = tc;
          for (int j = 0; j < nptot; ++j) {
            float tmp = pcur[a0 * nptot + j]; pcur[a0 * nptot + j] = pcur[a1 * nptot + j]; pcur[a1 * nptot + j] = tmp;
          }
          if (!burn) n_exch[a0] += 1;
        }
      }
    }
    if (!burn && pout) {
      for (int ib = 0; ib < nbeta; ++ib) {
        const size_t me = (size_t)ib * nseq + q;
        for (int j = 0; j < nptot; ++j)
          if (invar[j] >= 0) pout[(((size_t)nbeta * invar[j] + ib) * nseq + q) * niter + itr] = pcur[me * nptot + j];
        tout[me * niter + itr] = tcur[me];
      }
      chi2[(size_t)q * niter + itr] = chi2c[q];
    }
  }
  return nev;
}
