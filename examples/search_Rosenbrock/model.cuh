// Rosenbrock valley for the search_Rosenbrock project (one datum).
//   datm = (1 - x)^2 + 100 (y - x^2)^2, compared with dat[0] = 0, so the
//   chi^2 is the squared valley height (minimum 0 at x = y = 1).
// Contract: see cuda_mcmc_amd/csrc/ccalcmodel.h.
__device__ int model(CCalcModel *mc, int id, int ipt, int imd, int ipass) {
  if (id != 0) return 0;
  const float x = mc->pcur[ipt];
  const float y = mc->pcur[ipt + 1];
  const float a = 1.0f - x;
  const float b = y - x * x;
  mc->datm[imd] = a * a + 100.0f * b * b;
  return 1;
}
