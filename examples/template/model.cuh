// Template for a new project: compute one model value for datum id of the
// chain whose parameters start at mc->pcur[ipt], store it in mc->datm[imd]
// and return 1 (or 2 if the value is an angle, 0 if this pass produces no
// value for id).  mc->coor, mc->idat, mc->icoor, mc->dat ... are available.
// Contract: see cuda_mcmc_amd/csrc/ccalcmodel.h.
__device__ int model(CCalcModel *mc, int id, int ipt, int imd, int ipass) {
  return 1;
}
