// Two nonlinear equations in (x, y) for the solvsys project:
//   id 0:  x^P1 + y^P2      with P1 = coor[0], P2 = coor[1]
//   id 1:  -K x + y         with K  = coor[2]
// compared with the right-hand sides in dat[].
// Contract: see cuda_mcmc_amd/csrc/ccalcmodel.h.
__device__ int model(CCalcModel *mc, int id, int ipt, int imd, int ipass) {
  const float x = mc->pcur[ipt];
  const float y = mc->pcur[ipt + 1];
  if (id == 0)
    mc->datm[imd] = powf(x, mc->coor[0]) + powf(y, mc->coor[1]);
  else
    mc->datm[imd] = -mc->coor[2] * x + y;
  return 1;
}
