// Ring + Gaussian visibility model of the sgra_model_fit project written
// against the per-datum model() contract (generic path).  The library built
// for this project also carries the fused native kernels for imodel 1 and 2;
// this file is what the generic path runs (e.g. MHRE_FORCE_GENERIC=1).
//
//   idat = [nvis, nchi2m, cpExt];  dat = [amplitudes (nvis), closures (ncph)]
//   pass 0, id <  nvis: datm[imd] = |V(u_id, v_id)|, phase into the datm tail
//                       datm[nchi2m + ist*nvis + id]
//   pass 1, id >= nvis: closure phase wrap(ph_a + ph_b - ph_c) of the triplet
//                       icoor[3j..3j+2] (cpExt = 0), or of three model
//                       evaluations at (u1,v1), (u2,v2), -(u1+u2,v1+v2) taken
//                       from coor[2 nvis + 4j ..] (cpExt = 1); returns 2.
#include "eht_model.h"

__device__ inline float2 mhre_user_vis(CCalcModel *mc, int ipt, float u, float v) {
  const float *p = mc->pcur + ipt;
  if (mc->imodel == 2) {
    const mhre::Sgra13::Inv q = mhre::Sgra13::prepare(p);
    return mhre::Sgra13::vis(q, u, v, true);
  }
  const mhre::Sgra9::Inv q = mhre::Sgra9::prepare(p);
  return mhre::Sgra9::vis(q, u, v, true);
}

__device__ int model(CCalcModel *mc, int id, int ipt, int imd, int ipass) {
  if (mc->imodel != 1 && mc->imodel != 2) return -1;
  const int nvis = mc->idat[0];
  const int ist = imd / mc->ndat;
  // the phase tail of this chain must lie inside datm[ndatm]
  const bool tail = (long long)mc->nchi2m + (long long)(ist + 1) * nvis <= (long long)mc->ndatm;
  float *phase = mc->datm + mc->nchi2m + ist * nvis;
  if (ipass == 0 && id < nvis) {
    const float2 V = mhre_user_vis(mc, ipt, mc->coor[2 * id], mc->coor[2 * id + 1]);
    mc->datm[imd] = sqrtf(V.x * V.x + V.y * V.y);
    if (tail) phase[id] = atan2f(V.y, V.x);
    return 1;
  }
  if (ipass == 1 && id >= nvis) {
    const int j = id - nvis;
    float cph;
    if (mc->idat[2] == 0) {
      if (!tail) return -1;   // closures need the stored phases
      cph = phase[mc->icoor[3 * j]] + phase[mc->icoor[3 * j + 1]] - phase[mc->icoor[3 * j + 2]];
    } else {
      const float *c = mc->coor + 2 * nvis + 4 * j;
      const float2 a = mhre_user_vis(mc, ipt, c[0], c[1]);
      const float2 b = mhre_user_vis(mc, ipt, c[2], c[3]);
      const float2 e = mhre_user_vis(mc, ipt, -(c[0] + c[2]), -(c[1] + c[3]));
      cph = atan2f(a.y, a.x) + atan2f(b.y, b.x) - atan2f(e.y, e.x);
    }
    mc->datm[imd] = atan2f(sinf(cph), cosf(cph));
    return 2;
  }
  return 0;
}
