// The user-model contract of the reference: the object a project's model.cuh
// receives in
//     __device__ int model(CCalcModel *mc, int id, int ipt, int imd, int ipass);
// (template/model.cuh:21, src/gpu_mcmc.cu:299).  Field names, types and
// meanings are those of src/mcmcjob.cuh:30-87 that model code reads, so
// existing model.cuh files compile against this header unchanged:
//   id     datum index into dat[] / std2r[]
//   ipt    index of the chain's first parameter in pcur[] (ist * nptot)
//   imd    through index ist * ndat + id into datm[] (and chi2m[])
//   ipass  0 .. npass-1
// where ist = nseq * ibeta + iseq.  Return 0 = no result for this datum,
// 1 = value, 2 = angle (the residual is wrapped), -1 = error.
#pragma once

class CCalcModel {
 public:
  int imodel;
  int ncoor;
  int ndat;
  int nicoor;
  int nidat;
  int ndatm;
  int nptot;
  int nbeta;
  int nseq;
  int nchi2m;   // nbeta * nseq * ndat

  float *coor;   // [ncoor]
  float *dat;    // [ndat]
  float *std2r;  // [ndat]
  int *icoor;    // [nicoor]
  int *idat;     // [nidat]
  float *datm;   // [ndatm]
  float *chi2m;  // [nbeta][nseq][ndat]
  float *pcur;   // [nbeta][nseq][nptot]
  int *flag;     // [nbeta][nseq]
  float *chi2c;  // [nbeta][nseq]
};

// Constants model.cuh files of the reference use without defining them
// (src/mcmcjob.cuh:9-20 defines them for the model's translation unit).
#ifndef MHRE_NO_MODEL_MACROS
#ifndef pi
#define pi (3.1415926535897931f)
#endif
#ifndef pi4
#define pi4 (12.566370614359172f)
#endif
#ifndef min
#define min(a, b) ((a) < (b) ? (a) : (b))
#endif
#ifndef max
#define max(a, b) ((a) > (b) ? (a) : (b))
#endif
#endif  // MHRE_NO_MODEL_MACROS
