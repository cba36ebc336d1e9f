// fct.h -- the along-stream weights of the FCT bounds, shared by the
// per-point and the tiled tracer kernels.
#pragma once
#include "common.h"

namespace shp {

// get_fact_fct (COMMON.F90:2676-2734); fact(a,b,c) -> f[a + 2*((b+1) + 3*(c+1))]
__device__ __forceinline__ void get_fact(double* f, double rw, double ru, double rv) {
  const double ugev = fsign(0.5, __builtin_fabs(ru) - __builtin_fabs(rv)) + 0.5;
  const double ugew = fsign(0.5, __builtin_fabs(ru) - __builtin_fabs(rw)) + 0.5;
  const double vgew = fsign(0.5, __builtin_fabs(rv) - __builtin_fabs(rw)) + 0.5;
  const double uzero = fsign(0.5, __builtin_fabs(ru) - EPS_RP) - 0.5;
  const double vzero = fsign(0.5, __builtin_fabs(rv) - EPS_RP) - 0.5;
  const double wzero = fsign(0.5, __builtin_fabs(rw) - EPS_RP) - 0.5;
  const double sign_uv = fsign(0.5, ru * rv) + 0.5;
  const double sign_uw = fsign(0.5, ru * rw) + 0.5;
  const double sign_vw = fsign(0.5, rv * rw) + 0.5;
  const double wu = __builtin_fabs(rw / (ru + uzero) * (1.0 + uzero));
  const double vu = __builtin_fabs(rv / (ru + uzero) * (1.0 + uzero));
  const double uv = __builtin_fabs(ru / (rv + vzero) * (1.0 + vzero));
  const double wv = __builtin_fabs(rw / (rv + vzero) * (1.0 + vzero));
  const double uw = __builtin_fabs(ru / (rw + wzero) * (1.0 + wzero));
  const double vw = __builtin_fabs(rv / (rw + wzero) * (1.0 + wzero));
  const double umax = ugev * ugew * (1.0 + uzero);
  const double vmax = (1.0 - ugev) * vgew;
  const double wmax = 1.0 - ugev * ugew - vmax;
#define FCT_(a_, b_, c_) f[(a_) + 2 * (((b_) + 1) + 3 * ((c_) + 1))]
  FCT_(0, 0, 0) = -ugev * ugew * uzero;
  FCT_(1, 0, 0) = wmax * (1.0 - uw) * (1.0 - vw);
  FCT_(0, 1, 0) = umax * (1.0 - vu) * (1.0 - wu);
  FCT_(0, 0, 1) = vmax * (1.0 - uv) * (1.0 - wv);
  const double c3 = (umax * vu * wu + vmax * uv * wv + wmax * uw * vw);
  FCT_(1, 1, 1) = sign_uv * sign_uw * c3;
  FCT_(1, -1, 1) = (1.0 - sign_uv) * (1.0 - sign_uw) * c3;
  FCT_(1, 1, -1) = (1.0 - sign_uv) * sign_uw * c3;
  FCT_(1, -1, -1) = sign_uv * (1.0 - sign_uw) * c3;
  const double cuw = (ugew * wu * (1.0 - vu) + (1.0 - ugew) * uw * (1.0 - vw));
  FCT_(1, 1, 0) = sign_uw * (1.0 - vmax) * cuw;
  FCT_(1, -1, 0) = (1.0 - sign_uw) * (1.0 - vmax) * cuw;
  const double cvw = (vgew * wv * (1.0 - uv) + (1.0 - vgew) * vw * (1.0 - uw));
  FCT_(1, 0, 1) = sign_vw * (1.0 - umax) * cvw;
  FCT_(1, 0, -1) = (1.0 - sign_vw) * (1.0 - umax) * cvw;
  const double cuv = (ugev * vu * (1.0 - wu) + (1.0 - ugev) * uv * (1.0 - wv));
  FCT_(0, 1, 1) = sign_uv * (1.0 - wmax) * cuv;
  FCT_(0, 1, -1) = (1.0 - sign_uv) * (1.0 - wmax) * cuv;
#undef FCT_
}

}  // namespace shp
