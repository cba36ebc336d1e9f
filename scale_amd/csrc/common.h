// common.h -- device layout, constants and the UD3/CD4/UD1 face-value
// stencils shared by every HIP kernel of the HEVI short step and the FCT
// tracer step (gfx950).
//
// Device layout: a 3-D field (KA,IA,JA) of the reference is stored i-fastest,
// element (k,i,j) at i + IA*(j + JA*k) (0-based), so a wavefront of 64 lanes
// over consecutive i reads 512 contiguous bytes per load.  Multi-component
// fields append the component as the slowest index (GSQRT component c at
// offset c*KA*IA*JA; num_diff(v,d) at offset (v + 5*d)*KA*IA*JA), MAPF(i,j,n,c)
// at i + IA*(j + JA*(n + 2*c)).
//
// Arithmetic mirrors the reference expression by expression; the library is
// compiled with -ffp-contract=off so no multiply-add is fused.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace shp {

// staggering ids (scale_atmos_grid_cartesC_index.F90:90-101), 0-based
enum { I_XYZ = 0, I_XYW, I_UYW, I_XVW, I_UYZ, I_XVZ, I_UVZ };
enum { I_XY = 0, I_UY, I_XV, I_UV };
enum { ZDIR = 0, XDIR = 1, YDIR = 2 };
enum { I_DENS = 0, I_MOMZ, I_MOMX, I_MOMY, I_RHOT };

// scale_const.F90:46,55-56,88,162
constexpr double GRAV = 9.80665;
constexpr double RDRY = 287.04;
constexpr double CPDRY = 1004.64;
constexpr double P00 = 100000.0;
constexpr double EPS_RP = 2.220446049250313080847263336181640625e-16;

// flux coefficients (scale_atmos_dyn_fvm_flux_ud3.F90:80-89, _cd4.F90:84-88, _ud1.F90:84-86)
constexpr double F1 = 0.5;
constexpr double F2 = 0.5;
constexpr double F31 = -1.0 / 12.0;
constexpr double F32 = 7.0 / 12.0;
constexpr double F33 = 3.0 / 12.0;
constexpr double F41 = 7.0 / 12.0;
constexpr double F42 = -1.0 / 12.0;

// scale_atmos_dyn_fvm_flux_ud5.F90:92-96, scale_atmos_dyn_fvm_flux_cd6.F90:91-93
constexpr double F51 = 1.0 / 60.0;
constexpr double F52 = -8.0 / 60.0;
constexpr double F53 = 37.0 / 60.0;
constexpr double F54 = -5.0 / 60.0;
constexpr double F55 = 10.0 / 60.0;
constexpr double F61 = 37.0 / 60.0;
constexpr double F62 = -8.0 / 60.0;
constexpr double F63 = 1.0 / 60.0;
constexpr int UD1 = 1, CD2 = 2, UD3 = 3, CD4 = 4, UD5 = 5, CD6 = 6;
constexpr int UD3KOREN = 31;  // UD3 with the Koren (1993) limiter (tracer scheme UD3KOREN1993)

struct Dims {
  int KA, IA, JA;
  int KS, KE, IS, IE, JS, JE;  // 0-based, inclusive
  size_t n3;                   // KA*IA*JA
  size_t pl;                   // IA*JA
};

__host__ __device__ inline size_t IX(const Dims& d, int k, int i, int j) {
  return (size_t)i + (size_t)d.IA * (size_t)j + d.pl * (size_t)k;
}

// Fortran sign(a,b)
__device__ __forceinline__ double fsign(double a, double b) { return __builtin_copysign(__builtin_fabs(a), b); }
__device__ __forceinline__ float fsign(float a, float b) { return __builtin_copysignf(__builtin_fabsf(a), b); }
__device__ __forceinline__ double fabs_t(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float fabs_t(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double fmax2(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fmin2(double a, double b) { return a < b ? a : b; }

// The face values below are written for a real type T: double reproduces the
// reference's FP64 build; float is the FP32 arithmetic of the mixed-precision
// path (the reference's SINGLE build, scale_precision.F90:37-43), with the
// coefficients rounded to float once.
// Koren (1993) limiter of the UD3KOREN1993 scheme
// (scale_atmos_dyn_fvm_flux_ud3Koren1993.F90:1690-1710)
template <typename T>
__device__ __forceinline__ T koren_phi(T v1, T v2, T v3) {
  const T E = T(EPS_RP);
  const T z1 = E - fsign(E, fabs_t(v1 - v2) - E);
  const T z2 = E - fsign(E, fabs_t(v2 - v3) - E);
  const T r2 = T(2) * (v1 - v2 + z1 * z2) / (v2 - v3 + z2);
  const T m = ((T(1) + r2) / T(3) < T(2)) ? (T(1) + r2) / T(3) : T(2);
  const T m2 = (r2 < m) ? r2 : m;
  return (T(0) > m2) ? T(0) : m2;
}

// interior face between q0 = val(idx) and q1 = val(idx+1) (ERB flux(), order S)
template <int S, typename T>
__device__ __forceinline__ T fv_int(T qm1, T q0, T q1, T q2, T vel) {
  if constexpr (S == UD3KOREN)
    return (q0 + T(0.5) * koren_phi(q1, q0, qm1) * (q0 - qm1)) * (T(0.5) + fsign(T(0.5), vel))
         + (q1 + T(0.5) * koren_phi(q0, q1, q2) * (q1 - q2)) * (T(0.5) - fsign(T(0.5), vel));
  else if constexpr (S == UD3)
    return (T(F31) * (q2 + qm1) + T(F32) * (q1 + q0)) - (T(F31) * (q2 - qm1) + T(F33) * (q1 - q0)) * fsign(T(1), vel);
  else if constexpr (S == CD4)
    return T(F41) * (q1 + q0) + T(F42) * (q2 + qm1);
  else if constexpr (S == CD2)  // scale_atmos_dyn_fvm_flux_cd2.F90:168-171
    return T(F2) * (q1 + q0);
  else
    return T(F1) * (q1 + q0) - fsign(T(F1), vel) * (q1 - q0);
}
// lowest face (val(KS), val(KS+1), val(KS+2)) -- scale_atmos_dyn_fvm_flux_ud3.F90:225-229
template <int S, typename T>
__device__ __forceinline__ T fv_lo(T q0, T q1, T q2, T vel) {
  if constexpr (S == UD3KOREN)  // ud3Koren1993.F90:234-241
    return q0 * (T(0.5) + fsign(T(0.5), vel))
         + (q1 + T(0.5) * koren_phi(q0, q1, q2) * (q1 - q2)) * (T(0.5) - fsign(T(0.5), vel));
  else if constexpr (S == UD3)
    return T(F2) * (q1 + q0) * (T(0.5) + fsign(T(0.5), vel)) + (T(2) * q0 + T(5) * q1 - q2) / T(6) * (T(0.5) - fsign(T(0.5), vel));
  else if constexpr (S == UD1)  // the UD1 face has no boundary form (UD1.F90:135-171)
    return fv_int<UD1, T>(T(0), q0, q1, T(0), vel);
  else
    return T(F2) * (q1 + q0);
}
// highest face (val(KE-2), val(KE-1), val(KE)) -- scale_atmos_dyn_fvm_flux_ud3.F90:231-235
template <int S, typename T>
__device__ __forceinline__ T fv_hi(T qm1, T q0, T q1, T vel) {
  if constexpr (S == UD3KOREN)  // ud3Koren1993.F90:242-249 (the limiter argument order of the reference)
    return (q0 + T(0.5) * koren_phi(qm1, q0, q1) * (q0 - q1)) * (T(0.5) + fsign(T(0.5), vel))
         + q1 * (T(0.5) - fsign(T(0.5), vel));
  else if constexpr (S == UD3)
    return (T(2) * q1 + T(5) * q0 - qm1) / T(6) * (T(0.5) + fsign(T(0.5), vel)) + T(F2) * (q1 + q0) * (T(0.5) - fsign(T(0.5), vel));
  else if constexpr (S == UD1)
    return fv_int<UD1, T>(T(0), q0, q1, T(0), vel);
  else
    return T(F2) * (q1 + q0);
}

// Six-point faces of the UD5 and CD6 tracer schemes (the lateral halo is 3).
// Interior face between q0 = val(idx) and q1 = val(idx+1)
// (scale_atmos_dyn_fvm_flux_ud5.F90:236-242, scale_atmos_dyn_fvm_flux_cd6.F90:195-199)
template <int S>
__device__ __forceinline__ double fv6_int(double qm2, double qm1, double q0, double q1, double q2, double q3, double vel) {
  if constexpr (S == UD5)
    return (F51 * (q3 + qm2) + F52 * (q2 + qm1) + F53 * (q1 + q0))
         - (F51 * (q3 - qm2) + F54 * (q2 - qm1) + F55 * (q1 - q0)) * fsign(1.0, vel);
  else
    return F61 * (q1 + q0) + F62 * (q2 + qm1) + F63 * (q3 + qm2);
}
// the faces next to the walls; v holds val(KS..) for KS and KS+1, val(KE-4..KE)
// for KE-2 and val(KE-2..KE) for KE-1 (ud5.F90:270-302, cd6.F90:217-234)
template <int S>
__device__ __forceinline__ double fv6_ks(const double* v, double vel) {
  if constexpr (S == UD5)
    return F2 * (v[1] + v[0]) * (0.5 + fsign(0.5, vel)) + (2.0 * v[0] + 5.0 * v[1] - v[2]) / 6.0 * (0.5 - fsign(0.5, vel));
  else
    return F2 * (v[1] + v[0]);
}
template <int S>
__device__ __forceinline__ double fv6_ke1(const double* v, double vel) {
  if constexpr (S == UD5)
    return (2.0 * v[2] + 5.0 * v[1] - v[0]) / 6.0 * (0.5 + fsign(0.5, vel)) + F2 * (v[2] + v[1]) * (0.5 - fsign(0.5, vel));
  else
    return F2 * (v[2] + v[1]);
}
template <int S>
__device__ __forceinline__ double fv6_ks1(const double* v, double vel) {
  if constexpr (S == UD5)
    return (2.0 * v[2] + 5.0 * v[1] - v[0]) / 6.0 * (0.5 + fsign(0.5, vel))
         + (-3.0 * v[0] + 27.0 * v[1] + 47.0 * v[2] - 13.0 * v[3] + 2.0 * v[4]) / 60.0 * (0.5 - fsign(0.5, vel));
  else
    return F41 * (v[2] + v[1]) + F42 * (v[3] + v[0]);
}
template <int S>
__device__ __forceinline__ double fv6_ke2(const double* v, double vel) {
  if constexpr (S == UD5)
    return (-3.0 * v[4] + 27.0 * v[3] + 47.0 * v[2] - 13.0 * v[1] + 2.0 * v[0]) / 60.0 * (0.5 + fsign(0.5, vel))
         + (2.0 * v[2] + 5.0 * v[3] - v[4]) / 6.0 * (0.5 - fsign(0.5, vel));
  else
    return F41 * (v[3] + v[2]) + F42 * (v[4] + v[1]);
}

// Metric accessors.  TERRAIN=false is the flat Cartesian grid (GSQRT=1,
// J13G=J23G=0, MAPF=1): loads are skipped and the constants fold, which
// reproduces the general expressions exactly (up to the sign of exact zeros).
template <bool TERRAIN>
struct Met {
  const double* G;
  const double* J13;
  const double* J23;
  const double* MAPF;
  size_t n3;
  size_t pl;
  int IA;
  __device__ __forceinline__ double gs(int c, size_t p) const {
    if constexpr (TERRAIN) return G[p + n3 * (size_t)c]; else return 1.0;
  }
  __device__ __forceinline__ double j13(int c, size_t p) const {
    if constexpr (TERRAIN) return J13[p + n3 * (size_t)c]; else return 0.0;
  }
  __device__ __forceinline__ double j23(int c, size_t p) const {
    if constexpr (TERRAIN) return J23[p + n3 * (size_t)c]; else return 0.0;
  }
  // MAPF(i,j,n,c) with n = 0,1 (Fortran 1,2)
  __device__ __forceinline__ double mapf(int i, int j, int n, int c) const {
    if constexpr (TERRAIN) return MAPF[(size_t)i + (size_t)IA * (size_t)j + pl * (size_t)(n + 2 * c)];
    else return 1.0;
  }
};

// one-dimensional grid arrays, all indexed with the 0-based level/column
// most levels (KA) of a mixed-precision grid: the kernels keep the level
// profiles of the FP32 deviations in LDS
constexpr int KA_DEV_MAX = 512;

struct Grid1D {
  const double *CDZ, *FDZ, *RCDZ, *RFDZ;
  const double *CDX, *FDX, *RCDX, *RFDX;
  const double *CDY, *FDY, *RCDY, *RFDY;
};

// XCD-aware block index: hardware hands consecutive workgroup ids to the 8
// XCDs round-robin; remap so that each XCD walks one contiguous range of the
// logical tile order (neighbouring tiles then share that XCD's L2).
__device__ __forceinline__ unsigned xcd_block(unsigned bid, unsigned nblocks) {
  constexpr unsigned NXCD = 8;
  const unsigned per = (nblocks + NXCD - 1) / NXCD;
  const unsigned xcd = bid % NXCD, loc = bid / NXCD;
  const unsigned full = nblocks % NXCD == 0 ? NXCD : nblocks % NXCD;  // XCDs that own 'per' blocks
  unsigned base = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
  return base + loc;
}

}  // namespace shp
