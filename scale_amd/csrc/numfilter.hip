// numfilter.hip -- hyperdiffusion fluxes num_diff(KA,IA,JA,5,3) on gfx950.
//
// Reference: ATMOS_DYN_numfilter_coef with ND_ORDER = 1 and ND_USE_RS,
//   scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90:385-985,
//   calc_diff3 :1652-1800, coefficient tables filter_setup :92-319.
// Blocks of 64 (i) x TJ (j) columns march k.  At each level the block
// evaluates the five differenced fields (DENS-REF_dens, w, u, v,
// theta-REF_pott) once per tile cell (2-cell halo) into LDS for the x and y
// differences; each column keeps the fields at k-2..k+2 in registers for the
// z differences.  The periodic halo fill of the 15 outputs follows as a
// separate launch.  R is the storage type of the prognostics and of num_diff
// (double, or float for the mixed-precision path, where the differences from
// the reference state are taken in double).
#include <stdlib.h>

#include <vector>

#include "common.h"
#include "shp_internal.h"
#include "batch.h"

namespace shp {

template <typename R>
struct NfArgsT {
  Dims d;
  const R *DENS, *MOMZ, *MOMX, *MOMY, *RHOT;
  const double *REF_dens, *REF_pott;
  const double *CNZ3, *CNX3, *CNY3;  // (3,KA,2), (3,IA,2), (3,JA,2)
  const double *ndc;                 // nd_cdz(KA) nd_fdz(KA) nd_cdx(IA) nd_fdx(IA) nd_cdy(JA) nd_fdy(JA)
  double sfc;
  R* num_diff;
  // FP32: DENS and RHOT stored as deviations from the level profiles offD, offR
  const double *offD = nullptr, *offR = nullptr;
};

// coefficient tables of calc_diff3 (CNZ3/CNX3/CNY3 of filter_setup)
template <typename R>
struct Nf {
  const NfArgsT<R>& a;
  const Dims& d;
  __device__ Nf(const NfArgsT<R>& a_) : a(a_), d(a_.d) {}
  __device__ __forceinline__ double cz(int n, int k, int b) const { return a.CNZ3[(n - 1) + 3 * ((size_t)k + (size_t)d.KA * (b - 1))]; }
  __device__ __forceinline__ double cx(int n, int i, int b) const { return a.CNX3[(n - 1) + 3 * ((size_t)i + (size_t)d.IA * (b - 1))]; }
  __device__ __forceinline__ double cy(int n, int j, int b) const { return a.CNY3[(n - 1) + 3 * ((size_t)j + (size_t)d.JA * (b - 1))]; }
};

// a field minus its reference-state value (COMMON.F90:545, 882), in double
template <typename R>
__device__ __forceinline__ R phi_ref(R v, double ref) {
  if constexpr (sizeof(R) == sizeof(double)) return v - ref;
  else return R((double)v - ref);
}

// one tile (logical index lb of a grid ni_tiles tiles wide)
template <int TJ, typename R>
__device__ __forceinline__ void numfilter_tile_body(const NfArgsT<R>& a, unsigned lb, unsigned ni_tiles) {
  constexpr int TI = 64, CW = TI + 4, RW = TJ + 4, NT = TI * TJ;
  __shared__ R sD[2][RW + 1][CW + 1];  // DENS at k and k+1, one more cell east and north
  __shared__ R sP[5][RW][CW];          // the differenced fields at level k
  const Dims& d = a.d;
  const Nf<R> n(a);
  const int i0 = d.IS + (int)(lb % ni_tiles) * TI;
  const int j0 = d.JS + (int)(lb / ni_tiles) * TJ;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TI * ty;
  const int i = i0 + tx, j = j0 + ty;
  const bool own = (i <= d.IE) && (j <= d.JE);
  const int KS = d.KS, KE = d.KE;
  const size_t n3 = d.n3;
  const double* nd_cdz = a.ndc;
  const double* nd_fdz = a.ndc + d.KA;
  const double* nd_cdx = a.ndc + 2 * d.KA;
  const double* nd_fdx = a.ndc + 2 * d.KA + d.IA;
  const double* nd_cdy = a.ndc + 2 * d.KA + 2 * d.IA;
  const double* nd_fdy = a.ndc + 2 * d.KA + 2 * d.IA + d.JA;
  const size_t p0 = own ? IX(d, 0, i, j) : 0;
  auto out = [&](int v, int dir, int k) -> R& { return a.num_diff[p0 + d.pl * (size_t)k + n3 * (size_t)(v + 5 * dir)]; };
  // full density and rho*theta at level k (the stored values in FP64)
  constexpr bool DEV = sizeof(R) == sizeof(float);
  __shared__ double sOff[DEV ? 2 * KA_DEV_MAX : 1];  // the level profiles
  if constexpr (DEV) {
    for (int q = tid; q < d.KA; q += NT) { sOff[q] = a.offD[q]; sOff[KA_DEV_MAX + q] = a.offR[q]; }
    __syncthreads();
  }
  auto fD = [&](R v, int k) -> R { if constexpr (DEV) return R((double)v + sOff[k]); else return v; };
  auto fT = [&](R v, int k) -> R { if constexpr (DEV) return R((double)v + sOff[KA_DEV_MAX + k]); else return v; };
  auto fillD = [&](R (*dst)[CW + 1], int k) {
    for (int q = tid; q < (RW + 1) * (CW + 1); q += NT) {
      const int r = q / (CW + 1), c = q - r * (CW + 1);
      const int ii = i0 - 2 + c, jj = j0 - 2 + r;
      dst[r][c] = (ii < d.IA && jj < d.JA) ? fD(a.DENS[IX(d, k, ii, jj)], k) : R(1);
    }
  };
  // the five fields at (k, ii, jj) for the own column (COMMON.F90:545, 631, 715, 798, 882)
  auto phis = [&](int k, R* ph) {
    const size_t q = p0 + d.pl * (size_t)k;
    const R D = fD(a.DENS[q], k);
    ph[0] = phi_ref<R>(D, a.REF_dens[q]);
    ph[4] = phi_ref<R>(fT(a.RHOT[q], k) / D, a.REF_pott[q]);
    ph[1] = R(2) * a.MOMZ[q] / (fD(a.DENS[q + d.pl], k + 1) + D);
    ph[2] = R(2) * a.MOMX[q] / (fD(a.DENS[q + 1], k) + D);
    ph[3] = R(2) * a.MOMY[q] / (fD(a.DENS[q + d.IA], k) + D);
  };

  // the z coefficient tables in LDS (CNZ3 (3,KA,2), then nd_cdz, nd_fdz):
  // the level's arithmetic then issues no global load
  extern __shared__ double sZt[];
  for (int q = tid; q < 6 * d.KA; q += NT) sZt[q] = a.CNZ3[q];
  for (int q = tid; q < d.KA; q += NT) { sZt[6 * d.KA + q] = nd_cdz[q]; sZt[7 * d.KA + q] = nd_fdz[q]; }
  // the own column's x and y coefficients, fixed over k
  const int ic = own ? i : d.IS, jc = own ? j : d.JS;
  const R cxa1 = R(n.cx(1, ic + 1, 1)), cxa2 = R(n.cx(2, ic + 1, 1)), cxa3 = R(n.cx(3, ic + 1, 1)), cxa4 = R(n.cx(1, ic, 1));
  const R cxb1 = R(n.cx(1, ic, 2)), cxb2 = R(n.cx(2, ic, 2)), cxb3 = R(n.cx(3, ic, 2)), cxb4 = R(n.cx(1, ic - 1, 2));
  const R cya1 = R(n.cy(1, jc + 1, 1)), cya2 = R(n.cy(2, jc + 1, 1)), cya3 = R(n.cy(3, jc + 1, 1)), cya4 = R(n.cy(1, jc, 1));
  const R cyb1 = R(n.cy(1, jc, 2)), cyb2 = R(n.cy(2, jc, 2)), cyb3 = R(n.cy(3, jc, 2)), cyb4 = R(n.cy(1, jc - 1, 2));
  const R ncdx = R(nd_cdx[ic]), nfdx = R(nd_fdx[ic]), ncdy = R(nd_cdy[jc]), nfdy = R(nd_fdy[jc]);
  __syncthreads();

  // levels outside KS..KE are zero (KS-1 gets its DENS/RHOT z fluxes below)
  if (own) {
    for (int k = 0; k < d.KA; ++k) {
      if (k >= KS && k <= KE) continue;
      for (int q = 0; q < 15; ++q) out(q % 5, q / 5, k) = R(0);
    }
  }
  // z windows: w[f][m] = field f at level k-2+m
  R w[5][5];
  for (int f = 0; f < 5; ++f) for (int m = 0; m < 5; ++m) w[f][m] = R(0);
  if (own) {
    R ph[5];
    for (int m = 3; m < 5; ++m) {  // shifted down once at the first level
      phis(KS - 3 + m, ph);
      for (int f = 0; f < 5; ++f) w[f][m] = ph[f];
    }
  }
  int s0 = 0;
  fillD(sD[0], KS);

  // The global loads of level k, issued during the arithmetic of level k-1:
  // the raw fields of this thread's tile cells (the cells of the block's own
  // columns excepted: their five fields at level k are the middle of those
  // columns' z windows, from the same loads two levels earlier), DENS at k+1
  // for the ring, and the raw fields of the own column at k+2.
  constexpr int NC = (RW * CW + NT - 1) / NT;               // tile cells per thread
  constexpr int NF = ((RW + 1) * (CW + 1) + NT - 1) / NT;   // DENS ring cells per thread
  auto owned = [&](int r, int c) {  // tile cell (r, c) is the own column of a thread of this block
    return r >= 2 && r < TJ + 2 && c >= 2 && c < TI + 2 && i0 - 2 + c <= d.IE && j0 - 2 + r <= d.JE;
  };
  R rZ[NC], rX[NC], rY[NC], rR[NC];
  double rRD[NC], rRP[NC];
  R rDn[NF];
  R oD = 1, oDz = 1, oDx = 1, oDy = 1, oR = 0, oZ = 0, oX = 0, oY = 0;
  double oRD = 0.0, oRP = 0.0;
  auto issue = [&](int k) {
#pragma unroll
    for (int h = 0; h < NC; ++h) {
      const int q = tid + h * NT;
      const int r = q / CW, c = q - r * CW;
      const int ii = i0 - 2 + c, jj = j0 - 2 + r;
      const bool ok = q < RW * CW && ii < d.IA && jj < d.JA && !owned(r, c);
      const size_t g = ok ? IX(d, k, ii, jj) : 0;
      rZ[h] = ok ? a.MOMZ[g] : R(0); rX[h] = ok ? a.MOMX[g] : R(0); rY[h] = ok ? a.MOMY[g] : R(0);
      rR[h] = ok ? fT(a.RHOT[g], k) : R(0); rRD[h] = ok ? a.REF_dens[g] : 0.0; rRP[h] = ok ? a.REF_pott[g] : R(0);
    }
#pragma unroll
    for (int h = 0; h < NF; ++h) {
      const int q = tid + h * NT;
      const int r = q / (CW + 1), c = q - r * (CW + 1);
      const int ii = i0 - 2 + c, jj = j0 - 2 + r;
      const bool ok = q < (RW + 1) * (CW + 1) && ii < d.IA && jj < d.JA;
      rDn[h] = ok ? fD(a.DENS[IX(d, k + 1, ii, jj)], k + 1) : R(1);
    }
    if (own && k + 2 <= KE) {
      const size_t q = p0 + d.pl * (size_t)(k + 2);
      oD = a.DENS[q]; oDz = a.DENS[q + d.pl]; oDx = a.DENS[q + 1]; oDy = a.DENS[q + d.IA];
      oR = a.RHOT[q]; oZ = a.MOMZ[q]; oX = a.MOMX[q]; oY = a.MOMY[q];
      oRD = a.REF_dens[q]; oRP = a.REF_pott[q];
    }
  };
  if (KS <= KE) issue(KS);
  for (int k = KS; k <= KE; ++k) {
    const int s1 = s0 ^ 1;
#pragma unroll
    for (int h = 0; h < NF; ++h) {
      const int q = tid + h * NT;
      if (q >= (RW + 1) * (CW + 1)) continue;
      const int r = q / (CW + 1), c = q - r * (CW + 1);
      sD[s1][r][c] = rDn[h];
    }
    if (own) {
      // the five fields at k+2 (COMMON.F90:545, 631, 715, 798, 882), as phis computes them
      R ph[5] = {0, 0, 0, 0, 0};
      if (k + 2 <= KE) {
        const int kk = k + 2;
        const R D = fD(oD, kk);
        ph[0] = phi_ref<R>(D, oRD);
        ph[4] = phi_ref<R>(fT(oR, kk) / D, oRP);
        ph[1] = R(2) * oZ / (fD(oDz, kk + 1) + D);
        ph[2] = R(2) * oX / (fD(oDx, kk) + D);
        ph[3] = R(2) * oY / (fD(oDy, kk) + D);
      }
      for (int f = 0; f < 5; ++f) { w[f][0] = w[f][1]; w[f][1] = w[f][2]; w[f][2] = w[f][3]; w[f][3] = w[f][4]; w[f][4] = ph[f]; }
    }
    __syncthreads();
    if (own)
      for (int f = 0; f < 5; ++f) sP[f][ty + 2][tx + 2] = w[f][2];
#pragma unroll
    for (int h = 0; h < NC; ++h) {
      const int q = tid + h * NT;
      if (q >= RW * CW) continue;
      const int r = q / CW, c = q - r * CW;
      const int ii = i0 - 2 + c, jj = j0 - 2 + r;
      if (owned(r, c)) continue;
      if (ii >= d.IA || jj >= d.JA) { for (int f = 0; f < 5; ++f) sP[f][r][c] = R(0); continue; }
      const R D = sD[s0][r][c];
      sP[0][r][c] = phi_ref<R>(D, rRD[h]);
      sP[1][r][c] = R(2) * rZ[h] / (sD[s1][r][c] + D);
      sP[2][r][c] = R(2) * rX[h] / (sD[s0][r][c + 1] + D);
      sP[3][r][c] = R(2) * rY[h] / (sD[s0][r + 1][c] + D);
      sP[4][r][c] = phi_ref<R>(rR[h] / D, rRP[h]);
    }
    __syncthreads();
    // the next level's loads, in flight through this level's arithmetic
    if (k < KE) issue(k + 1);
    if (own) {
      const int r = ty + 2, c = tx + 2;
      auto Dk = [&](int di, int dj) { return sD[s0][r + dj][c + di]; };
      auto Dk1 = [&](int di, int dj) { return sD[s1][r + dj][c + di]; };
      auto cz = [&](int nn, int kk, int b) { return R(sZt[(nn - 1) + 3 * (kk + d.KA * (b - 1))]); };
      auto ncdz = [&](int kk) { return R(sZt[6 * d.KA + kk]); };
      auto nfdz = [&](int kk) { return R(sZt[7 * d.KA + kk]); };
      // calc_diff3 ZDIR KO = 0 at level k in KS..KE (COMMON.F90:1668-1697);
      // w[f][m] holds level k-2+m
      auto dz0 = [&](int f) {
        const R* P = w[f];
        if (k == KS)
          return (+cz(1, KS + 1, 1) * P[4] - cz(2, KS + 1, 1) * P[3]
                  + cz(3, KS + 1, 1) * P[2] - cz(1, KS, 1) * P[3]);
        if (k == KE - 1)
          return (+cz(1, KE, 1) * P[2] - cz(2, KE, 1) * P[3]
                  + cz(3, KE, 1) * P[2] - cz(1, KE - 1, 1) * P[1]);
        if (k == KE)
          return -(+cz(1, KE, 1) * P[1] - cz(2, KE, 1) * P[2]
                   + cz(3, KE, 1) * P[1] - cz(1, KE - 1, 1) * P[0]);
        return (+cz(1, k + 1, 1) * P[4] - cz(2, k + 1, 1) * P[3]
                + cz(3, k + 1, 1) * P[2] - cz(1, k, 1) * P[1]);
      };
      // calc_diff3 ZDIR KO = 1 at level k in KS+1..KE-1 (COMMON.F90:1699-1731)
      auto dz1 = [&](int f) {
        const R* P = w[f];
        if (k == KS + 1)
          return (+cz(1, KS + 1, 2) * P[3] - cz(2, KS + 1, 2) * P[2]
                  + cz(3, KS + 1, 2) * P[1] - cz(1, KS, 2) * P[2]);
        if (k == KE - 1)
          return (-cz(2, KE - 1, 2) * P[2] + cz(3, KE - 1, 2) * P[1]
                  - cz(1, KE - 2, 2) * P[0]);
        return (+cz(1, k, 2) * P[3] - cz(2, k, 2) * P[2]
                + cz(3, k, 2) * P[1] - cz(1, k - 1, 2) * P[0]);
      };
      // calc_diff3 XDIR / YDIR (COMMON.F90:1733-1800)
      auto dx = [&](int f, int io) {
        if (io == 0)
          return (+cxa1 * sP[f][r][c + 2] - cxa2 * sP[f][r][c + 1]
                  + cxa3 * sP[f][r][c] - cxa4 * sP[f][r][c - 1]);
        return (+cxb1 * sP[f][r][c + 1] - cxb2 * sP[f][r][c]
                + cxb3 * sP[f][r][c - 1] - cxb4 * sP[f][r][c - 2]);
      };
      auto dy = [&](int f, int jo) {
        if (jo == 0)
          return (+cya1 * sP[f][r + 2][c] - cya2 * sP[f][r + 1][c]
                  + cya3 * sP[f][r][c] - cya4 * sP[f][r - 1][c]);
        return (+cyb1 * sP[f][r + 1][c] - cyb2 * sP[f][r][c]
                + cyb3 * sP[f][r - 1][c] - cyb4 * sP[f][r - 2][c]);
      };
      auto sfc = [&](int kk, R v) {  // surface factor rows (COMMON.F90:593-594)
        if (kk == KS) return v * R(a.sfc);
        if (kk == KS + 1) return v * (R(1) + R(a.sfc)) * R(0.5);
        return v;
      };
      const bool kinm = k <= KE - 1;
      // density (COMMON.F90:552-621)
      const R dz0d = dz0(0);
      out(I_DENS, ZDIR, k) = dz0d * ncdz(k);
      if (k == KS) out(I_DENS, ZDIR, KS - 1) = (-dz0d) * ncdz(KS - 1);
      out(I_DENS, XDIR, k) = sfc(k, dx(0, 0) * ncdx);
      out(I_DENS, YDIR, k) = sfc(k, dy(0, 0) * ncdy);
      // z-momentum (COMMON.F90:637-707)
      out(I_MOMZ, ZDIR, k) = (k >= KS + 1 && k <= KE - 1) ? dz1(1) * nfdz(k) * Dk(0, 0) : R(0);
      out(I_MOMZ, XDIR, k) = kinm ? sfc(k, dx(1, 0) * ncdx * R(0.25) * (Dk1(1, 0) + Dk1(0, 0) + Dk(1, 0) + Dk(0, 0))) : R(0);
      out(I_MOMZ, YDIR, k) = kinm ? sfc(k, dy(1, 0) * ncdy * R(0.25) * (Dk1(0, 1) + Dk1(0, 0) + Dk(0, 1) + Dk(0, 0))) : R(0);
      // x-momentum (COMMON.F90:721-790)
      out(I_MOMX, ZDIR, k) = kinm ? dz0(2) * ncdz(k) * R(0.25) * (Dk1(1, 0) + Dk1(0, 0) + Dk(1, 0) + Dk(0, 0)) : R(0);
      out(I_MOMX, XDIR, k) = sfc(k, dx(2, 1) * nfdx * Dk(0, 0));
      out(I_MOMX, YDIR, k) = sfc(k, dy(2, 0) * ncdy * R(0.25) * (Dk(1, 1) + Dk(1, 0) + Dk(0, 1) + Dk(0, 0)));
      // y-momentum (COMMON.F90:804-873)
      out(I_MOMY, ZDIR, k) = kinm ? dz0(3) * ncdz(k) * R(0.25) * (Dk1(0, 1) + Dk1(0, 0) + Dk(0, 1) + Dk(0, 0)) : R(0);
      out(I_MOMY, XDIR, k) = sfc(k, dx(3, 0) * ncdx * R(0.25) * (Dk(1, 1) + Dk(0, 1) + Dk(1, 0) + Dk(0, 0)));
      out(I_MOMY, YDIR, k) = sfc(k, dy(3, 1) * nfdy * Dk(0, 0));
      // rho*theta (COMMON.F90:888-962)
      const R dz04 = dz0(4);
      if (kinm) out(I_RHOT, ZDIR, k) = dz04 * ncdz(k) * R(0.5) * (Dk1(0, 0) + Dk(0, 0));
      else out(I_RHOT, ZDIR, k) = dz04 * ncdz(KE) * Dk(0, 0);
      if (k == KS) out(I_RHOT, ZDIR, KS - 1) = (-dz04) * ncdz(KS - 1) * Dk(0, 0);
      out(I_RHOT, XDIR, k) = sfc(k, dx(4, 0) * ncdx * R(0.5) * (Dk(1, 0) + Dk(0, 0)));
      out(I_RHOT, YDIR, k) = sfc(k, dy(4, 0) * ncdy * R(0.5) * (Dk(0, 1) + Dk(0, 0)));
    }
    __syncthreads();
    s0 = s1;
  }
}

// FP32 with 8 rows: held to 128 VGPRs, four waves per SIMD (at 134 VGPRs,
// three, the 512x512x128 filter takes 1.50 ms instead of 1.06)
template <int TJ, typename R>
constexpr int nf_min_waves() { return (TJ == 8 && sizeof(R) == sizeof(float)) ? 4 : 1; }

template <int TJ, typename R>
__global__ void __launch_bounds__(64 * TJ) __attribute__((amdgpu_waves_per_eu(nf_min_waves<TJ, R>())))
numfilter_tile(NfArgsT<R> a, unsigned ni_tiles, unsigned nblocks) {
  numfilter_tile_body<TJ, R>(a, xcd_block(blockIdx.x, nblocks), ni_tiles);
}

// the filters of many members in one launch (tables as hevi_tile_batch)
template <int TJ, typename R>
__global__ void __launch_bounds__(64 * TJ) __attribute__((amdgpu_waves_per_eu(nf_min_waves<TJ, R>())))
numfilter_tile_batch(const NfArgsT<R>* __restrict__ tab,
                                                              const unsigned* __restrict__ first,
                                                              const unsigned* __restrict__ ni, int nmem, unsigned total) {
  for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
    const unsigned lt = xcd_block(t, total);
    const int lo = batch_member(first, nmem, lt);
    numfilter_tile_body<TJ, R>(tab[lo], lt - first[lo], ni[lo]);
    __syncthreads();
  }
}

template <typename R>
static hipError_t numfilter_t(const Ctx& c, R* num_diff, const R* const st[5], const double* REF_dens,
                              const double* REF_pott, hipStream_t s, const double* offD = nullptr,
                              const double* offR = nullptr) {
  NfArgsT<R> a;
  a.offD = offD; a.offR = offR;
  a.d = c.d;
  a.DENS = st[0]; a.MOMZ = st[1]; a.MOMX = st[2]; a.MOMY = st[3]; a.RHOT = st[4];
  a.REF_dens = REF_dens; a.REF_pott = REF_pott;
  a.CNZ3 = c.CNZ3; a.CNX3 = c.CNX3; a.CNY3 = c.CNY3; a.ndc = c.ndc; a.sfc = c.nd_sfc_fact;
  a.num_diff = num_diff;
  const Dims& d = c.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1;
  // tile height: 8 rows when that gives every CU a block, 4 otherwise (at
  // 512x512x128 the filter takes 1.95 ms at 8 rows, 1.96 at 4 and 4.04 at 16
  // in FP64, whose 1024-thread blocks are held to 128 VGPRs and spill; FP32
  // 1.11, 1.79 and 1.18 ms); the context's knob (scale_hip_tune_set), else
  // SCALE_HIP_NF_TJ = 4, 8 or 16 in the environment, overrides
  static const int ncu = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      n = 256;
    return n > 0 ? n : 256;
  }();
  static const int TJenv = [] {
    const char* e = getenv("SCALE_HIP_NF_TJ");
    const int v = e ? atoi(e) : 0;
    return (v == 4 || v == 8 || v == 16) ? v : 0;
  }();
  const unsigned ni = (imax + 63) / 64;
  int TJ = c.nf_tj ? c.nf_tj : TJenv;
  if (!TJ) {
    TJ = 4;
    if (ni * (unsigned)((jmax + 7) / 8) >= (unsigned)ncu) TJ = 8;
  }
  const unsigned nj = (jmax + TJ - 1) / TJ;
  const unsigned nb = ni * nj;
  const size_t lds = 8 * (size_t)d.KA * sizeof(double);  // the z coefficient tables
  prof_start(c, "numfilter", s);
  if (TJ == 16) hipLaunchKernelGGL((numfilter_tile<16, R>), dim3(nb), dim3(64, 16), lds, s, a, ni, nb);
  else if (TJ == 8) hipLaunchKernelGGL((numfilter_tile<8, R>), dim3(nb), dim3(64, 8), lds, s, a, ni, nb);
  else hipLaunchKernelGGL((numfilter_tile<4, R>), dim3(nb), dim3(64, 4), lds, s, a, ni, nb);
  prof_stop(c, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  R* arrs[15];
  for (int q = 0; q < 15; ++q) arrs[q] = num_diff + d.n3 * (size_t)q;
  return halo_fill(c, arrs, 15, s);
}

// ---- ensemble tables (batch.h) ----------------------------------------------------
template <typename R>
static hipError_t nf_batch_build_t(const NfIn<R>* in, int n, int tj, BatchBuf& b) {
  if (n <= 0 || (tj != 4 && tj != 8)) return hipErrorInvalidValue;
  std::vector<NfArgsT<R>> args((size_t)n);
  std::vector<unsigned> first((size_t)n + 1, 0), ni((size_t)n, 0);
  int ka_max = 0;
  for (int m = 0; m < n; ++m) {
    const Ctx& c = *in[m].c;
    NfArgsT<R>& a = args[(size_t)m];
    a.offD = in[m].offD; a.offR = in[m].offR;
    a.d = c.d;
    a.DENS = in[m].st[0]; a.MOMZ = in[m].st[1]; a.MOMX = in[m].st[2]; a.MOMY = in[m].st[3]; a.RHOT = in[m].st[4];
    a.REF_dens = in[m].REF_dens; a.REF_pott = in[m].REF_pott;
    a.CNZ3 = c.CNZ3; a.CNX3 = c.CNX3; a.CNY3 = c.CNY3; a.ndc = c.ndc; a.sfc = c.nd_sfc_fact;
    a.num_diff = in[m].num_diff;
    const Dims& d = c.d;
    ni[(size_t)m] = (unsigned)((d.IE - d.IS + 1 + 63) / 64);
    first[(size_t)m + 1] = first[(size_t)m] + ni[(size_t)m] * (unsigned)((d.JE - d.JS + 1 + tj - 1) / tj);
    ka_max = d.KA > ka_max ? d.KA : ka_max;
  }
  const hipError_t e = batch_upload(args, first, ni, b);
  b.tj = tj; b.fp32 = sizeof(R) == sizeof(float); b.ka_max = ka_max;
  return e;
}
hipError_t numfilter_batch_build(const NfIn<double>* in, int n, int tj, BatchBuf& b) { return nf_batch_build_t(in, n, tj, b); }
hipError_t numfilter_batch_build(const NfIn<float>* in, int n, int tj, BatchBuf& b) { return nf_batch_build_t(in, n, tj, b); }

template <int TJ, typename R>
static hipError_t nf_batch_launch_t(const BatchBuf& b, int bpc, int ncu, hipStream_t s) {
  const NfArgsT<R>* tab;
  const unsigned *first, *ni;
  batch_tables(b, tab, first, ni);
  const unsigned g = batch_grid(b.total, bpc, ncu);
  const size_t lds = 8 * (size_t)b.ka_max * sizeof(double);
  hipLaunchKernelGGL((numfilter_tile_batch<TJ, R>), dim3(g), dim3(64, TJ), lds, s, tab, first, ni, b.nmem, b.total);
  return hipGetLastError();
}
hipError_t numfilter_batch_launch(const BatchBuf& b, int bpc, int ncu, hipStream_t s) {
  if (!b.dev) return hipErrorInvalidValue;
  if (b.fp32) return b.tj == 8 ? nf_batch_launch_t<8, float>(b, bpc, ncu, s) : nf_batch_launch_t<4, float>(b, bpc, ncu, s);
  return b.tj == 8 ? nf_batch_launch_t<8, double>(b, bpc, ncu, s) : nf_batch_launch_t<4, double>(b, bpc, ncu, s);
}

hipError_t numfilter(const Ctx& c, double* num_diff, const double* const st[5], const double* REF_dens,
                     const double* REF_pott, hipStream_t s) {
  return numfilter_t<double>(c, num_diff, st, REF_dens, REF_pott, s);
}

// ---------------------------------------------------------------------------
// Tracer filter num_diff_q(KA,IA,JA,3): ATMOS_DYN_numfilter_coef_q with
// ND_ORDER = 1 (COMMON.F90:989-1200).  The third differences (calc_diff3,
// :1652-1800) of the tracer, or of QTRC - REF_qv for QV (:1085-1098), times
// the density at the flux point (:1117-1184).  It runs once per tracer per
// large step, so one thread per point of the interior columns (all levels)
// reads its stencil through the caches.
struct NfqArgs {
  Dims d;
  const double *DENS, *Q, *REF;      // REF: REF_qv or nullptr
  const double *CNZ3, *CNX3, *CNY3;
  const double* ndq;                 // nd_coef_cdz(KA) nd_coef_cdx(IA) nd_coef_cdy(JA) for ND_COEF_Q and dt
  double sfc;
  double* out;                       // num_diff_q, 3 fields
};

__global__ void __launch_bounds__(256) numfilter_q_kernel(NfqArgs a) {
  const Dims& d = a.d;
  const int i = d.IS + (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = d.JS + (int)blockIdx.z;
  if (i > d.IE || k >= d.KA) return;
  const int KS = d.KS, KE = d.KE;
  auto cz = [&](int n, int kk, int b) { return a.CNZ3[(n - 1) + 3 * ((size_t)kk + (size_t)d.KA * (b - 1))]; };
  auto cx = [&](int n, int ii, int b) { return a.CNX3[(n - 1) + 3 * ((size_t)ii + (size_t)d.IA * (b - 1))]; };
  auto cy = [&](int n, int jj, int b) { return a.CNY3[(n - 1) + 3 * ((size_t)jj + (size_t)d.JA * (b - 1))]; };
  auto phi = [&](int kk, int ii, int jj) {
    const size_t q = IX(d, kk, ii, jj);
    return a.REF ? a.Q[q] - a.REF[q] : a.Q[q];
  };
  auto D = [&](int kk, int ii, int jj) { return a.DENS[IX(d, kk, ii, jj)]; };
  auto dz = [&](int kk) -> double {
    if (kk == KS)
      return cz(1, KS + 1, 1) * phi(KS + 2, i, j) - cz(2, KS + 1, 1) * phi(KS + 1, i, j) + cz(3, KS + 1, 1) * phi(KS, i, j)
           - cz(1, KS, 1) * phi(KS + 1, i, j);
    if (kk == KE - 1)
      return cz(1, KE, 1) * phi(KE - 1, i, j) - cz(2, KE, 1) * phi(KE, i, j) + cz(3, KE, 1) * phi(KE - 1, i, j)
           - cz(1, KE - 1, 1) * phi(KE - 2, i, j);
    return cz(1, kk + 1, 1) * phi(kk + 2, i, j) - cz(2, kk + 1, 1) * phi(kk + 1, i, j) + cz(3, kk + 1, 1) * phi(kk, i, j)
         - cz(1, kk, 1) * phi(kk - 1, i, j);
  };
  const double* cdz = a.ndq;
  const double* cdx = a.ndq + d.KA;
  const double* cdy = a.ndq + d.KA + d.IA;
  const size_t p = IX(d, k, i, j), n3 = d.n3;
  double z = 0.0;
  if (k >= KS && k <= KE - 1) z = dz(k) * cdz[k] * 0.5 * (D(k + 1, i, j) + D(k, i, j));
  else if (k == KS - 1) z = (-dz(KS)) * cdz[KS - 1] * D(KS, i, j);
  else if (k == KE) z = (-dz(KE - 1)) * cdz[KE] * D(KE, i, j);
  double x = 0.0, y = 0.0;
  if (k >= KS && k <= KE) {
    const double wx = cx(1, i + 1, 1) * phi(k, i + 2, j) - cx(2, i + 1, 1) * phi(k, i + 1, j) + cx(3, i + 1, 1) * phi(k, i, j)
                    - cx(1, i, 1) * phi(k, i - 1, j);
    const double wy = cy(1, j + 1, 1) * phi(k, i, j + 2) - cy(2, j + 1, 1) * phi(k, i, j + 1) + cy(3, j + 1, 1) * phi(k, i, j)
                    - cy(1, j, 1) * phi(k, i, j - 1);
    x = wx * cdx[i] * 0.5 * (D(k, i + 1, j) + D(k, i, j));
    y = wy * cdy[j] * 0.5 * (D(k, i, j + 1) + D(k, i, j));
    if (k == KS) { x = x * a.sfc; y = y * a.sfc; }
    else if (k == KS + 1) { x = x * (1.0 + a.sfc) * 0.5; y = y * (1.0 + a.sfc) * 0.5; }
  }
  a.out[p] = z;
  a.out[p + n3] = x;
  a.out[p + 2 * n3] = y;
}

hipError_t numfilter_q(const Ctx& c, double* num_diff_q, const double* DENS, const double* QTRC, const double* REF_qv,
                       hipStream_t s) {
  const Dims& d = c.d;
  NfqArgs a;
  a.d = d; a.DENS = DENS; a.Q = QTRC; a.REF = REF_qv;
  a.CNZ3 = c.CNZ3; a.CNX3 = c.CNX3; a.CNY3 = c.CNY3; a.ndq = c.ndqc; a.sfc = c.ndq_sfc; a.out = num_diff_q;
  const dim3 grid((d.IE - d.IS + 1 + 63) / 64, (d.KA + 3) / 4, d.JE - d.JS + 1);
  prof_start(c, "numfilter_q", s);
  hipLaunchKernelGGL(numfilter_q_kernel, grid, dim3(64, 4), 0, s, a);
  prof_stop(c, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  double* arrs[3] = {num_diff_q, num_diff_q + d.n3, num_diff_q + 2 * d.n3};
  return halo_fill(c, arrs, 3, s);
}
hipError_t numfilter(const Ctx& c, float* num_diff, const float* const st[5], const double* REF_dens,
                     const double* REF_pott, hipStream_t s, const double* offD, const double* offR) {
  return numfilter_t<float>(c, num_diff, st, REF_dens, REF_pott, s, offD, offR);
}

}  // namespace shp
