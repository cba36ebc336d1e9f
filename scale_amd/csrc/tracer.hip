// tracer.hip -- flux-form tracer advection with the FCT limiter on gfx950.
//
// Reference: ATMOS_DYN_Tstep_tracer_fvm_heve,
//   scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_tracer_fvm_heve.F90:72-242,
// and ATMOS_DYN_fct / get_fact_fct,
//   scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90:1922-2734.
//
// Without FCT a stage is one kernel: each thread evaluates the high-order
// fluxes on the six faces of its cell and writes the updated mixing ratio.
// With FCT a stage is three kernels around one halo exchange:
//   fct_philo   low-order (UD1) provisional field phi_lo on the interior plus
//               a one-cell ring (COMMON.F90:2033-2045)
//   fct_ratio   anti-diffusive fluxes, P+/P-, the along-stream (or 7-point)
//               bounds Q+/Q- and the ratios R+/R- (COMMON.F90:1988-2547)
//   [periodic fill of R+/R-]                       (COMMON.F90:2558-2561)
//   fct_update  limited anti-diffusive fluxes and the flux-form update
//               (COMMON.F90:2570-2659, tracer_fvm_heve.F90:197-205)
// The high- and low-order face fluxes are recomputed where needed instead of
// being stored; only phi_lo and R+/R- go through memory.
#include "common.h"
#include "fct.h"
#include "shp_internal.h"
#include "batch.h"

namespace shp {

struct TrArgs {
  Dims d;
  const double *QTRC, *QTRC0, *RHOQ_t, *DENS0, *DENS, *mflx, *nd;
  size_t nds;                     // distance between the z, x and y components of nd
  const double *GSQRT, *MAPF_XY;  // GSQRT(KA,IA,JA,7); MAPF(IA,JA,2) I_XY slice
  const double *RCDZ, *RCDX, *RCDY;
  double dtl;
  double* QTRCo;
  double *phi_lo, *rjpls, *rjmns;  // scratch
};

template <int S, bool TERRAIN>
struct TrOps {
  const TrArgs& a;
  const Dims& d;
  __device__ TrOps(const TrArgs& a_) : a(a_), d(a_.d) {}
  __device__ __forceinline__ size_t ix(int k, int i, int j) const { return IX(d, k, i, j); }
  __device__ __forceinline__ double gs(int c, size_t p) const {
    if constexpr (TERRAIN) return a.GSQRT[p + d.n3 * (size_t)c]; else return 1.0;
  }
  __device__ __forceinline__ double m1(int i, int j) const {
    if constexpr (TERRAIN) return a.MAPF_XY[(size_t)i + (size_t)d.IA * j]; else return 1.0;
  }
  __device__ __forceinline__ double m2(int i, int j) const {
    if constexpr (TERRAIN) return a.MAPF_XY[(size_t)i + (size_t)d.IA * j + d.pl]; else return 1.0;
  }
  // num_diff and the source term: a null array stands for zeros (the model
  // without a tracer filter or tracer tendencies), added as the literal 0.0,
  // which gives the same sums as zeros loaded from memory
  __device__ __forceinline__ double ndv(size_t q) const { return a.nd ? a.nd[q] : 0.0; }
  __device__ __forceinline__ double rqv(size_t q) const { return a.RHOQ_t ? a.RHOQ_t[q] : 0.0; }
  __device__ __forceinline__ double mZ(size_t p) const { return a.mflx[p]; }
  __device__ __forceinline__ double mX(size_t p) const { return a.mflx[p + d.n3]; }
  __device__ __forceinline__ double mY(size_t p) const { return a.mflx[p + 2 * d.n3]; }

  // high-order fluxes on QTRC (fluxZ/X/Y_XYZ of the tracer scheme)
  __device__ double hZ(int kk, int i, int j) const {
    if (kk == d.KS - 1 || kk == d.KE) return 0.0;
    const double* Q = a.QTRC;
    size_t p = ix(kk, i, j);
    double vel = mZ(p), fv;
    if constexpr (S == UD5 || S == CD6) {
      // levels are d.pl apart in this layout; v[] gathers the column values a face form reads.
      // The reference assigns the interior, then KS, KE-1, KS+1 and KE-2: the last one wins.
      const size_t sk = d.pl;
      double v[5];
      auto col = [&](int k0, int n) { for (int m = 0; m < n; ++m) v[m] = Q[ix(k0 + m, i, j)]; };
      if (kk == d.KE - 2) { col(d.KE - 4, 5); fv = fv6_ke2<S>(v, vel); }
      else if (kk == d.KS + 1) { col(d.KS, 5); fv = fv6_ks1<S>(v, vel); }
      else if (kk == d.KE - 1) { col(d.KE - 2, 3); fv = fv6_ke1<S>(v, vel); }
      else if (kk == d.KS) { col(d.KS, 3); fv = fv6_ks<S>(v, vel); }
      else fv = fv6_int<S>(Q[p - 2 * sk], Q[p - sk], Q[p], Q[p + sk], Q[p + 2 * sk], Q[p + 3 * sk], vel);
      return vel * fv + gs(I_XYW, p) * ndv(p);
    }
    if (kk == d.KS) fv = fv_lo<S>(Q[p], Q[ix(kk + 1, i, j)], Q[ix(kk + 2, i, j)], vel);
    else if (kk == d.KE - 1) fv = fv_hi<S>(Q[ix(kk - 1, i, j)], Q[p], Q[ix(kk + 1, i, j)], vel);
    else fv = fv_int<S>(Q[ix(kk - 1, i, j)], Q[p], Q[ix(kk + 1, i, j)], Q[ix(kk + 2, i, j)], vel);
    if constexpr (S == UD1) return vel * fv;  // the UD1 routines take no num_diff (UD1.F90:135-249)
    return vel * fv + gs(I_XYW, p) * ndv(p);
  }
  __device__ double hX(int k, int ii, int j) const {
    const double* Q = a.QTRC;
    size_t p = ix(k, ii, j);
    double vel = mX(p);
    if constexpr (S == UD1) return vel * fv_int<UD1>(0.0, Q[p], Q[ix(k, ii + 1, j)], 0.0, vel);
    if constexpr (S == UD5 || S == CD6)
      return vel * fv6_int<S>(Q[ix(k, ii - 2, j)], Q[ix(k, ii - 1, j)], Q[p], Q[ix(k, ii + 1, j)], Q[ix(k, ii + 2, j)],
                              Q[ix(k, ii + 3, j)], vel) + gs(I_UYZ, p) * ndv(p + a.nds);
    return vel * fv_int<S>(Q[ix(k, ii - 1, j)], Q[p], Q[ix(k, ii + 1, j)], Q[ix(k, ii + 2, j)], vel) + gs(I_UYZ, p) * ndv(p + a.nds);
  }
  __device__ double hY(int k, int i, int jj) const {
    const double* Q = a.QTRC;
    size_t p = ix(k, i, jj);
    double vel = mY(p);
    if constexpr (S == UD1) return vel * fv_int<UD1>(0.0, Q[p], Q[ix(k, i, jj + 1)], 0.0, vel);
    if constexpr (S == UD5 || S == CD6)
      return vel * fv6_int<S>(Q[ix(k, i, jj - 2)], Q[ix(k, i, jj - 1)], Q[p], Q[ix(k, i, jj + 1)], Q[ix(k, i, jj + 2)],
                              Q[ix(k, i, jj + 3)], vel) + gs(I_XVZ, p) * ndv(p + 2 * a.nds);
    return vel * fv_int<S>(Q[ix(k, i, jj - 1)], Q[p], Q[ix(k, i, jj + 1)], Q[ix(k, i, jj + 2)], vel) + gs(I_XVZ, p) * ndv(p + 2 * a.nds);
  }
  // low-order UD1 fluxes on QTRC0 (scale_atmos_dyn_fvm_flux_ud1.F90:135-249)
  __device__ double lZ(int kk, int i, int j) const {
    if (kk == d.KS - 1 || kk == d.KE) return 0.0;
    const double* Q = a.QTRC0;
    size_t p = ix(kk, i, j);
    double vel = mZ(p);
    return vel * fv_int<UD1>(0.0, Q[p], Q[ix(kk + 1, i, j)], 0.0, vel);
  }
  __device__ double lX(int k, int ii, int j) const {
    const double* Q = a.QTRC0;
    size_t p = ix(k, ii, j);
    double vel = mX(p);
    return vel * fv_int<UD1>(0.0, Q[p], Q[ix(k, ii + 1, j)], 0.0, vel);
  }
  __device__ double lY(int k, int i, int jj) const {
    const double* Q = a.QTRC0;
    size_t p = ix(k, i, jj);
    double vel = mY(p);
    return vel * fv_int<UD1>(0.0, Q[p], Q[ix(k, i, jj + 1)], 0.0, vel);
  }
  // anti-diffusive fluxes qflx_hi - qflx_lo (COMMON.F90:1988-2024)
  __device__ double aZ(int kk, int i, int j) const {
    if (kk == d.KS - 1 || kk == d.KE) return 0.0;
    return hZ(kk, i, j) - lZ(kk, i, j);
  }
  __device__ double aX(int k, int ii, int j) const { return hX(k, ii, j) - lX(k, ii, j); }
  __device__ double aY(int k, int i, int jj) const { return hY(k, i, jj) - lY(k, i, jj); }
};

// ---- no FCT: one kernel (tracer_fvm_heve.F90:226-231) ----------------------
// Each pass is a body over the (64 i) x (4 k) x (1 j) tile lb of a grid
// ni_tiles x nk_tiles tiles across, run by its kernel once per block or, for
// an ensemble, by tracer_batch over a table of members.
template <int S, bool TERRAIN>
__device__ __forceinline__ void tracer_nofct_body(const TrArgs& a, unsigned lb, unsigned ni_tiles, unsigned nk_tiles) {
  const Dims& d = a.d;
  TrOps<S, TERRAIN> o(a);
  const unsigned it = lb % ni_tiles, kt = (lb / ni_tiles) % nk_tiles, jt = lb / (ni_tiles * nk_tiles);
  const int i = d.IS + (int)(it * 64 + threadIdx.x);
  const int k = d.KS + (int)(kt * 4 + threadIdx.y);
  const int j = d.JS + (int)jt;
  if (i > d.IE || k > d.KE || j > d.JE) return;
  const size_t p = IX(d, k, i, j);
  a.QTRCo[p] = (a.QTRC0[p] * a.DENS0[p]
                + a.dtl * (-((o.hZ(k, i, j) - o.hZ(k - 1, i, j)) * a.RCDZ[k]
                             + (o.hX(k, i, j) - o.hX(k, i - 1, j)) * a.RCDX[i]
                             + (o.hY(k, i, j) - o.hY(k, i, j - 1)) * a.RCDY[j])
                           * o.m1(i, j) * o.m2(i, j) / o.gs(I_XYZ, p)
                           + o.rqv(p)))
             / a.DENS[p];
}
template <int S, bool TERRAIN>
__global__ void __launch_bounds__(256) tracer_nofct(TrArgs a, unsigned ni_tiles, unsigned nk_tiles, unsigned nblocks) {
  tracer_nofct_body<S, TERRAIN>(a, xcd_block(blockIdx.x, nblocks), ni_tiles, nk_tiles);
}

// ---- FCT pass 1: phi_lo on the interior plus a one-cell ring ----------------
template <bool TERRAIN>
__device__ __forceinline__ void fct_philo_body(const TrArgs& a, unsigned lb, unsigned ni_tiles, unsigned nk_tiles) {
  const Dims& d = a.d;
  TrOps<UD1, TERRAIN> o(a);
  const unsigned it = lb % ni_tiles, kt = (lb / ni_tiles) % nk_tiles, jt = lb / (ni_tiles * nk_tiles);
  const int i = d.IS - 1 + (int)(it * 64 + threadIdx.x);
  const int k = d.KS + (int)(kt * 4 + threadIdx.y);
  const int j = d.JS - 1 + (int)jt;
  if (i > d.IE + 1 || k > d.KE || j > d.JE + 1) return;
  const size_t p = IX(d, k, i, j);
  a.phi_lo[p] = (a.QTRC0[p] * a.DENS0[p]
                 + a.dtl * (-((o.lZ(k, i, j) - o.lZ(k - 1, i, j)) * a.RCDZ[k]
                              + (o.lX(k, i, j) - o.lX(k, i - 1, j)) * a.RCDX[i]
                              + (o.lY(k, i, j) - o.lY(k, i, j - 1)) * a.RCDY[j])
                            * o.m1(i, j) * o.m2(i, j) / o.gs(I_XYZ, p)))
              / a.DENS[p];
}
template <bool TERRAIN>
__global__ void __launch_bounds__(256) fct_philo(TrArgs a, unsigned ni_tiles, unsigned nk_tiles, unsigned nblocks) {
  fct_philo_body<TERRAIN>(a, xcd_block(blockIdx.x, nblocks), ni_tiles, nk_tiles);
}

// upstream (qa) / downstream (qb) 14-point combinations (COMMON.F90:2096-2155)
__device__ __forceinline__ double fct_qa(const double* f, const double* P, const Dims& d, int k1, int k0, int i, int j) {
  return f[17] * P[IX(d, k1, i + 1, j + 1)] + f[16] * P[IX(d, k0, i + 1, j + 1)] + f[15] * P[IX(d, k1, i, j + 1)]
       + f[14] * P[IX(d, k0, i, j + 1)] + f[13] * P[IX(d, k1, i - 1, j + 1)] + f[11] * P[IX(d, k1, i + 1, j)]
       + f[10] * P[IX(d, k0, i + 1, j)] + f[9] * P[IX(d, k1, i, j)] + f[7] * P[IX(d, k1, i - 1, j)]
       + f[5] * P[IX(d, k1, i + 1, j - 1)] + f[4] * P[IX(d, k0, i + 1, j - 1)] + f[3] * P[IX(d, k1, i, j - 1)]
       + f[1] * P[IX(d, k1, i - 1, j - 1)] + f[8] * P[IX(d, k0, i, j)];
}
__device__ __forceinline__ double fct_qb(const double* f, const double* P, const Dims& d, int k1, int k0, int i, int j) {
  return f[17] * P[IX(d, k1, i - 1, j - 1)] + f[16] * P[IX(d, k0, i - 1, j - 1)] + f[15] * P[IX(d, k1, i, j - 1)]
       + f[14] * P[IX(d, k0, i, j - 1)] + f[13] * P[IX(d, k1, i + 1, j - 1)] + f[11] * P[IX(d, k1, i - 1, j)]
       + f[10] * P[IX(d, k0, i - 1, j)] + f[9] * P[IX(d, k1, i, j)] + f[7] * P[IX(d, k1, i + 1, j)]
       + f[5] * P[IX(d, k1, i - 1, j + 1)] + f[4] * P[IX(d, k0, i - 1, j - 1)] + f[3] * P[IX(d, k1, i, j - 1)]
       + f[1] * P[IX(d, k1, i + 1, j + 1)] + f[8] * P[IX(d, k0, i, j)];
}

// ---- FCT pass 2: R+ and R- ----------------------------------------------------
template <int S, bool TERRAIN, bool VECT>
__device__ __forceinline__ void fct_ratio_body(const TrArgs& a, unsigned lb, unsigned ni_tiles, unsigned nk_tiles) {
  const Dims& d = a.d;
  TrOps<S, TERRAIN> o(a);
  const unsigned it = lb % ni_tiles, kt = (lb / ni_tiles) % nk_tiles, jt = lb / (ni_tiles * nk_tiles);
  const int i = d.IS + (int)(it * 64 + threadIdx.x);
  const int k = d.KS + (int)(kt * 4 + threadIdx.y);
  const int j = d.JS + (int)jt;
  if (i > d.IE || k > d.KE || j > d.JE) return;
  const int KS = d.KS, KE = d.KE;
  const size_t p = IX(d, k, i, j);
  const double dt = a.dtl;
  const double az = o.aZ(k, i, j), azm = o.aZ(k - 1, i, j);
  const double ax = o.aX(k, i, j), axm = o.aX(k, i - 1, j);
  const double ay = o.aY(k, i, j), aym = o.aY(k, i, j - 1);
  const double mm = o.m1(i, j) * o.m2(i, j);
  const double g = o.gs(I_XYZ, p);
  const double pjpls = dt * ((fmax2(0.0, azm) - fmin2(0.0, az)) * a.RCDZ[k]
                           + (fmax2(0.0, axm) - fmin2(0.0, ax)) * a.RCDX[i]
                           + (fmax2(0.0, aym) - fmin2(0.0, ay)) * a.RCDY[j])
                     * o.m1(i, j) * o.m2(i, j) / g;
  const double pjmns = dt * ((fmax2(0.0, az) - fmin2(0.0, azm)) * a.RCDZ[k]
                           + (fmax2(0.0, ax) - fmin2(0.0, axm)) * a.RCDX[i]
                           + (fmax2(0.0, ay) - fmin2(0.0, aym)) * a.RCDY[j])
                     * o.m1(i, j) * o.m2(i, j) / g;
  (void)mm;
  const double* PI = a.QTRC0;
  const double* PL = a.phi_lo;
  const double pin = PI[p], plo = PL[p];
  double qmax, qmin;
  if constexpr (VECT) {
    double rw;
    int ka, kb;
    if (k == KS) { rw = (o.mZ(p)) * a.RCDZ[KS]; ka = KS + 1; kb = KS; }
    else if (k == KE) { rw = (o.mZ(IX(d, KE - 1, i, j))) * a.RCDZ[KE]; ka = KE; kb = KE - 1; }
    else { rw = (o.mZ(p) + o.mZ(IX(d, k - 1, i, j))) * a.RCDZ[k]; ka = k + 1; kb = k - 1; }
    const double ru = (o.mX(p) + o.mX(IX(d, k, i - 1, j))) * a.RCDX[i];
    const double rv = (o.mY(p) + o.mY(IX(d, k, i, j - 1))) * a.RCDY[j];
    double f[18];
    get_fact(f, rw, ru, rv);
    const double qa_in = fct_qa(f, PI, d, ka, k, i, j);
    const double qb_in = fct_qb(f, PI, d, kb, k, i, j);
    const double qa_lo = fct_qa(f, PL, d, ka, k, i, j);
    const double qb_lo = fct_qb(f, PL, d, kb, k, i, j);
    qmax = fmax2(fmax2(fmax2(fmax2(fmax2(pin, qa_in), qb_in), plo), qa_lo), qb_lo);
    qmin = fmin2(fmin2(fmin2(fmin2(fmin2(pin, qa_in), qb_in), plo), qa_lo), qb_lo);
  } else {  // 7-point bounds (COMMON.F90:2312-2515)
    const int kp = k == KE ? k : k + 1, km = k == KS ? k : k - 1;
    qmax = pin; qmin = pin;
    const double* arrs[2] = {PI, PL};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const double* P = arrs[s];
      const double v[7] = {P[p], P[IX(d, kp, i, j)], P[IX(d, km, i, j)], P[IX(d, k, i + 1, j)], P[IX(d, k, i - 1, j)],
                           P[IX(d, k, i, j + 1)], P[IX(d, k, i, j - 1)]};
#pragma unroll
      for (int q = 0; q < 7; ++q) { qmax = fmax2(qmax, v[q]); qmin = fmin2(qmin, v[q]); }
    }
  }
  const double qjpls = (qmax - plo) * a.DENS[p];
  const double qjmns = (plo - qmin) * a.DENS[p];
  double zerosw = 0.5 - fsign(0.5, pjpls - EPS_RP);
  a.rjpls[p] = fmin2(1.0, qjpls * (1.0 - zerosw) / (pjpls - zerosw));
  zerosw = 0.5 - fsign(0.5, pjmns - EPS_RP);
  a.rjmns[p] = fmin2(1.0, qjmns * (1.0 - zerosw) / (pjmns - zerosw));
}
template <int S, bool TERRAIN, bool VECT>
__global__ void __launch_bounds__(256) fct_ratio(TrArgs a, unsigned ni_tiles, unsigned nk_tiles, unsigned nblocks) {
  fct_ratio_body<S, TERRAIN, VECT>(a, xcd_block(blockIdx.x, nblocks), ni_tiles, nk_tiles);
}

// ---- FCT pass 3: limited anti-diffusive fluxes and the update ------------------
// THETA: the HEVE theta correction RHOT_RK += dtrk div(anti) (HEVE.F90:1459-1474)
// instead of the tracer update
template <int S, bool TERRAIN, bool THETA>
__device__ __forceinline__ void fct_update_body(const TrArgs& a, unsigned lb, unsigned ni_tiles, unsigned nk_tiles) {
  const Dims& d = a.d;
  TrOps<S, TERRAIN> o(a);
  const unsigned it = lb % ni_tiles, kt = (lb / ni_tiles) % nk_tiles, jt = lb / (ni_tiles * nk_tiles);
  const int i = d.IS + (int)(it * 64 + threadIdx.x);
  const int k = d.KS + (int)(kt * 4 + threadIdx.y);
  const int j = d.JS + (int)jt;
  if (i > d.IE || k > d.KE || j > d.JE) return;
  const size_t p = IX(d, k, i, j);
  const double* RP = a.rjpls;
  const double* RM = a.rjmns;
  auto limZ = [&](int kk) -> double {  // face kk between levels kk and kk+1
    if (kk == d.KS - 1 || kk == d.KE) return 0.0;
    const double an = o.aZ(kk, i, j);
    const double dirsw = 0.5 + fsign(0.5, an);
    return an * (1.0 - fmin2(RP[IX(d, kk + 1, i, j)], RM[IX(d, kk, i, j)]) * (dirsw)
                     - fmin2(RP[IX(d, kk, i, j)], RM[IX(d, kk + 1, i, j)]) * (1.0 - dirsw));
  };
  auto limX = [&](int ii) -> double {
    const double an = o.aX(k, ii, j);
    const double dirsw = 0.5 + fsign(0.5, an);
    return an * (1.0 - fmin2(RP[IX(d, k, ii + 1, j)], RM[IX(d, k, ii, j)]) * (dirsw)
                     - fmin2(RP[IX(d, k, ii, j)], RM[IX(d, k, ii + 1, j)]) * (1.0 - dirsw));
  };
  auto limY = [&](int jj) -> double {
    const double an = o.aY(k, i, jj);
    const double dirsw = 0.5 + fsign(0.5, an);
    return an * (1.0 - fmin2(RP[IX(d, k, i, jj + 1)], RM[IX(d, k, i, jj)]) * (dirsw)
                     - fmin2(RP[IX(d, k, i, jj)], RM[IX(d, k, i, jj + 1)]) * (1.0 - dirsw));
  };
  if constexpr (THETA) {
    a.QTRCo[p] = a.QTRCo[p]
               + a.dtl * ((limZ(k) - limZ(k - 1)) * a.RCDZ[k]
                          + (limX(i) - limX(i - 1)) * a.RCDX[i]
                          + (limY(j) - limY(j - 1)) * a.RCDY[j])
                 * o.m1(i, j) * o.m2(i, j) / o.gs(I_XYZ, p);
    return;
  }
  a.QTRCo[p] = (a.QTRC0[p] * a.DENS0[p]
                + a.dtl * (-((o.hZ(k, i, j) - limZ(k) - o.hZ(k - 1, i, j) + limZ(k - 1)) * a.RCDZ[k]
                             + (o.hX(k, i, j) - limX(i) - o.hX(k, i - 1, j) + limX(i - 1)) * a.RCDX[i]
                             + (o.hY(k, i, j) - limY(j) - o.hY(k, i, j - 1) + limY(j - 1)) * a.RCDY[j])
                           * o.m1(i, j) * o.m2(i, j) / o.gs(I_XYZ, p)
                           + o.rqv(p)))
             / a.DENS[p];
}
template <int S, bool TERRAIN, bool THETA>
__global__ void __launch_bounds__(256) fct_update(TrArgs a, unsigned ni_tiles, unsigned nk_tiles, unsigned nblocks) {
  fct_update_body<S, TERRAIN, THETA>(a, xcd_block(blockIdx.x, nblocks), ni_tiles, nk_tiles);
}

// ---- ensembles: one launch of a pass over the tiles of many members --------
// tab[m] is member m's TrArgs, first[m] its first logical tile, nk[m] its
// tile counts (ni in the low 16 bits, nk in the high ones); flat terrain
enum { TRB_NOFCT = 0, TRB_PHILO = 1, TRB_RATIO = 2, TRB_UPDATE = 3 };
template <int KIND, int S, bool VECT>
__global__ void __launch_bounds__(256) tracer_batch(const TrArgs* __restrict__ tab, const unsigned* __restrict__ first,
                                                    const unsigned* __restrict__ nk, int nmem, unsigned total) {
  const unsigned lt = xcd_block(blockIdx.x, total);
  const int m = batch_member(first, nmem, lt);
  const unsigned lb = lt - first[m], ni_t = nk[m] & 0xffffu, nk_t = nk[m] >> 16;
  if constexpr (KIND == TRB_NOFCT) tracer_nofct_body<S, false>(tab[m], lb, ni_t, nk_t);
  else if constexpr (KIND == TRB_PHILO) fct_philo_body<false>(tab[m], lb, ni_t, nk_t);
  else if constexpr (KIND == TRB_RATIO) fct_ratio_body<S, false, VECT>(tab[m], lb, ni_t, nk_t);
  else fct_update_body<S, false, false>(tab[m], lb, ni_t, nk_t);
}

// kinds: the stage without FCT, or the FCT passes (phi_lo on its ring tiles,
// then R+/R- and the update on the interior tiles, which share one table)
hipError_t tracer_batch_build(const Ctx* const* cs, const TracerIO* ios, int n, bool philo, BatchBuf& b) {
  if (n <= 0) return hipErrorInvalidValue;
  std::vector<TrArgs> args((size_t)n);
  std::vector<unsigned> first((size_t)n + 1, 0), nk((size_t)n, 0);
  const int scheme = cs[0]->scheme_tracer, along = ios[0].flag_along;
  for (int m = 0; m < n; ++m) {
    const Ctx& c = *cs[m];
    const TracerIO& io = ios[m];
    if (c.terrain || c.scheme_tracer != scheme || io.flag_along != along || c.scheme_tracer == UD5 ||
        c.scheme_tracer == CD6)
      return hipErrorNotSupported;
    TrArgs& a = args[(size_t)m];
    a.d = c.d;
    a.QTRC = io.QTRC; a.QTRC0 = io.QTRC0; a.RHOQ_t = io.RHOQ_t; a.DENS0 = io.DENS0; a.DENS = io.DENS;
    a.mflx = io.mflx_hi; a.nd = io.num_diff; a.nds = c.d.n3; a.GSQRT = c.GSQRT; a.MAPF_XY = c.MAPF;
    a.RCDZ = c.g1.RCDZ; a.RCDX = c.g1.RCDX; a.RCDY = c.g1.RCDY;
    a.dtl = io.dtl; a.QTRCo = io.QTRCo;
    a.phi_lo = c.scr[0]; a.rjpls = c.scr[1]; a.rjmns = c.scr[2];
    const Dims& d = c.d;
    const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1, kmax = d.KE - d.KS + 1;
    const unsigned ni_t = philo ? (imax + 2 + 63) / 64 : (imax + 63) / 64, nk_t = (kmax + 3) / 4;
    const unsigned nj_t = philo ? jmax + 2 : jmax;
    if (ni_t > 0xffffu || nk_t > 0xffffu) return hipErrorNotSupported;
    nk[(size_t)m] = ni_t | (nk_t << 16);
    first[(size_t)m + 1] = first[(size_t)m] + ni_t * nk_t * nj_t;
  }
  const hipError_t e = batch_upload(args, first, nk, b);
  b.scheme = scheme; b.tj = along;
  return e;
}

template <int KIND, int S>
static hipError_t tracer_batch_t(const BatchBuf& b, hipStream_t s) {
  const TrArgs* tab;
  const unsigned *first, *nk;
  batch_tables(b, tab, first, nk);
  if (b.tj) hipLaunchKernelGGL((tracer_batch<KIND, S, true>), dim3(b.total), dim3(64, 4), 0, s, tab, first, nk, b.nmem, b.total);
  else hipLaunchKernelGGL((tracer_batch<KIND, S, false>), dim3(b.total), dim3(64, 4), 0, s, tab, first, nk, b.nmem, b.total);
  return hipGetLastError();
}
template <int KIND>
static hipError_t tracer_batch_k(const BatchBuf& b, hipStream_t s) {
  switch (b.scheme) {
    case UD3: return tracer_batch_t<KIND, UD3>(b, s);
    case UD3KOREN: return tracer_batch_t<KIND, UD3KOREN>(b, s);
    case CD2: return tracer_batch_t<KIND, CD2>(b, s);
    case UD1: return tracer_batch_t<KIND, UD1>(b, s);
    default: return tracer_batch_t<KIND, CD4>(b, s);
  }
}
hipError_t tracer_batch_launch(const BatchBuf& b, int kind, hipStream_t s) {
  if (!b.dev) return hipErrorInvalidValue;
  switch (kind) {
    case TRB_NOFCT: return tracer_batch_k<TRB_NOFCT>(b, s);
    case TRB_PHILO: return tracer_batch_k<TRB_PHILO>(b, s);
    case TRB_RATIO: return tracer_batch_k<TRB_RATIO>(b, s);
    default: return tracer_batch_k<TRB_UPDATE>(b, s);
  }
}

// ---------------------------------------------------------------------------
template <int S, bool TERRAIN, bool THETA = false>
static hipError_t tracer_t(const Ctx& c, const TrArgs& a, bool fct, bool along, hipStream_t s) {
  const Dims& d = c.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1, kmax = d.KE - d.KS + 1;
  const unsigned ni = (imax + 63) / 64, nk = (kmax + 3) / 4;
  const unsigned nb = ni * nk * (unsigned)jmax;
  if (!fct) {
    prof_start(c, "tracer_nofct", s);
    hipLaunchKernelGGL((tracer_nofct<S, TERRAIN>), dim3(nb), dim3(64, 4), 0, s, a, ni, nk, nb);
    prof_stop(c, s);
    return hipGetLastError();
  }
  const unsigned ni2 = (imax + 2 + 63) / 64;
  const unsigned nb2 = ni2 * nk * (unsigned)(jmax + 2);
  prof_start(c, "fct_philo", s);
  hipLaunchKernelGGL((fct_philo<TERRAIN>), dim3(nb2), dim3(64, 4), 0, s, a, ni2, nk, nb2);
  prof_stop(c, s);
  prof_start(c, "fct_ratio", s);
  if (along) hipLaunchKernelGGL((fct_ratio<S, TERRAIN, true>), dim3(nb), dim3(64, 4), 0, s, a, ni, nk, nb);
  else hipLaunchKernelGGL((fct_ratio<S, TERRAIN, false>), dim3(nb), dim3(64, 4), 0, s, a, ni, nk, nb);
  prof_stop(c, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  double* r[2] = {a.rjpls, a.rjmns};
  e = halo_fill(c, r, 2, s);
  if (e != hipSuccess) return e;
  prof_start(c, "fct_update", s);
  hipLaunchKernelGGL((fct_update<S, TERRAIN, THETA>), dim3(nb), dim3(64, 4), 0, s, a, ni, nk, nb);
  prof_stop(c, s);
  return hipGetLastError();
}

// caller-supplied metrics and work arrays: always the general-metric kernels
// (with GSQRT = MAPF = 1 their products and quotients by 1.0 are exact, so
// the results equal the flat kernels' bit for bit)
hipError_t tracer_stage_ext(const Ctx& c, const TracerIO& io, const double* GSQRT, const double* MAPF_XY,
                            const double* RCDZ, const double* RCDX, const double* RCDY, double* work3, hipStream_t s) {
  TrArgs a;
  a.d = c.d;
  a.QTRC = io.QTRC; a.QTRC0 = io.QTRC0; a.RHOQ_t = io.RHOQ_t; a.DENS0 = io.DENS0; a.DENS = io.DENS;
  a.mflx = io.mflx_hi; a.nd = io.num_diff; a.nds = c.d.n3; a.GSQRT = GSQRT; a.MAPF_XY = MAPF_XY;
  a.RCDZ = RCDZ; a.RCDX = RCDX; a.RCDY = RCDY;
  a.dtl = io.dtl; a.QTRCo = io.QTRCo;
  a.phi_lo = work3; a.rjpls = work3 + c.d.n3; a.rjmns = work3 + 2 * c.d.n3;
  const bool fct = io.flag_fct != 0, along = io.flag_along != 0;
  switch (c.scheme_tracer) {
    case UD3: return tracer_t<UD3, true>(c, a, fct, along, s);
    case UD3KOREN: return tracer_t<UD3KOREN, true>(c, a, fct, along, s);
    case CD2: return tracer_t<CD2, true>(c, a, fct, along, s);
    case UD1: return tracer_t<UD1, true>(c, a, fct, along, s);
    case UD5: return tracer_t<UD5, true>(c, a, fct, along, s);
    case CD6: return tracer_t<CD6, true>(c, a, fct, along, s);
    default: return tracer_t<CD4, true>(c, a, fct, along, s);
  }
}

hipError_t tracer_stage(const Ctx& c, const TracerIO& io, hipStream_t s) {
  // flat terrain with a two-cell scheme and a grid with blocks enough: the
  // LDS-tiled kernels (tracer_tile.hip), with FCT or without
  if (io.flag_fct && c.tracer_tile && tracer_tile_supported(c)) return tracer_stage_tile(c, io, c.scr[1], c.scr[2], s);
  if (!io.flag_fct && c.tracer_tile && tracer_tile_supported(c)) return tracer_nofct_tile(c, io, s);
  TrArgs a;
  a.d = c.d;
  a.QTRC = io.QTRC; a.QTRC0 = io.QTRC0; a.RHOQ_t = io.RHOQ_t; a.DENS0 = io.DENS0; a.DENS = io.DENS;
  a.mflx = io.mflx_hi; a.nd = io.num_diff; a.nds = c.d.n3; a.GSQRT = c.GSQRT; a.MAPF_XY = c.MAPF;
  a.RCDZ = c.g1.RCDZ; a.RCDX = c.g1.RCDX; a.RCDY = c.g1.RCDY;
  a.dtl = io.dtl; a.QTRCo = io.QTRCo;
  a.phi_lo = c.scr[0]; a.rjpls = c.scr[1]; a.rjmns = c.scr[2];
  const bool fct = io.flag_fct != 0, along = io.flag_along != 0;
  if (c.scheme_tracer == UD3) return c.terrain ? tracer_t<UD3, true>(c, a, fct, along, s) : tracer_t<UD3, false>(c, a, fct, along, s);
  if (c.scheme_tracer == UD3KOREN)
    return c.terrain ? tracer_t<UD3KOREN, true>(c, a, fct, along, s) : tracer_t<UD3KOREN, false>(c, a, fct, along, s);
  if (c.scheme_tracer == CD2) return c.terrain ? tracer_t<CD2, true>(c, a, fct, along, s) : tracer_t<CD2, false>(c, a, fct, along, s);
  if (c.scheme_tracer == UD1) return c.terrain ? tracer_t<UD1, true>(c, a, fct, along, s) : tracer_t<UD1, false>(c, a, fct, along, s);
  if (c.scheme_tracer == UD5) return c.terrain ? tracer_t<UD5, true>(c, a, fct, along, s) : tracer_t<UD5, false>(c, a, fct, along, s);
  if (c.scheme_tracer == CD6) return c.terrain ? tracer_t<CD6, true>(c, a, fct, along, s) : tracer_t<CD6, false>(c, a, fct, along, s);
  return c.terrain ? tracer_t<CD4, true>(c, a, fct, along, s) : tracer_t<CD4, false>(c, a, fct, along, s);
}

// ---------------------------------------------------------------------------
// FCT of the HEVE theta flux (HEVE.F90:1388-1475).  The limiter is the
// tracer's with theta as the advected field: the high-order fluxes are the
// stage's (scheme of the dynamics on the current theta with num_diff(RHOT)),
// the low-order ones UD1 on the start-of-stage theta, the densities DENS0 and
// DENS_RK, and the update corrects RHOT_RK.  POTT keeps the reference's
// composition: the current theta on the whole halo, overwritten by the
// start-of-stage theta on IS-1..IE+2, JS-1..JE+2 (HEVE.F90:360-372, 1403-1409).
__global__ void __launch_bounds__(256) heve_pott(Dims d, const double* RHOT, const double* DENS, const double* RHOT0,
                                                 const double* DENS0, double* POTTc, double* POTT0) {
  const int i = d.IS - 2 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = d.KS + (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = d.JS - 2 + (int)blockIdx.z;
  if (i > d.IE + 2 || k > d.KE || j > d.JE + 2) return;
  const size_t p = IX(d, k, i, j);
  const double pc = RHOT[p] / DENS[p];
  POTTc[p] = pc;
  POTT0[p] = (i >= d.IS - 1 && j >= d.JS - 1) ? RHOT0[p] / DENS0[p] : pc;
}

hipError_t heve_fct_theta(const StageIO& io, hipStream_t s) {
  const Ctx& c = *io.ctx;
  const Dims& d = c.d;
  double* mf[3] = {io.mflx_hi, io.mflx_hi + d.n3, io.mflx_hi + 2 * d.n3};
  hipError_t e = halo_fill(c, mf, 3, s);
  if (e != hipSuccess) return e;
  double* dk[1] = {io.o[0]};
  e = halo_fill(c, dk, 1, s);
  if (e != hipSuccess) return e;
  TrArgs a;
  a.d = d;
  double* POTTc = c.scr[3];
  double* POTT0 = c.scr[4];
  {
    const int ni = d.IE - d.IS + 5, nk = d.KE - d.KS + 1, nj = d.JE - d.JS + 5;
    dim3 grid((ni + 63) / 64, (nk + 3) / 4, nj);
    hipLaunchKernelGGL(heve_pott, grid, dim3(64, 4), 0, s, d, io.s[4], io.s[0], io.s0[4], io.s0[0], POTTc, POTT0);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  a.QTRC = POTTc; a.QTRC0 = POTT0; a.RHOQ_t = nullptr; a.DENS0 = io.s0[0]; a.DENS = io.o[0];
  a.mflx = io.mflx_hi; a.nd = io.num_diff + d.n3 * I_RHOT; a.nds = 5 * d.n3;
  a.GSQRT = c.GSQRT; a.MAPF_XY = c.MAPF;
  a.RCDZ = c.g1.RCDZ; a.RCDX = c.g1.RCDX; a.RCDY = c.g1.RCDY;
  a.dtl = io.dtrk; a.QTRCo = io.o[4];
  a.phi_lo = c.scr[0]; a.rjpls = c.scr[1]; a.rjmns = c.scr[2];
  const bool along = io.fct_along != 0;
  if (c.scheme == UD3)
    return c.terrain ? tracer_t<UD3, true, true>(c, a, true, along, s) : tracer_t<UD3, false, true>(c, a, true, along, s);
  return c.terrain ? tracer_t<CD4, true, true>(c, a, true, along, s) : tracer_t<CD4, false, true>(c, a, true, along, s);
}

// ---------------------------------------------------------------------------
// ATMOS_DYN_fct on stored fluxes, for the HEVE momentum FCT
// (HEVE.F90:776-810, 1006-1057, 1259-1309): the advected field is a velocity
// at a staggered point, its densities the averages of DENS0 and DENS_RK with
// the neighbour (k+dk, i+di, j+dj) (at KE the vertical average repeats KE-1,
// HEVE.F90:757-773), and the low-order and anti-diffusive fluxes come from
// arrays.  Same three passes as the tracer FCT.
struct ArrFctArgs {
  Dims d;
  const double *phi, *D0, *D1;
  int dk, di, dj;
  const double *lo, *anti, *mflx;
  const double *RDZ, *RDX, *RDY;
  const double *G, *MAPF;  // GSQRT component and MAPF(:,:,:,c) of the staggered point
  double dt;
  double *phi_lo, *rjpls, *rjmns;
  double* out;
  int kmax_upd;  // the update covers KS..kmax_upd
};

template <bool TERRAIN>
struct AfOps {
  const ArrFctArgs& a;
  const Dims& d;
  __device__ AfOps(const ArrFctArgs& a_) : a(a_), d(a_.d) {}
  __device__ __forceinline__ size_t ix(int k, int i, int j) const { return IX(d, k, i, j); }
  __device__ __forceinline__ double gs(size_t p) const { if constexpr (TERRAIN) return a.G[p]; else return 1.0; }
  __device__ __forceinline__ double m1(int i, int j) const {
    if constexpr (TERRAIN) return a.MAPF[(size_t)i + (size_t)d.IA * j]; else return 1.0;
  }
  __device__ __forceinline__ double m2(int i, int j) const {
    if constexpr (TERRAIN) return a.MAPF[(size_t)i + (size_t)d.IA * j + d.pl]; else return 1.0;
  }
  __device__ __forceinline__ double dav(const double* D, int k, int i, int j) const {
    const int kk = (a.dk && k == d.KE) ? d.KE - 1 : k;
    return 0.5 * (D[ix(kk, i, j)] + D[ix(kk + a.dk, i + a.di, j + a.dj)]);
  }
  __device__ __forceinline__ double lZ(size_t p) const { return a.lo[p]; }
  __device__ __forceinline__ double lX(size_t p) const { return a.lo[p + d.n3]; }
  __device__ __forceinline__ double lY(size_t p) const { return a.lo[p + 2 * d.n3]; }
  __device__ __forceinline__ double aZ(size_t p) const { return a.anti[p]; }
  __device__ __forceinline__ double aX(size_t p) const { return a.anti[p + d.n3]; }
  __device__ __forceinline__ double aY(size_t p) const { return a.anti[p + 2 * d.n3]; }
};

template <bool TERRAIN>
__global__ void __launch_bounds__(256) afct_philo(ArrFctArgs a) {
  const Dims& d = a.d;
  AfOps<TERRAIN> o(a);
  const int i = d.IS - 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = d.KS + (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = d.JS - 1 + (int)blockIdx.z;
  if (i > d.IE + 1 || k > d.KE || j > d.JE + 1) return;
  const size_t p = IX(d, k, i, j);
  a.phi_lo[p] = (a.phi[p] * o.dav(a.D0, k, i, j)
                 + a.dt * (-((o.lZ(p) - o.lZ(IX(d, k - 1, i, j))) * a.RDZ[k]
                             + (o.lX(p) - o.lX(IX(d, k, i - 1, j))) * a.RDX[i]
                             + (o.lY(p) - o.lY(IX(d, k, i, j - 1))) * a.RDY[j])
                           * o.m1(i, j) * o.m2(i, j) / o.gs(p)))
              / o.dav(a.D1, k, i, j);
}

template <bool TERRAIN, bool VECT>
__global__ void __launch_bounds__(256) afct_ratio(ArrFctArgs a) {
  const Dims& d = a.d;
  AfOps<TERRAIN> o(a);
  const int i = d.IS + (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = d.KS + (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = d.JS + (int)blockIdx.z;
  if (i > d.IE || k > d.KE || j > d.JE) return;
  const int KS = d.KS, KE = d.KE;
  const size_t p = IX(d, k, i, j);
  const double dt = a.dt;
  const double az = o.aZ(p), azm = o.aZ(IX(d, k - 1, i, j));
  const double ax = o.aX(p), axm = o.aX(IX(d, k, i - 1, j));
  const double ay = o.aY(p), aym = o.aY(IX(d, k, i, j - 1));
  const double g = o.gs(p);
  const double pjpls = dt * ((fmax2(0.0, azm) - fmin2(0.0, az)) * a.RDZ[k]
                           + (fmax2(0.0, axm) - fmin2(0.0, ax)) * a.RDX[i]
                           + (fmax2(0.0, aym) - fmin2(0.0, ay)) * a.RDY[j])
                     * o.m1(i, j) * o.m2(i, j) / g;
  const double pjmns = dt * ((fmax2(0.0, az) - fmin2(0.0, azm)) * a.RDZ[k]
                           + (fmax2(0.0, ax) - fmin2(0.0, axm)) * a.RDX[i]
                           + (fmax2(0.0, ay) - fmin2(0.0, aym)) * a.RDY[j])
                     * o.m1(i, j) * o.m2(i, j) / g;
  const double* PI = a.phi;
  const double* PL = a.phi_lo;
  const double pin = PI[p], plo = PL[p];
  double qmax, qmin;
  if constexpr (VECT) {
    const double* mZ = a.mflx; const double* mX = a.mflx + d.n3; const double* mY = a.mflx + 2 * d.n3;
    double rw;
    int ka, kb;
    if (k == KS) { rw = (mZ[p]) * a.RDZ[KS]; ka = KS + 1; kb = KS; }
    else if (k == KE) { rw = (mZ[IX(d, KE - 1, i, j)]) * a.RDZ[KE]; ka = KE; kb = KE - 1; }
    else { rw = (mZ[p] + mZ[IX(d, k - 1, i, j)]) * a.RDZ[k]; ka = k + 1; kb = k - 1; }
    const double ru = (mX[p] + mX[IX(d, k, i - 1, j)]) * a.RDX[i];
    const double rv = (mY[p] + mY[IX(d, k, i, j - 1)]) * a.RDY[j];
    double f[18];
    get_fact(f, rw, ru, rv);
    const double qa_in = fct_qa(f, PI, d, ka, k, i, j);
    const double qb_in = fct_qb(f, PI, d, kb, k, i, j);
    const double qa_lo = fct_qa(f, PL, d, ka, k, i, j);
    const double qb_lo = fct_qb(f, PL, d, kb, k, i, j);
    qmax = fmax2(fmax2(fmax2(fmax2(fmax2(pin, qa_in), qb_in), plo), qa_lo), qb_lo);
    qmin = fmin2(fmin2(fmin2(fmin2(fmin2(pin, qa_in), qb_in), plo), qa_lo), qb_lo);
  } else {
    const int kp = k == KE ? k : k + 1, km = k == KS ? k : k - 1;
    qmax = pin; qmin = pin;
    const double* arrs[2] = {PI, PL};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const double* P = arrs[s];
      const double v[7] = {P[p], P[IX(d, kp, i, j)], P[IX(d, km, i, j)], P[IX(d, k, i + 1, j)], P[IX(d, k, i - 1, j)],
                           P[IX(d, k, i, j + 1)], P[IX(d, k, i, j - 1)]};
#pragma unroll
      for (int q = 0; q < 7; ++q) { qmax = fmax2(qmax, v[q]); qmin = fmin2(qmin, v[q]); }
    }
  }
  const double D = o.dav(a.D1, k, i, j);
  const double qjpls = (qmax - plo) * D;
  const double qjmns = (plo - qmin) * D;
  double zerosw = 0.5 - fsign(0.5, pjpls - EPS_RP);
  a.rjpls[p] = fmin2(1.0, qjpls * (1.0 - zerosw) / (pjpls - zerosw));
  zerosw = 0.5 - fsign(0.5, pjmns - EPS_RP);
  a.rjmns[p] = fmin2(1.0, qjmns * (1.0 - zerosw) / (pjmns - zerosw));
}

// the limited anti-diffusive fluxes and the momentum update: WFORM is the
// MOMZ form dt ((aZ) RFDZ + ((aX) RCDX + (aY) RCDY) M1 M2) / G (HEVE.F90:795-803),
// otherwise dt ((aZ) RCDZ + (aX) RFDX + (aY) RCDY) M1 M2 / G (:1035-1042, 1284-1291)
template <bool TERRAIN, bool WFORM>
__global__ void __launch_bounds__(256) afct_update(ArrFctArgs a) {
  const Dims& d = a.d;
  AfOps<TERRAIN> o(a);
  const int i = d.IS + (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = d.KS + (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = d.JS + (int)blockIdx.z;
  if (i > d.IE || k > a.kmax_upd || j > d.JE) return;
  const size_t p = IX(d, k, i, j);
  const double* RP = a.rjpls;
  const double* RM = a.rjmns;
  auto limZ = [&](int kk) -> double {
    if (kk == d.KS - 1 || kk == d.KE) return 0.0;
    const double an = o.aZ(IX(d, kk, i, j));
    const double dirsw = 0.5 + fsign(0.5, an);
    return an * (1.0 - fmin2(RP[IX(d, kk + 1, i, j)], RM[IX(d, kk, i, j)]) * (dirsw)
                     - fmin2(RP[IX(d, kk, i, j)], RM[IX(d, kk + 1, i, j)]) * (1.0 - dirsw));
  };
  auto limX = [&](int ii) -> double {
    const double an = o.aX(IX(d, k, ii, j));
    const double dirsw = 0.5 + fsign(0.5, an);
    return an * (1.0 - fmin2(RP[IX(d, k, ii + 1, j)], RM[IX(d, k, ii, j)]) * (dirsw)
                     - fmin2(RP[IX(d, k, ii, j)], RM[IX(d, k, ii + 1, j)]) * (1.0 - dirsw));
  };
  auto limY = [&](int jj) -> double {
    const double an = o.aY(IX(d, k, i, jj));
    const double dirsw = 0.5 + fsign(0.5, an);
    return an * (1.0 - fmin2(RP[IX(d, k, i, jj + 1)], RM[IX(d, k, i, jj)]) * (dirsw)
                     - fmin2(RP[IX(d, k, i, jj)], RM[IX(d, k, i, jj + 1)]) * (1.0 - dirsw));
  };
  if constexpr (WFORM) {
    a.out[p] = a.out[p]
             + a.dt * ((limZ(k) - limZ(k - 1)) * a.RDZ[k]
                       + ((limX(i) - limX(i - 1)) * a.RDX[i] + (limY(j) - limY(j - 1)) * a.RDY[j])
                         * o.m1(i, j) * o.m2(i, j))
               / o.gs(p);
  } else {
    a.out[p] = a.out[p]
             + a.dt * (((limZ(k) - limZ(k - 1)) * a.RDZ[k] + (limX(i) - limX(i - 1)) * a.RDX[i]
                        + (limY(j) - limY(j - 1)) * a.RDY[j]))
               * o.m1(i, j) * o.m2(i, j) / o.gs(p);
  }
}

template <bool TERRAIN>
static hipError_t afct_t(const Ctx& c, const ArrFctArgs& a, bool along, bool wform, hipStream_t s) {
  const Dims& d = c.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1, kmax = d.KE - d.KS + 1;
  const dim3 blk(64, 4);
  const dim3 g2((imax + 2 + 63) / 64, (kmax + 3) / 4, jmax + 2), g1((imax + 63) / 64, (kmax + 3) / 4, jmax);
  prof_start(c, "mfct_limit", s);
  hipLaunchKernelGGL(afct_philo<TERRAIN>, g2, blk, 0, s, a);
  if (along) hipLaunchKernelGGL((afct_ratio<TERRAIN, true>), g1, blk, 0, s, a);
  else hipLaunchKernelGGL((afct_ratio<TERRAIN, false>), g1, blk, 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  double* r[2] = {a.rjpls, a.rjmns};
  e = halo_fill(c, r, 2, s);
  if (e != hipSuccess) return e;
  if (wform) hipLaunchKernelGGL((afct_update<TERRAIN, true>), g1, blk, 0, s, a);
  else hipLaunchKernelGGL((afct_update<TERRAIN, false>), g1, blk, 0, s, a);
  prof_stop(c, s);
  return hipGetLastError();
}

hipError_t fct_momentum_limit(const Ctx& c, const MomFct& f, hipStream_t s) {
  ArrFctArgs a;
  a.d = c.d;
  a.phi = f.vel; a.D0 = f.D0; a.D1 = f.D1;
  a.dk = f.dk; a.di = f.di; a.dj = f.dj;
  a.lo = f.lo; a.anti = f.anti; a.mflx = f.mflx;
  a.RDZ = f.RDZ; a.RDX = f.RDX; a.RDY = f.RDY;
  a.G = f.G; a.MAPF = f.MAPF;
  a.dt = f.dt;
  a.phi_lo = f.phi_lo; a.rjpls = f.rjpls; a.rjmns = f.rjmns;
  a.out = f.out;
  a.kmax_upd = f.dk ? c.d.KE - 1 : c.d.KE;
  const bool wform = f.dk != 0;
  return c.terrain ? afct_t<true>(c, a, f.along, wform, s) : afct_t<false>(c, a, f.along, wform, s);
}


// ---------------------------------------------------------------------------
// Adjoint of one tracer stage with the FCT limiter frozen.
//
// With the mass fluxes and densities held fixed, the stage is linear in
// QTRC, QTRC0, RHOQ_t and num_diff once the limiter's ratios R+/R- (and with
// them the factor W of each anti-diffusive flux, limited = W (hi - lo)) keep
// their forward values: the face flux of the update is (1 - W) hi + W lo
// (COMMON.F90:2570-2659, tracer_fvm_heve.F90:197-205).  The adjoint runs the
// forward's first two FCT passes for R+/R-, then
//   tradj_faces   the adjoint of each face flux, split into its high-order
//                 part (1 - W) F_b and low-order part W F_b, and the
//                 adjoint of num_diff (GSQRT times the high-order part)
//   tradj_gather  for each point, the sum over the faces whose stencil reads
//                 it of (face adjoint) x vel x (the face form's coefficient
//                 of that point), plus the direct terms of the update
// The coefficient of a point in a face form is the form evaluated on the
// unit vector at that point: the forms are linear in the values for a fixed
// velocity (UD3KOREN1993's limiter is not, and is rejected).
struct TrAdjArgs {
  TrArgs f;            // the forward's arguments; f.rjpls/f.rjmns hold R+/R-
  const double* Qo_b;  // adjoint of QTRCo, read at the interior
  double *hb[3], *lb[3];                 // face adjoints (z, x, y): high- and low-order parts
  double *Q_b, *Q0_b, *RQ_b, *ND_b;      // adjoints of QTRC, QTRC0, RHOQ_t, num_diff (3 planes)
};

// z face kk between levels kk and kk+1: the coefficient of level k in the
// high-order face form (the branches of TrOps::hZ)
template <int S>
__device__ __forceinline__ double tradj_zcoef(const Dims& d, int kk, int k, double vel) {
  auto e = [&](int m) { return m == k ? 1.0 : 0.0; };
  if constexpr (S == UD5 || S == CD6) {
    double v[5];
    auto col = [&](int k0, int n) { for (int m = 0; m < n; ++m) v[m] = e(k0 + m); };
    if (kk == d.KE - 2) { col(d.KE - 4, 5); return fv6_ke2<S>(v, vel); }
    if (kk == d.KS + 1) { col(d.KS, 5); return fv6_ks1<S>(v, vel); }
    if (kk == d.KE - 1) { col(d.KE - 2, 3); return fv6_ke1<S>(v, vel); }
    if (kk == d.KS) { col(d.KS, 3); return fv6_ks<S>(v, vel); }
    return fv6_int<S>(e(kk - 2), e(kk - 1), e(kk), e(kk + 1), e(kk + 2), e(kk + 3), vel);
  } else {
    if (kk == d.KS) return fv_lo<S>(e(kk), e(kk + 1), e(kk + 2), vel);
    if (kk == d.KE - 1) return fv_hi<S>(e(kk - 1), e(kk), e(kk + 1), vel);
    return fv_int<S>(e(kk - 1), e(kk), e(kk + 1), e(kk + 2), vel);
  }
}
// horizontal face ii between points ii and ii+1: the coefficient of point i
template <int S>
__device__ __forceinline__ double tradj_hcoef(int ii, int i, double vel) {
  auto e = [&](int m) { return m == i ? 1.0 : 0.0; };
  if constexpr (S == UD1) return fv_int<UD1>(0.0, e(ii), e(ii + 1), 0.0, vel);
  else if constexpr (S == UD5 || S == CD6) return fv6_int<S>(e(ii - 2), e(ii - 1), e(ii), e(ii + 1), e(ii + 2), e(ii + 3), vel);
  else return fv_int<S>(e(ii - 1), e(ii), e(ii + 1), e(ii + 2), vel);
}
// the UD1 (low-order) face: coefficient of point i in face ii
__device__ __forceinline__ double tradj_lcoef(int ii, int i, double vel) {
  return fv_int<UD1>(0.0, ii == i ? 1.0 : 0.0, ii + 1 == i ? 1.0 : 0.0, 0.0, vel);
}

template <int S, bool TERRAIN, bool FCT>
__global__ void __launch_bounds__(256) tradj_faces(TrAdjArgs b) {
  const TrArgs& a = b.f;
  const Dims& d = a.d;
  TrOps<S, TERRAIN> o(a);
  const int i = (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = (int)blockIdx.z;
  if (i >= d.IA || k >= d.KA || j >= d.JA) return;
  const size_t p = IX(d, k, i, j), n3 = d.n3;
  const double* RP = a.rjpls;
  const double* RM = a.rjmns;
  // dt m1 m2 / (GSQRT DENS) QTRCo_b at an interior point, zero elsewhere: the
  // adjoint of the flux difference there is -+ this times RCDZ/RCDX/RCDY
  auto A = [&](int kk, int ii, int jj) -> double {
    if (kk < d.KS || kk > d.KE || ii < d.IS || ii > d.IE || jj < d.JS || jj > d.JE) return 0.0;
    const size_t q = IX(d, kk, ii, jj);
    return a.dtl * b.Qo_b[q] * o.m1(ii, jj) * o.m2(ii, jj) / (o.gs(I_XYZ, q) * a.DENS[q]);
  };
  // the frozen factor W of a face's limited anti-diffusive flux W (hi - lo)
  auto W = [&](double an, size_t pl, size_t pu) -> double {
    const double dirsw = 0.5 + fsign(0.5, an);
    return 1.0 - fmin2(RP[pu], RM[pl]) * dirsw - fmin2(RP[pl], RM[pu]) * (1.0 - dirsw);
  };
  const bool in_i = i >= d.IS && i <= d.IE, in_j = j >= d.JS && j <= d.JE, in_k = k >= d.KS && k <= d.KE;
  double h[3] = {0.0, 0.0, 0.0}, l[3] = {0.0, 0.0, 0.0};
  const bool v[3] = {in_i && in_j && k >= d.KS && k <= d.KE - 1, in_k && in_j && i >= d.IS - 1 && i <= d.IE,
                     in_k && in_i && j >= d.JS - 1 && j <= d.JE};
  // z face k (KS..KE-1)
  if (v[0]) {
    const double fb = -A(k, i, j) * a.RCDZ[k] + A(k + 1, i, j) * a.RCDZ[k + 1];
    double w = 0.0;
    if constexpr (FCT) w = W(o.aZ(k, i, j), p, IX(d, k + 1, i, j));
    h[0] = (1.0 - w) * fb; l[0] = w * fb;
  }
  // x face i (IS-1..IE)
  if (v[1]) {
    const double fb = -A(k, i, j) * a.RCDX[i] + A(k, i + 1, j) * a.RCDX[i + 1];
    double w = 0.0;
    if constexpr (FCT) w = W(o.aX(k, i, j), p, IX(d, k, i + 1, j));
    h[1] = (1.0 - w) * fb; l[1] = w * fb;
  }
  // y face j (JS-1..JE)
  if (v[2]) {
    const double fb = -A(k, i, j) * a.RCDY[j] + A(k, i, j + 1) * a.RCDY[j + 1];
    double w = 0.0;
    if constexpr (FCT) w = W(o.aY(k, i, j), p, IX(d, k, i, j + 1));
    h[2] = (1.0 - w) * fb; l[2] = w * fb;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    b.hb[c][p] = h[c];
    if constexpr (FCT) b.lb[c][p] = l[c];
  }
  // num_diff enters the high-order fluxes as + GSQRT num_diff (not for UD1)
  if constexpr (S == UD1) {
    b.ND_b[p] = 0.0; b.ND_b[p + n3] = 0.0; b.ND_b[p + 2 * n3] = 0.0;
  } else {
    b.ND_b[p] = v[0] ? o.gs(I_XYW, p) * h[0] : 0.0;
    b.ND_b[p + n3] = v[1] ? o.gs(I_UYZ, p) * h[1] : 0.0;
    b.ND_b[p + 2 * n3] = v[2] ? o.gs(I_XVZ, p) * h[2] : 0.0;
  }
}

template <int S, bool TERRAIN, bool FCT>
__global__ void __launch_bounds__(256) tradj_gather(TrAdjArgs b) {
  const TrArgs& a = b.f;
  const Dims& d = a.d;
  TrOps<S, TERRAIN> o(a);
  const int i = (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = (int)blockIdx.z;
  if (i >= d.IA || k >= d.KA || j >= d.JA) return;
  const size_t p = IX(d, k, i, j);
  const bool in_i = i >= d.IS && i <= d.IE, in_j = j >= d.JS && j <= d.JE, in_k = k >= d.KS && k <= d.KE;
  // the faces whose high-order form can read point m: m-3 .. m+2
  constexpr int R0 = (S == UD5 || S == CD6) ? 3 : (S == UD1 ? 1 : 2);
  constexpr int R1 = (S == UD5 || S == CD6) ? 2 : (S == UD1 ? 0 : 1);
  double qb = 0.0, q0b = 0.0;
  if (in_i && in_j) {
    for (int kk = max(k - R0, d.KS); kk <= min(k + R1, d.KE - 1); ++kk) {
      const size_t f = IX(d, kk, i, j);
      const double vel = o.mZ(f);
      qb += b.hb[0][f] * vel * tradj_zcoef<S>(d, kk, k, vel);
      if constexpr (FCT) if (kk >= k - 1) q0b += b.lb[0][f] * vel * tradj_lcoef(kk, k, vel);
    }
  }
  if (in_k && in_j) {
    for (int ii = max(i - R0, d.IS - 1); ii <= min(i + R1, d.IE); ++ii) {
      const size_t f = IX(d, k, ii, j);
      const double vel = o.mX(f);
      qb += b.hb[1][f] * vel * tradj_hcoef<S>(ii, i, vel);
      if constexpr (FCT) if (ii >= i - 1) q0b += b.lb[1][f] * vel * tradj_lcoef(ii, i, vel);
    }
  }
  if (in_k && in_i) {
    for (int jj = max(j - R0, d.JS - 1); jj <= min(j + R1, d.JE); ++jj) {
      const size_t f = IX(d, k, i, jj);
      const double vel = o.mY(f);
      qb += b.hb[2][f] * vel * tradj_hcoef<S>(jj, j, vel);
      if constexpr (FCT) if (jj >= j - 1) q0b += b.lb[2][f] * vel * tradj_lcoef(jj, j, vel);
    }
  }
  double rqb = 0.0;
  if (in_i && in_j && in_k) {
    const double ob = b.Qo_b[p] / a.DENS[p];
    q0b += ob * a.DENS0[p];
    rqb = a.dtl * ob;
  }
  b.Q_b[p] = qb;
  b.Q0_b[p] = q0b;
  b.RQ_b[p] = rqb;
}

template <int S, bool TERRAIN>
static hipError_t tracer_adj_t(const Ctx& c, TrAdjArgs& b, bool fct, bool along, hipStream_t s) {
  const Dims& d = c.d;
  const TrArgs& a = b.f;
  if (fct) {  // R+/R- of the forward (its first two passes and the halo fill)
    const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1, kmax = d.KE - d.KS + 1;
    const unsigned ni = (imax + 63) / 64, nk = (kmax + 3) / 4;
    const unsigned nb = ni * nk * (unsigned)jmax;
    const unsigned ni2 = (imax + 2 + 63) / 64;
    const unsigned nb2 = ni2 * nk * (unsigned)(jmax + 2);
    hipLaunchKernelGGL((fct_philo<TERRAIN>), dim3(nb2), dim3(64, 4), 0, s, a, ni2, nk, nb2);
    if (along) hipLaunchKernelGGL((fct_ratio<S, TERRAIN, true>), dim3(nb), dim3(64, 4), 0, s, a, ni, nk, nb);
    else hipLaunchKernelGGL((fct_ratio<S, TERRAIN, false>), dim3(nb), dim3(64, 4), 0, s, a, ni, nk, nb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    double* r[2] = {a.rjpls, a.rjmns};
    e = halo_fill(c, r, 2, s);
    if (e != hipSuccess) return e;
  }
  const dim3 g((d.IA + 63) / 64, (d.KA + 3) / 4, d.JA), blk(64, 4);
  prof_start(c, "tracer_adj", s);
  if (fct) {
    hipLaunchKernelGGL((tradj_faces<S, TERRAIN, true>), g, blk, 0, s, b);
    hipLaunchKernelGGL((tradj_gather<S, TERRAIN, true>), g, blk, 0, s, b);
  } else {
    hipLaunchKernelGGL((tradj_faces<S, TERRAIN, false>), g, blk, 0, s, b);
    hipLaunchKernelGGL((tradj_gather<S, TERRAIN, false>), g, blk, 0, s, b);
  }
  prof_stop(c, s);
  return hipGetLastError();
}

hipError_t tracer_stage_adj(const Ctx& c, const TracerIO& io, const TracerAdjIO& adj, hipStream_t s) {
  if (c.scheme_tracer == UD3KOREN) return hipErrorInvalidValue;
  TrAdjArgs b;
  TrArgs& a = b.f;
  a.d = c.d;
  a.QTRC = io.QTRC; a.QTRC0 = io.QTRC0; a.RHOQ_t = nullptr; a.DENS0 = io.DENS0; a.DENS = io.DENS;
  a.mflx = io.mflx_hi; a.nd = io.num_diff; a.nds = c.d.n3; a.GSQRT = c.GSQRT; a.MAPF_XY = c.MAPF;
  a.RCDZ = c.g1.RCDZ; a.RCDX = c.g1.RCDX; a.RCDY = c.g1.RCDY;
  a.dtl = io.dtl; a.QTRCo = nullptr;
  // work: phi_lo, R+, R- in scr[0..2]; the face adjoints in scr[3..5] and
  // (phi_lo being dead by then) scr[0], scr[6], scr[7]
  a.phi_lo = c.scr[0]; a.rjpls = c.scr[1]; a.rjmns = c.scr[2];
  b.hb[0] = c.scr[3]; b.hb[1] = c.scr[4]; b.hb[2] = c.scr[5];
  b.lb[0] = c.scr[0]; b.lb[1] = c.scr[6]; b.lb[2] = c.scr[7];
  b.Qo_b = adj.QTRCo_b; b.Q_b = adj.QTRC_b; b.Q0_b = adj.QTRC0_b; b.RQ_b = adj.RHOQ_t_b; b.ND_b = adj.num_diff_b;
  const bool fct = io.flag_fct != 0, along = io.flag_along != 0;
  switch (c.scheme_tracer) {
    case UD3: return c.terrain ? tracer_adj_t<UD3, true>(c, b, fct, along, s) : tracer_adj_t<UD3, false>(c, b, fct, along, s);
    case CD2: return c.terrain ? tracer_adj_t<CD2, true>(c, b, fct, along, s) : tracer_adj_t<CD2, false>(c, b, fct, along, s);
    case UD1: return c.terrain ? tracer_adj_t<UD1, true>(c, b, fct, along, s) : tracer_adj_t<UD1, false>(c, b, fct, along, s);
    case UD5: return c.terrain ? tracer_adj_t<UD5, true>(c, b, fct, along, s) : tracer_adj_t<UD5, false>(c, b, fct, along, s);
    case CD6: return c.terrain ? tracer_adj_t<CD6, true>(c, b, fct, along, s) : tracer_adj_t<CD6, false>(c, b, fct, along, s);
    default: return c.terrain ? tracer_adj_t<CD4, true>(c, b, fct, along, s) : tracer_adj_t<CD4, false>(c, b, fct, along, s);
  }
}

}  // namespace shp
