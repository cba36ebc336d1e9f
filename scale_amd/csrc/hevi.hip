// hevi.hip -- one RK stage of the HEVI finite-volume short step on gfx950.
//
// Reference: ATMOS_DYN_Tstep_short_fvm_hevi,
//   scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short_fvm_hevi.F90:128-1089
// with the flux stencils of scale_atmos_dyn_fvm_flux_{ud3,cd4}.F90.
//
// One kernel per stage.  Both kernels give each thread one column (i,j) and
// march k upward through the explicit part of every point (pressure, mass and
// theta fluxes, the explicit z fluxes, the tendencies Sr/St/Sw, the face value
// PT, the complete MOMX/MOMY updates; HEVI.F90:346-657, 816-1047) and, one
// level behind, the tridiagonal rows and the Thomas forward sweep
// (HEVI.F90:688-727, 1092-1115); marching back down they substitute into
// MOMZ/DENS/RHOT and add the implicit parts of the z fluxes
// (HEVI.F90:728-804, 1116-1127).
//   hevi_tile   flat terrain (GSQRT = 1, J13G = J23G = 0, MAPF = 1): blocks
//               of 64 x 8 columns stage each level in LDS and evaluate every
//               x and y face once;
//   hevi_march  general metrics: neighbours come through the caches and each
//               point evaluates the faces it needs.
#include <cstdlib>
#include <vector>
#include "common.h"
#include "shp_internal.h"
#include "batch.h"

namespace shp {

// Arguments of a stage.  R is the type of the prognostic, tendency, flux and
// num_diff arrays: double, or float for the mixed-precision path, which keeps
// the reference state, the pressure terms (DPRES0, RT2P, REF_*) and the
// vertical solve in double; the solve's coefficients e, f and the face
// value PT pass from the upward to the downward sweep through scratch in R.
template <typename R>
struct HeviArgsT {
  const Ctx* ctx;  // host side only
  Dims d;
  Grid1D g;
  const double *GSQRT, *J13G, *J23G, *MAPF, *PHI;
  double J33G;
  // inputs (device layout)
  const R *DENS0, *MOMZ0, *MOMX0, *MOMY0, *RHOT0;
  const R *DENS, *MOMZ, *MOMX, *MOMY, *RHOT;
  const R *DENS_t, *MOMZ_t, *MOMX_t, *MOMY_t, *RHOT_t;
  const double *DPRES0, *RT2P, *CORIOLI;
  const R* num_diff;
  const double* wdamp_coef;
  const R* DDIV;
  const double *REF_dens, *REF_rhot;
  double divdmp_coef, dtrk;
  int has_coriolis;
  // outputs
  R *DENS_RK, *MOMZ_RK, *MOMX_RK, *MOMY_RK, *RHOT_RK, *mflx_hi, *tflx_hi;
  // scratch (n3 each): in FP32 the Thomas coefficients are rounded to float
  // between the sweeps (PT is a float value already), like the prognostics
  R *PT, *e, *f;
  // Sw between the two launches of the split terrain stage (FP64 only)
  double* Sw = nullptr;
  // optional records for the adjoint: St and the solved C
  double *trSt = nullptr, *trCp = nullptr;
  // mixed precision: DENS and RHOT (and their stage values) are stored as
  // deviations from the level profiles offD(k), offR(k), kept in double
  const double *offD = nullptr, *offR = nullptr;
  // RK4 combination fused into the last stage (hevi_tile with CMB)
  const R *c1[5], *c2[5], *c3[5];
  const R *cm[3], *ct[3];
  R* st[5];
  R *mflx_av, *tflx_av;
};
using HeviArgs = HeviArgsT<double>;

// RK4 combination (tinteg_short_rk4.F90:560-660): (2 r1 + 4 r2 + 2 r3 + r4 - 3 s0) / 6
template <typename R>
__device__ __forceinline__ R rk4c(R r1, R r2, R r3, R r4, R s0) {
  return (r1 * R(2) + r2 * R(4) + r3 * R(2) + r4 - s0 * R(3)) / R(6);
}
// (f1 + 2 f2 + 2 f3 + f4) / 6
template <typename R>
__device__ __forceinline__ R rk4f(R f1, R f2, R f3, R f4) {
  return (f1 + f2 * R(2) + f3 * R(2) + f4) / R(6);
}

template <int S, bool TERRAIN, bool DIVDMP>
struct Ops {
  const HeviArgs& a;
  const Dims& d;
  Met<TERRAIN> m;
  __device__ Ops(const HeviArgs& a_) : a(a_), d(a_.d) {
    m.G = a.GSQRT; m.J13 = a.J13G; m.J23 = a.J23G; m.MAPF = a.MAPF; m.n3 = d.n3; m.pl = d.pl; m.IA = d.IA;
  }
  __device__ __forceinline__ size_t ix(int k, int i, int j) const { return IX(d, k, i, j); }
  __device__ __forceinline__ double ND(int v, int dir, size_t p) const { return a.num_diff[p + d.n3 * (size_t)(v + 5 * dir)]; }
  __device__ __forceinline__ double D(int k, int i, int j) const { return a.DENS[ix(k, i, j)]; }
  __device__ __forceinline__ double POTT(int k, int i, int j) const { size_t p = ix(k, i, j); return a.RHOT[p] / a.DENS[p]; }
  // DPRES (HEVI.F90:355-358)
  __device__ __forceinline__ double DPRES(int k, int i, int j) const {
    if (k == d.KS - 1) return a.DPRES0[ix(k, i, j)] - a.DENS[ix(d.KS, i, j)] * (a.PHI[ix(d.KS - 1, i, j)] - a.PHI[ix(d.KS + 1, i, j)]);
    if (k == d.KE + 1) return a.DPRES0[ix(k, i, j)] - a.DENS[ix(d.KE, i, j)] * (a.PHI[ix(d.KE + 1, i, j)] - a.PHI[ix(d.KE - 1, i, j)]);
    size_t p = ix(k, i, j);
    return a.DPRES0[p] + a.RT2P[p] * (a.RHOT[p] - a.REF_rhot[p]);
  }
  // F2H(k,p,c) (HEVI.F90:62 / UD3.F90:75) with GSQRT component c at (i,j)
  __device__ __forceinline__ double F2H(int k, int pp, int c, int i, int j) const {
    const double* CDZ = a.g.CDZ;
    return CDZ[k + pp - 1] * m.gs(c, ix(k + pp - 1, i, j)) / (CDZ[k] * m.gs(c, ix(k, i, j)) + CDZ[k + 1] * m.gs(c, ix(k + 1, i, j)));
  }

  // ---- continuity (HEVI.F90:369-438) ----
  __device__ double mflxX(int k, int i, int j) const {
    size_t p = ix(k, i, j);
    return m.gs(I_UYZ, p) / m.mapf(i, j, 1, I_UY) * (a.MOMX[p] + ND(I_DENS, XDIR, p));
  }
  __device__ double mflxY(int k, int i, int j) const {
    size_t p = ix(k, i, j);
    return m.gs(I_XVZ, p) / m.mapf(i, j, 0, I_XV) * (a.MOMY[p] + ND(I_DENS, YDIR, p));
  }
  __device__ double mflxZ(int k, int i, int j) const {  // explicit part, k in KS..KE-1
    size_t p = ix(k, i, j);
    double t = 0.0;
    if constexpr (TERRAIN) {
      const double* X = a.MOMX; const double* Y = a.MOMY;
      t = m.j13(I_XYW, p) * 0.25 / m.mapf(i, j, 1, I_XY)
            * (X[ix(k + 1, i, j)] + X[ix(k + 1, i - 1, j)] + X[ix(k, i, j)] + X[ix(k, i - 1, j)])
        + m.j23(I_XYW, p) * 0.25 / m.mapf(i, j, 0, I_XY)
            * (Y[ix(k + 1, i, j)] + Y[ix(k + 1, i, j - 1)] + Y[ix(k, i, j)] + Y[ix(k, i, j - 1)])
        + m.gs(I_XYW, p) / (m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY)) * ND(I_DENS, ZDIR, p);
      return t;
    } else {
      return m.gs(I_XYW, p) / (m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY)) * ND(I_DENS, ZDIR, p);
    }
  }
  __device__ double mflxZf(int k, int i, int j) const {  // with the zero end faces
    if (k == d.KS - 1 || k == d.KE) return 0.0;
    return mflxZ(k, i, j);
  }

  // ---- theta fluxes: fluxZ/X/Y_XYZ on POTT (UD3.F90:157-354) ----
  __device__ double tflxZ(int k, int i, int j) const {  // explicit part, face k
    if (k == d.KS - 1 || k == d.KE) return 0.0;
    size_t p = ix(k, i, j);
    double vel = mflxZ(k, i, j);
    double fv;
    if (k == d.KS) fv = fv_lo<S>(POTT(k, i, j), POTT(k + 1, i, j), POTT(k + 2, i, j), vel);
    else if (k == d.KE - 1) fv = fv_hi<S>(POTT(k - 1, i, j), POTT(k, i, j), POTT(k + 1, i, j), vel);
    else fv = fv_int<S>(POTT(k - 1, i, j), POTT(k, i, j), POTT(k + 1, i, j), POTT(k + 2, i, j), vel);
    return vel * fv + m.gs(I_XYW, p) * ND(I_RHOT, ZDIR, p);
  }
  // the same with the column's theta window w = POTT(k-1), POTT(k), POTT(k+1), POTT(k+2)
  __device__ double tflxZw(int k, int i, int j, double vel, const double* w) const {
    if (k == d.KS - 1 || k == d.KE) return 0.0;
    size_t p = ix(k, i, j);
    double fv;
    if (k == d.KS) fv = fv_lo<S>(w[1], w[2], w[3], vel);
    else if (k == d.KE - 1) fv = fv_hi<S>(w[0], w[1], w[2], vel);
    else fv = fv_int<S>(w[0], w[1], w[2], w[3], vel);
    return vel * fv + m.gs(I_XYW, p) * ND(I_RHOT, ZDIR, p);
  }
  __device__ double PTfw(int k, int i, int j, const double* w) const {
    const double vel = a.MOMZ[ix(k, i, j)];
    if (k == d.KS) return fv_lo<S>(w[1], w[2], w[3], vel);
    if (k == d.KE - 1) return fv_hi<S>(w[0], w[1], w[2], vel);
    return fv_int<S>(w[0], w[1], w[2], w[3], vel);
  }
  __device__ double tflxX(int k, int i, int j, double vel) const {
    size_t p = ix(k, i, j);
    return vel * fv_int<S>(POTT(k, i - 1, j), POTT(k, i, j), POTT(k, i + 1, j), POTT(k, i + 2, j), vel)
         + m.gs(I_UYZ, p) * ND(I_RHOT, XDIR, p);
  }
  __device__ double tflxY(int k, int i, int j, double vel) const {
    size_t p = ix(k, i, j);
    return vel * fv_int<S>(POTT(k, i, j - 1), POTT(k, i, j), POTT(k, i, j + 1), POTT(k, i, j + 2), vel)
         + m.gs(I_XVZ, p) * ND(I_RHOT, YDIR, p);
  }

  // ---- momentum z fluxes at stored index kk (UD3.F90:358-748) ----
  __device__ double qZ_XYW(int kk, int i, int j) const {
    if (kk == d.KS - 1 || kk >= d.KE - 1) return 0.0;
    const double* W = a.MOMZ;
    const int k = kk + 1;  // loop index of the reference
    double vel = (0.5 * (W[ix(k - 1, i, j)] + W[ix(k, i, j)])) / D(k, i, j);
    double fv;
    if (kk == d.KS) fv = fv_lo<S>(W[ix(kk, i, j)], W[ix(kk + 1, i, j)], W[ix(kk + 2, i, j)], vel);
    else fv = fv_int<S>(W[ix(kk - 1, i, j)], W[ix(kk, i, j)], W[ix(kk + 1, i, j)], W[ix(kk + 2, i, j)], vel);
    size_t pk = ix(k, i, j);
    return a.J33G * vel * fv + m.gs(I_XYZ, pk) * ND(I_MOMZ, ZDIR, pk);
  }
  // J13 (dir 0, MOMX) / J23 (dir 1, MOMY) fluxes at XYW (UD3.F90:457-593), TERRAIN only
  __device__ double qJ_XYW(int dir, int kk, int i, int j) const {
    if (kk == d.KS - 1 || kk >= d.KE - 1) return 0.0;
    const double* W = a.MOMZ;
    const double* M = dir == 0 ? a.MOMX : a.MOMY;
    const int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
    const double mp = m.mapf(i, j, dir == 0 ? 1 : 0, I_XY);
    double vel;
    if (kk == d.KS) {
      vel = ((0.5 * (M[ix(kk + 1, i, j)] + M[ix(kk + 1, i - di, j - dj)])) / D(kk + 1, i, j)
           + (0.5 * (M[ix(kk, i, j)] + M[ix(kk, i - di, j - dj)])) / D(kk, i, j)) * 0.5;
      vel = vel * (dir == 0 ? m.j13(I_XYZ, ix(kk + 1, i, j)) : m.j23(I_XYZ, ix(kk + 1, i, j)));
      return vel / mp * fv_lo<S>(W[ix(kk, i, j)], W[ix(kk + 1, i, j)], W[ix(kk + 2, i, j)], vel);
    }
    const int k = kk + 1;
    vel = (0.5 * (M[ix(k, i, j)] + M[ix(k, i - di, j - dj)])) / D(k, i, j);
    vel = vel * (dir == 0 ? m.j13(I_XYZ, ix(k, i, j)) : m.j23(I_XYZ, ix(k, i, j)));
    return vel / mp * fv_int<S>(W[ix(kk - 1, i, j)], W[ix(kk, i, j)], W[ix(kk + 1, i, j)], W[ix(kk + 2, i, j)], vel);
  }
  // X (dir 0) / Y (dir 1) fluxes at XYW, face (i,j), level k <= KE-1 (UD3.F90:598-748)
  __device__ double qH_XYW(int dir, int k, int i, int j) const {
    if (k == d.KE) return 0.0;
    const int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
    const int c = dir == 0 ? I_UYW : I_XVW;
    const double* M = dir == 0 ? a.MOMX : a.MOMY;
    const double* W = a.MOMZ;
    const double f1 = F2H(k, 1, c, i, j), f2 = F2H(k, 2, c, i, j);
    double vel = (f1 * M[ix(k + 1, i, j)] + f2 * M[ix(k, i, j)])
               / (f1 * 0.5 * (D(k + 1, i, j) + D(k + 1, i + di, j + dj)) + f2 * 0.5 * (D(k, i, j) + D(k, i + di, j + dj)));
    size_t p = ix(k, i, j);
    const double mp = dir == 0 ? m.mapf(i, j, 1, I_UY) : m.mapf(i, j, 0, I_XV);
    return m.gs(c, p) / mp * vel
         * fv_int<S>(W[ix(k, i - di, j - dj)], W[p], W[ix(k, i + di, j + dj)], W[ix(k, i + 2 * di, j + 2 * dj)], vel)
         + m.gs(c, p) * ND(I_MOMZ, dir == 0 ? XDIR : YDIR, p);
  }

  // ---- vertical face value of theta for the implicit solve (valueW_Z) ----
  __device__ double PTf(int k, int i, int j) const {
    const double vel = a.MOMZ[ix(k, i, j)];
    if (k == d.KS) return fv_lo<S>(POTT(k, i, j), POTT(k + 1, i, j), POTT(k + 2, i, j), vel);
    if (k == d.KE - 1) return fv_hi<S>(POTT(k - 1, i, j), POTT(k, i, j), POTT(k + 1, i, j), vel);
    return fv_int<S>(POTT(k - 1, i, j), POTT(k, i, j), POTT(k + 1, i, j), POTT(k + 2, i, j), vel);
  }

  // ---- momentum x fluxes (UD3.F90:753-1167) ----
  // z flux at (u,y,w), stored level kk; stag 0: (u,y), stag 1: (x,v)
  __device__ double qZ_UV(int stag, int kk, int i, int j) const {
    if (kk == d.KS - 1 || kk == d.KE) return 0.0;
    const int di = stag == 0 ? 1 : 0, dj = stag == 0 ? 0 : 1;
    const int c = stag == 0 ? I_UYW : I_XVW;
    const double* W = a.MOMZ;
    const double* V = stag == 0 ? a.MOMX : a.MOMY;
    double vel = (0.5 * (W[ix(kk, i, j)] + W[ix(kk, i + di, j + dj)]))
               / (F2H(kk, 1, c, i, j) * 0.5 * (D(kk + 1, i, j) + D(kk + 1, i + di, j + dj))
                + F2H(kk, 2, c, i, j) * 0.5 * (D(kk, i, j) + D(kk, i + di, j + dj)));
    double fv;
    if (kk == d.KS) fv = fv_lo<S>(V[ix(kk, i, j)], V[ix(kk + 1, i, j)], V[ix(kk + 2, i, j)], vel);
    else if (kk == d.KE - 1) fv = fv_hi<S>(V[ix(kk - 1, i, j)], V[ix(kk, i, j)], V[ix(kk + 1, i, j)], vel);
    else fv = fv_int<S>(V[ix(kk - 1, i, j)], V[ix(kk, i, j)], V[ix(kk + 1, i, j)], V[ix(kk + 2, i, j)], vel);
    size_t p = ix(kk, i, j);
    return a.J33G * vel * fv + m.gs(c, p) * ND(stag == 0 ? I_MOMX : I_MOMY, ZDIR, p);
  }
  // J13/J23 fluxes at (u,y,z) or (x,v,z), TERRAIN only (UD3.F90:849-1039, 1269-1459)
  __device__ double qJ_UV(int stag, int jdir, int kk, int i, int j) const {
    if (kk == d.KS - 1 || kk == d.KE) return 0.0;
    const int di = stag == 0 ? 1 : 0, dj = stag == 0 ? 0 : 1;
    const int c = stag == 0 ? I_UYZ : I_XVZ;
    const int cw = stag == 0 ? I_UYW : I_XVW;
    const bool quad = (stag == 0 && jdir == 1) || (stag == 1 && jdir == 0);
    const int qi = stag == 0 ? 1 : -1, qj = stag == 0 ? -1 : 1;
    const double* M = jdir == 0 ? a.MOMX : a.MOMY;
    const double* V = stag == 0 ? a.MOMX : a.MOMY;
    const double mp = m.mapf(i, j, jdir == 0 ? 1 : 0, stag == 0 ? I_UY : I_XV);
    auto fmom = [&](double fac, int kx) -> double {
      if (quad) return fac * 0.25 * (M[ix(kx, i, j)] + M[ix(kx, i + qi, j)] + M[ix(kx, i, j + qj)] + M[ix(kx, i + qi, j + qj)]);
      return fac * M[ix(kx, i, j)];
    };
    const double f1 = F2H(kk, 1, c, i, j), f2 = F2H(kk, 2, c, i, j);
    double vel = (fmom(f1, kk + 1) + fmom(f2, kk))
               / (f1 * 0.5 * (D(kk + 1, i, j) + D(kk + 1, i + di, j + dj)) + f2 * 0.5 * (D(kk, i, j) + D(kk, i + di, j + dj)));
    vel = vel * (jdir == 0 ? m.j13(cw, ix(kk, i, j)) : m.j23(cw, ix(kk, i, j)));
    double fv;
    if (kk == d.KS) fv = fv_lo<S>(V[ix(kk, i, j)], V[ix(kk + 1, i, j)], V[ix(kk + 2, i, j)], vel);
    else if (kk == d.KE - 1) fv = fv_hi<S>(V[ix(kk - 1, i, j)], V[ix(kk, i, j)], V[ix(kk + 1, i, j)], vel);
    else fv = fv_int<S>(V[ix(kk - 1, i, j)], V[ix(kk, i, j)], V[ix(kk + 1, i, j)], V[ix(kk + 2, i, j)], vel);
    return vel / mp * fv;
  }
  // fluxX_UYZ stored at (k,ii,j): loop index ii+1 (UD3.F90:1043-1097)
  __device__ double qX_UYZ(int k, int ii, int j) const {
    const int i = ii + 1;
    const double* M = a.MOMX;
    double vel = (0.5 * (M[ix(k, i, j)] + M[ix(k, i - 1, j)])) / (D(k, i, j));
    size_t p = ix(k, i, j);
    return m.gs(I_XYZ, p) / m.mapf(i, j, 1, I_XY) * vel
         * fv_int<S>(M[ix(k, i - 2, j)], M[ix(k, i - 1, j)], M[p], M[ix(k, i + 1, j)], vel)
         + m.gs(I_XYZ, p) * ND(I_MOMX, XDIR, p);
  }
  // fluxY_UYZ at (k,i,jj) (UD3.F90:1101-1150)
  __device__ double qY_UYZ(int k, int i, int j) const {
    const double* Y = a.MOMY; const double* X = a.MOMX;
    double vel = (0.5 * (Y[ix(k, i, j)] + Y[ix(k, i + 1, j)]))
               / (0.25 * (D(k, i, j) + D(k, i + 1, j) + D(k, i, j + 1) + D(k, i + 1, j + 1)));
    size_t p = ix(k, i, j);
    return m.gs(I_UVZ, p) / m.mapf(i, j, 0, I_UV) * vel
         * fv_int<S>(X[ix(k, i, j - 1)], X[p], X[ix(k, i, j + 1)], X[ix(k, i, j + 2)], vel)
         + m.gs(I_UVZ, p) * ND(I_MOMX, YDIR, p);
  }
  // fluxX_XVZ at (k,ii,j) (UD3.F90:1463-1512)
  __device__ double qX_XVZ(int k, int i, int j) const {
    const double* Y = a.MOMY; const double* X = a.MOMX;
    double vel = (0.5 * (X[ix(k, i, j)] + X[ix(k, i, j + 1)]))
               / (0.25 * (D(k, i, j) + D(k, i + 1, j) + D(k, i, j + 1) + D(k, i + 1, j + 1)));
    size_t p = ix(k, i, j);
    return m.gs(I_UVZ, p) / m.mapf(i, j, 1, I_UV) * vel
         * fv_int<S>(Y[ix(k, i - 1, j)], Y[p], Y[ix(k, i + 1, j)], Y[ix(k, i + 2, j)], vel)
         + m.gs(I_UVZ, p) * ND(I_MOMY, XDIR, p);
  }
  // fluxY_XVZ stored at (k,i,jj): loop index jj+1 (UD3.F90:1536-1587)
  __device__ double qY_XVZ(int k, int i, int jj) const {
    const int j = jj + 1;
    const double* Y = a.MOMY;
    double vel = (0.5 * (Y[ix(k, i, j)] + Y[ix(k, i, j - 1)])) / (D(k, i, j));
    size_t p = ix(k, i, j);
    return m.gs(I_XYZ, p) / m.mapf(i, j, 0, I_XY) * vel
         * fv_int<S>(Y[ix(k, i, j - 2)], Y[ix(k, i, j - 1)], Y[p], Y[ix(k, i, j + 1)], vel)
         + m.gs(I_XYZ, p) * ND(I_MOMY, YDIR, p);
  }
};

// ---------------------------------------------------------------------------
// Explicit part of one point (k,i,j): mass and theta fluxes in x and y, the
// explicit part of the z fluxes, the tendencies Sr/St/Sw, the face value PT
// of the vertical solve, and the complete MOMX/MOMY updates
// (HEVI.F90:346-657, 816-1047).  Fluxes needed on both faces of a cell are
// evaluated with the same expression by every point that uses them, so
// neighbours see identical values.
struct PointTend {
  double Sr, St, Sw, PT;                                  // tendencies and PT for the solve
  double mX, mXm, mY, mYm, mZ, mZw, tX, tXm, tY, tYm, tZ;  // fluxes to store
  double MX, MY;                                          // MOMX_RK, MOMY_RK
};

// Fluxes through the lower face of the current level, carried from the level
// below (where they were the upper face), and the column's theta window
// POTT(k-1..k+2).  Every face flux at KS-1 is zero.
struct ZCarry {
  double mZ = 0.0, tZ = 0.0, qW = 0.0, qWJ13 = 0.0, qWJ23 = 0.0;
  double qU = 0.0, qUJ13 = 0.0, qUJ23 = 0.0, qV = 0.0, qVJ13 = 0.0, qVJ23 = 0.0;
  double pw[4];
};

// MOMX and MOMY tendencies before divergence damping: advection, pressure
// gradient (with its terrain part) and the Coriolis and metric terms, which
// HEVI and HEVE share (HEVI.F90:816-927, 936-1047; HEVE.F90:744-945,
// 993-1193).  Each carries its z fluxes from the level below in zc.
template <int S, bool TERRAIN, bool DIVDMP>
__device__ __forceinline__ double momx_core(const Ops<S, TERRAIN, DIVDMP>& o, const HeviArgs& a, int k, int i, int j,
                                            ZCarry& zc) {
  const Dims& d = a.d;
  const size_t p = IX(d, k, i, j);
  const double* RCDX = a.g.RCDX; const double* RCDY = a.g.RCDY; const double* RCDZ = a.g.RCDZ;
  const double* RFDX = a.g.RFDX; const double* RFDY = a.g.RFDY;
  const auto& m = o.m;
  const double qu = o.qZ_UV(0, k, i, j);
  double sv = qu - zc.qU;
  zc.qU = qu;
  if constexpr (TERRAIN) {
    const double q13 = o.qJ_UV(0, 0, k, i, j), q23 = o.qJ_UV(0, 1, k, i, j);
    sv = sv + q13 - zc.qUJ13 + q23 - zc.qUJ23;
    zc.qUJ13 = q13; zc.qUJ23 = q23;
  }
  double advcv = -(sv) * RCDZ[k];
  double advch = -((o.qX_UYZ(k, i, j) - o.qX_UYZ(k, i - 1, j)) * RFDX[i]
                 + (o.qY_UYZ(k, i, j) - o.qY_UYZ(k, i, j - 1)) * RCDY[j])
               * m.mapf(i, j, 0, I_UY) * m.mapf(i, j, 1, I_UY);
  double pgh = (m.gs(I_XYZ, IX(d, k, i + 1, j)) * o.DPRES(k, i + 1, j) - m.gs(I_XYZ, p) * o.DPRES(k, i, j)) * RFDX[i];
  if constexpr (TERRAIN) {
    pgh = pgh + (m.j13(I_UYW, p)
                   * 0.5 * (o.F2H(k, 1, I_UYZ, i, j) * (o.DPRES(k + 1, i + 1, j) + o.DPRES(k + 1, i, j))
                          + o.F2H(k, 2, I_UYZ, i, j) * (o.DPRES(k, i + 1, j) + o.DPRES(k, i, j)))
                 - m.j13(I_UYW, IX(d, k - 1, i, j))
                   * 0.5 * (o.F2H(k, 1, I_UYZ, i, j) * (o.DPRES(k, i + 1, j) + o.DPRES(k, i, j))
                          + o.F2H(k, 2, I_UYZ, i, j) * (o.DPRES(k - 1, i + 1, j) + o.DPRES(k - 1, i, j))))
                  * RCDZ[k];
  }
  const double pg = pgh * m.mapf(i, j, 0, I_UY);
  double t = (advcv + advch - pg) / m.gs(I_UYZ, p);
  const double* Y = a.MOMY;
  if (a.has_coriolis || TERRAIN) {
    double cf = 0.125 * (a.CORIOLI[(size_t)(i + 1) + (size_t)d.IA * j] + a.CORIOLI[(size_t)i + (size_t)d.IA * j])
              * (Y[IX(d, k, i + 1, j)] + Y[p] + Y[IX(d, k, i + 1, j - 1)] + Y[IX(d, k, i, j - 1)]);
    if constexpr (TERRAIN) {
      const double sy = Y[p] + Y[IX(d, k, i, j - 1)] + Y[IX(d, k, i + 1, j)] + Y[IX(d, k, i + 1, j - 1)];
      cf = cf + 0.25 * m.mapf(i, j, 0, I_UY) * m.mapf(i, j, 1, I_UY) * sy
              * (sy * 0.25 * (1.0 / m.mapf(i + 1, j, 1, I_XY) - 1.0 / m.mapf(i, j, 1, I_XY)) * RFDX[i]
                 - a.MOMX[p] * (1.0 / m.mapf(i, j, 0, I_UV) - 1.0 / m.mapf(i, j - 1, 0, I_UV)) * RCDY[j])
              * 2.0 / (o.D(k, i + 1, j) + o.D(k, i, j));
    }
    t = t + cf;
  }
  return t;
}

template <int S, bool TERRAIN, bool DIVDMP>
__device__ __forceinline__ double momy_core(const Ops<S, TERRAIN, DIVDMP>& o, const HeviArgs& a, int k, int i, int j,
                                            ZCarry& zc) {
  const Dims& d = a.d;
  const size_t p = IX(d, k, i, j);
  const double* RCDX = a.g.RCDX; const double* RCDY = a.g.RCDY; const double* RCDZ = a.g.RCDZ;
  const double* RFDX = a.g.RFDX; const double* RFDY = a.g.RFDY;
  const auto& m = o.m;
  const double qv = o.qZ_UV(1, k, i, j);
  double sv = qv - zc.qV;
  zc.qV = qv;
  if constexpr (TERRAIN) {
    const double q13 = o.qJ_UV(1, 0, k, i, j), q23 = o.qJ_UV(1, 1, k, i, j);
    sv = sv + q13 - zc.qVJ13 + q23 - zc.qVJ23;
    zc.qVJ13 = q13; zc.qVJ23 = q23;
  }
  double advcv = -(sv) * RCDZ[k];
  double advch = -((o.qX_XVZ(k, i, j) - o.qX_XVZ(k, i - 1, j)) * RCDX[i]
                 + (o.qY_XVZ(k, i, j) - o.qY_XVZ(k, i, j - 1)) * RFDY[j])
               * m.mapf(i, j, 0, I_XV) * m.mapf(i, j, 1, I_XV);
  double pgh = (m.gs(I_XYZ, IX(d, k, i, j + 1)) * o.DPRES(k, i, j + 1) - m.gs(I_XYZ, p) * o.DPRES(k, i, j)) * RFDY[j];
  if constexpr (TERRAIN) {
    pgh = pgh + (m.j23(I_XVW, p)
                   * 0.5 * (o.F2H(k, 1, I_XVZ, i, j) * (o.DPRES(k + 1, i, j + 1) + o.DPRES(k + 1, i, j))
                          + o.F2H(k, 2, I_XVZ, i, j) * (o.DPRES(k, i, j + 1) + o.DPRES(k, i, j)))
                 - m.j23(I_XVW, IX(d, k - 1, i, j))
                   * 0.5 * (o.F2H(k - 1, 1, I_XVZ, i, j) * (o.DPRES(k, i, j + 1) + o.DPRES(k, i, j))
                          + o.F2H(k - 1, 2, I_XVZ, i, j) * (o.DPRES(k - 1, i, j + 1) + o.DPRES(k - 1, i, j))))
                  * RCDZ[k];
  }
  const double pg = pgh * m.mapf(i, j, 1, I_XV);
  double t = (advcv + advch - pg) / m.gs(I_XVZ, p);
  const double* X = a.MOMX;
  if (a.has_coriolis || TERRAIN) {
    double cf = -0.125 * (a.CORIOLI[(size_t)i + (size_t)d.IA * (j + 1)] + a.CORIOLI[(size_t)i + (size_t)d.IA * j])
              * (X[IX(d, k, i, j + 1)] + X[p] + X[IX(d, k, i - 1, j + 1)] + X[IX(d, k, i - 1, j)]);
    if constexpr (TERRAIN) {
      const double sx = X[p] + X[IX(d, k, i - 1, j)] + X[IX(d, k, i, j + 1)] + X[IX(d, k, i - 1, j + 1)];
      cf = cf - 0.25 * m.mapf(i, j, 0, I_XV) * m.mapf(i, j, 1, I_XV) * sx
              * (a.MOMY[p] * (1.0 / m.mapf(i, j, 1, I_UV) - 1.0 / m.mapf(i - 1, j, 1, I_UV)) * RCDX[i]
                 - 0.25 * sx * (1.0 / m.mapf(i, j + 1, 0, I_XY) - 1.0 / m.mapf(i, j, 0, I_XY)) * RFDY[j])
              * 2.0 / (o.D(k, i, j + 1) + o.D(k, i, j));
    }
    t = t + cf;
  }
  return t;
}

// PART selects the sections evaluated (P_ALL for the whole point): the
// continuity and theta fluxes with Sr and St (P_CT), the explicit part of
// MOMZ with PT (P_W), MOMX (P_U) and MOMY (P_V); each section carries its
// own z fluxes in zc.
enum { P_CT = 1, P_W = 2, P_U = 4, P_V = 8, P_ALL = 15 };
template <int S, bool TERRAIN, bool DIVDMP, int PART = P_ALL>
__device__ __forceinline__ PointTend explicit_point(const Ops<S, TERRAIN, DIVDMP>& o, const HeviArgs& a, int k, int i, int j,
                                                    ZCarry& zc) {
  const Dims& d = a.d;
  PointTend r;
  r.Sw = 0.0; r.PT = 0.0;
  const size_t p = IX(d, k, i, j);
  const double* RCDX = a.g.RCDX; const double* RCDY = a.g.RCDY; const double* RCDZ = a.g.RCDZ;
  const double* RFDX = a.g.RFDX; const double* RFDY = a.g.RFDY; const double* RFDZ = a.g.RFDZ;
  const auto& m = o.m;

  // ---- continuity (HEVI.F90:369-475) ----
  if constexpr ((PART & P_CT) != 0) {
    const double mX = o.mflxX(k, i, j), mXm = o.mflxX(k, i - 1, j);
    const double mY = o.mflxY(k, i, j), mYm = o.mflxY(k, i, j - 1);
    r.mX = mX; r.mXm = mXm; r.mY = mY; r.mYm = mYm;
    const double mZ = o.mflxZf(k, i, j), mZm = zc.mZ;
    zc.mZ = mZ;
    r.mZ = mZ;
    r.mZw = (i == d.IS && k <= d.KE - 1) ? o.mflxZ(k, i - 1, j) : 0.0;
    {
      double advcv = -(mZ - mZm) * RCDZ[k];
      double advch = -((mX - mXm) * RCDX[i] + (mY - mYm) * RCDY[j]);
      r.Sr = (advcv + advch) * m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY) / m.gs(I_XYZ, p) + a.DENS_t[p];
    }

    // ---- thermodynamics (HEVI.F90:588-657) ----
    {
      const double tX = o.tflxX(k, i, j, mX), tXm = o.tflxX(k, i - 1, j, mXm);
      const double tY = o.tflxY(k, i, j, mY), tYm = o.tflxY(k, i, j - 1, mYm);
      const double tZ = o.tflxZw(k, i, j, mZ, zc.pw), tZm = zc.tZ;
      zc.tZ = tZ;
      r.tX = tX; r.tXm = tXm; r.tY = tY; r.tYm = tYm; r.tZ = tZ;
      double advcv = -(tZ - tZm) * RCDZ[k];
      double advch = -((tX - tXm) * RCDX[i] + (tY - tYm) * RCDY[j]);
      r.St = (advcv + advch) * m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY) / m.gs(I_XYZ, p) + a.RHOT_t[p];
    }
  }

  // ---- momentum z, explicit part (HEVI.F90:486-577) and PT (HEVI.F90:691) ----
  if ((PART & P_W) != 0 && k <= d.KE - 1) {
    const double qw = o.qZ_XYW(k, i, j);
    double sv = qw - zc.qW;
    zc.qW = qw;
    if constexpr (TERRAIN) {
      const double q13 = o.qJ_XYW(0, k, i, j), q23 = o.qJ_XYW(1, k, i, j);
      sv = sv + q13 - zc.qWJ13 + q23 - zc.qWJ23;
      zc.qWJ13 = q13; zc.qWJ23 = q23;
    }
    double advcv = -(sv) * RFDZ[k];
    double advch = -((o.qH_XYW(0, k, i, j) - o.qH_XYW(0, k, i - 1, j)) * RCDX[i]
                   + (o.qH_XYW(1, k, i, j) - o.qH_XYW(1, k, i, j - 1)) * RCDY[j])
                 * m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY);
    double wdmp = -a.wdamp_coef[k] * a.MOMZ0[p];
    double sw = (advcv + advch) / m.gs(I_XYW, p) + wdmp;
    if constexpr (DIVDMP) sw = sw + a.divdmp_coef / a.dtrk * (a.DDIV[IX(d, k + 1, i, j)] - a.DDIV[p]) * a.g.FDZ[k];
    r.Sw = sw + a.MOMZ_t[p];
    r.PT = o.PTfw(k, i, j, zc.pw);
  }

  // ---- momentum x (HEVI.F90:816-927), iee = min(IIE,IEH) = IE for periodic x ----
  if constexpr ((PART & P_U) != 0) {
    double t = momx_core(o, a, k, i, j, zc);
    if constexpr (DIVDMP) {
      t = t + a.divdmp_coef / a.dtrk * (a.DDIV[IX(d, k, i + 1, j)] / m.mapf(i + 1, j, 1, I_XY) - a.DDIV[p] / m.mapf(i, j, 0, I_XY))
                * m.mapf(i, j, 0, I_UY) * m.mapf(i, j, 1, I_UY) * a.g.FDX[i];
    }
    r.MX = a.MOMX0[p] + a.dtrk * (t + a.MOMX_t[p]);
  }

  // ---- momentum y (HEVI.F90:936-1047) ----
  if constexpr ((PART & P_V) != 0) {
    double t = momy_core(o, a, k, i, j, zc);
    if constexpr (DIVDMP) {
      t = t + a.divdmp_coef / a.dtrk * (a.DDIV[IX(d, k, i, j + 1)] / m.mapf(i, j + 1, 0, I_XY) - a.DDIV[p] / m.mapf(i, j, 0, I_XY))
                * m.mapf(i, j, 0, I_XV) * m.mapf(i, j, 1, I_XV) * a.g.FDY[j];
    }
    r.MY = a.MOMY0[p] + a.dtrk * (t + a.MOMY_t[p]);
  }
  return r;
}

// The stores of one point, issued after all of its loads: the mass and
// theta fluxes (with the extra faces i = IS-1, j = JS-1 and the zero z faces
// KS-1 and KE, HEVI.F90:370-438, 596-622) and MOMX_RK/MOMY_RK.
template <int PART = P_ALL>
__device__ __forceinline__ void store_point(const HeviArgs& a, int k, int i, int j, const PointTend& r) {
  const Dims& d = a.d;
  const size_t p = IX(d, k, i, j);
  if constexpr ((PART & P_CT) != 0) {
    double* mfZ = a.mflx_hi; double* mfX = a.mflx_hi + d.n3; double* mfY = a.mflx_hi + 2 * d.n3;
    double* tfZ = a.tflx_hi; double* tfX = a.tflx_hi + d.n3; double* tfY = a.tflx_hi + 2 * d.n3;
    mfX[p] = r.mX;
    mfY[p] = r.mY;
    tfX[p] = r.tX;
    tfY[p] = r.tY;
    if (i == d.IS) { mfX[p - 1] = r.mXm; tfX[p - 1] = r.tXm; }
    if (j == d.JS) { mfY[p - d.IA] = r.mYm; tfY[p - d.IA] = r.tYm; }
    if (k <= d.KE - 1) { mfZ[p] = r.mZ; tfZ[p] = r.tZ; }
    else { mfZ[p] = 0.0; tfZ[p] = 0.0; }
    if (k == d.KS) { mfZ[p - d.pl] = 0.0; tfZ[p - d.pl] = 0.0; }
    if (i == d.IS) {  // mflx_hi(Z) is also defined on the i = IS-1 column (HEVI.F90:370)
      mfZ[p - 1] = r.mZw;
      if (k == d.KS) mfZ[p - 1 - d.pl] = 0.0;
    }
  }
  if constexpr ((PART & P_U) != 0) a.MOMX_RK[p] = r.MX;
  if constexpr ((PART & P_V) != 0) a.MOMY_RK[p] = r.MY;
}

// ---------------------------------------------------------------------------
// One thread per column (i,j), marching k upward through the explicit part
// and the forward sweep of the Thomas algorithm (HEVI.F90:688-727,
// 1092-1115), then downward through the back substitution and the implicit
// updates (HEVI.F90:728-804, 1116-1127).  Row r of the tridiagonal system
// needs the tendencies of level r+1, so the sweep runs one level behind the
// explicit part.  Between the two passes the column keeps e, f and PT in
// scratch and Sr/St in the DENS_RK/RHOT_RK outputs; the explicit z fluxes
// wait in mflx_hi/tflx_hi for their implicit parts.
template <int S, bool TERRAIN, bool DIVDMP>
__global__ void __launch_bounds__(256) hevi_march(HeviArgs a, unsigned ni_tiles, unsigned nblocks) {
  const Dims& d = a.d;
  Ops<S, TERRAIN, DIVDMP> o(a);
  const unsigned lb = xcd_block(blockIdx.x, nblocks);
  const int i = d.IS + (int)((lb % ni_tiles) * 64 + threadIdx.x);
  const int j = d.JS + (int)((lb / ni_tiles) * blockDim.y + threadIdx.y);
  if (i > d.IE || j > d.JE) return;
  const auto& m = o.m;
  const double* CDZ = a.g.CDZ; const double* RCDZ = a.g.RCDZ; const double* RFDZ = a.g.RFDZ;
  const double dtrk = a.dtrk, J33G = a.J33G;
  const int KS = d.KS, KE = d.KE;
  auto px = [&](int k) { return IX(d, k, i, j); };
  auto dpres = [&](int k) { size_t q = px(k); return a.DPRES0[q] + a.RT2P[q] * (a.RHOT[q] - a.REF_rhot[q]); };
  auto Af = [&](int k, double rt) { return dtrk * dtrk * J33G * RCDZ[k] * rt * J33G / m.gs(I_XYZ, px(k)); };
  auto f2h = [&](int k, int pp) {
    return CDZ[k + pp - 1] * m.gs(I_XYZ, px(k + pp - 1)) / (CDZ[k] * m.gs(I_XYZ, px(k)) + CDZ[k + 1] * m.gs(I_XYZ, px(k + 1)));
  };
  double* Sr_ = a.DENS_RK;  // Sr/St wait in the outputs until the back pass
  double* St_ = a.RHOT_RK;

  // ---- upward: explicit part at level k, sweep row k-1 ----
  double rt0 = 0.0, dp0 = 0.0, st0 = 0.0, sr0 = 0.0, d0 = 0.0, rd0 = 0.0, A0 = 0.0, pt0 = 0.0, ptm = 0.0, sw0 = 0.0;
  double e_prev = 0.0, f_prev = 0.0, C_last = 0.0;
  ZCarry zc;
  zc.pw[1] = o.POTT(KS - 1, i, j); zc.pw[2] = o.POTT(KS, i, j); zc.pw[3] = o.POTT(KS + 1, i, j);
  for (int k = KS; k <= KE; ++k) {
    zc.pw[0] = zc.pw[1]; zc.pw[1] = zc.pw[2]; zc.pw[2] = zc.pw[3];
    zc.pw[3] = o.POTT(k + 2, i, j);
    const PointTend t = explicit_point<S, TERRAIN, DIVDMP>(o, a, k, i, j, zc);
    const size_t q1 = px(k);
    const double rt1 = a.RT2P[q1], dp1 = dpres(k), d1 = a.DENS[q1], rd1 = a.REF_dens[q1], A1 = Af(k, rt1);
    const double pt1 = (k <= KE - 1) ? t.PT : 0.0;
    double ek = 0.0, fk = 0.0;
    if (k > KS) {
      const int r = k - 1;
      const size_t q = px(r);
      const double gw = m.gs(I_XYW, q);
      const double B = GRAV * (dtrk * dtrk) * J33G / (CDZ[r + 1] + CDZ[r]);
      const double F1v = -(pt1 * RFDZ[r] * A1 + B) / gw;
      const double F2v = 1.0 + (pt0 * RFDZ[r] * (A1 + A0)) / gw;
      const double F3v = -(ptm * RFDZ[r] * A0 - B) / gw;
      const double pg = -(dp1 + rt1 * dtrk * t.St - dp0 - rt0 * dtrk * st0) * RFDZ[r] * J33G / gw
                      - GRAV * (f2h(r, 1) * (d1 - rd1 + t.Sr * dtrk) + f2h(r, 2) * (d0 - rd0 + sr0 * dtrk));
      const double Ck = a.MOMZ[q] + dtrk * (pg + sw0);
      if (r == KS) {
        const double rdenom = 1.0 / F2v;
        ek = -F1v * rdenom;
        fk = Ck * rdenom;
      } else if (r < KE - 1) {
        const double rdenom = 1.0 / (F2v + F3v * e_prev);
        ek = -F1v * rdenom;
        fk = (Ck - F3v * f_prev) * rdenom;
      } else {
        C_last = (Ck - F3v * f_prev) / (F2v + F3v * e_prev);
      }
      e_prev = ek; f_prev = fk;
      ptm = pt0;
    }
    // stores of this level, after all of its loads
    store_point(a, k, i, j, t);
    Sr_[q1] = t.Sr;
    St_[q1] = t.St;
    if (a.trSt) a.trSt[q1] = t.St;
    if (k <= KE - 1) a.PT[q1] = t.PT;
    if (k > KS && k - 1 < KE - 1) { a.e[px(k - 1)] = ek; a.f[px(k - 1)] = fk; }
    rt0 = rt1; dp0 = dp1; st0 = t.St; sr0 = t.Sr; d0 = d1; rd0 = rd1; A0 = A1; pt0 = pt1; sw0 = t.Sw;
  }

  // ---- downward: back substitution and updates (HEVI.F90:728-790) ----
  double* mfZ = a.mflx_hi; double* tfZ = a.tflx_hi;
  const double M12 = m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY);
  double Cn = C_last;                         // C(k+1)
  double ptn = a.PT[px(KE - 1)];              // PT(k+1)
  {
    const size_t q = px(KE);
    double advcv = Cn * J33G * RCDZ[KE] / m.gs(I_XYZ, q);
    a.DENS_RK[q] = a.DENS0[q] + dtrk * (advcv + Sr_[q]);
    advcv = Cn * ptn * J33G * RCDZ[KE] / m.gs(I_XYZ, q);
    a.RHOT_RK[q] = a.RHOT0[q] + dtrk * (advcv + St_[q]);
    a.MOMZ_RK[q] = 0.0;
  }
  {
    const size_t q = px(KE - 1);
    mfZ[q] = mfZ[q] + J33G * Cn / (M12);
    tfZ[q] = tfZ[q] + J33G * Cn * ptn / (M12);
    a.MOMZ_RK[q] = a.MOMZ0[q] + (Cn - a.MOMZ[q]);
    if (a.trCp) a.trCp[q] = Cn;
  }
  for (int k = KE - 2; k >= KS; --k) {
    const size_t q = px(k), q1 = px(k + 1);
    const double Ck = a.e[q] * Cn + a.f[q];
    const double ptk = a.PT[q];
    mfZ[q] = mfZ[q] + J33G * Ck / (M12);
    tfZ[q] = tfZ[q] + J33G * Ck * ptk / (M12);
    a.MOMZ_RK[q] = a.MOMZ0[q] + (Ck - a.MOMZ[q]);
    if (a.trCp) a.trCp[q] = Ck;
    // level k+1 (KS+1..KE-1) now has both C(k+1) and C(k)
    double advcv = -(Cn - Ck) * J33G * RCDZ[k + 1] / m.gs(I_XYZ, q1);
    a.DENS_RK[q1] = a.DENS0[q1] + dtrk * (advcv + Sr_[q1]);
    advcv = -(Cn * ptn - Ck * ptk) * J33G * RCDZ[k + 1] / m.gs(I_XYZ, q1);
    a.RHOT_RK[q1] = a.RHOT0[q1] + dtrk * (advcv + St_[q1]);
    Cn = Ck; ptn = ptk;
  }
  {
    const size_t q = px(KS);
    double advcv = -Cn * J33G * RCDZ[KS] / m.gs(I_XYZ, q);
    a.DENS_RK[q] = a.DENS0[q] + dtrk * (advcv + Sr_[q]);
    advcv = -Cn * ptn * J33G * RCDZ[KS] / m.gs(I_XYZ, q);
    a.RHOT_RK[q] = a.RHOT0[q] + dtrk * (advcv + St_[q]);
    a.MOMZ_RK[px(KS - 1)] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// The same stage split in launches, the terrain stage's default
// (hevi_march runs one thread per column: 16K columns at 128x128 are 256
// waves, and the whole point needs 256 VGPRs).  hevi_expl: the explicit part
// (the sections PART of it), one thread per segment of L levels of a column,
// the carried z fluxes of the face below the segment computed afresh (the
// carry of hevi_march at that face); it stores the mass and theta fluxes,
// MOMX_RK/MOMY_RK, Sr/St (in DENS_RK/RHOT_RK), PT and Sw.  hevi_solve: one
// thread per column, the tridiagonal system, its forward sweep and the back
// substitution of hevi_march from those arrays.  The expressions and their
// order are hevi_march's: the results are its bits.
template <int S, bool TERRAIN, bool DIVDMP, int PART>
__global__ void __launch_bounds__(256) hevi_expl(HeviArgs a, unsigned ni_tiles, unsigned ncols, unsigned nblocks, int L) {
  const Dims& d = a.d;
  Ops<S, TERRAIN, DIVDMP> o(a);
  const unsigned lb = xcd_block(blockIdx.x, nblocks);
  const unsigned col = lb % ncols, seg = lb / ncols;
  const int i = d.IS + (int)((col % ni_tiles) * 64 + threadIdx.x);
  const int j = d.JS + (int)((col / ni_tiles) * blockDim.y + threadIdx.y);
  const int k0 = d.KS + (int)seg * L;
  if (i > d.IE || j > d.JE || k0 > d.KE) return;
  const int k1 = min(d.KE, k0 + L - 1);
  constexpr bool PW = (PART & (P_CT | P_W)) != 0;  // the sections that read POTT around k
  ZCarry zc;
  if (k0 > d.KS) {
    const int kb = k0 - 1;  // the face below the segment, as level kb of hevi_march leaves it
    if constexpr ((PART & P_CT) != 0) {
      double pw[4] = {o.POTT(kb - 1, i, j), o.POTT(kb, i, j), o.POTT(kb + 1, i, j), o.POTT(kb + 2, i, j)};
      zc.mZ = o.mflxZf(kb, i, j);
      zc.tZ = o.tflxZw(kb, i, j, zc.mZ, pw);
    }
    if constexpr ((PART & P_W) != 0) {
      zc.qW = o.qZ_XYW(kb, i, j);
      if constexpr (TERRAIN) { zc.qWJ13 = o.qJ_XYW(0, kb, i, j); zc.qWJ23 = o.qJ_XYW(1, kb, i, j); }
    }
    if constexpr ((PART & P_U) != 0) {
      zc.qU = o.qZ_UV(0, kb, i, j);
      if constexpr (TERRAIN) { zc.qUJ13 = o.qJ_UV(0, 0, kb, i, j); zc.qUJ23 = o.qJ_UV(0, 1, kb, i, j); }
    }
    if constexpr ((PART & P_V) != 0) {
      zc.qV = o.qZ_UV(1, kb, i, j);
      if constexpr (TERRAIN) { zc.qVJ13 = o.qJ_UV(1, 0, kb, i, j); zc.qVJ23 = o.qJ_UV(1, 1, kb, i, j); }
    }
  }
  if constexpr (PW) { zc.pw[1] = o.POTT(k0 - 1, i, j); zc.pw[2] = o.POTT(k0, i, j); zc.pw[3] = o.POTT(k0 + 1, i, j); }
  for (int k = k0; k <= k1; ++k) {
    if constexpr (PW) {
      zc.pw[0] = zc.pw[1]; zc.pw[1] = zc.pw[2]; zc.pw[2] = zc.pw[3];
      zc.pw[3] = o.POTT(k + 2, i, j);
    }
    const PointTend t = explicit_point<S, TERRAIN, DIVDMP, PART>(o, a, k, i, j, zc);
    const size_t q = IX(d, k, i, j);
    store_point<PART>(a, k, i, j, t);
    if constexpr ((PART & P_CT) != 0) {
      a.DENS_RK[q] = t.Sr;
      a.RHOT_RK[q] = t.St;
      if (a.trSt) a.trSt[q] = t.St;
    }
    if constexpr ((PART & P_W) != 0) {
      if (k <= d.KE - 1) { a.PT[q] = t.PT; a.Sw[q] = t.Sw; }
    }
  }
}

template <int S, bool TERRAIN, bool DIVDMP>
__global__ void __launch_bounds__(256) hevi_solve(HeviArgs a, unsigned ni_tiles, unsigned nblocks) {
  const Dims& d = a.d;
  Ops<S, TERRAIN, DIVDMP> o(a);
  const unsigned lb = xcd_block(blockIdx.x, nblocks);
  const int i = d.IS + (int)((lb % ni_tiles) * 64 + threadIdx.x);
  const int j = d.JS + (int)((lb / ni_tiles) * blockDim.y + threadIdx.y);
  if (i > d.IE || j > d.JE) return;
  const auto& m = o.m;
  const double* CDZ = a.g.CDZ; const double* RCDZ = a.g.RCDZ; const double* RFDZ = a.g.RFDZ;
  const double dtrk = a.dtrk, J33G = a.J33G;
  const int KS = d.KS, KE = d.KE;
  auto px = [&](int k) { return IX(d, k, i, j); };
  auto dpres = [&](int k) { size_t q = px(k); return a.DPRES0[q] + a.RT2P[q] * (a.RHOT[q] - a.REF_rhot[q]); };
  auto Af = [&](int k, double rt) { return dtrk * dtrk * J33G * RCDZ[k] * rt * J33G / m.gs(I_XYZ, px(k)); };
  auto f2h = [&](int k, int pp) {
    return CDZ[k + pp - 1] * m.gs(I_XYZ, px(k + pp - 1)) / (CDZ[k] * m.gs(I_XYZ, px(k)) + CDZ[k + 1] * m.gs(I_XYZ, px(k + 1)));
  };
  double* Sr_ = a.DENS_RK;
  double* St_ = a.RHOT_RK;

  // ---- upward: row k-1 of the system once level k is known; the loads of
  //      level k+1 are issued before the work of level k ----
  struct Up {
    double rt, dp0, rh, rr, dn, rd, pt, st, sr, sw, mz;
  };
  auto up_load = [&](int k, Up& L) {
    const size_t q = px(k);
    L.rt = a.RT2P[q]; L.dp0 = a.DPRES0[q]; L.rh = a.RHOT[q]; L.rr = a.REF_rhot[q];
    L.dn = a.DENS[q]; L.rd = a.REF_dens[q];
    L.st = St_[q]; L.sr = Sr_[q];
    const bool w = k <= KE - 1;
    L.pt = w ? a.PT[q] : 0.0;
    L.sw = w ? a.Sw[q] : 0.0;
    L.mz = w ? a.MOMZ[q] : 0.0;
  };
  double rt0 = 0.0, dp0 = 0.0, st0 = 0.0, sr0 = 0.0, d0 = 0.0, rd0 = 0.0, A0 = 0.0, pt0 = 0.0, ptm = 0.0, sw0 = 0.0, mz0 = 0.0;
  double e_prev = 0.0, f_prev = 0.0, C_last = 0.0;
  Up cur{}, nxt{};
  up_load(KS, cur);
  for (int k = KS; k <= KE; ++k) {
    if (k < KE) up_load(k + 1, nxt);
    const double rt1 = cur.rt, dp1 = cur.dp0 + cur.rt * (cur.rh - cur.rr), d1 = cur.dn, rd1 = cur.rd, A1 = Af(k, rt1);
    const double pt1 = cur.pt, st1 = cur.st, sr1 = cur.sr, sw1 = cur.sw;
    double ek = 0.0, fk = 0.0;
    if (k > KS) {
      const int r = k - 1;
      const size_t q = px(r);
      const double gw = m.gs(I_XYW, q);
      const double B = GRAV * (dtrk * dtrk) * J33G / (CDZ[r + 1] + CDZ[r]);
      const double F1v = -(pt1 * RFDZ[r] * A1 + B) / gw;
      const double F2v = 1.0 + (pt0 * RFDZ[r] * (A1 + A0)) / gw;
      const double F3v = -(ptm * RFDZ[r] * A0 - B) / gw;
      const double pg = -(dp1 + rt1 * dtrk * st1 - dp0 - rt0 * dtrk * st0) * RFDZ[r] * J33G / gw
                      - GRAV * (f2h(r, 1) * (d1 - rd1 + sr1 * dtrk) + f2h(r, 2) * (d0 - rd0 + sr0 * dtrk));
      const double Ck = mz0 + dtrk * (pg + sw0);
      if (r == KS) {
        const double rdenom = 1.0 / F2v;
        ek = -F1v * rdenom;
        fk = Ck * rdenom;
      } else if (r < KE - 1) {
        const double rdenom = 1.0 / (F2v + F3v * e_prev);
        ek = -F1v * rdenom;
        fk = (Ck - F3v * f_prev) * rdenom;
      } else {
        C_last = (Ck - F3v * f_prev) / (F2v + F3v * e_prev);
      }
      e_prev = ek; f_prev = fk;
      ptm = pt0;
      if (r < KE - 1) { a.e[q] = ek; a.f[q] = fk; }
    }
    rt0 = rt1; dp0 = dp1; st0 = st1; sr0 = sr1; d0 = d1; rd0 = rd1; A0 = A1; pt0 = pt1; sw0 = sw1; mz0 = cur.mz;
    cur = nxt;
  }

  // ---- downward: back substitution and updates (HEVI.F90:728-790) ----
  double* mfZ = a.mflx_hi; double* tfZ = a.tflx_hi;
  const double M12 = m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY);
  double Cn = C_last;
  double ptn = a.PT[px(KE - 1)];
  {
    const size_t q = px(KE);
    double advcv = Cn * J33G * RCDZ[KE] / m.gs(I_XYZ, q);
    a.DENS_RK[q] = a.DENS0[q] + dtrk * (advcv + Sr_[q]);
    advcv = Cn * ptn * J33G * RCDZ[KE] / m.gs(I_XYZ, q);
    a.RHOT_RK[q] = a.RHOT0[q] + dtrk * (advcv + St_[q]);
    a.MOMZ_RK[q] = 0.0;
  }
  {
    const size_t q = px(KE - 1);
    mfZ[q] = mfZ[q] + J33G * Cn / (M12);
    tfZ[q] = tfZ[q] + J33G * Cn * ptn / (M12);
    a.MOMZ_RK[q] = a.MOMZ0[q] + (Cn - a.MOMZ[q]);
    if (a.trCp) a.trCp[q] = Cn;
  }
  // the loads of level k-1 (and of level k's Sr/St/DENS0/RHOT0) are issued
  // before the work of level k
  struct Dn {
    double e, f, pt, mz0, mz, mf, tf, d0, r0, sr, st;
  };
  auto dn_load = [&](int k, Dn& L) {
    const size_t q = px(k), q1 = px(k + 1);
    L.e = a.e[q]; L.f = a.f[q]; L.pt = a.PT[q]; L.mz0 = a.MOMZ0[q]; L.mz = a.MOMZ[q];
    L.mf = mfZ[q]; L.tf = tfZ[q];
    L.d0 = a.DENS0[q1]; L.r0 = a.RHOT0[q1]; L.sr = Sr_[q1]; L.st = St_[q1];
  };
  Dn dc{}, dn{};
  if (KE - 2 >= KS) dn_load(KE - 2, dc);
  for (int k = KE - 2; k >= KS; --k) {
    if (k > KS) dn_load(k - 1, dn);
    const size_t q = px(k), q1 = px(k + 1);
    const double Ck = dc.e * Cn + dc.f;
    const double ptk = dc.pt;
    mfZ[q] = dc.mf + J33G * Ck / (M12);
    tfZ[q] = dc.tf + J33G * Ck * ptk / (M12);
    a.MOMZ_RK[q] = dc.mz0 + (Ck - dc.mz);
    if (a.trCp) a.trCp[q] = Ck;
    double advcv = -(Cn - Ck) * J33G * RCDZ[k + 1] / m.gs(I_XYZ, q1);
    a.DENS_RK[q1] = dc.d0 + dtrk * (advcv + dc.sr);
    advcv = -(Cn * ptn - Ck * ptk) * J33G * RCDZ[k + 1] / m.gs(I_XYZ, q1);
    a.RHOT_RK[q1] = dc.r0 + dtrk * (advcv + dc.st);
    Cn = Ck; ptn = ptk;
    dc = dn;
  }
  {
    const size_t q = px(KS);
    double advcv = -Cn * J33G * RCDZ[KS] / m.gs(I_XYZ, q);
    a.DENS_RK[q] = a.DENS0[q] + dtrk * (advcv + Sr_[q]);
    advcv = -Cn * ptn * J33G * RCDZ[KS] / m.gs(I_XYZ, q);
    a.RHOT_RK[q] = a.RHOT0[q] + dtrk * (advcv + St_[q]);
    a.MOMZ_RK[px(KS - 1)] = 0.0;
  }
}

// levels per segment of hevi_expl; SCALE_HIP_TERRAIN_SEG overrides the
// default 8 (0: the single-launch hevi_march)
static int terrain_seg_levels() {
  const char* e = getenv("SCALE_HIP_TERRAIN_SEG");
  return e ? atoi(e) : 8;
}

// the sections of hevi_expl per launch: SCALE_HIP_TERRAIN_PARTS lists the
// section masks of the launches ("15": one launch; "3,12"; "1,2,4,8").  One
// launch per section is the default: with the whole point in one launch a
// thread needs 256 VGPRs (one wave per SIMD); per section 150-256, with two
// to three waves for the momenta.  Measured at 128x128x64 / 512x512x128
// (bell mountain, UD3): 1.07 / 25.4 ms per stage in one launch, 0.95 / 22.4
// in two, 0.76 / 15.9 in four; hevi_march 2.92 / 27.6.
static std::vector<int> terrain_parts() {
  const char* e = getenv("SCALE_HIP_TERRAIN_PARTS");
  std::vector<int> v;
  for (const char* c = e ? e : "1,2,4,8"; *c;) {
    char* end = nullptr;
    const long m = strtol(c, &end, 10);
    if (end == c) break;
    v.push_back((int)m);
    c = *end == ',' ? end + 1 : end;
  }
  return v;
}

template <int S, bool TERRAIN, bool DIVDMP>
static void launch_split(const HeviArgs& a, hipStream_t s) {
  const Dims& d = a.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1, kmax = d.KE - d.KS + 1;
  constexpr unsigned TJ = 4;
  const unsigned ni = (imax + 63) / 64, ncols = ni * ((jmax + TJ - 1) / TJ);
  const int L = terrain_seg_levels();
  if (L <= 0) {
    hipLaunchKernelGGL((hevi_march<S, TERRAIN, DIVDMP>), dim3(ncols), dim3(64, TJ), 0, s, a, ni, ncols);
    return;
  }
  const unsigned nseg = (unsigned)((kmax + L - 1) / L), nb = ncols * nseg;
  const dim3 g(nb), b(64, TJ);
  for (int part : terrain_parts()) {
    switch (part) {
      case P_ALL: hipLaunchKernelGGL((hevi_expl<S, TERRAIN, DIVDMP, P_ALL>), g, b, 0, s, a, ni, ncols, nb, L); break;
      case P_CT | P_W: hipLaunchKernelGGL((hevi_expl<S, TERRAIN, DIVDMP, P_CT | P_W>), g, b, 0, s, a, ni, ncols, nb, L); break;
      case P_U | P_V: hipLaunchKernelGGL((hevi_expl<S, TERRAIN, DIVDMP, P_U | P_V>), g, b, 0, s, a, ni, ncols, nb, L); break;
      case P_CT: hipLaunchKernelGGL((hevi_expl<S, TERRAIN, DIVDMP, P_CT>), g, b, 0, s, a, ni, ncols, nb, L); break;
      case P_W: hipLaunchKernelGGL((hevi_expl<S, TERRAIN, DIVDMP, P_W>), g, b, 0, s, a, ni, ncols, nb, L); break;
      case P_U: hipLaunchKernelGGL((hevi_expl<S, TERRAIN, DIVDMP, P_U>), g, b, 0, s, a, ni, ncols, nb, L); break;
      case P_V: hipLaunchKernelGGL((hevi_expl<S, TERRAIN, DIVDMP, P_V>), g, b, 0, s, a, ni, ncols, nb, L); break;
      default: break;
    }
  }
  hipLaunchKernelGGL((hevi_solve<S, TERRAIN, DIVDMP>), dim3(ncols), dim3(64, TJ), 0, s, a, ni, ncols);
}

// Flat grids whose 64 x 8 tiles do not give every CU a block (hevi_tile at
// 128x128 is 32 blocks on 256 CUs) take the split launches instead; the RK4
// combination then runs as its own kernel.  SCALE_HIP_FLAT_SPLIT=0 keeps the
// tile kernel, =1 the split launches, whatever the grid.
bool hevi_flat_tile(const Ctx& c) {
  const char* e = getenv("SCALE_HIP_FLAT_SPLIT");
  if (e && *e) return atoi(e) == 0;
  const Dims& d = c.d;
  const unsigned nt = (unsigned)((d.IE - d.IS + 1 + 63) / 64) * (unsigned)((d.JE - d.JS + 1 + 7) / 8);
  return nt >= (unsigned)c.ncu;
}

// ---------------------------------------------------------------------------
// HEVE: the fully explicit short step (ATMOS_DYN_Tstep_short_fvm_heve,
// scale_atmos_dyn_tstep_short_fvm_heve.F90:110-1414).  Same column march as
// hevi_march without the solve: the vertical mass flux carries J33G MOMZ
// (HEVE.F90:468-478) and upwinds the theta flux with it, MOMZ takes its
// pressure gradient and buoyancy explicitly (HEVE.F90:660-677), and the
// horizontal divergence damping has the HEVE form FDX (DDIV(i+1) - DDIV(i))
// (HEVE.F90:933, 1182).
template <int S, bool TERRAIN, bool DIVDMP>
__device__ __forceinline__ double heve_mflxZ(const Ops<S, TERRAIN, DIVDMP>& o, const HeviArgs& a, int k, int i, int j) {
  const Dims& d = a.d;
  const auto& m = o.m;
  const size_t p = IX(d, k, i, j);
  double r = a.J33G * a.MOMZ[p] / (m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY));
  if constexpr (TERRAIN) {
    const double* X = a.MOMX; const double* Y = a.MOMY;
    r = r + m.j13(I_XYW, p) * 0.25
                * (X[IX(d, k + 1, i, j)] + X[IX(d, k + 1, i - 1, j)] + X[p] + X[IX(d, k, i - 1, j)])
                / m.mapf(i, j, 1, I_XY)
          + m.j23(I_XYW, p) * 0.25
                * (Y[IX(d, k + 1, i, j)] + Y[IX(d, k + 1, i, j - 1)] + Y[p] + Y[IX(d, k, i, j - 1)])
                / m.mapf(i, j, 0, I_XY);
  }
  return r + m.gs(I_XYW, p) * o.ND(I_DENS, ZDIR, p) / (m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY));
}

template <int S, bool TERRAIN, bool DIVDMP>
__global__ void __launch_bounds__(256) heve_march(HeviArgs a, unsigned ni_tiles, unsigned nblocks) {
  const Dims& d = a.d;
  Ops<S, TERRAIN, DIVDMP> o(a);
  const unsigned lb = xcd_block(blockIdx.x, nblocks);
  const int i = d.IS + (int)((lb % ni_tiles) * 64 + threadIdx.x);
  const int j = d.JS + (int)((lb / ni_tiles) * blockDim.y + threadIdx.y);
  if (i > d.IE || j > d.JE) return;
  const auto& m = o.m;
  const double* RCDX = a.g.RCDX; const double* RCDY = a.g.RCDY; const double* RCDZ = a.g.RCDZ;
  const double* RFDZ = a.g.RFDZ;
  const double dtrk = a.dtrk;
  const int KS = d.KS, KE = d.KE;
  double* mfZ = a.mflx_hi; double* mfX = a.mflx_hi + d.n3; double* mfY = a.mflx_hi + 2 * d.n3;
  double* tfZ = a.tflx_hi; double* tfX = a.tflx_hi + d.n3; double* tfY = a.tflx_hi + 2 * d.n3;
  ZCarry zc;
  zc.pw[1] = o.POTT(KS - 1, i, j); zc.pw[2] = o.POTT(KS, i, j); zc.pw[3] = o.POTT(KS + 1, i, j);
  for (int k = KS; k <= KE; ++k) {
    zc.pw[0] = zc.pw[1]; zc.pw[1] = zc.pw[2]; zc.pw[2] = zc.pw[3];
    zc.pw[3] = o.POTT(k + 2, i, j);
    const size_t p = IX(d, k, i, j);

    // ---- continuity (HEVE.F90:452-572) ----
    const double mX = o.mflxX(k, i, j), mXm = o.mflxX(k, i - 1, j);
    const double mY = o.mflxY(k, i, j), mYm = o.mflxY(k, i, j - 1);
    const double mZ = (k <= KE - 1) ? heve_mflxZ(o, a, k, i, j) : 0.0, mZm = zc.mZ;
    zc.mZ = mZ;
    double DR;
    {
      double advcv = -(mZ - mZm) * RCDZ[k];
      double advch = -(mX - mXm) * RCDX[i] - (mY - mYm) * RCDY[j];
      DR = a.DENS0[p] + dtrk * ((advcv + advch) * m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY) / m.gs(I_XYZ, p)
                                + a.DENS_t[p]);
    }

    // ---- thermodynamics (HEVE.F90:1316-1375) ----
    const double tX = o.tflxX(k, i, j, mX), tXm = o.tflxX(k, i - 1, j, mXm);
    const double tY = o.tflxY(k, i, j, mY), tYm = o.tflxY(k, i, j - 1, mYm);
    const double tZ = o.tflxZw(k, i, j, mZ, zc.pw), tZm = zc.tZ;
    zc.tZ = tZ;
    double RR;
    {
      double advcv = -(tZ - tZm) * RCDZ[k];
      double advch = -(tX - tXm) * RCDX[i] - (tY - tYm) * RCDY[j];
      RR = a.RHOT0[p] + dtrk * ((advcv + advch) * m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY) / m.gs(I_XYZ, p)
                                + a.RHOT_t[p]);
    }

    // ---- momentum z (HEVE.F90:588-700) ----
    double WR = 0.0;
    if (k <= KE - 1) {
      const double qw = o.qZ_XYW(k, i, j);
      double sv = qw - zc.qW;
      zc.qW = qw;
      if constexpr (TERRAIN) {
        const double q13 = o.qJ_XYW(0, k, i, j), q23 = o.qJ_XYW(1, k, i, j);
        sv = sv + q13 - zc.qWJ13 + q23 - zc.qWJ23;
        zc.qWJ13 = q13; zc.qWJ23 = q23;
      }
      const double pgf = a.J33G * (o.DPRES(k + 1, i, j) - o.DPRES(k, i, j)) * RFDZ[k];
      const double buoy = GRAV * m.gs(I_XYW, p)
                        * (o.F2H(k, 1, I_XYZ, i, j) * (o.D(k + 1, i, j) - a.REF_dens[IX(d, k + 1, i, j)])
                         + o.F2H(k, 2, I_XYZ, i, j) * (o.D(k, i, j) - a.REF_dens[p]));
      double advcv = -(sv) * RFDZ[k];
      double advch = -((o.qH_XYW(0, k, i, j) - o.qH_XYW(0, k, i - 1, j)) * RCDX[i]
                     + (o.qH_XYW(1, k, i, j) - o.qH_XYW(1, k, i, j - 1)) * RCDY[j])
                   * m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY);
      double t = (advcv + advch - pgf - buoy) / m.gs(I_XYW, p) + (-a.wdamp_coef[k] * a.MOMZ0[p]);
      if constexpr (DIVDMP) t = t + a.divdmp_coef / dtrk * a.g.FDZ[k] * (a.DDIV[IX(d, k + 1, i, j)] - a.DDIV[p]);
      WR = a.MOMZ0[p] + dtrk * (t + a.MOMZ_t[p]);
    }

    // ---- momentum x and y (HEVE.F90:744-945, 993-1193) ----
    double tx = momx_core(o, a, k, i, j, zc);
    if constexpr (DIVDMP) tx = tx + a.divdmp_coef / dtrk * a.g.FDX[i] * (a.DDIV[IX(d, k, i + 1, j)] - a.DDIV[p]);
    double ty = momy_core(o, a, k, i, j, zc);
    if constexpr (DIVDMP) ty = ty + a.divdmp_coef / dtrk * a.g.FDY[j] * (a.DDIV[IX(d, k, i, j + 1)] - a.DDIV[p]);

    // ---- stores, after every load of this level ----
    mfX[p] = mX; mfY[p] = mY; tfX[p] = tX; tfY[p] = tY;
    if (i == d.IS) { mfX[p - 1] = mXm; tfX[p - 1] = tXm; }
    if (j == d.JS) { mfY[p - d.IA] = mYm; tfY[p - d.IA] = tYm; }
    mfZ[p] = mZ; tfZ[p] = tZ;  // both 0 at KE
    if (k == KS) { mfZ[p - d.pl] = 0.0; tfZ[p - d.pl] = 0.0; a.MOMZ_RK[p - d.pl] = 0.0; }
    a.DENS_RK[p] = DR;
    a.RHOT_RK[p] = RR;
    a.MOMZ_RK[p] = WR;  // 0 at KE (HEVE.F90:697-698)
    a.MOMX_RK[p] = a.MOMX0[p] + dtrk * (tx + a.MOMX_t[p]);
    a.MOMY_RK[p] = a.MOMY0[p] + dtrk * (ty + a.MOMY_t[p]);
  }
}

// Division in the tile kernel: exact in FP64; in FP32 the hardware
// reciprocal (1 ulp) times the numerator, within the FP32 path's own rounding
template <typename R>
__device__ __forceinline__ R rdiv(R x, R y) {
  if constexpr (sizeof(R) == sizeof(float)) return x * __builtin_amdgcn_rcpf(y);
  else return x / y;
}

// ---------------------------------------------------------------------------
// Flat-terrain stage with LDS tiles: blocks of 64 (i) x TJ (j) columns march
// k together.  At each level the block
//   (a) stages DENS/MOMX/MOMY/MOMZ and POTT = RHOT/DENS in LDS with a 2-cell
//       halo, in a three-level ring (k, k+1 and k+2: the faces read k and
//       k+1, the own column's 4-level windows take k+2 from it), and DPRES
//       at level k;
//   (b) evaluates every x face and every y face of the tile once (mass and
//       theta fluxes, the x/y fluxes of MOMZ, MOMX and MOMY) into LDS;
//   (c) lets each column difference its faces, carry its z fluxes from the
//       level below, update MOMX/MOMY and run the Thomas forward sweep one
//       level behind, as hevi_march does; the back substitution follows.
// With GSQRT = 1, J13G = J23G = 0 and MAPF = 1 every metric factor of the
// general expressions is a multiplication or division by 1.0 (or an addition
// of 0.0), so the expressions below, written without them, give the same
// IEEE results (HEVI.F90:346-1047; UD3.F90/CD4.F90 flux routines).
// One tile (logical tile index lb of a grid ni_tiles tiles wide) of the
// stage; the kernels below run it once per block or, for an ensemble, over a
// table of members.  A thread that owns no column returns early from the
// back substitution, after the last barrier.
template <int S, bool DIVDMP, int TJ, typename R, bool CMB>
__device__ __forceinline__ void hevi_tile_body(const HeviArgsT<R>& a, unsigned lb, unsigned ni_tiles) {
  constexpr int TI = 64, CW = TI + 4, RW = TJ + 4, NT = TI * TJ;
  __shared__ R sD[3][RW][CW], sMX[3][RW][CW], sMY[3][RW][CW];  // ring slot = level % 3
  __shared__ R sPO[3][RW][CW], sW[3][RW][CW];
  __shared__ double sDP[TJ + 1][TI + 1];
  // x faces: position x = 0..TI is the face east of column i0-1+x
  __shared__ R xM[TJ][TI + 1], xT[TJ][TI + 1], xW[TJ][TI + 1], xU[TJ][TI + 1], xV[TJ][TI + 1];
  // y faces: position y = 0..TJ is the face north of row j0-1+y
  __shared__ R yM[TJ + 1][TI], yT[TJ + 1][TI], yW[TJ + 1][TI], yU[TJ + 1][TI], yV[TJ + 1][TI];

  const Dims& d = a.d;
  const int i0 = d.IS + (int)(lb % ni_tiles) * TI;
  const int j0 = d.JS + (int)(lb / ni_tiles) * TJ;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TI * ty;
  const int i = i0 + tx, j = j0 + ty;
  const bool own = (i <= d.IE) && (j <= d.JE);
  const int KS = d.KS, KE = d.KE;
  const size_t n3 = d.n3;
  const double* CDZ = a.g.CDZ; const double* RCDZ = a.g.RCDZ; const double* RFDZ = a.g.RFDZ;
  const double* RCDX = a.g.RCDX; const double* RCDY = a.g.RCDY; const double* RFDX = a.g.RFDX; const double* RFDY = a.g.RFDY;
  const double dtrk = a.dtrk, J33G = a.J33G;
  const R* ND = a.num_diff;  // component (v, dir) at ND + n3*(v + 5*dir)
  auto nd = [&](int v, int dir, size_t q) { return ND[q + n3 * (size_t)(v + 5 * dir)]; };
  auto gix = [&](int k, int ii, int jj) { return IX(d, k, ii, jj); };
  // full values of the stored DENS / RHOT at level k, and deviations to store
  // (the identity in FP64, where the arrays hold the full values)
  constexpr bool DEV = sizeof(R) == sizeof(float);
  // the level profiles staged in LDS (read every level by every thread)
  __shared__ double sOff[DEV ? 2 * KA_DEV_MAX : 1];
  // FP32: the 1-D z arrays in LDS too, so that the level's arithmetic issues
  // no global load behind the next level's prefetch (sZ rows: CDZ, RCDZ,
  // RFDZ, wdamp_coef)
  __shared__ double sZ[DEV ? 4 * KA_DEV_MAX : 1];
  if constexpr (DEV) {
    for (int q = tid; q < d.KA; q += NT) {
      sOff[q] = a.offD[q]; sOff[KA_DEV_MAX + q] = a.offR[q];
      sZ[q] = CDZ[q]; sZ[KA_DEV_MAX + q] = RCDZ[q]; sZ[3 * KA_DEV_MAX + q] = a.wdamp_coef[q];
      if (q < d.KA - 1) sZ[2 * KA_DEV_MAX + q] = RFDZ[q];
    }
    __syncthreads();
  }
  auto zCDZ = [&](int k) -> double { if constexpr (DEV) return sZ[k]; else return CDZ[k]; };
  auto zRCDZ = [&](int k) -> double { if constexpr (DEV) return sZ[KA_DEV_MAX + k]; else return RCDZ[k]; };
  auto zRFDZ = [&](int k) -> double { if constexpr (DEV) return sZ[2 * KA_DEV_MAX + k]; else return RFDZ[k]; };
  auto zWD = [&](int k) -> double { if constexpr (DEV) return sZ[3 * KA_DEV_MAX + k]; else return a.wdamp_coef[k]; };
  // the own column's x/y metrics and Coriolis factors, fixed over k
  const bool ownc = (i <= d.IE) && (j <= d.JE);
  const double hRCDX = ownc ? RCDX[i] : 0.0, hRCDY = ownc ? RCDY[j] : 0.0;
  const double hRFDX = ownc ? RFDX[i] : 0.0, hRFDY = ownc ? RFDY[j] : 0.0;
  double corx = 0.0, cory = 0.0;
  if (ownc && a.has_coriolis) {
    corx = 0.125 * (a.CORIOLI[(size_t)(i + 1) + (size_t)d.IA * j] + a.CORIOLI[(size_t)i + (size_t)d.IA * j]);
    cory = -0.125 * (a.CORIOLI[(size_t)i + (size_t)d.IA * (j + 1)] + a.CORIOLI[(size_t)i + (size_t)d.IA * j]);
  }
  auto fullDd = [&](R v, int k) -> double { if constexpr (DEV) return (double)v + sOff[k]; else return (double)v; };
  auto fullTd = [&](R v, int k) -> double { if constexpr (DEV) return (double)v + sOff[KA_DEV_MAX + k]; else return (double)v; };
  auto fullD = [&](R v, int k) -> R { if constexpr (DEV) return R(fullDd(v, k)); else return v; };
  auto fullT = [&](R v, int k) -> R { if constexpr (DEV) return R(fullTd(v, k)); else return v; };
  auto devD = [&](double v, int k) -> R { if constexpr (DEV) return R(v - sOff[k]); else return R(v); };
  auto devT = [&](double v, int k) -> R { if constexpr (DEV) return R(v - sOff[KA_DEV_MAX + k]); else return R(v); };

  // (a) the first two levels of the ring; tile cell (r, c) is column i0-2+c,
  // row j0-2+r
  for (int lv = KS; lv <= KS + 1; ++lv) {
    const int sl = lv % 3;
    for (int q = tid; q < RW * CW; q += NT) {
      const int r = q / CW, c = q - r * CW;
      const int ii = i0 - 2 + c, jj = j0 - 2 + r;
      const bool in = ii < d.IA && jj < d.JA;
      const size_t g = in ? gix(lv, ii, jj) : 0;
      const R vD = in ? a.DENS[g] : R(0), vX = in ? a.MOMX[g] : R(0), vY = in ? a.MOMY[g] : R(0);
      const R vW = in ? a.MOMZ[g] : R(0), vR = in ? a.RHOT[g] : R(0);
      sD[sl][r][c] = in ? fullD(vD, lv) : vD;
      sMX[sl][r][c] = vX; sMY[sl][r][c] = vY; sW[sl][r][c] = vW;
      sPO[sl][r][c] = in ? rdiv(fullT(vR, lv), fullD(vD, lv)) : R(0);
    }
  }
  __syncthreads();

  // own-column windows at k-1..k+2 and carried z-face fluxes
  const size_t p0 = own ? gix(0, i, j) : 0;
  auto at = [&](const R* f, int k) { return f[p0 + d.pl * (size_t)k]; };
  R pw[4] = {0, 0, 0, 0}, ww[4] = {0, 0, 0, 0}, xw[4] = {0, 0, 0, 0}, yw[4] = {0, 0, 0, 0};
  if (own) {
    for (int q = 1; q < 4; ++q) {
      const int k = KS - 2 + q;
      pw[q] = rdiv(fullT(at(a.RHOT, k), k), fullD(at(a.DENS, k), k));
      ww[q] = at(a.MOMZ, k); xw[q] = at(a.MOMX, k); yw[q] = at(a.MOMY, k);
    }
  }
  R cmZ = 0, ctZ = 0, cqW = 0, cqU = 0, cqV = 0;  // fluxes through face k-1 (zero at KS-1)
  double rt0 = 0.0, dp0 = 0.0, st0 = 0.0, sr0 = 0.0, d0 = 0.0, rd0 = 0.0, A0 = 0.0, pt0 = 0.0, ptm = 0.0, sw0 = 0.0;
  double e_prev = 0.0, f_prev = 0.0, C_last = 0.0;

  // Every other global load of level k goes into these registers before any
  // LDS write of the level, so that the level waits on memory once: num_diff
  // of the faces this thread evaluates, the staged tiles and the DPRES
  // inputs.  PIPE (FP32): the loads of level k+1 are issued right after the
  // staging of level k, behind the column's own loads, and stay in flight
  // through the face and column work of level k.
  constexpr bool PIPE = DEV;
  constexpr int NX = (TJ * (TI + 1) + NT - 1) / NT, NY = ((TJ + 1) * TI + NT - 1) / NT;
  constexpr int NC = (RW * CW + NT - 1) / NT, NP = ((TJ + 1) * (TI + 1) + NT - 1) / NT;
  R fx[NX][5], fy[NY][5];
  R tD[NC], tX[NC], tY[NC], tW[NC], tR[NC];
  double pD0[NP], pRT[NP], pRR[NP];
  R pR[NP];
  // PIPE: loads at invalid positions go to element 0 and are discarded by a
  // select, so that every load is issued unconditionally, with no branch
  // around it that would hide from the wait counters how many are in flight.
  auto issue = [&](int k) {
#pragma unroll
    for (int h = 0; h < NX; ++h) {
      const int q = tid + h * NT;
      const int row = q / (TI + 1), x = q - row * (TI + 1);
      const int ii = i0 - 1 + x, jj = j0 + row;
      const bool ok = q < TJ * (TI + 1) && ii <= d.IE && jj <= d.JE;
      const size_t g = ok ? gix(k, ii, jj) : 0;
      if constexpr (PIPE) {
        const R v0 = nd(I_DENS, XDIR, g), v1 = nd(I_RHOT, XDIR, g), v2 = nd(I_MOMZ, XDIR, g);
        const R v3 = nd(I_MOMX, XDIR, g + 1), v4 = nd(I_MOMY, XDIR, g);
        fx[h][0] = ok ? v0 : R(0);
        fx[h][1] = ok ? v1 : R(0);
        fx[h][2] = (ok && k < KE) ? v2 : R(0);
        fx[h][3] = ok ? v3 : R(0);
        fx[h][4] = ok ? v4 : R(0);
      } else {
        fx[h][0] = ok ? nd(I_DENS, XDIR, g) : R(0);
        fx[h][1] = ok ? nd(I_RHOT, XDIR, g) : R(0);
        fx[h][2] = (ok && k < KE) ? nd(I_MOMZ, XDIR, g) : R(0);
        fx[h][3] = ok ? nd(I_MOMX, XDIR, g + 1) : R(0);
        fx[h][4] = ok ? nd(I_MOMY, XDIR, g) : R(0);
      }
    }
#pragma unroll
    for (int h = 0; h < NY; ++h) {
      const int q = tid + h * NT;
      const int y = q / TI, col = q - y * TI;
      const int ii = i0 + col, jj = j0 - 1 + y;
      const bool ok = q < (TJ + 1) * TI && ii <= d.IE && jj <= d.JE;
      const size_t g = ok ? gix(k, ii, jj) : 0;
      if constexpr (PIPE) {
        const R v0 = nd(I_DENS, YDIR, g), v1 = nd(I_RHOT, YDIR, g), v2 = nd(I_MOMZ, YDIR, g);
        const R v3 = nd(I_MOMX, YDIR, g), v4 = nd(I_MOMY, YDIR, g + d.IA);
        fy[h][0] = ok ? v0 : R(0);
        fy[h][1] = ok ? v1 : R(0);
        fy[h][2] = (ok && k < KE) ? v2 : R(0);
        fy[h][3] = ok ? v3 : R(0);
        fy[h][4] = ok ? v4 : R(0);
      } else {
        fy[h][0] = ok ? nd(I_DENS, YDIR, g) : R(0);
        fy[h][1] = ok ? nd(I_RHOT, YDIR, g) : R(0);
        fy[h][2] = (ok && k < KE) ? nd(I_MOMZ, YDIR, g) : R(0);
        fy[h][3] = ok ? nd(I_MOMX, YDIR, g) : R(0);
        fy[h][4] = ok ? nd(I_MOMY, YDIR, g + d.IA) : R(0);
      }
    }
#pragma unroll
    for (int h = 0; h < NC; ++h) {
      const int q = tid + h * NT;
      const int r = q / CW, c = q - r * CW;
      const int ii = i0 - 2 + c, jj = j0 - 2 + r;
      const bool ok = q < RW * CW && ii < d.IA && jj < d.JA && k + 2 < d.KA;
      const size_t g = ok ? gix(k + 2, ii, jj) : 0;  // the ring's new level
      if constexpr (PIPE) {
        const R v0 = a.DENS[g], v1 = a.MOMX[g], v2 = a.MOMY[g], v3 = a.MOMZ[g], v4 = a.RHOT[g];
        tD[h] = ok ? v0 : R(0);
        tX[h] = ok ? v1 : R(0);
        tY[h] = ok ? v2 : R(0);
        tW[h] = ok ? v3 : R(0);
        tR[h] = ok ? v4 : R(0);
      } else {
        tD[h] = ok ? a.DENS[g] : R(0);
        tX[h] = ok ? a.MOMX[g] : R(0);
        tY[h] = ok ? a.MOMY[g] : R(0);
        tW[h] = ok ? a.MOMZ[g] : R(0);
        tR[h] = ok ? a.RHOT[g] : R(0);
      }
    }
#pragma unroll
    for (int h = 0; h < NP; ++h) {
      const int q = tid + h * NT;
      const int r = q / (TI + 1), c = q - r * (TI + 1);
      const int ii = i0 + c, jj = j0 + r;
      const bool ok = q < (TJ + 1) * (TI + 1) && ii < d.IA && jj < d.JA;
      const size_t g = ok ? gix(k, ii, jj) : 0;
      if constexpr (PIPE) {
        const double v0 = a.DPRES0[g], v1 = a.RT2P[g], v2 = a.REF_rhot[g];
        const R v3 = a.RHOT[g];
        pD0[h] = ok ? v0 : 0.0;
        pRT[h] = ok ? v1 : 0.0;
        pRR[h] = ok ? v2 : 0.0;
        pR[h] = ok ? v3 : R(0);
      } else {
        pD0[h] = ok ? a.DPRES0[g] : 0.0;
        pRT[h] = ok ? a.RT2P[g] : 0.0;
        pRR[h] = ok ? a.REF_rhot[g] : 0.0;
        pR[h] = ok ? a.RHOT[g] : R(0);
      }
    }
  };
  if constexpr (PIPE) issue(KS);

  for (int k = KS; k <= KE; ++k) {
    const int s0 = k % 3, s1 = (k + 1) % 3, s2 = (k + 2) % 3;  // ring slots of levels k, k+1, k+2
    if constexpr (!PIPE) issue(k);
    // ---- (a) staging ----
#pragma unroll
    for (int h = 0; h < NX; ++h) {
      const int q = tid + h * NT;
      if (q >= TJ * (TI + 1)) continue;
      const int row = q / (TI + 1), x = q - row * (TI + 1);
      xM[row][x] = fx[h][0]; xT[row][x] = fx[h][1]; xW[row][x] = fx[h][2]; xU[row][x] = fx[h][3]; xV[row][x] = fx[h][4];
    }
#pragma unroll
    for (int h = 0; h < NY; ++h) {
      const int q = tid + h * NT;
      if (q >= (TJ + 1) * TI) continue;
      const int y = q / TI, col = q - y * TI;
      yM[y][col] = fy[h][0]; yT[y][col] = fy[h][1]; yW[y][col] = fy[h][2]; yU[y][col] = fy[h][3]; yV[y][col] = fy[h][4];
    }
#pragma unroll
    for (int h = 0; h < NC; ++h) {
      const int q = tid + h * NT;
      if (q >= RW * CW) continue;
      const int r = q / CW, c = q - r * CW;
      const int ii = i0 - 2 + c, jj = j0 - 2 + r;
      const bool in = ii < d.IA && jj < d.JA;
      sD[s2][r][c] = in ? fullD(tD[h], k + 2) : tD[h];
      sMX[s2][r][c] = tX[h]; sMY[s2][r][c] = tY[h]; sW[s2][r][c] = tW[h];
      sPO[s2][r][c] = in ? rdiv(fullT(tR[h], k + 2), fullD(tD[h], k + 2)) : R(0);
    }
#pragma unroll
    for (int h = 0; h < NP; ++h) {
      const int q = tid + h * NT;
      if (q >= (TJ + 1) * (TI + 1)) continue;
      const int r = q / (TI + 1), c = q - r * (TI + 1);
      sDP[r][c] = pD0[h] + pRT[h] * (fullTd(pR[h], k) - pRR[h]);
    }
    __syncthreads();
    // ---- the column's own loads of this level, in flight during the face
    // evaluation ----
    // (PIPE: issued by every thread, at the first interior column when the
    // thread has none, and the next level's prefetch likewise, so that no
    // branch hides from the wait counters how many loads are in flight)
    const size_t gk = own ? gix(k, i, j) : (PIPE ? gix(k, d.IS, d.JS) : 0);
    R o_nZD = 0, o_nZT = 0, o_nZW = 0, o_nZX = 0, o_nZY = 0;
    R o_Dt = 0, o_Rt = 0, o_Wt = 0, o_Xt = 0, o_Yt = 0, o_X0 = 0, o_Y0 = 0, o_Z0 = 0;
    double o_RT = 0.0, o_RD = 0.0;
    R o_D = 1;
    const double wdc = zWD(k);
    R o_nDw = 0;  // mflx_hi(Z) of the i = IS-1 column
    if (own || PIPE) {
      o_nZD = nd(I_DENS, ZDIR, gk); o_nZT = nd(I_RHOT, ZDIR, gk); o_nZW = nd(I_MOMZ, ZDIR, gk + d.pl);
      o_nZX = nd(I_MOMX, ZDIR, gk); o_nZY = nd(I_MOMY, ZDIR, gk);
      o_Dt = a.DENS_t[gk]; o_Rt = a.RHOT_t[gk]; o_Wt = a.MOMZ_t[gk]; o_Xt = a.MOMX_t[gk]; o_Yt = a.MOMY_t[gk];
      o_X0 = a.MOMX0[gk]; o_Y0 = a.MOMY0[gk]; o_RT = a.RT2P[gk]; o_RD = a.REF_dens[gk];
      if constexpr (PIPE) {
        const R z0 = a.MOMZ0[gk];
        o_Z0 = wdc != 0.0 ? z0 : R(0);
      } else {
        if (wdc != 0.0) o_Z0 = a.MOMZ0[gk];
      }
      if constexpr (DEV) o_D = a.DENS[gk];
      if constexpr (PIPE) {
        const R w = nd(I_DENS, ZDIR, gk - 1);
        o_nDw = (i == d.IS && k <= KE - 1) ? w : R(0);
      } else {
        if (i == d.IS && k <= KE - 1) o_nDw = nd(I_DENS, ZDIR, gk - 1);
      }
    }
    if constexpr (PIPE) {
      // the next level's loads go out after the own loads (the wait for those
      // then leaves the prefetch in flight), at KE too (level KE+1 exists and
      // its values go unused; the ring's level KE+3 does not and reads element
      // 0 instead): no scheduling across this point
      __builtin_amdgcn_sched_barrier(0);
      issue(k + 1);
      __builtin_amdgcn_sched_barrier(0);
    }
    // ---- (b) faces ----
    const double cz0 = zCDZ(k), cz1 = zCDZ(k + 1);
    const R f1 = R(cz0 / (cz0 + cz1));  // F2H(k,1)
    const R f2 = R(cz1 / (cz0 + cz1));  // F2H(k,2)
    auto D0 = [&](int r, int c) { return sD[s0][r][c]; };
    auto D1 = [&](int r, int c) { return sD[s1][r][c]; };
    for (int q = tid; q < TJ * (TI + 1); q += NT) {
      const int row = q / (TI + 1), x = q - row * (TI + 1);
      const int ii = i0 - 1 + x, jj = j0 + row;
      if (ii > d.IE || jj > d.JE) continue;
      const R ndq[5] = {xM[row][x], xT[row][x], xW[row][x], xU[row][x], xV[row][x]};
      const int r = row + 2, c = x + 1;  // tile cell of column ii
      const size_t g = gix(k, ii, jj);
      const R vm = sMX[s0][r][c] + ndq[0];
      xM[row][x] = vm;
      xT[row][x] = vm * fv_int<S>(sPO[s0][r][c - 1], sPO[s0][r][c], sPO[s0][r][c + 1], sPO[s0][r][c + 2], vm) + ndq[1];
      R qw = 0;
      if (k < KE) {
        const R vel = rdiv((f1 * sMX[s1][r][c] + f2 * sMX[s0][r][c]),
                           (f1 * R(0.5) * (D1(r, c) + D1(r, c + 1)) + f2 * R(0.5) * (D0(r, c) + D0(r, c + 1))));
        qw = vel * fv_int<S>(sW[s0][r][c - 1], sW[s0][r][c], sW[s0][r][c + 1], sW[s0][r][c + 2], vel) + ndq[2];
      }
      xW[row][x] = qw;
      {  // fluxX_UYZ stored at ii (loop index ii+1)
        const R vel = rdiv((R(0.5) * (sMX[s0][r][c + 1] + sMX[s0][r][c])), D0(r, c + 1));
        xU[row][x] = vel * fv_int<S>(sMX[s0][r][c - 1], sMX[s0][r][c], sMX[s0][r][c + 1], sMX[s0][r][c + 2], vel)
                   + ndq[3];
      }
      {  // fluxX_XVZ at (ii, jj)
        const R vel = rdiv((R(0.5) * (sMX[s0][r][c] + sMX[s0][r + 1][c])),
                      (R(0.25) * (D0(r, c) + D0(r, c + 1) + D0(r + 1, c) + D0(r + 1, c + 1))));
        xV[row][x] = vel * fv_int<S>(sMY[s0][r][c - 1], sMY[s0][r][c], sMY[s0][r][c + 1], sMY[s0][r][c + 2], vel)
                   + ndq[4];
      }
      if constexpr (CMB) {  // averaged over the stages at the interior faces this tile owns
        if (x > 0) {
          a.mflx_av[g + n3] = rk4f(a.cm[0][g + n3], a.cm[1][g + n3], a.cm[2][g + n3], vm);
          a.tflx_av[g + n3] = rk4f(a.ct[0][g + n3], a.ct[1][g + n3], a.ct[2][g + n3], xT[row][x]);
        }
      } else if (x > 0 || i0 == d.IS) {  // faces IS-1 and those owned by this tile
        a.mflx_hi[g + n3] = vm;
        a.tflx_hi[g + n3] = xT[row][x];
      }
    }
    for (int q = tid; q < (TJ + 1) * TI; q += NT) {
      const int y = q / TI, col = q - y * TI;
      const int ii = i0 + col, jj = j0 - 1 + y;
      if (ii > d.IE || jj > d.JE) continue;
      const R ndq[5] = {yM[y][col], yT[y][col], yW[y][col], yU[y][col], yV[y][col]};
      const int r = y + 1, c = col + 2;  // tile cell of row jj
      const size_t g = gix(k, ii, jj);
      const R vm = sMY[s0][r][c] + ndq[0];
      yM[y][col] = vm;
      yT[y][col] = vm * fv_int<S>(sPO[s0][r - 1][c], sPO[s0][r][c], sPO[s0][r + 1][c], sPO[s0][r + 2][c], vm) + ndq[1];
      R qw = 0;
      if (k < KE) {
        const R vel = rdiv((f1 * sMY[s1][r][c] + f2 * sMY[s0][r][c]),
                           (f1 * R(0.5) * (D1(r, c) + D1(r + 1, c)) + f2 * R(0.5) * (D0(r, c) + D0(r + 1, c))));
        qw = vel * fv_int<S>(sW[s0][r - 1][c], sW[s0][r][c], sW[s0][r + 1][c], sW[s0][r + 2][c], vel) + ndq[2];
      }
      yW[y][col] = qw;
      {  // fluxY_UYZ at (ii, jj)
        const R vel = rdiv((R(0.5) * (sMY[s0][r][c] + sMY[s0][r][c + 1])),
                      (R(0.25) * (D0(r, c) + D0(r, c + 1) + D0(r + 1, c) + D0(r + 1, c + 1))));
        yU[y][col] = vel * fv_int<S>(sMX[s0][r - 1][c], sMX[s0][r][c], sMX[s0][r + 1][c], sMX[s0][r + 2][c], vel)
                   + ndq[3];
      }
      {  // fluxY_XVZ stored at jj (loop index jj+1)
        const R vel = rdiv((R(0.5) * (sMY[s0][r + 1][c] + sMY[s0][r][c])), D0(r + 1, c));
        yV[y][col] = vel * fv_int<S>(sMY[s0][r - 1][c], sMY[s0][r][c], sMY[s0][r + 1][c], sMY[s0][r + 2][c], vel)
                   + ndq[4];
      }
      if constexpr (CMB) {
        if (y > 0) {
          a.mflx_av[g + 2 * n3] = rk4f(a.cm[0][g + 2 * n3], a.cm[1][g + 2 * n3], a.cm[2][g + 2 * n3], vm);
          a.tflx_av[g + 2 * n3] = rk4f(a.ct[0][g + 2 * n3], a.ct[1][g + 2 * n3], a.ct[2][g + 2 * n3], yT[y][col]);
        }
      } else if (y > 0 || j0 == d.JS) {
        a.mflx_hi[g + 2 * n3] = vm;
        a.tflx_hi[g + 2 * n3] = yT[y][col];
      }
    }
    __syncthreads();

    // own-column windows move up one level, level k+2 from the ring
    if (own) {
      const int r = ty + 2, c = tx + 2;
      pw[0] = pw[1]; pw[1] = pw[2]; pw[2] = pw[3]; pw[3] = sPO[s2][r][c];
      ww[0] = ww[1]; ww[1] = ww[2]; ww[2] = ww[3]; ww[3] = sW[s2][r][c];
      xw[0] = xw[1]; xw[1] = xw[2]; xw[2] = xw[3]; xw[3] = sMX[s2][r][c];
      yw[0] = yw[1]; yw[1] = yw[2]; yw[2] = yw[3]; yw[3] = sMY[s2][r][c];
    }

    // ---- (c) the column ----
    if (own) {
      const int r = ty + 2, c = tx + 2;  // tile cell of the own column
      const size_t g = gix(k, i, j);
      const R rcdz = R(zRCDZ(k)), rcdx = R(hRCDX), rcdy = R(hRCDY), rfdx = R(hRFDX), rfdy = R(hRFDY);
      // continuity and theta (HEVI.F90:369-475, 588-657)
      const R mZ = (k <= KE - 1) ? o_nZD : R(0);
      const R Sr = (-(mZ - cmZ) * rcdz + -((xM[ty][tx + 1] - xM[ty][tx]) * rcdx + (yM[ty + 1][tx] - yM[ty][tx]) * rcdy))
                 + o_Dt;
      R tZ = 0;
      if (k <= KE - 1) {
        R fv;
        if (k == KS) fv = fv_lo<S>(pw[1], pw[2], pw[3], mZ);
        else if (k == KE - 1) fv = fv_hi<S>(pw[0], pw[1], pw[2], mZ);
        else fv = fv_int<S>(pw[0], pw[1], pw[2], pw[3], mZ);
        tZ = mZ * fv + o_nZT;
      }
      const R St = (-(tZ - ctZ) * rcdz + -((xT[ty][tx + 1] - xT[ty][tx]) * rcdx + (yT[ty + 1][tx] - yT[ty][tx]) * rcdy))
                 + o_Rt;
      // momentum z, explicit part (HEVI.F90:486-577) and PT
      R Sw = 0, PT = 0;
      if (k <= KE - 1) {
        R qw = 0;
        if (k < KE - 1) {
          const R vel = rdiv((R(0.5) * (ww[1] + ww[2])), D1(r, c));
          const R fv = (k == KS) ? fv_lo<S>(ww[1], ww[2], ww[3], vel) : fv_int<S>(ww[0], ww[1], ww[2], ww[3], vel);
          qw = R(J33G) * vel * fv + o_nZW;
        }
        const R advcv = -(qw - cqW) * R(zRFDZ(k));
        cqW = qw;
        const R advch = -((xW[ty][tx + 1] - xW[ty][tx]) * rcdx + (yW[ty + 1][tx] - yW[ty][tx]) * rcdy);
        const R wdmp = -R(wdc) * o_Z0;
        R sw = (advcv + advch) + wdmp;
        if constexpr (DIVDMP) sw = sw + R(a.divdmp_coef / dtrk) * (a.DDIV[g + d.pl] - a.DDIV[g]) * R(a.g.FDZ[k]);
        Sw = sw + o_Wt;
        const R vel = ww[1];
        if (k == KS) PT = fv_lo<S>(pw[1], pw[2], pw[3], vel);
        else if (k == KE - 1) PT = fv_hi<S>(pw[0], pw[1], pw[2], vel);
        else PT = fv_int<S>(pw[0], pw[1], pw[2], pw[3], vel);
      }
      // momentum x (HEVI.F90:816-927); the pressure gradient in double
      R MXn;
      {
        R qu = 0;
        if (k < KE) {
          const R vel = rdiv((R(0.5) * (ww[1] + sW[s0][r][c + 1])),
                        (f1 * R(0.5) * (D1(r, c) + D1(r, c + 1)) + f2 * R(0.5) * (D0(r, c) + D0(r, c + 1))));
          R fv;
          if (k == KS) fv = fv_lo<S>(xw[1], xw[2], xw[3], vel);
          else if (k == KE - 1) fv = fv_hi<S>(xw[0], xw[1], xw[2], vel);
          else fv = fv_int<S>(xw[0], xw[1], xw[2], xw[3], vel);
          qu = R(J33G) * vel * fv + o_nZX;
        }
        const R advcv = -(qu - cqU) * rcdz;
        cqU = qu;
        const R advch = -((xU[ty][tx + 1] - xU[ty][tx]) * rfdx + (yU[ty + 1][tx] - yU[ty][tx]) * rcdy);
        const R pg = R((sDP[ty][tx + 1] - sDP[ty][tx]) * hRFDX);
        R t = advcv + advch - pg;
        if (a.has_coriolis) {
          const R cf = R(corx)
                     * (sMY[s0][r][c + 1] + sMY[s0][r][c] + sMY[s0][r - 1][c + 1] + sMY[s0][r - 1][c]);
          t = t + cf;
        }
        if constexpr (DIVDMP) t = t + R(a.divdmp_coef / dtrk) * (a.DDIV[g + 1] - a.DDIV[g]) * R(a.g.FDX[i]);
        MXn = o_X0 + R(dtrk) * (t + o_Xt);
      }
      // momentum y (HEVI.F90:936-1047)
      R MYn;
      {
        R qv = 0;
        if (k < KE) {
          const R vel = rdiv((R(0.5) * (ww[1] + sW[s0][r + 1][c])),
                        (f1 * R(0.5) * (D1(r, c) + D1(r + 1, c)) + f2 * R(0.5) * (D0(r, c) + D0(r + 1, c))));
          R fv;
          if (k == KS) fv = fv_lo<S>(yw[1], yw[2], yw[3], vel);
          else if (k == KE - 1) fv = fv_hi<S>(yw[0], yw[1], yw[2], vel);
          else fv = fv_int<S>(yw[0], yw[1], yw[2], yw[3], vel);
          qv = R(J33G) * vel * fv + o_nZY;
        }
        const R advcv = -(qv - cqV) * rcdz;
        cqV = qv;
        const R advch = -((xV[ty][tx + 1] - xV[ty][tx]) * rcdx + (yV[ty + 1][tx] - yV[ty][tx]) * rfdy);
        const R pg = R((sDP[ty + 1][tx] - sDP[ty][tx]) * hRFDY);
        R t = advcv + advch - pg;
        if (a.has_coriolis) {
          const R cf = R(cory)
                     * (sMX[s0][r + 1][c] + sMX[s0][r][c] + sMX[s0][r + 1][c - 1] + sMX[s0][r][c - 1]);
          t = t + cf;
        }
        if constexpr (DIVDMP) t = t + R(a.divdmp_coef / dtrk) * (a.DDIV[g + d.IA] - a.DDIV[g]) * R(a.g.FDY[j]);
        MYn = o_Y0 + R(dtrk) * (t + o_Yt);
      }
      // tridiagonal row k-1 and forward sweep in double (HEVI.F90:688-727, 1092-1115)
      double d1;
      if constexpr (DEV) d1 = fullDd(o_D, k);  // the exact stored value, not the FP32 tile copy
      else d1 = (double)D0(r, c);
      const double rt1 = o_RT, dp1 = sDP[ty][tx], rd1 = o_RD;
      const double St1 = (double)St, Sr1 = (double)Sr;
      const double A1 = dtrk * dtrk * J33G * zRCDZ(k) * rt1 * J33G;
      const double pt1 = (k <= KE - 1) ? (double)PT : 0.0;
      double ek = 0.0, fk = 0.0;
      if (k > KS) {
        const int rr = k - 1;
        const double c0 = zCDZ(rr), c1 = zCDZ(rr + 1), rf = zRFDZ(rr);
        const double B = GRAV * (dtrk * dtrk) * J33G / (c1 + c0);
        const double F1v = -(pt1 * rf * A1 + B);
        const double F2v = 1.0 + (pt0 * rf * (A1 + A0));
        const double F3v = -(ptm * rf * A0 - B);
        const double h1 = c0 / (c0 + c1), h2 = c1 / (c0 + c1);
        const double pg = -(dp1 + rt1 * dtrk * St1 - dp0 - rt0 * dtrk * st0) * rf * J33G
                        - GRAV * (h1 * (d1 - rd1 + Sr1 * dtrk) + h2 * (d0 - rd0 + sr0 * dtrk));
        const double Ck = (double)ww[0] + dtrk * (pg + sw0);
        if (rr == KS) {
          const double rdenom = 1.0 / F2v;
          ek = -F1v * rdenom;
          fk = Ck * rdenom;
        } else if (rr < KE - 1) {
          const double rdenom = 1.0 / (F2v + F3v * e_prev);
          ek = -F1v * rdenom;
          fk = (Ck - F3v * f_prev) * rdenom;
        } else {
          C_last = (Ck - F3v * f_prev) / (F2v + F3v * e_prev);
        }
        e_prev = ek; f_prev = fk;
        ptm = pt0;
      }
      // stores of this level
      if constexpr (CMB) {
        const R mx = rk4c(a.c1[2][g], a.c2[2][g], a.c3[2][g], MXn, o_X0);
        const R my = rk4c(a.c1[3][g], a.c2[3][g], a.c3[3][g], MYn, o_Y0);
        a.st[2][g] = mx;
        a.st[3][g] = my;
        if (k == KS || k == KE) {  // ghost levels copy the combined KS / KE value
          const int lo = (k == KS) ? 0 : KE + 1, hi = (k == KS) ? KS - 1 : d.KA - 1;
          for (int kk = lo; kk <= hi; ++kk) { a.st[2][p0 + d.pl * (size_t)kk] = mx; a.st[3][p0 + d.pl * (size_t)kk] = my; }
        }
      } else {
        a.MOMX_RK[g] = MXn;
        a.MOMY_RK[g] = MYn;
      }
      a.DENS_RK[g] = Sr;  // Sr/St wait in the outputs until the back pass
      a.RHOT_RK[g] = St;
      if constexpr (sizeof(R) == sizeof(double)) {
        if (a.trSt) a.trSt[g] = St;
      }
      // the explicit z mass flux is num_diff itself: the back pass reads it
      // there again instead of a copy stored here
      if (k <= KE - 1) { a.tflx_hi[g] = tZ; a.PT[g] = PT; }
      else { a.mflx_hi[g] = R(0); a.tflx_hi[g] = R(0); }
      if (k == KS) { a.mflx_hi[g - d.pl] = R(0); a.tflx_hi[g - d.pl] = R(0); }
      if (i == d.IS) {  // mflx_hi(Z) on the i = IS-1 column (HEVI.F90:370)
        a.mflx_hi[g - 1] = (k <= KE - 1) ? o_nDw : R(0);
        if (k == KS) a.mflx_hi[g - 1 - d.pl] = R(0);
      }
      if (k > KS && k - 1 < KE - 1) { a.e[g - d.pl] = R(ek); a.f[g - d.pl] = R(fk); }
      cmZ = mZ; ctZ = tZ;
      rt0 = rt1; dp0 = dp1; st0 = St1; sr0 = Sr1; d0 = d1; rd0 = rd1; A0 = A1; pt0 = pt1; sw0 = (double)Sw;
    }
    __syncthreads();
  }
  if (!own) return;

  // ---- downward: back substitution and updates in double (HEVI.F90:728-790) ----
  R* mfZ = a.mflx_hi; R* tfZ = a.tflx_hi;
  auto px = [&](int k) { return p0 + d.pl * (size_t)k; };
  // final value of level q: stored as the stage output, or combined with
  // stages 1-3 into the state (plus the ghost levels at KS and KE)
  auto put_state = [&](int v, size_t q, R val, R s0v, int k) {
    if constexpr (CMB) {
      const R c = rk4c(a.c1[v][q], a.c2[v][q], a.c3[v][q], val, s0v);
      a.st[v][q] = c;
      if (k == KS || k == KE) {
        const int lo = (k == KS) ? 0 : KE + 1, hi = (k == KS) ? KS - 1 : d.KA - 1;
        for (int kk = lo; kk <= hi; ++kk) {
          R gv = (v == 1) ? R(0) : c;
          if constexpr (DEV) {  // the ghost level holds the same full value, its own deviation
            if (v == 0) gv = devD(fullDd(c, k), kk);
            if (v == 4) gv = devT(fullTd(c, k), kk);
          }
          a.st[v][p0 + d.pl * (size_t)kk] = gv;
        }
      }
    } else {
      R* o = v == 0 ? a.DENS_RK : (v == 1 ? a.MOMZ_RK : a.RHOT_RK);
      o[q] = val;
    }
  };
  auto put_zflux = [&](size_t q, R mv, R tv) {
    if constexpr (CMB) {
      a.mflx_av[q] = rk4f(a.cm[0][q], a.cm[1][q], a.cm[2][q], mv);
      a.tflx_av[q] = rk4f(a.ct[0][q], a.ct[1][q], a.ct[2][q], tv);
    } else {
      mfZ[q] = mv; tfZ[q] = tv;
    }
  };
  double Cn = C_last;
  double ptn = (double)a.PT[px(KE - 1)];
  {
    const size_t q = px(KE);
    const R d0 = a.DENS0[q], r0 = a.RHOT0[q];
    double advcv = Cn * J33G * RCDZ[KE];
    const R dn = devD(fullDd(d0, KE) + dtrk * (advcv + (double)a.DENS_RK[q]), KE);
    advcv = Cn * ptn * J33G * RCDZ[KE];
    const R rn = devT(fullTd(r0, KE) + dtrk * (advcv + (double)a.RHOT_RK[q]), KE);
    put_state(0, q, dn, d0, KE);
    put_state(4, q, rn, r0, KE);
    if constexpr (CMB) {
      a.st[1][q] = R(0);  // MOMZ(KE) keeps the last stage's value, 0
      for (int kk = KE + 1; kk < d.KA; ++kk) a.st[1][p0 + d.pl * (size_t)kk] = R(0);
      put_zflux(q, mfZ[q], tfZ[q]);
    } else {
      a.MOMZ_RK[q] = R(0);
    }
  }
  {
    const size_t q = px(KE - 1);
    const R z0 = a.MOMZ0[q];
    put_zflux(q, R((double)nd(I_DENS, ZDIR, q) + J33G * Cn), R((double)tfZ[q] + J33G * Cn * ptn));
    put_state(1, q, R((double)z0 + (Cn - (double)a.MOMZ[q])), z0, KE - 1);
    if constexpr (sizeof(R) == sizeof(double)) {
      if (a.trCp) a.trCp[q] = Cn;
    }
  }
  // Levels KE-2 .. KS: each level's loads (its own and the level above's
  // DENS0/RHOT0/Sr/St, with the combination its stages 1-3) go in chunks of
  // DCH levels, the next chunk's issued (clamped to KS, unconditionally)
  // before the current chunk's work: one wait per chunk instead of one per
  // level.  A prefetched level is written only after its loads, by the
  // level that consumes them.
  // (measured at 512x512x128: the mixed short step 13.67 -> 13.00 ms, the
  // FP64 one unchanged; 8 / 4 levels in FP64 spill and run 27 % slower)
  constexpr int DCH = CMB ? 2 : 4;
  struct DnV {
    R e, f, pt, z0, ndz, tf, mz, d0, r0, drk, rrk;
    R cz[3], cd[3], cr[3], cm[3], ct[3];  // the combination: stages 1-3 of MOMZ(k), DENS/RHOT(k+1), the z fluxes
  };
  auto dload = [&](int k, DnV& L) {
    const size_t q = px(max(k, KS)), q1 = q + d.pl;
    L.e = a.e[q]; L.f = a.f[q]; L.pt = a.PT[q]; L.z0 = a.MOMZ0[q]; L.ndz = nd(I_DENS, ZDIR, q);
    L.tf = tfZ[q]; L.mz = a.MOMZ[q];
    L.d0 = a.DENS0[q1]; L.r0 = a.RHOT0[q1]; L.drk = a.DENS_RK[q1]; L.rrk = a.RHOT_RK[q1];
    if constexpr (CMB) {
      const R* const* cs[3] = {a.c1, a.c2, a.c3};
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        L.cz[n] = cs[n][1][q]; L.cd[n] = cs[n][0][q1]; L.cr[n] = cs[n][4][q1];
        L.cm[n] = a.cm[n][q]; L.ct[n] = a.ct[n][q];
      }
    }
  };
  // put_state / put_zflux on prefetched stage values
  auto put_state_p = [&](int v, size_t q, R val, R s0v, int k, const R* cv) {
    if constexpr (CMB) {
      const R c = rk4c(cv[0], cv[1], cv[2], val, s0v);
      a.st[v][q] = c;
      if (k == KS) {
        for (int kk = 0; kk <= KS - 1; ++kk) {
          R gv = (v == 1) ? R(0) : c;
          if constexpr (DEV) {
            if (v == 0) gv = devD(fullDd(c, k), kk);
            if (v == 4) gv = devT(fullTd(c, k), kk);
          }
          a.st[v][p0 + d.pl * (size_t)kk] = gv;
        }
      }
    } else {
      R* o = v == 0 ? a.DENS_RK : (v == 1 ? a.MOMZ_RK : a.RHOT_RK);
      o[q] = val;
    }
  };
  auto put_zflux_p = [&](size_t q, R mv, R tv, const DnV& L) {
    if constexpr (CMB) {
      a.mflx_av[q] = rk4f(L.cm[0], L.cm[1], L.cm[2], mv);
      a.tflx_av[q] = rk4f(L.ct[0], L.ct[1], L.ct[2], tv);
    } else {
      mfZ[q] = mv; tfZ[q] = tv;
    }
  };
  DnV cur[DCH], nxt[DCH];
#pragma unroll
  for (int c = 0; c < DCH; ++c) dload(KE - 2 - c, cur[c]);
  for (int kc = KE - 2; kc >= KS; kc -= DCH) {
    if (kc - DCH >= KS) {
#pragma unroll
      for (int c = 0; c < DCH; ++c) dload(kc - DCH - c, nxt[c]);
    }
#pragma unroll
    for (int c = 0; c < DCH; ++c) {
      const int k = kc - c;
      if (k < KS) break;
      const DnV& L = cur[c];
      const size_t q = px(k), q1 = px(k + 1);
      const double Ck = (double)L.e * Cn + (double)L.f;
      const double ptk = (double)L.pt;
      const R z0 = L.z0;
      put_zflux_p(q, R((double)L.ndz + J33G * Ck), R((double)L.tf + J33G * Ck * ptk), L);
      put_state_p(1, q, R((double)z0 + (Ck - (double)L.mz)), z0, k, L.cz);
      if constexpr (sizeof(R) == sizeof(double)) {
        if (a.trCp) a.trCp[q] = Ck;
      }
      const R d0 = L.d0, r0 = L.r0;
      double advcv = -(Cn - Ck) * J33G * RCDZ[k + 1];
      put_state_p(0, q1, devD(fullDd(d0, k + 1) + dtrk * (advcv + (double)L.drk), k + 1), d0, k + 1, L.cd);
      advcv = -(Cn * ptn - Ck * ptk) * J33G * RCDZ[k + 1];
      put_state_p(4, q1, devT(fullTd(r0, k + 1) + dtrk * (advcv + (double)L.rrk), k + 1), r0, k + 1, L.cr);
      Cn = Ck; ptn = ptk;
    }
#pragma unroll
    for (int c = 0; c < DCH; ++c) cur[c] = nxt[c];
  }
  {
    const size_t q = px(KS);
    const R d0 = a.DENS0[q], r0 = a.RHOT0[q];
    double advcv = -Cn * J33G * RCDZ[KS];
    put_state(0, q, devD(fullDd(d0, KS) + dtrk * (advcv + (double)a.DENS_RK[q]), KS), d0, KS);
    advcv = -Cn * ptn * J33G * RCDZ[KS];
    put_state(4, q, devT(fullTd(r0, KS) + dtrk * (advcv + (double)a.RHOT_RK[q]), KS), r0, KS);
    if constexpr (!CMB) a.MOMZ_RK[px(KS - 1)] = R(0);
  }
}

template <int S, bool DIVDMP, int TJ, typename R, bool CMB>
__global__ void __launch_bounds__(64 * TJ) hevi_tile(HeviArgsT<R> a, unsigned ni_tiles, unsigned nblocks) {
  hevi_tile_body<S, DIVDMP, TJ, R, CMB>(a, xcd_block(blockIdx.x, nblocks), ni_tiles);
}

// The stages of many members (an ensemble of grids of any sizes) in one
// launch: tab[m] are member m's arguments, its tiles are the logical tiles
// first[m] .. first[m+1]-1 of total, ni[m] tiles wide.  The blocks walk the
// tiles with a stride of the grid size (persistent when the grid is smaller
// than total); a grid size that is a multiple of 8 keeps each XCD on one
// contiguous range of logical tiles.
template <int S, bool DIVDMP, int TJ, typename R, bool CMB>
__global__ void __launch_bounds__(64 * TJ) hevi_tile_batch(const HeviArgsT<R>* __restrict__ tab,
                                                         const unsigned* __restrict__ first,
                                                         const unsigned* __restrict__ ni, int nmem, unsigned total) {
  for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
    const unsigned lt = xcd_block(t, total);
    const int lo = batch_member(first, nmem, lt);
    hevi_tile_body<S, DIVDMP, TJ, R, CMB>(tab[lo], lt - first[lo], ni[lo]);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
template <int S, bool TERRAIN, bool DIVDMP>
static hipError_t launch_t(const HeviArgs& a, hipStream_t s, bool cmb) {
  const Dims& d = a.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1;
  const Ctx& c = *a.ctx;
  const unsigned ni = (imax + 63) / 64;
  prof_start(c, "hevi_stage", s);
  if constexpr (!TERRAIN) {
    if (!cmb && !hevi_flat_tile(c)) {
      launch_split<S, false, DIVDMP>(a, s);
      prof_stop(c, s);
      return hipGetLastError();
    }
    // flat terrain: LDS tiles of 64 x 8 columns
    constexpr int TJ = 8;
    const unsigned nb = ni * ((jmax + TJ - 1) / TJ);
    if (cmb) hipLaunchKernelGGL((hevi_tile<S, DIVDMP, TJ, double, true>), dim3(nb), dim3(64, TJ), 0, s, a, ni, nb);
    else hipLaunchKernelGGL((hevi_tile<S, DIVDMP, TJ, double, false>), dim3(nb), dim3(64, TJ), 0, s, a, ni, nb);
  } else {
    if (cmb) { prof_stop(c, s); return hipErrorNotSupported; }
    launch_split<S, true, DIVDMP>(a, s);
  }
  prof_stop(c, s);
  return hipGetLastError();
}

template <typename R>
static void fill_args(HeviArgsT<R>& a, const StageIOT<R>& io) {
  const Ctx& c = *io.ctx;
  a.ctx = &c;
  a.d = c.d;
  a.g = c.g1;
  a.GSQRT = c.GSQRT; a.J13G = c.J13G; a.J23G = c.J23G; a.MAPF = c.MAPF; a.PHI = c.PHI; a.J33G = c.J33G;
  a.DENS0 = io.s0[0]; a.MOMZ0 = io.s0[1]; a.MOMX0 = io.s0[2]; a.MOMY0 = io.s0[3]; a.RHOT0 = io.s0[4];
  a.DENS = io.s[0]; a.MOMZ = io.s[1]; a.MOMX = io.s[2]; a.MOMY = io.s[3]; a.RHOT = io.s[4];
  a.DENS_t = io.t[0]; a.MOMZ_t = io.t[1]; a.MOMX_t = io.t[2]; a.MOMY_t = io.t[3]; a.RHOT_t = io.t[4];
  a.DPRES0 = io.DPRES0; a.RT2P = io.RT2P; a.CORIOLI = io.CORIOLI; a.num_diff = io.num_diff;
  a.wdamp_coef = io.wdamp_coef; a.DDIV = io.DDIV; a.REF_dens = io.REF_dens; a.REF_rhot = io.REF_rhot;
  a.divdmp_coef = io.divdmp_coef; a.dtrk = io.dtrk; a.has_coriolis = io.has_coriolis;
  a.DENS_RK = io.o[0]; a.MOMZ_RK = io.o[1]; a.MOMX_RK = io.o[2]; a.MOMY_RK = io.o[3]; a.RHOT_RK = io.o[4];
  a.mflx_hi = io.mflx_hi; a.tflx_hi = io.tflx_hi;
  a.PT = (R*)c.scr[3]; a.e = (R*)c.scr[4]; a.f = (R*)c.scr[5]; a.Sw = c.scr[6];
  for (int v = 0; v < 5; ++v) { a.c1[v] = io.c1[v]; a.c2[v] = io.c2[v]; a.c3[v] = io.c3[v]; a.st[v] = io.st[v]; }
  for (int n = 0; n < 3; ++n) { a.cm[n] = io.cm[n]; a.ct[n] = io.ct[n]; }
  a.mflx_av = io.mflx_av; a.tflx_av = io.tflx_av;
  a.offD = io.offD; a.offR = io.offR;
  if constexpr (sizeof(R) == sizeof(double)) {
    if (io.trSt) { a.PT = io.trPT; a.trSt = io.trSt; a.trCp = io.trCp; }
  }
}

// ---- ensemble tables ---------------------------------------------------------
void batch_free(BatchBuf& b) {
  if (b.dev) (void)hipFree(b.dev);
  b = BatchBuf{};
}

template <typename R>
static hipError_t hevi_batch_build_t(const StageIOT<R>* ios, int n, BatchBuf& b) {
  if (n <= 0) return hipErrorInvalidValue;
  std::vector<HeviArgsT<R>> args((size_t)n);
  std::vector<unsigned> first((size_t)n + 1, 0), ni((size_t)n, 0);
  const int scheme = ios[0].ctx->scheme, cmb = ios[0].combine;
  for (int m = 0; m < n; ++m) {
    const StageIOT<R>& io = ios[m];
    const Ctx& c = *io.ctx;
    if (c.terrain || io.divdmp_coef != 0.0 || c.scheme != scheme || io.combine != cmb) return hipErrorNotSupported;
    fill_args(args[(size_t)m], io);
    const Dims& d = c.d;
    ni[(size_t)m] = (unsigned)((d.IE - d.IS + 1 + 63) / 64);
    first[(size_t)m + 1] = first[(size_t)m] + ni[(size_t)m] * (unsigned)((d.JE - d.JS + 1 + 7) / 8);
  }
  hipError_t e = batch_upload(args, first, ni, b);
  b.scheme = scheme; b.cmb = cmb; b.fp32 = sizeof(R) == sizeof(float); b.tj = 8;
  return e;
}
hipError_t hevi_batch_build(const StageIOT<double>* ios, int n, BatchBuf& b) { return hevi_batch_build_t(ios, n, b); }
hipError_t hevi_batch_build(const StageIOT<float>* ios, int n, BatchBuf& b) { return hevi_batch_build_t(ios, n, b); }

template <int S, typename R, bool CMB>
static hipError_t hevi_batch_launch_t(const BatchBuf& b, int bpc, int ncu, hipStream_t s) {
  const HeviArgsT<R>* tab;
  const unsigned *first, *ni;
  batch_tables(b, tab, first, ni);
  hipLaunchKernelGGL((hevi_tile_batch<S, false, 8, R, CMB>), dim3(batch_grid(b.total, bpc, ncu)), dim3(64, 8), 0, s,
                     tab, first, ni, b.nmem, b.total);
  return hipGetLastError();
}
hipError_t hevi_batch_launch(const BatchBuf& b, int bpc, int ncu, hipStream_t s) {
  if (!b.dev) return hipErrorInvalidValue;
  if (b.fp32) {
    if (b.scheme == UD3) return b.cmb ? hevi_batch_launch_t<UD3, float, true>(b, bpc, ncu, s) : hevi_batch_launch_t<UD3, float, false>(b, bpc, ncu, s);
    return b.cmb ? hevi_batch_launch_t<CD4, float, true>(b, bpc, ncu, s) : hevi_batch_launch_t<CD4, float, false>(b, bpc, ncu, s);
  }
  if (b.scheme == UD3) return b.cmb ? hevi_batch_launch_t<UD3, double, true>(b, bpc, ncu, s) : hevi_batch_launch_t<UD3, double, false>(b, bpc, ncu, s);
  return b.cmb ? hevi_batch_launch_t<CD4, double, true>(b, bpc, ncu, s) : hevi_batch_launch_t<CD4, double, false>(b, bpc, ncu, s);
}

hipError_t hevi_stage(const StageIO& io, hipStream_t s) {
  HeviArgs a;
  fill_args(a, io);
  const Ctx& c = *io.ctx;
  const bool div = io.divdmp_coef != 0.0;
  const bool ter = c.terrain;
  const bool cmb = io.combine != 0;
  if (c.scheme == UD3) {
    if (ter) return div ? launch_t<UD3, true, true>(a, s, cmb) : launch_t<UD3, true, false>(a, s, cmb);
    return div ? launch_t<UD3, false, true>(a, s, cmb) : launch_t<UD3, false, false>(a, s, cmb);
  }
  if (ter) return div ? launch_t<CD4, true, true>(a, s, cmb) : launch_t<CD4, true, false>(a, s, cmb);
  return div ? launch_t<CD4, false, true>(a, s, cmb) : launch_t<CD4, false, false>(a, s, cmb);
}

template <int S, bool TERRAIN, bool DIVDMP>
static hipError_t launch_heve(const HeviArgs& a, hipStream_t s) {
  const Dims& d = a.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1;
  constexpr unsigned TJ = 4;
  const unsigned ni = (imax + 63) / 64, nb = ni * ((jmax + TJ - 1) / TJ);
  prof_start(*a.ctx, "heve_stage", s);
  hipLaunchKernelGGL((heve_march<S, TERRAIN, DIVDMP>), dim3(nb), dim3(64, TJ), 0, s, a, ni, nb);
  prof_stop(*a.ctx, s);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// HEVE momentum FCT (FLAG_FCT_MOMENTUM; HEVE.F90:386-450, 716-810, 958-1057,
// 1207-1309).  After the stage kernel, for each momentum component C
// (0 MOMZ, 1 MOMX, 2 MOMY):
//   mfct_vel    its start-of-stage velocity on the interior (HEVE.F90:386-446),
//               followed by the periodic halo fill the exchange performs
//   mfct_flux   the UD1 fluxes of the start-of-stage momentum carried by the
//               current one (scale_atmos_dyn_fvm_flux_ud1.F90) on the interior
//               and the rings the low-order field reads, and the anti-diffusive
//               fluxes hi - lo on the interior; the high-order z flux enters
//               divided by the map factors and joined by its J13/J23 parts
//               (HEVE.F90:752-756, 990-996, 1241-1247)
//   fct_momentum_limit (tracer.hip): low-order field, ratios, limited fluxes
//               and the MOM_RK update.
// Work arrays: Ctx::mfct, MFCT_FIELDS fields of n3.
template <int C>
__global__ void __launch_bounds__(256) mfct_vel(Dims d, const double* M0, const double* D0, double* vel) {
  const int i = d.IS + (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = d.KS + (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = d.JS + (int)blockIdx.z;
  if (i > d.IE || k > d.KE || j > d.JE) return;
  const size_t p = IX(d, k, i, j);
  if constexpr (C == 0) vel[p] = (k <= d.KE - 1) ? 2.0 * M0[p] / (D0[IX(d, k + 1, i, j)] + D0[p]) : 0.0;
  else if constexpr (C == 1) vel[p] = 2.0 * M0[p] / (D0[IX(d, k, i + 1, j)] + D0[p]);
  else vel[p] = 2.0 * M0[p] / (D0[IX(d, k, i, j + 1)] + D0[p]);
}

__device__ __forceinline__ double ud1v(double q0, double q1, double vel) { return fv_int<UD1>(0.0, q0, q1, 0.0, vel); }

// the low-order fluxes: z at stored level kk, x at face ii, y at face jj
template <int S, bool TERRAIN, int C>
__device__ double mlo_z(const Ops<S, TERRAIN, false>& o, const HeviArgs& a, int kk, int i, int j) {
  const Dims& d = a.d;
  const double* W = a.MOMZ;
  if constexpr (C == 0) {  // fluxZ_XYW_ud1 (UD1.F90:299-375)
    if (kk == d.KS - 1 || kk >= d.KE - 1) return 0.0;
    const double* V = a.MOMZ0;
    const int k = kk + 1;
    const double vel = (0.5 * (W[IX(d, k - 1, i, j)] + W[IX(d, k, i, j)])) / o.D(k, i, j);
    return a.J33G * vel * ud1v(V[IX(d, kk, i, j)], V[IX(d, k, i, j)], vel);
  } else {  // fluxZ_UYZ_ud1 / fluxZ_XVZ_ud1 (UD1.F90:664-737, 979-1052)
    if (kk == d.KS - 1 || kk == d.KE) return 0.0;
    const double* V = C == 1 ? a.MOMX0 : a.MOMY0;
    const int di = C == 1 ? 1 : 0, dj = C == 1 ? 0 : 1;
    const int c = C == 1 ? I_UYW : I_XVZ;  // the GSQRT component the reference passes
    const double vel = (0.5 * (W[IX(d, kk, i, j)] + W[IX(d, kk, i + di, j + dj)]))
                     / (o.F2H(kk, 1, c, i, j) * 0.5 * (o.D(kk + 1, i, j) + o.D(kk + 1, i + di, j + dj))
                      + o.F2H(kk, 2, c, i, j) * 0.5 * (o.D(kk, i, j) + o.D(kk, i + di, j + dj)));
    return a.J33G * vel * ud1v(V[IX(d, kk, i, j)], V[IX(d, kk + 1, i, j)], vel);
  }
}

template <int S, bool TERRAIN, int C>
__device__ double mlo_h(const Ops<S, TERRAIN, false>& o, const HeviArgs& a, int dir, int k, int i, int j) {
  const Dims& d = a.d;
  const auto& m = o.m;
  if constexpr (C == 0) {  // fluxX_XYW_ud1 / fluxY_XYW_ud1 (UD1.F90:519-659)
    if (k == d.KE) return 0.0;
    const int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
    const int c = dir == 0 ? I_UYZ : I_XVZ;
    const double* M = dir == 0 ? a.MOMX : a.MOMY;
    const double* V = a.MOMZ0;
    const double f1 = o.F2H(k, 1, c, i, j), f2 = o.F2H(k, 2, c, i, j);
    const double vel = (f1 * M[IX(d, k + 1, i, j)] + f2 * M[IX(d, k, i, j)])
                     / (f1 * 0.5 * (o.D(k + 1, i, j) + o.D(k + 1, i + di, j + dj))
                      + f2 * 0.5 * (o.D(k, i, j) + o.D(k, i + di, j + dj)));
    const double mp = dir == 0 ? m.mapf(i, j, 1, I_UY) : m.mapf(i, j, 0, I_XV);
    return m.gs(c, IX(d, k, i, j)) / mp * vel * ud1v(V[IX(d, k, i, j)], V[IX(d, k, i + di, j + dj)], vel);
  } else {
    const bool own = (C == 1) == (dir == 0);  // the flux along the component's own direction
    const double* V = C == 1 ? a.MOMX0 : a.MOMY0;
    if (own) {  // fluxX_UYZ_ud1 / fluxY_XVZ_ud1 at loop index (i,j)+1, stored one cell back (UD1.F90:869-920, 1237-1288)
      const int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
      const int ip = i + di, jp = j + dj;
      const double* M = C == 1 ? a.MOMX : a.MOMY;
      const double vel = (0.5 * (M[IX(d, k, ip, jp)] + M[IX(d, k, i, j)])) / (o.D(k, ip, jp));
      const double mp = dir == 0 ? m.mapf(ip, jp, 1, I_UY) : m.mapf(ip, jp, 0, I_XY);
      return m.gs(I_XYZ, IX(d, k, ip, jp)) / mp * vel * ud1v(V[IX(d, k, i, j)], V[IX(d, k, ip, jp)], vel);
    }
    // cross fluxes fluxY_UYZ_ud1 / fluxX_XVZ_ud1 (UD1.F90:924-973, 1184-1233)
    const int si = C == 1 ? 1 : 0, sj = C == 1 ? 0 : 1;  // the component's staggering
    const int ci = dir == 0 ? 1 : 0, cj = dir == 0 ? 0 : 1;
    const double* M = dir == 0 ? a.MOMX : a.MOMY;
    const double vel = (0.5 * (M[IX(d, k, i, j)] + M[IX(d, k, i + si, j + sj)]))
                     / (0.25 * (o.D(k, i, j) + o.D(k, i + 1, j) + o.D(k, i, j + 1) + o.D(k, i + 1, j + 1)));
    const double mp = C == 1 ? m.mapf(i, j, 0, I_XV) : m.mapf(i, j, 1, I_XY);
    return m.gs(I_UVZ, IX(d, k, i, j)) / mp * vel * ud1v(V[IX(d, k, i, j)], V[IX(d, k, i + ci, j + cj)], vel);
  }
}

// the high-order fluxes of the stage, the z flux in the form ATMOS_DYN_fct takes
template <int S, bool TERRAIN, int C>
__device__ double mhi_z(const Ops<S, TERRAIN, false>& o, int kk, int i, int j) {
  const auto& m = o.m;
  if constexpr (C == 0) {
    double q = o.qZ_XYW(kk, i, j) / (m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY));
    if constexpr (TERRAIN) q = q + o.qJ_XYW(0, kk, i, j) + o.qJ_XYW(1, kk, i, j);
    return q;
  } else {
    constexpr int st = C == 1 ? 0 : 1;
    const int mc = C == 1 ? I_UY : I_XV;
    double q = o.qZ_UV(st, kk, i, j) / (m.mapf(i, j, 0, mc) * m.mapf(i, j, 1, mc));
    if constexpr (TERRAIN) q = q + o.qJ_UV(st, 0, kk, i, j) + o.qJ_UV(st, 1, kk, i, j);
    return q;
  }
}
template <int S, bool TERRAIN, int C>
__device__ double mhi_h(const Ops<S, TERRAIN, false>& o, int dir, int k, int i, int j) {
  if constexpr (C == 0) return o.qH_XYW(dir, k, i, j);
  else if constexpr (C == 1) return dir == 0 ? o.qX_UYZ(k, i, j) : o.qY_UYZ(k, i, j);
  else return dir == 0 ? o.qX_XVZ(k, i, j) : o.qY_XVZ(k, i, j);
}

// lo: z on columns IS-1..IE+1 x JS-1..JE+1, levels KS-1..KE; x on faces
// IS-2..IE+1 x JS-1..JE+1; y on IS-1..IE+1 x faces JS-2..JE+1 (levels KS..KE).
// anti: z on the interior columns (0 at KS-1 and KE), x on faces IS-1..IE,
// y on faces JS-1..JE.
template <int S, bool TERRAIN, int C>
__global__ void __launch_bounds__(256) mfct_flux(HeviArgs a, double* lo, double* anti) {
  const Dims& d = a.d;
  Ops<S, TERRAIN, false> o(a);
  const int i = d.IS - 2 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int k = d.KS - 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int j = d.JS - 2 + (int)blockIdx.z;
  if (i > d.IE + 1 || k > d.KE || j > d.JE + 1) return;
  const size_t p = IX(d, k, i, j), n3 = d.n3;
  const bool in_i = i >= d.IS && i <= d.IE, in_j = j >= d.JS && j <= d.JE;
  const bool ri = i >= d.IS - 1, rj = j >= d.JS - 1;
  if (ri && rj) {
    const double l = mlo_z<S, TERRAIN, C>(o, a, k, i, j);
    lo[p] = l;
    if (in_i && in_j) anti[p] = (k >= d.KS && k <= d.KE - 1) ? mhi_z<S, TERRAIN, C>(o, k, i, j) - l : 0.0;
  }
  if (k < d.KS) return;
  if (rj) {
    const double l = mlo_h<S, TERRAIN, C>(o, a, 0, k, i, j);
    lo[p + n3] = l;
    if (in_j && ri && i <= d.IE) anti[p + n3] = mhi_h<S, TERRAIN, C>(o, 0, k, i, j) - l;
  }
  if (ri) {
    const double l = mlo_h<S, TERRAIN, C>(o, a, 1, k, i, j);
    lo[p + 2 * n3] = l;
    if (in_i && rj && j <= d.JE) anti[p + 2 * n3] = mhi_h<S, TERRAIN, C>(o, 1, k, i, j) - l;
  }
}

template <int S, bool TERRAIN>
static hipError_t heve_fct_momentum_t(const HeviArgs& a, const StageIO& io, hipStream_t s) {
  const Ctx& c = *io.ctx;
  const Dims& d = c.d;
  const size_t n3 = d.n3;
  double* W = c.mfct;
  double *vel = W, *lo = W + n3, *anti = W + 4 * n3, *phi_lo = W + 7 * n3, *rp = W + 8 * n3, *rm = W + 9 * n3;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1, kmax = d.KE - d.KS + 1;
  const dim3 blk(64, 4);
  const dim3 gi((imax + 63) / 64, (kmax + 3) / 4, jmax), gf((imax + 4 + 63) / 64, (kmax + 2 + 3) / 4, jmax + 4);
  double* dk[1] = {io.o[0]};
  hipError_t e = halo_fill(c, dk, 1, s);  // DENS_RK (HEVE.F90:746-747)
  for (int C = 0; C < 3 && e == hipSuccess; ++C) {
    const double* M0 = io.s0[1 + C];
    prof_start(c, "mfct_flux", s);
    if (C == 0) hipLaunchKernelGGL(mfct_vel<0>, gi, blk, 0, s, d, M0, io.s0[0], vel);
    else if (C == 1) hipLaunchKernelGGL(mfct_vel<1>, gi, blk, 0, s, d, M0, io.s0[0], vel);
    else hipLaunchKernelGGL(mfct_vel<2>, gi, blk, 0, s, d, M0, io.s0[0], vel);
    if ((e = hipGetLastError()) != hipSuccess) break;
    double* v[1] = {vel};
    if ((e = halo_fill(c, v, 1, s)) != hipSuccess) break;
    if (C == 0) hipLaunchKernelGGL((mfct_flux<S, TERRAIN, 0>), gf, blk, 0, s, a, lo, anti);
    else if (C == 1) hipLaunchKernelGGL((mfct_flux<S, TERRAIN, 1>), gf, blk, 0, s, a, lo, anti);
    else hipLaunchKernelGGL((mfct_flux<S, TERRAIN, 2>), gf, blk, 0, s, a, lo, anti);
    prof_stop(c, s);
    if ((e = hipGetLastError()) != hipSuccess) break;
    MomFct f;
    f.vel = vel; f.D0 = io.s0[0]; f.D1 = io.o[0];
    f.dk = C == 0; f.di = C == 1; f.dj = C == 2;
    f.lo = lo; f.anti = anti; f.mflx = io.mflx_hi;
    f.RDZ = C == 0 ? c.g1.RFDZ : c.g1.RCDZ;
    f.RDX = C == 1 ? c.g1.RFDX : c.g1.RCDX;
    f.RDY = C == 2 ? c.g1.RFDY : c.g1.RCDY;
    const int gc = C == 0 ? I_XYW : C == 1 ? I_UYZ : I_XVZ;
    const int mc = C == 0 ? I_XY : C == 1 ? I_UY : I_XV;
    f.G = c.GSQRT + n3 * (size_t)gc;
    f.MAPF = c.MAPF + d.pl * 2 * (size_t)mc;
    f.dt = io.dtrk; f.along = io.fct_along != 0;
    f.phi_lo = phi_lo; f.rjpls = rp; f.rjmns = rm;
    f.out = io.o[1 + C];
    e = fct_momentum_limit(c, f, s);
  }
  return e;
}

hipError_t heve_stage(const StageIO& io, hipStream_t s) {
  HeviArgs a;
  fill_args(a, io);
  const Ctx& c = *io.ctx;
  if (io.combine) return hipErrorNotSupported;
  if (io.fct_momentum && !c.mfct) return hipErrorInvalidValue;
  const bool div = io.divdmp_coef != 0.0;
  const bool ter = c.terrain;
  hipError_t e;
  if (c.scheme == UD3) {
    if (ter) e = div ? launch_heve<UD3, true, true>(a, s) : launch_heve<UD3, true, false>(a, s);
    else e = div ? launch_heve<UD3, false, true>(a, s) : launch_heve<UD3, false, false>(a, s);
  } else {
    if (ter) e = div ? launch_heve<CD4, true, true>(a, s) : launch_heve<CD4, true, false>(a, s);
    else e = div ? launch_heve<CD4, false, true>(a, s) : launch_heve<CD4, false, false>(a, s);
  }
  if (e == hipSuccess && io.fct_momentum) {
    if (c.scheme == UD3) e = ter ? heve_fct_momentum_t<UD3, true>(a, io, s) : heve_fct_momentum_t<UD3, false>(a, io, s);
    else e = ter ? heve_fct_momentum_t<CD4, true>(a, io, s) : heve_fct_momentum_t<CD4, false>(a, io, s);
  }
  if (e != hipSuccess || !io.fct_t) return e;
  return heve_fct_theta(io, s);
}

// FP64 stage recording PT, St and the solved vertical momentum C for the
// adjoint: hevi_tile on flat grids whose tiles fill the device, the split
// launches (hevi_expl sections + hevi_solve) otherwise and with terrain
hipError_t hevi_stage_traced(const StageIO& io, double* PT, double* St, double* Cp, hipStream_t s) {
  HeviArgs a;
  fill_args(a, io);
  const Ctx& c = *io.ctx;
  if (io.divdmp_coef != 0.0 || io.combine) return hipErrorNotSupported;
  a.PT = PT; a.trSt = St; a.trCp = Cp;
  const Dims& d = c.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1;
  const unsigned ni = (imax + 63) / 64;
  if (c.terrain) {
    if (c.scheme == UD3) launch_split<UD3, true, false>(a, s);
    else launch_split<CD4, true, false>(a, s);
    return hipGetLastError();
  }
  if (!hevi_flat_tile(c)) {
    if (c.scheme == UD3) launch_split<UD3, false, false>(a, s);
    else launch_split<CD4, false, false>(a, s);
    return hipGetLastError();
  }
  constexpr int TJ = 8;
  const unsigned nb = ni * ((jmax + TJ - 1) / TJ);
  if (c.scheme == UD3) hipLaunchKernelGGL((hevi_tile<UD3, false, TJ, double, false>), dim3(nb), dim3(64, TJ), 0, s, a, ni, nb);
  else hipLaunchKernelGGL((hevi_tile<CD4, false, TJ, double, false>), dim3(nb), dim3(64, TJ), 0, s, a, ni, nb);
  return hipGetLastError();
}

hipError_t hevi_stage(const StageIOT<float>& io, hipStream_t s) {
  HeviArgsT<float> a;
  fill_args(a, io);
  const Ctx& c = *io.ctx;
  if (c.terrain || io.divdmp_coef != 0.0) return hipErrorNotSupported;
  const Dims& d = c.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1;
  constexpr int TJ = 8;
  const unsigned ni = (imax + 63) / 64, nb = ni * ((jmax + TJ - 1) / TJ);
  prof_start(c, "hevi_stage", s);
  const bool cmb = io.combine != 0;
  if (c.scheme == UD3) {
    if (cmb) hipLaunchKernelGGL((hevi_tile<UD3, false, TJ, float, true>), dim3(nb), dim3(64, TJ), 0, s, a, ni, nb);
    else hipLaunchKernelGGL((hevi_tile<UD3, false, TJ, float, false>), dim3(nb), dim3(64, TJ), 0, s, a, ni, nb);
  } else {
    if (cmb) hipLaunchKernelGGL((hevi_tile<CD4, false, TJ, float, true>), dim3(nb), dim3(64, TJ), 0, s, a, ni, nb);
    else hipLaunchKernelGGL((hevi_tile<CD4, false, TJ, float, false>), dim3(nb), dim3(64, TJ), 0, s, a, ni, nb);
  }
  prof_stop(c, s);
  return hipGetLastError();
}

}  // namespace shp
