// hevi.hip -- one RK stage of the HEVI finite-volume short step on gfx950.
//
// Reference: ATMOS_DYN_Tstep_short_fvm_hevi,
//   scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short_fvm_hevi.F90:128-1089
// with the flux stencils of scale_atmos_dyn_fvm_flux_{ud3,cd4}.F90.
//
// Two kernels per stage:
//   hevi_explicit  one thread per interior point (i fastest, 64-lane rows):
//                  mass fluxes X/Y, theta fluxes X/Y, the explicit part of the
//                  Z fluxes, the tendencies Sr/St/Sw, the vertical-solve
//                  face value PT, and the complete MOMX/MOMY updates
//                  (HEVI.F90:346-657, 816-1047).
//   hevi_solve     one thread per column marching k: tridiagonal bands,
//                  right-hand side, Thomas forward sweep and back
//                  substitution into MOMZ/DENS/RHOT and the implicit parts of
//                  the Z fluxes (HEVI.F90:688-804, 1092-1127).
// Fluxes needed on both faces of a cell are evaluated by the two threads that
// own them with the same expression, so both see identical values.
#include "common.h"
#include "shp_internal.h"

namespace shp {

struct HeviArgs {
  const Ctx* ctx;  // host side only
  Dims d;
  Grid1D g;
  const double *GSQRT, *J13G, *J23G, *MAPF, *PHI;
  double J33G;
  // inputs (device layout)
  const double *DENS0, *MOMZ0, *MOMX0, *MOMY0, *RHOT0;
  const double *DENS, *MOMZ, *MOMX, *MOMY, *RHOT;
  const double *DENS_t, *MOMZ_t, *MOMX_t, *MOMY_t, *RHOT_t;
  const double *DPRES0, *RT2P, *CORIOLI, *num_diff, *wdamp_coef, *DDIV, *REF_dens, *REF_rhot;
  double divdmp_coef, dtrk;
  int has_coriolis;
  // outputs
  double *DENS_RK, *MOMZ_RK, *MOMX_RK, *MOMY_RK, *RHOT_RK, *mflx_hi, *tflx_hi;
  // scratch (n3 each)
  double *Sr, *St, *Sw, *PT, *e, *f;
};

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
template <int S, bool TERRAIN, bool DIVDMP>
__global__ void __launch_bounds__(256) hevi_explicit(HeviArgs a, unsigned ni_tiles, unsigned nk_tiles, unsigned nblocks) {
  const Dims& d = a.d;
  Ops<S, TERRAIN, DIVDMP> o(a);
  const unsigned lb = xcd_block(blockIdx.x, nblocks);
  const unsigned it = lb % ni_tiles, kt = (lb / ni_tiles) % nk_tiles, jt = lb / (ni_tiles * nk_tiles);
  const int i = d.IS + (int)(it * 64 + threadIdx.x);
  const int k = d.KS + (int)(kt * 4 + threadIdx.y);
  const int j = d.JS + (int)jt;
  if (i > d.IE || k > d.KE || j > d.JE) return;
  const size_t p = IX(d, k, i, j);
  const double* RCDX = a.g.RCDX; const double* RCDY = a.g.RCDY; const double* RCDZ = a.g.RCDZ;
  const double* RFDX = a.g.RFDX; const double* RFDY = a.g.RFDY; const double* RFDZ = a.g.RFDZ;
  const auto& m = o.m;
  double* mfZ = a.mflx_hi; double* mfX = a.mflx_hi + d.n3; double* mfY = a.mflx_hi + 2 * d.n3;
  double* tfZ = a.tflx_hi; double* tfX = a.tflx_hi + d.n3; double* tfY = a.tflx_hi + 2 * d.n3;

  // ---- continuity (HEVI.F90:369-475) ----
  const double mX = o.mflxX(k, i, j), mXm = o.mflxX(k, i - 1, j);
  const double mY = o.mflxY(k, i, j), mYm = o.mflxY(k, i, j - 1);
  mfX[p] = mX;
  mfY[p] = mY;
  if (i == d.IS) mfX[IX(d, k, i - 1, j)] = mXm;
  if (j == d.JS) mfY[IX(d, k, i, j - 1)] = mYm;
  const double mZ = o.mflxZf(k, i, j), mZm = o.mflxZf(k - 1, i, j);
  if (k <= d.KE - 1) mfZ[p] = mZ;
  if (k == d.KS) mfZ[IX(d, k - 1, i, j)] = 0.0;
  if (k == d.KE) mfZ[p] = 0.0;
  if (i == d.IS) {  // mflx_hi(Z) is also defined on the i = IS-1 column (HEVI.F90:370)
    if (k <= d.KE - 1) mfZ[IX(d, k, i - 1, j)] = o.mflxZ(k, i - 1, j);
    else mfZ[IX(d, k, i - 1, j)] = 0.0;
    if (k == d.KS) mfZ[IX(d, k - 1, i - 1, j)] = 0.0;
  }
  const double M12 = m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY);
  {
    double advcv = -(mZ - mZm) * RCDZ[k];
    double advch = -((mX - mXm) * RCDX[i] + (mY - mYm) * RCDY[j]);
    a.Sr[p] = (advcv + advch) * m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY) / m.gs(I_XYZ, p) + a.DENS_t[p];
  }

  // ---- thermodynamics (HEVI.F90:588-657) ----
  {
    const double tX = o.tflxX(k, i, j, mX), tXm = o.tflxX(k, i - 1, j, mXm);
    const double tY = o.tflxY(k, i, j, mY), tYm = o.tflxY(k, i, j - 1, mYm);
    const double tZ = o.tflxZ(k, i, j), tZm = o.tflxZ(k - 1, i, j);
    tfX[p] = tX;
    tfY[p] = tY;
    if (i == d.IS) tfX[IX(d, k, i - 1, j)] = tXm;
    if (j == d.JS) tfY[IX(d, k, i, j - 1)] = tYm;
    if (k <= d.KE - 1) tfZ[p] = tZ;
    if (k == d.KS) tfZ[IX(d, k - 1, i, j)] = 0.0;
    if (k == d.KE) tfZ[p] = 0.0;
    double advcv = -(tZ - tZm) * RCDZ[k];
    double advch = -((tX - tXm) * RCDX[i] + (tY - tYm) * RCDY[j]);
    a.St[p] = (advcv + advch) * m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY) / m.gs(I_XYZ, p) + a.RHOT_t[p];
  }

  // ---- momentum z, explicit part (HEVI.F90:486-577) and PT (HEVI.F90:691) ----
  if (k <= d.KE - 1) {
    double sv = o.qZ_XYW(k, i, j) - o.qZ_XYW(k - 1, i, j);
    if constexpr (TERRAIN) {
      sv = sv + o.qJ_XYW(0, k, i, j) - o.qJ_XYW(0, k - 1, i, j) + o.qJ_XYW(1, k, i, j) - o.qJ_XYW(1, k - 1, i, j);
    }
    double advcv = -(sv) * RFDZ[k];
    double advch = -((o.qH_XYW(0, k, i, j) - o.qH_XYW(0, k, i - 1, j)) * RCDX[i]
                   + (o.qH_XYW(1, k, i, j) - o.qH_XYW(1, k, i, j - 1)) * RCDY[j])
                 * m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY);
    double wdmp = -a.wdamp_coef[k] * a.MOMZ0[p];
    double sw = (advcv + advch) / m.gs(I_XYW, p) + wdmp;
    if constexpr (DIVDMP) sw = sw + a.divdmp_coef / a.dtrk * (a.DDIV[IX(d, k + 1, i, j)] - a.DDIV[p]) * a.g.FDZ[k];
    a.Sw[p] = sw + a.MOMZ_t[p];
    a.PT[p] = o.PTf(k, i, j);
  }
  (void)M12;

  // ---- momentum x (HEVI.F90:816-927), iee = min(IIE,IEH) = IE for periodic x ----
  {
    double sv = o.qZ_UV(0, k, i, j) - o.qZ_UV(0, k - 1, i, j);
    if constexpr (TERRAIN) {
      sv = sv + o.qJ_UV(0, 0, k, i, j) - o.qJ_UV(0, 0, k - 1, i, j) + o.qJ_UV(0, 1, k, i, j) - o.qJ_UV(0, 1, k - 1, i, j);
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
    if constexpr (DIVDMP) {
      t = t + a.divdmp_coef / a.dtrk * (a.DDIV[IX(d, k, i + 1, j)] / m.mapf(i + 1, j, 1, I_XY) - a.DDIV[p] / m.mapf(i, j, 0, I_XY))
                * m.mapf(i, j, 0, I_UY) * m.mapf(i, j, 1, I_UY) * a.g.FDX[i];
    }
    a.MOMX_RK[p] = a.MOMX0[p] + a.dtrk * (t + a.MOMX_t[p]);
  }

  // ---- momentum y (HEVI.F90:936-1047) ----
  {
    double sv = o.qZ_UV(1, k, i, j) - o.qZ_UV(1, k - 1, i, j);
    if constexpr (TERRAIN) {
      sv = sv + o.qJ_UV(1, 0, k, i, j) - o.qJ_UV(1, 0, k - 1, i, j) + o.qJ_UV(1, 1, k, i, j) - o.qJ_UV(1, 1, k - 1, i, j);
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
    if constexpr (DIVDMP) {
      t = t + a.divdmp_coef / a.dtrk * (a.DDIV[IX(d, k, i, j + 1)] / m.mapf(i, j + 1, 0, I_XY) - a.DDIV[p] / m.mapf(i, j, 0, I_XY))
                * m.mapf(i, j, 0, I_XV) * m.mapf(i, j, 1, I_XV) * a.g.FDY[j];
    }
    a.MOMY_RK[p] = a.MOMY0[p] + a.dtrk * (t + a.MOMY_t[p]);
  }
}

// ---------------------------------------------------------------------------
// vertical implicit solve, one thread per column (HEVI.F90:688-804, 1092-1127)
template <bool TERRAIN>
__global__ void __launch_bounds__(256) hevi_solve(HeviArgs a, unsigned ni_tiles, unsigned nblocks) {
  const Dims& d = a.d;
  const unsigned lb = xcd_block(blockIdx.x, nblocks);
  const int i = d.IS + (int)((lb % ni_tiles) * 256 + threadIdx.x);
  const int j = d.JS + (int)(lb / ni_tiles);
  if (i > d.IE || j > d.JE) return;
  Met<TERRAIN> m;
  m.G = a.GSQRT; m.J13 = a.J13G; m.J23 = a.J23G; m.MAPF = a.MAPF; m.n3 = d.n3; m.pl = d.pl; m.IA = d.IA;
  const double* CDZ = a.g.CDZ; const double* RCDZ = a.g.RCDZ; const double* RFDZ = a.g.RFDZ;
  const double dtrk = a.dtrk, J33G = a.J33G;
  const int KS = d.KS, KE = d.KE;
  auto px = [&](int k) { return IX(d, k, i, j); };
  auto dpres = [&](int k) { size_t q = px(k); return a.DPRES0[q] + a.RT2P[q] * (a.RHOT[q] - a.REF_rhot[q]); };
  auto Af = [&](int k, double rt) { return dtrk * dtrk * J33G * RCDZ[k] * rt * J33G / m.gs(I_XYZ, px(k)); };
  auto f2h = [&](int k, int pp) {
    return CDZ[k + pp - 1] * m.gs(I_XYZ, px(k + pp - 1)) / (CDZ[k] * m.gs(I_XYZ, px(k)) + CDZ[k + 1] * m.gs(I_XYZ, px(k + 1)));
  };
  // level k values
  double rt0 = a.RT2P[px(KS)], dp0 = dpres(KS), st0 = a.St[px(KS)], sr0 = a.Sr[px(KS)];
  double d0 = a.DENS[px(KS)], rd0 = a.REF_dens[px(KS)], A0 = Af(KS, rt0), pt0 = a.PT[px(KS)], ptm = 0.0;
  double e_prev = 0.0, f_prev = 0.0, C_last = 0.0;
  for (int k = KS; k <= KE - 1; ++k) {
    const size_t q = px(k), q1 = px(k + 1);
    const double rt1 = a.RT2P[q1], dp1 = dpres(k + 1), st1 = a.St[q1], sr1 = a.Sr[q1];
    const double d1 = a.DENS[q1], rd1 = a.REF_dens[q1], A1 = Af(k + 1, rt1);
    const double pt1 = (k + 1 <= KE - 1) ? a.PT[q1] : 0.0;
    const double gw = m.gs(I_XYW, q);
    const double B = GRAV * (dtrk * dtrk) * J33G / (CDZ[k + 1] + CDZ[k]);
    const double F1v = -(pt1 * RFDZ[k] * A1 + B) / gw;
    const double F2v = 1.0 + (pt0 * RFDZ[k] * (A1 + A0)) / gw;
    const double F3v = -(ptm * RFDZ[k] * A0 - B) / gw;
    const double pg = -(dp1 + rt1 * dtrk * st1 - dp0 - rt0 * dtrk * st0) * RFDZ[k] * J33G / gw
                    - GRAV * (f2h(k, 1) * (d1 - rd1 + sr1 * dtrk) + f2h(k, 2) * (d0 - rd0 + sr0 * dtrk));
    const double Ck = a.MOMZ[q] + dtrk * (pg + a.Sw[q]);
    if (k == KS) {
      const double rdenom = 1.0 / F2v;
      e_prev = -F1v * rdenom;
      f_prev = Ck * rdenom;
      a.e[q] = e_prev; a.f[q] = f_prev;
    } else if (k < KE - 1) {
      const double rdenom = 1.0 / (F2v + F3v * e_prev);
      const double ek = -F1v * rdenom;
      const double fk = (Ck - F3v * f_prev) * rdenom;
      e_prev = ek; f_prev = fk;
      a.e[q] = ek; a.f[q] = fk;
    } else {
      C_last = (Ck - F3v * f_prev) / (F2v + F3v * e_prev);
    }
    rt0 = rt1; dp0 = dp1; st0 = st1; sr0 = sr1; d0 = d1; rd0 = rd1; A0 = A1; ptm = pt0; pt0 = pt1;
  }
  // back substitution and updates (HEVI.F90:728-790)
  double* mfZ = a.mflx_hi; double* tfZ = a.tflx_hi;
  const double M12 = m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY);
  double Cn = C_last;                         // C(k+1)
  double ptn = a.PT[px(KE - 1)];              // PT(k+1)
  {
    const size_t q = px(KE);
    double advcv = Cn * J33G * RCDZ[KE] / m.gs(I_XYZ, q);
    a.DENS_RK[q] = a.DENS0[q] + dtrk * (advcv + a.Sr[q]);
    advcv = Cn * ptn * J33G * RCDZ[KE] / m.gs(I_XYZ, q);
    a.RHOT_RK[q] = a.RHOT0[q] + dtrk * (advcv + a.St[q]);
    a.MOMZ_RK[q] = 0.0;
  }
  {
    const size_t q = px(KE - 1);
    mfZ[q] = mfZ[q] + J33G * Cn / (M12);
    tfZ[q] = tfZ[q] + J33G * Cn * ptn / (M12);
    a.MOMZ_RK[q] = a.MOMZ0[q] + (Cn - a.MOMZ[q]);
  }
  for (int k = KE - 2; k >= KS; --k) {
    const size_t q = px(k), q1 = px(k + 1);
    const double Ck = a.e[q] * Cn + a.f[q];
    const double ptk = a.PT[q];
    mfZ[q] = mfZ[q] + J33G * Ck / (M12);
    tfZ[q] = tfZ[q] + J33G * Ck * ptk / (M12);
    a.MOMZ_RK[q] = a.MOMZ0[q] + (Ck - a.MOMZ[q]);
    // level k+1 (KS+1..KE-1) now has both C(k+1) and C(k)
    double advcv = -(Cn - Ck) * J33G * RCDZ[k + 1] / m.gs(I_XYZ, q1);
    a.DENS_RK[q1] = a.DENS0[q1] + dtrk * (advcv + a.Sr[q1]);
    advcv = -(Cn * ptn - Ck * ptk) * J33G * RCDZ[k + 1] / m.gs(I_XYZ, q1);
    a.RHOT_RK[q1] = a.RHOT0[q1] + dtrk * (advcv + a.St[q1]);
    Cn = Ck; ptn = ptk;
  }
  {
    const size_t q = px(KS);
    double advcv = -Cn * J33G * RCDZ[KS] / m.gs(I_XYZ, q);
    a.DENS_RK[q] = a.DENS0[q] + dtrk * (advcv + a.Sr[q]);
    advcv = -Cn * ptn * J33G * RCDZ[KS] / m.gs(I_XYZ, q);
    a.RHOT_RK[q] = a.RHOT0[q] + dtrk * (advcv + a.St[q]);
    a.MOMZ_RK[px(KS - 1)] = 0.0;
  }
}

// ---------------------------------------------------------------------------
template <int S, bool TERRAIN, bool DIVDMP>
static hipError_t launch_t(const HeviArgs& a, hipStream_t s) {
  const Dims& d = a.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1, kmax = d.KE - d.KS + 1;
  const unsigned ni = (imax + 63) / 64, nk = (kmax + 3) / 4;
  const unsigned nb = ni * nk * (unsigned)jmax;
  const Ctx& c = *a.ctx;
  prof_start(c, "hevi_explicit", s);
  hipLaunchKernelGGL((hevi_explicit<S, TERRAIN, DIVDMP>), dim3(nb), dim3(64, 4), 0, s, a, ni, nk, nb);
  prof_stop(c, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const unsigned ns = (imax + 255) / 256;
  const unsigned nbs = ns * (unsigned)jmax;
  prof_start(c, "hevi_solve", s);
  hipLaunchKernelGGL((hevi_solve<TERRAIN>), dim3(nbs), dim3(256), 0, s, a, ns, nbs);
  prof_stop(c, s);
  return hipGetLastError();
}

hipError_t hevi_stage(const StageIO& io, hipStream_t s) {
  HeviArgs a;
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
  a.Sr = c.scr[0]; a.St = c.scr[1]; a.Sw = c.scr[2]; a.PT = c.scr[3]; a.e = c.scr[4]; a.f = c.scr[5];
  const bool div = io.divdmp_coef > 0.0;
  const bool ter = c.terrain;
  if (c.scheme == UD3) {
    if (ter) return div ? launch_t<UD3, true, true>(a, s) : launch_t<UD3, true, false>(a, s);
    return div ? launch_t<UD3, false, true>(a, s) : launch_t<UD3, false, false>(a, s);
  }
  if (ter) return div ? launch_t<CD4, true, true>(a, s) : launch_t<CD4, true, false>(a, s);
  return div ? launch_t<CD4, false, true>(a, s) : launch_t<CD4, false, false>(a, s);
}

}  // namespace shp
