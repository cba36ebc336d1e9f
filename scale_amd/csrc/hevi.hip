// hevi.hip -- one RK stage of the HEVI finite-volume short step on gfx950.
//
// Reference: ATMOS_DYN_Tstep_short_fvm_hevi,
//   scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short_fvm_hevi.F90:128-1089
// with the flux stencils of scale_atmos_dyn_fvm_flux_{ud3,cd4}.F90.
//
// One kernel per stage, hevi_march: one thread per column (i,j), blocks of
// 64 (i) x 4 (j) columns.  Marching k upward it evaluates the explicit part
// of each point (mass and theta fluxes, the explicit z fluxes, the tendencies
// Sr/St/Sw, the face value PT, the complete MOMX/MOMY updates;
// HEVI.F90:346-657, 816-1047) and, one level behind, the tridiagonal rows and
// the Thomas forward sweep (HEVI.F90:688-727, 1092-1115); marching back down
// it substitutes into MOMZ/DENS/RHOT and adds the implicit parts of the z
// fluxes (HEVI.F90:728-804, 1116-1127).
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
  double *PT, *e, *f;
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
struct PointTend { double Sr, St, Sw, PT; };

// Fluxes through the lower face of the current level, carried from the level
// below (where they were the upper face), and the column's theta window
// POTT(k-1..k+2).  Every face flux at KS-1 is zero.
struct ZCarry {
  double mZ = 0.0, tZ = 0.0, qW = 0.0, qWJ13 = 0.0, qWJ23 = 0.0;
  double qU = 0.0, qUJ13 = 0.0, qUJ23 = 0.0, qV = 0.0, qVJ13 = 0.0, qVJ23 = 0.0;
  double pw[4];
};

template <int S, bool TERRAIN, bool DIVDMP>
__device__ __forceinline__ PointTend explicit_point(const Ops<S, TERRAIN, DIVDMP>& o, const HeviArgs& a, int k, int i, int j,
                                                    ZCarry& zc) {
  const Dims& d = a.d;
  PointTend r;
  r.Sw = 0.0; r.PT = 0.0;
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
  const double mZ = o.mflxZf(k, i, j), mZm = zc.mZ;
  zc.mZ = mZ;
  if (k <= d.KE - 1) mfZ[p] = mZ;
  if (k == d.KS) mfZ[IX(d, k - 1, i, j)] = 0.0;
  if (k == d.KE) mfZ[p] = 0.0;
  if (i == d.IS) {  // mflx_hi(Z) is also defined on the i = IS-1 column (HEVI.F90:370)
    if (k <= d.KE - 1) mfZ[IX(d, k, i - 1, j)] = o.mflxZ(k, i - 1, j);
    else mfZ[IX(d, k, i - 1, j)] = 0.0;
    if (k == d.KS) mfZ[IX(d, k - 1, i - 1, j)] = 0.0;
  }
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
    tfX[p] = tX;
    tfY[p] = tY;
    if (i == d.IS) tfX[IX(d, k, i - 1, j)] = tXm;
    if (j == d.JS) tfY[IX(d, k, i, j - 1)] = tYm;
    if (k <= d.KE - 1) tfZ[p] = tZ;
    if (k == d.KS) tfZ[IX(d, k - 1, i, j)] = 0.0;
    if (k == d.KE) tfZ[p] = 0.0;
    double advcv = -(tZ - tZm) * RCDZ[k];
    double advch = -((tX - tXm) * RCDX[i] + (tY - tYm) * RCDY[j]);
    r.St = (advcv + advch) * m.mapf(i, j, 0, I_XY) * m.mapf(i, j, 1, I_XY) / m.gs(I_XYZ, p) + a.RHOT_t[p];
  }

  // ---- momentum z, explicit part (HEVI.F90:486-577) and PT (HEVI.F90:691) ----
  if (k <= d.KE - 1) {
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
  {
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
    if constexpr (DIVDMP) {
      t = t + a.divdmp_coef / a.dtrk * (a.DDIV[IX(d, k, i + 1, j)] / m.mapf(i + 1, j, 1, I_XY) - a.DDIV[p] / m.mapf(i, j, 0, I_XY))
                * m.mapf(i, j, 0, I_UY) * m.mapf(i, j, 1, I_UY) * a.g.FDX[i];
    }
    a.MOMX_RK[p] = a.MOMX0[p] + a.dtrk * (t + a.MOMX_t[p]);
  }

  // ---- momentum y (HEVI.F90:936-1047) ----
  {
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
    if constexpr (DIVDMP) {
      t = t + a.divdmp_coef / a.dtrk * (a.DDIV[IX(d, k, i, j + 1)] / m.mapf(i, j + 1, 0, I_XY) - a.DDIV[p] / m.mapf(i, j, 0, I_XY))
                * m.mapf(i, j, 0, I_XV) * m.mapf(i, j, 1, I_XV) * a.g.FDY[j];
    }
    a.MOMY_RK[p] = a.MOMY0[p] + a.dtrk * (t + a.MOMY_t[p]);
  }
  return r;
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
    Sr_[q1] = t.Sr;
    St_[q1] = t.St;
    if (k <= KE - 1) a.PT[q1] = t.PT;
    const double rt1 = a.RT2P[q1], dp1 = dpres(k), d1 = a.DENS[q1], rd1 = a.REF_dens[q1], A1 = Af(k, rt1);
    const double pt1 = (k <= KE - 1) ? t.PT : 0.0;
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
        e_prev = -F1v * rdenom;
        f_prev = Ck * rdenom;
        a.e[q] = e_prev; a.f[q] = f_prev;
      } else if (r < KE - 1) {
        const double rdenom = 1.0 / (F2v + F3v * e_prev);
        const double ek = -F1v * rdenom;
        const double fk = (Ck - F3v * f_prev) * rdenom;
        e_prev = ek; f_prev = fk;
        a.e[q] = ek; a.f[q] = fk;
      } else {
        C_last = (Ck - F3v * f_prev) / (F2v + F3v * e_prev);
      }
      ptm = pt0;
    }
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
template <int S, bool TERRAIN, bool DIVDMP>
static hipError_t launch_t(const HeviArgs& a, hipStream_t s) {
  const Dims& d = a.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1;
  constexpr unsigned TJ = 4;
  const unsigned ni = (imax + 63) / 64, nj = (jmax + TJ - 1) / TJ;
  const unsigned nb = ni * nj;
  const Ctx& c = *a.ctx;
  prof_start(c, "hevi_stage", s);
  hipLaunchKernelGGL((hevi_march<S, TERRAIN, DIVDMP>), dim3(nb), dim3(64, TJ), 0, s, a, ni, nb);
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
  a.PT = c.scr[3]; a.e = c.scr[4]; a.f = c.scr[5];
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
