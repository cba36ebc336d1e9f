// tracer_tile.hip -- the FCT tracer stage on LDS tiles for flat terrain and
// the tracer schemes whose faces read two cells on each side (UD3, CD4,
// UD3KOREN1993, UD1, CD2).  Same results as the per-point kernels of
// tracer.hip, bit for bit: every expression is the same, only the places the
// operands come from differ.
//
// Reference: ATMOS_DYN_Tstep_tracer_fvm_heve,
//   scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_tracer_fvm_heve.F90:72-242,
// and ATMOS_DYN_fct, scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90:1922-2734.
//
// A block owns 64 (i) x TJ (j) columns and marches k upward.  Per level it
// stages the planes it reads at neighbours in LDS (with a ring of one or two
// cells), evaluates every x and y face of the tile once into LDS, carries the
// z faces from the level below in registers, then finishes its cells.
//   trt_ratio   FCT pass 1: the low-order field phi_lo on the tile and a
//               one-cell ring (kept in LDS, three levels), the anti-diffusive
//               fluxes, P+/P-, the along-stream or 7-point bounds and R+/R-
//   [periodic fill of R+/R-]
//   trt_update  FCT pass 2: limited anti-diffusive fluxes and the update
#include "common.h"
#include "fct.h"
#include "shp_internal.h"
#include "batch.h"

namespace shp {

struct TtArgs {  // nd and RHOQ_t may be null: zeros (as in TrArgs)
  Dims d;
  const double *QTRC, *QTRC0, *RHOQ_t, *DENS0, *DENS, *mflx, *nd;
  size_t nds;  // distance between the z, x and y components of nd
  const double *RCDZ, *RCDX, *RCDY;
  double dtl;
  double* QTRCo;
  double *rjpls, *rjmns;
};

// Plane (k) of F on the rows j0m.. and columns i0m.. of an H x W LDS array;
// points outside the array extents (partial tiles) read as 0 and are never used.
template <int H, int W>
__device__ __forceinline__ void tt_load(double (*s)[W], const double* F, const Dims& d, int k, int i0m, int j0m, int tid,
                                        int nthr) {
  const size_t base = d.pl * (size_t)k;
  for (int t = tid; t < H * W; t += nthr) {
    const int r = t / W, c = t - r * W;
    const int i = i0m + c, j = j0m + r;
    s[r][c] = (i < d.IA && j < d.JA) ? F[base + (size_t)i + (size_t)d.IA * j] : 0.0;
  }
}

// high-order face value times velocity plus num_diff (TrOps::hX/hY/hZ of the
// per-point kernels, flat terrain); q holds the four cells around the face
template <int S>
__device__ __forceinline__ double tt_hflux(double vel, double qm1, double q0, double q1, double q2, double nd) {
  if constexpr (S == UD1) return vel * fv_int<UD1>(0.0, q0, q1, 0.0, vel);
  return vel * fv_int<S>(qm1, q0, q1, q2, vel) + 1.0 * nd;
}
__device__ __forceinline__ double tt_lflux(double vel, double q0, double q1) {
  return vel * fv_int<UD1>(0.0, q0, q1, 0.0, vel);
}

// vertical high-order flux through face kk of a column; w holds val(kk-1..kk+2)
template <int S>
__device__ __forceinline__ double tt_hz(const Dims& d, int kk, double vel, const double* w, double nd) {
  if (kk == d.KS - 1 || kk == d.KE) return 0.0;
  double fv;
  if (kk == d.KS) fv = fv_lo<S>(w[1], w[2], w[3], vel);
  else if (kk == d.KE - 1) fv = fv_hi<S>(w[0], w[1], w[2], vel);
  else fv = fv_int<S>(w[0], w[1], w[2], w[3], vel);
  if constexpr (S == UD1) return vel * fv;
  return vel * fv + 1.0 * nd;
}

// ---- FCT pass 1: phi_lo (ring of one), anti-diffusive fluxes, R+/R- ------------
// each pass is a body over tile lb of a grid ni_tiles tiles across, run by
// its kernel once per block or, for an ensemble, by trt_batch over a table
// of members
template <int S, int TJ, bool VECT>
__device__ __forceinline__ void trt_ratio_body(const TtArgs& a, unsigned lb, unsigned ni_tiles) {
  constexpr int W2 = 68, H2 = TJ + 4, W1 = 66, H1 = TJ + 2;
  __shared__ double sQ0[5][H2][W2];  // QTRC0, two-cell ring, levels k-2..k+2 (slot = level % 5)
  __shared__ double sQ[H2][W2];      // QTRC at level k, two-cell ring
  __shared__ double sPL[3][H1][W1];  // phi_lo, one-cell ring, levels k-1..k+1 (slot = level % 3)
  __shared__ double sLZ[H1][W1];     // low-order flux through the top face of the ring cell's last phi_lo level
  __shared__ double sAX[TJ][65];     // anti-diffusive x fluxes through faces i0-1 .. i0+63
  __shared__ double sAY[TJ + 1][64]; // anti-diffusive y fluxes through faces j0-1 .. j0+TJ-1
  const Dims& d = a.d;
  const int KS = d.KS, KE = d.KE;
  const int i0 = d.IS + 64 * (int)(lb % ni_tiles), j0 = d.JS + TJ * (int)(lb / ni_tiles);
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + 64 * ty, nthr = 64 * TJ;
  const int i = i0 + tx, j = j0 + ty;
  const bool act = i <= d.IE && j <= d.JE;
  const double* mZ = a.mflx;
  const double* mX = a.mflx + d.n3;
  const double* mY = a.mflx + 2 * d.n3;
  const double dt = a.dtl;

  // phi_lo of the ring cells at level k1 (COMMON.F90:2033-2045); the flux
  // through their bottom face is sLZ, the one through the top face replaces it
  // phi_lo of ring cell t (row-major over the H1 x W1 ring) at level k1
  auto ring_cell = [&](int k1, int t) {
    const double(*Q1)[W2] = sQ0[k1 % 5];
    const double(*Q2)[W2] = sQ0[(k1 + 1) % 5];
    const size_t pk = d.pl * (size_t)k1;
    const int r = t / W1, c = t - r * W1;
    const int ii = i0 - 1 + c, jj = j0 - 1 + r;
    if (ii > d.IE + 1 || jj > d.JE + 1) return;
    const size_t p = pk + (size_t)ii + (size_t)d.IA * jj;
    const int c2 = c + 1, r2 = r + 1;
    const double lzu = (k1 == KE) ? 0.0 : tt_lflux(mZ[p], Q1[r2][c2], Q2[r2][c2]);
    const double lzd = sLZ[r][c];
    const double lx = tt_lflux(mX[p], Q1[r2][c2], Q1[r2][c2 + 1]);
    const double lxm = tt_lflux(mX[p - 1], Q1[r2][c2 - 1], Q1[r2][c2]);
    const double ly = tt_lflux(mY[p], Q1[r2][c2], Q1[r2 + 1][c2]);
    const double lym = tt_lflux(mY[p - d.IA], Q1[r2 - 1][c2], Q1[r2][c2]);
    sPL[k1 % 3][r][c] = (Q1[r2][c2] * a.DENS0[p]
                         + dt * (-((lzu - lzd) * a.RCDZ[k1] + (lx - lxm) * a.RCDX[ii] + (ly - lym) * a.RCDY[jj])))
                      / a.DENS[p];
    sLZ[r][c] = lzu;
  };
  auto ring_philo = [&](int k1) {
    for (int t = tid; t < H1 * W1; t += nthr) ring_cell(k1, t);
  };

  tt_load<H2, W2>(sQ0[KS % 5], a.QTRC0, d, KS, i0 - 2, j0 - 2, tid, nthr);
  tt_load<H2, W2>(sQ0[(KS + 1) % 5], a.QTRC0, d, KS + 1, i0 - 2, j0 - 2, tid, nthr);
  for (int t = tid; t < H1 * W1; t += nthr) sLZ[t / W1][t % W1] = 0.0;
  __syncthreads();
  ring_philo(KS);

  const size_t pc = (size_t)(act ? i : d.IS) + (size_t)d.IA * (act ? j : d.JS);
  double w[4];
  for (int m = 0; m < 4; ++m) w[m] = a.QTRC[pc + d.pl * (size_t)(KS - 1 + m)];
  double azm = 0.0;  // anti-diffusive flux through the face below
  for (int k = KS; k <= KE; ++k) {
    if (k + 2 <= KE) tt_load<H2, W2>(sQ0[(k + 2) % 5], a.QTRC0, d, k + 2, i0 - 2, j0 - 2, tid, nthr);
    tt_load<H2, W2>(sQ, a.QTRC, d, k, i0 - 2, j0 - 2, tid, nthr);
    __syncthreads();
    {
      // one pass over the level's x faces, y faces and (below the top) the
      // ring cells of phi_lo at k+1: the block's threads take them in turn,
      // so no pass runs with a few threads while the rest wait
      const double(*Q0)[W2] = sQ0[k % 5];
      const size_t pk = d.pl * (size_t)k;
      constexpr int NXF = TJ * 65, NYF = (TJ + 1) * 64;
      const int nwork = NXF + NYF + (k + 1 <= KE ? H1 * W1 : 0);
      for (int t = tid; t < nwork; t += nthr) {
        if (t < NXF) {
          const int r = t / 65, f = t - r * 65;
          const int ii = i0 - 1 + f, jj = j0 + r;
          if (ii <= d.IE && jj <= d.JE) {
            const size_t p = pk + (size_t)ii + (size_t)d.IA * jj;
            const int c = f + 1, rr = r + 2;
            const double vel = mX[p];
            sAX[r][f] = tt_hflux<S>(vel, sQ[rr][c - 1], sQ[rr][c], sQ[rr][c + 1], sQ[rr][c + 2], (a.nd ? a.nd[p + a.nds] : 0.0))
                      - tt_lflux(vel, Q0[rr][c], Q0[rr][c + 1]);
          }
        } else if (t < NXF + NYF) {
          const int u = t - NXF;
          const int r = u / 64, c0 = u - r * 64;
          const int ii = i0 + c0, jj = j0 - 1 + r;
          if (ii <= d.IE && jj <= d.JE) {
            const size_t p = pk + (size_t)ii + (size_t)d.IA * jj;
            const int c = c0 + 2, rr = r + 1;
            const double vel = mY[p];
            sAY[r][c0] = tt_hflux<S>(vel, sQ[rr - 1][c], sQ[rr][c], sQ[rr + 1][c], sQ[rr + 2][c], (a.nd ? a.nd[p + 2 * a.nds] : 0.0))
                       - tt_lflux(vel, Q0[rr][c], Q0[rr + 1][c]);
          }
        } else {
          ring_cell(k + 1, t - NXF - NYF);
        }
      }
    }
    __syncthreads();
    if (act) {
      const size_t p = d.pl * (size_t)k + pc;
      const int c2 = tx + 2, r2 = ty + 2;  // the cell in the two-cell-ring arrays
      const int c1 = tx + 1, r1 = ty + 1;  // ... and in the one-cell-ring arrays
      const double(*Q0k)[W2] = sQ0[k % 5];
      double az = 0.0;
      if (k != KE) {
        const double vel = mZ[p];
        az = tt_hz<S>(d, k, vel, w, a.nd ? a.nd[p] : 0.0) - tt_lflux(vel, Q0k[r2][c2], sQ0[(k + 1) % 5][r2][c2]);
      }
      const double ax = sAX[ty][tx + 1], axm = sAX[ty][tx];
      const double ay = sAY[ty + 1][tx], aym = sAY[ty][tx];
      const double pjpls = dt * ((fmax2(0.0, azm) - fmin2(0.0, az)) * a.RCDZ[k]
                               + (fmax2(0.0, axm) - fmin2(0.0, ax)) * a.RCDX[i]
                               + (fmax2(0.0, aym) - fmin2(0.0, ay)) * a.RCDY[j]);
      const double pjmns = dt * ((fmax2(0.0, az) - fmin2(0.0, azm)) * a.RCDZ[k]
                               + (fmax2(0.0, ax) - fmin2(0.0, axm)) * a.RCDX[i]
                               + (fmax2(0.0, ay) - fmin2(0.0, aym)) * a.RCDY[j]);
      const double pin = Q0k[r2][c2], plo = sPL[k % 3][r1][c1];
      double qmax, qmin;
      if constexpr (VECT) {
        double rw;
        int ka, kb;
        if (k == KS) { rw = (mZ[p]) * a.RCDZ[KS]; ka = KS + 1; kb = KS; }
        else if (k == KE) { rw = (mZ[p - d.pl]) * a.RCDZ[KE]; ka = KE; kb = KE - 1; }
        else { rw = (mZ[p] + mZ[p - d.pl]) * a.RCDZ[k]; ka = k + 1; kb = k - 1; }
        const double ru = (mX[p] + mX[p - 1]) * a.RCDX[i];
        const double rv = (mY[p] + mY[p - d.IA]) * a.RCDY[j];
        double f[18];
        get_fact(f, rw, ru, rv);
        // upstream (qa) / downstream (qb) 14-point combinations (COMMON.F90:2096-2155),
        // in the order of fct_qa / fct_qb; P(l, dj, di) is level l at (j+dj, i+di)
        auto qa = [&](auto P, int k1, int k0) {
          return f[17] * P(k1, 1, 1) + f[16] * P(k0, 1, 1) + f[15] * P(k1, 1, 0) + f[14] * P(k0, 1, 0)
               + f[13] * P(k1, 1, -1) + f[11] * P(k1, 0, 1) + f[10] * P(k0, 0, 1) + f[9] * P(k1, 0, 0)
               + f[7] * P(k1, 0, -1) + f[5] * P(k1, -1, 1) + f[4] * P(k0, -1, 1) + f[3] * P(k1, -1, 0)
               + f[1] * P(k1, -1, -1) + f[8] * P(k0, 0, 0);
        };
        auto qb = [&](auto P, int k1, int k0) {
          return f[17] * P(k1, -1, -1) + f[16] * P(k0, -1, -1) + f[15] * P(k1, -1, 0) + f[14] * P(k0, -1, 0)
               + f[13] * P(k1, -1, 1) + f[11] * P(k1, 0, -1) + f[10] * P(k0, 0, -1) + f[9] * P(k1, 0, 0)
               + f[7] * P(k1, 0, 1) + f[5] * P(k1, 1, -1) + f[4] * P(k0, -1, -1) + f[3] * P(k1, -1, 0)
               + f[1] * P(k1, 1, 1) + f[8] * P(k0, 0, 0);
        };
        auto PI = [&](int l, int dj, int di) { return sQ0[l % 5][r2 + dj][c2 + di]; };
        auto PL = [&](int l, int dj, int di) { return sPL[l % 3][r1 + dj][c1 + di]; };
        const double qa_in = qa(PI, ka, k);
        const double qb_in = qb(PI, kb, k);
        const double qa_lo = qa(PL, ka, k);
        const double qb_lo = qb(PL, kb, k);
        qmax = fmax2(fmax2(fmax2(fmax2(fmax2(pin, qa_in), qb_in), plo), qa_lo), qb_lo);
        qmin = fmin2(fmin2(fmin2(fmin2(fmin2(pin, qa_in), qb_in), plo), qa_lo), qb_lo);
      } else {  // 7-point bounds (COMMON.F90:2312-2515)
        const int kp = k == KE ? k : k + 1, km = k == KS ? k : k - 1;
        qmax = pin; qmin = pin;
        {
          const double v[7] = {Q0k[r2][c2], sQ0[kp % 5][r2][c2], sQ0[km % 5][r2][c2], Q0k[r2][c2 + 1], Q0k[r2][c2 - 1],
                               Q0k[r2 + 1][c2], Q0k[r2 - 1][c2]};
#pragma unroll
          for (int q = 0; q < 7; ++q) { qmax = fmax2(qmax, v[q]); qmin = fmin2(qmin, v[q]); }
        }
        {
          const double(*L)[W1] = sPL[k % 3];
          const double v[7] = {L[r1][c1], sPL[kp % 3][r1][c1], sPL[km % 3][r1][c1], L[r1][c1 + 1], L[r1][c1 - 1],
                               L[r1 + 1][c1], L[r1 - 1][c1]};
#pragma unroll
          for (int q = 0; q < 7; ++q) { qmax = fmax2(qmax, v[q]); qmin = fmin2(qmin, v[q]); }
        }
      }
      const double D = a.DENS[p];
      const double qjpls = (qmax - plo) * D;
      const double qjmns = (plo - qmin) * D;
      double zerosw = 0.5 - fsign(0.5, pjpls - EPS_RP);
      a.rjpls[p] = fmin2(1.0, qjpls * (1.0 - zerosw) / (pjpls - zerosw));
      zerosw = 0.5 - fsign(0.5, pjmns - EPS_RP);
      a.rjmns[p] = fmin2(1.0, qjmns * (1.0 - zerosw) / (pjmns - zerosw));
      azm = az;
    }
    w[0] = w[1]; w[1] = w[2]; w[2] = w[3];
    if (k + 3 < d.KA) w[3] = a.QTRC[pc + d.pl * (size_t)(k + 3)];
  }
}

template <int S, int TJ, bool VECT>
__global__ void __launch_bounds__(64 * TJ) trt_ratio(TtArgs a, unsigned ni_tiles, unsigned nblocks) {
  trt_ratio_body<S, TJ, VECT>(a, xcd_block(blockIdx.x, nblocks), ni_tiles);
}

// ---- FCT pass 2: limited anti-diffusive fluxes and the update ------------------
// Software-pipelined: every global load of level k+1 (R+/R- of level k+2) is
// issued into registers right after the level-k barrier, so that it is in
// flight during the face and column work of level k instead of being waited
// for at the start of level k+1.  Each load is issued unconditionally (an
// invalid position reads element 0 and is discarded by a select), so that no
// branch hides from the wait counters how many loads are outstanding.  This
// holds 176 VGPRs (one block per CU) and beats the same pass without the
// prefetch at two blocks per CU by 15 %; the ratio pass, heavier in
// arithmetic, keeps two blocks per CU without it (with its two staged planes
// loaded a level ahead it ran 18 % slower held to two blocks per CU, 6 VGPRs
// spilled, and 26 % slower at one block per CU).
// With LIM false the same march is the stage without FCT (tracer_fvm_heve.F90:
// 226-231): the high-order fluxes alone, no limiter arrays staged or read.
template <int S, int TJ, bool LIM = true>
__device__ __forceinline__ void trt_update_body(const TtArgs& a, unsigned lb, unsigned ni_tiles) {
  constexpr int W2 = 68, H2 = TJ + 4, W1 = 66, H1 = TJ + 2, NT = 64 * TJ;
  constexpr int NQ = (H2 * W2 + NT - 1) / NT, NR = (H1 * W1 + NT - 1) / NT;
  constexpr int NX = (TJ * 65 + NT - 1) / NT, NY = ((TJ + 1) * 64 + NT - 1) / NT;
  __shared__ double sQ[H2][W2];      // QTRC at level k, two-cell ring
  __shared__ double sQ0[H1][W1];     // QTRC0 at level k, one-cell ring
  __shared__ double sRP[3][H1][W1];  // R+ and R-, one-cell ring (slot = level % 3)
  __shared__ double sRM[3][H1][W1];
  __shared__ double sHX[TJ][65], sLX[TJ][65];
  __shared__ double sHY[TJ + 1][64], sLY[TJ + 1][64];
  const Dims& d = a.d;
  const int KS = d.KS, KE = d.KE;
  const int i0 = d.IS + 64 * (int)(lb % ni_tiles), j0 = d.JS + TJ * (int)(lb / ni_tiles);
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + 64 * ty;
  const int i = i0 + tx, j = j0 + ty;
  const bool act = i <= d.IE && j <= d.JE;
  const double* mZ = a.mflx;
  const double* mX = a.mflx + d.n3;
  const double* mY = a.mflx + 2 * d.n3;
  auto lim = [](double an, double rpb, double rma, double rpa, double rmb) {
    const double dirsw = 0.5 + fsign(0.5, an);
    return an * (1.0 - fmin2(rpb, rma) * (dirsw) - fmin2(rpa, rmb) * (1.0 - dirsw));
  };
  const size_t pc = (size_t)(act ? i : d.IS) + (size_t)d.IA * (act ? j : d.JS);
  // prefetch registers
  double pQ[NQ], pQ0[NR], pRP[NR], pRM[NR], pXm[NX], pXn[NX], pYm[NY], pYn[NY];
  double pZ = 0.0, pN = 0.0, pD0 = 0.0, pRQ = 0.0, pD = 1.0;
  auto cell = [&](int q, int W, int off, int k, bool inb) -> size_t {  // plane index of tile cell q
    const int r = q / W, c = q - r * W;
    const int ii = i0 - off + c, jj = j0 - off + r;
    return (inb && ii < d.IA && jj < d.JA) ? d.pl * (size_t)k + (size_t)ii + (size_t)d.IA * jj : 0;
  };
  auto issue = [&](int k) {
#pragma unroll
    for (int h = 0; h < NQ; ++h) {
      const int q = tid + h * NT;
      const size_t g = cell(q, W2, 2, k, q < H2 * W2);
      const double v = a.QTRC[g];
      pQ[h] = g ? v : 0.0;
    }
    if constexpr (LIM) {
      const bool r1 = k + 1 <= KE;
#pragma unroll
      for (int h = 0; h < NR; ++h) {
        const int q = tid + h * NT;
        const size_t g0 = cell(q, W1, 1, k, q < H1 * W1);
        const size_t g1 = cell(q, W1, 1, k + 1, r1 && q < H1 * W1);
        const double v0 = a.QTRC0[g0], vp = a.rjpls[g1], vm = a.rjmns[g1];
        pQ0[h] = g0 ? v0 : 0.0;
        pRP[h] = g1 ? vp : 0.0;
        pRM[h] = g1 ? vm : 0.0;
      }
    }
    const size_t pk = d.pl * (size_t)k;
#pragma unroll
    for (int h = 0; h < NX; ++h) {
      const int t = tid + h * NT;
      const int r = t / 65, f = t - r * 65;
      const int ii = i0 - 1 + f, jj = j0 + r;
      const bool ok = t < TJ * 65 && ii <= d.IE && jj <= d.JE;
      const size_t p = ok ? pk + (size_t)ii + (size_t)d.IA * jj : 0;
      const double vm = mX[p], vn = a.nd ? a.nd[p + a.nds] : 0.0;
      pXm[h] = ok ? vm : 0.0;
      pXn[h] = ok ? vn : 0.0;
    }
#pragma unroll
    for (int h = 0; h < NY; ++h) {
      const int t = tid + h * NT;
      const int r = t / 64, c0 = t - r * 64;
      const int ii = i0 + c0, jj = j0 - 1 + r;
      const bool ok = t < (TJ + 1) * 64 && ii <= d.IE && jj <= d.JE;
      const size_t p = ok ? pk + (size_t)ii + (size_t)d.IA * jj : 0;
      const double vm = mY[p], vn = a.nd ? a.nd[p + 2 * a.nds] : 0.0;
      pYm[h] = ok ? vm : 0.0;
      pYn[h] = ok ? vn : 0.0;
    }
    const size_t p = pk + pc;
    const double vz = mZ[p], vn = a.nd ? a.nd[p] : 0.0, v0 = a.DENS0[p], vq = a.RHOQ_t ? a.RHOQ_t[p] : 0.0, vd = a.DENS[p];
    pZ = (k != KE) ? vz : 0.0;
    pN = (k != KE) ? vn : 0.0;
    pD0 = v0; pRQ = vq; pD = vd;
  };

  if constexpr (LIM) {
    tt_load<H1, W1>(sRP[KS % 3], a.rjpls, d, KS, i0 - 1, j0 - 1, tid, NT);
    tt_load<H1, W1>(sRM[KS % 3], a.rjmns, d, KS, i0 - 1, j0 - 1, tid, NT);
  }
  double w[4];
  for (int m = 0; m < 4; ++m) w[m] = a.QTRC[pc + d.pl * (size_t)(KS - 1 + m)];
  double q0c = a.QTRC0[pc + d.pl * (size_t)KS], q0n = a.QTRC0[pc + d.pl * (size_t)(KS + 1)];
  double hzm = 0.0, lzm = 0.0;
  issue(KS);
  for (int k = KS; k <= KE; ++k) {
    // (a) this level's tiles from the prefetch registers
#pragma unroll
    for (int h = 0; h < NQ; ++h) {
      const int q = tid + h * NT;
      if (q < H2 * W2) sQ[q / W2][q % W2] = pQ[h];
    }
    if constexpr (LIM) {
#pragma unroll
      for (int h = 0; h < NR; ++h) {
        const int q = tid + h * NT;
        if (q < H1 * W1) {
          sQ0[q / W1][q % W1] = pQ0[h];
          if (k + 1 <= KE) { sRP[(k + 1) % 3][q / W1][q % W1] = pRP[h]; sRM[(k + 1) % 3][q / W1][q % W1] = pRM[h]; }
        }
      }
    }
    double cXm[NX], cXn[NX], cYm[NY], cYn[NY];
#pragma unroll
    for (int h = 0; h < NX; ++h) { cXm[h] = pXm[h]; cXn[h] = pXn[h]; }
#pragma unroll
    for (int h = 0; h < NY; ++h) { cYm[h] = pYm[h]; cYn[h] = pYn[h]; }
    const double cZ = pZ, cN = pN, cD0 = pD0, cRQ = pRQ, cD = pD;
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
    if (k + 1 <= KE) issue(k + 1);
    __builtin_amdgcn_sched_barrier(0);
    // (b) faces
    {
      const double(*RP)[W1] = sRP[k % 3];
      const double(*RM)[W1] = sRM[k % 3];
#pragma unroll
      for (int h = 0; h < NX; ++h) {
        const int t = tid + h * NT;
        const int r = t / 65, f = t - r * 65;
        const int ii = i0 - 1 + f, jj = j0 + r;
        if (t < TJ * 65 && ii <= d.IE && jj <= d.JE) {
          const int c = f + 1, rr = r + 2;
          const int c1 = f, r1 = r + 1;
          const double vel = cXm[h];
          const double hf = tt_hflux<S>(vel, sQ[rr][c - 1], sQ[rr][c], sQ[rr][c + 1], sQ[rr][c + 2], cXn[h]);
          sHX[r][f] = hf;
          if constexpr (LIM) {
            const double an = hf - tt_lflux(vel, sQ0[r1][c1], sQ0[r1][c1 + 1]);
            sLX[r][f] = lim(an, RP[r1][c1 + 1], RM[r1][c1], RP[r1][c1], RM[r1][c1 + 1]);
          }
        }
      }
#pragma unroll
      for (int h = 0; h < NY; ++h) {
        const int t = tid + h * NT;
        const int r = t / 64, c0 = t - r * 64;
        const int ii = i0 + c0, jj = j0 - 1 + r;
        if (t < (TJ + 1) * 64 && ii <= d.IE && jj <= d.JE) {
          const int c = c0 + 2, rr = r + 1;
          const int c1 = c0 + 1, r1 = r;
          const double vel = cYm[h];
          const double hf = tt_hflux<S>(vel, sQ[rr - 1][c], sQ[rr][c], sQ[rr + 1][c], sQ[rr + 2][c], cYn[h]);
          sHY[r][c0] = hf;
          if constexpr (LIM) {
            const double an = hf - tt_lflux(vel, sQ0[r1][c1], sQ0[r1 + 1][c1]);
            sLY[r][c0] = lim(an, RP[r1 + 1][c1], RM[r1][c1], RP[r1][c1], RM[r1 + 1][c1]);
          }
        }
      }
    }
    __syncthreads();
    // (c) the column
    if (act) {
      const size_t p = d.pl * (size_t)k + pc;
      const int c1 = tx + 1, r1 = ty + 1;
      double hz = 0.0, lz = 0.0;
      if (k != KE) {
        hz = tt_hz<S>(d, k, cZ, w, cN);
        if constexpr (LIM) {
          const double an = hz - tt_lflux(cZ, q0c, q0n);
          lz = lim(an, sRP[(k + 1) % 3][r1][c1], sRM[k % 3][r1][c1], sRP[k % 3][r1][c1], sRM[(k + 1) % 3][r1][c1]);
        }
      }
      if constexpr (LIM) {
        a.QTRCo[p] = (q0c * cD0
                      + a.dtl * (-((hz - lz - hzm + lzm) * a.RCDZ[k]
                                   + (sHX[ty][tx + 1] - sLX[ty][tx + 1] - sHX[ty][tx] + sLX[ty][tx]) * a.RCDX[i]
                                   + (sHY[ty + 1][tx] - sLY[ty + 1][tx] - sHY[ty][tx] + sLY[ty][tx]) * a.RCDY[j])
                                 + cRQ))
                   / cD;
      } else {  // tracer_nofct's expression; its metric factors are 1.0 here
        a.QTRCo[p] = (q0c * cD0
                      + a.dtl * (-((hz - hzm) * a.RCDZ[k] + (sHX[ty][tx + 1] - sHX[ty][tx]) * a.RCDX[i]
                                   + (sHY[ty + 1][tx] - sHY[ty][tx]) * a.RCDY[j])
                                 + cRQ))
                   / cD;
      }
      hzm = hz;
      lzm = lz;
    }
    w[0] = w[1]; w[1] = w[2]; w[2] = w[3];
    if (k + 3 < d.KA) w[3] = a.QTRC[pc + d.pl * (size_t)(k + 3)];
    q0c = q0n;
    if (k + 2 < d.KA) q0n = a.QTRC0[pc + d.pl * (size_t)(k + 2)];
    __syncthreads();
  }
}

template <int S, int TJ>
__global__ void __launch_bounds__(64 * TJ) trt_nofct(TtArgs a, unsigned ni_tiles, unsigned nblocks) {
  trt_update_body<S, TJ, false>(a, xcd_block(blockIdx.x, nblocks), ni_tiles);
}

template <int S, int TJ>
__global__ void __launch_bounds__(64 * TJ) trt_update(TtArgs a, unsigned ni_tiles, unsigned nblocks) {
  trt_update_body<S, TJ>(a, xcd_block(blockIdx.x, nblocks), ni_tiles);
}

// both passes over the tiles of many members (batch.h tables; ni[m] the
// member's tiles across)
template <int KIND, int S, int TJ, bool VECT>
__global__ void __launch_bounds__(64 * TJ) trt_batch(const TtArgs* __restrict__ tab, const unsigned* __restrict__ first,
                                                     const unsigned* __restrict__ ni, int nmem, unsigned total) {
  for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
    const unsigned lt = xcd_block(t, total);
    const int m = batch_member(first, nmem, lt);
    if constexpr (KIND == 0) trt_ratio_body<S, TJ, VECT>(tab[m], lt - first[m], ni[m]);
    else trt_update_body<S, TJ>(tab[m], lt - first[m], ni[m]);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
template <int S, int TJR, int TJU>
static hipError_t trt_launch(const Ctx& c, const TtArgs& a, bool along, hipStream_t s) {
  const Dims& d = c.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1;
  const unsigned ni = (imax + 63) / 64, nbr = ni * ((jmax + TJR - 1) / TJR), nbu = ni * ((jmax + TJU - 1) / TJU);
  prof_start(c, "fct_ratio", s);
  if (along) hipLaunchKernelGGL((trt_ratio<S, TJR, true>), dim3(nbr), dim3(64, TJR), 0, s, a, ni, nbr);
  else hipLaunchKernelGGL((trt_ratio<S, TJR, false>), dim3(nbr), dim3(64, TJR), 0, s, a, ni, nbr);
  prof_stop(c, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  double* r[2] = {a.rjpls, a.rjmns};
  e = halo_fill(c, r, 2, s);
  if (e != hipSuccess) return e;
  prof_start(c, "fct_update", s);
  hipLaunchKernelGGL((trt_update<S, TJU>), dim3(nbu), dim3(64, TJU), 0, s, a, ni, nbu);
  prof_stop(c, s);
  return hipGetLastError();
}

// The tiles pay off where the column march has blocks enough for every CU
// (two each, the LDS limit); smaller grids keep the per-point kernels, whose
// launches spread over k as well.
constexpr int TT_TJ = 8;   // rows of the update tile (and of the size criterion)
constexpr int TT_TJR = 8;   // rows of the ratio tile (16 rows: 1.94 ms, 4 rows: 2.39, against 1.53 at 8 rows, 512x512x128)
bool tracer_tile_supported(const Ctx& c) {
  const int st = c.scheme_tracer;
  if (c.terrain || !(st == UD3 || st == CD4 || st == UD3KOREN || st == UD1 || st == CD2)) return false;
  int ncu = 256;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess) ncu = 256;
  const Dims& d = c.d;
  const long nb = (long)((d.IE - d.IS + 1 + 63) / 64) * ((d.JE - d.JS + 1 + TT_TJ - 1) / TT_TJ);
  return c.tracer_tile == 2 || nb >= 2L * ncu;
}

// the stage without FCT on the same tiles (8 rows, as the update)
template <int S>
static hipError_t trn_launch(const Ctx& c, const TtArgs& a, hipStream_t s) {
  const Dims& d = c.d;
  const int imax = d.IE - d.IS + 1, jmax = d.JE - d.JS + 1;
  const unsigned ni = (imax + 63) / 64, nb = ni * ((jmax + TT_TJ - 1) / TT_TJ);
  prof_start(c, "tracer_nofct", s);
  hipLaunchKernelGGL((trt_nofct<S, TT_TJ>), dim3(nb), dim3(64, TT_TJ), 0, s, a, ni, nb);
  prof_stop(c, s);
  return hipGetLastError();
}

hipError_t tracer_nofct_tile(const Ctx& c, const TracerIO& io, hipStream_t s) {
  TtArgs a;
  a.d = c.d;
  a.QTRC = io.QTRC; a.QTRC0 = io.QTRC0; a.RHOQ_t = io.RHOQ_t; a.DENS0 = io.DENS0; a.DENS = io.DENS;
  a.mflx = io.mflx_hi; a.nd = io.num_diff; a.nds = c.d.n3;
  a.RCDZ = c.g1.RCDZ; a.RCDX = c.g1.RCDX; a.RCDY = c.g1.RCDY;
  a.dtl = io.dtl; a.QTRCo = io.QTRCo;
  a.rjpls = nullptr; a.rjmns = nullptr;
  if (io.flag_fct) return hipErrorNotSupported;
  switch (c.scheme_tracer) {
    case UD3: return trn_launch<UD3>(c, a, s);
    case CD4: return trn_launch<CD4>(c, a, s);
    case UD3KOREN: return trn_launch<UD3KOREN>(c, a, s);
    case UD1: return trn_launch<UD1>(c, a, s);
    case CD2: return trn_launch<CD2>(c, a, s);
    default: return hipErrorNotSupported;
  }
}

hipError_t tracer_stage_tile(const Ctx& c, const TracerIO& io, double* rjpls, double* rjmns, hipStream_t s) {
  TtArgs a;
  a.d = c.d;
  a.QTRC = io.QTRC; a.QTRC0 = io.QTRC0; a.RHOQ_t = io.RHOQ_t; a.DENS0 = io.DENS0; a.DENS = io.DENS;
  a.mflx = io.mflx_hi; a.nd = io.num_diff; a.nds = c.d.n3;
  a.RCDZ = c.g1.RCDZ; a.RCDX = c.g1.RCDX; a.RCDY = c.g1.RCDY;
  a.dtl = io.dtl; a.QTRCo = io.QTRCo;
  a.rjpls = rjpls; a.rjmns = rjmns;
  const bool along = io.flag_along != 0;
  if (!io.flag_fct) return hipErrorNotSupported;
  switch (c.scheme_tracer) {
    case UD3: return trt_launch<UD3, TT_TJR, TT_TJ>(c, a, along, s);
    case CD4: return trt_launch<CD4, TT_TJR, TT_TJ>(c, a, along, s);
    case UD3KOREN: return trt_launch<UD3KOREN, TT_TJR, TT_TJ>(c, a, along, s);
    case UD1: return trt_launch<UD1, TT_TJR, TT_TJ>(c, a, along, s);
    case CD2: return trt_launch<CD2, TT_TJR, TT_TJ>(c, a, along, s);
    default: return hipErrorNotSupported;
  }
}

// the FCT passes of every member on tiles (the members' stage-3 arguments;
// flat terrain, one two-cell scheme, one bound kind)
hipError_t trt_batch_build(const Ctx* const* cs, const TracerIO* ios, int n, BatchBuf& b) {
  if (n <= 0) return hipErrorInvalidValue;
  std::vector<TtArgs> args((size_t)n);
  std::vector<unsigned> first((size_t)n + 1, 0), ni((size_t)n, 0);
  const int scheme = cs[0]->scheme_tracer, along = ios[0].flag_along;
  for (int m = 0; m < n; ++m) {
    const Ctx& c = *cs[m];
    const TracerIO& io = ios[m];
    const int st = c.scheme_tracer;
    if (c.terrain || st != scheme || io.flag_along != along || !(st == UD3 || st == CD4 || st == UD3KOREN || st == UD1 || st == CD2))
      return hipErrorNotSupported;
    TtArgs& a = args[(size_t)m];
    a.d = c.d;
    a.QTRC = io.QTRC; a.QTRC0 = io.QTRC0; a.RHOQ_t = io.RHOQ_t; a.DENS0 = io.DENS0; a.DENS = io.DENS;
    a.mflx = io.mflx_hi; a.nd = io.num_diff; a.nds = c.d.n3;
    a.RCDZ = c.g1.RCDZ; a.RCDX = c.g1.RCDX; a.RCDY = c.g1.RCDY;
    a.dtl = io.dtl; a.QTRCo = io.QTRCo;
    a.rjpls = c.scr[1]; a.rjmns = c.scr[2];
    const Dims& d = c.d;
    ni[(size_t)m] = (unsigned)((d.IE - d.IS + 1 + 63) / 64);
    first[(size_t)m + 1] = first[(size_t)m] + ni[(size_t)m] * (unsigned)((d.JE - d.JS + 1 + TT_TJ - 1) / TT_TJ);
  }
  const hipError_t e = batch_upload(args, first, ni, b);
  b.scheme = scheme; b.tj = along;
  return e;
}

template <int KIND, int S>
static hipError_t trt_batch_t(const BatchBuf& b, hipStream_t s) {
  const TtArgs* tab;
  const unsigned *first, *ni;
  batch_tables(b, tab, first, ni);
  if (KIND == 0 && b.tj)
    hipLaunchKernelGGL((trt_batch<0, S, TT_TJ, true>), dim3(b.total), dim3(64, TT_TJ), 0, s, tab, first, ni, b.nmem, b.total);
  else if (KIND == 0)
    hipLaunchKernelGGL((trt_batch<0, S, TT_TJ, false>), dim3(b.total), dim3(64, TT_TJ), 0, s, tab, first, ni, b.nmem, b.total);
  else
    hipLaunchKernelGGL((trt_batch<1, S, TT_TJ, false>), dim3(b.total), dim3(64, TT_TJ), 0, s, tab, first, ni, b.nmem, b.total);
  return hipGetLastError();
}
template <int KIND>
static hipError_t trt_batch_k(const BatchBuf& b, hipStream_t s) {
  switch (b.scheme) {
    case UD3: return trt_batch_t<KIND, UD3>(b, s);
    case UD3KOREN: return trt_batch_t<KIND, UD3KOREN>(b, s);
    case CD2: return trt_batch_t<KIND, CD2>(b, s);
    case UD1: return trt_batch_t<KIND, UD1>(b, s);
    default: return trt_batch_t<KIND, CD4>(b, s);
  }
}
hipError_t trt_batch_launch(const BatchBuf& b, int kind, hipStream_t s) {
  if (!b.dev) return hipErrorInvalidValue;
  return kind == 0 ? trt_batch_k<0>(b, s) : trt_batch_k<1>(b, s);
}

}  // namespace shp
