/*
 * scale_oracle_adj.c -- reverse-mode derivative (adjoint) of the oracle's
 * HEVI stage, numerical filter and large step, for flat terrain.
 *
 * TEST INFRASTRUCTURE ONLY (see scale_oracle.h): the checker of the HIP
 * adjoint.  The reference has no adjoint; this file is the transpose of the
 * Jacobian of the oracle's own forward functions in scale_oracle.c, derived
 * statement by statement from them.  It is pinned by finite differences of
 * those forward functions (tests/test_adjoint_oracle.py), not by the
 * reference.
 *
 * The upwind switch sign(vel) of the UD3 face values is held fixed: the
 * forward is differentiable wherever no face velocity (and no vertical
 * momentum used for the face potential temperature) is exactly zero, and
 * there this is its derivative.
 *
 * Scope: flat terrain (GSQRT = 1, J13G = J23G = 0, MAPF = 1, J33G = 1), in
 * which the terrain terms of the forward are exact zeros with zero
 * derivative; UD3 or CD4; divdmp_coef = 0; ND_ORDER = 1 with the reference
 * state; one short step per large step; periodic lateral boundaries.
 *
 * Reference files (relative to the reference root):
 *   HEVI   = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short_fvm_hevi.F90
 *   COMMON = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90
 *   RK4    = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tinteg_short_rk4.F90
 *   LARGE  = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_large_fvm_heve.F90
 */
#include "scale_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { ZDIR = 0, XDIR, YDIR };
enum { I_DENS = 0, I_MOMZ, I_MOMX, I_MOMY, I_RHOT };

static const double GRAV = 9.80665;
static const double Rdry = 287.04;
static const double CPdry = 1004.64;
static const double P00 = 100000.0;
static const double F2 = 0.5;
static const double F31 = -1.0 / 12.0;
static const double F32 = 7.0 / 12.0;
static const double F33 = 3.0 / 12.0;
static const double F41 = 7.0 / 12.0;
static const double F42 = -1.0 / 12.0;

#define GRID_LOCALS                                                        \
  const int KA = g->KA, IA = g->IA, JA = g->JA;                            \
  const int KS = g->KS, KE = g->KE, IS = g->IS, IE = g->IE, JS = g->JS, JE = g->JE; \
  const size_t N3 = (size_t)KA * IA * JA;                                  \
  (void)KA; (void)IA; (void)JA; (void)KS; (void)KE; (void)IS; (void)IE; (void)JS; (void)JE; (void)N3

#define I3(k, i, j) ((size_t)(k) + (size_t)KA * ((size_t)(i) + (size_t)IA * (size_t)(j)))

static double *zalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }

static int is_flat(const orc_grid *g, const orc_metric *m) {
  GRID_LOCALS;
  if (m->J33G != 1.0) return 0;
  for (size_t q = 0; q < 7 * N3; ++q)
    if (m->GSQRT[q] != 1.0 || m->J13G[q] != 0.0 || m->J23G[q] != 0.0) return 0;
  for (size_t q = 0; q < (size_t)IA * JA * 8; ++q)
    if (m->MAPF[q] != 1.0) return 0;
  return 1;
}

/* d(face value)/d(qm1, q0, q1, q2) for a fixed sign of vel; kind 0 interior
 * (fv_int), 1 lowest face (fv_lo: q0..q2), 2 highest face (fv_hi: qm1..q1) */
static void fv_weights(int s, int kind, double vel, double w[4]) {
  const double sg = copysign(1.0, vel), sh = copysign(0.5, vel);
  w[0] = w[1] = w[2] = w[3] = 0.0;
  if (s == ORC_CD4) {
    if (kind == 0) { w[0] = F42; w[1] = F41; w[2] = F41; w[3] = F42; }
    else { w[1] = F2; w[2] = F2; }
    return;
  }
  if (kind == 0) {
    w[0] = F31 + sg * F31; w[1] = F32 + sg * F33; w[2] = F32 - sg * F33; w[3] = F31 - sg * F31;
  } else if (kind == 1) {
    w[1] = F2 * (0.5 + sh) + (2.0 / 6.0) * (0.5 - sh);
    w[2] = F2 * (0.5 + sh) + (5.0 / 6.0) * (0.5 - sh);
    w[3] = -(1.0 / 6.0) * (0.5 - sh);
  } else {
    w[0] = -(1.0 / 6.0) * (0.5 + sh);
    w[1] = (5.0 / 6.0) * (0.5 + sh) + F2 * (0.5 - sh);
    w[2] = (2.0 / 6.0) * (0.5 + sh) + F2 * (0.5 - sh);
  }
}

/* Face value of q at the face whose q0 is q[p] along stride st, and its
 * adjoint: qb[p + (m-1)*st] += scale * dfv/dq_m.  Returns the face value. */
static double face_adj(int s, int kind, const double *q, double *qb, size_t p, size_t st, double vel, double scale) {
  double w[4];
  fv_weights(s, kind, vel, w);
  const double qm1 = kind == 1 ? 0.0 : q[p - st];
  const double q0 = q[p], q1 = q[p + st];
  const double q2 = kind == 2 ? 0.0 : q[p + 2 * st];
  if (kind != 1) qb[p - st] += scale * w[0];
  qb[p] += scale * w[1];
  qb[p + st] += scale * w[2];
  if (kind != 2) qb[p + 2 * st] += scale * w[3];
  return orc_face_value(s, kind, qm1, q0, q1, q2, vel);
}

static int zkind(int k, int KS, int KE) { return k == KS ? 1 : (k == KE - 1 ? 2 : 0); }

/* ---- HEVI stage (HEVI.F90:128-1089), reverse of scale_oracle.c's forward -- */
int orc_adj_tstep_short_hevi(const orc_grid *g, int s, const orc_metric *m,
    const double *DENS0, const double *MOMZ0, const double *MOMX0, const double *MOMY0, const double *RHOT0,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef,
    const double *REF_dens, const double *REF_rhot, double dtrk,
    const double *DENS_RK_b, const double *MOMZ_RK_b, const double *MOMX_RK_b, const double *MOMY_RK_b, const double *RHOT_RK_b,
    double *DENS0_b, double *MOMZ0_b, double *MOMX0_b, double *MOMY0_b, double *RHOT0_b,
    double *DENS_b, double *MOMZ_b, double *MOMX_b, double *MOMY_b, double *RHOT_b,
    double *num_diff_b, double *DPRES0_b, double *RT2P_b, double *REF_rhot_b) {
  GRID_LOCALS;
  if ((s != ORC_UD3 && s != ORC_CD4) || !is_flat(g, m) || KA > 512) return -1;
  const double *CDZ = m->CDZ, *RCDZ = m->RCDZ, *RCDX = m->RCDX, *RCDY = m->RCDY;
  const double *RFDZ = m->RFDZ, *RFDX = m->RFDX, *RFDY = m->RFDY;
  const double dt = dtrk;
  const size_t n3 = N3, sI = (size_t)KA, sJ = (size_t)KA * IA;
#define ND(k, i, j, v, d) num_diff[I3(k, i, j) + n3 * ((size_t)(v) + 5 * (size_t)(d))]
#define NDB(k, i, j, v, d) num_diff_b[I3(k, i, j) + n3 * ((size_t)(v) + 5 * (size_t)(d))]
#define COR(i, j) CORIOLI[(size_t)(i) + (size_t)IA * (size_t)(j)]

  /* forward pass with its intermediates recorded */
  orc_hevi_trace tr = {zalloc(n3), zalloc(n3), zalloc(n3), zalloc(n3), zalloc(n3), zalloc(n3), zalloc(n3)};
  double *o[5], *mf = zalloc(3 * n3), *tf = zalloc(3 * n3), *DDIV = zalloc(n3), *wd0 = zalloc(KA);
  for (int v = 0; v < 5; ++v) o[v] = zalloc(n3);
  orc_set_hevi_trace(&tr);
  int rc = orc_tstep_short_hevi(g, s, o[0], o[1], o[2], o[3], o[4], mf, tf, DENS0, MOMZ0, MOMX0, MOMY0, RHOT0,
                                DENS, MOMZ, MOMX, MOMY, RHOT, DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t,
                                DPRES0, RT2P, CORIOLI, num_diff, wdamp_coef ? wdamp_coef : wd0, 0.0, DDIV, m,
                                REF_dens, REF_rhot, dtrk);
  orc_set_hevi_trace(NULL);
  const double *POTT = tr.POTT, *St = tr.St;

  double *POTTb = zalloc(n3), *Stb = zalloc(n3), *Srb = zalloc(n3), *Swb = zalloc(n3), *DPRESb = zalloc(n3);
  double *qZb = zalloc(n3), *qXb = zalloc(n3), *qYb = zalloc(n3);
  double *tZb = zalloc(n3), *tXb = zalloc(n3), *tYb = zalloc(n3);
  double *mZb = zalloc(n3), *mXb = zalloc(n3), *mYb = zalloc(n3);

  /* ---- momentum y (HEVI.F90:936-1047) ---- */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double b = MOMY_RK_b[p];
        MOMY0_b[p] += b;
        const double T = dt * b;
        qZb[p] -= T * RCDZ[k]; qZb[p - 1] += T * RCDZ[k];
        qXb[p] -= T * RCDX[i]; qXb[p - sI] += T * RCDX[i];
        qYb[p] -= T * RFDY[j]; qYb[p - sJ] += T * RFDY[j];
        DPRESb[p + sJ] -= T * RFDY[j]; DPRESb[p] += T * RFDY[j];
        const double c = -0.125 * (COR(i, j + 1) + COR(i, j)) * T;
        MOMX_b[p + sJ] += c; MOMX_b[p] += c; MOMX_b[p - sI + sJ] += c; MOMX_b[p - sI] += c;
      }
  /* fluxZ_XVZ (UD3.F90:1173-1265) */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE - 1; ++k) {
        const size_t p = I3(k, i, j);
        const double f1 = CDZ[k] / (CDZ[k] + CDZ[k + 1]), f2 = CDZ[k + 1] / (CDZ[k] + CDZ[k + 1]);
        const double den = f1 * 0.5 * (DENS[p + 1] + DENS[p + 1 + sJ]) + f2 * 0.5 * (DENS[p] + DENS[p + sJ]);
        const double vel = 0.5 * (MOMZ[p] + MOMZ[p + sJ]) / den;
        const double fb = qZb[p];
        NDB(k, i, j, I_MOMY, ZDIR) += fb;
        const double velb = fb * face_adj(s, zkind(k, KS, KE), MOMY, MOMY_b, p, 1, vel, fb * vel);
        MOMZ_b[p] += 0.5 * velb / den; MOMZ_b[p + sJ] += 0.5 * velb / den;
        const double denb = -velb * vel / den;
        DENS_b[p + 1] += 0.5 * f1 * denb; DENS_b[p + 1 + sJ] += 0.5 * f1 * denb;
        DENS_b[p] += 0.5 * f2 * denb; DENS_b[p + sJ] += 0.5 * f2 * denb;
      }
  /* fluxX_XVZ (UD3.F90:1463-1512) */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS - 1; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double den = 0.25 * (DENS[p] + DENS[p + sI] + DENS[p + sJ] + DENS[p + sI + sJ]);
        const double vel = 0.5 * (MOMX[p] + MOMX[p + sJ]) / den;
        const double fb = qXb[p];
        NDB(k, i, j, I_MOMY, XDIR) += fb;
        const double velb = fb * face_adj(s, 0, MOMY, MOMY_b, p, sI, vel, fb * vel);
        MOMX_b[p] += 0.5 * velb / den; MOMX_b[p + sJ] += 0.5 * velb / den;
        const double denb = -velb * vel / den;
        DENS_b[p] += 0.25 * denb; DENS_b[p + sI] += 0.25 * denb; DENS_b[p + sJ] += 0.25 * denb; DENS_b[p + sI + sJ] += 0.25 * denb;
      }
  /* fluxY_XVZ (UD3.F90:1536-1587): flux(k,i,j-1) at loop index j */
  for (int j = JS; j <= JE + 1; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j), q = p - sJ;
        const double vel = 0.5 * (MOMY[p] + MOMY[q]) / DENS[p];
        const double fb = qYb[q];
        NDB(k, i, j, I_MOMY, YDIR) += fb;
        const double velb = fb * face_adj(s, 0, MOMY, MOMY_b, q, sJ, vel, fb * vel);
        MOMY_b[p] += 0.5 * velb / DENS[p]; MOMY_b[q] += 0.5 * velb / DENS[p];
        DENS_b[p] -= velb * vel / DENS[p];
      }

  /* ---- momentum x (HEVI.F90:816-927) ---- */
  memset(qZb, 0, n3 * sizeof(double)); memset(qXb, 0, n3 * sizeof(double)); memset(qYb, 0, n3 * sizeof(double));
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double b = MOMX_RK_b[p];
        MOMX0_b[p] += b;
        const double T = dt * b;
        qZb[p] -= T * RCDZ[k]; qZb[p - 1] += T * RCDZ[k];
        qXb[p] -= T * RFDX[i]; qXb[p - sI] += T * RFDX[i];
        qYb[p] -= T * RCDY[j]; qYb[p - sJ] += T * RCDY[j];
        DPRESb[p + sI] -= T * RFDX[i]; DPRESb[p] += T * RFDX[i];
        const double c = 0.125 * (COR(i + 1, j) + COR(i, j)) * T;
        MOMY_b[p + sI] += c; MOMY_b[p] += c; MOMY_b[p + sI - sJ] += c; MOMY_b[p - sJ] += c;
      }
  /* fluxZ_UYZ (UD3.F90:753-845) */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE - 1; ++k) {
        const size_t p = I3(k, i, j);
        const double f1 = CDZ[k] / (CDZ[k] + CDZ[k + 1]), f2 = CDZ[k + 1] / (CDZ[k] + CDZ[k + 1]);
        const double den = f1 * 0.5 * (DENS[p + 1] + DENS[p + 1 + sI]) + f2 * 0.5 * (DENS[p] + DENS[p + sI]);
        const double vel = 0.5 * (MOMZ[p] + MOMZ[p + sI]) / den;
        const double fb = qZb[p];
        NDB(k, i, j, I_MOMX, ZDIR) += fb;
        const double velb = fb * face_adj(s, zkind(k, KS, KE), MOMX, MOMX_b, p, 1, vel, fb * vel);
        MOMZ_b[p] += 0.5 * velb / den; MOMZ_b[p + sI] += 0.5 * velb / den;
        const double denb = -velb * vel / den;
        DENS_b[p + 1] += 0.5 * f1 * denb; DENS_b[p + 1 + sI] += 0.5 * f1 * denb;
        DENS_b[p] += 0.5 * f2 * denb; DENS_b[p + sI] += 0.5 * f2 * denb;
      }
  /* fluxX_UYZ (UD3.F90:1043-1097): flux(k,i-1,j) at loop index i */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE + 1; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j), q = p - sI;
        const double vel = 0.5 * (MOMX[p] + MOMX[q]) / DENS[p];
        const double fb = qXb[q];
        NDB(k, i, j, I_MOMX, XDIR) += fb;
        const double velb = fb * face_adj(s, 0, MOMX, MOMX_b, q, sI, vel, fb * vel);
        MOMX_b[p] += 0.5 * velb / DENS[p]; MOMX_b[q] += 0.5 * velb / DENS[p];
        DENS_b[p] -= velb * vel / DENS[p];
      }
  /* fluxY_UYZ (UD3.F90:1101-1150) */
  for (int j = JS - 1; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double den = 0.25 * (DENS[p] + DENS[p + sI] + DENS[p + sJ] + DENS[p + sI + sJ]);
        const double vel = 0.5 * (MOMY[p] + MOMY[p + sI]) / den;
        const double fb = qYb[p];
        NDB(k, i, j, I_MOMX, YDIR) += fb;
        const double velb = fb * face_adj(s, 0, MOMX, MOMX_b, p, sJ, vel, fb * vel);
        MOMY_b[p] += 0.5 * velb / den; MOMY_b[p + sI] += 0.5 * velb / den;
        const double denb = -velb * vel / den;
        DENS_b[p] += 0.25 * denb; DENS_b[p + sI] += 0.25 * denb; DENS_b[p + sJ] += 0.25 * denb; DENS_b[p + sI + sJ] += 0.25 * denb;
      }

  /* ---- vertical implicit block (HEVI.F90:688-804) ---- */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      double Cpb[512], PTb[512], Ab[512], A[512], Cb[512], cp[512], dp[512];
      const size_t c0 = I3(0, i, j);
      const double *PT = tr.PT + c0, *F1 = tr.F1 + c0, *F2v = tr.F2 + c0, *F3 = tr.F3 + c0, *Cp = tr.Cp + c0;
      for (int k = 0; k < KA; ++k) { Cpb[k] = 0.0; PTb[k] = 0.0; Ab[k] = 0.0; }
      for (int k = KS; k <= KE - 1; ++k) {
        const double b = MOMZ_RK_b[c0 + k];
        MOMZ0_b[c0 + k] += b; Cpb[k] += b; MOMZ_b[c0 + k] -= b;
      }
      for (int k = KS; k <= KE; ++k) {
        const size_t p = c0 + k;
        const double bd = DENS_RK_b[p], bt = RHOT_RK_b[p];
        DENS0_b[p] += bd; Srb[p] += dt * bd;
        RHOT0_b[p] += bt; Stb[p] += dt * bt;
        const double r = dt * RCDZ[k];
        if (k <= KE - 1) { /* -C(k)*RCDZ, -C(k)*PT(k)*RCDZ */
          Cpb[k] -= r * bd;
          Cpb[k] -= r * PT[k] * bt; PTb[k] -= r * Cp[k] * bt;
        }
        if (k >= KS + 1) { /* +C(k-1)*RCDZ, +C(k-1)*PT(k-1)*RCDZ */
          Cpb[k - 1] += r * bd;
          Cpb[k - 1] += r * PT[k - 1] * bt; PTb[k - 1] += r * Cp[k - 1] * bt;
        }
      }
      /* C = M^-1 Crhs  =>  Crhs_b = M^-T C_b, with M(k,k-1)=F3(k), M(k,k)=F2(k), M(k,k+1)=F1(k) */
      {
        double den = F2v[KS];
        cp[KS] = (KS < KE - 1 ? F3[KS + 1] : 0.0) / den;
        dp[KS] = Cpb[KS] / den;
        for (int k = KS + 1; k <= KE - 1; ++k) {
          den = F2v[k] - F1[k - 1] * cp[k - 1];
          cp[k] = (k < KE - 1 ? F3[k + 1] : 0.0) / den;
          dp[k] = (Cpb[k] - F1[k - 1] * dp[k - 1]) / den;
        }
        Cb[KE - 1] = dp[KE - 1];
        for (int k = KE - 2; k >= KS; --k) Cb[k] = dp[k] - cp[k] * Cb[k + 1];
      }
      for (int k = KS; k <= KE; ++k) A[k] = dt * dt * RCDZ[k] * RT2P[c0 + k];
      for (int k = KS; k <= KE - 1; ++k) {
        /* band adjoints: M_b = -C_b C^T */
        const double F2b = -Cb[k] * Cp[k];
        PTb[k] += F2b * RFDZ[k] * (A[k + 1] + A[k]);
        Ab[k + 1] += F2b * PT[k] * RFDZ[k]; Ab[k] += F2b * PT[k] * RFDZ[k];
        if (k <= KE - 2) {
          const double F1b = -Cb[k] * Cp[k + 1];
          PTb[k + 1] -= F1b * RFDZ[k] * A[k + 1]; Ab[k + 1] -= F1b * PT[k + 1] * RFDZ[k];
        }
        if (k >= KS + 1) {
          const double F3b = -Cb[k] * Cp[k - 1];
          PTb[k - 1] -= F3b * RFDZ[k] * A[k]; Ab[k] -= F3b * PT[k - 1] * RFDZ[k];
        }
        /* C(k) = MOMZ + dt*(pg + Sw) */
        const size_t p = c0 + k;
        MOMZ_b[p] += Cb[k];
        Swb[p] += dt * Cb[k];
        const double pgb = dt * Cb[k];
        const double a = -pgb * RFDZ[k];
        DPRESb[p + 1] += a; RT2P_b[p + 1] += a * dt * St[p + 1]; Stb[p + 1] += a * RT2P[p + 1] * dt;
        DPRESb[p] -= a; RT2P_b[p] -= a * dt * St[p]; Stb[p] -= a * RT2P[p] * dt;
        const double bb = -GRAV * pgb;
        const double f1 = CDZ[k] / (CDZ[k] + CDZ[k + 1]), f2 = CDZ[k + 1] / (CDZ[k] + CDZ[k + 1]);
        DENS_b[p + 1] += bb * f1; Srb[p + 1] += bb * f1 * dt;
        DENS_b[p] += bb * f2; Srb[p] += bb * f2 * dt;
      }
      for (int k = KS; k <= KE; ++k) RT2P_b[c0 + k] += Ab[k] * dt * dt * RCDZ[k];
      /* PT = valueW_Z(MOMZ, POTT) (UD3.F90:98-153) */
      for (int k = KS; k <= KE - 1; ++k)
        (void)face_adj(s, zkind(k, KS, KE), POTT, POTTb, c0 + k, 1, MOMZ[c0 + k], PTb[k]);
    }

  /* ---- St (HEVI.F90:634-657) ---- */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double b = Stb[p];
        tZb[p] -= b * RCDZ[k]; tZb[p - 1] += b * RCDZ[k];
        tXb[p] -= b * RCDX[i]; tXb[p - sI] += b * RCDX[i];
        tYb[p] -= b * RCDY[j]; tYb[p - sJ] += b * RCDY[j];
      }
  /* theta fluxes fluxZ/X/Y_XYZ (UD3.F90:157-354) with mflx before the solve */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE - 1; ++k) {
        const size_t p = I3(k, i, j);
        const double fb = tZb[p], vel = ND(k, i, j, I_DENS, ZDIR);
        NDB(k, i, j, I_RHOT, ZDIR) += fb;
        mZb[p] += fb * face_adj(s, zkind(k, KS, KE), POTT, POTTb, p, 1, vel, fb * vel);
      }
  for (int j = JS; j <= JE; ++j)
    for (int i = IS - 1; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double fb = tXb[p], vel = MOMX[p] + ND(k, i, j, I_DENS, XDIR);
        NDB(k, i, j, I_RHOT, XDIR) += fb;
        mXb[p] += fb * face_adj(s, 0, POTT, POTTb, p, sI, vel, fb * vel);
      }
  for (int j = JS - 1; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double fb = tYb[p], vel = MOMY[p] + ND(k, i, j, I_DENS, YDIR);
        NDB(k, i, j, I_RHOT, YDIR) += fb;
        mYb[p] += fb * face_adj(s, 0, POTT, POTTb, p, sJ, vel, fb * vel);
      }

  /* ---- Sw (HEVI.F90:538-577) and the MOMZ fluxes (HEVI.F90:486-525) ---- */
  memset(qZb, 0, n3 * sizeof(double)); memset(qXb, 0, n3 * sizeof(double)); memset(qYb, 0, n3 * sizeof(double));
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE - 1; ++k) {
        const size_t p = I3(k, i, j);
        const double b = Swb[p];
        qZb[p] -= b * RFDZ[k]; qZb[p - 1] += b * RFDZ[k];
        qXb[p] -= b * RCDX[i]; qXb[p - sI] += b * RCDX[i];
        qYb[p] -= b * RCDY[j]; qYb[p - sJ] += b * RCDY[j];
        if (wdamp_coef) MOMZ0_b[p] -= wdamp_coef[k] * b;
      }
  /* fluxZ_XYW (UD3.F90:358-452): face m at loop index m+1 */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int mm = KS; mm <= KE - 2; ++mm) {
        const size_t p = I3(mm, i, j);
        const double vel = 0.5 * (MOMZ[p] + MOMZ[p + 1]) / DENS[p + 1];
        const double fb = qZb[p];
        NDB(mm + 1, i, j, I_MOMZ, ZDIR) += fb;
        const double velb = fb * face_adj(s, mm == KS ? 1 : 0, MOMZ, MOMZ_b, p, 1, vel, fb * vel);
        MOMZ_b[p] += 0.5 * velb / DENS[p + 1]; MOMZ_b[p + 1] += 0.5 * velb / DENS[p + 1];
        DENS_b[p + 1] -= velb * vel / DENS[p + 1];
      }
  /* fluxX_XYW / fluxY_XYW (UD3.F90:598-748) */
  for (int dir = 0; dir < 2; ++dir) {
    const size_t sd = dir == 0 ? sI : sJ;
    const double *mom = dir == 0 ? MOMX : MOMY;
    double *momb = dir == 0 ? MOMX_b : MOMY_b;
    double *qb = dir == 0 ? qXb : qYb;
    const int i0 = dir == 0 ? IS - 1 : IS, j0 = dir == 0 ? JS : JS - 1;
    for (int j = j0; j <= JE; ++j)
      for (int i = i0; i <= IE; ++i)
        for (int k = KS; k <= KE - 1; ++k) {
          const size_t p = I3(k, i, j);
          const double f1 = CDZ[k] / (CDZ[k] + CDZ[k + 1]), f2 = CDZ[k + 1] / (CDZ[k] + CDZ[k + 1]);
          const double num = f1 * mom[p + 1] + f2 * mom[p];
          const double den = f1 * 0.5 * (DENS[p + 1] + DENS[p + 1 + sd]) + f2 * 0.5 * (DENS[p] + DENS[p + sd]);
          const double vel = num / den;
          const double fb = qb[p];
          NDB(k, i, j, I_MOMZ, dir == 0 ? XDIR : YDIR) += fb;
          const double velb = fb * face_adj(s, 0, MOMZ, MOMZ_b, p, sd, vel, fb * vel);
          const double numb = velb / den, denb = -velb * vel / den;
          momb[p + 1] += f1 * numb; momb[p] += f2 * numb;
          DENS_b[p + 1] += 0.5 * f1 * denb; DENS_b[p + 1 + sd] += 0.5 * f1 * denb;
          DENS_b[p] += 0.5 * f2 * denb; DENS_b[p + sd] += 0.5 * f2 * denb;
        }
  }

  /* ---- Sr (HEVI.F90:453-475) ---- */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double b = Srb[p];
        mZb[p] -= b * RCDZ[k]; mZb[p - 1] += b * RCDZ[k];
        mXb[p] -= b * RCDX[i]; mXb[p - sI] += b * RCDX[i];
        mYb[p] -= b * RCDY[j]; mYb[p - sJ] += b * RCDY[j];
      }
  /* ---- mass fluxes before the solve (HEVI.F90:369-438) ---- */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS - 1; i <= IE; ++i)
      for (int k = KS; k <= KE - 1; ++k) NDB(k, i, j, I_DENS, ZDIR) += mZb[I3(k, i, j)];
  for (int j = JS; j <= JE; ++j)
    for (int i = IS - 1; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        MOMX_b[p] += mXb[p]; NDB(k, i, j, I_DENS, XDIR) += mXb[p];
      }
  for (int j = JS - 1; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        MOMY_b[p] += mYb[p]; NDB(k, i, j, I_DENS, YDIR) += mYb[p];
      }
  /* ---- POTT = RHOT / DENS (HEVI.F90:588-598) ---- */
  for (int j = JS - 2; j <= JE + 2; ++j)
    for (int i = IS - 2; i <= IE + 2; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        RHOT_b[p] += POTTb[p] / DENS[p];
        DENS_b[p] -= POTTb[p] * POTT[p] / DENS[p];
      }
  /* ---- DPRES (HEVI.F90:346-360); the ghost levels are read only by terrain terms ---- */
  for (int j = JS; j <= JE + 1; ++j)
    for (int i = IS; i <= IE + 1; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double b = DPRESb[p];
        DPRES0_b[p] += b;
        RT2P_b[p] += b * (RHOT[p] - REF_rhot[p]);
        RHOT_b[p] += b * RT2P[p];
        REF_rhot_b[p] -= b * RT2P[p];
      }

  free(POTTb); free(Stb); free(Srb); free(Swb); free(DPRESb);
  free(qZb); free(qXb); free(qYb); free(tZb); free(tXb); free(tYb); free(mZb); free(mXb); free(mYb);
  free(tr.POTT); free(tr.St); free(tr.PT); free(tr.F1); free(tr.F2); free(tr.F3); free(tr.Cp);
  for (int v = 0; v < 5; ++v) free(o[v]);
  free(mf); free(tf); free(DDIV); free(wd0);
  return rc;
#undef ND
#undef NDB
#undef COR
}

/* ---- periodic halo fill (COMM_vars8), reverse of orc_halo_periodic ------- */
void orc_adj_halo_periodic(const orc_grid *g, double *ab) {
  GRID_LOCALS;
  const int IMAX = IE - IS + 1, JMAX = JE - JS + 1;
  for (int j = JA - 1; j >= JE + 1; --j)
    for (int i = 0; i < IA; ++i)
      for (int k = 0; k < KA; ++k) { ab[I3(k, i, j - JMAX)] += ab[I3(k, i, j)]; ab[I3(k, i, j)] = 0.0; }
  for (int j = JS - 1; j >= 0; --j)
    for (int i = 0; i < IA; ++i)
      for (int k = 0; k < KA; ++k) { ab[I3(k, i, j + JMAX)] += ab[I3(k, i, j)]; ab[I3(k, i, j)] = 0.0; }
  for (int j = JS; j <= JE; ++j) {
    for (int i = IA - 1; i >= IE + 1; --i)
      for (int k = 0; k < KA; ++k) { ab[I3(k, i - IMAX, j)] += ab[I3(k, i, j)]; ab[I3(k, i, j)] = 0.0; }
    for (int i = IS - 1; i >= 0; --i)
      for (int k = 0; k < KA; ++k) { ab[I3(k, i + IMAX, j)] += ab[I3(k, i, j)]; ab[I3(k, i, j)] = 0.0; }
  }
}

/* ---- numerical filter (COMMON.F90:385-985), reverse ---------------------- */

/* reverse of calc_diff3 (COMMON.F90:1652-1800): d*b consumed, phib accumulated */
static void diff3_adj(const orc_grid *g, const double *cnz, const double *cnx, const double *cny,
                      double *dZb, double *dXb, double *dYb, double *phib, int KO, int IO, int JO) {
  GRID_LOCALS;
#define CNZ3(a, k, b) cnz[(a)-1 + 3 * ((size_t)(k) + (size_t)KA * ((b)-1))]
#define CNX3(a, i, b) cnx[(a)-1 + 3 * ((size_t)(i) + (size_t)IA * ((b)-1))]
#define CNY3(a, j, b) cny[(a)-1 + 3 * ((size_t)(j) + (size_t)JA * ((b)-1))]
  const size_t sI = (size_t)KA, sJ = (size_t)KA * IA;
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      double *db = dZb + I3(0, i, j), *pb = phib + I3(0, i, j);
      if (KO == 0) {
        db[KE - 2] -= db[KE + 1]; db[KE - 1] -= db[KE]; db[KE + 1] = db[KE] = db[KE + 2] = 0.0;
        {
          const double b = db[KE - 1];
          pb[KE - 1] += CNZ3(1, KE, 1) * b; pb[KE] -= CNZ3(2, KE, 1) * b;
          pb[KE - 1] += CNZ3(3, KE, 1) * b; pb[KE - 2] -= CNZ3(1, KE - 1, 1) * b;
        }
        db[KS + 1] -= db[KS - 2]; db[KS] -= db[KS - 1]; db[KS - 2] = db[KS - 1] = 0.0;
        {
          const double b = db[KS];
          pb[KS + 2] += CNZ3(1, KS + 1, 1) * b; pb[KS + 1] -= CNZ3(2, KS + 1, 1) * b;
          pb[KS] += CNZ3(3, KS + 1, 1) * b; pb[KS + 1] -= CNZ3(1, KS, 1) * b;
        }
        for (int k = KS + 1; k <= KE - 2; ++k) {
          const double b = db[k];
          pb[k + 2] += CNZ3(1, k + 1, 1) * b; pb[k + 1] -= CNZ3(2, k + 1, 1) * b;
          pb[k] += CNZ3(3, k + 1, 1) * b; pb[k - 1] -= CNZ3(1, k, 1) * b;
        }
      } else {
        db[KE - 1] -= db[KE + 2]; db[KE] -= db[KE + 1]; db[KE + 1] = db[KE + 2] = 0.0;
        {
          const double b = db[KE];
          pb[KE - 1] += CNZ3(1, KE, 2) * b; pb[KE - 1] += CNZ3(3, KE, 2) * b; pb[KE - 2] -= CNZ3(1, KE - 1, 2) * b;
        }
        {
          const double b = db[KE - 1];
          pb[KE - 1] -= CNZ3(2, KE - 1, 2) * b; pb[KE - 2] += CNZ3(3, KE - 1, 2) * b; pb[KE - 3] -= CNZ3(1, KE - 2, 2) * b;
        }
        db[KS + 3] -= db[KS - 2]; db[KS + 2] -= db[KS - 1]; db[KS + 1] -= db[KS];
        db[KS - 2] = db[KS - 1] = db[KS] = 0.0;
        {
          const double b = db[KS + 1];
          pb[KS + 2] += CNZ3(1, KS + 1, 2) * b; pb[KS + 1] -= CNZ3(2, KS + 1, 2) * b;
          pb[KS] += CNZ3(3, KS + 1, 2) * b; pb[KS + 1] -= CNZ3(1, KS, 2) * b;
        }
        for (int k = KS + 2; k <= KE - 2; ++k) {
          const double b = db[k];
          pb[k + 1] += CNZ3(1, k, 2) * b; pb[k] -= CNZ3(2, k, 2) * b;
          pb[k - 1] += CNZ3(3, k, 2) * b; pb[k - 2] -= CNZ3(1, k - 1, 2) * b;
        }
      }
      for (int k = KS; k <= KE - KO; ++k) {
        const size_t p = I3(k, i, j);
        const double bx = dXb[p], by = dYb[p];
        if (IO == 0) {
          phib[p + 2 * sI] += CNX3(1, i + 1, 1) * bx; phib[p + sI] -= CNX3(2, i + 1, 1) * bx;
          phib[p] += CNX3(3, i + 1, 1) * bx; phib[p - sI] -= CNX3(1, i, 1) * bx;
        } else {
          phib[p + sI] += CNX3(1, i, 2) * bx; phib[p] -= CNX3(2, i, 2) * bx;
          phib[p - sI] += CNX3(3, i, 2) * bx; phib[p - 2 * sI] -= CNX3(1, i - 1, 2) * bx;
        }
        if (JO == 0) {
          phib[p + 2 * sJ] += CNY3(1, j + 1, 1) * by; phib[p + sJ] -= CNY3(2, j + 1, 1) * by;
          phib[p] += CNY3(3, j + 1, 1) * by; phib[p - sJ] -= CNY3(1, j, 1) * by;
        } else {
          phib[p + sJ] += CNY3(1, j, 2) * by; phib[p] -= CNY3(2, j, 2) * by;
          phib[p - sJ] += CNY3(3, j, 2) * by; phib[p - 2 * sJ] -= CNY3(1, j - 1, 2) * by;
        }
      }
    }
#undef CNZ3
#undef CNX3
#undef CNY3
}

static inline double pow4(double x) { double x2 = x * x; return x2 * x2; }

int orc_adj_numfilter_coef(const orc_grid *g, const orc_metric *m, double DT,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *REF_dens, const double *REF_pott, double ND_COEF, double ND_SFC_FACT,
    double *num_diff_b, double *DENS_b, double *MOMZ_b, double *MOMX_b, double *MOMY_b, double *RHOT_b) {
  GRID_LOCALS;
  const size_t n3 = N3, sI = (size_t)KA, sJ = (size_t)KA * IA;
  const double *CDZ = m->CDZ, *FDZ = m->FDZ, *CDX = m->CDX, *FDX = m->FDX, *CDY = m->CDY, *FDY = m->FDY;
  double *cnz = zalloc((size_t)6 * KA), *cnx = zalloc((size_t)6 * IA), *cny = zalloc((size_t)6 * JA);
  orc_filter_tables(g, m, cnz, cnx, cny);
  const double DIFF4 = ND_COEF / ((double)(1 << 4) * DT);
  double *nd_cdz = zalloc(KA), *nd_fdz = zalloc(KA), *nd_cdx = zalloc(IA), *nd_fdx = zalloc(IA), *nd_cdy = zalloc(JA), *nd_fdy = zalloc(JA);
  for (int k = KS - 1; k <= KE; ++k) nd_cdz[k] = DIFF4 * pow4(CDZ[k]);
  for (int k = KS + 1; k <= KE - 1; ++k) nd_fdz[k] = DIFF4 * pow4(FDZ[k]);
  for (int i = IS; i <= IE; ++i) { nd_cdx[i] = DIFF4 * pow4(CDX[i]); nd_fdx[i] = DIFF4 * pow4(FDX[i]); }
  for (int j = JS; j <= JE; ++j) { nd_cdy[j] = DIFF4 * pow4(CDY[j]); nd_fdy[j] = DIFF4 * pow4(FDY[j]); }
#define NDB(k, i, j, v, d) num_diff_b[I3(k, i, j) + n3 * ((size_t)(v) + 5 * (size_t)(d))]
#define D_(k, i, j) DENS[I3(k, i, j)]
#define DB_(k, i, j) DENS_b[I3(k, i, j)]
  for (int q = 0; q < 15; ++q) orc_adj_halo_periodic(g, num_diff_b + n3 * (size_t)q);
  /* surface factor rows (COMMON.F90:593-594 and the like), X and Y directions */
  for (int v = 0; v < 5; ++v)
    for (int d = XDIR; d <= YDIR; ++d)
      for (int j = JS; j <= JE; ++j)
        for (int i = IS; i <= IE; ++i) {
          NDB(KS, i, j, v, d) *= ND_SFC_FACT;
          NDB(KS + 1, i, j, v, d) *= (1.0 + ND_SFC_FACT) * 0.5;
        }
  double *fld = zalloc(n3), *w = zalloc(3 * n3), *wb = zalloc(3 * n3), *fb = zalloc(n3);
  double *wZ = w, *wX = w + n3, *wY = w + 2 * n3, *wZb = wb, *wXb = wb + n3, *wYb = wb + 2 * n3;

  /* density (COMMON.F90:541-621) */
  memset(wb, 0, 3 * n3 * sizeof(double)); memset(fb, 0, n3 * sizeof(double));
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS - 1; k <= KE; ++k) wZb[I3(k, i, j)] += NDB(k, i, j, I_DENS, ZDIR) * nd_cdz[k];
      for (int k = KS; k <= KE; ++k) {
        wXb[I3(k, i, j)] += NDB(k, i, j, I_DENS, XDIR) * nd_cdx[i];
        wYb[I3(k, i, j)] += NDB(k, i, j, I_DENS, YDIR) * nd_cdy[j];
      }
    }
  diff3_adj(g, cnz, cnx, cny, wZb, wXb, wYb, fb, 0, 0, 0);
  for (int j = JS - 1; j <= JE + 2; ++j)
    for (int i = IS - 1; i <= IE + 2; ++i)
      for (int k = KS; k <= KE; ++k) DENS_b[I3(k, i, j)] += fb[I3(k, i, j)];

  /* z-momentum (COMMON.F90:629-707) */
  for (int j = JS - 2; j <= JE + 2; ++j)
    for (int i = IS - 2; i <= IE + 2; ++i)
      for (int k = KS; k <= KE - 1; ++k) fld[I3(k, i, j)] = 2.0 * MOMZ[I3(k, i, j)] / (DENS[I3(k + 1, i, j)] + DENS[I3(k, i, j)]);
  orc_calc_diff3(g, m, w, fld, 1, 0, 0);
  memset(wb, 0, 3 * n3 * sizeof(double)); memset(fb, 0, n3 * sizeof(double));
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS + 1; k <= KE - 1; ++k) {
        const size_t p = I3(k, i, j);
        const double b = NDB(k, i, j, I_MOMZ, ZDIR);
        wZb[p] += b * nd_fdz[k] * DENS[p];
        DENS_b[p] += b * wZ[p] * nd_fdz[k];
      }
      for (int k = KS; k <= KE - 1; ++k) {
        const size_t p = I3(k, i, j);
        double b = NDB(k, i, j, I_MOMZ, XDIR);
        wXb[p] += b * nd_cdx[i] * 0.25 * (D_(k + 1, i + 1, j) + D_(k + 1, i, j) + D_(k, i + 1, j) + D_(k, i, j));
        double c = b * wX[p] * nd_cdx[i] * 0.25;
        DB_(k + 1, i + 1, j) += c; DB_(k + 1, i, j) += c; DB_(k, i + 1, j) += c; DB_(k, i, j) += c;
        b = NDB(k, i, j, I_MOMZ, YDIR);
        wYb[p] += b * nd_cdy[j] * 0.25 * (D_(k + 1, i, j + 1) + D_(k + 1, i, j) + D_(k, i, j + 1) + D_(k, i, j));
        c = b * wY[p] * nd_cdy[j] * 0.25;
        DB_(k + 1, i, j + 1) += c; DB_(k + 1, i, j) += c; DB_(k, i, j + 1) += c; DB_(k, i, j) += c;
      }
    }
  diff3_adj(g, cnz, cnx, cny, wZb, wXb, wYb, fb, 1, 0, 0);
  for (int j = JS - 2; j <= JE + 2; ++j)
    for (int i = IS - 2; i <= IE + 2; ++i)
      for (int k = KS; k <= KE - 1; ++k) {
        const size_t p = I3(k, i, j);
        const double ds = DENS[p + 1] + DENS[p];
        MOMZ_b[p] += fb[p] * 2.0 / ds;
        const double dsb = -fb[p] * fld[p] / ds;
        DENS_b[p + 1] += dsb; DENS_b[p] += dsb;
      }

  /* x-momentum (COMMON.F90:713-790) */
  for (int j = JS - 2; j <= JE + 2; ++j)
    for (int i = IS - 2; i <= IE + 1; ++i)
      for (int k = KS; k <= KE; ++k) fld[I3(k, i, j)] = 2.0 * MOMX[I3(k, i, j)] / (DENS[I3(k, i + 1, j)] + DENS[I3(k, i, j)]);
  orc_calc_diff3(g, m, w, fld, 0, 1, 0);
  memset(wb, 0, 3 * n3 * sizeof(double)); memset(fb, 0, n3 * sizeof(double));
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS; k <= KE - 1; ++k) {
        const size_t p = I3(k, i, j);
        const double b = NDB(k, i, j, I_MOMX, ZDIR);
        wZb[p] += b * nd_cdz[k] * 0.25 * (D_(k + 1, i + 1, j) + D_(k + 1, i, j) + D_(k, i + 1, j) + D_(k, i, j));
        const double c = b * wZ[p] * nd_cdz[k] * 0.25;
        DB_(k + 1, i + 1, j) += c; DB_(k + 1, i, j) += c; DB_(k, i + 1, j) += c; DB_(k, i, j) += c;
      }
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        double b = NDB(k, i, j, I_MOMX, XDIR);
        wXb[p] += b * nd_fdx[i] * DENS[p];
        DENS_b[p] += b * wX[p] * nd_fdx[i];
        b = NDB(k, i, j, I_MOMX, YDIR);
        wYb[p] += b * nd_cdy[j] * 0.25 * (D_(k, i + 1, j + 1) + D_(k, i + 1, j) + D_(k, i, j + 1) + D_(k, i, j));
        const double c = b * wY[p] * nd_cdy[j] * 0.25;
        DB_(k, i + 1, j + 1) += c; DB_(k, i + 1, j) += c; DB_(k, i, j + 1) += c; DB_(k, i, j) += c;
      }
    }
  diff3_adj(g, cnz, cnx, cny, wZb, wXb, wYb, fb, 0, 1, 0);
  for (int j = JS - 2; j <= JE + 2; ++j)
    for (int i = IS - 2; i <= IE + 1; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double ds = DENS[p + sI] + DENS[p];
        MOMX_b[p] += fb[p] * 2.0 / ds;
        const double dsb = -fb[p] * fld[p] / ds;
        DENS_b[p + sI] += dsb; DENS_b[p] += dsb;
      }

  /* y-momentum (COMMON.F90:796-873) */
  for (int j = JS - 2; j <= JE + 1; ++j)
    for (int i = IS - 2; i <= IE + 2; ++i)
      for (int k = KS; k <= KE; ++k) fld[I3(k, i, j)] = 2.0 * MOMY[I3(k, i, j)] / (DENS[I3(k, i, j + 1)] + DENS[I3(k, i, j)]);
  orc_calc_diff3(g, m, w, fld, 0, 0, 1);
  memset(wb, 0, 3 * n3 * sizeof(double)); memset(fb, 0, n3 * sizeof(double));
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS; k <= KE - 1; ++k) {
        const size_t p = I3(k, i, j);
        const double b = NDB(k, i, j, I_MOMY, ZDIR);
        wZb[p] += b * nd_cdz[k] * 0.25 * (D_(k + 1, i, j + 1) + D_(k + 1, i, j) + D_(k, i, j + 1) + D_(k, i, j));
        const double c = b * wZ[p] * nd_cdz[k] * 0.25;
        DB_(k + 1, i, j + 1) += c; DB_(k + 1, i, j) += c; DB_(k, i, j + 1) += c; DB_(k, i, j) += c;
      }
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        double b = NDB(k, i, j, I_MOMY, XDIR);
        wXb[p] += b * nd_cdx[i] * 0.25 * (D_(k, i + 1, j + 1) + D_(k, i, j + 1) + D_(k, i + 1, j) + D_(k, i, j));
        const double c = b * wX[p] * nd_cdx[i] * 0.25;
        DB_(k, i + 1, j + 1) += c; DB_(k, i, j + 1) += c; DB_(k, i + 1, j) += c; DB_(k, i, j) += c;
        b = NDB(k, i, j, I_MOMY, YDIR);
        wYb[p] += b * nd_fdy[j] * DENS[p];
        DENS_b[p] += b * wY[p] * nd_fdy[j];
      }
    }
  diff3_adj(g, cnz, cnx, cny, wZb, wXb, wYb, fb, 0, 0, 1);
  for (int j = JS - 2; j <= JE + 1; ++j)
    for (int i = IS - 2; i <= IE + 2; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double ds = DENS[p + sJ] + DENS[p];
        MOMY_b[p] += fb[p] * 2.0 / ds;
        const double dsb = -fb[p] * fld[p] / ds;
        DENS_b[p + sJ] += dsb; DENS_b[p] += dsb;
      }

  /* rho*theta (COMMON.F90:880-962) */
  for (int j = JS - 1; j <= JE + 2; ++j)
    for (int i = IS - 1; i <= IE + 2; ++i)
      for (int k = KS; k <= KE; ++k) fld[I3(k, i, j)] = RHOT[I3(k, i, j)] / DENS[I3(k, i, j)] - REF_pott[I3(k, i, j)];
  orc_calc_diff3(g, m, w, fld, 0, 0, 0);
  memset(wb, 0, 3 * n3 * sizeof(double)); memset(fb, 0, n3 * sizeof(double));
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS; k <= KE - 1; ++k) {
        const size_t p = I3(k, i, j);
        const double b = NDB(k, i, j, I_RHOT, ZDIR);
        wZb[p] += b * nd_cdz[k] * 0.5 * (DENS[p + 1] + DENS[p]);
        const double c = b * wZ[p] * nd_cdz[k] * 0.5;
        DENS_b[p + 1] += c; DENS_b[p] += c;
      }
      {
        const size_t p = I3(KS - 1, i, j);
        const double b = NDB(KS - 1, i, j, I_RHOT, ZDIR);
        wZb[p] += b * nd_cdz[KS - 1] * DENS[p + 1];
        DENS_b[p + 1] += b * wZ[p] * nd_cdz[KS - 1];
      }
      {
        const size_t p = I3(KE, i, j);
        const double b = NDB(KE, i, j, I_RHOT, ZDIR);
        wZb[p] += b * nd_cdz[KE] * DENS[p];
        DENS_b[p] += b * wZ[p] * nd_cdz[KE];
      }
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        double b = NDB(k, i, j, I_RHOT, XDIR);
        wXb[p] += b * nd_cdx[i] * 0.5 * (DENS[p + sI] + DENS[p]);
        double c = b * wX[p] * nd_cdx[i] * 0.5;
        DENS_b[p + sI] += c; DENS_b[p] += c;
        b = NDB(k, i, j, I_RHOT, YDIR);
        wYb[p] += b * nd_cdy[j] * 0.5 * (DENS[p + sJ] + DENS[p]);
        c = b * wY[p] * nd_cdy[j] * 0.5;
        DENS_b[p + sJ] += c; DENS_b[p] += c;
      }
    }
  diff3_adj(g, cnz, cnx, cny, wZb, wXb, wYb, fb, 0, 0, 0);
  for (int j = JS - 1; j <= JE + 2; ++j)
    for (int i = IS - 1; i <= IE + 2; ++i)
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        RHOT_b[p] += fb[p] / DENS[p];
        DENS_b[p] -= fb[p] * RHOT[p] / (DENS[p] * DENS[p]);
      }
  memset(num_diff_b, 0, 15 * n3 * sizeof(double));
#undef NDB
#undef D_
#undef DB_
  free(fld); free(w); free(wb); free(fb); free(cnz); free(cnx); free(cny);
  free(nd_cdz); free(nd_fdz); free(nd_cdx); free(nd_fdx); free(nd_cdy); free(nd_fdy);
  return 0;
}

/* ---- large step (LARGE.F90:377-1127) with RK4 (RK4.F90:300-660), reverse - */
int orc_adj_large_step(const orc_grid *g, const orc_metric *m, const orc_config *c, const orc_state *st,
    double *const xb[5]) {
  GRID_LOCALS;
  if (c->tinteg_short != 4 || c->short_type != 0 || c->divdmp_coef != 0.0 || c->nd_order != 1 || !c->nd_use_rs) return -1;
  if ((int)ceil((c->dtl - 2.220446049250313e-16) / c->dts) != 1) return -1;
  if (!is_flat(g, m)) return -1;
  const size_t n3 = N3;
  const double dt = c->dtl;
  const double *x[5] = {st->DENS, st->MOMZ, st->MOMX, st->MOMY, st->RHOT};
  /* forward recomputation: RT2P, DPRES0, REF_rhot (LARGE.F90:443-471) */
  const double CVdry = CPdry - Rdry, CPovCV = CPdry / CVdry;
  double *RT2P = zalloc(n3), *DPRES0 = zalloc(n3), *REF_rhot = zalloc(n3);
  for (int j = 0; j < JA; ++j)
    for (int i = 0; i < IA; ++i) {
      for (int k = KS; k <= KE; ++k) {
        const size_t p = I3(k, i, j);
        const double PRES = P00 * pow(Rdry * st->RHOT[p] / P00, CPovCV);
        RT2P[p] = CPovCV * PRES / st->RHOT[p];
        DPRES0[p] = PRES - st->REF_pres[p];
        REF_rhot[p] = st->RHOT[p];
      }
      DPRES0[I3(KS - 1, i, j)] = DPRES0[I3(KS + 1, i, j)] + (st->REF_pres[I3(KS + 1, i, j)] - st->REF_pres[I3(KS - 1, i, j)]);
      DPRES0[I3(KE + 1, i, j)] = DPRES0[I3(KE - 1, i, j)] + (st->REF_pres[I3(KE - 1, i, j)] - st->REF_pres[I3(KE + 1, i, j)]);
    }
  double *Dt[5];
  const double *tp[5] = {st->DENS_tp, st->MOMZ_tp, st->MOMX_tp, st->MOMY_tp, st->RHOT_tp};
  for (int v = 0; v < 5; ++v) {
    Dt[v] = zalloc(n3);
    for (int j = JS; j <= JE; ++j)
      for (int i = IS; i <= IE; ++i)
        for (int k = KS; k <= (v == 1 ? KE - 1 : KE); ++k) Dt[v][I3(k, i, j)] = tp[v][I3(k, i, j)] + 0.0;
    orc_halo_periodic(g, Dt[v]);
  }
  double *ND = zalloc(15 * n3), *NDb = zalloc(15 * n3);
  if (c->nd_coef != 0.0)
    orc_numfilter_coef(g, ND, x[0], x[1], x[2], x[3], x[4], m, dt, st->REF_dens, st->REF_pott, c->nd_coef, 1, c->nd_sfc_fact, 1);
  /* stage outputs R1..R3 (RK4.F90:378-560) */
  double *R[3][5], *Rb[4][5], *S0b[5];
  double *mf = zalloc(3 * n3), *tf = zalloc(3 * n3), *DDIV = zalloc(n3);
  const double dtrk[4] = {dt / 2.0, dt / 2.0, dt, dt};
  for (int v = 0; v < 5; ++v) {
    for (int r = 0; r < 3; ++r) R[r][v] = zalloc(n3);
    for (int r = 0; r < 4; ++r) Rb[r][v] = zalloc(n3);
    S0b[v] = zalloc(n3);
  }
  int rc = 0;
  for (int r = 0; r < 3; ++r) {
    const double *const *in = r == 0 ? (const double *const *)x : (const double *const *)R[r - 1];
    rc |= orc_tstep_short_hevi(g, c->scheme, R[r][0], R[r][1], R[r][2], R[r][3], R[r][4], mf, tf,
                               x[0], x[1], x[2], x[3], x[4], in[0], in[1], in[2], in[3], in[4],
                               Dt[0], Dt[1], Dt[2], Dt[3], Dt[4], DPRES0, RT2P, st->CORIOLI, ND, st->wdamp_coef,
                               0.0, DDIV, m, st->REF_dens, REF_rhot, dtrk[r]);
    for (int v = 0; v < 5; ++v) orc_halo_periodic(g, R[r][v]);
  }
  /* reverse: halo fill and ghost levels of the new state (LARGE.F90:908-927) */
  for (int v = 0; v < 5; ++v) orc_adj_halo_periodic(g, xb[v]);
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int v = 0; v < 5; ++v) {
        double *b = xb[v] + I3(0, i, j);
        for (int k = KE + 1; k < KA; ++k) { if (v != 1) b[KE] += b[k]; b[k] = 0.0; }
        for (int k = 0; k <= KS - 1; ++k) { if (v != 1) b[KS] += b[k]; b[k] = 0.0; }
      }
  /* combination (RK4.F90:562-660) */
  for (int v = 0; v < 5; ++v) {
    const int kend = v == 1 ? KE - 1 : KE;
    for (int j = JS; j <= JE; ++j)
      for (int i = IS; i <= IE; ++i)
        for (int k = KS; k <= kend; ++k) {
          const size_t p = I3(k, i, j);
          const double b = xb[v][p];
          Rb[0][v][p] += b * (2.0 / 6.0);
          Rb[1][v][p] += b * (4.0 / 6.0);
          Rb[2][v][p] += b * (2.0 / 6.0);
          Rb[3][v][p] += b * (1.0 / 6.0);
          S0b[v][p] -= b * (3.0 / 6.0);
        }
  }
  /* stages 4..1 */
  for (int r = 3; r >= 0; --r) {
    if (r < 3)
      for (int v = 0; v < 5; ++v) orc_adj_halo_periodic(g, Rb[r][v]);
    const double *const *in = r == 0 ? (const double *const *)x : (const double *const *)R[r - 1];
    double **inb = r == 0 ? S0b : Rb[r - 1];
    double *DPRES0_b = zalloc(n3), *RT2P_b = zalloc(n3), *REFrt_b = zalloc(n3);
    rc |= orc_adj_tstep_short_hevi(g, c->scheme, m, x[0], x[1], x[2], x[3], x[4], in[0], in[1], in[2], in[3], in[4],
                                   Dt[0], Dt[1], Dt[2], Dt[3], Dt[4], DPRES0, RT2P, st->CORIOLI, ND, st->wdamp_coef,
                                   st->REF_dens, REF_rhot, dtrk[r],
                                   Rb[r][0], Rb[r][1], Rb[r][2], Rb[r][3], Rb[r][4],
                                   S0b[0], S0b[1], S0b[2], S0b[3], S0b[4], inb[0], inb[1], inb[2], inb[3], inb[4],
                                   NDb, DPRES0_b, RT2P_b, REFrt_b);
    /* RT2P, DPRES0, REF_rhot as functions of RHOT at the step start */
    for (int j = 0; j < JA; ++j)
      for (int i = 0; i < IA; ++i) {
        DPRES0_b[I3(KE - 1, i, j)] += DPRES0_b[I3(KE + 1, i, j)];
        DPRES0_b[I3(KS + 1, i, j)] += DPRES0_b[I3(KS - 1, i, j)];
        for (int k = KS; k <= KE; ++k) {
          const size_t p = I3(k, i, j);
          S0b[4][p] += RT2P_b[p] * RT2P[p] * (CPovCV - 1.0) / st->RHOT[p] + DPRES0_b[p] * RT2P[p] + REFrt_b[p];
        }
      }
    free(DPRES0_b); free(RT2P_b); free(REFrt_b);
  }
  if (c->nd_coef != 0.0)
    orc_adj_numfilter_coef(g, m, dt, x[0], x[1], x[2], x[3], x[4], st->REF_dens, st->REF_pott, c->nd_coef, c->nd_sfc_fact,
                           NDb, S0b[0], S0b[1], S0b[2], S0b[3], S0b[4]);
  for (int v = 0; v < 5; ++v) memcpy(xb[v], S0b[v], n3 * sizeof(double));
  for (int v = 0; v < 5; ++v) {
    for (int r = 0; r < 3; ++r) free(R[r][v]);
    for (int r = 0; r < 4; ++r) free(Rb[r][v]);
    free(S0b[v]); free(Dt[v]);
  }
  free(RT2P); free(DPRES0); free(REF_rhot); free(ND); free(NDb); free(mf); free(tf); free(DDIV);
  return rc;
}
