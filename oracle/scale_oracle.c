/*
 * scale_oracle.c -- CPU restatement of the SCALE-RM (5.3.6) HEVI short step,
 * the FVM UD3/CD4/UD1 flux families, the numerical filter and FCT tracer
 * advection, in the reference's own Fortran array layout.
 *
 * TEST INFRASTRUCTURE ONLY (see scale_oracle.h).  Never linked by the product.
 *
 * Reference files (paths relative to the reference root):
 *   HEVI   = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short_fvm_hevi.F90
 *   UD3    = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_fvm_flux_ud3.F90
 *   CD4    = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_fvm_flux_cd4.F90
 *   UD1    = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_fvm_flux_ud1.F90
 *   ERB    = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_fvm_flux_udcd.F90.erb
 *   COMMON = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90
 *   TRACER = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_tracer_fvm_heve.F90
 *   RK4    = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tinteg_short_rk4.F90
 *   TRK3   = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tinteg_tracer_rk3.F90
 *   RK3    = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tinteg_short_rk3.F90
 *   TEUL   = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tinteg_tracer_euler.F90
 *   LARGE  = scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_large_fvm_heve.F90
 *   COMM   = scalelib/src/communication/scale_comm_cartesC.F90
 *
 * The UD3/CD4 files are generated from ERB; the three generated files share
 * one structure (nh = 1 halo-order for UD3 and CD4, nh = 0 for UD1) and differ
 * only in the face-value formula, so the face values are factored into
 * fv_int/fv_lo/fv_hi below and every flux routine is written once.
 */
#include "scale_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- indices ------------------------------------------------------------ */
enum { I_XYZ = 0, I_XYW, I_UYW, I_XVW, I_UYZ, I_XVZ, I_UVZ };   /* index.F90:90-96 */
enum { I_XY = 0, I_UY, I_XV, I_UV };                             /* index.F90:98-101 */
enum { ZDIR = 0, XDIR, YDIR };                                   /* index.F90:32-34 */
enum { I_DENS = 0, I_MOMZ, I_MOMX, I_MOMY, I_RHOT };             /* scale_index.F90:28-32 */

/* scale_const.F90:46,55-56,88,162,169 */
static const double GRAV = 9.80665;
static const double Rdry = 287.04;
static const double CPdry = 1004.64;
static const double P00 = 100000.0;
#define EPS_RP 2.220446049250313080847263336181640625e-16 /* epsilon(1.0_DP) */

/* UD3.F90:80-89, CD4.F90:84-88, UD1.F90:84-86 */
static const double F1 = 0.5;
static const double F2 = 0.5;
static const double F31 = -1.0 / 12.0;
static const double F32 = 7.0 / 12.0;
static const double F33 = 3.0 / 12.0;
static const double F41 = 7.0 / 12.0;
static const double F42 = -1.0 / 12.0;
/* ud5.F90:92-96, cd6.F90:91-93 */
static const double F51 = 1.0 / 60.0;
static const double F52 = -8.0 / 60.0;
static const double F53 = 37.0 / 60.0;
static const double F54 = -5.0 / 60.0;
static const double F55 = 10.0 / 60.0;
static const double F61 = 37.0 / 60.0;
static const double F62 = -8.0 / 60.0;
static const double F63 = 1.0 / 60.0;

#define GRID_LOCALS                                                        \
  const int KA = g->KA, IA = g->IA, JA = g->JA;                            \
  const int KS = g->KS, KE = g->KE, IS = g->IS, IE = g->IE, JS = g->JS, JE = g->JE; \
  const size_t N3 = (size_t)KA * IA * JA;                                  \
  (void)KA; (void)IA; (void)JA; (void)KS; (void)KE; (void)IS; (void)IE; (void)JS; (void)JE; (void)N3

#define I3(k, i, j) ((size_t)(k) + (size_t)KA * ((size_t)(i) + (size_t)IA * (size_t)(j)))
#define MAP(mp, i, j, n) (mp)[(size_t)(i) + (size_t)IA * ((size_t)(j) + (size_t)JA * (size_t)(n))]

static int g_threads = 1;
/* optional record of the HEVI stage's intermediates, used by the adjoint checker */
static orc_hevi_trace *g_trace = NULL;
void orc_set_hevi_trace(orc_hevi_trace *t) { g_trace = t; }
#define OMP_FOR _Pragma("omp parallel for num_threads(g_threads) schedule(static)")
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

static double *alloc_nan(size_t n) {
  double *p = (double *)malloc(n * sizeof(double));
  for (size_t q = 0; q < n; ++q) p[q] = NAN; /* UNDEF poisoning, HEVI.F90:298-322 */
  return p;
}

/* Fortran sign(a,b) = |a| with the sign bit of b (gfortran honours -0.0). */
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }
/* the FCT ratios R+/R-: plain values here; the tangent-linear build (dual.h)
 * holds them constant, the frozen limiter of the tracer adjoint */
#ifndef ORC_FROZEN
#define ORC_FROZEN(x) (x)
#endif
static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fmin2(double a, double b) { return a < b ? a : b; }

/* ---- face values (ERB:59-203) ------------------------------------------- */
/* Koren (1993) limiter, ud3Koren1993.F90:1690-1710 */
static inline double koren_phi(double v1, double v2, double v3) {
  const double zerosw1 = EPS_RP - fsign(EPS_RP, fabs(v1 - v2) - EPS_RP);
  const double zerosw2 = EPS_RP - fsign(EPS_RP, fabs(v2 - v3) - EPS_RP);
  const double r2 = 2.0 * (v1 - v2 + zerosw1 * zerosw2) / (v2 - v3 + zerosw2);
  return fmax2(0.0, fmin2(r2, fmin2((1.0 + r2) / 3.0, 2.0)));
}
/* interior face between q0 = val(idx) and q1 = val(idx+1); qm1 = val(idx-1), q2 = val(idx+2) */
static inline double fv_int(int s, double qm1, double q0, double q1, double q2, double vel) {
  if (s == ORC_UD3KOREN) /* ud3Koren1993.F90:201-209 (z), 299-307 (x), 356-364 (y) */
    return (q0 + 0.5 * koren_phi(q1, q0, qm1) * (q0 - qm1)) * (0.5 + fsign(0.5, vel))
         + (q1 + 0.5 * koren_phi(q0, q1, q2) * (q1 - q2)) * (0.5 - fsign(0.5, vel));
  if (s == ORC_UD3) /* UD3.F90:194-197 */
    return (F31 * (q2 + qm1) + F32 * (q1 + q0)) - (F31 * (q2 - qm1) + F33 * (q1 - q0)) * fsign(1.0, vel);
  if (s == ORC_CD4) /* CD4.F90:190-192 */
    return F41 * (q1 + q0) + F42 * (q2 + qm1);
  if (s == ORC_CD2) /* cd2.F90:168-171 */
    return F2 * (q1 + q0);
  /* UD1.F90:153-154 */
  return F1 * (q1 + q0) - fsign(F1, vel) * (q1 - q0);
}
/* lowest face: q0 = val(KS), q1 = val(KS+1), q2 = val(KS+2) (UD3.F90:225-229, CD4.F90:218-219) */
static inline double fv_lo(int s, double q0, double q1, double q2, double vel) {
  if (s == ORC_UD3KOREN) /* ud3Koren1993.F90:234-241 */
    return q0 * (0.5 + fsign(0.5, vel)) + (q1 + 0.5 * koren_phi(q0, q1, q2) * (q1 - q2)) * (0.5 - fsign(0.5, vel));
  if (s == ORC_UD3)
    return F2 * (q1 + q0) * (0.5 + fsign(0.5, vel)) + (2.0 * q0 + 5.0 * q1 - q2) / 6.0 * (0.5 - fsign(0.5, vel));
  if (s == ORC_UD1) return fv_int(s, 0.0, q0, q1, 0.0, vel); /* no boundary form (UD1.F90:135-171) */
  return F2 * (q1 + q0);
}
/* highest face: qm1 = val(KE-2), q0 = val(KE-1), q1 = val(KE) (UD3.F90:231-235, CD4.F90:221-222) */
static inline double fv_hi(int s, double qm1, double q0, double q1, double vel) {
  if (s == ORC_UD3KOREN) /* ud3Koren1993.F90:242-249, limiter arguments as in the reference */
    return (q0 + 0.5 * koren_phi(qm1, q0, q1) * (q0 - q1)) * (0.5 + fsign(0.5, vel)) + q1 * (0.5 - fsign(0.5, vel));
  if (s == ORC_UD3)
    return (2.0 * q1 + 5.0 * q0 - qm1) / 6.0 * (0.5 + fsign(0.5, vel)) + F2 * (q1 + q0) * (0.5 - fsign(0.5, vel));
  if (s == ORC_UD1) return fv_int(s, 0.0, q0, q1, 0.0, vel);
  return F2 * (q1 + q0);
}

/* exported for the closed-form tests: kind 0 interior, 1 lowest face, 2 highest face */
double orc_face_value(int s, int kind, double qm1, double q0, double q1, double q2, double vel) {
  if (kind == 1) return fv_lo(s, q0, q1, q2, vel);
  if (kind == 2) return fv_hi(s, qm1, q0, q1, vel);
  return fv_int(s, qm1, q0, q1, q2, vel);
}

/* ---- six-point faces of UD5 and CD6 ---------------------------------------- */
/* interior face between q0 = val(idx) and q1 = val(idx+1) (ud5.F90:236-242, cd6.F90:195-199) */
static inline double fv6_int(int s, double qm2, double qm1, double q0, double q1, double q2, double q3, double vel) {
  if (s == ORC_UD5)
    return (F51 * (q3 + qm2) + F52 * (q2 + qm1) + F53 * (q1 + q0))
         - (F51 * (q3 - qm2) + F54 * (q2 - qm1) + F55 * (q1 - q0)) * fsign(1.0, vel);
  return F61 * (q1 + q0) + F62 * (q2 + qm1) + F63 * (q3 + qm2);
}
/* face KS from v = val(KS..KS+2) (ud5.F90:270-275, cd6.F90:217-220) */
static inline double fv6_ks(int s, const double *v, double vel) {
  if (s == ORC_UD5)
    return F2 * (v[1] + v[0]) * (0.5 + fsign(0.5, vel)) + (2.0 * v[0] + 5.0 * v[1] - v[2]) / 6.0 * (0.5 - fsign(0.5, vel));
  return F2 * (v[1] + v[0]);
}
/* face KE-1 from v = val(KE-2..KE) (ud5.F90:276-281, cd6.F90:221-224) */
static inline double fv6_ke1(int s, const double *v, double vel) {
  if (s == ORC_UD5)
    return (2.0 * v[2] + 5.0 * v[1] - v[0]) / 6.0 * (0.5 + fsign(0.5, vel)) + F2 * (v[2] + v[1]) * (0.5 - fsign(0.5, vel));
  return F2 * (v[2] + v[1]);
}
/* face KS+1 from v = val(KS..KS+4) (ud5.F90:283-292, cd6.F90:225-229) */
static inline double fv6_ks1(int s, const double *v, double vel) {
  if (s == ORC_UD5)
    return (2.0 * v[2] + 5.0 * v[1] - v[0]) / 6.0 * (0.5 + fsign(0.5, vel))
         + (-3.0 * v[0] + 27.0 * v[1] + 47.0 * v[2] - 13.0 * v[3] + 2.0 * v[4]) / 60.0 * (0.5 - fsign(0.5, vel));
  return F41 * (v[2] + v[1]) + F42 * (v[3] + v[0]);
}
/* face KE-2 from v = val(KE-4..KE) (ud5.F90:293-302, cd6.F90:230-234) */
static inline double fv6_ke2(int s, const double *v, double vel) {
  if (s == ORC_UD5)
    return (-3.0 * v[4] + 27.0 * v[3] + 47.0 * v[2] - 13.0 * v[1] + 2.0 * v[0]) / 60.0 * (0.5 + fsign(0.5, vel))
         + (2.0 * v[2] + 5.0 * v[3] - v[4]) / 6.0 * (0.5 - fsign(0.5, vel));
  return F41 * (v[3] + v[2]) + F42 * (v[4] + v[1]);
}
double orc_face_value6(int s, int kind, const double *q, double vel) {
  switch (kind) {
    case 1: return fv6_ks(s, q, vel);
    case 2: return fv6_ke1(s, q + 3, vel);
    case 3: return fv6_ks1(s, q, vel);
    case 4: return fv6_ke2(s, q + 1, vel);
    default: return fv6_int(s, q[0], q[1], q[2], q[3], q[4], q[5], vel);
  }
}
static inline int six_point(int s) { return s == ORC_UD5 || s == ORC_CD6; }

/* F2H with the flux routine's own GSQRT argument (UD3.F90:75) */
#define F2H(G, k, p, i, j) (CDZ[(k) + (p)-1] * (G)[I3((k) + (p)-1, i, j)] / (CDZ[k] * (G)[I3(k, i, j)] + CDZ[(k) + 1] * (G)[I3((k) + 1, i, j)]))

/* ---- flux family (UD3.F90 / CD4.F90 / UD1.F90, 19 routines) -------------- */

/* valueW_Z (UD3.F90:98-153, CD4.F90:97-146): column arrays */
static void flux_valueW_Z(const orc_grid *g, int s, double *valW, const double *mflx, const double *val) {
  GRID_LOCALS;
  for (int k = KS + 1; k <= KE - 2; ++k) valW[k] = fv_int(s, val[k - 1], val[k], val[k + 1], val[k + 2], mflx[k]);
  valW[KS] = fv_lo(s, val[KS], val[KS + 1], val[KS + 2], mflx[KS]);
  valW[KE - 1] = fv_hi(s, val[KE - 2], val[KE - 1], val[KE], mflx[KE - 1]);
}

/* fluxZ_XYZ (UD3.F90:157-248, UD1.F90:135-171) */
static void fluxZ_XYZ(const orc_grid *g, int s, double *flux, const double *mflx, const double *val,
                      const double *G, const double *nd, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  if (s == ORC_UD1) {
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= IIE; ++i) {
        for (int k = KS; k <= KE - 1; ++k) {
          double vel = mflx[I3(k, i, j)];
          flux[I3(k, i, j)] = vel * fv_int(s, 0.0, val[I3(k, i, j)], val[I3(k + 1, i, j)], 0.0, vel);
        }
        flux[I3(KS - 1, i, j)] = 0.0;
        flux[I3(KE, i, j)] = 0.0;
      }
    return;
  }
  if (six_point(s)) { /* ud5.F90:194-321, cd6.F90:170-276: interior, then KS, KE-1, KS+1, KE-2 in this order */
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= IIE; ++i) {
        const double *v = val + I3(0, i, j);
        double *f = flux + I3(0, i, j);
        const double *m = mflx + I3(0, i, j), *gg = G + I3(0, i, j), *n = nd + I3(0, i, j);
        for (int k = KS + 2; k <= KE - 3; ++k)
          f[k] = m[k] * fv6_int(s, v[k - 2], v[k - 1], v[k], v[k + 1], v[k + 2], v[k + 3], m[k]) + gg[k] * n[k];
        f[KS - 1] = 0.0;
        f[KS] = m[KS] * fv6_ks(s, v + KS, m[KS]) + gg[KS] * n[KS];
        f[KE - 1] = m[KE - 1] * fv6_ke1(s, v + KE - 2, m[KE - 1]) + gg[KE - 1] * n[KE - 1];
        f[KS + 1] = m[KS + 1] * fv6_ks1(s, v + KS, m[KS + 1]) + gg[KS + 1] * n[KS + 1];
        f[KE - 2] = m[KE - 2] * fv6_ke2(s, v + KE - 4, m[KE - 2]) + gg[KE - 2] * n[KE - 2];
        f[KE] = 0.0;
      }
    return;
  }
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      for (int k = KS + 1; k <= KE - 2; ++k) {
        double vel = mflx[I3(k, i, j)];
        flux[I3(k, i, j)] = vel * fv_int(s, val[I3(k - 1, i, j)], val[I3(k, i, j)], val[I3(k + 1, i, j)], val[I3(k + 2, i, j)], vel)
                          + G[I3(k, i, j)] * nd[I3(k, i, j)];
      }
      flux[I3(KS - 1, i, j)] = 0.0;
      double vel = mflx[I3(KS, i, j)];
      flux[I3(KS, i, j)] = vel * fv_lo(s, val[I3(KS, i, j)], val[I3(KS + 1, i, j)], val[I3(KS + 2, i, j)], vel)
                         + G[I3(KS, i, j)] * nd[I3(KS, i, j)];
      vel = mflx[I3(KE - 1, i, j)];
      flux[I3(KE - 1, i, j)] = vel * fv_hi(s, val[I3(KE - 2, i, j)], val[I3(KE - 1, i, j)], val[I3(KE, i, j)], vel)
                             + G[I3(KE - 1, i, j)] * nd[I3(KE - 1, i, j)];
      flux[I3(KE, i, j)] = 0.0;
    }
}

/* fluxX_XYZ (UD3.F90:252-301, UD1.F90:175-210) */
static void fluxX_XYZ(const orc_grid *g, int s, double *flux, const double *mflx, const double *val,
                      const double *G, const double *nd, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS - 1; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double vel = mflx[I3(k, i, j)];
        if (s == ORC_UD1)
          flux[I3(k, i, j)] = vel * fv_int(s, 0.0, val[I3(k, i, j)], val[I3(k, i + 1, j)], 0.0, vel);
        else if (six_point(s)) /* ud5.F90:348-372, cd6.F90:290-310 */
          flux[I3(k, i, j)] = vel * fv6_int(s, val[I3(k, i - 2, j)], val[I3(k, i - 1, j)], val[I3(k, i, j)],
                                            val[I3(k, i + 1, j)], val[I3(k, i + 2, j)], val[I3(k, i + 3, j)], vel)
                            + G[I3(k, i, j)] * nd[I3(k, i, j)];
        else
          flux[I3(k, i, j)] = vel * fv_int(s, val[I3(k, i - 1, j)], val[I3(k, i, j)], val[I3(k, i + 1, j)], val[I3(k, i + 2, j)], vel)
                            + G[I3(k, i, j)] * nd[I3(k, i, j)];
      }
}

/* fluxY_XYZ (UD3.F90:305-354, UD1.F90:214-249) */
static void fluxY_XYZ(const orc_grid *g, int s, double *flux, const double *mflx, const double *val,
                      const double *G, const double *nd, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  OMP_FOR
  for (int j = JJS - 1; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double vel = mflx[I3(k, i, j)];
        if (s == ORC_UD1)
          flux[I3(k, i, j)] = vel * fv_int(s, 0.0, val[I3(k, i, j)], val[I3(k, i, j + 1)], 0.0, vel);
        else if (six_point(s)) /* ud5.F90:408-432, cd6.F90:347-367 */
          flux[I3(k, i, j)] = vel * fv6_int(s, val[I3(k, i, j - 2)], val[I3(k, i, j - 1)], val[I3(k, i, j)],
                                            val[I3(k, i, j + 1)], val[I3(k, i, j + 2)], val[I3(k, i, j + 3)], vel)
                            + G[I3(k, i, j)] * nd[I3(k, i, j)];
        else
          flux[I3(k, i, j)] = vel * fv_int(s, val[I3(k, i, j - 1)], val[I3(k, i, j)], val[I3(k, i, j + 1)], val[I3(k, i, j + 2)], vel)
                            + G[I3(k, i, j)] * nd[I3(k, i, j)];
      }
}

/* fluxZ_XYW (UD3.F90:358-452): flux(k-1) is written at loop index k */
static void fluxZ_XYW(const orc_grid *g, int s, double *flux, const double *mom, const double *val, const double *DENS,
                      const double *G, double J33G, const double *nd, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      for (int k = KS + 2; k <= KE - 1; ++k) {
        double vel = (0.5 * (mom[I3(k - 1, i, j)] + mom[I3(k, i, j)])) / DENS[I3(k, i, j)];
        flux[I3(k - 1, i, j)] = J33G * vel * fv_int(s, val[I3(k - 2, i, j)], val[I3(k - 1, i, j)], val[I3(k, i, j)], val[I3(k + 1, i, j)], vel)
                              + G[I3(k, i, j)] * nd[I3(k, i, j)];
      }
      flux[I3(KS - 1, i, j)] = 0.0;
      double vel = (0.5 * (mom[I3(KS, i, j)] + mom[I3(KS + 1, i, j)])) / DENS[I3(KS + 1, i, j)];
      flux[I3(KS, i, j)] = J33G * vel * fv_lo(s, val[I3(KS, i, j)], val[I3(KS + 1, i, j)], val[I3(KS + 2, i, j)], vel)
                         + G[I3(KS + 1, i, j)] * nd[I3(KS + 1, i, j)];
      flux[I3(KE - 1, i, j)] = 0.0;
      flux[I3(KE, i, j)] = 0.0;
    }
}

/* fluxJ13_XYW (UD3.F90:457-523) and fluxJ23_XYW (UD3.F90:527-593).
 * dir = 0: J13 (mom = MOMX, averaged i,i-1; MAPF(:,:,2));
 * dir = 1: J23 (mom = MOMY, averaged j,j-1; MAPF(:,:,1)). */
static void fluxJ_XYW(const orc_grid *g, int s, int dir, double *flux, const double *mom, const double *val, const double *DENS,
                      const double *J, const double *MAPF, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  const int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
  const int mn = dir == 0 ? 1 : 0;
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      const double mp = MAP(MAPF, i, j, mn);
      for (int k = KS + 2; k <= KE - 1; ++k) {
        double vel = (0.5 * (mom[I3(k, i, j)] + mom[I3(k, i - di, j - dj)])) / DENS[I3(k, i, j)];
        vel = vel * J[I3(k, i, j)];
        flux[I3(k - 1, i, j)] = vel / mp * fv_int(s, val[I3(k - 2, i, j)], val[I3(k - 1, i, j)], val[I3(k, i, j)], val[I3(k + 1, i, j)], vel);
      }
      flux[I3(KS - 1, i, j)] = 0.0;
      double vel = ((0.5 * (mom[I3(KS + 1, i, j)] + mom[I3(KS + 1, i - di, j - dj)])) / DENS[I3(KS + 1, i, j)]
                  + (0.5 * (mom[I3(KS, i, j)] + mom[I3(KS, i - di, j - dj)])) / DENS[I3(KS, i, j)]) * 0.5;
      vel = vel * J[I3(KS + 1, i, j)];
      flux[I3(KS, i, j)] = vel / mp * fv_lo(s, val[I3(KS, i, j)], val[I3(KS + 1, i, j)], val[I3(KS + 2, i, j)], vel);
      flux[I3(KE - 1, i, j)] = 0.0;
    }
}

/* fluxX_XYW (UD3.F90:598-671) and fluxY_XYW (UD3.F90:675-748) */
static void fluxH_XYW(const orc_grid *g, int s, int dir, double *flux, const double *mom, const double *val, const double *DENS,
                      const double *G, const double *MAPF, const double *nd, const double *CDZ,
                      int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  const int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
  const int mn = dir == 0 ? 1 : 0;
  OMP_FOR
  for (int j = JJS - dj; j <= JJE; ++j)
    for (int i = IIS - di; i <= IIE; ++i) {
      for (int k = KS; k <= KE - 1; ++k) {
        const double f1 = F2H(G, k, 1, i, j), f2 = F2H(G, k, 2, i, j);
        double vel = (f1 * mom[I3(k + 1, i, j)] + f2 * mom[I3(k, i, j)])
                   / (f1 * 0.5 * (DENS[I3(k + 1, i, j)] + DENS[I3(k + 1, i + di, j + dj)])
                    + f2 * 0.5 * (DENS[I3(k, i, j)] + DENS[I3(k, i + di, j + dj)]));
        flux[I3(k, i, j)] = G[I3(k, i, j)] / MAP(MAPF, i, j, mn) * vel
                          * fv_int(s, val[I3(k, i - di, j - dj)], val[I3(k, i, j)], val[I3(k, i + di, j + dj)], val[I3(k, i + 2 * di, j + 2 * dj)], vel)
                          + G[I3(k, i, j)] * nd[I3(k, i, j)];
      }
      flux[I3(KE, i, j)] = 0.0;
    }
}

/* fluxZ_UYZ (UD3.F90:753-845) and fluxZ_XVZ (UD3.F90:1173-1265) */
static void fluxZ_UVZ(const orc_grid *g, int s, int dir, double *flux, const double *mom, const double *val, const double *DENS,
                      const double *G, double J33G, const double *nd, const double *CDZ,
                      int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  const int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
#define VEL_ZUV(kk)                                                                                   \
  ((0.5 * (mom[I3(kk, i, j)] + mom[I3(kk, i + di, j + dj)]))                                          \
   / (F2H(G, kk, 1, i, j) * 0.5 * (DENS[I3((kk) + 1, i, j)] + DENS[I3((kk) + 1, i + di, j + dj)])     \
    + F2H(G, kk, 2, i, j) * 0.5 * (DENS[I3(kk, i, j)] + DENS[I3(kk, i + di, j + dj)])))
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      for (int k = KS + 1; k <= KE - 2; ++k) {
        double vel = VEL_ZUV(k);
        flux[I3(k, i, j)] = J33G * vel * fv_int(s, val[I3(k - 1, i, j)], val[I3(k, i, j)], val[I3(k + 1, i, j)], val[I3(k + 2, i, j)], vel)
                          + G[I3(k, i, j)] * nd[I3(k, i, j)];
      }
      flux[I3(KS - 1, i, j)] = 0.0;
      double vel = VEL_ZUV(KS);
      flux[I3(KS, i, j)] = J33G * vel * fv_lo(s, val[I3(KS, i, j)], val[I3(KS + 1, i, j)], val[I3(KS + 2, i, j)], vel)
                         + G[I3(KS, i, j)] * nd[I3(KS, i, j)];
      vel = VEL_ZUV(KE - 1);
      flux[I3(KE - 1, i, j)] = J33G * vel * fv_hi(s, val[I3(KE - 2, i, j)], val[I3(KE - 1, i, j)], val[I3(KE, i, j)], vel)
                             + G[I3(KE - 1, i, j)] * nd[I3(KE - 1, i, j)];
      flux[I3(KE, i, j)] = 0.0;
    }
#undef VEL_ZUV
}

/* fluxJ13_UYZ (UD3.F90:849-936), fluxJ23_UYZ (UD3.F90:940-1039),
 * fluxJ13_XVZ (UD3.F90:1269-1368), fluxJ23_XVZ (UD3.F90:1372-1459).
 * stag = 0: at (u,y,z) [dens averaged i,i+1], stag = 1: at (x,v,z) [j,j+1].
 * quad = 1 when mom lives at the other horizontal staggering (4-point mean). */
static void fluxJ_UVZ(const orc_grid *g, int s, int stag, int jdir, double *flux, const double *mom, const double *val,
                      const double *DENS, const double *G, const double *J, const double *MAPF, const double *CDZ,
                      int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  const int di = stag == 0 ? 1 : 0, dj = stag == 0 ? 0 : 1;
  const int quad = (stag == 0 && jdir == 1) || (stag == 1 && jdir == 0);
  /* 4-point means: UYZ<-XVW: (i,j),(i+1,j),(i,j-1),(i+1,j-1); XVZ<-UYW: (i,j),(i-1,j),(i,j+1),(i-1,j+1) */
  const int qi = stag == 0 ? 1 : -1, qj = stag == 0 ? -1 : 1;
  const int mn = jdir == 0 ? 1 : 0; /* J13 -> MAPF(:,:,2), J23 -> MAPF(:,:,1) */
#define FMOM(f, kk)                                                                                            \
  (quad ? (f) * 0.25 * (mom[I3(kk, i, j)] + mom[I3(kk, i + qi, j)] + mom[I3(kk, i, j + qj)] + mom[I3(kk, i + qi, j + qj)]) \
        : (f) * mom[I3(kk, i, j)])
#define VEL_J(kk)                                                                              \
  ((FMOM(F2H(G, kk, 1, i, j), (kk) + 1) + FMOM(F2H(G, kk, 2, i, j), kk))                       \
   / (F2H(G, kk, 1, i, j) * 0.5 * (DENS[I3((kk) + 1, i, j)] + DENS[I3((kk) + 1, i + di, j + dj)]) \
    + F2H(G, kk, 2, i, j) * 0.5 * (DENS[I3(kk, i, j)] + DENS[I3(kk, i + di, j + dj)])))
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      const double mp = MAP(MAPF, i, j, mn);
      for (int k = KS + 1; k <= KE - 2; ++k) {
        double vel = VEL_J(k);
        vel = vel * J[I3(k, i, j)];
        flux[I3(k, i, j)] = vel / mp * fv_int(s, val[I3(k - 1, i, j)], val[I3(k, i, j)], val[I3(k + 1, i, j)], val[I3(k + 2, i, j)], vel);
      }
      flux[I3(KS - 1, i, j)] = 0.0;
      double vel = VEL_J(KS);
      vel = vel * J[I3(KS, i, j)];
      flux[I3(KS, i, j)] = vel / mp * fv_lo(s, val[I3(KS, i, j)], val[I3(KS + 1, i, j)], val[I3(KS + 2, i, j)], vel);
      vel = VEL_J(KE - 1);
      vel = vel * J[I3(KE - 1, i, j)];
      flux[I3(KE - 1, i, j)] = vel / mp * fv_hi(s, val[I3(KE - 2, i, j)], val[I3(KE - 1, i, j)], val[I3(KE, i, j)], vel);
      flux[I3(KE, i, j)] = 0.0;
    }
#undef FMOM
#undef VEL_J
}

/* fluxX_UYZ (UD3.F90:1043-1097): flux(k,i-1,j) written at loop index i */
static void fluxX_UYZ(const orc_grid *g, int s, double *flux, const double *mom, const double *val, const double *DENS,
                      const double *G, const double *MAPF, const double *nd, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE + 1; ++i)
      for (int k = KS; k <= KE; ++k) {
        double vel = (0.5 * (mom[I3(k, i, j)] + mom[I3(k, i - 1, j)])) / (DENS[I3(k, i, j)]);
        flux[I3(k, i - 1, j)] = G[I3(k, i, j)] / MAP(MAPF, i, j, 1) * vel
                              * fv_int(s, val[I3(k, i - 2, j)], val[I3(k, i - 1, j)], val[I3(k, i, j)], val[I3(k, i + 1, j)], vel)
                              + G[I3(k, i, j)] * nd[I3(k, i, j)];
      }
}

/* fluxY_XVZ (UD3.F90:1536-1587): flux(k,i,j-1) written at loop index j */
static void fluxY_XVZ(const orc_grid *g, int s, double *flux, const double *mom, const double *val, const double *DENS,
                      const double *G, const double *MAPF, const double *nd, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  OMP_FOR
  for (int j = JJS; j <= JJE + 1; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double vel = (0.5 * (mom[I3(k, i, j)] + mom[I3(k, i, j - 1)])) / (DENS[I3(k, i, j)]);
        flux[I3(k, i, j - 1)] = G[I3(k, i, j)] / MAP(MAPF, i, j, 0) * vel
                              * fv_int(s, val[I3(k, i, j - 2)], val[I3(k, i, j - 1)], val[I3(k, i, j)], val[I3(k, i, j + 1)], vel)
                              + G[I3(k, i, j)] * nd[I3(k, i, j)];
      }
}

/* fluxY_UYZ (UD3.F90:1101-1150): at (u,v,z) */
static void fluxY_UYZ(const orc_grid *g, int s, double *flux, const double *mom, const double *val, const double *DENS,
                      const double *G, const double *MAPF, const double *nd, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  OMP_FOR
  for (int j = JJS - 1; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double vel = (0.5 * (mom[I3(k, i, j)] + mom[I3(k, i + 1, j)]))
                   / (0.25 * (DENS[I3(k, i, j)] + DENS[I3(k, i + 1, j)] + DENS[I3(k, i, j + 1)] + DENS[I3(k, i + 1, j + 1)]));
        flux[I3(k, i, j)] = G[I3(k, i, j)] / MAP(MAPF, i, j, 0) * vel
                          * fv_int(s, val[I3(k, i, j - 1)], val[I3(k, i, j)], val[I3(k, i, j + 1)], val[I3(k, i, j + 2)], vel)
                          + G[I3(k, i, j)] * nd[I3(k, i, j)];
      }
}

/* fluxX_XVZ (UD3.F90:1463-1512): at (u,v,z) */
static void fluxX_XVZ(const orc_grid *g, int s, double *flux, const double *mom, const double *val, const double *DENS,
                      const double *G, const double *MAPF, const double *nd, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS - 1; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double vel = (0.5 * (mom[I3(k, i, j)] + mom[I3(k, i, j + 1)]))
                   / (0.25 * (DENS[I3(k, i, j)] + DENS[I3(k, i + 1, j)] + DENS[I3(k, i, j + 1)] + DENS[I3(k, i + 1, j + 1)]));
        flux[I3(k, i, j)] = G[I3(k, i, j)] / MAP(MAPF, i, j, 1) * vel
                          * fv_int(s, val[I3(k, i - 1, j)], val[I3(k, i, j)], val[I3(k, i + 1, j)], val[I3(k, i + 2, j)], vel)
                          + G[I3(k, i, j)] * nd[I3(k, i, j)];
      }
}

/* ---- Thomas solver (HEVI.F90:1092-1127) --------------------------------- */
void orc_solve_direct(const orc_grid *g, double *C, const double *F1v, const double *F2v, const double *F3v) {
  GRID_LOCALS;
  const int KMAX = KE - KS + 1;
  double e[512], f[512]; /* e(1..KMAX-2) at e[1..]; KA <= 512 is checked by the callers */
  double *Cf = C - 1; /* Fortran C(1..KMAX-1) */
  double rdenom = 1.0 / F2v[KS];
  e[1] = -F1v[KS] * rdenom;
  f[1] = Cf[1] * rdenom;
  for (int k = 2; k <= KMAX - 2; ++k) {
    rdenom = 1.0 / (F2v[k + KS - 1] + F3v[k + KS - 1] * e[k - 1]);
    e[k] = -F1v[k + KS - 1] * rdenom;
    f[k] = (Cf[k] - F3v[k + KS - 1] * f[k - 1]) * rdenom;
  }
  Cf[KMAX - 1] = (Cf[KMAX - 1] - F3v[KE - 1] * f[KMAX - 2]) / (F2v[KE - 1] + F3v[KE - 1] * e[KMAX - 2]);
  for (int k = KMAX - 2; k >= 1; --k) Cf[k] = e[k] * Cf[k + 1] + f[k];
}

/* ---- HEVI stage (HEVI.F90:128-1089) -------------------------------------- */
int orc_tstep_short_hevi(const orc_grid *g, int s,
    double *DENS_RK, double *MOMZ_RK, double *MOMX_RK, double *MOMY_RK, double *RHOT_RK,
    double *mflx_hi, double *tflx_hi,
    const double *DENS0, const double *MOMZ0, const double *MOMX0, const double *MOMY0, const double *RHOT0,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef, double divdmp_coef, const double *DDIV,
    const orc_metric *m, const double *REF_dens, const double *REF_rhot, double dtrk) {
  GRID_LOCALS;
  if (s != ORC_UD3 && s != ORC_CD4) return -1;
  const double *CDZ = m->CDZ, *FDZ = m->FDZ, *FDX = m->FDX, *FDY = m->FDY;
  const double *RCDZ = m->RCDZ, *RCDX = m->RCDX, *RCDY = m->RCDY;
  const double *RFDZ = m->RFDZ, *RFDX = m->RFDX, *RFDY = m->RFDY;
  const double *PHI = m->PHI;
  const double J33G = m->J33G;
  const double *GS = m->GSQRT, *J13 = m->J13G, *J23 = m->J23G, *MP = m->MAPF;
#define GSQRT(k, i, j, c) GS[I3(k, i, j) + N3 * (c)]
#define J13G(k, i, j, c) J13[I3(k, i, j) + N3 * (c)]
#define J23G(k, i, j, c) J23[I3(k, i, j) + N3 * (c)]
#define MAPF(i, j, n, c) MP[(size_t)(i) + (size_t)IA * ((size_t)(j) + (size_t)JA * ((size_t)(n) + 2 * (size_t)(c)))]
#define ND(k, i, j, v, d) num_diff[I3(k, i, j) + N3 * ((size_t)(v) + 5 * (size_t)(d))]
#define F2HH(k, p, c) (CDZ[(k) + (p)-1] * GSQRT((k) + (p)-1, i, j, c) / (CDZ[k] * GSQRT(k, i, j, c) + CDZ[(k) + 1] * GSQRT((k) + 1, i, j, c)))
  const size_t n3 = N3;
  double *mfZ = mflx_hi, *mfX = mflx_hi + n3, *mfY = mflx_hi + 2 * n3;
  double *tfZ = tflx_hi, *tfX = tflx_hi + n3, *tfY = tflx_hi + 2 * n3;
  double *POTT = alloc_nan(n3), *DPRES = alloc_nan(n3);
  double *qflx_hi = alloc_nan(3 * n3), *qflx_J13 = alloc_nan(n3), *qflx_J23 = alloc_nan(n3);
  double *Sr = alloc_nan(n3), *Sw = alloc_nan(n3), *St = alloc_nan(n3);
  double *qZ = qflx_hi, *qX = qflx_hi + n3, *qY = qflx_hi + 2 * n3;
  const int IFS_OFF = 1, JFS_OFF = 1; /* BND_W = BND_S = .false. (HEVI.F90:332-335) */
  const int IEH = IE, JEH = JE;
  const int IIS = IS, IIE = IE, JJS = JS, JJE = JE; /* IBLOCK=IMAX, JBLOCK=JMAX */

  /* hevi_pres (HEVI.F90:346-360) */
  OMP_FOR
  for (int j = JJS; j <= JJE + 1; ++j)
    for (int i = IIS; i <= IIE + 1; ++i) {
      for (int k = KS; k <= KE; ++k) DPRES[I3(k, i, j)] = DPRES0[I3(k, i, j)] + RT2P[I3(k, i, j)] * (RHOT[I3(k, i, j)] - REF_rhot[I3(k, i, j)]);
      DPRES[I3(KS - 1, i, j)] = DPRES0[I3(KS - 1, i, j)] - DENS[I3(KS, i, j)] * (PHI[I3(KS - 1, i, j)] - PHI[I3(KS + 1, i, j)]);
      DPRES[I3(KE + 1, i, j)] = DPRES0[I3(KE + 1, i, j)] - DENS[I3(KE, i, j)] * (PHI[I3(KE + 1, i, j)] - PHI[I3(KE - 1, i, j)]);
    }

  /* hevi_mflx_z (HEVI.F90:369-392) */
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS - 1; i <= IIE; ++i) {
      mfZ[I3(KS - 1, i, j)] = 0.0;
      for (int k = KS; k <= KE - 1; ++k)
        mfZ[I3(k, i, j)] = J13G(k, i, j, I_XYW) * 0.25 / MAPF(i, j, 1, I_XY)
                             * (MOMX[I3(k + 1, i, j)] + MOMX[I3(k + 1, i - 1, j)] + MOMX[I3(k, i, j)] + MOMX[I3(k, i - 1, j)])
                         + J23G(k, i, j, I_XYW) * 0.25 / MAPF(i, j, 0, I_XY)
                             * (MOMY[I3(k + 1, i, j)] + MOMY[I3(k + 1, i, j - 1)] + MOMY[I3(k, i, j)] + MOMY[I3(k, i, j - 1)])
                         + GSQRT(k, i, j, I_XYW) / (MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY)) * ND(k, i, j, I_DENS, ZDIR);
      mfZ[I3(KE, i, j)] = 0.0;
    }

  /* hevi_mflx_x (HEVI.F90:400-417) */
  {
    int iss = (IIS - 1 > IS - IFS_OFF) ? IIS - 1 : IS - IFS_OFF;
    int iee = IIE < IEH ? IIE : IEH;
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = iss; i <= iee; ++i)
        for (int k = KS; k <= KE; ++k)
          mfX[I3(k, i, j)] = GSQRT(k, i, j, I_UYZ) / MAPF(i, j, 1, I_UY) * (MOMX[I3(k, i, j)] + ND(k, i, j, I_DENS, XDIR));
  }
  /* (HEVI.F90:426-438) */
  {
    int jss = (JJS - 1 > JS - JFS_OFF) ? JJS - 1 : JS - JFS_OFF;
    int jee = JJE < JEH ? JJE : JEH;
    OMP_FOR
    for (int j = jss; j <= jee; ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE; ++k)
          mfY[I3(k, i, j)] = GSQRT(k, i, j, I_XVZ) / MAPF(i, j, 0, I_XV) * (MOMY[I3(k, i, j)] + ND(k, i, j, I_DENS, YDIR));
  }

  /* hevi_sr (HEVI.F90:453-475) */
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double advcv = -(mfZ[I3(k, i, j)] - mfZ[I3(k - 1, i, j)]) * RCDZ[k];
        double advch = -((mfX[I3(k, i, j)] - mfX[I3(k, i - 1, j)]) * RCDX[i]
                       + (mfY[I3(k, i, j)] - mfY[I3(k, i, j - 1)]) * RCDY[j]);
        Sr[I3(k, i, j)] = (advcv + advch) * MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY) / GSQRT(k, i, j, I_XYZ) + DENS_t[I3(k, i, j)];
      }

  /* momentum z fluxes (HEVI.F90:486-525) */
  fluxZ_XYW(g, s, qZ, MOMZ, MOMZ, DENS, GS + n3 * I_XYZ, J33G, num_diff + n3 * (I_MOMZ + 5 * ZDIR), IIS, IIE, JJS, JJE);
  fluxJ_XYW(g, s, 0, qflx_J13, MOMX, MOMZ, DENS, J13 + n3 * I_XYZ, MP + (size_t)IA * JA * 2 * I_XY, IIS, IIE, JJS, JJE);
  fluxJ_XYW(g, s, 1, qflx_J23, MOMY, MOMZ, DENS, J23 + n3 * I_XYZ, MP + (size_t)IA * JA * 2 * I_XY, IIS, IIE, JJS, JJE);
  fluxH_XYW(g, s, 0, qX, MOMX, MOMZ, DENS, GS + n3 * I_UYW, MP + (size_t)IA * JA * 2 * I_UY, num_diff + n3 * (I_MOMZ + 5 * XDIR), CDZ, IIS, IIE, JJS, JJE);
  fluxH_XYW(g, s, 1, qY, MOMY, MOMZ, DENS, GS + n3 * I_XVW, MP + (size_t)IA * JA * 2 * I_XV, num_diff + n3 * (I_MOMZ + 5 * YDIR), CDZ, IIS, IIE, JJS, JJE);

  /* hevi_sw (HEVI.F90:538-577) */
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE - 1; ++k) {
        double advcv = -(qZ[I3(k, i, j)] - qZ[I3(k - 1, i, j)]
                       + qflx_J13[I3(k, i, j)] - qflx_J13[I3(k - 1, i, j)]
                       + qflx_J23[I3(k, i, j)] - qflx_J23[I3(k - 1, i, j)]) * RFDZ[k];
        double advch = -((qX[I3(k, i, j)] - qX[I3(k, i - 1, j)]) * RCDX[i]
                       + (qY[I3(k, i, j)] - qY[I3(k, i, j - 1)]) * RCDY[j])
                     * MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY);
        double wdmp = -wdamp_coef[k] * MOMZ0[I3(k, i, j)];
        double div = divdmp_coef / dtrk * (DDIV[I3(k + 1, i, j)] - DDIV[I3(k, i, j)]) * FDZ[k];
        Sw[I3(k, i, j)] = (advcv + advch) / GSQRT(k, i, j, I_XYW) + wdmp + div + MOMZ_t[I3(k, i, j)];
      }

  /* POTT (HEVI.F90:588-598) */
  OMP_FOR
  for (int j = JJS - 2; j <= JJE + 2; ++j)
    for (int i = IIS - 2; i <= IIE + 2; ++i)
      for (int k = KS; k <= KE; ++k) POTT[I3(k, i, j)] = RHOT[I3(k, i, j)] / DENS[I3(k, i, j)];
  if (g_trace) memcpy(g_trace->POTT, POTT, n3 * sizeof(double));

  /* theta fluxes (HEVI.F90:604-622) */
  fluxZ_XYZ(g, s, tfZ, mfZ, POTT, GS + n3 * I_XYW, num_diff + n3 * (I_RHOT + 5 * ZDIR), IIS, IIE, JJS, JJE);
  fluxX_XYZ(g, s, tfX, mfX, POTT, GS + n3 * I_UYZ, num_diff + n3 * (I_RHOT + 5 * XDIR), IIS, IIE, JJS, JJE);
  fluxY_XYZ(g, s, tfY, mfY, POTT, GS + n3 * I_XVZ, num_diff + n3 * (I_RHOT + 5 * YDIR), IIS, IIE, JJS, JJE);

  /* hevi_st (HEVI.F90:634-657) */
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double advcv = -(tfZ[I3(k, i, j)] - tfZ[I3(k - 1, i, j)]) * RCDZ[k];
        double advch = -((tfX[I3(k, i, j)] - tfX[I3(k, i - 1, j)]) * RCDX[i]
                       + (tfY[I3(k, i, j)] - tfY[I3(k, i, j - 1)]) * RCDY[j]);
        St[I3(k, i, j)] = (advcv + advch) * MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY) / GSQRT(k, i, j, I_XYZ) + RHOT_t[I3(k, i, j)];
      }

  if (g_trace) memcpy(g_trace->St, St, n3 * sizeof(double));

  /* hevi_solver (HEVI.F90:688-804) */
#pragma omp parallel for collapse(2) num_threads(g_threads) schedule(static)
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      double A[512], F1v[512], F2v[512], F3v[512], PT[512], C[512];
      double Mcol[512], Pcol[512];
      for (int k = 0; k < KA; ++k) { Mcol[k] = MOMZ[I3(k, i, j)]; Pcol[k] = POTT[I3(k, i, j)]; }
      flux_valueW_Z(g, s, PT, Mcol, Pcol);
      for (int k = KS; k <= KE; ++k)
        A[k] = dtrk * dtrk * J33G * RCDZ[k] * RT2P[I3(k, i, j)] * J33G / GSQRT(k, i, j, I_XYZ);
      double B = GRAV * (dtrk * dtrk) * J33G / (CDZ[KS + 1] + CDZ[KS]);
      F1v[KS] = -(PT[KS + 1] * RFDZ[KS] * A[KS + 1] + B) / GSQRT(KS, i, j, I_XYW);
      F2v[KS] = 1.0 + (PT[KS] * RFDZ[KS] * (A[KS + 1] + A[KS])) / GSQRT(KS, i, j, I_XYW);
      for (int k = KS + 1; k <= KE - 2; ++k) {
        B = GRAV * (dtrk * dtrk) * J33G / (CDZ[k + 1] + CDZ[k]);
        F1v[k] = -(PT[k + 1] * RFDZ[k] * A[k + 1] + B) / GSQRT(k, i, j, I_XYW);
        F2v[k] = 1.0 + (PT[k] * RFDZ[k] * (A[k + 1] + A[k])) / GSQRT(k, i, j, I_XYW);
        F3v[k] = -(PT[k - 1] * RFDZ[k] * A[k] - B) / GSQRT(k, i, j, I_XYW);
      }
      B = GRAV * (dtrk * dtrk) * J33G / (CDZ[KE] + CDZ[KE - 1]);
      F2v[KE - 1] = 1.0 + (PT[KE - 1] * RFDZ[KE - 1] * (A[KE] + A[KE - 1])) / GSQRT(KE - 1, i, j, I_XYW);
      F3v[KE - 1] = -(PT[KE - 2] * RFDZ[KE - 1] * A[KE - 1] - B) / GSQRT(KE - 1, i, j, I_XYW);
      for (int k = KS; k <= KE - 1; ++k) {
        double pg = -(DPRES[I3(k + 1, i, j)] + RT2P[I3(k + 1, i, j)] * dtrk * St[I3(k + 1, i, j)]
                      - DPRES[I3(k, i, j)] - RT2P[I3(k, i, j)] * dtrk * St[I3(k, i, j)])
                      * RFDZ[k] * J33G / GSQRT(k, i, j, I_XYW)
                    - GRAV
                      * (F2HH(k, 1, I_XYZ) * (DENS[I3(k + 1, i, j)] - REF_dens[I3(k + 1, i, j)] + Sr[I3(k + 1, i, j)] * dtrk)
                       + F2HH(k, 2, I_XYZ) * (DENS[I3(k, i, j)] - REF_dens[I3(k, i, j)] + Sr[I3(k, i, j)] * dtrk));
        C[k - KS] = MOMZ[I3(k, i, j)] + dtrk * (pg + Sw[I3(k, i, j)]);
      }
      orc_solve_direct(g, C, F1v, F2v, F3v);
      if (g_trace)
        for (int k = KS; k <= KE - 1; ++k) {
          const size_t p = I3(k, i, j);
          g_trace->PT[p] = PT[k];
          g_trace->F2[p] = F2v[k];
          g_trace->F1[p] = k <= KE - 2 ? F1v[k] : 0.0;
          g_trace->F3[p] = k >= KS + 1 ? F3v[k] : 0.0;
          g_trace->Cp[p] = C[k - KS];
        }
      /* C(k-KS+1) in Fortran is C[k-KS] here */
      for (int k = KS; k <= KE - 1; ++k) {
        const double ck = C[k - KS];
        mfZ[I3(k, i, j)] = mfZ[I3(k, i, j)] + J33G * ck / (MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY));
        tfZ[I3(k, i, j)] = tfZ[I3(k, i, j)] + J33G * ck * PT[k] / (MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY));
        MOMZ_RK[I3(k, i, j)] = MOMZ0[I3(k, i, j)] + (ck - MOMZ[I3(k, i, j)]);
      }
      MOMZ_RK[I3(KS - 1, i, j)] = 0.0;
      MOMZ_RK[I3(KE, i, j)] = 0.0;
      double advcv = -C[0] * J33G * RCDZ[KS] / GSQRT(KS, i, j, I_XYZ);
      DENS_RK[I3(KS, i, j)] = DENS0[I3(KS, i, j)] + dtrk * (advcv + Sr[I3(KS, i, j)]);
      advcv = -C[0] * PT[KS] * J33G * RCDZ[KS] / GSQRT(KS, i, j, I_XYZ);
      RHOT_RK[I3(KS, i, j)] = RHOT0[I3(KS, i, j)] + dtrk * (advcv + St[I3(KS, i, j)]);
      for (int k = KS + 1; k <= KE - 1; ++k) {
        advcv = -(C[k - KS] - C[k - KS - 1]) * J33G * RCDZ[k] / GSQRT(k, i, j, I_XYZ);
        DENS_RK[I3(k, i, j)] = DENS0[I3(k, i, j)] + dtrk * (advcv + Sr[I3(k, i, j)]);
        advcv = -(C[k - KS] * PT[k] - C[k - KS - 1] * PT[k - 1]) * J33G * RCDZ[k] / GSQRT(k, i, j, I_XYZ);
        RHOT_RK[I3(k, i, j)] = RHOT0[I3(k, i, j)] + dtrk * (advcv + St[I3(k, i, j)]);
      }
      advcv = C[KE - KS - 1] * J33G * RCDZ[KE] / GSQRT(KE, i, j, I_XYZ);
      DENS_RK[I3(KE, i, j)] = DENS0[I3(KE, i, j)] + dtrk * (advcv + Sr[I3(KE, i, j)]);
      advcv = C[KE - KS - 1] * PT[KE - 1] * J33G * RCDZ[KE] / GSQRT(KE, i, j, I_XYZ);
      RHOT_RK[I3(KE, i, j)] = RHOT0[I3(KE, i, j)] + dtrk * (advcv + St[I3(KE, i, j)]);
    }

  /* momentum x (HEVI.F90:816-927) */
  fluxZ_UVZ(g, s, 0, qZ, MOMZ, MOMX, DENS, GS + n3 * I_UYW, J33G, num_diff + n3 * (I_MOMX + 5 * ZDIR), CDZ, IIS, IIE, JJS, JJE);
  fluxJ_UVZ(g, s, 0, 0, qflx_J13, MOMX, MOMX, DENS, GS + n3 * I_UYZ, J13 + n3 * I_UYW, MP + (size_t)IA * JA * 2 * I_UY, CDZ, IIS, IIE, JJS, JJE);
  fluxJ_UVZ(g, s, 0, 1, qflx_J23, MOMY, MOMX, DENS, GS + n3 * I_UYZ, J23 + n3 * I_UYW, MP + (size_t)IA * JA * 2 * I_UY, CDZ, IIS, IIE, JJS, JJE);
  fluxX_UYZ(g, s, qX, MOMX, MOMX, DENS, GS + n3 * I_XYZ, MP + (size_t)IA * JA * 2 * I_XY, num_diff + n3 * (I_MOMX + 5 * XDIR), IIS, IIE, JJS, JJE);
  fluxY_UYZ(g, s, qY, MOMY, MOMX, DENS, GS + n3 * I_UVZ, MP + (size_t)IA * JA * 2 * I_UV, num_diff + n3 * (I_MOMX + 5 * YDIR), IIS, IIE, JJS, JJE);
  {
    int iee = IIE < IEH ? IIE : IEH;
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= iee; ++i)
        for (int k = KS; k <= KE; ++k) {
          double advcv = -(qZ[I3(k, i, j)] - qZ[I3(k - 1, i, j)]
                         + qflx_J13[I3(k, i, j)] - qflx_J13[I3(k - 1, i, j)]
                         + qflx_J23[I3(k, i, j)] - qflx_J23[I3(k - 1, i, j)]) * RCDZ[k];
          double advch = -((qX[I3(k, i, j)] - qX[I3(k, i - 1, j)]) * RFDX[i]
                         + (qY[I3(k, i, j)] - qY[I3(k, i, j - 1)]) * RCDY[j])
                       * MAPF(i, j, 0, I_UY) * MAPF(i, j, 1, I_UY);
          double pg = ((GSQRT(k, i + 1, j, I_XYZ) * DPRES[I3(k, i + 1, j)]
                      - GSQRT(k, i, j, I_XYZ) * DPRES[I3(k, i, j)]) * RFDX[i]
                     + (J13G(k, i, j, I_UYW)
                        * 0.5 * (F2HH(k, 1, I_UYZ) * (DPRES[I3(k + 1, i + 1, j)] + DPRES[I3(k + 1, i, j)])
                               + F2HH(k, 2, I_UYZ) * (DPRES[I3(k, i + 1, j)] + DPRES[I3(k, i, j)]))
                      - J13G(k - 1, i, j, I_UYW)
                        * 0.5 * (F2HH(k, 1, I_UYZ) * (DPRES[I3(k, i + 1, j)] + DPRES[I3(k, i, j)])
                               + F2HH(k, 2, I_UYZ) * (DPRES[I3(k - 1, i + 1, j)] + DPRES[I3(k - 1, i, j)])))
                       * RCDZ[k])
                    * MAPF(i, j, 0, I_UY);
          double cf = 0.125 * (CORIOLI[(size_t)(i + 1) + (size_t)IA * j] + CORIOLI[(size_t)i + (size_t)IA * j])
                        * (MOMY[I3(k, i + 1, j)] + MOMY[I3(k, i, j)] + MOMY[I3(k, i + 1, j - 1)] + MOMY[I3(k, i, j - 1)])
                    + 0.25 * MAPF(i, j, 0, I_UY) * MAPF(i, j, 1, I_UY)
                        * (MOMY[I3(k, i, j)] + MOMY[I3(k, i, j - 1)] + MOMY[I3(k, i + 1, j)] + MOMY[I3(k, i + 1, j - 1)])
                        * ((MOMY[I3(k, i, j)] + MOMY[I3(k, i, j - 1)] + MOMY[I3(k, i + 1, j)] + MOMY[I3(k, i + 1, j - 1)]) * 0.25
                             * (1.0 / MAPF(i + 1, j, 1, I_XY) - 1.0 / MAPF(i, j, 1, I_XY)) * RFDX[i]
                           - MOMX[I3(k, i, j)]
                             * (1.0 / MAPF(i, j, 0, I_UV) - 1.0 / MAPF(i, j - 1, 0, I_UV)) * RCDY[j])
                        * 2.0 / (DENS[I3(k, i + 1, j)] + DENS[I3(k, i, j)]);
          double div = divdmp_coef / dtrk * (DDIV[I3(k, i + 1, j)] / MAPF(i + 1, j, 1, I_XY) - DDIV[I3(k, i, j)] / MAPF(i, j, 0, I_XY))
                     * MAPF(i, j, 0, I_UY) * MAPF(i, j, 1, I_UY) * FDX[i];
          MOMX_RK[I3(k, i, j)] = MOMX0[I3(k, i, j)]
                               + dtrk * ((advcv + advch - pg) / GSQRT(k, i, j, I_UYZ) + cf + div + MOMX_t[I3(k, i, j)]);
        }
  }

  /* momentum y (HEVI.F90:936-1047) */
  fluxZ_UVZ(g, s, 1, qZ, MOMZ, MOMY, DENS, GS + n3 * I_XVW, J33G, num_diff + n3 * (I_MOMY + 5 * ZDIR), CDZ, IIS, IIE, JJS, JJE);
  fluxJ_UVZ(g, s, 1, 0, qflx_J13, MOMX, MOMY, DENS, GS + n3 * I_XVZ, J13 + n3 * I_XVW, MP + (size_t)IA * JA * 2 * I_XV, CDZ, IIS, IIE, JJS, JJE);
  fluxJ_UVZ(g, s, 1, 1, qflx_J23, MOMY, MOMY, DENS, GS + n3 * I_XVZ, J23 + n3 * I_XVW, MP + (size_t)IA * JA * 2 * I_XV, CDZ, IIS, IIE, JJS, JJE);
  fluxX_XVZ(g, s, qX, MOMX, MOMY, DENS, GS + n3 * I_UVZ, MP + (size_t)IA * JA * 2 * I_UV, num_diff + n3 * (I_MOMY + 5 * XDIR), IIS, IIE, JJS, JJE);
  fluxY_XVZ(g, s, qY, MOMY, MOMY, DENS, GS + n3 * I_XYZ, MP + (size_t)IA * JA * 2 * I_XY, num_diff + n3 * (I_MOMY + 5 * YDIR), IIS, IIE, JJS, JJE);
  {
    int jee = JJE < JEH ? JJE : JEH;
    OMP_FOR
    for (int j = JJS; j <= jee; ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE; ++k) {
          double advcv = -(qZ[I3(k, i, j)] - qZ[I3(k - 1, i, j)]
                         + qflx_J13[I3(k, i, j)] - qflx_J13[I3(k - 1, i, j)]
                         + qflx_J23[I3(k, i, j)] - qflx_J23[I3(k - 1, i, j)]) * RCDZ[k];
          double advch = -((qX[I3(k, i, j)] - qX[I3(k, i - 1, j)]) * RCDX[i]
                         + (qY[I3(k, i, j)] - qY[I3(k, i, j - 1)]) * RFDY[j])
                       * MAPF(i, j, 0, I_XV) * MAPF(i, j, 1, I_XV);
          double pg = ((GSQRT(k, i, j + 1, I_XYZ) * DPRES[I3(k, i, j + 1)]
                      - GSQRT(k, i, j, I_XYZ) * DPRES[I3(k, i, j)]) * RFDY[j]
                     + (J23G(k, i, j, I_XVW)
                        * 0.5 * (F2HH(k, 1, I_XVZ) * (DPRES[I3(k + 1, i, j + 1)] + DPRES[I3(k + 1, i, j)])
                               + F2HH(k, 2, I_XVZ) * (DPRES[I3(k, i, j + 1)] + DPRES[I3(k, i, j)]))
                      - J23G(k - 1, i, j, I_XVW)
                        * 0.5 * (F2HH(k - 1, 1, I_XVZ) * (DPRES[I3(k, i, j + 1)] + DPRES[I3(k, i, j)])
                               + F2HH(k - 1, 2, I_XVZ) * (DPRES[I3(k - 1, i, j + 1)] + DPRES[I3(k - 1, i, j)])))
                       * RCDZ[k])
                    * MAPF(i, j, 1, I_XV);
          double cf = -0.125 * (CORIOLI[(size_t)i + (size_t)IA * (j + 1)] + CORIOLI[(size_t)i + (size_t)IA * j])
                        * (MOMX[I3(k, i, j + 1)] + MOMX[I3(k, i, j)] + MOMX[I3(k, i - 1, j + 1)] + MOMX[I3(k, i - 1, j)])
                    - 0.25 * MAPF(i, j, 0, I_XV) * MAPF(i, j, 1, I_XV)
                        * (MOMX[I3(k, i, j)] + MOMX[I3(k, i - 1, j)] + MOMX[I3(k, i, j + 1)] + MOMX[I3(k, i - 1, j + 1)])
                        * (MOMY[I3(k, i, j)]
                             * (1.0 / MAPF(i, j, 1, I_UV) - 1.0 / MAPF(i - 1, j, 1, I_UV)) * RCDX[i]
                           - 0.25 * (MOMX[I3(k, i, j)] + MOMX[I3(k, i - 1, j)] + MOMX[I3(k, i, j + 1)] + MOMX[I3(k, i - 1, j + 1)])
                             * (1.0 / MAPF(i, j + 1, 0, I_XY) - 1.0 / MAPF(i, j, 0, I_XY)) * RFDY[j])
                        * 2.0 / (DENS[I3(k, i, j + 1)] + DENS[I3(k, i, j)]);
          double div = divdmp_coef / dtrk * (DDIV[I3(k, i, j + 1)] / MAPF(i, j + 1, 0, I_XY) - DDIV[I3(k, i, j)] / MAPF(i, j, 0, I_XY))
                     * MAPF(i, j, 0, I_XV) * MAPF(i, j, 1, I_XV) * FDY[j];
          MOMY_RK[I3(k, i, j)] = MOMY0[I3(k, i, j)]
                               + dtrk * ((advcv + advch - pg) / GSQRT(k, i, j, I_XVZ) + cf + div + MOMY_t[I3(k, i, j)]);
        }
  }
  free(POTT); free(DPRES); free(qflx_hi); free(qflx_J13); free(qflx_J23);
  free(Sr); free(Sw); free(St);
  return 0;
#undef GSQRT
#undef J13G
#undef J23G
#undef MAPF
#undef ND
#undef F2HH
}

/* ---- HEVE stage (scale_atmos_dyn_tstep_short_fvm_heve.F90:109-1516) -----
 * The fully explicit short step with its optional FCT of the momentum and
 * theta fluxes.  HEVE = scalelib/src/atmos-rm/dynamics/
 * scale_atmos_dyn_tstep_short_fvm_heve.F90.  The flux routines are the ones
 * the HEVI stage uses; the differences are the explicit vertical mass flux,
 * the explicit pressure-gradient and buoyancy terms of MOMZ and the
 * divergence-damping forms. */
/* ---- UD1 momentum fluxes of the HEVE momentum FCT (UD1.F90:299-1288) ------
 * F1 (q1 + q0) - sign(F1, vel) (q1 - q0), without num_diff; the face
 * velocities as in the UD3 routines.  Loop ranges follow the reference,
 * called with IIS-1..IIE+1, JJS-1..JJE+1. */
static double ud1f(double q0, double q1, double vel) { return fv_int(ORC_UD1, 0.0, q0, q1, 0.0, vel); }

/* fluxZ_XYW_ud1 (UD1.F90:299-375): z-index added by -1 */
static void fluxZ_XYW_ud1(const orc_grid *g, double *flux, const double *mom, const double *val, const double *DENS,
                          double J33G, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      for (int k = KS + 1; k <= KE - 1; ++k) {
        const double vel = (0.5 * (mom[I3(k - 1, i, j)] + mom[I3(k, i, j)])) / DENS[I3(k, i, j)];
        flux[I3(k - 1, i, j)] = J33G * vel * ud1f(val[I3(k - 1, i, j)], val[I3(k, i, j)], vel);
      }
      flux[I3(KS - 1, i, j)] = 0.0;
      flux[I3(KE - 1, i, j)] = 0.0;
      flux[I3(KE, i, j)] = 0.0;
    }
}

/* fluxX_XYW_ud1 (dir 0, UD1.F90:519-587) and fluxY_XYW_ud1 (dir 1, :591-659) */
static void fluxH_XYW_ud1(const orc_grid *g, int dir, double *flux, const double *mom, const double *val,
                          const double *DENS, const double *G, const double *MAPF, const double *CDZ,
                          int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  const int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
  const int mn = dir == 0 ? 1 : 0;
  OMP_FOR
  for (int j = JJS - dj; j <= JJE; ++j)
    for (int i = IIS - di; i <= IIE; ++i) {
      for (int k = KS; k <= KE - 1; ++k) {
        const double f1 = F2H(G, k, 1, i, j), f2 = F2H(G, k, 2, i, j);
        const double vel = (f1 * mom[I3(k + 1, i, j)] + f2 * mom[I3(k, i, j)])
                         / (f1 * 0.5 * (DENS[I3(k + 1, i, j)] + DENS[I3(k + 1, i + di, j + dj)])
                          + f2 * 0.5 * (DENS[I3(k, i, j)] + DENS[I3(k, i + di, j + dj)]));
        flux[I3(k, i, j)] = G[I3(k, i, j)] / MAP(MAPF, i, j, mn) * vel
                          * ud1f(val[I3(k, i, j)], val[I3(k, i + di, j + dj)], vel);
      }
      flux[I3(KE, i, j)] = 0.0;
    }
}

/* fluxZ_UYZ_ud1 (stag 0, UD1.F90:664-737) and fluxZ_XVZ_ud1 (stag 1, :979-1052) */
static void fluxZ_UVZ_ud1(const orc_grid *g, int stag, double *flux, const double *mom, const double *val,
                          const double *DENS, const double *G, double J33G, const double *CDZ,
                          int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  const int di = stag == 0 ? 1 : 0, dj = stag == 0 ? 0 : 1;
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      for (int k = KS; k <= KE - 1; ++k) {
        const double vel = (0.5 * (mom[I3(k, i, j)] + mom[I3(k, i + di, j + dj)]))
                         / (F2H(G, k, 1, i, j) * 0.5 * (DENS[I3(k + 1, i, j)] + DENS[I3(k + 1, i + di, j + dj)])
                          + F2H(G, k, 2, i, j) * 0.5 * (DENS[I3(k, i, j)] + DENS[I3(k, i + di, j + dj)]));
        flux[I3(k, i, j)] = J33G * vel * ud1f(val[I3(k, i, j)], val[I3(k + 1, i, j)], vel);
      }
      flux[I3(KS - 1, i, j)] = 0.0;
      flux[I3(KE, i, j)] = 0.0;
    }
}

/* the flux along the momentum's own direction, stored one cell back:
 * fluxX_UYZ_ud1 (dir 0, UD1.F90:869-920) and fluxY_XVZ_ud1 (dir 1, :1237-1288) */
static void fluxS_ud1(const orc_grid *g, int dir, double *flux, const double *mom, const double *val,
                      const double *DENS, const double *G, const double *MAPF, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  const int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
  const int mn = dir == 0 ? 1 : 0;
  OMP_FOR
  for (int j = JJS; j <= JJE + dj; ++j)
    for (int i = IIS; i <= IIE + di; ++i)
      for (int k = KS; k <= KE; ++k) {
        const double vel = (0.5 * (mom[I3(k, i, j)] + mom[I3(k, i - di, j - dj)])) / (DENS[I3(k, i, j)]);
        flux[I3(k, i - di, j - dj)] = G[I3(k, i, j)] / MAP(MAPF, i, j, mn) * vel
                                    * ud1f(val[I3(k, i - di, j - dj)], val[I3(k, i, j)], vel);
      }
}

/* the cross flux: fluxY_UYZ_ud1 (stag 0, UD1.F90:924-973; faces j, j+1 of the
 * u point) and fluxX_XVZ_ud1 (stag 1, :1184-1233; faces i, i+1 of the v point) */
static void fluxC_ud1(const orc_grid *g, int stag, double *flux, const double *mom, const double *val,
                      const double *DENS, const double *G, const double *MAPF, int IIS, int IIE, int JJS, int JJE) {
  GRID_LOCALS;
  const int di = stag == 0 ? 1 : 0, dj = stag == 0 ? 0 : 1;  /* the momentum's own staggering */
  const int ci = stag == 0 ? 0 : 1, cj = stag == 0 ? 1 : 0;  /* the flux direction */
  const int mn = stag == 0 ? 0 : 1;
  OMP_FOR
  for (int j = JJS - cj; j <= JJE; ++j)
    for (int i = IIS - ci; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        const double vel = (0.5 * (mom[I3(k, i, j)] + mom[I3(k, i + di, j + dj)]))
                         / (0.25 * (DENS[I3(k, i, j)] + DENS[I3(k, i + 1, j)] + DENS[I3(k, i, j + 1)] + DENS[I3(k, i + 1, j + 1)]));
        flux[I3(k, i, j)] = G[I3(k, i, j)] / MAP(MAPF, i, j, mn) * vel
                          * ud1f(val[I3(k, i, j)], val[I3(k, i + ci, j + cj)], vel);
      }
}

int orc_tstep_short_heve(const orc_grid *g, int s,
    double *DENS_RK, double *MOMZ_RK, double *MOMX_RK, double *MOMY_RK, double *RHOT_RK,
    double *mflx_hi, double *tflx_hi,
    const double *DENS0, const double *MOMZ0, const double *MOMX0, const double *MOMY0, const double *RHOT0,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef, double divdmp_coef, const double *DDIV,
    int flag_fct_momentum, int flag_fct_t, int flag_fct_along_stream,
    const orc_metric *m, const double *REF_dens, const double *REF_rhot, double dtrk) {
  GRID_LOCALS;
  if (s != ORC_UD3 && s != ORC_CD4) return -1;
  const double *CDZ = m->CDZ, *FDZ = m->FDZ, *FDX = m->FDX, *FDY = m->FDY;
  const double *RCDZ = m->RCDZ, *RCDX = m->RCDX, *RCDY = m->RCDY;
  const double *RFDZ = m->RFDZ, *RFDX = m->RFDX, *RFDY = m->RFDY;
  const double *PHI = m->PHI;
  const double J33G = m->J33G;
  const double *GS = m->GSQRT, *J13 = m->J13G, *J23 = m->J23G, *MP = m->MAPF;
#define GSQRT(k, i, j, c) GS[I3(k, i, j) + N3 * (c)]
#define J13G(k, i, j, c) J13[I3(k, i, j) + N3 * (c)]
#define J23G(k, i, j, c) J23[I3(k, i, j) + N3 * (c)]
#define MAPF(i, j, n, c) MP[(size_t)(i) + (size_t)IA * ((size_t)(j) + (size_t)JA * ((size_t)(n) + 2 * (size_t)(c)))]
#define ND(k, i, j, v, d) num_diff[I3(k, i, j) + N3 * ((size_t)(v) + 5 * (size_t)(d))]
#define F2HH(k, p, c) (CDZ[(k) + (p)-1] * GSQRT((k) + (p)-1, i, j, c) / (CDZ[k] * GSQRT(k, i, j, c) + CDZ[(k) + 1] * GSQRT((k) + 1, i, j, c)))
  const size_t n3 = N3;
  double *mfZ = mflx_hi, *mfX = mflx_hi + n3, *mfY = mflx_hi + 2 * n3;
  double *tfZ = tflx_hi, *tfX = tflx_hi + n3, *tfY = tflx_hi + 2 * n3;
  double *POTT = alloc_nan(n3), *DPRES = alloc_nan(n3);
  double *qflx_hi = alloc_nan(3 * n3), *qflx_J13 = alloc_nan(n3), *qflx_J23 = alloc_nan(n3);
  double *qZ = qflx_hi, *qX = qflx_hi + n3, *qY = qflx_hi + 2 * n3;
  const int IFS_OFF = 1, JFS_OFF = 1;
  const int IEH = IE, JEH = JE;
  const int IIS = IS, IIE = IE, JJS = JS, JJE = JE;

  /* DPRES and POTT (HEVE.F90:343-366) */
  OMP_FOR
  for (int j = JJS - 2; j <= JJE + 2; ++j)
    for (int i = IIS - 2; i <= IIE + 2; ++i) {
      for (int k = KS; k <= KE; ++k) DPRES[I3(k, i, j)] = DPRES0[I3(k, i, j)] + RT2P[I3(k, i, j)] * (RHOT[I3(k, i, j)] - REF_rhot[I3(k, i, j)]);
      DPRES[I3(KS - 1, i, j)] = DPRES0[I3(KS - 1, i, j)] - DENS[I3(KS, i, j)] * (PHI[I3(KS - 1, i, j)] - PHI[I3(KS + 1, i, j)]);
      DPRES[I3(KE + 1, i, j)] = DPRES0[I3(KE + 1, i, j)] - DENS[I3(KE, i, j)] * (PHI[I3(KE + 1, i, j)] - PHI[I3(KE - 1, i, j)]);
      for (int k = KS; k <= KE; ++k) POTT[I3(k, i, j)] = RHOT[I3(k, i, j)] / DENS[I3(k, i, j)];
    }

  /* momentum FCT (HEVE.F90:386-450): the start-of-stage velocities with
   * their periodic halos, the limiter's advected quantity */
  double *VELZ = NULL, *VELX = NULL, *VELY = NULL, *mlo = NULL, *manti = NULL, *D0u = NULL, *D1u = NULL;
  if (flag_fct_momentum) {
    VELZ = alloc_nan(n3); VELX = alloc_nan(n3); VELY = alloc_nan(n3);
    mlo = alloc_nan(3 * n3); manti = alloc_nan(3 * n3); D0u = alloc_nan(n3); D1u = alloc_nan(n3);
    OMP_FOR
    for (int j = JS - 1; j <= JE + 2; ++j)
      for (int i = IS - 1; i <= IE + 2; ++i) {
        for (int k = KS; k <= KE - 1; ++k)
          VELZ[I3(k, i, j)] = 2.0 * MOMZ0[I3(k, i, j)] / (DENS0[I3(k + 1, i, j)] + DENS0[I3(k, i, j)]);
        VELZ[I3(KE, i, j)] = 0.0;
      }
    OMP_FOR
    for (int j = JS - 1; j <= JE + 2; ++j)
      for (int i = IS - 2; i <= IE + 1; ++i)
        for (int k = KS; k <= KE; ++k)
          VELX[I3(k, i, j)] = 2.0 * MOMX0[I3(k, i, j)] / (DENS0[I3(k, i + 1, j)] + DENS0[I3(k, i, j)]);
    OMP_FOR
    for (int j = JS - 2; j <= JE + 1; ++j)
      for (int i = IS - 1; i <= IE + 2; ++i)
        for (int k = KS; k <= KE; ++k)
          VELY[I3(k, i, j)] = 2.0 * MOMY0[I3(k, i, j)] / (DENS0[I3(k, i, j + 1)] + DENS0[I3(k, i, j)]);
    orc_halo_periodic(g, VELZ); orc_halo_periodic(g, VELX); orc_halo_periodic(g, VELY);
  }
  /* densities averaged to the momentum's position (di, dj, dk: the
   * neighbour), from DENS0 and the halo-exchanged DENS_RK (HEVE.F90:757-773,
   * 1000-1010, 1250-1260); at KE the vertical average repeats KE-1 */
#define MOM_UVW_DENS(di, dj, dk)                                                                          \
  do {                                                                                                    \
    for (int j = JS - 1; j <= JE + 1; ++j)                                                                \
      for (int i = IS - 1; i <= IE + 1; ++i) {                                                            \
        for (int k = KS; k <= KE - (dk); ++k) {                                                           \
          D0u[I3(k, i, j)] = 0.5 * (DENS0[I3(k, i, j)] + DENS0[I3(k + (dk), i + (di), j + (dj))]);       \
          D1u[I3(k, i, j)] = 0.5 * (DENS_RK[I3(k, i, j)] + DENS_RK[I3(k + (dk), i + (di), j + (dj))]);   \
        }                                                                                                 \
        if (dk) { D0u[I3(KE, i, j)] = D0u[I3(KE - 1, i, j)]; D1u[I3(KE, i, j)] = D1u[I3(KE - 1, i, j)]; } \
      }                                                                                                   \
  } while (0)
  /* the high-order z flux as ATMOS_DYN_fct takes it: over the map factors,
   * plus the J13/J23 fluxes (HEVE.F90:752-756, 990-996, 1241-1247) */
#define MOM_HI_Z(c)                                                                                     \
  do {                                                                                                  \
    for (int j = JS; j <= JE; ++j)                                                                      \
      for (int i = IS; i <= IE; ++i)                                                                    \
        for (int k = KS; k <= KE; ++k)                                                                  \
          qZ[I3(k, i, j)] = qZ[I3(k, i, j)] / (MAPF(i, j, 0, c) * MAPF(i, j, 1, c))                     \
                          + qflx_J13[I3(k, i, j)] + qflx_J23[I3(k, i, j)];                              \
  } while (0)

  /* mass fluxes (HEVE.F90:452-520) */
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      for (int k = KS; k <= KE - 1; ++k)
        mfZ[I3(k, i, j)] = J33G * MOMZ[I3(k, i, j)] / (MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY))
                         + J13G(k, i, j, I_XYW) * 0.25 * (MOMX[I3(k + 1, i, j)] + MOMX[I3(k + 1, i - 1, j)]
                                                        + MOMX[I3(k, i, j)] + MOMX[I3(k, i - 1, j)])
                           / MAPF(i, j, 1, I_XY)
                         + J23G(k, i, j, I_XYW) * 0.25 * (MOMY[I3(k + 1, i, j)] + MOMY[I3(k + 1, i, j - 1)]
                                                        + MOMY[I3(k, i, j)] + MOMY[I3(k, i, j - 1)])
                           / MAPF(i, j, 0, I_XY)
                         + GSQRT(k, i, j, I_XYW) * ND(k, i, j, I_DENS, ZDIR) / (MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY));
      mfZ[I3(KS - 1, i, j)] = 0.0;
      mfZ[I3(KE, i, j)] = 0.0;
    }
  {
    int iss = IIS - IFS_OFF, iee = IIE < IEH ? IIE : IEH;
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = iss; i <= iee; ++i)
        for (int k = KS; k <= KE; ++k)
          mfX[I3(k, i, j)] = GSQRT(k, i, j, I_UYZ) / MAPF(i, j, 1, I_UY) * (MOMX[I3(k, i, j)] + ND(k, i, j, I_DENS, XDIR));
  }
  {
    int jss = JJS - JFS_OFF, jee = JJE < JEH ? JJE : JEH;
    OMP_FOR
    for (int j = jss; j <= jee; ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE; ++k)
          mfY[I3(k, i, j)] = GSQRT(k, i, j, I_XVZ) / MAPF(i, j, 0, I_XV) * (MOMY[I3(k, i, j)] + ND(k, i, j, I_DENS, YDIR));
  }
  /* DENS_RK (HEVE.F90:547-572) */
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double advcv = -(mfZ[I3(k, i, j)] - mfZ[I3(k - 1, i, j)]) * RCDZ[k];
        double advch = -(mfX[I3(k, i, j)] - mfX[I3(k, i - 1, j)]) * RCDX[i]
                       - (mfY[I3(k, i, j)] - mfY[I3(k, i, j - 1)]) * RCDY[j];
        DENS_RK[I3(k, i, j)] = DENS0[I3(k, i, j)]
                             + dtrk * ((advcv + advch) * MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY) / GSQRT(k, i, j, I_XYZ)
                                       + DENS_t[I3(k, i, j)]);
      }

  if (flag_fct_momentum) orc_halo_periodic(g, DENS_RK); /* HEVE.F90:746-747 */

  /* MOMZ (HEVE.F90:588-700) */
  fluxZ_XYW(g, s, qZ, MOMZ, MOMZ, DENS, GS + n3 * I_XYZ, J33G, num_diff + n3 * (I_MOMZ + 5 * ZDIR), IIS, IIE, JJS, JJE);
  fluxJ_XYW(g, s, 0, qflx_J13, MOMX, MOMZ, DENS, J13 + n3 * I_XYZ, MP + (size_t)IA * JA * 2 * I_XY, IIS, IIE, JJS, JJE);
  fluxJ_XYW(g, s, 1, qflx_J23, MOMY, MOMZ, DENS, J23 + n3 * I_XYZ, MP + (size_t)IA * JA * 2 * I_XY, IIS, IIE, JJS, JJE);
  fluxH_XYW(g, s, 0, qX, MOMX, MOMZ, DENS, GS + n3 * I_UYW, MP + (size_t)IA * JA * 2 * I_UY, num_diff + n3 * (I_MOMZ + 5 * XDIR), CDZ, IIS, IIE, JJS, JJE);
  fluxH_XYW(g, s, 1, qY, MOMY, MOMZ, DENS, GS + n3 * I_XVW, MP + (size_t)IA * JA * 2 * I_XV, num_diff + n3 * (I_MOMZ + 5 * YDIR), CDZ, IIS, IIE, JJS, JJE);
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      for (int k = KS; k <= KE - 1; ++k) {
        const double pgf = J33G * (DPRES[I3(k + 1, i, j)] - DPRES[I3(k, i, j)]) * RFDZ[k];
        const double buoy = GRAV * GSQRT(k, i, j, I_XYW)
                          * (F2HH(k, 1, I_XYZ) * (DENS[I3(k + 1, i, j)] - REF_dens[I3(k + 1, i, j)])
                           + F2HH(k, 2, I_XYZ) * (DENS[I3(k, i, j)] - REF_dens[I3(k, i, j)]));
        double advcv = -(qZ[I3(k, i, j)] - qZ[I3(k - 1, i, j)]
                       + qflx_J13[I3(k, i, j)] - qflx_J13[I3(k - 1, i, j)]
                       + qflx_J23[I3(k, i, j)] - qflx_J23[I3(k - 1, i, j)]) * RFDZ[k];
        double advch = -((qX[I3(k, i, j)] - qX[I3(k, i - 1, j)]) * RCDX[i]
                       + (qY[I3(k, i, j)] - qY[I3(k, i, j - 1)]) * RCDY[j])
                     * MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY);
        double wdamp = -wdamp_coef[k] * MOMZ0[I3(k, i, j)];
        double div = divdmp_coef / dtrk * FDZ[k] * (DDIV[I3(k + 1, i, j)] - DDIV[I3(k, i, j)]);
        MOMZ_RK[I3(k, i, j)] = MOMZ0[I3(k, i, j)]
                             + dtrk * ((advcv + advch - pgf - buoy) / GSQRT(k, i, j, I_XYW) + wdamp + div + MOMZ_t[I3(k, i, j)]);
      }
      MOMZ_RK[I3(KS - 1, i, j)] = 0.0;
      MOMZ_RK[I3(KE, i, j)] = 0.0;
    }

  /* FCT of the MOMZ fluxes (HEVE.F90:716-810) */
  if (flag_fct_momentum) {
    fluxZ_XYW_ud1(g, mlo, MOMZ, MOMZ0, DENS, J33G, IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    fluxH_XYW_ud1(g, 0, mlo + n3, MOMX, MOMZ0, DENS, GS + n3 * I_UYZ, MP + (size_t)IA * JA * 2 * I_UY, CDZ,
                  IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    fluxH_XYW_ud1(g, 1, mlo + 2 * n3, MOMY, MOMZ0, DENS, GS + n3 * I_XVZ, MP + (size_t)IA * JA * 2 * I_XV, CDZ,
                  IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    MOM_HI_Z(I_XY);
    MOM_UVW_DENS(0, 0, 1);
    orc_fct(g, manti, VELZ, D0u, D1u, qflx_hi, mlo, mflx_hi, RFDZ, RCDX, RCDY, GS + n3 * I_XYW,
            MP + (size_t)IA * JA * 2 * I_XY, dtrk, flag_fct_along_stream);
    const double *aZ = manti, *aX = manti + n3, *aY = manti + 2 * n3;
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE - 1; ++k)
          MOMZ_RK[I3(k, i, j)] = MOMZ_RK[I3(k, i, j)]
                               + dtrk * ((aZ[I3(k, i, j)] - aZ[I3(k - 1, i, j)]) * RFDZ[k]
                                         + ((aX[I3(k, i, j)] - aX[I3(k, i - 1, j)]) * RCDX[i]
                                            + (aY[I3(k, i, j)] - aY[I3(k, i, j - 1)]) * RCDY[j])
                                           * MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY))
                                 / GSQRT(k, i, j, I_XYW);
  }

  /* MOMX (HEVE.F90:744-945) */
  fluxZ_UVZ(g, s, 0, qZ, MOMZ, MOMX, DENS, GS + n3 * I_UYW, J33G, num_diff + n3 * (I_MOMX + 5 * ZDIR), CDZ, IIS, IIE, JJS, JJE);
  fluxJ_UVZ(g, s, 0, 0, qflx_J13, MOMX, MOMX, DENS, GS + n3 * I_UYZ, J13 + n3 * I_UYW, MP + (size_t)IA * JA * 2 * I_UY, CDZ, IIS, IIE, JJS, JJE);
  fluxJ_UVZ(g, s, 0, 1, qflx_J23, MOMY, MOMX, DENS, GS + n3 * I_UYZ, J23 + n3 * I_UYW, MP + (size_t)IA * JA * 2 * I_UY, CDZ, IIS, IIE, JJS, JJE);
  fluxX_UYZ(g, s, qX, MOMX, MOMX, DENS, GS + n3 * I_XYZ, MP + (size_t)IA * JA * 2 * I_XY, num_diff + n3 * (I_MOMX + 5 * XDIR), IIS, IIE, JJS, JJE);
  fluxY_UYZ(g, s, qY, MOMY, MOMX, DENS, GS + n3 * I_UVZ, MP + (size_t)IA * JA * 2 * I_UV, num_diff + n3 * (I_MOMX + 5 * YDIR), IIS, IIE, JJS, JJE);
  {
    int iee = IIE < IEH ? IIE : IEH;
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= iee; ++i)
        for (int k = KS; k <= KE; ++k) {
          double advcv = -(qZ[I3(k, i, j)] - qZ[I3(k - 1, i, j)]
                         + qflx_J13[I3(k, i, j)] - qflx_J13[I3(k - 1, i, j)]
                         + qflx_J23[I3(k, i, j)] - qflx_J23[I3(k - 1, i, j)]) * RCDZ[k];
          double advch = -((qX[I3(k, i, j)] - qX[I3(k, i - 1, j)]) * RFDX[i]
                         + (qY[I3(k, i, j)] - qY[I3(k, i, j - 1)]) * RCDY[j])
                       * MAPF(i, j, 0, I_UY) * MAPF(i, j, 1, I_UY);
          double pgf = ((GSQRT(k, i + 1, j, I_XYZ) * DPRES[I3(k, i + 1, j)]
                       - GSQRT(k, i, j, I_XYZ) * DPRES[I3(k, i, j)]) * RFDX[i]
                      + (J13G(k, i, j, I_UYW)
                         * 0.5 * (F2HH(k, 1, I_UYZ) * (DPRES[I3(k + 1, i + 1, j)] + DPRES[I3(k + 1, i, j)])
                                + F2HH(k, 2, I_UYZ) * (DPRES[I3(k, i + 1, j)] + DPRES[I3(k, i, j)]))
                       - J13G(k - 1, i, j, I_UYW)
                         * 0.5 * (F2HH(k, 1, I_UYZ) * (DPRES[I3(k, i + 1, j)] + DPRES[I3(k, i, j)])
                                + F2HH(k, 2, I_UYZ) * (DPRES[I3(k - 1, i + 1, j)] + DPRES[I3(k - 1, i, j)])))
                        * RCDZ[k])
                     * MAPF(i, j, 0, I_UY);
          double cor = 0.125 * (CORIOLI[(size_t)(i + 1) + (size_t)IA * j] + CORIOLI[(size_t)i + (size_t)IA * j])
                         * (MOMY[I3(k, i + 1, j)] + MOMY[I3(k, i, j)] + MOMY[I3(k, i + 1, j - 1)] + MOMY[I3(k, i, j - 1)])
                     + 0.25 * MAPF(i, j, 0, I_UY) * MAPF(i, j, 1, I_UY)
                         * (MOMY[I3(k, i, j)] + MOMY[I3(k, i, j - 1)] + MOMY[I3(k, i + 1, j)] + MOMY[I3(k, i + 1, j - 1)])
                         * ((MOMY[I3(k, i, j)] + MOMY[I3(k, i, j - 1)] + MOMY[I3(k, i + 1, j)] + MOMY[I3(k, i + 1, j - 1)]) * 0.25
                              * (1.0 / MAPF(i + 1, j, 1, I_XY) - 1.0 / MAPF(i, j, 1, I_XY)) * RFDX[i]
                            - MOMX[I3(k, i, j)]
                              * (1.0 / MAPF(i, j, 0, I_UV) - 1.0 / MAPF(i, j - 1, 0, I_UV)) * RCDY[j])
                         * 2.0 / (DENS[I3(k, i + 1, j)] + DENS[I3(k, i, j)]);
          double div = divdmp_coef / dtrk * FDX[i] * (DDIV[I3(k, i + 1, j)] - DDIV[I3(k, i, j)]);
          MOMX_RK[I3(k, i, j)] = MOMX0[I3(k, i, j)]
                               + dtrk * ((advcv + advch - pgf) / GSQRT(k, i, j, I_UYZ) + cor + div + MOMX_t[I3(k, i, j)]);
        }
  }

  /* FCT of the MOMX fluxes (HEVE.F90:958-1057) */
  if (flag_fct_momentum) {
    fluxZ_UVZ_ud1(g, 0, mlo, MOMZ, MOMX0, DENS, GS + n3 * I_UYW, J33G, CDZ, IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    fluxS_ud1(g, 0, mlo + n3, MOMX, MOMX0, DENS, GS + n3 * I_XYZ, MP + (size_t)IA * JA * 2 * I_UY,
              IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    fluxC_ud1(g, 0, mlo + 2 * n3, MOMY, MOMX0, DENS, GS + n3 * I_UVZ, MP + (size_t)IA * JA * 2 * I_XV,
              IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    MOM_HI_Z(I_UY);
    MOM_UVW_DENS(1, 0, 0);
    orc_fct(g, manti, VELX, D0u, D1u, qflx_hi, mlo, mflx_hi, RCDZ, RFDX, RCDY, GS + n3 * I_UYZ,
            MP + (size_t)IA * JA * 2 * I_UY, dtrk, flag_fct_along_stream);
    const double *aZ = manti, *aX = manti + n3, *aY = manti + 2 * n3;
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= (IIE < IEH ? IIE : IEH); ++i)
        for (int k = KS; k <= KE; ++k)
          MOMX_RK[I3(k, i, j)] = MOMX_RK[I3(k, i, j)]
                               + dtrk * (((aZ[I3(k, i, j)] - aZ[I3(k - 1, i, j)]) * RCDZ[k]
                                          + (aX[I3(k, i, j)] - aX[I3(k, i - 1, j)]) * RFDX[i]
                                          + (aY[I3(k, i, j)] - aY[I3(k, i, j - 1)]) * RCDY[j]))
                                 * MAPF(i, j, 0, I_UY) * MAPF(i, j, 1, I_UY) / GSQRT(k, i, j, I_UYZ);
  }

  /* MOMY (HEVE.F90:993-1193) */
  fluxZ_UVZ(g, s, 1, qZ, MOMZ, MOMY, DENS, GS + n3 * I_XVW, J33G, num_diff + n3 * (I_MOMY + 5 * ZDIR), CDZ, IIS, IIE, JJS, JJE);
  fluxJ_UVZ(g, s, 1, 0, qflx_J13, MOMX, MOMY, DENS, GS + n3 * I_XVZ, J13 + n3 * I_XVW, MP + (size_t)IA * JA * 2 * I_XV, CDZ, IIS, IIE, JJS, JJE);
  fluxJ_UVZ(g, s, 1, 1, qflx_J23, MOMY, MOMY, DENS, GS + n3 * I_XVZ, J23 + n3 * I_XVW, MP + (size_t)IA * JA * 2 * I_XV, CDZ, IIS, IIE, JJS, JJE);
  fluxX_XVZ(g, s, qX, MOMX, MOMY, DENS, GS + n3 * I_UVZ, MP + (size_t)IA * JA * 2 * I_UV, num_diff + n3 * (I_MOMY + 5 * XDIR), IIS, IIE, JJS, JJE);
  fluxY_XVZ(g, s, qY, MOMY, MOMY, DENS, GS + n3 * I_XYZ, MP + (size_t)IA * JA * 2 * I_XY, num_diff + n3 * (I_MOMY + 5 * YDIR), IIS, IIE, JJS, JJE);
  {
    int jee = JJE < JEH ? JJE : JEH;
    OMP_FOR
    for (int j = JJS; j <= jee; ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE; ++k) {
          double advcv = -(qZ[I3(k, i, j)] - qZ[I3(k - 1, i, j)]
                         + qflx_J13[I3(k, i, j)] - qflx_J13[I3(k - 1, i, j)]
                         + qflx_J23[I3(k, i, j)] - qflx_J23[I3(k - 1, i, j)]) * RCDZ[k];
          double advch = -((qX[I3(k, i, j)] - qX[I3(k, i - 1, j)]) * RCDX[i]
                         + (qY[I3(k, i, j)] - qY[I3(k, i, j - 1)]) * RFDY[j])
                       * MAPF(i, j, 0, I_XV) * MAPF(i, j, 1, I_XV);
          double pgf = ((GSQRT(k, i, j + 1, I_XYZ) * DPRES[I3(k, i, j + 1)]
                       - GSQRT(k, i, j, I_XYZ) * DPRES[I3(k, i, j)]) * RFDY[j]
                      + (J23G(k, i, j, I_XVW)
                         * 0.5 * (F2HH(k, 1, I_XVZ) * (DPRES[I3(k + 1, i, j + 1)] + DPRES[I3(k + 1, i, j)])
                                + F2HH(k, 2, I_XVZ) * (DPRES[I3(k, i, j + 1)] + DPRES[I3(k, i, j)]))
                       - J23G(k - 1, i, j, I_XVW)
                         * 0.5 * (F2HH(k - 1, 1, I_XVZ) * (DPRES[I3(k, i, j + 1)] + DPRES[I3(k, i, j)])
                                + F2HH(k - 1, 2, I_XVZ) * (DPRES[I3(k - 1, i, j + 1)] + DPRES[I3(k - 1, i, j)])))
                        * RCDZ[k])
                     * MAPF(i, j, 1, I_XV);
          double cor = -0.125 * (CORIOLI[(size_t)i + (size_t)IA * (j + 1)] + CORIOLI[(size_t)i + (size_t)IA * j])
                         * (MOMX[I3(k, i, j + 1)] + MOMX[I3(k, i, j)] + MOMX[I3(k, i - 1, j + 1)] + MOMX[I3(k, i - 1, j)])
                     - 0.25 * MAPF(i, j, 0, I_XV) * MAPF(i, j, 1, I_XV)
                         * (MOMX[I3(k, i, j)] + MOMX[I3(k, i - 1, j)] + MOMX[I3(k, i, j + 1)] + MOMX[I3(k, i - 1, j + 1)])
                         * (MOMY[I3(k, i, j)]
                              * (1.0 / MAPF(i, j, 1, I_UV) - 1.0 / MAPF(i - 1, j, 1, I_UV)) * RCDX[i]
                            - 0.25 * (MOMX[I3(k, i, j)] + MOMX[I3(k, i - 1, j)] + MOMX[I3(k, i, j + 1)] + MOMX[I3(k, i - 1, j + 1)])
                              * (1.0 / MAPF(i, j + 1, 0, I_XY) - 1.0 / MAPF(i, j, 0, I_XY)) * RFDY[j])
                         * 2.0 / (DENS[I3(k, i, j)] + DENS[I3(k, i, j + 1)]);
          double div = divdmp_coef / dtrk * FDY[j] * (DDIV[I3(k, i, j + 1)] - DDIV[I3(k, i, j)]);
          MOMY_RK[I3(k, i, j)] = MOMY0[I3(k, i, j)]
                               + dtrk * ((advcv + advch - pgf) / GSQRT(k, i, j, I_XVZ) + cor + div + MOMY_t[I3(k, i, j)]);
        }
  }

  /* FCT of the MOMY fluxes (HEVE.F90:1207-1309) */
  if (flag_fct_momentum) {
    fluxZ_UVZ_ud1(g, 1, mlo, MOMZ, MOMY0, DENS, GS + n3 * I_XVZ, J33G, CDZ, IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    fluxC_ud1(g, 1, mlo + n3, MOMX, MOMY0, DENS, GS + n3 * I_UVZ, MP + (size_t)IA * JA * 2 * I_XY,
              IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    fluxS_ud1(g, 1, mlo + 2 * n3, MOMY, MOMY0, DENS, GS + n3 * I_XYZ, MP + (size_t)IA * JA * 2 * I_XY,
              IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    MOM_HI_Z(I_XV);
    MOM_UVW_DENS(0, 1, 0);
    orc_fct(g, manti, VELY, D0u, D1u, qflx_hi, mlo, mflx_hi, RCDZ, RCDX, RFDY, GS + n3 * I_XVZ,
            MP + (size_t)IA * JA * 2 * I_XV, dtrk, flag_fct_along_stream);
    const double *aZ = manti, *aX = manti + n3, *aY = manti + 2 * n3;
    OMP_FOR
    for (int j = JJS; j <= (JJE < JEH ? JJE : JEH); ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE; ++k)
          MOMY_RK[I3(k, i, j)] = MOMY_RK[I3(k, i, j)]
                               + dtrk * (((aZ[I3(k, i, j)] - aZ[I3(k - 1, i, j)]) * RCDZ[k]
                                          + (aX[I3(k, i, j)] - aX[I3(k, i - 1, j)]) * RCDX[i]
                                          + (aY[I3(k, i, j)] - aY[I3(k, i, j - 1)]) * RFDY[j]))
                                 * MAPF(i, j, 0, I_XV) * MAPF(i, j, 1, I_XV) / GSQRT(k, i, j, I_XVZ);
    free(VELZ); free(VELX); free(VELY); free(mlo); free(manti); free(D0u); free(D1u);
  }
#undef MOM_UVW_DENS
#undef MOM_HI_Z

  /* RHOT (HEVE.F90:1316-1375) */
  fluxZ_XYZ(g, s, tfZ, mfZ, POTT, GS + n3 * I_XYW, num_diff + n3 * (I_RHOT + 5 * ZDIR), IIS, IIE, JJS, JJE);
  fluxX_XYZ(g, s, tfX, mfX, POTT, GS + n3 * I_UYZ, num_diff + n3 * (I_RHOT + 5 * XDIR), IIS, IIE, JJS, JJE);
  fluxY_XYZ(g, s, tfY, mfY, POTT, GS + n3 * I_XVZ, num_diff + n3 * (I_RHOT + 5 * YDIR), IIS, IIE, JJS, JJE);
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double advcv = -(tfZ[I3(k, i, j)] - tfZ[I3(k - 1, i, j)]) * RCDZ[k];
        double advch = -(tfX[I3(k, i, j)] - tfX[I3(k, i - 1, j)]) * RCDX[i]
                       - (tfY[I3(k, i, j)] - tfY[I3(k, i, j - 1)]) * RCDY[j];
        RHOT_RK[I3(k, i, j)] = RHOT0[I3(k, i, j)]
                             + dtrk * ((advcv + advch) * MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY) / GSQRT(k, i, j, I_XYZ)
                                       + RHOT_t[I3(k, i, j)]);
      }

  /* FCT of the theta flux (HEVE.F90:1388-1475): halo exchange of mflx_hi and
   * DENS_RK, low-order UD1 fluxes of the start-of-stage theta, ATMOS_DYN_fct
   * and the anti-diffusive correction of RHOT_RK.  As in the reference, the
   * start-of-stage theta overwrites POTT on IS-1..IE+2, JS-1..JE+2 only
   * (HEVE.F90:1403-1409), so the i = IS-2 and j = JS-2 cells the UD1 fluxes
   * at IS-2 / JS-2 read keep the theta of the current state. */
  if (flag_fct_t) {
    for (int c = 0; c < 3; ++c) orc_halo_periodic(g, mflx_hi + n3 * c);
    orc_halo_periodic(g, DENS_RK);
    double *POTT0 = POTT, *tflx_lo = alloc_nan(3 * n3), *tflx_anti = alloc_nan(3 * n3);
    OMP_FOR
    for (int j = JS - 1; j <= JE + 2; ++j)
      for (int i = IS - 1; i <= IE + 2; ++i)
        for (int k = KS; k <= KE; ++k) POTT0[I3(k, i, j)] = RHOT0[I3(k, i, j)] / DENS0[I3(k, i, j)];
    const double *ndR = num_diff + n3 * (I_RHOT + 5 * ZDIR);
    fluxZ_XYZ(g, ORC_UD1, tflx_lo, mfZ, POTT0, GS + n3 * I_XYZ, ndR, IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    fluxX_XYZ(g, ORC_UD1, tflx_lo + n3, mfX, POTT0, GS + n3 * I_UYZ, num_diff + n3 * (I_RHOT + 5 * XDIR),
              IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    fluxY_XYZ(g, ORC_UD1, tflx_lo + 2 * n3, mfY, POTT0, GS + n3 * I_XVZ, num_diff + n3 * (I_RHOT + 5 * YDIR),
              IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    orc_fct(g, tflx_anti, POTT0, DENS0, DENS_RK, tflx_hi, tflx_lo, mflx_hi, RCDZ, RCDX, RCDY, GS + n3 * I_XYZ,
            MP + (size_t)IA * JA * 2 * I_XY, dtrk, flag_fct_along_stream);
    const double *aZ = tflx_anti, *aX = tflx_anti + n3, *aY = tflx_anti + 2 * n3;
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE; ++k)
          RHOT_RK[I3(k, i, j)] = RHOT_RK[I3(k, i, j)]
                               + dtrk * ((aZ[I3(k, i, j)] - aZ[I3(k - 1, i, j)]) * RCDZ[k]
                                         + (aX[I3(k, i, j)] - aX[I3(k, i - 1, j)]) * RCDX[i]
                                         + (aY[I3(k, i, j)] - aY[I3(k, i, j - 1)]) * RCDY[j])
                                 * MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY) / GSQRT(k, i, j, I_XYZ);
    free(tflx_lo); free(tflx_anti);
  }
  free(POTT); free(DPRES); free(qflx_hi); free(qflx_J13); free(qflx_J23);
  return 0;
#undef GSQRT
#undef J13G
#undef J23G
#undef MAPF
#undef ND
#undef F2HH
}

/* ---- periodic halo fill (COMM_vars8 on one periodic rank) ---------------- */
void orc_halo_periodic(const orc_grid *g, double *a) {
  GRID_LOCALS;
  const int IMAX = IE - IS + 1, JMAX = JE - JS + 1;
  /* west/east faces for interior j, then south/north rows for all i (incl. corners) */
  OMP_FOR
  for (int j = JS; j <= JE; ++j) {
    for (int i = 0; i < IS; ++i)
      for (int k = 0; k < KA; ++k) a[I3(k, i, j)] = a[I3(k, i + IMAX, j)];
    for (int i = IE + 1; i < IA; ++i)
      for (int k = 0; k < KA; ++k) a[I3(k, i, j)] = a[I3(k, i - IMAX, j)];
  }
  OMP_FOR
  for (int j = 0; j < JS; ++j)
    for (int i = 0; i < IA; ++i)
      for (int k = 0; k < KA; ++k) a[I3(k, i, j)] = a[I3(k, i, j + JMAX)];
  for (int j = JE + 1; j < JA; ++j)
    for (int i = 0; i < IA; ++i)
      for (int k = 0; k < KA; ++k) a[I3(k, i, j)] = a[I3(k, i, j - JMAX)];
}

/* ---- numerical filter (COMMON.F90:385-985, calc_diff3 :1652-1800) ------- */
/* CNZ3/CNX3/CNY3 tables of ATMOS_DYN_filter_setup (COMMON.F90:92-319) */
typedef struct {
  double *CNZ3; /* (3,KA,2): [a + 3*(k + KA*b)] */
  double *CNX3; /* (3,IA,2) */
  double *CNY3; /* (3,JA,2) */
} cn_tables;

static void filter_setup(const orc_grid *g, const orc_metric *m, cn_tables *t) {
  GRID_LOCALS;
  const double *CDZ = m->CDZ, *FDZ = m->FDZ, *CDX = m->CDX, *FDX = m->FDX, *CDY = m->CDY, *FDY = m->FDY;
  t->CNZ3 = alloc_nan((size_t)3 * KA * 2);
  t->CNX3 = alloc_nan((size_t)3 * IA * 2);
  t->CNY3 = alloc_nan((size_t)3 * JA * 2);
#define CNZ3(a, k, b) t->CNZ3[(a)-1 + 3 * ((size_t)(k) + (size_t)KA * ((b)-1))]
#define CNX3(a, i, b) t->CNX3[(a)-1 + 3 * ((size_t)(i) + (size_t)IA * ((b)-1))]
#define CNY3(a, j, b) t->CNY3[(a)-1 + 3 * ((size_t)(j) + (size_t)JA * ((b)-1))]
  for (int k = KS + 1; k <= KE - 1; ++k) {
    CNZ3(1, k, 1) = 1.0 / (FDZ[k] * CDZ[k] * FDZ[k - 1]);
    CNZ3(2, k, 1) = 1.0 / (FDZ[k] * CDZ[k] * FDZ[k - 1]) + 1.0 / (FDZ[k - 1] * CDZ[k] * FDZ[k - 1]) + 1.0 / (FDZ[k - 1] * CDZ[k - 1] * FDZ[k - 1]);
  }
  for (int k = KS + 2; k <= KE; ++k)
    CNZ3(3, k, 1) = 1.0 / (FDZ[k - 1] * CDZ[k] * FDZ[k - 1]) + 1.0 / (FDZ[k - 1] * CDZ[k - 1] * FDZ[k - 1]) + 1.0 / (FDZ[k - 1] * CDZ[k - 1] * FDZ[k - 2]);
  CNZ3(1, KS - 1, 1) = 1.0 / (FDZ[KS] * CDZ[KS + 1] * FDZ[KS]);
  CNZ3(1, KS, 1) = 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]);
  CNZ3(2, KS, 1) = 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS + 1] * FDZ[KS]);
  CNZ3(3, KS, 1) = 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS + 1] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS + 1] * FDZ[KS + 1]);
  CNZ3(3, KS + 1, 1) = 1.0 / (FDZ[KS] * CDZ[KS + 1] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]) + 1.0 / (FDZ[KS] * CDZ[KS] * FDZ[KS]);
  CNZ3(1, KE, 1) = 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]);
  CNZ3(2, KE, 1) = 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE - 1] * FDZ[KE - 1]);
  CNZ3(1, KE + 1, 1) = 1.0 / (FDZ[KE - 2] * CDZ[KE - 1] * FDZ[KE - 1]);
  CNZ3(2, KE + 1, 1) = 1.0 / (FDZ[KE - 2] * CDZ[KE - 1] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE - 1] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]);
  CNZ3(3, KE + 1, 1) = 1.0 / (FDZ[KE - 1] * CDZ[KE + 1] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]) + 1.0 / (FDZ[KE - 1] * CDZ[KE] * FDZ[KE - 1]);
  for (int k = KS + 1; k <= KE - 1; ++k) {
    CNZ3(1, k, 2) = 1.0 / (CDZ[k + 1] * FDZ[k] * CDZ[k]);
    CNZ3(2, k, 2) = 1.0 / (CDZ[k + 1] * FDZ[k] * CDZ[k]) + 1.0 / (CDZ[k] * FDZ[k] * CDZ[k]) + 1.0 / (CDZ[k] * FDZ[k - 1] * CDZ[k]);
    CNZ3(3, k, 2) = 1.0 / (CDZ[k] * FDZ[k] * CDZ[k]) + 1.0 / (CDZ[k] * FDZ[k - 1] * CDZ[k]) + 1.0 / (CDZ[k] * FDZ[k - 1] * CDZ[k - 1]);
  }
  CNZ3(1, KS - 1, 2) = 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS]);
  CNZ3(1, KS, 2) = 1.0 / (CDZ[KS + 1] * FDZ[KS] * CDZ[KS]);
  CNZ3(2, KS, 2) = 1.0 / (CDZ[KS + 1] * FDZ[KS] * CDZ[KS]) + 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS]) + 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS]);
  CNZ3(3, KS, 2) = 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS]) + 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS]) + 1.0 / (CDZ[KS] * FDZ[KS] * CDZ[KS + 1]);
  CNZ3(1, KE, 2) = 1.0 / (CDZ[KE - 1] * FDZ[KE - 1] * CDZ[KE]);
  CNZ3(2, KE, 2) = 1.0 / (CDZ[KE - 1] * FDZ[KE - 1] * CDZ[KE]) + 1.0 / (CDZ[KE] * FDZ[KE - 1] * CDZ[KE]) + 1.0 / (CDZ[KE] * FDZ[KE - 1] * CDZ[KE]);
  CNZ3(3, KE, 2) = 1.0 / (CDZ[KE] * FDZ[KE - 1] * CDZ[KE]) + 1.0 / (CDZ[KE] * FDZ[KE - 1] * CDZ[KE]) + 1.0 / (CDZ[KE] * FDZ[KE - 1] * CDZ[KE - 1]);
  CNZ3(1, KE + 1, 2) = 1.0 / (CDZ[KE - 2] * FDZ[KE - 2] * CDZ[KE - 1]);
  CNZ3(2, KE + 1, 2) = 1.0 / (CDZ[KE - 2] * FDZ[KE - 2] * CDZ[KE - 1]) + 1.0 / (CDZ[KE - 1] * FDZ[KE - 2] * CDZ[KE - 1]) + 1.0 / (CDZ[KE - 1] * FDZ[KE - 1] * CDZ[KE - 1]);
  CNZ3(3, KE + 1, 2) = 1.0 / (CDZ[KE - 1] * FDZ[KE - 2] * CDZ[KE - 1]) + 1.0 / (CDZ[KE - 1] * FDZ[KE - 1] * CDZ[KE - 1]) + 1.0 / (CDZ[KE - 1] * FDZ[KE - 1] * CDZ[KE]);
  /* x */
  CNX3(1, IS - 1, 1) = 1.0 / (FDX[IS - 1] * CDX[IS - 1] * FDX[IS - 2]);
  for (int i = IS; i <= IE + 1; ++i) {
    CNX3(1, i, 1) = 1.0 / (FDX[i] * CDX[i] * FDX[i - 1]);
    CNX3(2, i, 1) = 1.0 / (FDX[i] * CDX[i] * FDX[i - 1]) + 1.0 / (FDX[i - 1] * CDX[i] * FDX[i - 1]) + 1.0 / (FDX[i - 1] * CDX[i - 1] * FDX[i - 1]);
    CNX3(3, i, 1) = 1.0 / (FDX[i - 1] * CDX[i] * FDX[i - 1]) + 1.0 / (FDX[i - 1] * CDX[i - 1] * FDX[i - 1]) + 1.0 / (FDX[i - 1] * CDX[i - 1] * FDX[i - 2]);
  }
  for (int i = IS - 1; i <= IE + 1; ++i) {
    CNX3(1, i, 2) = 1.0 / (CDX[i + 1] * FDX[i] * CDX[i]);
    CNX3(2, i, 2) = 1.0 / (CDX[i + 1] * FDX[i] * CDX[i]) + 1.0 / (CDX[i] * FDX[i] * CDX[i]) + 1.0 / (CDX[i] * FDX[i - 1] * CDX[i]);
    CNX3(3, i, 2) = 1.0 / (CDX[i] * FDX[i] * CDX[i]) + 1.0 / (CDX[i] * FDX[i - 1] * CDX[i]) + 1.0 / (CDX[i] * FDX[i - 1] * CDX[i - 1]);
  }
  /* y */
  CNY3(1, JS - 1, 1) = 1.0 / (FDY[JS - 1] * CDY[JS - 1] * FDY[JS - 2]);
  OMP_FOR
  for (int j = JS; j <= JE + 1; ++j) {
    CNY3(1, j, 1) = 1.0 / (FDY[j] * CDY[j] * FDY[j - 1]);
    CNY3(2, j, 1) = 1.0 / (FDY[j] * CDY[j] * FDY[j - 1]) + 1.0 / (FDY[j - 1] * CDY[j] * FDY[j - 1]) + 1.0 / (FDY[j - 1] * CDY[j - 1] * FDY[j - 1]);
    CNY3(3, j, 1) = 1.0 / (FDY[j - 1] * CDY[j] * FDY[j - 1]) + 1.0 / (FDY[j - 1] * CDY[j - 1] * FDY[j - 1]) + 1.0 / (FDY[j - 1] * CDY[j - 1] * FDY[j - 2]);
  }
  OMP_FOR
  for (int j = JS - 1; j <= JE + 1; ++j) {
    CNY3(1, j, 2) = 1.0 / (CDY[j + 1] * FDY[j] * CDY[j]);
    CNY3(2, j, 2) = 1.0 / (CDY[j + 1] * FDY[j] * CDY[j]) + 1.0 / (CDY[j] * FDY[j] * CDY[j]) + 1.0 / (CDY[j] * FDY[j - 1] * CDY[j]);
    CNY3(3, j, 2) = 1.0 / (CDY[j] * FDY[j] * CDY[j]) + 1.0 / (CDY[j] * FDY[j - 1] * CDY[j]) + 1.0 / (CDY[j] * FDY[j - 1] * CDY[j - 1]);
  }
}

/* calc_diff3 (COMMON.F90:1652-1800): diff(KA,IA,JA,3) over interior columns */
static void calc_diff3(const orc_grid *g, const cn_tables *t, double *diff, const double *phi, int KO, int IO, int JO) {
  GRID_LOCALS;
  double *dZ = diff, *dX = diff + N3, *dY = diff + 2 * N3;
  const int KEE = KE - KO;
  if (KO == 0) {
    OMP_FOR
    for (int j = JS; j <= JE; ++j)
      for (int i = IS; i <= IE; ++i) {
        for (int k = KS + 1; k <= KE - 2; ++k)
          dZ[I3(k, i, j)] = (+CNZ3(1, k + 1, 1) * phi[I3(k + 2, i, j)] - CNZ3(2, k + 1, 1) * phi[I3(k + 1, i, j)]
                             + CNZ3(3, k + 1, 1) * phi[I3(k, i, j)] - CNZ3(1, k, 1) * phi[I3(k - 1, i, j)]);
        dZ[I3(KS, i, j)] = (+CNZ3(1, KS + 1, 1) * phi[I3(KS + 2, i, j)] - CNZ3(2, KS + 1, 1) * phi[I3(KS + 1, i, j)]
                            + CNZ3(3, KS + 1, 1) * phi[I3(KS, i, j)] - CNZ3(1, KS, 1) * phi[I3(KS + 1, i, j)]);
        dZ[I3(KS - 1, i, j)] = -dZ[I3(KS, i, j)];
        dZ[I3(KS - 2, i, j)] = -dZ[I3(KS + 1, i, j)];
        dZ[I3(KE - 1, i, j)] = (+CNZ3(1, KE, 1) * phi[I3(KE - 1, i, j)] - CNZ3(2, KE, 1) * phi[I3(KE, i, j)]
                                + CNZ3(3, KE, 1) * phi[I3(KE - 1, i, j)] - CNZ3(1, KE - 1, 1) * phi[I3(KE - 2, i, j)]);
        dZ[I3(KE, i, j)] = -dZ[I3(KE - 1, i, j)];
        dZ[I3(KE + 1, i, j)] = -dZ[I3(KE - 2, i, j)];
        dZ[I3(KE + 2, i, j)] = 0.0;
      }
  } else {
    OMP_FOR
    for (int j = JS; j <= JE; ++j)
      for (int i = IS; i <= IE; ++i) {
        for (int k = KS + 2; k <= KE - 2; ++k)
          dZ[I3(k, i, j)] = (+CNZ3(1, k, 2) * phi[I3(k + 1, i, j)] - CNZ3(2, k, 2) * phi[I3(k, i, j)]
                             + CNZ3(3, k, 2) * phi[I3(k - 1, i, j)] - CNZ3(1, k - 1, 2) * phi[I3(k - 2, i, j)]);
        dZ[I3(KS + 1, i, j)] = (+CNZ3(1, KS + 1, 2) * phi[I3(KS + 2, i, j)] - CNZ3(2, KS + 1, 2) * phi[I3(KS + 1, i, j)]
                                + CNZ3(3, KS + 1, 2) * phi[I3(KS, i, j)] - CNZ3(1, KS, 2) * phi[I3(KS + 1, i, j)]);
        dZ[I3(KS, i, j)] = -dZ[I3(KS + 1, i, j)];
        dZ[I3(KS - 1, i, j)] = -dZ[I3(KS + 2, i, j)];
        dZ[I3(KS - 2, i, j)] = -dZ[I3(KS + 3, i, j)];
        dZ[I3(KE - 1, i, j)] = (-CNZ3(2, KE - 1, 2) * phi[I3(KE - 1, i, j)] + CNZ3(3, KE - 1, 2) * phi[I3(KE - 2, i, j)]
                                - CNZ3(1, KE - 2, 2) * phi[I3(KE - 3, i, j)]);
        dZ[I3(KE, i, j)] = (+CNZ3(1, KE, 2) * phi[I3(KE - 1, i, j)] + CNZ3(3, KE, 2) * phi[I3(KE - 1, i, j)]
                            - CNZ3(1, KE - 1, 2) * phi[I3(KE - 2, i, j)]);
        dZ[I3(KE + 1, i, j)] = -dZ[I3(KE, i, j)];
        dZ[I3(KE + 2, i, j)] = -dZ[I3(KE - 1, i, j)];
      }
  }
  OMP_FOR
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS; k <= KEE; ++k) {
        if (IO == 0)
          dX[I3(k, i, j)] = (+CNX3(1, i + 1, 1) * phi[I3(k, i + 2, j)] - CNX3(2, i + 1, 1) * phi[I3(k, i + 1, j)]
                             + CNX3(3, i + 1, 1) * phi[I3(k, i, j)] - CNX3(1, i, 1) * phi[I3(k, i - 1, j)]);
        else
          dX[I3(k, i, j)] = (+CNX3(1, i, 2) * phi[I3(k, i + 1, j)] - CNX3(2, i, 2) * phi[I3(k, i, j)]
                             + CNX3(3, i, 2) * phi[I3(k, i - 1, j)] - CNX3(1, i - 1, 2) * phi[I3(k, i - 2, j)]);
        if (JO == 0)
          dY[I3(k, i, j)] = (+CNY3(1, j + 1, 1) * phi[I3(k, i, j + 2)] - CNY3(2, j + 1, 1) * phi[I3(k, i, j + 1)]
                             + CNY3(3, j + 1, 1) * phi[I3(k, i, j)] - CNY3(1, j, 1) * phi[I3(k, i, j - 1)]);
        else
          dY[I3(k, i, j)] = (+CNY3(1, j, 2) * phi[I3(k, i, j + 1)] - CNY3(2, j, 2) * phi[I3(k, i, j)]
                             + CNY3(3, j, 2) * phi[I3(k, i, j - 1)] - CNY3(1, j - 1, 2) * phi[I3(k, i, j - 2)]);
      }
      for (int k = 0; k <= KS - 1; ++k) { dX[I3(k, i, j)] = 0.0; dY[I3(k, i, j)] = 0.0; }
      for (int k = KE + 1; k < KA; ++k) { dX[I3(k, i, j)] = 0.0; dY[I3(k, i, j)] = 0.0; }
    }
#undef CNZ3
#undef CNX3
#undef CNY3
}

void orc_filter_tables(const orc_grid *g, const orc_metric *m, double *cnz3, double *cnx3, double *cny3) {
  GRID_LOCALS;
  cn_tables t;
  filter_setup(g, m, &t);
  memcpy(cnz3, t.CNZ3, (size_t)3 * KA * 2 * sizeof(double));
  memcpy(cnx3, t.CNX3, (size_t)3 * IA * 2 * sizeof(double));
  memcpy(cny3, t.CNY3, (size_t)3 * JA * 2 * sizeof(double));
  free(t.CNZ3); free(t.CNX3); free(t.CNY3);
}

void orc_calc_diff3(const orc_grid *g, const orc_metric *m, double *diff, const double *phi, int KO, int IO, int JO) {
  cn_tables t;
  filter_setup(g, m, &t);
  calc_diff3(g, &t, diff, phi, KO, IO, JO);
  free(t.CNZ3); free(t.CNX3); free(t.CNY3);
}

static inline double pow4(double x) { double x2 = x * x; return x2 * x2; }

int orc_numfilter_coef(const orc_grid *g, double *num_diff,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const orc_metric *m, double DT, const double *REF_dens, const double *REF_pott,
    double ND_COEF, int nd_order, double ND_SFC_FACT, int ND_USE_RS) {
  GRID_LOCALS;
  if (nd_order != 1 || !ND_USE_RS) return -1;
  const double *CDZ = m->CDZ, *FDZ = m->FDZ, *CDX = m->CDX, *FDX = m->FDX, *CDY = m->CDY, *FDY = m->FDY;
  cn_tables t;
  filter_setup(g, m, &t);
  /* nd_order4 = 4 -> x**4 by binary exponentiation ((x*x)*(x*x)), as a Fortran real**integer power */
  const double DIFF4 = ND_COEF / ((double)(1 << 4) * DT);
  double *nd_cdz = alloc_nan(KA), *nd_fdz = alloc_nan(KA), *nd_cdx = alloc_nan(IA), *nd_fdx = alloc_nan(IA);
  double *nd_cdy = alloc_nan(JA), *nd_fdy = alloc_nan(JA);
  for (int k = KS - 1; k <= KE; ++k) nd_cdz[k] = DIFF4 * pow4(CDZ[k]);
  for (int k = KS + 1; k <= KE - 1; ++k) nd_fdz[k] = DIFF4 * pow4(FDZ[k]);
  for (int i = IS; i <= IE; ++i) { nd_cdx[i] = DIFF4 * pow4(CDX[i]); nd_fdx[i] = DIFF4 * pow4(FDX[i]); }
  for (int j = JS; j <= JE; ++j) { nd_cdy[j] = DIFF4 * pow4(CDY[j]); nd_fdy[j] = DIFF4 * pow4(FDY[j]); }
  double *work = alloc_nan(3 * N3), *fld = alloc_nan(N3);
  double *wZ = work, *wX = work + N3, *wY = work + 2 * N3;
#define ND(k, i, j, v, d) num_diff[I3(k, i, j) + N3 * ((size_t)(v) + 5 * (size_t)(d))]
#define D_(k, i, j) DENS[I3(k, i, j)]
#define SFC(v, d)                                                      \
  for (int j = JS; j <= JE; ++j)                                       \
    for (int i = IS; i <= IE; ++i) {                                   \
      ND(KS, i, j, v, d) = ND(KS, i, j, v, d) * ND_SFC_FACT;           \
      ND(KS + 1, i, j, v, d) = ND(KS + 1, i, j, v, d) * (1.0 + ND_SFC_FACT) * 0.5; \
    }
  /* density (COMMON.F90:541-621) */
  OMP_FOR
  for (int j = JS - 1; j <= JE + 2; ++j)
    for (int i = IS - 1; i <= IE + 2; ++i)
      for (int k = KS; k <= KE; ++k) fld[I3(k, i, j)] = DENS[I3(k, i, j)] - REF_dens[I3(k, i, j)];
  calc_diff3(g, &t, work, fld, 0, 0, 0);
  OMP_FOR
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS - 1; k <= KE; ++k) ND(k, i, j, I_DENS, ZDIR) = wZ[I3(k, i, j)] * nd_cdz[k];
      for (int k = 0; k <= KS - 2; ++k) ND(k, i, j, I_DENS, ZDIR) = 0.0;
      for (int k = KE + 1; k < KA; ++k) ND(k, i, j, I_DENS, ZDIR) = 0.0;
      for (int k = KS; k <= KE; ++k) ND(k, i, j, I_DENS, XDIR) = wX[I3(k, i, j)] * nd_cdx[i];
      for (int k = KS; k <= KE; ++k) ND(k, i, j, I_DENS, YDIR) = wY[I3(k, i, j)] * nd_cdy[j];
      for (int k = 0; k <= KS - 1; ++k) { ND(k, i, j, I_DENS, XDIR) = 0.0; ND(k, i, j, I_DENS, YDIR) = 0.0; }
      for (int k = KE + 1; k < KA; ++k) { ND(k, i, j, I_DENS, XDIR) = 0.0; ND(k, i, j, I_DENS, YDIR) = 0.0; }
    }
  SFC(I_DENS, XDIR);
  SFC(I_DENS, YDIR);
  /* z-momentum (COMMON.F90:629-707) */
  OMP_FOR
  for (int j = JS - 2; j <= JE + 2; ++j)
    for (int i = IS - 2; i <= IE + 2; ++i)
      for (int k = KS; k <= KE - 1; ++k) fld[I3(k, i, j)] = 2.0 * MOMZ[I3(k, i, j)] / (DENS[I3(k + 1, i, j)] + DENS[I3(k, i, j)]);
  calc_diff3(g, &t, work, fld, 1, 0, 0);
  OMP_FOR
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS + 1; k <= KE - 1; ++k) ND(k, i, j, I_MOMZ, ZDIR) = wZ[I3(k, i, j)] * nd_fdz[k] * DENS[I3(k, i, j)];
      for (int k = 0; k <= KS; ++k) ND(k, i, j, I_MOMZ, ZDIR) = 0.0;
      for (int k = KE; k < KA; ++k) ND(k, i, j, I_MOMZ, ZDIR) = 0.0;
      for (int k = KS; k <= KE - 1; ++k)
        ND(k, i, j, I_MOMZ, XDIR) = wX[I3(k, i, j)] * nd_cdx[i] * 0.25 * (D_(k + 1, i + 1, j) + D_(k + 1, i, j) + D_(k, i + 1, j) + D_(k, i, j));
      for (int k = KS; k <= KE - 1; ++k)
        ND(k, i, j, I_MOMZ, YDIR) = wY[I3(k, i, j)] * nd_cdy[j] * 0.25 * (D_(k + 1, i, j + 1) + D_(k + 1, i, j) + D_(k, i, j + 1) + D_(k, i, j));
      for (int k = 0; k <= KS - 1; ++k) { ND(k, i, j, I_MOMZ, XDIR) = 0.0; ND(k, i, j, I_MOMZ, YDIR) = 0.0; }
      for (int k = KE; k < KA; ++k) { ND(k, i, j, I_MOMZ, XDIR) = 0.0; ND(k, i, j, I_MOMZ, YDIR) = 0.0; }
    }
  SFC(I_MOMZ, XDIR);
  SFC(I_MOMZ, YDIR);
  /* x-momentum (COMMON.F90:713-790) */
  OMP_FOR
  for (int j = JS - 2; j <= JE + 2; ++j)
    for (int i = IS - 2; i <= IE + 1; ++i)
      for (int k = KS; k <= KE; ++k) fld[I3(k, i, j)] = 2.0 * MOMX[I3(k, i, j)] / (DENS[I3(k, i + 1, j)] + DENS[I3(k, i, j)]);
  calc_diff3(g, &t, work, fld, 0, 1, 0);
  OMP_FOR
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS; k <= KE - 1; ++k)
        ND(k, i, j, I_MOMX, ZDIR) = wZ[I3(k, i, j)] * nd_cdz[k] * 0.25 * (D_(k + 1, i + 1, j) + D_(k + 1, i, j) + D_(k, i + 1, j) + D_(k, i, j));
      for (int k = 0; k <= KS - 1; ++k) ND(k, i, j, I_MOMX, ZDIR) = 0.0;
      for (int k = KE; k < KA; ++k) ND(k, i, j, I_MOMX, ZDIR) = 0.0;
      for (int k = KS; k <= KE; ++k) ND(k, i, j, I_MOMX, XDIR) = wX[I3(k, i, j)] * nd_fdx[i] * DENS[I3(k, i, j)];
      for (int k = KS; k <= KE; ++k)
        ND(k, i, j, I_MOMX, YDIR) = wY[I3(k, i, j)] * nd_cdy[j] * 0.25 * (D_(k, i + 1, j + 1) + D_(k, i + 1, j) + D_(k, i, j + 1) + D_(k, i, j));
      for (int k = 0; k <= KS - 1; ++k) { ND(k, i, j, I_MOMX, XDIR) = 0.0; ND(k, i, j, I_MOMX, YDIR) = 0.0; }
      for (int k = KE + 1; k < KA; ++k) { ND(k, i, j, I_MOMX, XDIR) = 0.0; ND(k, i, j, I_MOMX, YDIR) = 0.0; }
    }
  SFC(I_MOMX, XDIR);
  SFC(I_MOMX, YDIR);
  /* y-momentum (COMMON.F90:796-873) */
  OMP_FOR
  for (int j = JS - 2; j <= JE + 1; ++j)
    for (int i = IS - 2; i <= IE + 2; ++i)
      for (int k = KS; k <= KE; ++k) fld[I3(k, i, j)] = 2.0 * MOMY[I3(k, i, j)] / (DENS[I3(k, i, j + 1)] + DENS[I3(k, i, j)]);
  calc_diff3(g, &t, work, fld, 0, 0, 1);
  OMP_FOR
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS; k <= KE - 1; ++k)
        ND(k, i, j, I_MOMY, ZDIR) = wZ[I3(k, i, j)] * nd_cdz[k] * 0.25 * (D_(k + 1, i, j + 1) + D_(k + 1, i, j) + D_(k, i, j + 1) + D_(k, i, j));
      for (int k = 0; k <= KS - 1; ++k) ND(k, i, j, I_MOMY, ZDIR) = 0.0;
      for (int k = KE; k < KA; ++k) ND(k, i, j, I_MOMY, ZDIR) = 0.0;
      for (int k = KS; k <= KE; ++k)
        ND(k, i, j, I_MOMY, XDIR) = wX[I3(k, i, j)] * nd_cdx[i] * 0.25 * (D_(k, i + 1, j + 1) + D_(k, i, j + 1) + D_(k, i + 1, j) + D_(k, i, j));
      for (int k = KS; k <= KE; ++k) ND(k, i, j, I_MOMY, YDIR) = wY[I3(k, i, j)] * nd_fdy[j] * DENS[I3(k, i, j)];
      for (int k = 0; k <= KS - 1; ++k) { ND(k, i, j, I_MOMY, XDIR) = 0.0; ND(k, i, j, I_MOMY, YDIR) = 0.0; }
      for (int k = KE + 1; k < KA; ++k) { ND(k, i, j, I_MOMY, XDIR) = 0.0; ND(k, i, j, I_MOMY, YDIR) = 0.0; }
    }
  SFC(I_MOMY, XDIR);
  SFC(I_MOMY, YDIR);
  /* rho*theta (COMMON.F90:880-962) */
  OMP_FOR
  for (int j = JS - 1; j <= JE + 2; ++j)
    for (int i = IS - 1; i <= IE + 2; ++i)
      for (int k = KS; k <= KE; ++k) fld[I3(k, i, j)] = RHOT[I3(k, i, j)] / DENS[I3(k, i, j)] - REF_pott[I3(k, i, j)];
  calc_diff3(g, &t, work, fld, 0, 0, 0);
  OMP_FOR
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS; k <= KE - 1; ++k)
        ND(k, i, j, I_RHOT, ZDIR) = wZ[I3(k, i, j)] * nd_cdz[k] * 0.5 * (D_(k + 1, i, j) + D_(k, i, j));
      for (int k = 0; k <= KS - 2; ++k) ND(k, i, j, I_RHOT, ZDIR) = 0.0;
      ND(KS - 1, i, j, I_RHOT, ZDIR) = wZ[I3(KS - 1, i, j)] * nd_cdz[KS - 1] * DENS[I3(KS, i, j)];
      ND(KE, i, j, I_RHOT, ZDIR) = wZ[I3(KE, i, j)] * nd_cdz[KE] * DENS[I3(KE, i, j)];
      for (int k = KE + 1; k < KA; ++k) ND(k, i, j, I_RHOT, ZDIR) = 0.0;
      for (int k = KS; k <= KE; ++k) ND(k, i, j, I_RHOT, XDIR) = wX[I3(k, i, j)] * nd_cdx[i] * 0.5 * (D_(k, i + 1, j) + D_(k, i, j));
      for (int k = KS; k <= KE; ++k) ND(k, i, j, I_RHOT, YDIR) = wY[I3(k, i, j)] * nd_cdy[j] * 0.5 * (D_(k, i, j + 1) + D_(k, i, j));
      for (int k = 0; k <= KS - 1; ++k) { ND(k, i, j, I_RHOT, XDIR) = 0.0; ND(k, i, j, I_RHOT, YDIR) = 0.0; }
      for (int k = KE + 1; k < KA; ++k) { ND(k, i, j, I_RHOT, XDIR) = 0.0; ND(k, i, j, I_RHOT, YDIR) = 0.0; }
    }
  SFC(I_RHOT, XDIR);
  SFC(I_RHOT, YDIR);
  for (int q = 0; q < 15; ++q) orc_halo_periodic(g, num_diff + N3 * (size_t)q);
#undef ND
#undef D_
#undef SFC
  free(work); free(fld); free(nd_cdz); free(nd_fdz); free(nd_cdx); free(nd_fdx); free(nd_cdy); free(nd_fdy);
  free(t.CNZ3); free(t.CNX3); free(t.CNY3);
  return 0;
}

/* ---- tracer filter (ATMOS_DYN_numfilter_coef_q, COMMON.F90:989-1200) ----
 * ND_ORDER = 1.  The differenced field is QTRC itself for a tracer that is
 * not QV (:1108-1112) and QTRC - REF_qv for QV with ND_USE_RS (:1085-1098);
 * QV without the reference state (the 19-point smoothing, :1049-1077) is
 * not supported (returns -1).  Density weights and surface rows as
 * :1117-1184, then the halo exchange of the three components (:1186-1196). */
int orc_numfilter_coef_q(const orc_grid *g, double *num_diff_q, const double *DENS, const double *QTRC, int is_qv,
    const orc_metric *m, double DT, const double *REF_qv, double ND_COEF, int nd_order, double ND_SFC_FACT, int ND_USE_RS) {
  GRID_LOCALS;
  if (nd_order != 1 || (is_qv && !ND_USE_RS)) return -1;
  const double *CDZ = m->CDZ, *CDX = m->CDX, *CDY = m->CDY;
  cn_tables t;
  filter_setup(g, m, &t);
  const double DIFF4 = ND_COEF / ((double)(1 << 4) * DT);
  double *nd_cdz = alloc_nan(KA), *nd_cdx = alloc_nan(IA), *nd_cdy = alloc_nan(JA);
  for (int k = KS - 1; k <= KE; ++k) nd_cdz[k] = DIFF4 * pow4(CDZ[k]);
  for (int i = IS; i <= IE; ++i) nd_cdx[i] = DIFF4 * pow4(CDX[i]);
  for (int j = JS; j <= JE; ++j) nd_cdy[j] = DIFF4 * pow4(CDY[j]);
  double *work = alloc_nan(3 * N3), *fld = alloc_nan(N3);
  double *wZ = work, *wX = work + N3, *wY = work + 2 * N3;
  const double *phi = QTRC;
  if (is_qv) {
    OMP_FOR
    for (int j = JS - 1; j <= JE + 2; ++j)
      for (int i = IS - 1; i <= IE + 2; ++i)
        for (int k = KS; k <= KE; ++k) fld[I3(k, i, j)] = QTRC[I3(k, i, j)] - REF_qv[I3(k, i, j)];
    phi = fld;
  }
  calc_diff3(g, &t, work, phi, 0, 0, 0);
  double *qZ = num_diff_q, *qX = num_diff_q + N3, *qY = num_diff_q + 2 * N3;
  OMP_FOR
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = KS; k <= KE - 1; ++k)
        qZ[I3(k, i, j)] = wZ[I3(k, i, j)] * nd_cdz[k] * 0.5 * (DENS[I3(k + 1, i, j)] + DENS[I3(k, i, j)]);
      for (int k = 0; k <= KS - 2; ++k) qZ[I3(k, i, j)] = 0.0;
      qZ[I3(KS - 1, i, j)] = wZ[I3(KS - 1, i, j)] * nd_cdz[KS - 1] * DENS[I3(KS, i, j)];
      qZ[I3(KE, i, j)] = wZ[I3(KE, i, j)] * nd_cdz[KE] * DENS[I3(KE, i, j)];
      for (int k = KE + 1; k < KA; ++k) qZ[I3(k, i, j)] = 0.0;
      for (int k = KS; k <= KE; ++k) {
        qX[I3(k, i, j)] = wX[I3(k, i, j)] * nd_cdx[i] * 0.5 * (DENS[I3(k, i + 1, j)] + DENS[I3(k, i, j)]);
        qY[I3(k, i, j)] = wY[I3(k, i, j)] * nd_cdy[j] * 0.5 * (DENS[I3(k, i, j + 1)] + DENS[I3(k, i, j)]);
      }
      for (int k = 0; k <= KS - 1; ++k) { qX[I3(k, i, j)] = 0.0; qY[I3(k, i, j)] = 0.0; }
      qX[I3(KS, i, j)] = qX[I3(KS, i, j)] * ND_SFC_FACT;
      qX[I3(KS + 1, i, j)] = qX[I3(KS + 1, i, j)] * (1.0 + ND_SFC_FACT) * 0.5;
      qY[I3(KS, i, j)] = qY[I3(KS, i, j)] * ND_SFC_FACT;
      qY[I3(KS + 1, i, j)] = qY[I3(KS + 1, i, j)] * (1.0 + ND_SFC_FACT) * 0.5;
      for (int k = KE + 1; k < KA; ++k) { qX[I3(k, i, j)] = 0.0; qY[I3(k, i, j)] = 0.0; }
    }
  for (int n = 0; n < 3; ++n) orc_halo_periodic(g, num_diff_q + N3 * (size_t)n);
  free(work); free(fld); free(nd_cdz); free(nd_cdx); free(nd_cdy);
  free(t.CNZ3); free(t.CNX3); free(t.CNY3);
  return 0;
}

/* ---- divergence for divergence damping (COMMON.F90:1437-1515) ----------- */
void orc_divergence(const orc_grid *g, double *DDIV, const double *MOMZ, const double *MOMX, const double *MOMY, const orc_metric *m) {
  GRID_LOCALS;
  const double *GS = m->GSQRT, *J13 = m->J13G, *J23 = m->J23G, *MP = m->MAPF;
  const double *RCDZ = m->RCDZ, *RCDX = m->RCDX, *RCDY = m->RCDY, *RFDZ = m->RFDZ, *FDZ = m->FDZ;
  const double J33G = m->J33G;
#define GSQRT(k, i, j, c) GS[I3(k, i, j) + N3 * (c)]
#define J13G(k, i, j, c) J13[I3(k, i, j) + N3 * (c)]
#define J23G(k, i, j, c) J23[I3(k, i, j) + N3 * (c)]
#define MAPF(i, j, n, c) MP[(size_t)(i) + (size_t)IA * ((size_t)(j) + (size_t)JA * ((size_t)(n) + 2 * (size_t)(c)))]
#define HDIV(kk)                                                                                          \
  MAPF(i, j, 0, I_XY) * MAPF(i, j, 1, I_XY)                                                               \
      * ((MOMX[I3(kk, i, j)] * GSQRT(kk, i, j, I_UYZ) / MAPF(i, j, 1, I_UY)                              \
          - MOMX[I3(kk, i - 1, j)] * GSQRT(kk, i - 1, j, I_UYZ) / MAPF(i - 1, j, 1, I_UY)) * RCDX[i]     \
         + (MOMY[I3(kk, i, j)] * GSQRT(kk, i, j, I_XVZ) / MAPF(i, j, 0, I_XV)                            \
          - MOMY[I3(kk, i, j - 1)] * GSQRT(kk, i, j - 1, I_XVZ) / MAPF(i, j - 1, 0, I_XV)) * RCDY[j])
  OMP_FOR
  for (int j = JS; j <= JE + 1; ++j)
    for (int i = IS; i <= IE + 1; ++i) {
      for (int k = KS - 1; k <= KE + 1; ++k) {
        if (k - 1 < 0 || k + 1 >= KA) continue;
        DDIV[I3(k, i, j)] = J33G * (MOMZ[I3(k, i, j)] - MOMZ[I3(k - 1, i, j)]) * RCDZ[k]
                          + ((MOMX[I3(k + 1, i, j)] + MOMX[I3(k + 1, i - 1, j)]) * J13G(k + 1, i, j, I_XYW)
                             - (MOMX[I3(k - 1, i, j)] + MOMX[I3(k - 1, i - 1, j)]) * J13G(k - 1, i, j, I_XYW)
                             + (MOMY[I3(k + 1, i, j)] + MOMY[I3(k + 1, i, j - 1)]) * J23G(k + 1, i, j, I_XYW)
                             - (MOMY[I3(k - 1, i, j)] + MOMY[I3(k - 1, i, j - 1)]) * J23G(k - 1, i, j, I_XYW)) / (FDZ[k] + FDZ[k - 1])
                          + HDIV(k);
      }
      DDIV[I3(KS, i, j)] = J33G * (MOMZ[I3(KS, i, j)]) * RCDZ[KS]
                         + ((MOMX[I3(KS + 1, i, j)] + MOMX[I3(KS + 1, i - 1, j)]) * J13G(KS + 1, i, j, I_XYW)
                            - (MOMX[I3(KS - 1, i, j)] + MOMX[I3(KS, i - 1, j)]) * J13G(KS, i, j, I_XYW)
                            + (MOMY[I3(KS + 1, i, j)] + MOMY[I3(KS + 1, i, j - 1)]) * J23G(KS + 1, i, j, I_XYW)
                            - (MOMY[I3(KS, i, j)] + MOMY[I3(KS, i, j - 1)]) * J23G(KS, i, j, I_XYW)) * RFDZ[KS]
                         + HDIV(KS);
      DDIV[I3(KE, i, j)] = J33G * (-MOMZ[I3(KE - 1, i, j)]) * RCDZ[KE]
                         + ((MOMX[I3(KE, i, j)] + MOMX[I3(KE, i - 1, j)]) * J13G(KE, i, j, I_XYW)
                            - (MOMX[I3(KE - 1, i, j)] + MOMX[I3(KE - 1, i - 1, j)]) * J13G(KE - 1, i, j, I_XYW)
                            + (MOMY[I3(KE, i, j)] + MOMY[I3(KE, i, j - 1)]) * J23G(KE, i, j, I_XYW)
                            - (MOMY[I3(KE - 1, i, j)] + MOMY[I3(KE - 1, i, j - 1)]) * J23G(KE - 1, i, j, I_XYW)) * RFDZ[KE]
                         + HDIV(KE);
    }
#undef HDIV
#undef GSQRT
#undef J13G
#undef J23G
#undef MAPF
}

/* ---- Rayleigh damping coefficient (COMMON.F90:323-381) ------------------- */
void orc_wdamp_setup(const orc_grid *g, double *wdamp_coef, double wdamp_tau, double wdamp_height, const double *FZ) {
  GRID_LOCALS;
  /* FZ(0:KA): FZ(kF) = FZ[kF]; for C level k (= kF-1) the face above is FZ[k+1] */
  const double PI = 4.0 * atan(1.0);
  if (wdamp_height < 0.0) { for (int k = 0; k < KA; ++k) wdamp_coef[k] = 0.0; return; }
  if (FZ[KE + 1] - wdamp_height < EPS_RP) { for (int k = 0; k < KA; ++k) wdamp_coef[k] = 0.0; return; }
  const double alpha = 1.0 / wdamp_tau;
  for (int k = KS; k <= KE; ++k) {
    double sw = 0.5 + fsign(0.5, FZ[k + 1] - wdamp_height);
    wdamp_coef[k] = alpha * sw * 0.5 * (1.0 - cos(PI * (FZ[k + 1] - wdamp_height) / (FZ[KE + 1] - wdamp_height)));
  }
  for (int k = 0; k <= KS - 1; ++k) wdamp_coef[k] = wdamp_coef[KS];
  for (int k = KE + 1; k < KA; ++k) wdamp_coef[k] = wdamp_coef[KE];
}

/* ---- FCT (COMMON.F90:1922-2734) ------------------------------------------ */
void orc_get_fact_fct(double *fact, double rw, double ru, double rv) {
  double ugev = fsign(0.5, fabs(ru) - fabs(rv)) + 0.5;
  double ugew = fsign(0.5, fabs(ru) - fabs(rw)) + 0.5;
  double vgew = fsign(0.5, fabs(rv) - fabs(rw)) + 0.5;
  double uzero = fsign(0.5, fabs(ru) - EPS_RP) - 0.5;
  double vzero = fsign(0.5, fabs(rv) - EPS_RP) - 0.5;
  double wzero = fsign(0.5, fabs(rw) - EPS_RP) - 0.5;
  double sign_uv = fsign(0.5, ru * rv) + 0.5;
  double sign_uw = fsign(0.5, ru * rw) + 0.5;
  double sign_vw = fsign(0.5, rv * rw) + 0.5;
  double wu = fabs(rw / (ru + uzero) * (1.0 + uzero));
  double vu = fabs(rv / (ru + uzero) * (1.0 + uzero));
  double uv = fabs(ru / (rv + vzero) * (1.0 + vzero));
  double wv = fabs(rw / (rv + vzero) * (1.0 + vzero));
  double uw = fabs(ru / (rw + wzero) * (1.0 + wzero));
  double vw = fabs(rv / (rw + wzero) * (1.0 + wzero));
  double umax = ugev * ugew * (1.0 + uzero);
  double vmax = (1.0 - ugev) * vgew;
  double wmax = 1.0 - ugev * ugew - vmax;
#define FACT(a, b, c) fact[(a) + 2 * (((b) + 1) + 3 * ((c) + 1))]
  for (int q = 0; q < 18; ++q) fact[q] = 0.0;
  FACT(0, 0, 0) = -ugev * ugew * uzero;
  FACT(1, 0, 0) = wmax * (1.0 - uw) * (1.0 - vw);
  FACT(0, 1, 0) = umax * (1.0 - vu) * (1.0 - wu);
  FACT(0, 0, 1) = vmax * (1.0 - uv) * (1.0 - wv);
  double c3 = (umax * vu * wu + vmax * uv * wv + wmax * uw * vw);
  FACT(1, 1, 1) = sign_uv * sign_uw * c3;
  FACT(1, -1, 1) = (1.0 - sign_uv) * (1.0 - sign_uw) * c3;
  FACT(1, 1, -1) = (1.0 - sign_uv) * sign_uw * c3;
  FACT(1, -1, -1) = sign_uv * (1.0 - sign_uw) * c3;
  double cuw = (ugew * wu * (1.0 - vu) + (1.0 - ugew) * uw * (1.0 - vw));
  FACT(1, 1, 0) = sign_uw * (1.0 - vmax) * cuw;
  FACT(1, -1, 0) = (1.0 - sign_uw) * (1.0 - vmax) * cuw;
  double cvw = (vgew * wv * (1.0 - uv) + (1.0 - vgew) * vw * (1.0 - uw));
  FACT(1, 0, 1) = sign_vw * (1.0 - umax) * cvw;
  FACT(1, 0, -1) = (1.0 - sign_vw) * (1.0 - umax) * cvw;
  double cuv = (ugev * vu * (1.0 - wu) + (1.0 - ugev) * uv * (1.0 - wv));
  FACT(0, 1, 1) = sign_uv * (1.0 - wmax) * cuv;
  FACT(0, 1, -1) = (1.0 - sign_uv) * (1.0 - wmax) * cuv;
}

/* 14-point upstream/downstream sums of COMMON.F90:2096-2155.  ka/kb are the
 * level offsets the reference uses for the "1" vertical index of qa/qb
 * (k+1 / k-1 in the interior, clamped at KS and KE); kc is level k. */
static inline double fct_qa(const double *f, const double *phi, size_t (*ix)(int, int, int, const orc_grid *),
                            const orc_grid *g, int k1, int k0, int i, int j) {
  return f[1 + 2 * (2 + 3 * 2)] * phi[ix(k1, i + 1, j + 1, g)]    /* fact(1, 1, 1) */
       + f[0 + 2 * (2 + 3 * 2)] * phi[ix(k0, i + 1, j + 1, g)]    /* fact(0, 1, 1) */
       + f[1 + 2 * (1 + 3 * 2)] * phi[ix(k1, i, j + 1, g)]        /* fact(1, 0, 1) */
       + f[0 + 2 * (1 + 3 * 2)] * phi[ix(k0, i, j + 1, g)]        /* fact(0, 0, 1) */
       + f[1 + 2 * (0 + 3 * 2)] * phi[ix(k1, i - 1, j + 1, g)]    /* fact(1,-1, 1) */
       + f[1 + 2 * (2 + 3 * 1)] * phi[ix(k1, i + 1, j, g)]        /* fact(1, 1, 0) */
       + f[0 + 2 * (2 + 3 * 1)] * phi[ix(k0, i + 1, j, g)]        /* fact(0, 1, 0) */
       + f[1 + 2 * (1 + 3 * 1)] * phi[ix(k1, i, j, g)]            /* fact(1, 0, 0) */
       + f[1 + 2 * (0 + 3 * 1)] * phi[ix(k1, i - 1, j, g)]        /* fact(1,-1, 0) */
       + f[1 + 2 * (2 + 3 * 0)] * phi[ix(k1, i + 1, j - 1, g)]    /* fact(1, 1,-1) */
       + f[0 + 2 * (2 + 3 * 0)] * phi[ix(k0, i + 1, j - 1, g)]    /* fact(0, 1,-1) */
       + f[1 + 2 * (1 + 3 * 0)] * phi[ix(k1, i, j - 1, g)]        /* fact(1, 0,-1) */
       + f[1 + 2 * (0 + 3 * 0)] * phi[ix(k1, i - 1, j - 1, g)]    /* fact(1,-1,-1) */
       + f[0 + 2 * (1 + 3 * 1)] * phi[ix(k0, i, j, g)];           /* fact(0, 0, 0) */
}
/* qb uses the mirrored stencil with the reference's own index pattern
 * (COMMON.F90:2111-2124): note fact(0,1,-1) and fact(1,0,-1) both look at j-1. */
static inline double fct_qb(const double *f, const double *phi, size_t (*ix)(int, int, int, const orc_grid *),
                            const orc_grid *g, int k1, int k0, int i, int j) {
  return f[1 + 2 * (2 + 3 * 2)] * phi[ix(k1, i - 1, j - 1, g)]    /* fact(1, 1, 1) */
       + f[0 + 2 * (2 + 3 * 2)] * phi[ix(k0, i - 1, j - 1, g)]    /* fact(0, 1, 1) */
       + f[1 + 2 * (1 + 3 * 2)] * phi[ix(k1, i, j - 1, g)]        /* fact(1, 0, 1) */
       + f[0 + 2 * (1 + 3 * 2)] * phi[ix(k0, i, j - 1, g)]        /* fact(0, 0, 1) */
       + f[1 + 2 * (0 + 3 * 2)] * phi[ix(k1, i + 1, j - 1, g)]    /* fact(1,-1, 1) */
       + f[1 + 2 * (2 + 3 * 1)] * phi[ix(k1, i - 1, j, g)]        /* fact(1, 1, 0) */
       + f[0 + 2 * (2 + 3 * 1)] * phi[ix(k0, i - 1, j, g)]        /* fact(0, 1, 0) */
       + f[1 + 2 * (1 + 3 * 1)] * phi[ix(k1, i, j, g)]            /* fact(1, 0, 0) */
       + f[1 + 2 * (0 + 3 * 1)] * phi[ix(k1, i + 1, j, g)]        /* fact(1,-1, 0) */
       + f[1 + 2 * (2 + 3 * 0)] * phi[ix(k1, i - 1, j + 1, g)]    /* fact(1, 1,-1) */
       + f[0 + 2 * (2 + 3 * 0)] * phi[ix(k0, i - 1, j - 1, g)]    /* fact(0, 1,-1) */
       + f[1 + 2 * (1 + 3 * 0)] * phi[ix(k1, i, j - 1, g)]        /* fact(1, 0,-1) */
       + f[1 + 2 * (0 + 3 * 0)] * phi[ix(k1, i + 1, j + 1, g)]    /* fact(1,-1,-1) */
       + f[0 + 2 * (1 + 3 * 1)] * phi[ix(k0, i, j, g)];           /* fact(0, 0, 0) */
}
static size_t ix3f(int k, int i, int j, const orc_grid *g) { const int KA = g->KA, IA = g->IA; return I3(k, i, j); }

/* Test hook of the frozen-limiter checks: when set, orc_fct copies its
 * ratios R+ then R- (2 x KA*IA*JA, before their halo fill) out to
 * orc_fct_save_r, and replaces them by orc_fct_fixed_r. */
double *orc_fct_save_r = NULL;
const double *orc_fct_fixed_r = NULL;

void orc_fct(const orc_grid *g, double *qflx_anti,
    const double *phi_in, const double *DENS0, const double *DENS,
    const double *qflx_hi, const double *qflx_lo, const double *mflx_hi,
    const double *RDZ, const double *RDX, const double *RDY,
    const double *GSQRT, const double *MAPF, double dt, int flag_vect) {
  GRID_LOCALS;
  const double EPSILON = EPS_RP;
  const double *hZ = qflx_hi, *hX = qflx_hi + N3, *hY = qflx_hi + 2 * N3;
  const double *lZ = qflx_lo, *lX = qflx_lo + N3, *lY = qflx_lo + 2 * N3;
  const double *mZ = mflx_hi, *mX = mflx_hi + N3, *mY = mflx_hi + 2 * N3;
  double *aZ = qflx_anti, *aX = qflx_anti + N3, *aY = qflx_anti + 2 * N3;
  double *phi_lo = alloc_nan(N3), *pjpls = alloc_nan(N3), *pjmns = alloc_nan(N3);
  double *qjpls = alloc_nan(N3), *qjmns = alloc_nan(N3), *rjpls = alloc_nan(N3), *rjmns = alloc_nan(N3);
  for (size_t q = 0; q < 3 * N3; ++q) qflx_anti[q] = -9.9999e30; /* UNDEF (COMMON.F90:1970) */
  const int IIS = IS, IIE = IE, JJS = JS, JJE = JE;
#define M1(i, j) MAPF[(size_t)(i) + (size_t)IA * (size_t)(j)]
#define M2(i, j) MAPF[(size_t)(i) + (size_t)IA * ((size_t)(j) + (size_t)JA)]
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      for (int k = KS; k <= KE - 1; ++k) aZ[I3(k, i, j)] = hZ[I3(k, i, j)] - lZ[I3(k, i, j)];
      aZ[I3(KS - 1, i, j)] = 0.0;
      aZ[I3(KE, i, j)] = 0.0;
    }
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS - 1; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) aX[I3(k, i, j)] = hX[I3(k, i, j)] - lX[I3(k, i, j)];
  for (int j = JJS - 1; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) aY[I3(k, i, j)] = hY[I3(k, i, j)] - lY[I3(k, i, j)];
  /* update monotone scheme (COMMON.F90:2033-2045) */
  OMP_FOR
  for (int j = JJS - 1; j <= JJE + 1; ++j)
    for (int i = IIS - 1; i <= IIE + 1; ++i)
      for (int k = KS; k <= KE; ++k)
        phi_lo[I3(k, i, j)] = (phi_in[I3(k, i, j)] * DENS0[I3(k, i, j)]
                               + dt * (-((lZ[I3(k, i, j)] - lZ[I3(k - 1, i, j)]) * RDZ[k]
                                         + (lX[I3(k, i, j)] - lX[I3(k, i - 1, j)]) * RDX[i]
                                         + (lY[I3(k, i, j)] - lY[I3(k, i, j - 1)]) * RDY[j])
                                       * M1(i, j) * M2(i, j) / GSQRT[I3(k, i, j)]))
                            / DENS[I3(k, i, j)];
  /* P+ / P- (COMMON.F90:2051-2077) */
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        pjpls[I3(k, i, j)] = dt * ((fmax2(0.0, aZ[I3(k - 1, i, j)]) - fmin2(0.0, aZ[I3(k, i, j)])) * RDZ[k]
                                 + (fmax2(0.0, aX[I3(k, i - 1, j)]) - fmin2(0.0, aX[I3(k, i, j)])) * RDX[i]
                                 + (fmax2(0.0, aY[I3(k, i, j - 1)]) - fmin2(0.0, aY[I3(k, i, j)])) * RDY[j])
                           * M1(i, j) * M2(i, j) / GSQRT[I3(k, i, j)];
        pjmns[I3(k, i, j)] = dt * ((fmax2(0.0, aZ[I3(k, i, j)]) - fmin2(0.0, aZ[I3(k - 1, i, j)])) * RDZ[k]
                                 + (fmax2(0.0, aX[I3(k, i, j)]) - fmin2(0.0, aX[I3(k, i - 1, j)])) * RDX[i]
                                 + (fmax2(0.0, aY[I3(k, i, j)]) - fmin2(0.0, aY[I3(k, i, j - 1)])) * RDY[j])
                           * M1(i, j) * M2(i, j) / GSQRT[I3(k, i, j)];
      }
  /* allowable range (COMMON.F90:2085-2515) */
  if (flag_vect) {
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE; ++k) {
          double fact[18];
          double rw, ru, rv;
          int ka, kb; /* vertical index used for "1" in qa / qb */
          if (k == KS) {
            rw = (mZ[I3(KS, i, j)]) * RDZ[KS];
            ka = KS + 1; kb = KS;
          } else if (k == KE) {
            rw = (mZ[I3(KE - 1, i, j)]) * RDZ[KE];
            ka = KE; kb = KE - 1;
          } else {
            rw = (mZ[I3(k, i, j)] + mZ[I3(k - 1, i, j)]) * RDZ[k];
            ka = k + 1; kb = k - 1;
          }
          ru = (mX[I3(k, i, j)] + mX[I3(k, i - 1, j)]) * RDX[i];
          rv = (mY[I3(k, i, j)] + mY[I3(k, i, j - 1)]) * RDY[j];
          orc_get_fact_fct(fact, rw, ru, rv);
          double qa_in = fct_qa(fact, phi_in, ix3f, g, ka, k, i, j);
          double qb_in = fct_qb(fact, phi_in, ix3f, g, kb, k, i, j);
          double qa_lo = fct_qa(fact, phi_lo, ix3f, g, ka, k, i, j);
          double qb_lo = fct_qb(fact, phi_lo, ix3f, g, kb, k, i, j);
          double pin = phi_in[I3(k, i, j)], plo = phi_lo[I3(k, i, j)];
          double qmax = fmax2(fmax2(fmax2(fmax2(fmax2(pin, qa_in), qb_in), plo), qa_lo), qb_lo);
          double qmin = fmin2(fmin2(fmin2(fmin2(fmin2(pin, qa_in), qb_in), plo), qa_lo), qb_lo);
          qjpls[I3(k, i, j)] = (qmax - plo) * DENS[I3(k, i, j)];
          qjmns[I3(k, i, j)] = (plo - qmin) * DENS[I3(k, i, j)];
        }
  } else {
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE; ++k) {
          const int kp = k == KE ? k : k + 1, km = k == KS ? k : k - 1;
          double qmax = phi_in[I3(k, i, j)], qmin = phi_in[I3(k, i, j)];
          const double *arr[2] = {phi_in, phi_lo};
          for (int a = 0; a < 2; ++a) {
            const double *p = arr[a];
            double v[7] = {p[I3(k, i, j)], p[I3(kp, i, j)], p[I3(km, i, j)], p[I3(k, i + 1, j)], p[I3(k, i - 1, j)], p[I3(k, i, j + 1)], p[I3(k, i, j - 1)]};
            for (int q = 0; q < 7; ++q) { qmax = fmax2(qmax, v[q]); qmin = fmin2(qmin, v[q]); }
          }
          qjpls[I3(k, i, j)] = (qmax - phi_lo[I3(k, i, j)]) * DENS[I3(k, i, j)];
          qjmns[I3(k, i, j)] = (phi_lo[I3(k, i, j)] - qmin) * DENS[I3(k, i, j)];
        }
  }
  /* limitation factors (COMMON.F90:2522-2547) */
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double zerosw = 0.5 - fsign(0.5, pjpls[I3(k, i, j)] - EPSILON);
        rjpls[I3(k, i, j)] = ORC_FROZEN(fmin2(1.0, qjpls[I3(k, i, j)] * (1.0 - zerosw) / (pjpls[I3(k, i, j)] - zerosw)));
        zerosw = 0.5 - fsign(0.5, pjmns[I3(k, i, j)] - EPSILON);
        rjmns[I3(k, i, j)] = ORC_FROZEN(fmin2(1.0, qjmns[I3(k, i, j)] * (1.0 - zerosw) / (pjmns[I3(k, i, j)] - zerosw)));
      }
  if (orc_fct_fixed_r) {
    memcpy(rjpls, orc_fct_fixed_r, N3 * sizeof(double));
    memcpy(rjmns, orc_fct_fixed_r + N3, N3 * sizeof(double));
  }
  if (orc_fct_save_r) {
    memcpy(orc_fct_save_r, rjpls, N3 * sizeof(double));
    memcpy(orc_fct_save_r + N3, rjmns, N3 * sizeof(double));
  }
  orc_halo_periodic(g, rjpls);
  orc_halo_periodic(g, rjmns);
  /* limited anti-diffusive fluxes (COMMON.F90:2570-2659) */
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i) {
      for (int k = KS; k <= KE - 1; ++k) {
        double dirsw = 0.5 + fsign(0.5, aZ[I3(k, i, j)]);
        aZ[I3(k, i, j)] = aZ[I3(k, i, j)]
                        * (1.0 - fmin2(rjpls[I3(k + 1, i, j)], rjmns[I3(k, i, j)]) * (dirsw)
                               - fmin2(rjpls[I3(k, i, j)], rjmns[I3(k + 1, i, j)]) * (1.0 - dirsw));
      }
      aZ[I3(KS - 1, i, j)] = 0.0;
      aZ[I3(KE, i, j)] = 0.0;
    }
  OMP_FOR
  for (int j = JJS; j <= JJE; ++j)
    for (int i = IIS - 1; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double dirsw = 0.5 + fsign(0.5, aX[I3(k, i, j)]);
        aX[I3(k, i, j)] = aX[I3(k, i, j)]
                        * (1.0 - fmin2(rjpls[I3(k, i + 1, j)], rjmns[I3(k, i, j)]) * (dirsw)
                               - fmin2(rjpls[I3(k, i, j)], rjmns[I3(k, i + 1, j)]) * (1.0 - dirsw));
      }
  OMP_FOR
  for (int j = JJS - 1; j <= JJE; ++j)
    for (int i = IIS; i <= IIE; ++i)
      for (int k = KS; k <= KE; ++k) {
        double dirsw = 0.5 + fsign(0.5, aY[I3(k, i, j)]);
        aY[I3(k, i, j)] = aY[I3(k, i, j)]
                        * (1.0 - fmin2(rjpls[I3(k, i, j + 1)], rjmns[I3(k, i, j)]) * (dirsw)
                               - fmin2(rjpls[I3(k, i, j)], rjmns[I3(k, i, j + 1)]) * (1.0 - dirsw));
      }
#undef M1
#undef M2
  free(phi_lo); free(pjpls); free(pjmns); free(qjpls); free(qjmns); free(rjpls); free(rjmns);
}

/* ---- tracer step (TRACER:72-242) ----------------------------------------- */
int orc_tstep_tracer_heve(const orc_grid *g, int s,
    double *QTRCo, const double *QTRC, const double *QTRC0, const double *RHOQ_t,
    const double *DENS0, const double *DENS, const double *mflx_hi, const double *num_diff,
    const double *GSQRT, const double *MAPF, const double *CDZ,
    const double *RCDZ, const double *RCDX, const double *RCDY,
    double dtl, int flag_fct, int flag_fct_along_stream) {
  GRID_LOCALS;
  (void)CDZ;
  if (s != ORC_UD3 && s != ORC_CD4 && s != ORC_UD3KOREN && s != ORC_UD1 && s != ORC_CD2 && s != ORC_UD5 && s != ORC_CD6) return -1;
  double *qflx_hi = alloc_nan(3 * N3), *qflx_lo = alloc_nan(3 * N3), *qflx_anti = NULL;
  const double *mZ = mflx_hi, *mX = mflx_hi + N3, *mY = mflx_hi + 2 * N3;
  const double *GXYW = GSQRT + N3 * I_XYW, *GUYZ = GSQRT + N3 * I_UYZ, *GXVZ = GSQRT + N3 * I_XVZ, *GXYZ = GSQRT + N3 * I_XYZ;
  const int IIS = IS, IIE = IE, JJS = JS, JJE = JE;
  fluxZ_XYZ(g, s, qflx_hi, mZ, QTRC, GXYW, num_diff, IIS, IIE, JJS, JJE);
  fluxX_XYZ(g, s, qflx_hi + N3, mX, QTRC, GUYZ, num_diff + N3, IIS, IIE, JJS, JJE);
  fluxY_XYZ(g, s, qflx_hi + 2 * N3, mY, QTRC, GXVZ, num_diff + 2 * N3, IIS, IIE, JJS, JJE);
  if (flag_fct) {
    fluxZ_XYZ(g, ORC_UD1, qflx_lo, mZ, QTRC0, GXYW, num_diff, IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    fluxX_XYZ(g, ORC_UD1, qflx_lo + N3, mX, QTRC0, GUYZ, num_diff + N3, IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    fluxY_XYZ(g, ORC_UD1, qflx_lo + 2 * N3, mY, QTRC0, GXVZ, num_diff + 2 * N3, IIS - 1, IIE + 1, JJS - 1, JJE + 1);
    qflx_anti = alloc_nan(3 * N3);
    orc_fct(g, qflx_anti, QTRC0, DENS0, DENS, qflx_hi, qflx_lo, mflx_hi, RCDZ, RCDX, RCDY, GXYZ, MAPF, dtl, flag_fct_along_stream);
    const double *hZ = qflx_hi, *hX = qflx_hi + N3, *hY = qflx_hi + 2 * N3;
    const double *aZ = qflx_anti, *aX = qflx_anti + N3, *aY = qflx_anti + 2 * N3;
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE; ++k)
          QTRCo[I3(k, i, j)] = (QTRC0[I3(k, i, j)] * DENS0[I3(k, i, j)]
                                + dtl * (-((hZ[I3(k, i, j)] - aZ[I3(k, i, j)] - hZ[I3(k - 1, i, j)] + aZ[I3(k - 1, i, j)]) * RCDZ[k]
                                           + (hX[I3(k, i, j)] - aX[I3(k, i, j)] - hX[I3(k, i - 1, j)] + aX[I3(k, i - 1, j)]) * RCDX[i]
                                           + (hY[I3(k, i, j)] - aY[I3(k, i, j)] - hY[I3(k, i, j - 1)] + aY[I3(k, i, j - 1)]) * RCDY[j])
                                         * MAPF[(size_t)i + (size_t)IA * j] * MAPF[(size_t)i + (size_t)IA * (j + JA)] / GXYZ[I3(k, i, j)]
                                         + RHOQ_t[I3(k, i, j)]))
                             / DENS[I3(k, i, j)];
    free(qflx_anti);
  } else {
    const double *hZ = qflx_hi, *hX = qflx_hi + N3, *hY = qflx_hi + 2 * N3;
    OMP_FOR
    for (int j = JJS; j <= JJE; ++j)
      for (int i = IIS; i <= IIE; ++i)
        for (int k = KS; k <= KE; ++k)
          QTRCo[I3(k, i, j)] = (QTRC0[I3(k, i, j)] * DENS0[I3(k, i, j)]
                                + dtl * (-((hZ[I3(k, i, j)] - hZ[I3(k - 1, i, j)]) * RCDZ[k]
                                           + (hX[I3(k, i, j)] - hX[I3(k, i - 1, j)]) * RCDX[i]
                                           + (hY[I3(k, i, j)] - hY[I3(k, i, j - 1)]) * RCDY[j])
                                         * MAPF[(size_t)i + (size_t)IA * j] * MAPF[(size_t)i + (size_t)IA * (j + JA)] / GXYZ[I3(k, i, j)]
                                         + RHOQ_t[I3(k, i, j)]))
                             / DENS[I3(k, i, j)];
  }
  free(qflx_hi);
  free(qflx_lo);
  return 0;
}

/* ---- tracer RK3 (TRK3:92-208) -------------------------------------------- */
int orc_tinteg_tracer_rk3(const orc_grid *g, int s,
    double *QTRC, const double *QTRC0, const double *RHOQ_t,
    const double *DENS0, const double *DENS, const double *mflx_hi, const double *num_diff,
    const double *GSQRT, const double *MAPF, const double *CDZ,
    const double *RCDZ, const double *RCDX, const double *RCDY,
    double dtl, int flag_fct, int flag_fct_along_stream) {
  GRID_LOCALS;
  double *DENS_RK = alloc_nan(N3), *RK1 = alloc_nan(N3), *RK2 = alloc_nan(N3);
  OMP_FOR
  for (int j = JS - 1; j <= JE + 1; ++j)
    for (int i = IS - 1; i <= IE + 1; ++i)
      for (int k = KS; k <= KE; ++k) DENS_RK[I3(k, i, j)] = DENS0[I3(k, i, j)] + (DENS[I3(k, i, j)] - DENS0[I3(k, i, j)]) / 3.0;
  int rc = orc_tstep_tracer_heve(g, s, RK1, QTRC, QTRC0, RHOQ_t, DENS0, DENS_RK, mflx_hi, num_diff, GSQRT, MAPF, CDZ, RCDZ, RCDX, RCDY,
                                 dtl / 3.0, 0, flag_fct_along_stream);
  orc_halo_periodic(g, RK1);
  OMP_FOR
  for (int j = JS - 1; j <= JE + 1; ++j)
    for (int i = IS - 1; i <= IE + 1; ++i)
      for (int k = KS; k <= KE; ++k) DENS_RK[I3(k, i, j)] = DENS0[I3(k, i, j)] + (DENS[I3(k, i, j)] - DENS0[I3(k, i, j)]) * 0.5;
  rc |= orc_tstep_tracer_heve(g, s, RK2, RK1, QTRC0, RHOQ_t, DENS0, DENS_RK, mflx_hi, num_diff, GSQRT, MAPF, CDZ, RCDZ, RCDX, RCDY,
                              dtl / 2.0, 0, flag_fct_along_stream);
  orc_halo_periodic(g, RK2);
  rc |= orc_tstep_tracer_heve(g, s, QTRC, RK2, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff, GSQRT, MAPF, CDZ, RCDZ, RCDX, RCDY,
                              dtl, flag_fct, flag_fct_along_stream);
  free(DENS_RK); free(RK1); free(RK2);
  return rc;
}

/* ---- tracer Euler (TEUL:80-133) ------------------------------------------- */
/* One stage of dtl with the end-of-step density; the reference's output
 * aliases its input QTRC (TEUL:120-121), which its stage reads only before
 * the update loop, so the stage writes a separate field here and copies it. */
int orc_tinteg_tracer_euler(const orc_grid *g, int s,
    double *QTRC, const double *QTRC0, const double *RHOQ_t,
    const double *DENS0, const double *DENS, const double *mflx_hi, const double *num_diff,
    const double *GSQRT, const double *MAPF, const double *CDZ,
    const double *RCDZ, const double *RCDX, const double *RCDY,
    double dtl, int flag_fct, int flag_fct_along_stream) {
  GRID_LOCALS;
  double *out = alloc_nan(N3);
  memcpy(out, QTRC, N3 * sizeof(double));
  int rc = orc_tstep_tracer_heve(g, s, out, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff, GSQRT, MAPF, CDZ, RCDZ,
                                 RCDX, RCDY, dtl, flag_fct, flag_fct_along_stream);
  memcpy(QTRC, out, N3 * sizeof(double));
  free(out);
  return rc;
}

/* ---- RK3 short step (RK3:188-576) ----------------------------------------- */
int orc_tinteg_short_rk3(const orc_grid *g, int s, int heve, int ws2002,
    double *DENS, double *MOMZ, double *MOMX, double *MOMY, double *RHOT,
    double *mflx_hi, double *tflx_hi,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef, double divdmp_coef, const double *DDIV,
    const orc_metric *m, const double *REF_dens, const double *REF_rhot, double dt) {
  GRID_LOCALS;
  double *S0[5], *R[2][5], *mf1 = alloc_nan(3 * N3), *tf1 = alloc_nan(3 * N3);
  double *mf2 = alloc_nan(3 * N3), *tf2 = alloc_nan(3 * N3);
  double *cur[5] = {DENS, MOMZ, MOMX, MOMY, RHOT};
  for (int v = 0; v < 5; ++v) {
    S0[v] = alloc_nan(N3);
    memcpy(S0[v], cur[v], N3 * sizeof(double)); /* DENS0 = DENS ... (RK3:306-317) */
    for (int r = 0; r < 2; ++r) R[r][v] = alloc_nan(N3);
  }
  /* fact_dt1, fact_dt2 (RK3:141-166); dtrk = dt * fact (RK3:359, :412) */
  const double fact_dt1 = 1.0 / 3.0, fact_dt2 = ws2002 ? 1.0 / 2.0 : 2.0 / 3.0;
  const double dtrk_s[3] = {dt * fact_dt1, dt * fact_dt2, dt};
  double *mo_s[3] = {mf1, mf2, mflx_hi}, *to_s[3] = {tf1, tf2, tflx_hi};
  int rc = 0;
  for (int st = 0; st < 3; ++st) {
    double **in = st == 0 ? S0 : R[st - 1];
    double **out = st == 2 ? cur : R[st];
    if (heve)
      rc |= orc_tstep_short_heve(g, s, out[0], out[1], out[2], out[3], out[4], mo_s[st], to_s[st],
                                 S0[0], S0[1], S0[2], S0[3], S0[4], in[0], in[1], in[2], in[3], in[4],
                                 DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, DPRES0, RT2P, CORIOLI,
                                 num_diff, wdamp_coef, divdmp_coef, DDIV, 0, 0, 1, m, REF_dens, REF_rhot, dtrk_s[st]);
    else
      rc |= orc_tstep_short_hevi(g, s, out[0], out[1], out[2], out[3], out[4], mo_s[st], to_s[st],
                                 S0[0], S0[1], S0[2], S0[3], S0[4], in[0], in[1], in[2], in[3], in[4],
                                 DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, DPRES0, RT2P, CORIOLI,
                                 num_diff, wdamp_coef, divdmp_coef, DDIV, m, REF_dens, REF_rhot, dtrk_s[st]);
    if (st < 2)
      for (int v = 0; v < 5; ++v) orc_halo_periodic(g, out[v]);
  }
  if (!ws2002) {
    /* Heun's combination (RK3:505-571); MOMZ on KS..KE-1 */
    for (int v = 0; v < 5; ++v) {
      const int kend = v == 1 ? KE - 1 : KE;
      OMP_FOR
      for (int j = JS; j <= JE; ++j)
        for (int i = IS; i <= IE; ++i)
          for (int k = KS; k <= kend; ++k) {
            size_t p = I3(k, i, j);
            cur[v][p] = (R[0][v][p] * 3.0 + cur[v][p] * 3.0 - S0[v][p] * 2.0) / 4.0;
          }
    }
    for (int n = 0; n < 3; ++n)
      for (int j = JS; j <= JE; ++j)
        for (int i = IS; i <= IE; ++i)
          for (int k = KS; k <= KE; ++k) {
            size_t p = I3(k, i, j) + N3 * n;
            mflx_hi[p] = (mf1[p] + mflx_hi[p] * 3.0) / 4.0;
            tflx_hi[p] = (tf1[p] + tflx_hi[p] * 3.0) / 4.0;
          }
  }
  for (int v = 0; v < 5; ++v) {
    free(S0[v]);
    for (int r = 0; r < 2; ++r) free(R[r][v]);
  }
  free(mf1); free(tf1); free(mf2); free(tf2);
  return rc;
}

/* ---- RK4 short step (RK4:260-660) ----------------------------------------- */
int orc_tinteg_short_rk4(const orc_grid *g, int s, int heve,
    double *DENS, double *MOMZ, double *MOMX, double *MOMY, double *RHOT,
    double *mflx_hi, double *tflx_hi,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef, double divdmp_coef, const double *DDIV,
    const orc_metric *m, const double *REF_dens, const double *REF_rhot, double dt) {
  GRID_LOCALS;
  double *S0[5], *R[3][5], *mf[3], *tf[3];
  double *cur[5] = {DENS, MOMZ, MOMX, MOMY, RHOT};
  for (int v = 0; v < 5; ++v) {
    S0[v] = alloc_nan(N3);
    memcpy(S0[v], cur[v], N3 * sizeof(double)); /* DENS0 = DENS ... (RK4.F90:326-336) */
    for (int r = 0; r < 3; ++r) R[r][v] = alloc_nan(N3);
  }
  for (int r = 0; r < 3; ++r) { mf[r] = alloc_nan(3 * N3); tf[r] = alloc_nan(3 * N3); }
  const double dtrk_s[4] = {dt / 2.0, dt / 2.0, dt, dt};
  int rc = 0;
  for (int st = 0; st < 4; ++st) {
    double **in = st == 0 ? S0 : R[st - 1];
    double **out = st == 3 ? cur : R[st];
    double *mo = st == 3 ? mflx_hi : mf[st], *to = st == 3 ? tflx_hi : tf[st];
    if (heve)
      rc |= orc_tstep_short_heve(g, s, out[0], out[1], out[2], out[3], out[4], mo, to,
                                 S0[0], S0[1], S0[2], S0[3], S0[4], in[0], in[1], in[2], in[3], in[4],
                                 DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, DPRES0, RT2P, CORIOLI,
                                 num_diff, wdamp_coef, divdmp_coef, DDIV, 0, 0, 1, m, REF_dens, REF_rhot, dtrk_s[st]);
    else
      rc |= orc_tstep_short_hevi(g, s, out[0], out[1], out[2], out[3], out[4], mo, to,
                                 S0[0], S0[1], S0[2], S0[3], S0[4], in[0], in[1], in[2], in[3], in[4],
                                 DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, DPRES0, RT2P, CORIOLI,
                                 num_diff, wdamp_coef, divdmp_coef, DDIV, m, REF_dens, REF_rhot, dtrk_s[st]);
    if (st < 3)
      for (int v = 0; v < 5; ++v) orc_halo_periodic(g, out[v]);
  }
  /* combination (RK4.F90:562-660) */
  for (int v = 0; v < 5; ++v) {
    const int kend = v == 1 ? KE - 1 : KE;
    OMP_FOR
    for (int j = JS; j <= JE; ++j)
      for (int i = IS; i <= IE; ++i)
        for (int k = KS; k <= kend; ++k) {
          size_t p = I3(k, i, j);
          cur[v][p] = (R[0][v][p] * 2.0 + R[1][v][p] * 4.0 + R[2][v][p] * 2.0 + cur[v][p] - S0[v][p] * 3.0) / 6.0;
        }
  }
  for (int n = 0; n < 3; ++n)
    for (int j = JS; j <= JE; ++j)
      for (int i = IS; i <= IE; ++i)
        for (int k = KS; k <= KE; ++k) {
          size_t p = I3(k, i, j) + N3 * n;
          mflx_hi[p] = (mf[0][p] + mf[1][p] * 2.0 + mf[2][p] * 2.0 + mflx_hi[p]) / 6.0;
          tflx_hi[p] = (tf[0][p] + tf[1][p] * 2.0 + tf[2][p] * 2.0 + tflx_hi[p]) / 6.0;
        }
  for (int v = 0; v < 5; ++v) {
    free(S0[v]);
    for (int r = 0; r < 3; ++r) free(R[r][v]);
  }
  for (int r = 0; r < 3; ++r) { free(mf[r]); free(tf[r]); }
  return rc;
}

/* ---- large step (LARGE:377-1127) ------------------------------------------ */
int orc_large_step(const orc_grid *g, const orc_metric *m, const orc_config *c, orc_state *st) {
  GRID_LOCALS;
  if (c->tinteg_short != ORC_TINTEG_RK4 && c->tinteg_short != ORC_TINTEG_RK3 && c->tinteg_short != ORC_TINTEG_RK3WS2002)
    return -1;
  if (c->tinteg_tracer != ORC_TINTEG_TRACER_RK3WS2002 && c->tinteg_tracer != ORC_TINTEG_TRACER_EULER) return -1;
  const double dtl = c->dtl;
  const int nstep = (int)ceil((dtl - EPS_RP) / c->dts);
  const double dts = dtl / nstep;
  double *RT2P = alloc_nan(N3), *DPRES0 = alloc_nan(N3), *REF_rhot = alloc_nan(N3);
  double *DENS00 = alloc_nan(N3), *mflx_av = alloc_nan(3 * N3);
  double *num_diff = alloc_nan(15 * N3), *DDIV = alloc_nan(N3);
  double *Dt[5];
  const double *tp[5] = {st->DENS_tp, st->MOMZ_tp, st->MOMX_tp, st->MOMY_tp, st->RHOT_tp};
  memcpy(DENS00, st->DENS, N3 * sizeof(double));
  for (size_t q = 0; q < 3 * N3; ++q) mflx_av[q] = 0.0;
  /* RT2P / DPRES0 / REF_rhot (LARGE.F90:443-471), dry form (QV = 0) */
  const double CVdry = CPdry - Rdry;
  const double CPovCV = CPdry / CVdry;
  OMP_FOR
  for (int j = 0; j < JA; ++j)
    for (int i = 0; i < IA; ++i) {
      for (int k = KS; k <= KE; ++k) {
        size_t p = I3(k, i, j);
        double PRES = P00 * pow(Rdry * st->RHOT[p] / P00, CPovCV);
        RT2P[p] = CPovCV * PRES / st->RHOT[p];
        DPRES0[p] = PRES - st->REF_pres[p];
        REF_rhot[p] = st->RHOT[p];
      }
      DPRES0[I3(KS - 1, i, j)] = DPRES0[I3(KS + 1, i, j)] + (st->REF_pres[I3(KS + 1, i, j)] - st->REF_pres[I3(KS - 1, i, j)]);
      DPRES0[I3(KE + 1, i, j)] = DPRES0[I3(KE - 1, i, j)] + (st->REF_pres[I3(KE - 1, i, j)] - st->REF_pres[I3(KE + 1, i, j)]);
    }
  /* tendencies with DAMP_alpha = 0: X_t = X_tp + 0 (LARGE.F90:614-839), zero ghost levels, halo */
  for (int v = 0; v < 5; ++v) {
    Dt[v] = alloc_nan(N3);
    OMP_FOR
    for (int j = JS; j <= JE; ++j)
      for (int i = IS; i <= IE; ++i)
        for (int k = 0; k < KA; ++k) {
          const int lo = v == 1 ? KS : KS, hi = v == 1 ? KE - 1 : KE;
          Dt[v][I3(k, i, j)] = (k >= lo && k <= hi) ? tp[v][I3(k, i, j)] + 0.0 : 0.0;
        }
    orc_halo_periodic(g, Dt[v]);
  }
  int rc = 0;
  for (int step = 0; step < nstep; ++step) {
    if (c->nd_coef == 0.0) {
      for (size_t q = 0; q < 15 * N3; ++q) num_diff[q] = 0.0;
    } else {
      rc |= orc_numfilter_coef(g, num_diff, st->DENS, st->MOMZ, st->MOMX, st->MOMY, st->RHOT, m, dts,
                               st->REF_dens, st->REF_pott, c->nd_coef, c->nd_order, c->nd_sfc_fact, c->nd_use_rs);
    }
    if (c->divdmp_coef > 0.0) orc_divergence(g, DDIV, st->MOMZ, st->MOMX, st->MOMY, m);
    else for (size_t q = 0; q < N3; ++q) DDIV[q] = 0.0;
    if (c->tinteg_short == ORC_TINTEG_RK4)
      rc |= orc_tinteg_short_rk4(g, c->scheme, c->short_type == 1, st->DENS, st->MOMZ, st->MOMX, st->MOMY, st->RHOT, st->mflx_hi, st->tflx_hi,
                                 Dt[0], Dt[1], Dt[2], Dt[3], Dt[4], DPRES0, RT2P, st->CORIOLI, num_diff, st->wdamp_coef,
                                 c->divdmp_coef, DDIV, m, st->REF_dens, REF_rhot, dts);
    else
      rc |= orc_tinteg_short_rk3(g, c->scheme, c->short_type == 1, c->tinteg_short == ORC_TINTEG_RK3WS2002,
                                 st->DENS, st->MOMZ, st->MOMX, st->MOMY, st->RHOT, st->mflx_hi, st->tflx_hi,
                                 Dt[0], Dt[1], Dt[2], Dt[3], Dt[4], DPRES0, RT2P, st->CORIOLI, num_diff, st->wdamp_coef,
                                 c->divdmp_coef, DDIV, m, st->REF_dens, REF_rhot, dts);
    /* ghost levels (LARGE.F90:908-927) */
    OMP_FOR
    for (int j = JS; j <= JE; ++j)
      for (int i = IS; i <= IE; ++i) {
        for (int k = 0; k <= KS - 1; ++k) {
          st->DENS[I3(k, i, j)] = st->DENS[I3(KS, i, j)];
          st->MOMZ[I3(k, i, j)] = 0.0;
          st->MOMX[I3(k, i, j)] = st->MOMX[I3(KS, i, j)];
          st->MOMY[I3(k, i, j)] = st->MOMY[I3(KS, i, j)];
          st->RHOT[I3(k, i, j)] = st->RHOT[I3(KS, i, j)];
        }
        for (int k = KE + 1; k < KA; ++k) {
          st->DENS[I3(k, i, j)] = st->DENS[I3(KE, i, j)];
          st->MOMZ[I3(k, i, j)] = 0.0;
          st->MOMX[I3(k, i, j)] = st->MOMX[I3(KE, i, j)];
          st->MOMY[I3(k, i, j)] = st->MOMY[I3(KE, i, j)];
          st->RHOT[I3(k, i, j)] = st->RHOT[I3(KE, i, j)];
        }
      }
    double *pv[5] = {st->DENS, st->MOMZ, st->MOMX, st->MOMY, st->RHOT};
    for (int v = 0; v < 5; ++v) orc_halo_periodic(g, pv[v]);
    for (int n = 0; n < 3; ++n)
      for (int j = JS; j <= JE; ++j)
        for (int i = IS; i <= IE; ++i)
          for (int k = KS; k <= KE; ++k) mflx_av[I3(k, i, j) + N3 * n] = mflx_av[I3(k, i, j) + N3 * n] + st->mflx_hi[I3(k, i, j) + N3 * n];
  }
  /* tracers (LARGE.F90:1016-1125) */
  if (st->QA > 0) {
    for (int n = 0; n < 3; ++n) {
      for (int j = JS; j <= JE; ++j)
        for (int i = IS; i <= IE; ++i)
          for (int k = KS; k <= KE; ++k) st->mflx_hi[I3(k, i, j) + N3 * n] = mflx_av[I3(k, i, j) + N3 * n] / nstep;
      orc_halo_periodic(g, st->mflx_hi + N3 * n);
    }
    double *ndq = alloc_nan(3 * N3);
    for (size_t q = 0; q < 3 * N3; ++q) ndq[q] = 0.0; /* ND_COEF_Q = 0 (driver default): zeros (LARGE.F90:1053-1054) */
    const double *MAPF_XY = m->MAPF; /* MAPF(:,:,:,I_XY) */
    for (int iq = 0; iq < st->QA; ++iq) {
      double *q = st->QTRC + N3 * iq;
      /* passive tracers (none is QV): LARGE.F90:1049-1061 with DENS00 */
      if (c->nd_coef_q != 0.0)
        rc |= orc_numfilter_coef_q(g, ndq, DENS00, q, 0, m, dtl, NULL, c->nd_coef_q, c->nd_order, c->nd_sfc_fact, c->nd_use_rs);
      if (c->tinteg_tracer == ORC_TINTEG_TRACER_EULER)
        rc |= orc_tinteg_tracer_euler(g, c->scheme_tracer, q, q, st->RHOQ_tp + N3 * iq, DENS00, st->DENS, st->mflx_hi, ndq,
                                      m->GSQRT, MAPF_XY, m->CDZ, m->RCDZ, m->RCDX, m->RCDY, dtl,
                                      c->flag_fct_tracer, c->flag_fct_along_stream);
      else
        rc |= orc_tinteg_tracer_rk3(g, c->scheme_tracer, q, q, st->RHOQ_tp + N3 * iq, DENS00, st->DENS, st->mflx_hi, ndq,
                                    m->GSQRT, MAPF_XY, m->CDZ, m->RCDZ, m->RCDX, m->RCDY, dtl,
                                    c->flag_fct_tracer, c->flag_fct_along_stream);
      orc_halo_periodic(g, q);
    }
    free(ndq);
  }
  for (int v = 0; v < 5; ++v) free(Dt[v]);
  free(RT2P); free(DPRES0); free(REF_rhot); free(DENS00); free(mflx_av); free(num_diff); free(DDIV);
  return rc;
}

/* ---- terrain-following metrics (REAL:391-545, METRIC:200-345) ------------- *
 * REAL = scalelib/src/atmosphere/grid/scale_atmos_grid_cartesC_real.F90,
 * METRIC = scalelib/src/atmosphere/grid/scale_atmos_grid_cartesC_metric.F90.
 * Levels are 0-based; face index kf = 0..KA is the reference's FZ(kf), the
 * top face of level kf-1. */
void orc_terrain_metrics(const orc_grid *g, const double *Zsfc, const double *CZ, const double *FZ,
    const double *RCDZ, const double *RFDZ, const double *RCDX, const double *RFDX,
    const double *RCDY, const double *RFDY,
    double *GSQRT, double *J13G, double *J23G, double *PHI) {
  GRID_LOCALS;
  const size_t KF = (size_t)KA + 1;
  const double Htop = FZ[KE + 1] - FZ[KS]; /* FZ(KE) - FZ(KS-1) (REAL:410) */
#define ZS(i, j) Zsfc[(size_t)(i) + (size_t)IA * (size_t)(j)]
  /* surface heights at the x, u, v and uv points (REAL:413-455, 478-545) */
  double *zx = malloc(sizeof(double) * IA * JA), *zu = malloc(sizeof(double) * IA * JA);
  double *zv = malloc(sizeof(double) * IA * JA), *zuv = malloc(sizeof(double) * IA * JA);
#define P2(i, j) ((size_t)(i) + (size_t)IA * (size_t)(j))
  for (int j = 0; j < JA; ++j)
    for (int i = 0; i < IA; ++i) {
      zx[P2(i, j)] = ZS(i, j);
      zu[P2(i, j)] = i < IA - 1 ? (ZS(i, j) + ZS(i + 1, j)) * 0.5 : ZS(IA - 1, j);
      zv[P2(i, j)] = j < JA - 1 ? (ZS(i, j) + ZS(i, j + 1)) * 0.5 : ZS(i, JA - 1);
      if (i < IA - 1 && j < JA - 1) zuv[P2(i, j)] = (ZS(i, j) + ZS(i + 1, j) + ZS(i, j + 1) + ZS(i + 1, j + 1)) * 0.25;
      else if (j < JA - 1) zuv[P2(i, j)] = (ZS(IA - 1, j) + ZS(IA - 1, j + 1)) * 0.5;
      else if (i < IA - 1) zuv[P2(i, j)] = (ZS(i, JA - 1) + ZS(i + 1, JA - 1)) * 0.5;
      else zuv[P2(i, j)] = ZS(IA - 1, JA - 1);
    }
  /* Gal-Chen heights ( Htop - Zs ) / Htop * Z + Zs of the cell centres and faces */
  double *cz[4], *fz[4];
  const double *zs[4] = {zx, zu, zv, zuv};
  for (int s = 0; s < 4; ++s) {
    cz[s] = malloc(sizeof(double) * N3);
    fz[s] = malloc(sizeof(double) * KF * IA * JA);
    for (int j = 0; j < JA; ++j)
      for (int i = 0; i < IA; ++i) {
        const double Zs = zs[s][P2(i, j)];
        for (int k = 0; k < KA; ++k) cz[s][I3(k, i, j)] = (Htop - Zs) / Htop * CZ[k] + Zs;
        for (int kf = 0; kf <= KA; ++kf) fz[s][(size_t)kf + KF * P2(i, j)] = (Htop - Zs) / Htop * FZ[kf] + Zs;
      }
  }
  enum { X_ = 0, U_ = 1, V_ = 2, UV_ = 3 };
#define CZS(s, k, i, j) cz[s][I3(k, i, j)]
#define FZS(s, kf, i, j) fz[s][(size_t)(kf) + KF * P2(i, j)]
#define GS(k, i, j, c) GSQRT[I3(k, i, j) + N3 * (c)]
#define J13(k, i, j, c) J13G[I3(k, i, j) + N3 * (c)]
#define J23(k, i, j, c) J23G[I3(k, i, j) + N3 * (c)]
  /* G^1/2 (METRIC:231-279) */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i) {
      for (int k = 0; k < KA; ++k) {
        GS(k, i, j, I_XYZ) = (FZS(X_, k + 1, i, j) - FZS(X_, k, i, j)) * RCDZ[k];
        GS(k, i, j, I_UYZ) = (FZS(U_, k + 1, i, j) - FZS(U_, k, i, j)) * RCDZ[k];
        GS(k, i, j, I_XVZ) = (FZS(V_, k + 1, i, j) - FZS(V_, k, i, j)) * RCDZ[k];
        GS(k, i, j, I_UVZ) = (FZS(UV_, k + 1, i, j) - FZS(UV_, k, i, j)) * RCDZ[k];
      }
      for (int k = 0; k < KA - 1; ++k) {
        GS(k, i, j, I_XYW) = (CZS(X_, k + 1, i, j) - CZS(X_, k, i, j)) * RFDZ[k];
        GS(k, i, j, I_UYW) = (CZS(U_, k + 1, i, j) - CZS(U_, k, i, j)) * RFDZ[k];
        GS(k, i, j, I_XVW) = (CZS(V_, k + 1, i, j) - CZS(V_, k, i, j)) * RFDZ[k];
      }
      GS(KA - 1, i, j, I_XYW) = GS(KA - 2, i, j, I_XYW);
      GS(KA - 1, i, j, I_UYW) = GS(KA - 2, i, j, I_UYW);
      GS(KA - 1, i, j, I_XVW) = GS(KA - 2, i, j, I_XVW);
    }
  /* J13G, J23G (METRIC:297-323) */
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = 0; k < KA; ++k) {
        J13(k, i, j, I_XYZ) = -(CZS(U_, k, i, j) - CZS(U_, k, i - 1, j)) * RCDX[i];
        J13(k, i, j, I_XYW) = -(FZS(U_, k + 1, i, j) - FZS(U_, k + 1, i - 1, j)) * RCDX[i];
        J13(k, i, j, I_UYW) = -(FZS(X_, k + 1, i + 1, j) - FZS(X_, k + 1, i, j)) * RFDX[i];
        J13(k, i, j, I_XVW) = -(FZS(UV_, k + 1, i, j) - FZS(UV_, k + 1, i - 1, j)) * RCDX[i];
        J13(k, i, j, I_UYZ) = -(CZS(X_, k, i + 1, j) - CZS(X_, k, i, j)) * RFDX[i];
        J13(k, i, j, I_XVZ) = -(CZS(UV_, k, i, j) - CZS(UV_, k, i - 1, j)) * RCDX[i];
        J13(k, i, j, I_UVZ) = -(CZS(V_, k, i + 1, j) - CZS(V_, k, i, j)) * RFDX[i];
        J23(k, i, j, I_XYZ) = -(CZS(V_, k, i, j) - CZS(V_, k, i, j - 1)) * RCDY[j];
        J23(k, i, j, I_XYW) = -(FZS(V_, k + 1, i, j) - FZS(V_, k + 1, i, j - 1)) * RCDY[j];
        J23(k, i, j, I_XVW) = -(FZS(X_, k + 1, i, j + 1) - FZS(X_, k + 1, i, j)) * RFDY[j];
        J23(k, i, j, I_UYW) = -(FZS(UV_, k + 1, i, j) - FZS(UV_, k + 1, i, j - 1)) * RCDY[j];
        J23(k, i, j, I_XVZ) = -(CZS(X_, k, i, j + 1) - CZS(X_, k, i, j)) * RFDY[j];
        J23(k, i, j, I_UYZ) = -(CZS(UV_, k, i, j) - CZS(UV_, k, i, j - 1)) * RCDY[j];
        J23(k, i, j, I_UVZ) = -(CZS(U_, k, i, j + 1) - CZS(U_, k, i, j)) * RFDY[j];
      }
  for (int c = 0; c < 7; ++c) {
    orc_halo_periodic(g, GSQRT + N3 * c);
    orc_halo_periodic(g, J13G + N3 * c);
    orc_halo_periodic(g, J23G + N3 * c);
  }
  /* PHI = GRAV * CZ(real) (REAL:549) */
  for (size_t q = 0; q < N3; ++q) PHI[q] = GRAV * cz[X_][q];
  for (int s = 0; s < 4; ++s) { free(cz[s]); free(fz[s]); }
  free(zx); free(zu); free(zv); free(zuv);
#undef ZS
#undef P2
#undef CZS
#undef FZS
#undef GS
#undef J13
#undef J23
}

/* ---- FP64 budget ---------------------------------------------------------- */
double orc_mass_sum(const orc_grid *g, const orc_metric *m, const double *a) {
  GRID_LOCALS;
  double s = 0.0;
  for (int j = JS; j <= JE; ++j)
    for (int i = IS; i <= IE; ++i)
      for (int k = KS; k <= KE; ++k)
        s += a[I3(k, i, j)] * m->CDZ[k] * m->CDX[i] * m->CDY[j] * m->GSQRT[I3(k, i, j)]
           / (m->MAPF[(size_t)i + (size_t)IA * j] * m->MAPF[(size_t)i + (size_t)IA * (j + JA)]);
  return s;
}
