/*
 * scale_oracle.h -- CPU restatement of SCALE-RM's HEVI short step, FVM flux
 * stencils, numerical filter and FCT tracer advection.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path (scale_amd, libscalehip.so)
 * never links, loads or calls it.
 *
 * Layout: every 3-D array is the reference's Fortran array (KA,IA,JA) with k
 * fastest, addressed here with 0-based indices: element (k,i,j) lives at
 * k + KA*(i + IA*j).  Fortran index kF = kC + 1, so KS = KHALO (=2),
 * KE = KS+KMAX-1, and likewise for i and j.  1-D arrays (CDZ(KA), FDZ(KA-1),
 * ...) shift the same way: CDZ(kF) = cdz[kC].  Multi-component arrays keep
 * the Fortran component order: GSQRT(KA,IA,JA,7) component c (0-based) at
 * offset c*KA*IA*JA; MAPF(IA,JA,2,4) entry (i,j,n,c) at
 * i + IA*(j + JA*(n + 2*c)); num_diff(KA,IA,JA,5,3) entry (k,i,j,v,d) at
 * IX3 + N3*(v + 5*d).
 *
 * Arithmetic follows the reference expression by expression (same grouping,
 * left-to-right evaluation, sign() as copysign) and is compiled with
 * -ffp-contract=off so that no fused multiply-add changes a rounding.
 *
 * Parity status of each function is stated in the header comment of its
 * definition in scale_oracle.c and in the repository's design notes.
 */
#ifndef SCALE_ORACLE_H
#define SCALE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* flux scheme identifiers (the order of the upwind/central scheme) */
#define ORC_UD1 1
#define ORC_CD2 2 /* tracer scheme: second-order central (cd2.F90) */
#define ORC_UD3 3
#define ORC_CD4 4
#define ORC_UD5 5 /* tracer scheme: fifth-order upwind (ud5.F90), needs IHALO = JHALO = 3 */
#define ORC_CD6 6 /* tracer scheme: sixth-order central (cd6.F90), needs IHALO = JHALO = 3 */
#define ORC_UD3KOREN 31 /* UD3KOREN1993: tracer scheme, UD3 with the Koren limiter */

typedef struct {
  int KA, IA, JA;
  int KS, KE, IS, IE, JS, JE; /* 0-based, inclusive */
} orc_grid;

typedef struct {
  const double *CDZ, *FDZ, *RCDZ, *RFDZ; /* (KA), (KA-1), (KA), (KA-1) */
  const double *CDX, *FDX, *RCDX, *RFDX; /* (IA), (IA-1), (IA), (IA-1) */
  const double *CDY, *FDY, *RCDY, *RFDY; /* (JA), (JA-1), (JA), (JA-1) */
  const double *PHI;                     /* (KA,IA,JA)   geopotential   */
  const double *GSQRT, *J13G, *J23G;     /* (KA,IA,JA,7)                */
  const double *MAPF;                    /* (IA,JA,2,4)                 */
  double J33G;
} orc_metric;

/* ATMOS_DYN_Tstep_short_fvm_hevi (HEVI.F90:128-1089), periodic lateral
 * boundaries only (BND_* false).  Returns 0, or -1 for unsupported options. */
int orc_tstep_short_hevi(const orc_grid *g, int scheme,
    double *DENS_RK, double *MOMZ_RK, double *MOMX_RK, double *MOMY_RK, double *RHOT_RK,
    double *mflx_hi, double *tflx_hi,
    const double *DENS0, const double *MOMZ0, const double *MOMX0, const double *MOMY0, const double *RHOT0,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef, double divdmp_coef, const double *DDIV,
    const orc_metric *m, const double *REF_dens, const double *REF_rhot,
    double dtrk);

/* ATMOS_DYN_Tstep_short_fvm_heve (scale_atmos_dyn_tstep_short_fvm_heve.F90:109-1516)
 * with the momentum FCT (flag_fct_momentum, :386-450, 716-810, 958-1057,
 * 1207-1309) and the theta FCT (flag_fct_t, :1388-1475), periodic lateral
 * boundaries; -1 for a scheme other than UD3/CD4. */
int orc_tstep_short_heve(const orc_grid *g, int scheme,
    double *DENS_RK, double *MOMZ_RK, double *MOMX_RK, double *MOMY_RK, double *RHOT_RK,
    double *mflx_hi, double *tflx_hi,
    const double *DENS0, const double *MOMZ0, const double *MOMX0, const double *MOMY0, const double *RHOT0,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef, double divdmp_coef, const double *DDIV,
    int flag_fct_momentum, int flag_fct_t, int flag_fct_along_stream,
    const orc_metric *m, const double *REF_dens, const double *REF_rhot, double dtrk);

/* ATMOS_DYN_numfilter_coef (COMMON.F90:385-985), ND_ORDER = 1 only,
 * followed by the periodic halo fill of the 15 num_diff fields. */
int orc_numfilter_coef(const orc_grid *g, double *num_diff,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const orc_metric *m, double dt, const double *REF_dens, const double *REF_pott,
    double nd_coef, int nd_order, double nd_sfc_fact, int nd_use_rs);

/* ATMOS_DYN_numfilter_coef_q (COMMON.F90:989-1200), ND_ORDER = 1, with the
 * halo fill of num_diff_q(KA,IA,JA,3).  is_qv selects QTRC - REF_qv (which
 * needs ND_USE_RS); returns -1 for unsupported options. */
int orc_numfilter_coef_q(const orc_grid *g, double *num_diff_q, const double *DENS, const double *QTRC, int is_qv,
    const orc_metric *m, double dt, const double *REF_qv, double nd_coef, int nd_order, double nd_sfc_fact, int nd_use_rs);

/* ATMOS_DYN_divergence (COMMON.F90:1437-1515) */
void orc_divergence(const orc_grid *g, double *DDIV,
    const double *MOMZ, const double *MOMX, const double *MOMY, const orc_metric *m);

/* ATMOS_DYN_wdamp_setup (COMMON.F90:323-381); FZ has KA+1 entries FZ(0:KA) */
void orc_wdamp_setup(const orc_grid *g, double *wdamp_coef, double wdamp_tau,
    double wdamp_height, const double *FZ);

/* ATMOS_DYN_Tstep_tracer_fvm_heve (tracer_fvm_heve.F90:72-242);
 * MAPF is the I_XY slice (IA,JA,2).  QTRCo may alias QTRC0. */
int orc_tstep_tracer_heve(const orc_grid *g, int scheme,
    double *QTRCo, const double *QTRC, const double *QTRC0, const double *RHOQ_t,
    const double *DENS0, const double *DENS, const double *mflx_hi, const double *num_diff,
    const double *GSQRT, const double *MAPF_XY, const double *CDZ,
    const double *RCDZ, const double *RCDX, const double *RCDY,
    double dtl, int flag_fct, int flag_fct_along_stream);

/* ATMOS_DYN_fct (COMMON.F90:1922-2671) incl. its R+/R- halo exchange. */
void orc_fct(const orc_grid *g, double *qflx_anti,
    const double *phi_in, const double *DENS0, const double *DENS,
    const double *qflx_hi, const double *qflx_lo, const double *mflx_hi,
    const double *RDZ, const double *RDX, const double *RDY,
    const double *GSQRT_XYZ, const double *MAPF_XY, double dt, int flag_vect);

/* test hook of the frozen-limiter checks (both NULL by default): orc_fct
 * copies R+ then R- (2 x KA*IA*JA, before their halo fill) out to
 * orc_fct_save_r and replaces them by orc_fct_fixed_r when these are set */
extern double *orc_fct_save_r;
extern const double *orc_fct_fixed_r;

/* get_fact_fct (COMMON.F90:2676-2734): fact(0:1,-1:1,-1:1) returned as
 * fact[a + 2*((b+1) + 3*(c+1))] for fact(a,b,c). */
void orc_get_fact_fct(double *fact, double rw, double ru, double rv);

/* periodic COMM_vars8 on one rank (COMM.F90:345-397,1600-2043): all KA
 * levels of the IHALO/JHALO-wide lateral halo, faces and corners. */
void orc_halo_periodic(const orc_grid *g, double *a);

/* Thomas algorithm of HEVI.F90:1092-1127 on KMAX-1 unknowns.  F1/F2/F3 are
 * indexed like the reference (KA entries, rows KS..KE-1); C has KMAX-1. */
void orc_solve_direct(const orc_grid *g, double *C, const double *F1, const double *F2, const double *F3);

/* ---- whole-model state for multi-step integrations ---------------------- */

typedef struct {
  /* prognostics (KA,IA,JA) with valid halos */
  double *DENS, *MOMZ, *MOMX, *MOMY, *RHOT;
  double *QTRC;        /* (KA,IA,JA,QA) */
  int QA;
  /* reference state (KA,IA,JA) incl. ghost levels and halos */
  const double *REF_dens, *REF_pott, *REF_pres;
  /* physics tendencies (zero in every workload here but kept as inputs) */
  const double *DENS_tp, *MOMZ_tp, *MOMX_tp, *MOMY_tp, *RHOT_tp, *RHOQ_tp;
  const double *CORIOLI;    /* (1,IA,JA) */
  const double *wdamp_coef; /* (KA) */
  /* outputs of the last large step */
  double *mflx_hi, *tflx_hi; /* (KA,IA,JA,3) */
} orc_state;

typedef struct {
  int scheme;          /* dynamics flux scheme (ORC_UD3 or ORC_CD4) */
  int scheme_tracer;   /* tracer flux scheme */
  int tinteg_short;    /* ATMOS_DYN_TINTEG_SHORT_TYPE: ORC_TINTEG_RK4, _RK3 (Heun) or _RK3WS2002 */
  double dtl;          /* large step */
  double dts;          /* short step (dtl/dts rounded up as in LARGE.F90:379-382) */
  double nd_coef, nd_sfc_fact, nd_coef_q;
  int nd_order, nd_use_rs;
  double divdmp_coef;
  int flag_fct_tracer, flag_fct_along_stream;
  int short_type;      /* ATMOS_DYN_TYPE: 0 = FVM-HEVI, 1 = FVM-HEVE */
  int tinteg_tracer;   /* ATMOS_DYN_TINTEG_TRACER_TYPE: ORC_TINTEG_TRACER_RK3WS2002 or _EULER */
} orc_config;

/* Terrain-following metrics from the surface height Zsfc(IA,JA) (periodic
 * halo filled): the Gal-Chen heights of ATMOS_GRID_CARTESC_REAL_calc_Z
 * (scalelib/src/atmosphere/grid/scale_atmos_grid_cartesC_real.F90:391-545) at
 * the eight staggered positions, then G^1/2, J13G and J23G of
 * ATMOS_GRID_CARTESC_METRIC_terrainfollowing
 * (scale_atmos_grid_cartesC_metric.F90:200-345) on the interior columns
 * with their periodic halos, and PHI = GRAV * CZ(real).  CZ(KA) and FZ(0:KA)
 * (KA+1 values) are the flat grid's heights; RCDZ..RFDY as in orc_metric. */
void orc_terrain_metrics(const orc_grid *g, const double *Zsfc, const double *CZ, const double *FZ,
    const double *RCDZ, const double *RFDZ, const double *RCDX, const double *RFDX,
    const double *RCDY, const double *RFDY,
    double *GSQRT, double *J13G, double *J23G, double *PHI);

/* time integrators (scale_atmos_dyn_tinteg_short.F90, _tinteg_tracer.F90) */
enum { ORC_TINTEG_RK3 = 3, ORC_TINTEG_RK4 = 4, ORC_TINTEG_RK3WS2002 = 32 };
enum { ORC_TINTEG_TRACER_RK3WS2002 = 0, ORC_TINTEG_TRACER_EULER = 1 };

/* ATMOS_DYN_Tstep_large_fvm_heve on-path subset (LARGE.F90:377-1127) for a
 * dry run with QA passive tracers (moist constants with QV=0 reduce to the
 * dry ones): prep of RT2P/DPRES0/REF_rhot, tendencies (DAMP_alpha=0),
 * numfilter, divergence, RK4 short steps, ghost-level fill, halo fill,
 * mflx average and the tracer RK3 loop.  Advances st by one large step. */
int orc_large_step(const orc_grid *g, const orc_metric *m, const orc_config *c, orc_state *st);

/* One full RK4 short step (tinteg_short_rk4.F90:260-660) on caller buffers;
 * heve = 1 runs orc_tstep_short_heve as the stage instead of the HEVI one. */
int orc_tinteg_short_rk4(const orc_grid *g, int scheme, int heve,
    double *DENS, double *MOMZ, double *MOMX, double *MOMY, double *RHOT,
    double *mflx_hi, double *tflx_hi,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef, double divdmp_coef, const double *DDIV,
    const orc_metric *m, const double *REF_dens, const double *REF_rhot, double dt);

/* One RK3 short step (tinteg_short_rk3.F90:188-576) on caller buffers:
 * ws2002 = 0 is Heun's method (stages dt/3, 2dt/3, dt, then the combination
 * (3 r1 + 3 r3 - 2 phi_n)/4), ws2002 = 1 Wicher and Skamarock (2002)
 * (dt/3, dt/2, dt; phi_{n+1} = r3). */
int orc_tinteg_short_rk3(const orc_grid *g, int scheme, int heve, int ws2002,
    double *DENS, double *MOMZ, double *MOMX, double *MOMY, double *RHOT,
    double *mflx_hi, double *tflx_hi,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef, double divdmp_coef, const double *DDIV,
    const orc_metric *m, const double *REF_dens, const double *REF_rhot, double dt);

/* tinteg_tracer_euler.F90:80-133: one tracer stage of dtl into QTRC */
int orc_tinteg_tracer_euler(const orc_grid *g, int scheme,
    double *QTRC, const double *QTRC0, const double *RHOQ_t,
    const double *DENS0, const double *DENS, const double *mflx_hi, const double *num_diff,
    const double *GSQRT, const double *MAPF_XY, const double *CDZ,
    const double *RCDZ, const double *RCDX, const double *RCDY,
    double dtl, int flag_fct, int flag_fct_along_stream);

/* tinteg_tracer_rk3.F90:92-208 (RK3WS2002) */
int orc_tinteg_tracer_rk3(const orc_grid *g, int scheme,
    double *QTRC, const double *QTRC0, const double *RHOQ_t,
    const double *DENS0, const double *DENS, const double *mflx_hi, const double *num_diff,
    const double *GSQRT, const double *MAPF_XY, const double *CDZ,
    const double *RCDZ, const double *RCDX, const double *RCDY,
    double dtl, int flag_fct, int flag_fct_along_stream);

/* FP64 budgets over the interior: sum(a(k,i,j) * CDZ*CDX*CDY * GSQRT_XYZ / (MAPF1*MAPF2)),
 * summed k-fastest then i then j in one fixed order (no threading). */
double orc_mass_sum(const orc_grid *g, const orc_metric *m, const double *a);

/* face value of the flux families: kind 0 interior (val(idx-1..idx+2)),
 * 1 lowest face (val(KS..KS+2) in q0..q2), 2 highest face (val(KE-2..KE) in qm1..q1) */
double orc_face_value(int scheme, int kind, double qm1, double q0, double q1, double q2, double vel);
/* six-point face values of UD5/CD6: kind 0 interior face between q[2] and q[3]
 * (q = val(k-2..k+3)); kind 1 the face KS, 3 the face KS+1 (q = val(KS..KS+5));
 * kind 2 the face KE-1, 4 the face KE-2 (q = val(KE-5..KE)). */
double orc_face_value6(int s, int kind, const double *q, double vel);

/* ---- adjoint checker (scale_oracle_adj.c) -------------------------------- */

/* Intermediates of one orc_tstep_short_hevi call, recorded when a trace is
 * set: POTT and St (3-D), and per interior column the face potential
 * temperature PT(KS..KE-1), the bands F1/F2/F3 and the solved vertical
 * momentum C (rows KS..KE-1, stored at k). */
typedef struct {
  double *POTT, *St, *PT, *F1, *F2, *F3, *Cp;
} orc_hevi_trace;
void orc_set_hevi_trace(orc_hevi_trace *t);

/* filter_setup tables CNZ3(3,KA,2), CNX3(3,IA,2), CNY3(3,JA,2) (COMMON.F90:92-319) */
void orc_filter_tables(const orc_grid *g, const orc_metric *m, double *cnz3, double *cnx3, double *cny3);
/* calc_diff3 (COMMON.F90:1652-1800) on one field; diff has 3 components */
void orc_calc_diff3(const orc_grid *g, const orc_metric *m, double *diff, const double *phi, int KO, int IO, int JO);

/* Adjoint of orc_tstep_short_hevi for flat terrain, UD3/CD4 and
 * divdmp_coef = 0.  Given the stage's inputs and the adjoints of its five
 * prognostic outputs (read only, at the points the stage writes), adds the
 * adjoints of its inputs into the *_b arrays (state0, state, num_diff(15),
 * DPRES0, RT2P, REF_rhot).  Returns 0, or -1 for unsupported options. */
int orc_adj_tstep_short_hevi(const orc_grid *g, int scheme, const orc_metric *m,
    const double *DENS0, const double *MOMZ0, const double *MOMX0, const double *MOMY0, const double *RHOT0,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef,
    const double *REF_dens, const double *REF_rhot, double dtrk,
    const double *DENS_RK_b, const double *MOMZ_RK_b, const double *MOMX_RK_b, const double *MOMY_RK_b, const double *RHOT_RK_b,
    double *DENS0_b, double *MOMZ0_b, double *MOMX0_b, double *MOMY0_b, double *RHOT0_b,
    double *DENS_b, double *MOMZ_b, double *MOMX_b, double *MOMY_b, double *RHOT_b,
    double *num_diff_b, double *DPRES0_b, double *RT2P_b, double *REF_rhot_b);

/* Adjoint of orc_numfilter_coef (ND_ORDER = 1, reference state, with its
 * halo fill): given num_diff_b (15 fields, modified: consumed), adds the
 * adjoints of DENS..RHOT. */
int orc_adj_numfilter_coef(const orc_grid *g, const orc_metric *m, double dt,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *REF_dens, const double *REF_pott, double nd_coef, double nd_sfc_fact,
    double *num_diff_b, double *DENS_b, double *MOMZ_b, double *MOMX_b, double *MOMY_b, double *RHOT_b);

/* Adjoint of one orc_large_step without tracers (one short step per large
 * step, flat terrain, divdmp_coef = 0).  st holds the state at the start of
 * the step; xb[0..4] (DENS..RHOT, whole arrays) hold the adjoint of the
 * state after the step on entry and the adjoint of the state before it on
 * return. */
int orc_adj_large_step(const orc_grid *g, const orc_metric *m, const orc_config *c, const orc_state *st,
    double *const xb[5]);

/* adjoint of orc_halo_periodic: halo adjoints are added to their sources and cleared */
void orc_adj_halo_periodic(const orc_grid *g, double *ab);

/* threads used by the OpenMP loops (1 = serial) */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
