/*
 * scale_hip.h -- C ABI of the MI355X (gfx950) HEVI short step and FCT tracer
 * advection for SCALE-RM.
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative code on failure;
 *    scale_hip_last_error() returns a message for the calling thread.
 *  - Array arguments of the drop-in entry points are host pointers to
 *    contiguous Fortran arrays with the reference's shapes, e.g. DENS(KA,IA,JA)
 *    (k fastest), num_diff(KA,IA,JA,5,3), GSQRT(KA,IA,JA,7), MAPF(IA,JA,2,4).
 *  - Scalars are passed by reference, as a Fortran caller passes them, so the
 *    iso_c_binding interface needs no VALUE attributes.  LOGICAL arguments are
 *    int32_t (0 = .false., nonzero = .true.).
 *  - Grid extents are the module variables of
 *    scalelib/src/atmosphere/grid/scale_atmos_grid_cartesC_index.F90:256-265,
 *    given once to scale_hip_init with the reference's 1-based values.
 *  - The library keeps its device buffers in the context; calls are synchronous
 *    at return except the device-resident model calls, which are ordered on
 *    the context's stream (scale_hip_synchronize waits for them).
 *  - Lateral boundaries must be periodic (BND_* = .false.); other values are
 *    rejected with an error.
 */
#ifndef SCALE_HIP_H
#define SCALE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct scale_hip_ctx scale_hip_ctx;

/* flux scheme ids (ATMOS_DYN_FVM_FLUX_TYPE / _TRACER_TYPE) */
#define SCALE_HIP_UD3 3
#define SCALE_HIP_CD4 4
/* tracers only: UD3 with the Koren (1993) limiter, the reference's default
 * ATMOS_DYN_FVM_FLUX_TRACER_TYPE (scale-rm/src/atmos/mod_atmos_dyn_driver.F90:52),
 * scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_fvm_flux_ud3Koren1993.F90 */
#define SCALE_HIP_UD3KOREN1993 31
/* tracers only: first-order upwind and second-order central
 * (scale_atmos_dyn_fvm_flux_ud1.F90, scale_atmos_dyn_fvm_flux_cd2.F90) */
#define SCALE_HIP_UD1 1
#define SCALE_HIP_CD2 2
#define SCALE_HIP_UD5 5 /* tracer only; needs IHALO = JHALO = 3 and KMAX >= 5 */
#define SCALE_HIP_CD6 6 /* tracer only; needs IHALO = JHALO = 3 and KMAX >= 5 */

typedef struct {
  int32_t KA, IA, JA;             /* array extents                           */
  int32_t KS, KE, IS, IE, JS, JE; /* interior bounds, Fortran 1-based         */
  int32_t KHALO, IHALO, JHALO;    /* KHALO = 2; IHALO = JHALO = 2, or 3 for the UD5/CD6 tracer
                                      schemes (filter_setup needs at least 2, COMMON.F90:112) */
  int32_t flux_scheme;            /* dynamics: SCALE_HIP_UD3 or SCALE_HIP_CD4 */
  int32_t flux_scheme_tracer;     /* tracers:  SCALE_HIP_UD3, _CD4, _UD3KOREN1993, _UD1, _CD2, _UD5 or _CD6 */
  int32_t device;                 /* HIP device ordinal                       */
} scale_hip_grid;

/* Creates a context on grid->device (the setup role of ATMOS_DYN_setup,
 * scalelib/src/atmos-rm/dynamics/scale_atmos_dyn.F90:81-255). */
int scale_hip_init(const scale_hip_grid *grid, scale_hip_ctx **ctx);
int scale_hip_finalize(scale_hip_ctx *ctx);
const char *scale_hip_last_error(void);
const char *scale_hip_version(void);
int scale_hip_synchronize(scale_hip_ctx *ctx);

/* One RK stage of the HEVI short step.  Same dummies, order and meaning as
 * the abstract interface `short` of ATMOS_DYN_Tstep_short,
 * scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short.F90:38-128,
 * implemented by ATMOS_DYN_Tstep_short_fvm_hevi,
 * scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short_fvm_hevi.F90:128-1089.
 * PROG_RK/PROG0/PROG are unused (VA = 0 for HEVI) and may be NULL; the FCT
 * flags are unused by HEVI (HEVI.F90:224-226). */
int scale_hip_tstep_short_hevi(scale_hip_ctx *ctx,
    double *DENS_RK, double *MOMZ_RK, double *MOMX_RK, double *MOMY_RK, double *RHOT_RK,
    double *PROG_RK, double *mflx_hi, double *tflx_hi,
    const double *DENS0, const double *MOMZ0, const double *MOMX0, const double *MOMY0, const double *RHOT0,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *PROG0, const double *PROG,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef, const double *divdmp_coef, const double *DDIV,
    const int32_t *FLAG_FCT_MOMENTUM, const int32_t *FLAG_FCT_T, const int32_t *FLAG_FCT_ALONG_STREAM,
    const double *CDZ, const double *FDZ, const double *FDX, const double *FDY,
    const double *RCDZ, const double *RCDX, const double *RCDY, const double *RFDZ, const double *RFDX, const double *RFDY,
    const double *PHI, const double *GSQRT, const double *J13G, const double *J23G, const double *J33G, const double *MAPF,
    const double *REF_dens, const double *REF_rhot,
    const int32_t *BND_W, const int32_t *BND_E, const int32_t *BND_S, const int32_t *BND_N,
    const double *dtrk, const int32_t *last);

/* One RK stage of the fully explicit HEVE short step.  Same dummies as
 * scale_hip_tstep_short_hevi (the interface `short`,
 * scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short.F90:38-128),
 * implemented by ATMOS_DYN_Tstep_short_fvm_heve,
 * scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short_fvm_heve.F90:109-1516.
 * FLAG_FCT_MOMENTUM limits the momentum fluxes with ATMOS_DYN_fct on the
 * start-of-stage velocities (HEVE.F90:386-450, 716-810, 958-1057, 1207-1309;
 * DENS_RK then also gets its periodic halo) and FLAG_FCT_T the theta flux
 * (HEVE.F90:1388-1475; DENS_RK and mflx_hi then get their periodic halos),
 * as the reference's exchanges leave them.  The first call with
 * FLAG_FCT_MOMENTUM allocates ten KA*IA*JA work arrays, kept until
 * scale_hip_finalize. */
int scale_hip_tstep_short_heve(scale_hip_ctx *ctx,
    double *DENS_RK, double *MOMZ_RK, double *MOMX_RK, double *MOMY_RK, double *RHOT_RK,
    double *PROG_RK, double *mflx_hi, double *tflx_hi,
    const double *DENS0, const double *MOMZ0, const double *MOMX0, const double *MOMY0, const double *RHOT0,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *PROG0, const double *PROG,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef, const double *divdmp_coef, const double *DDIV,
    const int32_t *FLAG_FCT_MOMENTUM, const int32_t *FLAG_FCT_T, const int32_t *FLAG_FCT_ALONG_STREAM,
    const double *CDZ, const double *FDZ, const double *FDX, const double *FDY,
    const double *RCDZ, const double *RCDX, const double *RCDY, const double *RFDZ, const double *RFDX, const double *RFDY,
    const double *PHI, const double *GSQRT, const double *J13G, const double *J23G, const double *J33G, const double *MAPF,
    const double *REF_dens, const double *REF_rhot,
    const int32_t *BND_W, const int32_t *BND_E, const int32_t *BND_S, const int32_t *BND_N,
    const double *dtrk, const int32_t *last);

/* One RK stage of flux-form tracer advection with optional FCT.  Same dummies
 * as the abstract interface `step` of ATMOS_DYN_Tstep_tracer,
 * scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_tracer.F90:34-64,
 * implemented by ATMOS_DYN_Tstep_tracer_fvm_heve,
 * scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_tracer_fvm_heve.F90:72-242
 * (limiter ATMOS_DYN_fct, scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90:1922-2734).
 * MAPF is the (IA,JA,2) I_XY slice; QTRCo may alias QTRC0. */
int scale_hip_tstep_tracer_heve(scale_hip_ctx *ctx,
    double *QTRCo, const double *QTRC, const double *QTRC0, const double *RHOQ_t,
    const double *DENS0, const double *DENS, const double *mflx_hi, const double *num_diff,
    const double *GSQRT, const double *MAPF, const double *CDZ,
    const double *RCDZ, const double *RCDX, const double *RCDY,
    const double *dtl, const int32_t *FLAG_FCT_TRACER, const int32_t *FLAG_FCT_ALONG_STREAM);

/* Adjoint of one scale_hip_tstep_tracer_heve call with the FCT limiter frozen.
 * The reference has no adjoint; this is the transpose of the stage's
 * Jacobian in QTRC, QTRC0, RHOQ_t and num_diff (KA,IA,JA,3) with the mass
 * fluxes and densities held fixed and the limiter's ratios R+/R-
 * (scale_atmos_dyn_common.F90:2522-2547) held at their forward values, so the
 * limited anti-diffusive flux of each face is a fixed multiple of hi - lo.
 * Takes the forward's inputs (same meaning and layout; RHOQ_t is not needed)
 * and QTRCo_b, the adjoint of QTRCo, read at the points the stage writes
 * (KS..KE, IS..IE, JS..JE); overwrites QTRC_b, QTRC0_b, RHOQ_t_b and
 * num_diff_b on the whole array (halo points included: the stage reads QTRC
 * and QTRC0 there).  UD3KOREN1993 (a limiter in the face form) is rejected. */
int scale_hip_tstep_tracer_heve_adj(scale_hip_ctx *ctx,
    const double *QTRCo_b, double *QTRC_b, double *QTRC0_b, double *RHOQ_t_b, double *num_diff_b,
    const double *QTRC, const double *QTRC0,
    const double *DENS0, const double *DENS, const double *mflx_hi, const double *num_diff,
    const double *GSQRT, const double *MAPF, const double *CDZ,
    const double *RCDZ, const double *RCDX, const double *RCDY,
    const double *dtl, const int32_t *FLAG_FCT_TRACER, const int32_t *FLAG_FCT_ALONG_STREAM);

/* The tracer stage of scale_hip_tstep_tracer_heve on device memory, with no
 * allocation of any kind.  This entry point has no counterpart in the
 * reference: it is for callers that cannot allocate while they step (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same dummies in the same order, but every array is a device
 * pointer in the library's device layout (element (k,i,j) at
 * i + IA*(j + JA*k), components of mflx_hi/num_diff/GSQRT as consecutive
 * KA*IA*JA planes; MAPF(IA,JA,2) i fastest; RCDZ(KA), RCDX(IA), RCDY(JA));
 * scalars stay host pointers.  QTRCo must not alias QTRC, QTRC0, DENS0 or
 * DENS (rejected).  The FCT limiter's three work arrays live in a fixed
 * device array of the library, so with FLAG_FCT_TRACER the grid is limited
 * to KA*IA*JA <= 4194304 (larger grids are rejected with an error).  The
 * call is ordered on the context's stream and returns without waiting
 * (scale_hip_synchronize waits).  All calls on one GPU share that array:
 * calls that overlap on one GPU (two contexts, two host threads) race. */
int scale_hip_tstep_tracer_heve_noalloc(scale_hip_ctx *ctx,
    double *QTRCo, const double *QTRC, const double *QTRC0, const double *RHOQ_t,
    const double *DENS0, const double *DENS, const double *mflx_hi, const double *num_diff,
    const double *GSQRT, const double *MAPF, const double *CDZ,
    const double *RCDZ, const double *RCDX, const double *RCDY,
    const double *dtl, const int32_t *FLAG_FCT_TRACER, const int32_t *FLAG_FCT_ALONG_STREAM);

/* Hyperdiffusion fluxes num_diff(KA,IA,JA,5,3) including their halo exchange:
 * ATMOS_DYN_numfilter_coef, scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90:385-985
 * (tables of ATMOS_DYN_filter_setup, :92-319).  ND_ORDER = 1 and
 * ND_USE_RS = .true. are supported; other values are rejected. */
int scale_hip_numfilter_coef(scale_hip_ctx *ctx, double *num_diff,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *CDZ, const double *CDX, const double *CDY, const double *FDZ, const double *FDX, const double *FDY,
    const double *DT, const double *REF_dens, const double *REF_pott,
    const double *ND_COEF, const int32_t *ND_ORDER, const double *ND_SFC_FACT, const int32_t *ND_USE_RS);

/* Hyperdiffusion fluxes of one tracer num_diff_q(KA,IA,JA,3) including their
 * halo exchange: ATMOS_DYN_numfilter_coef_q,
 * scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90:989-1200, same
 * dummies.  ND_ORDER = 1; QV (is_qv) needs ND_USE_RS = .true.  The filter
 * tables come from the grid given to an earlier scale_hip_numfilter_coef call
 * or to the model. */
int scale_hip_numfilter_coef_q(scale_hip_ctx *ctx, double *num_diff_q,
    const double *DENS, const double *QTRC, const int32_t *is_qv,
    const double *CDZ, const double *CDX, const double *CDY, const double *dt,
    const double *REF_qv, const int32_t *iq,
    const double *ND_COEF, const int32_t *ND_ORDER, const double *ND_SFC_FACT, const int32_t *ND_USE_RS);

/* Rayleigh damping coefficient wdamp_coef(KA) from the face heights FZ(0:KA)
 * (KA+1 values), the damping time scale and the height above which it acts:
 * ATMOS_DYN_wdamp_setup, scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_common.F90:323-381.
 * Host computation (no device call); grid supplies KA, KS, KE. */
int scale_hip_wdamp_setup(const scale_hip_grid *grid, double *wdamp_coef, const double *wdamp_tau,
                          const double *wdamp_height, const double *FZ);

/* ---- device-resident integration --------------------------------------------
 * The state stays in HBM between calls; host arrays cross PCIe only in
 * scale_hip_model_set / _get.  One large step reproduces the on-path subset of
 * ATMOS_DYN_Tstep_large_fvm_heve,
 * scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_large_fvm_heve.F90:377-1127:
 * RT2P/DPRES0/REF_rhot, tendencies (DAMP_alpha = 0), numerical filter, RK4
 * short steps (tinteg_short_rk4.F90:260-660), ghost levels, halos, the mass
 * flux average and the RK3 tracer loop (tinteg_tracer_rk3.F90:92-208). */
typedef struct {
  double dtl;          /* TIME_DT_ATMOS_DYN * large/short ratio: large step [s] */
  double dts;          /* short step [s]; nstep = ceiling((dtl-eps)/dts)       */
  double nd_coef;      /* ATMOS_DYN_NUMERICAL_DIFF_COEF (1e-4 by default)      */
  double nd_sfc_fact;  /* ATMOS_DYN_NUMERICAL_DIFF_sfc_fact (1.0)              */
  double divdmp_coef;  /* ATMOS_DYN_divdmp_coef (0 by default; >= 0; FP64)    */
  int32_t nd_order;    /* ATMOS_DYN_NUMERICAL_DIFF_order (1)                   */
  int32_t fct_tracer;  /* ATMOS_DYN_FLAG_FCT_TRACER                            */
  int32_t fct_along_stream; /* ATMOS_DYN_FLAG_FCT_along_stream (1)             */
  int32_t mixed_precision;  /* 0: FP64 throughout. 1: prognostics, fluxes and
                               tendencies in FP32 (the reference's SINGLE build),
                               with the reference state, the pressure terms, the
                               vertical solve and the budgets in FP64; flat
                               terrain only.  Tracers advance in FP64 on the
                               FP64 copies of the end-of-step density and of
                               the averaged mass fluxes                       */
  double nd_coef_q;    /* ATMOS_DYN_NUMERICAL_DIFF_COEF_TRACER (0 by default:
                          num_diff_q = 0, LARGE.F90:1053-1054)                 */
  int32_t short_type;  /* ATMOS_DYN_TYPE: 0 = FVM-HEVI, 1 = FVM-HEVE (the
                          explicit stage of scale_hip_tstep_short_heve; FP64
                          only, combined by a separate RK4 kernel)             */
  int32_t tinteg_short;  /* ATMOS_DYN_TINTEG_SHORT_TYPE: SCALE_HIP_RK4 (or 0),
                            SCALE_HIP_RK3 (Heun) or SCALE_HIP_RK3WS2002
                            (scale_atmos_dyn_tinteg_short_rk3.F90:188-576)    */
  int32_t tinteg_tracer; /* ATMOS_DYN_TINTEG_TRACER_TYPE: SCALE_HIP_TRACER_RK3WS2002
                            (0) or SCALE_HIP_TRACER_EULER
                            (scale_atmos_dyn_tinteg_tracer_euler.F90:80-133)  */
} scale_hip_dyn_config;

/* time integrator ids of scale_hip_dyn_config */
#define SCALE_HIP_RK3 3
#define SCALE_HIP_RK4 4
#define SCALE_HIP_RK3WS2002 32
#define SCALE_HIP_TRACER_RK3WS2002 0
#define SCALE_HIP_TRACER_EULER 1

/* field ids for scale_hip_model_set/_get/_mass */
enum {
  SCALE_HIP_DENS = 0, SCALE_HIP_MOMZ = 1, SCALE_HIP_MOMX = 2, SCALE_HIP_MOMY = 3, SCALE_HIP_RHOT = 4,
  SCALE_HIP_REF_DENS = 10, SCALE_HIP_REF_POTT = 11, SCALE_HIP_REF_PRES = 12,
  SCALE_HIP_CORIOLI = 13,    /* (1,IA,JA) */
  SCALE_HIP_WDAMP_COEF = 14, /* (KA)      */
  SCALE_HIP_MFLX_HI = 20, SCALE_HIP_TFLX_HI = 21, /* (KA,IA,JA,3), get only */
  SCALE_HIP_DENS_TP = 30, SCALE_HIP_MOMZ_TP = 31, SCALE_HIP_MOMX_TP = 32, SCALE_HIP_MOMY_TP = 33, SCALE_HIP_RHOT_TP = 34,
  SCALE_HIP_QTRC = 100,      /* + tracer index */
  SCALE_HIP_RHOQ_TP = 200    /* + tracer index */
};

/* Grid and metric arrays of scale_atmos_grid_cartesC(_metric); shapes as in
 * the Tstep_short interface plus CDX(IA), CDY(JA). */
int scale_hip_set_metrics(scale_hip_ctx *ctx,
    const double *CDZ, const double *FDZ, const double *RCDZ, const double *RFDZ,
    const double *CDX, const double *FDX, const double *RCDX, const double *RFDX,
    const double *CDY, const double *FDY, const double *RCDY, const double *RFDY,
    const double *PHI, const double *GSQRT, const double *J13G, const double *J23G,
    const double *J33G, const double *MAPF);

/* Terrain-following metrics computed on the device from the surface height,
 * in place of scale_hip_set_metrics: the Gal-Chen heights
 * z = (Htop - Zsfc)/Htop * zeta + Zsfc of the cell centres CZ(KA) and faces
 * FZ(0:KA) (KA+1 values) at the x, u, v and uv points, Htop = FZ(KE) -
 * FZ(KS-1), and from them GSQRT, J13G, J23G (7 components each, periodic
 * halos), PHI = GRAV * z and J33G = 1
 * (scalelib/src/atmosphere/grid/scale_atmos_grid_cartesC_real.F90:391-563,
 *  scalelib/src/atmosphere/grid/scale_atmos_grid_cartesC_metric.F90:200-345).
 * Zsfc(IA,JA) with its halo filled; MAPF(IA,JA,2,4) as given. */
int scale_hip_set_terrain(scale_hip_ctx *ctx,
    const double *CDZ, const double *FDZ, const double *RCDZ, const double *RFDZ,
    const double *CDX, const double *FDX, const double *RCDX, const double *RFDX,
    const double *CDY, const double *FDY, const double *RCDY, const double *RFDY,
    const double *CZ, const double *FZ, const double *Zsfc, const double *MAPF);
/* The context's metrics back in the reference's layout: GSQRT, J13G, J23G
 * (KA,IA,JA,7) and PHI (KA,IA,JA); any may be NULL. */
int scale_hip_get_metrics(scale_hip_ctx *ctx, double *GSQRT, double *J13G, double *J23G, double *PHI);
int scale_hip_model_create(scale_hip_ctx *ctx, const scale_hip_dyn_config *cfg, int32_t QA);
int scale_hip_model_set(scale_hip_ctx *ctx, int32_t field, const double *host);
int scale_hip_model_get(scale_hip_ctx *ctx, int32_t field, double *host);
/* n large steps (dynamics + tracers) */
int scale_hip_model_large_step(scale_hip_ctx *ctx, int32_t n);
/* large-step preparation only (RT2P, DPRES0, REF_rhot, tendencies) */
int scale_hip_model_prepare(scale_hip_ctx *ctx);
/* n short steps: numerical filter + 4 HEVI stages + RK4 combination + halos,
 * using the preparation of the last scale_hip_model_prepare */
int scale_hip_model_short_step(scale_hip_ctx *ctx, int32_t n);
/* Restart (checkpoint/resume) of the device-resident model: the variables of
 * the reference's atmosphere restart, DENS, MOMZ, MOMX, MOMY, RHOT and the
 * QA tracers (scale-rm/src/atmos/mod_atmos_vars.F90:985-1093, 3409-3483,
 * written through FILE_CARTESC as NetCDF there), here as one binary file:
 * 8 bytes "SCLHIPR1", int32 KA, IA, JA, QA, then each field as a FP64
 * (KA,IA,JA) array in Fortran order with its halo.  A run restarted from a
 * file written after step n continues bit for bit as the run that went on
 * (the reference's restart/500m test, runA against runB1 + runB2).  Reading
 * checks the grid and the number of tracers; FP64 models only (a model in
 * mixed precision is rejected). */
int scale_hip_model_restart_write(scale_hip_ctx *ctx, const char *path);
int scale_hip_model_restart_read(scale_hip_ctx *ctx, const char *path);
/* FP64 sum over the interior of field(k,i,j) [* DENS if weight_dens] * CDZ*CDX*CDY*GSQRT/(MAPF1*MAPF2) */
int scale_hip_model_mass(scale_hip_ctx *ctx, int32_t field, int32_t weight_dens, double *out);
/* device pointer of a resident field (device layout: i fastest, then j, then k) */
int scale_hip_model_device_ptr(scale_hip_ctx *ctx, int32_t field, double **dptr);
/* the HIP stream every model call is ordered on */
int scale_hip_stream(scale_hip_ctx *ctx, void **stream);

/* Named kernel timers: hipEvents recorded around every launch on the context's
 * stream (the counterpart of PROF_rapstart/PROF_rapend,
 * scalelib/src/common/scale_prof.F90:156-253).  Names: hevi_stage,
 * numfilter, numfilter_q, halo_fill, rk4_combine, tracer_nofct, fct_philo, fct_ratio,
 * fct_update, mfct_flux, mfct_limit, hbm_copy, model_large_step,
 * model_forward_checkpointed, model_adjoint.  on = 1 times every launch;
 * on = 2 only the outermost timer of nested ones (model_large_step,
 * model_forward_checkpointed, model_adjoint: two events per call, nothing
 * recorded between the kernels inside).  Enabling resets the accumulators. */
int scale_hip_profile_enable(scale_hip_ctx *ctx, int32_t on);
int scale_hip_profile_query(scale_hip_ctx *ctx, const char *name, int64_t *count, double *total_ms);

/* ---- adjoint --------------------------------------------------------------
 * No counterpart in the reference.  The transpose of the Jacobian of the
 * forward above (with the UD3 upwind switch sign(vel) held fixed), for flat
 * or terrain-following metrics (held fixed), FP64, divdmp_coef = 0 and one
 * short step per large step; other configurations are rejected.  Adjoint arrays have the shapes of the fields
 * they belong to (Fortran layout, host memory).
 *
 * Adjoint of one scale_hip_tstep_short_hevi call: given the stage inputs
 * (the same arrays as the forward call) and the adjoints *_RK_b of its five
 * prognostic outputs, returns the adjoints of its inputs (overwritten):
 * state0, state, num_diff(KA,IA,JA,5,3), DPRES0, RT2P and REF_rhot. */
int scale_hip_tstep_short_hevi_adj(scale_hip_ctx *ctx,
    const double *DENS_RK_b, const double *MOMZ_RK_b, const double *MOMX_RK_b, const double *MOMY_RK_b, const double *RHOT_RK_b,
    double *DENS0_b, double *MOMZ0_b, double *MOMX0_b, double *MOMY0_b, double *RHOT0_b,
    double *DENS_b, double *MOMZ_b, double *MOMX_b, double *MOMY_b, double *RHOT_b,
    double *num_diff_b, double *DPRES0_b, double *RT2P_b, double *REF_rhot_b,
    const double *DENS0, const double *MOMZ0, const double *MOMX0, const double *MOMY0, const double *RHOT0,
    const double *DENS, const double *MOMZ, const double *MOMX, const double *MOMY, const double *RHOT,
    const double *DENS_t, const double *MOMZ_t, const double *MOMX_t, const double *MOMY_t, const double *RHOT_t,
    const double *DPRES0, const double *RT2P, const double *CORIOLI,
    const double *num_diff, const double *wdamp_coef,
    const double *CDZ, const double *FDZ, const double *FDX, const double *FDY,
    const double *RCDZ, const double *RCDX, const double *RCDY, const double *RFDZ, const double *RFDX, const double *RFDY,
    const double *PHI, const double *GSQRT, const double *J13G, const double *J23G, const double *J33G, const double *MAPF,
    const double *REF_dens, const double *REF_rhot, const double *dtrk);

/* Forward/adjoint pair of the device-resident model.  The forward runs n
 * large steps (as scale_hip_model_large_step) and keeps the state every
 * `every` steps in HBM.  The adjoint takes the adjoint of the five
 * prognostics after that run (DENS_b .. RHOT_b, whole arrays) and returns in
 * place their adjoint at its start, recomputing each segment of `every` steps
 * that the forward kept no tape of (SCALE_HIP_TUNE_ADJ_TAPE below) from its
 * checkpoint.  The model state is left at the end of the run. */
int scale_hip_model_forward_checkpointed(scale_hip_ctx *ctx, int32_t n, int32_t every);
int scale_hip_model_adjoint(scale_hip_ctx *ctx, double *DENS_b, double *MOMZ_b, double *MOMX_b, double *MOMY_b, double *RHOT_b);

/* The same pair for a batch of n member contexts (each with its model, of
 * any grid size, flat or terrain, all on one device), advanced together:
 * each member runs on its own stream, the forwards step by step and the
 * reverse sweeps segment by segment across the members, so that their
 * kernels overlap, and the members split the device memory for tapes.  Each
 * member's results equal its own model_forward_checkpointed/model_adjoint
 * pair (up to the order of the adjoint's atomic sums).  DENS_b .. RHOT_b are
 * arrays of n host pointers, one field per member, overwritten as in
 * scale_hip_model_adjoint.  No counterpart in the reference (which runs
 * members as separate jobs, scale-rm/src/scale-rm.F90:180-184). */
int scale_hip_batch_forward_checkpointed(scale_hip_ctx *const *members, int32_t n, int32_t nsteps, int32_t every);
int scale_hip_batch_adjoint(scale_hip_ctx *const *members, int32_t n, double *const *DENS_b, double *const *MOMZ_b,
                            double *const *MOMX_b, double *const *MOMY_b, double *const *RHOT_b);

/* Launch-shape and memory knobs, with no counterpart in the reference (whose
 * loops take IBLOCK/JBLOCK from PARAM_ATMOS_DYN; here the tiles are chosen
 * from the grid and the CU count unless a knob is set).  Results do not
 * depend on them: every tile shape computes the same operations in the same
 * order.
 *   SCALE_HIP_TUNE_NF_TJ: rows of the numerical-filter tile, 4, 8 or 16
 *   (0 = chosen from the grid).
 *   SCALE_HIP_TUNE_TRACER_TILE: the FCT tracer stage on LDS tiles (flat
 *   terrain; UD3, CD4, UD3KOREN1993, UD1 and CD2): 1 (default) where the grid
 *   gives every CU two tiles, 2 always, 0 never (per-point kernels).
 *   SCALE_HIP_TUNE_ADJ_TAPE: the number of trailing segments (of `every`
 *   steps) whose tapes (the step states and the 45 fields per step the
 *   reverse step reads, 400 B per grid point and step) the checkpointed
 *   forward keeps in device memory, so that the adjoint reverses them without
 *   a recording forward: n >= 1 at most n, 0 (default) as many as half the
 *   device memory free holds when the forward first runs with its n and
 *   every.  The segments
 *   without a tape are recomputed from their checkpoints as before.
 * Unknown knobs and values outside the listed ones are rejected. */
enum { SCALE_HIP_TUNE_NF_TJ = 1, SCALE_HIP_TUNE_TRACER_TILE = 2, SCALE_HIP_TUNE_ADJ_TAPE = 3 };
int scale_hip_tune_set(scale_hip_ctx *ctx, int32_t knob, int32_t value);
int scale_hip_tune_get(scale_hip_ctx *ctx, int32_t knob, int32_t *value);

/* HBM probe with no counterpart in the reference: `reps` timed streaming
 * copies of n doubles between two scratch buffers (allocated and freed by the
 * call); *avg_ms is the mean time of one copy, which moves 16*n bytes.  Gives
 * the achievable copy bandwidth and a kernel of known traffic to calibrate
 * the FETCH_SIZE / WRITE_SIZE counters of the stencil kernels against. */
int scale_hip_hbm_copy_probe(scale_hip_ctx *ctx, int64_t n, int32_t reps, double *avg_ms);

/* Bandwidth probes on two fresh buffers of n doubles, timed by events over
 * reps launches after one warm-up (average ms per launch): kind 0 the copy
 * above, 1 the copy with non-temporal loads and stores, 2 read-only (a
 * block-sum reduction), 3 write-only (a fill), 4 hipMemcpyAsync device to
 * device.  No counterpart in the reference: it measures the device. */
int scale_hip_hbm_probe(scale_hip_ctx *ctx, int32_t kind, int64_t n, int32_t reps, double *avg_ms);

/* Ensembles: member contexts, each with its model created and its state set,
 * of any grid sizes on one device, stepped together.  The numerical filter
 * and each of the four HEVI stages of a short step run as one launch over
 * the tiles of every member (a member table on the device, blocks walking
 * the tiles); the other launches of a large step stay per member.  The
 * members must share the precision (FP64 or mixed), the dynamics scheme and
 * the short steps per large step, with RK4 short steps on flat terrain and
 * no divergence damping; each member's results are bit-identical to its own
 * scale_hip_model_large_step.  The reference runs ensemble members as
 * separate jobs of one program (the bulk-job communicator split,
 * scale-rm/src/scale-rm.F90:180-184); this has no other counterpart there.
 * The member contexts stay owned by the caller and must outlive the
 * ensemble; their fields are read with scale_hip_model_get after
 * scale_hip_ensemble_synchronize.
 *   set_graph: capture one large step of the whole ensemble in a hipGraph
 *     on the next call and replay it from then on (1), or launch directly (0).
 *   set_shape: persistent blocks per CU of the batched stage (0..2; 0 = one
 *     block per tile) and filter (0..4), and the filter's tile rows (4 or 8).
 *   tune: the autotuner: times `steps` large steps of every candidate shape
 *     after one warm-up step, keeps the fastest, and returns the members to
 *     their state before the call; *best_ms is the fastest time per large
 *     step, report (len bytes, may be NULL) lists every candidate. */
typedef struct scale_hip_ensemble scale_hip_ensemble;
int scale_hip_ensemble_create(scale_hip_ctx *const *members, int32_t n, scale_hip_ensemble **ens);
int scale_hip_ensemble_destroy(scale_hip_ensemble *ens);
int scale_hip_ensemble_large_step(scale_hip_ensemble *ens, int32_t n);
int scale_hip_ensemble_synchronize(scale_hip_ensemble *ens);
int scale_hip_ensemble_set_graph(scale_hip_ensemble *ens, int32_t on);
int scale_hip_ensemble_set_shape(scale_hip_ensemble *ens, int32_t stage_bpc, int32_t nf_bpc, int32_t nf_tj);
int scale_hip_ensemble_get_shape(scale_hip_ensemble *ens, int32_t *stage_bpc, int32_t *nf_bpc, int32_t *nf_tj);
int scale_hip_ensemble_tune(scale_hip_ensemble *ens, int32_t steps, double *best_ms, char *report, int32_t len);

#ifdef __cplusplus
}
#endif
#endif
