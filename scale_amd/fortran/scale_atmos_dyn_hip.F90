!-------------------------------------------------------------------------------
!> module scale_atmos_dyn_hip
!!
!! Fortran drop-in for the MI355X HEVI and HEVE short steps and FCT tracer advection.
!!
!! ATMOS_DYN_Tstep_short_fvm_hevi_hip and ATMOS_DYN_Tstep_short_fvm_heve_hip have the dummy list, intents and meaning
!! of the abstract interface `short` of ATMOS_DYN_Tstep_short
!! (scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_short.F90:38-128) and
!! ATMOS_DYN_Tstep_tracer_fvm_heve_hip those of `step` of ATMOS_DYN_Tstep_tracer
!! (scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_tracer.F90:34-64), so a
!! host binds them to the procedure pointers in place of the CPU schemes.  The
!! arrays go through iso_c_binding to libscalehip.so (include/scale_hip.h);
!! a nonzero status is fatal, as LOG_ERROR + PRC_abort are in the reference.
!!
!! ATMOS_DYN_Tstep_large_fvm_heve_hip has the dummies of the abstract
!! interface `large` of ATMOS_DYN_Tstep_large
!! (scalelib/src/atmos-rm/dynamics/scale_atmos_dyn_tstep_large.F90:34-178): one
!! large step of the device-resident model (filter, short steps, mass-flux
!! average, tracers), one host-to-device copy of the state in and one back
!! per call instead of one per stage.
!-------------------------------------------------------------------------------
module scale_atmos_dyn_hip
  use iso_c_binding
  implicit none
  private

  public :: ATMOS_DYN_hip_setup
  public :: ATMOS_DYN_hip_finalize
  public :: ATMOS_DYN_Tstep_short_fvm_hevi_hip
  public :: ATMOS_DYN_Tstep_short_fvm_heve_hip
  public :: ATMOS_DYN_Tstep_tracer_fvm_heve_hip
  public :: ATMOS_DYN_numfilter_coef_hip
  public :: ATMOS_DYN_numfilter_coef_q_hip
  public :: ATMOS_DYN_Tstep_large_fvm_heve_hip
  public :: ATMOS_DYN_hip_setup_large

  integer, parameter :: DP = c_double

  type, bind(C) :: scale_hip_grid
     integer(c_int32_t) :: KA, IA, JA
     integer(c_int32_t) :: KS, KE, IS, IE, JS, JE
     integer(c_int32_t) :: KHALO, IHALO, JHALO
     integer(c_int32_t) :: flux_scheme, flux_scheme_tracer
     integer(c_int32_t) :: device
  end type scale_hip_grid

  type, bind(C) :: scale_hip_dyn_config
     real(c_double) :: dtl, dts, nd_coef, nd_sfc_fact, divdmp_coef
     integer(c_int32_t) :: nd_order, fct_tracer, fct_along_stream, mixed_precision
     real(c_double) :: nd_coef_q
     integer(c_int32_t) :: short_type, tinteg_short, tinteg_tracer
  end type scale_hip_dyn_config

  type(c_ptr), save :: ctx = c_null_ptr
  integer, save, public :: KA, IA, JA, VA = 0
  ! the large step: tracer count, integrators (scale_hip_dyn_config ids) and
  ! precision, set by ATMOS_DYN_hip_setup_large; the model is created on the
  ! first call with the step's own settings
  integer, save, public :: QA = 0, BND_QA_max = 0
  integer, save :: large_short_type = 0, large_tinteg_short = 4, large_tinteg_tracer = 0, large_mixed = 0
  logical, save :: model_ready = .false.
  real(DP), save :: model_dtl = -1.0_DP, model_dts = -1.0_DP

  interface
     integer(c_int) function c_init(g, h) bind(C, name="scale_hip_init")
       import :: scale_hip_grid, c_ptr, c_int
       type(scale_hip_grid), intent(in) :: g
       type(c_ptr), intent(out) :: h
     end function c_init
     integer(c_int) function c_finalize(h) bind(C, name="scale_hip_finalize")
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function c_finalize
     type(c_ptr) function c_last_error() bind(C, name="scale_hip_last_error")
       import :: c_ptr
     end function c_last_error
     integer(c_int) function c_tstep_short(h, &
          DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, PROG_RK, mflx_hi, tflx_hi, &
          DENS0, MOMZ0, MOMX0, MOMY0, RHOT0, DENS, MOMZ, MOMX, MOMY, RHOT, &
          DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, PROG0, PROG, &
          DPRES0, RT2P, CORIOLI, num_diff, wdamp_coef, divdmp_coef, DDIV, &
          FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_ALONG_STREAM, &
          CDZ, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY, &
          PHI, GSQRT, J13G, J23G, J33G, MAPF, REF_dens, REF_rhot, &
          BND_W, BND_E, BND_S, BND_N, dtrk, last) bind(C, name="scale_hip_tstep_short_hevi")
       import :: c_ptr, c_int, c_double, c_int32_t
       type(c_ptr), value :: h
       real(c_double) :: DENS_RK(*), MOMZ_RK(*), MOMX_RK(*), MOMY_RK(*), RHOT_RK(*), PROG_RK(*), mflx_hi(*), tflx_hi(*)
       real(c_double), intent(in) :: DENS0(*), MOMZ0(*), MOMX0(*), MOMY0(*), RHOT0(*)
       real(c_double), intent(in) :: DENS(*), MOMZ(*), MOMX(*), MOMY(*), RHOT(*)
       real(c_double), intent(in) :: DENS_t(*), MOMZ_t(*), MOMX_t(*), MOMY_t(*), RHOT_t(*), PROG0(*), PROG(*)
       real(c_double), intent(in) :: DPRES0(*), RT2P(*), CORIOLI(*), num_diff(*), wdamp_coef(*), divdmp_coef, DDIV(*)
       integer(c_int32_t), intent(in) :: FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_ALONG_STREAM
       real(c_double), intent(in) :: CDZ(*), FDZ(*), FDX(*), FDY(*), RCDZ(*), RCDX(*), RCDY(*), RFDZ(*), RFDX(*), RFDY(*)
       real(c_double), intent(in) :: PHI(*), GSQRT(*), J13G(*), J23G(*), J33G, MAPF(*), REF_dens(*), REF_rhot(*)
       integer(c_int32_t), intent(in) :: BND_W, BND_E, BND_S, BND_N, last
       real(c_double), intent(in) :: dtrk
     end function c_tstep_short
     integer(c_int) function c_tstep_short_heve(h, &
          DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, PROG_RK, mflx_hi, tflx_hi, &
          DENS0, MOMZ0, MOMX0, MOMY0, RHOT0, DENS, MOMZ, MOMX, MOMY, RHOT, &
          DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, PROG0, PROG, &
          DPRES0, RT2P, CORIOLI, num_diff, wdamp_coef, divdmp_coef, DDIV, &
          FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_ALONG_STREAM, &
          CDZ, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY, &
          PHI, GSQRT, J13G, J23G, J33G, MAPF, REF_dens, REF_rhot, &
          BND_W, BND_E, BND_S, BND_N, dtrk, last) bind(C, name="scale_hip_tstep_short_heve")
       import :: c_ptr, c_int, c_double, c_int32_t
       type(c_ptr), value :: h
       real(c_double) :: DENS_RK(*), MOMZ_RK(*), MOMX_RK(*), MOMY_RK(*), RHOT_RK(*), PROG_RK(*), mflx_hi(*), tflx_hi(*)
       real(c_double), intent(in) :: DENS0(*), MOMZ0(*), MOMX0(*), MOMY0(*), RHOT0(*)
       real(c_double), intent(in) :: DENS(*), MOMZ(*), MOMX(*), MOMY(*), RHOT(*)
       real(c_double), intent(in) :: DENS_t(*), MOMZ_t(*), MOMX_t(*), MOMY_t(*), RHOT_t(*), PROG0(*), PROG(*)
       real(c_double), intent(in) :: DPRES0(*), RT2P(*), CORIOLI(*), num_diff(*), wdamp_coef(*), divdmp_coef, DDIV(*)
       integer(c_int32_t), intent(in) :: FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_ALONG_STREAM
       real(c_double), intent(in) :: CDZ(*), FDZ(*), FDX(*), FDY(*), RCDZ(*), RCDX(*), RCDY(*), RFDZ(*), RFDX(*), RFDY(*)
       real(c_double), intent(in) :: PHI(*), GSQRT(*), J13G(*), J23G(*), J33G, MAPF(*), REF_dens(*), REF_rhot(*)
       integer(c_int32_t), intent(in) :: BND_W, BND_E, BND_S, BND_N, last
       real(c_double), intent(in) :: dtrk
     end function c_tstep_short_heve
     integer(c_int) function c_tstep_tracer(h, QTRCo, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff, &
          GSQRT, MAPF, CDZ, RCDZ, RCDX, RCDY, dtl, FLAG_FCT_TRACER, FLAG_FCT_ALONG_STREAM) &
          bind(C, name="scale_hip_tstep_tracer_heve")
       import :: c_ptr, c_int, c_double, c_int32_t
       type(c_ptr), value :: h
       real(c_double) :: QTRCo(*)
       real(c_double), intent(in) :: QTRC(*), QTRC0(*), RHOQ_t(*), DENS0(*), DENS(*), mflx_hi(*), num_diff(*)
       real(c_double), intent(in) :: GSQRT(*), MAPF(*), CDZ(*), RCDZ(*), RCDX(*), RCDY(*), dtl
       integer(c_int32_t), intent(in) :: FLAG_FCT_TRACER, FLAG_FCT_ALONG_STREAM
     end function c_tstep_tracer
     integer(c_int) function c_numfilter(h, num_diff, DENS, MOMZ, MOMX, MOMY, RHOT, CDZ, CDX, CDY, FDZ, FDX, FDY, &
          DT, REF_dens, REF_pott, ND_COEF, ND_ORDER, ND_SFC_FACT, ND_USE_RS) bind(C, name="scale_hip_numfilter_coef")
       import :: c_ptr, c_int, c_double, c_int32_t
       type(c_ptr), value :: h
       real(c_double) :: num_diff(*)
       real(c_double), intent(in) :: DENS(*), MOMZ(*), MOMX(*), MOMY(*), RHOT(*)
       real(c_double), intent(in) :: CDZ(*), CDX(*), CDY(*), FDZ(*), FDX(*), FDY(*), DT, REF_dens(*), REF_pott(*)
       real(c_double), intent(in) :: ND_COEF, ND_SFC_FACT
       integer(c_int32_t), intent(in) :: ND_ORDER, ND_USE_RS
     end function c_numfilter
     integer(c_int) function c_set_metrics(h, CDZ, FDZ, RCDZ, RFDZ, CDX, FDX, RCDX, RFDX, CDY, FDY, RCDY, RFDY, &
          PHI, GSQRT, J13G, J23G, J33G, MAPF) bind(C, name="scale_hip_set_metrics")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: CDZ(*), FDZ(*), RCDZ(*), RFDZ(*), CDX(*), FDX(*), RCDX(*), RFDX(*)
       real(c_double), intent(in) :: CDY(*), FDY(*), RCDY(*), RFDY(*), PHI(*), GSQRT(*), J13G(*), J23G(*), J33G, MAPF(*)
     end function c_set_metrics
     integer(c_int) function c_model_create(h, cfg, nq) bind(C, name="scale_hip_model_create")
       import :: c_ptr, c_int, c_int32_t, scale_hip_dyn_config
       type(c_ptr), value :: h
       type(scale_hip_dyn_config), intent(in) :: cfg
       integer(c_int32_t), value :: nq
     end function c_model_create
     integer(c_int) function c_model_set(h, field, a) bind(C, name="scale_hip_model_set")
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: field
       real(c_double), intent(in) :: a(*)
     end function c_model_set
     integer(c_int) function c_model_get(h, field, a) bind(C, name="scale_hip_model_get")
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: field
       real(c_double) :: a(*)
     end function c_model_get
     integer(c_int) function c_model_large_step(h, n) bind(C, name="scale_hip_model_large_step")
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: h
       integer(c_int32_t), value :: n
     end function c_model_large_step
     integer(c_int) function c_synchronize(h) bind(C, name="scale_hip_synchronize")
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function c_synchronize
     integer(c_int) function c_numfilter_q(h, num_diff_q, DENS, QTRC, is_qv, CDZ, CDX, CDY, dt, REF_qv, iq, &
          ND_COEF, ND_ORDER, ND_SFC_FACT, ND_USE_RS) bind(C, name="scale_hip_numfilter_coef_q")
       import :: c_ptr, c_int, c_double, c_int32_t
       type(c_ptr), value :: h
       real(c_double) :: num_diff_q(*)
       real(c_double), intent(in) :: DENS(*), QTRC(*), CDZ(*), CDX(*), CDY(*), dt, REF_qv(*)
       real(c_double), intent(in) :: ND_COEF, ND_SFC_FACT
       integer(c_int32_t), intent(in) :: is_qv, iq, ND_ORDER, ND_USE_RS
     end function c_numfilter_q
  end interface

contains

  !> create the device context; extents and 1-based bounds as in scale_atmos_grid_cartesC_index
  subroutine ATMOS_DYN_hip_setup(KA_in, IA_in, JA_in, KS, KE, IS, IE, JS, JE, flux_scheme, flux_scheme_tracer, device)
    integer, intent(in) :: KA_in, IA_in, JA_in, KS, KE, IS, IE, JS, JE
    integer, intent(in) :: flux_scheme, flux_scheme_tracer, device
    type(scale_hip_grid) :: g
    KA = KA_in; IA = IA_in; JA = JA_in
    g = scale_hip_grid(KA, IA, JA, KS, KE, IS, IE, JS, JE, KS - 1, IS - 1, JS - 1, flux_scheme, flux_scheme_tracer, device)
    call check(c_init(g, ctx), "ATMOS_DYN_hip_setup")
  end subroutine ATMOS_DYN_hip_setup

  subroutine ATMOS_DYN_hip_finalize
    integer(c_int) :: ierr
    if (c_associated(ctx)) ierr = c_finalize(ctx)
    ctx = c_null_ptr
    model_ready = .false.
  end subroutine ATMOS_DYN_hip_finalize

  !> settings of the large step the reference takes from other modules: the
  !> tracer count (scale_tracer's QA), ATMOS_DYN_TYPE (0 FVM-HEVI, 1 FVM-HEVE),
  !> ATMOS_DYN_TINTEG_SHORT_TYPE (4 RK4, 3 RK3, 32 RK3WS2002),
  !> ATMOS_DYN_TINTEG_TRACER_TYPE (0 RK3WS2002, 1 EULER), and mixed precision
  !> (1: the reference's SINGLE build with the FP64 vertical solve)
  subroutine ATMOS_DYN_hip_setup_large(QA_in, short_type, tinteg_short, tinteg_tracer, mixed, BND_QA_in)
    integer, intent(in) :: QA_in, short_type, tinteg_short, tinteg_tracer, mixed, BND_QA_in
    QA = QA_in
    BND_QA_max = BND_QA_in
    large_short_type = short_type
    large_tinteg_short = tinteg_short
    large_tinteg_tracer = tinteg_tracer
    large_mixed = mixed
  end subroutine ATMOS_DYN_hip_setup_large

  !> one large step with the dummies of the abstract interface `large`
  !> (scale_atmos_dyn_tstep_large.F90:34-178) on the device-resident model.
  !> Supported: periodic boundaries, DAMP_alpha_* = 0 (no boundary damping),
  !> USE_AVERAGE = .false., FLAG_TRACER_SPLIT_TEND = .false., dry dynamics
  !> (I_QV < 1: the tracers are passive), ND_ORDER = 1 with ND_USE_RS; other
  !> values stop with an error.  num_diff and num_diff_q stay on the device
  !> (the reference returns them for history output only); QTRC0 is the
  !> tracer on entry, as the reference's driver passes it.
  subroutine ATMOS_DYN_Tstep_large_fvm_heve_hip( &
       DENS, MOMZ, MOMX, MOMY, RHOT, QTRC, PROG,             &
       DENS_av, MOMZ_av, MOMX_av, MOMY_av, RHOT_av, QTRC_av, &
       mflx_hi, tflx_hi,                                     &
       num_diff, num_diff_q,                                 &
       QTRC0,                                                &
       DENS_tp, MOMZ_tp, MOMX_tp, MOMY_tp, RHOT_tp, RHOQ_tp, &
       CORIOLI,                                              &
       CDZ, CDX, CDY, FDZ, FDX, FDY,                         &
       RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY,                   &
       PHI, GSQRT,                                           &
       J13G, J23G, J33G, MAPF,                               &
       AQ_R, AQ_CV, AQ_CP, AQ_MASS,                          &
       REF_dens, REF_pott, REF_qv, REF_pres,                 &
       BND_W, BND_E, BND_S, BND_N,                           &
       ND_COEF, ND_COEF_Q, ND_ORDER, ND_SFC_FACT, ND_USE_RS, &
       BND_QA, BND_SMOOTHER_FACT,                            &
       DAMP_DENS,       DAMP_VELZ,       DAMP_VELX,          &
       DAMP_VELY,       DAMP_POTT,       DAMP_QTRC,          &
       DAMP_alpha_DENS, DAMP_alpha_VELZ, DAMP_alpha_VELX,    &
       DAMP_alpha_VELY, DAMP_alpha_POTT, DAMP_alpha_QTRC,    &
       wdamp_coef,                                           &
       divdmp_coef,                                          &
       FLAG_TRACER_SPLIT_TEND,                               &
       FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_TRACER,       &
       FLAG_FCT_ALONG_STREAM,                                &
       USE_AVERAGE,                                          &
       I_QV,                                                 &
       DTL, DTS, Llast                                       )
    real(DP), intent(inout) :: DENS(KA,IA,JA), MOMZ(KA,IA,JA), MOMX(KA,IA,JA), MOMY(KA,IA,JA), RHOT(KA,IA,JA)
    real(DP), intent(inout) :: QTRC(KA,IA,JA,QA), PROG(KA,IA,JA,VA)
    real(DP), intent(inout) :: DENS_av(KA,IA,JA), MOMZ_av(KA,IA,JA), MOMX_av(KA,IA,JA), MOMY_av(KA,IA,JA)
    real(DP), intent(inout) :: RHOT_av(KA,IA,JA), QTRC_av(KA,IA,JA,QA)
    real(DP), intent(out)   :: mflx_hi(KA,IA,JA,3), tflx_hi(KA,IA,JA,3)
    real(DP), intent(out)   :: num_diff(KA,IA,JA,5,3), num_diff_q(KA,IA,JA,3)
    real(DP), intent(in)    :: QTRC0(KA,IA,JA,QA)
    real(DP), intent(in)    :: DENS_tp(KA,IA,JA), MOMZ_tp(KA,IA,JA), MOMX_tp(KA,IA,JA), MOMY_tp(KA,IA,JA)
    real(DP), intent(in)    :: RHOT_tp(KA,IA,JA), RHOQ_tp(KA,IA,JA,QA)
    real(DP), intent(in)    :: CORIOLI(IA,JA)
    real(DP), intent(in)    :: CDZ(KA), CDX(IA), CDY(JA), FDZ(KA-1), FDX(IA-1), FDY(JA-1)
    real(DP), intent(in)    :: RCDZ(KA), RCDX(IA), RCDY(JA), RFDZ(KA-1), RFDX(IA-1), RFDY(JA-1)
    real(DP), intent(in)    :: PHI(KA,IA,JA), GSQRT(KA,IA,JA,7), J13G(KA,IA,JA,7), J23G(KA,IA,JA,7), J33G
    real(DP), intent(in)    :: MAPF(IA,JA,2,4)
    real(DP), intent(in)    :: AQ_R(QA), AQ_CV(QA), AQ_CP(QA), AQ_MASS(QA)
    real(DP), intent(in)    :: REF_dens(KA,IA,JA), REF_pott(KA,IA,JA), REF_qv(KA,IA,JA), REF_pres(KA,IA,JA)
    logical,  intent(in)    :: BND_W, BND_E, BND_S, BND_N
    real(DP), intent(in)    :: ND_COEF, ND_COEF_Q, ND_SFC_FACT
    integer,  intent(in)    :: ND_ORDER
    logical,  intent(in)    :: ND_USE_RS
    integer,  intent(in)    :: BND_QA
    real(DP), intent(in)    :: BND_SMOOTHER_FACT
    real(DP), intent(in)    :: DAMP_DENS(KA,IA,JA), DAMP_VELZ(KA,IA,JA), DAMP_VELX(KA,IA,JA), DAMP_VELY(KA,IA,JA)
    real(DP), intent(in)    :: DAMP_POTT(KA,IA,JA), DAMP_QTRC(KA,IA,JA,BND_QA)
    real(DP), intent(in)    :: DAMP_alpha_DENS(KA,IA,JA), DAMP_alpha_VELZ(KA,IA,JA), DAMP_alpha_VELX(KA,IA,JA)
    real(DP), intent(in)    :: DAMP_alpha_VELY(KA,IA,JA), DAMP_alpha_POTT(KA,IA,JA), DAMP_alpha_QTRC(KA,IA,JA,BND_QA)
    real(DP), intent(in)    :: wdamp_coef(KA), divdmp_coef
    logical,  intent(in)    :: FLAG_TRACER_SPLIT_TEND, FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_TRACER
    logical,  intent(in)    :: FLAG_FCT_ALONG_STREAM, USE_AVERAGE
    integer,  intent(in)    :: I_QV
    real(DP), intent(in)    :: DTL, DTS
    logical,  intent(in)    :: Llast
    type(scale_hip_dyn_config) :: cfg
    integer :: iq
    character(len=*), parameter :: where = "ATMOS_DYN_Tstep_large_fvm_heve_hip"

    if (BND_W .or. BND_E .or. BND_S .or. BND_N) call stop_with(where, "only periodic lateral boundaries are supported")
    if (USE_AVERAGE) call stop_with(where, "USE_AVERAGE = .true. is not supported")
    if (FLAG_TRACER_SPLIT_TEND) call stop_with(where, "FLAG_TRACER_SPLIT_TEND = .true. is not supported")
    if (I_QV > 0) call stop_with(where, "moist dynamics (I_QV > 0) is not supported: the tracers are passive")
    if (ND_ORDER /= 1 .or. .not. ND_USE_RS) call stop_with(where, "ND_ORDER = 1 with ND_USE_RS is required")
    if (any(DAMP_alpha_DENS /= 0.0_DP) .or. any(DAMP_alpha_VELZ /= 0.0_DP) .or. any(DAMP_alpha_VELX /= 0.0_DP) .or. &
        any(DAMP_alpha_VELY /= 0.0_DP) .or. any(DAMP_alpha_POTT /= 0.0_DP)) &
         call stop_with(where, "boundary damping (DAMP_alpha /= 0) is not supported")
    if (BND_QA > 0) then
       if (any(DAMP_alpha_QTRC /= 0.0_DP)) call stop_with(where, "boundary damping of tracers is not supported")
    end if

    if (.not. model_ready) then
       call check(c_set_metrics(ctx, CDZ, FDZ, RCDZ, RFDZ, CDX, FDX, RCDX, RFDX, CDY, FDY, RCDY, RFDY, &
            PHI, GSQRT, J13G, J23G, J33G, MAPF), where)
       cfg = scale_hip_dyn_config(DTL, DTS, ND_COEF, ND_SFC_FACT, divdmp_coef, int(ND_ORDER, c_int32_t), &
            l2i(FLAG_FCT_TRACER), l2i(FLAG_FCT_ALONG_STREAM), int(large_mixed, c_int32_t), ND_COEF_Q, &
            int(large_short_type, c_int32_t), int(large_tinteg_short, c_int32_t), int(large_tinteg_tracer, c_int32_t))
       call check(c_model_create(ctx, cfg, int(QA, c_int32_t)), where)
       call check(c_model_set(ctx, 10_c_int32_t, REF_dens), where)
       call check(c_model_set(ctx, 11_c_int32_t, REF_pott), where)
       call check(c_model_set(ctx, 12_c_int32_t, REF_pres), where)
       call check(c_model_set(ctx, 13_c_int32_t, CORIOLI), where)
       call check(c_model_set(ctx, 14_c_int32_t, wdamp_coef), where)
       model_ready = .true.
       model_dtl = DTL
       model_dts = DTS
    else if (DTL /= model_dtl .or. DTS /= model_dts) then
       call stop_with(where, "DTL and DTS must not change between calls")
    end if

    ! the state, the tracers and the physics tendencies in
    call check(c_model_set(ctx, 0_c_int32_t, DENS), where)
    call check(c_model_set(ctx, 1_c_int32_t, MOMZ), where)
    call check(c_model_set(ctx, 2_c_int32_t, MOMX), where)
    call check(c_model_set(ctx, 3_c_int32_t, MOMY), where)
    call check(c_model_set(ctx, 4_c_int32_t, RHOT), where)
    call check(c_model_set(ctx, 30_c_int32_t, DENS_tp), where)
    call check(c_model_set(ctx, 31_c_int32_t, MOMZ_tp), where)
    call check(c_model_set(ctx, 32_c_int32_t, MOMX_tp), where)
    call check(c_model_set(ctx, 33_c_int32_t, MOMY_tp), where)
    call check(c_model_set(ctx, 34_c_int32_t, RHOT_tp), where)
    do iq = 1, QA
       call check(c_model_set(ctx, int(99 + iq, c_int32_t), QTRC(:,:,:,iq)), where)
       call check(c_model_set(ctx, int(199 + iq, c_int32_t), RHOQ_tp(:,:,:,iq)), where)
    end do

    call check(c_model_large_step(ctx, 1_c_int32_t), where)

    ! and back
    call check(c_model_get(ctx, 0_c_int32_t, DENS), where)
    call check(c_model_get(ctx, 1_c_int32_t, MOMZ), where)
    call check(c_model_get(ctx, 2_c_int32_t, MOMX), where)
    call check(c_model_get(ctx, 3_c_int32_t, MOMY), where)
    call check(c_model_get(ctx, 4_c_int32_t, RHOT), where)
    call check(c_model_get(ctx, 20_c_int32_t, mflx_hi), where)
    call check(c_model_get(ctx, 21_c_int32_t, tflx_hi), where)
    do iq = 1, QA
       call check(c_model_get(ctx, int(99 + iq, c_int32_t), QTRC(:,:,:,iq)), where)
    end do
    call check(c_synchronize(ctx), where)
  end subroutine ATMOS_DYN_Tstep_large_fvm_heve_hip

  subroutine stop_with(where, msg)
    character(len=*), intent(in) :: where, msg
    write(*,'(A,A,A,A)') 'xxx [', where, '] ', msg
    error stop 1
  end subroutine stop_with

  !> HEVI stage with the dummies of the abstract interface `short`
  subroutine ATMOS_DYN_Tstep_short_fvm_hevi_hip( &
       DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, &
       PROG_RK,                                     &
       mflx_hi, tflx_hi,                            &
       DENS0,   MOMZ0,   MOMX0,   MOMY0,   RHOT0,   &
       DENS,    MOMZ,    MOMX,    MOMY,    RHOT,    &
       DENS_t,  MOMZ_t,  MOMX_t,  MOMY_t,  RHOT_t,  &
       PROG0, PROG,                                 &
       DPRES0, RT2P, CORIOLI,                       &
       num_diff, wdamp_coef, divdmp_coef, DDIV,     &
       FLAG_FCT_MOMENTUM, FLAG_FCT_T,               &
       FLAG_FCT_ALONG_STREAM,                       &
       CDZ, FDZ, FDX, FDY,                          &
       RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY,          &
       PHI, GSQRT, J13G, J23G, J33G, MAPF,          &
       REF_dens, REF_rhot,                          &
       BND_W, BND_E, BND_S, BND_N,                  &
       dtrk, last                                   )
    real(DP), intent(out)   :: DENS_RK(KA,IA,JA), MOMZ_RK(KA,IA,JA), MOMX_RK(KA,IA,JA), MOMY_RK(KA,IA,JA), RHOT_RK(KA,IA,JA)
    real(DP), intent(out)   :: PROG_RK(KA,IA,JA,VA)
    real(DP), intent(inout) :: mflx_hi(KA,IA,JA,3)
    real(DP), intent(out)   :: tflx_hi(KA,IA,JA,3)
    real(DP), intent(in), target :: DENS0(KA,IA,JA), MOMZ0(KA,IA,JA), MOMX0(KA,IA,JA), MOMY0(KA,IA,JA), RHOT0(KA,IA,JA)
    real(DP), intent(in)    :: DENS(KA,IA,JA), MOMZ(KA,IA,JA), MOMX(KA,IA,JA), MOMY(KA,IA,JA), RHOT(KA,IA,JA)
    real(DP), intent(in)    :: DENS_t(KA,IA,JA), MOMZ_t(KA,IA,JA), MOMX_t(KA,IA,JA), MOMY_t(KA,IA,JA), RHOT_t(KA,IA,JA)
    real(DP), intent(in)    :: PROG0(KA,IA,JA,VA), PROG(KA,IA,JA,VA)
    real(DP), intent(in)    :: DPRES0(KA,IA,JA), RT2P(KA,IA,JA), CORIOLI(1,IA,JA)
    real(DP), intent(in)    :: num_diff(KA,IA,JA,5,3), wdamp_coef(KA), divdmp_coef, DDIV(KA,IA,JA)
    logical,  intent(in)    :: FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_ALONG_STREAM
    real(DP), intent(in)    :: CDZ(KA), FDZ(KA-1), FDX(IA-1), FDY(JA-1)
    real(DP), intent(in)    :: RCDZ(KA), RCDX(IA), RCDY(JA), RFDZ(KA-1), RFDX(IA-1), RFDY(JA-1)
    real(DP), intent(in)    :: PHI(KA,IA,JA), GSQRT(KA,IA,JA,7), J13G(KA,IA,JA,7), J23G(KA,IA,JA,7), J33G
    real(DP), intent(in)    :: MAPF(IA,JA,2,4), REF_dens(KA,IA,JA), REF_rhot(KA,IA,JA)
    logical,  intent(in)    :: BND_W, BND_E, BND_S, BND_N
    real(DP), intent(in)    :: dtrk
    logical,  intent(in)    :: last
    real(DP) :: dummy(1)
    call check(c_tstep_short(ctx, DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, dummy, mflx_hi, tflx_hi, &
         DENS0, MOMZ0, MOMX0, MOMY0, RHOT0, DENS, MOMZ, MOMX, MOMY, RHOT, &
         DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, dummy, dummy, &
         DPRES0, RT2P, CORIOLI, num_diff, wdamp_coef, divdmp_coef, DDIV, &
         l2i(FLAG_FCT_MOMENTUM), l2i(FLAG_FCT_T), l2i(FLAG_FCT_ALONG_STREAM), &
         CDZ, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY, &
         PHI, GSQRT, J13G, J23G, J33G, MAPF, REF_dens, REF_rhot, &
         l2i(BND_W), l2i(BND_E), l2i(BND_S), l2i(BND_N), dtrk, l2i(last)), "ATMOS_DYN_Tstep_short_fvm_hevi_hip")
    ! HEVI has no extra prognostics (VA_FVM_HEVI = 0, HEVI.F90:72): PROG_RK is empty
  end subroutine ATMOS_DYN_Tstep_short_fvm_hevi_hip

  !> explicit HEVE stage with the dummies of the abstract interface `short`
  !! (scale_atmos_dyn_tstep_short_fvm_heve.F90:109-1516), with its momentum and theta FCT
  subroutine ATMOS_DYN_Tstep_short_fvm_heve_hip( &
       DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, &
       PROG_RK,                                     &
       mflx_hi, tflx_hi,                            &
       DENS0,   MOMZ0,   MOMX0,   MOMY0,   RHOT0,   &
       DENS,    MOMZ,    MOMX,    MOMY,    RHOT,    &
       DENS_t,  MOMZ_t,  MOMX_t,  MOMY_t,  RHOT_t,  &
       PROG0, PROG,                                 &
       DPRES0, RT2P, CORIOLI,                       &
       num_diff, wdamp_coef, divdmp_coef, DDIV,     &
       FLAG_FCT_MOMENTUM, FLAG_FCT_T,               &
       FLAG_FCT_ALONG_STREAM,                       &
       CDZ, FDZ, FDX, FDY,                          &
       RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY,          &
       PHI, GSQRT, J13G, J23G, J33G, MAPF,          &
       REF_dens, REF_rhot,                          &
       BND_W, BND_E, BND_S, BND_N,                  &
       dtrk, last                                   )
    real(DP), intent(out)   :: DENS_RK(KA,IA,JA), MOMZ_RK(KA,IA,JA), MOMX_RK(KA,IA,JA), MOMY_RK(KA,IA,JA), RHOT_RK(KA,IA,JA)
    real(DP), intent(out)   :: PROG_RK(KA,IA,JA,VA)
    real(DP), intent(inout) :: mflx_hi(KA,IA,JA,3)
    real(DP), intent(out)   :: tflx_hi(KA,IA,JA,3)
    real(DP), intent(in), target :: DENS0(KA,IA,JA), MOMZ0(KA,IA,JA), MOMX0(KA,IA,JA), MOMY0(KA,IA,JA), RHOT0(KA,IA,JA)
    real(DP), intent(in)    :: DENS(KA,IA,JA), MOMZ(KA,IA,JA), MOMX(KA,IA,JA), MOMY(KA,IA,JA), RHOT(KA,IA,JA)
    real(DP), intent(in)    :: DENS_t(KA,IA,JA), MOMZ_t(KA,IA,JA), MOMX_t(KA,IA,JA), MOMY_t(KA,IA,JA), RHOT_t(KA,IA,JA)
    real(DP), intent(in)    :: PROG0(KA,IA,JA,VA), PROG(KA,IA,JA,VA)
    real(DP), intent(in)    :: DPRES0(KA,IA,JA), RT2P(KA,IA,JA), CORIOLI(1,IA,JA)
    real(DP), intent(in)    :: num_diff(KA,IA,JA,5,3), wdamp_coef(KA), divdmp_coef, DDIV(KA,IA,JA)
    logical,  intent(in)    :: FLAG_FCT_MOMENTUM, FLAG_FCT_T, FLAG_FCT_ALONG_STREAM
    real(DP), intent(in)    :: CDZ(KA), FDZ(KA-1), FDX(IA-1), FDY(JA-1)
    real(DP), intent(in)    :: RCDZ(KA), RCDX(IA), RCDY(JA), RFDZ(KA-1), RFDX(IA-1), RFDY(JA-1)
    real(DP), intent(in)    :: PHI(KA,IA,JA), GSQRT(KA,IA,JA,7), J13G(KA,IA,JA,7), J23G(KA,IA,JA,7), J33G
    real(DP), intent(in)    :: MAPF(IA,JA,2,4), REF_dens(KA,IA,JA), REF_rhot(KA,IA,JA)
    logical,  intent(in)    :: BND_W, BND_E, BND_S, BND_N
    real(DP), intent(in)    :: dtrk
    logical,  intent(in)    :: last
    real(DP) :: dummy(1)
    call check(c_tstep_short_heve(ctx, DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, dummy, mflx_hi, tflx_hi, &
         DENS0, MOMZ0, MOMX0, MOMY0, RHOT0, DENS, MOMZ, MOMX, MOMY, RHOT, &
         DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t, dummy, dummy, &
         DPRES0, RT2P, CORIOLI, num_diff, wdamp_coef, divdmp_coef, DDIV, &
         l2i(FLAG_FCT_MOMENTUM), l2i(FLAG_FCT_T), l2i(FLAG_FCT_ALONG_STREAM), &
         CDZ, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY, &
         PHI, GSQRT, J13G, J23G, J33G, MAPF, REF_dens, REF_rhot, &
         l2i(BND_W), l2i(BND_E), l2i(BND_S), l2i(BND_N), dtrk, l2i(last)), "ATMOS_DYN_Tstep_short_fvm_heve_hip")
    ! HEVE has no extra prognostics (VA_FVM_HEVE = 0, HEVE.F90:60): PROG_RK is empty
  end subroutine ATMOS_DYN_Tstep_short_fvm_heve_hip

  !> tracer stage with the dummies of the interface `step`
  subroutine ATMOS_DYN_Tstep_tracer_fvm_heve_hip( &
       QTRCo, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff, GSQRT, MAPF, &
       CDZ, RCDZ, RCDX, RCDY, dtl, FLAG_FCT_TRACER, FLAG_FCT_ALONG_STREAM)
    real(DP), intent(inout) :: QTRCo(KA,IA,JA)
    real(DP), intent(in) :: QTRC(KA,IA,JA), QTRC0(KA,IA,JA), RHOQ_t(KA,IA,JA), DENS0(KA,IA,JA), DENS(KA,IA,JA)
    real(DP), intent(in) :: mflx_hi(KA,IA,JA,3), num_diff(KA,IA,JA,3), GSQRT(KA,IA,JA,7), MAPF(IA,JA,2)
    real(DP), intent(in) :: CDZ(KA), RCDZ(KA), RCDX(IA), RCDY(JA), dtl
    logical, intent(in) :: FLAG_FCT_TRACER, FLAG_FCT_ALONG_STREAM
    call check(c_tstep_tracer(ctx, QTRCo, QTRC, QTRC0, RHOQ_t, DENS0, DENS, mflx_hi, num_diff, GSQRT, MAPF, &
         CDZ, RCDZ, RCDX, RCDY, dtl, l2i(FLAG_FCT_TRACER), l2i(FLAG_FCT_ALONG_STREAM)), "ATMOS_DYN_Tstep_tracer_fvm_heve_hip")
  end subroutine ATMOS_DYN_Tstep_tracer_fvm_heve_hip

  !> ATMOS_DYN_numfilter_coef (scale_atmos_dyn_common.F90:385-985), ND_ORDER = 1
  subroutine ATMOS_DYN_numfilter_coef_hip(num_diff, DENS, MOMZ, MOMX, MOMY, RHOT, CDZ, CDX, CDY, FDZ, FDX, FDY, DT, &
       REF_dens, REF_pott, ND_COEF, ND_ORDER, ND_SFC_FACT, ND_USE_RS)
    real(DP), intent(out) :: num_diff(KA,IA,JA,5,3)
    real(DP), intent(in) :: DENS(KA,IA,JA), MOMZ(KA,IA,JA), MOMX(KA,IA,JA), MOMY(KA,IA,JA), RHOT(KA,IA,JA)
    real(DP), intent(in) :: CDZ(KA), CDX(IA), CDY(JA), FDZ(KA-1), FDX(IA-1), FDY(JA-1), DT
    real(DP), intent(in) :: REF_dens(KA,IA,JA), REF_pott(KA,IA,JA), ND_COEF, ND_SFC_FACT
    integer, intent(in) :: ND_ORDER
    logical, intent(in) :: ND_USE_RS
    call check(c_numfilter(ctx, num_diff, DENS, MOMZ, MOMX, MOMY, RHOT, CDZ, CDX, CDY, FDZ, FDX, FDY, DT, &
         REF_dens, REF_pott, ND_COEF, int(ND_ORDER, c_int32_t), ND_SFC_FACT, l2i(ND_USE_RS)), "ATMOS_DYN_numfilter_coef_hip")
  end subroutine ATMOS_DYN_numfilter_coef_hip

  !> ATMOS_DYN_numfilter_coef_q (scale_atmos_dyn_common.F90:989-1200), ND_ORDER = 1
  subroutine ATMOS_DYN_numfilter_coef_q_hip(num_diff_q, DENS, QTRC, is_qv, CDZ, CDX, CDY, dt, REF_qv, iq, &
       ND_COEF, ND_ORDER, ND_SFC_FACT, ND_USE_RS)
    real(DP), intent(out) :: num_diff_q(KA,IA,JA,3)
    real(DP), intent(in) :: DENS(KA,IA,JA), QTRC(KA,IA,JA)
    logical, intent(in) :: is_qv
    real(DP), intent(in) :: CDZ(KA), CDX(IA), CDY(JA), dt, REF_qv(KA,IA,JA)
    integer, intent(in) :: iq
    real(DP), intent(in) :: ND_COEF, ND_SFC_FACT
    integer, intent(in) :: ND_ORDER
    logical, intent(in) :: ND_USE_RS
    call check(c_numfilter_q(ctx, num_diff_q, DENS, QTRC, l2i(is_qv), CDZ, CDX, CDY, dt, REF_qv, int(iq, c_int32_t), &
         ND_COEF, int(ND_ORDER, c_int32_t), ND_SFC_FACT, l2i(ND_USE_RS)), "ATMOS_DYN_numfilter_coef_q_hip")
  end subroutine ATMOS_DYN_numfilter_coef_q_hip

  integer(c_int32_t) function l2i(l)
    logical, intent(in) :: l
    l2i = 0
    if (l) l2i = 1
  end function l2i

  subroutine check(ierr, where)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (ierr == 0) return
    call c_f_pointer(c_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
       if (msg(n+1) == c_null_char) exit
       n = n + 1
    end do
    write(*,'(A,A,A,I0,A,1024A1)') 'xxx [', where, '] libscalehip error ', ierr, ': ', msg(1:n)
    error stop 1
  end subroutine check

end module scale_atmos_dyn_hip
