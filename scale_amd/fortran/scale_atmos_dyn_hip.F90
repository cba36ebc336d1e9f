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

  integer, parameter :: DP = c_double

  type, bind(C) :: scale_hip_grid
     integer(c_int32_t) :: KA, IA, JA
     integer(c_int32_t) :: KS, KE, IS, IE, JS, JE
     integer(c_int32_t) :: KHALO, IHALO, JHALO
     integer(c_int32_t) :: flux_scheme, flux_scheme_tracer
     integer(c_int32_t) :: device
  end type scale_hip_grid

  type(c_ptr), save :: ctx = c_null_ptr
  integer, save, public :: KA, IA, JA, VA = 0

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
  end subroutine ATMOS_DYN_hip_finalize

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
