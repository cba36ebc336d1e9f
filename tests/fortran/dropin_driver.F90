!-------------------------------------------------------------------------------
! Test driver of the Fortran drop-in module scale_atmos_dyn_hip: reads one
! short-step stage's inputs from a stream file written by
! tests/test_gpu_fortran.py, calls ATMOS_DYN_Tstep_short_fvm_hevi_hip (or,
! with short = 1, ATMOS_DYN_Tstep_short_fvm_heve_hip) exactly as
! ATMOS_DYN_tinteg_short calls its procedure pointer
! (tinteg_short_rk4.F90:286-311), then one ATMOS_DYN_Tstep_tracer_fvm_heve_hip
! stage, and writes the outputs.
!
! usage: dropin_driver <input.bin> <output.bin>
! input:  int32 KA IA JA KS KE IS IE JS JE scheme scheme_tracer fct short
!         real64 arrays in the order of the READ statements below
!-------------------------------------------------------------------------------
program dropin_driver
  use iso_fortran_env, only: real64, int32
  use scale_atmos_dyn_hip
  implicit none

  integer(int32) :: hdr(13)
  integer :: KS, KE, IS, IE, JS, JE
  character(len=512) :: fin, fout
  real(real64), allocatable :: DENS_RK(:,:,:), MOMZ_RK(:,:,:), MOMX_RK(:,:,:), MOMY_RK(:,:,:), RHOT_RK(:,:,:)
  real(real64), allocatable :: PROG_RK(:,:,:,:), PROG0(:,:,:,:), PROG(:,:,:,:)
  real(real64), allocatable :: mflx_hi(:,:,:,:), tflx_hi(:,:,:,:)
  real(real64), allocatable :: DENS0(:,:,:), MOMZ0(:,:,:), MOMX0(:,:,:), MOMY0(:,:,:), RHOT0(:,:,:)
  real(real64), allocatable :: DENS(:,:,:), MOMZ(:,:,:), MOMX(:,:,:), MOMY(:,:,:), RHOT(:,:,:)
  real(real64), allocatable :: DENS_t(:,:,:), MOMZ_t(:,:,:), MOMX_t(:,:,:), MOMY_t(:,:,:), RHOT_t(:,:,:)
  real(real64), allocatable :: DPRES0(:,:,:), RT2P(:,:,:), CORIOLI(:,:,:), num_diff(:,:,:,:,:), wdamp_coef(:), DDIV(:,:,:)
  real(real64), allocatable :: CDZ(:), FDZ(:), FDX(:), FDY(:), RCDZ(:), RCDX(:), RCDY(:), RFDZ(:), RFDX(:), RFDY(:)
  real(real64), allocatable :: PHI(:,:,:), GSQRT(:,:,:,:), J13G(:,:,:,:), J23G(:,:,:,:), MAPF(:,:,:,:)
  real(real64), allocatable :: REF_dens(:,:,:), REF_rhot(:,:,:)
  real(real64), allocatable :: QTRCo(:,:,:), QTRC(:,:,:), RHOQ_t(:,:,:), nd_q(:,:,:,:), MAPF_XY(:,:,:)
  real(real64) :: J33G, divdmp_coef, dtrk, dtl
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  KA = hdr(1); IA = hdr(2); JA = hdr(3)
  KS = hdr(4); KE = hdr(5); IS = hdr(6); IE = hdr(7); JS = hdr(8); JE = hdr(9)

  allocate(DENS_RK(KA,IA,JA), MOMZ_RK(KA,IA,JA), MOMX_RK(KA,IA,JA), MOMY_RK(KA,IA,JA), RHOT_RK(KA,IA,JA))
  allocate(PROG_RK(KA,IA,JA,0), PROG0(KA,IA,JA,0), PROG(KA,IA,JA,0))
  allocate(mflx_hi(KA,IA,JA,3), tflx_hi(KA,IA,JA,3))
  allocate(DENS0(KA,IA,JA), MOMZ0(KA,IA,JA), MOMX0(KA,IA,JA), MOMY0(KA,IA,JA), RHOT0(KA,IA,JA))
  allocate(DENS(KA,IA,JA), MOMZ(KA,IA,JA), MOMX(KA,IA,JA), MOMY(KA,IA,JA), RHOT(KA,IA,JA))
  allocate(DENS_t(KA,IA,JA), MOMZ_t(KA,IA,JA), MOMX_t(KA,IA,JA), MOMY_t(KA,IA,JA), RHOT_t(KA,IA,JA))
  allocate(DPRES0(KA,IA,JA), RT2P(KA,IA,JA), CORIOLI(1,IA,JA), num_diff(KA,IA,JA,5,3), wdamp_coef(KA), DDIV(KA,IA,JA))
  allocate(CDZ(KA), FDZ(KA-1), FDX(IA-1), FDY(JA-1), RCDZ(KA), RCDX(IA), RCDY(JA), RFDZ(KA-1), RFDX(IA-1), RFDY(JA-1))
  allocate(PHI(KA,IA,JA), GSQRT(KA,IA,JA,7), J13G(KA,IA,JA,7), J23G(KA,IA,JA,7), MAPF(IA,JA,2,4))
  allocate(REF_dens(KA,IA,JA), REF_rhot(KA,IA,JA))
  allocate(QTRCo(KA,IA,JA), QTRC(KA,IA,JA), RHOQ_t(KA,IA,JA), nd_q(KA,IA,JA,3), MAPF_XY(IA,JA,2))

  read(u) DENS0, MOMZ0, MOMX0, MOMY0, RHOT0
  read(u) DENS, MOMZ, MOMX, MOMY, RHOT
  read(u) DENS_t, MOMZ_t, MOMX_t, MOMY_t, RHOT_t
  read(u) DPRES0, RT2P, CORIOLI, num_diff, wdamp_coef, divdmp_coef, DDIV
  read(u) CDZ, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY
  read(u) PHI, GSQRT, J13G, J23G, J33G, MAPF, REF_dens, REF_rhot, dtrk
  read(u) QTRCo, QTRC, RHOQ_t, nd_q, dtl
  close(u)
  MAPF_XY(:,:,:) = MAPF(:,:,:,1)

  call ATMOS_DYN_hip_setup(KA, IA, JA, KS, KE, IS, IE, JS, JE, hdr(10), hdr(11), 0)

  mflx_hi(:,:,:,:) = 0.0_real64
  if (hdr(13) == 1) then
     call ATMOS_DYN_Tstep_short_fvm_heve_hip( &
          DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, &
          PROG_RK,                                     &
          mflx_hi, tflx_hi,                            &
          DENS0,   MOMZ0,   MOMX0,   MOMY0,   RHOT0,   &
          DENS,    MOMZ,    MOMX,    MOMY,    RHOT,    &
          DENS_t,  MOMZ_t,  MOMX_t,  MOMY_t,  RHOT_t,  &
          PROG0, PROG,                                 &
          DPRES0, RT2P, CORIOLI,                       &
          num_diff, wdamp_coef, divdmp_coef, DDIV,     &
          .false., .false., .true.,                    &
          CDZ, FDZ, FDX, FDY,                          &
          RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY,          &
          PHI, GSQRT, J13G, J23G, J33G, MAPF,          &
          REF_dens, REF_rhot,                          &
          .false., .false., .false., .false.,          &
          dtrk, .false.                                )
  else
     call ATMOS_DYN_Tstep_short_fvm_hevi_hip( &
          DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, &
          PROG_RK,                                     &
          mflx_hi, tflx_hi,                            &
          DENS0,   MOMZ0,   MOMX0,   MOMY0,   RHOT0,   &
          DENS,    MOMZ,    MOMX,    MOMY,    RHOT,    &
          DENS_t,  MOMZ_t,  MOMX_t,  MOMY_t,  RHOT_t,  &
          PROG0, PROG,                                 &
          DPRES0, RT2P, CORIOLI,                       &
          num_diff, wdamp_coef, divdmp_coef, DDIV,     &
          .false., .false., .true.,                    &
          CDZ, FDZ, FDX, FDY,                          &
          RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY,          &
          PHI, GSQRT, J13G, J23G, J33G, MAPF,          &
          REF_dens, REF_rhot,                          &
          .false., .false., .false., .false.,          &
          dtrk, .false.                                )
  end if

  ! tracer stage driven by the mass flux just computed, after the periodic
  ! halo exchange the large step applies to mflx_hi (LARGE.F90:1026-1031)
  do u = 1, 3
     call halo_periodic(mflx_hi(:,:,:,u))
  end do
  call ATMOS_DYN_Tstep_tracer_fvm_heve_hip( &
       QTRCo, QTRC, QTRC, RHOQ_t, DENS0, DENS, mflx_hi, nd_q, GSQRT, MAPF_XY, &
       CDZ, RCDZ, RCDX, RCDY, dtl, hdr(12) /= 0, .true.)

  call ATMOS_DYN_hip_finalize

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) DENS_RK, MOMZ_RK, MOMX_RK, MOMY_RK, RHOT_RK, mflx_hi, tflx_hi, QTRCo
  close(u)

contains

  ! periodic lateral halo of width IS-1 on one rank, faces then corners
  subroutine halo_periodic(a)
    real(real64), intent(inout) :: a(KA,IA,JA)
    integer :: h
    h = IS - 1
    a(:,1:h,JS:JE)     = a(:,IE-h+1:IE,JS:JE)
    a(:,IE+1:IA,JS:JE) = a(:,IS:IS+h-1,JS:JE)
    h = JS - 1
    a(:,:,1:h)     = a(:,:,JE-h+1:JE)
    a(:,:,JE+1:JA) = a(:,:,JS:JS+h-1)
  end subroutine halo_periodic

end program dropin_driver
