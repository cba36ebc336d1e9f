!-------------------------------------------------------------------------------
! Test driver of the large-step drop-in ATMOS_DYN_Tstep_large_fvm_heve_hip:
! reads a case from a stream file written by tests/test_gpu_fortran.py, calls
! the procedure n times exactly as ATMOS_DYN_driver calls the `large`
! procedure pointer (scale-rm/src/atmos/mod_atmos_dyn_driver.F90), with zero
! physics tendencies and no boundary damping, and writes the state and the
! tracers after them.
!
! usage: large_driver <input.bin> <output.bin>
! input:  int32 KA IA JA KS KE IS IE JS JE scheme scheme_tracer QA nsteps fct
!         real64 arrays in the order of the READ statements below
!-------------------------------------------------------------------------------
program large_driver
  use iso_fortran_env, only: real64, int32
  use scale_atmos_dyn_hip
  implicit none

  integer(int32) :: hdr(14)
  integer :: KS, KE, IS, IE, JS, JE, nsteps, n
  character(len=512) :: fin, fout
  real(real64), allocatable :: DENS(:,:,:), MOMZ(:,:,:), MOMX(:,:,:), MOMY(:,:,:), RHOT(:,:,:)
  real(real64), allocatable :: QTRC(:,:,:,:), PROG(:,:,:,:), QTRC0(:,:,:,:)
  real(real64), allocatable :: DENS_av(:,:,:), MOMZ_av(:,:,:), MOMX_av(:,:,:), MOMY_av(:,:,:), RHOT_av(:,:,:)
  real(real64), allocatable :: QTRC_av(:,:,:,:), mflx_hi(:,:,:,:), tflx_hi(:,:,:,:), num_diff(:,:,:,:,:), num_diff_q(:,:,:,:)
  real(real64), allocatable :: zero3(:,:,:), RHOQ_tp(:,:,:,:), CORIOLI(:,:)
  real(real64), allocatable :: CDZ(:), CDX(:), CDY(:), FDZ(:), FDX(:), FDY(:)
  real(real64), allocatable :: RCDZ(:), RCDX(:), RCDY(:), RFDZ(:), RFDX(:), RFDY(:)
  real(real64), allocatable :: PHI(:,:,:), GSQRT(:,:,:,:), J13G(:,:,:,:), J23G(:,:,:,:), MAPF(:,:,:,:)
  real(real64), allocatable :: AQ(:), REF_dens(:,:,:), REF_pott(:,:,:), REF_qv(:,:,:), REF_pres(:,:,:)
  real(real64), allocatable :: DAMP_QTRC(:,:,:,:), wdamp_coef(:)
  real(real64) :: J33G, DTL, DTS, ND_COEF
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  KS = hdr(4); KE = hdr(5); IS = hdr(6); IE = hdr(7); JS = hdr(8); JE = hdr(9)
  nsteps = hdr(13)
  call ATMOS_DYN_hip_setup(hdr(1), hdr(2), hdr(3), KS, KE, IS, IE, JS, JE, hdr(10), hdr(11), 0)
  call ATMOS_DYN_hip_setup_large(hdr(12), 0, 4, 0, 0, 0)

  allocate(DENS(KA,IA,JA), MOMZ(KA,IA,JA), MOMX(KA,IA,JA), MOMY(KA,IA,JA), RHOT(KA,IA,JA))
  allocate(QTRC(KA,IA,JA,QA), PROG(KA,IA,JA,0), QTRC0(KA,IA,JA,QA))
  allocate(DENS_av(KA,IA,JA), MOMZ_av(KA,IA,JA), MOMX_av(KA,IA,JA), MOMY_av(KA,IA,JA), RHOT_av(KA,IA,JA))
  allocate(QTRC_av(KA,IA,JA,QA), mflx_hi(KA,IA,JA,3), tflx_hi(KA,IA,JA,3), num_diff(KA,IA,JA,5,3), num_diff_q(KA,IA,JA,3))
  allocate(zero3(KA,IA,JA), RHOQ_tp(KA,IA,JA,QA), CORIOLI(IA,JA))
  allocate(CDZ(KA), CDX(IA), CDY(JA), FDZ(KA-1), FDX(IA-1), FDY(JA-1))
  allocate(RCDZ(KA), RCDX(IA), RCDY(JA), RFDZ(KA-1), RFDX(IA-1), RFDY(JA-1))
  allocate(PHI(KA,IA,JA), GSQRT(KA,IA,JA,7), J13G(KA,IA,JA,7), J23G(KA,IA,JA,7), MAPF(IA,JA,2,4))
  allocate(AQ(QA), REF_dens(KA,IA,JA), REF_pott(KA,IA,JA), REF_qv(KA,IA,JA), REF_pres(KA,IA,JA))
  allocate(DAMP_QTRC(KA,IA,JA,0), wdamp_coef(KA))

  read(u) DENS, MOMZ, MOMX, MOMY, RHOT, QTRC
  read(u) REF_dens, REF_pott, REF_pres, CORIOLI, wdamp_coef
  read(u) CDZ, CDX, CDY, FDZ, FDX, FDY, RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY
  read(u) PHI, GSQRT, J13G, J23G, J33G, MAPF
  read(u) DTL, DTS, ND_COEF
  close(u)
  zero3 = 0.0_real64
  RHOQ_tp = 0.0_real64
  AQ = 0.0_real64
  REF_qv = 0.0_real64

  do n = 1, nsteps
     QTRC0 = QTRC
     call ATMOS_DYN_Tstep_large_fvm_heve_hip( &
          DENS, MOMZ, MOMX, MOMY, RHOT, QTRC, PROG,             &
          DENS_av, MOMZ_av, MOMX_av, MOMY_av, RHOT_av, QTRC_av, &
          mflx_hi, tflx_hi, num_diff, num_diff_q, QTRC0,        &
          zero3, zero3, zero3, zero3, zero3, RHOQ_tp,           &
          CORIOLI,                                              &
          CDZ, CDX, CDY, FDZ, FDX, FDY,                         &
          RCDZ, RCDX, RCDY, RFDZ, RFDX, RFDY,                   &
          PHI, GSQRT, J13G, J23G, J33G, MAPF,                   &
          AQ, AQ, AQ, AQ,                                       &
          REF_dens, REF_pott, REF_qv, REF_pres,                 &
          .false., .false., .false., .false.,                   &
          ND_COEF, 0.0_real64, 1, 1.0_real64, .true.,           &
          0, 1.0_real64,                                        &
          zero3, zero3, zero3, zero3, zero3, DAMP_QTRC,         &
          zero3, zero3, zero3, zero3, zero3, DAMP_QTRC,         &
          wdamp_coef, 0.0_real64,                               &
          .false., .false., .false., hdr(14) /= 0, .true.,      &
          .false., 0, DTL, DTS, n == nsteps)
  end do

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) DENS, MOMZ, MOMX, MOMY, RHOT, QTRC
  close(u)
  call ATMOS_DYN_hip_finalize
end program large_driver
