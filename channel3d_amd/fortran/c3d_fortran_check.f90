!-----------------------------------------------------------------------
! c3d_fortran_check -- drives the library from Fortran the way Channel3d's
! ChannelIterate does (src/CFD_2nd/m_ChannelSystem.f90:91-130), through
! both forms of m_channel3d_gpu:
!   1. one RK substep (ns=1) with the per-routine wrappers on host arrays,
!      in the reference's call order and with its scratch aliasing
!      (RhsX/prsrc => RealArr1, RhsY => RealArr2, RhsZ/prphiHalo => RealHalo);
!   2. nsteps time steps on the device-resident state, then clcStat, the
!      stats file <fields_out>.stats<nsteps> and the restart file
!      <fields_out>.RestartForCheck<nsteps>;
!   3. a fresh context that reads the restart file back;
!   4. the LES scheme (LES_type, FilterType): Clc_SGS_Vis and one substep
!      with the per-routine wrappers, then nsteps device-resident steps.
! usage: c3d_fortran_check <case.txt> <fields_in.bin> <fields_out.bin>
!   case.txt: nxc nyc nzc / xlx yly zlz / xnu dt / nsteps / LES_type FilterType
!   fields_in.bin (stream): yp(0:nyc+1), ux, uy, uz, pressure (halo layout)
!   fields_out.bin (stream): ux, uy, uz, pressure after 1., after 2., after 3.;
!     then nut, and ux, uy, uz, pressure after each half of 4.
!-----------------------------------------------------------------------
program c3d_fortran_check
  use iso_c_binding
  use channel3d_c
  use m_channel3d_gpu
  implicit none
  type(c3d_params) :: p
  integer :: nxc, nyc, nzc, nsteps, u, n, les_type, filter_type
  real(c_double), allocatable :: yp(:)
  real(c_double), allocatable, dimension(:,:,:) :: ux, uy, uz, pressure, RealHalo
  real(c_double), allocatable, dimension(:,:,:) :: ux0, uy0, uz0, pr0, nut
  real(c_double), allocatable, dimension(:,:,:) :: HistXOld, HistYOld, HistZOld, RealArr1, RealArr2
  real(c_double), allocatable, dimension(:,:) :: uy_ym, duy_ym
  real(c_double) :: divmax1, cflmp, divmax2, vmaxabs(3), SimTime
  character(len=512) :: fcase, fin, fout

  call get_command_argument(1, fcase)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  open(newunit=u, file=trim(fcase), status='old', action='read')
  read(u, *) nxc, nyc, nzc
  p%nxc = nxc; p%nyc = nyc; p%nzc = nzc
  read(u, *) p%xlx, p%yly, p%zlz
  read(u, *) p%xnu, p%dt
  read(u, *) nsteps
  read(u, *) les_type, filter_type
  close(u)
  p%bc = [C3D_BC_PERIOD, C3D_BC_PERIOD, C3D_BC_NOSLIP, C3D_BC_NOSLIP, C3D_BC_PERIOD, C3D_BC_PERIOD]
  p%uxbc = 0.0d0; p%uybc = 0.0d0; p%uzbc = 0.0d0; p%gravity = 0.0d0
  p%is_ux_const = 1; p%ubulk = 1.0d0; p%ybulk1 = 0.0d0; p%ybulk2 = p%yly
  p%ischeme = 3; p%precision = C3D_FP64; p%order = 2; p%u_crf = 0.0d0

  allocate(yp(0:nyc+1))
  allocate(ux(0:nxc+2, 0:nyc+2, 0:nzc+2), uy(0:nxc+2, 0:nyc+2, 0:nzc+2), &
           uz(0:nxc+2, 0:nyc+2, 0:nzc+2), pressure(0:nxc+2, 0:nyc+2, 0:nzc+2), &
           RealHalo(0:nxc+2, 0:nyc+2, 0:nzc+2), ux0(0:nxc+2, 0:nyc+2, 0:nzc+2), &
           uy0(0:nxc+2, 0:nyc+2, 0:nzc+2), uz0(0:nxc+2, 0:nyc+2, 0:nzc+2), &
           pr0(0:nxc+2, 0:nyc+2, 0:nzc+2), nut(0:nxc+2, 0:nyc+2, 0:nzc+2))
  allocate(HistXOld(nxc, nyc, nzc), HistYOld(nxc, nyc, nzc), HistZOld(nxc, nyc, nzc), &
           RealArr1(nxc, nyc, nzc), RealArr2(nxc, nyc, nzc))
  allocate(uy_ym(nxc, nzc), duy_ym(nxc, nzc))
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) yp, ux0, uy0, uz0, pr0
  close(u)

  ! ---- 1. per-routine wrappers on host arrays, one substep -------------
  call gpu_init(p, yp)
  ux = ux0; uy = uy0; uz = uz0; pressure = pr0
  HistXOld = 0; HistYOld = 0; HistZOld = 0; RealArr1 = 0; RealArr2 = 0; RealHalo = 0
  uy_ym = 0; duy_ym = 0
  call gpu_PMcoeUpdate(1)
  call gpu_Update_uy_ym(uy_ym, duy_ym)
  call gpu_clcRhsX(ux, uy, uz, RealArr1, HistXOld, pressure)
  call gpu_clcRhsY(ux, uy, uz, RealArr2, HistYOld, pressure)
  call gpu_clcRhsZ(ux, uy, uz, RealHalo, HistZOld, pressure)
  call gpu_clcU1Hat(ux, RealArr1)
  call gpu_clcU2Hat(uy, RealArr2, duy_ym)
  call gpu_clcU3Hat(uz, RealHalo)
  call gpu_SetBC_and_UpdateHaloForPrSrc(ux, uy, uz, uy_ym)
  call gpu_clcPrSrc(ux, uy, uz, RealArr1, pressure, divmax1)
  call gpu_clcPPE(RealArr1, RealHalo)
  call gpu_SetBC_and_UpdateHalo_pr(RealHalo)
  call gpu_FluidVelUpdate(RealHalo, ux, uy, uz)
  call gpu_PressureUpdate(pressure, RealHalo)
  call gpu_SetBC_and_UpdateHalo(ux, uy, uz, uy_ym)
  call gpu_SetBC_and_UpdateHalo_pr(pressure)
  call gpu_CheckDivergence(ux, uy, uz, divmax2)
  write(*, '(A,ES12.4,A,ES12.4)') "per-routine substep: divmax1 ", divmax1, "  divmax2 ", divmax2
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) ux, uy, uz, pressure
  call gpu_finalize()

  ! ---- 2. device-resident time steps ------------------------------------
  call gpu_init(p, yp)
  HistXOld = 0; HistYOld = 0; HistZOld = 0
  call gpu_upload_state(ux0, uy0, uz0, pr0, HistXOld, HistYOld, HistZOld)
  do n = 1, nsteps
    call gpu_ChannelIterate(cflmp, divmax1, divmax2, vmaxabs, SimTime)
  end do
  write(*, '(A,I0,A,ES12.4,A,ES12.4,A,F10.6)') "device-resident ", nsteps, " steps: cfl ", cflmp, &
      "  divmax2 ", divmax2, "  SimTime ", SimTime
  call gpu_download_state(ux, uy, uz, pressure, HistXOld, HistYOld, HistZOld)
  write(u) ux, uy, uz, pressure
  call gpu_clcStat()
  call gpu_write_stat(trim(fout)//".", nsteps, 1)
  call gpu_write_restart(nsteps, trim(fout)//".", "Check", .false.)
  call gpu_finalize()

  ! ---- 3. a fresh context from the restart file --------------------------
  call gpu_init(p, yp)
  call gpu_read_restart(nsteps, trim(fout)//".", "Check", .false.)
  call gpu_download_state(ux, uy, uz, pressure, HistXOld, HistYOld, HistZOld)
  write(u) ux, uy, uz, pressure
  call gpu_finalize()

  ! ---- 4. LES: the scheme of m_TScheme.f90:179-206 behind the same calls ---
  call gpu_init(p, yp)
  call gpu_set_les(les_type, filter_type)
  ux = ux0; uy = uy0; uz = uz0; pressure = pr0; nut = 0
  HistXOld = 0; HistYOld = 0; HistZOld = 0; RealArr1 = 0; RealArr2 = 0; RealHalo = 0
  uy_ym = 0; duy_ym = 0
  call gpu_Clc_SGS_Vis(ux, uy, uz, nut)
  write(u) nut
  call gpu_PMcoeUpdate(1)
  call gpu_Update_uy_ym(uy_ym, duy_ym)
  call gpu_clcRhsX(ux, uy, uz, RealArr1, HistXOld, pressure)
  call gpu_clcRhsY(ux, uy, uz, RealArr2, HistYOld, pressure)
  call gpu_clcRhsZ(ux, uy, uz, RealHalo, HistZOld, pressure)
  call gpu_clcU1Hat(ux, RealArr1)
  call gpu_clcU2Hat(uy, RealArr2, duy_ym)
  call gpu_clcU3Hat(uz, RealHalo)
  call gpu_SetBC_and_UpdateHaloForPrSrc(ux, uy, uz, uy_ym)
  call gpu_clcPrSrc(ux, uy, uz, RealArr1, pressure, divmax1)
  call gpu_clcPPE(RealArr1, RealHalo)
  call gpu_SetBC_and_UpdateHalo_pr(RealHalo)
  call gpu_FluidVelUpdate(RealHalo, ux, uy, uz)
  call gpu_PressureUpdate(pressure, RealHalo)
  call gpu_SetBC_and_UpdateHalo(ux, uy, uz, uy_ym)
  call gpu_SetBC_and_UpdateHalo_pr(pressure)
  write(*, '(A,I0,A,ES12.4)') "LES_type ", les_type, " per-routine substep: divmax1 ", divmax1
  write(u) ux, uy, uz, pressure
  call gpu_finalize()
  call gpu_init(p, yp)
  call gpu_set_les(les_type, filter_type)
  HistXOld = 0; HistYOld = 0; HistZOld = 0
  call gpu_upload_state(ux0, uy0, uz0, pr0, HistXOld, HistYOld, HistZOld)
  do n = 1, nsteps
    call gpu_ChannelIterate(cflmp, divmax1, divmax2, vmaxabs, SimTime)
  end do
  write(*, '(A,I0,A,ES12.4)') "LES device-resident ", nsteps, " steps: divmax2 ", divmax2
  call gpu_download_state(ux, uy, uz, pressure, HistXOld, HistYOld, HistZOld)
  write(u) ux, uy, uz, pressure
  close(u)
  call gpu_finalize()
end program c3d_fortran_check
