!-----------------------------------------------------------------------
! m_channel3d_gpu -- the MI355X time-step core behind Channel3d's Fortran
! call sites.
!
! Two ways in:
!  * per routine, host arrays: gpu_clcRhsX(ux,uy,uz,RhsX,HistXOld,pressure)
!    and the rest carry the signatures of the procedure pointers that
!    InitTimeScheme binds (src/CFD_2nd/m_TScheme.f90:28-30, 147-173) and of
!    clcPPE (src/CFD_2nd/m_Poisson.f90:16), so a driver can point them here
!    (clcRhsX => gpu_clcRhsX ...).  Every call uploads its array arguments
!    to device mirrors and downloads what the routine writes: correct at
!    every call site, PCIe-bound, meant for bring-up and validation;
!  * device resident: gpu_upload_state once, gpu_ChannelIterate per time
!    step (the loop of src/CFD_2nd/m_ChannelSystem.f90:73-130 on the GPU,
!    no host traffic), gpu_download_state when the host needs the fields.
! Array arguments are the reference's own explicit-shape FP64 arrays
! (halo, interior or wall layout, m_Variables.f90:52-118), passed by address.
!-----------------------------------------------------------------------
module m_channel3d_gpu
  use iso_c_binding
  use channel3d_c
  implicit none
  private

  type(c_ptr), save :: ctx = c_null_ptr
  ! device mirrors of argument slots: halo arrays (u, v, w, p, two scratch),
  ! interior arrays (three), wall planes (two)
  type(c_ptr), save :: dh(6), di(3), dw(2)

  public :: gpu_init, gpu_finalize, gpu_context
  public :: gpu_PMcoeUpdate, gpu_Update_uy_ym
  public :: gpu_clcRhsX, gpu_clcRhsY, gpu_clcRhsZ
  public :: gpu_clcU1Hat, gpu_clcU2Hat, gpu_clcU3Hat
  public :: gpu_SetBC_and_UpdateHaloForPrSrc, gpu_SetBC_and_UpdateHalo_pr, gpu_SetBC_and_UpdateHalo
  public :: gpu_clcPrSrc, gpu_clcPPE, gpu_FluidVelUpdate, gpu_PressureUpdate
  public :: gpu_CalcMaxCFL, gpu_CheckDivergence, gpu_CalcVmax, gpu_CalcUxAver
  public :: gpu_upload_state, gpu_download_state, gpu_ChannelIterate
  public :: gpu_write_restart, gpu_read_restart, gpu_clcStat, gpu_write_stat, gpu_append_prgrad
  public :: gpu_set_les, gpu_Clc_SGS_Vis

contains

  ! InitMeshAndMetries + InitPoissonSolver + InitTimeScheme in one call;
  ! yp is the reference's yp(0:nyp)
  subroutine gpu_init(p, yp)
    type(c3d_params), intent(in) :: p
    real(c_double), intent(in) :: yp(0:)
    integer :: k
    call c3d_check(c3d_create(p, yp, ctx), "c3d_create")
    do k = 1, 6
      call c3d_check(c3d_field_alloc(ctx, C3D_HALO, dh(k)), "c3d_field_alloc")
    end do
    do k = 1, 3
      call c3d_check(c3d_field_alloc(ctx, C3D_INTERIOR, di(k)), "c3d_field_alloc")
    end do
    do k = 1, 2
      call c3d_check(c3d_field_alloc(ctx, C3D_WALL, dw(k)), "c3d_field_alloc")
    end do
  end subroutine gpu_init

  subroutine gpu_finalize()
    if (c_associated(ctx)) call c3d_check(c3d_destroy(ctx), "c3d_destroy")
    ctx = c_null_ptr
  end subroutine gpu_finalize

  type(c_ptr) function gpu_context()
    gpu_context = ctx
  end function gpu_context

  ! ------------------------------------------------------------ helpers --
  subroutine up(layout, d, a)
    integer(c_int), intent(in) :: layout
    type(c_ptr), intent(in) :: d
    real(c_double), intent(in) :: a(*)
    call c3d_check(c3d_upload(ctx, layout, d, a), "c3d_upload")
  end subroutine up

  subroutine down(layout, d, a)
    integer(c_int), intent(in) :: layout
    type(c_ptr), intent(in) :: d
    real(c_double), intent(inout) :: a(*)
    call c3d_check(c3d_download(ctx, layout, d, a), "c3d_download")
  end subroutine down

  ! ------------------------------------- per routine, on host arrays --
  ! m_Parameters.f90:173-186
  subroutine gpu_PMcoeUpdate(ns)
    integer, intent(in) :: ns
    call c3d_check(c3d_PMcoeUpdate(ctx, int(ns, c_int)), "PMcoeUpdate")
  end subroutine gpu_PMcoeUpdate

  ! m_FlowType_Channel.f90:120-130
  subroutine gpu_Update_uy_ym(uy_ym, duy_ym)
    real(c_double), intent(inout) :: uy_ym(*), duy_ym(*)
    call up(C3D_WALL, dw(1), uy_ym)
    call up(C3D_WALL, dw(2), duy_ym)
    call c3d_check(c3d_Update_uy_ym(ctx, dw(1), dw(2)), "Update_uy_ym")
    call down(C3D_WALL, dw(1), uy_ym)
    call down(C3D_WALL, dw(2), duy_ym)
  end subroutine gpu_Update_uy_ym

  ! m_TSchemeFIMP_inc.f90:7-75 (and :80-129, :134-182 below)
  subroutine gpu_clcRhsX(ux, uy, uz, RhsX, HistXOld, pressure)
    real(c_double), intent(in) :: ux(*), uy(*), uz(*), pressure(*)
    real(c_double), intent(inout) :: RhsX(*), HistXOld(*)
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call up(C3D_HALO, dh(4), pressure); call up(C3D_INTERIOR, di(2), HistXOld)
    call c3d_check(c3d_clcRhsX(ctx, dh(1), dh(2), dh(3), di(1), di(2), dh(4)), "clcRhsX")
    call down(C3D_INTERIOR, di(1), RhsX); call down(C3D_INTERIOR, di(2), HistXOld)
  end subroutine gpu_clcRhsX

  subroutine gpu_clcRhsY(ux, uy, uz, RhsY, HistYOld, pressure)
    real(c_double), intent(in) :: ux(*), uy(*), uz(*), pressure(*)
    real(c_double), intent(inout) :: RhsY(*), HistYOld(*)
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call up(C3D_HALO, dh(4), pressure); call up(C3D_INTERIOR, di(2), HistYOld)
    call c3d_check(c3d_clcRhsY(ctx, dh(1), dh(2), dh(3), di(1), di(2), dh(4)), "clcRhsY")
    call down(C3D_INTERIOR, di(1), RhsY); call down(C3D_INTERIOR, di(2), HistYOld)
  end subroutine gpu_clcRhsY

  ! RhsZ is a halo array in the reference (RealHalo)
  subroutine gpu_clcRhsZ(ux, uy, uz, RhsZ, HistZOld, pressure)
    real(c_double), intent(in) :: ux(*), uy(*), uz(*), pressure(*)
    real(c_double), intent(inout) :: RhsZ(*), HistZOld(*)
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call up(C3D_HALO, dh(4), pressure); call up(C3D_INTERIOR, di(2), HistZOld)
    call up(C3D_HALO, dh(5), RhsZ)
    call c3d_check(c3d_clcRhsZ(ctx, dh(1), dh(2), dh(3), dh(5), di(2), dh(4)), "clcRhsZ")
    call down(C3D_HALO, dh(5), RhsZ); call down(C3D_INTERIOR, di(2), HistZOld)
  end subroutine gpu_clcRhsZ

  ! m_TSchemeFIMP_inc.f90:263-347 / 636-729 / 1035-1128 (_010 variants)
  subroutine gpu_clcU1Hat(ux, RhsX)
    real(c_double), intent(inout) :: ux(*), RhsX(*)
    call up(C3D_HALO, dh(1), ux); call up(C3D_INTERIOR, di(1), RhsX)
    call c3d_check(c3d_clcU1Hat(ctx, dh(1), di(1)), "clcU1Hat")
    call down(C3D_HALO, dh(1), ux)
  end subroutine gpu_clcU1Hat

  subroutine gpu_clcU2Hat(uy, RhsY, duy_ym)
    real(c_double), intent(inout) :: uy(*), RhsY(*)
    real(c_double), intent(in) :: duy_ym(*)
    call up(C3D_HALO, dh(2), uy); call up(C3D_INTERIOR, di(1), RhsY); call up(C3D_WALL, dw(2), duy_ym)
    call c3d_check(c3d_clcU2Hat(ctx, dh(2), di(1), dw(2)), "clcU2Hat")
    call down(C3D_HALO, dh(2), uy)
  end subroutine gpu_clcU2Hat

  subroutine gpu_clcU3Hat(uz, RhsZ)
    real(c_double), intent(inout) :: uz(*), RhsZ(*)
    call up(C3D_HALO, dh(3), uz); call up(C3D_HALO, dh(5), RhsZ)
    call c3d_check(c3d_clcU3Hat(ctx, dh(3), dh(5)), "clcU3Hat")
    call down(C3D_HALO, dh(3), uz)
  end subroutine gpu_clcU3Hat

  ! m_BC_and_Halo.f90:349-405 / 410-456 / 232-344
  subroutine gpu_SetBC_and_UpdateHaloForPrSrc(ux, uy, uz, uy_ym)
    real(c_double), intent(inout) :: ux(*), uy(*), uz(*)
    real(c_double), intent(in) :: uy_ym(*)
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call up(C3D_WALL, dw(1), uy_ym)
    call c3d_check(c3d_SetBC_and_UpdateHaloForPrSrc(ctx, dh(1), dh(2), dh(3), dw(1)), &
                   "SetBC_and_UpdateHaloForPrSrc")
    call down(C3D_HALO, dh(1), ux); call down(C3D_HALO, dh(2), uy); call down(C3D_HALO, dh(3), uz)
  end subroutine gpu_SetBC_and_UpdateHaloForPrSrc

  subroutine gpu_SetBC_and_UpdateHalo_pr(pressure)
    real(c_double), intent(inout) :: pressure(*)
    call up(C3D_HALO, dh(4), pressure)
    call c3d_check(c3d_SetBC_and_UpdateHalo_pr(ctx, dh(4)), "SetBC_and_UpdateHalo_pr")
    call down(C3D_HALO, dh(4), pressure)
  end subroutine gpu_SetBC_and_UpdateHalo_pr

  subroutine gpu_SetBC_and_UpdateHalo(ux, uy, uz, uy_ym)
    real(c_double), intent(inout) :: ux(*), uy(*), uz(*)
    real(c_double), intent(in) :: uy_ym(*)
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call up(C3D_WALL, dw(1), uy_ym)
    call c3d_check(c3d_SetBC_and_UpdateHalo(ctx, dh(1), dh(2), dh(3), dw(1)), "SetBC_and_UpdateHalo")
    call down(C3D_HALO, dh(1), ux); call down(C3D_HALO, dh(2), uy); call down(C3D_HALO, dh(3), uz)
  end subroutine gpu_SetBC_and_UpdateHalo

  ! m_TScheme.f90:246-276
  subroutine gpu_clcPrSrc(ux, uy, uz, prsrc, pressure, divmax)
    real(c_double), intent(in) :: ux(*), uy(*), uz(*)
    real(c_double), intent(inout) :: prsrc(*), pressure(*)
    real(c_double), intent(out) :: divmax
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call up(C3D_HALO, dh(4), pressure)
    call c3d_check(c3d_clcPrSrc(ctx, dh(1), dh(2), dh(3), di(3), dh(4), divmax), "clcPrSrc")
    call down(C3D_INTERIOR, di(3), prsrc); call down(C3D_HALO, dh(4), pressure)
  end subroutine gpu_clcPrSrc

  ! m_Poisson.f90:567-688
  subroutine gpu_clcPPE(prsrc, prphiHalo)
    real(c_double), intent(inout) :: prsrc(*), prphiHalo(*)
    call up(C3D_INTERIOR, di(3), prsrc); call up(C3D_HALO, dh(6), prphiHalo)
    call c3d_check(c3d_clcPPE(ctx, di(3), dh(6)), "clcPPE")
    call down(C3D_HALO, dh(6), prphiHalo)
  end subroutine gpu_clcPPE

  ! m_TScheme.f90:281-306
  subroutine gpu_FluidVelUpdate(prphiHalo, ux, uy, uz)
    real(c_double), intent(in) :: prphiHalo(*)
    real(c_double), intent(inout) :: ux(*), uy(*), uz(*)
    call up(C3D_HALO, dh(6), prphiHalo)
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call c3d_check(c3d_FluidVelUpdate(ctx, dh(6), dh(1), dh(2), dh(3)), "FluidVelUpdate")
    call down(C3D_HALO, dh(1), ux); call down(C3D_HALO, dh(2), uy); call down(C3D_HALO, dh(3), uz)
  end subroutine gpu_FluidVelUpdate

  ! m_TSchemeFIMP_inc.f90:1371-1386
  subroutine gpu_PressureUpdate(pressure, prphiHalo)
    real(c_double), intent(inout) :: pressure(*)
    real(c_double), intent(in) :: prphiHalo(*)
    call up(C3D_HALO, dh(4), pressure); call up(C3D_HALO, dh(6), prphiHalo)
    call c3d_check(c3d_PressureUpdate(ctx, dh(4), dh(6)), "PressureUpdate")
    call down(C3D_HALO, dh(4), pressure)
  end subroutine gpu_PressureUpdate

  ! ---------------------------------------------------------- LES layer --
  ! InitFilterCoe (m_LESModels.f90:19-42) and the LES procedure pointers of
  ! InitTimeScheme (m_TScheme.f90:179-206): after this call gpu_clcRhsX/Y/Z,
  ! gpu_clcU1/2/3Hat, gpu_clcPrSrc and gpu_PressureUpdate are the *_PIMP_LES
  ! routines and clcPrSrcOther, reading the eddy viscosity the last
  ! gpu_Clc_SGS_Vis left on the device (the reference's module array nut)
  subroutine gpu_set_les(LES_type, FilterType)
    integer, intent(in) :: LES_type, FilterType
    call c3d_check(c3d_set_les(ctx, int(LES_type, c_int), int(FilterType, c_int)), "c3d_set_les")
  end subroutine gpu_set_les

  ! Clc_SGS_Vis (m_LESModels.f90:47-62); nut is the halo array
  subroutine gpu_Clc_SGS_Vis(ux, uy, uz, nut)
    real(c_double), intent(in) :: ux(*), uy(*), uz(*)
    real(c_double), intent(inout) :: nut(*)
    type(c_ptr) :: d
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call c3d_check(c3d_state_field(ctx, C3D_NUT, d), "state")
    call c3d_check(c3d_Clc_SGS_Vis(ctx, dh(1), dh(2), dh(3), d), "Clc_SGS_Vis")
    call down(C3D_HALO, d, nut)
  end subroutine gpu_Clc_SGS_Vis

  ! m_Tools.f90:21-104, 250-270
  subroutine gpu_CalcMaxCFL(ux, uy, uz, uddxmax)
    real(c_double), intent(in) :: ux(*), uy(*), uz(*)
    real(c_double), intent(out) :: uddxmax
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call c3d_check(c3d_CalcMaxCFL(ctx, dh(1), dh(2), dh(3), uddxmax), "CalcMaxCFL")
  end subroutine gpu_CalcMaxCFL

  subroutine gpu_CheckDivergence(ux, uy, uz, divmax)
    real(c_double), intent(in) :: ux(*), uy(*), uz(*)
    real(c_double), intent(out) :: divmax
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call c3d_check(c3d_CheckDivergence(ctx, dh(1), dh(2), dh(3), divmax), "CheckDivergence")
  end subroutine gpu_CheckDivergence

  function gpu_CalcVmax(ux, uy, uz) result(vmaxabs)
    real(c_double), intent(in) :: ux(*), uy(*), uz(*)
    real(c_double) :: vmaxabs(3)
    call up(C3D_HALO, dh(1), ux); call up(C3D_HALO, dh(2), uy); call up(C3D_HALO, dh(3), uz)
    call c3d_check(c3d_CalcVmax(ctx, dh(1), dh(2), dh(3), vmaxabs), "CalcVmax")
  end function gpu_CalcVmax

  function gpu_CalcUxAver(ux) result(res)
    real(c_double), intent(in) :: ux(*)
    real(c_double) :: res
    call up(C3D_HALO, dh(1), ux)
    call c3d_check(c3d_CalcUxAver(ctx, dh(1), res), "CalcUxAver")
  end function gpu_CalcUxAver

  ! ------------------------------------------------- device resident --
  subroutine gpu_upload_state(ux, uy, uz, pressure, HistXOld, HistYOld, HistZOld)
    real(c_double), intent(in) :: ux(*), uy(*), uz(*), pressure(*)
    real(c_double), intent(in) :: HistXOld(*), HistYOld(*), HistZOld(*)
    type(c_ptr) :: d
    call c3d_check(c3d_state_field(ctx, C3D_UX, d), "state"); call up(C3D_HALO, d, ux)
    call c3d_check(c3d_state_field(ctx, C3D_UY, d), "state"); call up(C3D_HALO, d, uy)
    call c3d_check(c3d_state_field(ctx, C3D_UZ, d), "state"); call up(C3D_HALO, d, uz)
    call c3d_check(c3d_state_field(ctx, C3D_PR, d), "state"); call up(C3D_HALO, d, pressure)
    call c3d_check(c3d_state_field(ctx, C3D_HISTX, d), "state"); call up(C3D_INTERIOR, d, HistXOld)
    call c3d_check(c3d_state_field(ctx, C3D_HISTY, d), "state"); call up(C3D_INTERIOR, d, HistYOld)
    call c3d_check(c3d_state_field(ctx, C3D_HISTZ, d), "state"); call up(C3D_INTERIOR, d, HistZOld)
  end subroutine gpu_upload_state

  subroutine gpu_download_state(ux, uy, uz, pressure, HistXOld, HistYOld, HistZOld)
    real(c_double), intent(inout) :: ux(*), uy(*), uz(*), pressure(*)
    real(c_double), intent(inout) :: HistXOld(*), HistYOld(*), HistZOld(*)
    type(c_ptr) :: d
    call c3d_check(c3d_state_field(ctx, C3D_UX, d), "state"); call down(C3D_HALO, d, ux)
    call c3d_check(c3d_state_field(ctx, C3D_UY, d), "state"); call down(C3D_HALO, d, uy)
    call c3d_check(c3d_state_field(ctx, C3D_UZ, d), "state"); call down(C3D_HALO, d, uz)
    call c3d_check(c3d_state_field(ctx, C3D_PR, d), "state"); call down(C3D_HALO, d, pressure)
    call c3d_check(c3d_state_field(ctx, C3D_HISTX, d), "state"); call down(C3D_INTERIOR, d, HistXOld)
    call c3d_check(c3d_state_field(ctx, C3D_HISTY, d), "state"); call down(C3D_INTERIOR, d, HistYOld)
    call c3d_check(c3d_state_field(ctx, C3D_HISTZ, d), "state"); call down(C3D_INTERIOR, d, HistZOld)
  end subroutine gpu_download_state

  ! one time step of ChannelIterate (m_ChannelSystem.f90:73-130) on the
  ! device-resident state; returns what the reference logs
  subroutine gpu_ChannelIterate(cflmp, divmax1, divmax2, vmaxabs, SimTime)
    real(c_double), intent(out) :: cflmp, divmax1, divmax2, vmaxabs(3), SimTime
    real(c_double) :: d(9)
    call c3d_check(c3d_step(ctx), "c3d_step")
    call c3d_check(c3d_step_diagnostics(ctx, d), "c3d_step_diagnostics")
    cflmp = d(1); divmax1 = d(2); divmax2 = d(3); vmaxabs = d(4:6); SimTime = d(9)
  end subroutine gpu_ChannelIterate

  ! ------------------------------------------- restart files, statistics --
  ! write_restart / read_restart (src/CFD_2nd/m_IOAndVisu.f90:352-558) on the
  ! device-resident state; with_hist adds HistX/Y/ZOld (the AB2 restart)
  subroutine gpu_write_restart(ntime, RestartDir, RunName, with_hist)
    integer, intent(in) :: ntime
    character(*), intent(in) :: RestartDir, RunName
    logical, intent(in) :: with_hist
    character(len=512) :: chFile
    type(c_ptr) :: f(7)
    write(chFile, "(A,I10.10)") trim(RestartDir)//"RestartFor"//trim(RunName), ntime
    call state_ptrs(f, with_hist)
    call c3d_check(c3d_write_restart(ctx, trim(chFile)//c_null_char, f(1), f(2), f(3), f(4), &
                                     f(5), f(6), f(7)), "c3d_write_restart")
  end subroutine gpu_write_restart

  ! read_restart, then what ChannelInitialize does after it
  ! (m_ChannelSystem.f90:54-62): wall values, ghosts and halos
  subroutine gpu_read_restart(ntime, RestartDir, RunName, with_hist)
    integer, intent(in) :: ntime
    character(*), intent(in) :: RestartDir, RunName
    logical, intent(in) :: with_hist
    character(len=512) :: chFile
    type(c_ptr) :: f(7), ym, dym
    write(chFile, "(A,I10.10)") trim(RestartDir)//"RestartFor"//trim(RunName), ntime
    call state_ptrs(f, with_hist)
    call c3d_check(c3d_read_restart(ctx, trim(chFile)//c_null_char, f(1), f(2), f(3), f(4), &
                                    f(5), f(6), f(7)), "c3d_read_restart")
    call c3d_check(c3d_state_field(ctx, C3D_UY_YM, ym), "state")
    call c3d_check(c3d_state_field(ctx, C3D_DUY_YM, dym), "state")
    call c3d_check(c3d_Update_uy_ym(ctx, ym, dym), "Update_uy_ym")
    call c3d_check(c3d_SetBC_and_UpdateHalo(ctx, f(1), f(2), f(3), ym), "SetBC_and_UpdateHalo")
    call c3d_check(c3d_SetBC_and_UpdateHalo_pr(ctx, f(4)), "SetBC_and_UpdateHalo_pr")
  end subroutine gpu_read_restart

  subroutine state_ptrs(f, with_hist)
    type(c_ptr), intent(out) :: f(7)
    logical, intent(in) :: with_hist
    call c3d_check(c3d_state_field(ctx, C3D_UX, f(1)), "state")
    call c3d_check(c3d_state_field(ctx, C3D_UY, f(2)), "state")
    call c3d_check(c3d_state_field(ctx, C3D_UZ, f(3)), "state")
    call c3d_check(c3d_state_field(ctx, C3D_PR, f(4)), "state")
    f(5:7) = c_null_ptr
    if (with_hist) then
      call c3d_check(c3d_state_field(ctx, C3D_HISTX, f(5)), "state")
      call c3d_check(c3d_state_field(ctx, C3D_HISTY, f(6)), "state")
      call c3d_check(c3d_state_field(ctx, C3D_HISTZ, f(7)), "state")
    end if
  end subroutine state_ptrs

  ! clcStat_CH on the device-resident state (m_FlowType_Channel.f90:175-318)
  subroutine gpu_clcStat()
    type(c_ptr) :: f(7)
    call state_ptrs(f, .false.)
    call c3d_check(c3d_clcStat_CH(ctx, f(1), f(2), f(3), f(4)), "c3d_clcStat_CH")
  end subroutine gpu_clcStat

  ! the stats file of the samples since the last write (:330-361)
  subroutine gpu_write_stat(ResultsDir, itime, ivstats)
    character(*), intent(in) :: ResultsDir
    integer, intent(in) :: itime, ivstats
    character(len=512) :: filename
    write(filename, "(A,I10.10)") trim(ResultsDir)//'stats', itime
    call c3d_check(c3d_write_stat(ctx, trim(filename)//c_null_char, int(itime, c_int), &
                                  int(ivstats, c_int)), "c3d_write_stat")
  end subroutine gpu_write_stat

  ! one line of the PrGrad file (:317-326)
  subroutine gpu_append_prgrad(ResultsDir, ilast, itime)
    character(*), intent(in) :: ResultsDir
    integer, intent(in) :: ilast, itime
    character(len=512) :: filename
    write(filename, '(A,I10.10)') trim(ResultsDir)//"PrGrad", ilast
    call c3d_check(c3d_append_prgrad(ctx, trim(filename)//c_null_char, int(itime, c_int)), &
                   "c3d_append_prgrad")
  end subroutine gpu_append_prgrad

end module m_channel3d_gpu
