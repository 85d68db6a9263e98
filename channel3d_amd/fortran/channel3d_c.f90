!-----------------------------------------------------------------------
! channel3d_c -- ISO_C_BINDING interface to libchannel3d (include/channel3d.h)
!
! One interface per C entry point, same names.  Field arguments are device
! handles (type(c_ptr)) obtained from c3d_field_alloc / c3d_state_field;
! host arrays cross with c3d_upload / c3d_download in the Channel3d layouts
!   C3D_HALO      real(RK)(0:nxc+2, 0:nyc+2, 0:nzc+2)   (m_Variables.f90:58-61)
!   C3D_INTERIOR  real(RK)(1:nxc, 1:nyc, 1:nzc)          (m_Variables.f90:91-95)
!   C3D_WALL      real(RK)(1:nxc, 1:nzc)                 (m_Variables.f90:96-97)
! c3d_check turns a non-zero status into a stop with the library's message,
! the counterpart of the reference's CheckForError(ErrT_Abort, ...)
! (src/CFD_2nd/m_LogInfo.f90:162-168).
!-----------------------------------------------------------------------
module channel3d_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: C3D_OK = 0
  integer(c_int), parameter :: C3D_BC_PERIOD = 0, C3D_BC_NOSLIP = -1, C3D_BC_FREESLIP = -2
  integer(c_int), parameter :: C3D_FP64 = 0, C3D_FP32 = 1
  integer(c_int), parameter :: C3D_HALO = 0, C3D_INTERIOR = 1, C3D_WALL = 2
  ! state fields (m_Variables.f90:10-24)
  integer(c_int), parameter :: C3D_UX = 0, C3D_UY = 1, C3D_UZ = 2, C3D_PR = 3
  integer(c_int), parameter :: C3D_HISTX = 4, C3D_HISTY = 5, C3D_HISTZ = 6
  integer(c_int), parameter :: C3D_ARR1 = 7, C3D_ARR2 = 8, C3D_HALOARR = 9
  integer(c_int), parameter :: C3D_UY_YM = 10, C3D_DUY_YM = 11
  integer(c_int), parameter :: C3D_NUT = 12

  ! the module globals the reference reads (m_Parameters.f90:21-63,
  ! m_TScheme.f90:16-17); layout identical to struct c3d_params
  type, bind(C) :: c3d_params
    integer(c_int) :: nxc, nyc, nzc
    real(c_double) :: xlx, yly, zlz
    real(c_double) :: xnu
    real(c_double) :: dt
    integer(c_int) :: bc(6)
    real(c_double) :: uxbc(6), uybc(6), uzbc(6)
    real(c_double) :: gravity(3)
    integer(c_int) :: is_ux_const
    real(c_double) :: ubulk, ybulk1, ybulk2
    integer(c_int) :: ischeme
    integer(c_int) :: precision
    integer(c_int) :: order          ! 2 (src/CFD_2nd) or 4 (src/CFD_4th)
    real(c_double) :: u_crf          ! uCRF of the 4th-order code
  end type c3d_params

  interface
    integer(c_int) function c3d_create(p, yp, ctx) bind(C, name="c3d_create")
      import :: c_int, c_double, c_ptr, c3d_params
      type(c3d_params), intent(in) :: p
      real(c_double), intent(in) :: yp(*)
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function c3d_destroy(ctx) bind(C, name="c3d_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function c3d_last_error() bind(C, name="c3d_last_error")
      import :: c_ptr
    end function
    integer(c_int) function c3d_sync(ctx) bind(C, name="c3d_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c3d_mesh_coords(nyc, yly, nsection, section_length, nyc_section, &
        stret_type, stret_option, cstret, yp_out) bind(C, name="c3d_mesh_coords")
      import :: c_int, c_double
      integer(c_int), value :: nyc, nsection
      real(c_double), value :: yly
      real(c_double), intent(in) :: section_length(*), cstret(*)
      integer(c_int), intent(in) :: nyc_section(*), stret_type(*), stret_option(*)
      real(c_double), intent(out) :: yp_out(*)
    end function
    integer(c_int) function c3d_get_metric(ctx, id, lo, hi, out) bind(C, name="c3d_get_metric")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: id, lo, hi
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function c3d_set_dt(ctx, dt) bind(C, name="c3d_set_dt")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function c3d_field_alloc(ctx, layout, dptr) bind(C, name="c3d_field_alloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: layout
      type(c_ptr), intent(out) :: dptr
    end function
    integer(c_int) function c3d_field_free(ctx, dptr) bind(C, name="c3d_field_free")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dptr
    end function
    integer(c_int) function c3d_upload(ctx, layout, dptr, host) bind(C, name="c3d_upload")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, dptr
      integer(c_int), value :: layout
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function c3d_download(ctx, layout, dptr, host) bind(C, name="c3d_download")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, dptr
      integer(c_int), value :: layout
      real(c_double), intent(inout) :: host(*)
    end function
    integer(c_int) function c3d_field_copy(ctx, layout, dst, src) bind(C, name="c3d_field_copy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dst, src
      integer(c_int), value :: layout
    end function
    integer(c_int) function c3d_state_field(ctx, id, dptr) bind(C, name="c3d_state_field")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      type(c_ptr), intent(out) :: dptr
    end function

    ! ---- reference routines (m_Parameters.f90, m_FlowType_Channel.f90,
    !      m_TSchemeFIMP_inc.f90, m_TScheme.f90, m_Poisson.f90, m_BC_and_Halo.f90)
    integer(c_int) function c3d_PMcoeUpdate(ctx, ns) bind(C, name="c3d_PMcoeUpdate")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ns
    end function
    integer(c_int) function c3d_Update_uy_ym(ctx, uy_ym, duy_ym) bind(C, name="c3d_Update_uy_ym")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, uy_ym, duy_ym
    end function
    integer(c_int) function c3d_clcRhsX(ctx, ux, uy, uz, rhs, hist, pr) bind(C, name="c3d_clcRhsX")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, rhs, hist, pr
    end function
    integer(c_int) function c3d_clcRhsY(ctx, ux, uy, uz, rhs, hist, pr) bind(C, name="c3d_clcRhsY")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, rhs, hist, pr
    end function
    integer(c_int) function c3d_clcRhsZ(ctx, ux, uy, uz, rhs, hist, pr) bind(C, name="c3d_clcRhsZ")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, rhs, hist, pr
    end function
    ! 4th-order contexts: clcRhsXYZ (src/CFD_4th/m_TScheme.f90:189-438)
    integer(c_int) function c3d_clcRhsXYZ(ctx, ux, uy, uz, rx, ry, rz, hx, hy, hz, pr) &
        bind(C, name="c3d_clcRhsXYZ")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, rx, ry, rz, hx, hy, hz, pr
    end function
    integer(c_int) function c3d_clcU1Hat(ctx, ux, rhs) bind(C, name="c3d_clcU1Hat")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, rhs
    end function
    integer(c_int) function c3d_clcU2Hat(ctx, uy, rhs, duy_ym) bind(C, name="c3d_clcU2Hat")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, uy, rhs, duy_ym
    end function
    integer(c_int) function c3d_clcU3Hat(ctx, uz, rhs) bind(C, name="c3d_clcU3Hat")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, uz, rhs
    end function
    integer(c_int) function c3d_SetBC_and_UpdateHaloForPrSrc(ctx, ux, uy, uz, uy_ym) &
        bind(C, name="c3d_SetBC_and_UpdateHaloForPrSrc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, uy_ym
    end function
    integer(c_int) function c3d_SetBC_and_UpdateHalo_pr(ctx, pr) bind(C, name="c3d_SetBC_and_UpdateHalo_pr")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, pr
    end function
    integer(c_int) function c3d_SetBC_and_UpdateHalo(ctx, ux, uy, uz, uy_ym) &
        bind(C, name="c3d_SetBC_and_UpdateHalo")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, uy_ym
    end function
    integer(c_int) function c3d_clcPrSrc(ctx, ux, uy, uz, prsrc, pr, divmax) bind(C, name="c3d_clcPrSrc")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, prsrc, pr
      real(c_double), intent(out) :: divmax
    end function
    integer(c_int) function c3d_clcPPE(ctx, prsrc, phi) bind(C, name="c3d_clcPPE")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, prsrc, phi
    end function
    integer(c_int) function c3d_FluidVelUpdate(ctx, phi, ux, uy, uz) bind(C, name="c3d_FluidVelUpdate")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, phi, ux, uy, uz
    end function
    integer(c_int) function c3d_PressureUpdate(ctx, pr, phi) bind(C, name="c3d_PressureUpdate")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, pr, phi
    end function
    integer(c_int) function c3d_CalcMaxCFL(ctx, ux, uy, uz, uddxmax) bind(C, name="c3d_CalcMaxCFL")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz
      real(c_double), intent(out) :: uddxmax
    end function
    integer(c_int) function c3d_CheckDivergence(ctx, ux, uy, uz, divmax) bind(C, name="c3d_CheckDivergence")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz
      real(c_double), intent(out) :: divmax
    end function
    integer(c_int) function c3d_CalcVmax(ctx, ux, uy, uz, vmax3) bind(C, name="c3d_CalcVmax")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz
      real(c_double), intent(out) :: vmax3(3)
    end function
    integer(c_int) function c3d_CalcUxAver(ctx, ux, uxaver) bind(C, name="c3d_CalcUxAver")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, ux
      real(c_double), intent(out) :: uxaver
    end function

    ! ---- fused, device-resident (m_ChannelSystem.f90:73-130)
    integer(c_int) function c3d_substep(ctx, ns) bind(C, name="c3d_substep")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ns
    end function
    integer(c_int) function c3d_step(ctx) bind(C, name="c3d_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c3d_step_diagnostics(ctx, out9) bind(C, name="c3d_step_diagnostics")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out9(9)
    end function
    integer(c_int) function c3d_project(ctx) bind(C, name="c3d_project")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c3d_ppe(ctx, prsrc, phi) bind(C, name="c3d_ppe")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, prsrc, phi
    end function
    integer(c_size_t) function c3d_device_bytes(ctx) bind(C, name="c3d_device_bytes")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c3d_set_ppe_refine(ctx, sweeps) bind(C, name="c3d_set_ppe_refine")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: sweeps
    end function
    integer(c_int) function c3d_set_les_closure(ctx, cl) bind(C, name="c3d_set_les_closure")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, cl
    end function
    integer(c_int) function c3d_set_graph(ctx, on) bind(C, name="c3d_set_graph")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function

    ! ---- 4th-order spectra (SpectraOptions, CFD_4th m_FlowCase.f90:644-938)
    integer(c_int) function c3d_clcSpectra1D(ctx, ux, uy, uz, pr, j1, j2) &
        bind(C, name="c3d_clcSpectra1D")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, pr
      integer(c_int), value :: j1, j2
    end function
    integer(c_int) function c3d_write_spectra(ctx, path_x, path_z) bind(C, name="c3d_write_spectra")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path_x(*), path_z(*)
    end function
    integer(c_int) function c3d_clcSpectra2D(ctx, ux, uy, uz, pr, j1) &
        bind(C, name="c3d_clcSpectra2D")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, pr
      integer(c_int), value :: j1
    end function
    integer(c_int) function c3d_write_spectra2d(ctx, path) bind(C, name="c3d_write_spectra2d")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
    end function

    ! ---- LES layer (LesOptions, m_LESModels.f90, m_TScheme.f90:179-206)
    integer(c_int) function c3d_set_les(ctx, les_type, filter_type) bind(C, name="c3d_set_les")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: les_type, filter_type
    end function
    integer(c_int) function c3d_Clc_SGS_Vis(ctx, ux, uy, uz, nut) bind(C, name="c3d_Clc_SGS_Vis")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, nut
    end function

    ! ---- restart files (m_IOAndVisu.f90:352-558); path: NUL-terminated
    integer(c_int) function c3d_write_restart(ctx, path, ux, uy, uz, pr, hx, hy, hz) &
        bind(C, name="c3d_write_restart")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx, ux, uy, uz, pr, hx, hy, hz
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function c3d_read_restart(ctx, path, ux, uy, uz, pr, hx, hy, hz) &
        bind(C, name="c3d_read_restart")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx, ux, uy, uz, pr, hx, hy, hz
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function c3d_restart_name(restart_dir, run_name, ntime, out, out_len) &
        bind(C, name="c3d_restart_name")
      import :: c_int, c_char, c_size_t
      character(kind=c_char), intent(in) :: restart_dir(*), run_name(*)
      integer(c_int), value :: ntime
      character(kind=c_char), intent(out) :: out(*)
      integer(c_size_t), value :: out_len
    end function

    ! ---- channel statistics (m_FlowType_Channel.f90:135-363)
    integer(c_int) function c3d_clcStat_CH(ctx, ux, uy, uz, pr) bind(C, name="c3d_clcStat_CH")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, pr
    end function
    integer(c_int) function c3d_stat_get(ctx, sumstat, nfstime, prgradsum) &
        bind(C, name="c3d_stat_get")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sumstat(*)
      integer(c_int), intent(out) :: nfstime
      real(c_double), intent(out) :: prgradsum
    end function
    integer(c_int) function c3d_stat_reset(ctx) bind(C, name="c3d_stat_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c3d_write_stat(ctx, path, itime, ivstats) bind(C, name="c3d_write_stat")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: itime, ivstats
    end function
    ! ---- 4th-order statistics (CFD_4th m_FlowCase.f90:306-633)
    integer(c_int) function c3d_clcStat(ctx, ux, uy, uz, pr) bind(C, name="c3d_clcStat")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, pr
    end function
    integer(c_int) function c3d_stat_set_grad(ctx, on) bind(C, name="c3d_stat_set_grad")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function c3d_stat_get_grad(ctx, sumgrad) bind(C, name="c3d_stat_get_grad")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sumgrad(*)
    end function
    integer(c_int) function c3d_write_vgrad(ctx, path) bind(C, name="c3d_write_vgrad")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function c3d_append_prgrad(ctx, path, itime) bind(C, name="c3d_append_prgrad")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: itime
    end function
  end interface

contains

  ! the library's last error message as a Fortran string
  function c3d_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    p = c3d_last_error()
    if (.not. c_associated(p)) then
      msg = ""
      return
    end if
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function c3d_error_message

  ! abort on a non-zero status, naming the call site
  subroutine c3d_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= C3D_OK) then
      write(*, '(A)') "channel3d: "//trim(where)//" failed: "//c3d_error_message()
      error stop 1
    end if
  end subroutine c3d_check

end module channel3d_c
