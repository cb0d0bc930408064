!==================================================================================================
! lbm_rb -- iso_c_binding interface to the gfx950 solver (include/lbm_rb.h).
!
! The reference's hot path is the body of its time loop (2DRB/3DRB.F90:225-264): eight
! argument-less subroutines over module globals.  A Fortran driver replaces those calls with
! the C entry points below; host arrays keep the reference's Fortran layout, so f(0:18,nx,ny,nz)
! and T(nx,ny,nz) are passed as they are.
!==================================================================================================
module lbm_rb
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LBM_OK = 0
  integer(c_int), parameter :: LBM_VEL_NOSLIP = 0, LBM_VEL_PERIODIC = 1
  integer(c_int), parameter :: LBM_T_CONST = 0, LBM_T_ADIABATIC = 1, LBM_T_PERIODIC = 2
  integer(c_int), parameter :: LBM_FP64 = 0, LBM_FP32_SHIFT = 1, LBM_FP16_SHIFT = 2
  integer(c_int), parameter :: LBM_LAYOUT_TWO_ARRAY = 0, LBM_LAYOUT_AA = 1, LBM_LAYOUT_AUTO = 2

  type, bind(C) :: lbm_config
    integer(c_int32_t) :: dims, nx, ny, nz
    real(c_double) :: rayleigh, prandtl, mach, thot, tcold
    integer(c_int32_t) :: side_heated, vel_x, vel_z, temp_x, temp_y, temp_z, use_g
    integer(c_int32_t) :: precision, device
    integer(c_int32_t) :: legacy_thermal = 0
    integer(c_int32_t) :: layout = 0          ! LBM_LAYOUT_*
  end type lbm_config

  type, bind(C) :: lbm_params
    real(c_double) :: L, tref, tauf, viscosity, diffusivity, gbeta, time_unit, velocity_unit
    real(c_double) :: s_e, s_eps, s_nu, s_pi, s_q, s_m
    real(c_double) :: taug, q_k, q_nu, sg
    real(c_double) :: para_a
  end type lbm_params

  type, bind(C) :: lbm_diag
    real(c_double) :: nu_vol, re_vol, nu_global, nu_hot, nu_cold, nu_middle
    real(c_double) :: nu_hot_max, nu_hot_max_pos, nu_hot_min, nu_hot_min_pos
    real(c_double) :: mass, heat, kinetic
  end type lbm_diag

  type, bind(C) :: lbm_psi
    real(c_double) :: psi_center_coarse, psi_center_fine, psi_abs_max, x_at_max, y_at_max
  end type lbm_psi

  type, bind(C) :: lbm_midplane
    real(c_double) :: u_max, u_pos1, u_pos2, u_plane
    real(c_double) :: v_max, v_pos1, v_pos2, v_plane
    real(c_double) :: w_max, w_pos1, w_pos2, w_plane
  end type lbm_midplane

  interface
    integer(c_int) function lbm_create(cfg, h) bind(C, name="lbm_create")
      import :: c_int, c_ptr, lbm_config
      type(lbm_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function lbm_destroy(h) bind(C, name="lbm_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function lbm_init_exact(h, T0) bind(C, name="lbm_init_exact")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, T0
    end function
    integer(c_int) function lbm_init_seeded(h, seed, amp) bind(C, name="lbm_init_seeded")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), value :: seed
      real(c_double), value :: amp
    end function
    integer(c_int) function lbm_step(h, nsteps) bind(C, name="lbm_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function lbm_sync(h) bind(C, name="lbm_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function lbm_get_params(h, p) bind(C, name="lbm_get_params")
      import :: c_int, c_ptr, lbm_params
      type(c_ptr), value :: h
      type(lbm_params), intent(out) :: p
    end function
    integer(c_int) function lbm_macros(h, rho, u, v, w, T) bind(C, name="lbm_macros")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, rho, u, v, w, T
    end function
    integer(c_int) function lbm_check(h, error_u, error_t) bind(C, name="lbm_check")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: error_u, error_t
    end function
    integer(c_int) function lbm_diagnostics(h, d) bind(C, name="lbm_diagnostics")
      import :: c_int, c_ptr, lbm_diag
      type(c_ptr), value :: h
      type(lbm_diag), intent(out) :: d
    end function
    integer(c_int) function lbm_run_diagnostics(h, nsteps, every, d) bind(C, name="lbm_run_diagnostics")
      import :: c_int, c_int32_t, c_ptr, lbm_diag
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps, every
      type(lbm_diag), intent(out) :: d(*)
    end function
    integer(c_int) function lbm_midplane_extrema(h, m) bind(C, name="lbm_midplane_extrema")
      import :: c_int, c_ptr, lbm_midplane
      type(c_ptr), value :: h
      type(lbm_midplane), intent(out) :: m
    end function
    integer(c_int) function lbm_midplane_slice(h, axis, u, v, w, t, psi, vort) &
         bind(C, name="lbm_midplane_slice")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: axis
      type(c_ptr), value :: u, v, w, t, psi, vort
    end function
    integer(c_int) function lbm_psi_vorticity_2d(h, psi, vort, st) bind(C, name="lbm_psi_vorticity_2d")
      import :: c_int, c_ptr, lbm_psi
      type(c_ptr), value :: h, psi, vort
      type(lbm_psi), intent(out) :: st
    end function
    integer(c_int) function lbm_gather(h, f, g, bprev) bind(C, name="lbm_gather")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, f, g, bprev
    end function
    integer(c_int) function lbm_scatter(h, f, g) bind(C, name="lbm_scatter")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, f, g
    end function
    type(c_ptr) function lbm_last_error_c(h) bind(C, name="lbm_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function
  end interface

contains

  ! Reference convention (2DRB/3DRB.F90:379-388): on failure write the message to the log
  ! unit and stop.
  subroutine lbm_require(rc, what, h, logunit)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: logunit
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    integer :: n
    if (rc == LBM_OK) return
    text = ''
    if (c_associated(h)) then
      call c_f_pointer(lbm_last_error_c(h), msg, [256])
      n = 0
      do while (n < 256)
        if (msg(n + 1) == c_null_char) exit
        n = n + 1
        text(n:n) = msg(n)
      enddo
    endif
    write(logunit, *) 'Error in ', what, ': code ', rc, ' ', trim(text)
    write(*, *) 'Error in ', what, ': code ', rc, ' ', trim(text)
    stop 1
  end subroutine lbm_require

end module lbm_rb
