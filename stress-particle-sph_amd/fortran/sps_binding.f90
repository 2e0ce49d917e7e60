!-------------------------------------------------------------------------------
! sps_binding - iso_c_binding interface to the C-ABI of include/sps.h
!
! The derived types mirror the C structs field for field (bind(C) gives them
! the companion C processor's layout); the interfaces bind the entry points by
! their C names.  Arrays cross the boundary as plain addresses: the caller
! passes contiguous Fortran arrays in the reference's column-major layout.
!-------------------------------------------------------------------------------
module sps_binding
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: SPS_ABI_VERSION = 4
  integer, parameter :: SPS_MAX_TCURVE = 64
  integer, parameter :: SPS_MAX_BCS = 32, SPS_MAX_CURVES = 16
  integer(c_int), parameter :: SPS_OK = 0
  integer(c_int32_t), parameter :: SPS_MODE_ORDERED = 0, SPS_MODE_FAST = 1

  type, bind(C) :: sps_params_t
     integer(c_int32_t) :: abi_version, nnode, nstress, ndummy, ndummy2, nint_vars
     integer(c_int32_t) :: ndimn, nstre, skf
     integer(c_int32_t) :: sp_sph, inside_approach, cspm, update_x, xsph, cont_density
     integer(c_int32_t) :: art_stress, dummy_nodes, sph_shift, no_bcs, bforce_gate, mode
     integer(c_int32_t) :: ic_grav, tcurve_npts
     real(c_double) :: props(20), pi, dx, alpha, beta, damping, cspm_det_tol
     real(c_double) :: D11, D22, D12, D33, D41, D42
     real(c_double) :: xmin_domain(2), xmax_domain(2), cgrav(2), ft_grav
     real(c_double) :: tcurve_t(SPS_MAX_TCURVE)
     real(c_float) :: tcurve_f(SPS_MAX_TCURVE)
     integer(c_int32_t) :: npoints, vel_vector, shift_update, reserved0
     real(c_double) :: r_x, r_y, disp_tol
  end type sps_params_t

  type, bind(C) :: sps_state_t
     type(c_ptr) :: x = c_null_ptr, vel = c_null_ptr, stress = c_null_ptr
     type(c_ptr) :: mass = c_null_ptr, rho = c_null_ptr, hsml = c_null_ptr
     type(c_ptr) :: itype = c_null_ptr, if_out_domain = c_null_ptr
     type(c_ptr) :: internal_vars = c_null_ptr, f_drucker = c_null_ptr
     type(c_ptr) :: x00 = c_null_ptr, displ = c_null_ptr
     type(c_ptr) :: wall_position = c_null_ptr, horizontal_or_not = c_null_ptr
     type(c_ptr) :: grad_u = c_null_ptr, ddev_strn = c_null_ptr
     type(c_ptr) :: f_bound = c_null_ptr, art_visc = c_null_ptr
     type(c_ptr) :: x_10 = c_null_ptr, disp_10 = c_null_ptr
  end type sps_state_t

  type, bind(C) :: sps_bcs_t
     integer(c_int32_t) :: no_bcs, ifsigman, ncurves, reserved0
     type(c_ptr) :: bc_list = c_null_ptr, bc_info = c_null_ptr, bc_or_not = c_null_ptr
     type(c_ptr) :: curve_npts = c_null_ptr, curve_t = c_null_ptr, curve_f = c_null_ptr
  end type sps_bcs_t

  type, bind(C) :: sps_stats_t
     integer(c_int64_t) :: n_pairs, m_pairs, n_by_type(10)
     integer(c_int32_t) :: ndivx, ndivy
     integer(c_int64_t) :: steps, pairs_total
  end type sps_stats_t

  interface
     integer(c_int) function sps_create(params, h) bind(C, name="sps_create")
       import :: c_int, c_ptr, sps_params_t
       type(sps_params_t), intent(in) :: params
       type(c_ptr), intent(out) :: h
     end function sps_create

     integer(c_int) function sps_set_state(h, st) bind(C, name="sps_set_state")
       import :: c_int, c_ptr, sps_state_t
       type(c_ptr), value :: h
       type(sps_state_t), intent(in) :: st
     end function sps_set_state

     integer(c_int) function sps_set_bcs(h, b) bind(C, name="sps_set_bcs")
       import :: c_int, c_ptr, sps_bcs_t
       type(c_ptr), value :: h
       type(sps_bcs_t), intent(in) :: b
     end function sps_set_bcs

     integer(c_int) function sps_get_surface(h, bc_or_not, normal) bind(C, name="sps_get_surface")
       import :: c_int, c_ptr
       type(c_ptr), value :: h, bc_or_not, normal
     end function sps_get_surface

     integer(c_int) function sps_step(h, nsteps, dt, time_sph, itimestep_sph) bind(C, name="sps_step")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
       real(c_double), value :: dt
       real(c_double), intent(inout) :: time_sph
       integer(c_int32_t), intent(inout) :: itimestep_sph
     end function sps_step

     integer(c_int) function sps_get_state(h, st) bind(C, name="sps_get_state")
       import :: c_int, c_ptr, sps_state_t
       type(c_ptr), value :: h
       type(sps_state_t), intent(in) :: st
     end function sps_get_state

     integer(c_int) function sps_get_pairs(h, cap, pair_i, pair_j, pint_type, w, dwdx, dwdy, n) &
          bind(C, name="sps_get_pairs")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int64_t), value :: cap
       type(c_ptr), value :: pair_i, pair_j, pint_type, w, dwdx, dwdy
       integer(c_int64_t), intent(out) :: n
     end function sps_get_pairs

     integer(c_int) function sps_get_stats(h, stats) bind(C, name="sps_get_stats")
       import :: c_int, c_ptr, sps_stats_t
       type(c_ptr), value :: h
       type(sps_stats_t), intent(out) :: stats
     end function sps_get_stats

     type(c_ptr) function sps_last_error() bind(C, name="sps_last_error")
       import :: c_ptr
     end function sps_last_error

     type(c_ptr) function sps_backend() bind(C, name="sps_backend")
       import :: c_ptr
     end function sps_backend

     subroutine sps_destroy(h) bind(C, name="sps_destroy")
       import :: c_ptr
       type(c_ptr), value :: h
     end subroutine sps_destroy
  end interface

contains

  ! C string -> Fortran string
  function sps_c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    interface
       integer(c_size_t) function c_strlen(str) bind(C, name="strlen")
         import :: c_ptr, c_size_t
         type(c_ptr), value :: str
       end function c_strlen
    end interface
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    s = transfer(chars(1:n), s)
  end function sps_c_string

  ! Stop with the library's message on a failed call, as the reference stops
  ! on its own errors (code/2_SPH_main_2018.f90:1476-1477).
  subroutine sps_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= SPS_OK) then
       write(*, '(A,A,A,I0,A,A)') 'sps: ', where, ' failed (', rc, '): ', sps_c_string(sps_last_error())
       stop 1
    end if
  end subroutine sps_check

end module sps_binding
