!-------------------------------------------------------------------------------
! sps_glue - drop-in replacement of the reference's time_integration
!
! A maintainer adds this file to the reference's build (after its modules)
! and replaces `call time_integration` in the driver
! (code/1_SPH_2018.f90:82) by `call sps_time_integration`.  Everything else
! of the reference stays: its driver loop, its input.txt/.dat/.pts readers
! (Init_sph, code/2_SPH_main_2018.f90:32-46) and its CSV/GiD writers.
!
! On its first call the glue packs the problem constants and the state that
! the readers left in the reference's module globals
! (code/5_SPH_global_vars_2018.f90:27-182) into the C-ABI (include/sps.h),
! creates a handle and uploads the state.  Every call then advances one
! step on the device and downloads the state back into the same globals, so
! the writers and the driver see what time_integration would have left.
! With sps_glue_sync_every_step set to .false. the download is skipped and a
! driver that needs the state only at output times calls sps_glue_sync; with
! sps_glue_sync_at_output also set, the glue downloads on exactly the steps
! after which the reference driver writes: those whose single-precision
! timers, which the driver advances before calling time_integration, reach
! their periods (code/1_SPH_2018.f90:72-100, 6_SPH_time_vars_2018.f90:31-36).
! The state stays on the device in between.
!-------------------------------------------------------------------------------
module sps_glue
  use, intrinsic :: iso_c_binding
  use sps_binding
  use variable_types
  use SPH_time_vars_2018
  use SPH_global_vars_2018
  use SPH_material_vars_2018, only: ic_grav, tcurve_grav, cgrav, ft_grav, dampingTG, bc_list, ic_tcurve
  implicit none
  private
  public :: sps_time_integration, sps_glue_sync, sps_glue_finish

  type(c_ptr), save :: handle = c_null_ptr
  ! fork of the reference the host was built from: code/ (0) or the Bui et
  ! al. (2008) fork (1) -- wall-force gate and CSPM determinant threshold
  integer, save, public :: sps_glue_fork = 0
  logical, save, public :: sps_glue_sync_every_step = .true.
  logical, save, public :: sps_glue_sync_at_output = .false.

contains

  subroutine sps_time_integration
    real(c_double) :: t
    integer(c_int32_t) :: it
    if (.not. c_associated(handle)) call open_handle()
    ! the driver has already advanced itimestep_sph (1_SPH_2018.f90:80) and
    ! advances time_sph after the call (:83): step from the previous counter
    t = time_sph
    it = int(itimestep_sph - 1, c_int32_t)
    call sps_check(sps_step(handle, 1_c_int32_t, real(dt_sph, c_double), t, it), 'sps_step')
    call neighbour_statistics()
    if (sps_glue_sync_every_step) then
       call sps_glue_sync()
    else if (sps_glue_sync_at_output) then
       if (t_plot_reset >= time_plot .or. t_print_reset >= time_print) call sps_glue_sync()
    end if
  end subroutine sps_time_integration

  subroutine sps_glue_sync
    if (.not. c_associated(handle)) return
    call get_arrays(x, vel, stress, internal_vars, f_drucker, displ, If_Out_Domain, grad_u, &
                    Ddev_strn, f_bound, art_visc, x_10, disp_10, .not. inside_approach)
    ! BC_or_not and the normals of the last free-surface detection (run
    ! every step with the position update, 2_...:150-154; OutputRes writes
    ! surface_points from them, 3_...:2910-2919)
    if (no_bcs > 0 .or. update_x) call get_surface(BC_or_not, normal)
  end subroutine sps_glue_sync

  ! The neighbour statistics grid_find_NEW keeps and writes to output_file
  ! every print_step steps (code/2_SPH_main_2018.f90:1240, 1355-1356,
  ! 1405-1431; nothing else reads them): each particle's pair count of the
  ! step's search (countiac),
  ! the largest and smallest with the first particle that has them, the
  ! average, niac (set to 0 by the search) and the particles with none.  The
  ! counts come from the library's pair list; the lines are written with the
  ! reference's list-directed statements, so output.txt is the reference's.
  subroutine neighbour_statistics
    integer(c_int64_t) :: n, k
    integer(c_int32_t), allocatable, target :: pi_(:), pj_(:)
    integer, allocatable :: cnt(:)
    integer :: i, sumiac, maxiac, miniac, noiac, maxp, minp
    niac = 0
    if (mod(itimestep_sph, print_step) /= 0) return
    if (sps_get_pairs(handle, 0_c_int64_t, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                      c_null_ptr, n) /= 0 .and. n == 0) call sps_check(-1_c_int, 'sps_get_pairs')
    allocate(pi_(max(n, 1_c_int64_t)), pj_(max(n, 1_c_int64_t)), cnt(ntotal2))
    call sps_check(sps_get_pairs(handle, n, c_loc(pi_), c_loc(pj_), c_null_ptr, c_null_ptr, c_null_ptr, &
                                 c_null_ptr, n), 'sps_get_pairs')
    cnt = 0
    do k = 1, n
       cnt(pi_(k)) = cnt(pi_(k)) + 1
       cnt(pj_(k)) = cnt(pj_(k)) + 1
    end do
    sumiac = 0
    maxiac = 0
    miniac = 1000
    noiac  = 0
    maxp = 0
    minp = 0
    do i = 1, ntotal2
       sumiac = sumiac + cnt(i)
       if (cnt(i) > maxiac) then
          maxiac = cnt(i)
          maxp = i
       end if
       if (cnt(i) < miniac) then
          miniac = cnt(i)
          minp = i
       end if
       if (cnt(i) == 0) noiac = noiac + 1
    end do
    write(output_file,*) ' >> Statistics: interactions per particle:', time, dt_sph
    write(output_file,*)'**** Particle:',maxp, ' maximal interactions:',maxiac
    write(output_file,*)'**** Particle:',minp, ' minimal interactions:',miniac
    write(output_file,*)'**** Average :',real(sumiac)/real(ntotal2)
    write(output_file,*)'**** Total pairs : ',niac
    write(output_file,*)'**** Particles with no interactions:',noiac
  end subroutine neighbour_statistics

  subroutine sps_glue_finish
    if (c_associated(handle)) call sps_destroy(handle)
    handle = c_null_ptr
  end subroutine sps_glue_finish

  subroutine open_handle
    type(sps_params_t) :: p
    integer :: k, n, st
    real, save :: no_walls(1) = 0.0
    character(len=16) :: mode
    ! SPS_GLUE_SYNC=output keeps the state on the device between the steps
    ! after which the driver writes (a driver left as the reference has it
    ! cannot set the module variables itself); =step copies it back every step
    call get_environment_variable('SPS_GLUE_SYNC', mode, status=st)
    if (st == 0 .and. trim(mode) == 'output') then
       sps_glue_sync_every_step = .false.
       sps_glue_sync_at_output = .true.
    else if (st == 0 .and. trim(mode) == 'step') then
       sps_glue_sync_every_step = .true.
    end if
    p%abi_version = SPS_ABI_VERSION
    p%nnode = nnode; p%nstress = nstress
    p%ndummy = merge(ndummy, 0, dummy_nodes)
    p%ndummy2 = merge(ndummy2, 0, dummy_nodes)
    p%nint_vars = nint_Vars
    p%ndimn = ndimn; p%nstre = nstre; p%skf = skf
    p%sp_sph = merge(1, 0, SP_SPH); p%inside_approach = merge(1, 0, inside_approach)
    p%cspm = merge(1, 0, cspm); p%update_x = merge(1, 0, update_x); p%xsph = merge(1, 0, XSPH)
    p%cont_density = merge(1, 0, cont_density); p%art_stress = merge(1, 0, art_stress)
    p%dummy_nodes = merge(1, 0, dummy_nodes)
    ! the outside approach's controls are read only for it (3_...:371-378)
    p%sph_shift = 0; p%npoints = npoints; p%vel_vector = 0; p%shift_update = 1; p%reserved0 = 0
    p%r_x = 0.0_c_double; p%r_y = 0.0_c_double; p%disp_tol = 0.0_c_double
    if (.not. inside_approach) then
       p%sph_shift = merge(1, 0, SPH_shift); p%vel_vector = merge(1, 0, vel_vector)
       p%shift_update = shift_update; p%r_x = r_x; p%r_y = r_y; p%disp_tol = disp_tol
    end if
    p%no_bcs = no_bcs
    p%bforce_gate = sps_glue_fork
    p%mode = SPS_MODE_ORDERED
    p%props = 0.0_c_double
    p%props(1:20) = props(1, 1:20)
    p%pi = pi; p%dx = dx; p%alpha = alpha; p%beta = beta; p%damping = dampingTG
    if (sps_glue_fork == 1) then
       p%cspm_det_tol = real(1e-07, c_double)
    else
       p%cspm_det_tol = real(1e-03, c_double)
    end if
    p%D11 = D11; p%D22 = D22; p%D12 = D12; p%D33 = D33; p%D41 = D41; p%D42 = D42
    p%xmin_domain = Xmin_Domain(1:2); p%xmax_domain = Xmax_Domain(1:2)
    p%ic_grav = ic_grav
    p%cgrav = 0.0_c_double; p%ft_grav = 0.0_c_double; p%tcurve_npts = 0
    p%tcurve_t = 0.0_c_double; p%tcurve_f = 0.0_c_float
    if (ic_grav == 1) then
       p%cgrav = cgrav(1:2); p%ft_grav = ft_grav
       n = nptstcurves(tcurve_grav)
       p%tcurve_npts = n
       do k = 1, min(n, SPS_MAX_TCURVE)
          p%tcurve_t(k) = ttcurves(tcurve_grav, k)
          p%tcurve_f(k) = ftcurves(tcurve_grav, k)
       end do
    end if
    call sps_check(sps_create(p, handle), 'sps_create')
    if (dummy_nodes) then
       call set_arrays(x, vel, stress, mass, rho, hsml, itype, If_Out_Domain, internal_vars, &
                       f_drucker, x00, displ, wall_position, horizontal_or_not, .true., &
                       x_10, disp_10, .not. inside_approach)
    else
       call set_arrays(x, vel, stress, mass, rho, hsml, itype, If_Out_Domain, internal_vars, &
                       f_drucker, x00, displ, no_walls, no_walls, .false., &
                       x_10, disp_10, .not. inside_approach)
    end if
    if (no_bcs > 0) call set_bcs(bc_info, BC_or_not)
  end subroutine open_handle

  ! Prescribed BCs: bc_list as read, the nodes' columns of bc_info,
  ! BC_or_not of every particle, and the time curves repacked one after another
  ! (ttcurves(ntcurves, mpts) holds a curve's points with stride ntcurves).
  subroutine set_bcs(info_, bcn_)
    integer, target, intent(in) :: info_(*), bcn_(*)
    type(sps_bcs_t) :: b
    integer(c_int32_t), target :: npts(SPS_MAX_CURVES)
    real(c_double), target :: ct(SPS_MAX_TCURVE, SPS_MAX_CURVES)
    real(c_float), target :: cf(SPS_MAX_TCURVE, SPS_MAX_CURVES)
    real(c_double), allocatable, target :: bl(:,:)
    integer :: c, k, nc
    npts = 0; ct = 0.0_c_double; cf = 0.0_c_float
    nc = 0
    if (ic_tcurve == 1) nc = min(ntcurves, SPS_MAX_CURVES)
    do c = 1, nc
       npts(c) = min(nptstcurves(c), SPS_MAX_TCURVE)
       do k = 1, npts(c)
          ct(k, c) = ttcurves(c, k)
          cf(k, c) = ftcurves(c, k)
       end do
    end do
    allocate(bl(8, no_bcs))
    bl = bc_list(1:8, 1:no_bcs)
    b%no_bcs = no_bcs; b%ifsigman = ifsigman; b%ncurves = nc; b%reserved0 = 0
    b%bc_list = c_loc(bl); b%bc_info = c_loc(info_); b%bc_or_not = c_loc(bcn_)
    b%curve_npts = c_loc(npts); b%curve_t = c_loc(ct); b%curve_f = c_loc(cf)
    call sps_check(sps_set_bcs(handle, b), 'sps_set_bcs')
    deallocate(bl)
  end subroutine set_bcs

  ! The module arrays arrive as assumed-size TARGET dummies, so their
  ! addresses are valid for the duration of the C call made here.
  subroutine set_arrays(x_, vel_, stress_, mass_, rho_, hsml_, itype_, ood_, ivars_, fdr_, x00_, &
                        displ_, wall_, horiz_, with_walls, x10_, d10_, outside)
    real(DP), target, intent(in) :: x_(*), vel_(*), stress_(*), mass_(*), rho_(*), hsml_(*)
    real(DP), target, intent(in) :: ivars_(*), fdr_(*), x00_(*), displ_(*), x10_(*), d10_(*)
    integer, target, intent(in) :: itype_(*), ood_(*)
    real, target, intent(in) :: wall_(*), horiz_(*)
    logical, intent(in) :: with_walls, outside
    type(sps_state_t) :: st
    st%x = c_loc(x_); st%vel = c_loc(vel_); st%stress = c_loc(stress_)
    st%mass = c_loc(mass_); st%rho = c_loc(rho_); st%hsml = c_loc(hsml_)
    st%itype = c_loc(itype_); st%if_out_domain = c_loc(ood_)
    st%internal_vars = c_loc(ivars_); st%f_drucker = c_loc(fdr_)
    st%x00 = c_loc(x00_); st%displ = c_loc(displ_)
    if (with_walls) then
       st%wall_position = c_loc(wall_); st%horizontal_or_not = c_loc(horiz_)
    end if
    if (outside) then
       st%x_10 = c_loc(x10_); st%disp_10 = c_loc(d10_)
    end if
    call sps_check(sps_set_state(handle, st), 'sps_set_state')
  end subroutine set_arrays

  subroutine get_arrays(x_, vel_, stress_, ivars_, fdr_, displ_, ood_, gradu_, ddev_, fb_, av_, &
                        x10_, d10_, outside)
    real(DP), target, intent(inout) :: x_(*), vel_(*), stress_(*), ivars_(*), fdr_(*), displ_(*)
    real(DP), target, intent(inout) :: gradu_(*), ddev_(*), fb_(*), av_(*), x10_(*), d10_(*)
    logical, intent(in) :: outside
    integer, target, intent(inout) :: ood_(*)
    type(sps_state_t) :: st
    st%x = c_loc(x_); st%vel = c_loc(vel_); st%stress = c_loc(stress_)
    st%internal_vars = c_loc(ivars_); st%f_drucker = c_loc(fdr_); st%displ = c_loc(displ_)
    st%if_out_domain = c_loc(ood_); st%grad_u = c_loc(gradu_); st%ddev_strn = c_loc(ddev_)
    st%f_bound = c_loc(fb_); st%art_visc = c_loc(av_)
    if (outside) then
       st%x_10 = c_loc(x10_); st%disp_10 = c_loc(d10_)
    end if
    call sps_check(sps_get_state(handle, st), 'sps_get_state')
  end subroutine get_arrays

  ! BC_or_not (ntotal) and normal (2,ntotal) as the last free-surface
  ! detection left them, for the driver's surface_points writer.
  subroutine get_surface(bcn_, nrm_)
    integer, target, intent(inout) :: bcn_(*)
    real(DP), target, intent(inout) :: nrm_(*)
    call sps_check(sps_get_surface(handle, c_loc(bcn_), c_loc(nrm_)), 'sps_get_surface')
  end subroutine get_surface

end module sps_glue
