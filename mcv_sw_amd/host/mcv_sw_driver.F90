! Fortran host program of the GPU MCV shallow-water model: the host keeps the
! namelist, mesh and projection setup, initial state, history cadence and
! diagnostics of MCV_SW.F90:30-82, and hands the time integration to the
! MI355X library through the iso_c_binding interfaces of mcv_hip_binding.
!
! namelist.input groups:
!   &time_settings  run_days, run_hours, run_minutes, run_seconds, dt,
!                   history_interval, integral_scheme   (parameters_mod.F90)
!   &case_select    case_num: 2, 5, 6 (test_case_mod.F90) or 0, the balanced
!                   jet of &jet_settings
!   &domain         dx (degrees; 90/dx cells per panel edge), dy = dx
!   &gpu_settings   dof (3 = MCV3, 4 = MCV4; the reference's -DMCV3/-DMCV4),
!                   precision ('FP64' or 'FP32C'), device, dump_file,
!                   history_file (default 'mcv_output.nc', the reference's
!                   ncFile of output_mod.F90:10; '' writes no history)
!   &jet_settings   u0, h0, lmax, seed, rms_h
! History: the reference's netCDF file and cadence (MCV_SW.F90:49-50,62-70),
! the first record from the initial arrays, later ones from the device.
! The final state (phi, u, v on (nh, nh, 6), stream access) is written to
! dump_file when it is set.
module mcv_hip_binding
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: MCV_ABI_VERSION = 1
  integer(c_int32_t), parameter :: MCV_RK3_TVD = 3, MCV_RK4 = 4
  integer(c_int32_t), parameter :: MCV_FP64 = 0, MCV_FP32C = 1

  type, bind(C) :: mcv_config
    integer(c_int32_t) :: abi_version, dof, n, scheme, precision, device
    real(c_double) :: dt, dx
  end type mcv_config

  type, bind(C) :: mcv_mesh
    type(c_ptr) :: sqrtG, matrixG, matrixIG, matrixA, matrixIA, f, phi_s, ghost_coef, ghost_iref
  end type mcv_mesh

  type, bind(C) :: mcv_state
    type(c_ptr) :: phi, phiG, u, v, contraU, contraV
  end type mcv_state

  type, bind(C) :: mcv_history_info
    integer(c_int32_t) :: dof, n, xhalo, yhalo, case_num
    real(c_double) :: dx, dy, dt
  end type mcv_history_info

  type, bind(C) :: mcv_history_fields
    type(c_ptr) :: u, v, contraU, contraV, phi, phi_t, zonal, meridional, uC, vC, phiC
  end type mcv_history_fields

  interface
    integer(c_int) function mcv_hip_init(cfg, mesh, initial, h) bind(C, name='mcv_hip_init')
      import :: c_int, c_ptr, mcv_config, mcv_mesh, mcv_state
      type(mcv_config), intent(in) :: cfg
      type(mcv_mesh), intent(in) :: mesh
      type(mcv_state), intent(in) :: initial
      type(c_ptr), intent(out) :: h
    end function mcv_hip_init

    integer(c_int) function mcv_hip_step(h, nsteps) bind(C, name='mcv_hip_step')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps
    end function mcv_hip_step

    integer(c_int) function mcv_hip_sync(h, out, mass, energy) bind(C, name='mcv_hip_sync')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      type(c_ptr), value :: out  ! c_loc of an mcv_state, or c_null_ptr
      real(c_double), intent(out) :: mass, energy
    end function mcv_hip_sync

    integer(c_int) function mcv_history_init(path, info, lonC, latC, lonP, latP, areaCell, &
                                             phi_s) bind(C, name='mcv_history_init')
      import :: c_int, c_char, c_double, mcv_history_info
      character(kind=c_char), intent(in) :: path(*)
      type(mcv_history_info), intent(in) :: info
      real(c_double), intent(in) :: lonC(*), latC(*), lonP(*), latP(*), areaCell(*), phi_s(*)
    end function mcv_history_init

    integer(c_int) function mcv_history_write(path, time_slot, dof, n, rec) &
        bind(C, name='mcv_history_write')
      import :: c_int, c_int32_t, c_char, mcv_history_fields
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: time_slot, dof, n
      type(mcv_history_fields), intent(in) :: rec
    end function mcv_history_write

    integer(c_int) function mcv_hip_history_write(h, path, time_slot) &
        bind(C, name='mcv_hip_history_write')
      import :: c_int, c_int32_t, c_char, c_ptr
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: time_slot
    end function mcv_hip_history_write

    integer(c_int) function mcv_hip_free(h) bind(C, name='mcv_hip_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mcv_hip_free

    type(c_ptr) function mcv_hip_last_error() bind(C, name='mcv_hip_last_error')
      import :: c_ptr
    end function mcv_hip_last_error
  end interface

contains

  ! Stops with the library's message, as the reference stops on errors.
  subroutine mcv_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: k
    if (rc == 0) return
    call c_f_pointer(mcv_hip_last_error(), msg, [512])
    text = ''
    do k = 1, 512
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    enddo
    print *, what//' failed: '//trim(text)
    stop 1
  end subroutine mcv_check

end module mcv_hip_binding

program mcv_sw_gpu
  use iso_c_binding
  use mcv_hip_binding
  use mcv_host_lib
  implicit none

  real(c_double), parameter :: radius = 6371220.0_c_double, omega = 7.292e-5_c_double
  real(c_double), parameter :: gravity = 9.80616_c_double
  real(c_double), parameter :: d2r = 2.0_c_double * asin(1.0_c_double) / 180.0_c_double

  ! namelist variables (reference names and defaults, parameters_mod.F90:113-130)
  integer :: run_days = 1, run_hours = 0, run_minutes = 0, run_seconds = 0
  real(c_double) :: dt = 300.0_c_double, dx = 2.0_c_double, dy = 2.0_c_double
  integer :: history_interval = 360, xhalo = 1, yhalo = 1
  character(len=16) :: integral_scheme = 'RK4'
  integer :: case_num = 2
  integer :: dof = 3, device = 0
  character(len=16) :: precision = 'FP64'
  character(len=256) :: dump_file = '', history_file = 'mcv_output.nc'
  real(c_double) :: u0 = 20.0_c_double, h0 = 5960.0_c_double, rms_h = 30.0_c_double
  integer :: lmax = 8
  integer(c_int64_t) :: seed = 0
  namelist /time_settings/ run_days, run_hours, run_minutes, run_seconds, dt, &
                           history_interval, integral_scheme
  namelist /case_select/ case_num
  namelist /domain/ dx, dy, xhalo, yhalo
  namelist /gpu_settings/ dof, precision, device, dump_file, history_file
  namelist /jet_settings/ u0, h0, lmax, seed, rms_h

  integer :: n, H, npv, nh, np, ios, it, nsteps, pending, output_idx, total_output_num
  integer(c_int) :: rc
  integer :: total_run_time
  real(c_double), allocatable, target :: xP(:), yP(:), lonP(:), latP(:), sqrtG(:), f(:)
  real(c_double), allocatable, target :: G(:, :), IG(:, :), A(:, :), IA(:, :)
  real(c_double), allocatable, target :: coef(:, :), phi_s(:)
  integer(c_int32_t), allocatable, target :: iref(:, :)
  real(c_double), allocatable, target :: phi(:), zonal(:), merid(:), phiG(:), u(:), v(:)
  real(c_double), allocatable, target :: cu(:), cv(:)
  real(c_double), allocatable, target :: lonC(:), latC(:), areaCell(:), phi_t(:), zero_c(:)
  type(mcv_config) :: cfg
  type(mcv_mesh) :: mesh
  type(mcv_state), target :: st
  type(mcv_history_info) :: hinfo
  type(mcv_history_fields) :: hrec
  type(c_ptr) :: solve
  real(c_double) :: mass0, energy0, mass, energy, mass_old, energy_old
  logical :: history
  integer(8) :: t_start, t_end, t_rate

  call system_clock(t_start, t_rate)
  print *, 'Start the MCV Shallow Water model (MI355X time integration)'
  open(1, file='namelist.input', status='old')
  read(1, nml=time_settings)
  read(1, nml=case_select)
  read(1, nml=domain)
  read(1, nml=gpu_settings, iostat=ios)
  if (ios > 0) stop 'bad &gpu_settings'
  rewind(1)
  if (case_num == 0) then
    read(1, nml=jet_settings, iostat=ios)
    if (ios > 0) stop 'bad &jet_settings'
  endif
  close(1)

  ! parameters_mod.F90:131-160
  total_run_time = run_days * 86400 + run_hours * 3600 + run_minutes * 60 + run_seconds
  if (90.0_c_double / dx - int(90.0_c_double / dx) /= 0.) &
    stop '90 divide dx must be integer, choose another dx'
  n = int(90.0_c_double / dx)
  H = dof - 1
  npv = n * H + 1
  nh = npv + 2 * H
  np = nh * nh * 6
  nsteps = int(total_run_time / dt)
  if (trim(adjustl(integral_scheme)) /= 'RK3_TVD' .and. trim(adjustl(integral_scheme)) /= 'RK4') &
    stop 'Unknown integral scheme, please select from RK3_TVD or RK4 ...'

  ! mesh and ghost tables (mesh_mod.F90, ghost_mod.f90)
  allocate(xP(np), yP(np), lonP(np), latP(np), sqrtG(np), f(np), phi_s(np))
  allocate(G(4, np), IG(4, np), A(4, np), IA(4, np), coef(npv, H), iref(npv, H))
  if (mcv_host_mesh(dof, n, radius, omega, xP, yP, lonP, latP, sqrtG, G, IG, A, IA, f) /= 0) &
    stop 'mesh setup failed'
  if (mcv_host_ghost(dof, n, coef, iref) /= 0) stop 'Ghost cell location error'

  ! initial state (test_case_mod.F90)
  allocate(phi(np), zonal(np), merid(np), phiG(np), u(np), v(np), cu(np), cv(np))
  if (case_num == 0) then
    if (mcv_host_jet(np, radius, omega, gravity, u0, h0, lmax, seed, rms_h, lonP, latP, &
                     phi, zonal, merid, phi_s) /= 0) stop 'jet setup failed'
  else
    if (mcv_host_case(case_num, np, radius, omega, gravity, lonP, latP, phi, zonal, merid, &
                      phi_s) /= 0) stop 'Unknown case_num'
  endif
  if (mcv_host_handoff(np, sqrtG, G, IA, phi, zonal, merid, phiG, u, v, cu, cv) /= 0) &
    stop 'state handoff failed'

  print *, 'The MCV order is ', dof
  print *, 'Temporal integration scheme is '//trim(adjustl(integral_scheme))
  print *, 'Arithmetic is '//trim(adjustl(precision))

  cfg%abi_version = MCV_ABI_VERSION
  cfg%dof = dof
  cfg%n = n
  cfg%scheme = merge(MCV_RK4, MCV_RK3_TVD, trim(adjustl(integral_scheme)) == 'RK4')
  cfg%precision = merge(MCV_FP32C, MCV_FP64, trim(adjustl(precision)) == 'FP32C')
  cfg%device = device
  cfg%dt = dt
  cfg%dx = dx * d2r
  mesh = mcv_mesh(c_loc(sqrtG), c_loc(G), c_loc(IG), c_loc(A), c_loc(IA), c_loc(f), &
                  c_loc(phi_s), c_loc(coef), c_loc(iref))
  st = mcv_state(c_loc(phi), c_loc(phiG), c_loc(u), c_loc(v), c_loc(cu), c_loc(cv))
  call mcv_check(mcv_hip_init(cfg, mesh, st, solve), 'mcv_hip_init')

  ! history cadence of MCV_SW.F90:48-77
  output_idx = 1
  total_output_num = int(nsteps * dt / history_interval)
  history = len_trim(history_file) > 0
  if (history) then
    ! history_init + history_write_stat(stat(old), 1) (MCV_SW.F90:49-50); the
    ! reference's cell averages are not formed before the first record
    allocate(lonC((n + 2) * (n + 2) * 6), latC((n + 2) * (n + 2) * 6), &
             areaCell((n + 2) * (n + 2) * 6), zero_c((n + 2) * (n + 2) * 6), phi_t(np))
    if (mcv_host_cells(dof, n, radius, lonC, latC, areaCell) /= 0) stop 'cell setup failed'
    hinfo = mcv_history_info(dof, n, xhalo, yhalo, case_num, dx * d2r, dy * d2r, dt)
    call mcv_check(mcv_history_init(trim(history_file)//c_null_char, hinfo, lonC, latC, lonP, &
                                    latP, areaCell, phi_s), 'history_init')
    phi_t = phi + phi_s
    zero_c = 0.0_c_double
    hrec = mcv_history_fields(c_loc(u), c_loc(v), c_loc(cu), c_loc(cv), c_loc(phi), &
                              c_loc(phi_t), c_loc(zonal), c_loc(merid), c_loc(zero_c), &
                              c_loc(zero_c), c_loc(zero_c))
    call mcv_check(mcv_history_write(trim(history_file)//c_null_char, 1, int(dof, c_int32_t), &
                                     int(n, c_int32_t), hrec), &
                   'history_write_stat')
  endif
  call mcv_check(mcv_hip_sync(solve, c_null_ptr, mass0, energy0), 'mcv_hip_sync')
  print *, 'output index/total, MCR, ECR :', output_idx - 1, '/', total_output_num, ' ', 0., 0.
  pending = 0
  do it = 1, nsteps
    pending = pending + 1
    if (mod(it * dt, real(history_interval, c_double)) == 0. .and. &
        it * dt >= history_interval) then
      ! the energy ratio is the one of stat(old), the state before this step
      ! (MCV_SW.F90:72)
      if (pending > 1) call mcv_check(mcv_hip_step(solve, int(pending - 1, c_int32_t)), &
                                      'mcv_hip_step')
      call mcv_check(mcv_hip_sync(solve, c_null_ptr, mass_old, energy_old), 'mcv_hip_sync')
      call mcv_check(mcv_hip_step(solve, 1_c_int32_t), 'mcv_hip_step')
      pending = 0
      output_idx = output_idx + 1
      if (history) call mcv_check(mcv_hip_history_write(solve, trim(history_file)//c_null_char, &
                                  int(output_idx, c_int32_t)), 'history_write_stat')
      call mcv_check(mcv_hip_sync(solve, c_null_ptr, mass, energy), 'mcv_hip_sync')
      print *, 'output index/total, MCR, ECR :', output_idx - 1, '/', total_output_num, ' ', &
               (mass - mass0) / mass0, (energy_old - energy0) / energy0
    endif
  enddo
  if (pending > 0) call mcv_check(mcv_hip_step(solve, int(pending, c_int32_t)), 'mcv_hip_step')
  call mcv_check(mcv_hip_sync(solve, c_loc(st), mass, energy), 'mcv_hip_sync')
  if (len_trim(dump_file) > 0) then
    open(2, file=trim(dump_file), access='stream', form='unformatted', status='replace')
    write(2) phi, u, v
    close(2)
  endif
  rc = mcv_hip_free(solve)
  call system_clock(t_end)
  print *, 'It took ', dble(t_end - t_start) / dble(t_rate), ' seconds to run this program'
end program mcv_sw_gpu
