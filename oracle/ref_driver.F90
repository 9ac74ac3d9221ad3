! Test-infrastructure driver for the reference build (oracle/_ref).
!
! This program is NOT part of the product.  It links the reference's own
! Fortran modules, compiled unmodified from the reference's src/ directory
! (constants, parameters, projection, mesh, stat, tend, ghost,
! spatial_operators, diag, test_case), and uses them to write golden vectors
! and to time the reference CPU path.
!
! time_scheme_mod.F90 cannot be compiled here: it imports output_mod, which
! needs the netCDF-Fortran library that this image lacks.  Its orchestration
! (time_scheme_mod.F90:13-142: RK4, RK3_TVD, update_stat*, correct_bdy_ghost)
! is therefore restated below, expression by expression, and it calls the
! reference's compiled spatial_operator, fill_ghost, unify_bdy_stat and wind
! conversions.  Everything numerical below the time loop is reference code.
!
! Inputs (current directory):
!   namelist.input  reference namelist (parameters_mod.F90:86-99)
!   driver.nml      &driver nsteps, ic_file, out_dir, dump_steps, history_every,
!                   dump_mesh, dump_states
! Outputs (out_dir): mesh.bin, ghost.bin, cells.bin, state_NNNNNNN.bin (raw
! float64, Fortran order, halo included), via_NNNNNNN.bin (cell averages of
! history steps), and one JSON timing line on stdout.
module refdrv_scheme
  use parameters_mod
  use stat_mod
  use tend_mod
  use mesh_mod
  use ghost_mod
  use spatial_operators_mod
  implicit none

contains

  ! time_scheme_mod.F90:125-142 ("Scheme 2")
  subroutine drv_boundary(s)
    type(stat_field), intent(inout) :: s
    call convert_ghost_wind_P2SP (s)
    call fill_ghost              (s)
    call unify_bdy_stat          (s)
    call convert_ghost_wind_SP2P (s)
    call convert_wind_cov2contrav(s)
  end subroutine drv_boundary

  ! time_scheme_mod.F90:81-85: phi from phiG, boundary pipeline, phiG from phi.
  subroutine drv_finish_stage(s)
    type(stat_field), intent(inout) :: s
    s%phi = s%phiG / mesh%sqrtG
    call drv_boundary(s)
    s%phiG = s%phi * mesh%sqrtG
  end subroutine drv_finish_stage

  ! time_scheme_mod.F90:71-87: Y = Yb + inc*k on the interior.
  subroutine drv_euler(sn, sb, td, inc)
    type(stat_field), intent(inout) :: sn
    type(stat_field), intent(in   ) :: sb
    type(tend_field), intent(in   ) :: td
    real            , intent(in   ) :: inc
    sn%phiG(ids:ide,jds:jde,ifs:ife) = sb%phiG(ids:ide,jds:jde,ifs:ife) + inc * td%phiG(ids:ide,jds:jde,ifs:ife)
    sn%u   (ids:ide,jds:jde,ifs:ife) = sb%u   (ids:ide,jds:jde,ifs:ife) + inc * td%u   (ids:ide,jds:jde,ifs:ife)
    sn%v   (ids:ide,jds:jde,ifs:ife) = sb%v   (ids:ide,jds:jde,ifs:ife) + inc * td%v   (ids:ide,jds:jde,ifs:ife)
    call drv_finish_stage(sn)
  end subroutine drv_euler

  ! time_scheme_mod.F90:89-105: Y = 3/4 Yn + 1/4 Y1 + 1/4 dt k.
  subroutine drv_ssp_mid(sn, sb, s1, td)
    type(stat_field), intent(inout) :: sn
    type(stat_field), intent(in   ) :: sb, s1
    type(tend_field), intent(in   ) :: td
    sn%phiG(ids:ide,jds:jde,ifs:ife) = 0.75 * sb%phiG(ids:ide,jds:jde,ifs:ife) + 0.25 * s1%phiG(ids:ide,jds:jde,ifs:ife) + 0.25 * dt * td%phiG(ids:ide,jds:jde,ifs:ife)
    sn%u   (ids:ide,jds:jde,ifs:ife) = 0.75 * sb%u   (ids:ide,jds:jde,ifs:ife) + 0.25 * s1%u   (ids:ide,jds:jde,ifs:ife) + 0.25 * dt * td%u   (ids:ide,jds:jde,ifs:ife)
    sn%v   (ids:ide,jds:jde,ifs:ife) = 0.75 * sb%v   (ids:ide,jds:jde,ifs:ife) + 0.25 * s1%v   (ids:ide,jds:jde,ifs:ife) + 0.25 * dt * td%v   (ids:ide,jds:jde,ifs:ife)
    call drv_finish_stage(sn)
  end subroutine drv_ssp_mid

  ! time_scheme_mod.F90:107-123: Y = Yn/3 + 2/3 Y2 + 2/3 dt k.
  subroutine drv_ssp_last(sn, sb, s2, td)
    type(stat_field), intent(inout) :: sn
    type(stat_field), intent(in   ) :: sb, s2
    type(tend_field), intent(in   ) :: td
    sn%phiG(ids:ide,jds:jde,ifs:ife) = sb%phiG(ids:ide,jds:jde,ifs:ife) / 3. + 2./3. * s2%phiG(ids:ide,jds:jde,ifs:ife) + 2./3. * dt * td%phiG(ids:ide,jds:jde,ifs:ife)
    sn%u   (ids:ide,jds:jde,ifs:ife) = sb%u   (ids:ide,jds:jde,ifs:ife) / 3. + 2./3. * s2%u   (ids:ide,jds:jde,ifs:ife) + 2./3. * dt * td%u   (ids:ide,jds:jde,ifs:ife)
    sn%v   (ids:ide,jds:jde,ifs:ife) = sb%v   (ids:ide,jds:jde,ifs:ife) / 3. + 2./3. * s2%v   (ids:ide,jds:jde,ifs:ife) + 2./3. * dt * td%v   (ids:ide,jds:jde,ifs:ife)
    call drv_finish_stage(sn)
  end subroutine drv_ssp_last

  ! time_scheme_mod.F90:13-44
  subroutine drv_rk4(snew, sold)
    type(stat_field), intent(inout) :: snew
    type(stat_field), intent(in   ) :: sold
    call copyStat(stat(-1), sold)
    call spatial_operator(stat(-1), tend(-1))
    call drv_euler(stat(-2), stat(-1), tend(-1), 0.5 * dt)
    call spatial_operator(stat(-2), tend(-2))
    call drv_euler(stat(-3), stat(-1), tend(-2), 0.5 * dt)
    call spatial_operator(stat(-3), tend(-3))
    call drv_euler(stat(-4), stat(-1), tend(-3), dt)
    call spatial_operator(stat(-4), tend(-4))
    tend(1)%phiG = (tend(-1)%phiG + 2. * tend(-2)%phiG + 2. * tend(-3)%phiG + tend(-4)%phiG) / 6.
    tend(1)%u    = (tend(-1)%u    + 2. * tend(-2)%u    + 2. * tend(-3)%u    + tend(-4)%u   ) / 6.
    tend(1)%v    = (tend(-1)%v    + 2. * tend(-2)%v    + 2. * tend(-3)%v    + tend(-4)%v   ) / 6.
    call drv_euler(snew, sold, tend(1), dt)
  end subroutine drv_rk4

  ! time_scheme_mod.F90:46-69
  subroutine drv_rk3(snew, sold)
    type(stat_field), intent(inout) :: snew
    type(stat_field), intent(in   ) :: sold
    call copyStat(stat(-1), sold)
    call spatial_operator(stat(-1), tend(-1))
    call drv_euler(stat(-2), stat(-1), tend(-1), dt)
    call spatial_operator(stat(-2), tend(-2))
    call drv_ssp_mid(stat(-3), stat(-1), stat(-2), tend(-2))
    call spatial_operator(stat(-3), tend(-3))
    call drv_ssp_last(snew, stat(-1), stat(-3), tend(-3))
  end subroutine drv_rk3

  ! output_mod.F90:230-248 (addFillValue), restated because output_mod needs netCDF.
  subroutine drv_fill_corners(s)
    type(stat_field), intent(inout) :: s
    s%u(ips:ids-1,jps:jds-1,:) = FillValue
    s%v(ips:ids-1,jps:jds-1,:) = FillValue
    s%u(ide+1:ipe,jps:jds-1,:) = FillValue
    s%v(ide+1:ipe,jps:jds-1,:) = FillValue
    s%u(ips:ids-1,jde+1:jpe,:) = FillValue
    s%v(ips:ids-1,jde+1:jpe,:) = FillValue
    s%u(ide+1:ipe,jde+1:jpe,:) = FillValue
    s%v(ide+1:ipe,jde+1:jpe,:) = FillValue
  end subroutine drv_fill_corners

end module refdrv_scheme

program ref_driver
  use parameters_mod
  use stat_mod
  use tend_mod
  use mesh_mod
  use ghost_mod
  use test_case_mod
  use spatial_operators_mod
  use diag_mod
  use refdrv_scheme
  implicit none

  integer            :: nsteps_drv, history_every, it, k, u
  logical            :: dump_mesh, dump_states, do_dump
  character(len=512) :: ic_file, out_dir
  integer            :: dump_steps(64)
  integer(8)         :: c0, c1, crate, d0, d1, dump_ticks
  real(8)            :: total_mass, total_energy

  namelist /driver/ nsteps_drv, ic_file, out_dir, dump_steps, history_every, dump_mesh, &
                    dump_states

  nsteps_drv    = 1
  ic_file       = ''
  out_dir       = '.'
  dump_steps    = -1
  history_every = 0
  dump_mesh     = .true.
  dump_states   = .true.

  call initParameters
  call initMesh
  call initStat
  call initTend
  call initGhost

  open(newunit=u, file='driver.nml', status='old')
  read(u, nml=driver)
  close(u)

  if (len_trim(ic_file) > 0) then
    ! Injected initial state: phi, zonal, meridional, phi_s on all points.
    ! case_num=0 makes initTestCase skip its case routines and run only its
    ! tail (test_case_mod.F90:30-43) on the injected fields.
    open(newunit=u, file=trim(ic_file), access='stream', form='unformatted', status='old')
    read(u) stat(0)%phi
    read(u) stat(0)%zonal_wind
    read(u) stat(0)%meridional_wind
    read(u) mesh%phi_s
    close(u)
    case_num = 0
  endif
  call initTestCase

  if (dump_mesh) then
    open(newunit=u, file=trim(out_dir)//'/mesh.bin', access='stream', form='unformatted', status='replace')
    write(u) mesh%xP, mesh%yP, mesh%lonP, mesh%latP, mesh%sqrtG
    write(u) mesh%matrixG, mesh%matrixIG, mesh%matrixA, mesh%matrixIA
    write(u) mesh%f, mesh%phi_s
    close(u)
    open(newunit=u, file=trim(out_dir)//'/ghost.bin', access='stream', form='unformatted', status='replace')
    write(u) ghost%coef
    write(u) ghost%iref
    close(u)
    ! cell geometry of the history file (mesh_mod.F90:165-214)
    open(newunit=u, file=trim(out_dir)//'/cells.bin', access='stream', form='unformatted', status='replace')
    write(u) mesh%lonC, mesh%latC, mesh%areaCell
    close(u)
  endif

  if (dump_states) call dump_state(0, stat(0))

  dump_ticks = 0
  call system_clock(c0, crate)
  do it = 1, nsteps_drv
    if (trim(adjustl(integral_scheme)) == 'RK3_TVD') call drv_rk3(stat(1), stat(0))
    if (trim(adjustl(integral_scheme)) == 'RK4'    ) call drv_rk4(stat(1), stat(0))

    ! Output-time side effects of MCV_SW.F90:62-74 on stat(new).
    if (history_every > 0) then
      if (mod(it, history_every) == 0) then
        call drv_fill_corners(stat(1))
        call convert_wind_P2SP(stat(1))
        call calc_VIA(stat(1)%uC  , stat(1)%u  )
        call calc_VIA(stat(1)%vC  , stat(1)%v  )
        call calc_VIA(stat(1)%phiC, stat(1)%phi)
      endif
    endif

    do_dump = (it == nsteps_drv) .and. dump_states
    do k = 1, size(dump_steps)
      if (dump_steps(k) == it) do_dump = .true.
    enddo
    if (do_dump) then
      call system_clock(d0)
      call dump_state(it, stat(1))
      call system_clock(d1)
      dump_ticks = dump_ticks + (d1 - d0)
    endif

    call switch_stat(stat(0), stat(1))
  enddo
  call system_clock(c1)

  write(*,'(a,i0,a,es22.15,a,i0,a)') '{"ref_steps": ', nsteps_drv, ', "ref_seconds": ', &
       dble(c1 - c0 - dump_ticks) / dble(crate), ', "dof": ', DOF, '}'

contains

  subroutine dump_state(step, s)
    integer         , intent(in) :: step
    type(stat_field), intent(in) :: s
    character(len=600) :: fname
    integer :: uu
    call calc_total_mass  (total_mass  , s)
    call calc_total_energy(total_energy, s)
    write(fname,'(a,a,i7.7,a)') trim(out_dir), '/state_', step, '.bin'
    open(newunit=uu, file=trim(fname), access='stream', form='unformatted', status='replace')
    write(uu) s%phi, s%phiG, s%u, s%v, s%contraU, s%contraV, s%zonal_wind, s%meridional_wind
    write(uu) total_mass, total_energy
    close(uu)
    ! cell averages of a history step (calc_VIA above)
    if (history_every > 0 .and. step > 0) then
      if (mod(step, history_every) == 0) then
        write(fname,'(a,a,i7.7,a)') trim(out_dir), '/via_', step, '.bin'
        open(newunit=uu, file=trim(fname), access='stream', form='unformatted', status='replace')
        write(uu) s%uC, s%vC, s%phiC
        close(uu)
      endif
    endif
  end subroutine dump_state

end program ref_driver
