! Host-side setup of an MCV shallow-water solve, as the host program owns it:
! cubed-sphere geometry (mesh_mod.F90:57-214, projection_mod.F90:8-322),
! ghost-interpolation tables (ghost_mod.f90:21-79), initial states
! (test_case_mod.F90: Williamson cases 2, 5, 6; a balanced jet with seeded
! spherical-harmonic perturbations) and the state handoff of
! test_case_mod.F90:30-39.
!
! Everything is exported with bind(C) so the Fortran host program and the
! Python package (ctypes) call the same code.  Arrays are the reference
! layout, flattened: scalar fields (nh, nh, 6), 2x2 matrices (2, 2, nh, nh, 6)
! with nh = n*(dof-1) + 1 + 2*(dof-1).
!
! The geometry keeps the reference's expressions and kinds (REAL(16) where the
! reference's -fdefault-real-8 build promotes dble()/d0 literals), so with the
! same compiler it reproduces the reference's arrays bit for bit.
module mcv_host_lib
  use iso_c_binding
  implicit none
  private

  integer, parameter :: dp = c_double
  integer, parameter :: qp = selected_real_kind(33)
  real(dp), parameter :: pi = 2.0_dp * asin(1.0_dp)

  public :: mcv_host_mesh, mcv_host_ghost, mcv_host_case, mcv_host_jet, mcv_host_handoff, &
            mcv_host_cells

  ! State of L'Ecuyer's MRG32k3a combined multiple-recursive generator
  ! (Operations Research 47(1), 1999), kept in integer(8) with no overflow.
  type mrg_state
    integer(8) :: s1(3), s2(3)
  end type mrg_state

contains

  ! ---------------------------------------------------------------------------
  ! Geometry
  ! ---------------------------------------------------------------------------

  ! Local angles (x, y) on panel p -> longitude, latitude (projection_mod.F90:8-38)
  subroutine panel_to_sphere(x, y, p, lon, lat)
    real(dp), intent(in)  :: x, y
    integer,  intent(in)  :: p
    real(dp), intent(out) :: lon, lat
    real(dp) :: a, b
    select case (p)
    case (1:4)
      lon = real(real(x, qp) + real(p - 1, qp) * real(pi, qp) / 2.0_qp, dp)
      lat = atan2(tan(y) * cos(x), 1.0_dp)
    case (5)
      a = tan(x)
      b = tan(y)
      lon = atan2(a, -b)
      lat = atan2(1.0_dp, sqrt(a*a + b*b))
    case default
      a = tan(x)
      b = tan(y)
      lon = atan2(a, b)
      lat = -atan2(1.0_dp, sqrt(a*a + b*b))
    end select
  end subroutine panel_to_sphere

  ! Metric tensor G, its inverse IG and the Jacobian sqrt(G) at (x, y)
  ! (projection_mod.F90:150-211).  M = (m11, m21, m12, m22).
  subroutine metric(x, y, radius, G, IG, jab)
    real(dp), intent(in)  :: x, y, radius
    real(dp), intent(out) :: G(4), IG(4), jab
    real(dp) :: rho, s
    rho = sqrt(1. + tan(x)**2. + tan(y)**2.)
    s = radius**2 / (rho**4. * cos(x)**2. * cos(y)**2.)
    G(1) = s * (1. + tan(x)**2.)
    G(2) = s * (-tan(x) * tan(y))
    G(3) = G(2)
    G(4) = s * (1. + tan(y)**2.)
    s = (rho**2. * cos(x)**2. * cos(y)**2.) / radius**2
    IG(1) = s * (1. + tan(y)**2.)
    IG(2) = s * (tan(x) * tan(y))
    IG(3) = IG(2)
    IG(4) = s * (1. + tan(x)**2.)
    rho = sqrt(1 + tan(x)**2 + tan(y)**2)
    jab = 1. / (cos(x)**2 * cos(y)**2 * rho**3)
    jab = jab * radius**2
  end subroutine metric

  ! Sphere <-> panel vector bases A (contravariant -> sphere) and IA
  ! (projection_mod.F90:213-322).  M = (m11, m21, m12, m22).
  subroutine bases(lon, lat, p, radius, A, IA)
    real(dp), intent(in)  :: lon, lat, radius
    integer,  intent(in)  :: p
    real(dp), intent(out) :: A(4), IA(4)
    real(dp) :: al, sa, ca, st, ct, t
    if (p <= 4) then
      al = real(real(lon, qp) - real(p - 1, qp) * real(pi, qp) / 2.0_qp, dp)
    else
      al = lon
    endif
    sa = sin(al)
    ca = cos(al)
    st = sin(lat)
    ct = cos(lat)
    if (p <= 4) then
      t = ct*ct*ca*ca + st*st
      IA(1) = 1./ct
      IA(2) = sa*st/t
      IA(3) = 0.
      IA(4) = ca/t
      A(1) = ct
      A(2) = -st*ct*sa/ca
      A(3) = 0.
      A(4) = ca*ct*ct + st*st/ca
    else if (p == 5) then
      t = st + sa*sa*ct*ct/st
      IA(1) = ca/t
      IA(3) = -sa/st/t
      t = st + ca*ca*ct*ct/st
      IA(2) = sa/t
      IA(4) = ca/st/t
      t = 1. + sa*sa*ct*ct/st/st
      A(1) = ca*st*t
      A(2) = -st*st*sa*t
      t = 1. + ca*ca*ct*ct/st/st
      A(3) = sa*st*t
      A(4) = ca*st*st*t
    else
      t = st + sa*sa*ct*ct/st
      IA(1) = -ca/t
      IA(3) = sa/st/t
      t = st + ca*ca*ct*ct/st
      IA(2) = sa/t
      IA(4) = ca/st/t
      t = 1. + sa*sa*ct*ct/st/st
      A(1) = -ca*st*t
      A(2) = st*st*sa*t
      t = 1. + ca*ca*ct*ct/st/st
      A(3) = sa*st*t
      A(4) = ca*st*st*t
    endif
    IA = IA / radius
    A = A * radius
  end subroutine bases

  ! Point geometry of the whole mesh (mesh_mod.F90:132-163).
  integer(c_int) function mcv_host_mesh(dof, n, radius, omega, xP, yP, lonP, latP, sqrtG, &
                                        G, IG, A, IA, f) bind(C, name='mcv_host_mesh')
    integer(c_int), value :: dof, n
    real(c_double), value :: radius, omega
    real(c_double), intent(out) :: xP(*), yP(*), lonP(*), latP(*), sqrtG(*), f(*)
    real(c_double), intent(out) :: G(4, *), IG(4, *), A(4, *), IA(4, *)
    integer :: H, npv, nh, p, i, j, k
    real(dp) :: d2r, dx, xmin
    mcv_host_mesh = -1
    if ((dof /= 3 .and. dof /= 4) .or. n < 2) return
    H = dof - 1
    npv = n * H + 1
    nh = npv + 2 * H
    d2r = pi / 180.
    dx = (90.0_dp / real(n, dp)) * d2r
    xmin = -45.0_dp * d2r
    k = 0
    do p = 1, 6
      do j = 1 - H, npv + H
        do i = 1 - H, npv + H
          k = k + 1
          xP(k) = (i - 1) * dx / (dof - 1) + xmin
          yP(k) = (j - 1) * dx / (dof - 1) + xmin
          call panel_to_sphere(xP(k), yP(k), p, lonP(k), latP(k))
          call metric(xP(k), yP(k), radius, G(:, k), IG(:, k), sqrtG(k))
          call bases(lonP(k), latP(k), p, radius, A(:, k), IA(:, k))
        enddo
      enddo
    enddo
    call coriolis(k, omega, latP, f)
    mcv_host_mesh = 0
  end function mcv_host_mesh

  ! Cell geometry of the history file (mesh_mod.F90:165-214): cell-centre
  ! longitude and latitude of cells 0..n+1 (one halo cell), centre angles
  ! halfway between the cell's first and last point angles; and the cell
  ! areas, zero on the halo ring.  An equiangular cell is a spherical
  ! quadrilateral whose area is its spherical excess, the sum of its four
  ! interior angles minus 2 pi; at the grid vertex (a, b) the grid lines meet
  ! at acos(-sin a sin b), and the interior angles at the cell's (+1, 0) and
  ! (0, +1) vertices are the supplements.  The vertex angles are formed in
  ! REAL(16) and rounded once, as the reference's dble() is under
  ! -fdefault-real-8.  Arrays (n+2, n+2, 6).
  integer(c_int) function mcv_host_cells(dof, n, radius, lonC, latC, areaCell) &
      bind(C, name='mcv_host_cells')
    integer(c_int), value :: dof, n
    real(c_double), value :: radius
    real(c_double), intent(out) :: lonC(0:n+1, 0:n+1, 6), latC(0:n+1, 0:n+1, 6)
    real(c_double), intent(out) :: areaCell(0:n+1, 0:n+1, 6)
    integer :: H, p, ic, jc, k, k1, k2
    real(dp) :: d2r, dx, xmin, xs, xe, ys, ye, xc, yc
    real(dp), allocatable :: gp(:), ang(:, :), da(:, :)
    mcv_host_cells = -1
    if ((dof /= 3 .and. dof /= 4) .or. n < 2) return
    H = dof - 1
    d2r = pi / 180.
    dx = (90.0_dp / real(n, dp)) * d2r
    xmin = -45.0_dp * d2r
    do p = 1, 6
      do jc = 0, n + 1
        do ic = 0, n + 1
          xs = ((ic - 1) * H) * dx / (dof - 1) + xmin
          xe = (ic * H) * dx / (dof - 1) + xmin
          ys = ((jc - 1) * H) * dx / (dof - 1) + xmin
          ye = (jc * H) * dx / (dof - 1) + xmin
          xc = 0.5 * (xs + xe)
          yc = 0.5 * (ys + ye)
          call panel_to_sphere(xc, yc, p, lonC(ic, jc, p), latC(ic, jc, p))
        enddo
      enddo
    enddo
    allocate(gp(n + 1), ang(n + 1, n + 1), da(n, n))
    do k = 1, n + 1
      gp(k) = real(real(-0.25_dp * pi, qp) + (real(pi, qp) / real(2 * n, qp)) * real(k - 1, qp), dp)
    enddo
    do k2 = 1, n + 1
      do k1 = 1, n + 1
        ang(k1, k2) = acos(-sin(gp(k1)) * sin(gp(k2)))
      enddo
    enddo
    do k2 = 1, n
      do k1 = 1, n
        da(k1, k2) = -2.0_dp * pi + ang(k1, k2) + (pi - ang(k1 + 1, k2)) + &
                     (pi - ang(k1, k2 + 1)) + ang(k1 + 1, k2 + 1)
      enddo
    enddo
    da = da * (radius * radius)
    areaCell = 0.0_dp
    do p = 1, 6
      areaCell(1:n, 1:n, p) = da
    enddo
    mcv_host_cells = 0
  end function mcv_host_cells

  ! f = 2 Omega sin(lat) (mesh_mod.F90:160), in a pass of its own.
  subroutine coriolis(np, omega, lat, f)
    integer,  intent(in)  :: np
    real(dp), intent(in)  :: omega, lat(np)
    real(dp), intent(out) :: f(np)
    integer :: k
    do k = 1, np
      f(k) = 2. * omega * sin(lat(k))
    enddo
  end subroutine coriolis

  ! Ghost-point tables (ghost_mod.f90:21-79): panel 1's north halo rows
  ! projected onto panel 5; the same table serves every panel edge.
  integer(c_int) function mcv_host_ghost(dof, n, coef, iref) bind(C, name='mcv_host_ghost')
    integer(c_int), value :: dof, n
    real(c_double), intent(out) :: coef(n*(dof-1)+1, dof-1)
    integer(c_int), intent(out) :: iref(n*(dof-1)+1, dof-1)
    integer :: H, npv, k, s, r
    real(dp) :: d2r, dx, xmin, x, y, lon, lat, X5
    real(dp), allocatable :: xraw(:), gx(:, :), gy(:, :)
    mcv_host_ghost = -1
    if ((dof /= 3 .and. dof /= 4) .or. n < 2) return
    H = dof - 1
    npv = n * H + 1
    d2r = pi / 180.
    dx = (90.0_dp / real(n, dp)) * d2r
    xmin = -45.0_dp * d2r
    allocate(xraw(n + 1), gx(npv, H), gy(npv, H))
    do k = 1, H
      do s = 1, npv
        x = (s - 1) * dx / (dof - 1) + xmin
        y = (npv + k - 1) * dx / (dof - 1) + xmin
        call panel_to_sphere(x, y, 1, lon, lat)
        call sphere_to_panel5(lon, lat, gx(s, k), gy(s, k))
      enddo
    enddo
    do r = 1, n + 1
      xraw(r) = ((r - 1) * H) * dx / (dof - 1) + xmin
    enddo
    do k = 1, H
      r = 1
      do s = 1, npv
        X5 = gx(s, k)
        do while (r <= n + 1)
          if (.not. (X5 > xraw(r))) exit
          r = r + 1
        enddo
        if (r < 2 .or. r > n + 1) then
          mcv_host_ghost = -3
          return
        endif
        if (.not. (X5 >= xraw(r-1) .and. X5 <= xraw(r))) then
          mcv_host_ghost = -3   ! ghost_mod.f90:73-74 'Ghost cell location error'
          return
        endif
        coef(s, k) = X5 - xraw(r-1)
        iref(s, k) = r - 1
      enddo
    enddo
    mcv_host_ghost = 0
  end function mcv_host_ghost

  ! Panel-5 local angles of a sphere point (projection_mod.F90:52-54)
  subroutine sphere_to_panel5(lon, lat, x, y)
    real(dp), intent(in)  :: lon, lat
    real(dp), intent(out) :: x, y
    integer, parameter :: p = 5
    x = atan((-1.)**(p+1) * sin(lon) / tan(lat))
    y = atan(-cos(lon) / tan(lat))
  end subroutine sphere_to_panel5

  ! ---------------------------------------------------------------------------
  ! Initial states
  ! ---------------------------------------------------------------------------

  ! Williamson test cases of test_case_mod.F90 (2: steady zonal flow, 5:
  ! isolated mountain, 6: Rossby-Haurwitz wave), on np points.
  integer(c_int) function mcv_host_case(case_num, np, radius, omega, gravity, lonP, latP, &
                                        phi, zonal, merid, phi_s) bind(C, name='mcv_host_case')
    integer(c_int), value :: case_num, np
    real(c_double), value :: radius, omega, gravity
    real(c_double), intent(in)  :: lonP(np), latP(np)
    real(c_double), intent(out) :: phi(np), zonal(np), merid(np), phi_s(np)
    mcv_host_case = 0
    select case (case_num)
    case (2)
      call case2(np, radius, omega, latP, phi, zonal, merid, phi_s)
    case (5)
      call case5(np, radius, omega, gravity, lonP, latP, phi, zonal, merid, phi_s)
    case (6)
      call case6(np, radius, omega, gravity, lonP, latP, phi, zonal, merid, phi_s)
    case default
      mcv_host_case = -1
    end select
  end function mcv_host_case

  ! test_case_mod.F90:52-81
  subroutine case2(np, radius, omega, lat, phi, u, v, hs)
    integer,  intent(in)  :: np
    real(dp), intent(in)  :: radius, omega, lat(np)
    real(dp), intent(out) :: phi(np), u(np), v(np), hs(np)
    real(dp) :: u0, gh0
    integer :: k
    gh0 = 29400.
    u0 = 2. * pi * radius / (12. * 86400.)
    do k = 1, np
      phi(k) = gh0 - (radius * omega * u0 + u0**2 / 2.) * sin(lat(k))**2
      u(k) = u0 * cos(lat(k))
    enddo
    v = 0.
    hs = 0.
  end subroutine case2

  ! test_case_mod.F90:84-130 (its u0**2/2.d0 is REAL(16) in the reference build)
  subroutine case5(np, radius, omega, gravity, lonP, lat, phi, u, v, hs)
    integer,  intent(in)  :: np
    real(dp), intent(in)  :: radius, omega, gravity, lonP(np), lat(np)
    real(dp), intent(out) :: phi(np), u(np), v(np), hs(np)
    real(dp), parameter :: hs0 = 2000., u0 = 20., alpha = 0.
    real(dp) :: gh0, rr, lc, tc, lon, r, s
    real(qp) :: c
    integer :: k
    gh0 = 5960. * gravity
    rr = pi / 9.
    lc = 1.5 * pi
    tc = pi / 6.
    do k = 1, np
      lon = lonP(k)
      if (lon < 0) lon = 2. * pi + lon
      r = sqrt(min(rr**2, (lon - lc)**2 + (lat(k) - tc)**2))
      hs(k) = gravity * hs0 * (1. - r / rr)
      u(k) = u0 * (cos(lat(k)) * cos(alpha) + cos(lon) * sin(lat(k)) * sin(alpha))
      v(k) = -u0 * sin(lon) * sin(alpha)
      c = real(radius * omega * u0, qp) + real(u0**2, qp) / 2.0_qp
      s = -cos(lon) * cos(lat(k)) * sin(alpha) + sin(lat(k)) * cos(alpha)
      phi(k) = real(real(gh0, qp) - c * real(s**2, qp) - real(hs(k), qp), dp)
    enddo
  end subroutine case5

  ! Rossby-Haurwitz wave of wavenumber 4 (Williamson et al. 1992, J. Comput.
  ! Phys. 102, test case 6, eqs. 143-146), the state of test_case_mod.F90:133-191.
  ! With c = cos(lat), s = sin(lat), a the radius, w = K = 7.848e-6 1/s, R = 4:
  !   u   = a w c + a K c^3 (R s^2 - c^2) cos(R lon)
  !   v   = -a K R c^3 s sin(R lon)
  !   g h = g h0 + a^2 (A + B cos(R lon) + C cos(2 R lon))
  !   A   = w/2 (2 Omega + w) c^2 + K^2/4 c^(2R) ((R+1) c^2 + 2R^2 - R - 2 - 2R^2/c^2)
  !   B   = 2 (Omega + w) K / ((R+1)(R+2)) c^R (R^2 + 2R + 2 - (R+1)^2 c^2)
  !   C   = K^2/4 c^(2R) ((R+1) c^2 - (R+2))
  ! Powers of c are formed by repeated squaring; the c^(2R) c^(-2) term of A is
  ! written as c^(2R-2).  The reference evaluates the same functions in another
  ! order, so the two states agree to a few ulps, not bit for bit.
  subroutine case6(np, radius, omega, gravity, lonP, lat, phi, u, v, hs)
    integer,  intent(in)  :: np
    real(dp), intent(in)  :: radius, omega, gravity, lonP(np), lat(np)
    real(dp), intent(out) :: phi(np), u(np), v(np), hs(np)
    real(dp), parameter :: w = 7.848e-6_dp, h0 = 8000._dp
    integer,  parameter :: nw = 4
    real(dp), parameter :: rw = real(nw, dp)
    real(dp) :: c, s, c2, c3, c4, c6, c8, cr, sr, c2r, ga, gb, gc, kk
    integer :: k
    kk = 0.25_dp * w * w
    do k = 1, np
      c = cos(lat(k))
      s = sin(lat(k))
      c2 = c * c
      c3 = c2 * c
      c4 = c2 * c2
      c6 = c4 * c2
      c8 = c4 * c4
      cr = cos(rw * lonP(k))
      sr = sin(rw * lonP(k))
      u(k) = radius * w * (c + c3 * (rw * s * s - c2) * cr)
      v(k) = -radius * w * rw * c3 * s * sr
      c2r = cos(2._dp * rw * lonP(k))
      ga = 0.5_dp * w * (2._dp * omega + w) * c2 + &
           kk * (c8 * ((rw + 1._dp) * c2 + (2._dp * rw * rw - rw - 2._dp)) - 2._dp * rw * rw * c6)
      gb = 2._dp * (omega + w) * w / ((rw + 1._dp) * (rw + 2._dp)) * c4 * &
           ((rw * rw + 2._dp * rw + 2._dp) - (rw + 1._dp)**2 * c2)
      gc = kk * c8 * ((rw + 1._dp) * c2 - (rw + 2._dp))
      phi(k) = gravity * h0 + radius * radius * (ga + gb * cr + gc * c2r)
    enddo
    hs = 0.
  end subroutine case6

  subroutine mrg_seed(st, seed)
    type(mrg_state), intent(out) :: st
    integer(8), intent(in) :: seed
    integer(8), parameter :: m1 = 4294967087_8, m2 = 4294944443_8
    integer :: k
    real(dp) :: u
    st%s1 = [modulo(12345_8 + seed, m1), modulo(23456_8 + 7_8 * seed, m1), &
             modulo(34567_8 + 13_8 * seed, m1)]
    st%s2 = [modulo(45678_8 + 17_8 * seed, m2), modulo(56789_8 + 19_8 * seed, m2), &
             modulo(67890_8 + 23_8 * seed, m2)]
    do k = 1, 32
      u = mrg_next(st)
    enddo
  end subroutine mrg_seed

  real(dp) function mrg_next(st)
    type(mrg_state), intent(inout) :: st
    integer(8), parameter :: m1 = 4294967087_8, m2 = 4294944443_8
    integer(8), parameter :: a12 = 1403580_8, a13n = 810728_8, a21 = 527612_8, a23n = 1370589_8
    real(dp), parameter :: norm = 2.328306549295727688e-10_dp
    integer(8) :: p1, p2
    p1 = modulo(a12 * st%s1(2) - a13n * st%s1(1), m1)
    st%s1 = [st%s1(2), st%s1(3), p1]
    p2 = modulo(a21 * st%s2(3) - a23n * st%s2(1), m2)
    st%s2 = [st%s2(2), st%s2(3), p2]
    if (p1 > p2) then
      mrg_next = real(p1 - p2, dp) * norm
    else
      mrg_next = real(p1 - p2 + m1, dp) * norm
    endif
  end function mrg_next

  ! Balanced zonal jet (the isolated-mountain flow of test_case_mod.F90:116-120
  ! without the mountain) plus a seeded height perturbation
  !   h' = A * sum_{l=1..lmax, m=-l..l} c_lm Y_lm,  c_lm ~ N(0, 1),
  ! Y_lm the real orthonormal spherical harmonics, c_lm drawn in the order
  ! (l = 1..lmax, m = -l..l) from MRG32k3a through Box-Muller, A chosen so
  ! that the sphere-mean rms of h' is rms_h exactly (sum c^2 / (4 pi) = 1/A^2 rms^2),
  ! and the matching wind from the streamfunction psi = g h' F(lat) with
  ! F = sin(lat) / (2 Omega (sin^2(lat) + sin^2(10 deg))): geostrophic away
  ! from the equator, smooth and non-divergent across it.
  integer(c_int) function mcv_host_jet(np, radius, omega, gravity, u0, h0, lmax, seed, rms_h, &
                                       lonP, latP, phi, zonal, merid, phi_s) &
                                       bind(C, name='mcv_host_jet')
    integer(c_int), value :: np, lmax
    integer(c_int64_t), value :: seed
    real(c_double), value :: radius, omega, gravity, u0, h0, rms_h
    real(c_double), intent(in)  :: lonP(np), latP(np)
    real(c_double), intent(out) :: phi(np), zonal(np), merid(np), phi_s(np)
    real(dp), allocatable :: c(:, :), p(:, :), q(:, :), dp_(:, :)
    type(mrg_state) :: st
    real(dp) :: u1, u2, r, ss, amp, s0, x, cl, lon, hp, hl, ht, F, dF, sl2, trig, dtrig
    real(dp) :: a_lm, b_lm, z(2)
    integer :: l, m, k, iz, ma
    mcv_host_jet = -1
    if (lmax < 1 .or. lmax > 200 .or. np < 1) return
    allocate(c(0:lmax, -lmax:lmax), p(0:lmax, 0:lmax), q(0:lmax, 0:lmax), dp_(0:lmax, 0:lmax))
    c = 0.
    call mrg_seed(st, int(seed, 8))
    ss = 0.
    iz = 3
    do l = 1, lmax
      do m = -l, l
        if (iz > 2) then
          u1 = mrg_next(st)
          u2 = mrg_next(st)
          r = sqrt(-2.0_dp * log(u1))
          z = [r * cos(2.0_dp * pi * u2), r * sin(2.0_dp * pi * u2)]
          iz = 1
        endif
        c(l, m) = z(iz)
        iz = iz + 1
        ss = ss + c(l, m)**2
      enddo
    enddo
    amp = 0.
    if (ss > 0.) amp = rms_h / sqrt(ss / (4.0_dp * pi))
    s0 = sin(10.0_dp * pi / 180.0_dp)
    do k = 1, np
      lon = lonP(k)
      x = sin(latP(k))
      cl = cos(latP(k))
      call legendre(lmax, x, cl, p, q, dp_)
      hp = 0.
      hl = 0.   ! (1/cos lat) d h'/d lon
      ht = 0.   ! d h'/d lat
      do l = 1, lmax
        do m = -l, l
          ma = abs(m)
          if (m == 0) then
            trig = 1.
            dtrig = 0.
          else if (m > 0) then
            trig = sqrt(2.0_dp) * cos(ma * lon)
            dtrig = -sqrt(2.0_dp) * ma * sin(ma * lon)
          else
            trig = sqrt(2.0_dp) * sin(ma * lon)
            dtrig = sqrt(2.0_dp) * ma * cos(ma * lon)
          endif
          hp = hp + c(l, m) * p(l, ma) * trig
          if (m /= 0) hl = hl + c(l, m) * q(l, ma) * dtrig
          ht = ht + c(l, m) * dp_(l, ma) * trig
        enddo
      enddo
      hp = amp * hp
      hl = amp * hl
      ht = amp * ht
      sl2 = x * x + s0 * s0
      F = x / (2.0_dp * omega * sl2)
      dF = cl * (s0 * s0 - x * x) / (2.0_dp * omega * sl2 * sl2)
      phi(k) = gravity * h0 - (radius * omega * u0 + u0**2 / 2.0_dp) * x**2 + gravity * hp
      zonal(k) = u0 * cl - gravity / radius * (ht * F + hp * dF)
      merid(k) = gravity / radius * F * hl
    enddo
    phi_s = 0.
    mcv_host_jet = 0
  end function mcv_host_jet

  ! Fully normalised associated Legendre functions p(l, m) = N_lm P_l^m(x),
  ! N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!), x = sin(lat), no Condon-Shortley
  ! phase; q(l, m) = p(l, m)/cos(lat) for m >= 1; dp(l, m) = d p(l, m)/d lat.
  subroutine legendre(lmax, x, cl, p, q, dpl)
    integer,  intent(in)  :: lmax
    real(dp), intent(in)  :: x, cl
    real(dp), intent(out) :: p(0:lmax, 0:lmax), q(0:lmax, 0:lmax), dpl(0:lmax, 0:lmax)
    integer :: l, m
    real(dp) :: a, b
    p = 0.
    q = 0.
    dpl = 0.
    p(0, 0) = sqrt(1.0_dp / (4.0_dp * pi))
    do m = 1, lmax
      if (m == 1) then
        q(1, 1) = p(0, 0) * sqrt(3.0_dp / 2.0_dp)
      else
        q(m, m) = q(m-1, m-1) * sqrt((2.0_dp * m + 1.0_dp) / (2.0_dp * m)) * cl
      endif
      p(m, m) = q(m, m) * cl
    enddo
    do m = 0, lmax
      if (m + 1 <= lmax) then
        p(m+1, m) = x * sqrt(2.0_dp * m + 3.0_dp) * p(m, m)
        if (m >= 1) q(m+1, m) = x * sqrt(2.0_dp * m + 3.0_dp) * q(m, m)
      endif
      do l = m + 2, lmax
        a = sqrt((4.0_dp * l * l - 1.0_dp) / (real(l, dp)**2 - real(m, dp)**2))
        b = sqrt((real(l - 1, dp)**2 - real(m, dp)**2) / (4.0_dp * real(l - 1, dp)**2 - 1.0_dp))
        p(l, m) = a * (x * p(l-1, m) - b * p(l-2, m))
        if (m >= 1) q(l, m) = a * (x * q(l-1, m) - b * q(l-2, m))
      enddo
    enddo
    do l = 1, lmax
      dpl(l, 0) = sqrt(real(l, dp) * real(l + 1, dp)) * p(l, 1)
      do m = 1, l
        dpl(l, m) = -l * x * q(l, m)
        if (l - 1 >= m) dpl(l, m) = dpl(l, m) + &
            sqrt((2.0_dp * l + 1.0_dp) * real(l - m, dp) * real(l + m, dp) / (2.0_dp * l - 1.0_dp)) * q(l-1, m)
      enddo
    enddo
  end subroutine legendre

  ! test_case_mod.F90:30-39: phiG = phi*sqrtG, covariant winds (G*IA) and
  ! contravariant winds IA applied to the sphere winds, on np points.
  integer(c_int) function mcv_host_handoff(np, sqrtG, G, IA, phi, zonal, merid, phiG, u, v, &
                                           contraU, contraV) bind(C, name='mcv_host_handoff')
    integer(c_int), value :: np
    real(c_double), intent(in)  :: sqrtG(np), G(2, 2, np), IA(2, 2, np)
    real(c_double), intent(in)  :: phi(np), zonal(np), merid(np)
    real(c_double), intent(out) :: phiG(np), u(np), v(np), contraU(np), contraV(np)
    real(dp) :: M(2, 2)
    integer :: k
    do k = 1, np
      phiG(k) = phi(k) * sqrtG(k)
      M = matmul(G(:, :, k), IA(:, :, k))
      u(k) = M(1, 1) * zonal(k) + M(1, 2) * merid(k)
      v(k) = M(2, 1) * zonal(k) + M(2, 2) * merid(k)
      contraU(k) = IA(1, 1, k) * zonal(k) + IA(1, 2, k) * merid(k)
      contraV(k) = IA(2, 1, k) * zonal(k) + IA(2, 2, k) * merid(k)
    enddo
    mcv_host_handoff = 0
  end function mcv_host_handoff

end module mcv_host_lib
