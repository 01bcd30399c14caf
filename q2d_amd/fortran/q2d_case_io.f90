! q2d_case_io.f90 — the case-directory contract of the Q2D host, for a Fortran host of the device step.
!
! Q2D reads a case from its directory and writes its results there; its own readers and writers sit in
! libraries that are not part of its tree (FiNeR for input.ini, ORION for Tecplot files, FLINT for the gas
! tables).  This module restates the part of that contract a run of the explicit step needs, so that the
! device solver runs a reference case directory unchanged:
!   input.ini         [Q2D-Parameters] [Q2D-Physics] [Q2D-Numerics] [Q2D-Sources] [Q2D-IO] keys
!                     (Register_Sim_Param.f90:26-29, Register_Physics.f90:24, Register_Numerics.f90:31-75,
!                     Register_Sources.f90:33), bound as Assign_Setup.f90:43-62 binds them
!   INPUT/phase.txt   species and molecular weight; INPUT/thermo.dat the species' table (T, cp, h, s, dcp)
!   INPUT/bc.txt      5-integer records b i j f type, and the property line of the codes that have one
!                     (IO_BC.f90:103-141 and :163-420)
!   INPUT/ic.tec      Tecplot ASCII BLOCK zones: nodal x, y and cell-centred rho1, u, v, p (IO_Solution.f90:38-129);
!                     a restart (newrun = false) reads OUTPUT/field.tec instead, and its SOLUTIONTIME
!   OUTPUT/field.tec  nodal x, y and cell-centred rho(1), u, v, p, T, g, R with SOLUTIONTIME
!                     (IO_Solution.f90:234-326)
!   OUTPUT/residual-history.dat   '(I8,E20.10,4E20.10)' per written step: iteration, time, residuals
!                     (Wrap_Setup.f90:114-127, Wrap_Postprocess.f90:67-73)
! and computes the set-up geometry from the nodes (Setup_Metrics, Mod_Metrics.f90:11-148: Lib_Metrics.f90:
! 41-189 for the cells and faces, :285-728 for the ghost cells of the boundary records; the axisymmetric
! radii of Lib_Source_Axisymmetry.f90:19-37).  Arithmetic keeps the reference's operation order.
!
! Gas: one species of a table with constant cp (h = cp T), the tables the reference's cases use; the gas
! constant is R_u / W with R_u = 8314.51.  Other tables, Navier-Stokes transport tables, area-variation
! files and multi-species cases are reported as unsupported.
module q2d_case_io
  use iso_c_binding
  use q2d_gpu
  implicit none
  private

  integer, parameter, public :: GC = 2
  real(c_double), parameter, public :: R_UNIVERSAL = 8314.51d0

  type, public :: case_config
    real(c_double) :: time_threshold = 1d30, res_threshold = 1d-10, cfl = 0.5d0, vnn = 0.3d0, irs_beta = 0d0
    integer :: iter_threshold = 1000000000, ramp = 0, n_rk = 1, res_diter = 1, ini_diter = 10000
    logical :: time_accurate = .false., newrun = .true.
    character(len=64) :: equations = '', space = '', limiter = '', riemann = 'HLLC', intvars = 'cons'
    character(len=64) :: source = 'none', time_scheme = 'euler'
  end type

  type, public :: block_data
    integer :: ni = 0, nj = 0
    real(c_double), allocatable :: node(:,:,:)       ! (2, 0:ni, 0:nj)
    real(c_double), allocatable :: P(:,:,:)          ! (4, 1-gc:ni+gc, 1-gc:nj+gc)
    real(c_double), allocatable :: dl(:,:,:), M(:,:,:), sur(:,:)
    real(c_double), allocatable :: fi(:,:,:), fj(:,:,:)   ! (3, 0:ni, 1:nj), (3, 1:ni, 0:nj)
    real(c_double), allocatable :: sa(:,:), sb(:,:)  ! (1:ni, 1:nj) axis r_ave, 1/r
  end type

  public :: read_input_ini, read_gas, read_tables, read_bc_file, read_tec_field, setup_geometry, fill_initial_ghosts
  public :: write_field, lower, fail_stop

contains

  subroutine fail_stop(msg)
    character(len=*), intent(in) :: msg
    print '(a)', 'q2d_fcase: '//trim(msg)
    error stop 2
  end subroutine fail_stop

  pure function lower(s) result(t)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: t
    integer :: k
    t = s
    do k = 1, len(s)
      if (s(k:k) >= 'A' .and. s(k:k) <= 'Z') t(k:k) = achar(iachar(s(k:k)) + 32)
    end do
  end function lower

  logical function parse_logical(v)
    character(len=*), intent(in) :: v
    character(len=len(v)) :: w
    w = lower(adjustl(v))
    parse_logical = (w(1:1) == 't' .or. w(1:2) == '.t')
  end function parse_logical

  ! input.ini: '[section]' lines, 'key = value' lines, ';' or '#' comments.
  subroutine read_input_ini(path, cfg)
    character(len=*), intent(in) :: path
    type(case_config), intent(inout) :: cfg
    integer :: u, ios, eq, c
    character(len=512) :: line, sec, key, val
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call fail_stop('cannot open '//trim(path))
    sec = ''
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      c = scan(line, ';#')
      if (c > 0) line = line(1:c-1)
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '[') then
        sec = line(2:index(line, ']')-1)
        cycle
      end if
      eq = index(line, '=')
      if (eq == 0) cycle
      key = adjustl(line(1:eq-1))
      val = adjustl(line(eq+1:))
      select case (trim(sec))
      case ('Q2D-Parameters')
        select case (trim(key))
        case ('time-threshold'); read(val, *) cfg%time_threshold
        case ('res-threshold');  read(val, *) cfg%res_threshold
        case ('iter-threshold'); read(val, *) cfg%iter_threshold
        case ('newrun');         cfg%newrun = parse_logical(val)
        end select
      case ('Q2D-Physics')
        if (trim(key) == 'equations') cfg%equations = val
      case ('Q2D-Numerics')
        select case (trim(key))
        case ('cfl');                   read(val, *) cfg%cfl
        case ('vnn');                   read(val, *) cfg%vnn
        case ('cfl-rise-threshold');    read(val, *) cfg%ramp
        case ('time-accurate');         cfg%time_accurate = parse_logical(val)
        case ('time-scheme');           cfg%time_scheme = val
        case ('integration-variables'); cfg%intvars = val
        case ('irs-beta');              read(val, *) cfg%irs_beta
        case ('space-reconstruction');  cfg%space = val
        case ('flux-limiter');          cfg%limiter = val
        case ('riemann-solver');        cfg%riemann = val
        end select
      case ('Q2D-Sources')
        if (trim(key) == 'active') cfg%source = val
      case ('Q2D-IO')
        if (trim(key) == 'res-diter') read(val, *) cfg%res_diter
        if (trim(key) == 'ini-diter') read(val, *) cfg%ini_diter
      end select
    end do
    close(u)
    ! time-scheme: 'euler' or RKn (Assign_Setup.f90:55-59)
    if (trim(cfg%time_scheme) == 'euler') then
      cfg%n_rk = 1
    else
      read(cfg%time_scheme(3:3), *, iostat=ios) cfg%n_rk
      if (ios /= 0) call fail_stop('unknown time-scheme '//trim(cfg%time_scheme))
    end if
  end subroutine read_input_ini

  ! INPUT/phase.txt (one species) and INPUT/thermo.dat; returns R and cp of a constant-cp table.  With
  ! tables = .true. a varying cp is accepted (table mode reads the table itself, read_tables).
  subroutine read_gas(dir, R, cp, tables)
    character(len=*), intent(in) :: dir
    real(c_double), intent(out) :: R, cp
    logical, intent(in) :: tables
    integer :: u, ios, n
    character(len=256) :: line, name
    real(c_double) :: W, T, c, h, cp0
    open(newunit=u, file=trim(dir)//'/INPUT/phase.txt', status='old', action='read', iostat=ios)
    if (ios /= 0) call fail_stop('cannot open INPUT/phase.txt')
    read(u, '(a)') line
    read(u, *, iostat=ios) name, W
    if (ios /= 0) call fail_stop('INPUT/phase.txt: no species line')
    read(u, '(a)', iostat=ios) line
    if (ios == 0 .and. len_trim(line) > 0) call fail_stop('INPUT/phase.txt: one species only on this path')
    close(u)
    open(newunit=u, file=trim(dir)//'/INPUT/thermo.dat', status='old', action='read', iostat=ios)
    if (ios /= 0) call fail_stop('cannot open INPUT/thermo.dat')
    n = 0
    cp0 = 0d0
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (scan(line(1:1), '0123456789.-') == 0) cycle          ! TITLE / VARIABLES / ZONE / I= lines
      read(line, *, iostat=ios) T, c, h
      if (ios /= 0) cycle
      n = n + 1
      if (n == 1) cp0 = c
      if (.not. tables .and. (abs(c - cp0) > 1d-12 * abs(cp0) .or. abs(h - cp0 * T) > 1d-9 * abs(cp0 * T))) &
        call fail_stop('INPUT/thermo.dat: cp varies with T (the closed form needs a constant cp; run with --tables)')
    end do
    close(u)
    if (n == 0) call fail_stop('INPUT/thermo.dat: no table rows')
    R = R_UNIVERSAL / W
    cp = cp0
  end subroutine read_gas

  ! The property planes of table mode (q2d_case_desc%thermo): h, cp, mu, k with node k at T = k K, k = 0 ..
  ! n-1, from the rows T = 1, 2, ... of INPUT/thermo.dat (T, Cp, Enthalpy, ...) and, where it exists,
  ! INPUT/transport.dat (T, Viscosity, Conductivity; the planes cover the shorter file); node 0 is
  ! extrapolated linearly from nodes 1 and 2.  Plane q occupies tab(q*n+1 : q*n+n).
  subroutine read_tables(dir, tab, n)
    character(len=*), intent(in) :: dir
    real(c_double), allocatable, intent(out) :: tab(:)
    integer, intent(out) :: n
    real(c_double), allocatable :: th(:,:), tr(:,:)
    integer :: nth, ntr, m, q
    logical :: have_tr
    call read_rows(dir//'/INPUT/thermo.dat', 3, th, nth)
    inquire(file=dir//'/INPUT/transport.dat', exist=have_tr)
    m = nth
    if (have_tr) then
      call read_rows(dir//'/INPUT/transport.dat', 3, tr, ntr)
      m = min(nth, ntr)
      if (any(tr(1, 1:m) /= th(1, 1:m))) call fail_stop('thermo.dat and transport.dat list different temperatures')
    end if
    if (m < 2) call fail_stop('INPUT/thermo.dat: fewer than two table rows')
    do q = 1, m
      if (th(1, q) /= real(q, c_double)) call fail_stop('property tables must list T = 1, 2, 3, ... K')
    end do
    n = m + 1
    allocate(tab(4 * n))
    tab = 0d0
    tab(2:n) = th(3, 1:m)                          ! h
    tab(n + 2:2 * n) = th(2, 1:m)                  ! cp
    if (have_tr) then
      tab(2 * n + 2:3 * n) = tr(2, 1:m)            ! mu
      tab(3 * n + 2:4 * n) = tr(3, 1:m)            ! k
    end if
    do q = 0, 3
      tab(q * n + 1) = 2d0 * tab(q * n + 2) - tab(q * n + 3)
    end do
  end subroutine read_tables

  ! the numeric rows of a Tecplot POINT table: the first ncol values of each, as columns of v(ncol, nrow)
  subroutine read_rows(path, ncol, v, nrow)
    character(len=*), intent(in) :: path
    integer, intent(in) :: ncol
    real(c_double), allocatable, intent(out) :: v(:,:)
    integer, intent(out) :: nrow
    integer :: u, ios, pass
    character(len=512) :: line
    real(c_double) :: x(8)
    do pass = 1, 2
      open(newunit=u, file=path, status='old', action='read', iostat=ios)
      if (ios /= 0) call fail_stop('cannot open '//path)
      nrow = 0
      do
        read(u, '(a)', iostat=ios) line
        if (ios /= 0) exit
        line = adjustl(line)
        if (len_trim(line) == 0) cycle
        if (scan(line(1:1), '0123456789.-+') == 0) cycle
        read(line, *, iostat=ios) x(1:ncol)
        if (ios /= 0) cycle
        nrow = nrow + 1
        if (pass == 2) v(:, nrow) = x(1:ncol)
      end do
      close(u)
      if (pass == 1) allocate(v(ncol, max(1, nrow)))
    end do
  end subroutine read_rows

  ! a direction token of a property line: 'normal' -> huge (Compute_Inflow_Direction), else a number
  real(c_double) function dir_token(tok)
    character(len=*), intent(in) :: tok
    integer :: ios
    if (index(lower(tok), 'normal') > 0) then
      dir_token = huge(1d0)
    else
      read(tok, *, iostat=ios) dir_token
      if (ios /= 0) dir_token = 0d0
    end if
  end function dir_token

  ! INPUT/bc.txt (IO_BC.f90:163-420) into records ordered by block, then face (stable).
  subroutine read_bc_file(path, nrec, recs)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nrec
    type(q2d_bc_entry), allocatable, intent(out) :: recs(:)
    type(q2d_bc_entry), allocatable :: tmp(:)
    integer :: u, ios, k, h(5), n, b, f
    real(c_double) :: a(3)
    character(len=64) :: t1, t2
    character(len=512) :: line
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call fail_stop('cannot open '//trim(path))
    allocate(tmp(nrec))
    do k = 1, nrec
      read(u, *, iostat=ios) h
      if (ios /= 0) call fail_stop('INPUT/bc.txt: fewer records than boundary cells')
      tmp(k)%b = h(1); tmp(k)%i = h(2); tmp(k)%j = h(3); tmp(k)%f = h(4); tmp(k)%type = h(5)
      tmp(k)%bs = 0; tmp(k)%is = 0; tmp(k)%js = 0; tmp(k)%fs = 0
      tmp(k)%d11 = 1; tmp(k)%d12 = 0; tmp(k)%d21 = 0; tmp(k)%d22 = 1
      tmp(k)%case_id = 0; tmp(k)%reserved = 0
      tmp(k)%mach = 0d0; tmp(k)%T = 0d0; tmp(k)%p = 0d0; tmp(k)%rel_fac = 1d0; tmp(k)%alpha = 0d0
      select case (h(5))
      case (101, 201)
        read(u, *) tmp(k)%bs, tmp(k)%is, tmp(k)%js, tmp(k)%fs, tmp(k)%d11, tmp(k)%d12, tmp(k)%d21, tmp(k)%d22
      case (302)
        read(u, *) tmp(k)%T
      case (405)                                        ! mach, T, p, alpha, beta, rel_fac, massf...
        read(u, '(a)') line
        read(line, *) a(1), a(2), a(3), t1, t2, tmp(k)%rel_fac
        tmp(k)%mach = a(1); tmp(k)%T = a(2); tmp(k)%p = a(3); tmp(k)%alpha = dir_token(t1)
      case (406)                                        ! p_back, rel_fac
        read(u, *) tmp(k)%p, tmp(k)%rel_fac
      case (200, 300, 400)
      case (103, 301, 303:309, 401:404, 407, 420, 501, 502)
        read(u, '(a)') line                             ! a code of one property line that this path rejects
      case default
        call fail_stop('INPUT/bc.txt: unsupported record code')
      end select
    end do
    close(u)
    ! stable order by (block, face), the order BC_Fluxes visits the records (Mod_BC_Fluxes.f90:79-86)
    allocate(recs(nrec))
    n = 0
    do b = 1, maxval(tmp%b)
      do f = 1, 4
        do k = 1, nrec
          if (tmp(k)%b == b .and. tmp(k)%f == f) then
            n = n + 1
            recs(n) = tmp(k)
          end if
        end do
      end do
    end do
    if (n /= nrec) call fail_stop('INPUT/bc.txt: record block or face out of range')
  end subroutine read_bc_file

  ! Tecplot ASCII BLOCK file: zones of nodal x, y and cell-centred variables; the first four cell
  ! variables are taken as rho, u, v, p.  Allocates the blocks' nodes and interior state.
  subroutine read_tec_field(path, blk, soltime)
    character(len=*), intent(in) :: path
    type(block_data), allocatable, intent(inout) :: blk(:)
    real(c_double), intent(out) :: soltime
    integer :: u, ios, nz, z, ni1, nj1, nv, k, i, j, p
    character(len=4096) :: vline, zline
    real(c_double), allocatable :: xs(:), ys(:), cv(:,:)
    type(block_data), allocatable :: tmp(:)
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call fail_stop('cannot open '//trim(path))
    read(u, '(a)') vline
    nv = 0
    do k = 1, len_trim(vline)
      if (vline(k:k) == '"') nv = nv + 1
    end do
    nv = nv / 2 - 2                                     ! cell-centred variables after x, y
    if (nv < 4) call fail_stop(trim(path)//': fewer than four cell variables')
    if (nv > 4 .and. index(vline, '"rho2"') > 0) call fail_stop(trim(path)//': one species only on this path')
    soltime = 0d0
    nz = 0
    allocate(tmp(0))
    do
      read(u, '(a)', iostat=ios) zline
      if (ios /= 0) exit
      if (index(zline, 'ZONE') == 0) cycle
      ni1 = zone_int(zline, 'I='); nj1 = zone_int(zline, 'J=')
      p = index(zline, 'SOLUTIONTIME=')
      if (p > 0) read(zline(p+13:), *, iostat=ios) soltime
      nz = nz + 1
      call grow(tmp, nz)
      tmp(nz)%ni = ni1 - 1; tmp(nz)%nj = nj1 - 1
      allocate(xs(ni1 * nj1), ys(ni1 * nj1), cv((ni1 - 1) * (nj1 - 1), nv))
      read(u, *) xs
      read(u, *) ys
      read(u, *) cv
      allocate(tmp(nz)%node(2, 0:ni1-1, 0:nj1-1))
      do j = 0, nj1 - 1
        do i = 0, ni1 - 1
          tmp(nz)%node(1, i, j) = xs(j * ni1 + i + 1)
          tmp(nz)%node(2, i, j) = ys(j * ni1 + i + 1)
        end do
      end do
      allocate(tmp(nz)%P(4, 1-GC:ni1-1+GC, 1-GC:nj1-1+GC))
      tmp(nz)%P = 0d0
      do j = 1, nj1 - 1
        do i = 1, ni1 - 1
          do k = 1, 4
            tmp(nz)%P(k, i, j) = cv((j - 1) * (ni1 - 1) + i, k)
          end do
        end do
      end do
      deallocate(xs, ys, cv)
    end do
    close(u)
    if (nz == 0) call fail_stop(trim(path)//': no zone')
    call move_alloc(tmp, blk)
  contains
    integer function zone_int(s, tag)
      character(len=*), intent(in) :: s, tag
      integer :: q, e
      q = index(s, tag)
      if (q == 0) call fail_stop(trim(path)//': zone without '//tag)
      e = q + len(tag)
      read(s(e:), *) zone_int
    end function zone_int
    subroutine grow(a, n)
      type(block_data), allocatable, intent(inout) :: a(:)
      integer, intent(in) :: n
      type(block_data), allocatable :: b(:)
      integer :: q
      allocate(b(n))
      do q = 1, size(a)
        call move_alloc(a(q)%node, b(q)%node)
        call move_alloc(a(q)%P, b(q)%P)
        b(q)%ni = a(q)%ni; b(q)%nj = a(q)%nj
      end do
      call move_alloc(b, a)
    end subroutine grow
  end subroutine read_tec_field

  ! Ghost layers of an initial state: the nearest interior cell (the first ghost fill overwrites the face
  ! ghosts; corners keep it).
  subroutine fill_initial_ghosts(b)
    type(block_data), intent(inout) :: b
    integer :: i, j
    do j = 1 - GC, b%nj + GC
      do i = 1 - GC, b%ni + GC
        b%P(:, i, j) = b%P(:, min(max(i, 1), b%ni), min(max(j, 1), b%nj))
      end do
    end do
  end subroutine fill_initial_ghosts

  ! Compute_Metric_Tensor — Lib_Metrics.f90:125-189 for one cell from its nodes (i-1,j-1), (i-1,j), (i,j-1), (i,j).
  subroutine metric_cell(b, i, j, N1, N2, N3, N4)
    type(block_data), intent(inout) :: b
    integer, intent(in) :: i, j
    real(c_double), intent(in) :: N1(2), N2(2), N3(2), N4(2)
    real(c_double) :: A1(2), A2(2), det, d1(2), d2(2)
    A1 = ((N4 - N2) + N3) - N1
    A2 = ((N4 - N3) + N2) - N1
    A1 = 0.5d0 * A1
    A2 = 0.5d0 * A2
    det = A1(1) * A2(2) - A1(2) * A2(1)
    if (det == 0d0) call fail_stop('Compute_Metric_Tensor: jacobian det=0')
    b%M(1, i, j) = A2(2) / det
    b%M(2, i, j) = (-A1(2)) / det
    b%M(3, i, j) = (-A2(1)) / det
    b%M(4, i, j) = A1(1) / det
    b%dl(1, i, j) = sqrt(A1(1) * A1(1) + A1(2) * A1(2))
    b%dl(2, i, j) = sqrt(A2(1) * A2(1) + A2(2) * A2(2))
    d1 = N4 - N1
    d2 = N3 - N2
    b%sur(i, j) = 0.5d0 * abs(d1(1) * d2(2) - d1(2) * d2(1))
    if (b%sur(i, j) <= 0d0) call fail_stop('Compute_Metric_Tensor: negative cell surface')
  end subroutine metric_cell

  ! Compute_Norm_Area — Lib_Metrics.f90:41-122: unit face normals with the global orientation sign, int_l.
  subroutine norm_area(b)
    type(block_data), intent(inout) :: b
    real(c_double) :: d1(2), d2(2), d3(2), sx, sy, signi, signj, A
    integer :: i, j
    d1 = b%node(:, 1, 1) - b%node(:, 1, 0)
    sx = d1(2); sy = -d1(1)
    d1 = 0.5d0 * (b%node(:, 1, 1) + b%node(:, 1, 0))
    d2 = 0.5d0 * (b%node(:, 0, 1) + b%node(:, 0, 0))
    d3 = d1 - d2
    signi = sign(1d0, d3(1) * sx + d3(2) * sy)
    d1 = b%node(:, 0, 1) - b%node(:, 1, 1)
    sx = d1(2); sy = -d1(1)
    d1 = 0.5d0 * (b%node(:, 1, 1) + b%node(:, 0, 1))
    d2 = 0.5d0 * (b%node(:, 1, 0) + b%node(:, 0, 0))
    d3 = d1 - d2
    signj = sign(1d0, d3(1) * sx + d3(2) * sy)
    do j = 1, b%nj
      do i = 0, b%ni
        d1 = b%node(:, i, j) - b%node(:, i, j - 1)
        d3(1) = d1(2); d3(2) = -d1(1)
        A = sqrt(d3(1) * d3(1) + d3(2) * d3(2))
        if (A == 0d0) then
          b%fi(1:2, i, j) = 0d0
        else
          b%fi(1, i, j) = d3(1) / A * signi
          b%fi(2, i, j) = d3(2) / A * signi
        end if
        b%fi(3, i, j) = A
      end do
    end do
    do j = 0, b%nj
      do i = 1, b%ni
        d1 = b%node(:, i - 1, j) - b%node(:, i, j)
        d3(1) = d1(2); d3(2) = -d1(1)
        A = sqrt(d3(1) * d3(1) + d3(2) * d3(2))
        if (A == 0d0) then
          b%fj(1:2, i, j) = 0d0
        else
          b%fj(1, i, j) = d3(1) / A * signj
          b%fj(2, i, j) = d3(2) / A * signj
        end if
        b%fj(3, i, j) = A
      end do
    end do
  end subroutine norm_area

  ! BC_Extrapolate_Metrics — Lib_Metrics.f90:285-395 (quadratic node extrapolation; linear across a block
  ! two cells high or less)
  subroutine extrapolate_metrics(b, Im, Jm, Fm)
    type(block_data), intent(inout) :: b
    integer, intent(in) :: Im, Jm, Fm
    real(c_double) :: Na(2, -2:2), Nb(2, -2:2)
    integer :: k
    select case (Fm)
    case (1)
      do k = 0, 2
        Na(:, k) = b%node(:, Im + k - 1, Jm - 1); Nb(:, k) = b%node(:, Im + k - 1, Jm)
      end do
      do k = -1, -2, -1
        Na(:, k) = 3d0 * Na(:, k + 1) - 3d0 * Na(:, k + 2) + Na(:, k + 3)
        Nb(:, k) = 3d0 * Nb(:, k + 1) - 3d0 * Nb(:, k + 2) + Nb(:, k + 3)
        call metric_cell(b, Im + k, Jm, Na(:, k), Nb(:, k), Na(:, k + 1), Nb(:, k + 1))
      end do
    case (2)
      do k = 0, -2, -1
        Na(:, k) = b%node(:, Im + k, Jm - 1); Nb(:, k) = b%node(:, Im + k, Jm)
      end do
      do k = 1, 2
        Na(:, k) = 3d0 * Na(:, k - 1) - 3d0 * Na(:, k - 2) + Na(:, k - 3)
        Nb(:, k) = 3d0 * Nb(:, k - 1) - 3d0 * Nb(:, k - 2) + Nb(:, k - 3)
        call metric_cell(b, Im + k, Jm, Na(:, k - 1), Nb(:, k - 1), Na(:, k), Nb(:, k))
      end do
    case (3)
      if (b%nj > 2) then
        do k = 0, 2
          Na(:, k) = b%node(:, Im - 1, Jm + k - 1); Nb(:, k) = b%node(:, Im, Jm + k - 1)
        end do
        do k = -1, -2, -1
          Na(:, k) = 3d0 * Na(:, k + 1) - 3d0 * Na(:, k + 2) + Na(:, k + 3)
          Nb(:, k) = 3d0 * Nb(:, k + 1) - 3d0 * Nb(:, k + 2) + Nb(:, k + 3)
          call metric_cell(b, Im, Jm + k, Na(:, k), Na(:, k + 1), Nb(:, k), Nb(:, k + 1))
        end do
      else
        do k = 0, 1
          Na(:, k) = b%node(:, Im - 1, Jm + k - 1); Nb(:, k) = b%node(:, Im, Jm + k - 1)
        end do
        do k = -1, -2, -1
          Na(:, k) = 2d0 * Na(:, k + 1) - Na(:, k + 2)
          Nb(:, k) = 2d0 * Nb(:, k + 1) - Nb(:, k + 2)
          call metric_cell(b, Im, Jm + k, Na(:, k), Na(:, k + 1), Nb(:, k), Nb(:, k + 1))
        end do
      end if
    case (4)
      if (b%nj > 2) then
        do k = 0, -2, -1
          Na(:, k) = b%node(:, Im - 1, Jm + k); Nb(:, k) = b%node(:, Im, Jm + k)
        end do
        do k = 1, 2
          Na(:, k) = 3d0 * Na(:, k - 1) - 3d0 * Na(:, k - 2) + Na(:, k - 3)
          Nb(:, k) = 3d0 * Nb(:, k - 1) - 3d0 * Nb(:, k - 2) + Nb(:, k - 3)
          call metric_cell(b, Im, Jm + k, Na(:, k - 1), Na(:, k), Nb(:, k - 1), Nb(:, k))
        end do
      else
        do k = 0, -1, -1
          Na(:, k) = b%node(:, Im - 1, Jm + k); Nb(:, k) = b%node(:, Im, Jm + k)
        end do
        do k = 1, 2
          Na(:, k) = 2d0 * Na(:, k - 1) - Na(:, k - 2)
          Nb(:, k) = 2d0 * Nb(:, k - 1) - Nb(:, k - 2)
          call metric_cell(b, Im, Jm + k, Na(:, k - 1), Na(:, k), Nb(:, k - 1), Nb(:, k))
        end do
      end if
    end select
  end subroutine extrapolate_metrics

  ! BC_Symmetry_Metrics — Lib_Metrics.f90:398-441: the ghost cells mirror the interior cells' metrics.
  subroutine symmetry_metrics(b, Im, Jm, Fm)
    type(block_data), intent(inout) :: b
    integer, intent(in) :: Im, Jm, Fm
    integer :: k, di, dj, si, sj
    do k = 1, 2
      select case (Fm)
      case (1); di = Im - k; dj = Jm; si = Im + k - 1; sj = Jm
      case (2); di = Im + k; dj = Jm; si = Im - k + 1; sj = Jm
      case (3); di = Im; dj = Jm - k; si = Im; sj = Jm + k - 1
      case default; di = Im; dj = Jm + k; si = Im; sj = Jm - k + 1
      end select
      b%M(:, di, dj) = b%M(:, si, sj)
      b%dl(:, di, dj) = b%dl(:, si, sj)
      b%sur(di, dj) = b%sur(si, sj)
    end do
  end subroutine symmetry_metrics

  ! BC_Connect_Metrics (101) / BC_Connect_Metrics_Periodic (201) — Lib_Metrics.f90:444-728: the ghost
  ! cells take the donor block's nodes (translated across a periodic seam).
  subroutine connect_metrics(bm, Im, Jm, Fm, bs, Is, Js, Fs, d11, d21, periodic)
    type(block_data), intent(inout) :: bm
    type(block_data), intent(in) :: bs
    integer, intent(in) :: Im, Jm, Fm, Is, Js, Fs, d11, d21
    logical, intent(in) :: periodic
    real(c_double) :: N(2, 4, 0:2), T(2), T2(2)
    integer :: g, ms, ds, II, JJ, III, JJJ
    integer, parameter :: g1(2, 4) = reshape([0, 0, -1, 0, 0, 0, 0, -1], [2, 4])
    integer, parameter :: g2(2, 4) = reshape([1, 0, -2, 0, 0, 1, 0, -2], [2, 4])
    N = 0d0
    select case (Fm)
    case (1); N(:, 1, 0) = bm%node(:, Im - 1, Jm - 1); N(:, 2, 0) = bm%node(:, Im - 1, Jm)
    case (2); N(:, 3, 0) = bm%node(:, Im, Jm - 1);     N(:, 4, 0) = bm%node(:, Im, Jm)
    case (3); N(:, 1, 0) = bm%node(:, Im - 1, Jm - 1); N(:, 3, 0) = bm%node(:, Im, Jm - 1)
    case default; N(:, 2, 0) = bm%node(:, Im - 1, Jm); N(:, 4, 0) = bm%node(:, Im, Jm)
    end select
    ms = merge(2, merge(4, merge(3, 4, Fm == 3), Fm == 2), Fm == 1)
    ds = merge(1, merge(3, merge(1, 2, Fm == 3), Fm == 2), Fm == 1)
    T = 0d0
    if (periodic) then
      call donor(0, II, JJ, III, JJJ)
      T = N(:, ms, 0) - bs%node(:, II, JJ)
      T2 = N(:, ds, 0) - bs%node(:, III, JJJ)
      if (maxval(abs(T - T2)) > 1d-10 * (1d0 + maxval(abs(T)))) &
        call fail_stop('BC_Connect_Metrics_Periodic: non-rigid periodic seam')
    end if
    do g = 1, 2
      call donor(g, II, JJ, III, JJJ)
      N(:, ms, g) = bs%node(:, II, JJ)
      N(:, ds, g) = bs%node(:, III, JJJ)
      if (periodic) then
        N(:, ms, g) = N(:, ms, g) + T
        N(:, ds, g) = N(:, ds, g) + T
      end if
      select case (Fm)
      case (1); call metric_cell(bm, Im - g, Jm, N(:, 1, g), N(:, 2, g), N(:, 1, g - 1), N(:, 2, g - 1))
      case (2); call metric_cell(bm, Im + g, Jm, N(:, 3, g - 1), N(:, 4, g - 1), N(:, 3, g), N(:, 4, g))
      case (3); call metric_cell(bm, Im, Jm - g, N(:, 1, g), N(:, 1, g - 1), N(:, 3, g), N(:, 3, g - 1))
      case default; call metric_cell(bm, Im, Jm + g, N(:, 2, g - 1), N(:, 2, g), N(:, 4, g - 1), N(:, 4, g))
      end select
    end do
  contains
    subroutine donor(g, II, JJ, III, JJJ)
      integer, intent(in) :: g
      integer, intent(out) :: II, JJ, III, JJJ
      integer :: Is1, Js1
      Is1 = Is + (GC - g) * g1(1, Fs) + (g - 1) * g2(1, Fs)
      Js1 = Js + (GC - g) * g1(2, Fs) + (g - 1) * g2(2, Fs)
      if (Fs <= 2) then
        II = Is1; JJ = floor_div(2 * Js1 + d11 + d21, 2)
        III = II; JJJ = JJ - d11
      else
        II = floor_div(2 * Is1 + d11 + d21, 2); JJ = Js1
        III = II - d11; JJJ = JJ
      end if
    end subroutine donor
    integer function floor_div(a, q)
      integer, intent(in) :: a, q
      floor_div = floor(real(a, c_double) / real(q, c_double))
    end function floor_div
  end subroutine connect_metrics

  ! Setup_Metrics (Mod_Metrics.f90:11-148): cells, faces, then the ghost cells of every record; the
  ! axisymmetric radii when the case has the axisymmetric source.
  subroutine setup_geometry(blk, recs, axisym)
    type(block_data), intent(inout) :: blk(:)
    type(q2d_bc_entry), intent(in) :: recs(:)
    logical, intent(in) :: axisym
    integer :: b, i, j, k
    real(c_double) :: r
    do b = 1, size(blk)
      associate (x => blk(b))
        allocate(x%dl(2, 1-GC:x%ni+GC, 1-GC:x%nj+GC), x%M(4, 1-GC:x%ni+GC, 1-GC:x%nj+GC))
        allocate(x%sur(1-GC:x%ni+GC, 1-GC:x%nj+GC), x%fi(3, 0:x%ni, 1:x%nj), x%fj(3, 1:x%ni, 0:x%nj))
        x%dl = 0d0; x%M = 0d0; x%sur = 0d0
        do j = 1, x%nj
          do i = 1, x%ni
            call metric_cell(x, i, j, x%node(:, i - 1, j - 1), x%node(:, i - 1, j), x%node(:, i, j - 1), x%node(:, i, j))
          end do
        end do
        call norm_area(x)
      end associate
    end do
    do k = 1, size(recs)
      select case (recs(k)%type)
      case (101, 201)
        call connect_metrics(blk(recs(k)%b), recs(k)%i, recs(k)%j, recs(k)%f, blk(recs(k)%bs), recs(k)%is, recs(k)%js, &
                             recs(k)%fs, recs(k)%d11, recs(k)%d21, recs(k)%type == 201)
      case (300)
        call symmetry_metrics(blk(recs(k)%b), recs(k)%i, recs(k)%j, recs(k)%f)
      case default
        call extrapolate_metrics(blk(recs(k)%b), recs(k)%i, recs(k)%j, recs(k)%f)
      end select
    end do
    if (axisym) then                                    ! Precompute_Axisymmetric_Geometry
      do b = 1, size(blk)
        associate (x => blk(b))
          allocate(x%sa(x%ni, x%nj), x%sb(x%ni, x%nj))
          do j = 1, x%nj
            do i = 1, x%ni
              r = 0.25d0 * (x%node(2, i, j) + x%node(2, i - 1, j) + x%node(2, i, j - 1) + x%node(2, i - 1, j - 1))
              x%sa(i, j) = r
              x%sb(i, j) = 1d0 / max(r, 1d-12)
            end do
          end do
        end associate
      end do
    end if
  end subroutine setup_geometry

  ! OUTPUT/field.tec: nodal x, y; cell-centred rho(1), u, v, p, T, g, R; SOLUTIONTIME.
  subroutine write_field(path, blk, soltime, R, cp)
    character(len=*), intent(in) :: path
    type(block_data), intent(in) :: blk(:)
    real(c_double), intent(in) :: soltime, R, cp
    integer :: u, b, i, j, k
    character(len=32) :: tbuf
    open(newunit=u, file=path, status='replace', action='write')
    write(u, '(a)') ' VARIABLES ="x" "y" "rho(1)" "u" "v" "p" "T" "g" "R"'
    write(tbuf, '(es23.15e3)') soltime
    do b = 1, size(blk)
      associate (x => blk(b))
        write(u, '(a,i0,a,i0,a,i0,a)') ' ZONE  T = Block', b, ', I=', x%ni + 1, ', J=', x%nj + 1, &
          ', K=1, DATAPACKING=BLOCK, VARLOCATION=([1-2]=NODAL,[3-9]=CELLCENTERED), SOLUTIONTIME='//trim(adjustl(tbuf))
        write(u, '(es24.15e3)') ((x%node(1, i, j), i = 0, x%ni), j = 0, x%nj)
        write(u, '(es24.15e3)') ((x%node(2, i, j), i = 0, x%ni), j = 0, x%nj)
        do k = 1, 4
          write(u, '(es24.15e3)') ((x%P(k, i, j), i = 1, x%ni), j = 1, x%nj)
        end do
        write(u, '(es24.15e3)') ((x%P(4, i, j) / (x%P(1, i, j) * R), i = 1, x%ni), j = 1, x%nj)
        write(u, '(es24.15e3)') ((cp / (cp - R), i = 1, x%ni), j = 1, x%nj)
        write(u, '(es24.15e3)') ((R, i = 1, x%ni), j = 1, x%nj)
      end associate
    end do
    close(u)
  end subroutine write_field

end module q2d_case_io
