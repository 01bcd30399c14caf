! q2d_fcase.f90 — a Fortran host that runs a Q2D case directory on the device explicit step.
!
! The call sequence of Q2D's driver around its time loop (main.f90:48 -> Wrap_Solve.f90:19-32, set-up as
! Wrap_Setup.f90:114-152, output as Wrap_Postprocess.f90:67-113), with Explicit_Step replaced by the
! library through module q2d_gpu: read input.ini, INPUT/bc.txt and INPUT/ic.tec (OUTPUT/field.tec on a
! restart), set the geometry up, hand the arrays to q2d_gpu_create with c_loc, step until the device's
! TODO test ends the run (iter-, res- or time-threshold), appending residual-history.dat every res-diter
! steps, and write OUTPUT/field.tec with its SOLUTIONTIME.
!
! usage: q2d_fcase <case directory> [--mixed] [--tables] [--max-steps N] [--dump <file>]
!   --mixed      the mixed arithmetic form (FP32 faces, FP64 state) instead of plain FP64
!   --tables     table-mode thermodynamics: the property planes of INPUT/thermo.dat (and transport.dat),
!                interpolated from node int(T), as FLINT's native tabulated mode (plain FP64 form only);
!                without it the closed forms of a constant-cp gas (a varying cp is refused)
!   --max-steps  stop after N steps even if no threshold is reached
!   --dump       write the problem the host built (case-file format of q2d_fdriver) and stop before the
!                device is touched
program q2d_fcase
  use iso_c_binding
  use q2d_gpu
  use q2d_case_io
  implicit none

  type(case_config) :: cfg
  type(block_data), allocatable, target :: blk(:)
  type(q2d_bc_entry), allocatable, target :: recs(:)
  type(q2d_block_desc), allocatable, target :: desc(:)
  type(q2d_case_desc), target :: cs(1)
  type(q2d_step_out) :: out(1)
  type(q2d_problem) :: prob
  type(c_ptr) :: h
  character(len=1024) :: dir, arg, dump
  real(c_double) :: R, cp, soltime, t0, t1
  integer :: nb, b, nrec, rc, ndev, k, u, iter_general, max_steps, nsteps, chunk
  logical :: mixed, tables
  integer(c_int64_t) :: c0, c1, crate
  real(c_double), allocatable, target :: tab(:)
  integer :: ntab

  call get_command_argument(1, dir)
  if (len_trim(dir) == 0) call fail_stop('usage: q2d_fcase <case directory> [--mixed] [--max-steps N] [--dump file]')
  mixed = .false.; tables = .false.; max_steps = huge(1); dump = ''
  k = 2
  do while (k <= command_argument_count())
    call get_command_argument(k, arg)
    select case (trim(arg))
    case ('--mixed'); mixed = .true.
    case ('--tables'); tables = .true.
    case ('--max-steps'); k = k + 1; call get_command_argument(k, arg); read(arg, *) max_steps
    case ('--dump'); k = k + 1; call get_command_argument(k, dump)
    case default; call fail_stop('unknown option '//trim(arg))
    end select
    k = k + 1
  end do

  ! ---- configuration, gas, initial state, boundary records, geometry ----
  call read_input_ini(trim(dir)//'/input.ini', cfg)
  if (lower(trim(cfg%equations)) /= 'euler') &
    call fail_stop('equations = '//trim(cfg%equations)//': this host runs the Euler equations (transport tables are not read)')
  if (trim(cfg%source) /= 'none' .and. trim(cfg%source) /= 'axisymmetric') &
    call fail_stop('[Q2D-Sources] active = '//trim(cfg%source)//' is not read by this host')
  call read_gas(trim(dir), R, cp, tables)
  ntab = 0
  if (tables) call read_tables(trim(dir), tab, ntab)
  if (cfg%newrun) then
    call read_tec_field(trim(dir)//'/INPUT/ic.tec', blk, soltime)
    soltime = 0d0
  else
    call read_tec_field(trim(dir)//'/OUTPUT/field.tec', blk, soltime)   ! restart (IO_Solution.f90:58-96)
  end if
  nb = size(blk)
  nrec = 0
  do b = 1, nb
    nrec = nrec + 2 * blk(b)%ni + 2 * blk(b)%nj                        ! Allocate_BC (IO_BC.f90:57-71)
    call fill_initial_ghosts(blk(b))
  end do
  call read_bc_file(trim(dir)//'/INPUT/bc.txt', nrec, recs)
  call setup_geometry(blk, recs, trim(cfg%source) == 'axisymmetric')

  ! ---- the case descriptor (Assign_Setup.f90:43-62) ----
  cs(1)%scheme%n_rk = cfg%n_rk
  cs(1)%scheme%limiter = limiter_id(cfg)
  cs(1)%scheme%riemann = riemann_id(cfg%riemann)
  cs(1)%scheme%time_accurate = merge(1, 0, cfg%time_accurate)
  cs(1)%scheme%model = 0
  cs(1)%scheme%source = merge(Q2D_SRC_AXI, Q2D_SRC_NONE, trim(cfg%source) == 'axisymmetric')
  cs(1)%scheme%precision = merge(Q2D_PREC_MIXED, Q2D_PREC_FP64, mixed)
  cs(1)%scheme%ramp_iter = cfg%ramp
  cs(1)%scheme%cfl = cfg%cfl; cs(1)%scheme%vnn = cfg%vnn
  cs(1)%scheme%gas_R = R; cs(1)%scheme%gas_cp = cp
  cs(1)%scheme%mu_ref = 0d0; cs(1)%scheme%T_ref = 1d0; cs(1)%scheme%S_suth = 0d0; cs(1)%scheme%Pr = 0.72d0
  cs(1)%scheme%irs_beta = cfg%irs_beta
  cs(1)%scheme%space_sd = merge(1, 0, trim(cfg%space) == 'MUSCL-SD')
  cs(1)%scheme%integration_vars = merge(1, 0, trim(cfg%intvars) == 'prim')
  cs(1)%itermax = cfg%iter_threshold; cs(1)%reserved = 0
  cs(1)%res_threshold = cfg%res_threshold; cs(1)%time_threshold = cfg%time_threshold
  cs(1)%res_drop = 0d0
  cs(1)%thermo = 0; cs(1)%thermo_n = ntab; cs(1)%reserved2 = 0
  if (tables) cs(1)%thermo = transfer(c_loc(tab), cs(1)%thermo)
  ! domain%time at set-up (Wrap_Setup.f90:129-152)
  iter_general = 0
  if (cfg%time_accurate) then
    cs(1)%time0 = merge(0d0, soltime, cfg%newrun)
  else
    cs(1)%time0 = -1d0
    if (.not. cfg%newrun) iter_general = int(soltime)
  end if

  allocate(desc(nb))
  do b = 1, nb
    desc(b)%ni = blk(b)%ni; desc(b)%nj = blk(b)%nj; desc(b)%case_id = 0; desc(b)%reserved = 0
    desc(b)%P = c_loc(blk(b)%P); desc(b)%dl = c_loc(blk(b)%dl); desc(b)%M = c_loc(blk(b)%M)
    desc(b)%sur = c_loc(blk(b)%sur); desc(b)%face_i = c_loc(blk(b)%fi); desc(b)%face_j = c_loc(blk(b)%fj)
    desc(b)%node = c_loc(blk(b)%node)
    desc(b)%src_a = c_null_ptr; desc(b)%src_b = c_null_ptr
    if (allocated(blk(b)%sa)) desc(b)%src_a = c_loc(blk(b)%sa)
    if (allocated(blk(b)%sb)) desc(b)%src_b = c_loc(blk(b)%sb)
  end do
  prob%nblocks = nb; prob%nbc = nrec; prob%ncases = 1; prob%reserved = 0
  prob%blocks = c_loc(desc); prob%bc = c_loc(recs); prob%cases = c_loc(cs)

  if (len_trim(dump) > 0) then
    call write_dump(trim(dump))
    print '(a,i0,a,i0,a)', 'q2d_fcase: problem of ', nb, ' block(s) and ', nrec, ' records written to '//trim(dump)
  else
    call run_on_device()
  end if

contains

  ! ---- device solver and time loop ----
  subroutine run_on_device()
  rc = q2d_device_count(ndev)
  if (rc /= Q2D_OK .or. ndev < 1) then
    print '(a)', 'q2d_fcase: no device: '//q2d_error_message()
    error stop 3
  end if
  rc = q2d_gpu_create(prob, h)
  if (rc /= Q2D_OK) then
    print '(a,i0,a)', 'q2d_fcase: create failed (', rc, '): '//q2d_error_message()
    error stop 4
  end if
  call execute_command_line('mkdir -p '//trim(dir)//'/OUTPUT')
  open(newunit=u, file=trim(dir)//'/OUTPUT/residual-history.dat', status='unknown', position='append', action='write')
  chunk = max(1, cfg%res_diter)
  nsteps = 0
  call system_clock(c0, crate)
  do
    k = min(chunk, max_steps - nsteps)
    ! stop at the next multiple of ini-diter, where the reference re-reads input.ini
    k = min(k, max(1, cfg%ini_diter) - mod(nsteps, max(1, cfg%ini_diter)))
    if (k == 1) then
      rc = q2d_gpu_step(h, out)
    else
      rc = q2d_gpu_run(h, int(k, c_int32_t), out)
    end if
    if (rc /= Q2D_OK) then
      ! the reference's stops (Lib_Newstate.f90:327-331, Lib_Convective.f90:49-54)
      print '(a,3(1x,i0))', 'Integration failed at b, i, j:', out(1)%fail_b, out(1)%fail_i, out(1)%fail_j
      print '(a)', 'q2d_fcase: '//q2d_error_message()
      error stop 5
    end if
    nsteps = out(1)%iter
    if (mod(nsteps, max(1, cfg%res_diter)) == 0 .or. out(1)%todo == 3 .or. nsteps >= max_steps) &
      write(u, '(I8,E20.10,4E20.10)') iter_general + out(1)%iter, out(1)%time, out(1)%residual
    if (out(1)%todo == 3 .or. nsteps >= max_steps) exit
    if (mod(nsteps, max(1, cfg%ini_diter)) == 0) call reread_runtime()   ! Wrap_Postprocess.f90:288-291
  end do
  call system_clock(c1)
  close(u)
  do b = 1, nb
    rc = q2d_gpu_download_state(h, int(b - 1, c_int32_t), blk(b)%P)
    if (rc /= Q2D_OK) error stop 6
  end do
  rc = q2d_gpu_destroy(h)
  if (cfg%time_accurate) then
    t1 = out(1)%time
  else
    t1 = real(iter_general + out(1)%iter, c_double)     ! steady: SOLUTIONTIME carries the iteration
  end if
  call write_field(trim(dir)//'/OUTPUT/field.tec', blk, t1, R, cp)
  t0 = real(c1 - c0, c_double) / real(crate, c_double)
  print '(a,i0,a,es14.6,a,f10.3,a)', 'q2d_fcase: ', out(1)%iter, ' steps, time ', out(1)%time, ', ', t0, ' s'
  end subroutine run_on_device

  ! Read_Inifile_Runtime (Read_Ini.f90:65-88): input.ini again, its runtime keys to the device
  subroutine reread_runtime()
    type(case_config) :: c2
    type(q2d_runtime) :: rt
    call read_input_ini(trim(dir)//'/input.ini', c2)
    rc = q2d_gpu_get_runtime(h, 0_c_int32_t, rt)
    if (rc /= Q2D_OK) error stop 7
    rt%cfl = c2%cfl; rt%vnn = c2%vnn; rt%ramp_iter = c2%ramp; rt%itermax = c2%iter_threshold
    rt%res_threshold = c2%res_threshold; rt%time_threshold = c2%time_threshold
    rt%limiter = limiter_id(c2); rt%riemann = riemann_id(c2%riemann)
    rt%space_sd = merge(1, 0, trim(c2%space) == 'MUSCL-SD')
    rc = q2d_gpu_set_runtime(h, 0_c_int32_t, rt)
    if (rc /= Q2D_OK) then
      print '(a,i0,a)', 'q2d_fcase: runtime keys refused (', rc, '): '//q2d_error_message()
      error stop 7
    end if
    cfg%res_diter = c2%res_diter; cfg%ini_diter = c2%ini_diter
  end subroutine reread_runtime

  integer function limiter_id(c)
    type(case_config), intent(in) :: c
    if (trim(c%space) == 'first-order') then
      limiter_id = Q2D_LIM_IORD
      return
    end if
    select case (trim(c%limiter))                        ! Mod_Space.f90:44-78; unknown -> IORD
    case ('minmod'); limiter_id = Q2D_LIM_MINMOD
    case ('vanleer'); limiter_id = Q2D_LIM_VANLEER
    case ('vanalbada'); limiter_id = Q2D_LIM_VANALBADA
    case ('mc'); limiter_id = Q2D_LIM_MC
    case ('superbee'); limiter_id = Q2D_LIM_SUPERBEE
    case ('LIMO3'); limiter_id = Q2D_LIM_LIMO3
    case default; limiter_id = Q2D_LIM_IORD
    end select
  end function limiter_id

  integer function riemann_id(name)
    character(len=*), intent(in) :: name
    select case (trim(name))                             ! Mod_Riemann.f90:25-109
    case ('HLLC'); riemann_id = Q2D_RIEMANN_HLLC
    case ('HLLE'); riemann_id = Q2D_RIEMANN_HLLE
    case ('HLLEM'); riemann_id = Q2D_RIEMANN_HLLEM
    case ('HLLE++'); riemann_id = Q2D_RIEMANN_HLLEPP
    case ('HLLC+'); riemann_id = Q2D_RIEMANN_HLLCP
    case default
      riemann_id = -1
      call fail_stop('riemann-solver = '//trim(name)//' is not on this path')
    end select
  end function riemann_id

  ! The problem in q2d_fdriver's case-file layout (q2d_amd/fortran_io.py): header with the struct sizes,
  ! the case and record structs' bytes, then each block's arrays.
  subroutine write_dump(path)
    character(len=*), intent(in) :: path
    integer :: v, q, flags
    type(q2d_case_desc) :: cw(1)
    cw = cs
    cw(1)%thermo = 0                                   ! the tables follow the blocks
    open(newunit=v, file=path, access='stream', form='unformatted', status='replace')
    write(v) int([1362248754, nb, nrec, 1, 0], c_int32_t), int(c_sizeof(cs(1)), c_int32_t), &
             int(c_sizeof(recs(1)), c_int32_t), int(c_sizeof(out(1)), c_int32_t)
    write(v) cw
    write(v) recs
    do q = 1, nb
      flags = 1
      if (allocated(blk(q)%sa)) flags = flags + 2
      if (allocated(blk(q)%sb)) flags = flags + 4
      write(v) int([blk(q)%ni, blk(q)%nj, 0, flags], c_int32_t)
      write(v) blk(q)%P, blk(q)%dl, blk(q)%M, blk(q)%sur, blk(q)%fi, blk(q)%fj, blk(q)%node
      if (allocated(blk(q)%sa)) write(v) blk(q)%sa
      if (allocated(blk(q)%sb)) write(v) blk(q)%sb
    end do
    if (tables) write(v) tab
    close(v)
  end subroutine write_dump

end program q2d_fcase
