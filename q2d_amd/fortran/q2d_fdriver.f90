! q2d_fdriver.f90 — a Fortran host of the device explicit step, through module q2d_gpu only.
!
! Reads a problem in the reference's allocation layout from a stream file, hands the arrays to the
! library with c_loc (no repacking), runs nsteps explicit steps, and writes every block's state and
! the per-case step results.  This is the call sequence Q2D's driver would make in place of its
! Explicit_Step loop (Wrap_Solve.f90:19-32).
!
! usage: q2d_fdriver <case file> <output file>
!   case file (native byte order, stream access):
!     int32  magic (= 1362248754), nblocks, nbc, ncases, nsteps, and the byte sizes of the C structs
!            q2d_case_desc, q2d_bc_entry and q2d_step_out the writer used (checked against this host's
!            bind(C) types before anything else is read)
!     ncases x q2d_case_desc, nbc x q2d_bc_entry           (the C structs' bytes)
!     per block: int32 ni, nj, case_id, flags (1 node, 2 src_a, 4 src_b), then
!       P(4,ni+4,nj+4), dl(2,ni+4,nj+4), M(4,ni+4,nj+4), sur(ni+4,nj+4),
!       face_i(3,ni+1,nj), face_j(3,ni,nj+1) [, node(2,ni+1,nj+1)] [, src_a(ni,nj)] [, src_b(ni,nj)]
!     per case with thermo_n > 0 (table mode): its property planes, 4 x thermo_n doubles (the file's
!       thermo pointers are null; this host points them at its own copies)
!   output: per block P(4,ni+4,nj+4); then ncases x q2d_step_out
program q2d_fdriver
  use iso_c_binding
  use q2d_gpu
  implicit none

  type :: block_arrays
    integer :: ni, nj, case_id, flags
    real(c_double), allocatable :: P(:,:,:), dl(:,:,:), M(:,:,:), sur(:,:), fi(:,:,:), fj(:,:,:)
    real(c_double), allocatable :: node(:,:,:), sa(:,:), sb(:,:)
  end type
  type :: case_tables
    real(c_double), allocatable :: tab(:)
  end type
  type(case_tables), allocatable, target :: tabs(:)
  integer :: c

  integer(c_int32_t) :: hdr(8)
  type(block_arrays), allocatable, target :: blk(:)
  type(q2d_block_desc), allocatable, target :: desc(:)
  type(q2d_bc_entry), allocatable, target :: bc(:)
  type(q2d_case_desc), allocatable, target :: cs(:)
  type(q2d_step_out), allocatable :: out(:)
  type(q2d_problem) :: prob
  type(c_ptr) :: h
  character(len=1024) :: fin, fout
  integer :: u, b, rc, ndev, dims(4)
  type(q2d_case_desc) :: cs1
  type(q2d_bc_entry) :: bc1
  type(q2d_step_out) :: out1

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  if (hdr(1) /= 1362248754) then
    print '(a)', 'q2d_fdriver: not a case file of this layout version (magic mismatch)'
    error stop 2
  end if
  if ((hdr(4) > 0 .and. hdr(6) /= c_sizeof(cs1)) .or. hdr(7) /= c_sizeof(bc1) .or. hdr(8) /= c_sizeof(out1)) then
    print '(a,3(1x,i0),a,3(1x,i0))', 'q2d_fdriver: descriptor layout mismatch: file has struct sizes', hdr(6:8), &
      ', this host', c_sizeof(cs1), c_sizeof(bc1), c_sizeof(out1)
    error stop 2
  end if
  allocate(blk(hdr(2)), desc(hdr(2)), bc(max(1, hdr(3))), cs(hdr(4)), out(hdr(4)))
  read(u) cs
  if (hdr(3) > 0) read(u) bc(1:hdr(3))
  do b = 1, hdr(2)
    read(u) dims
    associate (x => blk(b))
      x%ni = dims(1); x%nj = dims(2); x%case_id = dims(3); x%flags = dims(4)
      allocate(x%P(4, x%ni+4, x%nj+4), x%dl(2, x%ni+4, x%nj+4), x%M(4, x%ni+4, x%nj+4), x%sur(x%ni+4, x%nj+4))
      allocate(x%fi(3, x%ni+1, x%nj), x%fj(3, x%ni, x%nj+1))
      read(u) x%P, x%dl, x%M, x%sur, x%fi, x%fj
      if (iand(x%flags, 1) /= 0) then
        allocate(x%node(2, x%ni+1, x%nj+1)); read(u) x%node
      end if
      if (iand(x%flags, 2) /= 0) then
        allocate(x%sa(x%ni, x%nj)); read(u) x%sa
      end if
      if (iand(x%flags, 4) /= 0) then
        allocate(x%sb(x%ni, x%nj)); read(u) x%sb
      end if
    end associate
  end do
  allocate(tabs(hdr(4)))
  do c = 1, hdr(4)
    cs(c)%thermo = 0
    if (cs(c)%thermo_n > 0) then
      allocate(tabs(c)%tab(4 * cs(c)%thermo_n))
      read(u) tabs(c)%tab
      cs(c)%thermo = transfer(c_loc(tabs(c)%tab), cs(c)%thermo)
    end if
  end do
  close(u)

  do b = 1, hdr(2)
    desc(b)%ni = blk(b)%ni; desc(b)%nj = blk(b)%nj; desc(b)%case_id = blk(b)%case_id; desc(b)%reserved = 0
    desc(b)%P = c_loc(blk(b)%P); desc(b)%dl = c_loc(blk(b)%dl); desc(b)%M = c_loc(blk(b)%M)
    desc(b)%sur = c_loc(blk(b)%sur); desc(b)%face_i = c_loc(blk(b)%fi); desc(b)%face_j = c_loc(blk(b)%fj)
    desc(b)%node = c_null_ptr; desc(b)%src_a = c_null_ptr; desc(b)%src_b = c_null_ptr
    if (allocated(blk(b)%node)) desc(b)%node = c_loc(blk(b)%node)
    if (allocated(blk(b)%sa)) desc(b)%src_a = c_loc(blk(b)%sa)
    if (allocated(blk(b)%sb)) desc(b)%src_b = c_loc(blk(b)%sb)
  end do
  prob%nblocks = hdr(2); prob%nbc = hdr(3); prob%ncases = hdr(4); prob%reserved = 0
  prob%blocks = c_loc(desc); prob%bc = c_loc(bc); prob%cases = c_loc(cs)

  rc = q2d_device_count(ndev)
  if (rc /= Q2D_OK .or. ndev < 1) then
    print '(a)', 'q2d_fdriver: no device: '//q2d_error_message()
    error stop 3
  end if
  rc = q2d_gpu_create(prob, h)
  if (rc /= Q2D_OK) then
    print '(a,i0,a)', 'q2d_fdriver: create failed (', rc, '): '//q2d_error_message()
    error stop 4
  end if
  rc = q2d_gpu_run(h, hdr(5), out)
  if (rc /= Q2D_OK) then
    print '(a,i0,a)', 'q2d_fdriver: run failed (', rc, '): '//q2d_error_message()
    error stop 5
  end if
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  do b = 1, hdr(2)
    rc = q2d_gpu_download_state(h, int(b - 1, c_int32_t), blk(b)%P)
    if (rc /= Q2D_OK) error stop 6
    write(u) blk(b)%P
  end do
  write(u) out
  close(u)
  rc = q2d_gpu_destroy(h)
  print '(a,i0,a,i0)', 'q2d_fdriver: ', hdr(5), ' steps, iter ', out(1)%iter
end program q2d_fdriver
