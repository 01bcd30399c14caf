! q2d_gpu_mod.f90 — ISO_C_BINDING interface of the MI355X explicit step (include/q2d_gpu.h).
!
! A Fortran host (Q2D's driver, Wrap_Solve.f90:19-32) uses this module to hand its own allocations
! (Mod_Allocate_Data.f90:100-126) to the device solver through c_loc, without repacking, and to call
! one Explicit_Step (Mod_Explicit.f90:10-174) per q2d_gpu_step.  Every derived type below is
! interoperable (bind(C)) and mirrors the C struct of the same name field by field.
module q2d_gpu
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: Q2D_OK = 0, Q2D_E_ARG = -1, Q2D_E_HIP = -2, Q2D_E_NOMEM = -3, &
                                       Q2D_E_STATE = -4, Q2D_E_INTERFACE = -5, Q2D_E_UNSUPPORTED = -6, &
                                       Q2D_E_NODEVICE = -7
  integer(c_int), parameter, public :: Q2D_LIM_IORD = 0, Q2D_LIM_MINMOD = 1, Q2D_LIM_VANLEER = 2, &
                                       Q2D_LIM_VANALBADA = 3, Q2D_LIM_MC = 4, Q2D_LIM_SUPERBEE = 5, &
                                       Q2D_LIM_LIMO3 = 6
  integer(c_int), parameter, public :: Q2D_RIEMANN_HLLC = 0, Q2D_RIEMANN_HLLE = 1, Q2D_RIEMANN_HLLEM = 2, Q2D_RIEMANN_HLLEPP = 3, &
                                       Q2D_RIEMANN_HLLCP = 4
  integer(c_int), parameter, public :: Q2D_SRC_NONE = 0, Q2D_SRC_AXI = 1, Q2D_SRC_AREA = 2
  integer(c_int), parameter, public :: Q2D_PREC_FP64 = 0, Q2D_PREC_MIXED = 1

  type, bind(C), public :: q2d_scheme
    integer(c_int32_t) :: n_rk, limiter, riemann, time_accurate
    integer(c_int32_t) :: model, source, precision, ramp_iter
    real(c_double)     :: cfl, vnn
    real(c_double)     :: gas_R, gas_cp
    real(c_double)     :: mu_ref, T_ref, S_suth, Pr
    real(c_double)     :: irs_beta
    integer(c_int32_t) :: space_sd, integration_vars
  end type

  type, bind(C), public :: q2d_block_desc
    integer(c_int32_t) :: ni, nj, case_id, reserved
    type(c_ptr)        :: P, dl, M, sur, face_i, face_j, node, src_a, src_b
  end type

  type, bind(C), public :: q2d_bc_entry
    integer(c_int32_t) :: b, i, j, f, type
    integer(c_int32_t) :: bs, is, js, fs, d11, d12, d21, d22
    integer(c_int32_t) :: case_id
    integer(c_int32_t) :: reserved(2)
    real(c_double)     :: mach, T, p, rel_fac, alpha
  end type

  type, bind(C), public :: q2d_case_desc
    type(q2d_scheme)   :: scheme
    integer(c_int32_t) :: itermax, reserved
    real(c_double)     :: res_threshold, time_threshold
    real(c_double)     :: res_drop
    real(c_double)     :: time0
    integer(c_intptr_t) :: thermo             ! table mode: address of planes h, cp, mu, k (4 x thermo_n), else 0
                                              ! (an address integer, so that the type stays readable by
                                              ! unformatted I/O: transfer(c_loc(tab), 0_c_intptr_t))
    integer(c_int32_t) :: thermo_n, reserved2
  end type

  type, bind(C), public :: q2d_problem
    integer(c_int32_t) :: nblocks, nbc, ncases, reserved
    type(c_ptr)        :: blocks, bc, cases
  end type

  type, bind(C), public :: q2d_step_out
    real(c_double)     :: dtglobal, time
    real(c_double)     :: residual(4)
    integer(c_int32_t) :: iter, todo, status
    integer(c_int32_t) :: fail_b, fail_i, fail_j
    integer(c_int32_t) :: reserved(2)
  end type

  ! runtime keys of one case (input.ini re-read every ini-diter steps, Read_Ini.f90:65-88)
  type, bind(C), public :: q2d_runtime
    real(c_double)     :: cfl, vnn
    integer(c_int32_t) :: ramp_iter, itermax
    real(c_double)     :: res_threshold, time_threshold, res_drop
    integer(c_int32_t) :: limiter, riemann, space_sd, reserved
  end type

  public :: q2d_version, q2d_last_error, q2d_device_count
  public :: q2d_gpu_create, q2d_gpu_destroy, q2d_gpu_step, q2d_gpu_run
  public :: q2d_gpu_set_runtime, q2d_gpu_get_runtime
  public :: q2d_gpu_download_state, q2d_gpu_upload_state, q2d_gpu_download_dt
  public :: q2d_error_message

  interface
    type(c_ptr) function q2d_version() bind(C, name="q2d_version")
      import :: c_ptr
    end function
    type(c_ptr) function q2d_last_error() bind(C, name="q2d_last_error")
      import :: c_ptr
    end function
    integer(c_int) function q2d_device_count(count) bind(C, name="q2d_device_count")
      import :: c_int
      integer(c_int), intent(out) :: count
    end function
    integer(c_int) function q2d_gpu_create(prob, handle) bind(C, name="q2d_gpu_create")
      import :: c_int, c_ptr, q2d_problem
      type(q2d_problem), intent(in) :: prob
      type(c_ptr), intent(out) :: handle
    end function
    integer(c_int) function q2d_gpu_destroy(handle) bind(C, name="q2d_gpu_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function
    integer(c_int) function q2d_gpu_step(handle, out) bind(C, name="q2d_gpu_step")
      import :: c_int, c_ptr, q2d_step_out
      type(c_ptr), value :: handle
      type(q2d_step_out), intent(out) :: out(*)
    end function
    integer(c_int) function q2d_gpu_run(handle, nsteps, out) bind(C, name="q2d_gpu_run")
      import :: c_int, c_int32_t, c_ptr, q2d_step_out
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: nsteps
      type(q2d_step_out), intent(out) :: out(*)
    end function
    integer(c_int) function q2d_gpu_set_runtime(handle, case_id, rt) bind(C, name="q2d_gpu_set_runtime")
      import :: c_int, c_int32_t, c_ptr, q2d_runtime
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: case_id
      type(q2d_runtime), intent(in) :: rt
    end function
    integer(c_int) function q2d_gpu_get_runtime(handle, case_id, rt) bind(C, name="q2d_gpu_get_runtime")
      import :: c_int, c_int32_t, c_ptr, q2d_runtime
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: case_id
      type(q2d_runtime), intent(out) :: rt
    end function
    integer(c_int) function q2d_gpu_download_state(handle, block, P) bind(C, name="q2d_gpu_download_state")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: block
      real(c_double), intent(out) :: P(*)
    end function
    integer(c_int) function q2d_gpu_upload_state(handle, block, P) bind(C, name="q2d_gpu_upload_state")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: block
      real(c_double), intent(in) :: P(*)
    end function
    integer(c_int) function q2d_gpu_download_dt(handle, block, dtlocal) bind(C, name="q2d_gpu_download_dt")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: block
      real(c_double), intent(out) :: dtlocal(*)
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! The library's message for the last failing call, as a Fortran string.
  function q2d_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, k
    p = q2d_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do k = 1, n
      msg(k:k) = chars(k)
    end do
  end function q2d_error_message

end module q2d_gpu
