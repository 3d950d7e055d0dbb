! hprmat_hip_solver.f90 - Fortran binding of libhprmat_hip's reference ABI.
!
! The 15 C entry points that HPRMAT's Fortran code binds by name
! (src/gpu_solver_interface.F90:42-162) are declared here with
! bind(C, name=...) interface blocks, and wrapped by module procedures with
! the names and argument lists of that module's public procedures
! (src/gpu_solver_interface.F90:20-34, :165-560):
!   gpu_solver_init(device_id, ierr), gpu_solver_finalize(ierr),
!   gpu_host_unregister(), gpu_solve_mixed / _hybrid / _tf32 / _auto /
!   _multi_mixed(A, B, n, nrhs, max_refine, tol, ierr),
!   gpu_solve_double / _multi(A, B, n, nrhs, ierr), gpu_is_available(),
!   gpu_get_device_count(), gpu_get_info(str), gpu_multi_init(ngpu, ierr),
!   gpu_multi_finalize(ierr).
!
! Arguments are passed by reference, A and B are complex(c_double_complex)
! column-major with leading dimension n, A is read only and B returns the
! solution (left untouched on failure).  ierr is the library's info code:
! 0 success, -1 initialisation failed, -2 allocation failed, > 0 zero pivot
! (1-based column), 99 refinement did not converge.  One difference from the
! reference's wrappers, which re-solve with LAPACK ZGESV when the GPU call
! fails: a failed call here returns its info code (or -1 when the library
! reports failure without one) and leaves B as it was, so a caller never
! receives a CPU result from a GPU entry point.
!
! Build: amdflang -c hprmat_hip_solver.f90; link with -lhprmat_hip.
module hprmat_hip_solver
  use iso_c_binding
  implicit none
  private

  public :: gpu_solver_init, gpu_solver_finalize, gpu_host_unregister
  public :: gpu_solve_mixed, gpu_solve_hybrid, gpu_solve_tf32, gpu_solve_auto
  public :: gpu_solve_double, gpu_solve_multi, gpu_solve_multi_mixed
  public :: gpu_is_available, gpu_get_device_count, gpu_get_info
  public :: gpu_multi_init, gpu_multi_finalize

  interface
    integer(c_int) function c_gpu_solver_init(device_id) bind(C, name="gpu_solver_init_")
      import :: c_int
      integer(c_int), intent(in) :: device_id
    end function
    integer(c_int) function c_gpu_solver_finalize() bind(C, name="gpu_solver_finalize_")
      import :: c_int
    end function
    integer(c_int) function c_gpu_host_unregister() bind(C, name="gpu_host_unregister_")
      import :: c_int
    end function
    integer(c_int) function c_gpu_solve_mixed(A, B, n, nrhs, max_refine, tol, info) &
        bind(C, name="gpu_solve_mixed_")
      import :: c_int, c_double, c_double_complex
      complex(c_double_complex), intent(in) :: A(*)
      complex(c_double_complex), intent(inout) :: B(*)
      integer(c_int), intent(in) :: n, nrhs, max_refine
      real(c_double), intent(in) :: tol
      integer(c_int), intent(out) :: info
    end function
    integer(c_int) function c_gpu_solve_hybrid(A, B, n, nrhs, max_refine, tol, info) &
        bind(C, name="gpu_solve_hybrid_")
      import :: c_int, c_double, c_double_complex
      complex(c_double_complex), intent(in) :: A(*)
      complex(c_double_complex), intent(inout) :: B(*)
      integer(c_int), intent(in) :: n, nrhs, max_refine
      real(c_double), intent(in) :: tol
      integer(c_int), intent(out) :: info
    end function
    integer(c_int) function c_gpu_solve_tf32(A, B, n, nrhs, max_refine, tol, info) &
        bind(C, name="gpu_solve_tf32_")
      import :: c_int, c_double, c_double_complex
      complex(c_double_complex), intent(in) :: A(*)
      complex(c_double_complex), intent(inout) :: B(*)
      integer(c_int), intent(in) :: n, nrhs, max_refine
      real(c_double), intent(in) :: tol
      integer(c_int), intent(out) :: info
    end function
    integer(c_int) function c_gpu_solve_auto(A, B, n, nrhs, max_refine, tol, info) &
        bind(C, name="gpu_solve_auto_")
      import :: c_int, c_double, c_double_complex
      complex(c_double_complex), intent(in) :: A(*)
      complex(c_double_complex), intent(inout) :: B(*)
      integer(c_int), intent(in) :: n, nrhs, max_refine
      real(c_double), intent(in) :: tol
      integer(c_int), intent(out) :: info
    end function
    integer(c_int) function c_gpu_solve_double(A, B, n, nrhs, info) &
        bind(C, name="gpu_solve_double_")
      import :: c_int, c_double_complex
      complex(c_double_complex), intent(in) :: A(*)
      complex(c_double_complex), intent(inout) :: B(*)
      integer(c_int), intent(in) :: n, nrhs
      integer(c_int), intent(out) :: info
    end function
    integer(c_int) function c_gpu_is_available() bind(C, name="gpu_is_available_")
      import :: c_int
    end function
    integer(c_int) function c_gpu_get_device_count() bind(C, name="gpu_get_device_count_")
      import :: c_int
    end function
    subroutine c_gpu_get_info(info_str, len) bind(C, name="gpu_get_info_")
      import :: c_char, c_int
      character(kind=c_char), intent(out) :: info_str(*)
      integer(c_int), intent(in) :: len
    end subroutine
    integer(c_int) function c_gpu_multi_init(ngpu) bind(C, name="gpu_multi_init_")
      import :: c_int
      integer(c_int), intent(inout) :: ngpu
    end function
    integer(c_int) function c_gpu_multi_finalize() bind(C, name="gpu_multi_finalize_")
      import :: c_int
    end function
    integer(c_int) function c_gpu_solve_multi(A, B, n, nrhs, info) &
        bind(C, name="gpu_solve_multi_")
      import :: c_int, c_double_complex
      complex(c_double_complex), intent(in) :: A(*)
      complex(c_double_complex), intent(inout) :: B(*)
      integer(c_int), intent(in) :: n, nrhs
      integer(c_int), intent(out) :: info
    end function
    integer(c_int) function c_gpu_solve_multi_mixed(A, B, n, nrhs, max_refine, tol, info) &
        bind(C, name="gpu_solve_multi_mixed_")
      import :: c_int, c_double, c_double_complex
      complex(c_double_complex), intent(in) :: A(*)
      complex(c_double_complex), intent(inout) :: B(*)
      integer(c_int), intent(in) :: n, nrhs, max_refine
      real(c_double), intent(in) :: tol
      integer(c_int), intent(out) :: info
    end function
  end interface

contains

  ! The library's info code, or -1 when the call failed without one.
  integer function outcome(ret, info)
    integer(c_int), intent(in) :: ret, info
    outcome = info
    if (ret /= 0 .and. info == 0) outcome = -1
  end function outcome

  subroutine gpu_solver_init(device_id, ierr)
    integer, intent(in) :: device_id
    integer, intent(out) :: ierr
    integer(c_int) :: dev
    dev = device_id
    ierr = c_gpu_solver_init(dev)
  end subroutine gpu_solver_init

  subroutine gpu_solver_finalize(ierr)
    integer, intent(out) :: ierr
    ierr = c_gpu_solver_finalize()
  end subroutine gpu_solver_finalize

  subroutine gpu_host_unregister()
    integer(c_int) :: ret
    ret = c_gpu_host_unregister()
  end subroutine gpu_host_unregister

  subroutine gpu_solve_mixed(A, B, n, nrhs, max_refine, tol, ierr)
    integer, intent(in) :: n, nrhs, max_refine
    complex(c_double_complex), intent(in) :: A(n, n)
    complex(c_double_complex), intent(inout) :: B(n, nrhs)
    real(c_double), intent(in) :: tol
    integer, intent(out) :: ierr
    integer(c_int) :: cn, cr, cm, info, ret
    cn = n; cr = nrhs; cm = max_refine; info = 0
    ret = c_gpu_solve_mixed(A, B, cn, cr, cm, tol, info)
    ierr = outcome(ret, info)
  end subroutine gpu_solve_mixed

  subroutine gpu_solve_hybrid(A, B, n, nrhs, max_refine, tol, ierr)
    integer, intent(in) :: n, nrhs, max_refine
    complex(c_double_complex), intent(in) :: A(n, n)
    complex(c_double_complex), intent(inout) :: B(n, nrhs)
    real(c_double), intent(in) :: tol
    integer, intent(out) :: ierr
    integer(c_int) :: cn, cr, cm, info, ret
    cn = n; cr = nrhs; cm = max_refine; info = 0
    ret = c_gpu_solve_hybrid(A, B, cn, cr, cm, tol, info)
    ierr = outcome(ret, info)
  end subroutine gpu_solve_hybrid

  subroutine gpu_solve_tf32(A, B, n, nrhs, max_refine, tol, ierr)
    integer, intent(in) :: n, nrhs, max_refine
    complex(c_double_complex), intent(in) :: A(n, n)
    complex(c_double_complex), intent(inout) :: B(n, nrhs)
    real(c_double), intent(in) :: tol
    integer, intent(out) :: ierr
    integer(c_int) :: cn, cr, cm, info, ret
    cn = n; cr = nrhs; cm = max_refine; info = 0
    ret = c_gpu_solve_tf32(A, B, cn, cr, cm, tol, info)
    ierr = outcome(ret, info)
  end subroutine gpu_solve_tf32

  subroutine gpu_solve_auto(A, B, n, nrhs, max_refine, tol, ierr)
    integer, intent(in) :: n, nrhs, max_refine
    complex(c_double_complex), intent(in) :: A(n, n)
    complex(c_double_complex), intent(inout) :: B(n, nrhs)
    real(c_double), intent(in) :: tol
    integer, intent(out) :: ierr
    integer(c_int) :: cn, cr, cm, info, ret
    cn = n; cr = nrhs; cm = max_refine; info = 0
    ret = c_gpu_solve_auto(A, B, cn, cr, cm, tol, info)
    ierr = outcome(ret, info)
  end subroutine gpu_solve_auto

  subroutine gpu_solve_multi_mixed(A, B, n, nrhs, max_refine, tol, ierr)
    integer, intent(in) :: n, nrhs, max_refine
    complex(c_double_complex), intent(in) :: A(n, n)
    complex(c_double_complex), intent(inout) :: B(n, nrhs)
    real(c_double), intent(in) :: tol
    integer, intent(out) :: ierr
    integer(c_int) :: cn, cr, cm, info, ret
    cn = n; cr = nrhs; cm = max_refine; info = 0
    ret = c_gpu_solve_multi_mixed(A, B, cn, cr, cm, tol, info)
    ierr = outcome(ret, info)
  end subroutine gpu_solve_multi_mixed

  subroutine gpu_solve_double(A, B, n, nrhs, ierr)
    integer, intent(in) :: n, nrhs
    complex(c_double_complex), intent(in) :: A(n, n)
    complex(c_double_complex), intent(inout) :: B(n, nrhs)
    integer, intent(out) :: ierr
    integer(c_int) :: cn, cr, info, ret
    cn = n; cr = nrhs; info = 0
    ret = c_gpu_solve_double(A, B, cn, cr, info)
    ierr = outcome(ret, info)
  end subroutine gpu_solve_double

  subroutine gpu_solve_multi(A, B, n, nrhs, ierr)
    integer, intent(in) :: n, nrhs
    complex(c_double_complex), intent(in) :: A(n, n)
    complex(c_double_complex), intent(inout) :: B(n, nrhs)
    integer, intent(out) :: ierr
    integer(c_int) :: cn, cr, info, ret
    cn = n; cr = nrhs; info = 0
    ret = c_gpu_solve_multi(A, B, cn, cr, info)
    ierr = outcome(ret, info)
  end subroutine gpu_solve_multi

  logical function gpu_is_available()
    gpu_is_available = c_gpu_is_available() /= 0
  end function gpu_is_available

  integer function gpu_get_device_count()
    gpu_get_device_count = c_gpu_get_device_count()
  end function gpu_get_device_count

  subroutine gpu_get_info(info_str)
    character(len=*), intent(out) :: info_str
    character(kind=c_char) :: buf(256)
    integer(c_int) :: blen
    integer :: i
    blen = 256
    buf = c_null_char
    call c_gpu_get_info(buf, blen)
    info_str = ''
    do i = 1, min(255, len(info_str))
      if (buf(i) == c_null_char) exit
      info_str(i:i) = buf(i)
    end do
  end subroutine gpu_get_info

  subroutine gpu_multi_init(ngpu, ierr)
    integer, intent(inout) :: ngpu
    integer, intent(out) :: ierr
    integer(c_int) :: cg
    cg = ngpu
    ierr = c_gpu_multi_init(cg)
    ngpu = cg
  end subroutine gpu_multi_init

  subroutine gpu_multi_finalize(ierr)
    integer, intent(out) :: ierr
    ierr = c_gpu_multi_finalize()
  end subroutine gpu_multi_finalize

end module hprmat_hip_solver
