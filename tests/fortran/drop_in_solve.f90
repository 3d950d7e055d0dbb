! Drop-in check of libhprmat_hip from Fortran, through the ISO_C_BINDING
! module hprmat_amd/fortran/hprmat_hip_solver.f90 (the bind(C) names of
! src/gpu_solver_interface.F90:42-162).
!
! Usage: drop_in_solve <input> <output>
!   input  (stream, native endian): int32 n, int32 nrhs, complex128 A(n, n),
!          complex128 B(n, nrhs), both column-major
!   output: for each of the solvers double, mixed, hybrid, tf32, auto, multi,
!          multi_mixed (in that order): int32 ierr, complex128 X(n, nrhs)
! Without a GPU it prints the device count and "no GPU" and writes nothing.
program drop_in_solve
  use iso_c_binding
  use hprmat_hip_solver
  implicit none
  character(len=512) :: fin, fout
  character(len=256) :: info
  integer :: n, nrhs, ierr, k, ngpu
  complex(c_double_complex), allocatable :: A(:, :), B(:, :), X(:, :)
  real(c_double) :: tol
  integer, parameter :: nsolvers = 7
  character(len=12), parameter :: names(nsolvers) = [character(len=12) :: &
      'double', 'mixed', 'hybrid', 'tf32', 'auto', 'multi', 'multi_mixed']

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  print '(a, i0)', 'devices ', gpu_get_device_count()
  if (.not. gpu_is_available()) then
    print '(a)', 'no GPU'
    stop
  end if
  call gpu_solver_init(0, ierr)
  if (ierr /= 0) then
    print '(a, i0)', 'gpu_solver_init failed: ', ierr
    error stop 2
  end if
  call gpu_get_info(info)
  print '(a, a)', 'device ', trim(info)
  ngpu = 4
  call gpu_multi_init(ngpu, ierr)
  print '(a, i0, a, i0)', 'multi_init ', ierr, ' ngpu ', ngpu

  open(10, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(10) n, nrhs
  allocate(A(n, n), B(n, nrhs), X(n, nrhs))
  read(10) A, B
  close(10)
  print '(a, i0, a, i0)', 'n ', n, ' nrhs ', nrhs

  tol = 1.0d-13
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  do k = 1, nsolvers
    X = B
    select case (k)
    case (1)
      call gpu_solve_double(A, X, n, nrhs, ierr)
    case (2)
      call gpu_solve_mixed(A, X, n, nrhs, 10, tol, ierr)
    case (3)
      call gpu_solve_hybrid(A, X, n, nrhs, 10, tol, ierr)
    case (4)
      call gpu_solve_tf32(A, X, n, nrhs, 10, tol, ierr)
    case (5)
      call gpu_solve_auto(A, X, n, nrhs, 10, tol, ierr)
    case (6)
      call gpu_solve_multi(A, X, n, nrhs, ierr)
    case (7)
      call gpu_solve_multi_mixed(A, X, n, nrhs, 10, tol, ierr)
    end select
    print '(a, a, a, i0)', 'solver ', trim(names(k)), ' ierr ', ierr
    write(11) int(ierr, c_int32_t), X
  end do
  close(11)
  call gpu_multi_finalize(ierr)
  call gpu_host_unregister()
  call gpu_solver_finalize(ierr)
  print '(a, i0)', 'finalize ', ierr
end program drop_in_solve
