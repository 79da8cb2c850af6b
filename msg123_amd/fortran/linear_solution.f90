! linear_solution -- drop-in replacement of the reference's linear solver
! module (Linear_solution.f90).  Same module name and public names
! (solve_linalg, :26-58, and in_iter): the level-1 system array_var(1) is
! loaded on the GPU (msg123_load_system) and solved by right-preconditioned
! BiCGStab with the D-ILU(0) preconditioner (msg123_solve_linalg) to the
! tolerance errtol the Newton step set; in_iter and conv_flag are set as
! solve_bicgs and check_insol set them.  Only the D-ILU path (precon_type 0):
! the AMG path's side effects on the level-1 system are not copied back.
module linear_solution
  use iso_c_binding, only: c_int32_t, c_double
  use kind_module, only: I4, DP
  use msg123_gpu_api, only: msg123_load_system, msg123_solve_linalg
  use msg123_gpu_context, only: gpu_ctx, gpu_check
  implicit none
  private
  public :: solve_linalg
  integer(I4), public :: in_iter

contains

  subroutine solve_linalg(init_norm, inx, last_norm)
    use initial_module, only: errtol, precon_type
    use prep_calculation, only: conv_flag
    use allocate_solution, only: array_var
    real(DP), intent(in) :: init_norm
    real(DP), intent(inout) :: inx(:)
    real(DP), intent(out) :: last_norm
    integer(c_int32_t) :: it, conv
    real(c_double) :: rn
    if (precon_type /= 0) stop 'GPU solve_linalg: D-ILU preconditioner (precon_type 0) only'
    call gpu_check(msg123_load_system(gpu_ctx(), array_var(1)%rhs, array_var(1)%dmat,&
                                      array_var(1)%lumat), 'msg123_load_system')
    call gpu_check(msg123_solve_linalg(gpu_ctx(), real(init_norm, c_double),&
                                       real(errtol, c_double), inx, rn, it, conv),&
                   'msg123_solve_linalg')
    in_iter = it
    conv_flag = conv
    last_norm = rn
  end subroutine solve_linalg

end module linear_solution
