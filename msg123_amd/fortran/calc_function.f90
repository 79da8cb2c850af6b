! calc_function -- drop-in replacement of the reference's residual module
! (Calc_function.f90).  Same module name and public names.  calc_func
! (:24-101) and calc_vecjacf (:515-621), the two the Newton step calls, run
! on the GPU through libmsg123_gpu.so (msg123_calc_func,
! msg123_calc_vecjacf); the time level they read is handed over on every
! call and their side effects on srat_new, rel_perm and surf_rati are copied
! back.  The other public procedures (calc_mass and the boundary terms the
! output module reads) are the reference's own, re-exported from its module
! compiled under the name ref_calc_function.
module calc_function
  use iso_c_binding, only: c_int32_t, c_double
  use kind_module, only: I4, DP
  use msg123_gpu_api, only: msg123_calc_func, msg123_calc_vecjacf
  use msg123_gpu_context, only: gpu_ctx, gpu_load_level, gpu_store_side, gpu_check
  use ref_calc_function, only: calc_mass, func_rechterm, func_wellterm, func_surfterm,&
                               func_riveterm, func_laketerm, func_sealterm
  implicit none
  private
  public :: calc_func, calc_mass, calc_vecjacf, func_rechterm, func_wellterm
  public :: func_surfterm, func_riveterm, func_laketerm, func_sealterm

contains

  subroutine calc_func(infx, funcv)
    real(DP), intent(inout) :: infx(:)
    real(DP), intent(out) :: funcv(:)
    call gpu_load_level()
    call gpu_check(msg123_calc_func(gpu_ctx(), infx, funcv), 'msg123_calc_func')
    call gpu_store_side()
  end subroutine calc_func

  subroutine calc_vecjacf(vjlevel, injx, injvec, outjvec)
    integer(I4), intent(in) :: vjlevel
    real(DP), intent(in) :: injvec(:)
    real(DP), intent(inout) :: injx(:)
    real(DP), intent(out) :: outjvec(:)
    if (vjlevel /= 1) stop 'GPU calc_vecjacf: level 1 only'
    call gpu_load_level()
    call gpu_check(msg123_calc_vecjacf(gpu_ctx(), injx, injvec, outjvec), 'msg123_calc_vecjacf')
    call gpu_store_side()
  end subroutine calc_vecjacf

end module calc_function
