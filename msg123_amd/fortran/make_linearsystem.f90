! make_linearsystem -- drop-in replacement of the reference's Jacobian
! module (Make_linearsystem.f90).  Same module name and public procedure
! (make_matvec, :24-68): rhs = -F(head_new) and the finite-difference
! Jacobian are formed on the GPU (msg123_make_matvec) and written to the
! level-1 system array_var(1)%rhs / dmat / lumat in the reference's CSR
! order; srat_new, rel_perm and surf_rati are left as calc_func leaves them.
! With precon_type = 1 the AMG hierarchy is then built by make_amgmat, as
! the reference does.
module make_linearsystem
  use kind_module, only: I4, DP
  use msg123_gpu_api, only: msg123_make_matvec
  use msg123_gpu_context, only: gpu_ctx, gpu_load_level, gpu_store_side, gpu_check
  implicit none
  private
  public :: make_matvec

contains

  subroutine make_matvec()
    use initial_module, only: precon_type
    use set_cell, only: ncalc
    use allocate_solution, only: array_var, crs_index
    use make_amg_matrix, only: make_amgmat
    integer(I4) :: nconn
    nconn = crs_index(1)%offind(ncalc)
    call gpu_load_level()
    call gpu_check(msg123_make_matvec(gpu_ctx(), array_var(1)%rhs, array_var(1)%dmat,&
                                      array_var(1)%lumat), 'msg123_make_matvec')
    call gpu_store_side()
    if (precon_type == 1) call make_amgmat()
  end subroutine make_matvec

end module make_linearsystem
