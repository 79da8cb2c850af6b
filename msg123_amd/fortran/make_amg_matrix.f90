! make_amg_matrix -- drop-in replacement of the reference's AMG set-up module
! (Make_amg_matrix.f90).  Same module name and public procedure
! (make_amgmat, :20-65); the hierarchy is built by msg123_amg_setup in
! libmsg123_gpu.so (host code) and copied into the reference's level
! storage: crs_index(l), array_var(l), pro_var(l), res_var(l) for
! l = 2..nlevel, nlevel and amg_setflag, exactly as the reference fills them.
module make_amg_matrix
  use iso_c_binding, only: c_int32_t, c_double, c_float, c_f_pointer
  use kind_module, only: I4, DP
  use constval_module, only: DZERO
  use allocate_solution, only: crs_index, array_var, pro_var, res_var
  use msg123_gpu_api, only: msg123_amg, msg123_amg_level, msg123_amg_setup, msg123_amg_free,&
                            MSG123_OK
  implicit none
  private
  public :: make_amgmat

contains

  subroutine make_amgmat()
    use initial_module, only: nlevel, amg_nlevel, amg_theta, jac_omega
    use set_cell, only: amg_setflag
    type(msg123_amg) :: amg
    type(msg123_amg_level), pointer :: lv(:)
    integer(I4) :: l, n1, nnz1, rc

    amg_setflag = 0
    n1 = crs_index(1)%unknow
    nnz1 = crs_index(1)%offind(n1)
    rc = msg123_amg_setup(crs_index(1)%offind(0:n1), crs_index(1)%offrow(1:nnz1),&
                          array_var(1)%dmat(1:n1), array_var(1)%lumat(1:nnz1),&
                          int(n1, c_int32_t), int(amg_nlevel, c_int32_t),&
                          real(amg_theta, c_float), real(jac_omega, c_float), amg)
    if (rc /= MSG123_OK) stop 'msg123_amg_setup failed (prolongation width exceeded?)'
    call c_f_pointer(amg%levels, lv, [amg%nlevel])
    nlevel = amg%nlevel
    do l = 2, nlevel
      call take_level(l, lv(l))
    end do
    call msg123_amg_free(amg)
    amg_setflag = 1
  end subroutine make_amgmat

  subroutine take_level(l, src)
    integer(I4), intent(in) :: l
    type(msg123_amg_level), intent(in) :: src
    integer(c_int32_t), pointer :: ip(:)
    real(c_double), pointer :: rp(:)
    integer(I4) :: n, nnz, nf, np

    n = src%n ; nnz = src%nnz ; nf = src%nfine
    crs_index(l)%unknow = n
    crs_index(l)%lunum = nnz
    allocate(crs_index(l)%offind(0:n), crs_index(l)%offrow(nnz))
    call c_f_pointer(src%offind, ip, [n + 1]) ; crs_index(l)%offind(0:n) = ip
    if (nnz > 0) then
      call c_f_pointer(src%offrow, ip, [nnz]) ; crs_index(l)%offrow(:) = ip
    end if
    allocate(array_var(l)%dmat(n), array_var(l)%lumat(nnz))
    allocate(array_var(l)%x(n), array_var(l)%rhs(n))
    if (n > 0) then
      call c_f_pointer(src%dmat, rp, [n]) ; array_var(l)%dmat(:) = rp
    end if
    if (nnz > 0) then
      call c_f_pointer(src%lumat, rp, [nnz]) ; array_var(l)%lumat(:) = rp
    end if
    array_var(l)%x(:) = DZERO ; array_var(l)%rhs(:) = DZERO

    allocate(pro_var(l)%pindex(0:nf))
    call c_f_pointer(src%pindex, ip, [nf + 1]) ; pro_var(l)%pindex(0:nf) = ip
    np = pro_var(l)%pindex(nf)
    allocate(pro_var(l)%poffrow(np), pro_var(l)%pval(np))
    allocate(res_var(l)%rindex(0:n), res_var(l)%roffrow(np), res_var(l)%rval(np))
    call c_f_pointer(src%rindex, ip, [n + 1]) ; res_var(l)%rindex(0:n) = ip
    if (np > 0) then
      call c_f_pointer(src%poffrow, ip, [np]) ; pro_var(l)%poffrow(:) = ip
      call c_f_pointer(src%pval, rp, [np]) ; pro_var(l)%pval(:) = rp
      call c_f_pointer(src%roffrow, ip, [np]) ; res_var(l)%roffrow(:) = ip
      call c_f_pointer(src%rval, rp, [np]) ; res_var(l)%rval(:) = rp
    end if
  end subroutine take_level

end module make_amg_matrix
