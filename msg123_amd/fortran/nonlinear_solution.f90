! nonlinear_solution -- drop-in replacement of the reference's Newton module
! (Nonlinear_solution.f90).  Same module name and public procedure
! (calc_numsol, :33-294, called by Msg.f90's time loop through
! time_module); the Newton-Krylov step runs on the GPU through
! libmsg123_gpu.so.  On entry the reference's module state (heads, old
! saturations, surface terms, recharge, dt) is handed to the device context;
! on return heads, saturations, relative permeabilities, surface head,
! out_iter, conv_flag and in_iter are written back, the convergence log is
! written through the reference's own unit, and the restart file as the
! reference writes it at convergence.
!
! Scope: what the library supports -- a structured box of active cells
! (grid type 7), transient runs (sim_type 1), no MPI, no wells, rivers,
! lakes or sea cells; other decks stop with a message.
module nonlinear_solution
  use iso_c_binding, only: c_ptr, c_int32_t, c_double
  use kind_module, only: I4, SP, DP
  use msg123_gpu_api
  use msg123_gpu_context, only: gpu_ctx, gpu_load_level, gpu_check
  implicit none
  private
  public :: calc_numsol

contains

  subroutine calc_numsol()
    use initial_module, only: st_sim, st_out_step
    use prep_calculation, only: delt, conv_flag, out_iter, form_switch, current_t
    use set_cell, only: ncalc, ncals
    use allocate_solution, only: head_new, head_old, srat_old, srat_new, rel_perm,&
                                 surf_head, surf_rati
    use calc_boundary, only: calc_rech, rech2cals
    use linear_solution, only: in_iter
    use check_condition, only: st_out_fnum
    use time_module, only: now_time
    type(msg123_newton_stats) :: stats
    type(msg123_iter_rec), allocatable :: recs(:)
    type(c_ptr) :: ctx
    integer(c_int32_t) :: rc, nrec, why

    form_switch = 1
    call gpu_load_level()
    ctx = gpu_ctx()
    rc = msg123_calc_numsol(ctx, stats)
    if (rc /= MSG123_OK .and. rc /= MSG123_ENAN) call gpu_check(rc, 'msg123_calc_numsol')
    rc = msg123_get_state(ctx, head_new, srat_new, rel_perm, surf_head, surf_rati)
    call gpu_check(rc, 'msg123_get_state')
    out_iter = stats%out_iter
    conv_flag = stats%conv_flag
    in_iter = stats%in_iter_last

    ! convergence log through the reference's unit
    allocate(recs(max(stats%out_iter, 1)))
    rc = msg123_iter_log(ctx, recs, int(size(recs), c_int32_t), nrec, why)
    call gpu_check(rc, 'msg123_iter_log')
    call write_log(st_out_fnum%conv, now_time, trim(st_sim%cal_unit), delt, recs,&
                   min(nrec, int(size(recs), c_int32_t)), why)
    if (conv_flag == 1) then
      if (st_out_step%rest == 0.0_SP) then
        call write_restart(st_out_fnum%rest, now_time)
      else if (mod(current_t, st_out_step%rest) == 0) then
        call write_restart(st_out_fnum%rest, now_time)
      end if
    end if
  end subroutine calc_numsol

  subroutine write_log(u, t, unit_name, dt, recs, nrec, why)
    integer(I4), intent(in) :: u
    real(SP), intent(in) :: t
    character(*), intent(in) :: unit_name
    real(DP), intent(in) :: dt
    type(msg123_iter_rec), intent(in) :: recs(:)
    integer(c_int32_t), intent(in) :: nrec, why
    character(64) :: cell
    integer(I4) :: k
    write(u, '(//1x,"CURRENT TIME : ",es12.5,1x,"(",a,")",20x,"TIME STEP : ",es12.5,1x,"(SEC)")')&
      t, unit_name, dt
    write(u, '(1x,84("-"))')
    write(u, '(1x,a)') " OUTER INNER  BACK     MAXIMUM           MAXIMUM    DIAGONAL  RIGHT HAND     UNKNOWN"
    write(u, '(1x,a)') "                        CHANGE              CELL      MATRIX      VECTOR       VALUE"
    write(u, '(1x,84("-"))')
    do k = 1, nrec
      write(cell, '("(",i0,",",i0,",",i0,")")') recs(k)%cell_i, recs(k)%cell_j, recs(k)%cell_k
      write(u, '(1x,3(i6),es12.3,a18,4(es12.3))') recs(k)%out_iter, recs(k)%in_iter,&
        recs(k)%back_iter, recs(k)%max_change, trim(adjustl(cell)), recs(k)%dmat,&
        recs(k)%rhs, recs(k)%head
    end do
    select case (why)
    case (MSG123_END_BACKTRACK)
      write(u, '(1x,a)') "Stop due to maximum value or change in backtracking"
    case (MSG123_END_MAXITER)
      write(u, '(1x,a)') "Stop due to maximum number of nonlinear iteration"
    case (MSG123_END_DIVERGED)
      write(u, '(1x,a)') "Didn't converge due to maximum value or change"
    end select
  end subroutine write_log

  ! restart record: time, then the heads in metres (unformatted, rewound)
  subroutine write_restart(u, t)
    use set_cell, only: ncalc
    use read_input, only: len_scal
    use allocate_solution, only: head_new
    integer(I4), intent(in) :: u
    real(SP), intent(in) :: t
    integer(I4) :: i
    rewind(u)
    write(u) real(t, kind=DP)
    write(u) (head_new(i)*len_scal, i = 1, ncalc)
  end subroutine write_restart

end module nonlinear_solution
