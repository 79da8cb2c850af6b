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
  use iso_c_binding, only: c_ptr, c_null_ptr, c_associated, c_loc, c_int32_t, c_double,&
                           c_float
  use kind_module, only: I4, SP, DP
  use msg123_gpu_api
  implicit none
  private
  public :: calc_numsol

  type(c_ptr), save :: ctx = c_null_ptr
  type(msg123_mesh), save, target :: mesh
  ! host copies the mesh descriptor points to (kept for the run)
  integer(c_int32_t), allocatable, save, target :: m_offind(:), m_offrow(:), m_dir(:)
  real(c_double), allocatable, save, target :: m_hydf(:), m_abyd(:), m_vol(:), m_top(:)
  real(c_float), allocatable, save, target :: m_reta(:), m_retn(:), m_resi(:), m_poro(:), m_ss(:)
  real(c_double), allocatable, save, target :: m_hsurf(:), m_asurf(:), m_stop(:), m_sbott(:)

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
    real(DP), allocatable :: rech_top(:)
    integer(c_int32_t) :: rc, nrec, why
    integer(I4) :: i

    if (st_sim%sim_type /= 1) stop 'GPU calc_numsol: only transient runs (sim_type 1)'
    if (.not. c_associated(ctx)) call open_context()
    form_switch = 1
    ! recharge volume per top cell (calc_boundary keeps it per recharge cell)
    allocate(rech_top(ncals))
    rech_top(:) = 0.0_DP
    do i = 1, size(calc_rech)
      rech_top(rech2cals(i)) = rech_top(rech2cals(i)) + calc_rech(i)
    end do
    rc = msg123_set_state(ctx, head_new, head_old, srat_old, surf_head, surf_rati, rech_top,&
                          real(delt, c_double))
    call check(rc, 'msg123_set_state')
    rc = msg123_calc_numsol(ctx, stats)
    if (rc /= MSG123_OK .and. rc /= MSG123_ENAN) call check(rc, 'msg123_calc_numsol')
    rc = msg123_get_state(ctx, head_new, srat_new, rel_perm, surf_head, surf_rati)
    call check(rc, 'msg123_get_state')
    out_iter = stats%out_iter
    conv_flag = stats%conv_flag
    in_iter = stats%in_iter_last

    ! convergence log through the reference's unit
    allocate(recs(max(stats%out_iter, 1)))
    rc = msg123_iter_log(ctx, recs, int(size(recs), c_int32_t), nrec, why)
    call check(rc, 'msg123_iter_log')
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

  subroutine open_context()
    use initial_module, only: st_grid, maxout_iter, maxinn_iter, criteria, precon_type, newper
    use read_input, only: len_scal
    use set_cell, only: ncalc, ncals
    use allocate_solution, only: crs_index, dir_conn, hydf_conn, abyd_conn
    use make_cell, only: cell_vol, cell_top
    use assign_calc, only: read_reta, read_retn, read_resi, read_poro, read_ss, surf_bott
    use set_condition, only: hydf_surf, abyd_surf
    use prep_calculation, only: surf_top
    type(msg123_params) :: prm
    integer(I4) :: nconn
    integer(c_int32_t) :: rc

    nconn = crs_index(1)%offind(ncalc)
    if (ncalc /= st_grid%nx*st_grid%ny*st_grid%nz .or. ncals /= st_grid%nx*st_grid%ny)&
      stop 'GPU calc_numsol: the deck must be a full structured box'
    m_offind = crs_index(1)%offind(0:ncalc) ; m_offrow = crs_index(1)%offrow(1:nconn)
    m_dir = dir_conn(1:nconn) ; m_hydf = hydf_conn(1:nconn) ; m_abyd = abyd_conn(1:nconn)
    m_vol = cell_vol(1:ncalc) ; m_top = cell_top(1:ncalc)
    m_reta = read_reta(1:ncalc) ; m_retn = read_retn(1:ncalc) ; m_resi = read_resi(1:ncalc)
    m_poro = read_poro(1:ncalc) ; m_ss = read_ss(1:ncalc)
    m_hsurf = hydf_surf(1:ncals) ; m_asurf = abyd_surf(1:ncals)
    m_stop = surf_top(1:ncals) ; m_sbott = surf_bott(1:ncals)
    mesh%nx = st_grid%nx ; mesh%ny = st_grid%ny ; mesh%nz = st_grid%nz
    mesh%ncalc = ncalc ; mesh%ncals = ncals ; mesh%nconn = nconn
    mesh%offind = c_loc(m_offind) ; mesh%offrow = c_loc(m_offrow) ; mesh%dir_conn = c_loc(m_dir)
    mesh%hydf_conn = c_loc(m_hydf) ; mesh%abyd_conn = c_loc(m_abyd)
    mesh%cell_vol = c_loc(m_vol) ; mesh%cell_top = c_loc(m_top)
    mesh%reta = c_loc(m_reta) ; mesh%retn = c_loc(m_retn) ; mesh%resi = c_loc(m_resi)
    mesh%poro = c_loc(m_poro) ; mesh%ss = c_loc(m_ss)
    mesh%hydf_surf = c_loc(m_hsurf) ; mesh%abyd_surf = c_loc(m_asurf)
    mesh%surf_top = c_loc(m_stop) ; mesh%surf_bott = c_loc(m_sbott)
    mesh%len_scal = real(len_scal, c_double) ; mesh%newper = newper
    prm%maxout_iter = maxout_iter ; prm%maxinn_iter = maxinn_iter
    prm%criteria = real(criteria, c_double)
    prm%precon = MSG123_PRECON_DILU
    if (precon_type == 1) prm%precon = MSG123_PRECON_AMG
    prm%consistent_jacobian = 0
    rc = msg123_ctx_create(mesh, prm, ctx)
    call check(rc, 'msg123_ctx_create')
  end subroutine open_context

  subroutine check(rc, what)
    integer(c_int32_t), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= MSG123_OK) then
      write(*, '(a,a,i0)') what, ' failed with code ', rc
      stop 'GPU calc_numsol: library error'
    end if
  end subroutine check

end module nonlinear_solution
