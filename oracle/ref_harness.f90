! ref_harness -- run the reference solver on a deck and dump FP64 state.
!
! Test infrastructure only.  This program is compiled against the reference
! Fortran modules (built from their sources under the reference tree by
! oracle/Makefile into oracle/_ref/) and replaces the reference's main
! program Msg.f90 (Msg.f90:43-104): it performs the same set-up calls and
! the same time loop, and additionally writes FP64 data the reference's own
! outputs do not carry (its head output is float32, Write_module.f90:171):
!
!   static.bin   -- grid, CSR pattern, face coefficients, soil parameters
!   assem1.bin   -- F, dmat, lumat at the first Newton iteration of step 1,
!                   plus the first linear solve's correction and counts
!   wet.bin      -- F, dmat, lumat at a fully saturated, ponded state
!   steps.bin    -- per accepted step: time, dt, Newton count, FP64 heads
!   timing.txt   -- with MSG123_TIME_NEWTON=k set: wall seconds of the first
!                   step's calc_numsol capped at k Newton iterations
!   steptime.txt -- wall seconds of every calc_numsol call of the time loop
!                   (accepted or not) and its Newton count
!
! MSG123_MAX_STEPS=k stops the time loop after k accepted steps.
!
! With a file bterms.bin in the run directory, river, lake, well and sea
! boundary terms are set directly in the reference's module variables after
! its set-up (the arrays its input readers would have built: rive2cals,
! rive_head, rive_bott, abyd_rive; the lake equivalents; one well whose
! connections well_conn carry calc_well; the sea connections seal2calc,
! seal2seal, hydf_seal, abyd_seal with read_seal), so the reference's own
! calc_func, make_matrix and calc_mass evaluate them.  calc_well is the
! fixed per-cell value given: well_num is 0 while update_tstep runs, which
! would otherwise recompute it from well heads the deck does not have.
!
! All records are raw little-endian stream data; oracle/refdump.py reads them.
program ref_harness
  use kind_module, only: I4, SP, DP
  use initial_module, only: init_msg, errtol, st_sim
  use read_input, only: read_main_file, len_scal, len_scal_inv
  use set_cell, only: set_cell_info, ncalc, ncals
  use prep_calculation, only: prepare_calc, conv_flag, current_t, delt, out_iter, surf_top,&
                              form_switch
  use set_boundary, only: set_bound, rech_num
  use allocate_solution, only: allocate_solvar, crs_index, dir_conn, hydf_conn, abyd_conn,&
                               array_var, head_new, head_old, srat_new, srat_old, rel_perm,&
                               head_change, surf_head
  use time_module, only: update_tstep, now_time
  use nonlinear_solution, only: calc_numsol
  use write_output, only: write_outf
  use check_simulation, only: check_lastts, lasttime_flag
  use make_cell, only: cell_vol, cell_top
  use assign_calc, only: read_reta, read_retn, read_resi, read_poro, read_ss, surf_bott
  use set_condition, only: hydf_surf, abyd_surf, rech_area
  use calc_boundary, only: rech2cals, calc_rech
  use make_linearsystem, only: make_matvec
  use linear_solution, only: solve_linalg, in_iter
  use calc_simulation, only: calc_l2norm2
  use calc_function, only: calc_func
  use initial_module, only: newper, precon_type, nlevel, maxout_iter
  use allocate_solution, only: pro_var, res_var
  use set_boundary, only: rive_num, lake_num, well_num, abyd_rive, abyd_lake
  use calc_boundary, only: rive2cals, rive_head, rive_bott, lake2cals, lake_head, lake_bott
  use set_condition, only: well_index, well_conn, nseal
  use assign_boundary, only: calc_well, read_seal
  use allocate_solution, only: seal2calc, seal2seal, hydf_seal, abyd_seal

  implicit none
  integer(I4) :: sta_value(8), u, ut, nconn, nstep, i, ktime, maxsteps
  integer(I4) :: nwell_bt
  logical :: dumped, have_bt
  integer(8) :: c0, c1, crate
  character(32) :: envv
  real(DP) :: l2b, lastn
  real(DP), allocatable :: dh(:), f(:), hsave(:)

  call DATE_AND_TIME(values = sta_value)
  call init_msg(sta_value)
  call read_main_file()
  call set_cell_info()
  call prepare_calc()
  call set_bound()
  call allocate_solvar()

  nwell_bt = 0
  inquire(file='bterms.bin', exist=have_bt)
  if (have_bt) call read_bterms()

  nconn = crs_index(1)%offind(ncalc)
  open(newunit=u, file='static.bin', access='stream', form='unformatted', status='replace')
  write(u) ncalc, ncals, nconn, rech_num
  write(u) real(len_scal, DP), real(len_scal_inv, DP), newper
  write(u) crs_index(1)%offind(0:ncalc)
  write(u) crs_index(1)%offrow(1:nconn)
  write(u) dir_conn(1:nconn)
  write(u) hydf_conn(1:nconn)
  write(u) abyd_conn(1:nconn)
  write(u) cell_vol(1:ncalc)
  write(u) cell_top(1:ncalc)
  write(u) read_reta(1:ncalc), read_retn(1:ncalc), read_resi(1:ncalc)
  write(u) read_poro(1:ncalc), read_ss(1:ncalc)
  write(u) hydf_surf(1:ncals), abyd_surf(1:ncals), surf_top(1:ncals), surf_bott(1:ncals)
  write(u) rech_area(1:ncals)
  close(u)

  allocate(dh(ncalc), f(ncalc), hsave(ncalc))

  ! timing mode (MSG123_TIME_NEWTON=k): the first step's calc_numsol capped at
  ! k Newton iterations, wall-clocked; set-up and input reading are excluded
  call get_environment_variable('MSG123_TIME_NEWTON', envv, status=i)
  if (i == 0) then
    read(envv, *) ktime
    call tstep()
    maxout_iter = ktime
    call system_clock(c0, crate)
    call calc_numsol()
    call system_clock(c1)
    open(newunit=u, file='timing.txt', status='replace')
    write(u, '(es16.8,1x,i0,1x,i0,1x,i0)') real(c1 - c0, DP)/real(crate, DP), out_iter,&
      in_iter, ncalc
    close(u)
    stop
  end if

  maxsteps = 0
  call get_environment_variable('MSG123_MAX_STEPS', envv, status=i)
  if (i == 0) read(envv, *) maxsteps

  open(newunit=u, file='steps.bin', access='stream', form='unformatted', status='replace')
  open(newunit=ut, file='steptime.txt', status='replace')
  nstep = 0
  dumped = .false.
  tstep_loop: do
    call tstep()
    ! the first attempt of the first step only (a failed first attempt is
    ! retried with dt/2, which would pass the test again)
    if (nstep == 0 .and. current_t == real(delt, SP) .and. .not. dumped) then
      call dump_first()
      dumped = .true.
    end if
    call system_clock(c0, crate)
    call calc_numsol()
    call system_clock(c1)
    write(ut, '(es16.8,1x,i0,1x,i0)') real(c1 - c0, DP)/real(crate, DP), out_iter, conv_flag
    if (conv_flag == 1) then
      nstep = nstep + 1
      write(u) real(current_t, DP), delt, out_iter
      write(u) head_new(1:ncalc)
      call check_lastts()
      call write_outf(now_time)
      if (lasttime_flag == 1) then
        exit tstep_loop
      end if
      if (maxsteps > 0 .and. nstep >= maxsteps) exit tstep_loop
    end if
  end do tstep_loop
  close(u)
  close(ut)

  ! AMG hierarchy (Make_amg_matrix.f90) of the Jacobian at the final state
  if (precon_type == 1) then
    call free_levels()
    call make_matvec()
    call dump_amg()
    call free_levels()
  end if

  ! fully saturated and ponded state: every head 2 cm above the top surface
  hsave(:) = head_new(1:ncalc)
  do i = 1, ncalc
    head_new(i) = surf_top(1) + 0.02_DP*real(len_scal_inv, DP)
  end do
  call make_matvec()
  open(newunit=u, file='wet.bin', access='stream', form='unformatted', status='replace')
  write(u) head_new(1:ncalc), surf_head(1:ncals)
  write(u) array_var(1)%rhs(1:ncalc), array_var(1)%dmat(1:ncalc), array_var(1)%lumat(1:nconn)
  close(u)
  head_new(1:ncalc) = hsave(:)

contains

  ! update_tstep with the injected well hidden from it (see the header)
  subroutine tstep()
    if (nwell_bt > 0) well_num = 0
    call update_tstep()
    if (nwell_bt > 0) well_num = 1
  end subroutine tstep

  subroutine read_bterms()
    integer(I4) :: v, nr, nl, nw, nse, nsv
    open(newunit=v, file='bterms.bin', access='stream', form='unformatted', status='old')
    read(v) nr
    if (allocated(rive2cals)) deallocate(rive2cals)
    if (allocated(rive_head)) deallocate(rive_head)
    if (allocated(rive_bott)) deallocate(rive_bott)
    if (allocated(abyd_rive)) deallocate(abyd_rive)
    allocate(rive2cals(nr), rive_head(nr), rive_bott(nr), abyd_rive(nr))
    read(v) rive2cals, rive_head, rive_bott, abyd_rive
    rive_num = nr
    read(v) nl
    if (allocated(lake2cals)) deallocate(lake2cals)
    if (allocated(lake_head)) deallocate(lake_head)
    if (allocated(lake_bott)) deallocate(lake_bott)
    if (allocated(abyd_lake)) deallocate(abyd_lake)
    allocate(lake2cals(nl), lake_head(nl), lake_bott(nl), abyd_lake(nl))
    read(v) lake2cals, lake_head, lake_bott, abyd_lake
    lake_num = nl
    read(v) nw
    if (allocated(well_index)) deallocate(well_index)
    if (allocated(well_conn)) deallocate(well_conn)
    if (allocated(calc_well)) deallocate(calc_well)
    allocate(well_index(0:1), well_conn(nw), calc_well(ncalc))
    read(v) well_conn, calc_well
    well_index(0) = 0 ; well_index(1) = nw
    nwell_bt = nw
    well_num = 0
    if (nw > 0) well_num = 1
    read(v) nse
    if (allocated(seal2calc)) deallocate(seal2calc)
    if (allocated(seal2seal)) deallocate(seal2seal)
    if (allocated(hydf_seal)) deallocate(hydf_seal)
    if (allocated(abyd_seal)) deallocate(abyd_seal)
    allocate(seal2calc(nse), seal2seal(nse), hydf_seal(nse), abyd_seal(nse))
    read(v) seal2calc, seal2seal, hydf_seal, abyd_seal
    nseal = nse
    read(v) nsv
    if (allocated(read_seal)) deallocate(read_seal)
    allocate(read_seal(nsv))
    read(v) read_seal
    close(v)
  end subroutine read_bterms

  ! levels >= 2 are rebuilt by every make_matvec; free them first, as
  ! reset_matvec does (Nonlinear_solution.f90:312-321)
  subroutine free_levels()
    integer(I4) :: l
    do l = 2, nlevel
      if (allocated(array_var(l)%dmat)) deallocate(array_var(l)%dmat)
      if (allocated(array_var(l)%rhs)) deallocate(array_var(l)%rhs)
      if (allocated(array_var(l)%lumat)) deallocate(array_var(l)%lumat)
      if (allocated(array_var(l)%x)) deallocate(array_var(l)%x)
      if (allocated(crs_index(l)%offrow)) deallocate(crs_index(l)%offrow)
      if (allocated(crs_index(l)%offind)) deallocate(crs_index(l)%offind)
      if (allocated(pro_var(l)%pindex)) deallocate(pro_var(l)%pindex)
      if (allocated(pro_var(l)%poffrow)) deallocate(pro_var(l)%poffrow)
      if (allocated(pro_var(l)%pval)) deallocate(pro_var(l)%pval)
      if (allocated(res_var(l)%rindex)) deallocate(res_var(l)%rindex)
      if (allocated(res_var(l)%roffrow)) deallocate(res_var(l)%roffrow)
      if (allocated(res_var(l)%rval)) deallocate(res_var(l)%rval)
    end do
  end subroutine free_levels

  subroutine dump_amg()
    integer(I4) :: v, l, nf, nc, nnz
    open(newunit=v, file='amg.bin', access='stream', form='unformatted', status='replace')
    write(v) nlevel
    write(v) head_new(1:ncalc)
    do l = 1, nlevel
      nc = crs_index(l)%unknow
      nnz = crs_index(l)%offind(nc)
      write(v) nc, nnz
      write(v) crs_index(l)%offind(0:nc)
      write(v) crs_index(l)%offrow(1:nnz)
      write(v) array_var(l)%dmat(1:nc)
      write(v) array_var(l)%lumat(1:nnz)
      if (l >= 2) then
        nf = crs_index(l-1)%unknow
        write(v) pro_var(l)%pindex(nf)
        write(v) pro_var(l)%pindex(0:nf)
        write(v) pro_var(l)%poffrow(1:pro_var(l)%pindex(nf))
        write(v) pro_var(l)%pval(1:pro_var(l)%pindex(nf))
        write(v) res_var(l)%rindex(nc)
        write(v) res_var(l)%rindex(0:nc)
        write(v) res_var(l)%roffrow(1:res_var(l)%rindex(nc))
        write(v) res_var(l)%rval(1:res_var(l)%rindex(nc))
      end if
    end do
    close(v)
  end subroutine dump_amg

  subroutine dump_first()
    integer(I4) :: v
    ! calc_numsol switches to BiCGStab and the full FD Jacobian before its first
    ! assembly (Nonlinear_solution.f90:79); do the same here
    form_switch = 1
    call make_matvec()
    call calc_l2norm2(1, array_var(1)%rhs, l2b)
    open(newunit=v, file='assem1.bin', access='stream', form='unformatted', status='replace')
    write(v) delt, calc_rech(1:rech_num)
    write(v) rech2cals(1:rech_num)
    write(v) head_new(1:ncalc), head_old(1:ncalc), srat_old(1:ncalc)
    write(v) array_var(1)%rhs(1:ncalc), array_var(1)%dmat(1:ncalc), array_var(1)%lumat(1:nconn)
    write(v) srat_new(1:ncalc), rel_perm(1:ncalc)
    ! first Newton iteration's linear solve: x0 = 0, errtol = eta = 0.5
    dh(:) = 0.0_DP
    errtol = 0.5_DP
    conv_flag = 0
    call solve_linalg(l2b, dh, lastn)
    write(v) in_iter, conv_flag
    write(v) l2b, lastn
    write(v) dh(1:ncalc)
    close(v)
    conv_flag = 0
  end subroutine dump_first

end program ref_harness
