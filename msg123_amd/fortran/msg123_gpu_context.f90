! msg123_gpu_context -- the device context the drop-in modules share.
!
! gpu_ctx() creates the context on first use from the reference's own set-up
! arrays (crs_index(1), dir_conn, hydf_conn, abyd_conn, cell_vol, cell_top,
! read_*, the surface coefficients), so every shim (nonlinear_solution,
! calc_function, make_linearsystem, linear_solution) works on the same
! device copy of the level-1 system.  gpu_load_level() hands the time level
! calc_numsol and its pieces read (head_new, head_old, srat_old, surf_head,
! surf_rati, the recharge volume per top cell, delt) to the device, and
! gpu_store_side() copies back what calc_func leaves in module state
! (srat_new, rel_perm, surf_rati).
!
! gpu_load_level() also hands over the river, lake, well and sea arrays the
! reference's set-up and time module keep (msg123_set_bterms), so their
! terms enter the device residual, Jacobian and mass balance.
!
! Scope: what the library supports -- a structured box of active cells
! (grid type 7), transient runs (sim_type 1), no MPI; other decks stop with
! a message.
module msg123_gpu_context
  use iso_c_binding, only: c_ptr, c_null_ptr, c_associated, c_loc, c_int32_t, c_double,&
                           c_float
  use kind_module, only: I4, DP
  use msg123_gpu_api
  implicit none
  private
  public :: gpu_ctx, gpu_load_level, gpu_store_side, gpu_check

  type(c_ptr), save :: ctx = c_null_ptr
  type(msg123_mesh), save, target :: mesh
  ! host copies the mesh descriptor points to (kept for the run)
  integer(c_int32_t), allocatable, save, target :: m_offind(:), m_offrow(:), m_dir(:)
  real(c_double), allocatable, save, target :: m_hydf(:), m_abyd(:), m_vol(:), m_top(:)
  real(c_float), allocatable, save, target :: m_reta(:), m_retn(:), m_resi(:), m_poro(:), m_ss(:)
  real(c_double), allocatable, save, target :: m_hsurf(:), m_asurf(:), m_stop(:), m_sbott(:)
  ! host copies of the boundary arrays handed to msg123_set_bterms
  integer(c_int32_t), allocatable, save, target :: b_rcell(:), b_lcell(:), b_wconn(:),&
                                                   b_scell(:), b_ssea(:)
  real(c_double), allocatable, save, target :: b_rhead(:), b_rbott(:), b_rabyd(:), b_lhead(:),&
                                               b_lbott(:), b_labyd(:), b_cwell(:), b_shydf(:),&
                                               b_sabyd(:)
  real(c_float), allocatable, save, target :: b_sread(:)
  logical, save :: b_set = .false.

contains

  type(c_ptr) function gpu_ctx()
    if (.not. c_associated(ctx)) call open_context()
    gpu_ctx = ctx
  end function gpu_ctx

  subroutine gpu_load_level()
    use prep_calculation, only: delt
    use set_cell, only: ncals
    use allocate_solution, only: head_new, head_old, srat_old, surf_head, surf_rati
    use calc_boundary, only: calc_rech, rech2cals
    real(DP), allocatable :: rech_top(:)
    integer(I4) :: i
    ! recharge volume per top cell (calc_boundary keeps it per recharge cell)
    allocate(rech_top(ncals))
    rech_top(:) = 0.0_DP
    do i = 1, size(calc_rech)
      rech_top(rech2cals(i)) = rech_top(rech2cals(i)) + calc_rech(i)
    end do
    call gpu_check(msg123_set_state(gpu_ctx(), head_new, head_old, srat_old, surf_head,&
                                    surf_rati, rech_top, real(delt, c_double)),&
                   'msg123_set_state')
    call gpu_load_bterms()
  end subroutine gpu_load_level

  ! rivers (rive2cals, rive_head, rive_bott, abyd_rive), lakes, the wells'
  ! connections with calc_well, and the sea connections with read_seal, as
  ! the time module left them for this step
  subroutine gpu_load_bterms()
    use set_cell, only: ncalc
    use set_boundary, only: rive_num, lake_num, well_num, abyd_rive, abyd_lake
    use calc_boundary, only: rive2cals, rive_head, rive_bott, lake2cals, lake_head, lake_bott
    use set_condition, only: well_index, well_conn, nseal
    use assign_boundary, only: calc_well, read_seal
    use allocate_solution, only: seal2calc, seal2seal, hydf_seal, abyd_seal
    type(msg123_bterms) :: bt
    integer(I4) :: nw
    nw = 0
    if (well_num > 0) nw = well_index(well_num)
    if (rive_num + lake_num + nw + nseal == 0) then
      if (b_set) call gpu_check(msg123_set_bterms(gpu_ctx(), bt), 'msg123_set_bterms')
      b_set = .false.
      return
    end if
    if (rive_num > 0) then
      b_rcell = rive2cals(1:rive_num) ; b_rhead = rive_head(1:rive_num)
      b_rbott = rive_bott(1:rive_num) ; b_rabyd = abyd_rive(1:rive_num)
      bt%nrive = rive_num
      bt%rive2cals = c_loc(b_rcell) ; bt%rive_head = c_loc(b_rhead)
      bt%rive_bott = c_loc(b_rbott) ; bt%abyd_rive = c_loc(b_rabyd)
    end if
    if (lake_num > 0) then
      b_lcell = lake2cals(1:lake_num) ; b_lhead = lake_head(1:lake_num)
      b_lbott = lake_bott(1:lake_num) ; b_labyd = abyd_lake(1:lake_num)
      bt%nlake = lake_num
      bt%lake2cals = c_loc(b_lcell) ; bt%lake_head = c_loc(b_lhead)
      bt%lake_bott = c_loc(b_lbott) ; bt%abyd_lake = c_loc(b_labyd)
    end if
    if (nw > 0) then
      b_wconn = well_conn(1:nw) ; b_cwell = calc_well(1:ncalc)
      bt%nwell_conn = nw
      bt%well_conn = c_loc(b_wconn) ; bt%calc_well = c_loc(b_cwell)
    end if
    if (nseal > 0) then
      b_scell = seal2calc(1:nseal) ; b_ssea = seal2seal(1:nseal)
      b_shydf = hydf_seal(1:nseal) ; b_sabyd = abyd_seal(1:nseal) ; b_sread = read_seal(:)
      bt%nseal = nseal ; bt%nseal_levels = size(read_seal)
      bt%seal2calc = c_loc(b_scell) ; bt%seal2seal = c_loc(b_ssea)
      bt%hydf_seal = c_loc(b_shydf) ; bt%abyd_seal = c_loc(b_sabyd)
      bt%read_seal = c_loc(b_sread)
    end if
    call gpu_check(msg123_set_bterms(gpu_ctx(), bt), 'msg123_set_bterms')
    b_set = .true.
  end subroutine gpu_load_bterms

  subroutine gpu_store_side()
    use set_cell, only: ncalc, ncals
    use allocate_solution, only: srat_new, rel_perm, surf_rati
    real(DP), allocatable :: h(:), sh(:)
    allocate(h(ncalc), sh(ncals))
    call gpu_check(msg123_get_state(gpu_ctx(), h, srat_new, rel_perm, sh, surf_rati),&
                   'msg123_get_state')
  end subroutine gpu_store_side

  subroutine open_context()
    use initial_module, only: st_sim, st_grid, maxout_iter, maxinn_iter, criteria, precon_type,&
                              newper
    use read_input, only: len_scal
    use set_cell, only: ncalc, ncals
    use allocate_solution, only: crs_index, dir_conn, hydf_conn, abyd_conn
    use make_cell, only: cell_vol, cell_top
    use assign_calc, only: read_reta, read_retn, read_resi, read_poro, read_ss, surf_bott
    use set_condition, only: hydf_surf, abyd_surf
    use prep_calculation, only: surf_top
    type(msg123_params) :: prm
    integer(I4) :: nconn

    if (st_sim%sim_type /= 1) stop 'GPU drop-in: only transient runs (sim_type 1)'
    nconn = crs_index(1)%offind(ncalc)
    if (ncalc /= st_grid%nx*st_grid%ny*st_grid%nz .or. ncals /= st_grid%nx*st_grid%ny)&
      stop 'GPU drop-in: the deck must be a full structured box'
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
    call gpu_check(msg123_ctx_create(mesh, prm, ctx), 'msg123_ctx_create')
  end subroutine open_context

  subroutine gpu_check(rc, what)
    integer(c_int32_t), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= MSG123_OK) then
      write(*, '(a,a,i0)') what, ' failed with code ', rc
      stop 'GPU drop-in: library error'
    end if
  end subroutine gpu_check

end module msg123_gpu_context
