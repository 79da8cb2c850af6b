! msg123_gpu_api -- iso_c_binding interface of libmsg123_gpu.so
! (include/msg123_gpu.h).  Derived types mirror the C structs member for
! member; every function returns the library's int code (0 = ok).
module msg123_gpu_api
  use iso_c_binding, only: c_int32_t, c_double, c_float, c_ptr, c_null_ptr, c_char
  implicit none
  private

  integer(c_int32_t), parameter, public :: MSG123_OK = 0, MSG123_EINVAL = -1,&
                                           MSG123_EDEVICE = -2, MSG123_ENAN = -3,&
                                           MSG123_ENOMEM = -4
  integer(c_int32_t), parameter, public :: MSG123_PRECON_DILU = 0, MSG123_PRECON_LINE32 = 1,&
                                           MSG123_PRECON_AMG = 2, MSG123_PRECON_AMG32 = 3,&
                                           MSG123_PRECON_MCILU32 = 4
  integer(c_int32_t), parameter, public :: MSG123_LAW_REFERENCE = 0, MSG123_LAW_LOGEXP = 1
  integer(c_int32_t), parameter, public :: MSG123_END_CONVERGED = 0, MSG123_END_BACKTRACK = 1,&
                                           MSG123_END_MAXITER = 2, MSG123_END_DIVERGED = 3,&
                                           MSG123_END_NAN = 4

  type, bind(C), public :: msg123_mesh
    integer(c_int32_t) :: nx, ny, nz, ncalc, ncals, nconn
    type(c_ptr) :: offind = c_null_ptr, offrow = c_null_ptr, dir_conn = c_null_ptr
    type(c_ptr) :: hydf_conn = c_null_ptr, abyd_conn = c_null_ptr
    type(c_ptr) :: cell_vol = c_null_ptr, cell_top = c_null_ptr
    type(c_ptr) :: reta = c_null_ptr, retn = c_null_ptr, resi = c_null_ptr
    type(c_ptr) :: poro = c_null_ptr, ss = c_null_ptr
    type(c_ptr) :: hydf_surf = c_null_ptr, abyd_surf = c_null_ptr
    type(c_ptr) :: surf_top = c_null_ptr, surf_bott = c_null_ptr
    real(c_double) :: len_scal, newper
  end type msg123_mesh

  type, bind(C), public :: msg123_params
    integer(c_int32_t) :: maxout_iter, maxinn_iter
    real(c_double) :: criteria
    integer(c_int32_t) :: precon, consistent_jacobian
  end type msg123_params

  type, bind(C), public :: msg123_newton_stats
    integer(c_int32_t) :: out_iter, conv_flag, in_iter_last, in_iter_total
    integer(c_int32_t) :: back_iter_total, func_evals
    real(c_double) :: max_change, l2_new
  end type msg123_newton_stats

  type, bind(C), public :: msg123_iter_rec
    integer(c_int32_t) :: out_iter, in_iter, back_iter, cell_i, cell_j, cell_k
    real(c_double) :: max_change, dmat, rhs, head
  end type msg123_iter_rec

  type, bind(C), public :: msg123_amg_level
    integer(c_int32_t) :: n, nnz, nfine
    type(c_ptr) :: offind, offrow, dmat, lumat
    type(c_ptr) :: pindex, poffrow, rindex, roffrow, pval, rval
  end type msg123_amg_level

  type, bind(C), public :: msg123_amg
    integer(c_int32_t) :: nlevel
    type(c_ptr) :: levels
  end type msg123_amg

  type, bind(C), public :: msg123_amg_opts
    integer(c_int32_t) :: max_levels = 12
    real(c_float) :: theta = 0.05, omega = 0.67, smooth_omega = 0.67
    integer(c_int32_t) :: npre = 2, npost = 2, coarse_max = 64, coarse_sweeps = 20
    integer(c_int32_t) :: complete_orphans = 1, project = 1
  end type msg123_amg_opts

  type, bind(C), public :: msg123_mass_balance
    real(c_double) :: storage, connection, recharge, surface, total
    real(c_double) :: residual, recharge_abs
    integer(c_int32_t) :: steps
    real(c_double) :: well, river, lake, sea
  end type msg123_mass_balance

  ! river, lake, well and sea arrays of the reference's module state
  ! (1-based cells), for msg123_set_bterms
  type, bind(C), public :: msg123_bterms
    integer(c_int32_t) :: nrive = 0
    type(c_ptr) :: rive2cals = c_null_ptr, rive_head = c_null_ptr, rive_bott = c_null_ptr,&
                   abyd_rive = c_null_ptr
    integer(c_int32_t) :: nlake = 0
    type(c_ptr) :: lake2cals = c_null_ptr, lake_head = c_null_ptr, lake_bott = c_null_ptr,&
                   abyd_lake = c_null_ptr
    integer(c_int32_t) :: nwell_conn = 0
    type(c_ptr) :: well_conn = c_null_ptr, calc_well = c_null_ptr
    integer(c_int32_t) :: nseal = 0
    type(c_ptr) :: seal2calc = c_null_ptr, seal2seal = c_null_ptr, hydf_seal = c_null_ptr,&
                   abyd_seal = c_null_ptr
    integer(c_int32_t) :: nseal_levels = 0
    type(c_ptr) :: read_seal = c_null_ptr
  end type msg123_bterms

  public :: msg123_ctx_create, msg123_ctx_destroy, msg123_set_state, msg123_get_state
  public :: msg123_calc_srat_rperm, msg123_calc_func, msg123_make_matvec, msg123_solve_linalg
  public :: msg123_load_system, msg123_calc_numsol, msg123_iter_log, msg123_amg_setup
  public :: msg123_amg_free, msg123_set_newton, msg123_set_law, msg123_set_amg_opts
  public :: msg123_set_bottom_drain, msg123_set_bterms, msg123_calc_mass_bterms
  public :: msg123_amg_stats, msg123_calc_mass, msg123_set_mass_balance, msg123_mass_balance_get
  public :: msg123_write_head, msg123_write_restart, msg123_read_restart
  public :: msg123_calc_vecjacf

  interface
    integer(c_int32_t) function msg123_ctx_create(mesh, params, ctx) bind(C)
      import :: c_int32_t, c_ptr, msg123_mesh, msg123_params
      type(msg123_mesh), intent(in) :: mesh
      type(msg123_params), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
    end function msg123_ctx_create

    subroutine msg123_ctx_destroy(ctx) bind(C)
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine msg123_ctx_destroy

    integer(c_int32_t) function msg123_set_state(ctx, head_new, head_old, srat_old, surf_head,&
                                                 surf_rati, calc_rech, delt) bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: head_new(*), head_old(*), srat_old(*), surf_head(*)
      real(c_double), intent(in) :: surf_rati(*), calc_rech(*)
      real(c_double), value :: delt
    end function msg123_set_state

    integer(c_int32_t) function msg123_get_state(ctx, head_new, srat_new, rel_perm, surf_head,&
                                                 surf_rati) bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: head_new(*), srat_new(*), rel_perm(*), surf_head(*)
      real(c_double), intent(out) :: surf_rati(*)
    end function msg123_get_state

    integer(c_int32_t) function msg123_calc_srat_rperm(ctx, pertur, pres, srat, rperm, sstor)&
                                                       bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: pertur
      real(c_double), intent(in) :: pres(*)
      real(c_double), intent(out) :: srat(*), rperm(*), sstor(*)
    end function msg123_calc_srat_rperm

    integer(c_int32_t) function msg123_calc_func(ctx, h, funcv) bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: h(*)
      real(c_double), intent(out) :: funcv(*)
    end function msg123_calc_func

    integer(c_int32_t) function msg123_calc_vecjacf(ctx, x, v, jv) bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: x(*), v(*)
      real(c_double), intent(out) :: jv(*)
    end function msg123_calc_vecjacf

    integer(c_int32_t) function msg123_make_matvec(ctx, rhs, dmat, lumat) bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rhs(*), dmat(*), lumat(*)
    end function msg123_make_matvec

    integer(c_int32_t) function msg123_solve_linalg(ctx, init_norm, errtol, dh, last_norm,&
                                                    in_iter, conv_flag) bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: init_norm, errtol
      real(c_double), intent(inout) :: dh(*)
      real(c_double), intent(out) :: last_norm
      integer(c_int32_t), intent(out) :: in_iter, conv_flag
    end function msg123_solve_linalg

    integer(c_int32_t) function msg123_load_system(ctx, rhs, dmat, lumat) bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rhs(*), dmat(*), lumat(*)
    end function msg123_load_system

    integer(c_int32_t) function msg123_calc_numsol(ctx, stats) bind(C)
      import :: c_int32_t, c_ptr, msg123_newton_stats
      type(c_ptr), value :: ctx
      type(msg123_newton_stats), intent(out) :: stats
    end function msg123_calc_numsol

    integer(c_int32_t) function msg123_iter_log(ctx, recs, cap, count, end_reason) bind(C)
      import :: c_int32_t, c_ptr, msg123_iter_rec
      type(c_ptr), value :: ctx
      type(msg123_iter_rec), intent(out) :: recs(*)
      integer(c_int32_t), value :: cap
      integer(c_int32_t), intent(out) :: count, end_reason
    end function msg123_iter_log

    integer(c_int32_t) function msg123_amg_setup(offind, offrow, dmat, lumat, n, amg_nlevel,&
                                                 amg_theta, jac_omega, amg) bind(C)
      import :: c_int32_t, c_double, c_float, msg123_amg
      integer(c_int32_t), intent(in) :: offind(*), offrow(*)
      real(c_double), intent(in) :: dmat(*), lumat(*)
      integer(c_int32_t), value :: n, amg_nlevel
      real(c_float), value :: amg_theta, jac_omega
      type(msg123_amg), intent(out) :: amg
    end function msg123_amg_setup

    subroutine msg123_amg_free(amg) bind(C)
      import :: msg123_amg
      type(msg123_amg), intent(inout) :: amg
    end subroutine msg123_amg_free

    integer(c_int32_t) function msg123_set_newton(ctx, eta_first, eta_max, algebraic_slope,&
                                                  beta_refine) bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: eta_first, eta_max
      integer(c_int32_t), value :: algebraic_slope, beta_refine
    end function msg123_set_newton

    integer(c_int32_t) function msg123_set_law(ctx, law) bind(C)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: law
    end function msg123_set_law

    ! cond(nx*ny): free-drainage conductance per column; c_null_ptr closes the bottom
    integer(c_int32_t) function msg123_set_bottom_drain(ctx, cond) bind(C)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: cond
    end function msg123_set_bottom_drain

    integer(c_int32_t) function msg123_set_bterms(ctx, bt) bind(C)
      import :: c_int32_t, c_ptr, msg123_bterms
      type(c_ptr), value :: ctx
      type(msg123_bterms), intent(in) :: bt
    end function msg123_set_bterms

    integer(c_int32_t) function msg123_calc_mass_bterms(ctx, out) bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(4)
    end function msg123_calc_mass_bterms

    ! file names end with c_null_char: trim(name)//c_null_char
    integer(c_int32_t) function msg123_write_head(ctx, path, append, time) bind(C)
      import :: c_int32_t, c_ptr, c_char, c_float
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: append
      real(c_float), value :: time
    end function msg123_write_head

    integer(c_int32_t) function msg123_write_restart(ctx, path, time) bind(C)
      import :: c_int32_t, c_ptr, c_char, c_float
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
      real(c_float), value :: time
    end function msg123_write_restart

    integer(c_int32_t) function msg123_read_restart(path, ncalc, len_scal, heads, time) bind(C)
      import :: c_int32_t, c_char, c_double, c_float
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: ncalc
      real(c_double), value :: len_scal
      real(c_double), intent(out) :: heads(*)
      real(c_float), intent(out) :: time
    end function msg123_read_restart

    integer(c_int32_t) function msg123_set_amg_opts(ctx, opts) bind(C)
      import :: c_int32_t, c_ptr, msg123_amg_opts
      type(c_ptr), value :: ctx
      type(msg123_amg_opts), intent(in) :: opts
    end function msg123_set_amg_opts

    integer(c_int32_t) function msg123_amg_stats(ctx, setups, fallbacks) bind(C)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: setups, fallbacks
    end function msg123_amg_stats

    integer(c_int32_t) function msg123_calc_mass(ctx, out) bind(C)
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(6)
    end function msg123_calc_mass

    integer(c_int32_t) function msg123_set_mass_balance(ctx, enable) bind(C)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
    end function msg123_set_mass_balance

    integer(c_int32_t) function msg123_mass_balance_get(ctx, mb) bind(C)
      import :: c_int32_t, c_ptr, msg123_mass_balance
      type(c_ptr), value :: ctx
      type(msg123_mass_balance), intent(out) :: mb
    end function msg123_mass_balance_get
  end interface

end module msg123_gpu_api
