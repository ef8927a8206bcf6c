!
! rbmpp_mod -- Fortran binding of include/rbmpp.h (iso_c_binding).
!
! A Fortran driver in the shape of the reference's main.f90 uses this module
! in place of its module-global sampler / estimator / SR-step routines:
!
!   update_slater_elements()            (main.f90:318)     -> rbmpp_set_params / rbmpp_set_gamma
!   make_sample(myrank,first,naccept)   (main.f90:323-327) -> rbmpp_make_sample
!   calc_phys_quantity_opt(...) + the
!   reductions of main.f90:347-367      (main.f90:333)     -> rbmpp_measure
!   include "SR_pdposv.f90"             (main.f90:375)     -> rbmpp_sr_step
!   calc_phys_quantity_after_opt(...)   (main.f90:518)     -> rbmpp_measure_energy
!   Calc_Correlation calc_phys_quantity_after_opt(...)     -> rbmpp_measure_correlation
!
! Arrays are passed in the reference's own layouts (column-major, 1-based
! index values in the tables), so the reference's arrays go through as they
! are.  Every function returns 0 or a negative RBMPP_E_* status; the message
! of the last failure is rbmpp_error_message(h).
!
module rbmpp_mod
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: RBMPP_OK = 0
  integer(c_int), parameter, public :: RBMPP_S_FP64 = 0, RBMPP_S_FP32 = 1, RBMPP_S_BF16X3 = 2
  integer(c_int), parameter, public :: RBMPP_PROJ_NORMAL = 0, RBMPP_PROJ_TRANSVERSE = 1, &
                                        RBMPP_PROJ_NORMAL_EXACT = 2
  integer(c_int), parameter, public :: RBMPP_SR_NO_GRAPH = 1
  integer(c_int), parameter, public :: RBMPP_SR_ASSEMBLED_REFINE = 2

  ! struct rbmpp_config (field order and kinds as in rbmpp.h)
  type, bind(C), public :: rbmpp_config
    integer(c_int32_t) :: L, n_sym, alphac, alphar, n_slater, spin_parity
    real(c_double)     :: J1, J2
    integer(c_int32_t) :: n_walkers, n_samples, n_warmup, n_warmup_first, n_update
    integer(c_int32_t) :: recompute_k, s_precision, fused_measure
    integer(c_int64_t) :: seed
    real(c_double)     :: p_restart, p_global, cond_guard
  end type

  type, bind(C), public :: rbmpp_sr_config
    real(c_double)     :: s_factor0, delta_tau, max_change, slater_abs_av
    integer(c_int32_t) :: refine_iters, flags
  end type

  type, bind(C), public :: rbmpp_tune_report
    integer(c_int32_t) :: s_tile, s_nsplit, s_default_tile, s_default_nsplit
    real(c_double)     :: s_build_s
    integer(c_int32_t) :: measure_grid, measure_configurations
    real(c_double)     :: measure_s
  end type

  integer(c_int), parameter, public :: RBMPP_TUNE_S_BUILD = 1, RBMPP_TUNE_MEASURE = 2

  type, bind(C), public :: rbmpp_sr_report
    real(c_double)     :: lambda, delw_max, scaling, sdiag_av, s_factor, residual, mean_abs_f
    integer(c_int32_t) :: k_nonzero, cg_iters
    real(c_double) :: precond_boost
  end type

  public :: rbmpp_create, rbmpp_destroy, rbmpp_last_error, rbmpp_nv
  public :: rbmpp_set_params, rbmpp_set_gamma, rbmpp_get_gamma, rbmpp_set_projection
  public :: rbmpp_make_sample, rbmpp_measure, rbmpp_sr_step, rbmpp_set_s_precision
  public :: rbmpp_measure_energy, rbmpp_measure_correlation
  public :: rbmpp_set_samples, rbmpp_get_samples, rbmpp_n_samples
  public :: rbmpp_get_walkers, rbmpp_get_walker_electrons, rbmpp_eval, rbmpp_ratio_exchange, rbmpp_local
  public :: rbmpp_get_smat, rbmpp_timings, rbmpp_get_measured, rbmpp_set_measured, rbmpp_local_noalloc
  public :: rbmpp_tune
  public :: rbmpp_batch_create, rbmpp_batch_make_sample, rbmpp_batch_sample_seconds
  public :: rbmpp_batch_last_error, rbmpp_batch_destroy
  public :: rbmpp_error_message

  interface
    integer(c_int) function rbmpp_create(cfg, w_map, para_mp_trans, orbital_idx, h) &
        bind(C, name="rbmpp_create")
      import :: c_int, c_int32_t, c_double, c_ptr, rbmpp_config
      type(rbmpp_config), intent(in) :: cfg
      integer(c_int32_t), intent(in) :: w_map(*), orbital_idx(*)
      real(c_double), intent(in) :: para_mp_trans(*)
      type(c_ptr), intent(out) :: h
    end function

    subroutine rbmpp_destroy(h) bind(C, name="rbmpp_destroy")
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine

    type(c_ptr) function rbmpp_last_error(h) bind(C, name="rbmpp_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function rbmpp_nv(h) bind(C, name="rbmpp_nv")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    ! complex arrays (Wcirr, bcirr) are passed as interleaved re/im reals, the
    ! memory layout of complex(8)
    integer(c_int) function rbmpp_set_params(h, wcirr, bcirr, wrirr, brirr, slater) &
        bind(C, name="rbmpp_set_params")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(in) :: wcirr(*), bcirr(*), wrirr(*), brirr(*), slater(*)
    end function

    integer(c_int) function rbmpp_set_gamma(h, gamma) bind(C, name="rbmpp_set_gamma")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(in) :: gamma(*)
    end function

    integer(c_int) function rbmpp_get_gamma(h, gamma) bind(C, name="rbmpp_get_gamma")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(out) :: gamma(*)
    end function

    integer(c_int) function rbmpp_make_sample(h, first_iteration, shift, naccept) &
        bind(C, name="rbmpp_make_sample")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: first_iteration
      real(c_double), value :: shift
      integer(c_int64_t), intent(out) :: naccept
    end function

    integer(c_int) function rbmpp_measure(h, shift, etot, etot2, norm, wf_av, gvec, ovec) &
        bind(C, name="rbmpp_measure")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), value :: shift
      real(c_double), intent(out) :: etot, etot2, norm, wf_av
      real(c_double), intent(out) :: gvec(*), ovec(*)
    end function

    integer(c_int) function rbmpp_sr_step(h, c, gamma, rep) bind(C, name="rbmpp_sr_step")
      import :: c_int, c_double, c_ptr, rbmpp_sr_config, rbmpp_sr_report
      type(c_ptr), value :: h
      type(rbmpp_sr_config), intent(in) :: c
      real(c_double), intent(inout) :: gamma(*)
      type(rbmpp_sr_report), intent(out) :: rep
    end function

    integer(c_int) function rbmpp_set_projection(h, projection) bind(C, name="rbmpp_set_projection")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: projection
    end function

    integer(c_int) function rbmpp_set_s_precision(h, prec) bind(C, name="rbmpp_set_s_precision")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: prec
    end function

    integer(c_int) function rbmpp_measure_energy(h, shift, etot, etot2, norm, wf_av) &
        bind(C, name="rbmpp_measure_energy")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), value :: shift
      real(c_double), intent(out) :: etot, etot2, norm, wf_av
    end function

    ! Calc_Correlation: chi_sp(2,Nchi), chi_vbs(Nchi,4), wf_sgn(2) = (re, im)
    integer(c_int) function rbmpp_measure_correlation(h, shift, etot, etot2, norm, wf_av, wf_sgn, &
        chi_sp, chi_vbs) bind(C, name="rbmpp_measure_correlation")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), value :: shift
      real(c_double), intent(out) :: etot, etot2, norm, wf_av, wf_sgn(2), chi_sp(*), chi_vbs(*)
    end function

    integer(c_int) function rbmpp_set_samples(h, n, x) bind(C, name="rbmpp_set_samples")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: x(*)
    end function

    integer(c_int) function rbmpp_get_samples(h, x, psi) bind(C, name="rbmpp_get_samples")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: x(*)
      real(c_double), intent(out) :: psi(*)
    end function

    integer(c_int64_t) function rbmpp_n_samples(h) bind(C, name="rbmpp_n_samples")
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function rbmpp_get_measured(h, psi, eloc, rw) bind(C, name="rbmpp_get_measured")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(out) :: psi(*), eloc(*), rw(*)
    end function

    integer(c_int) function rbmpp_get_walkers(h, x, pf, psi, invm, naccept) &
        bind(C, name="rbmpp_get_walkers")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: x(*)
      real(c_double), intent(out) :: pf(*), psi(*), invm(*)
      integer(c_int64_t), intent(out) :: naccept(*)
    end function

    integer(c_int) function rbmpp_get_walker_electrons(h, ele_idx) &
        bind(C, name="rbmpp_get_walker_electrons")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: ele_idx(*)
    end function

    integer(c_int) function rbmpp_tune(h, what, rep) bind(C, name="rbmpp_tune")
      import :: c_int, c_int32_t, c_ptr, rbmpp_tune_report
      type(c_ptr), value :: h
      integer(c_int32_t), value :: what
      type(rbmpp_tune_report), intent(out) :: rep
    end function

    ! several handles sampled by one launch (handles: array of c_ptr)
    integer(c_int) function rbmpp_batch_create(handles, n, b) bind(C, name="rbmpp_batch_create")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), intent(in) :: handles(*)
      integer(c_int32_t), value :: n
      type(c_ptr), intent(out) :: b
    end function

    integer(c_int) function rbmpp_batch_make_sample(b, first_iteration, shift, naccept) &
        bind(C, name="rbmpp_batch_make_sample")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: b
      integer(c_int32_t), intent(in) :: first_iteration(*)
      real(c_double), intent(in) :: shift(*)
      integer(c_int64_t), intent(out) :: naccept(*)
    end function

    real(c_double) function rbmpp_batch_sample_seconds(b) bind(C, name="rbmpp_batch_sample_seconds")
      import :: c_double, c_ptr
      type(c_ptr), value :: b
    end function

    type(c_ptr) function rbmpp_batch_last_error(b) bind(C, name="rbmpp_batch_last_error")
      import :: c_ptr
      type(c_ptr), value :: b
    end function

    subroutine rbmpp_batch_destroy(b) bind(C, name="rbmpp_batch_destroy")
      import :: c_ptr
      type(c_ptr), value :: b
    end subroutine

    ! a measurement loaded for rbmpp_sr_step (orows = c_null_ptr: synthetic)
    integer(c_int) function rbmpp_set_measured(h, n, psi, eloc, rw, orows, seed) &
        bind(C, name="rbmpp_set_measured")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: h, psi, eloc, rw, orows
      integer(c_int64_t), value :: n, seed
    end function

    ! rbmpp_local without O, staging in a fixed device array (no allocation)
    integer(c_int) function rbmpp_local_noalloc(h, n, x, eloc, psi) bind(C, name="rbmpp_local_noalloc")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: x(*)
      real(c_double), intent(out) :: eloc(*), psi(*)
    end function

    integer(c_int) function rbmpp_eval(h, n, x, psi, pf, invm) bind(C, name="rbmpp_eval")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: x(*)
      real(c_double), intent(out) :: psi(*), pf(*), invm(*)
    end function

    integer(c_int) function rbmpp_ratio_exchange(h, n, x, i1, i2, pf_new, psi_new) &
        bind(C, name="rbmpp_ratio_exchange")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: x(*), i1(*), i2(*)
      real(c_double), intent(out) :: pf_new(*), psi_new(*)
    end function

    integer(c_int) function rbmpp_local(h, n, x, eloc, psi, oloc) bind(C, name="rbmpp_local")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: x(*)
      real(c_double), intent(out) :: eloc(*), psi(*), oloc(*)
    end function

    integer(c_int) function rbmpp_get_smat(h, s) bind(C, name="rbmpp_get_smat")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(out) :: s(*)
    end function

    integer(c_int) function rbmpp_timings(h, t) bind(C, name="rbmpp_timings")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(out) :: t(*)
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! message of the last failure on h, as a Fortran string
  function rbmpp_error_message(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = rbmpp_last_error(h)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

end module rbmpp_mod
