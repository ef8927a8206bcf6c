!
! C-callable harness around the reference's own routines, compiled from the
! reference sources where they lie (see oracle/Makefile).  TEST
! INFRASTRUCTURE ONLY: it sets the reference's module state the way main.f90
! does (main.f90:85-141,170-202,239-243) and calls the reference subroutines
! on caller-given inputs, so tests can compare the CPU oracle with the
! reference on identical data.  calculate_MAll (matrix.f90) needs the PFAPACK
! binary and is not built; Pf and InvM are passed in from outside instead.
!
module ref_harness
  use iso_c_binding
  use mod_main
  use mod_RBM
  use mod_PP
  implicit none
contains

  subroutine ref_init(L_, P, ac, ar, nsl, J1_, J2_, sp, wmap, chi, orb) bind(C, name="ref_init")
    integer(c_int), value :: L_, P, ac, ar, nsl, sp
    real(c_double), value :: J1_, J2_
    integer(c_int), intent(in) :: wmap(*), orb(*)
    real(c_double), intent(in) :: chi(*)
    integer :: g, i, j
    L = L_; N = L*L; N_MP_Trans = P; alphac = ac; alphar = ar; N_Slater = nsl
    J1 = J1_; J2 = J2_; spin_parity = sp
    N_SP_Gauss_Leg = 2
    Mc = alphac * N_MP_Trans * Ncopy * 2
    Mr = alphar * N_MP_Trans * Ncopy * 2
    Nv = alphac*2*(N+1) + alphar*(N+1) + N_Slater
    Nv_RBM = Nv - N_Slater
    Nupdate = N
    N_QP_Full = N_SP_Gauss_Leg * N_MP_Trans
    Npair = N/2
    wf_av = 1d0
    shift = 0d0
    if (allocated(W_map)) then
      deallocate(W_map, para_mp_trans, para_spin_proj, Wcirr, Wrirr, bcirr, brirr)
      deallocate(Orbital_Idx, Orbital_Sgn, MP_Trans_Idx, MP_Trans_Sgn, Slater, Slater_Elm, subB_or_not)
      deallocate(SP_GL_Cos, SP_GL_Sin, SP_GL_CosCos, SP_GL_CosSin, SP_GL_SinSin)
    end if
    allocate(W_map(N,Ncopy,N_MP_Trans), para_mp_trans(N_MP_Trans))
    allocate(para_spin_proj(N_SP_Gauss_Leg))
    allocate(Wcirr(N,alphac), Wrirr(N,alphar), bcirr(alphac), brirr(alphar))
    allocate(Orbital_Idx(N,N), Orbital_Sgn(N,N), MP_Trans_Idx(N,N_MP_Trans), MP_Trans_Sgn(N,N_MP_Trans))
    allocate(Slater(N_Slater), Slater_Elm(2*N,2*N,N_QP_Full), subB_or_not(N))
    Wcirr = 0d0; Wrirr = 0d0; bcirr = 0d0; brirr = 0d0
    Orbital_Sgn = 1; MP_Trans_Sgn = 1; Slater_Elm = 0d0
    do g = 1, P
      para_mp_trans(g) = dcmplx(chi(2*g-1), chi(2*g))
      do i = 1, N
        W_map(i,1,g) = wmap((g-1)*N + i) + 1
        MP_Trans_Idx(i,g) = W_map(i,1,g)
      end do
    end do
    do i = 1, N
      do j = 1, N
        Orbital_Idx(i,j) = orb((i-1)*N + j) + 1
      end do
    end do
    call InitSPWeight(1)
    call define_sublattice()
#ifdef REF_CC
    Nchi = N*(N+1)/2
#endif
  end subroutine

  subroutine ref_set_params(wc, bc, wr, br, f) bind(C, name="ref_set_params")
    real(c_double), intent(in) :: wc(*), bc(*), wr(*), br(*), f(*)
    integer :: ff, iw
    do ff = 1, alphac
      bcirr(ff) = dcmplx(bc(2*ff-1), bc(2*ff))
      do iw = 1, N
        Wcirr(iw,ff) = dcmplx(wc(2*((ff-1)*N+iw)-1), wc(2*((ff-1)*N+iw)))
      end do
    end do
    do ff = 1, alphar
      brirr(ff) = br(ff)
      do iw = 1, N
        Wrirr(iw,ff) = wr((ff-1)*N + iw)
      end do
    end do
    Slater(1:N_Slater) = f(1:N_Slater)
    call update_slater_elements()
  end subroutine

  subroutine ref_get_slater_elm(out) bind(C, name="ref_get_slater_elm")
    real(c_double), intent(out) :: out(2*N,2*N,N_QP_Full)
    out = Slater_Elm
  end subroutine

  subroutine ref_get_sp(out) bind(C, name="ref_get_sp")
    real(c_double), intent(out) :: out(6,2)
    integer :: s
    do s = 1, 2
      out(1,s) = SP_GL_Cos(s); out(2,s) = SP_GL_Sin(s); out(3,s) = SP_GL_CosCos(s)
      out(4,s) = SP_GL_CosSin(s); out(5,s) = SP_GL_SinSin(s); out(6,s) = para_spin_proj(s)
    end do
  end subroutine

  subroutine ref_calc_theta(x, thc, thr) bind(C, name="ref_calc_theta")
    integer(c_int), intent(in) :: x(N)
    complex(c_double_complex), intent(out) :: thc(max(Mc,1))
    real(c_double), intent(out) :: thr(max(Mr,1))
    integer :: xx(N)
    xx = x
    call calc_theta(xx, thc, thr)
  end subroutine

  subroutine ref_update_theta(x, i1, i2, thc, thr) bind(C, name="ref_update_theta")
    integer(c_int), intent(in) :: x(N)
    integer(c_int), value :: i1, i2
    complex(c_double_complex), intent(inout) :: thc(max(Mc,1))
    real(c_double), intent(inout) :: thr(max(Mr,1))
    integer :: xx(N)
    xx = x
    call update_theta(xx, i1+1, i2+1, thc, thr)
  end subroutine

  subroutine ref_x_to_ele(x, idx, cfg) bind(C, name="ref_x_to_ele")
    integer(c_int), intent(in) :: x(N)
    integer(c_int), intent(out) :: idx(N), cfg(N)
    integer :: xx(N), ii(N), cc(N)
    xx = x
    call x_to_Ele_Idx_Ele_Cfg(xx, ii, cc)
    idx = ii - 1; cfg = cc - 1
  end subroutine

  subroutine ref_sign_pp(idx, sgn) bind(C, name="ref_sign_pp")
    integer(c_int), intent(in) :: idx(N)
    integer(c_int), intent(out) :: sgn(N_MP_Trans)
    integer :: ii(N), ss(N_MP_Trans)
    ii = idx + 1
    call sign_PP(ii, ss)
    sgn = ss
  end subroutine

  subroutine ref_amplitude(thc, thr, pf, sgn, psi_wo, psi) bind(C, name="ref_amplitude")
    complex(c_double_complex), intent(in) :: thc(max(Mc,1))
    real(c_double), intent(in) :: thr(max(Mr,1)), pf(N_QP_Full)
    integer(c_int), intent(in) :: sgn(N_MP_Trans)
    complex(c_double_complex), intent(out) :: psi_wo(N_QP_Full), psi(0:N_QP_Full)
    integer :: ss(N_MP_Trans)
    ss = sgn
    call calc_amplitude_RBM_PP(thc, thr, pf, ss, psi_wo, psi)
  end subroutine

  subroutine ref_pfm_two(msa, msb, tmp_idx, invm, pf_old, pf_new) bind(C, name="ref_pfm_two")
    integer(c_int), value :: msa, msb
    integer(c_int), intent(in) :: tmp_idx(N)
    real(c_double), intent(in) :: invm(N,N,N_QP_Full), pf_old(N_QP_Full)
    real(c_double), intent(out) :: pf_new(N_QP_Full)
    integer :: ii(N)
    ii = tmp_idx + 1
    call calculate_new_PfM_two(msa+1, msb+1, ii, invm, pf_old, pf_new)
  end subroutine

  subroutine ref_invm_two(msa, msb, idx, invm) bind(C, name="ref_invm_two")
    integer(c_int), value :: msa, msb
    integer(c_int), intent(in) :: idx(N)
    real(c_double), intent(inout) :: invm(N,N,N_QP_Full)
    integer :: ii(N)
    ii = idx + 1
    call calculate_new_InvM_two(msa+1, msb+1, ii, invm)
  end subroutine

  subroutine ref_eloc(x, psi0, thc, thr, idx, cfg, sgn, pf, invm, eloc) bind(C, name="ref_eloc")
    integer(c_int), intent(in) :: x(N), idx(N), cfg(N), sgn(N_MP_Trans)
    complex(c_double_complex), intent(in) :: psi0
    complex(c_double_complex), intent(in) :: thc(max(Mc,1))
    real(c_double), intent(in) :: thr(max(Mr,1)), pf(N_QP_Full), invm(N,N,N_QP_Full)
    complex(c_double_complex), intent(out) :: eloc
    integer :: xx(N), ii(N), cc(N), ss(N_MP_Trans)
#ifdef REF_CC
    complex(c_double_complex) :: sdots(N,4)
    xx = x; ii = idx + 1; cc = cfg + 1; ss = sgn
    call calc_local_energy_correlation(xx, psi0, thc, thr, ii, cc, ss, pf, invm, eloc, sdots)
#else
    xx = x; ii = idx + 1; cc = cfg + 1; ss = sgn
    call calc_local_energy_correlation(xx, psi0, thc, thr, ii, cc, ss, pf, invm, eloc)
#endif
  end subroutine

#ifdef REF_CC
  ! Calc_Correlation tree only: the per-bond SdotS and Chi_loc_x of one sample
  subroutine ref_eloc_sdots(x, psi0, thc, thr, idx, cfg, sgn, pf, invm, eloc, sdots) &
      bind(C, name="ref_eloc_sdots")
    integer(c_int), intent(in) :: x(N), idx(N), cfg(N), sgn(N_MP_Trans)
    complex(c_double_complex), intent(in) :: psi0
    complex(c_double_complex), intent(in) :: thc(max(Mc,1))
    real(c_double), intent(in) :: thr(max(Mr,1)), pf(N_QP_Full), invm(N,N,N_QP_Full)
    complex(c_double_complex), intent(out) :: eloc, sdots(N,4)
    integer :: xx(N), ii(N), cc(N), ss(N_MP_Trans)
    xx = x; ii = idx + 1; cc = cfg + 1; ss = sgn
    call calc_local_energy_correlation(xx, psi0, thc, thr, ii, cc, ss, pf, invm, eloc, sdots)
  end subroutine

  subroutine ref_local_correlation(x, psi0, thc, thr, idx, cfg, sgn, pf, invm, chi) &
      bind(C, name="ref_local_correlation")
    integer(c_int), intent(in) :: x(N), idx(N), cfg(N), sgn(N_MP_Trans)
    complex(c_double_complex), intent(in) :: psi0
    complex(c_double_complex), intent(in) :: thc(max(Mc,1))
    real(c_double), intent(in) :: thr(max(Mr,1)), pf(N_QP_Full), invm(N,N,N_QP_Full)
    complex(c_double_complex), intent(out) :: chi(2,Nchi)
    integer :: xx(N), ii(N), cc(N), ss(N_MP_Trans)
    xx = x; ii = idx + 1; cc = cfg + 1; ss = sgn
    call calc_local_correlation(xx, psi0, thc, thr, ii, cc, ss, pf, invm, chi)
  end subroutine
#endif

  subroutine ref_deriv_pf(idx, pf, invm, deriv) bind(C, name="ref_deriv_pf")
    integer(c_int), intent(in) :: idx(N)
    real(c_double), intent(in) :: pf(N_QP_Full), invm(N,N,N_QP_Full)
    real(c_double), intent(out) :: deriv(N_Slater,N_QP_Full)
    integer :: ii(N)
    ii = idx + 1
    call calculate_derivative_Pfaffian(deriv, ii, pf, invm)
  end subroutine

  subroutine ref_bond_map(i0, i1, i2) bind(C, name="ref_bond_map")
    integer(c_int), value :: i0, i1
    integer(c_int), intent(out) :: i2
    integer :: j
    call bond_map(i0+1, i1+1, j)
    i2 = j - 1
  end subroutine

  ! parameter.f90 (initialize_parameters, write_optimized_parameters) on the
  ! files of the current directory; the caller changes into the run directory
  subroutine ref_initialize_parameters() bind(C, name="ref_initialize_parameters")
    call initialize_parameters(0)
  end subroutine

  subroutine ref_write_optimized_parameters() bind(C, name="ref_write_optimized_parameters")
    call write_optimized_parameters(0)
  end subroutine

  ! the module's parameters and orbital table in the layouts of ref_set_params
  ! (orb(i*N + j + 1) = Orbital_Idx(i+1, j+1) - 1)
  subroutine ref_get_params(wc, bc, wr, br, f, orb) bind(C, name="ref_get_params")
    real(c_double), intent(out) :: wc(*), bc(*), wr(*), br(*), f(*)
    integer(c_int), intent(out) :: orb(*)
    integer :: ff, iw, i, j
    do ff = 1, alphac
      bc(2*ff-1) = dble(bcirr(ff)); bc(2*ff) = dimag(bcirr(ff))
      do iw = 1, N
        wc(2*((ff-1)*N+iw)-1) = dble(Wcirr(iw,ff)); wc(2*((ff-1)*N+iw)) = dimag(Wcirr(iw,ff))
      end do
    end do
    do ff = 1, alphar
      br(ff) = brirr(ff)
      do iw = 1, N
        wr((ff-1)*N + iw) = Wrirr(iw,ff)
      end do
    end do
    f(1:N_Slater) = Slater(1:N_Slater)
    do i = 1, N
      do j = 1, N
        orb((i-1)*N + j) = Orbital_Idx(i,j) - 1
      end do
    end do
  end subroutine

  subroutine ref_sgrnd(seed) bind(C, name="ref_sgrnd")
    integer(c_int), value :: seed
    call sgrnd(seed)
  end subroutine

  subroutine ref_grnd_n(n, out) bind(C, name="ref_grnd_n")
    integer(c_int), value :: n
    real(c_double), intent(out) :: out(n)
    real(8), external :: grnd
    integer :: i
    do i = 1, n
      out(i) = grnd()
    end do
  end subroutine

end module ref_harness
