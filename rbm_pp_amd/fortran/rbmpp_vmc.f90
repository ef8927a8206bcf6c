!
! rbmpp_vmc -- the optimisation driver of main.f90 on one GPU through the
! C ABI (rbmpp_mod).  Run in a directory that holds the reference's input
! files: RBM.input, sym_info.txt, orbitalidx.def, zqp_orbital_initial.dat and
! W.input (the reference draws W from a clock-seeded stream when W.input is
! absent, parameter.f90:21-41; this driver requires the file).
!
!   rbmpp_vmc [walkers [s_precision [recompute_k [projection [mode [seed]]]]]]
!
! seed is the key of the walkers' Philox streams (default 1).
!
! projection 1 runs the transverse_code variant: its spin-projection weights
! (transverse_code/spin_projection.f90:30-38) and its reweighting-offset
! factor 0.003 (transverse_code/main.f90:451) in place of 0.001.
!
! mode 2 reads the run directory and writes the parameters it read with the
! writer of optimized_W.txt and zqp_orbital_opt.dat, then stops (no device
! work): the reader and writer against parameter.f90's.
!
! mode 1 is the measurement run of Calc_Correlation (Calc_Correlation/
! main.f90:311 skips the optimisation, :499-640 measure): Nav iterations of
! energy, spin correlation Chi_sp and dimer correlation Chi_vbs with the
! parameters of the input files, written to fort.<60000+it> and
! fort.<70000+it> in the reference's formats, plus energy_after_opt.txt and
! Average_Energy.txt.
!
! Writes Energy_vs_Iteration.txt (main.f90:410), optimized_W.txt and
! zqp_orbital_opt.dat (parameter.f90:238-276), energy_after_opt.txt and
! Average_Energy.txt (main.f90:488-572).  Nsample of RBM.input is the total
! over walkers, as it is over MPI ranks in the reference.
!
program rbmpp_vmc
  use iso_c_binding
  use rbmpp_mod
  implicit none
  integer :: L, N, Nsample, spin_parity, alphac, alphar, Nstep, Nav, P, ncopy, nsl
  real(8) :: delta_tau, s_factor0, s_factor_scale, s_cut, J1, J2
  integer(c_int32_t), allocatable :: w_map(:,:), orb(:,:)
  real(c_double), allocatable :: chi(:,:), wc(:,:,:), bc(:,:), wr(:,:), br(:), slater(:)
  real(c_double), allocatable :: gvec(:), ovec(:), gam(:), gam_av(:)
  real(8), allocatable :: e_data(:), e2_data(:)
  type(rbmpp_config) :: cfg
  type(rbmpp_sr_config) :: src
  type(rbmpp_sr_report) :: rep
  type(c_ptr) :: h
  integer(c_int) :: rc
  integer(c_int64_t) :: naccept
  integer :: walkers, nspw, prec, kre, proj, mode, it, i, nv, nav_chk
  integer(c_int64_t) :: seed
  real(8) :: shift_factor
  real(8) :: etot, etot2, norm, wf_av, shift, shift_av, wf_av_list(10), acc, e_av, e_er, v_av, v_er
  character(len=64) :: arg

  walkers = 4096; prec = RBMPP_S_FP64; kre = 0; proj = RBMPP_PROJ_NORMAL; mode = 0; seed = 1_c_int64_t
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg); read(arg, *) walkers
  end if
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg); read(arg, *) prec
  end if
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg); read(arg, *) kre
  end if
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg); read(arg, *) proj
  end if
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg); read(arg, *) mode
  end if
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg); read(arg, *) seed
  end if
  shift_factor = merge(0.003d0, 0.001d0, proj == RBMPP_PROJ_TRANSVERSE)

  call read_inputs()
  nspw = (Nsample + walkers - 1) / walkers
  if (mode == 2) then
    ! the parameters as read (initialize_parameters) written back with the
    ! writer of write_optimized_parameters; no device work
    call write_optimized_parameters(packed_inputs())
    stop
  end if

  cfg%L = L; cfg%n_sym = P; cfg%alphac = alphac; cfg%alphar = alphar
  cfg%n_slater = nsl; cfg%spin_parity = spin_parity; cfg%J1 = J1; cfg%J2 = J2
  cfg%n_walkers = walkers; cfg%n_samples = nspw; cfg%n_warmup = -1; cfg%n_warmup_first = -1
  cfg%n_update = 0; cfg%recompute_k = kre; cfg%s_precision = prec; cfg%fused_measure = 0
  cfg%seed = seed; cfg%p_restart = 0.05d0; cfg%p_global = 3d-4; cfg%cond_guard = 1d4
  rc = rbmpp_create(cfg, w_map, chi, orb, h)
  call check(rc, "rbmpp_create")
  nv = rbmpp_nv(h)
  allocate(gvec(nv), ovec(nv), gam(nv), gam_av(nv), e_data(max(Nav, 1)), e2_data(max(Nav, 1)))
  rc = rbmpp_set_params(h, wc, bc, wr, br, slater)
  call check(rc, "rbmpp_set_params")
  if (proj /= RBMPP_PROJ_NORMAL) then
    rc = rbmpp_set_projection(h, proj)
    call check(rc, "rbmpp_set_projection")
  end if
  if (mode == 1) then
    call correlation_run()
    call rbmpp_destroy(h)
    stop
  end if

  ! optimisation (main.f90:300-455)
  src%s_factor0 = s_factor0; src%delta_tau = delta_tau; src%max_change = 0.04d0
  ! one refinement sweep in the captured solve; a mixed-precision S is then
  ! refined against the FP64 operator of the O rows until the residual is
  ! below 1e-13, which brings the step to the FP64 system's solution
  src%slater_abs_av = 0.25d0; src%refine_iters = 1; src%flags = 0
  shift = 0d0; shift_av = 0d0; wf_av_list = 0d0; gam_av = 0d0; nav_chk = 0
  open(unit=10, file='Energy_vs_Iteration.txt', status='unknown')
  do it = 1, Nstep
    write(6, '(a,I6)') 'Iteration :', it
    rc = rbmpp_make_sample(h, merge(1, 0, it == 1), shift, naccept)
    call check(rc, "rbmpp_make_sample")
    acc = dble(naccept) / dble(walkers * nspw) / dble(N)
    rc = rbmpp_measure(h, shift, etot, etot2, norm, wf_av, gvec, ovec)
    call check(rc, "rbmpp_measure")
    rc = rbmpp_get_gamma(h, gam)
    src%s_factor0 = s_factor0
    rc = rbmpp_sr_step(h, src, gam, rep)
    call check(rc, "rbmpp_sr_step")
    write(6, '(a,F20.10)')  '   Total Energy         :', etot
    write(6, '(a,F20.10)')  '   Energy Variance      :', (etot2 - etot*etot) / (etot*etot)
    write(6, '(a,F20.10)')  '   Acceptance Ratio     :', acc
    write(6, '(a,2F20.10)') '   Stabilization Factor :', rep%s_factor, s_cut
    write(6, '(a,E12.3,I6,E12.3)') '   SR residual, CG, boost:', rep%residual, rep%cg_iters, rep%precond_boost
    write(6, '(a,2E20.5)')  '   wf_av, shift         :', wf_av, shift
    write(6, '(a,F20.10)')  '   norm                 :', norm / dble(walkers * nspw)
    write(10, *) it, etot, (etot2 - etot*etot) / (etot*etot)
    flush(10)
    flush(6)
    if (it > Nstep - Nav) then
      nav_chk = nav_chk + 1
      shift_av = shift_av + shift
      gam_av = gam_av + gam
    end if
    s_factor0 = s_factor0 * s_factor_scale
    ! reweighting offset from the running mean of wf_av (main.f90:441-451)
    wf_av_list(1:9) = wf_av_list(2:10)
    wf_av_list(10) = wf_av
    shift = sum(wf_av_list) / 10d0 * shift_factor / dble(N) / dble(N)
  end do
  close(10)
  if (Nav <= 0) then
    call rbmpp_destroy(h)
    stop
  end if
  if (nav_chk /= Nav) stop 'Nav_chk /= Nav'

  ! averaged parameters (main.f90:457-486)
  gam_av = gam_av / dble(Nav)
  shift = shift_av / dble(Nav)
  rc = rbmpp_set_gamma(h, gam_av)
  call check(rc, "rbmpp_set_gamma")
  call write_optimized_parameters(gam_av)

  ! final energy (main.f90:488-572)
  open(unit=10, file='energy_after_opt.txt', status='unknown')
  rc = rbmpp_make_sample(h, 1, shift, naccept)
  call check(rc, "rbmpp_make_sample")
  do i = 1, 2
    rc = rbmpp_make_sample(h, 0, shift, naccept)
    call check(rc, "rbmpp_make_sample")
  end do
  do it = 1, Nav
    rc = rbmpp_make_sample(h, 0, shift, naccept)
    call check(rc, "rbmpp_make_sample")
    rc = rbmpp_measure_energy(h, shift, etot, etot2, norm, wf_av)
    call check(rc, "rbmpp_measure_energy")
    write(6, '(a,I6)')     'Iteration (Energy Measuement) :', it
    write(6, '(a,F20.10)') '   total energy         :', etot
    e_data(it) = etot
    e2_data(it) = etot2 - etot*etot
    write(10, *) it, etot, (etot2 - etot*etot) / (etot*etot)
    flush(10)
    flush(6)
  end do
  close(10)
  call calc_av_er(Nav, e_data, e_av, e_er)
  call calc_av_er(Nav, e2_data, v_av, v_er)
  v_av = v_av / (e_av*e_av)
  v_er = v_er / (e_av*e_av)
  open(unit=10, file='Average_Energy.txt', status='unknown')
  write(10, '(3F25.15)') e_av, e_er / dsqrt(dble(Nav-1)), e_er
  write(10, '(3E25.15)') v_av, v_er / dsqrt(dble(Nav-1)), v_er
  write(10, '(4F20.12)') e_av, e_er / dsqrt(dble(Nav-1)), v_av, v_er / dsqrt(dble(Nav-1))
  close(10)
  call rbmpp_destroy(h)

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= RBMPP_OK) then
      write(0, '(a,a,I6,a,a)') what, ' failed: ', code, ' ', rbmpp_error_message(h)
      stop 1
    end if
  end subroutine

  ! Calc_Correlation/main.f90:499-624
  subroutine correlation_run()
    integer :: nchi, j, u
    real(c_double) :: wf_sgn(2)
    real(c_double), allocatable :: chi_sp(:,:), chi_vbs(:,:)
    character(len=32) :: fname
    nchi = N * (N + 1) / 2
    allocate(chi_sp(2, nchi), chi_vbs(nchi, 4))
    open(unit=10, file='energy_after_opt.txt', status='unknown')
    ! warm-up sampling; the offset is 1e-5 of the mean |psi| (:515-523)
    rc = rbmpp_make_sample(h, 1, 0d0, naccept)
    call check(rc, "rbmpp_make_sample")
    rc = rbmpp_measure_energy(h, 0d0, etot, etot2, norm, wf_av)
    call check(rc, "rbmpp_measure_energy")
    shift = wf_av * 1d-5
    write(6, '(a,2E15.5)') 'wf_av, shift =', wf_av, shift
    do it = 1, 7
      rc = rbmpp_make_sample(h, 0, shift, naccept)
      call check(rc, "rbmpp_make_sample")
    end do
    do it = 1, Nav
      rc = rbmpp_make_sample(h, 0, shift, naccept)
      call check(rc, "rbmpp_make_sample")
      acc = dble(naccept) / dble(walkers * nspw) / dble(N)
      rc = rbmpp_measure_correlation(h, shift, etot, etot2, norm, wf_av, wf_sgn, chi_sp, chi_vbs)
      call check(rc, "rbmpp_measure_correlation")
      write(6, '(a,I6)')      'Iteration (Energy Measuement) :', it
      write(6, '(a,F20.10)')  '   total energy         :', etot
      write(6, '(a,F20.10)')  '   energy variance      :', (etot2 - etot*etot) / (etot*etot)
      write(6, '(a,F20.10)')  '   acceptance ratio     :', acc
      write(6, '(a,2E20.5)')  '   wf_av, shift         :', wf_av, shift
      write(6, '(a,F20.10)')  '   norm                 :', norm / dble(walkers * nspw)
      write(6, '(a,3F20.10)') '   wf_sgn               :', sqrt(wf_sgn(1)**2 + wf_sgn(2)**2), wf_sgn
      e_data(it) = etot
      e2_data(it) = etot2 - etot*etot
      write(fname, '(a,I0)') 'fort.', 60000 + it
      open(newunit=u, file=trim(fname), status='replace')
      write(u, '(a)') '# chi index,  SzSz,  SpSm'
      do i = 1, nchi
        write(u, '(I6,2F25.15)') i, chi_sp(1, i), chi_sp(2, i)
      end do
      close(u)
      write(fname, '(a,I0)') 'fort.', 70000 + it
      open(newunit=u, file=trim(fname), status='replace')
      write(u, '(a)') '# chi index, <SSSS> for (1,0), (0,1), (1,1), (1,-1)'
      do i = 1, nchi
        write(u, '(I6,4F25.15)') i, (chi_vbs(i, j), j = 1, 4)
      end do
      close(u)
      write(10, *) it, etot, (etot2 - etot*etot) / (etot*etot)
    end do
    close(10)
    if (Nav > 1) then
      call calc_av_er(Nav, e_data, e_av, e_er)
      call calc_av_er(Nav, e2_data, v_av, v_er)
      v_av = v_av / (e_av*e_av)
      v_er = v_er / (e_av*e_av)
      open(unit=10, file='Average_Energy.txt', status='unknown')
      write(10, '(3F25.15)') e_av, e_er / dsqrt(dble(Nav-1)), e_er
      write(10, '(3E25.15)') v_av, v_er / dsqrt(dble(Nav-1)), v_er
      write(10, '(4F20.12)') e_av, e_er / dsqrt(dble(Nav-1)), v_av, v_er / dsqrt(dble(Nav-1))
      close(10)
    end if
  end subroutine

  ! sub.f90:93-115
  subroutine calc_av_er(n_, d, av, er)
    integer, intent(in) :: n_
    real(8), intent(in) :: d(n_)
    real(8), intent(out) :: av, er
    av = sum(d) / dble(n_)
    er = dsqrt(sum((d - av)**2) / dble(n_))
  end subroutine

  ! the input files in the formats main.f90:75-141 and parameter.f90:44-141 read
  subroutine read_inputs()
    integer :: u, g, i, j, k, i1, i2, i3, i4, f, iw
    complex(8) :: z
    real(8) :: re, im, mean_abs
    character(len=256) :: line, tag

    open(newunit=u, file='RBM.input', status='old')
    read(u, *)
    read(u, *) L, Nsample, spin_parity
    read(u, *)
    read(u, *) alphac, alphar, Nstep, Nav, delta_tau
    read(u, *)
    read(u, *) s_factor0, s_factor_scale, s_cut
    read(u, *)
    read(u, *) J1, J2
    close(u)
    N = L * L

    open(newunit=u, file='sym_info.txt', status='old')
    read(u, *)
    read(u, *) P, ncopy
    if (ncopy /= 1) stop 'Ncopy /= 1'
    read(u, *)
    allocate(chi(2, P), w_map(N, P))
    do g = 1, P
      read(u, *) i1, z
      if (i1 /= g) stop 'i1/=mpidx'
      chi(1, g) = dble(z); chi(2, g) = aimag(z)
    end do
    read(u, *)
    do g = 1, P
      do i = 1, N
        read(u, *) i1, i2, i3, i4
        if (i1 /= g .or. i2 /= 1 .or. i3 /= i) stop 'W_map wrong'
        w_map(i, g) = i4
      end do
    end do
    close(u)

    open(newunit=u, file='orbitalidx.def', status='old')
    read(u, *)
    read(u, *) tag, nsl
    read(u, *)
    read(u, *)
    read(u, *)
    allocate(orb(N, N))
    do k = 1, N * N
      read(u, *) i, j, i1
      orb(i+1, j+1) = i1 + 1
    end do
    close(u)

    open(newunit=u, file='zqp_orbital_initial.dat', status='old')
    read(u, *)
    read(u, *)
    read(u, *)
    read(u, *)
    read(u, *)
    allocate(slater(nsl))
    do k = 1, nsl
      read(u, *) i, re, im
      if (i /= k - 1 .or. abs(im) > 1d-10) stop 'error in zqp_orbital'
      slater(k) = re
    end do
    close(u)
    ! mean |f| = 0.25 (parameter.f90:121-140), the reference's arithmetic:
    ! sequential sum, then f * 0.25 / mean
    mean_abs = 0d0
    do k = 1, nsl
      mean_abs = mean_abs + abs(slater(k))
    end do
    mean_abs = mean_abs / dble(nsl)
    do k = 1, nsl
      slater(k) = slater(k) * 0.25d0 / mean_abs
    end do

    allocate(wc(2, N, max(alphac, 1)), bc(2, max(alphac, 1)), wr(N, max(alphar, 1)), br(max(alphar, 1)))
    wc = 0d0; bc = 0d0; wr = 0d0; br = 0d0
    open(newunit=u, file='W.input', status='old')
    read(u, '(a)') line
    do f = 1, alphac
      read(u, '(a)') line
      read(line(7:), *) i1, bc(1, f), bc(2, f)
      do iw = 1, N
        read(u, *) i1, i2, wc(1, iw, f), wc(2, iw, f)
      end do
      read(u, '(a)') line
    end do
    do f = 1, alphar
      read(u, '(a)') line
      read(line(7:), *) i1, br(f), re
      do iw = 1, N
        read(u, *) i1, i2, wr(iw, f), re
      end do
      read(u, '(a)') line
    end do
    close(u)
  end subroutine

  ! the parameters read from the input files, packed as main.f90:263-294
  function packed_inputs() result(g)
    real(c_double), allocatable :: g(:)
    integer :: f, iw, k
    allocate(g(2*alphac*(N+1) + alphar*(N+1) + nsl))
    k = 0
    do f = 1, alphac
      g(k+1) = bc(1, f); g(k+2) = bc(2, f); k = k + 2
    end do
    do f = 1, alphac
      do iw = 1, N
        g(k+1) = wc(1, iw, f); g(k+2) = wc(2, iw, f); k = k + 2
      end do
    end do
    do f = 1, alphar
      g(k+1) = br(f); k = k + 1
    end do
    do f = 1, alphar
      do iw = 1, N
        g(k+1) = wr(iw, f); k = k + 1
      end do
    end do
    g(k+1:k+nsl) = slater(1:nsl)
  end function

  ! parameter.f90:238-276, from the packed gamma (main.f90:263-294)
  subroutine write_optimized_parameters(g)
    real(c_double), intent(in) :: g(:)
    integer :: u, f, iw, k
    k = 0
    open(newunit=u, file='optimized_W.txt', status='unknown')
    write(u, '(a)') '# index for independent neurons (f), index for interaction (iw), irreducible interaction Wirr(iw,f)'
    do f = 1, alphac
      write(u, '(a,I6,2F22.15)') '# bias', f, g(2*f-1), g(2*f)
      do iw = 1, N
        k = 2*alphac + 2*((f-1)*N + iw - 1)
        write(u, '(2I6,2F22.15)') f, iw, g(k+1), g(k+2)
      end do
      write(u, *)
    end do
    k = 2*alphac*(N+1)
    do f = 1, alphar
      write(u, '(a,I6,2F22.15)') '# bias', f, g(k+f), 0d0
      do iw = 1, N
        write(u, '(2I6,2F22.15)') f, iw, g(k + alphar + (f-1)*N + iw), 0d0
      end do
      write(u, *)
    end do
    close(u)
    k = 2*alphac*(N+1) + alphar*(N+1)
    open(newunit=u, file='zqp_orbital_opt.dat', status='unknown')
    write(u, *)
    write(u, '(a,I10)') ' Slater', nsl
    write(u, *)
    write(u, *)
    write(u, *)
    do f = 1, nsl
      write(u, '(I5,2e25.15)') f-1, g(k+f), 0d0
    end do
    close(u)
  end subroutine

end program rbmpp_vmc
