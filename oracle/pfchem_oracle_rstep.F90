! pfchem_oracle_rstep.F90 -- TEST INFRASTRUCTURE: a Fortran restatement of
! RStep / RReact for timing the CPU baseline.  Only tests/ and bench.py's
! cpu_baseline leg load it; no product path links or calls it.
!
! It follows the reference's per-cell operator-split step statement by
! statement, in the language the reference is written in, compiled -O3 the
! way a production build of the reference would be:
!   RStep                 reaction.F90:3564-3738 (sub-step loop, dt cuts,
!                         totals <= 1e-40 clamped and restored)
!   RReact                reaction.F90:3742-4055 (log-formulation Newton,
!                         convergence tests, max-iteration restore)
!   RTotal                reaction.F90:4958-5066 (aqueous complexes, dtotal)
!   RTAccumulation(Deriv) reaction.F90:5710-5848 (liquid phase)
!   RKineticMineral       reaction_mineral.F90:647-1078 (TST rate law,
!                         Arrhenius factor)
!   RSolve                reaction.F90:5457-5516 (row scaling, log-form
!                         column scaling, LU)
!   LUDecomposition / LUBackSubstitution  utility.F90:620-735
!   MineralUpdateKineticState  reaction_mineral.F90:1456-1525
!
! Scope: the networks of the cells/s workload -- aqueous primaries, aqueous
! complexes, kinetic minerals with the plain TST rate (optional activation
! energy), activity coefficients held fixed within the step (frequency OFF or
! TIMESTEP), isothermal log K.  orc_f90_step_batch returns -1 for anything
! else.  The C restatement (pfchem_oracle.c) covers the rest and is the
! checker; this file is checked against it (tests/test_oracle_f90.py).
module pfchem_oracle_rstep
  use, intrinsic :: iso_c_binding
  use, intrinsic :: ieee_arithmetic
  use pfchem_c
  implicit none
  private
  public :: orc_f90_step_batch

  real(c_double), parameter :: LOG_TO_LN = 2.30258509299d0   ! pflotran_constants.F90:84
  real(c_double), parameter :: IDEAL_GAS = 8.31446d0          ! pflotran_constants.F90:89
  real(c_double), parameter :: MIN_SAT = 1.0d-40              ! reactive_transport_aux.F90:19
  integer, parameter :: DEFAULT_MAX_CALLS = 10000

  ! network (read-only while a batch runs)
  integer :: naq, nc, cms, nm, mms
  integer(c_int32_t), pointer :: cid(:, :), mid(:, :)
  real(c_double), pointer :: cst(:, :), ch2o(:), clogk(:)
  real(c_double), pointer :: mst(:, :), mh2o(:), mlogk(:), mk(:), mea(:), mvm(:)
  logical :: has_ch2o, has_mh2o, has_mea
  ! options
  integer :: max_it, max_cuts, max_calls
  logical :: full_geochem, log_form, total_guess
  real(c_double) :: max_dlnc, tol_abs, tol_rel, tol_change
  ! cell state, SoA: field(cell, species)
  real(c_double), pointer :: pm(:, :), tt(:, :), sm(:, :), gp(:, :), gs(:, :)
  real(c_double), pointer :: vfa(:, :), ara(:, :), rta(:, :)
  real(c_double), pointer :: lnaw_a(:), por_a(:), sat_a(:), den_a(:), vol_a(:), temp_a(:)
  integer(c_int32_t), pointer :: nit_a(:), nsub_a(:), stat_a(:)
  logical :: has_sm, has_gp, has_gs, has_lnaw, has_rate, has_temp, has_nit, has_nsub, has_stat

contains

  ! RStep over cells 0 .. ncell-1 of `st` (in place), on `nthreads` threads
  ! (0: the OpenMP default).  Returns the number of failed cells, or -1 when
  ! the network or options are outside this restatement's scope.
  function orc_f90_step_batch(net, opt, st, ncell, dt, nthreads) &
      bind(C, name="orc_f90_step_batch") result(nfail)
    type(pfchem_network), intent(in) :: net
    type(pfchem_options), intent(in) :: opt
    type(pfchem_state), intent(in) :: st
    integer(c_int64_t), value :: ncell
    real(c_double), value :: dt
    integer(c_int), value :: nthreads
    integer(c_int64_t) :: nfail
    integer(c_int64_t) :: c, ld
    integer :: ierr, nsub, nit

    nfail = -1
    if (net%nimmobile /= 0 .or. net%ndecay_rxn /= 0 .or. c_associated(net%somdec)) return
    if (c_associated(net%kinmnrl_Temkin_const) .or. c_associated(net%kinmnrl_min_scale_factor) &
        .or. c_associated(net%kinmnrl_affinity_power) &
        .or. c_associated(net%kinmnrl_affinity_threshold) &
        .or. c_associated(net%kinmnrl_rate_limiter) .or. c_associated(net%kinmnrl_irreversible) &
        .or. c_associated(net%kinmnrl_num_prefactors)) return
    if (opt%act_coef_update_frequency == PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER) return
    if (opt%variable_ncomp /= 0) return
    if (opt%isothermal == 0 .and. (c_associated(net%eqcplx_logKcoef) .or. &
                                   c_associated(net%kinmnrl_logKcoef))) return
    if (net%naqcomp < 1 .or. net%naqcomp > 128) return

    naq = net%naqcomp
    nc = net%neqcplx
    cms = net%eqcplx_maxsp
    nm = net%nkinmnrl
    mms = net%kinmnrl_maxsp
    if (nc > 0) then
      call c_f_pointer(net%eqcplxspecid, cid, [cms + 1, nc])
      call c_f_pointer(net%eqcplxstoich, cst, [cms, nc])
      call c_f_pointer(net%eqcplx_logK, clogk, [nc])
      has_ch2o = c_associated(net%eqcplxh2ostoich)
      if (has_ch2o) call c_f_pointer(net%eqcplxh2ostoich, ch2o, [nc])
    else
      has_ch2o = .false.
    end if
    if (nm > 0) then
      call c_f_pointer(net%kinmnrlspecid, mid, [mms + 1, nm])
      call c_f_pointer(net%kinmnrlstoich, mst, [mms, nm])
      call c_f_pointer(net%kinmnrl_logK, mlogk, [nm])
      call c_f_pointer(net%kinmnrl_rate_constant, mk, [nm])
      call c_f_pointer(net%kinmnrl_molar_vol, mvm, [nm])
      has_mh2o = c_associated(net%kinmnrlh2ostoich)
      if (has_mh2o) call c_f_pointer(net%kinmnrlh2ostoich, mh2o, [nm])
      has_mea = c_associated(net%kinmnrl_activation_energy)
      if (has_mea) call c_f_pointer(net%kinmnrl_activation_energy, mea, [nm])
    else
      has_mh2o = .false.
      has_mea = .false.
    end if

    full_geochem = opt%use_full_geochemistry /= 0
    log_form = opt%use_log_formulation /= 0
    total_guess = opt%use_total_as_guess /= 0
    max_it = opt%maximum_reaction_iterations
    max_cuts = opt%maximum_reaction_cuts
    max_calls = opt%maximum_rreact_calls
    if (max_calls <= 0) max_calls = DEFAULT_MAX_CALLS
    max_dlnc = opt%max_dlnC_rreact
    tol_abs = opt%max_residual_tolerance
    tol_rel = opt%max_rel_residual_tolerance
    tol_change = opt%max_relative_change_tolerance

    ld = st%ld
    call c_f_pointer(st%pri_molal, pm, [ld, int(naq, c_int64_t)])
    call c_f_pointer(st%total, tt, [ld, int(naq, c_int64_t)])
    has_gp = c_associated(st%pri_act_coef)
    if (has_gp) call c_f_pointer(st%pri_act_coef, gp, [ld, int(naq, c_int64_t)])
    has_sm = c_associated(st%sec_molal) .and. nc > 0
    if (has_sm) call c_f_pointer(st%sec_molal, sm, [ld, int(nc, c_int64_t)])
    has_gs = c_associated(st%sec_act_coef) .and. nc > 0
    if (has_gs) call c_f_pointer(st%sec_act_coef, gs, [ld, int(nc, c_int64_t)])
    if (nm > 0) then
      call c_f_pointer(st%mnrl_volfrac, vfa, [ld, int(nm, c_int64_t)])
      call c_f_pointer(st%mnrl_area, ara, [ld, int(nm, c_int64_t)])
    end if
    has_rate = c_associated(st%mnrl_rate) .and. nm > 0
    if (has_rate) call c_f_pointer(st%mnrl_rate, rta, [ld, int(nm, c_int64_t)])
    has_lnaw = c_associated(st%ln_act_h2o)
    if (has_lnaw) call c_f_pointer(st%ln_act_h2o, lnaw_a, [ncell])
    call c_f_pointer(st%porosity, por_a, [ncell])
    call c_f_pointer(st%saturation, sat_a, [ncell])
    call c_f_pointer(st%den_kg, den_a, [ncell])
    call c_f_pointer(st%volume, vol_a, [ncell])
    has_temp = c_associated(st%temperature)
    if (has_temp) call c_f_pointer(st%temperature, temp_a, [ncell])
    has_nit = c_associated(st%num_iterations)
    if (has_nit) call c_f_pointer(st%num_iterations, nit_a, [ncell])
    has_nsub = c_associated(st%num_sub_steps)
    if (has_nsub) call c_f_pointer(st%num_sub_steps, nsub_a, [ncell])
    has_stat = c_associated(st%status)
    if (has_stat) call c_f_pointer(st%status, stat_a, [ncell])

    nfail = 0
    if (nthreads > 0) then
      !$omp parallel do schedule(dynamic, 64) num_threads(nthreads) &
      !$omp& private(ierr, nsub, nit) reduction(+:nfail)
      do c = 1, ncell
        call rstep_cell(c, dt, ierr, nsub, nit)
        if (ierr /= 0) nfail = nfail + 1
      end do
      !$omp end parallel do
    else
      !$omp parallel do schedule(dynamic, 64) private(ierr, nsub, nit) reduction(+:nfail)
      do c = 1, ncell
        call rstep_cell(c, dt, ierr, nsub, nit)
        if (ierr /= 0) nfail = nfail + 1
      end do
      !$omp end parallel do
    end if
  end function orc_f90_step_batch

  ! RStep for cell c (1-based column of the SoA arrays)
  subroutine rstep_cell(c, dt, ierr, nsub, nit)
    integer(c_int64_t), intent(in) :: c
    real(c_double), intent(in) :: dt
    integer, intent(out) :: ierr, nsub, nit
    real(c_double) :: x(naq), tot(naq), gam(naq), guess(naq), small_val(naq)
    real(c_double) :: dtot(naq, naq)
    real(c_double) :: sec(max(nc, 1)), gsec(max(nc, 1))
    real(c_double) :: vf(max(nm, 1)), area(max(nm, 1)), rate(max(nm, 1))
    real(c_double) :: lnaw, por, sat, den, vol, temp
    logical :: small(naq)
    real(c_double) :: target_time, cum, tran_dt
    integer :: num_const, num_cuts, last_reason, attempts, inner, reason, rc, i, k, m

    ! gather
    do i = 1, naq
      x(i) = pm(c, i)
      tot(i) = tt(c, i)
      gam(i) = 1.0d0
      if (has_gp) gam(i) = gp(c, i)
    end do
    do k = 1, nc
      sec(k) = 0.0d0
      gsec(k) = 1.0d0
      if (has_sm) sec(k) = sm(c, k)
      if (has_gs) gsec(k) = gs(c, k)
    end do
    do m = 1, nm
      vf(m) = vfa(c, m)
      area(m) = ara(c, m)
      rate(m) = 0.0d0
      if (has_rate) rate(m) = rta(c, m)
    end do
    lnaw = 0.0d0
    if (has_lnaw) lnaw = lnaw_a(c)
    por = por_a(c)
    sat = sat_a(c)
    den = den_a(c)
    vol = vol_a(c)
    temp = 25.0d0
    if (has_temp) temp = temp_a(c)

    ierr = 0
    nsub = 0
    nit = 0
    if (.not. full_geochem) then
      do i = 1, naq
        x(i) = tot(i) / den * 1.0d3
      end do
      call scatter()
      return
    end if
    ! totals <= 1e-40 are solved at 1e-40 and restored (reaction.F90:3617-3631)
    do i = 1, naq
      small(i) = tot(i) <= 1.0d-40
      if (small(i)) then
        small_val(i) = tot(i)
        tot(i) = 1.0d-40
      end if
    end do
    guess = x
    if (total_guess) guess = tot

    ! sub-step loop (reaction.F90:3674-3725)
    target_time = dt
    cum = 0.0d0
    tran_dt = dt
    num_const = 0
    num_cuts = 0
    last_reason = 0
    attempts = 0
    do
      if (cum >= target_time) exit
      attempts = attempts + 1
      if (attempts > max_calls) then
        ierr = last_reason
        if (ierr == 0) ierr = 1
        call scatter()
        return
      end if
      call rreact(tran_dt, inner, reason, rc)
      nit = nit + inner
      if (rc /= 0) then
        last_reason = reason
        num_cuts = num_cuts + 1
        if (num_cuts > max_cuts) then
          ierr = last_reason
          call scatter()
          return
        end if
        tran_dt = 0.5d0 * tran_dt
        num_const = 0
      else
        call update_kinetic_state(tran_dt)
        cum = cum + tran_dt
        nsub = nsub + 1
        num_const = num_const + 1
        guess = x
        if (num_const >= 4) then
          num_cuts = num_cuts - 1
          tran_dt = min(2.0d0 * tran_dt, target_time - cum)
        end if
      end if
    end do
    do i = 1, naq
      if (small(i)) tot(i) = small_val(i)
    end do
    call scatter()

  contains

    subroutine scatter()
      integer :: ii, kk, mm
      do ii = 1, naq
        pm(c, ii) = x(ii)
        tt(c, ii) = tot(ii)
      end do
      if (has_sm) then
        do kk = 1, nc
          sm(c, kk) = sec(kk)
        end do
      end if
      do mm = 1, nm
        vfa(c, mm) = vf(mm)
        if (has_rate) rta(c, mm) = rate(mm)
      end do
      if (has_nit) nit_a(c) = nit
      if (has_nsub) nsub_a(c) = nsub
      if (has_stat) stat_a(c) = ierr
    end subroutine scatter

    ! RTotal: totals, secondary molalities and dtotal at x
    subroutine total_at_x()
      real(c_double) :: ln_conc(naq), ln_act(naq), lnqk, tempreal, den_kg_per_l
      integer :: ii, jj, kk, ncs, jc
      den_kg_per_l = den * 1.0d-3
      do ii = 1, naq
        ln_conc(ii) = log(x(ii))
        ln_act(ii) = ln_conc(ii) + log(gam(ii))
        tot(ii) = x(ii)
      end do
      dtot = 0.0d0
      do ii = 1, naq
        dtot(ii, ii) = 1.0d0
      end do
      do kk = 1, nc
        lnqk = -clogk(kk) * LOG_TO_LN
        if (has_ch2o) then
          if (ch2o(kk) /= 0.0d0) lnqk = lnqk + ch2o(kk) * lnaw
        end if
        ncs = cid(1, kk)
        do ii = 1, ncs
          lnqk = lnqk + cst(ii, kk) * ln_act(cid(1 + ii, kk) + 1)
        end do
        sec(kk) = exp(lnqk) / gsec(kk)
        do ii = 1, ncs
          tot(cid(1 + ii, kk) + 1) = tot(cid(1 + ii, kk) + 1) + cst(ii, kk) * sec(kk)
        end do
        do jj = 1, ncs
          jc = cid(1 + jj, kk) + 1
          tempreal = cst(jj, kk) * exp(lnqk - ln_conc(jc)) / gsec(kk)
          do ii = 1, ncs
            dtot(cid(1 + ii, kk) + 1, jc) = dtot(cid(1 + ii, kk) + 1, jc) + cst(ii, kk) * tempreal
          end do
        end do
      end do
      do ii = 1, naq
        tot(ii) = tot(ii) * den_kg_per_l
      end do
      dtot = dtot * den_kg_per_l
    end subroutine total_at_x

    ! RTAccumulation (liquid phase)
    subroutine accumulation(res)
      real(c_double), intent(out) :: res(naq)
      real(c_double) :: psv_t
      integer :: ii
      res = 0.0d0
      if (sat < MIN_SAT) return
      psv_t = por * sat * 1000.0d0 * vol
      do ii = 1, naq
        res(ii) = psv_t * tot(ii)
      end do
    end subroutine accumulation

    ! RKineticMineral, TST rate law: rates, residual and Jacobian terms
    subroutine kinetic_mineral(res, jac, deriv)
      real(c_double), intent(inout) :: res(naq), jac(naq, naq)
      logical, intent(in) :: deriv
      real(c_double) :: ln_conc(naq), ln_act(naq)
      real(c_double) :: lnqk, qk, aff, sgn, arrhenius, ea, spr, im, im_const, dim_dqk
      real(c_double) :: dqk_dcj, dqk_dmj
      integer :: ii, jj, mm, ncs, jc
      do ii = 1, naq
        ln_conc(ii) = log(x(ii))
        ln_act(ii) = ln_conc(ii) + log(gam(ii))
      end do
      do mm = 1, nm
        rate(mm) = 0.0d0
      end do
      do mm = 1, nm
        ncs = mid(1, mm)
        lnqk = -mlogk(mm) * LOG_TO_LN
        if (has_mh2o) then
          if (mh2o(mm) /= 0.0d0) lnqk = lnqk + mh2o(mm) * lnaw
        end if
        do ii = 1, ncs
          lnqk = lnqk + mst(ii, mm) * ln_act(mid(1 + ii, mm) + 1)
        end do
        qk = exp(lnqk)
        aff = 1.0d0 - qk
        sgn = sign(1.0d0, aff)
        if (.not. (vf(mm) > 0.0d0 .or. sgn < 0.0d0)) cycle
        arrhenius = 1.0d0
        ea = 0.0d0
        if (has_mea) ea = mea(mm)
        if (ea > 0.0d0) arrhenius = exp(ea / IDEAL_GAS * &
                                        (1.0d0 / (25.0d0 + 273.15d0) - 1.0d0 / (temp + 273.15d0)))
        spr = mk(mm) * arrhenius
        im_const = -area(mm)
        im = im_const * sgn * abs(aff) * spr
        rate(mm) = im
        im_const = im_const * vol
        im = im * vol
        do ii = 1, ncs
          res(mid(1 + ii, mm) + 1) = res(mid(1 + ii, mm) + 1) + mst(ii, mm) * im
        end do
        if (.not. deriv) cycle
        dim_dqk = -im_const * spr
        do jj = 1, ncs
          jc = mid(1 + jj, mm) + 1
          dqk_dcj = mst(jj, mm) * qk * exp(-ln_conc(jc))
          dqk_dmj = dqk_dcj * den * 1.0d-3
          do ii = 1, ncs
            jac(mid(1 + ii, mm) + 1, jc) = jac(mid(1 + ii, mm) + 1, jc) + &
                                           mst(ii, mm) * dim_dqk * dqk_dmj
          end do
        end do
      end do
    end subroutine kinetic_mineral

    ! one RReact iterate's residual and Jacobian (reaction.F90:3899-3916)
    subroutine rreact_system(fixed, rdt, res, jac)
      real(c_double), intent(in) :: fixed(naq), rdt
      real(c_double), intent(out) :: res(naq), jac(naq, naq)
      real(c_double) :: psvd_t
      integer :: ii, jj
      call accumulation(res)
      jac = 0.0d0
      if (sat < MIN_SAT) then
        do ii = 1, naq
          jac(ii, ii) = 1.0d0
        end do
      else
        psvd_t = por * sat * 1000.0d0 * vol / rdt
        do jj = 1, naq
          do ii = 1, naq
            jac(ii, jj) = dtot(ii, jj) * psvd_t
          end do
        end do
      end if
      do ii = 1, naq
        res(ii) = (res(ii) - fixed(ii)) / rdt
      end do
      if (sat < MIN_SAT) return
      if (nm > 0) call kinetic_mineral(res, jac, .true.)
    end subroutine rreact_system

    ! RReact (reaction.F90:3742-4055); rc = 0 converged, else reason
    ! 1 (maximum iterations, totals restored) or 2 (singular Jacobian)
    subroutine rreact(rdt, niter, reason_out, rc_out)
      real(c_double), intent(in) :: rdt
      integer, intent(out) :: niter, reason_out, rc_out
      real(c_double) :: fixed(naq), init_tot(naq), res(naq), upd(naq), conc(naq)
      real(c_double) :: prev(naq), latest(naq), jac(naq, naq)
      real(c_double) :: two_norm_r, two_norm_r0, maxabs, rel, u, mrc, r, min_ratio, ratio
      logical :: have
      integer :: ii, lu_err
      niter = 0
      reason_out = 0
      rc_out = 0
      call accumulation(fixed)
      init_tot = tot
      x = guess
      two_norm_r0 = 0.0d0
      do
        niter = niter + 1
        call total_at_x()
        if (niter > max_it) then
          tot = init_tot
          reason_out = 1
          rc_out = 1
          exit
        end if
        call rreact_system(fixed, rdt, res, jac)
        two_norm_r = 0.0d0
        maxabs = 0.0d0
        do ii = 1, naq
          two_norm_r = two_norm_r + res(ii) * res(ii)
          if (abs(res(ii)) > maxabs) maxabs = abs(res(ii))
        end do
        two_norm_r = sqrt(two_norm_r)
        if (niter == 1) two_norm_r0 = two_norm_r
        rel = two_norm_r / two_norm_r0
        if (maxabs < tol_abs) exit
        if (rel < tol_rel) exit
        conc = x
        call rsolve(res, jac, conc, upd, lu_err)
        if (lu_err /= 0) then
          reason_out = 2
          rc_out = 1
          exit
        end if
        prev = conc
        if (log_form) then
          do ii = 1, naq
            u = sign(1.0d0, upd(ii)) * min(abs(upd(ii)), max_dlnc)
            latest(ii) = prev(ii) * exp(-u)
          end do
        else
          min_ratio = 1.0d308
          do ii = 1, naq
            if (prev(ii) <= upd(ii)) then
              ratio = abs(prev(ii) / upd(ii))
              if (ratio < min_ratio) min_ratio = ratio
            end if
          end do
          if (min_ratio < 1.0d0) upd = upd * min_ratio * 0.99d0
          latest = prev - upd
        end if
        ! maxval of the relative change, NaN entries skipped
        mrc = 0.0d0
        have = .false.
        do ii = 1, naq
          r = abs((latest(ii) - prev(ii)) / prev(ii))
          if (ieee_is_nan(r)) cycle
          if (.not. have .or. r > mrc) mrc = r
          have = .true.
        end do
        if (.not. have) mrc = ieee_value(mrc, ieee_quiet_nan)
        if (mrc < tol_change) exit
        x = latest
      end do
      if (rc_out == 0) call total_at_x()
    end subroutine rreact

    ! MineralUpdateKineticState: rates at the sub-step's answer, volume
    ! fractions advanced over the sub-step and clipped at 0
    subroutine update_kinetic_state(rdt)
      real(c_double), intent(in) :: rdt
      real(c_double) :: res(naq), jac(1, 1)
      integer :: mm
      if (nm == 0) return
      res = 0.0d0
      call kinetic_mineral(res, jac, .false.)
      do mm = 1, nm
        vf(mm) = vf(mm) + rate(mm) * mvm(mm) * rdt
        if (vf(mm) < 0.0d0) vf(mm) = 0.0d0
      end do
    end subroutine update_kinetic_state

  end subroutine rstep_cell

  ! RSolve: row scaling by the row's largest magnitude (at least 1), the
  ! log-form column scaling by the free molalities, then LU (utility.F90)
  subroutine rsolve(res, jac, conc, upd, err)
    real(c_double), intent(in) :: res(:), conc(:)
    real(c_double), intent(inout) :: jac(:, :)
    real(c_double), intent(out) :: upd(:)
    integer, intent(out) :: err
    real(c_double) :: rhs(size(res)), norm
    integer :: indx(size(res)), n, i, j
    n = size(res)
    do i = 1, n
      norm = 0.0d0
      do j = 1, n
        if (abs(jac(i, j)) > norm) norm = abs(jac(i, j))
      end do
      if (.not. (norm > 1.0d0)) norm = 1.0d0
      norm = 1.0d0 / norm
      rhs(i) = res(i) * norm
      do j = 1, n
        jac(i, j) = jac(i, j) * norm
      end do
    end do
    if (log_form) then
      do j = 1, n
        do i = 1, n
          jac(i, j) = jac(i, j) * conc(j)
        end do
      end do
    end if
    call lu_decomposition(jac, n, indx, err)
    if (err /= 0) return
    call lu_back_substitution(jac, n, indx, rhs)
    upd = rhs
  end subroutine rsolve

  ! LUDecomposition (utility.F90:620-690): Crout with implicit row scaling
  ! and partial pivoting; err = 1 for an all-zero row
  subroutine lu_decomposition(a, n, indx, err)
    integer, intent(in) :: n
    real(c_double), intent(inout) :: a(n, n)
    integer, intent(out) :: indx(n), err
    real(c_double), parameter :: tiny = 1.0d-20
    real(c_double) :: vv(n), aamax, sum, dum
    integer :: i, j, k, imax
    err = 0
    imax = 1
    do i = 1, n
      aamax = 0.0d0
      do j = 1, n
        if (abs(a(i, j)) > aamax) aamax = abs(a(i, j))
      end do
      if (aamax <= 0.0d0) then
        err = 1
        return
      end if
      vv(i) = 1.0d0 / aamax
    end do
    do j = 1, n
      do i = 1, j - 1
        sum = a(i, j)
        do k = 1, i - 1
          sum = sum - a(i, k) * a(k, j)
        end do
        a(i, j) = sum
      end do
      aamax = 0.0d0
      do i = j, n
        sum = a(i, j)
        do k = 1, j - 1
          sum = sum - a(i, k) * a(k, j)
        end do
        a(i, j) = sum
        dum = vv(i) * abs(sum)
        if (dum >= aamax) then
          imax = i
          aamax = dum
        end if
      end do
      if (j /= imax) then
        do k = 1, n
          dum = a(imax, k)
          a(imax, k) = a(j, k)
          a(j, k) = dum
        end do
        vv(imax) = vv(j)
      end if
      indx(j) = imax
      if (a(j, j) == 0.0d0) a(j, j) = tiny
      if (j /= n) then
        dum = 1.0d0 / a(j, j)
        do i = j + 1, n
          a(i, j) = a(i, j) * dum
        end do
      end if
    end do
  end subroutine lu_decomposition

  ! LUBackSubstitution (utility.F90:692-735), skipping leading zeros of b
  subroutine lu_back_substitution(a, n, indx, b)
    integer, intent(in) :: n, indx(n)
    real(c_double), intent(in) :: a(n, n)
    real(c_double), intent(inout) :: b(n)
    real(c_double) :: sum
    integer :: i, j, ii, ll
    ii = 0
    do i = 1, n
      ll = indx(i)
      sum = b(ll)
      b(ll) = b(i)
      if (ii /= 0) then
        do j = ii, i - 1
          sum = sum - a(i, j) * b(j)
        end do
      else if (sum /= 0.0d0) then
        ii = i
      end if
      b(i) = sum
    end do
    do i = n, 1, -1
      sum = b(i)
      do j = i + 1, n
        sum = sum - a(i, j) * b(j)
      end do
      b(i) = sum / a(i, i)
    end do
  end subroutine lu_back_substitution

end module pfchem_oracle_rstep
