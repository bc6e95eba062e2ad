! pfchem_rstep_driver.F90 -- exercises the Fortran host module as the OSRT
! loop would: reads a network (reaction_rt_type fields, 1-based ids as
! BasisInit leaves them) and per-cell auxvars from a text file, runs
! RStepBatch over all cells (and RStep on the first cell of a copy), and
! writes the results.  With --layout it prints the sizes and a few field
! offsets of the bind(C) types instead.
!
! usage: pfchem_rstep_driver input output
!        pfchem_rstep_driver --layout
program pfchem_rstep_driver
  use, intrinsic :: iso_c_binding
  use pfchem_c
  use pfchem_reaction_module
  implicit none

  character(len=512) :: arg1, arg2
  type(reaction_rt_type), target :: reaction
  type(option_type) :: option
  type(reactive_transport_auxvar_type), allocatable :: rt(:), rt1(:)
  type(global_auxvar_type), allocatable :: gl(:)
  type(material_auxvar_type), allocatable :: ma(:)
  integer, allocatable :: ids(:), nit(:), nsub(:), status(:)
  real(c_double), allocatable :: guess(:)
  integer :: naq, nimm, nc, cms, nm, mms, ndecay, h2o, flags(9), ncell, variant
  integer :: k, j, ierror, u, i1, i2, i3
  real(c_double) :: reals(7), dt
  real(c_double), allocatable :: row(:)

  call get_command_argument(1, arg1)
  if (trim(arg1) == '--layout') then
    call print_layout()
    stop
  end if
  call get_command_argument(2, arg2)

  open(newunit=u, file=trim(arg1), status='old', action='read')
  read(u, *) naq, nimm, nc, cms, nm, mms, ndecay, h2o
  read(u, *) flags
  read(u, *) reals
  reaction%naqcomp = naq
  reaction%ncomp = naq + nimm
  reaction%offset_immobile = naq
  reaction%use_full_geochemistry = flags(1) /= 0
  reaction%use_log_formulation = flags(2) /= 0
  reaction%use_total_as_guess = flags(3) /= 0
  reaction%use_activity_h2o = flags(4) /= 0
  reaction%act_coef_update_frequency = flags(5)
  reaction%act_coef_update_algorithm = flags(6)
  reaction%maximum_reaction_iterations = flags(7)
  reaction%maximum_reaction_cuts = flags(8)
  option%use_isothermal = flags(9) /= 0
  reaction%max_dlnC_rreact = reals(1)
  reaction%max_residual_tolerance = reals(2)
  reaction%max_rel_residual_tolerance = reals(3)
  reaction%max_relative_change_tolerance = reals(4)
  reaction%debyeA = reals(5)
  reaction%debyeB = reals(6)
  reaction%debyeBdot = reals(7)
  allocate(reaction%species_idx)
  reaction%species_idx%h2o_aq_id = h2o
  allocate(reaction%primary_spec_Z(naq), reaction%primary_spec_a0(naq))
  read(u, *) reaction%primary_spec_Z
  read(u, *) reaction%primary_spec_a0

  reaction%neqcplx = nc
  if (nc > 0) then
    allocate(reaction%eqcplxspecid(0:cms, nc), reaction%eqcplxstoich(cms, nc))
    allocate(reaction%eqcplxh2oid(nc), reaction%eqcplxh2ostoich(nc), reaction%eqcplx_logK(nc))
    allocate(reaction%eqcplx_Z(nc), reaction%eqcplx_a0(nc))
    do k = 1, nc
      read(u, *) reaction%eqcplxspecid(:, k)
      read(u, *) reaction%eqcplxstoich(:, k)
      read(u, *) reaction%eqcplxh2oid(k), reaction%eqcplxh2ostoich(k), reaction%eqcplx_logK(k), &
                 reaction%eqcplx_Z(k), reaction%eqcplx_a0(k)
    end do
  end if

  allocate(reaction%mineral)
  reaction%mineral%nkinmnrl = nm
  if (nm > 0) then
    associate(m => reaction%mineral)
      allocate(m%kinmnrlspecid(0:mms, nm), m%kinmnrlstoich(mms, nm), m%kinmnrlh2oid(nm))
      allocate(m%kinmnrlh2ostoich(nm), m%kinmnrl_logK(nm), m%kinmnrl_rate_constant(nm))
      allocate(m%kinmnrl_molar_vol(nm))
      do k = 1, nm
        read(u, *) m%kinmnrlspecid(:, k)
        read(u, *) m%kinmnrlstoich(:, k)
        read(u, *) m%kinmnrlh2oid(k), m%kinmnrlh2ostoich(k), m%kinmnrl_logK(k), &
                   m%kinmnrl_rate_constant(k), m%kinmnrl_molar_vol(k)
      end do
    end associate
  end if

  allocate(reaction%immobile)
  reaction%immobile%nimmobile = nimm
  reaction%immobile%ndecay_rxn = ndecay
  if (ndecay > 0) then
    allocate(reaction%immobile%decayspecid(ndecay), reaction%immobile%decay_rate_constant(ndecay))
    read(u, *) reaction%immobile%decayspecid
    read(u, *) reaction%immobile%decay_rate_constant
  end if

  read(u, *) ncell, dt, variant
  option%tran_dt = dt
  allocate(rt(ncell), gl(ncell), ma(ncell), ids(ncell), nit(ncell), nsub(ncell), status(ncell))
  allocate(row(3 * naq + nimm + 2 * nc + 2 * nm + 6))
  do j = 1, ncell
    ids(j) = j
    read(u, *) row
    allocate(rt(j)%pri_molal(naq), rt(j)%total(naq, 1), rt(j)%pri_act_coef(naq))
    allocate(rt(j)%immobile(max(nimm, 1)), rt(j)%sec_molal(max(nc, 1)), rt(j)%sec_act_coef(max(nc, 1)))
    allocate(rt(j)%mnrl_volfrac(max(nm, 1)), rt(j)%mnrl_area(max(nm, 1)), rt(j)%mnrl_rate(max(nm, 1)))
    allocate(gl(j)%sat(1), gl(j)%den_kg(1))
    i1 = 0
    rt(j)%pri_molal = row(i1 + 1:i1 + naq); i1 = i1 + naq
    rt(j)%total(:, 1) = row(i1 + 1:i1 + naq); i1 = i1 + naq
    rt(j)%pri_act_coef = row(i1 + 1:i1 + naq); i1 = i1 + naq
    if (nimm > 0) rt(j)%immobile(1:nimm) = row(i1 + 1:i1 + nimm)
    i1 = i1 + nimm
    if (nc > 0) then
      rt(j)%sec_molal(1:nc) = row(i1 + 1:i1 + nc)
      rt(j)%sec_act_coef(1:nc) = row(i1 + nc + 1:i1 + 2 * nc)
    end if
    i1 = i1 + 2 * nc
    if (nm > 0) then
      rt(j)%mnrl_volfrac(1:nm) = row(i1 + 1:i1 + nm)
      rt(j)%mnrl_area(1:nm) = row(i1 + nm + 1:i1 + 2 * nm)
      rt(j)%mnrl_rate(1:nm) = 0.d0
    end if
    i1 = i1 + 2 * nm
    rt(j)%ln_act_h2o = row(i1 + 1)
    ma(j)%porosity = row(i1 + 2)
    gl(j)%sat(1) = row(i1 + 3)
    gl(j)%den_kg(1) = row(i1 + 4)
    ma(j)%volume = row(i1 + 5)
    gl(j)%temp = row(i1 + 6)
  end do
  close(u)

  call PfchemSetVariant(variant)

  ! RStep on a deep copy of the first cell (the reference's per-cell entry)
  allocate(rt1(1))
  rt1(1) = rt(1)
  allocate(rt1(1)%pri_molal(naq), rt1(1)%total(naq, 1), rt1(1)%pri_act_coef(naq))
  allocate(rt1(1)%immobile(max(nimm, 1)), rt1(1)%sec_molal(max(nc, 1)), rt1(1)%sec_act_coef(max(nc, 1)))
  allocate(rt1(1)%mnrl_volfrac(max(nm, 1)), rt1(1)%mnrl_area(max(nm, 1)), rt1(1)%mnrl_rate(max(nm, 1)))
  rt1(1)%pri_molal = rt(1)%pri_molal
  rt1(1)%total = rt(1)%total
  rt1(1)%pri_act_coef = rt(1)%pri_act_coef
  rt1(1)%immobile = rt(1)%immobile
  rt1(1)%sec_molal = rt(1)%sec_molal
  rt1(1)%sec_act_coef = rt(1)%sec_act_coef
  rt1(1)%mnrl_volfrac = rt(1)%mnrl_volfrac
  rt1(1)%mnrl_area = rt(1)%mnrl_area
  rt1(1)%mnrl_rate = rt(1)%mnrl_rate
  allocate(guess(naq + nimm))
  guess(1:naq) = rt1(1)%pri_molal
  if (nimm > 0) guess(naq + 1:) = rt1(1)%immobile(1:nimm)
  call RStep(guess, rt1(1), gl(1), ma(1), i1, i2, i3, reaction, 1, option, ierror)
  if (ierror /= 0) write(*, '(a)') 'RStep: ' // PfchemLastError()

  call RStepBatch(reaction, rt, gl, ma, ids, option, ierror, nit, nsub, status)
  if (ierror /= 0) write(*, '(a)') 'RStepBatch ierror=1 ' // PfchemLastError()

  open(newunit=u, file=trim(arg2), status='replace', action='write')
  write(u, '(*(g0,:," "))') 'rstep', i1, i2, rt1(1)%pri_molal
  do j = 1, ncell
    write(u, '(*(g0,:," "))') status(j), nit(j), nsub(j), rt(j)%pri_molal, rt(j)%total(:, 1), &
      rt(j)%immobile(1:nimm), rt(j)%sec_molal(1:nc), rt(j)%mnrl_volfrac(1:nm), &
      rt(j)%mnrl_rate(1:nm)
  end do
  close(u)
  call PfchemFinalize()

contains

  subroutine print_layout()
    type(pfchem_network), target :: net
    type(pfchem_options), target :: opt
    type(pfchem_state), target :: st
    type(pfchem_somdec), target :: sd
    type(pfchem_launch), target :: la
    type(pfchem_constraint), target :: con
    type(pfchem_connections), target :: cn
    type(pfchem_transport_coefs), target :: tc
    integer(c_intptr_t) :: b
    b = transfer(c_loc(net), b)
    write(*, '(*(g0,:," "))') 'network', c_sizeof(net), &
      transfer(c_loc(net%neqcplx), b) - b, transfer(c_loc(net%nkinmnrl), b) - b, &
      transfer(c_loc(net%ndecay_rxn), b) - b, transfer(c_loc(net%debyeA), b) - b, &
      transfer(c_loc(net%somdec), b) - b, transfer(c_loc(net%kinmnrl_max_pref), b) - b, &
      transfer(c_loc(net%kinmnrl_pref_activation_energy), b) - b
    b = transfer(c_loc(opt), b)
    write(*, '(*(g0,:," "))') 'options', c_sizeof(opt), &
      transfer(c_loc(opt%max_dlnC_rreact), b) - b, &
      transfer(c_loc(opt%max_relative_change_tolerance), b) - b
    b = transfer(c_loc(st), b)
    write(*, '(*(g0,:," "))') 'state', c_sizeof(st), transfer(c_loc(st%temperature), b) - b, &
      transfer(c_loc(st%status), b) - b, transfer(c_loc(st%ncomp_active), b) - b
    b = transfer(c_loc(sd), b)
    write(*, '(*(g0,:," "))') 'somdec', c_sizeof(sd), transfer(c_loc(sd%x0eps), b) - b, &
      transfer(c_loc(sd%rate_constant), b) - b, transfer(c_loc(sd%inhibition_constant2), b) - b, &
      transfer(c_loc(sd%species_o2), b) - b
    write(*, '(*(g0,:," "))') 'launch', c_sizeof(la)
    b = transfer(c_loc(con), b)
    write(*, '(*(g0,:," "))') 'constraint', c_sizeof(con), transfer(c_loc(con%nrxn), b) - b, &
      transfer(c_loc(con%h_ion_id), b) - b, transfer(c_loc(con%max_dlnC), b) - b, &
      transfer(c_loc(con%free_ion_guess), b) - b
    b = transfer(c_loc(cn), b)
    write(*, '(*(g0,:," "))') 'connections', c_sizeof(cn), transfer(c_loc(cn%nbc), b) - b, &
      transfer(c_loc(cn%bc_cell), b) - b
    b = transfer(c_loc(tc), b)
    write(*, '(*(g0,:," "))') 'transport_coefs', c_sizeof(tc), &
      transfer(c_loc(tc%bc_area), b) - b, transfer(c_loc(tc%bc_total), b) - b
  end subroutine print_layout

end program pfchem_rstep_driver
