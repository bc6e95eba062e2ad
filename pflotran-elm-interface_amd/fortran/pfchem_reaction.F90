! pfchem_reaction.F90 -- Fortran host side of the MI355X reaction step.
!
! The module keeps the reference's names for the objects the OSRT loop hands
! to RStep (reaction_rt_type, reaction_aux.F90:123-311, with its mineral_type
! (reaction_mineral_aux.F90:82-140) and immobile_type
! (reaction_immobile_aux.F90:40-56) members; reactive_transport_auxvar_type,
! reactive_transport_aux.F90:21-74; global_auxvar_type, global_aux.F90:11-40;
! material_auxvar_type, material_aux.F90:53-70; option%tran_dt and
! option%use_isothermal, option.F90:104-108), restricted to the fields the
! reaction step reads or writes, and provides
!
!   RStep(guess, rt_auxvar, global_auxvar, material_auxvar, num_sub_steps,
!         num_iterations, num_kinetic_state_updates, reaction, natural_id,
!         option, ierror)             -- reaction.F90:3564-3566 signature
!   RStepBatch(reaction, rt_auxvars, global_auxvars, material_auxvars,
!              cell_ids, option, ierror[, num_iterations, num_sub_steps,
!              cell_status])
!
! RStepBatch replaces the per-cell loop of PMCSubsurfaceOSRTStepDT
! (pmc_subsurface_osrt.F90:349-378): it gathers the listed cells' auxvars into
! structure-of-arrays host buffers, runs pfchem_step_host, and scatters the
! results back.  ierror = 1 if any cell failed, as the reference's RStep
! ierror reduced with MAX over the cells (pmc_subsurface_osrt.F90:381-382).
! The network is uploaded on the first call for a reaction object and kept
! in a device context until PfchemFinalize.
module pfchem_reaction_module
  use, intrinsic :: iso_c_binding
  use pfchem_c
  implicit none
  private

  integer, parameter :: dp = c_double

  type, public :: species_idx_type
    integer :: h2o_aq_id = 0
  end type species_idx_type

  type, public :: mineral_type
    integer :: nkinmnrl = 0
    integer, pointer :: kinmnrlspecid(:,:) => null()       ! (0:maxsp, nkinmnrl)
    real(dp), pointer :: kinmnrlstoich(:,:) => null()      ! (maxsp, nkinmnrl)
    integer, pointer :: kinmnrlh2oid(:) => null()
    real(dp), pointer :: kinmnrlh2ostoich(:) => null()
    real(dp), pointer :: kinmnrl_logK(:) => null()
    real(dp), pointer :: kinmnrl_logKcoef(:,:) => null()   ! (5, nkinmnrl)
    real(dp), pointer :: kinmnrl_rate_constant(:) => null()
    real(dp), pointer :: kinmnrl_activation_energy(:) => null()
    real(dp), pointer :: kinmnrl_molar_vol(:) => null()
    real(dp), pointer :: kinmnrl_Temkin_const(:) => null()
    real(dp), pointer :: kinmnrl_min_scale_factor(:) => null()
    real(dp), pointer :: kinmnrl_affinity_power(:) => null()
    real(dp), pointer :: kinmnrl_affinity_threshold(:) => null()
    real(dp), pointer :: kinmnrl_rate_limiter(:) => null()
    integer, pointer :: kinmnrl_irreversible(:) => null()
    ! rate prefactors (PREFACTOR blocks)
    integer, pointer :: kinmnrl_num_prefactors(:) => null()          ! (nkinmnrl)
    integer, pointer :: kinmnrl_prefactor_id(:,:,:) => null()        ! (0:maxspec, maxpref, nkinmnrl)
    real(dp), pointer :: kinmnrl_pref_alpha(:,:,:) => null()         ! (maxspec, maxpref, nkinmnrl)
    real(dp), pointer :: kinmnrl_pref_beta(:,:,:) => null()
    real(dp), pointer :: kinmnrl_pref_atten_coef(:,:,:) => null()
    real(dp), pointer :: kinmnrl_pref_rate(:,:) => null()            ! (maxpref, nkinmnrl)
    real(dp), pointer :: kinmnrl_pref_activation_energy(:,:) => null()
  end type mineral_type

  type, public :: immobile_type
    integer :: nimmobile = 0
    integer :: ndecay_rxn = 0
    integer, pointer :: decayspecid(:) => null()
    real(dp), pointer :: decay_rate_constant(:) => null()
  end type immobile_type

  type, public :: reaction_rt_type
    integer :: ncomp = 0
    integer :: naqcomp = 0
    integer :: offset_immobile = 0
    logical :: use_full_geochemistry = .false.
    logical :: use_log_formulation = .false.
    logical :: use_total_as_guess = .false.
    logical :: use_activity_h2o = .false.
    integer :: act_coef_update_frequency = PFCHEM_ACT_COEF_FREQUENCY_OFF
    integer :: act_coef_update_algorithm = PFCHEM_ACT_COEF_ALGORITHM_LAG
    integer :: maximum_reaction_cuts = 10
    integer :: maximum_reaction_iterations = 20
    real(dp) :: max_dlnC_rreact = 5.d0
    real(dp) :: max_relative_change_tolerance = 1.d-6
    real(dp) :: max_residual_tolerance = 1.d-12
    real(dp) :: max_rel_residual_tolerance = 1.d-8
    real(dp) :: debyeA = 0.5114d0
    real(dp) :: debyeB = 0.3288d0
    real(dp) :: debyeBdot = 0.0410d0
    real(dp), pointer :: primary_spec_a0(:) => null()
    real(dp), pointer :: primary_spec_Z(:) => null()
    integer :: neqcplx = 0
    integer, pointer :: eqcplxspecid(:,:) => null()        ! (0:maxsp, neqcplx)
    real(dp), pointer :: eqcplxstoich(:,:) => null()       ! (maxsp, neqcplx)
    integer, pointer :: eqcplxh2oid(:) => null()
    real(dp), pointer :: eqcplxh2ostoich(:) => null()
    real(dp), pointer :: eqcplx_a0(:) => null()
    real(dp), pointer :: eqcplx_Z(:) => null()
    real(dp), pointer :: eqcplx_logK(:) => null()
    real(dp), pointer :: eqcplx_logKcoef(:,:) => null()    ! (5, neqcplx)
    type(species_idx_type), pointer :: species_idx => null()
    type(mineral_type), pointer :: mineral => null()
    type(immobile_type), pointer :: immobile => null()
    ! SomDec sandbox setup, serialized with 0-based ids (target owned by the
    ! caller; see include/pfchem.h for the field meanings)
    type(pfchem_somdec), pointer :: somdec => null()
  end type reaction_rt_type

  type, public :: reactive_transport_auxvar_type
    real(dp), pointer :: pri_molal(:) => null()
    real(dp), pointer :: total(:,:) => null()               ! (naqcomp, nphase)
    real(dp), pointer :: sec_molal(:) => null()
    real(dp), pointer :: mnrl_volfrac(:) => null()
    real(dp), pointer :: mnrl_area(:) => null()
    real(dp), pointer :: mnrl_rate(:) => null()
    real(dp), pointer :: pri_act_coef(:) => null()
    real(dp), pointer :: sec_act_coef(:) => null()
    real(dp) :: ln_act_h2o = 0.d0
    real(dp), pointer :: immobile(:) => null()
  end type reactive_transport_auxvar_type

  type, public :: global_auxvar_type
    real(dp) :: temp = 25.d0
    real(dp), pointer :: sat(:) => null()
    real(dp), pointer :: den_kg(:) => null()
  end type global_auxvar_type

  type, public :: material_auxvar_type
    real(dp) :: volume = 1.d0
    real(dp) :: porosity = 1.d0
  end type material_auxvar_type

  type, public :: option_type
    real(dp) :: tran_dt = 0.d0
    logical :: use_isothermal = .false.
    integer :: device = 0
  end type option_type

  public :: RStep, RStepBatch, PfchemSetVariant, PfchemSetLaunch, PfchemFinalize, &
            PfchemLastError

  ! device context of the current reaction object, and the network arrays
  ! handed to pfchem_create (0-based ids)
  type(c_ptr), save :: ctx = c_null_ptr
  type(reaction_rt_type), pointer, save :: ctx_reaction => null()
  integer(c_int), save :: kernel_variant = PFCHEM_KERNEL_BATCHED_MIXED
  character(len=:), allocatable, save :: last_message

contains

  subroutine PfchemSetVariant(variant)
    ! kernel used by RStep / RStepBatch (PFCHEM_KERNEL_*)
    integer, intent(in) :: variant
    kernel_variant = int(variant, c_int)
  end subroutine PfchemSetVariant

  subroutine PfchemSetLaunch(variant, lanes_per_cell, rows_per_lane, block_size, blocks_per_cu, &
                             ierror)
    ! team shape of a batched variant for the current context
    integer, intent(in) :: variant, lanes_per_cell, rows_per_lane, block_size, blocks_per_cu
    integer, intent(out) :: ierror
    type(pfchem_launch) :: cfg
    cfg%lanes_per_cell = lanes_per_cell
    cfg%rows_per_lane = rows_per_lane
    cfg%block_size = block_size
    cfg%blocks_per_cu = blocks_per_cu
    ierror = 1
    if (.not. c_associated(ctx)) return
    ierror = int(pfchem_set_launch(ctx, int(variant, c_int), cfg))
    if (ierror /= 0) last_message = pfchem_string(pfchem_last_error(ctx))
  end subroutine PfchemSetLaunch

  function PfchemLastError() result(msg)
    character(len=:), allocatable :: msg
    if (allocated(last_message)) then
      msg = last_message
    else
      msg = ''
    end if
  end function PfchemLastError

  subroutine PfchemFinalize()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = pfchem_destroy(ctx)
    ctx = c_null_ptr
    nullify(ctx_reaction)
  end subroutine PfchemFinalize

  ! ------------------------------------------------------------------
  ! network upload (ReactionCreate / BasisInit output -> pfchem_network)
  ! ------------------------------------------------------------------
  subroutine PfchemContext(reaction, option, ierror)
    type(reaction_rt_type), target, intent(in) :: reaction
    type(option_type), intent(in) :: option
    integer, intent(out) :: ierror

    type(pfchem_network) :: net
    type(pfchem_options) :: opt
    integer(c_int32_t), allocatable, target, save :: cspec(:,:), mspec(:,:), dspec(:), irrev(:)
    integer(c_int32_t), allocatable, target, save :: pnum(:), pid(:,:,:)
    real(dp), allocatable, target, save :: ch2o(:), mh2o(:)
    integer :: k, ms, nimm, mp, msp
    integer(c_int) :: rc

    ierror = 0
    if (c_associated(ctx) .and. associated(ctx_reaction, reaction)) return
    call PfchemFinalize()

    nimm = 0
    if (associated(reaction%immobile)) nimm = reaction%immobile%nimmobile
    net%naqcomp = reaction%naqcomp
    net%nimmobile = nimm
    net%h2o_aq_id = -1
    if (associated(reaction%species_idx)) net%h2o_aq_id = reaction%species_idx%h2o_aq_id - 1
    net%pad0 = 0
    net%primary_spec_Z = c_loc(reaction%primary_spec_Z)
    net%primary_spec_a0 = c_loc(reaction%primary_spec_a0)

    ! complexes: (0:maxsp, neqcplx) count + 1-based ids -> 0-based ids
    net%neqcplx = reaction%neqcplx
    net%eqcplx_maxsp = 0
    net%eqcplxspecid = c_null_ptr
    net%eqcplxstoich = c_null_ptr
    net%eqcplxh2ostoich = c_null_ptr
    net%eqcplx_logK = c_null_ptr
    net%eqcplx_logKcoef = c_null_ptr
    net%eqcplx_Z = c_null_ptr
    net%eqcplx_a0 = c_null_ptr
    if (allocated(cspec)) deallocate(cspec)
    if (allocated(ch2o)) deallocate(ch2o)
    if (reaction%neqcplx > 0) then
      ms = size(reaction%eqcplxstoich, 1)
      allocate(cspec(0:ms, reaction%neqcplx))
      cspec(0, :) = reaction%eqcplxspecid(0, 1:reaction%neqcplx)
      cspec(1:ms, :) = reaction%eqcplxspecid(1:ms, 1:reaction%neqcplx) - 1
      do k = 1, reaction%neqcplx
        cspec(cspec(0, k) + 1:ms, k) = 0
      end do
      net%eqcplx_maxsp = ms
      net%eqcplxspecid = c_loc(cspec)
      net%eqcplxstoich = c_loc(reaction%eqcplxstoich)
      if (associated(reaction%eqcplxh2oid)) then
        allocate(ch2o(reaction%neqcplx))
        ch2o = 0.d0
        where (reaction%eqcplxh2oid(1:reaction%neqcplx) > 0) &
          ch2o = reaction%eqcplxh2ostoich(1:reaction%neqcplx)
        net%eqcplxh2ostoich = c_loc(ch2o)
      end if
      net%eqcplx_logK = c_loc(reaction%eqcplx_logK)
      if (associated(reaction%eqcplx_logKcoef)) net%eqcplx_logKcoef = c_loc(reaction%eqcplx_logKcoef)
      net%eqcplx_Z = c_loc(reaction%eqcplx_Z)
      net%eqcplx_a0 = c_loc(reaction%eqcplx_a0)
    end if

    ! kinetic minerals
    net%nkinmnrl = 0
    net%kinmnrl_maxsp = 0
    net%kinmnrlspecid = c_null_ptr
    net%kinmnrlstoich = c_null_ptr
    net%kinmnrlh2ostoich = c_null_ptr
    net%kinmnrl_logK = c_null_ptr
    net%kinmnrl_logKcoef = c_null_ptr
    net%kinmnrl_rate_constant = c_null_ptr
    net%kinmnrl_activation_energy = c_null_ptr
    net%kinmnrl_molar_vol = c_null_ptr
    net%kinmnrl_Temkin_const = c_null_ptr
    net%kinmnrl_min_scale_factor = c_null_ptr
    net%kinmnrl_affinity_power = c_null_ptr
    net%kinmnrl_affinity_threshold = c_null_ptr
    net%kinmnrl_rate_limiter = c_null_ptr
    net%kinmnrl_irreversible = c_null_ptr
    net%kinmnrl_max_pref = 0
    net%kinmnrl_max_pref_spec = 0
    net%kinmnrl_num_prefactors = c_null_ptr
    net%kinmnrl_prefactor_id = c_null_ptr
    net%kinmnrl_pref_alpha = c_null_ptr
    net%kinmnrl_pref_beta = c_null_ptr
    net%kinmnrl_pref_atten_coef = c_null_ptr
    net%kinmnrl_pref_rate = c_null_ptr
    net%kinmnrl_pref_activation_energy = c_null_ptr
    if (allocated(pnum)) deallocate(pnum)
    if (allocated(pid)) deallocate(pid)
    if (allocated(mspec)) deallocate(mspec)
    if (allocated(mh2o)) deallocate(mh2o)
    if (allocated(irrev)) deallocate(irrev)
    if (associated(reaction%mineral)) then
      associate(m => reaction%mineral)
        if (m%nkinmnrl > 0) then
          ms = size(m%kinmnrlstoich, 1)
          allocate(mspec(0:ms, m%nkinmnrl))
          mspec(0, :) = m%kinmnrlspecid(0, 1:m%nkinmnrl)
          mspec(1:ms, :) = m%kinmnrlspecid(1:ms, 1:m%nkinmnrl) - 1
          do k = 1, m%nkinmnrl
            mspec(mspec(0, k) + 1:ms, k) = 0
          end do
          net%nkinmnrl = m%nkinmnrl
          net%kinmnrl_maxsp = ms
          net%kinmnrlspecid = c_loc(mspec)
          net%kinmnrlstoich = c_loc(m%kinmnrlstoich)
          if (associated(m%kinmnrlh2oid)) then
            allocate(mh2o(m%nkinmnrl))
            mh2o = 0.d0
            where (m%kinmnrlh2oid(1:m%nkinmnrl) > 0) mh2o = m%kinmnrlh2ostoich(1:m%nkinmnrl)
            net%kinmnrlh2ostoich = c_loc(mh2o)
          end if
          net%kinmnrl_logK = c_loc(m%kinmnrl_logK)
          if (associated(m%kinmnrl_logKcoef)) net%kinmnrl_logKcoef = c_loc(m%kinmnrl_logKcoef)
          net%kinmnrl_rate_constant = c_loc(m%kinmnrl_rate_constant)
          if (associated(m%kinmnrl_activation_energy)) &
            net%kinmnrl_activation_energy = c_loc(m%kinmnrl_activation_energy)
          net%kinmnrl_molar_vol = c_loc(m%kinmnrl_molar_vol)
          if (associated(m%kinmnrl_Temkin_const)) net%kinmnrl_Temkin_const = c_loc(m%kinmnrl_Temkin_const)
          if (associated(m%kinmnrl_min_scale_factor)) &
            net%kinmnrl_min_scale_factor = c_loc(m%kinmnrl_min_scale_factor)
          if (associated(m%kinmnrl_affinity_power)) &
            net%kinmnrl_affinity_power = c_loc(m%kinmnrl_affinity_power)
          if (associated(m%kinmnrl_affinity_threshold)) &
            net%kinmnrl_affinity_threshold = c_loc(m%kinmnrl_affinity_threshold)
          if (associated(m%kinmnrl_rate_limiter)) net%kinmnrl_rate_limiter = c_loc(m%kinmnrl_rate_limiter)
          if (associated(m%kinmnrl_irreversible)) then
            allocate(irrev(m%nkinmnrl))
            irrev = int(m%kinmnrl_irreversible(1:m%nkinmnrl), c_int32_t)
            net%kinmnrl_irreversible = c_loc(irrev)
          end if
          ! rate prefactors: same memory order; primary ids -> 0-based
          ! (negative ids name complexes and stay negative)
          if (associated(m%kinmnrl_num_prefactors)) then
            if (maxval(m%kinmnrl_num_prefactors(1:m%nkinmnrl)) > 0) then
              msp = size(m%kinmnrl_prefactor_id, 1) - 1
              mp = size(m%kinmnrl_prefactor_id, 2)
              allocate(pnum(m%nkinmnrl))
              pnum = int(m%kinmnrl_num_prefactors(1:m%nkinmnrl), c_int32_t)
              allocate(pid(0:msp, mp, m%nkinmnrl))
              pid = int(m%kinmnrl_prefactor_id(:, :, 1:m%nkinmnrl), c_int32_t)
              where (pid(1:msp, :, :) > 0) pid(1:msp, :, :) = pid(1:msp, :, :) - 1
              net%kinmnrl_max_pref = mp
              net%kinmnrl_max_pref_spec = msp
              net%kinmnrl_num_prefactors = c_loc(pnum)
              net%kinmnrl_prefactor_id = c_loc(pid)
              net%kinmnrl_pref_alpha = c_loc(m%kinmnrl_pref_alpha)
              net%kinmnrl_pref_beta = c_loc(m%kinmnrl_pref_beta)
              net%kinmnrl_pref_atten_coef = c_loc(m%kinmnrl_pref_atten_coef)
              net%kinmnrl_pref_rate = c_loc(m%kinmnrl_pref_rate)
              if (associated(m%kinmnrl_pref_activation_energy)) &
                net%kinmnrl_pref_activation_energy = c_loc(m%kinmnrl_pref_activation_energy)
            end if
          end if
        end if
      end associate
    end if

    ! immobile decay (immobile ids -> 0-based)
    net%ndecay_rxn = 0
    net%pad1 = 0
    net%decayspecid = c_null_ptr
    net%decay_rate_constant = c_null_ptr
    if (allocated(dspec)) deallocate(dspec)
    if (associated(reaction%immobile)) then
      if (reaction%immobile%ndecay_rxn > 0) then
        allocate(dspec(reaction%immobile%ndecay_rxn))
        dspec = int(reaction%immobile%decayspecid(1:reaction%immobile%ndecay_rxn) - 1, c_int32_t)
        net%ndecay_rxn = reaction%immobile%ndecay_rxn
        net%decayspecid = c_loc(dspec)
        net%decay_rate_constant = c_loc(reaction%immobile%decay_rate_constant)
      end if
    end if

    net%debyeA = reaction%debyeA
    net%debyeB = reaction%debyeB
    net%debyeBdot = reaction%debyeBdot
    net%somdec = c_null_ptr
    if (associated(reaction%somdec)) net%somdec = c_loc(reaction%somdec)

    opt%use_full_geochemistry = merge(1, 0, reaction%use_full_geochemistry)
    opt%use_log_formulation = merge(1, 0, reaction%use_log_formulation)
    opt%act_coef_update_frequency = reaction%act_coef_update_frequency
    opt%act_coef_update_algorithm = reaction%act_coef_update_algorithm
    opt%use_activity_h2o = merge(1, 0, reaction%use_activity_h2o)
    opt%use_total_as_guess = merge(1, 0, reaction%use_total_as_guess)
    opt%isothermal = merge(1, 0, option%use_isothermal)
    opt%maximum_reaction_iterations = reaction%maximum_reaction_iterations
    opt%maximum_reaction_cuts = reaction%maximum_reaction_cuts
    opt%variable_ncomp = 0
    opt%maximum_rreact_calls = 0
    opt%max_dlnC_rreact = reaction%max_dlnC_rreact
    opt%max_residual_tolerance = reaction%max_residual_tolerance
    opt%max_rel_residual_tolerance = reaction%max_rel_residual_tolerance
    opt%max_relative_change_tolerance = reaction%max_relative_change_tolerance

    rc = pfchem_create(net, opt, int(option%device, c_int), ctx)
    if (rc /= PFCHEM_OK) then
      last_message = 'pfchem_create: ' // pfchem_string(pfchem_strerror(rc))
      ctx = c_null_ptr
      ierror = 1
      return
    end if
    ctx_reaction => reaction
  end subroutine PfchemContext

  ! ------------------------------------------------------------------
  ! RStepBatch: the OSRT cell loop as one call
  ! ------------------------------------------------------------------
  subroutine RStepBatch(reaction, rt_auxvars, global_auxvars, material_auxvars, cell_ids, &
                        option, ierror, num_iterations, num_sub_steps, cell_status)
    type(reaction_rt_type), target, intent(in) :: reaction
    type(reactive_transport_auxvar_type), intent(inout) :: rt_auxvars(:)
    type(global_auxvar_type), intent(in) :: global_auxvars(:)
    type(material_auxvar_type), intent(in) :: material_auxvars(:)
    integer, intent(in) :: cell_ids(:)        ! indices into the auxvar arrays
    type(option_type), intent(in) :: option
    integer, intent(out) :: ierror
    integer, intent(out), optional :: num_iterations(:), num_sub_steps(:), cell_status(:)

    real(dp), allocatable, target :: pri(:,:), tot(:,:), imm(:,:), sec(:,:), gp(:,:), gs(:,:), &
                                     law(:), vf(:,:), area(:,:), rate(:,:), por(:), sat(:), &
                                     den(:), vol(:), temp(:)
    integer(c_int32_t), allocatable, target :: nit(:), nsub(:), status(:)
    type(pfchem_state) :: st
    integer :: n, naq, nimm, nc, nm, j, id
    integer(c_int) :: rc

    ierror = 0
    n = size(cell_ids)
    if (n == 0) return
    call PfchemContext(reaction, option, ierror)
    if (ierror /= 0) return
    naq = reaction%naqcomp
    nimm = 0
    if (associated(reaction%immobile)) nimm = reaction%immobile%nimmobile
    nc = reaction%neqcplx
    nm = 0
    if (associated(reaction%mineral)) nm = reaction%mineral%nkinmnrl

    ! SoA, field-major: element (k, cell) at k*n + cell (0-based) = (cell, k)
    allocate(pri(n, max(naq, 1)), tot(n, max(naq, 1)), gp(n, max(naq, 1)))
    allocate(imm(n, max(nimm, 1)), sec(n, max(nc, 1)), gs(n, max(nc, 1)))
    allocate(vf(n, max(nm, 1)), area(n, max(nm, 1)), rate(n, max(nm, 1)))
    allocate(law(n), por(n), sat(n), den(n), vol(n), temp(n), nit(n), nsub(n), status(n))
    do j = 1, n
      id = cell_ids(j)
      associate(rt => rt_auxvars(id), gl => global_auxvars(id), ma => material_auxvars(id))
        pri(j, 1:naq) = rt%pri_molal(1:naq)
        tot(j, 1:naq) = rt%total(1:naq, 1)
        gp(j, 1:naq) = rt%pri_act_coef(1:naq)
        if (nimm > 0) imm(j, 1:nimm) = rt%immobile(1:nimm)
        if (nc > 0) then
          sec(j, 1:nc) = rt%sec_molal(1:nc)
          gs(j, 1:nc) = rt%sec_act_coef(1:nc)
        end if
        if (nm > 0) then
          vf(j, 1:nm) = rt%mnrl_volfrac(1:nm)
          area(j, 1:nm) = rt%mnrl_area(1:nm)
          rate(j, 1:nm) = rt%mnrl_rate(1:nm)
        end if
        law(j) = rt%ln_act_h2o
        por(j) = ma%porosity
        sat(j) = gl%sat(1)
        den(j) = gl%den_kg(1)
        vol(j) = ma%volume
        temp(j) = gl%temp
      end associate
    end do

    st%ld = int(n, c_int64_t)
    st%pri_molal = c_loc(pri)
    st%total = c_loc(tot)
    st%pri_act_coef = c_loc(gp)
    st%immobile = c_null_ptr
    if (nimm > 0) st%immobile = c_loc(imm)
    st%sec_molal = c_null_ptr
    st%sec_act_coef = c_null_ptr
    if (nc > 0) then
      st%sec_molal = c_loc(sec)
      st%sec_act_coef = c_loc(gs)
    end if
    st%ln_act_h2o = c_loc(law)
    st%mnrl_volfrac = c_null_ptr
    st%mnrl_area = c_null_ptr
    st%mnrl_rate = c_null_ptr
    if (nm > 0) then
      st%mnrl_volfrac = c_loc(vf)
      st%mnrl_area = c_loc(area)
      st%mnrl_rate = c_loc(rate)
    end if
    st%porosity = c_loc(por)
    st%saturation = c_loc(sat)
    st%den_kg = c_loc(den)
    st%volume = c_loc(vol)
    st%temperature = c_loc(temp)
    st%somdec_f_t = c_null_ptr
    st%somdec_f_w = c_null_ptr
    st%num_iterations = c_loc(nit)
    st%num_sub_steps = c_loc(nsub)
    st%status = c_loc(status)
    st%ncomp_active = c_null_ptr

    rc = pfchem_step_host(ctx, int(n, c_int64_t), st, real(option%tran_dt, c_double), &
                          kernel_variant)
    if (rc /= PFCHEM_OK .and. rc /= PFCHEM_ERR_CELLS) then
      last_message = 'pfchem_step_host: ' // pfchem_string(pfchem_last_error(ctx))
      ierror = 1
      return
    end if
    if (rc == PFCHEM_ERR_CELLS) ierror = 1

    do j = 1, n
      id = cell_ids(j)
      associate(rt => rt_auxvars(id))
        rt%pri_molal(1:naq) = pri(j, 1:naq)
        rt%total(1:naq, 1) = tot(j, 1:naq)
        rt%pri_act_coef(1:naq) = gp(j, 1:naq)
        if (nimm > 0) rt%immobile(1:nimm) = imm(j, 1:nimm)
        if (nc > 0) then
          rt%sec_molal(1:nc) = sec(j, 1:nc)
          rt%sec_act_coef(1:nc) = gs(j, 1:nc)
        end if
        if (nm > 0) then
          rt%mnrl_volfrac(1:nm) = vf(j, 1:nm)
          rt%mnrl_rate(1:nm) = rate(j, 1:nm)
        end if
        rt%ln_act_h2o = law(j)
      end associate
    end do
    if (present(num_iterations)) num_iterations(1:n) = nit(1:n)
    if (present(num_sub_steps)) num_sub_steps(1:n) = nsub(1:n)
    if (present(cell_status)) cell_status(1:n) = status(1:n)
  end subroutine RStepBatch

  ! ------------------------------------------------------------------
  ! RStep: the reference's per-cell entry (reaction.F90:3564-3566)
  ! ------------------------------------------------------------------
  subroutine RStep(guess, rt_auxvar, global_auxvar, material_auxvar, num_sub_steps, &
                   num_iterations, num_kinetic_state_updates, reaction, natural_id, option, &
                   ierror)
    real(dp), intent(inout) :: guess(:)
    type(reactive_transport_auxvar_type), intent(inout) :: rt_auxvar
    type(global_auxvar_type), intent(in) :: global_auxvar
    type(material_auxvar_type), intent(in) :: material_auxvar
    integer, intent(out) :: num_sub_steps, num_iterations, num_kinetic_state_updates
    type(reaction_rt_type), target, intent(in) :: reaction
    integer, intent(in) :: natural_id
    type(option_type), intent(in) :: option
    integer, intent(out) :: ierror

    type(reactive_transport_auxvar_type) :: rt(1)
    type(global_auxvar_type) :: gl(1)
    type(material_auxvar_type) :: ma(1)
    integer :: nit(1), nsub(1), naq, nimm, nm

    naq = reaction%naqcomp
    nimm = 0
    if (associated(reaction%immobile)) nimm = reaction%immobile%nimmobile
    nm = 0
    if (associated(reaction%mineral)) nm = reaction%mineral%nkinmnrl
    ! the guess is the Newton starting point of the aqueous (and immobile)
    ! unknowns, as in RStep's guess argument
    rt(1) = rt_auxvar
    rt(1)%pri_molal(1:naq) = guess(1:naq)
    if (nimm > 0 .and. size(guess) >= naq + nimm) rt(1)%immobile(1:nimm) = guess(naq + 1:naq + nimm)
    gl(1) = global_auxvar
    ma(1) = material_auxvar
    call RStepBatch(reaction, rt, gl, ma, [1], option, ierror, nit, nsub)
    guess(1:naq) = rt_auxvar%pri_molal(1:naq)
    if (nimm > 0 .and. size(guess) >= naq + nimm) guess(naq + 1:naq + nimm) = rt_auxvar%immobile(1:nimm)
    num_iterations = nit(1)
    num_sub_steps = nsub(1)
    num_kinetic_state_updates = 0
    if (nm > 0 .and. ierror == 0) num_kinetic_state_updates = nsub(1)
  end subroutine RStep

end module pfchem_reaction_module
