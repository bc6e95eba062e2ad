! pfchem_c.F90 -- ISO_C_BINDING view of include/pfchem.h for Fortran hosts.
!
! bind(C) derived types mirror the C structs field by field (array members are
! type(c_ptr) obtained with c_loc; NULL means "not associated"), and the
! interfaces bind the C entry points by name.  Species ids in these structs
! are 0-based, as in C; pfchem_reaction_module converts from the reference's
! 1-based ids.
module pfchem_c
  use, intrinsic :: iso_c_binding
  implicit none

  ! return codes
  integer(c_int), parameter :: PFCHEM_OK = 0
  integer(c_int), parameter :: PFCHEM_ERR_INVALID = 1
  integer(c_int), parameter :: PFCHEM_ERR_HIP = 2
  integer(c_int), parameter :: PFCHEM_ERR_UNSUPPORTED = 3
  integer(c_int), parameter :: PFCHEM_ERR_NOMEM = 4
  integer(c_int), parameter :: PFCHEM_ERR_CELLS = 5

  ! per-cell status
  integer(c_int), parameter :: PFCHEM_CELL_OK = 0
  integer(c_int), parameter :: PFCHEM_CELL_MAX_CUTS_ITERATIONS = 1
  integer(c_int), parameter :: PFCHEM_CELL_MAX_CUTS_SINGULAR = 2
  integer(c_int), parameter :: PFCHEM_CELL_MAX_CUTS_ACTIVITY = 3
  integer(c_int), parameter :: PFCHEM_CELL_CONSTRAINT_SINGULAR = 6
  integer(c_int), parameter :: PFCHEM_CELL_CONSTRAINT_NONPOSITIVE = 7
  integer(c_int), parameter :: PFCHEM_CELL_CONSTRAINT_MAX_ITERATIONS = 8

  ! constraint types (transport_constraint_rt.F90:22-34)
  integer(c_int), parameter :: PFCHEM_CONSTRAINT_NULL = 0
  integer(c_int), parameter :: PFCHEM_CONSTRAINT_FREE = 1
  integer(c_int), parameter :: PFCHEM_CONSTRAINT_TOTAL = 2
  integer(c_int), parameter :: PFCHEM_CONSTRAINT_LOG = 3
  integer(c_int), parameter :: PFCHEM_CONSTRAINT_PH = 4
  integer(c_int), parameter :: PFCHEM_CONSTRAINT_MINERAL = 7
  integer(c_int), parameter :: PFCHEM_CONSTRAINT_GAS = 8
  integer(c_int), parameter :: PFCHEM_CONSTRAINT_CHARGE_BAL = 9

  ! activity coefficient options (the reaction_aux.F90:27-31 values)
  integer(c_int), parameter :: PFCHEM_ACT_COEF_FREQUENCY_OFF = 0
  integer(c_int), parameter :: PFCHEM_ACT_COEF_FREQUENCY_TIMESTEP = 1
  integer(c_int), parameter :: PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER = 2
  integer(c_int), parameter :: PFCHEM_ACT_COEF_ALGORITHM_LAG = 3
  integer(c_int), parameter :: PFCHEM_ACT_COEF_ALGORITHM_NEWTON = 4

  ! kernel variants
  integer(c_int), parameter :: PFCHEM_KERNEL_PLAIN_FP64 = 0
  integer(c_int), parameter :: PFCHEM_KERNEL_BATCHED_FP64 = 1
  integer(c_int), parameter :: PFCHEM_KERNEL_BATCHED_MIXED = 2

  integer, parameter :: PFCHEM_SOMDEC_MAX_DOWN = 4
  integer, parameter :: PFCHEM_SOMDEC_MAX_MONOD = 4
  integer, parameter :: PFCHEM_SOMDEC_MAX_INHIB = 4

  type, bind(C) :: pfchem_network
    integer(c_int32_t) :: naqcomp
    integer(c_int32_t) :: nimmobile
    integer(c_int32_t) :: h2o_aq_id
    integer(c_int32_t) :: pad0
    type(c_ptr) :: primary_spec_Z
    type(c_ptr) :: primary_spec_a0
    integer(c_int32_t) :: neqcplx
    integer(c_int32_t) :: eqcplx_maxsp
    type(c_ptr) :: eqcplxspecid
    type(c_ptr) :: eqcplxstoich
    type(c_ptr) :: eqcplxh2ostoich
    type(c_ptr) :: eqcplx_logK
    type(c_ptr) :: eqcplx_logKcoef
    type(c_ptr) :: eqcplx_Z
    type(c_ptr) :: eqcplx_a0
    integer(c_int32_t) :: nkinmnrl
    integer(c_int32_t) :: kinmnrl_maxsp
    type(c_ptr) :: kinmnrlspecid
    type(c_ptr) :: kinmnrlstoich
    type(c_ptr) :: kinmnrlh2ostoich
    type(c_ptr) :: kinmnrl_logK
    type(c_ptr) :: kinmnrl_logKcoef
    type(c_ptr) :: kinmnrl_rate_constant
    type(c_ptr) :: kinmnrl_activation_energy
    type(c_ptr) :: kinmnrl_molar_vol
    type(c_ptr) :: kinmnrl_Temkin_const
    type(c_ptr) :: kinmnrl_min_scale_factor
    type(c_ptr) :: kinmnrl_affinity_power
    type(c_ptr) :: kinmnrl_affinity_threshold
    type(c_ptr) :: kinmnrl_rate_limiter
    type(c_ptr) :: kinmnrl_irreversible
    integer(c_int32_t) :: ndecay_rxn
    integer(c_int32_t) :: pad1
    type(c_ptr) :: decayspecid
    type(c_ptr) :: decay_rate_constant
    real(c_double) :: debyeA
    real(c_double) :: debyeB
    real(c_double) :: debyeBdot
    type(c_ptr) :: somdec
    ! kinetic-mineral rate prefactors (mineral%kinmnrl_num_prefactors ...)
    integer(c_int32_t) :: kinmnrl_max_pref, kinmnrl_max_pref_spec
    type(c_ptr) :: kinmnrl_num_prefactors, kinmnrl_prefactor_id
    type(c_ptr) :: kinmnrl_pref_alpha, kinmnrl_pref_beta, kinmnrl_pref_atten_coef
    type(c_ptr) :: kinmnrl_pref_rate, kinmnrl_pref_activation_energy
  end type pfchem_network

  type, bind(C) :: pfchem_options
    integer(c_int32_t) :: use_full_geochemistry
    integer(c_int32_t) :: use_log_formulation
    integer(c_int32_t) :: act_coef_update_frequency
    integer(c_int32_t) :: act_coef_update_algorithm
    integer(c_int32_t) :: use_activity_h2o
    integer(c_int32_t) :: use_total_as_guess
    integer(c_int32_t) :: isothermal
    integer(c_int32_t) :: maximum_reaction_iterations
    integer(c_int32_t) :: maximum_reaction_cuts
    integer(c_int32_t) :: variable_ncomp
    integer(c_int32_t) :: maximum_rreact_calls = 0
    integer(c_int32_t) :: pad_options = 0
    real(c_double) :: max_dlnC_rreact
    real(c_double) :: max_residual_tolerance
    real(c_double) :: max_rel_residual_tolerance
    real(c_double) :: max_relative_change_tolerance
  end type pfchem_options

  type, bind(C) :: pfchem_somdec
    integer(c_int32_t) :: nrxn
    integer(c_int32_t) :: species_id_co2
    integer(c_int32_t) :: species_id_nh4
    integer(c_int32_t) :: species_id_no3
    integer(c_int32_t) :: species_id_n2o
    integer(c_int32_t) :: species_id_proton
    integer(c_int32_t) :: species_id_hr
    integer(c_int32_t) :: species_id_nmin
    integer(c_int32_t) :: species_id_nimm
    integer(c_int32_t) :: species_id_nimp
    integer(c_int32_t) :: species_id_ngasmin
    integer(c_int32_t) :: pad0
    real(c_double) :: x0eps
    real(c_double) :: n2o_frac_mineralization
    real(c_double) :: inhibition_nh4_no3
    type(c_ptr) :: rate_constant
    type(c_ptr) :: rate_decomposition
    type(c_ptr) :: rate_ad_factor
    type(c_ptr) :: upstream_c_id
    type(c_ptr) :: upstream_n_id
    type(c_ptr) :: upstream_hr_id
    type(c_ptr) :: upstream_nmin_id
    type(c_ptr) :: upstream_nimp_id
    type(c_ptr) :: upstream_nimm_id
    type(c_ptr) :: upstream_nc
    type(c_ptr) :: upstream_is_varycn
    type(c_ptr) :: n_downstream_pools
    type(c_ptr) :: downstream_c_id
    type(c_ptr) :: downstream_n_id
    type(c_ptr) :: downstream_stoich
    type(c_ptr) :: downstream_nc
    type(c_ptr) :: downstream_is_varycn
    type(c_ptr) :: mineral_c_stoich
    type(c_ptr) :: mineral_n_stoich
    type(c_ptr) :: temperature_response_function
    type(c_ptr) :: Q10
    type(c_ptr) :: Ea
    type(c_ptr) :: moisture_response_function
    type(c_ptr) :: Ox_response_function
    type(c_ptr) :: Ox_half_saturation
    type(c_ptr) :: n_monod
    type(c_ptr) :: monod_specid
    type(c_ptr) :: monod_half_saturation
    type(c_ptr) :: monod_threshold
    type(c_ptr) :: monod_pool_normalized
    type(c_ptr) :: n_inhibition
    type(c_ptr) :: inhibition_specid
    type(c_ptr) :: inhibition_type
    type(c_ptr) :: inhibition_constant
    type(c_ptr) :: inhibition_constant2
    ! oxygen, 1-based codes (aqueous id + 1, -(immobile id + 1), 0 = none)
    type(c_ptr) :: Ox_specid = c_null_ptr
    integer(c_int32_t) :: species_o2 = 0
    integer(c_int32_t) :: pad1 = 0
  end type pfchem_somdec

  type, bind(C) :: pfchem_state
    integer(c_int64_t) :: ld
    type(c_ptr) :: pri_molal
    type(c_ptr) :: total
    type(c_ptr) :: immobile
    type(c_ptr) :: sec_molal
    type(c_ptr) :: pri_act_coef
    type(c_ptr) :: sec_act_coef
    type(c_ptr) :: ln_act_h2o
    type(c_ptr) :: mnrl_volfrac
    type(c_ptr) :: mnrl_area
    type(c_ptr) :: mnrl_rate
    type(c_ptr) :: porosity
    type(c_ptr) :: saturation
    type(c_ptr) :: den_kg
    type(c_ptr) :: volume
    type(c_ptr) :: temperature
    type(c_ptr) :: somdec_f_t
    type(c_ptr) :: somdec_f_w
    type(c_ptr) :: num_iterations
    type(c_ptr) :: num_sub_steps
    type(c_ptr) :: status
    type(c_ptr) :: ncomp_active
  end type pfchem_state

  type, bind(C) :: pfchem_launch
    integer(c_int32_t) :: lanes_per_cell
    integer(c_int32_t) :: block_size
    integer(c_int32_t) :: blocks_per_cu
    integer(c_int32_t) :: rows_per_lane
  end type pfchem_launch

  ! one aqueous constraint (pfchem_constraint); 0-based ids
  type, bind(C) :: pfchem_constraint
    type(c_ptr) :: constraint_type
    type(c_ptr) :: constraint_rxn
    integer(c_int32_t) :: nrxn
    integer(c_int32_t) :: rxn_maxsp
    type(c_ptr) :: rxnspecid
    type(c_ptr) :: rxnstoich
    type(c_ptr) :: rxnh2ostoich
    type(c_ptr) :: rxn_logK
    type(c_ptr) :: rxn_logKcoef
    integer(c_int32_t) :: h_ion_id
    integer(c_int32_t) :: initialize_with_molality
    integer(c_int32_t) :: use_prev_soln_as_guess
    integer(c_int32_t) :: max_iterations
    real(c_double) :: max_dlnC
    type(c_ptr) :: free_ion_guess
  end type pfchem_constraint

  ! OSRT transport: connection topology (host arrays, 0-based cells) and the
  ! per-step coefficients (device arrays); pfchem.h pfchem_connections /
  ! pfchem_transport_coefs
  type, bind(C) :: pfchem_connections
    integer(c_int64_t) :: nconn
    type(c_ptr) :: id_up
    type(c_ptr) :: id_dn
    integer(c_int64_t) :: nbc
    type(c_ptr) :: bc_cell
  end type pfchem_connections

  type, bind(C) :: pfchem_transport_coefs
    type(c_ptr) :: area
    type(c_ptr) :: velocity
    type(c_ptr) :: coef_over_dist
    type(c_ptr) :: bc_area
    type(c_ptr) :: bc_velocity
    type(c_ptr) :: bc_coef_over_dist
    type(c_ptr) :: bc_total
  end type pfchem_transport_coefs

  interface
    integer(c_int) function pfchem_create(net, opt, device, ctx) bind(C, name="pfchem_create")
      import :: c_int, c_ptr, pfchem_network, pfchem_options
      type(pfchem_network), intent(in) :: net
      type(pfchem_options), intent(in) :: opt
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function pfchem_create

    integer(c_int) function pfchem_destroy(ctx) bind(C, name="pfchem_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function pfchem_destroy

    type(c_ptr) function pfchem_strerror(code) bind(C, name="pfchem_strerror")
      import :: c_int, c_ptr
      integer(c_int), value :: code
    end function pfchem_strerror

    type(c_ptr) function pfchem_last_error(ctx) bind(C, name="pfchem_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function pfchem_last_error

    integer(c_int) function pfchem_step(ctx, ncell, state, dt, variant, stream) &
        bind(C, name="pfchem_step")
      import :: c_int, c_int64_t, c_double, c_ptr, pfchem_state
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ncell
      type(pfchem_state), intent(in) :: state
      real(c_double), value :: dt
      integer(c_int), value :: variant
      type(c_ptr), value :: stream
    end function pfchem_step

    integer(c_int) function pfchem_step_host(ctx, ncell, state, dt, variant) &
        bind(C, name="pfchem_step_host")
      import :: c_int, c_int64_t, c_double, c_ptr, pfchem_state
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ncell
      type(pfchem_state), intent(in) :: state
      real(c_double), value :: dt
      integer(c_int), value :: variant
    end function pfchem_step_host

    ! GIRT reaction block of one iterate: res(i*ld+cell), jac((i*ncomp+j)*ld+cell)
    integer(c_int) function pfchem_residual_jacobian(ctx, ncell, state, fixed_total, &
        fixed_immobile, dt, res, jac, stream) bind(C, name="pfchem_residual_jacobian")
      import :: c_int, c_int64_t, c_double, c_ptr, pfchem_state
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ncell
      type(pfchem_state), intent(in) :: state
      type(c_ptr), value :: fixed_total
      type(c_ptr), value :: fixed_immobile
      real(c_double), value :: dt
      type(c_ptr), value :: res
      type(c_ptr), value :: jac
      type(c_ptr), value :: stream
    end function pfchem_residual_jacobian

    integer(c_int) function pfchem_equilibrate(ctx, ncell, state, con, conc, stream) &
        bind(C, name="pfchem_equilibrate")
      import :: c_int, c_int64_t, c_ptr, pfchem_state, pfchem_constraint
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ncell
      type(pfchem_state), intent(in) :: state
      type(pfchem_constraint), intent(in) :: con
      type(c_ptr), value :: conc
      type(c_ptr), value :: stream
    end function pfchem_equilibrate

    integer(c_int) function pfchem_transport_setup(ctx, ncell, conn) &
        bind(C, name="pfchem_transport_setup")
      import :: c_int, c_int64_t, c_ptr, pfchem_connections
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ncell
      type(pfchem_connections), intent(in) :: conn
    end function pfchem_transport_setup

    integer(c_int) function pfchem_transport_step(ctx, ncell, state, coefs, dt, rtol, &
        max_iterations, iterations, stream) bind(C, name="pfchem_transport_step")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, pfchem_state, &
        pfchem_transport_coefs
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ncell
      type(pfchem_state), intent(in) :: state
      type(pfchem_transport_coefs), intent(in) :: coefs
      real(c_double), value :: dt, rtol
      integer(c_int), value :: max_iterations
      integer(c_int32_t), intent(out) :: iterations
      type(c_ptr), value :: stream
    end function pfchem_transport_step

    integer(c_int) function pfchem_set_launch(ctx, variant, cfg) bind(C, name="pfchem_set_launch")
      import :: c_int, c_ptr, pfchem_launch
      type(c_ptr), value :: ctx
      integer(c_int), value :: variant
      type(pfchem_launch), intent(in) :: cfg
    end function pfchem_set_launch

    integer(c_size_t) function pfchem_c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function pfchem_c_strlen
  end interface

contains

  ! Fortran string of a NUL-terminated C string (pfchem_strerror /
  ! pfchem_last_error)
  function pfchem_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    n = int(pfchem_c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    do i = 1, n
      s(i:i) = chars(i)
    end do
  end function pfchem_string

end module pfchem_c
