"""ctypes mirror of include/pfchem.h (the C ABI of the chemistry step).

Field order and types must match the header exactly; tests check sizes
against the compiled library.
"""
import ctypes as C

# return codes
PFCHEM_OK = 0
PFCHEM_ERR_INVALID = 1
PFCHEM_ERR_HIP = 2
PFCHEM_ERR_UNSUPPORTED = 3
PFCHEM_ERR_NOMEM = 4
PFCHEM_ERR_CELLS = 5

# per-cell status
PFCHEM_CELL_OK = 0
PFCHEM_CELL_MAX_CUTS_ITERATIONS = 1
PFCHEM_CELL_MAX_CUTS_SINGULAR = 2
PFCHEM_CELL_MAX_CUTS_ACTIVITY = 3
PFCHEM_CELL_SENS_UNSUPPORTED = 4
PFCHEM_CELL_TOO_MANY_SUBSTEPS = 5
PFCHEM_CELL_CONSTRAINT_SINGULAR = 6
PFCHEM_CELL_CONSTRAINT_NONPOSITIVE = 7
PFCHEM_CELL_CONSTRAINT_MAX_ITERATIONS = 8

# constraint types (transport_constraint_rt.F90:22-34)
CONSTRAINT_NULL = 0
CONSTRAINT_FREE = 1
CONSTRAINT_TOTAL = 2
CONSTRAINT_LOG = 3
CONSTRAINT_PH = 4
CONSTRAINT_PE = 5
CONSTRAINT_EH = 6
CONSTRAINT_MINERAL = 7
CONSTRAINT_GAS = 8
CONSTRAINT_CHARGE_BAL = 9
CONSTRAINT_TOTAL_SORB = 10
CONSTRAINT_SUPERCRIT_CO2 = 11
CONSTRAINT_TOTAL_AQ_PLUS_SORB = 12

# activity coefficients (reaction_aux.F90:27-31)
ACT_COEF_FREQUENCY_OFF = 0
ACT_COEF_FREQUENCY_TIMESTEP = 1
ACT_COEF_FREQUENCY_NEWTON_ITER = 2
ACT_COEF_ALGORITHM_LAG = 3
ACT_COEF_ALGORITHM_NEWTON = 4

# kernels
KERNEL_PLAIN_FP64 = 0
KERNEL_BATCHED_FP64 = 1
KERNEL_BATCHED_MIXED = 2
KERNEL_NAMES = {
    KERNEL_PLAIN_FP64: "plain_fp64",
    KERNEL_BATCHED_FP64: "batched_fp64",
    KERNEL_BATCHED_MIXED: "batched_mixed",
}

SOMDEC_MAX_DOWN = 4
SOMDEC_MAX_MONOD = 4
SOMDEC_MAX_INHIB = 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Somdec(C.Structure):
    _fields_ = [
        ("nrxn", C.c_int32),
        ("species_id_co2", C.c_int32),
        ("species_id_nh4", C.c_int32),
        ("species_id_no3", C.c_int32),
        ("species_id_n2o", C.c_int32),
        ("species_id_proton", C.c_int32),
        ("species_id_hr", C.c_int32),
        ("species_id_nmin", C.c_int32),
        ("species_id_nimm", C.c_int32),
        ("species_id_nimp", C.c_int32),
        ("species_id_ngasmin", C.c_int32),
        ("pad0", C.c_int32),
        ("x0eps", C.c_double),
        ("n2o_frac_mineralization", C.c_double),
        ("inhibition_nh4_no3", C.c_double),
        ("rate_constant", _dp),
        ("rate_decomposition", _dp),
        ("rate_ad_factor", _dp),
        ("upstream_c_id", _ip),
        ("upstream_n_id", _ip),
        ("upstream_hr_id", _ip),
        ("upstream_nmin_id", _ip),
        ("upstream_nimp_id", _ip),
        ("upstream_nimm_id", _ip),
        ("upstream_nc", _dp),
        ("upstream_is_varycn", _ip),
        ("n_downstream_pools", _ip),
        ("downstream_c_id", _ip),
        ("downstream_n_id", _ip),
        ("downstream_stoich", _dp),
        ("downstream_nc", _dp),
        ("downstream_is_varycn", _ip),
        ("mineral_c_stoich", _dp),
        ("mineral_n_stoich", _dp),
        ("temperature_response_function", _ip),
        ("Q10", _dp),
        ("Ea", _dp),
        ("moisture_response_function", _ip),
        ("Ox_response_function", _ip),
        ("Ox_half_saturation", _dp),
        ("n_monod", _ip),
        ("monod_specid", _ip),
        ("monod_half_saturation", _dp),
        ("monod_threshold", _dp),
        ("monod_pool_normalized", _ip),
        ("n_inhibition", _ip),
        ("inhibition_specid", _ip),
        ("inhibition_type", _ip),
        ("inhibition_constant", _dp),
        ("inhibition_constant2", _dp),
        ("Ox_specid", _ip),
        ("species_o2", C.c_int32),
        ("pad1", C.c_int32),
    ]


class Network(C.Structure):
    _fields_ = [
        ("naqcomp", C.c_int32),
        ("nimmobile", C.c_int32),
        ("h2o_aq_id", C.c_int32),
        ("pad0", C.c_int32),
        ("primary_spec_Z", _dp),
        ("primary_spec_a0", _dp),
        ("neqcplx", C.c_int32),
        ("eqcplx_maxsp", C.c_int32),
        ("eqcplxspecid", _ip),
        ("eqcplxstoich", _dp),
        ("eqcplxh2ostoich", _dp),
        ("eqcplx_logK", _dp),
        ("eqcplx_logKcoef", _dp),
        ("eqcplx_Z", _dp),
        ("eqcplx_a0", _dp),
        ("nkinmnrl", C.c_int32),
        ("kinmnrl_maxsp", C.c_int32),
        ("kinmnrlspecid", _ip),
        ("kinmnrlstoich", _dp),
        ("kinmnrlh2ostoich", _dp),
        ("kinmnrl_logK", _dp),
        ("kinmnrl_logKcoef", _dp),
        ("kinmnrl_rate_constant", _dp),
        ("kinmnrl_activation_energy", _dp),
        ("kinmnrl_molar_vol", _dp),
        ("kinmnrl_Temkin_const", _dp),
        ("kinmnrl_min_scale_factor", _dp),
        ("kinmnrl_affinity_power", _dp),
        ("kinmnrl_affinity_threshold", _dp),
        ("kinmnrl_rate_limiter", _dp),
        ("kinmnrl_irreversible", _ip),
        ("ndecay_rxn", C.c_int32),
        ("pad1", C.c_int32),
        ("decayspecid", _ip),
        ("decay_rate_constant", _dp),
        ("debyeA", C.c_double),
        ("debyeB", C.c_double),
        ("debyeBdot", C.c_double),
        ("somdec", C.POINTER(Somdec)),
        ("kinmnrl_max_pref", C.c_int32),
        ("kinmnrl_max_pref_spec", C.c_int32),
        ("kinmnrl_num_prefactors", _ip),
        ("kinmnrl_prefactor_id", _ip),
        ("kinmnrl_pref_alpha", _dp),
        ("kinmnrl_pref_beta", _dp),
        ("kinmnrl_pref_atten_coef", _dp),
        ("kinmnrl_pref_rate", _dp),
        ("kinmnrl_pref_activation_energy", _dp),
    ]


class Options(C.Structure):
    _fields_ = [
        ("use_full_geochemistry", C.c_int32),
        ("use_log_formulation", C.c_int32),
        ("act_coef_update_frequency", C.c_int32),
        ("act_coef_update_algorithm", C.c_int32),
        ("use_activity_h2o", C.c_int32),
        ("use_total_as_guess", C.c_int32),
        ("isothermal", C.c_int32),
        ("maximum_reaction_iterations", C.c_int32),
        ("maximum_reaction_cuts", C.c_int32),
        ("variable_ncomp", C.c_int32),
        ("maximum_rreact_calls", C.c_int32),
        ("pad_options", C.c_int32),
        ("max_dlnC_rreact", C.c_double),
        ("max_residual_tolerance", C.c_double),
        ("max_rel_residual_tolerance", C.c_double),
        ("max_relative_change_tolerance", C.c_double),
    ]


class State(C.Structure):
    _fields_ = [
        ("ld", C.c_int64),
        ("pri_molal", C.c_void_p),
        ("total", C.c_void_p),
        ("immobile", C.c_void_p),
        ("sec_molal", C.c_void_p),
        ("pri_act_coef", C.c_void_p),
        ("sec_act_coef", C.c_void_p),
        ("ln_act_h2o", C.c_void_p),
        ("mnrl_volfrac", C.c_void_p),
        ("mnrl_area", C.c_void_p),
        ("mnrl_rate", C.c_void_p),
        ("porosity", C.c_void_p),
        ("saturation", C.c_void_p),
        ("den_kg", C.c_void_p),
        ("volume", C.c_void_p),
        ("temperature", C.c_void_p),
        ("somdec_f_t", C.c_void_p),
        ("somdec_f_w", C.c_void_p),
        ("num_iterations", C.c_void_p),
        ("num_sub_steps", C.c_void_p),
        ("status", C.c_void_p),
        ("ncomp_active", C.c_void_p),
    ]


class Constraint(C.Structure):
    _fields_ = [
        ("constraint_type", _ip),
        ("constraint_rxn", _ip),
        ("nrxn", C.c_int32),
        ("rxn_maxsp", C.c_int32),
        ("rxnspecid", _ip),
        ("rxnstoich", _dp),
        ("rxnh2ostoich", _dp),
        ("rxn_logK", _dp),
        ("rxn_logKcoef", _dp),
        ("h_ion_id", C.c_int32),
        ("initialize_with_molality", C.c_int32),
        ("use_prev_soln_as_guess", C.c_int32),
        ("max_iterations", C.c_int32),
        ("max_dlnC", C.c_double),
        ("free_ion_guess", _dp),
    ]


class Connections(C.Structure):
    _fields_ = [
        ("nconn", C.c_int64),
        ("id_up", C.c_void_p),
        ("id_dn", C.c_void_p),
        ("nbc", C.c_int64),
        ("bc_cell", C.c_void_p),
    ]


class TransportCoefs(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in
                ("area", "velocity", "coef_over_dist", "bc_area", "bc_velocity",
                 "bc_coef_over_dist", "bc_total")]


class Launch(C.Structure):
    _fields_ = [
        ("lanes_per_cell", C.c_int32),
        ("block_size", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("rows_per_lane", C.c_int32),
    ]


# (lanes per cell, rows per lane) team shapes compiled into libpfchem.so
TEAM_SHAPES = [(4, 1), (2, 4), (2, 8), (4, 4), (8, 2), (16, 1), (8, 4), (32, 1), (64, 1)]
