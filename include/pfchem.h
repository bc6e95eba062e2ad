/*
 * pfchem.h -- C ABI of the MI355X per-cell chemistry step (operator-split
 * reaction step of PFLOTRAN: RStep -> RReact -> RTotal / RActivityCoefficients
 * / RKineticMineral / RImmobileDecay / SomDec sandbox / RSolve).
 *
 * Plain C, plain pointers and sizes.  Every entry point returns an int status
 * (PFCHEM_OK == 0); pfchem_strerror() / pfchem_last_error() explain failures.
 * No entry point aborts the process: the reference's `stop` /
 * MPI_Abort paths become per-cell status codes the host acts on.
 *
 * Reference interfaces replaced (paths relative to the PFLOTRAN tree):
 *   pfchem_step        <- RStep(guess, rt_auxvar, global_auxvar,
 *                          material_auxvar, num_sub_steps, num_iterations,
 *                          num_kinetic_state_updates, reaction, natural_id,
 *                          option, ierror)   src/pflotran/reaction.F90:3564-3566,
 *                        called once per cell from the OSRT loop
 *                          src/pflotran/pmc_subsurface_osrt.F90:349-378;
 *                        batched: one call per local cell range.
 *   pfchem_update_activity <- RActivityCoefficients (reaction.F90:4368) as
 *                        called per cell by RTUpdateActivityCoefficients
 *                        (reactive_transport.F90:3441-3492) at the start of
 *                        an OSRT step (pm_osrt.F90:148-151).
 *   pfchem_residual / pfchem_residual_jacobian <- the per-cell reaction
 *                        block of the global implicit residual and Jacobian:
 *                        RTAccumulation + RReaction in RTResidualNonFlux
 *                        (reactive_transport.F90:2407, 2627) and
 *                        RTAccumulationDerivative + RReactionDerivative in
 *                        RTJacobianNonFlux (:3177, 3288), in the RReact form
 *                        of one Newton iterate (reaction.F90:3896-3916), FP64.
 *   pfchem_tangent / pfchem_adjoint: no reference counterpart (the
 *                        reference has no sensitivities); they differentiate
 *                        the converged RReact system (reaction.F90:3853-4049).
 *
 * Layout: every per-cell field is structure-of-arrays, FP64, field-major:
 * element (k, cell) of an array with per-cell length m lives at
 * ptr[k * ld + cell], 0 <= k < m, 0 <= cell < ncell <= ld.
 */
#ifndef PFCHEM_H
#define PFCHEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* return codes                                                        */
/* ------------------------------------------------------------------ */
enum {
  PFCHEM_OK = 0,
  PFCHEM_ERR_INVALID = 1,     /* bad argument / inconsistent network    */
  PFCHEM_ERR_HIP = 2,         /* a HIP runtime call failed              */
  PFCHEM_ERR_UNSUPPORTED = 3, /* feature/size not supported by a kernel */
  PFCHEM_ERR_NOMEM = 4,       /* device allocation failed               */
  PFCHEM_ERR_CELLS = 5        /* step ran; >= 1 cell reports failure    */
};

/* per-cell status written to pfchem_state.status (RStep ierror, refined) */
enum {
  PFCHEM_CELL_OK = 0,
  PFCHEM_CELL_MAX_CUTS_ITERATIONS = 1, /* cuts exhausted, last RReact hit
                                          maximum_reaction_iterations
                                          (reaction.F90:3869-3895, 3662)   */
  PFCHEM_CELL_MAX_CUTS_SINGULAR = 2,   /* cuts exhausted, last RReact had a
                                          zero Jacobian row (utility.F90:629)*/
  PFCHEM_CELL_MAX_CUTS_ACTIVITY = 3,   /* cuts exhausted, last RReact had an
                                          activity-coefficient failure
                                          (reaction.F90:4421-4431,4477-4486) */
  PFCHEM_CELL_SENS_UNSUPPORTED = 4,    /* pfchem_tangent/adjoint: the cell is
                                          dry (no sensitivity computed)      */
  PFCHEM_CELL_TOO_MANY_SUBSTEPS = 5,   /* pfchem_tangent/adjoint: more than 8
                                          accepted sub-steps to record       */
  PFCHEM_CELL_CONSTRAINT_SINGULAR = 6, /* pfchem_equilibrate: zero Jacobian
                                          row (reaction.F90:1916-1926)       */
  PFCHEM_CELL_CONSTRAINT_NONPOSITIVE = 7, /* pfchem_equilibrate: a free-ion
                                          molality <= 0 (reaction.F90:1954-1975) */
  PFCHEM_CELL_CONSTRAINT_MAX_ITERATIONS = 8 /* pfchem_equilibrate: no
                                          convergence within max_iterations
                                          (reaction.F90:2006-2016)           */
};

/* activity coefficient update frequency / algorithm
   (same values as reaction_aux.F90:27-31) */
enum {
  PFCHEM_ACT_COEF_FREQUENCY_OFF = 0,
  PFCHEM_ACT_COEF_FREQUENCY_TIMESTEP = 1,
  PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER = 2,
  PFCHEM_ACT_COEF_ALGORITHM_LAG = 3,
  PFCHEM_ACT_COEF_ALGORITHM_NEWTON = 4
};

/* kernel variants for pfchem_step */
enum {
  PFCHEM_KERNEL_PLAIN_FP64 = 0,   /* one cell per thread, FP64 throughout   */
  PFCHEM_KERNEL_BATCHED_FP64 = 1, /* persistent, work stealing, FP64 J/LU   */
  PFCHEM_KERNEL_BATCHED_MIXED = 2 /* persistent, work stealing, FP32 J/LU,
                                     FP64 residual / log K sums / update /
                                     convergence test                       */
};

/* ------------------------------------------------------------------ */
/* reaction network (host memory; copied into the context at create)   */
/* mirrors reaction_rt_type fields (reaction_aux.F90:123-311) and       */
/* mineral_type kinetic fields (reaction_mineral_aux.F90:82-140).       */
/* All species ids are 0-based.                                        */
/* ------------------------------------------------------------------ */
typedef struct pfchem_somdec pfchem_somdec; /* SomDec sandbox, see below */

typedef struct pfchem_network {
  int32_t naqcomp;    /* reaction%naqcomp; aqueous dofs are 0..naqcomp-1   */
  int32_t nimmobile;  /* reaction%immobile%nimmobile; dof naqcomp+i         */
  int32_t h2o_aq_id;  /* species_idx%h2o_aq_id (0-based) or -1              */
  int32_t pad0;
  const double* primary_spec_Z;  /* [naqcomp]                               */
  const double* primary_spec_a0; /* [naqcomp] (Angstrom)                    */

  /* aqueous equilibrium complexes (reaction%eqcplx*) */
  int32_t neqcplx;
  int32_t eqcplx_maxsp;          /* row length of the two arrays below      */
  const int32_t* eqcplxspecid;   /* [neqcplx][eqcplx_maxsp+1]; [k][0]=count  */
  const double* eqcplxstoich;    /* [neqcplx][eqcplx_maxsp]                 */
  const double* eqcplxh2ostoich; /* [neqcplx] or NULL (0 => no water)       */
  const double* eqcplx_logK;     /* [neqcplx] at the reference temperature  */
  const double* eqcplx_logKcoef; /* [neqcplx][5] or NULL (ReactionInterpolateLogK) */
  const double* eqcplx_Z;        /* [neqcplx]                               */
  const double* eqcplx_a0;       /* [neqcplx]                               */

  /* kinetic minerals (reaction%mineral%kinmnrl*) */
  int32_t nkinmnrl;
  int32_t kinmnrl_maxsp;
  const int32_t* kinmnrlspecid;      /* [nkinmnrl][kinmnrl_maxsp+1]         */
  const double* kinmnrlstoich;       /* [nkinmnrl][kinmnrl_maxsp]           */
  const double* kinmnrlh2ostoich;    /* [nkinmnrl] or NULL                  */
  const double* kinmnrl_logK;        /* [nkinmnrl]                          */
  const double* kinmnrl_logKcoef;    /* [nkinmnrl][5] or NULL               */
  const double* kinmnrl_rate_constant;      /* [nkinmnrl] mol/m^2/s         */
  const double* kinmnrl_activation_energy;  /* [nkinmnrl] J/mol, or NULL    */
  const double* kinmnrl_molar_vol;          /* [nkinmnrl] m^3/mol           */
  const double* kinmnrl_Temkin_const;       /* [nkinmnrl] or NULL (not associated) */
  const double* kinmnrl_min_scale_factor;   /* [nkinmnrl] or NULL           */
  const double* kinmnrl_affinity_power;     /* [nkinmnrl] or NULL           */
  const double* kinmnrl_affinity_threshold; /* [nkinmnrl] or NULL (=0)      */
  const double* kinmnrl_rate_limiter;       /* [nkinmnrl] or NULL (=0)      */
  const int32_t* kinmnrl_irreversible;      /* [nkinmnrl] or NULL (=0)      */

  /* first-order decay of immobile species (reaction%immobile%decay*) */
  int32_t ndecay_rxn;
  int32_t pad1;
  const int32_t* decayspecid;          /* [ndecay_rxn] immobile index       */
  const double* decay_rate_constant;   /* [ndecay_rxn] 1/s                  */

  /* Debye-Hueckel constants (reaction_aux.F90:486-488 defaults) */
  double debyeA, debyeB, debyeBdot;

  /* optional SomDec decomposition sandbox (reaction_sandbox_somdec.F90) */
  const pfchem_somdec* somdec;

  /* kinetic-mineral rate prefactors (PREFACTOR blocks, reaction_mineral.F90:
     820-870; mineral%kinmnrl_num_prefactors ...).  A mineral m with
     kinmnrl_num_prefactors[m] > 0 replaces its rate constant and activation
     energy by the sum over its prefactors p of
       pref_rate[m][p] * Arrhenius(pref_activation_energy[m][p]) *
       prod_s a_s^alpha[m][p][s] / (1 + atten_coef[m][p][s] a_s^beta[m][p][s])
     over the prefactor's species s (0-based primary ids; the reference's
     secondary-species prefactors are not supported).  NULL
     kinmnrl_num_prefactors: no mineral has prefactors.  The one-lane
     kernel and the sensitivity entry points do not take such networks. */
  int32_t kinmnrl_max_pref;        /* prefactors per mineral (array stride)   */
  int32_t kinmnrl_max_pref_spec;   /* species per prefactor (array stride)    */
  const int32_t* kinmnrl_num_prefactors;  /* [nkinmnrl]                       */
  const int32_t* kinmnrl_prefactor_id;    /* [nkinmnrl][max_pref][max_pref_spec+1];
                                             [m][p][0] = species count        */
  const double* kinmnrl_pref_alpha;       /* [nkinmnrl][max_pref][max_pref_spec] */
  const double* kinmnrl_pref_beta;        /* same                             */
  const double* kinmnrl_pref_atten_coef;  /* same                             */
  const double* kinmnrl_pref_rate;        /* [nkinmnrl][max_pref] mol/m^2/s   */
  const double* kinmnrl_pref_activation_energy; /* [nkinmnrl][max_pref] J/mol, or NULL */
} pfchem_network;

/* solver options: the reaction_rt_type control fields */
typedef struct pfchem_options {
  int32_t use_full_geochemistry;         /* reaction%use_full_geochemistry  */
  int32_t use_log_formulation;           /* LOG_FORMULATION                 */
  int32_t act_coef_update_frequency;     /* PFCHEM_ACT_COEF_FREQUENCY_*     */
  int32_t act_coef_update_algorithm;     /* PFCHEM_ACT_COEF_ALGORITHM_*     */
  int32_t use_activity_h2o;              /* ACTIVITY_H2O                    */
  int32_t use_total_as_guess;            /* USE_TOTAL_CONCENTRATION_AS_GUESS*/
  int32_t isothermal;                    /* option%use_isothermal           */
  int32_t maximum_reaction_iterations;   /* default 20                      */
  int32_t maximum_reaction_cuts;         /* default 10                      */
  int32_t variable_ncomp;                /* 1: pfchem_state.ncomp_active gives
                                            each cell's active primaries (see
                                            there); 0: every cell solves the
                                            whole network                   */
  int32_t maximum_rreact_calls;          /* RReact calls one RStep may make
                                            before the cell fails with
                                            status 1 (a bound on the time a
                                            stuck cell holds its lanes; the
                                            reference's loop has none);
                                            0 = 10000                       */
  int32_t pad_options;
  double max_dlnC_rreact;                /* default 5                       */
  double max_residual_tolerance;         /* default 1e-12                   */
  double max_rel_residual_tolerance;     /* default 1e-8                    */
  double max_relative_change_tolerance;  /* default 1e-6                    */
} pfchem_options;

/* ------------------------------------------------------------------ */
/* SomDec sandbox descriptor (Setup output of SomDecSetup,             */
/* reaction_sandbox_somdec.F90:987-1501), one entry per reaction.       */
/* Species ids are 0-based: aqueous ids index pri/total, immobile ids   */
/* index immobile.  -1 means "not present".                             */
/* ------------------------------------------------------------------ */
#define PFCHEM_SOMDEC_MAX_DOWN 4
#define PFCHEM_SOMDEC_MAX_MONOD 4
#define PFCHEM_SOMDEC_MAX_INHIB 4

struct pfchem_somdec {
  int32_t nrxn;
  int32_t species_id_co2;        /* aqueous id (ITYPE_AQUEOUS)              */
  int32_t species_id_nh4;        /* aqueous id                              */
  int32_t species_id_no3;        /* aqueous id or -1                        */
  int32_t species_id_n2o;        /* aqueous id or -1                        */
  int32_t species_id_proton;     /* aqueous id or -1 (for N2O pH factor)    */
  int32_t species_id_hr;         /* immobile id or -1 (tracking)            */
  int32_t species_id_nmin;       /* immobile id or -1                       */
  int32_t species_id_nimm;       /* immobile id or -1                       */
  int32_t species_id_nimp;       /* immobile id or -1                       */
  int32_t species_id_ngasmin;    /* immobile id or -1                       */
  int32_t pad0;
  double x0eps;                  /* this%x0eps                              */
  double n2o_frac_mineralization;/* this%n2o_frac_mineralization            */
  double inhibition_nh4_no3;     /* this%inhibition_nh4_no3                 */
  /* per reaction arrays [nrxn] */
  const double* rate_constant;       /* <0 => use rate_decomposition       */
  const double* rate_decomposition;
  const double* rate_ad_factor;
  const int32_t* upstream_c_id;      /* immobile id                        */
  const int32_t* upstream_n_id;      /* immobile id or -1                  */
  const int32_t* upstream_hr_id;
  const int32_t* upstream_nmin_id;
  const int32_t* upstream_nimp_id;
  const int32_t* upstream_nimm_id;
  const double* upstream_nc;         /* setup-time N:C (mol/mol)           */
  const int32_t* upstream_is_varycn;
  const int32_t* n_downstream_pools;
  const int32_t* downstream_c_id;    /* [nrxn][MAX_DOWN] immobile id       */
  const int32_t* downstream_n_id;    /* [nrxn][MAX_DOWN]                   */
  const double* downstream_stoich;   /* [nrxn][MAX_DOWN]                   */
  const double* downstream_nc;       /* [nrxn][MAX_DOWN] setup-time N:C    */
  const int32_t* downstream_is_varycn; /* [nrxn][MAX_DOWN]                 */
  const double* mineral_c_stoich;    /* setup-time                         */
  const double* mineral_n_stoich;    /* setup-time                         */
  /* abiotic factors per reaction */
  const int32_t* temperature_response_function; /* elm_rspfuncs ids, 0=off */
  const double* Q10;
  const double* Ea;
  const int32_t* moisture_response_function;
  const int32_t* Ox_response_function;
  const double* Ox_half_saturation;
  /* Monod / inhibition terms per reaction ([nrxn][MAX_*]) */
  const int32_t* n_monod;
  const int32_t* monod_specid;       /* aqueous id (>=0) or -(immobile id)-1 */
  const double* monod_half_saturation;
  const double* monod_threshold;
  const int32_t* monod_pool_normalized;
  const int32_t* n_inhibition;
  const int32_t* inhibition_specid;
  const int32_t* inhibition_type;
  const double* inhibition_constant;
  const double* inhibition_constant2;
  /* oxygen, as 1-based species codes (0 = none): aqueous id + 1 (> 0) or
     -(immobile id + 1) (< 0).  Ox_specid[r]: reaction r's oxidizer of the
     Monod oxygen response (Ox_response_function 1, rxn%Ox_species,
     reaction_sandbox_somdec.F90:1335-1339); NULL = none for every
     reaction.  species_o2: the species whose consumption (equal to the CO2
     release) is tracked (O2_SPECIES_NAME, :1400-1405, 2176-2186). */
  const int32_t* Ox_specid;
  int32_t species_o2;
  int32_t pad1;
};

/* ------------------------------------------------------------------ */
/* per-cell state (SoA, see layout note at the top)                    */
/* rt_auxvar / global_auxvar / material_auxvar fields used by RStep    */
/* (reactive_transport_aux.F90:21-74, global_aux.F90, material_aux).   */
/* ------------------------------------------------------------------ */
typedef struct pfchem_state {
  int64_t ld;                 /* stride between fields, >= ncell          */
  double* pri_molal;          /* [naqcomp]  in: guess (free ion molality,
                                 mol/kg); out: converged free molality     */
  double* total;              /* [naqcomp]  in: total after transport
                                 (mol/L); out: total at end of step        */
  double* immobile;           /* [nimmobile] in/out (mol/m^3 bulk)         */
  double* sec_molal;          /* [neqcplx]  in (LAG ionic strength) / out  */
  double* pri_act_coef;       /* [naqcomp]  in/out                         */
  double* sec_act_coef;       /* [neqcplx]  in/out                         */
  double* ln_act_h2o;         /* [1] in/out, may be NULL if unused         */
  double* mnrl_volfrac;       /* [nkinmnrl] in/out (m^3/m^3 bulk)          */
  const double* mnrl_area;    /* [nkinmnrl] in (m^2/m^3 bulk)              */
  double* mnrl_rate;          /* [nkinmnrl] out (mol/m^3 bulk/s)           */
  const double* porosity;     /* [1]                                       */
  const double* saturation;   /* [1] liquid saturation                     */
  const double* den_kg;       /* [1] liquid density kg/m^3                 */
  const double* volume;       /* [1] cell volume m^3                       */
  const double* temperature;  /* [1] C (non-isothermal logK, Arrhenius)    */
  const double* somdec_f_t;   /* [1] SomDec temperature scalar (ELM t_scalar)
                                 or NULL => response functions            */
  const double* somdec_f_w;   /* [1] SomDec moisture scalar (ELM w_scalar) */
  int32_t* num_iterations;    /* [1] out: Newton iterations (RStep)        */
  int32_t* num_sub_steps;     /* [1] out                                   */
  int32_t* status;            /* [1] out: PFCHEM_CELL_*                    */
  const int32_t* ncomp_active;/* [1] with options.variable_ncomp: the cell
                                 solves the network restricted to primaries
                                 0..ncomp_active-1 (clamped to [1, naqcomp]):
                                 the complexes and kinetic minerals all of
                                 whose species are among them; the other
                                 entries of the cell are left untouched.
                                 Networks with immobile species or the
                                 SomDec sandbox are not supported.  NULL
                                 otherwise.                                 */
} pfchem_state;

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
typedef struct pfchem_ctx pfchem_ctx;

/* Validate and copy the network to device `device`.  Replaces the
   per-run setup of reaction_rt_type (ReactionCreate, reaction_aux.F90:371;
   BasisInit, reaction_database.F90:812). */
int pfchem_create(const pfchem_network* net, const pfchem_options* opt,
                  int device, pfchem_ctx** out);
int pfchem_destroy(pfchem_ctx* ctx);
const char* pfchem_strerror(int code);
/* message of the last failure on this context (never NULL) */
const char* pfchem_last_error(const pfchem_ctx* ctx);

/* One operator-split reaction step of length dt (option%tran_dt) for
   ncell cells whose state pointers are DEVICE pointers.  Launches on
   `stream` (a hipStream_t, NULL = default stream) and returns without
   synchronizing.  Per-cell outcome goes to state->status; the call itself
   returns PFCHEM_OK if the launch succeeded.  Calls on one context may use
   different streams; steps that use the context's scratch lists (the
   one-lane kernel's hand-over list, the status array kept when
   state->status is NULL, the size buckets) run one after another on the
   device, whatever their streams.
   Replaces RStep (reaction.F90:3564) applied to every local cell. */
int pfchem_step(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                double dt, int variant, void* stream);

/* Same as pfchem_step for HOST pointers: stages the state through device
   buffers owned by the context, synchronizes, and returns PFCHEM_ERR_CELLS
   if any cell failed (the reference's rstep_error all-reduce MAX,
   pmc_subsurface_osrt.F90:381-382).  For a Fortran RStepBatch. */
int pfchem_step_host(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                     double dt, int variant);

/* RActivityCoefficients (reaction.F90:4368-4614) for every cell, device
   pointers; uses the algorithm in the options. */
int pfchem_update_activity(pfchem_ctx* ctx, int64_t ncell,
                           const pfchem_state* state, void* stream);

/* FP64 residual of the RReact system at the state's current free
   molalities and activity coefficients (reaction.F90:3896-3916):
   res[i*ld+cell] = (Acc(x) - Acc(total^k,imm^k))/dt + reaction terms, with
   the fixed accumulation taken from the totals `fixed_total` [naqcomp][ld]
   and `fixed_immobile` [nimmobile][ld] (SoA; NULL when the count is 0).
   Rows i < naqcomp are the aqueous primaries, then the immobile species.
   As RTAuxVarCompute does, the call rewrites state->total and
   state->sec_molal from the free molalities, and state->mnrl_rate.  Dry
   cells (saturation < 1e-40) give res = 0.  The SomDec sandbox's terms are
   included (per-cell stoichiometry from the Setup values).  Device
   pointers; launches on `stream` without synchronizing. */
int pfchem_residual(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                    const double* fixed_total, const double* fixed_immobile,
                    double dt, double* res, void* stream);

/* pfchem_residual plus the Jacobian block of the same iterate:
   jac[(i*ncomp + j)*ld + cell] = d res_i / d x_j (ncomp = naqcomp +
   nimmobile; aqueous columns per mol/kg water of free molality, immobile
   columns per mol/m^3 bulk; identity for a dry cell), the analytic
   RTAccumulationDerivative + RReaction derivative that RTJacobianNonFlux
   adds into the cell's diagonal block (reactive_transport.F90:3177, 3288).
   jac may be NULL.  Replaces the per-cell loop bodies of RTResidualNonFlux
   (reactive_transport.F90:2393-2430, 2599-2642) and RTJacobianNonFlux
   (:3172-3183, 3270-3300) for a batch of cells. */
int pfchem_residual_jacobian(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                             const double* fixed_total, const double* fixed_immobile,
                             double dt, double* res, double* jac, void* stream);

/* Launch configuration of the batched kernels.  lanes_per_cell = 1: the
   network-specialized kernel, one cell per lane (64 cells per wavefront),
   generated from the network and options at pfchem_create and compiled for
   the device at the first batched step; it takes aqueous networks of up to
   16 primaries (plus immobile species with first-order decay) without the
   SomDec sandbox or per-cell active counts, and is then the default.
   block_size must be 256 and blocks_per_cu caps the workgroups per CU (0:
   as many as its register use allows).  lanes_per_cell > 1: a team of
   lanes_per_cell lanes owns one cell (64 / lanes_per_cell cells per
   wavefront) and each lane owns rows_per_lane rows of the cell's system
   (lanes_per_cell * rows_per_lane >= ncomp).  Team shapes compiled in:
   4x1 2x4 2x8 4x4 8x2 16x1 8x4 32x1 64x1. */
typedef struct pfchem_launch {
  int32_t lanes_per_cell;   /* 1 (one-lane kernel), 2, 4, 8, 16, 32, 64  */
  int32_t block_size;       /* threads per workgroup (64, 128 or 256)     */
  int32_t blocks_per_cu;    /* persistent grid = blocks_per_cu * CUs (0:
                               the kernel's default)                     */
  int32_t rows_per_lane;    /* 1, 2, 4 or 8                               */
} pfchem_launch;
int pfchem_set_launch(pfchem_ctx* ctx, int variant, const pfchem_launch* cfg);
int pfchem_get_launch(const pfchem_ctx* ctx, int variant, pfchem_launch* cfg);

/* Launch-shape autotuner.  No reference counterpart: the reference runs one
   cell at a time on the CPU.  Times pfchem_step of `variant` (a batched
   variant) on the caller's cells (device pointers) for the launch shapes
   the context accepts: the one-lane kernel with 0, 1 and 2 workgroups per
   CU as cap; every compiled team shape with lanes x rows in
   [ncomp, max(4 ncomp, 16)] at 256-lane workgroups, 4 per CU; then the
   three fastest teams (those within 2x of the best shape so far) over
   workgroup sizes 64/128/256 and 1/2/4/8 workgroups per CU.  With per-cell active counts, each size bucket's team
   (16x1, 32x1, 64x1) is swept over workgroup size and grid in turn.  Every
   timing is the best of 3 steps (1 with PFCHEM_TUNE_QUICK) after an untimed
   one, each from the caller's state as given; the state is left as given.
   Everything runs on `stream` (NULL: the default stream); the call
   synchronizes it.
   The fastest shape whose cells all end with the starting shape's status
   and free molalities within 1e-6 relative becomes the context's shape for
   `variant` and is returned in `chosen`.
   cache_path (optional): JSON cache of choices keyed by device, network
   content and cell-count bucket (power of two).  An entry for this exact
   key and bucket is applied without measuring unless flags has
   PFCHEM_TUNE_FORCE; a measured choice replaces it.  pfchem_create applies
   the cache named by the environment variable PFCHEM_TUNE_CACHE (exact
   network first, else the same system-size bucket).  table_path
   (optional): text table of every candidate's time, difference and
   outcome.  Allocates device scratch for a copy of the state's output
   fields. */
/* Measurement hook (no counterpart in the reference; for benchmarks and
   profilers that time the dominant kernel on its own stream).  With timing
   on, every batched pfchem_step that runs the one-lane kernel records a
   pair of HIP events on its stream around that kernel's pass over every
   cell (the per-launch cursor clear included).  pfchem_pass_time waits
   for the last recorded pair and returns the pass's duration in ms;
   PFCHEM_ERR_INVALID when no step recorded one since timing was turned on. */
int pfchem_set_pass_timing(pfchem_ctx* ctx, int on);
int pfchem_pass_time(pfchem_ctx* ctx, float* ms);

#define PFCHEM_TUNE_FORCE 1
#define PFCHEM_TUNE_QUICK 2
int pfchem_autotune(pfchem_ctx* ctx, int variant, int64_t ncell, const pfchem_state* state,
                    double dt, const char* cache_path, const char* table_path, int flags,
                    pfchem_launch* chosen, void* stream);
/* The cache keys of a context: `key` (device, network content) and
   `bucket` (device, system size rounded up to a power of two, kernel
   family). */
int pfchem_tune_key(const pfchem_ctx* ctx, char* key, int key_len, char* bucket,
                    int bucket_len);
/* Look up a cache file without a device: returns 1 and the entry's launch
   shape for (key, else bucket; variant; nearest cell-count bucket), 0 if
   none, or a negative error. */
int pfchem_tune_cache_query(const char* cache_path, const char* key, const char* bucket,
                            int variant, int64_t ncell, pfchem_launch* out);

/* Generate and compile the network-specialized one-lane kernels of
   (net, opt) for gfx950 without a device, e.g. to check a network at set-up
   time (pfchem_step compiles them on the device at its first batched call).
   PFCHEM_ERR_UNSUPPORTED: the one-lane kernel does not take this network
   (the team kernels run it).  source_path (optional) receives the generated
   device source.  No reference counterpart: the reference interprets the
   network at run time (reaction_aux.F90:123-311). */
int pfchem_prepare_kernels(const pfchem_network* net, const pfchem_options* opt,
                           const char* source_path);

/* ------------------------------------------------------------------ */
/* sensitivities of one step                                           */
/* ------------------------------------------------------------------ */
/* Derivatives of the per-cell objective
     f = sum_n w_vf[n] * mnrl_volfrac_n(end) + sum_k w_sec[k] * sec_molal_k(end)
   (end of one RStep of length dt from the given state, which is read only)
   with respect to p = (log10 kinmnrl_rate_constant[0..nkinmnrl),
   log10 K of the complexes eqcplx_logK[0..neqcplx)), through the converged
   Newton system of every accepted sub-step (implicit-function theorem).
   The reference has no counterpart; the derivatives are those of RStep /
   RReact (reaction.F90:3564-4055) and MineralUpdateKineticState
   (reaction_mineral.F90:1456-1525).  jacobian: 1 = exact dF/dm,
   0 = as the reference assembles it (mineral block times rho/1000,
   reaction_mineral.F90:945).  Supported: no immobile species, plain TST
   kinetics (optional activation energy), activity coefficients fixed
   during the step (not PFCHEM_ACT_COEF_FREQUENCY_NEWTON_ITER), up to 32
   primaries and 16 minerals; otherwise PFCHEM_ERR_UNSUPPORTED.  state,
   f_cell, status: device pointers; weights and dp: host arrays. */

/* tangent along dp[nkinmnrl + neqcplx]: df_cell[cell] = df/dp . dp.
   f_cell (optional) receives f.  Asynchronous on `stream`. */
int pfchem_tangent(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state, double dt,
                   const double* w_vf, const double* w_sec, const double* dp, int jacobian,
                   double* df_cell, double* f_cell, int32_t* status, void* stream);

/* adjoint: grad[nkinmnrl + neqcplx] (host) = sum over cells with status 0 of
   df/dp; grad_cell (device, optional, [q * ncell + cell]) the per-cell
   gradients.  Synchronizes `stream` before returning. */
int pfchem_adjoint(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state, double dt,
                   const double* w_vf, const double* w_sec, int jacobian, double* grad,
                   double* grad_cell, double* f_cell, int32_t* status, void* stream);

/* ------------------------------------------------------------------ */
/* constraint speciation                                               */
/* ------------------------------------------------------------------ */
/* constraint types, same values as transport_constraint_rt.F90:22-34.
   Supported here: NULL, FREE, TOTAL, LOG, PH, MINERAL, GAS, CHARGE_BAL.
   PE/EH (redox), TOTAL_SORB / TOTAL_AQ_PLUS_SORB (sorption) and
   SUPERCRIT_CO2 (CO2 equation of state) give PFCHEM_ERR_UNSUPPORTED. */
enum {
  PFCHEM_CONSTRAINT_NULL = 0,
  PFCHEM_CONSTRAINT_FREE = 1,
  PFCHEM_CONSTRAINT_TOTAL = 2,
  PFCHEM_CONSTRAINT_LOG = 3,
  PFCHEM_CONSTRAINT_PH = 4,
  PFCHEM_CONSTRAINT_PE = 5,
  PFCHEM_CONSTRAINT_EH = 6,
  PFCHEM_CONSTRAINT_MINERAL = 7,
  PFCHEM_CONSTRAINT_GAS = 8,
  PFCHEM_CONSTRAINT_CHARGE_BAL = 9,
  PFCHEM_CONSTRAINT_TOTAL_SORB = 10,
  PFCHEM_CONSTRAINT_SUPERCRIT_CO2 = 11,
  PFCHEM_CONSTRAINT_TOTAL_AQ_PLUS_SORB = 12
};

/* One aqueous constraint (tran_constraint_rt_type, aq_species_constraint
   part) applied to many cells.  MINERAL and GAS constraints name a row of
   the reaction table below: the mineral's (mineral_reaction%mnrl*) or
   passive gas's (reaction%gas%paseq*) reaction written in the network's
   basis, same compressed layout as the complexes. */
typedef struct pfchem_constraint {
  const int32_t* constraint_type; /* [naqcomp] PFCHEM_CONSTRAINT_*          */
  const int32_t* constraint_rxn;  /* [naqcomp] row for MINERAL / GAS, else
                                     ignored (constraint_spec_id)           */
  int32_t nrxn;                   /* rows of the reaction table            */
  int32_t rxn_maxsp;
  const int32_t* rxnspecid;       /* [nrxn][rxn_maxsp+1]; [r][0] = count     */
  const double* rxnstoich;        /* [nrxn][rxn_maxsp]                      */
  const double* rxnh2ostoich;     /* [nrxn] or NULL                         */
  const double* rxn_logK;         /* [nrxn] (used when isothermal)          */
  const double* rxn_logKcoef;     /* [nrxn][5] or NULL (ReactionInterpolateLogK
                                     coefficients, non-isothermal runs)     */
  int32_t h_ion_id;               /* species_idx%h_ion_id: >= 0 H+ is that
                                     primary; -(k+1): H+ is complex k; only
                                     read when a PH constraint is present   */
  int32_t initialize_with_molality; /* reaction%initialize_with_molality   */
  int32_t use_prev_soln_as_guess; /* start from state->pri_molal           */
  int32_t max_iterations;         /* the reference stops at 10000          */
  double max_dlnC;                /* reaction%max_dlnC (default 5)          */
  const double* free_ion_guess;   /* [naqcomp] or NULL: 1e-9 (host array)   */
} pfchem_constraint;

/* ReactionEquilibrateConstraint (reaction.F90:1328-2117) on ncell cells:
   the free-ion molalities that satisfy the constraint, with
   conc[i * state->ld + cell] the constraint value of component i in that
   cell (device array; units as in the reference: mol/L for TOTAL unless
   initialize_with_molality, pH, log10 activity, log10 or partial pressure
   for GAS, a guess for MINERAL / CHARGE_BAL).  Writes pri_molal, total,
   sec_molal, pri_act_coef, sec_act_coef, ln_act_h2o, num_iterations and
   status (PFCHEM_CELL_OK or PFCHEM_CELL_CONSTRAINT_*); reads den_kg,
   temperature (non-isothermal) and, with use_prev_soln_as_guess,
   pri_molal.  The basis molarity of the reference is pri_molal *
   den_kg / 1000.  Mineral volume fractions and areas of a constraint are
   plain state values the caller writes.  No sorption, so the sorbed
   totals the reference computes afterwards do not arise.  Blocking:
   returns after the kernel finished on `stream`; PFCHEM_ERR_CELLS if a
   cell failed. */
int pfchem_equilibrate(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                       const pfchem_constraint* con, const double* conc, void* stream);

/* pfchem_equilibrate without any allocation in the call: no reference
   counterpart; for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   The constraint descriptor, its reaction table and the failure count are
   staged in a fixed 32 KiB device array of the library (one per GPU, held
   until the process ends) instead of context memory grown on demand, and
   the host image is built on the stack.  Same arguments, results and
   return codes as pfchem_equilibrate; a constraint whose staging exceeds
   32 KiB (about 190 reaction rows of 8 species) gives
   PFCHEM_ERR_UNSUPPORTED.  Every call on one GPU shares that array:
   overlapping calls on one GPU (other host threads or streams) race; each
   call clears its failure count and blocks until its kernel finished. */
int pfchem_equilibrate_static(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                              const pfchem_constraint* con, const double* conc, void* stream);

/* ------------------------------------------------------------------ */
/* OSRT transport step (the operator-split step before the reaction    */
/* step): RTCalculateTransportMatrix (reactive_transport.F90:1469-1703),*/
/* the boundary part of RTCalculateRHS_t1 (:1239-1465) and one linear  */
/* solve per aqueous component (pmc_subsurface_osrt.F90:286-333).      */
/* ------------------------------------------------------------------ */
/* Connection topology (host memory, read by pfchem_transport_setup):
   internal connections id_up -> id_dn (0-based cells; positive flux from
   up to dn) and boundary connections into cell bc_cell.  Cells without
   connections only accumulate. */
typedef struct pfchem_connections {
  int64_t nconn;
  const int64_t* id_up;      /* [nconn] */
  const int64_t* id_dn;      /* [nconn] */
  int64_t nbc;
  const int64_t* bc_cell;    /* [nbc]   */
} pfchem_connections;
/* Builds the sparse pattern of the transport matrix on the device (replaces
   a previous one).  ncell: cells of the states later passed. */
int pfchem_transport_setup(pfchem_ctx* ctx, int64_t ncell, const pfchem_connections* conn);

/* Per-step coefficients (device memory): the reference's connection area,
   Darcy flux (patch%internal_velocities / boundary_velocities, m/s) and
   transport coefficient over distance (patch%internal_tran_coefs /
   boundary_tran_coefs from RTUpdateTransportCoefs, m/s; the same for every
   component: species-dependent diffusion is refused by the reference's OSRT
   too), and the boundary totals (mol/L, [i * nbc + b]). */
typedef struct pfchem_transport_coefs {
  const double* area;               /* [nconn] m^2                       */
  const double* velocity;           /* [nconn] m/s, + from up to dn      */
  const double* coef_over_dist;     /* [nconn] m/s                       */
  const double* bc_area;            /* [nbc]                             */
  const double* bc_velocity;        /* [nbc] m/s, + into the cell        */
  const double* bc_coef_over_dist;  /* [nbc]                             */
  const double* bc_total;           /* [naqcomp * nbc]                   */
} pfchem_transport_coefs;
/* One implicit transport step of the state's aqueous totals (in place):
   per component  M c^{k+1} = psv c^k / dt + sum_bc T_up c_bc  with
   M = psv / dt + the TFluxCoef upwind advection / diffusion coefficients
   (transport.F90:495-565; psv = porosity sat 1000 volume, totals in mol/L);
   dry cells (saturation < 1e-40) have no fluxes and keep their totals (an
   empty row in the reference's matrix).  Solved on the device by
   BiCGSTAB with a Jacobi preconditioner (all components together) to a
   relative residual rtol (<= 0: 1e-12) within max_iterations (<= 0: 1000);
   iterations (host, optional) receives the iterations taken.  Source /
   sink terms are not supported.  Synchronizes `stream`;
   PFCHEM_ERR_CELLS: a component did not converge. */
int pfchem_transport_step(pfchem_ctx* ctx, int64_t ncell, const pfchem_state* state,
                          const pfchem_transport_coefs* coefs, double dt, double rtol,
                          int max_iterations, int32_t* iterations, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PFCHEM_H */
