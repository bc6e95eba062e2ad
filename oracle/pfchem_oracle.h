/*
 * pfchem_oracle.h -- FP64 CPU restatement of PFLOTRAN's operator-split
 * chemistry step, used ONLY as the checker of the HIP kernels (tests,
 * __graft_entry__.smoke, bench.py's cpu_baseline leg).  Never linked into
 * the product library.
 *
 * Every function cites the reference routine it restates (paths relative to
 * src/pflotran of the PFLOTRAN tree).  One intended divergence: RSolve always
 * back-substitutes (reaction.F90:5506-5512); the reference's RReact branch
 * (:5501-5505) returns an unwritten update.
 */
#ifndef PFCHEM_ORACLE_H
#define PFCHEM_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "pfchem.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXCOMP 96   /* naqcomp + nimmobile                          */
#define ORC_MAXCPLX 512
#define ORC_MAXMNRL 32
#define ORC_MAXRXN 32    /* SomDec reactions                             */
#define ORC_MAX_RREACT_ATTEMPTS 10000 /* RReact calls per RStep (safety bound) */

/* one cell's auxvars: rt_auxvar + global_auxvar + material_auxvar subset
   (reactive_transport_aux.F90:21-74) */
typedef struct orc_cell {
  double pri_molal[ORC_MAXCOMP];
  double total[ORC_MAXCOMP];
  double immobile[ORC_MAXCOMP];
  double sec_molal[ORC_MAXCPLX];
  double pri_act_coef[ORC_MAXCOMP];
  double sec_act_coef[ORC_MAXCPLX];
  double ln_act_h2o;
  double mnrl_volfrac[ORC_MAXMNRL];
  double mnrl_area[ORC_MAXMNRL];
  double mnrl_rate[ORC_MAXMNRL];
  double porosity, sat, den_kg, volume, temp;
  double f_t, f_w;          /* SomDec ELM scalars (NaN = use functions)   */
  int has_f_t, has_f_w;
  double dtotal[ORC_MAXCOMP * ORC_MAXCOMP]; /* (i,j) at [i*naq+j]          */
  double eqcplx_logK[ORC_MAXCPLX];          /* current (T-dependent) logK  */
  double kinmnrl_logK[ORC_MAXMNRL];
  /* SomDec per-cell stoichiometry (the reference keeps these in the shared
     sandbox object, reaction_sandbox_somdec.F90:1817-1858; here per cell) */
  double sd_upstream_nc[ORC_MAXRXN];
  double sd_downstream_nc[ORC_MAXRXN * PFCHEM_SOMDEC_MAX_DOWN];
  double sd_mineral_c_stoich[ORC_MAXRXN];
  double sd_mineral_n_stoich[ORC_MAXRXN];
  int activity_error;       /* option%ierror from RActivityCoefficients  */
} orc_cell;

/* per-cell record of the last accepted sub-step's Newton system, used to
   evaluate relative residuals of other solvers' answers */
typedef struct orc_substep_record {
  double dt;                          /* tran_dt of last sub-step          */
  double fixed_total[ORC_MAXCOMP];    /* total^k at its start (mol/L)      */
  double fixed_immobile[ORC_MAXCOMP];
  double guess[ORC_MAXCOMP];          /* x0 of its Newton                  */
  double volfrac[ORC_MAXMNRL];        /* mineral state during it           */
  double r0_norm;                     /* ||F(x0)||_2                       */
  double final_rel_residual;          /* ||F(x*)||/||F(x0)|| at exit       */
} orc_substep_record;

/* accepted sub-steps of one RStep: start totals, mineral volume fractions
   before the kinetic update, solution and sub-step size (for the adjoint) */
#define ORC_LOG_MAX_SUBSTEPS 64
typedef struct orc_substep_log {
  int n, overflow;
  double dt[ORC_LOG_MAX_SUBSTEPS];
  double t0[ORC_LOG_MAX_SUBSTEPS][ORC_MAXCOMP];
  double vf0[ORC_LOG_MAX_SUBSTEPS][ORC_MAXMNRL];
  double m[ORC_LOG_MAX_SUBSTEPS][ORC_MAXCOMP];
} orc_substep_log;
/* when non-NULL, orc_rstep fills it (thread-local) */
extern __thread orc_substep_log* orc_log;

/* ---- building blocks (exported for unit tests) ---- */
void orc_cell_defaults(const pfchem_network* net, orc_cell* c);
void orc_update_temp_dependent_coefs(const pfchem_network* net,
                                     const pfchem_options* opt, orc_cell* c);
void orc_activity_coefficients(const pfchem_network* net,
                               const pfchem_options* opt, orc_cell* c);
void orc_total(const pfchem_network* net, orc_cell* c);
void orc_accumulation(const pfchem_network* net, const orc_cell* c,
                      double* res);
void orc_accumulation_derivative(const pfchem_network* net, const orc_cell* c,
                                 double dt, double* jac);
void orc_reaction(const pfchem_network* net, orc_cell* c, double dt,
                  double* res, double* jac, int compute_derivative);
void orc_reaction_derivative_numerical(const pfchem_network* net,
                                       const orc_cell* c, double dt,
                                       double* res, double* jac);
int orc_lu_decomposition(double* a, int n, int* indx, int* d);
void orc_lu_back_substitution(const double* a, int n, const int* indx,
                              double* b);
int orc_rsolve(const double* res, double* jac, const double* conc,
               double* update, int n, int use_log_formulation);
/* full RReact system at the cell's current free molality/immobile:
   res = (Acc(x) - fixed)/dt + R(x), jac = dres/dx (molal units) */
void orc_rreact_system(const pfchem_network* net, const pfchem_options* opt,
                       orc_cell* c, const double* fixed_accum, double dt,
                       double* res, double* jac, int compute_derivative);
int orc_rreact(const pfchem_network* net, const pfchem_options* opt,
               orc_cell* c, double* guess, double dt, int* num_iterations,
               orc_substep_record* rec, int* fail_reason);
int orc_rstep(const pfchem_network* net, const pfchem_options* opt,
              orc_cell* c, double* guess, double dt, int* num_sub_steps,
              int* num_iterations, int* num_kinetic_state_updates,
              orc_substep_record* rec);
void orc_update_kinetic_state(const pfchem_network* net, orc_cell* c,
                              double dt);

/* ---- batch drivers over SoA state (host pointers, pfchem_state layout) ---- */
void orc_gather(const pfchem_network* net, const pfchem_state* s, int64_t cell,
                orc_cell* c);
void orc_scatter(const pfchem_network* net, const orc_cell* c,
                 const pfchem_state* s, int64_t cell);
/* RStep over cells [0, ncell) with `nthreads` OpenMP threads (<=0: all).
   If rec != NULL it receives one record per cell.  Returns number of failed
   cells. */
int64_t orc_step_batch(const pfchem_network* net, const pfchem_options* opt,
                       const pfchem_state* s, int64_t ncell, double dt,
                       int nthreads, orc_substep_record* rec);
/* activity coefficient update over cells (RTUpdateActivityCoefficients) */
void orc_update_activity_batch(const pfchem_network* net,
                               const pfchem_options* opt,
                               const pfchem_state* s, int64_t ncell,
                               int nthreads);
/* ||F(x)||_2 / r0 of a record's system at the free molality stored in s
   (pri_molal / immobile of the given cell). */
double orc_record_rel_residual(const pfchem_network* net,
                               const pfchem_options* opt,
                               const pfchem_state* s, int64_t cell,
                               const orc_substep_record* rec);
size_t orc_sizeof_record(void);

/* ---- constraint speciation (pfchem_oracle_constraint.c) ---- */
/* ReactionEquilibrateConstraint for one cell: conc[naqcomp] constraint
   values.  Returns PFCHEM_CELL_OK or PFCHEM_CELL_CONSTRAINT_*, -1 for a
   constraint type not restated. */
int orc_equilibrate(const pfchem_network* net, const pfchem_options* opt,
                    const pfchem_constraint* con, const double* conc, orc_cell* c,
                    int* num_iterations);
/* over cells [0, ncell); conc SoA with the state's ld; without
   use_prev_soln_as_guess every cell starts with unit activity
   coefficients.  Returns the number of failed cells. */
int64_t orc_equilibrate_batch(const pfchem_network* net, const pfchem_options* opt,
                              const pfchem_constraint* con, const pfchem_state* s,
                              const double* conc, int64_t ncell, int nthreads);

/* ---- sensitivities of one step (pfchem_oracle_adjoint.c) ----
   Objective of a cell: f = sum_m wv[m] VF_m(end) + sum_k ws[k] m_k(end)
   (end-of-step mineral volume fractions and secondary molalities).
   Parameters p = (log10 k_m for the nkinmnrl minerals, log10 K_k for the
   neqcplx complexes).  exact_jacobian = 0 uses the reference's analytic
   Jacobian as assembled (its mineral block carries rho/1000,
   reaction_mineral.F90:945); 1 the exact derivative.  Return 0, or
   1 step failed, 2 unsupported (immobile species, activity coefficients
   updated inside Newton, mineral affinity options), 3 too many sub-steps. */
int orc_cell_adjoint(const pfchem_network* net, const pfchem_options* opt,
                     orc_cell* c, double dt, const double* wv, const double* ws,
                     int exact_jacobian, double* f, double* grad);
int orc_cell_tangent(const pfchem_network* net, const pfchem_options* opt,
                     orc_cell* c, double dt, const double* wv, const double* ws,
                     const double* dp, int exact_jacobian, double* f, double* df);
/* over cells: grad_sum[nkinmnrl + neqcplx] += per-cell gradients; f[cell]
   and status[cell] per cell (may be NULL).  Returns failed cells. */
int64_t orc_adjoint_batch(const pfchem_network* net, const pfchem_options* opt,
                          const pfchem_state* s, int64_t ncell, double dt,
                          const double* wv, const double* ws, int exact_jacobian,
                          int nthreads, double* grad_sum, double* f, int32_t* status);
int64_t orc_tangent_batch(const pfchem_network* net, const pfchem_options* opt,
                          const pfchem_state* s, int64_t ncell, double dt,
                          const double* wv, const double* ws, const double* dp,
                          int exact_jacobian, int nthreads, double* df, double* f,
                          int32_t* status);
size_t orc_sizeof_cell(void);
void orc_cell_total(const pfchem_network* net, const pfchem_options* opt,
                    const pfchem_state* s, int64_t cell, double* total,
                    double* sec_molal, double* dtotal);

#ifdef __cplusplus
}
#endif
#endif
