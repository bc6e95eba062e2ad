/*
 * pfchem_oracle_somdec.c -- oracle restatement of the SomDec decomposition
 * sandbox (reaction_sandbox_somdec.F90).  TEST INFRASTRUCTURE ONLY.
 */
#include "pfchem_oracle.h"

void orc_somdec_react(const pfchem_network* net, orc_cell* c, double dt,
                      double* res, double* jac, int compute_derivative) {
  (void)net; (void)c; (void)dt; (void)res; (void)jac; (void)compute_derivative;
}
