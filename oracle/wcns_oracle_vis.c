/*
 * wcns_oracle_vis.c -- CPU oracle (TEST INFRASTRUCTURE ONLY): viscous terms.
 * Placeholder until the laminar viscous restatement lands.
 */
#include "wcns_oracle.h"

int o_spectvis(o_case *c, int nb) { (void)c; (void)nb; return -100; }
int o_rhs_vis(o_case *c) { (void)c; return -100; }
