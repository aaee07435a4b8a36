/*
 * FP64 CPU oracle of gridpointd's parameterizations: TEST INFRASTRUCTURE
 * ONLY (see exo_oracle.c).
 */
#define EXO_ORACLE_INTERNAL
#include "exo_oracle.h"

void oracle_physics(const oracle_member* m) { (void)m; }
