/*
 * oracle.c — single translation unit of the CPU oracle (test infrastructure
 * only): the Cartesian restatement and the simplex restatement share the
 * block-sparse arithmetic and the multigrid of cart.c.
 */
#include "cart.c"
#include "simplex.c"
