/*
 * Seeded permeability-field generators (host C, no GPU).
 *
 * These replace the reference's permeability-file readers
 * (permeabilities.c:57-186, 494-668, 673-774, unreachable in the reference
 * because main.c:341 leaves the file name unset) with generators that are a
 * pure function of (grid, parameters, seed). The homogeneous case matches
 * perm_homogeneous() (permeabilities.c:776-789) with a caller-chosen value.
 *
 * All return 0 on success, -1 on bad arguments, -2 on allocation failure.
 * Layout: K[s], s = i + j*nx + k*nx*ny (pindex(), solve_linearsystem_AMG.c:15-20).
 */
#ifndef HMFEM_FIELDS_H
#define HMFEM_FIELDS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int hmfem_field_homogeneous(int nx, int ny, int nz, double k, double *K);

/* ln K = mean_ln + sigma * g, g unit-variance with covariance exp(-r^2/corr_len^2). */
int hmfem_field_lognormal(int nx, int ny, int nz, double mean_ln, double sigma,
                          double corr_len, uint64_t seed, double *K);

/* Multiply K in place by 10^U(4,6) along `count` seeded conduit polylines. */
int hmfem_field_conduits(int nx, int ny, int nz, int count, uint64_t seed, double *K);

uint64_t hmfem_field_instance_seed(uint64_t seed, int instance);
double hmfem_field_uniform(uint64_t *state);

#ifdef __cplusplus
}
#endif

#endif
