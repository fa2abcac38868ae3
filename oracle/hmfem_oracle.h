/*
 * CPU restatement of the reference's pressure-solve path. TEST INFRASTRUCTURE:
 * only the tests, the smoke check and the benchmark's CPU-baseline leg may load
 * it, and only as the checker. The product path never calls it.
 *
 * Each function cites the reference function it restates. Arithmetic follows
 * the reference's operation order so that, built without contraction
 * (-ffp-contract=off), results match the reference build bit for bit.
 */
#ifndef HMFEM_ORACLE_H
#define HMFEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same member order as ELE / IND (structs.h:7-17, 38-48). */
typedef struct { double u, d, l, r, c, t, b; } oracle_ele;
typedef struct { int c, u, d, l, r, t, b; } oracle_ind;

/* Face order used throughout: l, r, u, d, t, b (FACE, structs.h:50-59). */
typedef struct {
    int nx, ny, nz;
    double hx, hy, hz;
    double visc;       /* viscosidade */
    int bc_type[6];    /* 0 Dirichlet, 1 Neumann, per face l, r, u, d, t, b */
} oracle_grid;

/* Per-face boundary values, indexed as the reference's Bound arrays:
   l, r: pyzindex = k*ny + j;  u, d: pxzindex = k*nx + i;  t, b: pxyindex = j*nx + i. */
typedef struct { const double *l, *r, *u, *d, *t, *b; } oracle_bound;

/* permeabilidade_arestas (permeabilities.c:196-492) */
void oracle_permeabilidade_arestas(const oracle_grid *g, const double *Kx, const double *Ky, const double *Kz,
                                   const double *WI, const double *KI, oracle_ele *perm, oracle_ind *indcol);

/* source_term (set_source_term.c:9-223); accumulates into b. */
void oracle_source_term(const oracle_grid *g, const oracle_bound *B, const oracle_ele *perm,
                        const double *pc, const double *pw, const double *KI, const double *WI, double *b);

/* Count-CSR build of solve_linsys (solve_linearsystem_AMG.c:57-143). Returns nnz
   (0 means the reference would stop with "No nonzero elements"). col/ele need 7n. */
long oracle_build_csr(int n, const oracle_ele *perm, const oracle_ind *indcol, int *cnt, int *col, double *ele);

/* prod_mat_vec (solve_linearsystem_AMG.c:160-194); res += A x. */
void oracle_prod_mat_vec(int n, const oracle_ele *perm, const oracle_ind *indcol, const double *x, double *res);

/* AMG hierarchy (amg.h:1135-1239) with the toolbox's greedy C/F selection,
   averaging interpolation and Galerkin triple product. */
typedef struct oracle_amg oracle_amg;
oracle_amg *oracle_amg_create(int n, const int *cnt, const int *col, const double *ele,
                              int max_level, int min_nodes, double epsilon);
void oracle_amg_destroy(oracle_amg *a);
int oracle_amg_levels(const oracle_amg *a);                 /* number of levels */
void oracle_amg_level_size(const oracle_amg *a, int l, int *rows, long *nnz);
/* Copies of level l's matrix (count-CSR) and of its interpolation to level l+1;
   pass NULL pointers to query sizes only. Return the number of entries. */
long oracle_amg_level_matrix(const oracle_amg *a, int l, int *cnt, int *col, double *ele);
long oracle_amg_level_interp(const oracle_amg *a, int l, int *pcnt, int *pcol, double *pele);
/* One V-cycle (amg.h:1261-1367, NOMPI Gauss-Seidel branch): u = M f. */
void oracle_amg_apply(oracle_amg *a, const double *f, double *u);

/* conjugate_gradient (pcg.h:33-75). Returns the loop counter (_loop). */
int oracle_pcg(int n, const int *cnt, const int *col, const double *ele, oracle_amg *amg,
               const double *f, double *u, double eps, int max_iterations);

/* compute (compute.cpp:7-50) with a caller-chosen stop; returns iterations (_loop). */
int oracle_compute(int n, int nnz, const int *cnt, const int *col, const double *ele,
                   const double *b, double *x, double eps);

/* fluxos (flux_reconstruction.c:9-279); writes u,d,l,r,t,b of v (not c). */
void oracle_fluxos(const oracle_grid *g, const double *p, const oracle_ele *perm, const oracle_bound *B, oracle_ele *v);

/* divvel_pontual (flux_reconstruction.c:281-319) without the print. */
void oracle_divvel_pontual(const oracle_grid *g, const oracle_ele *v, double *divmin, double *divmax);

#ifdef __cplusplus
}
#endif
#endif
