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

/* ---- the 1-D karst conduit model (karst.c, permeabilities.c:795-867,
   set_source_term.c:256-414, solve_linearsystem_AMG.c:200-291,
   flux_reconstruction.c:322-496) and its coupling state (main.c:197-514).
   The reference's matrices X[i][j] (i: cell or breccia, j: conduit) are flat
   row-major arrays X[i * count + j] with mmax rows (cells) or max_brec rows
   (breccias). */
typedef struct {
    int count;                  /* number_karst */
    int mmax;                   /* Mmax: cells of the longest conduit */
    int max_brec;               /* rows of the breccia matrices */
    const int *cells;           /* M[j].s */
    const int *sx, *sy, *sz;    /* M[j].sxi, .syi, .szi */
    const int *axis;            /* M[j].pos: 1 x, 2 y, 3 z */
    const int *bc_l, *bc_r;     /* BoundKarst->l[j], ->r[j]: 0 Dirichlet, 1 Neumann, 3 Robin */
    const double *cont;         /* cont_karst[side][j] at [side * count + j] */
    const double *ccont;        /* C_contorno[side][j] at [side * count + j] */
    const int *nbrec;           /* n_brechas[j] */
    const int *brec;            /* brechas[i][j] */
    const double *beta;         /* beta[i][j] (R^4 - 1, as main.c stores it) */
    const double *C;            /* C[i][j] */
    double rho;
} oracle_conduits;

/* Per-conduit arrays the model keeps (main.c's globals and locals). */
typedef struct {
    double *KS;                 /* K->S[i][j] */
    double *KI_c;               /* KI_c[i][j] */
    double *hs;                 /* hs[j] */
    double *p_aux;              /* p_aux[i][j]: reservoir pressure along the conduit */
    double *pc;                 /* pc[i][j]: conduit pressure (out of conduit()) */
    double *pc_old;             /* pc_old[i][j]: previous pass (read by source_term1d) */
    double *ucR_old, *ucL_old;  /* breccia velocities [i][j], updated by source_term1d */
    double *kl, *kr, *kc;       /* keff->l, r, c [i][j] */
    int *ind_l, *ind_c, *ind_r; /* indcol1[i].l, .c, .r (mmax entries, shared by the conduits) */
} oracle_karst;

/* conduit (karst.c:8-42): for each conduit permeabilidade_karst, source_term1d
   and the 1-D solve; the solve is solve_linsys_1d + compute at `eps1d`
   (the reference: 1e-8), or with direct = 1 an exact tridiagonal (Thomas)
   solve of the same matrix. The reservoir's grid, K, perm and p_old feed the
   non-linear Robin ends. */
void oracle_conduit(const oracle_conduits *cs, oracle_karst *st, const oracle_grid *g, const double *Kx,
                    const double *Ky, const double *Kz, const oracle_ele *perm, const double *p_old, double visc,
                    double eps1d, int direct);

/* erro_relativo (main.c:106-119): sqrt(sum (y - y_old)^2) / sqrt(sum y^2),
   summed in index order. */
double oracle_rel_change(int n, const double *y, const double *y_old);

/* fluxo1d (flux_reconstruction.c:322-496): conduit fluxes of x (= pc). */
void oracle_fluxo1d(const oracle_conduits *cs, const oracle_karst *st, const oracle_grid *g, const double *Kx,
                    const double *Ky, const double *Kz, const oracle_ele *perm, const double *p_old, double visc,
                    const double *x, double *ucl, double *ucr);

#ifdef __cplusplus
}
#endif
#endif
