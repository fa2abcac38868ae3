/*
 * libhmfem_gpu: MI355X pressure solver for hmfem's cell-centred Darcy model.
 *
 * C ABI only (plain pointers and sizes). Two layers:
 *
 *  1. Drop-in replacements for the reference's functions on the pressure
 *     path, same names, signatures and argument meaning. A host that keeps
 *     reservoir() (reservoir.c:8-57) links this library instead of the
 *     reference's permeabilities.c / set_source_term.c /
 *     solve_linearsystem_AMG.c / compute.cpp / flux_reconstruction.c.
 *     Host buffers in, host buffers out; the work runs on the GPU.
 *
 *  2. A device-resident API (hmfem_*) with all arrays in HBM, options for
 *     precision and stopping rule, and a status record.
 *
 * Types: if the reference's structs.h was included first (guard STRUCTS_H),
 * its definitions are used; otherwise this header defines layout-compatible
 * copies of the types the path needs.
 */
#ifndef HMFEM_GPU_H
#define HMFEM_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef STRUCTS_H
/* structs.h:7-17 */
typedef struct ELEstruct { double u, d, l, r, c, t, b; } ELE;
/* structs.h:27-36 */
typedef struct Boundary { double *u, *d, *l, *r, *t, *b; } Bound;
/* structs.h:38-48 */
typedef struct Index { int c, u, d, l, r, t, b; } IND;
/* structs.h:50-59 */
typedef struct Faces { int u, d, l, r, t, b; } FACE;
/* structs.h:71-79 */
typedef struct Permeab { double *X, *Y, *Z; double **S; int permeability; } Perm;
/* structs.h:83-88 */
typedef struct Grid { double x, y, z; } grid;
/* structs.h:93-104 */
typedef struct Dimension { int x, y, z, xy, t, px, pz, py; } Dim;
/* structs.h:19-25: per-conduit matrices [cell][conduit] */
typedef struct ELEstruct_K { double **l, **r, **c; } ELEK;
/* structs.h:61-67 */
typedef struct FacesKarst { int *l, *r; } FACEK;
/* structs.h:107-115 */
typedef struct Dimension_karst { int s, sxi, syi, szi, pos; } DimK;
#endif

/* ---------------------------------------------------------------- layer 1 */

/* Globals the reference's path reads or writes (variables.h). Defined here
   with default visibility: a host executable that defines them too wins by
   ELF interposition, so the library and the host share one copy. */
extern int iter_amg;    /* written by compute()/solve_linsys(): PCG loop count (compute.cpp:46) */
extern FACE boundary;   /* read by fluxos() (flux_reconstruction.c:104-270) */

/* solve_linearsystem_AMG.c:37 solve_linsys(); prototype functions.h:54.
   perm/indcol as filled by permeabilidade_arestas(), b the RHS, x in/out
   (initial guess in, pressure out). Stops as the reference's PCG does
   (pcg.h:57): when <M r, r> has fallen to 1e-16 of its value at the initial
   guess, so a warm start is tightened further (HMFEM_RTOL replaces 1e-8;
   HMFEM_STOP=residual stops at ||b - A x|| <= rtol ||b|| instead). Writes
   iter_amg (iterations + 1, as compute.cpp:46). Exits like the reference when
   the matrix has no entries. With HMFEM_REUSE_SETUP=1 in the environment,
   solve_linsys() and compute() reuse the previous call's hierarchy on the
   same grid (hmfem_options.reuse_setup); 1-D systems (the conduits') use a
   context of their own, so they do not evict the reservoir's. */
void solve_linsys(ELE *perm, IND *indcol, double *b, double *x, Dim N);

/* compute.cpp:7 compute(); prototype functions.h:53. Count-CSR: hdr[0] = n,
   hdr[1] = nnz, row i holds the next cnt[i] entries, 0-based columns, FP64
   values; the matrix must be symmetric. x in/out. Same stopping rule and
   iter_amg as solve_linsys(). Returns 0; a matrix the library cannot solve
   ends the process with a message (the reference's callers ignore the
   return value, solve_linearsystem_AMG.c:150, :283). */
int compute(int *hdr, int *cnt, int *col, double *ele, double *b, double *x);

/* permeabilities.c:196 permeabilidade_arestas(); functions.h:31. */
int permeabilidade_arestas(Perm *K, ELE *perm, IND *indcol, grid h, Dim N, FACE boundary_types,
                           double *WI, double *KI, double viscosidade);

/* set_source_term.c:9 source_term(); functions.h:45. Accumulates into b. */
void source_term(Bound *B, ELE *perm, grid h, Dim N, FACE boundary_types, double *pc_aux, double *pw,
                 double *KI, double *WI, double viscosidade, double *b);

/* flux_reconstruction.c:9 fluxos(); functions.h:50. Reads the global
   `boundary`; writes u,d,l,r,t,b of v; prints the divergence line. */
int fluxos(double *p, ELE *perm, ELE *v, Bound *B, grid h, Dim N);

/* flux_reconstruction.c:281 divvel_pontual(); functions.h:51. */
void divvel_pontual(ELE *v, grid h, Dim N);

/* solve_linearsystem_AMG.c:160 prod_mat_vec(); functions.h:56. res += A x. */
void prod_mat_vec(ELE *perm, IND *indcol, double *x, double *res, Dim N);

/* Output writers, byte for byte the reference's files, into the directory
   `dirname` (variables.h; main.c:343 sets it from argv; default "."). Host
   buffers; the lines are formatted on several threads. */
extern const char *dirname;
/* No counterpart in the reference: the output directory for hosts that do
   not define `dirname` (it takes precedence; the string must stay alive). */
void hmfem_set_output_dir(const char *dir);
/* posprocessing_output.c:200 save_pressure(); functions.h:67. Writes
   pressure_reservoir.dat; exits like the reference if it cannot open it. */
void save_pressure(double *x, Dim N, grid h);
/* posprocessing_output.c:236 save_flux(); functions.h:68. Writes
   flux_reservoir.dat (cell-centre velocities 0.5(r-l), 0.5(u-d), 0.5(t-b)). */
void save_flux(ELE *v, Dim N, grid h);
/* vtk.c:8 vts_StructuredGrid_file(). Writes campos.vts (velocity, pressure,
   cell centres). Returns -1 if it cannot open the file (the reference
   carries on writing to a null stream). */
int vts_StructuredGrid_file(ELE *v, double *x, Dim N, grid h);

/* posprocessing_output.c:437 save_pressure_1d(); functions.h:73. Writes
   pressure_karst_<j>.dat for each of the number_karst conduits (x[i][j]:
   cell i of conduit j), the cell centres along the conduit's direction
   M[j].pos (1 x, 2 y, 3 z; 4-7 the diagonals the reference's writer
   knows). */
extern int number_karst;    /* variables.h; the host's copy is read by ELF interposition */
void save_pressure_1d(double **x, DimK *M, double *hs, grid h);
/* posprocessing_output.c:536 save_flux_1d(); functions.h:72. Writes
   flux_karst_<j>.dat: cell-centre velocity 0.5 (uc->r - uc->l). */
void save_flux_1d(ELEK *uc, DimK *M, double *hs, grid h);

/* ---------------------------------------------------------------- layer 2 */

enum {
    HMFEM_OK = 0,
    HMFEM_NOT_CONVERGED = 1,
    HMFEM_ERR_ARG = -1,
    HMFEM_ERR_HIP = -2,
    HMFEM_ERR_NOMEM = -3,
    HMFEM_ERR_NOT_STRUCTURED = -4,
    HMFEM_ERR_BREAKDOWN = -5
};

enum { HMFEM_PREC_FP64 = 0, HMFEM_PREC_FP32 = 1, HMFEM_PREC_FP16 = 2 };

/* HMFEM_STOP_RESIDUAL: ||b - A x||_2 <= rtol ||b||_2.
   HMFEM_STOP_PRECONDITIONED: <M r, r> <= rtol^2 <M r0, r0> (pcg.h:57). */
enum { HMFEM_STOP_RESIDUAL = 0, HMFEM_STOP_PRECONDITIONED = 1 };
/* Smoother of the multi-block V-cycle levels (hmfem_options.smoother). */
enum { HMFEM_SMOOTH_JACOBI = 0, HMFEM_SMOOTH_BLOCK_GS = 1 };
/* Preconditioner of the CG (hmfem_options.precond). */
enum { HMFEM_PRECOND_AMG = 0, HMFEM_PRECOND_JACOBI = 1 };

typedef struct {
    int nx, ny, nz;
    double hx, hy, hz;
    double viscosidade;
    int bc_type[6];          /* faces l, r, u, d, t, b: 0 Dirichlet, 1 Neumann */
} hmfem_grid;

typedef struct {
    double rtol;             /* stopping tolerance (default 1e-8) */
    int max_iterations;      /* default 10000, as pcg.h's _max */
    int precision;           /* storage/arithmetic of the AMG hierarchy (default FP32). HMFEM_PREC_FP16:
                                FP32 hierarchy whose fine level the smoother and the restriction
                                read as FP16 couplings of the diagonally scaled operator plus an
                                FP32 scale (single solves with HMFEM_SMOOTH_BLOCK_GS only; the
                                batched and block solves use FP32) */
    int stop_rule;           /* default HMFEM_STOP_RESIDUAL */
    int nu;                  /* Jacobi sweeps before and after the coarse correction (default 3) */
    double omega;            /* Jacobi damping (default 0.95) */
    double ocf;              /* coarse-correction scaling; default 0: 2.2 in single and block solves,
                                2.0 in batched solves (measured best on each) */
    int coarse_max;          /* largest coarsest level, solved densely (default and maximum 80) */
    int check_every;         /* CG iterations queued between host convergence checks (default 4) */
    int tail_max;            /* levels of at most this many cells run in one single-workgroup
                                kernel (default 512; 0 disables) */
    int tail_gamma;          /* cycle inside that kernel: 1 V-cycle, 2 W-cycle (default 2) */
    int use_graph;           /* replay each CG iteration from a captured HIP graph (default 1) */
    int tail_lds;            /* run that kernel on LDS copies of its levels when they fit (default 1) */
    int smoother;            /* smoother of the multi-block levels: HMFEM_SMOOTH_JACOBI (nu damped Jacobi
                                sweeps) or HMFEM_SMOOTH_BLOCK_GS (gs_sweeps red-black Gauss-Seidel sweeps
                                inside 32x8x8 tiles, one pass per smoothing step; the default); the
                                levels of the single-workgroup kernel always use Jacobi */
    int gs_sweeps;           /* sweeps per smoothing step of HMFEM_SMOOTH_BLOCK_GS; default 0: two in single
                                and block solves, four in batched solves (measured best on each) */
    int autotune;            /* 1: hmfem_reservoir_device first runs hmfem_autotune_device for each new
                                grid shape / precision / tolerance of the context and then uses the
                                options it chose; hmfem_reservoir_batch_device likewise runs
                                hmfem_autotune_batch_device for each new batch shape and uses its
                                per-instance options (default 0) */
    int deflate;             /* 1: add a coarse correction on each 7-point-connected cluster of at least
                                8 cells whose diagonal exceeds deflate_ratio times the median diagonal
                                (karst conduits); off when there are none (default 1; single, block and
                                batched solves; in a batch each instance has its own threshold).
                                2: single solves use the balancing form instead (the V-cycle on
                                r - A Z c, its output projected onto the complement of the clusters'
                                modes: fewer iterations, a larger error at a given residual); block
                                and batched solves treat 2 as 1 */
    double deflate_ratio;    /* default 3000 */
    int precond;             /* HMFEM_PRECOND_AMG (default), or HMFEM_PRECOND_JACOBI: z = r / diag(A) in
                                FP64 with no hierarchy, the plain FP64 PCG that the AMG is measured
                                against (single solves; precision, smoother and deflate are ignored) */
    int reuse_setup;         /* 1: if the previous single solve on this context built its hierarchy (and
                                conduit correction) for the same grid and setup options, reuse them
                                instead of rebuilding: the preconditioner then follows the earlier
                                operator while the CG solves the current one (for repeated solves with
                                small operator changes, e.g. the fixed-point passes of reservoir());
                                default 0 */
    int semicoarsen;         /* 1: coarsen an axis only where its couplings are at least a quarter of the
                                strongest axis's (geometric means over the fine operator): anisotropic cells
                                or K get boxes of 2x1x1, 2x2x1, ... until the couplings even out; isotropic
                                problems keep 2x2x2 boxes (single solves; default 1) */
    int block_compact;       /* 1: a block solve through four lanes goes on with two once no right-hand
                                side is left to start and at most two lanes are busy (their vectors are
                                repacked; the kernels then move half the bytes per iteration); default 1 */
} hmfem_options;

typedef struct {
    int status;              /* HMFEM_OK, HMFEM_NOT_CONVERGED or an error */
    int iterations;          /* CG iterations performed */
    int levels;              /* hierarchy levels including the dense coarsest */
    double rel_residual;     /* ||b - A x|| / ||b|| at exit (recursively updated r) */
    double precond_ratio;    /* sqrt(<M r, r> / <M r0, r0>) at exit */
    double t_assemble_ms, t_setup_ms, t_solve_ms, t_flux_ms;
    double div_max, div_min; /* divvel_pontual() of the fluxes, when computed */
    double solve_bytes;      /* algorithmic HBM bytes of the solve phase */
    int deflation_clusters;  /* clusters of the conduit correction (0: none) */
    int deflation_cells;     /* cells in those clusters */
} hmfem_status;

typedef struct hmfem_context hmfem_context;

void hmfem_default_options(hmfem_options *opt);
int hmfem_create(hmfem_context **ctx);
void hmfem_destroy(hmfem_context *ctx);
/* The HIP stream (hipStream_t) every call of this context is ordered on. */
void *hmfem_stream(hmfem_context *ctx);

/* Assembly only (permeabilidade_arestas + source_term), device pointers.
   perm/indcol (AoS, may be NULL) and b (must be non-NULL, overwritten).
   KI, WI, pc, pw may be NULL (zero). bval[f] is face f's value array in the
   Bound layout (faces l, r, u, d, t, b). */
int hmfem_assemble_device(hmfem_context *ctx, const hmfem_grid *g, const double *Kx, const double *Ky,
                          const double *Kz, const double *KI, const double *WI, const double *pc,
                          const double *pw, const double *const bval[6], ELE *perm, IND *indcol, double *b);

/* The whole reservoir() step on the device: assembly, AMG setup, FCG solve
   (x in/out in p), then fluxes v (AoS, may be NULL) and their divergence. */
int hmfem_reservoir_device(hmfem_context *ctx, const hmfem_grid *g, const double *Kx, const double *Ky,
                           const double *Kz, const double *KI, const double *WI, const double *pc,
                           const double *pw, const double *const bval[6], double *p, ELE *v,
                           const hmfem_options *opt, hmfem_status *st);

/* Fluxes and divergence from a pressure (fluxos + divvel_pontual), device pointers. */
int hmfem_flux_device(hmfem_context *ctx, const hmfem_grid *g, const double *Kx, const double *Ky,
                      const double *Kz, const double *const bval[6], const double *p, ELE *v,
                      double *div_min, double *div_max);

/* No counterpart in the reference: hmfem_flux_device without any device
   allocation, for callers that cannot allocate (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc). Same arguments and
   results as hmfem_flux_device; its scratch (the flux kernel's per-block
   min/max partials) is a fixed 8 MiB device array per GPU, resolved when the
   context is created, so grids of more than 2^27 cells are rejected with
   HMFEM_ERR_ARG. Every call on one GPU shares that array: calls that overlap
   on one GPU (different contexts or streams) race. */
int hmfem_flux_device_noalloc(hmfem_context *ctx, const hmfem_grid *g, const double *Kx, const double *Ky,
                              const double *Kz, const double *const bval[6], const double *p, ELE *v,
                              double *div_min, double *div_max);

/* No counterpart in the reference (its hierarchy is internal to amg_solver,
   amg.h:1135-1239): the hierarchy of the context's last single AMG solve,
   for inspection and tests. level < 0: returns the number of levels (the
   last is solved densely). level >= 0: dims[3] receives its grid, fac[3] the
   box factors towards the next level, and e, wx, wy, wz (host arrays of
   nx*ny*nz, each may be NULL) its FP64 flux-form coefficients
   ((A u)_s = e_s u_s + sum_f w_f (u_s - u_nb), w the coupling to the +x/+y/+z
   neighbour); returns the number of levels. HMFEM_ERR_ARG if the context
   holds no hierarchy (none built yet, or a later plain-PCG, block or batch
   solve replaced it). */
int hmfem_hierarchy_level(hmfem_context *ctx, int level, int *dims, int *fac, double *e, double *wx, double *wy,
                          double *wz);
/* The dense FP64 inverse of that hierarchy's coarsest level (nc x nc,
   row-major, host; may be NULL to query nc), as the setup computed it on the
   matrix cores; returns nc or HMFEM_ERR_ARG. No counterpart in the reference,
   whose coarsest "solve" is a diagonal scaling (amg.h:1361-1362). */
int hmfem_hierarchy_coarse_inverse(hmfem_context *ctx, double *inv);

/* Solve on an already assembled structured operator given in AoS form (device
   perm/indcol as from hmfem_assemble_device), b and x device FP64. */
int hmfem_solve_structured_device(hmfem_context *ctx, int nx, int ny, int nz, const ELE *perm,
                                  const IND *indcol, const double *b, double *x,
                                  const hmfem_options *opt, hmfem_status *st);

/* Autotuner: solves the given problem (no sources, zero initial guess) with
   a fixed list of cycle variants (smoother and sweeps, coarse-correction
   scaling, tail cycle), twice each, and writes the fastest converged one
   (setup + solve time) into *opt; best_ms (may be NULL) receives its time.
   The tolerance, precision and stopping rule of *opt are kept. No
   counterpart in the reference (its cycle is fixed, amg.h:1135-1239). */
int hmfem_autotune_device(hmfem_context *ctx, const hmfem_grid *g, const double *Kx, const double *Ky,
                          const double *Kz, const double *const bval[6], hmfem_options *opt, double *best_ms);

/* Batch of independent problems solved together (every launch covers all
   of them): instance q is an nx*ny*nz grid given by dims[3q..3q+2] (host),
   with extents lengths[3q..3q+2] (host; NULL = unit cube). All instances
   share viscosidade, the face types and the face value arrays bval[f]
   (device, each at least as long as the largest face of the batch).
   Kx, Ky, Kz, p and v are the instances' arrays concatenated in order
   (device). Per-instance results go to the optional host arrays
   iterations, rel_residual, done_ms (time from the start of the solve
   phase at which the instance's stop test first held, resolution
   check_every iterations) and div_max. */
typedef struct {
    int status;              /* HMFEM_OK if every instance converged */
    int max_iterations;      /* largest iteration count over the instances */
    double t_assemble_ms, t_setup_ms, t_solve_ms, t_flux_ms;
    int deflation_clusters;  /* clusters of the conduit correction, all instances (0: none) */
    int deflation_cells;
} hmfem_batch_status;

int hmfem_reservoir_batch_device(hmfem_context *ctx, int ninst, const int *dims, const double *lengths,
                                 double viscosidade, const int bc_type[6], const double *const bval[6],
                                 const double *Kx, const double *Ky, const double *Kz, double *p, ELE *v,
                                 const hmfem_options *opt, hmfem_batch_status *st, int *iterations,
                                 double *rel_residual, double *done_ms, double *div_max);

/* Batch autotuner: block Gauss-Seidel sweeps and coarse-correction scaling
   for every instance (host arrays gs_sweeps[ninst], ocf[ninst]). Every
   candidate pair (sweeps 2-4, scaling 2.0-2.5) solves the whole batch from
   the initial guess p (device, kept; NULL: zero) twice; each instance takes
   the pair with the least iterations x cells x cost per cell-iteration of
   that pair (its batch time over the summed work), and the per-instance
   choice is kept if a batch run with it is faster than the best single pair,
   else every instance gets that pair. best_ms (may be NULL): the chosen
   setup + solve time. The other fields of *opt are kept. No counterpart in
   the reference (one cycle for every solve, amg.h:1135-1239). */
int hmfem_autotune_batch_device(hmfem_context *ctx, int ninst, const int *dims, const double *lengths,
                                double viscosidade, const int bc_type[6], const double *const bval[6],
                                const double *Kx, const double *Ky, const double *Kz, const double *p,
                                const hmfem_options *opt, int *gs_sweeps, double *ocf, double *best_ms);

/* Block solve: nrhs (1..8) right-hand sides of one operator, solved
   together (one hierarchy; every kernel reads the coefficients once for all
   right-hand sides; vectors interleaved [cell][right-hand side] inside).
   Right-hand side r has the face values bval[6 r + f] (device, Bound layout,
   faces l, r, u, d, t, b; the face TYPES, hence the matrix, are g->bc_type
   for all) plus, if q[r] is not NULL, the device array q[r] (n values) added
   to b (sources entered in b only, the matrix unchanged). p[r] (device, n) is
   the initial guess in and the pressure out; v[r] (device ELE[n], or NULL)
   receives the fluxes. Per right-hand side: iterations, rel_residual and
   div_max (host arrays of nrhs, or NULL). Uses the FP32 hierarchy and the
   block Gauss-Seidel smoother whatever opt says; the right-hand sides go
   through four lanes (two if there are only one or two) of one interleaved
   solve: the first four start together and each further one takes the first
   lane whose right-hand side has converged, so that no lane idles while
   work is left (the lane restarts its own CG; the hierarchy and the conduit
   correction, opt.deflate, are built once per call). No counterpart
   in the reference's API: it is the interleaved multi-RHS solve of
   generic_operator.h:335-391 behind the reservoir() step. */
typedef struct {
    int status;              /* HMFEM_OK if every right-hand side converged */
    int max_iterations;
    double t_assemble_ms, t_setup_ms, t_solve_ms, t_flux_ms;
    double footprint_bytes;  /* device memory the call used */
    int deflation_clusters;  /* clusters of the conduit correction (0: none) */
    int deflation_cells;
    int block_iterations;    /* iterations of the interleaved solve (every lane advances once per iteration) */
} hmfem_block_status;

int hmfem_reservoir_block_device(hmfem_context *ctx, const hmfem_grid *g, const double *Kx, const double *Ky,
                                 const double *Kz, int nrhs, const double *const *bval, const double *const *q,
                                 double *const *p, ELE *const *v, const hmfem_options *opt,
                                 hmfem_block_status *st, int *iterations, double *rel_residual, double *div_max);

/* The reference's 1-D karst conduit model (karst.c, main.c:197-300): each
   conduit is a line of cells along one axis of the reservoir grid, with its
   own permeability, exchanging fluid with the reservoir cells it runs
   through (karst index), with collapse breccias (a quadratic head loss
   between two of its cells) and Dirichlet (0), Neumann (1) or non-linear
   Robin (3) ends. Host arrays, laid out as the reference's data file
   (data_<nk>_<type>.dat) lists them: [j] per conduit, [side*count + j] per
   end (side 0 the start, 1 the far end), [i*count + j] per breccia
   (i < max_breccias). */
typedef struct {
    int count;                   /* number_karst */
    const int *cells;            /* M[j].s */
    const int *start;            /* [3*j .. 3*j+2]: M[j].sxi, syi, szi (first cell) */
    const int *axis;             /* M[j].pos: 1 x, 2 y, 3 z */
    const double *Kc;            /* conduit permeability */
    const int *bc_l, *bc_r;      /* BoundKarst->l[j], ->r[j] */
    const double *cont;          /* cont_karst[side][j]: end value (Robin: R^4) */
    const double *ccont;         /* C_contorno[side][j] */
    const double *karst_index;   /* KI */
    const double *area;          /* Ac */
    const int *breccias;         /* n_brechas[j] */
    int max_breccias;
    const int *breccia_cell;     /* brechas[i][j]: 0 <= cell < M[j].s - 1 */
    const double *breccia_beta;  /* beta[i][j] = R^4 - 1 (as main.c stores it) */
    const double *breccia_C;     /* C[i][j] */
    double rho;                  /* fluid density (main.c:130: 1000) */
} hmfem_conduit_set;

typedef struct {
    int status;                  /* HMFEM_OK, HMFEM_NOT_CONVERGED (pass cap reached before tol) or an error */
    int passes;                  /* fixed-point passes after the first reservoir solve (main.c's iter) */
    double change;               /* relative change of the reservoir pressure in the last pass */
    int iterations;              /* CG iterations of all reservoir solves */
    double t_reservoir_ms;       /* reservoir steps (assembly, setup, solve) */
    double t_conduit_ms;         /* conduit passes, exchange terms and convergence tests */
    double t_total_ms;
} hmfem_coupled_status;

/* The coupled reservoir / conduit model of main.c:395-530 on the device:
   a reservoir step without well and conduits (main.c:395), then the well
   (WI, pw: device arrays of n; NULL for none) and the conduits' karst index,
   a conduit pass, and the fixed-point loop (reservoir step warm-started from
   the previous pressure with the conduit pressures in its source term, then a
   conduit pass with the new reservoir pressure) until the reservoir pressure
   changes by at most tol between passes (main.c:457-514 with tol = 1e-8;
   the reference's max search over the per-conduit changes keeps the
   reservoir's) or max_passes passes; then fluxo1d() and fluxos(). p (device,
   n) is the first step's initial guess (main.c: zero) and the final
   pressure; pc, ucl, ucr (device, sum of cells: conduit j's cells follow
   conduit j-1's) receive the conduit pressure and left/right face fluxes;
   v (device ELE[n] or NULL) the reservoir fluxes; pc_first (device or NULL)
   the conduit pressure of the pass before the loop; pass_change (host,
   max_passes, or NULL) each pass's change. The conduit systems are solved
   exactly (parallel cyclic reduction), where the reference's compute() stops
   its AMG-PCG at 1e-8. opt applies to every reservoir step (the reference's
   stop rule, HMFEM_STOP_PRECONDITIONED, keeps the later warm-started passes
   iterating; reuse_setup = 1 reuses the hierarchy, exact after the first
   pass since the operator no longer changes); opt NULL: the defaults with
   those two. */
int hmfem_coupled_device(hmfem_context *ctx, const hmfem_grid *g, const double *Kx, const double *Ky,
                         const double *Kz, const double *const bval[6], const double *WI, const double *pw,
                         const hmfem_conduit_set *cs, int max_passes, double tol, double *p, double *pc,
                         double *ucl, double *ucr, ELE *v, double *pc_first, double *pass_change,
                         const hmfem_options *opt, hmfem_coupled_status *st);

/* Per-kernel timing: with profiling enabled, every fine-level and CG launch
   of the following calls is bracketed by HIP events on the context's stream
   (capacity: max_launches launches per call). hmfem_profile_read returns, per
   kernel class, the summed milliseconds, the launch count and the algorithmic
   HBM bytes of one launch. Classes: 0 CG SpMV+dot, 1 CG update+dot, 2 CG
   dots, 3 pre-smoothing, 4 residual+restriction, 5 post-smoothing (block GS:
   prolongation + all sweeps + CG dots; Jacobi: prolongation + one sweep),
   6 further Jacobi sweeps, 7 coarse levels. Launches that returned at once
   because the solve had already converged (shorter than a quarter of their
   class's median launch) are not counted. Returns the number of classes. */
int hmfem_profile_enable(hmfem_context *ctx, int max_launches);
int hmfem_profile_read(hmfem_context *ctx, double *ms, long *count, double *bytes, int nclasses);
void hmfem_profile_reset(hmfem_context *ctx);

/* Algorithmic HBM bytes per cell of the main kernels at a precision, for the
   roofline record: which = 0 CG SpMV+dot (with the previous iteration's x update, as on
   grids whose xy plane has at most 1024 tiles of 32x8 cells), 1 fused Jacobi sweep,
   2 residual+restriction. */
double hmfem_kernel_bytes_per_cell(int which, int precision);

#ifdef __cplusplus
}
#endif

#endif
