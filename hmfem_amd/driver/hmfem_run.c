/*
 * hmfem_run: a C host for the reservoir pressure step, the part of main.c
 * (main.c:123-514) that this library replaces, driven by a small text file
 * instead of main.c's hard-coded constants and argv.
 *
 *   hmfem_run <config> <output dir>
 *
 * It builds the grid and a seeded permeability field, then runs `passes`
 * reservoir() steps as reservoir.c:8-57 does them (permeabilidade_arestas,
 * create_boundary/set_boundary equivalents, source_term, solve_linsys with
 * the previous pass's pressure as the initial guess, fluxos,
 * divvel_pontual), printing the PCG loop count of each pass, and writes
 * pressure_reservoir.dat, flux_reservoir.dat and campos.vts into the output
 * directory (main.c:529-530, :725). Between passes the wells' productivity
 * index KI is scaled by `ki_growth` so that successive passes solve slightly
 * different systems, as the reference's fixed-point passes do; with
 * HMFEM_REUSE_SETUP=1 the later passes reuse the first pass's hierarchy.
 *
 * Config keys (one per line, `#` starts a comment; defaults in brackets):
 *   grid nx ny nz            [16 16 16]
 *   lengths Ix Iy Iz         [1 1 1]
 *   viscosity mu             [1]
 *   bc_type l r u d t b      [0 0 1 1 1 1]   0 Dirichlet, 1 Neumann
 *   bc_value l r u d t b     [1 0 0 0 0 0]   uniform value per face
 *   field homogeneous K      [homogeneous 1e-14]
 *   field lognormal mean_ln sigma corr seed
 *   conduits count seed      [0]
 *   well i j k KI pc         [none]          one source cell: KI and its pressure pc (WI = 0)
 *   ki_growth g              [1]
 *   passes n                 [1]
 *
 * With `karst_data` it runs main.c's coupled model instead (main.c:395-530):
 * the conduit set of a file in the reference's data_<nk>_<type>.dat format
 * (read as main.c:197-300 reads it), the well column, and the fixed-point
 * loop of reservoir steps and conduit passes on the device
 * (hmfem_coupled_device) until the reservoir pressure changes by at most
 * `coupling_tol` between passes; it prints main.c's closing lines
 * ("iterations number", "error") and writes, besides the three files above,
 * pressure_karst_<j>.dat and flux_karst_<j>.dat (main.c:526-527).
 *   karst_data count path    [none]          number_karst and the data file
 *   well_column i j k0 WI pw [none]          main.c:316-326, 398-402: pw on (i, j, k >= k0),
 *                                            WI from the second reservoir step on
 *   coupling_tol t           [1e-8]          main.c:127
 *   max_passes n             [1000]          main.c:128
 */
#include "hmfem_gpu.h"
#include "hmfem_fields.h"

#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nx, ny, nz;
    double I[3], mu;
    int bc_type[6];
    double bc_val[6];
    int lognormal;
    double k_hom, mean_ln, sigma, corr;
    unsigned long long seed;
    int conduits;
    unsigned long long conduit_seed;
    int well, wi, wj, wk;
    double wki, wpc, ki_growth;
    int passes;
    int karst;
    char karst_path[400];
    int column, ci, cj, ck0;
    double cwi, cpw, coupling_tol;
    int max_passes;
} Config;

static void defaults(Config *c)
{
    memset(c, 0, sizeof(*c));
    c->nx = c->ny = c->nz = 16;
    c->I[0] = c->I[1] = c->I[2] = 1.0;
    c->mu = 1.0;
    c->bc_type[2] = c->bc_type[3] = c->bc_type[4] = c->bc_type[5] = 1;
    c->bc_val[0] = 1.0;
    c->k_hom = 1e-14;
    c->ki_growth = 1.0;
    c->passes = 1;
    c->coupling_tol = 1e-8;
    c->max_passes = 1000;
}

static int read_config(const char *path, Config *c)
{
    FILE *f = fopen(path, "r");
    if (!f) { perror(path); return 1; }
    char line[512], key[64];
    int ln = 0;
    while (fgets(line, sizeof(line), f)) {
        ln++;
        char *h = strchr(line, '#');
        if (h) *h = 0;
        if (sscanf(line, "%63s", key) != 1) continue;
        const char *a = strstr(line, key) + strlen(key);
        int ok = 0;
        if (!strcmp(key, "grid")) ok = sscanf(a, "%d %d %d", &c->nx, &c->ny, &c->nz) == 3;
        else if (!strcmp(key, "lengths")) ok = sscanf(a, "%lf %lf %lf", &c->I[0], &c->I[1], &c->I[2]) == 3;
        else if (!strcmp(key, "viscosity")) ok = sscanf(a, "%lf", &c->mu) == 1;
        else if (!strcmp(key, "bc_type"))
            ok = sscanf(a, "%d %d %d %d %d %d", &c->bc_type[0], &c->bc_type[1], &c->bc_type[2], &c->bc_type[3],
                        &c->bc_type[4], &c->bc_type[5]) == 6;
        else if (!strcmp(key, "bc_value"))
            ok = sscanf(a, "%lf %lf %lf %lf %lf %lf", &c->bc_val[0], &c->bc_val[1], &c->bc_val[2], &c->bc_val[3],
                        &c->bc_val[4], &c->bc_val[5]) == 6;
        else if (!strcmp(key, "field")) {
            char kind[32];
            if (sscanf(a, "%31s", kind) == 1) {
                const char *b = strstr(a, kind) + strlen(kind);
                if (!strcmp(kind, "homogeneous")) { c->lognormal = 0; ok = sscanf(b, "%lf", &c->k_hom) == 1; }
                else if (!strcmp(kind, "lognormal")) {
                    c->lognormal = 1;
                    ok = sscanf(b, "%lf %lf %lf %llu", &c->mean_ln, &c->sigma, &c->corr, &c->seed) == 4;
                }
            }
        } else if (!strcmp(key, "conduits")) ok = sscanf(a, "%d %llu", &c->conduits, &c->conduit_seed) == 2;
        else if (!strcmp(key, "well")) {
            ok = sscanf(a, "%d %d %d %lf %lf", &c->wi, &c->wj, &c->wk, &c->wki, &c->wpc) == 5;
            c->well = ok;
        } else if (!strcmp(key, "ki_growth")) ok = sscanf(a, "%lf", &c->ki_growth) == 1;
        else if (!strcmp(key, "passes")) ok = sscanf(a, "%d", &c->passes) == 1;
        else if (!strcmp(key, "karst_data")) ok = sscanf(a, "%d %399s", &c->karst, c->karst_path) == 2 && c->karst > 0;
        else if (!strcmp(key, "well_column")) {
            ok = sscanf(a, "%d %d %d %lf %lf", &c->ci, &c->cj, &c->ck0, &c->cwi, &c->cpw) == 5;
            c->column = ok;
        } else if (!strcmp(key, "coupling_tol")) ok = sscanf(a, "%lf", &c->coupling_tol) == 1;
        else if (!strcmp(key, "max_passes")) ok = sscanf(a, "%d", &c->max_passes) == 1;
        if (!ok) { fprintf(stderr, "%s:%d: cannot read `%s`\n", path, ln, key); fclose(f); return 1; }
    }
    fclose(f);
    if (c->nx < 1 || c->ny < 1 || c->nz < 1 || c->passes < 1) { fprintf(stderr, "%s: bad sizes\n", path); return 1; }
    return 0;
}

/* A conduit set read as main.c:197-300 reads data_<nk>_<type>.dat. */
typedef struct {
    int nk, mb;
    int *cells, *start, *axis, *bc_l, *bc_r, *nbrec, *brec;
    double *Kc, *cont, *ccont, *ki, *area, *beta, *C;
} Karst;

static int read_karst(const char *path, int nk, Karst *k)
{
    FILE *f = fopen(path, "r");
    if (!f) { perror(path); return 1; }
    memset(k, 0, sizeof(*k));
    k->nk = nk;
    k->cells = calloc(nk, sizeof(int)); k->start = calloc(3 * nk, sizeof(int)); k->axis = calloc(nk, sizeof(int));
    k->bc_l = calloc(nk, sizeof(int)); k->bc_r = calloc(nk, sizeof(int)); k->nbrec = calloc(nk, sizeof(int));
    k->Kc = calloc(nk, sizeof(double)); k->cont = calloc(2 * nk, sizeof(double)); k->ccont = calloc(2 * nk, sizeof(double));
    k->ki = calloc(nk, sizeof(double)); k->area = calloc(nk, sizeof(double));
    int ok = 1, i, j;
    for (i = 0; i < nk; i++) ok &= fscanf(f, "%d", &k->cells[i]) == 1;
    for (j = 0; j < 3; j++)
        for (i = 0; i < nk; i++) ok &= fscanf(f, "%d", &k->start[3 * i + j]) == 1;
    for (i = 0; i < nk; i++) ok &= fscanf(f, "%d", &k->axis[i]) == 1;
    for (i = 0; i < nk; i++) ok &= fscanf(f, "%lf", &k->Kc[i]) == 1;
    for (i = 0; i < nk; i++) ok &= fscanf(f, "%d", &k->bc_l[i]) == 1;
    for (i = 0; i < nk; i++) ok &= fscanf(f, "%d", &k->bc_r[i]) == 1;
    for (i = 0; i < 2 * nk; i++) ok &= fscanf(f, "%lf", &k->cont[i]) == 1;
    for (i = 0; i < 2 * nk; i++) ok &= fscanf(f, "%lf", &k->ccont[i]) == 1;
    for (i = 0; i < nk; i++) ok &= fscanf(f, "%lf", &k->ki[i]) == 1;
    for (i = 0; i < nk; i++) ok &= fscanf(f, "%lf", &k->area[i]) == 1;
    for (i = 0; i < nk; i++) ok &= fscanf(f, "%d", &k->nbrec[i]) == 1;
    k->mb = 0;
    for (i = 0; i < nk; i++) if (k->nbrec[i] > k->mb) k->mb = k->nbrec[i];
    const int nb = k->mb * nk + 1;
    k->beta = calloc(nb, sizeof(double)); k->C = calloc(nb, sizeof(double)); k->brec = calloc(nb, sizeof(int));
    for (i = 0; i < k->mb * nk; i++) {
        ok &= fscanf(f, "%lf", &k->beta[i]) == 1;
        k->beta[i] = k->beta[i] - 1;   /* main.c:262: beta = R^4 - 1 */
    }
    for (i = 0; i < k->mb * nk; i++) ok &= fscanf(f, "%lf", &k->C[i]) == 1;
    for (i = 0; i < k->mb * nk; i++) ok &= fscanf(f, "%d", &k->brec[i]) == 1;
    fclose(f);
    if (!ok) fprintf(stderr, "%s: short conduit data file\n", path);
    return !ok;
}

#define HIPX(call) do { if ((call) != hipSuccess) { fprintf(stderr, "%s failed\n", #call); return 1; } } while (0)

static double *dev_copy(const double *h, size_t n)
{
    double *d = NULL;
    if (hipMalloc((void **)&d, n * sizeof(double) + 16) != hipSuccess) return NULL;
    if (h && hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return NULL;
    if (!h && hipMemset(d, 0, n * sizeof(double)) != hipSuccess) return NULL;
    return d;
}

/* main.c:395-530 through hmfem_coupled_device */
static int run_coupled(const Config *cf, Dim N, grid h, const double *K, double *const fv[6])
{
    Karst k;
    if (read_karst(cf->karst_path, cf->karst, &k)) return 2;
    const int n = N.t, nk = k.nk;
    int total = 0, mmax = 0;
    for (int j = 0; j < nk; j++) { total += k.cells[j]; if (k.cells[j] > mmax) mmax = k.cells[j]; }
    double *WI = calloc(n, sizeof(double)), *pw = calloc(n, sizeof(double));
    if (cf->column)
        for (int z = cf->ck0; z < N.z; z++) {
            const int s = cf->ci + cf->cj * N.x + z * N.xy;
            if (s < 0 || s >= n) { fprintf(stderr, "well column outside the grid\n"); return 2; }
            WI[s] = cf->cwi;
            pw[s] = cf->cpw;
        }
    hmfem_context *ctx;
    if (hmfem_create(&ctx) != HMFEM_OK) { fprintf(stderr, "no HIP device\n"); return 1; }
    const int fsz[6] = {N.y * N.z, N.y * N.z, N.x * N.z, N.x * N.z, N.x * N.y, N.x * N.y};
    const double *bv[6];
    for (int q = 0; q < 6; q++) bv[q] = dev_copy(fv[q], fsz[q]);
    double *dK = dev_copy(K, n), *dWI = dev_copy(WI, n), *dpw = dev_copy(pw, n), *dp = dev_copy(NULL, n);
    double *dpc = dev_copy(NULL, total), *dl = dev_copy(NULL, total), *dr = dev_copy(NULL, total);
    ELE *dv = NULL;
    HIPX(hipMalloc((void **)&dv, sizeof(ELE) * n));
    if (!dK || !dWI || !dpw || !dp || !dpc || !dl || !dr) { fprintf(stderr, "out of device memory\n"); return 1; }
    hmfem_grid g = {N.x, N.y, N.z, h.x, h.y, h.z, cf->mu, {0}};
    for (int q = 0; q < 6; q++) g.bc_type[q] = cf->bc_type[q];
    hmfem_conduit_set cs = {nk, k.cells, k.start, k.axis, k.Kc, k.bc_l, k.bc_r, k.cont, k.ccont, k.ki, k.area,
                            k.nbrec, k.mb, k.brec, k.beta, k.C, 1000.0};
    hmfem_options opt;
    hmfem_default_options(&opt);
    opt.stop_rule = HMFEM_STOP_PRECONDITIONED;   /* the reference's rule: warm passes keep iterating */
    opt.reuse_setup = 1;
    hmfem_coupled_status st;
    int rc = hmfem_coupled_device(ctx, &g, dK, dK, dK, bv, dWI, dpw, &cs, cf->max_passes, cf->coupling_tol, dp, dpc,
                                  dl, dr, dv, NULL, NULL, &opt, &st);
    if (rc < 0) { fprintf(stderr, "hmfem_coupled_device failed (%d)\n", rc); return 1; }
    printf("iterations number=%d\n", st.passes);   /* main.c:517-518 */
    printf("error=%e\n", st.change);
    printf("gpu time=%f ms (reservoir steps %f ms, CG iterations %d)\n", st.t_total_ms, st.t_reservoir_ms,
           st.iterations);
    fflush(stdout);
    double *p = malloc(sizeof(double) * n), *pc = malloc(sizeof(double) * total);
    double *ul = malloc(sizeof(double) * total), *ur = malloc(sizeof(double) * total);
    ELE *v = malloc(sizeof(ELE) * n);
    HIPX(hipMemcpy(p, dp, sizeof(double) * n, hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(pc, dpc, sizeof(double) * total, hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(ul, dl, sizeof(double) * total, hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(ur, dr, sizeof(double) * total, hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(v, dv, sizeof(ELE) * n, hipMemcpyDeviceToHost));
    /* the writers take the reference's [cell][conduit] matrices */
    double **mp = malloc(sizeof(double *) * mmax), **ml = malloc(sizeof(double *) * mmax),
           **mr = malloc(sizeof(double *) * mmax);
    for (int i = 0; i < mmax; i++) {
        mp[i] = calloc(nk, sizeof(double)); ml[i] = calloc(nk, sizeof(double)); mr[i] = calloc(nk, sizeof(double));
    }
    DimK *M = calloc(nk, sizeof(DimK));
    for (int j = 0, off = 0; j < nk; off += k.cells[j], j++) {
        M[j].s = k.cells[j]; M[j].sxi = k.start[3 * j]; M[j].syi = k.start[3 * j + 1];
        M[j].szi = k.start[3 * j + 2]; M[j].pos = k.axis[j];
        for (int i = 0; i < k.cells[j]; i++) { mp[i][j] = pc[off + i]; ml[i][j] = ul[off + i]; mr[i][j] = ur[off + i]; }
    }
    number_karst = nk;
    ELEK uc = {ml, mr, NULL};
    double *hs = calloc(nk, sizeof(double));
    save_pressure_1d(mp, M, hs, h);
    save_flux_1d(&uc, M, hs, h);
    save_pressure(p, N, h);
    save_flux(v, N, h);
    if (vts_StructuredGrid_file(v, p, N, h) != 0) return 1;
    hmfem_destroy(ctx);
    return st.status == HMFEM_OK ? 0 : 3;
}

int main(int argc, char **argv)
{
    if (argc < 3) { fprintf(stderr, "usage: %s <config> <output dir>\n", argv[0]); return 2; }
    Config cf;
    defaults(&cf);
    if (read_config(argv[1], &cf)) return 2;
    hmfem_set_output_dir(argv[2]);

    Dim N;
    memset(&N, 0, sizeof(N));
    N.x = cf.nx; N.y = cf.ny; N.z = cf.nz; N.xy = cf.nx * cf.ny; N.t = N.xy * cf.nz;
    grid h = {cf.I[0] / cf.nx, cf.I[1] / cf.ny, cf.I[2] / cf.nz};
    const int n = N.t;
    double *K = malloc(sizeof(double) * n);
    if (cf.lognormal) hmfem_field_lognormal(cf.nx, cf.ny, cf.nz, cf.mean_ln, cf.sigma, cf.corr, cf.seed, K);
    else hmfem_field_homogeneous(cf.nx, cf.ny, cf.nz, cf.k_hom, K);
    if (cf.conduits > 0) hmfem_field_conduits(cf.nx, cf.ny, cf.nz, cf.conduits, cf.conduit_seed, K);
    Perm P;
    memset(&P, 0, sizeof(P));
    P.X = P.Y = P.Z = K;
    /* boundary types and the face value arrays (set_boundaries.c:8-125: sized by face) */
    FACE ft = {cf.bc_type[2], cf.bc_type[3], cf.bc_type[0], cf.bc_type[1], cf.bc_type[4], cf.bc_type[5]};
    boundary = ft;
    const int fsz[6] = {N.y * N.z, N.y * N.z, N.x * N.z, N.x * N.z, N.x * N.y, N.x * N.y};   /* l r u d t b */
    double *fv[6];
    for (int q = 0; q < 6; q++) {
        fv[q] = malloc(sizeof(double) * fsz[q]);
        for (int s = 0; s < fsz[q]; s++) fv[q][s] = cf.bc_val[q];
    }
    if (cf.karst) return run_coupled(&cf, N, h, K, fv);
    Bound B = {fv[2], fv[3], fv[0], fv[1], fv[4], fv[5]};   /* u, d, l, r, t, b */
    double *KI = calloc(n, sizeof(double)), *WI = calloc(n, sizeof(double));
    double *pc = calloc(n, sizeof(double)), *pw = calloc(n, sizeof(double));
    if (cf.well) {
        const int s = cf.wi + cf.wj * N.x + cf.wk * N.xy;
        if (s < 0 || s >= n) { fprintf(stderr, "well outside the grid\n"); return 2; }
        KI[s] = cf.wki;
        pc[s] = cf.wpc;
    }
    ELE *perm = calloc(n, sizeof(ELE)), *v = calloc(n, sizeof(ELE));
    IND *indcol = calloc(n, sizeof(IND));
    double *b = malloc(sizeof(double) * n), *p = calloc(n, sizeof(double));

    for (int pass = 0; pass < cf.passes; pass++) {
        /* reservoir.c:8-57 */
        permeabilidade_arestas(&P, perm, indcol, h, N, boundary, WI, KI, cf.mu);
        memset(b, 0, sizeof(double) * n);
        source_term(&B, perm, h, N, boundary, pc, pw, KI, WI, cf.mu, b);
        solve_linsys(perm, indcol, b, p, N);
        const int its = iter_amg;
        fluxos(p, perm, v, &B, h, N);
        printf("pass %d: iter_amg %d\n", pass + 1, its);
        fflush(stdout);
        for (int s = 0; s < n; s++) KI[s] *= cf.ki_growth;
    }
    save_pressure(p, N, h);
    save_flux(v, N, h);
    if (vts_StructuredGrid_file(v, p, N, h) != 0) return 1;
    return 0;
}
