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
 */
#include "hmfem_gpu.h"
#include "hmfem_fields.h"

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
        if (!ok) { fprintf(stderr, "%s:%d: cannot read `%s`\n", path, ln, key); fclose(f); return 1; }
    }
    fclose(f);
    if (c->nx < 1 || c->ny < 1 || c->nz < 1 || c->passes < 1) { fprintf(stderr, "%s: bad sizes\n", path); return 1; }
    return 0;
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
