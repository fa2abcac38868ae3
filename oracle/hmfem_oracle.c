/*
 * CPU restatement of the reference's pressure-solve path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests, the smoke check and the
 * benchmark's CPU-baseline leg as the checker; the product never calls it.
 *
 * Parity pins (see tests/test_oracle.py):
 *   - assembly, RHS, count-CSR, AMG-PCG pressure, iteration counts and fluxes
 *     are compared bit for bit with the reference build (oracle/_ref) through
 *     the committed fixtures in tests/golden/;
 *   - closed forms: the 16^3 homogeneous case has p_i = 1 - (i+1/2)/16 and
 *     uniform flux K*lambda; hm() identities; Galerkin A_c = P^T A P by brute
 *     force; P row sums = 1.
 * Every function is pinned by one of these; none is "parity unpinned".
 *
 * Operation order follows the reference so that a build without floating
 * point contraction reproduces the reference's bits.
 */
#include "hmfem_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PIDX(g, i, j, k) ((i) + (j) * (g)->nx + (k) * (g)->nx * (g)->ny)

/* hm (permeabilities.c:188-193) */
static double hm(double a, double b) { return 2 * a * b / (a + b); }

void oracle_permeabilidade_arestas(const oracle_grid *g, const double *Kx, const double *Ky, const double *Kz,
                                   const double *WI, const double *KI, oracle_ele *perm, oracle_ind *indcol)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, nxy = nx * ny;
    const double hx = g->hx, hy = g->hy, hz = g->hz, mu = g->visc;
    int i, j, k, s;
    /* interior faces (permeabilities.c:216-293) */
    for (k = 0; k < nz; k++)
        for (j = 0; j < ny; j++)
            for (i = 0; i < nx; i++) {
                s = PIDX(g, i, j, k);
                if (j < ny - 1) { perm[s].u = hm(Ky[s], Ky[s + nx]); perm[s].u = perm[s].u / (hy * hy * mu); indcol[s].u = s + nx; }
                if (i < nx - 1) { perm[s].r = hm(Kx[s], Kx[s + 1]); perm[s].r = perm[s].r / (hx * hx * mu); indcol[s].r = s + 1; }
                if (j > 0) { perm[s].d = hm(Ky[s], Ky[s - nx]); perm[s].d = perm[s].d / (hy * hy * mu); indcol[s].d = s - nx; }
                if (i > 0) { perm[s].l = hm(Kx[s], Kx[s - 1]); perm[s].l = perm[s].l / (hx * hx * mu); indcol[s].l = s - 1; }
                if (k < nz - 1) { perm[s].t = hm(Kz[s], Kz[s + nxy]); perm[s].t = perm[s].t / (hz * hz * mu); indcol[s].t = s + nxy; }
                if (k > 0) { perm[s].b = hm(Kz[s], Kz[s - nxy]); perm[s].b = perm[s].b / (hz * hz * mu); indcol[s].b = s - nxy; }
            }
    /* boundary faces (permeabilities.c:297-477): Dirichlet 2K/(h^2 mu), Neumann 0 */
    const int *bc = g->bc_type;
    for (k = 0; k < nz; k++)
        for (j = 0; j < ny; j++) {
            s = PIDX(g, 0, j, k);
            if (bc[0] == 0) perm[s].l = 2.0 * Kx[s] / (hx * hx * mu);
            if (bc[0] == 1) perm[s].l = 0;
            if (bc[0] == 0 || bc[0] == 1) indcol[s].l = -1;
        }
    if (bc[0] == 1 && bc[1] == 1 && bc[2] == 1 && bc[3] == 1 && bc[4] == 1 && bc[5] == 1) {
        s = PIDX(g, 0, ny - 1, nz / 2); /* all-Neumann pin (permeabilities.c:328-333) */
        perm[s].l = 2.0 * Kx[s] / (hx * hx * mu);
        indcol[s].l = -1;
    }
    for (k = 0; k < nz; k++)
        for (j = 0; j < ny; j++) {
            s = PIDX(g, nx - 1, j, k);
            if (bc[1] == 0) perm[s].r = 2.0 * Kx[s] / (hx * hx * mu);
            if (bc[1] == 1) perm[s].r = 0;
            if (bc[1] == 0 || bc[1] == 1) indcol[s].r = -1;
        }
    for (k = 0; k < nz; k++)
        for (i = 0; i < nx; i++) {
            s = PIDX(g, i, 0, k);
            if (bc[3] == 0) perm[s].d = 2.0 * Ky[s] / (hy * hy * mu);
            if (bc[3] == 1) perm[s].d = 0;
            if (bc[3] == 0 || bc[3] == 1) indcol[s].d = -1;
        }
    for (k = 0; k < nz; k++)
        for (i = 0; i < nx; i++) {
            s = PIDX(g, i, ny - 1, k);
            if (bc[2] == 0) perm[s].u = 2.0 * Ky[s] / (hy * hy * mu);
            if (bc[2] == 1) perm[s].u = 0;
            if (bc[2] == 0 || bc[2] == 1) indcol[s].u = -1;
        }
    for (j = 0; j < ny; j++)
        for (i = 0; i < nx; i++) {
            s = PIDX(g, i, j, nz - 1);
            if (bc[4] == 0) perm[s].t = 2.0 * Kz[s] / (hz * hz * mu);
            if (bc[4] == 1) perm[s].t = 0;
            if (bc[4] == 0 || bc[4] == 1) indcol[s].t = -1;
        }
    for (j = 0; j < ny; j++)
        for (i = 0; i < nx; i++) {
            s = PIDX(g, i, j, 0);
            if (bc[5] == 0) perm[s].b = 2.0 * Kz[s] / (hz * hz * mu);
            if (bc[5] == 1) perm[s].b = 0;
            if (bc[5] == 0 || bc[5] == 1) indcol[s].b = -1;
        }
    /* centre (permeabilities.c:481-489) */
    for (s = 0; s < nxy * nz; s++) {
        perm[s].c = perm[s].l + perm[s].r + perm[s].u + perm[s].d + perm[s].t + perm[s].b
                  + KI[s] / (mu * hx * hy * hz) + WI[s] / (mu * hx * hy * hz);
        indcol[s].c = s;
    }
}

void oracle_source_term(const oracle_grid *g, const oracle_bound *B, const oracle_ele *perm,
                        const double *pc, const double *pw, const double *KI, const double *WI, double *b)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int *bc = g->bc_type;
    int i, j, k, s, l;
    /* case 1 LEFT (set_source_term.c:28-62) */
    for (j = 0; j < ny; j++)
        for (k = 0; k < nz; k++) {
            s = PIDX(g, 0, j, k); l = k * ny + j;
            if (bc[0] == 0) b[s] = b[s] + (perm[s].l) * B->l[l];
            if (bc[0] == 1) b[s] = b[s] + B->l[l] / g->hx;
        }
    if (bc[0] == 1 && bc[1] == 1 && bc[2] == 1 && bc[3] == 1 && bc[4] == 1 && bc[5] == 1) {
        s = PIDX(g, 0, ny - 1, nz / 2);
        b[s] = b[s] + (perm[s].l) * 1.0;
    }
    /* case 2 RIGHT (:66-90) */
    for (j = 0; j < ny; j++)
        for (k = 0; k < nz; k++) {
            s = PIDX(g, nx - 1, j, k); l = k * ny + j;
            if (bc[1] == 0) b[s] = b[s] + perm[s].r * B->r[l];
            if (bc[1] == 1) b[s] = b[s] - B->r[l] / g->hx;
        }
    /* case 3 DOWN (:94-118) */
    for (i = 0; i < nx; i++)
        for (k = 0; k < nz; k++) {
            s = PIDX(g, i, 0, k); l = k * nx + i;
            if (bc[3] == 0) b[s] = b[s] + perm[s].d * B->d[l];
            if (bc[3] == 1) b[s] = b[s] + B->d[l] / g->hy;
        }
    /* case 4 UPPER (:122-146) */
    for (i = 0; i < nx; i++)
        for (k = 0; k < nz; k++) {
            s = PIDX(g, i, ny - 1, k); l = k * nx + i;
            if (bc[2] == 0) b[s] = b[s] + perm[s].u * B->u[l];
            if (bc[2] == 1) b[s] = b[s] - B->u[l] / g->hy;
        }
    /* case 5 TOP (:150-174) */
    for (i = 0; i < nx; i++)
        for (j = 0; j < ny; j++) {
            s = PIDX(g, i, j, nz - 1); l = j * nx + i;
            if (bc[4] == 0) b[s] = b[s] + perm[s].t * B->t[l];
            if (bc[4] == 1) b[s] = b[s] - B->t[l] / g->hz;
        }
    /* case 6 BOTTOM (:178-202) */
    for (i = 0; i < nx; i++)
        for (j = 0; j < ny; j++) {
            s = PIDX(g, i, j, 0); l = j * nx + i;
            if (bc[5] == 0) b[s] = b[s] + perm[s].b * B->b[l];
            if (bc[5] == 1) b[s] = b[s] + B->b[l] / g->hz;
        }
    /* KI / WI sources (:210-220) */
    double Ve = g->hx * g->hy * g->hz;
    for (s = 0; s < nx * ny * nz; s++)
        b[s] = b[s] + (KI[s] * pc[s]) / (g->visc * Ve) + (WI[s] * pw[s]) / (g->visc * Ve);
}

long oracle_build_csr(int n, const oracle_ele *perm, const oracle_ind *indcol, int *cnt, int *col, double *ele)
{
    long l = 0;
    for (int s = 0; s < n; s++) {
        int c = 0;
        if (indcol[s].d != -1) { col[l] = indcol[s].d; ele[l] = -perm[s].d; l++; c++; }
        if (indcol[s].l != -1) { col[l] = indcol[s].l; ele[l] = -perm[s].l; l++; c++; }
        if (indcol[s].c != -1) { col[l] = indcol[s].c; ele[l] = perm[s].c; l++; c++; }
        if (indcol[s].r != -1) { col[l] = indcol[s].r; ele[l] = -perm[s].r; l++; c++; }
        if (indcol[s].u != -1) { col[l] = indcol[s].u; ele[l] = -perm[s].u; l++; c++; }
        if (indcol[s].t != -1) { col[l] = indcol[s].t; ele[l] = -perm[s].t; l++; c++; }
        if (indcol[s].b != -1) { col[l] = indcol[s].b; ele[l] = -perm[s].b; l++; c++; }
        cnt[s] = c;
    }
    return l;
}

void oracle_prod_mat_vec(int n, const oracle_ele *perm, const oracle_ind *indcol, const double *x, double *res)
{
    for (int s = 0; s < n; s++) {
        if (indcol[s].d != -1) res[s] = res[s] - perm[s].d * x[indcol[s].d];
        if (indcol[s].l != -1) res[s] = res[s] - perm[s].l * x[indcol[s].l];
        if (indcol[s].c != -1) res[s] = res[s] + perm[s].c * x[indcol[s].c];
        if (indcol[s].r != -1) res[s] = res[s] - perm[s].r * x[indcol[s].r];
        if (indcol[s].u != -1) res[s] = res[s] - perm[s].u * x[indcol[s].u];
        if (indcol[s].t != -1) res[s] = res[s] - perm[s].t * x[indcol[s].t];
        if (indcol[s].b != -1) res[s] = res[s] - perm[s].b * x[indcol[s].b];
    }
}

/* ------------------------------------------------------------------ AMG */

typedef struct {
    int n;
    long nnz;
    int *cnt, *dsp, *col;
    double *ele;
    double *dia;          /* inverted diagonal */
    /* interpolation to the next level (rows = n) */
    int *pcnt, *pcol;
    double *pele;
    double *u, *f, *s, *r;
} olevel;

struct oracle_amg {
    int nlev;
    olevel lev[32];
};

static void *xmalloc(size_t b)
{
    void *p = malloc(b ? b : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory (%zu bytes)\n", b); abort(); }
    return p;
}

static void level_alloc_vectors(olevel *L)
{
    L->u = (double *)calloc(L->n + 1, sizeof(double));
    L->f = (double *)calloc(L->n + 1, sizeof(double));
    L->s = (double *)calloc(L->n + 1, sizeof(double));
    L->r = (double *)calloc(L->n + 1, sizeof(double));
}

static void level_offsets(olevel *L)
{
    L->dsp = (int *)xmalloc(sizeof(int) * (L->n + 1));
    long t = 0;
    for (int i = 0; i < L->n; i++) { L->dsp[i] = (int)t; t += L->cnt[i]; }
    L->dsp[L->n] = (int)t;
}

/* sparse_multiply_transpose, live branch (amg.h:480-563): C = (A B)^T with
   C rows = bcols, columns = A rows in increasing order. */
static void spmt(int arows, const int *acnt, const int *acol, const double *aele,
                 int brows, const int *bcnt, const int *bcol, const double *bele, int bcols,
                 int **ccnt_o, int **ccol_o, double **cele_o, long *csize_o)
{
    int *bdsp = (int *)xmalloc(sizeof(int) * (brows + 1));
    long t = 0;
    for (int i = 0; i < brows; i++) { bdsp[i] = (int)t; t += bcnt[i]; }
    int *ccnt = (int *)calloc(bcols + 1, sizeof(int));
    int *clst = (int *)xmalloc(sizeof(int) * (bcols + 1));
    for (int i = 0; i < bcols; i++) clst[i] = -1;
    long csize = 0, m = 0;
    for (int i = 0; i < arows; i++)
        for (int j = 0; j < acnt[i]; j++) {
            int c = acol[m++];
            for (int k = bdsp[c]; k < bdsp[c] + bcnt[c]; k++) {
                int d = bcol[k];
                if (clst[d] != i) { clst[d] = i; ccnt[d]++; csize++; }
            }
        }
    for (int i = 0; i < bcols; i++) clst[i] = -1;
    int *cdsp = (int *)xmalloc(sizeof(int) * (bcols + 1));
    t = 0;
    for (int i = 0; i < bcols; i++) { cdsp[i] = (int)t; t += ccnt[i]; }
    int *ccol = (int *)xmalloc(sizeof(int) * (csize + 1));
    double *cele = (double *)xmalloc(sizeof(double) * (csize + 1));
    m = 0;
    for (int i = 0; i < arows; i++)
        for (int j = 0; j < acnt[i]; j++) {
            int c = acol[m];
            double s = aele[m];
            m++;
            for (int k = bdsp[c]; k < bdsp[c] + bcnt[c]; k++) {
                int d = bcol[k];
                int e = cdsp[d];
                double tt = s * bele[k];
                if (clst[d] != i) { clst[d] = i; ccol[e] = i; cele[e] = tt; cdsp[d] = e + 1; }
                else cele[e - 1] += tt;
            }
        }
    free(bdsp); free(clst); free(cdsp);
    *ccnt_o = ccnt; *ccol_o = ccol; *cele_o = cele; *csize_o = csize;
}

/* One coarsening step: parallel_coarsening (amg.h:34-226, serial path: phases
   (b) and (c) of node_selection, amg.h:1-32), parallel_interpolation
   (amg.h:228-406) and new_triple_product (amg.h:567-582). Fills L->dia with
   the (not yet inverted) diagonal, L->p* and the coarse matrix C. */
static void restrict_level(olevel *L, double epsilon, olevel *C)
{
    const int n = L->n;
    const int *cnt = L->cnt, *dsp = L->dsp, *col = L->col;
    const double *ele = L->ele;
    L->dia = (double *)calloc(n + 1, sizeof(double));
    for (int i = 0; i < n; i++)
        for (int j = dsp[i]; j < dsp[i] + cnt[i]; j++)
            if (col[j] == i) L->dia[i] = ele[j];
    char *strong = (char *)calloc(L->nnz + 1, 1);
    for (int i = 0; i < n; i++) {
        double asda = fabs(epsilon * L->dia[i]);
        for (int j = dsp[i]; j < dsp[i] + cnt[i]; j++)
            if (fabs(ele[j]) >= asda) strong[j] = 1;
    }
    char *sel = (char *)calloc(n + 1, 1);
    /* phase (b): nodes already C mark their strongly-dependent neighbours F */
    /* phase (c): greedy C/F in index order */
    for (int phase = 1; phase >= 0; phase--)
        for (int i = 0; i < n; i++) {
            if (sel[i] != phase) continue;
            if (phase == 0) sel[i] = 1;
            for (int j = dsp[i]; j < dsp[i] + cnt[i]; j++) {
                int c = col[j];
                if (sel[c] == 0)
                    for (int k = dsp[c]; k < dsp[c] + cnt[c]; k++)
                        if (col[k] == i && strong[k] != 0) { sel[c] = 2; break; }
            }
        }
    /* interpolation: C rows weight 1, F rows 1/#(strong C neighbours) */
    int *cidx = (int *)xmalloc(sizeof(int) * (n + 1));
    int nc = 0;
    for (int i = 0; i < n; i++) cidx[i] = sel[i] == 1 ? nc++ : -1;
    L->pcnt = (int *)calloc(n + 1, sizeof(int));
    long psize = 0;
    for (int i = 0; i < n; i++) {
        if (sel[i] == 1) L->pcnt[i] = 1;
        else
            for (int j = dsp[i]; j < dsp[i] + cnt[i]; j++)
                if (sel[col[j]] == 1 && strong[j] != 0) L->pcnt[i]++;
        psize += L->pcnt[i];
    }
    L->pcol = (int *)xmalloc(sizeof(int) * (psize + 1));
    L->pele = (double *)xmalloc(sizeof(double) * (psize + 1));
    long m = 0;
    for (int i = 0; i < n; i++) {
        if (sel[i] == 1) { L->pcol[m] = cidx[i]; L->pele[m] = 1.0; m++; }
        else {
            int mc = L->pcnt[i];
            for (int j = dsp[i]; j < dsp[i] + cnt[i]; j++)
                if (sel[col[j]] == 1 && strong[j] != 0) { L->pcol[m] = cidx[col[j]]; L->pele[m] = 1.0 / mc; m++; }
        }
    }
    /* Galerkin: T = (A P)^T, A_c = (T P)^T */
    int *tcnt, *tcol; double *tele; long tsize;
    spmt(n, cnt, col, ele, n, L->pcnt, L->pcol, L->pele, nc, &tcnt, &tcol, &tele, &tsize);
    int *ccnt, *ccol; double *cele; long csize;
    spmt(nc, tcnt, tcol, tele, n, L->pcnt, L->pcol, L->pele, nc, &ccnt, &ccol, &cele, &csize);
    free(tcnt); free(tcol); free(tele);
    C->n = nc; C->nnz = csize; C->cnt = ccnt; C->col = ccol; C->ele = cele;
    free(strong); free(sel); free(cidx);
}

oracle_amg *oracle_amg_create(int n, const int *cnt, const int *col, const double *ele,
                              int max_level, int min_nodes, double epsilon)
{
    oracle_amg *a = (oracle_amg *)calloc(1, sizeof(oracle_amg));
    if (max_level > 32) max_level = 32;
    olevel *L0 = &a->lev[0];
    L0->n = n;
    long nnz = 0;
    for (int i = 0; i < n; i++) nnz += cnt[i];
    L0->nnz = nnz;
    L0->cnt = (int *)xmalloc(sizeof(int) * (n + 1));
    L0->col = (int *)xmalloc(sizeof(int) * (nnz + 1));
    L0->ele = (double *)xmalloc(sizeof(double) * (nnz + 1));
    memcpy(L0->cnt, cnt, sizeof(int) * n);
    memcpy(L0->col, col, sizeof(int) * nnz);
    memcpy(L0->ele, ele, sizeof(double) * nnz);
    level_offsets(L0);
    level_alloc_vectors(L0);
    /* level loop (amg.h:1188-1215) */
    int l = 0, g = n;
    double rel = 0.0;
    while (l < max_level - 1 && g > min_nodes && rel < 1.0) {
        restrict_level(&a->lev[l], epsilon, &a->lev[l + 1]);
        for (int i = 0; i < a->lev[l].n; i++) a->lev[l].dia[i] = 1.0 / a->lev[l].dia[i];
        l++;
        level_offsets(&a->lev[l]);
        level_alloc_vectors(&a->lev[l]);
        int gold = g;
        g = a->lev[l].n;
        rel = (double)g / (double)gold;
    }
    /* coarsest: inverted diagonal (amg.h:1217-1226) */
    olevel *Lc = &a->lev[l];
    Lc->dia = (double *)calloc(Lc->n + 1, sizeof(double));
    for (long t = 0, i = 0; i < Lc->n; i++)
        for (int j = 0; j < Lc->cnt[i]; j++, t++)
            if (Lc->col[t] == i) Lc->dia[i] = Lc->ele[t];
    for (int i = 0; i < Lc->n; i++) Lc->dia[i] = 1.0 / Lc->dia[i];
    a->nlev = l + 1;
    return a;
}

void oracle_amg_destroy(oracle_amg *a)
{
    if (!a) return;
    for (int l = 0; l < a->nlev; l++) {
        olevel *L = &a->lev[l];
        free(L->cnt); free(L->dsp); free(L->col); free(L->ele); free(L->dia);
        free(L->pcnt); free(L->pcol); free(L->pele);
        free(L->u); free(L->f); free(L->s); free(L->r);
    }
    free(a);
}

int oracle_amg_levels(const oracle_amg *a) { return a->nlev; }

void oracle_amg_level_size(const oracle_amg *a, int l, int *rows, long *nnz)
{
    if (l < 0 || l >= a->nlev) { *rows = 0; *nnz = 0; return; }
    *rows = a->lev[l].n;
    *nnz = a->lev[l].nnz;
}

long oracle_amg_level_matrix(const oracle_amg *a, int l, int *cnt, int *col, double *ele)
{
    if (l < 0 || l >= a->nlev) return -1;
    const olevel *L = &a->lev[l];
    if (cnt) memcpy(cnt, L->cnt, sizeof(int) * L->n);
    if (col) memcpy(col, L->col, sizeof(int) * L->nnz);
    if (ele) memcpy(ele, L->ele, sizeof(double) * L->nnz);
    return L->nnz;
}

long oracle_amg_level_interp(const oracle_amg *a, int l, int *pcnt, int *pcol, double *pele)
{
    if (l < 0 || l >= a->nlev - 1) return -1;
    const olevel *L = &a->lev[l];
    long m = 0;
    for (int i = 0; i < L->n; i++) m += L->pcnt[i];
    if (pcnt) memcpy(pcnt, L->pcnt, sizeof(int) * L->n);
    if (pcol) memcpy(pcol, L->pcol, sizeof(int) * m);
    if (pele) memcpy(pele, L->pele, sizeof(double) * m);
    return m;
}

/* gauss_seidel_f, 3-argument form (amg.h:759-784): forward GS from zero and
   the residual r = f - A u by column updates (needs A symmetric). */
static void gs_forward_residual(olevel *L, const double *f, double *u, double *r)
{
    const int n = L->n;
    const int *col = L->col;
    const double *ele = L->ele;
    for (int i = 0; i < n; i++) r[i] = f[i];
    long m = 0;
    for (int i = 0; i < n; i++) {
        double s = L->dia[i] * r[i];
        u[i] = s;
        for (int j = 0; j < L->cnt[i]; j++) {
            int c = col[m];
            double t = ele[m];
            m++;
            r[c] -= t * s;
        }
    }
}

/* gauss_seidel_b, 3-argument form (amg.h:930-960): u += s, backward GS. */
static void gs_backward(olevel *L, const double *f, double *u, const double *s)
{
    const int n = L->n;
    for (int i = 0; i < n; i++) u[i] += s[i];
    long m = L->nnz;
    for (int i = n - 1; i >= 0; i--) {
        double acc = f[i];
        for (int j = 0; j < L->cnt[i]; j++) {
            m--;
            acc -= L->ele[m] * u[L->col[m]];
        }
        u[i] += acc * L->dia[i];
    }
}

/* simple_restriction (generic_operator.h:68-129) */
static void restriction(const olevel *L, const double *u, double *s, int ncoarse)
{
    for (int i = 0; i < ncoarse; i++) s[i] = 0.0;
    const int *rcol = L->pcol;
    for (int i = 0; i < L->n; i++) {
        int c = L->pcnt[i];
        if (c == 0) continue;
        if (c == 1) { s[*rcol++] += u[i]; continue; }
        double t = u[i] / (double)c;
        for (int j = 0; j < c; j++) s[*rcol++] += t;
    }
}

/* simple_prolongation (generic_operator.h:7-66) */
static void prolongation(const olevel *L, const double *u, double *s)
{
    const int *rcol = L->pcol;
    for (int i = 0; i < L->n; i++) {
        int c = L->pcnt[i];
        if (c == 0) continue;
        if (c == 1) { s[i] = u[*rcol++]; continue; }
        double t;
        if (c <= 4) {
            t = u[*rcol++];
            for (int j = 1; j < c; j++) t += u[*rcol++];
        } else {
            t = 0.0;
            for (int j = 0; j < c; j++) t += u[*rcol++];
        }
        s[i] = t / (double)c;
    }
}

/* multigrid (amg.h:1270-1367), NOMPI Gauss-Seidel branch 1317-1322. */
static void multigrid(oracle_amg *a, int l)
{
    olevel *L = &a->lev[l];
    if (l < a->nlev - 1) {
        olevel *C = &a->lev[l + 1];
        gs_forward_residual(L, L->f, L->u, L->r);
        restriction(L, L->r, C->f, C->n);
        multigrid(a, l + 1);
        prolongation(L, C->u, L->s);
        gs_backward(L, L->f, L->u, L->s);
    } else {
        for (int i = 0; i < L->n; i++) L->u[i] = L->dia[i] * L->f[i]; /* diagonal_solver, amg.h:1084-1096 */
    }
}

void oracle_amg_apply(oracle_amg *a, const double *f, double *u)
{
    memcpy(a->lev[0].f, f, sizeof(double) * a->lev[0].n);
    multigrid(a, 0);
    memcpy(u, a->lev[0].u, sizeof(double) * a->lev[0].n);
}

/* linear_operator (generic_operator.h:223-246) */
static void spmv(int n, const int *cnt, const int *col, const double *ele, const double *u, double *v)
{
    long m = 0;
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < cnt[i]; j++, m++) s += ele[m] * u[col[m]];
        v[i] = s;
    }
}

static double dot(int n, const double *x, const double *y)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

int oracle_pcg(int n, const int *cnt, const int *col, const double *ele, oracle_amg *amg,
               const double *f, double *u, double eps, int max_iterations)
{
    double *r = (double *)xmalloc(sizeof(double) * (n + 1));
    double *s = (double *)xmalloc(sizeof(double) * (n + 1));
    double *v = (double *)xmalloc(sizeof(double) * (n + 1));
    memcpy(r, f, sizeof(double) * n);
    memcpy(s, u, sizeof(double) * n);
    double alpha = 1.0, beta = 0.0, sigma, sigma_first, sigma_last, tau;
    spmv(n, cnt, col, ele, s, v);
    for (int i = 0; i < n; i++) r[i] -= v[i] * alpha;
    oracle_amg_apply(amg, r, v);
    sigma = dot(n, v, r);
    if (sigma < 0.0) sigma = fabs(sigma);
    for (int i = 0; i < n; i++) s[i] = s[i] * beta + v[i];
    sigma_last = sigma;
    sigma_first = sigma;
    if (sigma == 0.0) sigma_first = 1.0;
    tau = sigma / sigma_first;
    int loop = 0;
    while ((++loop < max_iterations) && (tau > eps * eps)) {
        spmv(n, cnt, col, ele, s, v);
        tau = dot(n, s, v);
        if (tau < 0.0) tau = fabs(tau);
        alpha = sigma / tau;
        for (int i = 0; i < n; i++) u[i] += s[i] * alpha;
        for (int i = 0; i < n; i++) r[i] -= v[i] * alpha;
        oracle_amg_apply(amg, r, v);
        sigma = dot(n, v, r);
        if (sigma < 0.0) sigma = fabs(sigma);
        beta = sigma / sigma_last;
        for (int i = 0; i < n; i++) s[i] = s[i] * beta + v[i];
        sigma_last = sigma;
        tau = sigma / sigma_first;
    }
    free(r); free(s); free(v);
    return loop;
}

int oracle_compute(int n, int nnz, const int *cnt, const int *col, const double *ele,
                   const double *b, double *x, double eps)
{
    (void)nnz;
    oracle_amg *amg = oracle_amg_create(n, cnt, col, ele, 16, 1, 0.0);
    int it = oracle_pcg(n, cnt, col, ele, amg, b, x, eps, 10000);
    oracle_amg_destroy(amg);
    return it;
}

void oracle_fluxos(const oracle_grid *g, const double *p, const oracle_ele *perm, const oracle_bound *B, oracle_ele *v)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz, nxy = nx * ny;
    const double hx = g->hx, hy = g->hy, hz = g->hz;
    const int *bc = g->bc_type;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int s = PIDX(g, i, j, k);
                /* interior faces (flux_reconstruction.c:30-97) */
                if (j < ny - 1) v[s].u = hy * perm[s].u * (p[s] - p[s + nx]);
                if (i < nx - 1) v[s].r = hx * perm[s].r * (p[s] - p[s + 1]);
                if (j > 0) v[s].d = hy * perm[s].d * (p[s] - p[s - nx]);
                if (i > 0) v[s].l = hx * perm[s].l * (p[s] - p[s - 1]);
                if (k < nz - 1) v[s].t = hz * perm[s].t * (p[s] - p[s + nxy]);
                if (k > 0) v[s].b = hz * perm[s].b * (p[s] - p[s - nxy]);
                /* boundary faces (flux_reconstruction.c:101-273) */
                if (i == 0) {
                    int l = k * ny + j;
                    if (bc[0] == 0) v[s].l = hx * perm[s].l * (p[s] - B->l[l]);
                    if (bc[0] == 1) v[s].l = -B->l[l];
                }
                if (i == nx - 1) {
                    int l = k * ny + j;
                    if (bc[1] == 0) v[s].r = hx * perm[s].r * (p[s] - B->r[l]);
                    if (bc[1] == 1) v[s].r = B->r[l];
                }
                if (j == 0) {
                    int l = k * nx + i;
                    if (bc[3] == 0) v[s].d = hy * perm[s].d * (p[s] - B->d[l]);
                    if (bc[3] == 1) v[s].d = -B->d[l];
                }
                if (j == ny - 1) {
                    int l = k * nx + i;
                    if (bc[2] == 0) v[s].u = hy * perm[s].u * (p[s] - B->u[l]);
                    if (bc[2] == 1) v[s].u = B->u[l];
                }
                if (k == nz - 1) {
                    int l = j * nx + i;
                    if (bc[4] == 0) v[s].t = hz * perm[s].t * (p[s] - B->t[l]);
                    if (bc[4] == 1) v[s].t = B->t[l];
                }
                if (k == 0) {
                    int l = j * nx + i;
                    if (bc[5] == 0) v[s].b = hz * perm[s].b * (p[s] - B->b[l]);
                    if (bc[5] == 1) v[s].b = -B->b[l];
                }
            }
}

void oracle_divvel_pontual(const oracle_grid *g, const oracle_ele *v, double *divmin, double *divmax)
{
    double maxx = 0.0, minn = 100.0;
    int n = g->nx * g->ny * g->nz;
    for (int s = 0; s < n; s++) {
        double d = v[s].r / g->hx + v[s].l / g->hx + v[s].d / g->hy + v[s].u / g->hy + v[s].b / g->hz + v[s].t / g->hz;
        if (fabs(d) > maxx) maxx = fabs(d);
        if (fabs(d) < minn) minn = fabs(d);
    }
    *divmin = fabs(minn);
    *divmax = fabs(maxx);
}

/* ------------------------------------------------------------------ 1-D conduits */

#define KX(cs, i, j) ((i) * (cs)->count + (j))

/* permeabilidade_karst (permeabilities.c:795-867) */
static void karst_perm(const oracle_conduits *cs, oracle_karst *st, int j, double visc)
{
    const int m = cs->cells[j];
    const double hs = st->hs[j];
    for (int i = 0; i < m - 1; i++) {
        double t = hm(st->KS[KX(cs, i, j)], st->KS[KX(cs, i + 1, j)]);
        st->kr[KX(cs, i, j)] = t / (hs * hs * visc);
        st->ind_r[i] = i + 1;
    }
    for (int i = 1; i < m; i++) {
        double t = hm(st->KS[KX(cs, i, j)], st->KS[KX(cs, i - 1, j)]);
        st->kl[KX(cs, i, j)] = t / (hs * hs * visc);
        st->ind_l[i] = i - 1;
    }
    const double S0 = st->KS[KX(cs, 0, j)], Sm = st->KS[KX(cs, m - 1, j)];
    if (cs->bc_l[j] == 0) { st->kl[KX(cs, 0, j)] = 2 * (S0) / (hs * hs * visc); st->ind_l[0] = -1; }
    if (cs->bc_l[j] == 1) { st->kl[KX(cs, 0, j)] = 0; st->ind_l[0] = -1; }
    if (cs->bc_l[j] == 3) { st->kl[KX(cs, 0, j)] = 2 * S0 / (visc * hs * hs); st->ind_l[0] = -1; }
    if (cs->bc_r[j] == 0) { st->kr[KX(cs, m - 1, j)] = 2 * (Sm) / (hs * hs * visc); st->ind_r[m - 1] = -1; }
    if (cs->bc_r[j] == 1) { st->kr[KX(cs, m - 1, j)] = 0; st->ind_r[m - 1] = -1; }
    if (cs->bc_r[j] == 3) { st->kr[KX(cs, m - 1, j)] = 2 * Sm / (visc * hs * hs); st->ind_r[m - 1] = -1; }
    for (int i = 0; i < m; i++) {
        st->kc[KX(cs, i, j)] = st->kl[KX(cs, i, j)] + st->kr[KX(cs, i, j)] + st->KI_c[KX(cs, i, j)] / visc;
        st->ind_c[i] = i;
    }
}

/* The reservoir face transmissibility and the K-weighted mean pressure across
   the face between reservoir cells s (outside the conduit) and sa (its end),
   as the Robin branches of source_term1d and fluxo1d compute them. */
static void robin_face(const oracle_grid *g, const double *Kx, const double *Ky, const double *Kz,
                       const oracle_ele *perm, const double *p_old, int axis, int left, int s, int sa, double *u,
                       double *lw)
{
    const double *K = axis == 1 ? Kx : axis == 2 ? Ky : Kz;
    double hf = axis == 1 ? g->hx : axis == 2 ? g->hy : g->hz;
    double T;
    if (axis == 1) T = left ? perm[s].r : perm[s].l;
    else if (axis == 2) T = left ? perm[s].u : perm[s].d;
    else T = left ? perm[s].t : perm[s].b;
    *u = hf * T * (p_old[s] - p_old[sa]);
    *lw = (K[s] * p_old[s] + K[sa] * p_old[sa]) / (K[s] + K[sa]);
}

/* Cell of the reservoir at offset `off` along conduit j's axis from its start. */
static int conduit_cell(const oracle_conduits *cs, const oracle_grid *g, int j, int off)
{
    int i = cs->sx[j], jj = cs->sy[j], k = cs->sz[j];
    if (cs->axis[j] == 1) i += off;
    else if (cs->axis[j] == 2) jj += off;
    else k += off;
    return PIDX(g, i, jj, k);
}

/* source_term1d (set_source_term.c:256-414) */
static void karst_source(const oracle_conduits *cs, oracle_karst *st, int j, const oracle_grid *g, const double *Kx,
                         const double *Ky, const double *Kz, const oracle_ele *perm, const double *p_old, double visc,
                         double *f)
{
    const int m = cs->cells[j];
    const double hs = st->hs[j], rho = cs->rho;
    for (int i = 0; i < m; i++) f[i] = (st->KI_c[KX(cs, i, j)] * st->p_aux[KX(cs, i, j)]) / visc;
    for (int i = 0; i < cs->nbrec[j]; i++) {
        const int s = cs->brec[KX(cs, i, j)];
        const double cR = hm(st->KS[KX(cs, s, j)], st->KS[KX(cs, s + 1, j)]);
        const double Cb = cs->C[KX(cs, i, j)], beta = cs->beta[KX(cs, i, j)];
        double R;
        if (st->ucR_old[KX(cs, i, j)] < 0) R = -Cb * beta;
        else R = Cb * beta;
        const double uRo = st->ucR_old[KX(cs, i, j)], uLo = st->ucL_old[KX(cs, i, j)];
        const double pcs = st->pc_old[KX(cs, s, j)], pcs1 = st->pc_old[KX(cs, s + 1, j)];
        double uR = (cR / (visc * hs)) * (pcs - pcs1 - rho * uRo * uRo * R);
        double uL = (cR / (visc * hs)) * (pcs1 - pcs + rho * uLo * uLo * R);
        st->ucR_old[KX(cs, i, j)] = uR;
        st->ucL_old[KX(cs, i, j)] = uL;
        f[s] = f[s] + (cR / (visc * hs)) * (rho * uR * uR * R);
        f[s + 1] = f[s + 1] - (cR / (visc * hs)) * (rho * uL * uL * R);
    }
    /* left end */
    if (cs->bc_l[j] == 0) f[0] = f[0] + (st->kl[KX(cs, 0, j)]) * cs->cont[KX(cs, 0, j)];
    if (cs->bc_l[j] == 1) f[0] = f[0] + cs->cont[KX(cs, 0, j)] / (hs);
    if (cs->bc_l[j] == 3) {
        int s = conduit_cell(cs, g, j, -1), sa = conduit_cell(cs, g, j, 0);
        double u, lw, c;
        robin_face(g, Kx, Ky, Kz, perm, p_old, cs->axis[j], 1, s, sa, &u, &lw);
        if (u < 0) c = -cs->ccont[KX(cs, 0, j)] * (cs->cont[KX(cs, 0, j)] - 1);
        else c = cs->ccont[KX(cs, 0, j)] * (cs->cont[KX(cs, 0, j)] - 1);
        double font = -lw + rho * u * u * c;
        f[0] = f[0] - 2 * st->KS[KX(cs, 0, j)] / (visc * hs * hs) * font;
    }
    /* right end (the reference has no branch for a Robin right end along z;
       such conduits are rejected by the callers) */
    if (cs->bc_r[j] == 0) f[m - 1] = f[m - 1] + st->kr[KX(cs, m - 1, j)] * cs->cont[KX(cs, 1, j)];
    if (cs->bc_r[j] == 1) f[m - 1] = f[m - 1] - cs->cont[KX(cs, 1, j)] / (hs);
    if (cs->bc_r[j] == 3 && cs->axis[j] != 3) {
        int s = conduit_cell(cs, g, j, m), sa = conduit_cell(cs, g, j, m - 1);
        double u, lw, c;
        robin_face(g, Kx, Ky, Kz, perm, p_old, cs->axis[j], 0, s, sa, &u, &lw);
        if (u < 0) c = -cs->ccont[KX(cs, 1, j)] * (cs->cont[KX(cs, 1, j)] - 1);
        else c = cs->ccont[KX(cs, 1, j)] * (cs->cont[KX(cs, 1, j)] - 1);
        double font = -lw + rho * u * u * c;
        f[m - 1] = f[m - 1] - font * 2 * st->KS[KX(cs, m - 1, j)] / (visc * hs * hs);
    }
}

/* solve_linsys_1d (solve_linearsystem_AMG.c:200-291): rows in the order
   l, c, r, then compute(). */
static void karst_solve(const oracle_conduits *cs, const oracle_karst *st, int j, const double *f, double *x,
                        double eps)
{
    const int m = cs->cells[j];
    int *cnt = (int *)xmalloc(sizeof(int) * (m + 1));
    int *col = (int *)xmalloc(sizeof(int) * (3 * m + 1));
    double *ele = (double *)xmalloc(sizeof(double) * (3 * m + 1));
    int l = 0;
    for (int s = 0; s < m; s++) {
        cnt[s] = 0;
        if (st->ind_l[s] != -1) { col[l] = st->ind_l[s]; ele[l] = -st->kl[KX(cs, s, j)]; l++; cnt[s]++; }
        if (st->ind_c[s] != -1) { col[l] = st->ind_c[s]; ele[l] = st->kc[KX(cs, s, j)]; l++; cnt[s]++; }
        if (st->ind_r[s] != -1) { col[l] = st->ind_r[s]; ele[l] = -st->kr[KX(cs, s, j)]; l++; cnt[s]++; }
    }
    oracle_compute(m, l, cnt, col, ele, f, x, eps);
    free(cnt); free(col); free(ele);
}

/* Thomas algorithm on the same tridiagonal matrix (exact solve; the GPU's
   batched solve is checked against this). */
static void karst_solve_direct(const oracle_conduits *cs, const oracle_karst *st, int j, const double *f, double *x)
{
    const int m = cs->cells[j];
    double *cp = (double *)xmalloc(sizeof(double) * m), *dp = (double *)xmalloc(sizeof(double) * m);
    for (int i = 0; i < m; i++) {
        double a = i > 0 ? -st->kl[KX(cs, i, j)] : 0.0;
        double b = st->kc[KX(cs, i, j)];
        double c = i < m - 1 ? -st->kr[KX(cs, i, j)] : 0.0;
        double den = i > 0 ? b - a * cp[i - 1] : b;
        cp[i] = c / den;
        dp[i] = (i > 0 ? f[i] - a * dp[i - 1] : f[i]) / den;
    }
    for (int i = m - 1; i >= 0; i--) x[i] = i < m - 1 ? dp[i] - cp[i] * x[i + 1] : dp[i];
    free(cp); free(dp);
}

void oracle_conduit(const oracle_conduits *cs, oracle_karst *st, const oracle_grid *g, const double *Kx,
                    const double *Ky, const double *Kz, const oracle_ele *perm, const double *p_old, double visc,
                    double eps1d, int direct)
{
    for (int j = 0; j < cs->count; j++) {
        const int m = cs->cells[j];
        double *x1 = (double *)calloc(m, sizeof(double)), *f = (double *)calloc(m, sizeof(double));
        karst_perm(cs, st, j, visc);
        karst_source(cs, st, j, g, Kx, Ky, Kz, perm, p_old, visc, f);
        if (direct) karst_solve_direct(cs, st, j, f, x1);
        else karst_solve(cs, st, j, f, x1, eps1d);
        for (int i = 0; i < m; i++) st->pc[KX(cs, i, j)] = x1[i];
        free(f); free(x1);
    }
}

void oracle_fluxo1d(const oracle_conduits *cs, const oracle_karst *st, const oracle_grid *g, const double *Kx,
                    const double *Ky, const double *Kz, const oracle_ele *perm, const double *p_old, double visc,
                    const double *x, double *ucl, double *ucr)
{
    const double rho = cs->rho;
    for (int j = 0; j < cs->count; j++) {
        const int m = cs->cells[j];
        const double hs = st->hs[j];
        for (int i = 0; i < m - 1; i++)
            ucr[KX(cs, i, j)] = hs * st->kr[KX(cs, i, j)] * (x[KX(cs, i, j)] - x[KX(cs, i + 1, j)]);
        for (int i = 1; i < m; i++)
            ucl[KX(cs, i, j)] = hs * st->kl[KX(cs, i, j)] * (x[KX(cs, i, j)] - x[KX(cs, i - 1, j)]);
        for (int i = 0; i < cs->nbrec[j]; i++) {
            const double Cb = cs->C[KX(cs, i, j)], beta = cs->beta[KX(cs, i, j)];
            double c;
            if (st->ucR_old[KX(cs, i, j)] < 0) c = -Cb * beta;
            else c = Cb * beta;
            const int s = cs->brec[KX(cs, i, j)];
            const double cR = hm(st->KS[KX(cs, s, j)], st->KS[KX(cs, s + 1, j)]);
            const double uRo = st->ucR_old[KX(cs, i, j)], uLo = st->ucL_old[KX(cs, i, j)];
            ucr[KX(cs, s, j)] = (cR / (visc * hs)) * (x[KX(cs, s, j)] - x[KX(cs, s + 1, j)] - 0.5 * rho * uRo * uRo * c);
            ucl[KX(cs, s + 1, j)] =
                (cR / (visc * hs)) * (x[KX(cs, s + 1, j)] - x[KX(cs, s, j)] + 0.5 * rho * uLo * uLo * c);
        }
        /* left end; a Neumann end reports -cont_karst[1][j], as the reference does */
        if (cs->bc_l[j] == 0) ucl[KX(cs, 0, j)] = hs * st->kl[KX(cs, 0, j)] * (x[KX(cs, 0, j)] - cs->cont[KX(cs, 0, j)]);
        if (cs->bc_l[j] == 1) ucl[KX(cs, 0, j)] = -cs->cont[KX(cs, 1, j)];
        if (cs->bc_l[j] == 3) {
            int s = conduit_cell(cs, g, j, -1), sa = conduit_cell(cs, g, j, 0);
            double u, lw, c;
            robin_face(g, Kx, Ky, Kz, perm, p_old, cs->axis[j], 1, s, sa, &u, &lw);
            if (u < 0) c = -cs->ccont[KX(cs, 0, j)] * (cs->cont[KX(cs, 0, j)] - 1);
            else c = cs->ccont[KX(cs, 0, j)] * (cs->cont[KX(cs, 0, j)] - 1);
            ucl[KX(cs, 0, j)] = (2 * st->KS[KX(cs, 0, j)]) / (visc * hs) * (x[KX(cs, 0, j)] - lw + rho * u * u * c);
        }
        /* right end */
        if (cs->bc_r[j] == 0)
            ucr[KX(cs, m - 1, j)] = hs * st->kr[KX(cs, m - 1, j)] * (x[KX(cs, m - 1, j)] - cs->cont[KX(cs, 1, j)]);
        if (cs->bc_r[j] == 1) ucr[KX(cs, m - 1, j)] = cs->cont[KX(cs, 1, j)];
        if (cs->bc_r[j] == 3) {
            int s = conduit_cell(cs, g, j, m), sa = conduit_cell(cs, g, j, m - 1);
            double u, lw, c;
            robin_face(g, Kx, Ky, Kz, perm, p_old, cs->axis[j], 0, s, sa, &u, &lw);
            if (u < 0) c = -cs->ccont[KX(cs, 1, j)] * (cs->cont[KX(cs, 1, j)] - 1);
            else c = cs->ccont[KX(cs, 1, j)] * (cs->cont[KX(cs, 1, j)] - 1);
            ucr[KX(cs, m - 1, j)] =
                (2 * st->KS[KX(cs, m - 1, j)]) / (visc * hs) * (x[KX(cs, m - 1, j)] - lw + rho * u * u * c);
        }
    }
}

double oracle_rel_change(int n, const double *y, const double *y_old)
{
    double s1 = 0, s2 = 0;
    for (int i = 0; i < n; i++) {
        s1 = s1 + (y[i] - y_old[i]) * (y[i] - y_old[i]);
        s2 = s2 + y[i] * y[i];
    }
    return sqrt(s1) / sqrt(s2);
}
