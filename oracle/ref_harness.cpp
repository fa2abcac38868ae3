/*
 * Driver for the reference build (oracle/_ref). TEST INFRASTRUCTURE ONLY.
 *
 * Links the reference's own assembly, right-hand side, linear-solve and flux
 * sources (compiled in place from the reference tree by oracle/Makefile) and
 * its header-only AMG/PCG toolbox, and drives them on one problem read from a
 * binary file. It replaces only main.c's role (main.c:123-514): it defines
 * the globals that those sources read (variables.h) and sets them from the
 * input instead of from hard-coded constants.
 *
 * Outputs (binary, little-endian, into the given prefix):
 *   <prefix>.A      n, nnz, cnt[n], col[nnz], ele[nnz]  (row order d,l,c,r,u,t,b)
 *   <prefix>.b      b[n]
 *   <prefix>.p8     pressure from solve_linsys() (reference stop 1e-8)
 *   <prefix>.p12    pressure from the same AMG-PCG driven to `eps_tight`
 *   <prefix>.v      fluxos() on p12, 7 doubles per cell (u,d,l,r,c,t,b)
 *   <prefix>.json   iteration counts, level sizes, timings, divergence
 *
 *   <prefix>.p8w    a second solve_linsys() started from p8 (warm start, as
 *                   main.c:460 passes the previous pressure back in)
 *
 * With a third argument "time" it only times one reservoir() step (assembly,
 * solve_linsys at the reference's 1e-8, fluxos) and writes <prefix>.json;
 * "time <k>" stops the PCG after k iterations (a bounded sample of a large
 * step: the JSON then holds the setup time and the per-iteration time).
 * With "summary" it writes p8, p8w, p12 and the JSON (no matrix, fluxes or
 * level sizes): the large cases.
 * With "karst <data file> <max passes> [<dir>]" it runs the coupled reservoir/conduit
 * model as main.c:390-530 does: reservoir() without well or conduits, the
 * conduit set read from the data file (the reference's data_<nk>_<type>.dat
 * format, parsed as main.c:197-300 parses it), conduit(), then the
 * fixed-point loop of reservoir() and conduit() until the reservoir pressure
 * changes by at most 1e-8 between passes (or the pass cap), then fluxo1d()
 * and fluxos(). The input's WI and pw arrays are the well; its KI and pc are
 * ignored (the loop sets them). Outputs: <prefix>.karst.p (final reservoir
 * pressure), .p1 (after the loop's first pass), .pc0 (conduit pressure
 * before the loop), .pc1 (after the first pass), .pc (final), .ucl/.ucr
 * (fluxo1d), .ucR_old/.ucL_old (breccia velocities), .v (fluxos), and
 * <prefix>.karst.json (passes, per-pass change, iteration counts); with
 * <dir>, save_pressure_1d() and save_flux_1d() (posprocessing_output.c:
 * 437-659) write pressure_karst_<j>.dat and flux_karst_<j>.dat there.
 * With "write <dir>" it also runs the reference's output writers on the tight
 * pressure and its fluxes: save_pressure(), save_flux()
 * (posprocessing_output.c:200-296) and vts_StructuredGrid_file() (vtk.c:8-76)
 * write pressure_reservoir.dat, flux_reservoir.dat and campos.vts into <dir>.
 *
 * Input file layout (all little-endian):
 *   int32  nx, ny, nz
 *   f64    Ix, Iy, Iz, viscosidade, eps_tight
 *   int32  bc_type[6]   order l, r, u, d, t, b   (0 Dirichlet, 1 Neumann)
 *   f64    bc_val[6]    same order
 *   int32  has_sources  (1: KI, WI, pc, pw arrays follow the K arrays)
 *   f64    Kx[n], Ky[n], Kz[n] [, KI[n], WI[n], pc[n], pw[n]]
 */
#define NOMPI
#include "toolbox/toolbox.h"

#include "structs.h"
#include "variables.h"
#include "functions.h"

#include <sys/time.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>

/* Globals that the linked reference sources read or write (variables.h). */
Dim N;
grid I;
grid h;
Perm *P;
ELE *perm;
IND *indcol;
ELE *v;
Bound *B;
FACE boundary;
FACEK *BoundKarst;
IND *indcol1;
ELEK *uc;
Perm *KI;
DimK *M;
double *wi;
int *IDwi;
int Nwi;
int ix0, ixf, jy0, jyf;
double karst;
double viscosidade;
double **beta;
double rho;
int *n_brechas;
int **brechas;
double **pc_old;
double **ucL_old;
double **ucR_old;
double *hs;
double *Kc;
double **cont_karst;
int number_karst;
int Mmax;
double **pc;
double **C;
double **C_contorno;
double nxg, nyg, nzg;
char fnam[1024];
const char *fn;
const char *fnwi;
const char *dirname;
double time_total;
double time_amg;
int iter_amg;
int layer;
int lay;

static double now()
{
    struct timeval t;
    gettimeofday(&t, 0);
    return (double)t.tv_sec + 1e-6 * (double)t.tv_usec;
}

static void put(const char *prefix, const char *ext, const void *data, size_t bytes)
{
    char name[2048];
    snprintf(name, sizeof(name), "%s%s", prefix, ext);
    FILE *f = fopen(name, "wb");
    if (!f) { perror(name); exit(2); }
    if (bytes && fwrite(data, 1, bytes, f) != bytes) { perror(name); exit(2); }
    fclose(f);
}

static void get(FILE *f, void *dst, size_t bytes)
{
    if (fread(dst, 1, bytes, f) != bytes) { fprintf(stderr, "short input\n"); exit(2); }
}

static void put_matrix(const char *prefix, const char *ext, double **m, int rows, int cols)
{
    std::vector<double> flat((size_t)rows * cols);
    for (int i = 0; i < rows; i++)
        for (int j = 0; j < cols; j++) flat[(size_t)i * cols + j] = m[i][j];
    put(prefix, ext, flat.data(), flat.size() * sizeof(double));
}

static void fail(const char *what)
{
    fprintf(stderr, "karst input: %s\n", what);
    exit(2);
}

/* main.c:197-300 and :390-530 for one problem (see the header comment). */
static int run_karst(const char *datafile, int max_passes, const char *pre, Perm *K, Bound *Bb, const double *WIin,
                     const double *pwin, const char *writedir)
{
    const int n = N.t;
    FILE *arq = fopen(datafile, "r");
    if (!arq) fail("cannot open the data file");
    /* the first line holds number_karst values of M[].s: count them */
    char line[65536];
    if (!fgets(line, sizeof(line), arq)) fail("empty data file");
    int nk = 0;
    for (char *q = strtok(line, " \t\r\n"); q; q = strtok(nullptr, " \t\r\n")) nk++;
    rewind(arq);
    number_karst = nk;
    DimK *Mk = (DimK *)calloc(nk + 1, sizeof(DimK));
    Kc = create_vector(nk);
    FACEK BK;
    BK.l = create_int(nk);
    BK.r = create_int(nk);
    cont_karst = create_matrix(2, nk);
    C_contorno = create_matrix(2, nk);
    std::vector<double> karst_index(nk), Ac(nk);
    n_brechas = create_int(nk);
    int i, j;
#define RD(fmt, dst) do { if (fscanf(arq, fmt, dst) != 1) fail("short data file"); } while (0)
    for (i = 0; i < nk; i++) RD("%i", &Mk[i].s);
    for (i = 0; i < nk; i++) RD("%i", &Mk[i].sxi);
    for (i = 0; i < nk; i++) RD("%i", &Mk[i].syi);
    for (i = 0; i < nk; i++) RD("%i", &Mk[i].szi);
    for (i = 0; i < nk; i++) RD("%i", &Mk[i].pos);
    for (i = 0; i < nk; i++) RD("%lf", &Kc[i]);
    for (i = 0; i < nk; i++) RD("%i", &BK.l[i]);
    for (i = 0; i < nk; i++) RD("%i", &BK.r[i]);
    for (j = 0; j < 2; j++) for (i = 0; i < nk; i++) RD("%lf", &cont_karst[j][i]);
    for (j = 0; j < 2; j++) for (i = 0; i < nk; i++) RD("%lf", &C_contorno[j][i]);
    for (i = 0; i < nk; i++) RD("%lf", &karst_index[i]);
    for (i = 0; i < nk; i++) RD("%lf", &Ac[i]);
    for (i = 0; i < nk; i++) RD("%i", &n_brechas[i]);
    int max_brechas = -1;
    for (i = 0; i < nk; i++) if (n_brechas[i] >= max_brechas) max_brechas = n_brechas[i];
    ::beta = create_matrix(max_brechas, nk);
    for (j = 0; j < max_brechas; j++) for (i = 0; i < nk; i++) { RD("%lf", &::beta[j][i]); ::beta[j][i] = ::beta[j][i] - 1; }
    C = create_matrix(max_brechas, nk);
    for (j = 0; j < max_brechas; j++) for (i = 0; i < nk; i++) RD("%lf", &C[j][i]);
    brechas = create_matrix_int(max_brechas, nk);
    for (j = 0; j < max_brechas; j++) for (i = 0; i < nk; i++) RD("%d", &brechas[j][i]);
#undef RD
    fclose(arq);
    hs = create_vector(nk);
    for (i = 0; i < nk; i++) hs[i] = Mk[i].pos == 1 ? h.x : Mk[i].pos == 2 ? h.y : h.z;
    int mmax = 0;
    for (i = 0; i < nk; i++) if (Mk[i].s > mmax) mmax = Mk[i].s;
    Mmax = mmax;
    pc = create_matrix(Mmax, nk);
    pc_old = create_matrix(Mmax, nk);
    double **KI_c = create_matrix(Mmax, nk);
    double *KIv = create_vector(n);
    double *WI0 = create_vector(n);
    double *pw = create_vector(n);
    for (int s = 0; s < n; s++) pw[s] = pwin[s];
    K->S = create_matrix(Mmax, nk);
    for (j = 0; j < nk; j++) for (i = 0; i < Mmax; i++) K->S[i][j] = Kc[j];   /* set_perm_faces */
    ucL_old = create_matrix(max_brechas, nk);
    ucR_old = create_matrix(max_brechas, nk);
    ELE *pm = (ELE *)calloc(n + 1, sizeof(ELE));
    ELEK keff;
    keff.l = create_matrix(Mmax, nk); keff.r = create_matrix(Mmax, nk); keff.c = create_matrix(Mmax, nk);
    ELEK ucf;
    ucf.l = create_matrix(Mmax, nk); ucf.r = create_matrix(Mmax, nk); ucf.c = create_matrix(Mmax, nk);
    ELE *vv = (ELE *)calloc(n + 1, sizeof(ELE));
    IND *ind1 = (IND *)calloc(Mmax + 1, sizeof(IND));
    IND *ic = (IND *)calloc(n + 1, sizeof(IND));
    double *p = create_vector(n), *p_old = create_vector(n);
    std::vector<int> iters;

    /* main.c:395-396: first pass, no well (WI still zero) and no conduits */
    reservoir(N, K, WI0, KIv, pc_old, ic, Bb, pm, h, boundary, viscosidade, pw, Mk, p);
    iters.push_back(iter_amg);
    for (int s = 0; s < n; s++) p_old[s] = p[s];
    double *WI = create_vector(n);
    for (int s = 0; s < n; s++) WI[s] = WIin[s];
    double **p_aux = create_matrix(Mmax, nk);
    auto cell = [&](int jj, int ii) {
        int x = Mk[jj].sxi, y = Mk[jj].syi, z = Mk[jj].szi;
        if (Mk[jj].pos == 1) x += ii; else if (Mk[jj].pos == 2) y += ii; else z += ii;
        return x + y * N.x + z * N.xy;
    };
    for (j = 0; j < nk; j++)
        for (i = 0; i < Mk[j].s; i++) {
            int aux = cell(j, i);
            KIv[aux] = karst_index[j];
            KI_c[i][j] = karst_index[j] / (Ac[j] * hs[j]);
            p_aux[i][j] = p_old[aux];
            K->S[i][j] = Kc[j];
        }
    conduit(p_aux, cont_karst, KI_c, viscosidade, Mk, hs, ind1, &BK, &keff, K, pm, p_old, pc);
    put_matrix(pre, ".karst.pc0", pc, Mmax, nk);

    /* main.c:457-514; the loop's stop test reduces to the reservoir change
       (error1[number_karst]), the value the reference's max search keeps */
    double error = 1;
    int iter = 0;
    std::vector<double> errs;
    while (iter < max_passes && error > 1e-8) {
        for (i = 0; i < Mmax; i++) for (j = 0; j < nk; j++) pc_old[i][j] = pc[i][j];
        reservoir(N, K, WI, KIv, pc_old, ic, Bb, pm, h, boundary, viscosidade, pw, Mk, p);
        iters.push_back(iter_amg);
        double s1 = 0, s2 = 0;
        for (int s = 0; s < n; s++) { s1 = s1 + (p[s] - p_old[s]) * (p[s] - p_old[s]); s2 = s2 + p[s] * p[s]; }
        error = sqrt(s1) / sqrt(s2);
        errs.push_back(error);
        for (int s = 0; s < n; s++) p_old[s] = p[s];
        for (j = 0; j < nk; j++) for (i = 0; i < Mk[j].s; i++) p_aux[i][j] = p_old[cell(j, i)];
        conduit(p_aux, cont_karst, KI_c, viscosidade, Mk, hs, ind1, &BK, &keff, K, pm, p_old, pc);
        iter++;
        if (iter == 1) {
            put(pre, ".karst.p1", p, sizeof(double) * n);
            put_matrix(pre, ".karst.pc1", pc, Mmax, nk);
        }
    }
    fluxo1d(pc, &keff, &ucf, hs, Mk, cont_karst, &BK, K, pm, p_old);
    fluxos(p, pm, vv, Bb, h, N);
    if (writedir) {   /* main.c:526-527 */
        dirname = writedir;
        save_pressure_1d(pc, Mk, hs, h);
        save_flux_1d(&ucf, Mk, hs, h);
    }
    put(pre, ".karst.p", p, sizeof(double) * n);
    put_matrix(pre, ".karst.pc", pc, Mmax, nk);
    put_matrix(pre, ".karst.ucl", ucf.l, Mmax, nk);
    put_matrix(pre, ".karst.ucr", ucf.r, Mmax, nk);
    put_matrix(pre, ".karst.ucR_old", ucR_old, max_brechas, nk);
    put_matrix(pre, ".karst.ucL_old", ucL_old, max_brechas, nk);
    put_matrix(pre, ".karst.kc", keff.c, Mmax, nk);
    put(pre, ".karst.v", vv, sizeof(ELE) * n);
    char name[2048];
    snprintf(name, sizeof(name), "%s.karst.json", pre);
    FILE *f = fopen(name, "w");
    fprintf(f, "{\"number_karst\": %d, \"mmax\": %d, \"max_brec\": %d, \"passes\": %d, \"error\": %.17g,\n",
            nk, Mmax, max_brechas, iter, error);
    fprintf(f, " \"cells\": [");
    for (j = 0; j < nk; j++) fprintf(f, "%s%d", j ? ", " : "", Mk[j].s);
    fprintf(f, "], \"start\": [");
    for (j = 0; j < nk; j++) fprintf(f, "%s%d, %d, %d", j ? ", " : "", Mk[j].sxi, Mk[j].syi, Mk[j].szi);
    fprintf(f, "], \"axis\": [");
    for (j = 0; j < nk; j++) fprintf(f, "%s%d", j ? ", " : "", Mk[j].pos);
    fprintf(f, "], \"Kc\": [");
    for (j = 0; j < nk; j++) fprintf(f, "%s%.17g", j ? ", " : "", Kc[j]);
    fprintf(f, "], \"bc\": [");
    for (j = 0; j < nk; j++) fprintf(f, "%s%d, %d", j ? ", " : "", BK.l[j], BK.r[j]);
    fprintf(f, "], \"cont\": [");
    for (int sd = 0; sd < 2; sd++) for (j = 0; j < nk; j++) fprintf(f, "%s%.17g", sd + j ? ", " : "", cont_karst[sd][j]);
    fprintf(f, "], \"ccont\": [");
    for (int sd = 0; sd < 2; sd++) for (j = 0; j < nk; j++) fprintf(f, "%s%.17g", sd + j ? ", " : "", C_contorno[sd][j]);
    fprintf(f, "], \"karst_index\": [");
    for (j = 0; j < nk; j++) fprintf(f, "%s%.17g", j ? ", " : "", karst_index[j]);
    fprintf(f, "], \"area\": [");
    for (j = 0; j < nk; j++) fprintf(f, "%s%.17g", j ? ", " : "", Ac[j]);
    fprintf(f, "], \"nbrec\": [");
    for (j = 0; j < nk; j++) fprintf(f, "%s%d", j ? ", " : "", n_brechas[j]);
    fprintf(f, "], \"brec\": [");
    for (i = 0; i < max_brechas; i++) for (j = 0; j < nk; j++) fprintf(f, "%s%d", i + j ? ", " : "", brechas[i][j]);
    fprintf(f, "], \"beta\": [");
    for (i = 0; i < max_brechas; i++) for (j = 0; j < nk; j++) fprintf(f, "%s%.17g", i + j ? ", " : "", ::beta[i][j]);
    fprintf(f, "], \"C\": [");
    for (i = 0; i < max_brechas; i++) for (j = 0; j < nk; j++) fprintf(f, "%s%.17g", i + j ? ", " : "", C[i][j]);
    fprintf(f, "],\n \"pass_change\": [");
    for (size_t q = 0; q < errs.size(); q++) fprintf(f, "%s%.17g", q ? ", " : "", errs[q]);
    fprintf(f, "], \"iter_amg\": [");
    for (size_t q = 0; q < iters.size(); q++) fprintf(f, "%s%d", q ? ", " : "", iters[q]);
    fprintf(f, "]}\n");
    fclose(f);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3) { fprintf(stderr, "usage: %s input.bin out_prefix\n", argv[0]); return 2; }
    FILE *fi = fopen(argv[1], "rb");
    if (!fi) { perror(argv[1]); return 2; }
    int dims[3];
    double geo[5];
    int bct[6];
    double bcv[6];
    int has_src;
    get(fi, dims, sizeof(dims));
    get(fi, geo, sizeof(geo));
    get(fi, bct, sizeof(bct));
    get(fi, bcv, sizeof(bcv));
    get(fi, &has_src, sizeof(has_src));

    N.x = dims[0]; N.y = dims[1]; N.z = dims[2];
    N.xy = N.x * N.y; N.t = N.xy * N.z;
    N.px = N.x / 2; N.py = N.y / 2; N.pz = 0;
    I.x = geo[0]; I.y = geo[1]; I.z = geo[2];
    h.x = I.x / N.x; h.y = I.y / N.y; h.z = I.z / N.z;
    viscosidade = geo[3];
    double eps_tight = geo[4];
    boundary.l = bct[0]; boundary.r = bct[1]; boundary.u = bct[2];
    boundary.d = bct[3]; boundary.t = bct[4]; boundary.b = bct[5];
    number_karst = 0;
    Mmax = 1;
    int n = N.t;

    Perm Kp;
    Kp.X = create_vector(n); Kp.Y = create_vector(n); Kp.Z = create_vector(n);
    Kp.S = 0; Kp.permeability = 1;
    get(fi, Kp.X, sizeof(double) * n);
    get(fi, Kp.Y, sizeof(double) * n);
    get(fi, Kp.Z, sizeof(double) * n);
    double *KIv = create_vector(n), *WIv = create_vector(n), *pcv = create_vector(n), *pwv = create_vector(n);
    if (has_src) {
        get(fi, KIv, sizeof(double) * n);
        get(fi, WIv, sizeof(double) * n);
        get(fi, pcv, sizeof(double) * n);
        get(fi, pwv, sizeof(double) * n);
    }
    fclose(fi);

    /* Boundary value arrays, as main.c:376-387 + create_boundary(). */
    double val_bound[7];
    /* create_boundary order: [1]=l, [2]=r, [3]=u, [4]=d, [5]=t, [6]=b */
    val_bound[1] = bcv[0]; val_bound[2] = bcv[1]; val_bound[3] = bcv[2];
    val_bound[4] = bcv[3]; val_bound[5] = bcv[4]; val_bound[6] = bcv[5];
    Bound Bb;
    Bb.l = create_vector(n); Bb.r = create_vector(n); Bb.u = create_vector(n);
    Bb.d = create_vector(n); Bb.t = create_vector(n); Bb.b = create_vector(n);
    create_boundary(val_bound, &Bb, N, boundary);

    ELE *pm = (ELE *)calloc(n + 1, sizeof(ELE));
    IND *ic = (IND *)calloc(n + 1, sizeof(IND));
    ELE *vv = (ELE *)calloc(n + 1, sizeof(ELE));
    double *b = create_vector(n);
    double *p8 = create_vector(n);
    double *p12 = create_vector(n);

    double t0 = now();
    permeabilidade_arestas(&Kp, pm, ic, h, N, boundary, WIv, KIv, viscosidade);
    source_term(&Bb, pm, h, N, boundary, pcv, pwv, KIv, WIv, viscosidade, b);
    double t_assemble = now() - t0;

    const char *mode = argc > 3 ? argv[3] : "";
    const bool summary = strcmp(mode, "summary") == 0;
    rho = 1000;   /* main.c:130 */

    if (strcmp(mode, "karst") == 0) {
        if (argc < 6) { fprintf(stderr, "usage: %s input.bin out_prefix karst data_file max_passes\n", argv[0]); return 2; }
        return run_karst(argv[4], atoi(argv[5]), argv[2], &Kp, &Bb, WIv, pwv, argc > 6 ? argv[6] : nullptr);
    }

    if (strcmp(mode, "time") == 0 && argc > 4) {
        /* bounded timing sample: the parts of one reservoir() step at this
           size, with the PCG cut off after k iterations. The count-CSR and
           the toolbox objects are built as solve_linsys() and compute() build
           them (solve_linearsystem_AMG.c:57-150, compute.cpp:16-45). */
        int k = atoi(argv[4]);
        t0 = now();
        std::vector<int> cnt(n), col;
        std::vector<double> ele;
        col.reserve((size_t)7 * n); ele.reserve((size_t)7 * n);
        for (int s = 0; s < n; s++) {
            int c = 0;
            if (ic[s].d != -1) { col.push_back(ic[s].d); ele.push_back(-pm[s].d); c++; }
            if (ic[s].l != -1) { col.push_back(ic[s].l); ele.push_back(-pm[s].l); c++; }
            if (ic[s].c != -1) { col.push_back(ic[s].c); ele.push_back(pm[s].c); c++; }
            if (ic[s].r != -1) { col.push_back(ic[s].r); ele.push_back(-pm[s].r); c++; }
            if (ic[s].u != -1) { col.push_back(ic[s].u); ele.push_back(-pm[s].u); c++; }
            if (ic[s].t != -1) { col.push_back(ic[s].t); ele.push_back(-pm[s].t); c++; }
            if (ic[s].b != -1) { col.push_back(ic[s].b); ele.push_back(-pm[s].b); c++; }
            cnt[s] = c;
        }
        std::vector<int> nod(n);
        for (int i = 0; i < n; i++) nod[i] = i;
        std::vector<double> bv(b, b + n), xv(n, 0.0);
        double t_csr = now() - t0;
        t0 = now();
        communicator<int, double> com(nod);
        linear_operator<int, double> lin(cnt, col, ele);
        vector_product<int, double> sca(com);
        amg_solver<int, double> amg(nod, cnt, col, ele, 16, 1, 0.0, 1.0, true);
        double t_setup = now() - t0;
        t0 = now();
        /* _max = k + 1: the loop runs while ++_loop < _max (pcg.h:57), plus
           the initial residual and preconditioner application */
        conjugate_gradient<int, double> cg(lin, amg, sca, 1e-8, k + 1, false);
        cg(bv, xv);
        double t_pcg = now() - t0;
        int done = cg.iterations();
        for (int i = 0; i < n; i++) p8[i] = xv[i];
        t0 = now();
        fluxos(p8, pm, vv, &Bb, h, N);
        double t_flux = now() - t0;
        char name[2048];
        snprintf(name, sizeof(name), "%s.json", argv[2]);
        FILE *f = fopen(name, "w");
        fprintf(f, "{\"n\": %d, \"pcg_loop\": %d, \"t_assemble\": %.6f, \"t_csr\": %.6f, \"t_setup\": %.6f, "
                   "\"t_pcg\": %.6f, \"t_flux\": %.6f}\n", n, done, t_assemble, t_csr, t_setup, t_pcg, t_flux);
        fclose(f);
        return 0;
    }

    /* Reference seam at its own stopping rule (compute.cpp:41, 1e-8). */
    t0 = now();
    solve_linsys(pm, ic, b, p8, N);
    double t_solve8 = now() - t0;
    int iter8 = iter_amg;

    if (strcmp(mode, "time") == 0) {
        /* timing mode: one reservoir() step as the reference runs it
           (assembly + solve_linsys at 1e-8) plus fluxos(); no outputs but JSON */
        t0 = now();
        fluxos(p8, pm, vv, &Bb, h, N);
        double t_flux8 = now() - t0;
        char name[2048];
        snprintf(name, sizeof(name), "%s.json", argv[2]);
        FILE *f = fopen(name, "w");
        fprintf(f, "{\"n\": %d, \"iter_amg_1e-8\": %d, \"t_assemble\": %.6f, \"t_solve_1e-8\": %.6f, \"t_flux\": %.6f}\n",
                n, iter8, t_assemble, t_solve8, t_flux8);
        fclose(f);
        return 0;
    }

    /* A second pass started from p8, as main.c:460 hands the previous
       pressure back to solve_linsys(): the stop test is relative to this
       start's preconditioned residual (pcg.h:42-57). */
    double *p8w = create_vector(n);
    for (int i = 0; i < n; i++) p8w[i] = p8[i];
    solve_linsys(pm, ic, b, p8w, N);
    int iter8w = iter_amg;

    /* Count-CSR exactly as solve_linsys() builds it (solve_linearsystem_AMG.c:57-143). */
    std::vector<int> cnt(n), col;
    std::vector<double> ele;
    col.reserve((size_t)7 * n); ele.reserve((size_t)7 * n);
    for (int s = 0; s < n; s++) {
        int c = 0;
        if (ic[s].d != -1) { col.push_back(ic[s].d); ele.push_back(-pm[s].d); c++; }
        if (ic[s].l != -1) { col.push_back(ic[s].l); ele.push_back(-pm[s].l); c++; }
        if (ic[s].c != -1) { col.push_back(ic[s].c); ele.push_back(pm[s].c); c++; }
        if (ic[s].r != -1) { col.push_back(ic[s].r); ele.push_back(-pm[s].r); c++; }
        if (ic[s].u != -1) { col.push_back(ic[s].u); ele.push_back(-pm[s].u); c++; }
        if (ic[s].t != -1) { col.push_back(ic[s].t); ele.push_back(-pm[s].t); c++; }
        if (ic[s].b != -1) { col.push_back(ic[s].b); ele.push_back(-pm[s].b); c++; }
        cnt[s] = c;
    }
    int nnz = (int)col.size();

    /* Same solver objects as compute.cpp:31-45, stop at eps_tight. */
    std::vector<int> nod(n);
    for (int i = 0; i < n; i++) nod[i] = i;
    std::vector<double> bv(b, b + n), xv(n, 0.0);
    t0 = now();
    communicator<int, double> com(nod);
    linear_operator<int, double> lin(cnt, col, ele);
    vector_product<int, double> sca(com);
    amg_solver<int, double> amg(nod, cnt, col, ele, 16, 1, 0.0, 1.0, true);
    double t_setup = now() - t0;
    t0 = now();
    conjugate_gradient<int, double> cg(lin, amg, sca, eps_tight, 10000, false);
    cg(bv, xv);
    double t_pcg = now() - t0;
    int iter12 = cg.iterations();
    for (int i = 0; i < n; i++) p12[i] = xv[i];

    /* Level sizes: rerun the hierarchy loop of amg.h:1188-1215 with the
       reference's own amg_restriction, recording rows and nonzeros. */
    std::vector<int> lev_rows, lev_nnz;
    if (!summary) {
        std::vector<int> anod(nod), acnt(cnt), acol(col);
        std::vector<double> aele(ele);
        amg_restriction<int, double> restr;
        int l = 0;
        double rel = 0.0;
        int g = n;
        lev_rows.push_back(n); lev_nnz.push_back(nnz);
        while (l < 15 && g > 1 && rel < 1.0) {
            communicator<int, double> c(anod);
            c.assign(anod);
            std::vector<double> adia;
            std::vector<char> abou, rnod;
            std::vector<int> rcnt, rcol, cnod, ccnt, ccol;
            std::vector<double> rele, cele;
            restr(0.0, anod, acnt, acol, aele, adia, abou, c, rnod, rcnt, rcol, rele, cnod, ccnt, ccol, cele);
            anod.swap(cnod); acnt.swap(ccnt); acol.swap(ccol); aele.swap(cele);
            int gold = g;
            g = (int)acnt.size();
            rel = (double)g / (double)gold;
            lev_rows.push_back(g); lev_nnz.push_back((int)aele.size());
            l++;
        }
    }

    /* Fluxes of the tight solution (flux_reconstruction.c:9-279); fluxos()
       reads the global `boundary` and prints divvel_pontual()'s min/max. */
    t0 = now();
    fluxos(p12, pm, vv, &Bb, h, N);
    double t_flux = now() - t0;
    double divmax = 0.0, divmin = 100.0;
    for (int s = 0; s < n; s++) {
        double dv = vv[s].r / h.x + vv[s].l / h.x + vv[s].d / h.y + vv[s].u / h.y + vv[s].b / h.z + vv[s].t / h.z;
        if (fabs(dv) > divmax) divmax = fabs(dv);
        if (fabs(dv) < divmin) divmin = fabs(dv);
    }

    if (argc > 4 && strcmp(mode, "write") == 0) {
        dirname = argv[4];
        save_pressure(p12, N, h);
        save_flux(vv, N, h);
        vts_StructuredGrid_file(vv, p12, N, h);
    }
    const char *pre = argv[2];
    if (!summary) {
        char name[2048];
        snprintf(name, sizeof(name), "%s.A", pre);
        FILE *f = fopen(name, "wb");
        if (!f) { perror(name); return 2; }
        fwrite(&n, sizeof(int), 1, f);
        fwrite(&nnz, sizeof(int), 1, f);
        fwrite(cnt.data(), sizeof(int), n, f);
        fwrite(col.data(), sizeof(int), nnz, f);
        fwrite(ele.data(), sizeof(double), nnz, f);
        fclose(f);
    }
    put(pre, ".b", b, sizeof(double) * n);
    put(pre, ".p8", p8, sizeof(double) * n);
    put(pre, ".p8w", p8w, sizeof(double) * n);
    put(pre, ".p12", p12, sizeof(double) * n);
    if (!summary) put(pre, ".v", vv, sizeof(ELE) * n);
    {
        char name[2048];
        snprintf(name, sizeof(name), "%s.json", pre);
        FILE *f = fopen(name, "w");
        fprintf(f, "{\"n\": %d, \"nnz\": %d, \"iter_amg_1e-8\": %d, \"iter_amg_warm\": %d, \"iter_tight\": %d, \"eps_tight\": %.3e,\n",
                n, nnz, iter8, iter8w, iter12, eps_tight);
        fprintf(f, " \"levels_rows\": [");
        for (size_t i = 0; i < lev_rows.size(); i++) fprintf(f, "%s%d", i ? ", " : "", lev_rows[i]);
        fprintf(f, "], \"levels_nnz\": [");
        for (size_t i = 0; i < lev_nnz.size(); i++) fprintf(f, "%s%d", i ? ", " : "", lev_nnz[i]);
        fprintf(f, "],\n \"t_assemble\": %.6f, \"t_solve_1e-8\": %.6f, \"t_setup\": %.6f, \"t_pcg_tight\": %.6f, \"t_flux\": %.6f,\n", t_assemble, t_solve8, t_setup, t_pcg, t_flux);
        fprintf(f, " \"div_max\": %.17g, \"div_min\": %.17g}\n", divmax, divmin);
        fclose(f);
    }
    return 0;
}
