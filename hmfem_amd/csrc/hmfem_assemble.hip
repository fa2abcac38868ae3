// Transmissibility assembly, right-hand side and flux reconstruction on the
// GPU, one thread per cell. Built with floating-point contraction off so that
// every value matches the reference's FP64 expressions bit for bit:
//   permeabilidade_arestas  permeabilities.c:196-492
//   source_term             set_source_term.c:9-223
//   fluxos                  flux_reconstruction.c:9-279
//   divvel_pontual          flux_reconstruction.c:281-319
// Per cell the kernels read K (3 x 8 B), optional KI/WI/pc/pw, and write the
// stencil (4 x 8 B) and b (8 B): HBM-bound, ~64 algorithmic bytes per cell.
#include "hmfem_internal.h"

#include <string.h>

namespace hmfem {

struct GridArgs {
    int nx, ny, nz;
    double hx, hy, hz, mu;
    int bc[6];            // l, r, u, d, t, b
    int all_neumann;
    long pin;             // pindex(0, ny-1, nz/2) when all faces are Neumann
};

struct FaceVals {
    const double *l, *r, *u, *d, *t, *b;
};

// The six face transmissibilities of one cell as the reference stores them
// in perm[s] (boundary faces: Dirichlet 2K/(h^2 mu), Neumann 0, pinned cell
// of an all-Neumann problem 2K/(h^2 mu) on its left face).
struct Faces6 {
    double l, r, u, d, t, b;
    bool il, ir, iu, id, it, ib;   // face is interior (indcol != -1)
};

__device__ __forceinline__ double hm(double a, double b)
{
#pragma clang fp contract(off)
    return 2 * a * b / (a + b);
}

__device__ __forceinline__ Faces6 faces_from_K(const GridArgs &g, const double *Kx, const double *Ky,
                                               const double *Kz, long s, int i, int j, int k)
{
#pragma clang fp contract(off)
    const long nx = g.nx, nxy = (long)g.nx * g.ny;
    const double dx = g.hx * g.hx * g.mu, dy = g.hy * g.hy * g.mu, dz = g.hz * g.hz * g.mu;
    Faces6 f;
    f.ir = i < g.nx - 1; f.il = i > 0;
    f.iu = j < g.ny - 1; f.id = j > 0;
    f.it = k < g.nz - 1; f.ib = k > 0;
    double kx = Kx[s], ky = Ky[s], kz = Kz[s];
    // interior faces: hm(K_s, K_nb) / (h^2 mu)   (permeabilities.c:216-293)
    f.r = f.ir ? hm(kx, Kx[s + 1]) / dx : (g.bc[1] == 0 ? 2.0 * kx / dx : 0.0);
    f.l = f.il ? hm(kx, Kx[s - 1]) / dx : (g.bc[0] == 0 ? 2.0 * kx / dx : 0.0);
    f.u = f.iu ? hm(ky, Ky[s + nx]) / dy : (g.bc[2] == 0 ? 2.0 * ky / dy : 0.0);
    f.d = f.id ? hm(ky, Ky[s - nx]) / dy : (g.bc[3] == 0 ? 2.0 * ky / dy : 0.0);
    f.t = f.it ? hm(kz, Kz[s + nxy]) / dz : (g.bc[4] == 0 ? 2.0 * kz / dz : 0.0);
    f.b = f.ib ? hm(kz, Kz[s - nxy]) / dz : (g.bc[5] == 0 ? 2.0 * kz / dz : 0.0);
    if (g.all_neumann && s == g.pin) f.l = 2.0 * kx / dx;   // permeabilities.c:328-333
    return f;
}

__device__ __forceinline__ Faces6 faces_from_perm(const GridArgs &g, const ELE *perm, const IND *indcol,
                                                  long s, int i, int j, int k)
{
    ELE p = perm[s];
    Faces6 f;
    f.l = p.l; f.r = p.r; f.u = p.u; f.d = p.d; f.t = p.t; f.b = p.b;
    if (indcol) {
        IND c = indcol[s];
        f.il = c.l != -1; f.ir = c.r != -1; f.iu = c.u != -1;
        f.id = c.d != -1; f.it = c.t != -1; f.ib = c.b != -1;
    } else {
        f.il = i > 0; f.ir = i < g.nx - 1; f.iu = j < g.ny - 1;
        f.id = j > 0; f.it = k < g.nz - 1; f.ib = k > 0;
    }
    return f;
}

// b[s] in source_term's order of faces: L, pin, R, D, U, T, B, then sources.
__device__ __forceinline__ double rhs_cell(const GridArgs &g, const Faces6 &f, const FaceVals &B, double b0,
                                           long s, int i, int j, int k, const double *KI, const double *WI,
                                           const double *pc, const double *pw)
{
#pragma clang fp contract(off)
    double b = b0;
    if (i == 0) {
        long l = (long)k * g.ny + j;
        if (g.bc[0] == 0) b = b + (f.l) * B.l[l];
        if (g.bc[0] == 1) b = b + B.l[l] / g.hx;
    }
    if (g.all_neumann && s == g.pin) b = b + (f.l) * 1.0;
    if (i == g.nx - 1) {
        long l = (long)k * g.ny + j;
        if (g.bc[1] == 0) b = b + f.r * B.r[l];
        if (g.bc[1] == 1) b = b - B.r[l] / g.hx;
    }
    if (j == 0) {
        long l = (long)k * g.nx + i;
        if (g.bc[3] == 0) b = b + f.d * B.d[l];
        if (g.bc[3] == 1) b = b + B.d[l] / g.hy;
    }
    if (j == g.ny - 1) {
        long l = (long)k * g.nx + i;
        if (g.bc[2] == 0) b = b + f.u * B.u[l];
        if (g.bc[2] == 1) b = b - B.u[l] / g.hy;
    }
    if (k == g.nz - 1) {
        long l = (long)j * g.nx + i;
        if (g.bc[4] == 0) b = b + f.t * B.t[l];
        if (g.bc[4] == 1) b = b - B.t[l] / g.hz;
    }
    if (k == 0) {
        long l = (long)j * g.nx + i;
        if (g.bc[5] == 0) b = b + f.b * B.b[l];
        if (g.bc[5] == 1) b = b + B.b[l] / g.hz;
    }
    double Ve = g.hx * g.hy * g.hz;
    double ki = KI ? KI[s] : 0.0, wi = WI ? WI[s] : 0.0;
    double pcv = pc ? pc[s] : 0.0, pwv = pw ? pw[s] : 0.0;
    b = b + (ki * pcv) / (g.mu * Ve) + (wi * pwv) / (g.mu * Ve);
    return b;
}

__device__ __forceinline__ void cell_coords(const GridArgs &g, long s, int &i, int &j, int &k)
{
    unsigned nxy = (unsigned)g.nx * (unsigned)g.ny;
    unsigned us = (unsigned)s;
    k = (int)(us / nxy);
    unsigned rem = us - (unsigned)k * nxy;
    j = (int)(rem / (unsigned)g.nx);
    i = (int)(rem - (unsigned)j * (unsigned)g.nx);
}

// permeabilidade_arestas + source_term. Writes the flux-form stencil
// (e, wx, wy, wz), b, and optionally the reference's AoS perm/indcol.
__global__ void __launch_bounds__(kBlock)
k_assemble(GridArgs g, const double *__restrict__ Kx, const double *__restrict__ Ky, const double *__restrict__ Kz,
           const double *KI, const double *WI, const double *pc, const double *pw, FaceVals B,
           double *e, double *wx, double *wy, double *wz, double *b, ELE *perm, IND *indcol)
{
#pragma clang fp contract(off)
    long n = (long)g.nx * g.ny * g.nz;
    long s = (long)xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    int i, j, k;
    cell_coords(g, s, i, j, k);
    Faces6 f = faces_from_K(g, Kx, Ky, Kz, s, i, j, k);
    double vol = g.mu * g.hx * g.hy * g.hz;
    double ki = (KI ? KI[s] : 0.0) / vol;
    double wi = (WI ? WI[s] : 0.0) / vol;
    // permeabilities.c:485
    double c = f.l + f.r + f.u + f.d + f.t + f.b + ki + wi;
    if (e) {
        // diagonal excess: boundary faces and sources only (no cancellation)
        double x = (f.il ? 0.0 : f.l) + (f.ir ? 0.0 : f.r) + (f.iu ? 0.0 : f.u) + (f.id ? 0.0 : f.d)
                 + (f.it ? 0.0 : f.t) + (f.ib ? 0.0 : f.b) + ki + wi;
        e[s] = x;
        wx[s] = f.ir ? f.r : 0.0;
        wy[s] = f.iu ? f.u : 0.0;
        wz[s] = f.it ? f.t : 0.0;
    }
    if (b) b[s] = rhs_cell(g, f, B, 0.0, s, i, j, k, KI, WI, pc, pw);
    if (perm) {
        ELE p;
        p.u = f.u; p.d = f.d; p.l = f.l; p.r = f.r; p.c = c; p.t = f.t; p.b = f.b;
        perm[s] = p;
    }
    if (indcol) {
        long nx = g.nx, nxy = (long)g.nx * g.ny;
        IND q;
        q.c = (int)s;
        q.u = f.iu ? (int)(s + nx) : -1;
        q.d = f.id ? (int)(s - nx) : -1;
        q.l = f.il ? (int)(s - 1) : -1;
        q.r = f.ir ? (int)(s + 1) : -1;
        q.t = f.it ? (int)(s + nxy) : -1;
        q.b = f.ib ? (int)(s - nxy) : -1;
        if (g.all_neumann && s == g.pin) q.l = -1;
        indcol[s] = q;
    }
}

// source_term() from a given AoS perm/indcol; accumulates into b.
__global__ void __launch_bounds__(kBlock)
k_source_from_perm(GridArgs g, const ELE *perm, const IND *indcol, const double *KI, const double *WI,
                   const double *pc, const double *pw, FaceVals B, double *b)
{
    long n = (long)g.nx * g.ny * g.nz;
    long s = (long)xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    int i, j, k;
    cell_coords(g, s, i, j, k);
    Faces6 f = faces_from_perm(g, perm, indcol, s, i, j, k);
    b[s] = rhs_cell(g, f, B, b[s], s, i, j, k, KI, WI, pc, pw);
}

// Structured stencil from AoS perm/indcol (the solve_linsys() input). Also
// checks that the input is the 7-point Cartesian pattern of
// permeabilidade_arestas with symmetric couplings; clears *ok otherwise.
__global__ void __launch_bounds__(kBlock)
k_perm_to_stencil(int nx, int ny, int nz, const ELE *perm, const IND *indcol, double *e, double *wx,
                  double *wy, double *wz, int *ok)
{
    long n = (long)nx * ny * nz, nxy = (long)nx * ny;
    long s = (long)xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    unsigned us = (unsigned)s, unxy = (unsigned)nxy;
    int k = (int)(us / unxy);
    unsigned rem = us - (unsigned)k * unxy;
    int j = (int)(rem / (unsigned)nx), i = (int)(rem - (unsigned)j * (unsigned)nx);
    ELE p = perm[s];
    IND c = indcol[s];
    bool good = c.c == (int)s;
    good &= (i < nx - 1) ? c.r == (int)(s + 1) : c.r == -1;
    good &= (i > 0) ? c.l == (int)(s - 1) : c.l == -1;
    good &= (j < ny - 1) ? c.u == (int)(s + nx) : c.u == -1;
    good &= (j > 0) ? c.d == (int)(s - nx) : c.d == -1;
    good &= (k < nz - 1) ? c.t == (int)(s + nxy) : c.t == -1;
    // the pinned cell of an all-Neumann problem has l = -1 inside i > 0 never; i == 0 always
    good &= (k > 0) ? c.b == (int)(s - nxy) : c.b == -1;
    if (i < nx - 1) good &= perm[s + 1].l == p.r;
    if (j < ny - 1) good &= perm[s + nx].d == p.u;
    if (k < nz - 1) good &= perm[s + nxy].b == p.t;
    if (!good) *ok = 0;
    double sum = p.l + p.r + p.u + p.d + p.t + p.b;
    double bnd = (c.l == -1 ? p.l : 0.0) + (c.r == -1 ? p.r : 0.0) + (c.u == -1 ? p.u : 0.0)
               + (c.d == -1 ? p.d : 0.0) + (c.t == -1 ? p.t : 0.0) + (c.b == -1 ? p.b : 0.0);
    e[s] = bnd + (p.c - sum);
    wx[s] = c.r != -1 ? p.r : 0.0;
    wy[s] = c.u != -1 ? p.u : 0.0;
    wz[s] = c.t != -1 ? p.t : 0.0;
}

// Flux of one cell (flux_reconstruction.c:30-273) and its divergence
// (flux_reconstruction.c:299).
__device__ __forceinline__ double flux_cell_to(const GridArgs &g, const Faces6 &f, const FaceVals &B,
                                               const double *p, long s, int i, int j, int k, double *o)
{
#pragma clang fp contract(off)
    const long nx = g.nx, nxy = (long)g.nx * g.ny;
    double ps = p[s];
    double vu, vd, vl, vr, vt, vb;
    if (j < g.ny - 1) vu = g.hy * f.u * (ps - p[s + nx]);
    else { long l = (long)k * g.nx + i; vu = g.bc[2] == 0 ? g.hy * f.u * (ps - B.u[l]) : B.u[l]; }
    if (i < g.nx - 1) vr = g.hx * f.r * (ps - p[s + 1]);
    else { long l = (long)k * g.ny + j; vr = g.bc[1] == 0 ? g.hx * f.r * (ps - B.r[l]) : B.r[l]; }
    if (j > 0) vd = g.hy * f.d * (ps - p[s - nx]);
    else { long l = (long)k * g.nx + i; vd = g.bc[3] == 0 ? g.hy * f.d * (ps - B.d[l]) : -B.d[l]; }
    if (i > 0) vl = g.hx * f.l * (ps - p[s - 1]);
    else { long l = (long)k * g.ny + j; vl = g.bc[0] == 0 ? g.hx * f.l * (ps - B.l[l]) : -B.l[l]; }
    if (k < g.nz - 1) vt = g.hz * f.t * (ps - p[s + nxy]);
    else { long l = (long)j * g.nx + i; vt = g.bc[4] == 0 ? g.hz * f.t * (ps - B.t[l]) : B.t[l]; }
    if (k > 0) vb = g.hz * f.b * (ps - p[s - nxy]);
    else { long l = (long)j * g.nx + i; vb = g.bc[5] == 0 ? g.hz * f.b * (ps - B.b[l]) : -B.b[l]; }
    if (o) {   // u,d,l,r,c,t,b; c is left untouched
        o[0] = vu; o[1] = vd; o[2] = vl; o[3] = vr; o[5] = vt; o[6] = vb;
    }
    return vr / g.hx + vl / g.hx + vd / g.hy + vu / g.hy + vb / g.hz + vt / g.hz;
}

__device__ __forceinline__ double flux_cell(const GridArgs &g, const Faces6 &f, const FaceVals &B, const double *p,
                                            long s, int i, int j, int k, ELE *v)
{
    return flux_cell_to(g, f, B, p, s, i, j, k, v ? (double *)(v + s) : nullptr);
}

// The block's cells' ELE records, staged in LDS by flux_cell_to, written out
// as consecutive 8-byte values (every lane of a wave on one contiguous run;
// each thread's own record would be six stores 56 bytes apart), c skipped.
__device__ __forceinline__ void flux_store_block(const double *stage, long s0, long n, ELE *v)
{
    const long cnt = min((long)kBlock, n - s0);
    double *out = (double *)(v + s0);
    for (int q = threadIdx.x; q < cnt * 7; q += kBlock)
        if (q % 7 != 4) out[q] = stage[q];
}

// Max/min |div| of a block, written as one partial pair per block; the
// k_minmax_final kernel combines the partials (no inter-block fences).
__device__ void minmax_block(double mx, double mn, double *partial)
{
    __shared__ double smx[kBlock / 64], smn[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        mn = fmin(mn, __shfl_xor(mn, o));
    }
    int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) { smx[w] = mx; smn[w] = mn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; q++) { mx = fmax(mx, smx[q]); mn = fmin(mn, smn[q]); }
        partial[2 * blockIdx.x] = mx;
        partial[2 * blockIdx.x + 1] = mn;
    }
}

// out[0] = min |div|, out[1] = max |div| (divvel_pontual's minn / maxx).
__global__ void __launch_bounds__(kBlock)
k_minmax_final(const double *partial, int G, double *out)
{
    double mx = 0.0, mn = 100.0;
    for (int q = threadIdx.x; q < G; q += kBlock) {
        mx = fmax(mx, partial[2 * q]);
        mn = fmin(mn, partial[2 * q + 1]);
    }
    __shared__ double smx[kBlock / 64], smn[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        mn = fmin(mn, __shfl_xor(mn, o));
    }
    int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) { smx[w] = mx; smn[w] = mn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; q++) { mx = fmax(mx, smx[q]); mn = fmin(mn, smn[q]); }
        out[0] = fabs(mn);
        out[1] = fabs(mx);
    }
}

__global__ void __launch_bounds__(kBlock)
k_flux_from_K(GridArgs g, const double *Kx, const double *Ky, const double *Kz, FaceVals B, const double *p,
              ELE *v, double *partial)
{
    __shared__ double stage[kBlock * 7];
    long n = (long)g.nx * g.ny * g.nz;
    const long s0 = (long)xcd_block() * kBlock;
    long s = s0 + threadIdx.x;
    double mx = 0.0, mn = 100.0;
    if (s < n) {
        int i, j, k;
        cell_coords(g, s, i, j, k);
        Faces6 f = faces_from_K(g, Kx, Ky, Kz, s, i, j, k);
        double dv = fabs(flux_cell_to(g, f, B, p, s, i, j, k, v ? stage + 7 * threadIdx.x : nullptr));
        mx = dv; mn = dv;
    }
    minmax_block(mx, mn, partial);   // (its barrier also orders the LDS writes before the reads)
    if (v && s0 < n) flux_store_block(stage, s0, n, v);
}

__global__ void __launch_bounds__(kBlock)
k_flux_from_perm(GridArgs g, const ELE *perm, FaceVals B, const double *p, ELE *v, double *partial)
{
    long n = (long)g.nx * g.ny * g.nz;
    long s = (long)xcd_block() * kBlock + threadIdx.x;
    double mx = 0.0, mn = 100.0;
    if (s < n) {
        int i, j, k;
        cell_coords(g, s, i, j, k);
        Faces6 f = faces_from_perm(g, perm, nullptr, s, i, j, k);
        double dv = fabs(flux_cell(g, f, B, p, s, i, j, k, v));
        mx = dv; mn = dv;
    }
    minmax_block(mx, mn, partial);
}

// divvel_pontual() of an existing AoS v.
__global__ void __launch_bounds__(kBlock)
k_divergence(GridArgs g, const ELE *v, double *partial)
{
#pragma clang fp contract(off)
    long n = (long)g.nx * g.ny * g.nz;
    long s = (long)xcd_block() * kBlock + threadIdx.x;
    double mx = 0.0, mn = 100.0;
    if (s < n) {
        ELE q = v[s];
        double dv = fabs(q.r / g.hx + q.l / g.hx + q.d / g.hy + q.u / g.hy + q.b / g.hz + q.t / g.hz);
        mx = dv; mn = dv;
    }
    minmax_block(mx, mn, partial);
}

// prod_mat_vec(): res += A x with A from AoS perm/indcol, in the reference's
// order of terms (d, l, c, r, u, t, b).
__global__ void __launch_bounds__(kBlock)
k_prod_mat_vec(long n, const ELE *perm, const IND *indcol, const double *x, double *res)
{
#pragma clang fp contract(off)
    long s = (long)xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    ELE p = perm[s];
    IND c = indcol[s];
    double r = res[s];
    if (c.d != -1) r = r - p.d * x[c.d];
    if (c.l != -1) r = r - p.l * x[c.l];
    if (c.c != -1) r = r + p.c * x[c.c];
    if (c.r != -1) r = r - p.r * x[c.r];
    if (c.u != -1) r = r - p.u * x[c.u];
    if (c.t != -1) r = r - p.t * x[c.t];
    if (c.b != -1) r = r - p.b * x[c.b];
    res[s] = r;
}

static GridArgs make_grid(const hmfem_grid *h)
{
    GridArgs g;
    g.nx = h->nx; g.ny = h->ny; g.nz = h->nz;
    g.hx = h->hx; g.hy = h->hy; g.hz = h->hz; g.mu = h->viscosidade;
    bool alln = true;
    for (int f = 0; f < 6; f++) { g.bc[f] = h->bc_type[f]; alln = alln && h->bc_type[f] == 1; }
    g.all_neumann = alln;
    g.pin = 0 + (long)(h->ny - 1) * h->nx + (long)(h->nz / 2) * h->nx * h->ny;
    return g;
}

static FaceVals make_faces(const double *const bval[6])
{
    FaceVals B;
    B.l = bval[0]; B.r = bval[1]; B.u = bval[2]; B.d = bval[3]; B.t = bval[4]; B.b = bval[5];
    return B;
}

// ------------------------------------------------------------ host launchers

int launch_assemble(hipStream_t st, const hmfem_grid *h, const double *Kx, const double *Ky, const double *Kz,
                    const double *KI, const double *WI, const double *pc, const double *pw,
                    const double *const bval[6], double *e, double *wx, double *wy, double *wz, double *b,
                    ELE *perm, IND *indcol)
{
    GridArgs g = make_grid(h);
    long n = (long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_assemble, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, g, Kx, Ky, Kz, KI, WI, pc, pw,
                       make_faces(bval), e, wx, wy, wz, b, perm, indcol);
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

int launch_source_from_perm(hipStream_t st, const hmfem_grid *h, const ELE *perm, const IND *indcol,
                            const double *KI, const double *WI, const double *pc, const double *pw,
                            const double *const bval[6], double *b)
{
    GridArgs g = make_grid(h);
    long n = (long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_source_from_perm, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, g, perm, indcol, KI, WI,
                       pc, pw, make_faces(bval), b);
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

int launch_perm_to_stencil(hipStream_t st, int nx, int ny, int nz, const ELE *perm, const IND *indcol, double *e,
                           double *wx, double *wy, double *wz, int *ok)
{
    long n = (long)nx * ny * nz;
    hipLaunchKernelGGL(k_perm_to_stencil, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, nx, ny, nz, perm, indcol,
                       e, wx, wy, wz, ok);
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

int launch_flux(hipStream_t st, const hmfem_grid *h, const double *Kx, const double *Ky, const double *Kz,
                const ELE *perm, const double *const bval[6], const double *p, ELE *v, double *partial,
                unsigned int *counter, double *out)
{
    GridArgs g = make_grid(h);
    long n = (long)g.nx * g.ny * g.nz;
    int G = grid_for(n, kBlock);
    if (perm)
        hipLaunchKernelGGL(k_flux_from_perm, dim3(G), dim3(kBlock), 0, st, g, perm, make_faces(bval), p, v, partial);
    else
        hipLaunchKernelGGL(k_flux_from_K, dim3(G), dim3(kBlock), 0, st, g, Kx, Ky, Kz, make_faces(bval), p, v,
                           partial);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(kBlock), 0, st, partial, G, out);
    (void)counter;
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

int launch_divergence(hipStream_t st, const hmfem_grid *h, const ELE *v, double *partial, unsigned int *counter,
                      double *out)
{
    GridArgs g = make_grid(h);
    long n = (long)g.nx * g.ny * g.nz;
    int G = grid_for(n, kBlock);
    hipLaunchKernelGGL(k_divergence, dim3(G), dim3(kBlock), 0, st, g, v, partial);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(kBlock), 0, st, partial, G, out);
    (void)counter;
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

int launch_prod_mat_vec(hipStream_t st, long n, const ELE *perm, const IND *indcol, const double *x, double *res)
{
    hipLaunchKernelGGL(k_prod_mat_vec, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, n, perm, indcol, x, res);
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

// ------------------------------------------------------------ batched variants
//
// A batch concatenates independent grids; instance q occupies cells
// [off, off + n) of every array and blocks [blk0, blk0 + ceil(n / kBlock))
// of a launch, and bmap[block] = q. Blocks never straddle two instances.

struct BGrid {
    GridArgs g;
    int off, blk0, nblk;
};

__global__ void __launch_bounds__(kBlock)
kb_assemble(const BGrid *__restrict__ G, const int *__restrict__ bmap, const double *__restrict__ Kx,
            const double *__restrict__ Ky, const double *__restrict__ Kz, FaceVals B, double *e, double *wx,
            double *wy, double *wz, double *b)
{
#pragma clang fp contract(off)
    const BGrid D = G[bmap[blockIdx.x]];
    const GridArgs &g = D.g;
    long n = (long)g.nx * g.ny * g.nz;
    long s = (long)(blockIdx.x - D.blk0) * kBlock + threadIdx.x;
    if (s >= n) return;
    int i, j, k;
    cell_coords(g, s, i, j, k);
    const double *kx = Kx + D.off, *ky = Ky + D.off, *kz = Kz + D.off;
    Faces6 f = faces_from_K(g, kx, ky, kz, s, i, j, k);
    double x = (f.il ? 0.0 : f.l) + (f.ir ? 0.0 : f.r) + (f.iu ? 0.0 : f.u) + (f.id ? 0.0 : f.d)
             + (f.it ? 0.0 : f.t) + (f.ib ? 0.0 : f.b);
    long o = D.off + s;
    e[o] = x;
    wx[o] = f.ir ? f.r : 0.0;
    wy[o] = f.iu ? f.u : 0.0;
    wz[o] = f.it ? f.t : 0.0;
    b[o] = rhs_cell(g, f, B, 0.0, s, i, j, k, nullptr, nullptr, nullptr, nullptr);
}

__global__ void __launch_bounds__(kBlock)
kb_flux(const BGrid *__restrict__ G, const int *__restrict__ bmap, const double *__restrict__ Kx,
        const double *__restrict__ Ky, const double *__restrict__ Kz, FaceVals B, const double *p, ELE *v,
        double *partial)
{
    const BGrid D = G[bmap[blockIdx.x]];
    const GridArgs &g = D.g;
    long n = (long)g.nx * g.ny * g.nz;
    long s = (long)(blockIdx.x - D.blk0) * kBlock + threadIdx.x;
    double mx = 0.0, mn = 100.0;
    if (s < n) {
        int i, j, k;
        cell_coords(g, s, i, j, k);
        Faces6 f = faces_from_K(g, Kx + D.off, Ky + D.off, Kz + D.off, s, i, j, k);
        double dv = fabs(flux_cell(g, f, B, p + D.off, s, i, j, k, v ? v + D.off : nullptr));
        mx = dv; mn = dv;
    }
    minmax_block(mx, mn, partial);
}

// Per-instance max/min of the flux partials: one block per instance.
__global__ void __launch_bounds__(kBlock)
kb_minmax_final(const BGrid *__restrict__ G, const double *partial, double *out)
{
    const BGrid D = G[blockIdx.x];
    double mx = 0.0, mn = 100.0;
    for (int q = threadIdx.x; q < D.nblk; q += kBlock) {
        mx = fmax(mx, partial[2 * (D.blk0 + q)]);
        mn = fmin(mn, partial[2 * (D.blk0 + q) + 1]);
    }
    __shared__ double smx[kBlock / 64], smn[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        mn = fmin(mn, __shfl_xor(mn, o));
    }
    int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) { smx[w] = mx; smn[w] = mn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; q++) { mx = fmax(mx, smx[q]); mn = fmin(mn, smn[q]); }
        out[2 * blockIdx.x] = fabs(mn);
        out[2 * blockIdx.x + 1] = fabs(mx);
    }
}

// Batch layout of the fine grids: per-instance descriptors and block map on
// the device (in `arena`), uniform boundary types and values.
int batch_grids(hipStream_t st, int ninst, const int *dims, const double *lengths, double mu, const int bc[6],
                std::vector<char> &stage, void *d_grids, int *d_bmap, int *nblk_out)
{
    std::vector<BGrid> G(ninst);
    std::vector<int> bmap;
    long off = 0;
    for (int q = 0; q < ninst; q++) {
        hmfem_grid h;
        h.nx = dims[3 * q]; h.ny = dims[3 * q + 1]; h.nz = dims[3 * q + 2];
        double Lx = lengths ? lengths[3 * q] : 1.0, Ly = lengths ? lengths[3 * q + 1] : 1.0;
        double Lz = lengths ? lengths[3 * q + 2] : 1.0;
        h.hx = Lx / h.nx; h.hy = Ly / h.ny; h.hz = Lz / h.nz;
        h.viscosidade = mu;
        for (int f = 0; f < 6; f++) h.bc_type[f] = bc[f];
        G[q].g = make_grid(&h);
        long n = (long)h.nx * h.ny * h.nz;
        G[q].off = (int)off;
        G[q].blk0 = (int)bmap.size();
        G[q].nblk = (int)((n + kBlock - 1) / kBlock);
        for (int b = 0; b < G[q].nblk; b++) bmap.push_back(q);
        off += n;
    }
    *nblk_out = (int)bmap.size();
    size_t gb = sizeof(BGrid) * ninst, mb = sizeof(int) * bmap.size();
    stage.resize(gb + mb);
    memcpy(stage.data(), G.data(), gb);
    memcpy(stage.data() + gb, bmap.data(), mb);
    HMFEM_CHECK(hipMemcpyAsync(d_grids, stage.data(), gb, hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipMemcpyAsync(d_bmap, stage.data() + gb, mb, hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipStreamSynchronize(st));
    return HMFEM_OK;
}

size_t batch_grid_bytes(int ninst) { return sizeof(BGrid) * (size_t)ninst; }

int launch_batch_assemble(hipStream_t st, const void *d_grids, const int *d_bmap, int nblk, const double *Kx,
                          const double *Ky, const double *Kz, const double *const bval[6], double *e, double *wx,
                          double *wy, double *wz, double *b)
{
    hipLaunchKernelGGL(kb_assemble, dim3(nblk), dim3(kBlock), 0, st, (const BGrid *)d_grids, d_bmap, Kx, Ky, Kz,
                       make_faces(bval), e, wx, wy, wz, b);
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

int launch_batch_flux(hipStream_t st, int ninst, const void *d_grids, const int *d_bmap, int nblk, const double *Kx,
                      const double *Ky, const double *Kz, const double *const bval[6], const double *p, ELE *v,
                      double *partial, double *out)
{
    hipLaunchKernelGGL(kb_flux, dim3(nblk), dim3(kBlock), 0, st, (const BGrid *)d_grids, d_bmap, Kx, Ky, Kz,
                       make_faces(bval), p, v, partial);
    hipLaunchKernelGGL(kb_minmax_final, dim3(ninst), dim3(kBlock), 0, st, (const BGrid *)d_grids, partial, out);
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

} // namespace hmfem
