// AMG hierarchy, V-cycle and flexible CG for the structured 7-point operator.
//
// Reference counterparts (toolbox, header-only C++):
//   amg_solver setup        amg.h:1135-1239  (here: box aggregation + Galerkin sums)
//   multigrid V-cycle       amg.h:1270-1367  (here: damped Jacobi, scaled correction)
//   conjugate_gradient      pcg.h:13-76      (here: flexible CG, device scalars)
//   linear_operator SpMV    generic_operator.h:212-247
//   scalar_product/axpys    algorithm.h:7-324
//
// Hierarchy. Level l+1 aggregates boxes of fx*fy*fz (1 or 2 per axis) cells
// of level l. With piecewise-constant interpolation P the Galerkin product
// P^T A P of a flux-form operator is again a flux-form 7-point operator:
//   e_I = sum_{s in I} e_s,   w_IJ = sum of the fine couplings across I|J,
// so the "RAP" is a summation kernel (no SpGEMM) and every level keeps the
// implicit-index stencil layout. The coarsest level (<= 80 cells) is
// inverted densely in LDS in FP64.
//
// Precision: coefficients are accumulated in FP64; the V-cycle runs on an
// FP32 copy (or FP64 for the plain-FP64 solve); CG vectors, dots, alpha and
// beta are FP64 and stay on the device.
#include "hmfem_internal.h"

namespace hmfem {

// ---------------------------------------------------------------- operator

template <typename T>
struct OpArgs {
    const T *e, *wx, *wy, *wz;
    int nx, ny, nz;
};

template <typename T>
__device__ __forceinline__ OpArgs<T> op_of(const Stencil<T> &s, const Dims &d)
{
    OpArgs<T> o;
    o.e = s.e; o.wx = s.wx; o.wy = s.wy; o.wz = s.wz;
    o.nx = d.nx; o.ny = d.ny; o.nz = d.nz;
    return o;
}

struct Cell {
    int s;
    int i, j, k;
};

template <typename T>
__device__ __forceinline__ Cell cell_at(const OpArgs<T> &o, int s)
{
    Cell c;
    unsigned nxy = (unsigned)o.nx * (unsigned)o.ny;
    c.s = s;
    c.k = (int)((unsigned)s / nxy);
    unsigned rem = (unsigned)s - (unsigned)c.k * nxy;
    c.j = (int)(rem / (unsigned)o.nx);
    c.i = (int)(rem - (unsigned)c.j * (unsigned)o.nx);
    return c;
}

__device__ __forceinline__ Cell shifted(const Cell &c, int di, int dj, int dk, int nx, int nxy)
{
    Cell q;
    q.i = c.i + di; q.j = c.j + dj; q.k = c.k + dk;
    q.s = c.s + di + dj * nx + dk * nxy;
    return q;
}

// Diagonal of A at a cell: e + all six couplings.
template <typename T, typename C>
__device__ __forceinline__ C diag_at(const OpArgs<T> &o, const Cell &c)
{
    const int nx = o.nx, nxy = o.nx * o.ny;
    C d = (C)o.e[c.s] + (C)o.wx[c.s] + (C)o.wy[c.s] + (C)o.wz[c.s];
    if (c.i > 0) d += (C)o.wx[c.s - 1];
    if (c.j > 0) d += (C)o.wy[c.s - nx];
    if (c.k > 0) d += (C)o.wz[c.s - nxy];
    return d;
}

// (A u)_s in flux form and the diagonal, for a functor u(Cell) giving the
// value of u at a cell (so callers can build u on the fly).
template <typename T, typename C, typename U>
__device__ __forceinline__ C apply_at(const OpArgs<T> &o, const Cell &c, U u, C &diag)
{
    const int nx = o.nx, nxy = o.nx * o.ny;
    const int s = c.s;
    C us = u(c);
    C w, d, a;
    d = (C)o.e[s];
    a = d * us;
    if (c.i < o.nx - 1) { w = (C)o.wx[s]; d += w; a += w * (us - u(shifted(c, 1, 0, 0, nx, nxy))); }
    if (c.i > 0) { w = (C)o.wx[s - 1]; d += w; a += w * (us - u(shifted(c, -1, 0, 0, nx, nxy))); }
    if (c.j < o.ny - 1) { w = (C)o.wy[s]; d += w; a += w * (us - u(shifted(c, 0, 1, 0, nx, nxy))); }
    if (c.j > 0) { w = (C)o.wy[s - nx]; d += w; a += w * (us - u(shifted(c, 0, -1, 0, nx, nxy))); }
    if (c.k < o.nz - 1) { w = (C)o.wz[s]; d += w; a += w * (us - u(shifted(c, 0, 0, 1, nx, nxy))); }
    if (c.k > 0) { w = (C)o.wz[s - nxy]; d += w; a += w * (us - u(shifted(c, 0, 0, -1, nx, nxy))); }
    diag = d;
    return a;
}

// ---------------------------------------------------------------- V-cycle

// Pre-smoothing from a zero guess: one sweep gives u1 = w f/d; a second sweep
// u2 = u1 + w (f - A u1)/d evaluates u1 at the neighbours on the fly.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_smooth_zero(OpArgs<T> o, const FT *__restrict__ f, T *__restrict__ u, T omega, int sweeps)
{
    using C = T;
    int n = o.nx * o.ny * o.nz;
    int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    Cell c = cell_at(o, s);
    if (sweeps == 1) {
        C d = diag_at<T, C>(o, c);
        u[s] = omega * (C)f[s] / d;
        return;
    }
    auto u1 = [&](const Cell &q) -> C { return omega * (C)f[q.s] / diag_at<T, C>(o, q); };
    C d;
    C au = apply_at<T, C>(o, c, u1, d);
    C u1s = omega * (C)f[s] / d;
    u[s] = u1s + omega * ((C)f[s] - au) / d;
}

// One damped Jacobi sweep u_out = u + w (f - A u)/d.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_jacobi(OpArgs<T> o, const FT *__restrict__ f, const T *__restrict__ u, T *__restrict__ uo, T omega)
{
    using C = T;
    int n = o.nx * o.ny * o.nz;
    int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    Cell c = cell_at(o, s);
    C d;
    C au = apply_at<T, C>(o, c, [&](const Cell &q) -> C { return (C)u[q.s]; }, d);
    uo[s] = (C)u[s] + omega * ((C)f[s] - au) / d;
}

// Residual r = f - A u summed over each aggregate: one thread per coarse cell.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_resid_restrict(OpArgs<T> o, int fx, int fy, int fz, int cx, int cy, int cz, const FT *__restrict__ f,
                 const T *__restrict__ u, T *__restrict__ fc)
{
    using C = T;
    int nc = cx * cy * cz;
    int I = xcd_block() * kBlock + threadIdx.x;
    if (I >= nc) return;
    int ck = I / (cx * cy);
    int rem = I - ck * cx * cy;
    int cj = rem / cx, ci = rem - cj * cx;
    C acc = 0;
    auto uf = [&](const Cell &q) -> C { return (C)u[q.s]; };
    for (int dk = 0; dk < fz; dk++) {
        int k = ck * fz + dk;
        if (k >= o.nz) break;
        for (int dj = 0; dj < fy; dj++) {
            int j = cj * fy + dj;
            if (j >= o.ny) break;
            for (int di = 0; di < fx; di++) {
                int i = ci * fx + di;
                if (i >= o.nx) break;
                Cell c;
                c.i = i; c.j = j; c.k = k;
                c.s = i + j * o.nx + k * o.nx * o.ny;
                C d;
                C au = apply_at<T, C>(o, c, uf, d);
                acc += (C)f[c.s] - au;
            }
        }
    }
    fc[I] = acc;
}

// Coarse correction u' = u + ocf * u_c[agg] followed by one Jacobi sweep on u'.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_prolong_jacobi(OpArgs<T> o, int fx, int fy, int fz, int cx, int cy, const T *__restrict__ uc, T ocf,
                 const FT *__restrict__ f, const T *__restrict__ u, T *__restrict__ uo, T omega)
{
    using C = T;
    int n = o.nx * o.ny * o.nz;
    int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    Cell c = cell_at(o, s);
    auto up = [&](const Cell &q) -> C {
        int I = q.i / fx + cx * (q.j / fy + cy * (q.k / fz));
        return (C)u[q.s] + ocf * (C)uc[I];
    };
    C d;
    C au = apply_at<T, C>(o, c, up, d);
    uo[s] = up(c) + omega * ((C)f[s] - au) / d;
}

// Dense coarsest solve u = A^-1 f, one block.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_dense_apply(int n, const double *__restrict__ inv, const FT *__restrict__ f, T *__restrict__ u)
{
    __shared__ double fs[kCoarseMax];
    for (int q = threadIdx.x; q < n; q += kBlock) fs[q] = (double)f[q];
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += kBlock) {
        double acc = 0.0;
        const double *row = inv + (long)r * n;
        for (int q = 0; q < n; q++) acc += row[q] * fs[q];
        u[r] = (T)acc;
    }
}

// ---------------------------------------------------------------- setup

// Galerkin sums for box aggregates (FP64), writing FP64 and FP32 copies.
__global__ void __launch_bounds__(kBlock)
k_rap_box(OpArgs<double> o, int fx, int fy, int fz, int cx, int cy, int cz, double *e, double *wx, double *wy,
          double *wz, float *e32, float *wx32, float *wy32, float *wz32)
{
    int nc = cx * cy * cz;
    int I = xcd_block() * kBlock + threadIdx.x;
    if (I >= nc) return;
    int ck = I / (cx * cy);
    int rem = I - ck * cx * cy;
    int cj = rem / cx, ci = rem - cj * cx;
    int i0 = ci * fx, j0 = cj * fy, k0 = ck * fz;
    int i1 = min(i0 + fx, o.nx) - 1, j1 = min(j0 + fy, o.ny) - 1, k1 = min(k0 + fz, o.nz) - 1;
    const int nx = o.nx, nxy = o.nx * o.ny;
    double se = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = k0; k <= k1; k++)
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++) {
                int s = i + j * nx + k * nxy;
                se += o.e[s];
                if (i == i1) sx += o.wx[s];
                if (j == j1) sy += o.wy[s];
                if (k == k1) sz += o.wz[s];
            }
    e[I] = se; wx[I] = sx; wy[I] = sy; wz[I] = sz;
    e32[I] = (float)se; wx32[I] = (float)sx; wy32[I] = (float)sy; wz32[I] = (float)sz;
}

__global__ void __launch_bounds__(kBlock)
k_to_f32(long n, const double *e, const double *wx, const double *wy, const double *wz, float *e32, float *wx32,
         float *wy32, float *wz32)
{
    long s = (long)xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    e32[s] = (float)e[s]; wx32[s] = (float)wx[s]; wy32[s] = (float)wy[s]; wz32[s] = (float)wz[s];
}

// Dense inverse of the coarsest operator by Gauss-Jordan elimination in LDS
// (SPD, no pivoting needed), FP64, one block of 1024 threads.
__global__ void __launch_bounds__(1024)
k_dense_invert(OpArgs<double> o, double *inv, int *info)
{
    extern __shared__ double A[];            // n*n
    __shared__ double colk[kCoarseMax];
    const int n = o.nx * o.ny * o.nz;
    const int nx = o.nx, nxy = o.nx * o.ny;
    for (int q = threadIdx.x; q < n * n; q += blockDim.x) A[q] = 0.0;
    __syncthreads();
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
        Cell c = cell_at(o, s);
        double d = diag_at<double, double>(o, c);
        A[s * n + s] = d;
        if (c.i < o.nx - 1) A[s * n + s + 1] = -o.wx[s];
        if (c.i > 0) A[s * n + s - 1] = -o.wx[s - 1];
        if (c.j < o.ny - 1) A[s * n + s + nx] = -o.wy[s];
        if (c.j > 0) A[s * n + s - nx] = -o.wy[s - nx];
        if (c.k < o.nz - 1) A[s * n + s + nxy] = -o.wz[s];
        if (c.k > 0) A[s * n + s - nxy] = -o.wz[s - nxy];
    }
    __syncthreads();
    for (int k = 0; k < n; k++) {
        double piv = A[k * n + k];
        if (!(piv > 0.0)) { if (threadIdx.x == 0) *info = k + 1; return; }
        for (int i = threadIdx.x; i < n; i += blockDim.x) colk[i] = A[i * n + k];
        __syncthreads();
        double rp = 1.0 / piv;
        for (int j = threadIdx.x; j < n; j += blockDim.x) A[k * n + j] = (j == k ? 1.0 : A[k * n + j]) * rp;
        __syncthreads();
        for (int q = threadIdx.x; q < n * n; q += blockDim.x) {
            int i = q / n, j = q - i * n;
            if (i == k) continue;
            double a = (j == k) ? 0.0 : A[q];
            A[q] = a - colk[i] * A[k * n + j];
        }
        __syncthreads();
    }
    for (int q = threadIdx.x; q < n * n; q += blockDim.x) inv[q] = A[q];
    if (threadIdx.x == 0) *info = 0;
}

// ---------------------------------------------------------------- CG kernels

// Block sum of up to three FP64 values, then the last block to finish sums
// the per-block partials in a fixed order (deterministic) and calls fin().
template <int NV, typename F>
__device__ __forceinline__ void reduce_finish(double (&v)[NV], double *partial, unsigned int *counter, F fin)
{
    __shared__ double sh[kBlock / 64][NV];
    __shared__ int last;
    for (int q = 0; q < NV; q++)
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o);
    int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0)
        for (int q = 0; q < NV; q++) sh[w][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < NV; q++) {
            double t = sh[0][q];
            for (int x = 1; x < kBlock / 64; x++) t += sh[x][q];
            partial[(long)NV * blockIdx.x + q] = t;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        unsigned int prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    double t[NV];
    for (int q = 0; q < NV; q++) t[q] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock)
        for (int q = 0; q < NV; q++) t[q] += partial[(long)NV * b + q];
    for (int q = 0; q < NV; q++)
        for (int o = 32; o > 0; o >>= 1) t[q] += __shfl_xor(t[q], o);
    __syncthreads();
    if (lane == 0)
        for (int q = 0; q < NV; q++) sh[w][q] = t[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < NV; q++) {
            double a = sh[0][q];
            for (int x = 1; x < kBlock / 64; x++) a += sh[x][q];
            t[q] = a;
        }
        *counter = 0;
        fin(t);
    }
}

// r = b - A x, ||b||^2, ||r||^2.
__global__ void __launch_bounds__(kBlock)
k_cg_init(OpArgs<double> o, const double *__restrict__ b, const double *__restrict__ x, double *__restrict__ r,
          CgScalars *sc, double *partial)
{
    int n = o.nx * o.ny * o.nz;
    int base = xcd_block() * kBlock * kCellsPerThread + threadIdx.x;
    double v[2] = {0.0, 0.0};
    for (int q = 0; q < kCellsPerThread; q++) {
        int s = base + q * kBlock;
        if (s >= n) break;
        Cell c = cell_at(o, s);
        double d;
        double ax = apply_at<double, double>(o, c, [&](const Cell &t) { return x[t.s]; }, d);
        double bs = b[s];
        double rs = bs - ax;
        r[s] = rs;
        v[0] += bs * bs;
        v[1] += rs * rs;
    }
    reduce_finish<2>(v, partial, &sc->counter[0], [&](double *t) {
        sc->bb = t[0];
        sc->rr = t[1];
        sc->iter = 0;
        sc->status = 0;
        sc->done = (t[0] == 0.0 || t[1] <= sc->rtol2 * t[0]) ? 1 : 0;
        if (t[0] == 0.0) sc->rr = 0.0;
    });
}

// sigma = <z, r> after the first preconditioning.
template <typename ZT>
__global__ void __launch_bounds__(kBlock)
k_cg_first_dot(long n, const ZT *__restrict__ z, const double *__restrict__ r, CgScalars *sc, double *partial)
{
    if (sc->done) return;
    int base = xcd_block() * kBlock * kCellsPerThread + threadIdx.x;
    double v[1] = {0.0};
    for (int q = 0; q < kCellsPerThread; q++) {
        int s = base + q * kBlock;
        if (s >= n) break;
        v[0] += (double)z[s] * r[s];
    }
    reduce_finish<1>(v, partial, &sc->counter[1], [&](double *t) {
        double sg = fabs(t[0]);
        sc->sigma = sg;
        sc->sigma_first = sg == 0.0 ? 1.0 : sg;
        sc->sigma_last = sg;
        if (sg == 0.0) sc->done = 1;
    });
}

// s_out = z + beta s_in (s_out = z on the first iteration), v = A s_out,
// tau = <s_out, v>, alpha = sigma / tau.
template <typename ZT>
__global__ void __launch_bounds__(kBlock)
k_cg_spmv(OpArgs<double> o, const ZT *__restrict__ z, const double *__restrict__ sin, double *__restrict__ sout,
          double *__restrict__ v, int first, CgScalars *sc, double *partial)
{
    if (sc->done) return;
    int n = o.nx * o.ny * o.nz;
    const double beta = first ? 0.0 : sc->beta;
    int base = xcd_block() * kBlock * kCellsPerThread + threadIdx.x;
    double acc[1] = {0.0};
    auto sv = [&](const Cell &t) -> double { return first ? (double)z[t.s] : (double)z[t.s] + beta * sin[t.s]; };
    for (int q = 0; q < kCellsPerThread; q++) {
        int s = base + q * kBlock;
        if (s >= n) break;
        Cell c = cell_at(o, s);
        double d;
        double av = apply_at<double, double>(o, c, sv, d);
        double ss = sv(c);
        sout[s] = ss;
        v[s] = av;
        acc[0] += ss * av;
    }
    reduce_finish<1>(acc, partial, &sc->counter[2], [&](double *t) {
        double tau = fabs(t[0]);
        sc->tau = tau;
        if (tau == 0.0 || !(tau == tau)) { sc->done = 1; sc->status = HMFEM_ERR_BREAKDOWN; return; }
        sc->alpha = sc->sigma / tau;
    });
}

// x += alpha s, r -= alpha v, rr = ||r||^2; stopping test on the true-residual rule.
__global__ void __launch_bounds__(kBlock)
k_cg_update(long n, const double *__restrict__ s, const double *__restrict__ v, double *__restrict__ x,
            double *__restrict__ r, CgScalars *sc, double *partial)
{
    if (sc->done) return;
    const double alpha = sc->alpha;
    int base = xcd_block() * kBlock * kCellsPerThread + threadIdx.x;
    double acc[1] = {0.0};
    for (int q = 0; q < kCellsPerThread; q++) {
        int i = base + q * kBlock;
        if (i >= n) break;
        x[i] += alpha * s[i];
        double ri = r[i] - alpha * v[i];
        r[i] = ri;
        acc[0] += ri * ri;
    }
    reduce_finish<1>(acc, partial, &sc->counter[3], [&](double *t) {
        sc->rr = t[0];
        sc->iter += 1;
        if (sc->stop_rule == HMFEM_STOP_RESIDUAL && t[0] <= sc->rtol2 * sc->bb) sc->done = 1;
        else if (sc->iter >= sc->maxit) { sc->done = 1; sc->status = HMFEM_NOT_CONVERGED; }
    });
}

// sigma_new = <z, r>, gamma = <z, v>; Polak-Ribiere beta = -alpha gamma / sigma.
template <typename ZT>
__global__ void __launch_bounds__(kBlock)
k_cg_dots(long n, const ZT *__restrict__ z, const double *__restrict__ r, const double *__restrict__ v,
          CgScalars *sc, double *partial)
{
    if (sc->done) return;
    int base = xcd_block() * kBlock * kCellsPerThread + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    for (int q = 0; q < kCellsPerThread; q++) {
        int i = base + q * kBlock;
        if (i >= n) break;
        double zi = (double)z[i];
        acc[0] += zi * r[i];
        acc[1] += zi * v[i];
    }
    reduce_finish<2>(acc, partial, &sc->counter[4], [&](double *t) {
        double sg = fabs(t[0]);
        sc->gamma = t[1];
        sc->beta = -sc->alpha * t[1] / sc->sigma;
        sc->sigma_last = sc->sigma;
        sc->sigma = sg;
        if (sc->stop_rule == HMFEM_STOP_PRECONDITIONED && sg <= sc->rtol2 * sc->sigma_first) sc->done = 1;
        if (!(sg == sg)) { sc->done = 1; sc->status = HMFEM_ERR_BREAKDOWN; }
    });
}

// ---------------------------------------------------------------- host side

struct Hierarchy {
    int nlev = 0;
    Level lev[kMaxLevels];
    double *inv = nullptr;   // dense inverse of the coarsest level
    int nc = 0;
    int prec = HMFEM_PREC_FP32;
};

struct Workspace {
    // allocation sizes to reuse buffers across calls
    size_t bytes = 0;
    char *base = nullptr;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Carve the device memory of one hierarchy out of a single arena.
struct Arena {
    char *p;
    size_t used, cap;
    template <typename T>
    T *take(size_t count)
    {
        size_t b = align256(count * sizeof(T));
        if (used + b > cap) return nullptr;
        T *r = (T *)(p + used);
        used += b;
        return r;
    }
};

static void plan_levels(Dims fine, int coarse_max, Level *lev, int &nlev)
{
    nlev = 0;
    Dims d = fine;
    while (true) {
        Level &L = lev[nlev];
        L = Level();
        L.d = d;
        nlev++;
        if (d.n() <= coarse_max || nlev == kMaxLevels) break;
        L.fx = d.nx > 1 ? 2 : 1;
        L.fy = d.ny > 1 ? 2 : 1;
        L.fz = d.nz > 1 ? 2 : 1;
        Dims c{(d.nx + L.fx - 1) / L.fx, (d.ny + L.fy - 1) / L.fy, (d.nz + L.fz - 1) / L.fz};
        d = c;
    }
}

size_t hierarchy_bytes(Dims fine, int coarse_max, int prec)
{
    Level lev[kMaxLevels];
    int nlev;
    plan_levels(fine, coarse_max, lev, nlev);
    size_t b = 0;
    size_t vs = prec == HMFEM_PREC_FP64 ? 8 : 4;
    for (int l = 0; l < nlev; l++) {
        size_t n = lev[l].d.n();
        if (l > 0) b += 4 * align256(n * 8);            // FP64 coefficients (fine level is the caller's)
        b += 4 * align256(n * 4);                       // FP32 coefficients
        b += 3 * align256(n * vs);                      // f, u, t
    }
    b += align256((size_t)kCoarseMax * kCoarseMax * 8);
    return b + 4096;
}

// Build the hierarchy from a fine FP64 stencil (device), inside `arena`.
int build_hierarchy(hipStream_t st, Dims fine, const Stencil<double> &s0, int coarse_max, int prec, Arena &ar,
                    Hierarchy &H, int *d_info)
{
    if (coarse_max > kCoarseMax) coarse_max = kCoarseMax;
    if (coarse_max < 1) coarse_max = 1;
    plan_levels(fine, coarse_max, H.lev, H.nlev);
    H.prec = prec;
    size_t vs = prec == HMFEM_PREC_FP64 ? 8 : 4;
    for (int l = 0; l < H.nlev; l++) {
        Level &L = H.lev[l];
        size_t n = L.d.n();
        if (l == 0) L.s64 = s0;
        else {
            L.s64.e = ar.take<double>(n); L.s64.wx = ar.take<double>(n);
            L.s64.wy = ar.take<double>(n); L.s64.wz = ar.take<double>(n);
        }
        L.s32.e = ar.take<float>(n); L.s32.wx = ar.take<float>(n);
        L.s32.wy = ar.take<float>(n); L.s32.wz = ar.take<float>(n);
        L.f = ar.take<char>(n * vs); L.u = ar.take<char>(n * vs); L.t = ar.take<char>(n * vs);
        if (!L.s32.wz || !L.t || (l > 0 && !L.s64.wz)) return HMFEM_ERR_NOMEM;
    }
    H.nc = (int)H.lev[H.nlev - 1].d.n();
    if (H.nc > kCoarseMax) return HMFEM_ERR_ARG;   // too many levels would be needed
    H.inv = ar.take<double>((size_t)H.nc * H.nc);
    if (!H.inv) return HMFEM_ERR_NOMEM;

    {
        Level &L = H.lev[0];
        long n = L.d.n();
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, n, L.s64.e, L.s64.wx, L.s64.wy,
                           L.s64.wz, L.s32.e, L.s32.wx, L.s32.wy, L.s32.wz);
    }
    for (int l = 0; l + 1 < H.nlev; l++) {
        Level &L = H.lev[l], &C = H.lev[l + 1];
        OpArgs<double> o;
        o.e = L.s64.e; o.wx = L.s64.wx; o.wy = L.s64.wy; o.wz = L.s64.wz;
        o.nx = L.d.nx; o.ny = L.d.ny; o.nz = L.d.nz;
        long nc = C.d.n();
        hipLaunchKernelGGL(k_rap_box, dim3(grid_for(nc, kBlock)), dim3(kBlock), 0, st, o, L.fx, L.fy, L.fz, C.d.nx,
                           C.d.ny, C.d.nz, C.s64.e, C.s64.wx, C.s64.wy, C.s64.wz, C.s32.e, C.s32.wx, C.s32.wy,
                           C.s32.wz);
    }
    {
        Level &L = H.lev[H.nlev - 1];
        OpArgs<double> o;
        o.e = L.s64.e; o.wx = L.s64.wx; o.wy = L.s64.wy; o.wz = L.s64.wz;
        o.nx = L.d.nx; o.ny = L.d.ny; o.nz = L.d.nz;
        size_t lds = (size_t)H.nc * H.nc * sizeof(double);
        hipLaunchKernelGGL(k_dense_invert, dim3(1), dim3(1024), lds, st, o, H.inv, d_info);
    }
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

template <typename T>
static OpArgs<T> opargs(const Level &L);
template <>
OpArgs<float> opargs<float>(const Level &L)
{
    OpArgs<float> o;
    o.e = L.s32.e; o.wx = L.s32.wx; o.wy = L.s32.wy; o.wz = L.s32.wz;
    o.nx = L.d.nx; o.ny = L.d.ny; o.nz = L.d.nz;
    return o;
}
template <>
OpArgs<double> opargs<double>(const Level &L)
{
    OpArgs<double> o;
    o.e = L.s64.e; o.wx = L.s64.wx; o.wy = L.s64.wy; o.wz = L.s64.wz;
    o.nx = L.d.nx; o.ny = L.d.ny; o.nz = L.d.nz;
    return o;
}

struct CycleParams {
    int nu;
    double omega, ocf;
};

// Optional per-kernel timing: an event pair around each profiled launch.
struct Prof {
    hipEvent_t *ev = nullptr;
    int *type = nullptr;
    int used = 0, cap = 0;
    hipStream_t st = nullptr;
    void begin(int t)
    {
        if (!ev || used + 2 > cap) return;
        type[used / 2] = t;
        (void)hipEventRecord(ev[used], st);
    }
    void end()
    {
        if (!ev || used + 2 > cap) return;
        (void)hipEventRecord(ev[used + 1], st);
        used += 2;
    }
};

#define PROF_LAUNCH(P, T, ...)            \
    do {                                  \
        if (P) (P)->begin(T);             \
        hipLaunchKernelGGL(__VA_ARGS__);  \
        if (P) (P)->end();                \
    } while (0)

// Recursive V-cycle on level l. f0 is the level-0 right-hand side (FP64 CG
// residual) when l == 0; returns the buffer holding u_l.
template <typename T, typename FT>
static T *vcycle_level(hipStream_t st, Hierarchy &H, int l, const FT *f, const CycleParams &cp, Prof *P)
{
    Level &L = H.lev[l];
    long n = L.d.n();
    T *u = (T *)L.u, *t = (T *)L.t;
    const int pc = l == 0 ? 0 : 1;   // profile class: fine level vs coarse levels
    if (l == H.nlev - 1) {
        PROF_LAUNCH(P, HMFEM_PROF_COARSE, (k_dense_apply<T, FT>), dim3(1), dim3(kBlock), 0, st, H.nc, H.inv, f, u);
        return u;
    }
    OpArgs<T> o = opargs<T>(L);
    const T om = (T)cp.omega;
    int G = grid_for(n, kBlock);
    // pre-smoothing from zero
    PROF_LAUNCH(P, pc ? HMFEM_PROF_COARSE : HMFEM_PROF_SMOOTH0, (k_smooth_zero<T, FT>), dim3(G), dim3(kBlock), 0, st,
                o, f, u, om, cp.nu >= 2 ? 2 : 1);
    for (int q = 2; q < cp.nu; q++) {
        PROF_LAUNCH(P, pc ? HMFEM_PROF_COARSE : HMFEM_PROF_JACOBI, (k_jacobi<T, FT>), dim3(G), dim3(kBlock), 0, st, o,
                    f, u, t, om);
        T *tmp = u; u = t; t = tmp;
    }
    // residual and restriction
    Level &C = H.lev[l + 1];
    T *fc = (T *)C.f;
    long nc = C.d.n();
    PROF_LAUNCH(P, pc ? HMFEM_PROF_COARSE : HMFEM_PROF_RESTRICT, (k_resid_restrict<T, FT>), dim3(grid_for(nc, kBlock)),
                dim3(kBlock), 0, st, o, L.fx, L.fy, L.fz, C.d.nx, C.d.ny, C.d.nz, f, u, fc);
    T *uc = vcycle_level<T, T>(st, H, l + 1, fc, cp, P);
    // correction + post-smoothing
    PROF_LAUNCH(P, pc ? HMFEM_PROF_COARSE : HMFEM_PROF_PROLONG, (k_prolong_jacobi<T, FT>), dim3(G), dim3(kBlock), 0,
                st, o, L.fx, L.fy, L.fz, C.d.nx, C.d.ny, uc, (T)cp.ocf, f, u, t, om);
    { T *tmp = u; u = t; t = tmp; }
    for (int q = 1; q < cp.nu; q++) {
        PROF_LAUNCH(P, pc ? HMFEM_PROF_COARSE : HMFEM_PROF_JACOBI, (k_jacobi<T, FT>), dim3(G), dim3(kBlock), 0, st, o,
                    f, u, t, om);
        T *tmp = u; u = t; t = tmp;
    }
    // keep the level's buffers consistent with where the result lives
    L.u = u; L.t = t;
    return u;
}

struct SolveResult {
    int iterations, status;
    double rr, bb, sigma, sigma_first;
};

template <typename T>
static int run_cg(hipStream_t st, Hierarchy &H, const Stencil<double> &A, Dims d, const double *b, double *x,
                  const hmfem_options &opt, CgBuffers &B, CgScalars *h_sc, SolveResult &res, Prof *P)
{
    long n = d.n();
    OpArgs<double> o;
    o.e = A.e; o.wx = A.wx; o.wy = A.wy; o.wz = A.wz;
    o.nx = d.nx; o.ny = d.ny; o.nz = d.nz;
    CycleParams cp{opt.nu < 1 ? 1 : opt.nu, opt.omega, opt.ocf};
    int Gr = grid_for(n, kBlock * kCellsPerThread);

    // scalars: set the constant fields from the host
    CgScalars init{};
    init.rtol2 = opt.rtol * opt.rtol;
    init.maxit = opt.max_iterations;
    init.stop_rule = opt.stop_rule;
    HMFEM_CHECK(hipMemcpyAsync(B.sc, &init, sizeof(CgScalars), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_cg_init, dim3(Gr), dim3(kBlock), 0, st, o, b, x, B.r, B.sc, B.partial);
    T *z = vcycle_level<T, double>(st, H, 0, B.r, cp, P);
    hipLaunchKernelGGL(k_cg_first_dot<T>, dim3(Gr), dim3(kBlock), 0, st, n, (const T *)z, B.r, B.sc, B.partial);
    HMFEM_CHECK(hipGetLastError());

    double *s_in = B.s0, *s_out = B.s1;
    int first = 1;
    int queued = 0;
    const int chunk = opt.check_every < 1 ? 1 : opt.check_every;
    while (true) {
        for (int q = 0; q < chunk; q++) {
            PROF_LAUNCH(P, HMFEM_PROF_CG_SPMV, k_cg_spmv<T>, dim3(Gr), dim3(kBlock), 0, st, o, (const T *)z, s_in,
                        s_out, B.v, first, B.sc, B.partial);
            PROF_LAUNCH(P, HMFEM_PROF_CG_UPDATE, k_cg_update, dim3(Gr), dim3(kBlock), 0, st, n, s_out, B.v, x, B.r,
                        B.sc, B.partial);
            z = vcycle_level<T, double>(st, H, 0, B.r, cp, P);
            PROF_LAUNCH(P, HMFEM_PROF_CG_DOTS, k_cg_dots<T>, dim3(Gr), dim3(kBlock), 0, st, n, (const T *)z, B.r, B.v,
                        B.sc, B.partial);
            double *tmp = s_in; s_in = s_out; s_out = tmp;
            first = 0;
            queued++;
        }
        HMFEM_CHECK(hipGetLastError());
        HMFEM_CHECK(hipMemcpyAsync(h_sc, B.sc, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
        HMFEM_CHECK(hipStreamSynchronize(st));
        if (h_sc->done || queued >= opt.max_iterations + chunk) break;
    }
    res.iterations = h_sc->iter;
    res.status = h_sc->status;
    res.rr = h_sc->rr;
    res.bb = h_sc->bb;
    res.sigma = h_sc->sigma;
    res.sigma_first = h_sc->sigma_first;
    return HMFEM_OK;
}

int solve_structured(hipStream_t st, Dims d, const Stencil<double> &A, const double *b, double *x,
                     const hmfem_options &opt, char *arena, size_t arena_bytes, CgBuffers &B, CgScalars *h_sc,
                     int *d_info, hmfem_status *out, hipEvent_t ev[4], ProfileSink *sink)
{
    Prof prof, *P = nullptr;
    if (sink && sink->ev && sink->cap >= 2) {
        prof.ev = sink->ev; prof.type = sink->type; prof.cap = sink->cap; prof.st = st;
        P = &prof;
    }
    Arena ar{arena, 0, arena_bytes};
    Hierarchy H;
    int prec = opt.precision == HMFEM_PREC_FP64 ? HMFEM_PREC_FP64 : HMFEM_PREC_FP32;
    HMFEM_CHECK(hipEventRecord(ev[0], st));
    int rc = build_hierarchy(st, d, A, opt.coarse_max, prec, ar, H, d_info);
    if (rc) return rc;
    HMFEM_CHECK(hipEventRecord(ev[1], st));
    SolveResult res{};
    if (prec == HMFEM_PREC_FP64) rc = run_cg<double>(st, H, A, d, b, x, opt, B, h_sc, res, P);
    else rc = run_cg<float>(st, H, A, d, b, x, opt, B, h_sc, res, P);
    if (rc) return rc;
    HMFEM_CHECK(hipEventRecord(ev[2], st));
    HMFEM_CHECK(hipEventSynchronize(ev[2]));
    int info = 0;
    HMFEM_CHECK(hipMemcpy(&info, d_info, sizeof(int), hipMemcpyDeviceToHost));
    if (P) {
        for (int q = 0; q + 1 < P->used; q += 2) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, P->ev[q], P->ev[q + 1]);
            int t = P->type[q / 2];
            if (t >= 0 && t < HMFEM_PROF_TYPES) { sink->ms[t] += ms; sink->count[t] += 1; }
        }
        // bytes per launch of the fine-level kernels (algorithmic, actual cells)
        double n0 = (double)d.n();
        double vs = prec == HMFEM_PREC_FP64 ? 8.0 : 4.0;
        double c4 = 4.0 * vs;
        const Level &L1 = H.lev[H.nlev > 1 ? 1 : 0];
        double n1 = (double)L1.d.n();
        sink->bytes[HMFEM_PROF_CG_SPMV] = n0 * (32 + 8 + vs + 8 + 8);        // A64, s_in, z, s_out, v
        sink->bytes[HMFEM_PROF_CG_UPDATE] = n0 * (6 * 8);                    // x, s, r, v in; x, r out
        sink->bytes[HMFEM_PROF_CG_DOTS] = n0 * (vs + 16);                    // z, r, v
        sink->bytes[HMFEM_PROF_SMOOTH0] = n0 * (c4 + 8 + vs);                // A, f, u out
        sink->bytes[HMFEM_PROF_RESTRICT] = n0 * (c4 + 8 + vs) + n1 * vs;     // A, f, u; f_c out
        sink->bytes[HMFEM_PROF_PROLONG] = n0 * (c4 + 8 + 2 * vs) + n1 * vs;  // A, f, u, u_c; u out
        sink->bytes[HMFEM_PROF_JACOBI] = n0 * (c4 + 8 + 2 * vs);             // A, f, u; u out
    }
    if (out) {
        float t0 = 0, t1 = 0;
        (void)hipEventElapsedTime(&t0, ev[0], ev[1]);
        (void)hipEventElapsedTime(&t1, ev[1], ev[2]);
        out->t_setup_ms = t0;
        out->t_solve_ms = t1;
        out->iterations = res.iterations;
        out->levels = H.nlev;
        out->status = info ? HMFEM_ERR_BREAKDOWN : res.status;
        out->rel_residual = res.bb > 0 ? sqrt(res.rr / res.bb) : 0.0;
        out->precond_ratio = res.sigma_first > 0 ? sqrt(res.sigma / res.sigma_first) : 0.0;
        // algorithmic bytes of the solve phase (per cell, per iteration)
        double per = 0.0;
        double vs = prec == HMFEM_PREC_FP64 ? 8.0 : 4.0;
        double cs = vs;  // coefficient bytes
        for (int l = 0; l + 1 < H.nlev; l++) {
            double nl = (double)H.lev[l].d.n();
            double fb = l == 0 ? 8.0 : vs;
            double sweeps = 2.0 * opt.nu;  // pre + post
            // each sweep: 4 coefficients + f + u (read) + u (write)
            per += nl * (sweeps * (4 * cs + fb + 2 * vs) + (4 * cs + fb + vs) + vs);
        }
        double n0 = (double)d.n();
        per += n0 * (32 + 8 + vs + 8 + 8) + n0 * 48 + n0 * (vs + 16);
        out->solve_bytes = per * (res.iterations + 1);
    }
    return HMFEM_OK;
}

size_t cg_partial_count(long n) { return (size_t)2 * grid_for(n, kBlock * kCellsPerThread) + 64; }

} // namespace hmfem
