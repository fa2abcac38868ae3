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
// Cycle. Large levels run one multi-block kernel per step (pre-smoothing,
// residual + restriction, prolongation + smoothing). Levels of at most
// `tail_max` cells run inside a single one-workgroup "tail" kernel that
// executes a host-built list of steps (V- or W-cycle) with a barrier between
// steps, so a whole coarse sub-cycle costs one launch.
//
// Reductions. CG dot products are deterministic and need no inter-block
// fences: each reduction kernel writes one partial per block, and the next
// kernel that needs the value sums the partials in a fixed order itself.
//
// Precision: coefficients are accumulated in FP64; the V-cycle runs on an
// FP32 copy (or FP64 for the plain-FP64 solve); CG vectors, dots, alpha and
// beta are FP64 and stay on the device.
#include "hmfem_gs.h"

#include <string.h>

#include <algorithm>

#include <type_traits>
#include <vector>

namespace hmfem {

// ---------------------------------------------------------------- V-cycle, large levels

// Pre-smoothing from a zero guess: one sweep gives u1 = w f/d; a second sweep
// u2 = u1 + w (f - A u1)/d evaluates u1 at the neighbours on the fly.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_smooth_zero(OpArgs<T> o, const FT *__restrict__ f, T *__restrict__ u, T omega, int sweeps, const int *done)
{
    using C = T;
    if (*done) return;
    int n = o.nx * o.ny * o.nz;
    int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    Cell c = cell_at(o, s);
    const C ds = (C)o.di[s];
    if (sweeps == 1) {
        u[s] = omega * (C)f[s] * ds;
        return;
    }
    auto u1 = [&](const Cell &q) -> C { return omega * (C)f[q.s] * (C)o.di[q.s]; };
    C d;
    C au = apply_at<T, C>(o, c, u1, d);
    C u1s = omega * (C)f[s] * ds;
    u[s] = u1s + omega * ((C)f[s] - au) * ds;
}

// One damped Jacobi sweep u_out = u + w (f - A u)/d.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_jacobi(OpArgs<T> o, const FT *__restrict__ f, const T *__restrict__ u, T *__restrict__ uo, T omega,
         const int *done)
{
    using C = T;
    if (*done) return;
    int n = o.nx * o.ny * o.nz;
    int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    Cell c = cell_at(o, s);
    C d;
    C au = apply_at<T, C>(o, c, [&](const Cell &q) -> C { return (C)u[q.s]; }, d);
    uo[s] = (C)u[s] + omega * ((C)f[s] - au) * rcp(d);
}

// Residual r = f - A u summed over each aggregate: one thread per coarse cell.
// F = 2: all aggregation factors are 2 (loops unrolled); F = 0: run-time factors.
template <typename T, typename FT, int F>
__global__ void __launch_bounds__(kBlock)
k_resid_restrict(OpArgs<T> o, int fx, int fy, int fz, int cx, int cy, int cz, const FT *__restrict__ f,
                 const T *__restrict__ u, T *__restrict__ fc, const int *done)
{
    using C = T;
    if (*done) return;
    if (F) { fx = F; fy = F; fz = F; }
    int nc = cx * cy * cz;
    int I = xcd_block() * kBlock + threadIdx.x;
    if (I >= nc) return;
    int ck = I / (cx * cy);
    int rem = I - ck * cx * cy;
    int cj = rem / cx, ci = rem - cj * cx;
    C acc = 0;
    auto uf = [&](const Cell &q) -> C { return (C)u[q.s]; };
#pragma unroll
    for (int dk = 0; dk < (F ? F : 8); dk++) {
        if (dk >= fz) break;
        int k = ck * fz + dk;
        if (k >= o.nz) break;
#pragma unroll
        for (int dj = 0; dj < (F ? F : 8); dj++) {
            if (dj >= fy) break;
            int j = cj * fy + dj;
            if (j >= o.ny) break;
#pragma unroll
            for (int di = 0; di < (F ? F : 8); di++) {
                if (di >= fx) break;
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

// Residual f - A u summed over 2x2x2 boxes, marching along z: a block owns a
// kGsTx x kGsTy column tile (one thread per column) and the planes
// [kb, kb + zlen) (kb and zlen even). A thread keeps its column's u at k-1, k,
// k+1 and wz at k-1 in registers; the plane's u and its x/y ring go through a
// double-buffered LDS plane (one barrier per plane). The box sums are two
// planes in a register, then the x and y pairs by lane shuffles (lanes t^1,
// t^32), written by the lane of the box's first cell. Same operation order as
// apply_at, with a zero coupling where there is no neighbour.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_restrict_z(OpArgs<T> o, int zlen, int cx, int cy, int cz, const FT *__restrict__ f, const T *__restrict__ u,
             T *__restrict__ fc, const int *done, int eface)
{
    if (*done) return;
    __shared__ T sh[2][kGsTy + 2][kGsTx + 2];
    const int nx = o.nx, ny = o.ny, nz = o.nz;
    const long nxy = (long)nx * ny;
    const int ntx = (nx + kGsTx - 1) / kGsTx, nty = (ny + kGsTy - 1) / kGsTy;
    const int ntxy = ntx * nty, nzc = (nz + zlen - 1) / zlen;
    const int lb = xcd_block();
    if (lb >= ntxy * nzc) return;   // the whole block
    const int tile = lb % ntxy, zc = lb / ntxy;
    const int i0 = (tile % ntx) * kGsTx, j0 = (tile / ntx) * kGsTy;
    const int kb = zc * zlen, ke = min(nz, kb + zlen);
    const int tx = threadIdx.x % kGsTx, ty = threadIdx.x / kGsTx;
    const int i = i0 + tx, j = j0 + ty;
    const bool col = i < nx && j < ny;
    int rx = 0, ry = 0;
    const bool has_ring = threadIdx.x < kGsRing;
    if (has_ring) {
        int hz;
        gs_ring_cell(threadIdx.x, rx, ry, hz);
    }
    const int ri = i0 + rx - 1, rj = j0 + ry - 1;
    const bool ring_in = has_ring && ri >= 0 && rj >= 0 && ri < nx && rj < ny;
    const long cs = i + (long)j * nx, rs = ri + (long)rj * nx;
    // with eface, e is read only on the grid's faces (zero elsewhere)
    const bool eall = !eface || i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
    T um = (T)0, uc = (T)0, up = (T)0, wzm = (T)0;
    if (col) {
        if (kb > 0) { um = u[cs + (kb - 1) * nxy]; wzm = o.wz[cs + (kb - 1) * nxy]; }
        uc = u[cs + kb * nxy];
        if (kb + 1 < nz) up = u[cs + (kb + 1) * nxy];
    }
    T acc = (T)0;
    for (int kk = kb; kk < ke; kk++) {
        const int bf = kk & 1;
        sh[bf][ty + 1][tx + 1] = uc;
        if (has_ring) sh[bf][ry][rx] = ring_in ? u[rs + kk * nxy] : (T)0;
        T r = (T)0, e = (T)0, xp = (T)0, xm = (T)0, yp = (T)0, ym = (T)0, zp = (T)0, fq = (T)0, un = (T)0;
        if (col) {
            const long q = cs + kk * nxy;
            if (eall || kk == 0 || kk == nz - 1) e = o.e[q];
            if (i < nx - 1) xp = o.wx[q];
            if (i > 0) xm = o.wx[q - 1];
            if (j < ny - 1) yp = o.wy[q];
            if (j > 0) ym = o.wy[q - nx];
            if (kk < nz - 1) zp = o.wz[q];
            fq = (T)f[q];
            if (kk + 2 < nz) un = u[q + 2 * nxy];
        }
        __syncthreads();
        if (col) {
            T a = e * uc;
            a += xp * (uc - sh[bf][ty + 1][tx + 2]);
            a += xm * (uc - sh[bf][ty + 1][tx]);
            a += yp * (uc - sh[bf][ty + 2][tx + 1]);
            a += ym * (uc - sh[bf][ty][tx + 1]);
            a += zp * (uc - up);
            a += wzm * (uc - um);
            r = fq - a;
            um = uc; uc = up; up = un; wzm = zp;
        }
        acc += r;
        if (((kk - kb) & 1) || kk == ke - 1) {   // block-uniform
            acc += __shfl_xor(acc, 1);
            acc += __shfl_xor(acc, 32);
            const int ci = i >> 1, cj = j >> 1, ck = kk >> 1;
            if (!(tx & 1) && !(ty & 1) && ci < cx && cj < cy && ck < cz) fc[ci + (long)cx * (cj + (long)cy * ck)] = acc;
            acc = (T)0;
        }
    }
}

// Planes per block of k_restrict_z: even, at most 12, and enough blocks to
// fill the device (2048 when the level allows it). At most 12: 102 us per
// 256^3 launch against 106 at 16 and 8, 107 at 32 and 109 at 24 (at 192^3
// 12 and 16 are equal, at 512^3 12 takes 740 us against 726 for 16).
inline int restrict_zlen(const Dims &d)
{
    const long ntxy = (long)((d.nx + kGsTx - 1) / kGsTx) * ((d.ny + kGsTy - 1) / kGsTy);
    long zl = 12;
    while (zl > 2 && ntxy * ((d.nz + zl - 1) / zl) < 2048) zl -= 2;
    return (int)zl;
}

// Coarse correction u' = u + ocf * u_c[agg] followed by one Jacobi sweep on u'.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_prolong_jacobi(OpArgs<T> o, int fx, int fy, int fz, int cx, int cy, const T *__restrict__ uc, T ocf,
                 const FT *__restrict__ f, const T *__restrict__ u, T *__restrict__ uo, T omega, const int *done)
{
    using C = T;
    if (*done) return;
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
    uo[s] = up(c) + omega * ((C)f[s] - au) * rcp(d);
}

// Last smoothing sweep of the fine level fused with the CG dot products:
// z = u_out, partials of <z, r> and <z, v> (r is the fine right-hand side f).
// With uc != nullptr the sweep also applies the coarse correction first.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_final_sweep_dots(OpArgs<T> o, int fx, int fy, int cx, int cy, int fz, const T *__restrict__ uc, T ocf,
                   const double *__restrict__ f, const T *__restrict__ u, T *__restrict__ uo, T omega,
                   const double *__restrict__ v, int chunk, double *Pdot, const int *done)
{
    using C = T;
    if (*done) return;
    int n = o.nx * o.ny * o.nz;
    int lb, beg, end;
    chunk_range(n, chunk, lb, beg, end);
    double acc[2] = {0.0, 0.0};
    auto up = [&](const Cell &q) -> C {
        C x = (C)u[q.s];
        if (uc) x += ocf * (C)uc[q.i / fx + cx * (q.j / fy + cy * (q.k / fz))];
        return x;
    };
    for (int base = beg; base < end; base += kUnroll * kBlock) {
        C zz[kUnroll];
        double ff[kUnroll], vv[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; q++) {
            int s = base + q * kBlock + threadIdx.x;
            zz[q] = 0; ff[q] = 0.0; vv[q] = 0.0;
            if (s < end) {
                Cell c = cell_at(o, s);
                C d;
                C au = apply_at<T, C>(o, c, up, d);
                ff[q] = f[s];
                if (v) vv[q] = v[s];
                zz[q] = up(c) + omega * ((C)ff[q] - au) * rcp(d);
            }
        }
#pragma unroll
        for (int q = 0; q < kUnroll; q++) {
            int s = base + q * kBlock + threadIdx.x;
            if (s < end) {
                uo[s] = zz[q];
                acc[0] += (double)zz[q] * ff[q];
                acc[1] += (double)zz[q] * vv[q];
            }
        }
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) { Pdot[2 * lb] = acc[0]; Pdot[2 * lb + 1] = acc[1]; }
}

// Diagonal (Jacobi) preconditioner in FP64, z = r / diag(A), with the dot
// partials <z, r>, <z, v>: the plain FP64 PCG (no hierarchy).
__global__ void __launch_bounds__(kBlock)
k_jacobi_pc(OpArgs<double> o, const double *__restrict__ r, const double *__restrict__ v, double *__restrict__ z,
            int chunk, double *Pdot, const int *done)
{
    if (*done) return;
    const int n = o.nx * o.ny * o.nz;
    int lb, beg, end;
    chunk_range(n, chunk, lb, beg, end);
    double acc[2] = {0.0, 0.0};
    for (int s = beg + threadIdx.x; s < end; s += kBlock) {
        Cell c = cell_at(o, s);
        const double rs = r[s];
        const double zs = rs / diag_at<double, double>(o, c);
        z[s] = zs;
        acc[0] += zs * rs;
        if (v) acc[1] += zs * v[s];
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) { Pdot[2 * lb] = acc[0]; Pdot[2 * lb + 1] = acc[1]; }
}

// ---------------------------------------------------------------- block Gauss-Seidel smoother
//
// Each (logical) block smooths tiles [lb * tpb, (lb + 1) * tpb) of the level.
template <typename T, typename FT, int MODE, bool H16 = false>
__global__ void __launch_bounds__(kGsThreads)
k_gs_tile(OpArgs<T> o, int sweeps, const FT *__restrict__ f, const T *__restrict__ u, T *__restrict__ uo, int sx,
          int sy, int sz, int cx, int cy, const T *__restrict__ uc, T ocf, const double *__restrict__ v, int tpb,
          double *Pdot, const int *done, Fine16 h = Fine16{})
{
    if (*done) return;
    const int ntiles = gs_tiles(Dims{o.nx, o.ny, o.nz});
    const int lb = xcd_block();
    double acc[2] = {0.0, 0.0};
    // tiles lb, lb + G, lb + 2G, ...: the blocks work on G consecutive tiles
    // at a time, so a tile's ring is read while its neighbours' blocks have
    // the same lines in L2 (contiguous chunks of tpb tiles per block had 1.3x
    // the algorithmic HBM traffic at 256^3)
    (void)tpb;
    for (int t = lb; t < ntiles; t += gridDim.x)
        gs_tile_body<T, FT, MODE, H16>(o, sweeps, f, u, uo, sx, sy, sz, cx, cy, uc, ocf, v, t, acc, h);
    if (MODE == 2) {
        block_sum_n<2, kGsThreads>(acc);
        if (threadIdx.x == 0) { Pdot[2 * lb] = acc[0]; Pdot[2 * lb + 1] = acc[1]; }
    }
}

// u = A^-1 f for the dense coarsest level: groups of L lanes (a power of two
// dividing 64) share one row and combine their partial sums by shuffles.
template <typename T>
__device__ __forceinline__ void dense_gemv(const double *inv, const double *fs, T *u, int nc, int tid, int nthreads)
{
    int L = 16;
    while (L > 1 && L * nc > nthreads) L >>= 1;
    const int row = tid / L, lane = tid % L;
    double acc = 0.0;
    if (row < nc) {
        const double *r = inv + (long)row * nc;
        for (int j = lane; j < nc; j += L) acc += r[j] * fs[j];
    }
    for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (row < nc && lane == 0) u[row] = (T)acc;
}

// Dense coarsest solve u = A^-1 f, one block.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_dense_apply(int n, const double *__restrict__ inv, const FT *__restrict__ f, T *__restrict__ u, const int *done)
{
    __shared__ double fs[kCoarseMax];
    if (*done) return;
    for (int q = threadIdx.x; q < n; q += kBlock) fs[q] = (double)f[q];
    __syncthreads();
    dense_gemv<T>(inv, fs, u, n, threadIdx.x, kBlock);
}

// Dense solve with the CG dots, for a single-level hierarchy: block 0 does the
// work, the other blocks of the chunked launch write zero partials.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_dense_dots(int n, const double *__restrict__ inv, const double *__restrict__ f, const double *__restrict__ v,
             T *__restrict__ z, double *Pdot, const int *done)
{
    __shared__ double fs[kCoarseMax];
    if (*done) return;
    double acc[2] = {0.0, 0.0};
    const int lb = xcd_block();
    if (lb == 0) {
        for (int q = threadIdx.x; q < n; q += kBlock) fs[q] = f[q];
    }
    __syncthreads();
    if (lb == 0) {
        for (int r = threadIdx.x; r < n; r += kBlock) {
            double a = 0.0;
            const double *row = inv + (long)r * n;
            for (int q = 0; q < n; q++) a += row[q] * fs[q];
            T zr = (T)a;
            z[r] = zr;
            acc[0] += (double)zr * fs[r];
            if (v) acc[1] += (double)zr * v[r];
        }
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) { Pdot[2 * lb] = acc[0]; Pdot[2 * lb + 1] = acc[1]; }
}

// ---------------------------------------------------------------- tail kernel

enum TailOp { OP_SMOOTH0 = 0, OP_JACOBI, OP_RESTRICT, OP_PROLONG_JACOBI, OP_DENSE, OP_RESID, OP_SAVE, OP_ADDSAVED,
              OP_FINAL };


// Executes ops[0..nops) in order, each over all cells of its level with the
// 1024 threads of one workgroup, and a barrier after each op. Every thread
// runs the same op list, so the per-level "which buffer holds u" parity is
// kept in a register and agrees across threads.
template <typename T>
__global__ void __launch_bounds__(kTailThreads)
k_tail(const TailLevel<T> *__restrict__ Lg, const int4 *__restrict__ ops, int nops, const double *__restrict__ invg,
       int nc, T omega, T ocf, int nlev, int lds_inv_off, const int *done)
{
    using C = T;
    extern __shared__ char lds[];
    __shared__ TailLevel<T> Ls[kMaxTailLevels];
    __shared__ double fs[kCoarseMax];
    if (*done) return;
    const int tid = threadIdx.x;
    // LDS mode: copy coefficients, the top level's rhs and the dense inverse
    // into LDS and run every op on the copies.
    const bool in_lds = lds_inv_off >= 0;
    const TailLevel<T> *Lv = Lg;
    const double *inv = invg;
    if (in_lds) {
        if (tid < nlev) {
            TailLevel<T> L = Lg[tid];
            T *b = (T *)(lds + L.lds_off);
            const int n = L.n;
            TailLevel<T> M = L;
            M.o.e = b; M.o.wx = b + n; M.o.wy = b + 2 * n; M.o.wz = b + 3 * n; M.o.di = b + 4 * n;
            M.f = b + 5 * n; M.u = b + 6 * n; M.t = b + 7 * n; M.r = b + 8 * n; M.w = b + 9 * n;
            Ls[tid] = M;
        }
        for (int l = 0; l < nlev; l++) {
            const TailLevel<T> L = Lg[l];
            T *b = (T *)(lds + L.lds_off);
            const int n = L.n;
            for (int s = tid; s < n; s += kTailThreads) {
                b[s] = L.o.e[s]; b[n + s] = L.o.wx[s]; b[2 * n + s] = L.o.wy[s];
                b[3 * n + s] = L.o.wz[s]; b[4 * n + s] = L.o.di[s];
                if (l == 0) b[5 * n + s] = L.f[s];
            }
        }
        double *ilds = (double *)(lds + lds_inv_off);
        for (int q = tid; q < nc * nc; q += kTailThreads) ilds[q] = invg[q];
        __syncthreads();
        Lv = Ls;
        inv = ilds;
    }
    unsigned parity = 0;
    for (int q = 0; q < nops; q++) {
        const int4 op = ops[q];
        const int l = op.y;
        const TailLevel<T> L = Lv[l];
        const bool pl = (parity >> l) & 1u;
        T *cur = pl ? L.t : L.u;
        T *oth = pl ? L.u : L.t;
        const T *f = op.z ? L.r : L.f;
        const OpArgs<T> &o = L.o;
        switch (op.x) {
        case OP_SMOOTH0:
            for (int s = tid; s < L.n; s += kTailThreads) cur[s] = omega * (C)f[s] * (C)o.di[s];
            break;
        case OP_JACOBI:
            for (int s = tid; s < L.n; s += kTailThreads) {
                Cell c = cell_at(o, s);
                C d;
                C au = apply_at<T, C>(o, c, [&](const Cell &x) -> C { return (C)cur[x.s]; }, d);
                oth[s] = (C)cur[s] + omega * ((C)f[s] - au) * rcp(d);
            }
            parity ^= 1u << l;
            break;
        case OP_RESID:
            for (int s = tid; s < L.n; s += kTailThreads) {
                Cell c = cell_at(o, s);
                C d;
                C au = apply_at<T, C>(o, c, [&](const Cell &x) -> C { return (C)cur[x.s]; }, d);
                L.r[s] = (C)L.f[s] - au;
            }
            break;
        case OP_SAVE:
            for (int s = tid; s < L.n; s += kTailThreads) L.w[s] = cur[s];
            break;
        case OP_ADDSAVED:
            for (int s = tid; s < L.n; s += kTailThreads) cur[s] += L.w[s];
            break;
        case OP_FINAL:
            if (pl) {
                for (int s = tid; s < L.n; s += kTailThreads) L.u[s] = L.t[s];
                parity &= ~(1u << l);
            }
            break;
        case OP_RESTRICT: {
            const TailLevel<T> Cn = Lv[l + 1];
            T *dst = op.w ? Cn.r : Cn.f;
            const int cx = Cn.o.nx, cy = Cn.o.ny;
            for (int I = tid; I < Cn.n; I += kTailThreads) {
                int ck = I / (cx * cy);
                int rem = I - ck * cx * cy;
                int cj = rem / cx, ci = rem - cj * cx;
                C acc = 0;
                for (int dk = 0; dk < L.fz; dk++) {
                    int k = ck * L.fz + dk;
                    if (k >= o.nz) break;
                    for (int dj = 0; dj < L.fy; dj++) {
                        int j = cj * L.fy + dj;
                        if (j >= o.ny) break;
                        for (int di = 0; di < L.fx; di++) {
                            int i = ci * L.fx + di;
                            if (i >= o.nx) break;
                            Cell c;
                            c.i = i; c.j = j; c.k = k;
                            c.s = i + j * o.nx + k * o.nx * o.ny;
                            C d;
                            C au = apply_at<T, C>(o, c, [&](const Cell &x) -> C { return (C)cur[x.s]; }, d);
                            acc += (C)f[c.s] - au;
                        }
                    }
                }
                dst[I] = acc;
            }
            break;
        }
        case OP_PROLONG_JACOBI: {
            const TailLevel<T> Cn = Lv[l + 1];
            const T *uc = ((parity >> (l + 1)) & 1u) ? Cn.t : Cn.u;
            const int cx = Cn.o.nx, cy = Cn.o.ny;
            const int fx = L.fx, fy = L.fy, fz = L.fz;
            for (int s = tid; s < L.n; s += kTailThreads) {
                Cell c = cell_at(o, s);
                auto up = [&](const Cell &x) -> C {
                    return (C)cur[x.s] + ocf * (C)uc[x.i / fx + cx * (x.j / fy + cy * (x.k / fz))];
                };
                C d;
                C au = apply_at<T, C>(o, c, up, d);
                oth[s] = up(c) + omega * ((C)f[s] - au) * rcp(d);
            }
            parity ^= 1u << l;
            break;
        }
        case OP_DENSE:
            for (int s = tid; s < nc; s += kTailThreads) fs[s] = (double)f[s];
            __syncthreads();
            dense_gemv<T>(inv, fs, cur, nc, tid, kTailThreads);
            break;
        default:
            break;
        }
        __syncthreads();
    }
    if (in_lds) {
        const TailLevel<T> G0 = Lg[0];
        const T *src = Lv[0].u;   // OP_FINAL left the result in u
        for (int s = tid; s < G0.n; s += kTailThreads) G0.u[s] = src[s];
    }
}

// The same op list with every level resident in LDS: levels are addressed by
// element offsets from the LDS arena (so every access is a direct LDS access),
// and the op list is copied into LDS first. Layout of level l at
// A + base_l: e, wx, wy, wz, di, f, u, t, r, w (n_l values each).

template <typename T>
__global__ void __launch_bounds__(kTailThreads)
k_tail_lds(const TailLevel<T> *__restrict__ Lg, const int4 *__restrict__ opsg, int nops,
           const double *__restrict__ invg, int nc, T omega, T ocf, int nlev, int inv_off, const T *ftop, T *utop,
           int stride, const int *done)
{
    using C = T;
    extern __shared__ char lds[];
    __shared__ int dsc[kMaxTailLevels][8];   // base, n, nx, ny, nz, fx, fy, fz
    __shared__ int4 ops[kTailLdsOps];
    __shared__ double fs[kCoarseMax];
    if (*done) return;
    T *A = (T *)lds;
    double *inv_l = inv_off >= 0 ? (double *)(lds + inv_off) : nullptr;
    const double *inv = inv_l ? inv_l : invg;   // kTailInvGlobal: read from global memory
    const int tid = threadIdx.x;
    if (tid < nlev) {
        const TailLevel<T> &L = Lg[tid];
        dsc[tid][0] = L.lds_off / (int)sizeof(T); dsc[tid][1] = L.n;
        dsc[tid][2] = L.o.nx; dsc[tid][3] = L.o.ny; dsc[tid][4] = L.o.nz;
        dsc[tid][5] = L.fx; dsc[tid][6] = L.fy; dsc[tid][7] = L.fz;
    }
    for (int q = tid; q < nops; q += kTailThreads) ops[q] = opsg[q];
    if (inv_l)
        for (int q = tid; q < nc * nc; q += kTailThreads) inv_l[q] = invg[q];
    __syncthreads();
    for (int l = 0; l < nlev; l++) {
        const TailLevel<T> &L = Lg[l];
        T *b = A + dsc[l][0];
        const int n = dsc[l][1];
        for (int s = tid; s < n; s += kTailThreads) {
            b[s] = L.o.e[s]; b[n + s] = L.o.wx[s]; b[2 * n + s] = L.o.wy[s];
            b[3 * n + s] = L.o.wz[s]; b[4 * n + s] = L.o.di[s];
            if (l == 0) b[5 * n + s] = ftop ? ftop[(long)s * stride] : L.f[s];
        }
    }
    __syncthreads();

    auto level = [&](int l, OpArgs<T> &o, T *&f, T *&u, T *&t, T *&r, T *&w, int &n) {
        const int base = dsc[l][0];
        n = dsc[l][1];
        o.e = A + base; o.wx = A + base + n; o.wy = A + base + 2 * n; o.wz = A + base + 3 * n;
        o.di = A + base + 4 * n;
        o.nx = dsc[l][2]; o.ny = dsc[l][3]; o.nz = dsc[l][4];
        f = A + base + 5 * n; u = A + base + 6 * n; t = A + base + 7 * n; r = A + base + 8 * n;
        w = A + base + 9 * n;
    };
    unsigned parity = 0;
    for (int q = 0; q < nops; q++) {
        const int4 op = ops[q];
        const int l = op.y;
        OpArgs<T> o;
        T *Lf, *Lu, *Lt, *Lr, *Lw;
        int n;
        level(l, o, Lf, Lu, Lt, Lr, Lw, n);
        const bool pl = (parity >> l) & 1u;
        T *cur = pl ? Lt : Lu;
        T *oth = pl ? Lu : Lt;
        const T *f = op.z ? Lr : Lf;
        switch (op.x) {
        case OP_SMOOTH0:
            for (int s = tid; s < n; s += kTailThreads) cur[s] = omega * (C)f[s] * (C)o.di[s];
            break;
        case OP_JACOBI:
            for (int s = tid; s < n; s += kTailThreads) {
                Cell c = cell_at(o, s);
                C d;
                C au = apply_at<T, C>(o, c, [&](const Cell &x) -> C { return (C)cur[x.s]; }, d);
                oth[s] = (C)cur[s] + omega * ((C)f[s] - au) * rcp(d);
            }
            parity ^= 1u << l;
            break;
        case OP_RESID:
            for (int s = tid; s < n; s += kTailThreads) {
                Cell c = cell_at(o, s);
                C d;
                C au = apply_at<T, C>(o, c, [&](const Cell &x) -> C { return (C)cur[x.s]; }, d);
                Lr[s] = (C)Lf[s] - au;
            }
            break;
        case OP_SAVE:
            for (int s = tid; s < n; s += kTailThreads) Lw[s] = cur[s];
            break;
        case OP_ADDSAVED:
            for (int s = tid; s < n; s += kTailThreads) cur[s] += Lw[s];
            break;
        case OP_FINAL:
            if (pl) {
                for (int s = tid; s < n; s += kTailThreads) Lu[s] = Lt[s];
                parity &= ~(1u << l);
            }
            break;
        case OP_RESTRICT: {
            OpArgs<T> oc;
            T *Cf, *Cu, *Ct, *Cr, *Cw;
            int ncl;
            level(l + 1, oc, Cf, Cu, Ct, Cr, Cw, ncl);
            T *dst = op.w ? Cr : Cf;
            const int cx = oc.nx, cy = oc.ny;
            const int fx = dsc[l][5], fy = dsc[l][6], fz = dsc[l][7];
            for (int I = tid; I < ncl; I += kTailThreads) {
                int ck = I / (cx * cy);
                int rem = I - ck * cx * cy;
                int cj = rem / cx, ci = rem - cj * cx;
                C acc = 0;
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
                            C au = apply_at<T, C>(o, c, [&](const Cell &x) -> C { return (C)cur[x.s]; }, d);
                            acc += (C)f[c.s] - au;
                        }
                    }
                }
                dst[I] = acc;
            }
            break;
        }
        case OP_PROLONG_JACOBI: {
            OpArgs<T> oc;
            T *Cf, *Cu, *Ct, *Cr, *Cw;
            int ncl;
            level(l + 1, oc, Cf, Cu, Ct, Cr, Cw, ncl);
            const T *uc = ((parity >> (l + 1)) & 1u) ? Ct : Cu;
            const int cx = oc.nx, cy = oc.ny;
            const int fx = dsc[l][5], fy = dsc[l][6], fz = dsc[l][7];
            for (int s = tid; s < n; s += kTailThreads) {
                Cell c = cell_at(o, s);
                auto up = [&](const Cell &x) -> C {
                    return (C)cur[x.s] + ocf * (C)uc[x.i / fx + cx * (x.j / fy + cy * (x.k / fz))];
                };
                C d;
                C au = apply_at<T, C>(o, c, up, d);
                oth[s] = up(c) + omega * ((C)f[s] - au) * rcp(d);
            }
            parity ^= 1u << l;
            break;
        }
        case OP_DENSE:
            for (int s = tid; s < nc; s += kTailThreads) fs[s] = (double)f[s];
            __syncthreads();
            dense_gemv<T>(inv, fs, cur, nc, tid, kTailThreads);
            break;
        default:
            break;
        }
        __syncthreads();
    }
    {
        const TailLevel<T> &G0 = Lg[0];
        const T *src = A + dsc[0][0] + 6 * dsc[0][1];   // OP_FINAL left the result in u
        for (int s = tid; s < dsc[0][1]; s += kTailThreads) {
            if (utop) utop[(long)s * stride] = src[s];
            else G0.u[s] = src[s];
        }
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

// The fine level's FP32 copy and its inverse diagonal in one pass.
__global__ void __launch_bounds__(kBlock)
k_to_f32_dinv(OpArgs<double> o, float *e32, float *wx32, float *wy32, float *wz32, float *di32)
{
    const int n = o.nx * o.ny * o.nz;
    const int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    to_f32_dinv_cell(o, cell_at(o, s), e32, wx32, wy32, wz32, di32);
}

// The same pass, which also takes the conduit correction's histogram of the
// FP64 diagonal's bins (Deflation::kHistBins) and keeps each cell's bin.
// Block b covers cells [b kBlock kConvCells, (b + 1) kBlock kConvCells);
// runs of equal bins are counted in a register, then summed in LDS, and the
// block adds its non-empty bins to the global ones (few global atomics on
// the crowded bins). The cells are taken kConvBatch at a time, all their
// loads first.
constexpr int kConvCells = 16;
constexpr int kConvBatch = 4;

__global__ void __launch_bounds__(kBlock)
k_to_f32_dinv_hist(OpArgs<double> o, float *e32, float *wx32, float *wy32, float *wz32, float *di32,
                   unsigned *hist, unsigned short *bins)
{
    constexpr int NB = Deflation::kHistBins;
    __shared__ unsigned h[NB];
    for (int t = threadIdx.x; t < NB; t += kBlock) h[t] = 0;
    __syncthreads();
    const int n = o.nx * o.ny * o.nz;
    const int s0 = blockIdx.x * kBlock * kConvCells + threadIdx.x;
    int rb = 0;
    unsigned run = 0;
    for (int t0 = 0; t0 < kConvCells; t0 += kConvBatch) {
        Cell c[kConvBatch];
        ConvLoads L[kConvBatch];
#pragma unroll
        for (int u = 0; u < kConvBatch; u++) {
            const int s = s0 + (t0 + u) * kBlock;
            c[u] = cell_at(o, s < n ? s : n - 1);
            L[u] = conv_load(o, c[u]);
        }
#pragma unroll
        for (int u = 0; u < kConvBatch; u++) {
            if (s0 + (t0 + u) * kBlock >= n) break;
            const int b = conv_store(L[u], c[u], e32, wx32, wy32, wz32, di32, bins, NB);
            if (b != rb) {
                if (run) atomicAdd(&h[rb], run);
                rb = b;
                run = 0;
            }
            run++;
        }
    }
    if (run) atomicAdd(&h[rb], run);
    __syncthreads();
    for (int t = threadIdx.x; t < NB; t += kBlock)
        if (h[t]) atomicAdd(&hist[t], h[t]);
}

// sq = sqrt(d) of the fine level from the FP64 stencil.
__global__ void __launch_bounds__(kBlock)
k_fine16_scale(OpArgs<double> o, float *sq)
{
    const int n = o.nx * o.ny * o.nz;
    const int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    Cell c = cell_at(o, s);
    sq[s] = (float)sqrt(diag_at<double, double>(o, c));
}

// Scaled FP16 couplings w / sqrt(d_s d_nb) towards +x, +y, +z.
__global__ void __launch_bounds__(kBlock)
k_fine16_couplings(OpArgs<double> o, const float *__restrict__ sq, __half *hx, __half *hy, __half *hz)
{
    const int n = o.nx * o.ny * o.nz;
    const int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    Cell c = cell_at(o, s);
    const int nx = o.nx, nxy = o.nx * o.ny;
    const double ss = sq[s];
    hx[s] = __float2half((float)(c.i < o.nx - 1 ? o.wx[s] / (ss * (double)sq[s + 1]) : 0.0));
    hy[s] = __float2half((float)(c.j < o.ny - 1 ? o.wy[s] / (ss * (double)sq[s + nx]) : 0.0));
    hz[s] = __float2half((float)(c.k < o.nz - 1 ? o.wz[s] / (ss * (double)sq[s + nxy]) : 0.0));
}

// Inverse diagonal of a level, in its storage type.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_dinv(OpArgs<T> o, T *di)
{
    int n = o.nx * o.ny * o.nz;
    int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    Cell c = cell_at(o, s);
    di[s] = (T)(1.0 / (double)diag_at<T, double>(o, c));
}

// Dense inverse of the coarsest operator (SPD, n <= kCoarseMax) by block
// Gauss-Jordan elimination in LDS, FP64, one block of 1024 threads (16
// waves). The matrix is padded with the identity to N = 16 ceil(n / 16);
// for each 16-wide block column k: the diagonal block is inverted in place
// (scalar Gauss-Jordan, P = A_kk^-1), then
//   A_kj <- P A_kj, A_ij <- A_ij - A_ik A_kj, A_ik <- -A_ik P  (i, j != k)
// as 16x16 tile products on the matrix cores (v_mfma_f64_16x16x4f64, four
// K-steps of 4 per tile), one tile per wave at a time. At the end A holds A^-1.
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

// C + s X Y for 16x16 tiles X at (xr, xc) and Y at (yr, yc) of the N x N
// matrix M in LDS. Lane l supplies X[l % 16][k], Y[k][l % 16] with
// k = 4 kk + l / 16, and holds C[4 q + l / 16][l % 16] in register q (the
// instruction's lane layout, measured on the device).
__device__ __forceinline__ mfma_d4 tile_mma(const double *M, int N, int xr, int xc, int yr, int yc, mfma_d4 c,
                                            double s)
{
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
        const int k = 4 * kk + (l >> 4);
        const double a = s * M[(xr + (l & 15)) * N + xc + k];
        const double b = M[(yr + k) * N + yc + (l & 15)];
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    return c;
}

__device__ __forceinline__ mfma_d4 tile_load(const double *M, int N, int r0, int c0)
{
    const int l = threadIdx.x & 63;
    mfma_d4 c;
#pragma unroll
    for (int q = 0; q < 4; q++) c[q] = M[(r0 + 4 * q + (l >> 4)) * N + c0 + (l & 15)];
    return c;
}

__device__ __forceinline__ void tile_store(double *M, int N, int r0, int c0, mfma_d4 c)
{
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < 4; q++) M[(r0 + 4 * q + (l >> 4)) * N + c0 + (l & 15)] = c[q];
}

__global__ void __launch_bounds__(1024)
k_dense_invert(OpArgs<double> o, double *inv, int *info)
{
    extern __shared__ double A[];            // N*N
    __shared__ double colk[16], rowk[16];
    const int n = o.nx * o.ny * o.nz;
    const int N = (n + 15) / 16 * 16, T = N / 16;
    const int nx = o.nx, nxy = o.nx * o.ny;
    for (int q = threadIdx.x; q < N * N; q += blockDim.x) A[q] = (q / N == q % N && q / N >= n) ? 1.0 : 0.0;
    __syncthreads();
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
        Cell c = cell_at(o, s);
        double d = diag_at<double, double>(o, c);
        A[s * N + s] = d;
        if (c.i < o.nx - 1) A[s * N + s + 1] = -o.wx[s];
        if (c.i > 0) A[s * N + s - 1] = -o.wx[s - 1];
        if (c.j < o.ny - 1) A[s * N + s + nx] = -o.wy[s];
        if (c.j > 0) A[s * N + s - nx] = -o.wy[s - nx];
        if (c.k < o.nz - 1) A[s * N + s + nxy] = -o.wz[s];
        if (c.k > 0) A[s * N + s - nxy] = -o.wz[s - nxy];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    for (int kb = 0; kb < T; kb++) {
        const int k0 = 16 * kb;
        // P = A_kk^-1 in place: scalar Gauss-Jordan, one entry per thread of the first 256
        for (int p = 0; p < 16; p++) {
            const double piv = A[(k0 + p) * N + k0 + p];
            if (!(piv > 0.0)) { if (threadIdx.x == 0) *info = k0 + p + 1; return; }   // uniform
            const double rp = 1.0 / piv;
            if (threadIdx.x < 16) {
                const int t = threadIdx.x;
                colk[t] = A[(k0 + t) * N + k0 + p];
                rowk[t] = (t == p ? 1.0 : A[(k0 + p) * N + k0 + t]) * rp;
            }
            __syncthreads();
            if (threadIdx.x < 256) {
                const int i = threadIdx.x >> 4, j = threadIdx.x & 15;
                double &x = A[(k0 + i) * N + k0 + j];
                x = i == p ? rowk[j] : ((j == p ? 0.0 : x) - colk[i] * rowk[j]);
            }
            __syncthreads();
        }
        // row panel: A_kj <- P A_kj
        for (int jb = wave; jb < T; jb += nwaves) {
            if (jb == kb) continue;
            mfma_d4 c = {0.0, 0.0, 0.0, 0.0};
            c = tile_mma(A, N, k0, k0, k0, 16 * jb, c, 1.0);
            tile_store(A, N, k0, 16 * jb, c);
        }
        __syncthreads();
        // trailing tiles: A_ij <- A_ij - A_ik A_kj
        for (int t = wave; t < T * T; t += nwaves) {
            const int ib = t / T, jb = t - ib * T;
            if (ib == kb || jb == kb) continue;
            mfma_d4 c = tile_load(A, N, 16 * ib, 16 * jb);
            c = tile_mma(A, N, 16 * ib, k0, k0, 16 * jb, c, -1.0);
            tile_store(A, N, 16 * ib, 16 * jb, c);
        }
        __syncthreads();
        // column panel: A_ik <- -A_ik P
        for (int ib = wave; ib < T; ib += nwaves) {
            if (ib == kb) continue;
            mfma_d4 c = {0.0, 0.0, 0.0, 0.0};
            c = tile_mma(A, N, 16 * ib, k0, k0, k0, c, -1.0);   // the wave's own tile: read, then written
            tile_store(A, N, 16 * ib, k0, c);
        }
        __syncthreads();
    }
    for (int q = threadIdx.x; q < n * n; q += blockDim.x) inv[q] = A[(q / n) * N + q % n];
    if (threadIdx.x == 0) *info = 0;
}

// ---------------------------------------------------------------- CG kernels
//
// Iteration k:  SPMV_k (s_k = z_k + beta s_{k-1}, v = A s_k, partial tau),
//               UPDATE_k (alpha_k = sigma_k / tau, x += alpha s, r -= alpha v,
//               partial ||r||^2), V-cycle on r whose last sweep writes z_{k+1}
//               and partials of <z, r>, <z, v>.
// SPMV_k sums the partials of UPDATE_{k-1} and of the dots: it decides the
// stop and computes beta. alpha and sigma live in two slots by parity of k so
// that no kernel overwrites a value that its own blocks still read.

// r = b - A x; partials of ||b||^2 and ||r||^2.
__global__ void __launch_bounds__(kBlock)
k_cg_init(OpArgs<double> o, const double *__restrict__ b, const double *__restrict__ x, double *__restrict__ r,
          int chunk, double *Pinit)
{
    int n = o.nx * o.ny * o.nz;
    int lb, beg, end;
    chunk_range(n, chunk, lb, beg, end);
    double v[2] = {0.0, 0.0};
    for (int base = beg; base < end; base += kUnroll * kBlock) {
        double bs[kUnroll], rs[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; q++) {
            int s = base + q * kBlock + threadIdx.x;
            bs[q] = 0.0; rs[q] = 0.0;
            if (s < end) {
                Cell c = cell_at(o, s);
                double d;
                double ax = apply_at<double, double>(o, c, [&](const Cell &t) { return x[t.s]; }, d);
                bs[q] = b[s];
                rs[q] = bs[q] - ax;
            }
        }
#pragma unroll
        for (int q = 0; q < kUnroll; q++) {
            int s = base + q * kBlock + threadIdx.x;
            if (s < end) r[s] = rs[q];
            v[0] += bs[q] * bs[q];
            v[1] += rs[q] * rs[q];
        }
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) { Pinit[2 * lb] = v[0]; Pinit[2 * lb + 1] = v[1]; }
}

// Start of a CG iteration k (read from the device): the stopping tests on
// the partial sums of the previous kernels, beta (Polak-Ribiere form), and the
// scalars written by one thread. Returns true if the solve stops here.
template <int NT = kBlock>
__device__ __forceinline__ bool cg_iteration_scalars(int G, int Gdot, CgScalars *sc, const double *Pinit,
                                                     const double *Prr, const double *Pdot, int &k, double &beta)
{
    k = sc->kspmv;
    double dots[2];
    sum_partials_n<2, NT>(Pdot, Gdot, dots);
    double bb, rr;
    if (k == 0) {
        double ir[2];
        sum_partials_n<2, NT>(Pinit, G, ir);
        bb = ir[0];
        rr = ir[1];
    } else {
        double t[1];
        sum_partials_n<1, NT>(Prr, G, t);
        rr = t[0];
        bb = sc->bb;
    }
    const double sigma = fabs(dots[0]);
    int stop = 0, status = HMFEM_OK;
    beta = 0.0;
    double sigma_first = k == 0 ? (sigma == 0.0 ? 1.0 : sigma) : sc->sigma_first;
    if (k == 0) {
        // the reference stops at once only when <M r0, r0> = 0 (pcg.h:51-57); the residual rule also
        // when the start already meets it
        if (sigma == 0.0 || (sc->stop_rule == HMFEM_STOP_RESIDUAL && (bb == 0.0 || rr <= sc->rtol2 * bb))) stop = 1;
    } else {
        beta = -sc->alpha[(k - 1) & 1] * dots[1] / sc->sigma[(k - 1) & 1];
        if (sc->stop_rule == HMFEM_STOP_RESIDUAL && rr <= sc->rtol2 * bb) stop = 1;
        if (sc->stop_rule == HMFEM_STOP_PRECONDITIONED && sigma <= sc->rtol2 * sigma_first) stop = 1;
        if (!(sigma == sigma) || !(beta == beta)) { stop = 1; status = HMFEM_ERR_BREAKDOWN; }
        else if (!stop && k >= sc->maxit) { stop = 1; status = HMFEM_NOT_CONVERGED; }
    }
    if (xcd_block() == 0 && threadIdx.x == 0) {
        if (k == 0) { sc->bb = bb; sc->sigma_first = sigma_first; }
        sc->rr = rr;
        sc->sigma[k & 1] = sigma;
        sc->beta = beta;
        sc->iter = k;
        sc->kupd = k;
        if (stop) { sc->done = 1; sc->status = status; }
    }
    return stop != 0;
}

// The stopping tests and beta of CG iteration sc->kspmv (one block; the
// partials summed in the order every SpMV block used before, so the iterates
// are the same: with 1024 threads the 256^3 conduit case took 71 iterations
// instead of 69, rounding alone).
constexpr int kScalarThreads = kBlock;
__global__ void __launch_bounds__(kScalarThreads)
k_cg_scalars(int G, int Gdot, CgScalars *sc, const double *Pinit, const double *Prr, const double *Pdot)
{
    if (sc->done) return;
    int k;
    double beta;
    (void)cg_iteration_scalars<kScalarThreads>(G, Gdot, sc, Pinit, Prr, Pdot, k, beta);
}

template <typename ZT>
__global__ void __launch_bounds__(kBlock)
k_cg_spmv(OpArgs<double> o, const ZT *__restrict__ z, double *__restrict__ s0, double *__restrict__ s1,
          double *__restrict__ v, int chunk, int G, int Gdot, CgScalars *sc, const double *Pinit, const double *Prr,
          const double *Pdot, double *Ptau)
{
    if (sc->done) return;   // the iteration's scalars: k_cg_scalars, just before
    (void)Gdot; (void)Pinit; (void)Prr; (void)Pdot;
    const int n = o.nx * o.ny * o.nz;
    const int k = sc->kspmv;
    const double beta = sc->beta;
    const double *__restrict__ sin = (k & 1) ? s1 : s0;
    double *__restrict__ sout = (k & 1) ? s0 : s1;
    int lb, beg, end;
    chunk_range(n, chunk, lb, beg, end);
    double acc[1] = {0.0};
    auto sv = [&](const Cell &t) -> double { return k == 0 ? (double)z[t.s] : (double)z[t.s] + beta * sin[t.s]; };
    for (int base = beg; base < end; base += kUnroll * kBlock) {
        double ss[kUnroll], av[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; q++) {
            int s = base + q * kBlock + threadIdx.x;
            ss[q] = 0.0; av[q] = 0.0;
            if (s < end) {
                Cell c = cell_at(o, s);
                double d;
                av[q] = apply_at<double, double>(o, c, sv, d);
                ss[q] = sv(c);
            }
        }
#pragma unroll
        for (int q = 0; q < kUnroll; q++) {
            int s = base + q * kBlock + threadIdx.x;
            if (s < end) { sout[s] = ss[q]; v[s] = av[q]; }
            acc[0] += ss[q] * av[q];
        }
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) Ptau[lb] = acc[0];
}

// The CG SpMV marching along z: s = z + beta s_old, v = A s and <s, v>.
// A block owns a tile of kGsTx x kGsTy columns and the planes [k0, k0 + zlen);
// s of the tile and its x/y ring is kept in LDS for three planes (k-1, k,
// k+1), so each value of s_old and z is read once per block (the ring is
// also read by the neighbouring tiles, from L2), however large a plane is.
// One partial <s, v> per logical block (Ptau[lb]; blocks past the last tile
// write 0).
// It also applies the previous iteration's x += alpha s_old, which
// k_cg_update leaves to it (s_old is read here anyway): that saves reading s
// again and a read-modify-write pass over x per iteration. If the solve stops
// at this kernel, k_cg_finish_x applies the pending update instead.
constexpr int kSpmvSlots = 4;   // LDS planes of s (three need a second barrier per plane)
template <typename ZT>
__global__ void __launch_bounds__(kBlock)
k_cg_spmv_z(OpArgs<double> o, const ZT *__restrict__ z, double *__restrict__ s0, double *__restrict__ s1,
            double *__restrict__ v, double *__restrict__ x, int zlen, int G, int Gdot, CgScalars *sc,
            const double *Pinit, const double *Prr, const double *Pdot, double *Ptau, int eface)
{
    // the iteration's scalars come from k_cg_scalars (one block, just before):
    // uniform loads instead of every block summing the partials again
    if (sc->done) return;
    (void)G; (void)Gdot; (void)Pinit; (void)Prr; (void)Pdot;
    const int k = sc->kspmv;
    const double beta = sc->beta;
    const double *__restrict__ sin = (k & 1) ? s1 : s0;
    double *__restrict__ sout = (k & 1) ? s0 : s1;
    const double alpha_prev = k ? sc->alpha[(k - 1) & 1] : 0.0;
    __shared__ double sh[kSpmvSlots][kGsTy + 2][kGsTx + 2];
    const int nx = o.nx, ny = o.ny, nz = o.nz;
    const long nxy = (long)nx * ny;
    const int ntx = (nx + kGsTx - 1) / kGsTx, nty = (ny + kGsTy - 1) / kGsTy;
    const int ntxy = ntx * nty, nzc = (nz + zlen - 1) / zlen;
    const int lb = xcd_block();
    double acc[1] = {0.0};
    if (lb < ntxy * nzc) {
        const int tile = lb % ntxy, zc = lb / ntxy;
        const int i0 = (tile % ntx) * kGsTx, j0 = (tile / ntx) * kGsTy;
        const int kb = zc * zlen, ke = min(nz, kb + zlen);
        const int tx = threadIdx.x % kGsTx, ty = threadIdx.x / kGsTx;
        const int i = i0 + tx, j = j0 + ty;
        const bool col = i < nx && j < ny;
        // ring cell of this thread (threads 0 .. kGsRing-1), as in the smoother
        int rx = 0, ry = 0;
        const bool has_ring = threadIdx.x < kGsRing;
        if (has_ring) {
            int hz;
            gs_ring_cell(threadIdx.x, rx, ry, hz);
        }
        const int ri = i0 + rx - 1, rj = j0 + ry - 1;
        const bool ring_in = has_ring && ri >= 0 && rj >= 0 && ri < nx && rj < ny;
        // Loads run one plane ahead of the LDS writes and two planes ahead of
        // the arithmetic: raw z and s_old of plane kk + 2 and the coefficients
        // of plane kk + 1 are in flight while plane kk is computed.
        struct Raw {
            double zo, so, zr, sr, xo;
        };
        auto load_raw = [&](int kk) -> Raw {
            Raw r{0.0, 0.0, 0.0, 0.0, 0.0};
            if (kk >= 0 && kk < nz) {
                if (col) {
                    const long q = i + (long)j * nx + (long)kk * nxy;
                    r.zo = (double)z[q];
                    if (k) r.so = sin[q];
                    if (k && kk >= kb && kk < ke) r.xo = x[q];
                }
                if (ring_in) {
                    const long q = ri + (long)rj * nx + (long)kk * nxy;
                    r.zr = (double)z[q];
                    if (k) r.sr = sin[q];
                }
            }
            return r;
        };
        auto store_raw = [&](int kk, const Raw &r) {
            const int slot = (kk + kSpmvSlots) % kSpmvSlots;
            sh[slot][ty + 1][tx + 1] = r.zo + beta * r.so;   // beta = 0 and s_old unread at k = 0
            if (has_ring) sh[slot][ry][rx] = r.zr + beta * r.sr;
            if (col && k && kk >= kb && kk < ke) x[i + (long)j * nx + (long)kk * nxy] = r.xo + alpha_prev * r.so;
        };
        struct Coef {
            double e, xp, xm, yp, ym, zp;
        };
        // with eface, e is read only on the grid's faces (zero elsewhere)
        const bool eall = !eface || i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
        auto load_coef = [&](int kk) -> Coef {
            Coef c{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (col && kk < nz) {
                const long q = i + (long)j * nx + (long)kk * nxy;
                if (eall || kk == 0 || kk == nz - 1) c.e = o.e[q];
                c.xp = i < nx - 1 ? o.wx[q] : 0.0;
                c.xm = i > 0 ? o.wx[q - 1] : 0.0;
                c.yp = j < ny - 1 ? o.wy[q] : 0.0;
                c.ym = j > 0 ? o.wy[q - nx] : 0.0;
                c.zp = kk < nz - 1 ? o.wz[q] : 0.0;
            }
            return c;
        };
        store_raw(kb - 1, load_raw(kb - 1));
        store_raw(kb, load_raw(kb));
        Raw nxt = load_raw(kb + 1);
        Coef cf = load_coef(kb);
        double wz_lo = col && kb > 0 ? o.wz[i + (long)j * nx + (long)(kb - 1) * nxy] : 0.0;
        for (int kk = kb; kk < ke; kk++) {
            store_raw(kk + 1, nxt);
            nxt = load_raw(kk + 2);
            const Coef cn = load_coef(kk + 1);
            __syncthreads();
            const int c = (kk + kSpmvSlots) % kSpmvSlots, lo = (kk + kSpmvSlots - 1) % kSpmvSlots,
                      hi = (kk + 1) % kSpmvSlots;
            if (col) {
                const long q = i + (long)j * nx + (long)kk * nxy;
                const double us = sh[c][ty + 1][tx + 1];
                double a = cf.e * us;
                a += cf.xp * (us - sh[c][ty + 1][tx + 2]);
                a += cf.xm * (us - sh[c][ty + 1][tx]);
                a += cf.yp * (us - sh[c][ty + 2][tx + 1]);
                a += cf.ym * (us - sh[c][ty][tx + 1]);
                a += cf.zp * (us - sh[hi][ty + 1][tx + 1]);
                a += wz_lo * (us - sh[lo][ty + 1][tx + 1]);
                wz_lo = cf.zp;
                sout[q] = us;
                v[q] = a;
                acc[0] += us * a;
            }
            cf = cn;
            if (kSpmvSlots == 3) __syncthreads();   // four slots: the next store goes to a plane nobody reads
        }
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) Ptau[lb] = acc[0];
}

// Blocks of k_cg_spmv_z (each writes one partial of <s, v>). Before it took
// over the x update the kernel held 78 VGPRs (6 waves per SIMD) and measured,
// at 128^3 / 256^3: 1024 blocks 29.6 / 219 us, 512: 36.4 / 321, 1536:
// 30.5 / 214, 2048: 36.3 / 262 (two rounds of blocks); capping it at 64
// VGPRs spilled (43 / 360 us). With the x update it holds 84 (5 waves):
// 34.5 / 255 us, against 18.7 -> 11.6 / 158 -> 78 us saved in k_cg_update
// (loading x at the store instead, 82 VGPRs: 35.2 / 275 us; beta and
// alpha moved to scalar registers and z kept in FP32 until the LDS store,
// 78 VGPRs and 6 waves: 33.9 / 282 us, so more waves are slower here).
constexpr int kMaxSpmvBlocks = 1024;

// Planes per block of k_cg_spmv_z so that it needs at most maxb blocks (at
// least 4 planes each); 0 if a single plane has more tiles than that.
inline int spmv_zlen(const Dims &d, int maxb = kMaxSpmvBlocks)
{
    const long ntxy = (long)((d.nx + kGsTx - 1) / kGsTx) * ((d.ny + kGsTy - 1) / kGsTy);
    if (ntxy > maxb) return 0;
    long zl = 4;
    while (ntxy * ((d.nz + zl - 1) / zl) > maxb) zl++;
    return (int)zl;
}

inline int spmv_z_blocks(const Dims &d, int zl)
{
    const int ntxy = ((d.nx + kGsTx - 1) / kGsTx) * ((d.ny + kGsTy - 1) / kGsTy);
    return grid_for((long)ntxy * ((d.nz + zl - 1) / zl), 1);
}

// r -= alpha v and the partials of <r, r>; x += alpha s too unless `defer`
// (then the next k_cg_spmv_z, or k_cg_finish_x, applies it: sc->xpend names
// the pending iteration).
__global__ void __launch_bounds__(kBlock)
k_cg_update(int n, const double *__restrict__ s0, const double *__restrict__ s1, const double *__restrict__ v,
            double *__restrict__ x, double *__restrict__ r, int chunk, int Gtau, CgScalars *sc, const double *Ptau,
            double *Prr, int defer)
{
    if (sc->done) return;
    const int k = sc->kupd;
    const double *__restrict__ s = (k & 1) ? s0 : s1;
    double t[1];
    sum_partials<1>(Ptau, Gtau, t);
    const double tau = fabs(t[0]);
    if (tau == 0.0 || !(tau == tau)) {
        if (xcd_block() == 0 && threadIdx.x == 0) {
            sc->done = 1; sc->status = HMFEM_ERR_BREAKDOWN; sc->tau = tau;
            sc->xpend = -1;   // the previous update was applied by this iteration's SpMV
        }
        return;
    }
    const double alpha = sc->sigma[k & 1] / tau;
    if (xcd_block() == 0 && threadIdx.x == 0) {
        sc->alpha[k & 1] = alpha; sc->tau = tau; sc->kspmv = k + 1;
        if (defer) sc->xpend = k;
    }
    int lb, beg, end;
    chunk_range(n, chunk, lb, beg, end);
    double acc[1] = {0.0};
    for (int base = beg; base < end; base += kUnroll * kBlock) {
        double xs[kUnroll], ss[kUnroll], rs[kUnroll], vs[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; q++) {
            int i = base + q * kBlock + threadIdx.x;
            xs[q] = 0.0; ss[q] = 0.0; rs[q] = 0.0; vs[q] = 0.0;
            if (i < end) {
                rs[q] = r[i]; vs[q] = v[i];
                if (!defer) { xs[q] = x[i]; ss[q] = s[i]; }
            }
        }
#pragma unroll
        for (int q = 0; q < kUnroll; q++) {
            int i = base + q * kBlock + threadIdx.x;
            double ri = rs[q] - alpha * vs[q];
            if (i < end) {
                if (!defer) x[i] = xs[q] + alpha * ss[q];
                r[i] = ri;
            }
            acc[0] += ri * ri;
        }
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) Prr[lb] = acc[0];
}

// After the CG loop: the x += alpha s of iteration sc->xpend, if no SpMV
// applied it (the solve stopped at the SpMV that would have).
__global__ void __launch_bounds__(kBlock)
k_cg_finish_x(int n, const double *__restrict__ s0, const double *__restrict__ s1, double *__restrict__ x,
              const CgScalars *sc)
{
    const int k = sc->xpend;
    if (k < 0) return;
    const double *__restrict__ s = (k & 1) ? s0 : s1;
    const double alpha = sc->alpha[k & 1];
    for (long i = (long)xcd_block() * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock)
        x[i] = x[i] + alpha * s[i];
}

// ---------------------------------------------------------------- host side


static void plan_levels(Dims fine, int coarse_max, Level *lev, int &nlev, const PlanHint *hint = nullptr)
{
    nlev = 0;
    Dims d = fine;
    // coupling strength per axis, carried down the levels: a coarse face
    // between boxes sums the fine faces it covers, so a coarse x coupling is
    // about fy * fz fine ones (and so on)
    double m[3] = {1.0, 1.0, 1.0};
    const bool semi = hint && hint->on;
    if (semi)
        for (int a = 0; a < 3; a++) m[a] = hint->m[a];
    while (true) {
        Level &L = lev[nlev];
        L = Level();
        L.d = d;
        nlev++;
        if (d.n() <= coarse_max || nlev == kMaxLevels) break;
        // 2x2x2 boxes where the couplings are about isotropic: 4x4x4 boxes
        // below 32^3 (one level fewer) took 28 CG iterations instead of 16 at
        // 128^3 and 121 instead of 85 on the 256^3 conduit field
        const int ext[3] = {d.nx, d.ny, d.nz};
        double top = 0.0;
        for (int a = 0; a < 3; a++)
            if (ext[a] > 1) top = fmax(top, m[a]);
        int f[3];
        for (int a = 0; a < 3; a++) f[a] = ext[a] > 1 && (!semi || !(top > 0) || m[a] >= kSemiTheta * top) ? 2 : 1;
        L.fx = f[0]; L.fy = f[1]; L.fz = f[2];
        const double m0 = m[0] * f[1] * f[2], m1 = m[1] * f[0] * f[2], m2 = m[2] * f[0] * f[1];
        m[0] = m0; m[1] = m1; m[2] = m2;
        Dims c{(d.nx + L.fx - 1) / L.fx, (d.ny + L.fy - 1) / L.fy, (d.nz + L.fz - 1) / L.fz};
        d = c;
    }
}

// Per-block sums of ln w and counts of nonzero w per axis over every 16th
// cell (6 partials per block).
__global__ void __launch_bounds__(kBlock) k_axis_logmean(OpArgs<double> o, double *partial)
{
    const long n = (long)o.nx * o.ny * o.nz;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    // 64 consecutive cells out of every 1024 (whole cache lines; every 16th
    // cell alone fetched every line: 103 instead of 8 us at 256^3)
    for (long q = (long)blockIdx.x * kBlock + threadIdx.x;; q += (long)gridDim.x * kBlock) {
        const long s = (q >> 6) * 1024 + (q & 63);
        if (s >= n) break;
        const double w[3] = {o.wx[s], o.wy[s], o.wz[s]};
        for (int a = 0; a < 3; a++)
            if (w[a] > 0) { acc[2 * a] += log(w[a]); acc[2 * a + 1] += 1.0; }
    }
    block_sum<6>(acc);
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; q++) partial[6 * blockIdx.x + q] = acc[q];
}

int axis_strength(hipStream_t st, Dims d, const Stencil<double> &A, double *partial, PlanHint &hint)
{
    OpArgs<double> o;
    o.e = A.e; o.wx = A.wx; o.wy = A.wy; o.wz = A.wz; o.di = nullptr;
    o.nx = d.nx; o.ny = d.ny; o.nz = d.nz;
    hipLaunchKernelGGL(k_axis_logmean, dim3(kAxisBlocks), dim3(kBlock), 0, st, o, partial);
    HMFEM_CHECK(hipGetLastError());
    std::vector<double> h(6 * kAxisBlocks);
    HMFEM_CHECK(hipMemcpyAsync(h.data(), partial, h.size() * 8, hipMemcpyDeviceToHost, st));
    HMFEM_CHECK(hipStreamSynchronize(st));
    double sum[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < kAxisBlocks; b++)
        for (int q = 0; q < 6; q++) sum[q] += h[6 * b + q];
    hint.on = true;
    for (int a = 0; a < 3; a++) hint.m[a] = sum[2 * a + 1] > 0 ? exp(sum[2 * a] / sum[2 * a + 1]) : 0.0;
    return HMFEM_OK;
}

static int tail_start(const Level *lev, int nlev, long tail_max)
{
    if (nlev < 3 || tail_max <= 0) return -1;
    for (int l = 1; l < nlev - 1; l++)
        if (lev[l].d.n() <= tail_max) return (nlev - l <= kMaxTailLevels) ? l : -1;
    return -1;
}

static int clamp_coarse(int c) { return c > kCoarseMax ? kCoarseMax : (c < 1 ? 1 : c); }

size_t hierarchy_bytes(Dims fine, const hmfem_options &opt, const PlanHint *hint)
{
    Level lev[kMaxLevels];
    int nlev;
    plan_levels(fine, clamp_coarse(opt.coarse_max), lev, nlev, hint);
    size_t b = 0;
    size_t vs = opt.precision == HMFEM_PREC_FP64 ? 8 : 4;
    for (int l = 0; l < nlev; l++) {
        size_t n = lev[l].d.n();
        if (l > 0) b += 4 * align256(n * 8);            // FP64 coefficients (fine level is the caller's)
        if (vs == 8) b += align256(n * 8);              // FP64 inverse diagonal
        b += 5 * align256(n * 4);                       // FP32 coefficients and inverse diagonal
        b += 5 * align256(n * vs);                      // f, u, t, r, w
    }
    b += align256((size_t)kCoarseMax * kCoarseMax * 8);
    b += align256(sizeof(TailLevel<double>) * kMaxTailLevels) + align256(sizeof(int4) * kMaxTailOps);
    if (opt.precision == HMFEM_PREC_FP16) b += 3 * align256(fine.n() * 2) + align256(fine.n() * 4);
    return b + 8192;
}

// Op list of the tail: cycle(l) leaves u_l solving A_l u = rhs(fsel).
static void gen_tail(std::vector<int4> &ops, int l, int fsel, int last, int nu, int gamma)
{
    if (l == last) { ops.push_back(make_int4(OP_DENSE, l, fsel, 0)); return; }
    ops.push_back(make_int4(OP_SMOOTH0, l, fsel, 0));
    for (int q = 1; q < nu; q++) ops.push_back(make_int4(OP_JACOBI, l, fsel, 0));
    ops.push_back(make_int4(OP_RESTRICT, l, fsel, 0));
    gen_tail(ops, l + 1, 0, last, nu, gamma);
    if (gamma >= 2 && l + 1 != last) {
        ops.push_back(make_int4(OP_RESID, l + 1, 0, 0));
        ops.push_back(make_int4(OP_SAVE, l + 1, 0, 0));
        gen_tail(ops, l + 1, 1, last, nu, gamma);
        ops.push_back(make_int4(OP_ADDSAVED, l + 1, 0, 0));
    }
    ops.push_back(make_int4(OP_PROLONG_JACOBI, l, fsel, 0));
    for (int q = 1; q < nu; q++) ops.push_back(make_int4(OP_JACOBI, l, fsel, 0));
}


template <typename T>
static int upload_tail(hipStream_t st, Hierarchy &H, const hmfem_options &opt, std::vector<char> &stage,
                       long lds_limit)
{
    int last = H.nlev - 1;
    std::vector<TailLevel<T>> tl(H.nlev - H.tail0);
    size_t off = 0;
    for (int l = H.tail0; l < H.nlev; l++) {
        const Level &L = H.lev[l];
        TailLevel<T> &t = tl[l - H.tail0];
        t.o = opargs<T>(L);
        t.f = (T *)L.f; t.u = (T *)L.u; t.t = (T *)L.t; t.r = (T *)L.r; t.w = (T *)L.w;
        t.n = (int)L.d.n();
        t.fx = L.fx; t.fy = L.fy; t.fz = L.fz;
        t.lds_off = (int)off;
        // e, wx, wy, wz, di, f, u, t, and r, w (second right-hand side and
        // saved u of a repeated visit) except on the first tail level, which
        // the op list visits once
        const int arrays = l == H.tail0 ? 8 : 10;
        off += (arrays * (size_t)t.n * sizeof(T) + 15) / 16 * 16;
    }
    size_t inv_off = off;
    const size_t inv_bytes = (size_t)H.nc * H.nc * sizeof(double);
    if ((long)(off + inv_bytes) <= lds_limit) {
        H.tail_lds_bytes = (int)(off + inv_bytes);
        H.tail_lds_inv = (int)inv_off;
    } else if ((long)off <= lds_limit) {   // the levels in LDS, the dense inverse read from global memory
        H.tail_lds_bytes = (int)off;
        H.tail_lds_inv = kTailInvGlobal;
    } else {
        H.tail_lds_bytes = 0;
        H.tail_lds_inv = -1;
        for (auto &t : tl) t.lds_off = -1;
    }
    std::vector<int4> ops;
    int gamma = opt.tail_gamma < 1 ? 1 : (opt.tail_gamma > 2 ? 2 : opt.tail_gamma);
    gen_tail(ops, 0, 0, last - H.tail0, opt.nu < 1 ? 1 : opt.nu, gamma);
    ops.push_back(make_int4(OP_FINAL, 0, 0, 0));
    if ((int)ops.size() > kMaxTailOps) return HMFEM_ERR_ARG;
    H.tail_nops = (int)ops.size();
    size_t lb = sizeof(TailLevel<T>) * tl.size(), ob = sizeof(int4) * ops.size();
    stage.resize(lb + ob);
    memcpy(stage.data(), tl.data(), lb);
    memcpy(stage.data() + lb, ops.data(), ob);
    HMFEM_CHECK(hipMemcpyAsync(H.tail_levels, stage.data(), lb, hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipMemcpyAsync(H.tail_ops, stage.data() + lb, ob, hipMemcpyHostToDevice, st));
    return HMFEM_OK;
}

// Build the hierarchy from a fine FP64 stencil (device), inside `arena`.
int build_hierarchy(hipStream_t st, Dims fine, const Stencil<double> &s0, const hmfem_options &opt, Arena &ar,
                    Hierarchy &H, int *d_info, std::vector<char> &stage, long lds_limit, const PlanHint *hint,
                    Deflation *dh)
{
    plan_levels(fine, clamp_coarse(opt.coarse_max), H.lev, H.nlev, hint);
    int prec = opt.precision == HMFEM_PREC_FP64 ? HMFEM_PREC_FP64 : HMFEM_PREC_FP32;
    H.prec = prec;
    H.tail0 = tail_start(H.lev, H.nlev, opt.tail_max);
    size_t vs = prec == HMFEM_PREC_FP64 ? 8 : 4;
    for (int l = 0; l < H.nlev; l++) {
        Level &L = H.lev[l];
        size_t n = L.d.n();
        if (l == 0) L.s64 = s0;
        else {
            L.s64.e = ar.take<double>(n); L.s64.wx = ar.take<double>(n);
            L.s64.wy = ar.take<double>(n); L.s64.wz = ar.take<double>(n);
        }
        // a box sum of cells off the faces is off the faces: every level keeps eface
        L.s64.eface = L.s32.eface = s0.eface;
        if (prec == HMFEM_PREC_FP64) L.s64.di = ar.take<double>(n);
        L.s32.e = ar.take<float>(n); L.s32.wx = ar.take<float>(n);
        L.s32.wy = ar.take<float>(n); L.s32.wz = ar.take<float>(n); L.s32.di = ar.take<float>(n);
        L.f = ar.take<char>(n * vs); L.u = ar.take<char>(n * vs); L.t = ar.take<char>(n * vs);
        L.r = ar.take<char>(n * vs); L.w = ar.take<char>(n * vs);
        if (!L.s32.di || !L.w || (l > 0 && !L.s64.wz) || (prec == HMFEM_PREC_FP64 && !L.s64.di)) return HMFEM_ERR_NOMEM;
    }
    if (opt.precision == HMFEM_PREC_FP16 && H.nlev > 1) {
        const size_t n = H.lev[0].d.n();
        H.f16x = ar.take<__half>(n); H.f16y = ar.take<__half>(n); H.f16z = ar.take<__half>(n);
        H.f16sq = ar.take<float>(n);
        if (!H.f16sq) return HMFEM_ERR_NOMEM;
    }
    H.nc = (int)H.lev[H.nlev - 1].d.n();
    if (H.nc > kCoarseMax) return HMFEM_ERR_ARG;   // too many levels would be needed
    H.inv = ar.take<double>((size_t)H.nc * H.nc);
    H.tail_levels = ar.take<char>(sizeof(TailLevel<double>) * kMaxTailLevels);
    H.tail_ops = ar.take<int4>(kMaxTailOps);
    if (!H.inv || !H.tail_levels || !H.tail_ops) return HMFEM_ERR_NOMEM;

    {
        Level &L = H.lev[0];
        long n = L.d.n();
        if (prec == HMFEM_PREC_FP64)
            hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, n, L.s64.e, L.s64.wx,
                               L.s64.wy, L.s64.wz, L.s32.e, L.s32.wx, L.s32.wy, L.s32.wz);
        else if (dh && dh->pending && dh->h.bhist) {   // with the inverse diagonal and the histogram
            hipLaunchKernelGGL(k_to_f32_dinv_hist, dim3((unsigned)((n + kBlock * kConvCells - 1) / (kBlock * kConvCells))),
                               dim3(kBlock), 0, st, opargs<double>(L), L.s32.e, L.s32.wx, L.s32.wy, L.s32.wz,
                               L.s32.di, dh->h.bhist, dh->h.bbins);
            int rc = deflation_hist_queued(st, *dh);
            if (rc) return rc;
        } else {   // with the fine level's inverse diagonal
            hipLaunchKernelGGL(k_to_f32_dinv, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, opargs<double>(L),
                               L.s32.e, L.s32.wx, L.s32.wy, L.s32.wz, L.s32.di);
        }
        if (H.f16sq) {
            hipLaunchKernelGGL(k_fine16_scale, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, opargs<double>(L),
                               H.f16sq);
            hipLaunchKernelGGL(k_fine16_couplings, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st,
                               opargs<double>(L), (const float *)H.f16sq, H.f16x, H.f16y, H.f16z);
        }
    }
    for (int l = 0; l + 1 < H.nlev; l++) {
        Level &L = H.lev[l], &C = H.lev[l + 1];
        OpArgs<double> o = opargs<double>(L);
        long nc = C.d.n();
        hipLaunchKernelGGL(k_rap_box, dim3(grid_for(nc, kBlock)), dim3(kBlock), 0, st, o, L.fx, L.fy, L.fz, C.d.nx,
                           C.d.ny, C.d.nz, C.s64.e, C.s64.wx, C.s64.wy, C.s64.wz, C.s32.e, C.s32.wx, C.s32.wy,
                           C.s32.wz);
    }
    for (int l = 0; l < H.nlev; l++) {
        Level &L = H.lev[l];
        int G = grid_for(L.d.n(), kBlock);
        if (prec == HMFEM_PREC_FP64)
            hipLaunchKernelGGL(k_dinv<double>, dim3(G), dim3(kBlock), 0, st, opargs<double>(L), L.s64.di);
        else if (l > 0)
            hipLaunchKernelGGL(k_dinv<float>, dim3(G), dim3(kBlock), 0, st, opargs<float>(L), L.s32.di);
    }
    {
        Level &L = H.lev[H.nlev - 1];
        OpArgs<double> o = opargs<double>(L);
        const size_t N = (size_t)(H.nc + 15) / 16 * 16;   // padded to whole 16x16 tiles
        size_t lds = N * N * sizeof(double);
        hipLaunchKernelGGL(k_dense_invert, dim3(1), dim3(1024), lds, st, o, H.inv, d_info);
    }
    HMFEM_CHECK(hipGetLastError());
    if (H.tail0 > 0) {
        long lim = opt.tail_lds ? lds_limit : 0;
        int rc = prec == HMFEM_PREC_FP64 ? upload_tail<double>(st, H, opt, stage, lim)
                                         : upload_tail<float>(st, H, opt, stage, lim);
        if (rc) return rc;
    }
    return HMFEM_OK;
}

struct CycleParams {
    int nu;
    double omega, ocf;
    int smoother, gs;    // smoother of the multi-block levels and its sweeps
};

static int lg2f(int f) { return f == 2 ? 1 : 0; }

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

// Residual + restriction from level L to C: the z-marching kernel for 2x2x2
// aggregation, the general one otherwise.
template <typename T, typename FT>
static void launch_restrict(hipStream_t st, Prof *P, int cls, const Level &L, const Level &C, const FT *f,
                            const T *u, const int *done)
{
    OpArgs<T> o = opargs<T>(L);
    if (L.fx == 2 && L.fy == 2 && L.fz == 2) {
        const int zl = restrict_zlen(L.d);
        const int nb = ((L.d.nx + kGsTx - 1) / kGsTx) * ((L.d.ny + kGsTy - 1) / kGsTy) * ((L.d.nz + zl - 1) / zl);
        PROF_LAUNCH(P, cls, (k_restrict_z<T, FT>), dim3(grid_for(nb, 1)), dim3(kBlock), 0, st, o, zl, C.d.nx,
                    C.d.ny, C.d.nz, f, u, (T *)C.f, done, (int)L.s64.eface);
    } else {
        PROF_LAUNCH(P, cls, (k_resid_restrict<T, FT, 0>), dim3(grid_for(C.d.n(), kBlock)), dim3(kBlock), 0, st, o,
                    L.fx, L.fy, L.fz, C.d.nx, C.d.ny, C.d.nz, f, u, (T *)C.f, done);
    }
}


static Chunking chunking(long n)
{
    Chunking c;
    long g = (n + kBlock * kCellsPerThread - 1) / (kBlock * kCellsPerThread);
    if (g > kMaxReduceBlocks) g = kMaxReduceBlocks;
    if (g < kNumXcd) g = kNumXcd;
    g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
    long chunk = (n + g - 1) / g;
    chunk = (chunk + kUnroll * kBlock - 1) / (kUnroll * kBlock) * (kUnroll * kBlock);
    c.G = (int)g;
    c.chunk = (int)chunk;
    return c;
}

// Device buffers and state of one CG solve.
struct CgState {
    CgBuffers B;
    Chunking ch;
    double *Pinit, *Ptau, *Prr, *Pdot;
    const int *done;
};

// Recursive V-cycle on level l >= 1 (multi-block levels) down to the tail.
// The result is left in L.u.
template <typename T>
static T *vcycle_coarse(hipStream_t st, Hierarchy &H, int l, const CycleParams &cp, Prof *P, const int *done)
{
    Level &L = H.lev[l];
    T *u = (T *)L.u, *t = (T *)L.t;
    const T *f = (const T *)L.f;
    if (l == H.nlev - 1) {
        PROF_LAUNCH(P, HMFEM_PROF_COARSE, (k_dense_apply<T, T>), dim3(1), dim3(kBlock), 0, st, H.nc, H.inv, f, u,
                    done);
        return u;
    }
    if (H.tail0 > 0 && l == H.tail0 && tail_in_lds(H)) {
        PROF_LAUNCH(P, HMFEM_PROF_COARSE, k_tail_lds<T>, dim3(1), dim3(kTailThreads), H.tail_lds_bytes, st,
                    (const TailLevel<T> *)H.tail_levels, H.tail_ops, H.tail_nops, H.inv, H.nc, (T)cp.omega,
                    (T)cp.ocf, H.nlev - H.tail0, H.tail_lds_inv, (const T *)nullptr, (T *)nullptr, 1, done);
        return u;
    }
    if (H.tail0 > 0 && l == H.tail0) {
        PROF_LAUNCH(P, HMFEM_PROF_COARSE, k_tail<T>, dim3(1), dim3(kTailThreads), H.tail_lds_bytes, st,
                    (const TailLevel<T> *)H.tail_levels, H.tail_ops, H.tail_nops, H.inv, H.nc, (T)cp.omega,
                    (T)cp.ocf, H.nlev - H.tail0, H.tail_lds_inv, done);
        return u;
    }
    OpArgs<T> o = opargs<T>(L);
    const T om = (T)cp.omega;
    int G = grid_for(L.d.n(), kBlock);
    Level &C = H.lev[l + 1];
    if (cp.smoother == HMFEM_SMOOTH_BLOCK_GS) {
        const int GT = grid_for(gs_tiles(L.d), 1);
        // the coarse levels use the fine level's sweep count: 1, 2, 3, 4 sweeps
        // there gave 22, 18, 16, 16 CG iterations at 128^3 (3.92, 3.30, 3.08,
        // 3.13 ms) and 103, 83, 85, 88 on the 256^3 conduit field
        const int cgs = cp.gs;
        PROF_LAUNCH(P, HMFEM_PROF_COARSE, (k_gs_tile<T, T, 0>), dim3(GT), dim3(kGsThreads), 0, st, o, cgs, f,
                    (const T *)nullptr, u, 0, 0, 0, 0, 0, (const T *)nullptr, (T)0, (const double *)nullptr, 1,
                    (double *)nullptr, done);
        launch_restrict<T, T>(st, P, HMFEM_PROF_COARSE, L, C, f, (const T *)u, done);
        T *uc = vcycle_coarse<T>(st, H, l + 1, cp, P, done);
        PROF_LAUNCH(P, HMFEM_PROF_COARSE, (k_gs_tile<T, T, 1>), dim3(GT), dim3(kGsThreads), 0, st, o, cgs, f,
                    (const T *)u, t, lg2f(L.fx), lg2f(L.fy), lg2f(L.fz), C.d.nx, C.d.ny, (const T *)uc, (T)cp.ocf,
                    (const double *)nullptr, 1, (double *)nullptr, done);
        L.u = t; L.t = u;
        return t;
    }
    PROF_LAUNCH(P, HMFEM_PROF_COARSE, (k_smooth_zero<T, T>), dim3(G), dim3(kBlock), 0, st, o, f, u, om,
                cp.nu >= 2 ? 2 : 1, done);
    for (int q = 2; q < cp.nu; q++) {
        PROF_LAUNCH(P, HMFEM_PROF_COARSE, (k_jacobi<T, T>), dim3(G), dim3(kBlock), 0, st, o, f, u, t, om, done);
        T *tmp = u; u = t; t = tmp;
    }
    launch_restrict<T, T>(st, P, HMFEM_PROF_COARSE, L, C, f, (const T *)u, done);
    T *uc = vcycle_coarse<T>(st, H, l + 1, cp, P, done);
    PROF_LAUNCH(P, HMFEM_PROF_COARSE, (k_prolong_jacobi<T, T>), dim3(G), dim3(kBlock), 0, st, o, L.fx, L.fy, L.fz,
                C.d.nx, C.d.ny, uc, (T)cp.ocf, f, u, t, om, done);
    { T *tmp = u; u = t; t = tmp; }
    for (int q = 1; q < cp.nu; q++) {
        PROF_LAUNCH(P, HMFEM_PROF_COARSE, (k_jacobi<T, T>), dim3(G), dim3(kBlock), 0, st, o, f, u, t, om, done);
        T *tmp = u; u = t; t = tmp;
    }
    if (u != (T *)L.u) { L.t = L.u; L.u = u; }
    return u;
}

// Fine-level V-cycle: z = M r, with the CG dots fused into its last sweep.
// v is nullptr before the first iteration.
template <typename T>
static T *vcycle_fine(hipStream_t st, Hierarchy &H, const double *r, const double *v, const CycleParams &cp, Prof *P,
                      CgState &S)
{
    Level &L = H.lev[0];
    T *u = (T *)L.u, *t = (T *)L.t;
    const int *done = S.done;
    if (H.nlev == 1) {
        hipLaunchKernelGGL(k_dense_dots<T>, dim3(S.ch.G), dim3(kBlock), 0, st, H.nc, H.inv, r, v, u, S.Pdot, done);
        return u;
    }
    OpArgs<T> o = opargs<T>(L);
    const T om = (T)cp.omega;
    int G = grid_for(L.d.n(), kBlock);
    if (cp.smoother == HMFEM_SMOOTH_BLOCK_GS && H.f16sq) {
        // fine level stored in FP16 (scaled), FP32 arithmetic
        Level &C = H.lev[1];
        const int nt = gs_tiles(L.d);
        const Fine16 h{H.f16x, H.f16y, H.f16z, H.f16sq};
        PROF_LAUNCH(P, HMFEM_PROF_SMOOTH0, (k_gs_tile<T, double, 0, true>), dim3(grid_for(nt, 1)), dim3(kGsThreads), 0,
                    st, o, cp.gs, r, (const T *)nullptr, u, 0, 0, 0, 0, 0, (const T *)nullptr, (T)0,
                    (const double *)nullptr, 1, (double *)nullptr, done, h);
        if (L.fx == 2 && L.fy == 2 && L.fz == 2) {
            const int nb = ((L.d.nx + 31) / 32) * ((L.d.ny + 3) / 4) * ((L.d.nz + 1) / 2);
            PROF_LAUNCH(P, HMFEM_PROF_RESTRICT, (k_restrict_brick16<T, double>), dim3(grid_for(nb, 1)), dim3(kBlock),
                        0, st, L.d.nx, L.d.ny, L.d.nz, h, C.d.nx, C.d.ny, C.d.nz, r, (const T *)u, (T *)C.f, done);
        } else {
            launch_restrict<T, double>(st, P, HMFEM_PROF_RESTRICT, L, C, r, (const T *)u, done);
        }
        T *uc = vcycle_coarse<T>(st, H, 1, cp, P, done);
        const int tpb = (nt + S.ch.G - 1) / S.ch.G;
        PROF_LAUNCH(P, HMFEM_PROF_PROLONG, (k_gs_tile<T, double, 2, true>), dim3(S.ch.G), dim3(kGsThreads), 0, st, o,
                    cp.gs, r, (const T *)u, t, lg2f(L.fx), lg2f(L.fy), lg2f(L.fz), C.d.nx, C.d.ny, (const T *)uc,
                    (T)cp.ocf, v, tpb, S.Pdot, done, h);
        return t;
    }
    if (cp.smoother == HMFEM_SMOOTH_BLOCK_GS) {
        Level &C = H.lev[1];
        const int nt = gs_tiles(L.d);
        PROF_LAUNCH(P, HMFEM_PROF_SMOOTH0, (k_gs_tile<T, double, 0>), dim3(grid_for(nt, 1)), dim3(kGsThreads), 0, st, o,
                    cp.gs, r, (const T *)nullptr, u, 0, 0, 0, 0, 0, (const T *)nullptr, (T)0,
                    (const double *)nullptr, 1, (double *)nullptr, done);
        launch_restrict<T, double>(st, P, HMFEM_PROF_RESTRICT, L, C, r, (const T *)u, done);
        T *uc = vcycle_coarse<T>(st, H, 1, cp, P, done);
        // the dot partials must come from exactly S.ch.G blocks
        const int tpb = (nt + S.ch.G - 1) / S.ch.G;
        PROF_LAUNCH(P, HMFEM_PROF_PROLONG, (k_gs_tile<T, double, 2>), dim3(S.ch.G), dim3(kGsThreads), 0, st, o, cp.gs,
                    r, (const T *)u, t, lg2f(L.fx), lg2f(L.fy), lg2f(L.fz), C.d.nx, C.d.ny, (const T *)uc,
                    (T)cp.ocf, v, tpb, S.Pdot, done);
        return t;
    }
    PROF_LAUNCH(P, HMFEM_PROF_SMOOTH0, (k_smooth_zero<T, double>), dim3(G), dim3(kBlock), 0, st, o, r, u, om,
                cp.nu >= 2 ? 2 : 1, done);
    for (int q = 2; q < cp.nu; q++) {
        PROF_LAUNCH(P, HMFEM_PROF_JACOBI, (k_jacobi<T, double>), dim3(G), dim3(kBlock), 0, st, o, r, u, t, om, done);
        T *tmp = u; u = t; t = tmp;
    }
    Level &C = H.lev[1];
    launch_restrict<T, double>(st, P, HMFEM_PROF_RESTRICT, L, C, r, (const T *)u, done);
    T *uc = vcycle_coarse<T>(st, H, 1, cp, P, done);
    if (cp.nu == 1) {
        PROF_LAUNCH(P, HMFEM_PROF_PROLONG, k_final_sweep_dots<T>, dim3(S.ch.G), dim3(kBlock), 0, st, o, L.fx, L.fy,
                    C.d.nx, C.d.ny, L.fz, (const T *)uc, (T)cp.ocf, r, u, t, om, v, S.ch.chunk, S.Pdot, done);
        return t;
    }
    PROF_LAUNCH(P, HMFEM_PROF_PROLONG, (k_prolong_jacobi<T, double>), dim3(G), dim3(kBlock), 0, st, o, L.fx, L.fy,
                L.fz, C.d.nx, C.d.ny, uc, (T)cp.ocf, r, u, t, om, done);
    { T *tmp = u; u = t; t = tmp; }
    for (int q = 1; q + 1 < cp.nu; q++) {
        PROF_LAUNCH(P, HMFEM_PROF_JACOBI, (k_jacobi<T, double>), dim3(G), dim3(kBlock), 0, st, o, r, u, t, om, done);
        T *tmp = u; u = t; t = tmp;
    }
    PROF_LAUNCH(P, HMFEM_PROF_JACOBI, k_final_sweep_dots<T>, dim3(S.ch.G), dim3(kBlock), 0, st, o, 1, 1, 1, 1, 1,
                (const T *)nullptr, (T)0, r, u, t, om, v, S.ch.chunk, S.Pdot, done);
    return t;
}

struct SolveResult {
    int iterations, status;
    double rr, bb, sigma, sigma_first;
};

template <typename T>
static int run_cg(hipStream_t st, Hierarchy &H, const Stencil<double> &A, Dims d, const double *b, double *x,
                  const hmfem_options &opt, CgBuffers &B, CgScalars *h_sc, SolveResult &res, Prof *P,
                  GraphCache *graph, const Deflation *D, double *jz = nullptr)
{
    const int n = (int)d.n();
    OpArgs<double> o;
    o.e = A.e; o.wx = A.wx; o.wy = A.wy; o.wz = A.wz; o.di = nullptr;
    o.nx = d.nx; o.ny = d.ny; o.nz = d.nz;
    // gs_sweeps 0: two sweeps on a single grid (256^3 conduits: 69 CG iterations and 16.3 solves/s with two,
    // 80 and 14.1 with three; 128^3 lognormal: 17 and 287 against 16 and 290)
    CycleParams cp{opt.nu < 1 ? 1 : opt.nu, opt.omega, ocf_of(opt, kOcfSingle), opt.smoother,
                   opt.gs_sweeps < 1 ? kSweepsSingle : opt.gs_sweeps};
    CgState S;
    S.B = B;
    S.ch = chunking(n);
    const int G = S.ch.G;
    S.Pinit = B.partial;
    S.Prr = B.partial + 2 * kMaxReduceBlocks;
    S.Pdot = B.partial + 3 * kMaxReduceBlocks;                  // 2 (G + 1) <= 2 kMaxReduceBlocks + 2
    S.Ptau = B.partial + 5 * kMaxReduceBlocks + 64;             // up to kMaxSpmvBlocks
    S.done = &B.sc->done;

    CgScalars init{};
    init.rtol2 = opt.rtol * opt.rtol;
    init.maxit = opt.max_iterations;
    init.stop_rule = opt.stop_rule;
    init.xpend = -1;
    HMFEM_CHECK(hipMemcpyAsync(B.sc, &init, sizeof(CgScalars), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_cg_init, dim3(G), dim3(kBlock), 0, st, o, b, x, B.r, S.ch.chunk, S.Pinit);
    // z = M r: the V-cycle plus, where clusters of strong cells were found, the
    // additive cluster correction (its dot products are partial number G)
    const bool defl = D && D->nc > 0;
    const int Gdot = defl ? G + 1 : G;
    auto precond = [&](const double *vv, Prof *PP) -> T * {
        if (jz) {   // plain FP64 PCG: T is double
            PROF_LAUNCH(PP, HMFEM_PROF_JACOBI, k_jacobi_pc, dim3(G), dim3(kBlock), 0, st, o, B.r, vv, jz, S.ch.chunk,
                        S.Pdot, S.done);
            return (T *)jz;
        }
        if (defl) deflation_sums(st, *D, B.r, vv, S.Pdot + 2 * G, S.done);
        // the balancing form (opt.deflate 2): the V-cycle on r - A Z c1,
        // its output projected; the additive form (1, the default) adds Z c1
        // only (the balancing form took 64 instead of 69 CG iterations on the
        // 256^3 conduit case, but its iterate at the true relative residual
        // 1e-8 was 1.7e-6 from the converged pressure, against 1.0e-6)
        const bool bal = defl && opt.deflate == 2;
        if (bal) deflation_pre(st, *D, B.r, S.done);
        T *zz = vcycle_fine<T>(st, H, B.r, vv, cp, PP, S);
        if (bal) deflation_post(st, *D, A, B.r, zz, S.Pdot + 2 * G, S.done);
        if (defl) deflation_scatter(st, *D, zz, S.done);
        return zz;
    };
    T *z = precond(nullptr, P);
    HMFEM_CHECK(hipGetLastError());

    const int zlen = spmv_zlen(d);
    const int Gz = zlen > 0 ? spmv_z_blocks(d, zlen) : G;
    // One CG iteration; identical launches every iteration (the index and the
    // buffer parity are read from the device).
    auto iteration = [&](Prof *PP) -> T * {
        PROF_LAUNCH(PP, HMFEM_PROF_CG_DOTS, k_cg_scalars, dim3(1), dim3(kScalarThreads), 0, st, G, Gdot, B.sc,
                    (const double *)S.Pinit, (const double *)S.Prr, (const double *)S.Pdot);
        if (zlen > 0)
            PROF_LAUNCH(PP, HMFEM_PROF_CG_SPMV, k_cg_spmv_z<T>, dim3(Gz), dim3(kBlock), 0, st, o, (const T *)z, B.s0,
                        B.s1, B.v, x, zlen, G, Gdot, B.sc, S.Pinit, S.Prr, S.Pdot, S.Ptau, (int)A.eface);
        else
            PROF_LAUNCH(PP, HMFEM_PROF_CG_SPMV, k_cg_spmv<T>, dim3(G), dim3(kBlock), 0, st, o, (const T *)z, B.s0,
                        B.s1, B.v, S.ch.chunk, G, Gdot, B.sc, S.Pinit, S.Prr, S.Pdot, S.Ptau);
        PROF_LAUNCH(PP, HMFEM_PROF_CG_UPDATE, k_cg_update, dim3(G), dim3(kBlock), 0, st, n, B.s0, B.s1, B.v, x, B.r,
                    S.ch.chunk, zlen > 0 ? Gz : G, B.sc, S.Ptau, S.Prr, zlen > 0 ? 1 : 0);
        return precond(B.v, PP);
    };
    hipGraphExec_t exec = nullptr;
    if (graph && !P) {
        // key: everything the captured launches depend on
        std::vector<char> key;
        auto put = [&](const void *p, size_t nb) { key.insert(key.end(), (const char *)p, (const char *)p + nb); };
        put(&d, sizeof(d)); put(&opt, sizeof(opt)); put(&H.lev[0].s32.e, sizeof(void *));
        put(&A, sizeof(A)); put(&x, sizeof(x)); put(&B, sizeof(B)); put(&z, sizeof(z));
        put(&H.tail_ops, sizeof(void *)); put(&H.inv, sizeof(void *));
        if (defl) {   // the cluster correction's sizes and buffers
            put(&D->nc, sizeof(int)); put(&D->m, sizeof(int)); put(&D->cells, sizeof(void *));
            put(&D->comp, sizeof(void *)); put(&D->pb, sizeof(void *)); put(&D->einv, sizeof(void *));
            put(&D->np, sizeof(int)); put(&D->cp, sizeof(void *)); put(&D->c, sizeof(void *));
            put(&D->psr, sizeof(void *)); put(&D->alpha, sizeof(void *)); put(&D->rec_max, sizeof(int));
            put(&D->nbr, sizeof(void *)); put(&D->inst, sizeof(void *));
        }
        if (graph->exec && graph->key == key) exec = graph->exec;
        else {
            if (graph->exec) { (void)hipGraphExecDestroy(graph->exec); graph->exec = nullptr; }
            hipGraph_t g = nullptr;
            HMFEM_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            T *zc = iteration(nullptr);
            HMFEM_CHECK(hipStreamEndCapture(st, &g));
            HMFEM_CHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
            if (zc != z) { (void)hipGraphExecDestroy(exec); exec = nullptr; }   // z must be a fixed buffer
            else { graph->exec = exec; graph->key = key; }
        }
    }
    int k = 0;
    const int every = opt.check_every < 1 ? 1 : opt.check_every;
    while (true) {
        for (int q = 0; q < every; q++, k++) {
            if (exec) HMFEM_CHECK(hipGraphLaunch(exec, st));
            else z = iteration(P);
        }
        HMFEM_CHECK(hipGetLastError());
        HMFEM_CHECK(hipMemcpyAsync(h_sc, B.sc, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
        HMFEM_CHECK(hipStreamSynchronize(st));
        if (h_sc->done || k > opt.max_iterations + every) break;
    }
    if (zlen > 0) {   // the x update that no SpMV applied (the solve stopped at the SpMV that would have)
        hipLaunchKernelGGL(k_cg_finish_x, dim3(grid_for(std::min<long>(n, 4096L * kBlock), kBlock)), dim3(kBlock), 0,
                           st, n, B.s0, B.s1, x, B.sc);
        HMFEM_CHECK(hipGetLastError());
    }
    res.iterations = h_sc->iter;
    res.status = h_sc->done ? h_sc->status : HMFEM_NOT_CONVERGED;
    res.rr = h_sc->rr;
    res.bb = h_sc->bb;
    res.sigma = h_sc->sigma[h_sc->iter & 1];
    res.sigma_first = h_sc->sigma_first;
    return HMFEM_OK;
}

int hierarchy_level_copy(hipStream_t st, const SetupCache &cache, int l, int *dims, int *fac, double *e, double *wx,
                         double *wy, double *wz)
{
    if (!cache.valid || cache.hier.size() != sizeof(Hierarchy)) return HMFEM_ERR_ARG;
    Hierarchy H;
    memcpy((void *)&H, cache.hier.data(), sizeof(Hierarchy));
    if (l < 0) return H.nlev;
    if (l >= H.nlev) return HMFEM_ERR_ARG;
    const Level &L = H.lev[l];
    if (dims) { dims[0] = L.d.nx; dims[1] = L.d.ny; dims[2] = L.d.nz; }
    if (fac) { fac[0] = L.fx; fac[1] = L.fy; fac[2] = L.fz; }
    const size_t nb = (size_t)L.d.n() * sizeof(double);
    if (e) HMFEM_CHECK(hipMemcpyAsync(e, L.s64.e, nb, hipMemcpyDeviceToHost, st));
    if (wx) HMFEM_CHECK(hipMemcpyAsync(wx, L.s64.wx, nb, hipMemcpyDeviceToHost, st));
    if (wy) HMFEM_CHECK(hipMemcpyAsync(wy, L.s64.wy, nb, hipMemcpyDeviceToHost, st));
    if (wz) HMFEM_CHECK(hipMemcpyAsync(wz, L.s64.wz, nb, hipMemcpyDeviceToHost, st));
    HMFEM_CHECK(hipStreamSynchronize(st));
    return H.nlev;
}

int hierarchy_coarse_inverse(hipStream_t st, const SetupCache &cache, double *inv)
{
    if (!cache.valid || cache.hier.size() != sizeof(Hierarchy)) return HMFEM_ERR_ARG;
    Hierarchy H;
    memcpy((void *)&H, cache.hier.data(), sizeof(Hierarchy));
    if (inv) {
        HMFEM_CHECK(hipMemcpyAsync(inv, H.inv, (size_t)H.nc * H.nc * sizeof(double), hipMemcpyDeviceToHost, st));
        HMFEM_CHECK(hipStreamSynchronize(st));
    }
    return H.nc;
}

int solve_structured(hipStream_t st, Dims d, const Stencil<double> &A, const double *b, double *x,
                     const hmfem_options &opt, char *arena, size_t arena_bytes, CgBuffers &B, CgScalars *h_sc,
                     int *d_info, hmfem_status *out, hipEvent_t ev[4], ProfileSink *sink, std::vector<char> &stage,
                     GraphCache *graph, long lds_limit, Deflation *defl, char *(*dalloc)(void *, size_t, int),
                     void *dctx, SetupCache *cache, const PlanHint *hint)
{
    Prof prof, *P = nullptr;
    if (sink && sink->ev && sink->cap >= 2) {
        prof.ev = sink->ev; prof.type = sink->type; prof.cap = sink->cap; prof.st = st;
        P = &prof;
    }
    if (opt.precision == HMFEM_PREC_FP16 && opt.smoother != HMFEM_SMOOTH_BLOCK_GS) return HMFEM_ERR_ARG;
    Arena ar{arena, 0, arena_bytes};
    Hierarchy H;
    int prec = opt.precision == HMFEM_PREC_FP64 ? HMFEM_PREC_FP64 : HMFEM_PREC_FP32;
    HMFEM_CHECK(hipEventRecord(ev[0], st));
    int rc = HMFEM_OK;
    if (opt.precond == HMFEM_PRECOND_JACOBI) {   // plain FP64 PCG: no hierarchy, no correction
        if (defl) { defl->nc = 0; defl->pending = false; defl->gen++; }
        if (cache) cache->valid = false;   // z below overwrites the start of the arena
        double *jz = ar.take<double>(d.n());
        if (!jz) return HMFEM_ERR_NOMEM;
        HMFEM_CHECK(hipEventRecord(ev[1], st));
        SolveResult res{};
        rc = run_cg<double>(st, H, A, d, b, x, opt, B, h_sc, res, P, opt.use_graph ? graph : nullptr, nullptr, jz);
        if (rc) return rc;
        HMFEM_CHECK(hipEventRecord(ev[2], st));
        HMFEM_CHECK(hipEventSynchronize(ev[2]));
        if (out) {
            float t0 = 0, t1 = 0;
            (void)hipEventElapsedTime(&t0, ev[0], ev[1]);
            (void)hipEventElapsedTime(&t1, ev[1], ev[2]);
            out->t_setup_ms = t0;
            out->t_solve_ms = t1;
            out->iterations = res.iterations;
            out->levels = 0;
            out->status = res.status;
            out->rel_residual = res.bb > 0 ? sqrt(res.rr / res.bb) : 0.0;
            out->precond_ratio = res.sigma_first > 0 ? sqrt(res.sigma / res.sigma_first) : 0.0;
            out->deflation_clusters = 0;
            out->deflation_cells = 0;
            // per iteration: SpMV (A64 32 B, s in, z in, s out, v out; with the
            // x update also x in and out), update (r, v in; r out; without the
            // x update also x, s in and x out), Jacobi (A64, r, v in; z out)
            const double xb = spmv_zlen(d) > 0 ? 16.0 : 0.0;
            out->solve_bytes =
                (double)d.n() * ((32 + 4 * 8 + xb) + (48 - 1.5 * xb) + (32 + 3 * 8)) * (res.iterations + 1);
        }
        return HMFEM_OK;
    }
    // what the setup depends on, for opt.reuse_setup
    std::vector<char> key;
    {
        auto put = [&](const void *p, size_t nb) { key.insert(key.end(), (const char *)p, (const char *)p + nb); };
        const int ints[] = {opt.precision, opt.coarse_max, opt.tail_max, opt.tail_gamma, opt.tail_lds, opt.nu,
                            opt.smoother, opt.deflate};
        put(&d, sizeof(d)); put(&arena, sizeof(arena)); put(&arena_bytes, sizeof(arena_bytes));
        put(ints, sizeof(ints)); put(&opt.deflate_ratio, sizeof(double)); put(&lds_limit, sizeof(lds_limit));
        Level lev[kMaxLevels];   // the coarsening plan
        int nlev = 0;
        plan_levels(d, clamp_coarse(opt.coarse_max), lev, nlev, hint);
        for (int l = 0; l < nlev; l++) {
            const int f[3] = {lev[l].fx, lev[l].fy, lev[l].fz};
            put(f, sizeof(f));
        }
    }
    static_assert(std::is_trivially_copyable<Hierarchy>::value, "the hierarchy is cached as bytes");
    const bool reuse = opt.reuse_setup && cache && cache->valid && cache->key == key &&
                       cache->hier.size() == sizeof(Hierarchy) && (!defl || defl->gen == cache->dgen);
    if (reuse) {
        memcpy((void *)&H, cache->hier.data(), sizeof(Hierarchy));
    } else {
        if (cache) cache->valid = false;
        if (defl) {   // the histogram first: its copy to the host overlaps the hierarchy setup
            defl->nc = 0;
            defl->pending = false;
            defl->gen++;
            if (opt.deflate && dalloc) {
                // in the FP32 hierarchy the conversion pass takes the histogram
                rc = deflation_begin(st, d, A, *defl, dalloc, dctx, opt.precision != HMFEM_PREC_FP64);
                if (rc) return rc;
            }
        }
        rc = build_hierarchy(st, d, A, opt, ar, H, d_info, stage, lds_limit, hint, defl);
        if (rc) return rc;
        if (defl && defl->pending) {
            rc = deflation_finish(st, d, A, opt.deflate_ratio, *defl, dalloc, dctx, opt.deflate == 2);
            if (rc) return rc;
        }
        if (cache) {
            cache->key = key;
            cache->hier.assign((const char *)&H, (const char *)&H + sizeof(Hierarchy));
            cache->dgen = defl ? defl->gen : 0;
            cache->valid = true;
        }
    }
    HMFEM_CHECK(hipEventRecord(ev[1], st));
    SolveResult res{};
    GraphCache *gc = opt.use_graph ? graph : nullptr;
    if (prec == HMFEM_PREC_FP64) rc = run_cg<double>(st, H, A, d, b, x, opt, B, h_sc, res, P, gc, defl);
    else rc = run_cg<float>(st, H, A, d, b, x, opt, B, h_sc, res, P, gc, defl);
    if (rc) return rc;
    HMFEM_CHECK(hipEventRecord(ev[2], st));
    HMFEM_CHECK(hipEventSynchronize(ev[2]));
    int info = 0;
    HMFEM_CHECK(hipMemcpy(&info, d_info, sizeof(int), hipMemcpyDeviceToHost));
    if (P) {
        // Launches queued after the solve had converged (the host checks every
        // check_every iterations) return at once; they are left out: a launch
        // counts if it ran for at least a quarter of its class's median.
        std::vector<float> dur(P->used / 2, 0.f);
        std::vector<float> per[HMFEM_PROF_TYPES];
        float median[HMFEM_PROF_TYPES] = {};
        for (int q = 0; q + 1 < P->used; q += 2) {
            (void)hipEventElapsedTime(&dur[q / 2], P->ev[q], P->ev[q + 1]);
            int t = P->type[q / 2];
            if (t >= 0 && t < HMFEM_PROF_TYPES) per[t].push_back(dur[q / 2]);
        }
        for (int t = 0; t < HMFEM_PROF_TYPES; t++) {
            if (per[t].empty()) continue;
            std::sort(per[t].begin(), per[t].end());
            median[t] = per[t][per[t].size() / 2];
        }
        for (int q = 0; q + 1 < P->used; q += 2) {
            int t = P->type[q / 2];
            if (t < 0 || t >= HMFEM_PROF_TYPES || dur[q / 2] < 0.25f * median[t]) continue;
            sink->ms[t] += dur[q / 2];
            sink->count[t] += 1;
        }
        // algorithmic bytes per launch of the fine-level kernels (actual cells)
        double n0 = (double)d.n();
        double vs = prec == HMFEM_PREC_FP64 ? 8.0 : 4.0;
        double c4 = 4.0 * vs;
        double n1 = (double)H.lev[H.nlev > 1 ? 1 : 0].d.n();
        // A64, s_in, z, s_out, v (+ x in and out where the SpMV applies the x update)
        const double xb = spmv_zlen(d) > 0 ? 16.0 : 0.0;
        // with eface the z-marching SpMV and restriction read e on the face cells only
        const bool ez = A.eface && xb > 0;
        const double nin = (double)std::max(d.nx - 2, 0) * std::max(d.ny - 2, 0) * std::max(d.nz - 2, 0);
        const double e64 = ez ? 8.0 * (n0 - nin) : 8.0 * n0;
        const double er = A.eface && H.lev[0].fx == 2 && H.lev[0].fy == 2 && H.lev[0].fz == 2 ? vs * (n0 - nin) : vs * n0;
        sink->bytes[HMFEM_PROF_CG_SPMV] = e64 + n0 * (24 + 8 + vs + 8 + 8 + xb);
        sink->bytes[HMFEM_PROF_CG_UPDATE] = n0 * (48 - 1.5 * xb);           // r, v in; r out (+ x, s in; x out)
        sink->bytes[HMFEM_PROF_CG_DOTS] = 0.0;                               // one workgroup over the partial sums
        sink->bytes[HMFEM_PROF_SMOOTH0] = n0 * (c4 + 8 + vs);                // A, f, u out
        sink->bytes[HMFEM_PROF_RESTRICT] = er + n0 * (c4 - vs + 8 + vs) + n1 * vs;   // A, f, u; f_c out
        sink->bytes[HMFEM_PROF_PROLONG] = n0 * (c4 + 8 + 2 * vs) + n1 * vs;  // A, f, u, u_c; u out
        sink->bytes[HMFEM_PROF_JACOBI] = n0 * (c4 + 8 + 2 * vs + 8);         // A, f, u, v; z out (+dots)
        if (opt.smoother == HMFEM_SMOOTH_BLOCK_GS) {
            // pre: wx, wy, wz, di, f; u out. post: the same plus u, u_c, v; z out (+dots)
            sink->bytes[HMFEM_PROF_PROLONG] = n0 * (c4 + 8 + 2 * vs + 8) + n1 * vs;
            sink->bytes[HMFEM_PROF_JACOBI] = 0.0;
            if (H.f16sq) {   // FP16 couplings (6 B) and the FP32 scale (4 B) instead of 16 B
                sink->bytes[HMFEM_PROF_SMOOTH0] = n0 * (10 + 8 + 4);
                sink->bytes[HMFEM_PROF_RESTRICT] = n0 * (10 + 8 + 4) + n1 * 4;
                sink->bytes[HMFEM_PROF_PROLONG] = n0 * (10 + 8 + 8 + 8) + n1 * 4;
            }
        }
    }
    if (out) {
        float t0 = 0, t1 = 0;
        (void)hipEventElapsedTime(&t0, ev[0], ev[1]);
        (void)hipEventElapsedTime(&t1, ev[1], ev[2]);
        out->t_setup_ms = t0;
        out->t_solve_ms = t1;
        out->iterations = res.iterations;
        out->levels = H.nlev;
        out->deflation_clusters = defl ? defl->nc : 0;
        out->deflation_cells = defl && defl->nc ? defl->m : 0;
        out->status = info ? HMFEM_ERR_BREAKDOWN : res.status;
        out->rel_residual = res.bb > 0 ? sqrt(res.rr / res.bb) : 0.0;
        out->precond_ratio = res.sigma_first > 0 ? sqrt(res.sigma / res.sigma_first) : 0.0;
        // algorithmic HBM bytes of the solve phase: per CG iteration the
        // fine-level kernels (CG SpMV, update, V-cycle sweeps) plus the
        // multi-block coarse levels
        double vs = prec == HMFEM_PREC_FP64 ? 8.0 : 4.0, cs = vs;
        double per = 0.0;
        int lim = H.tail0 > 0 ? H.tail0 : H.nlev - 1;
        for (int l = 0; l < lim; l++) {
            double nl = (double)H.lev[l].d.n();
            double fb = l == 0 ? 8.0 : vs;
            if (opt.smoother == HMFEM_SMOOTH_BLOCK_GS) {
                // one pass each: pre-smoothing, residual + restriction, post-smoothing
                per += nl * (3 * (4 * cs + fb) + 4 * vs);
                continue;
            }
            double sweeps = 2.0 * opt.nu;
            per += nl * (sweeps * (4 * cs + fb + 2 * vs) + (4 * cs + fb + vs));
        }
        double n0 = (double)d.n();
        const double xb = spmv_zlen(d) > 0 ? 16.0 : 0.0;   // x update in the SpMV (24 B less in the update)
        per += n0 * (32 + 8 + vs + 8 + 8 + xb) + n0 * (48 - 1.5 * xb) + n0 * 8;
        out->solve_bytes = per * (res.iterations + 1);
    }
    return HMFEM_OK;
}

int launch_tail_strided(hipStream_t st, const Hierarchy &H, double omega, double ocf, const float *ftop,
                        float *utop, int stride, const int *done)
{
    if (H.prec != HMFEM_PREC_FP32 || H.tail0 <= 0 || !tail_in_lds(H))
        return HMFEM_ERR_ARG;
    hipLaunchKernelGGL(k_tail_lds<float>, dim3(1), dim3(kTailThreads), H.tail_lds_bytes, st,
                       (const TailLevel<float> *)H.tail_levels, H.tail_ops, H.tail_nops, H.inv, H.nc, (float)omega,
                       (float)ocf, H.nlev - H.tail0, H.tail_lds_inv, ftop, utop, stride, done);
    return HMFEM_OK;
}

size_t cg_partial_count(long n)
{
    (void)n;
    return (size_t)5 * kMaxReduceBlocks + 64 + kMaxSpmvBlocks;
}

} // namespace hmfem
