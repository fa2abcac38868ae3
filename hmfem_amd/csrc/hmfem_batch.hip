// Batched pressure solve: many independent grids of uneven sizes solved
// together, every launch covering all of them.
//
// Layout. Level l of the batch concatenates level l of every instance
// (instance q at cells [off, off + n) of each array). Instances are given the
// same number of levels: once an instance is at most `coarse_max` cells it is
// carried down with identity aggregation (factors 1), so each level holds
// every instance and one launch per V-cycle step serves the whole batch.
// Blocks never straddle instances; bmap[block] names the instance.
//
// CG. Each instance has its own scalars (alpha, beta, sigma, iteration
// index), its own stopping test and its own partial sums (the blocks of one
// instance sum only their own partials). Kernels skip converged instances, so
// a finished instance costs nothing while the others iterate. One captured
// HIP graph replays a CG iteration of the whole batch.
#include "hmfem_gs.h"

#include <string.h>

#include <algorithm>
#include <vector>

namespace hmfem {

struct BInst {
    int nx, ny, nz, n, off, blk0, nblk;   // this level
    int fx, fy, fz;                       // aggregation factors to the next level
    int tile0, ntile;                     // Gauss-Seidel tiles of this instance
    int zblk0, zG, zlen;                  // kb_restrict_z blocks (zlen 0: not 2x2x2, kb_resid_restrict)
    int gs;                               // the instance's own Gauss-Seidel sweeps (0: the batch's)
    float ocf;                            // the instance's own coarse-correction scaling (0: the batch's)
};

constexpr int kBatchZlen = 16;   // planes per block of the batch's z-marching SpMV
// Planes per block of the batch's residual-restriction: short columns give
// the small instances more blocks (96 instances of 32^3 to 96^3: 1887
// solves/s with 4, 1862 with 2, 1876 with 6, 1860 with 8, 1838 with 12,
// 1799 with 16, 1730 with 32; same iterations, as the coarse sums do not
// depend on it).
constexpr int kBatchRestrictZlen = 4;

struct CInst {                            // fine level, chunked launches (CG, final sweep)
    int nx, ny, nz, n, off, cblk0, cG, chunk;
    int tpb;                              // smoother tiles per chunk block
    int zblk0, zG, zlen;                  // z-marching SpMV: first block, blocks, planes per block
};

template <typename T>
struct LPtr {
    const T *e, *wx, *wy, *wz, *di;
};

template <typename T>
__device__ __forceinline__ OpArgs<T> bop(const LPtr<T> &A, int off, int nx, int ny, int nz)
{
    OpArgs<T> o;
    o.e = A.e + off; o.wx = A.wx + off; o.wy = A.wy + off; o.wz = A.wz + off;
    o.di = A.di ? A.di + off : nullptr;
    o.nx = nx; o.ny = ny; o.nz = nz;
    return o;
}

// ---------------------------------------------------------------- V-cycle kernels

template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
kb_smooth_zero(const BInst *__restrict__ D, const int *__restrict__ bmap, LPtr<T> A, const FT *__restrict__ f,
               T *__restrict__ u, T omega, int sweeps, const CgScalars *sc)
{
    using C = T;
    const int q = bmap[blockIdx.x];
    if (sc[q].done) return;
    const BInst I = D[q];
    const int s = (blockIdx.x - I.blk0) * kBlock + threadIdx.x;
    if (s >= I.n) return;
    const OpArgs<T> o = bop(A, I.off, I.nx, I.ny, I.nz);
    const FT *fl = f + I.off;
    T *ul = u + I.off;
    const C ds = (C)o.di[s];
    if (sweeps == 1) { ul[s] = omega * (C)fl[s] * ds; return; }
    Cell c = cell_at(o, s);
    auto u1 = [&](const Cell &x) -> C { return omega * (C)fl[x.s] * (C)o.di[x.s]; };
    C d;
    C au = apply_at<T, C>(o, c, u1, d);
    ul[s] = omega * (C)fl[s] * ds + omega * ((C)fl[s] - au) * ds;
}

template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
kb_jacobi(const BInst *__restrict__ D, const int *__restrict__ bmap, LPtr<T> A, const FT *__restrict__ f,
          const T *__restrict__ u, T *__restrict__ uo, T omega, const CgScalars *sc)
{
    using C = T;
    const int q = bmap[blockIdx.x];
    if (sc[q].done) return;
    const BInst I = D[q];
    const int s = (blockIdx.x - I.blk0) * kBlock + threadIdx.x;
    if (s >= I.n) return;
    const OpArgs<T> o = bop(A, I.off, I.nx, I.ny, I.nz);
    const FT *fl = f + I.off;
    const T *ul = u + I.off;
    Cell c = cell_at(o, s);
    C d;
    C au = apply_at<T, C>(o, c, [&](const Cell &x) -> C { return (C)ul[x.s]; }, d);
    uo[I.off + s] = (C)ul[s] + omega * ((C)fl[s] - au) * rcp(d);
}

// One thread per coarse cell of instance q (launched over the coarse level).
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
kb_resid_restrict(const BInst *__restrict__ D, const BInst *__restrict__ Dc, const int *__restrict__ bmapc, LPtr<T> A,
                  const FT *__restrict__ f, const T *__restrict__ u, T *__restrict__ fc, const CgScalars *sc)
{
    using C = T;
    const int q = bmapc[blockIdx.x];
    if (sc[q].done || D[q].zlen > 0) return;   // 2x2x2 instances: kb_restrict_z
    const BInst Ic = Dc[q];
    const int I = (blockIdx.x - Ic.blk0) * kBlock + threadIdx.x;
    if (I >= Ic.n) return;
    const BInst If = D[q];
    const OpArgs<T> o = bop(A, If.off, If.nx, If.ny, If.nz);
    const FT *fl = f + If.off;
    const T *ul = u + If.off;
    const int cx = Ic.nx, cy = Ic.ny;
    const int ck = I / (cx * cy);
    const int rem = I - ck * cx * cy;
    const int cj = rem / cx, ci = rem - cj * cx;
    C acc = 0;
    for (int dk = 0; dk < If.fz; dk++) {
        int k = ck * If.fz + dk;
        if (k >= o.nz) break;
        for (int dj = 0; dj < If.fy; dj++) {
            int j = cj * If.fy + dj;
            if (j >= o.ny) break;
            for (int di = 0; di < If.fx; di++) {
                int i = ci * If.fx + di;
                if (i >= o.nx) break;
                Cell c;
                c.i = i; c.j = j; c.k = k;
                c.s = i + j * o.nx + k * o.nx * o.ny;
                C d;
                C au = apply_at<T, C>(o, c, [&](const Cell &x) -> C { return (C)ul[x.s]; }, d);
                acc += (C)fl[c.s] - au;
            }
        }
    }
    fc[Ic.off + I] = acc;
}

// Residual and restriction f_c = P^T (f - A u) of the instances that
// aggregate 2x2x2 boxes, marching along z as the single-grid kernel: a block
// owns a 32 x 8 column tile of its instance and zlen (even) planes, keeps its
// column's u at k-1, k, k+1 in registers and the plane with its x/y ring in a
// double-buffered LDS plane, sums each box over two planes in a register and
// over x and y by lane shuffles. zmap[block] names the instance.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
kb_restrict_z(const BInst *__restrict__ D, const BInst *__restrict__ Dc, const int *__restrict__ zmap, LPtr<T> A,
              const FT *__restrict__ f, const T *__restrict__ u, T *__restrict__ fc, const CgScalars *sc)
{
    const int q = zmap[blockIdx.x];
    if (sc[q].done) return;
    const BInst If = D[q], Ic = Dc[q];
    const OpArgs<T> o = bop(A, If.off, If.nx, If.ny, If.nz);
    const FT *__restrict__ fl = f + If.off;
    const T *__restrict__ ul = u + If.off;
    T *__restrict__ fcl = fc + Ic.off;
    const int cx = Ic.nx, cy = Ic.ny, cz = Ic.nz, zlen = If.zlen;
    __shared__ T sh[2][kGsTy + 2][kGsTx + 2];
    const int nx = o.nx, ny = o.ny, nz = o.nz;
    const long nxy = (long)nx * ny;
    const int ntx = (nx + kGsTx - 1) / kGsTx, nty = (ny + kGsTy - 1) / kGsTy;
    const int ntxy = ntx * nty;
    const int lb = blockIdx.x - If.zblk0;
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
    T um = (T)0, uc = (T)0, up = (T)0, wzm = (T)0;
    if (col) {
        if (kb > 0) { um = ul[cs + (kb - 1) * nxy]; wzm = o.wz[cs + (kb - 1) * nxy]; }
        uc = ul[cs + kb * nxy];
        if (kb + 1 < nz) up = ul[cs + (kb + 1) * nxy];
    }
    T acc = (T)0;
    for (int kk = kb; kk < ke; kk++) {
        const int bf = kk & 1;
        sh[bf][ty + 1][tx + 1] = uc;
        if (has_ring) sh[bf][ry][rx] = ring_in ? ul[rs + kk * nxy] : (T)0;
        T r = (T)0, e = (T)0, xp = (T)0, xm = (T)0, yp = (T)0, ym = (T)0, zp = (T)0, fq = (T)0, un = (T)0;
        if (col) {
            const long g = cs + kk * nxy;
            e = o.e[g];
            if (i < nx - 1) xp = o.wx[g];
            if (i > 0) xm = o.wx[g - 1];
            if (j < ny - 1) yp = o.wy[g];
            if (j > 0) ym = o.wy[g - nx];
            if (kk < nz - 1) zp = o.wz[g];
            fq = (T)fl[g];
            if (kk + 2 < nz) un = ul[g + 2 * nxy];
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
            if (!(tx & 1) && !(ty & 1) && ci < cx && cj < cy && ck < cz) fcl[ci + (long)cx * (cj + (long)cy * ck)] = acc;
            acc = (T)0;
        }
    }
}

template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
kb_prolong_jacobi(const BInst *__restrict__ D, const BInst *__restrict__ Dc, const int *__restrict__ bmap, LPtr<T> A,
                  const T *__restrict__ uc, T ocf, const FT *__restrict__ f, const T *__restrict__ u,
                  T *__restrict__ uo, T omega, const CgScalars *sc)
{
    using C = T;
    const int q = bmap[blockIdx.x];
    if (sc[q].done) return;
    const BInst I = D[q];
    const int s = (blockIdx.x - I.blk0) * kBlock + threadIdx.x;
    if (s >= I.n) return;
    const BInst Ic = Dc[q];
    if (I.ocf > 0.f) ocf = (T)I.ocf;
    const OpArgs<T> o = bop(A, I.off, I.nx, I.ny, I.nz);
    const FT *fl = f + I.off;
    const T *ul = u + I.off;
    const T *ucl = uc + Ic.off;
    const int fx = I.fx, fy = I.fy, fz = I.fz, cx = Ic.nx, cy = Ic.ny;
    Cell c = cell_at(o, s);
    auto up = [&](const Cell &x) -> C {
        return (C)ul[x.s] + ocf * (C)ucl[x.i / fx + cx * (x.j / fy + cy * (x.k / fz))];
    };
    C d;
    C au = apply_at<T, C>(o, c, up, d);
    uo[I.off + s] = up(c) + omega * ((C)fl[s] - au) * rcp(d);
}

// Last fine sweep with the CG dots of each instance (chunked launch).
template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_final_sweep_dots(const CInst *__restrict__ CI, const int *__restrict__ cbmap, LPtr<T> A,
                    const double *__restrict__ f, const T *__restrict__ u, T *__restrict__ uo, T omega,
                    const double *__restrict__ v, double *Pdot, const CgScalars *sc)
{
    using C = T;
    const int q = cbmap[blockIdx.x];
    if (sc[q].done) return;
    const CInst I = CI[q];
    const OpArgs<T> o = bop(A, I.off, I.nx, I.ny, I.nz);
    const double *fl = f + I.off;
    const T *ul = u + I.off;
    T *uol = uo + I.off;
    const double *vl = v ? v + I.off : nullptr;
    const int lb = blockIdx.x - I.cblk0;
    const int beg = lb * I.chunk, end = min(I.n, beg + I.chunk);
    double acc[2] = {0.0, 0.0};
    for (int base = beg; base < end; base += kBlock) {
        int s = base + threadIdx.x;
        if (s < end) {
            Cell c = cell_at(o, s);
            C d;
            C au = apply_at<T, C>(o, c, [&](const Cell &x) -> C { return (C)ul[x.s]; }, d);
            double fs = fl[s];
            C z = (C)ul[s] + omega * ((C)fs - au) * rcp(d);
            uol[s] = z;
            acc[0] += (double)z * fs;
            if (vl) acc[1] += (double)z * vl[s];
        }
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) { Pdot[2 * blockIdx.x] = acc[0]; Pdot[2 * blockIdx.x + 1] = acc[1]; }
}

// Block Gauss-Seidel smoothing (MODE 0 from zero, MODE 1 after the coarse
// correction) of one tile per block; tmap gives the block's instance.
template <typename T, typename FT, int MODE>
__global__ void __launch_bounds__(kGsThreads)
kb_gs_tile(const BInst *__restrict__ D, const BInst *__restrict__ Dc, const int *__restrict__ tmap, LPtr<T> A,
           int sweeps, const FT *__restrict__ f, const T *__restrict__ u, T *__restrict__ uo,
           const T *__restrict__ uc, T ocf, const CgScalars *sc)
{
    const int q = tmap[blockIdx.x];
    if (sc[q].done) return;
    const BInst I = D[q];
    if (I.gs > 0) sweeps = I.gs;
    if (I.ocf > 0.f) ocf = (T)I.ocf;
    const OpArgs<T> o = bop(A, I.off, I.nx, I.ny, I.nz);
    double acc[2] = {0.0, 0.0};
    if (MODE == 0) {
        gs_tile_body<T, FT, 0>(o, sweeps, f + I.off, (const T *)nullptr, uo + I.off, 0, 0, 0, 0, 0,
                               (const T *)nullptr, (T)0, (const double *)nullptr, blockIdx.x - I.tile0, acc);
    } else {
        const BInst Ic = Dc[q];
        gs_tile_body<T, FT, 1>(o, sweeps, f + I.off, u + I.off, uo + I.off, I.fx == 2, I.fy == 2, I.fz == 2, Ic.nx,
                               Ic.ny, uc + Ic.off, ocf, (const double *)nullptr, blockIdx.x - I.tile0, acc);
    }
}

// Post-smoothing of the fine level with the CG dots of each instance, over
// the instance's chunk blocks (tiles [lb tpb, (lb + 1) tpb) of it each).
template <typename T>
__global__ void __launch_bounds__(kGsThreads)
kb_gs_tile_dots(const CInst *__restrict__ CI, const int *__restrict__ cbmap, const BInst *__restrict__ D,
                const BInst *__restrict__ Dc, LPtr<T> A, int sweeps, const double *__restrict__ f,
                const T *__restrict__ u, T *__restrict__ uo, const T *__restrict__ uc, T ocf,
                const double *__restrict__ v, double *Pdot, const CgScalars *sc)
{
    const int q = cbmap[blockIdx.x];
    if (sc[q].done) return;
    const CInst C = CI[q];
    const BInst I = D[q], Ic = Dc[q];
    if (I.gs > 0) sweeps = I.gs;
    if (I.ocf > 0.f) ocf = (T)I.ocf;
    const OpArgs<T> o = bop(A, I.off, I.nx, I.ny, I.nz);
    const int lb = blockIdx.x - C.cblk0;
    const int tbeg = lb * C.tpb, tend = min(I.ntile, tbeg + C.tpb);
    double acc[2] = {0.0, 0.0};
    for (int t = tbeg; t < tend; t++)
        gs_tile_body<T, double, 2>(o, sweeps, f + I.off, u + I.off, uo + I.off, I.fx == 2, I.fy == 2, I.fz == 2,
                                   Ic.nx, Ic.ny, uc + Ic.off, ocf, v ? v + I.off : nullptr, t, acc);
    block_sum_n<2, kGsThreads>(acc);
    if (threadIdx.x == 0) { Pdot[2 * blockIdx.x] = acc[0]; Pdot[2 * blockIdx.x + 1] = acc[1]; }
}

// Dense coarsest solve of instance q = blockIdx.x.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
kb_dense_apply(const BInst *__restrict__ D, const double *__restrict__ inv, const FT *__restrict__ f,
               T *__restrict__ u, const CgScalars *sc)
{
    __shared__ double fs[kCoarseMax];
    const int q = blockIdx.x;
    if (sc[q].done) return;
    const BInst I = D[q];
    const int n = I.n;
    for (int j = threadIdx.x; j < n; j += kBlock) fs[j] = (double)f[I.off + j];
    __syncthreads();
    const double *iv = inv + (long)q * kCoarseMax * kCoarseMax;
    int L = 16;
    while (L > 1 && L * n > kBlock) L >>= 1;
    const int row = threadIdx.x / L, lane = threadIdx.x % L;
    double acc = 0.0;
    if (row < n)
        for (int j = lane; j < n; j += L) acc += iv[row * n + j] * fs[j];
    for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (row < n && lane == 0) u[I.off + row] = (T)acc;
}

// ---------------------------------------------------------------- setup kernels

__global__ void __launch_bounds__(kBlock)
kb_rap(const BInst *__restrict__ D, const BInst *__restrict__ Dc, const int *__restrict__ bmapc, LPtr<double> A,
       double *e, double *wx, double *wy, double *wz, float *e32, float *wx32, float *wy32, float *wz32)
{
    const int q = bmapc[blockIdx.x];
    const BInst Ic = Dc[q];
    const int I = (blockIdx.x - Ic.blk0) * kBlock + threadIdx.x;
    if (I >= Ic.n) return;
    const BInst If = D[q];
    const OpArgs<double> o = bop(A, If.off, If.nx, If.ny, If.nz);
    const int cx = Ic.nx, cy = Ic.ny;
    const int ck = I / (cx * cy);
    const int rem = I - ck * cx * cy;
    const int cj = rem / cx, ci = rem - cj * cx;
    const int i0 = ci * If.fx, j0 = cj * If.fy, k0 = ck * If.fz;
    const int i1 = min(i0 + If.fx, o.nx) - 1, j1 = min(j0 + If.fy, o.ny) - 1, k1 = min(k0 + If.fz, o.nz) - 1;
    double se = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = k0; k <= k1; k++)
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++) {
                int s = i + j * o.nx + k * o.nx * o.ny;
                se += o.e[s];
                if (i == i1) sx += o.wx[s];
                if (j == j1) sy += o.wy[s];
                if (k == k1) sz += o.wz[s];
            }
    const int g = Ic.off + I;
    e[g] = se; wx[g] = sx; wy[g] = sy; wz[g] = sz;
    e32[g] = (float)se; wx32[g] = (float)sx; wy32[g] = (float)sy; wz32[g] = (float)sz;
}

__global__ void __launch_bounds__(kBlock)
kb_to_f32(long n, const double *e, const double *wx, const double *wy, const double *wz, float *e32, float *wx32,
          float *wy32, float *wz32)
{
    long s = (long)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    e32[s] = (float)e[s]; wx32[s] = (float)wx[s]; wy32[s] = (float)wy[s]; wz32[s] = (float)wz[s];
}

// The fine level's FP32 copy and its inverse diagonal in one pass, per instance.
__global__ void __launch_bounds__(kBlock)
kb_to_f32_dinv(const BInst *__restrict__ D, const int *__restrict__ bmap, LPtr<double> A, float *e32, float *wx32,
               float *wy32, float *wz32, float *di32)
{
    const int q = bmap[blockIdx.x];
    const BInst I = D[q];
    const int s = (blockIdx.x - I.blk0) * kBlock + threadIdx.x;
    if (s >= I.n) return;
    const OpArgs<double> o = bop(A, I.off, I.nx, I.ny, I.nz);
    to_f32_dinv_cell(o, cell_at(o, s), e32 + I.off, wx32 + I.off, wy32 + I.off, wz32 + I.off, di32 + I.off);
}

// The same pass, which also takes the conduit correction's per-instance
// histograms of the FP64 diagonal's bins (kHistBins per instance) and keeps
// each cell's bin. Block b covers the fine level's blocks [b kConvSub, (b + 1)
// kConvSub): the bins of the instance of its first block are summed in LDS
// (runs of equal bins counted in a register first), those of a later
// instance (a block that crosses an instance's end) go to the global bins
// directly. Per-launch atomics on the crowded bins stay few: one flush per
// 4096 cells instead of one per wave.
constexpr int kConvSub = 16;
constexpr int kConvBatch = 4;

__global__ void __launch_bounds__(kBlock)
kb_to_f32_dinv_hist(const BInst *__restrict__ D, const int *__restrict__ bmap, int nblk, LPtr<double> A,
                    float *e32, float *wx32, float *wy32, float *wz32, float *di32, unsigned *hist,
                    unsigned short *bins)
{
    constexpr int NB = Deflation::kHistBins;
    __shared__ unsigned h[NB];
    for (int t = threadIdx.x; t < NB; t += kBlock) h[t] = 0;
    __syncthreads();
    const int sb0 = blockIdx.x * kConvSub;
    const int q0 = bmap[sb0];
    int rq = q0, rb = 0;
    unsigned run = 0;
    auto flush = [&]() {
        if (!run) return;
        if (rq == q0) atomicAdd(&h[rb], run);
        else atomicAdd(&hist[(long)rq * NB + rb], run);
    };
    // kConvBatch blocks' cells at a time, all their loads first (a lane past
    // its instance's end, or a block past the last, loads a cell of its own
    // and stores nothing)
    for (int t0 = 0; t0 < kConvSub; t0 += kConvBatch) {
        Cell c[kConvBatch];
        ConvLoads L[kConvBatch];
        int qq[kConvBatch];
        bool live[kConvBatch];
#pragma unroll
        for (int u = 0; u < kConvBatch; u++) {
            const int sb = min(sb0 + t0 + u, nblk - 1);
            const int q = bmap[sb];
            const BInst I = D[q];
            const int s = (sb - I.blk0) * kBlock + threadIdx.x;
            live[u] = sb0 + t0 + u < nblk && s < I.n;
            qq[u] = q;
            const OpArgs<double> o = bop(A, I.off, I.nx, I.ny, I.nz);
            c[u] = cell_at(o, live[u] ? s : I.n - 1);
            c[u].s += I.off;   // batch index: the loads and stores below are on the whole batch's arrays
            L[u] = conv_load(bop(A, 0, I.nx, I.ny, I.nz), c[u]);
        }
#pragma unroll
        for (int u = 0; u < kConvBatch; u++) {
            if (!live[u]) continue;
            const int b = conv_store(L[u], c[u], e32, wx32, wy32, wz32, di32, bins, NB);
            if (qq[u] != rq || b != rb) {
                flush();
                rq = qq[u];
                rb = b;
                run = 0;
            }
            run++;
        }
    }
    flush();
    __syncthreads();
    for (int t = threadIdx.x; t < NB; t += kBlock)
        if (h[t]) atomicAdd(&hist[(long)q0 * NB + t], h[t]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
kb_dinv(const BInst *__restrict__ D, const int *__restrict__ bmap, LPtr<T> A, T *di)
{
    const int q = bmap[blockIdx.x];
    const BInst I = D[q];
    const int s = (blockIdx.x - I.blk0) * kBlock + threadIdx.x;
    if (s >= I.n) return;
    const OpArgs<T> o = bop(A, I.off, I.nx, I.ny, I.nz);
    Cell c = cell_at(o, s);
    di[I.off + s] = (T)(1.0 / (double)diag_at<T, double>(o, c));
}

// Dense inverse of every instance's coarsest level (block q = instance q).
__global__ void __launch_bounds__(1024)
kb_dense_invert(const BInst *__restrict__ D, LPtr<double> A, double *inv, int *info)
{
    __shared__ double M[kCoarseMax * kCoarseMax];
    __shared__ double colk[kCoarseMax];
    const int q = blockIdx.x;
    const BInst I = D[q];
    const OpArgs<double> o = bop(A, I.off, I.nx, I.ny, I.nz);
    const int n = I.n;
    const int nx = o.nx, nxy = o.nx * o.ny;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) M[t] = 0.0;
    __syncthreads();
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
        Cell c = cell_at(o, s);
        M[s * n + s] = diag_at<double, double>(o, c);
        if (c.i < o.nx - 1) M[s * n + s + 1] = -o.wx[s];
        if (c.i > 0) M[s * n + s - 1] = -o.wx[s - 1];
        if (c.j < o.ny - 1) M[s * n + s + nx] = -o.wy[s];
        if (c.j > 0) M[s * n + s - nx] = -o.wy[s - nx];
        if (c.k < o.nz - 1) M[s * n + s + nxy] = -o.wz[s];
        if (c.k > 0) M[s * n + s - nxy] = -o.wz[s - nxy];
    }
    __syncthreads();
    for (int k = 0; k < n; k++) {
        double piv = M[k * n + k];
        if (!(piv > 0.0)) { if (threadIdx.x == 0) info[q] = k + 1; return; }
        for (int i = threadIdx.x; i < n; i += blockDim.x) colk[i] = M[i * n + k];
        __syncthreads();
        double rp = 1.0 / piv;
        for (int j = threadIdx.x; j < n; j += blockDim.x) M[k * n + j] = (j == k ? 1.0 : M[k * n + j]) * rp;
        __syncthreads();
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
            int i = t / n, j = t - i * n;
            if (i == k) continue;
            double a = (j == k) ? 0.0 : M[t];
            M[t] = a - colk[i] * M[k * n + j];
        }
        __syncthreads();
    }
    double *iv = inv + (long)q * kCoarseMax * kCoarseMax;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) iv[t] = M[t];
    if (threadIdx.x == 0) info[q] = 0;
}

// ---------------------------------------------------------------- CG kernels

__global__ void __launch_bounds__(kBlock)
kb_cg_init(const CInst *__restrict__ CI, const int *__restrict__ cbmap, LPtr<double> A, const double *__restrict__ b,
           const double *__restrict__ x, double *__restrict__ r, double *Pinit)
{
    const int q = cbmap[blockIdx.x];
    const CInst I = CI[q];
    const OpArgs<double> o = bop(A, I.off, I.nx, I.ny, I.nz);
    const int lb = blockIdx.x - I.cblk0;
    const int beg = lb * I.chunk, end = min(I.n, beg + I.chunk);
    const double *xl = x + I.off, *bl = b + I.off;
    double *rl = r + I.off;
    double v[2] = {0.0, 0.0};
    for (int base = beg; base < end; base += kBlock) {
        int s = base + threadIdx.x;
        if (s < end) {
            Cell c = cell_at(o, s);
            double d;
            double ax = apply_at<double, double>(o, c, [&](const Cell &t) { return xl[t.s]; }, d);
            double bs = bl[s], rs = bs - ax;
            rl[s] = rs;
            v[0] += bs * bs;
            v[1] += rs * rs;
        }
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) { Pinit[2 * blockIdx.x] = v[0]; Pinit[2 * blockIdx.x + 1] = v[1]; }
}

// Per instance, marching along z (as the single-grid CG SpMV): a block owns
// a 32 x 8 column tile of its instance and zlen planes, keeps s = z + beta
// s_old of the tile and its x/y ring in LDS for three planes (loads two
// planes ahead of the arithmetic), writes s, v = A s and one partial <s, v>,
// and applies the previous iteration's x += alpha s_old of its cells.
// zmap[block] names the instance; its blocks are [zblk0, zblk0 + zG).
template <typename ZT>
__global__ void __launch_bounds__(kBlock)
kb_cg_spmv_z(const CInst *__restrict__ CI, const int *__restrict__ zmap, LPtr<double> A, const ZT *__restrict__ z,
             double *__restrict__ s0, double *__restrict__ s1, double *__restrict__ v, double *__restrict__ x,
             CgScalars *SC, const double *Pinit, const double *Prr, const double *Pdot, double *Ptau,
             const double *Pext)
{
    const int q = zmap[blockIdx.x];
    CgScalars *sc = SC + q;
    if (sc->done) return;
    const CInst I = CI[q];
    const int k = sc->kspmv;
    double dots[2];
    sum_partials<2>(Pdot + 2 * I.cblk0, I.cG, dots);
    if (Pext) {   // the conduit correction's share of <z, r> and <z, v>
        dots[0] += Pext[2 * q];
        dots[1] += Pext[2 * q + 1];
    }
    double bb, rr;
    if (k == 0) {
        double ir[2];
        sum_partials<2>(Pinit + 2 * I.cblk0, I.cG, ir);
        bb = ir[0]; rr = ir[1];
    } else {
        double t[1];
        sum_partials<1>(Prr + I.cblk0, I.cG, t);
        rr = t[0]; bb = sc->bb;
    }
    const double sigma = fabs(dots[0]);
    int stop = 0, status = HMFEM_OK;
    double beta = 0.0, sigma_first = k == 0 ? (sigma == 0.0 ? 1.0 : sigma) : sc->sigma_first;
    if (k == 0) {
        if (sigma == 0.0 || (sc->stop_rule == HMFEM_STOP_RESIDUAL && (bb == 0.0 || rr <= sc->rtol2 * bb))) stop = 1;
    } else {
        beta = -sc->alpha[(k - 1) & 1] * dots[1] / sc->sigma[(k - 1) & 1];
        if (sc->stop_rule == HMFEM_STOP_RESIDUAL && rr <= sc->rtol2 * bb) stop = 1;
        if (sc->stop_rule == HMFEM_STOP_PRECONDITIONED && sigma <= sc->rtol2 * sigma_first) stop = 1;
        if (!(sigma == sigma) || !(beta == beta)) { stop = 1; status = HMFEM_ERR_BREAKDOWN; }
        else if (!stop && k >= sc->maxit) { stop = 1; status = HMFEM_NOT_CONVERGED; }
    }
    if (blockIdx.x == I.zblk0 && threadIdx.x == 0) {
        if (k == 0) { sc->bb = bb; sc->sigma_first = sigma_first; }
        sc->rr = rr;
        sc->sigma[k & 1] = sigma;
        sc->beta = beta;
        sc->iter = k;
        sc->kupd = k;
        if (stop) { sc->done = 1; sc->status = status; }
    }
    if (stop) return;
    const double alpha_prev = k ? sc->alpha[(k - 1) & 1] : 0.0;
    const OpArgs<double> o = bop(A, I.off, I.nx, I.ny, I.nz);
    const ZT *__restrict__ zl = z + I.off;
    const double *__restrict__ sin = ((k & 1) ? s1 : s0) + I.off;
    double *__restrict__ sout = ((k & 1) ? s0 : s1) + I.off;
    double *__restrict__ vl = v + I.off;
    double *__restrict__ xl = x + I.off;
    __shared__ double sh[3][kGsTy + 2][kGsTx + 2];
    const int nx = I.nx, ny = I.ny, nz = I.nz, zlen = I.zlen;
    const long nxy = (long)nx * ny;
    const int ntx = (nx + kGsTx - 1) / kGsTx;
    const int ntxy = ntx * ((ny + kGsTy - 1) / kGsTy);
    const int lb = blockIdx.x - I.zblk0;
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
    struct Raw {
        double zo, so, zr, sr, xo;
    };
    auto load_raw = [&](int kk) -> Raw {
        Raw r{0.0, 0.0, 0.0, 0.0, 0.0};
        if (kk >= 0 && kk < nz) {
            if (col) {
                const long c = i + (long)j * nx + (long)kk * nxy;
                r.zo = (double)zl[c];
                if (k) r.so = sin[c];
                if (k && kk >= kb && kk < ke) r.xo = xl[c];
            }
            if (ring_in) {
                const long c = ri + (long)rj * nx + (long)kk * nxy;
                r.zr = (double)zl[c];
                if (k) r.sr = sin[c];
            }
        }
        return r;
    };
    auto store_raw = [&](int kk, const Raw &r) {
        const int slot = (kk + 3) % 3;
        sh[slot][ty + 1][tx + 1] = r.zo + beta * r.so;   // beta = 0 and s_old unread at k = 0
        if (has_ring) sh[slot][ry][rx] = r.zr + beta * r.sr;
        if (col && k && kk >= kb && kk < ke) xl[i + (long)j * nx + (long)kk * nxy] = r.xo + alpha_prev * r.so;
    };
    struct Coef {
        double e, xp, xm, yp, ym, zp;
    };
    auto load_coef = [&](int kk) -> Coef {
        Coef c{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (col && kk < nz) {
            const long g = i + (long)j * nx + (long)kk * nxy;
            c.e = o.e[g];
            c.xp = i < nx - 1 ? o.wx[g] : 0.0;
            c.xm = i > 0 ? o.wx[g - 1] : 0.0;
            c.yp = j < ny - 1 ? o.wy[g] : 0.0;
            c.ym = j > 0 ? o.wy[g - nx] : 0.0;
            c.zp = kk < nz - 1 ? o.wz[g] : 0.0;
        }
        return c;
    };
    double acc[1] = {0.0};
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
        const int c = (kk + 3) % 3, lo = (kk + 2) % 3, hi = (kk + 1) % 3;
        if (col) {
            const long g = i + (long)j * nx + (long)kk * nxy;
            const double us = sh[c][ty + 1][tx + 1];
            double a = cf.e * us;
            a += cf.xp * (us - sh[c][ty + 1][tx + 2]);
            a += cf.xm * (us - sh[c][ty + 1][tx]);
            a += cf.yp * (us - sh[c][ty + 2][tx + 1]);
            a += cf.ym * (us - sh[c][ty][tx + 1]);
            a += cf.zp * (us - sh[hi][ty + 1][tx + 1]);
            a += wz_lo * (us - sh[lo][ty + 1][tx + 1]);
            wz_lo = cf.zp;
            sout[g] = us;
            vl[g] = a;
            acc[0] += us * a;
        }
        cf = cn;
        __syncthreads();
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) Ptau[blockIdx.x] = acc[0];
}

// Per instance: r -= alpha v and the partials of <r, r>; x += alpha s is
// left to the next kb_cg_spmv_z (sc->xpend names the pending iteration).
__global__ void __launch_bounds__(kBlock)
kb_cg_update(const CInst *__restrict__ CI, const int *__restrict__ cbmap, const double *__restrict__ v,
             double *__restrict__ r, CgScalars *SC, const double *Ptau, double *Prr)
{
    const int q = cbmap[blockIdx.x];
    CgScalars *sc = SC + q;
    if (sc->done) return;
    const CInst I = CI[q];
    const int k = sc->kupd;
    double t[1];
    sum_partials<1>(Ptau + I.zblk0, I.zG, t);   // one partial per block of kb_cg_spmv_z
    const double tau = fabs(t[0]);
    if (tau == 0.0 || !(tau == tau)) {
        if (blockIdx.x == I.cblk0 && threadIdx.x == 0) {
            sc->done = 1; sc->status = HMFEM_ERR_BREAKDOWN;
            sc->xpend = -1;   // the previous update was applied by this iteration's SpMV
        }
        return;
    }
    const double alpha = sc->sigma[k & 1] / tau;
    if (blockIdx.x == I.cblk0 && threadIdx.x == 0) {
        sc->alpha[k & 1] = alpha; sc->tau = tau; sc->kspmv = k + 1; sc->xpend = k;
    }
    const double *vl = v + I.off;
    double *rl = r + I.off;
    const int lb = blockIdx.x - I.cblk0;
    const int beg = lb * I.chunk, end = min(I.n, beg + I.chunk);
    double acc[1] = {0.0};
    for (int base = beg; base < end; base += kBlock) {
        int i = base + threadIdx.x;
        if (i < end) {
            double ri = rl[i] - alpha * vl[i];
            rl[i] = ri;
            acc[0] += ri * ri;
        }
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) Prr[blockIdx.x] = acc[0];
}

// After the CG loop, per instance: the x += alpha s of iteration sc->xpend if
// no SpMV applied it (the instance stopped at the SpMV that would have).
__global__ void __launch_bounds__(kBlock)
kb_cg_finish_x(const CInst *__restrict__ CI, const int *__restrict__ cbmap, const double *__restrict__ s0,
               const double *__restrict__ s1, double *__restrict__ x, const CgScalars *SC)
{
    const int q = cbmap[blockIdx.x];
    const CgScalars *sc = SC + q;
    const int k = sc->xpend;
    if (k < 0) return;
    const CInst I = CI[q];
    const double *s = ((k & 1) ? s0 : s1) + I.off;
    const double alpha = sc->alpha[k & 1];
    double *xl = x + I.off;
    const int lb = blockIdx.x - I.cblk0;
    const int beg = lb * I.chunk, end = min(I.n, beg + I.chunk);
    for (int i = beg + threadIdx.x; i < end; i += kBlock) xl[i] = xl[i] + alpha * s[i];
}

// ---------------------------------------------------------------- host side

namespace {

struct BLevel {
    std::vector<BInst> h;
    BInst *d = nullptr;
    int *bmap = nullptr;
    int *zmap = nullptr;      // kb_restrict_z block -> instance
    int nzblk = 0, nold = 0;  // its blocks; instances left to kb_resid_restrict
    int nblk = 0;
    long ncell = 0;
    int *tmap = nullptr;      // smoother tile -> instance
    int ntiles = 0;
    LPtr<double> a64{};
    LPtr<float> a32{};
    double *e64 = nullptr, *wx64 = nullptr, *wy64 = nullptr, *wz64 = nullptr, *di64 = nullptr;
    float *e32 = nullptr, *wx32 = nullptr, *wy32 = nullptr, *wz32 = nullptr, *di32 = nullptr;
    void *f = nullptr, *u = nullptr, *t = nullptr;
};

struct Batch {
    int ninst = 0, nlev = 0;
    BLevel lev[kMaxLevels];
    std::vector<CInst> ch;
    CInst *dch = nullptr;
    int *cbmap = nullptr;
    int ncblk = 0;
    int *zmap = nullptr;       // kb_cg_spmv_z block -> instance
    int nzblk = 0;
    double *inv = nullptr;
    int *info = nullptr;
    double *Pinit = nullptr, *Ptau = nullptr, *Prr = nullptr, *Pdot = nullptr;
    double *r = nullptr, *s0 = nullptr, *s1 = nullptr, *v = nullptr;
    CgScalars *sc = nullptr;
    int *never = nullptr;   // a device int that stays 0 (the correction's kernels never skip)
};

size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

struct Carve {
    char *p;
    size_t used, cap;
    template <typename T>
    T *take(size_t count)
    {
        size_t b = al(count * sizeof(T));
        if (used + b > cap) return nullptr;
        T *r = (T *)(p + used);
        used += b;
        return r;
    }
};

// Level plan and sizes of a batch (host only).
void plan_batch(int ninst, const int *dims, int coarse_max, Batch &B)
{
    std::vector<Dims> cur(ninst);
    for (int q = 0; q < ninst; q++) cur[q] = Dims{dims[3 * q], dims[3 * q + 1], dims[3 * q + 2]};
    int depth = 1;
    for (int q = 0; q < ninst; q++) {
        Dims d = cur[q];
        int lv = 1;
        while (d.n() > coarse_max && lv < kMaxLevels) {
            d = Dims{(d.nx + (d.nx > 1)) / (1 + (d.nx > 1)), (d.ny + (d.ny > 1)) / (1 + (d.ny > 1)),
                     (d.nz + (d.nz > 1)) / (1 + (d.nz > 1))};
            lv++;
        }
        depth = std::max(depth, lv);
    }
    depth = std::max(depth, 2);   // the fine level always smooths (its last sweep carries the CG dots)
    B.ninst = ninst;
    B.nlev = depth;
    for (int l = 0; l < depth; l++) {
        BLevel &L = B.lev[l];
        L.h.assign(ninst, BInst{});
        long off = 0;
        int blk = 0;
        for (int q = 0; q < ninst; q++) {
            Dims d = cur[q];
            BInst &I = L.h[q];
            I.nx = d.nx; I.ny = d.ny; I.nz = d.nz;
            I.n = (int)d.n();
            I.off = (int)off;
            I.blk0 = blk;
            I.nblk = (I.n + kBlock - 1) / kBlock;
            bool coarsen = d.n() > coarse_max && l + 1 < depth;
            I.fx = coarsen && d.nx > 1 ? 2 : 1;
            I.fy = coarsen && d.ny > 1 ? 2 : 1;
            I.fz = coarsen && d.nz > 1 ? 2 : 1;
            I.tile0 = L.ntiles;
            I.ntile = gs_tiles(d);
            L.ntiles += I.ntile;
            off += I.n;
            blk += I.nblk;
            cur[q] = Dims{(d.nx + I.fx - 1) / I.fx, (d.ny + I.fy - 1) / I.fy, (d.nz + I.fz - 1) / I.fz};
        }
        L.ncell = off;
        L.nblk = blk;
    }
    // residual-restriction of the 2x2x2 instances of each level: z-marching
    // blocks of 32 x 8 columns and an even number of planes (at most
    // kBatchRestrictZlen; the coarse sums are the same for any even count)
    for (int l = 0; l < depth; l++) {
        BLevel &L = B.lev[l];
        int zb = 0, nold = 0;
        for (int q = 0; q < ninst; q++) {
            BInst &I = L.h[q];
            I.zblk0 = zb; I.zG = 0; I.zlen = 0;
            if (l + 1 >= depth) continue;
            if (I.fx == 2 && I.fy == 2 && I.fz == 2) {
                I.zlen = std::min(kBatchRestrictZlen, (I.nz + 1) & ~1);
                I.zG = ((I.nx + kGsTx - 1) / kGsTx) * ((I.ny + kGsTy - 1) / kGsTy) * ((I.nz + I.zlen - 1) / I.zlen);
                zb += I.zG;
            } else {
                nold++;
            }
        }
        L.nzblk = zb;
        L.nold = nold;
    }
    // chunked layout of the fine level
    B.ch.assign(ninst, CInst{});
    int cb = 0;
    for (int q = 0; q < ninst; q++) {
        const BInst &I = B.lev[0].h[q];
        CInst &c = B.ch[q];
        c.nx = I.nx; c.ny = I.ny; c.nz = I.nz; c.n = I.n; c.off = I.off;
        int G = (I.n + kBlock * kCellsPerThread - 1) / (kBlock * kCellsPerThread);
        G = std::max(1, std::min(G, kMaxReduceBlocks));
        int chunk = (I.n + G - 1) / G;
        chunk = (chunk + kBlock - 1) / kBlock * kBlock;
        G = (I.n + chunk - 1) / chunk;
        c.cblk0 = cb; c.cG = G; c.chunk = chunk;
        c.tpb = (I.ntile + G - 1) / G;
        cb += G;
    }
    B.ncblk = cb;
    // z-marching SpMV: 32 x 8 column tiles, kBatchZlen planes per block (at
    // 96^3, 216 blocks of 16 planes)
    int zb = 0;
    for (int q = 0; q < ninst; q++) {
        CInst &c = B.ch[q];
        const int ntxy = ((c.nx + kGsTx - 1) / kGsTx) * ((c.ny + kGsTy - 1) / kGsTy);
        c.zlen = std::min(c.nz, kBatchZlen);   // longer where that keeps the partials at most kMaxReduceBlocks
        while (c.zlen < c.nz && (long)ntxy * ((c.nz + c.zlen - 1) / c.zlen) > kMaxReduceBlocks) c.zlen++;
        c.zblk0 = zb;
        c.zG = ntxy * ((c.nz + c.zlen - 1) / c.zlen);
        zb += c.zG;
    }
    B.nzblk = zb;
}

size_t batch_bytes(const Batch &B, int prec)
{
    size_t vs = prec == HMFEM_PREC_FP64 ? 8 : 4, b = 0;
    for (int l = 0; l < B.nlev; l++) {
        size_t n = B.lev[l].ncell;
        if (l > 0) b += 4 * al(n * 8);
        if (vs == 8) b += al(n * 8);
        b += 5 * al(n * 4) + 3 * al(n * vs);
        b += al(sizeof(BInst) * B.ninst) + al(sizeof(int) * B.lev[l].nblk) + al(sizeof(int) * B.lev[l].ntiles) +
             al(sizeof(int) * B.lev[l].nzblk);
    }
    size_t n0 = B.lev[0].ncell;
    b += 4 * al(n0 * 8);                                      // r, s0, s1, v
    b += al(sizeof(CInst) * B.ninst) + al(sizeof(int) * B.ncblk) + al(sizeof(int) * B.nzblk) + al((size_t)B.nzblk * 8);
    b += al((size_t)B.ninst * kCoarseMax * kCoarseMax * 8) + al(sizeof(int) * B.ninst);
    b += 6 * al((size_t)B.ncblk * 8) + al(sizeof(CgScalars) * B.ninst);
    return b + 65536;
}

template <typename T>
LPtr<T> lp(const BLevel &L);
template <>
LPtr<float> lp<float>(const BLevel &L) { return LPtr<float>{L.e32, L.wx32, L.wy32, L.wz32, L.di32}; }
template <>
LPtr<double> lp<double>(const BLevel &L) { return LPtr<double>{L.e64, L.wx64, L.wy64, L.wz64, L.di64}; }

struct BCycle {
    int nu;
    double omega, ocf;
    int smoother, gs;
};

// f_c = P^T (f - A u) from level L to C: the z-marching kernel for the 2x2x2
// instances, the per-coarse-cell kernel for the others (if any).
template <typename T, typename FT>
void launch_restrict(hipStream_t st, const BLevel &L, const BLevel &C, LPtr<T> A, const FT *f, const T *u,
                     const CgScalars *sc)
{
    if (L.nzblk)
        hipLaunchKernelGGL((kb_restrict_z<T, FT>), dim3(L.nzblk), dim3(kBlock), 0, st, L.d, C.d, L.zmap, A, f, u,
                           (T *)C.f, sc);
    if (L.nold)
        hipLaunchKernelGGL((kb_resid_restrict<T, FT>), dim3(C.nblk), dim3(kBlock), 0, st, L.d, C.d, C.bmap, A, f, u,
                           (T *)C.f, sc);
}

template <typename T, typename FT>
T *vb_level(hipStream_t st, Batch &B, int l, const FT *f, const double *vdot, const BCycle &cp, double *Pdot)
{
    BLevel &L = B.lev[l];
    T *u = (T *)L.u, *t = (T *)L.t;
    const T om = (T)cp.omega;
    if (l == B.nlev - 1) {
        hipLaunchKernelGGL((kb_dense_apply<T, FT>), dim3(B.ninst), dim3(kBlock), 0, st, L.d, B.inv, f, u, B.sc);
        return u;
    }
    LPtr<T> A = lp<T>(L);
    if (cp.smoother == HMFEM_SMOOTH_BLOCK_GS) {
        BLevel &C = B.lev[l + 1];
        hipLaunchKernelGGL((kb_gs_tile<T, FT, 0>), dim3(L.ntiles), dim3(kGsThreads), 0, st, L.d, C.d, L.tmap, A,
                           cp.gs, f, (const T *)nullptr, u, (const T *)nullptr, (T)0, B.sc);
        launch_restrict<T, FT>(st, L, C, A, f, u, B.sc);
        T *uc = vb_level<T, T>(st, B, l + 1, (const T *)C.f, nullptr, cp, nullptr);
        if (l == 0) {
            hipLaunchKernelGGL(kb_gs_tile_dots<T>, dim3(B.ncblk), dim3(kGsThreads), 0, st, B.dch, B.cbmap, L.d, C.d,
                               A, cp.gs, (const double *)f, (const T *)u, t, (const T *)uc, (T)cp.ocf, vdot, Pdot,
                               B.sc);
            return t;
        }
        hipLaunchKernelGGL((kb_gs_tile<T, FT, 1>), dim3(L.ntiles), dim3(kGsThreads), 0, st, L.d, C.d, L.tmap, A,
                           cp.gs, f, (const T *)u, t, (const T *)uc, (T)cp.ocf, B.sc);
        L.t = u;
        L.u = t;
        return t;
    }
    hipLaunchKernelGGL((kb_smooth_zero<T, FT>), dim3(L.nblk), dim3(kBlock), 0, st, L.d, L.bmap, A, f, u, om,
                       cp.nu >= 2 ? 2 : 1, B.sc);
    for (int q = 2; q < cp.nu; q++) {
        hipLaunchKernelGGL((kb_jacobi<T, FT>), dim3(L.nblk), dim3(kBlock), 0, st, L.d, L.bmap, A, f, u, t, om, B.sc);
        std::swap(u, t);
    }
    BLevel &C = B.lev[l + 1];
    launch_restrict<T, FT>(st, L, C, A, f, u, B.sc);
    T *uc = vb_level<T, T>(st, B, l + 1, (const T *)C.f, nullptr, cp, nullptr);
    hipLaunchKernelGGL((kb_prolong_jacobi<T, FT>), dim3(L.nblk), dim3(kBlock), 0, st, L.d, C.d, L.bmap, A,
                       (const T *)uc, (T)cp.ocf, f, u, t, om, B.sc);
    std::swap(u, t);
    const int post = l == 0 ? cp.nu - 2 : cp.nu - 1;
    for (int q = 0; q < post; q++) {
        hipLaunchKernelGGL((kb_jacobi<T, FT>), dim3(L.nblk), dim3(kBlock), 0, st, L.d, L.bmap, A, f, u, t, om, B.sc);
        std::swap(u, t);
    }
    if (l == 0) {
        hipLaunchKernelGGL(kb_final_sweep_dots<T>, dim3(B.ncblk), dim3(kBlock), 0, st, B.dch, B.cbmap, A,
                           (const double *)f, (const T *)u, t, om, vdot, Pdot, B.sc);
        std::swap(u, t);
    }
    if (l > 0 && u != (T *)L.u) { L.t = L.u; L.u = u; }
    return u;
}

} // namespace

int batch_solve(hipStream_t st, int ninst, const int *dims, const Stencil<double> &A0, const double *b, double *x,
                const hmfem_options &opt_in, char *(*alloc)(void *, size_t), void *alloc_ctx, CgScalars *h_sc,
                GraphCache *graph, std::vector<char> &stage, hipEvent_t ev[4], int *iters_out,
                double *relres_out, double *done_ms_out, int *status_out, double *setup_ms, double *solve_ms,
                Deflation *defl, char *(*dalloc)(void *, size_t, int), void *dctx, const int *inst_gs,
                const double *inst_ocf)
{
    hmfem_options opt = opt_in;
    if (opt.nu < 2) opt.nu = 2;
    int coarse_max = opt.coarse_max > kCoarseMax || opt.coarse_max < 1 ? kCoarseMax : opt.coarse_max;
    const int prec = opt.precision == HMFEM_PREC_FP64 ? HMFEM_PREC_FP64 : HMFEM_PREC_FP32;
    Batch B;
    plan_batch(ninst, dims, coarse_max, B);
    // per-instance cycle options (the batch autotuner's), on every level
    for (int l = 0; l < B.nlev; l++)
        for (int q = 0; q < ninst; q++) {
            B.lev[l].h[q].gs = inst_gs && inst_gs[q] > 0 ? inst_gs[q] : 0;
            B.lev[l].h[q].ocf = inst_ocf && inst_ocf[q] > 0 ? (float)inst_ocf[q] : 0.f;
        }
    size_t need = batch_bytes(B, prec);
    char *arena = alloc(alloc_ctx, need);
    if (!arena) return HMFEM_ERR_NOMEM;
    Carve cv{arena, 0, need};
    const size_t vs = prec == HMFEM_PREC_FP64 ? 8 : 4;
    for (int l = 0; l < B.nlev; l++) {
        BLevel &L = B.lev[l];
        size_t n = L.ncell;
        if (l == 0) { L.e64 = A0.e; L.wx64 = A0.wx; L.wy64 = A0.wy; L.wz64 = A0.wz; }
        else {
            L.e64 = cv.take<double>(n); L.wx64 = cv.take<double>(n);
            L.wy64 = cv.take<double>(n); L.wz64 = cv.take<double>(n);
        }
        if (prec == HMFEM_PREC_FP64) L.di64 = cv.take<double>(n);
        L.e32 = cv.take<float>(n); L.wx32 = cv.take<float>(n); L.wy32 = cv.take<float>(n);
        L.wz32 = cv.take<float>(n); L.di32 = cv.take<float>(n);
        L.f = cv.take<char>(n * vs); L.u = cv.take<char>(n * vs); L.t = cv.take<char>(n * vs);
        L.d = cv.take<BInst>(B.ninst);
        L.bmap = cv.take<int>(L.nblk);
        L.tmap = cv.take<int>(L.ntiles);
        L.zmap = L.nzblk ? cv.take<int>(L.nzblk) : nullptr;
        if (!L.bmap || !L.tmap || !L.t || !L.di32 || (L.nzblk && !L.zmap)) return HMFEM_ERR_NOMEM;
    }
    size_t n0 = B.lev[0].ncell;
    B.r = cv.take<double>(n0); B.s0 = cv.take<double>(n0); B.s1 = cv.take<double>(n0); B.v = cv.take<double>(n0);
    B.dch = cv.take<CInst>(B.ninst);
    B.cbmap = cv.take<int>(B.ncblk);
    B.zmap = cv.take<int>(B.nzblk);
    B.inv = cv.take<double>((size_t)B.ninst * kCoarseMax * kCoarseMax);
    B.info = cv.take<int>(B.ninst);
    B.Pinit = cv.take<double>(2 * (size_t)B.ncblk);
    B.Ptau = cv.take<double>(B.nzblk);
    B.Prr = cv.take<double>(B.ncblk);
    B.Pdot = cv.take<double>(2 * (size_t)B.ncblk);
    B.sc = cv.take<CgScalars>(B.ninst);
    B.never = cv.take<int>(1);
    if (!B.sc || !B.Pdot || !B.inv || !B.never || !B.zmap || !B.Ptau) return HMFEM_ERR_NOMEM;

    // descriptors
    {
        size_t total = 0;
        for (int l = 0; l < B.nlev; l++)
            total += sizeof(BInst) * B.ninst + sizeof(int) * (B.lev[l].nblk + B.lev[l].ntiles + B.lev[l].nzblk);
        total += sizeof(CInst) * B.ninst + sizeof(int) * (B.ncblk + B.nzblk) + sizeof(CgScalars) * B.ninst;
        stage.resize(total);
        char *p = stage.data();
        for (int l = 0; l < B.nlev; l++) {
            BLevel &L = B.lev[l];
            std::vector<int> bm(L.nblk);
            for (int q = 0; q < B.ninst; q++)
                for (int k = 0; k < L.h[q].nblk; k++) bm[L.h[q].blk0 + k] = q;
            memcpy(p, L.h.data(), sizeof(BInst) * B.ninst);
            HMFEM_CHECK(hipMemcpyAsync(L.d, p, sizeof(BInst) * B.ninst, hipMemcpyHostToDevice, st));
            p += sizeof(BInst) * B.ninst;
            memcpy(p, bm.data(), sizeof(int) * L.nblk);
            HMFEM_CHECK(hipMemcpyAsync(L.bmap, p, sizeof(int) * L.nblk, hipMemcpyHostToDevice, st));
            p += sizeof(int) * L.nblk;
            std::vector<int> tm(L.ntiles);
            for (int q = 0; q < B.ninst; q++)
                for (int k = 0; k < L.h[q].ntile; k++) tm[L.h[q].tile0 + k] = q;
            memcpy(p, tm.data(), sizeof(int) * L.ntiles);
            HMFEM_CHECK(hipMemcpyAsync(L.tmap, p, sizeof(int) * L.ntiles, hipMemcpyHostToDevice, st));
            p += sizeof(int) * L.ntiles;
            if (L.nzblk) {
                std::vector<int> zm(L.nzblk);
                for (int q = 0; q < B.ninst; q++)
                    for (int k = 0; k < L.h[q].zG; k++) zm[L.h[q].zblk0 + k] = q;
                memcpy(p, zm.data(), sizeof(int) * L.nzblk);
                HMFEM_CHECK(hipMemcpyAsync(L.zmap, p, sizeof(int) * L.nzblk, hipMemcpyHostToDevice, st));
                p += sizeof(int) * L.nzblk;
            }
        }
        std::vector<int> cbm(B.ncblk);
        for (int q = 0; q < B.ninst; q++)
            for (int k = 0; k < B.ch[q].cG; k++) cbm[B.ch[q].cblk0 + k] = q;
        memcpy(p, B.ch.data(), sizeof(CInst) * B.ninst);
        HMFEM_CHECK(hipMemcpyAsync(B.dch, p, sizeof(CInst) * B.ninst, hipMemcpyHostToDevice, st));
        p += sizeof(CInst) * B.ninst;
        memcpy(p, cbm.data(), sizeof(int) * B.ncblk);
        HMFEM_CHECK(hipMemcpyAsync(B.cbmap, p, sizeof(int) * B.ncblk, hipMemcpyHostToDevice, st));
        p += sizeof(int) * B.ncblk;
        std::vector<int> zbm(B.nzblk);
        for (int q = 0; q < B.ninst; q++)
            for (int k = 0; k < B.ch[q].zG; k++) zbm[B.ch[q].zblk0 + k] = q;
        memcpy(p, zbm.data(), sizeof(int) * B.nzblk);
        HMFEM_CHECK(hipMemcpyAsync(B.zmap, p, sizeof(int) * B.nzblk, hipMemcpyHostToDevice, st));
        p += sizeof(int) * B.nzblk;
        std::vector<CgScalars> init(B.ninst);
        for (auto &c : init) {
            memset(&c, 0, sizeof(c));
            c.rtol2 = opt.rtol * opt.rtol;
            c.maxit = opt.max_iterations;
            c.stop_rule = opt.stop_rule;
            c.xpend = -1;
        }
        memcpy(p, init.data(), sizeof(CgScalars) * B.ninst);
        HMFEM_CHECK(hipMemcpyAsync(B.sc, p, sizeof(CgScalars) * B.ninst, hipMemcpyHostToDevice, st));
    }

    HMFEM_CHECK(hipEventRecord(ev[0], st));
    // conduit correction, first step: the per-instance histograms (in the
    // FP32 hierarchy, taken by the fine level's conversion pass) and their
    // copy to the host, which overlaps the rest of the hierarchy's setup
    const Deflation *D = nullptr;
    const bool fuse_hist = prec != HMFEM_PREC_FP64;
    if (defl) {
        defl->nc = 0;
        defl->pending = false;
        defl->gen++;
        if (opt.deflate && dalloc) {
            int rc = deflation_begin_batch(st, ninst, dims, A0, opt.deflate_ratio, *defl, dalloc, dctx, fuse_hist);
            if (rc) return rc;
        }
    }
    const bool hist_here = fuse_hist && defl && defl->pending;
    // hierarchy
    if (prec == HMFEM_PREC_FP64)
        hipLaunchKernelGGL(kb_to_f32, dim3((unsigned)((n0 + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, (long)n0,
                           B.lev[0].e64, B.lev[0].wx64, B.lev[0].wy64, B.lev[0].wz64, B.lev[0].e32, B.lev[0].wx32,
                           B.lev[0].wy32, B.lev[0].wz32);
    else if (hist_here)   // with the fine level's inverse diagonal and the conduit correction's histograms
        hipLaunchKernelGGL(kb_to_f32_dinv_hist, dim3((B.lev[0].nblk + kConvSub - 1) / kConvSub), dim3(kBlock), 0, st,
                           B.lev[0].d, B.lev[0].bmap, B.lev[0].nblk, lp<double>(B.lev[0]), B.lev[0].e32,
                           B.lev[0].wx32, B.lev[0].wy32, B.lev[0].wz32, B.lev[0].di32, defl->h.bhist,
                           defl->h.bbins);
    else   // with the fine level's inverse diagonal
        hipLaunchKernelGGL(kb_to_f32_dinv, dim3(B.lev[0].nblk), dim3(kBlock), 0, st, B.lev[0].d, B.lev[0].bmap,
                           lp<double>(B.lev[0]), B.lev[0].e32, B.lev[0].wx32, B.lev[0].wy32, B.lev[0].wz32,
                           B.lev[0].di32);
    if (hist_here) {
        int rc = deflation_hist_queued_batch(st, *defl);
        if (rc) return rc;
    }
    for (int l = 0; l + 1 < B.nlev; l++) {
        BLevel &L = B.lev[l], &C = B.lev[l + 1];
        hipLaunchKernelGGL(kb_rap, dim3(C.nblk), dim3(kBlock), 0, st, L.d, C.d, C.bmap, lp<double>(L), C.e64, C.wx64,
                           C.wy64, C.wz64, C.e32, C.wx32, C.wy32, C.wz32);
    }
    for (int l = 0; l < B.nlev; l++) {
        BLevel &L = B.lev[l];
        if (prec == HMFEM_PREC_FP64)
            hipLaunchKernelGGL(kb_dinv<double>, dim3(L.nblk), dim3(kBlock), 0, st, L.d, L.bmap, lp<double>(L),
                               L.di64);
        else if (l > 0)
            hipLaunchKernelGGL(kb_dinv<float>, dim3(L.nblk), dim3(kBlock), 0, st, L.d, L.bmap, lp<float>(L), L.di32);
    }
    hipLaunchKernelGGL(kb_dense_invert, dim3(B.ninst), dim3(1024), 0, st, B.lev[B.nlev - 1].d,
                       lp<double>(B.lev[B.nlev - 1]), B.inv, B.info);
    HMFEM_CHECK(hipGetLastError());
    // conduit correction, second step: per-instance thresholds and clusters
    if (defl && defl->pending) {
        int rc = deflation_finish_batch(st, A0, *defl, dalloc, dctx);
        if (rc) return rc;
    }
    if (defl && defl->nc > 0) D = defl;
    HMFEM_CHECK(hipMemsetAsync(B.never, 0, sizeof(int), st));
    HMFEM_CHECK(hipEventRecord(ev[1], st));

    // CG
    // gs_sweeps 0 / ocf 0: four sweeps at a scaling of 2.0 for a batch of small grids (96 uneven
    // instances, 32^3 to 96^3: 1803-1811 solves/s; three at 2.2 1655, three at 2.0 1733, four at 1.9
    // 1793, four at 2.1 1730, five at 2.0 1739)
    BCycle cp{opt.nu, opt.omega, ocf_of(opt, kOcfBatch), opt.smoother,
              opt.gs_sweeps < 1 ? kSweepsBatch : opt.gs_sweeps};
    LPtr<double> A64 = lp<double>(B.lev[0]);
    hipLaunchKernelGGL(kb_cg_init, dim3(B.ncblk), dim3(kBlock), 0, st, B.dch, B.cbmap, A64, b, x, B.r, B.Pinit);
    auto vcycle = [&](const double *vd) -> void * {
        if (D) deflation_sums_batch(st, *D, B.r, vd, B.never);
        if (prec == HMFEM_PREC_FP64) {
            double *zz = vb_level<double, double>(st, B, 0, B.r, vd, cp, B.Pdot);
            if (D) deflation_scatter(st, *D, zz, B.never);
            return zz;
        }
        float *zz = vb_level<float, double>(st, B, 0, B.r, vd, cp, B.Pdot);
        if (D) deflation_scatter(st, *D, zz, B.never);
        return zz;
    };
    const double *Pext = D ? D->pext : nullptr;
    void *z = vcycle(nullptr);
    auto iteration = [&]() -> void * {
        if (prec == HMFEM_PREC_FP64) {
            hipLaunchKernelGGL(kb_cg_spmv_z<double>, dim3(B.nzblk), dim3(kBlock), 0, st, B.dch, B.zmap, A64,
                               (const double *)z, B.s0, B.s1, B.v, x, B.sc, B.Pinit, B.Prr, B.Pdot, B.Ptau, Pext);
        } else {
            hipLaunchKernelGGL(kb_cg_spmv_z<float>, dim3(B.nzblk), dim3(kBlock), 0, st, B.dch, B.zmap, A64,
                               (const float *)z, B.s0, B.s1, B.v, x, B.sc, B.Pinit, B.Prr, B.Pdot, B.Ptau, Pext);
        }
        hipLaunchKernelGGL(kb_cg_update, dim3(B.ncblk), dim3(kBlock), 0, st, B.dch, B.cbmap, B.v, B.r, B.sc, B.Ptau,
                           B.Prr);
        return vcycle(B.v);
    };
    hipGraphExec_t exec = nullptr;
    if (graph && opt.use_graph) {
        if (graph->exec) { (void)hipGraphExecDestroy(graph->exec); graph->exec = nullptr; graph->key.clear(); }
        hipGraph_t g = nullptr;
        HMFEM_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        void *zc = iteration();
        HMFEM_CHECK(hipStreamEndCapture(st, &g));
        HMFEM_CHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        if (zc != z) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
    }
    const int every = opt.check_every < 1 ? 1 : opt.check_every;
    std::vector<double> done_at(ninst, -1.0);
    hipEvent_t tstart = ev[1];
    int k = 0, ndone = 0;
    while (true) {
        for (int q = 0; q < every; q++, k++) {
            if (exec) HMFEM_CHECK(hipGraphLaunch(exec, st));
            else z = iteration();
        }
        HMFEM_CHECK(hipMemcpyAsync(h_sc, B.sc, sizeof(CgScalars) * ninst, hipMemcpyDeviceToHost, st));
        HMFEM_CHECK(hipEventRecord(ev[2], st));
        HMFEM_CHECK(hipEventSynchronize(ev[2]));
        float now = 0.f;
        (void)hipEventElapsedTime(&now, tstart, ev[2]);
        ndone = 0;
        for (int q = 0; q < ninst; q++) {
            if (h_sc[q].done) {
                ndone++;
                if (done_at[q] < 0) done_at[q] = now;
            }
        }
        if (ndone == ninst || k > opt.max_iterations + every) break;
    }
    // the x updates that no SpMV applied (part of the solve phase)
    hipLaunchKernelGGL(kb_cg_finish_x, dim3(B.ncblk), dim3(kBlock), 0, st, B.dch, B.cbmap, B.s0, B.s1, x, B.sc);
    HMFEM_CHECK(hipGetLastError());
    HMFEM_CHECK(hipEventRecord(ev[2], st));
    HMFEM_CHECK(hipEventSynchronize(ev[2]));
    if (exec) (void)hipGraphExecDestroy(exec);
    float t0 = 0.f, t1 = 0.f;
    (void)hipEventElapsedTime(&t0, ev[0], ev[1]);
    (void)hipEventElapsedTime(&t1, ev[1], ev[2]);
    if (setup_ms) *setup_ms = t0;
    if (solve_ms) *solve_ms = t1;
    std::vector<int> info(ninst, 0);
    HMFEM_CHECK(hipMemcpy(info.data(), B.info, sizeof(int) * ninst, hipMemcpyDeviceToHost));
    for (int q = 0; q < ninst; q++) {
        if (iters_out) iters_out[q] = h_sc[q].iter;
        if (relres_out) relres_out[q] = h_sc[q].bb > 0 ? sqrt(h_sc[q].rr / h_sc[q].bb) : 0.0;
        if (done_ms_out) done_ms_out[q] = done_at[q];
        if (status_out) status_out[q] = info[q] ? HMFEM_ERR_BREAKDOWN : (h_sc[q].done ? h_sc[q].status : HMFEM_NOT_CONVERGED);
    }
    return HMFEM_OK;
}

} // namespace hmfem
