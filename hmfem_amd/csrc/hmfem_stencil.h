// Device helpers shared by the single-instance and batched solvers: the
// flux-form 7-point operator, deterministic block reductions, chunking.
#pragma once

#include "hmfem_internal.h"

namespace hmfem {

// ---------------------------------------------------------------- operator

template <typename T>
struct OpArgs {
    const T *e, *wx, *wy, *wz, *di;
    int nx, ny, nz;
};

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

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }

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

// The fine level's FP32 copy (e, wx, wy, wz) from the FP64 stencil and its
// inverse diagonal in one pass: the diagonal is summed from the FP32-rounded
// couplings in diag_at's order, so di is what diag_at<float, double> on the
// copy gives.
__device__ __forceinline__ void to_f32_dinv_cell(const OpArgs<double> &o, const Cell &c, float *e32, float *wx32,
                                                 float *wy32, float *wz32, float *di32)
{
    const int nx = o.nx, nxy = o.nx * o.ny, s = c.s;
    const float e = (float)o.e[s], x = (float)o.wx[s], y = (float)o.wy[s], z = (float)o.wz[s];
    e32[s] = e; wx32[s] = x; wy32[s] = y; wz32[s] = z;
    double d = (double)e + (double)x + (double)y + (double)z;
    if (c.i > 0) d += (double)(float)o.wx[s - 1];
    if (c.j > 0) d += (double)(float)o.wy[s - nx];
    if (c.k > 0) d += (double)(float)o.wz[s - nxy];
    di32[s] = (float)(1.0 / d);
}

// Bin of a diagonal value for the conduit correction's histograms:
// floor(log2 d) + 1100, clamped to [0, nbins) (d <= 0 or NaN: bin 0).
__device__ __forceinline__ int diag_bin(double d, int nbins)
{
    int b = d > 0 ? ilogb(d) + 1100 : 0;
    return b < 0 ? 0 : (b >= nbins ? nbins - 1 : b);
}

// The conversion pass with the conduit correction's histogram, split into its
// loads and the rest so that a thread can have several cells' loads in
// flight: every load is unconditional (a neighbour off the grid reads the
// cell itself and is not added), and conv_store adds in diag_at's order, so
// the FP64 diagonal's bin is that of diag_at<double, double> and di32 that of
// to_f32_dinv_cell.
struct ConvLoads {
    double e, x, y, z, xm, ym, zm;
};

__device__ __forceinline__ ConvLoads conv_load(const OpArgs<double> &o, const Cell &c)
{
    const int nx = o.nx, nxy = o.nx * o.ny, s = c.s;
    ConvLoads L;
    L.e = o.e[s]; L.x = o.wx[s]; L.y = o.wy[s]; L.z = o.wz[s];
    L.xm = o.wx[c.i > 0 ? s - 1 : s];
    L.ym = o.wy[c.j > 0 ? s - nx : s];
    L.zm = o.wz[c.k > 0 ? s - nxy : s];
    return L;
}

// Stores cell c's FP32 copy, inverse diagonal and bin; returns the bin.
__device__ __forceinline__ int conv_store(const ConvLoads &L, const Cell &c, float *e32, float *wx32, float *wy32,
                                          float *wz32, float *di32, unsigned short *bins, int nbins)
{
    const int s = c.s;
    double d = L.e + L.x + L.y + L.z;
    d = c.i > 0 ? d + L.xm : d;
    d = c.j > 0 ? d + L.ym : d;
    d = c.k > 0 ? d + L.zm : d;
    const int b = diag_bin(d, nbins);
    bins[s] = (unsigned short)b;
    const float e = (float)L.e, x = (float)L.x, y = (float)L.y, z = (float)L.z;
    e32[s] = e; wx32[s] = x; wy32[s] = y; wz32[s] = z;
    double d32 = (double)e + (double)x + (double)y + (double)z;
    d32 = c.i > 0 ? d32 + (double)(float)L.xm : d32;
    d32 = c.j > 0 ? d32 + (double)(float)L.ym : d32;
    d32 = c.k > 0 ? d32 + (double)(float)L.zm : d32;
    di32[s] = (float)(1.0 / d32);
    return b;
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

constexpr int kUnroll = 4;   // cells per thread in flight in chunked kernels

// ---------------------------------------------------------------- reductions

// Sum of NV values over the block in a fixed order; every thread gets the sums.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double sh[kBlock / 64][NV];
    for (int q = 0; q < NV; q++)
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o);
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0)
        for (int q = 0; q < NV; q++) sh[w][q] = v[q];
    __syncthreads();
    for (int q = 0; q < NV; q++) {
        double t = sh[0][q];
        for (int x = 1; x < kBlock / 64; x++) t += sh[x][q];
        v[q] = t;
    }
    __syncthreads();
}

// Sum of NV values over a block of NT threads in a fixed order; every thread
// gets the sums.
template <int NV, int NT>
__device__ __forceinline__ void block_sum_n(double (&v)[NV])
{
    __shared__ double sh[NT / 64][NV];
    for (int q = 0; q < NV; q++)
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o);
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0)
        for (int q = 0; q < NV; q++) sh[w][q] = v[q];
    __syncthreads();
    for (int q = 0; q < NV; q++) {
        double t = sh[0][q];
        for (int x = 1; x < NT / 64; x++) t += sh[x][q];
        v[q] = t;
    }
    __syncthreads();
}

// Sum of the per-block partials of a previous kernel (NV interleaved values
// per block, G blocks), in a fixed order; every thread gets the sums.
template <int NV>
__device__ __forceinline__ void sum_partials(const double *P, int G, double (&out)[NV])
{
    for (int q = 0; q < NV; q++) out[q] = 0.0;
    for (int b = threadIdx.x; b < G; b += kBlock)
        for (int q = 0; q < NV; q++) out[q] += P[(long)NV * b + q];
    block_sum<NV>(out);
}

// sum_partials for a block of NT threads
template <int NV, int NT>
__device__ __forceinline__ void sum_partials_n(const double *P, int G, double (&out)[NV])
{
    for (int q = 0; q < NV; q++) out[q] = 0.0;
    for (int b = threadIdx.x; b < G; b += NT)
        for (int q = 0; q < NV; q++) out[q] += P[(long)NV * b + q];
    block_sum_n<NV, NT>(out);
}

// Contiguous range of cells of this (logical) block in a chunked launch.
__device__ __forceinline__ void chunk_range(int n, int chunk, int &lb, int &beg, int &end)
{
    lb = xcd_block();
    beg = lb * chunk;
    end = min(n, beg + chunk);
}

} // namespace hmfem
