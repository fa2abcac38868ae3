// Block red-black Gauss-Seidel smoother on tiles, shared by the single
// instance and batched solvers.
#pragma once

#include "hmfem_hier.h"

namespace hmfem {

// The grid is cut into tiles of kGsTx x kGsTy x kGsTz cells, one workgroup
// of kGsThreads per tile, kGsCz cells of a tile column per thread. Inside a
// tile the smoother runs red-black Gauss-Seidel sweeps (colour (i+j+k) mod 2)
// while the values just outside the tile stay at what they were when the
// kernel started: a block-Jacobi smoother whose block solves are approximated
// by red-black sweeps. All sweeps of one smoothing step cost one pass over the
// level (coefficients in registers, the tile and its x/y ring in LDS), where
// damped Jacobi needs one pass per sweep. The post-smoother runs the colours
// in the opposite order.
//
// GS update of a cell: u_s = (f_s + sum_nb w u_nb) / d_s with d_s = e_s + sum w
// (the inverse diagonal di is precomputed), i.e. (A u)_s = f_s afterwards.
__host__ __device__ inline int gs_tiles(const Dims &d)
{
    return ((d.nx + kGsTx - 1) / kGsTx) * ((d.ny + kGsTy - 1) / kGsTy) * ((d.nz + kGsTz - 1) / kGsTz);
}

// MODE 0: pre-smoothing from u = 0 (u and uc unused), result in uo.
// MODE 1: post-smoothing from u + ocf * uc[aggregate] (computed on the fly,
//         also for the frozen ring), result in uo (uo must not alias u).
// MODE 2: MODE 1 on the fine level plus the CG dot partials <z, f>, <z, v>
//         of this block's tiles (Pdot[2 lb], Pdot[2 lb + 1]).
// sx, sy, sz: log2 of the aggregation factors (0 or 1) towards uc's level.
// gs_tile_body smooths tile t of the grid of o; MODE 2 adds its dot
// products to acc. Every thread of the workgroup calls it.
// The fine level stored in FP16: couplings of the symmetrically scaled
// operator D^-1/2 A D^-1/2 (unit diagonal, couplings w / sqrt(d_s d_nb) in
// (0, 1)) and the FP32 scale sq = sqrt(d). With it (H16) the smoother runs on
// scaled values u^ = sq u, f^ = f / sq, the same red-black update with a unit
// diagonal, and returns u = u^ / sq.
struct Fine16 {
    const __half *wx, *wy, *wz;
    const float *sq;
};

template <typename T, typename FT, int MODE, bool H16 = false>
__device__ __forceinline__ void gs_tile_body(const OpArgs<T> &o, int sweeps, const FT *__restrict__ f,
                                             const T *__restrict__ u, T *__restrict__ uo, int sx, int sy, int sz,
                                             int cx, int cy, const T *__restrict__ uc, T ocf,
                                             const double *__restrict__ v, int t, double (&acc)[2],
                                             Fine16 h = Fine16{})
{
    using C = T;
    constexpr int CZ = kGsCz, NTZ = kGsTz / kGsCz;
    constexpr int RPER = (kGsTz * kGsRing + kGsThreads - 1) / kGsThreads;
    __shared__ C sh[kGsTz][kGsTy + 2][kGsTx + 2];
    const int nx = o.nx, ny = o.ny, nz = o.nz;
    const long nxy = (long)nx * ny;
    const int ntx = (nx + kGsTx - 1) / kGsTx, nty = (ny + kGsTy - 1) / kGsTy;
    const int tx = threadIdx.x % kGsTx, ty = (threadIdx.x / kGsTx) % kGsTy, tz = threadIdx.x / (kGsTx * kGsTy);
    const int kbase = tz * CZ;   // first tile plane of this thread's cells

    // starting value of a cell (zero outside the grid)
    auto start = [&](int ii, int jj, int kk) -> C {
        if (MODE == 0 || ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz) return (C)0;
        long s = ii + (long)jj * nx + (long)kk * nxy;
        int I = (ii >> sx) + cx * ((jj >> sy) + cy * (kk >> sz));
        if constexpr (H16) return ((C)u[s] + ocf * (C)uc[I]) * (C)h.sq[s];
        return (C)u[s] + ocf * (C)uc[I];
    };

    const int ti = t % ntx, tj = (t / ntx) % nty, tk = t / (ntx * nty);
    const int i0 = ti * kGsTx, j0 = tj * kGsTy, k0 = tk * kGsTz;
    const int i = i0 + tx, j = j0 + ty;
    const bool col = i < nx && j < ny;
    const int kc = k0 + kbase;   // grid plane of this thread's first cell
    const long s0 = i + (long)j * nx + (long)kc * nxy;

    // this thread's coefficients (zero for cells outside the grid)
    C wxp[CZ], wxm[CZ], wyp[CZ], wym[CZ], wzp[CZ], dinv[CZ], ff[CZ], uu[CZ];
    double fd[MODE == 2 ? CZ : 1], vd[MODE == 2 ? CZ : 1];   // MODE 2: f and v for the dot products
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) {
        const bool act = col && kc + kk < nz;
        const long s = s0 + kk * nxy;
        double fv = act ? (double)f[s] : 0.0;
        if constexpr (H16) {
            wxp[kk] = act ? (C)__half2float(h.wx[s]) : (C)0;
            wxm[kk] = act && i > 0 ? (C)__half2float(h.wx[s - 1]) : (C)0;
            wyp[kk] = act ? (C)__half2float(h.wy[s]) : (C)0;
            wym[kk] = act && j > 0 ? (C)__half2float(h.wy[s - nx]) : (C)0;
            wzp[kk] = act ? (C)__half2float(h.wz[s]) : (C)0;
            dinv[kk] = act ? (C)h.sq[s] : (C)0;   // holds sq until the sweeps (unit diagonal there)
            ff[kk] = act ? (C)fv / dinv[kk] : (C)0;
        } else {
            wxp[kk] = act ? (C)o.wx[s] : (C)0;
            wxm[kk] = act && i > 0 ? (C)o.wx[s - 1] : (C)0;
            wyp[kk] = act ? (C)o.wy[s] : (C)0;
            wym[kk] = act && j > 0 ? (C)o.wy[s - nx] : (C)0;
            wzp[kk] = act ? (C)o.wz[s] : (C)0;
            dinv[kk] = act ? (C)o.di[s] : (C)0;
            ff[kk] = (C)fv;
        }
        if constexpr (MODE == 2) {
            fd[kk] = fv;
            vd[kk] = act && v ? v[s] : 0.0;   // loaded now, used after the sweeps
        }
    }
    C wzlo;
    if constexpr (H16) wzlo = col && kc > 0 && kc < nz ? (C)__half2float(h.wz[s0 - nxy]) : (C)0;
    else wzlo = col && kc > 0 && kc < nz ? (C)o.wz[s0 - nxy] : (C)0;
    // scaled mode: the cell's scale, and the unit diagonal of active cells
    C scale[H16 ? CZ : 1];
    if constexpr (H16) {
#pragma unroll
        for (int kk = 0; kk < CZ; kk++) {
            scale[kk] = dinv[kk];
            dinv[kk] = dinv[kk] > (C)0 ? (C)1 : (C)0;
        }
    }

    // starting values of this thread's cells, the tile's two z neighbours
    // of the column and RPER cells of the x/y ring, all loads issued
    // before the first use
    C ring[RPER];
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) uu[kk] = start(i, j, kc + kk);
    const C zlo = tz == 0 ? start(i, j, k0 - 1) : (C)0;
    const C zhi = tz == NTZ - 1 ? start(i, j, k0 + kGsTz) : (C)0;
#pragma unroll
    for (int m = 0; m < RPER; m++) {
        const int q = threadIdx.x + m * kGsThreads;
        int hx = 0, hy = 0, hz = 0;
        if (q < kGsTz * kGsRing) gs_ring_cell(q, hx, hy, hz);
        ring[m] = q < kGsTz * kGsRing ? start(i0 + hx - 1, j0 + hy - 1, k0 + hz) : (C)0;
    }
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) sh[kbase + kk][ty + 1][tx + 1] = uu[kk];
#pragma unroll
    for (int m = 0; m < RPER; m++) {
        const int q = threadIdx.x + m * kGsThreads;
        if (q < kGsTz * kGsRing) {
            int hx, hy, hz;
            gs_ring_cell(q, hx, hy, hz);
            sh[hz][hy][hx] = ring[m];
        }
    }
    __syncthreads();

    for (int sw = 0; sw < sweeps; sw++) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = MODE == 0 ? h : 1 - h;       // colour of this half sweep
            const int sel = (c + i + j + kc) & 1;      // cell 2m + sel of each pair has colour c
#pragma unroll
            for (int m = 0; m < CZ / 2; m++) {
                const int ka = 2 * m, kb = 2 * m + 1;
                const int kk = kbase + ka + sel;       // tile plane of the cell updated
                // z neighbours across the boundary between the threads of
                // a column come from LDS (updated in the previous half sweep)
                const C below = m > 0 ? uu[ka - 1] : (tz == 0 ? zlo : sh[kbase - 1][ty + 1][tx + 1]);
                const C above = kb + 1 < CZ ? uu[kb + 1] : (tz == NTZ - 1 ? zhi : sh[kbase + CZ][ty + 1][tx + 1]);
                const C uzm = sel ? uu[ka] : below;
                const C uzp = sel ? above : uu[kb];
                const C wzm_ = sel ? wzp[ka] : (m > 0 ? wzp[ka - 1] : wzlo);
                const C wzp_ = sel ? wzp[kb] : wzp[ka];
                const C a = (sel ? ff[kb] : ff[ka]) + (sel ? wxp[kb] : wxp[ka]) * sh[kk][ty + 1][tx + 2] +
                            (sel ? wxm[kb] : wxm[ka]) * sh[kk][ty + 1][tx] +
                            (sel ? wyp[kb] : wyp[ka]) * sh[kk][ty + 2][tx + 1] +
                            (sel ? wym[kb] : wym[ka]) * sh[kk][ty][tx + 1] + wzp_ * uzp + wzm_ * uzm;
                const C un = a * (sel ? dinv[kb] : dinv[ka]);
                uu[ka] = sel ? uu[ka] : un;
                uu[kb] = sel ? un : uu[kb];
                sh[kk][ty + 1][tx + 1] = un;
            }
            __syncthreads();
        }
    }

#pragma unroll
    for (int kk = 0; kk < CZ; kk++) {
        if (col && kc + kk < nz) {
            const long s = s0 + kk * nxy;
            C z = uu[kk];
            if constexpr (H16) z = z / scale[kk];
            uo[s] = z;
            if constexpr (MODE == 2) {
                acc[0] += (double)z * fd[kk];
                if (v) acc[1] += (double)z * vd[kk];
            }
        }
    }
}

// Residual f - A u summed over 2x2x2 aggregates on the FP16-stored fine level:
// (A u)_s = sq_s (sq_s u_s - sum_nb w^_nb sq_nb u_nb). Bricks of 32 x 4 x 2
// fine cells as in the FP32 kernel.
template <typename T, typename FT>
__global__ void __launch_bounds__(kBlock)
k_restrict_brick16(int nx, int ny, int nz, Fine16 h, int cx, int cy, int cz, const FT *__restrict__ f,
                   const T *__restrict__ u, T *__restrict__ fc, const int *done)
{
    using C = T;
    if (*done) return;
    __shared__ C upper[2][16];
    const int nbx = (nx + 31) / 32, nby = (ny + 3) / 4;
    const int b = xcd_block();
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int t = threadIdx.x;
    const int i = bx * 32 + (t & 31), j = by * 4 + ((t >> 5) & 3), k = bz * 2 + (t >> 7);
    C r = 0;
    if (i < nx && j < ny && k < nz) {
        const int nxy = nx * ny;
        const int s = i + j * nx + k * nxy;
        auto su = [&](int q) -> C { return (C)h.sq[q] * (C)u[q]; };
        C a = su(s);
        if (i < nx - 1) a -= (C)__half2float(h.wx[s]) * su(s + 1);
        if (i > 0) a -= (C)__half2float(h.wx[s - 1]) * su(s - 1);
        if (j < ny - 1) a -= (C)__half2float(h.wy[s]) * su(s + nx);
        if (j > 0) a -= (C)__half2float(h.wy[s - nx]) * su(s - nx);
        if (k < nz - 1) a -= (C)__half2float(h.wz[s]) * su(s + nxy);
        if (k > 0) a -= (C)__half2float(h.wz[s - nxy]) * su(s - nxy);
        r = (C)f[s] - (C)h.sq[s] * a;
    }
    r += __shfl_xor(r, 1);
    r += __shfl_xor(r, 32);
    const int lane = t & 63, cjl = (t >> 6) & 1, cil = (t & 31) >> 1;
    if ((t >> 7) == 1 && lane < 32 && !(lane & 1)) upper[cjl][cil] = r;
    __syncthreads();
    if ((t >> 7) == 0 && lane < 32 && !(lane & 1)) {
        const int ci = bx * 16 + cil, cj = by * 2 + cjl, ck = bz;
        if (ci < cx && cj < cy && ck < cz) fc[ci + cx * (cj + cy * ck)] = r + upper[cjl][cil];
    }
}

} // namespace hmfem
