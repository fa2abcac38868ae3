// Block solve: several right-hand sides of one operator share one hierarchy
// and one pass over the coefficients per kernel.
//
// Reference counterpart: the toolbox solves k interleaved right-hand sides
// with one operator (linear_operator8, generic_operator.h:335-391; BLAS-1 over
// packed vectors with stride k, algorithm.h:7-324). Here every vector of the
// block is stored interleaved [cell][NR] (FP64 CG vectors, FP32 V-cycle
// vectors), so one thread loads a cell's coefficients once and updates all NR
// right-hand sides; each right-hand side keeps its own CG scalars and stops on
// its own test (a finished one is frozen: alpha = 0, its search direction is
// zeroed).
//
// The V-cycle is the single-RHS cycle with the block red-black Gauss-Seidel
// smoother (tiles of 32 x 8 x 8 cells, NR values per cell in LDS), the
// residual-restriction marching along z, and the single-workgroup tail run
// once per right-hand side on strided views of the block vectors.
#include "hmfem_hier.h"

#include <string.h>

#include <vector>

namespace hmfem {

constexpr int kMaxRhs = 8;
// (The MODE 2 instance holds 205 VGPRs, one 512-thread block per CU. Two cells
// per thread with 1024-thread blocks, or a 128-VGPR cap, spill 272-580 bytes
// per lane; fewer right-hand sides per pass (two: 137 VGPRs, one: 101) with
// the dot partials summed per wave in LDS ran the 256^3 block solve in 854
// and 921 ms against 794.)
// z extent of a block-smoother tile: 32 x 8 x 8 tiles, two threads per
// column. At 256^3 with four right-hand sides a CG iteration takes 3.19 ms
// against 3.05 for 32 x 8 x 4 tiles (one thread per column), 3.82 (32 x 8 x 4,
// two per column) and 3.95 (32 x 8 x 8, four per column), but on the 512^3
// sigma=3 conduit field the thinner tiles (more frozen ring) need 10-15% more
// iterations, which costs more than they save.
constexpr int kBTz = 8;

struct BlockScalars {
    double bb[kMaxRhs], rr[kMaxRhs], sigma_first[kMaxRhs], beta[kMaxRhs], tau[kMaxRhs];
    double sigma[2][kMaxRhs], alpha[2][kMaxRhs];
    double rtol2;
    int done[kMaxRhs], status[kMaxRhs], iter[kMaxRhs];
    int k0[kMaxRhs];   // the block iteration at which the lane's right-hand side started
    int all_done, kspmv, kupd, maxit, stop_rule, nr;
};

// Couplings of a cell to its six neighbours (0 where there is none) and the
// neighbour indices (the cell itself where there is none).
template <typename T>
struct Nb {
    T e, w[6];
    int s, q[6];
};

template <typename T>
__device__ __forceinline__ Nb<T> nb_at(const OpArgs<T> &o, int s)
{
    Cell c = cell_at(o, s);
    const int nx = o.nx, nxy = o.nx * o.ny;
    Nb<T> b;
    b.s = s;
    b.e = o.e[s];
    const bool xp = c.i < o.nx - 1, xm = c.i > 0, yp = c.j < o.ny - 1, ym = c.j > 0, zp = c.k < o.nz - 1,
               zm = c.k > 0;
    b.w[0] = xp ? o.wx[s] : (T)0;       b.q[0] = xp ? s + 1 : s;
    b.w[1] = xm ? o.wx[s - 1] : (T)0;   b.q[1] = xm ? s - 1 : s;
    b.w[2] = yp ? o.wy[s] : (T)0;       b.q[2] = yp ? s + nx : s;
    b.w[3] = ym ? o.wy[s - nx] : (T)0;  b.q[3] = ym ? s - nx : s;
    b.w[4] = zp ? o.wz[s] : (T)0;       b.q[4] = zp ? s + nxy : s;
    b.w[5] = zm ? o.wz[s - nxy] : (T)0; b.q[5] = zm ? s - nxy : s;
    return b;
}

// (A u)_s for NR interleaved vectors given by a functor val(cell, r).
template <int NR, typename C, typename T, typename V>
__device__ __forceinline__ void apply_block(const Nb<T> &b, V val, C (&out)[NR])
{
    C us[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) { us[r] = val(b.s, r); out[r] = (C)b.e * us[r]; }
#pragma unroll
    for (int f = 0; f < 6; f++) {
#pragma unroll
        for (int r = 0; r < NR; r++) out[r] += (C)b.w[f] * (us[r] - val(b.q[f], r));
    }
}

// The NR values of one cell of an interleaved block vector ([cell][NR],
// 16-byte aligned for NR = 2 and 4) in vector loads.
template <int NR>
__device__ __forceinline__ void ld_lanes(const double *p, double (&v)[NR])
{
    static_assert(NR % 2 == 0, "pairs");
#pragma unroll
    for (int h = 0; h < NR / 2; h++) {
        const double2 t = reinterpret_cast<const double2 *>(p)[h];
        v[2 * h] = t.x;
        v[2 * h + 1] = t.y;
    }
}

template <int NR>
__device__ __forceinline__ void ld_lanes(const float *p, float (&v)[NR])
{
    static_assert(NR == 2 || NR == 4, "two or four lanes");
    if constexpr (NR == 4) {
        const float4 t = *reinterpret_cast<const float4 *>(p);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
        const float2 t = *reinterpret_cast<const float2 *>(p);
        v[0] = t.x; v[1] = t.y;
    }
}

// ---------------------------------------------------------------- CG kernels

// r = b - A x, partials (b.b, r.r) per right-hand side: P[lb][2 NR].
template <int NR>
__global__ void __launch_bounds__(kBlock)
k_b_init(OpArgs<double> o, const double *__restrict__ B, const double *__restrict__ X, double *__restrict__ R,
         float *__restrict__ R32, int chunk, double *P)
{
    const int n = o.nx * o.ny * o.nz;
    int lb, beg, end;
    chunk_range(n, chunk, lb, beg, end);
    double acc[2 * NR];
#pragma unroll
    for (int q = 0; q < 2 * NR; q++) acc[q] = 0.0;
    for (int s = beg + threadIdx.x; s < end; s += kBlock) {
        Nb<double> b = nb_at(o, s);
        double ax[NR];
        apply_block<NR, double>(b, [&](int q, int r) { return X[(long)q * NR + r]; }, ax);
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const double bv = B[(long)s * NR + r];
            const double rv = bv - ax[r];
            R[(long)s * NR + r] = rv;
            R32[(long)s * NR + r] = (float)rv;
            acc[r] += bv * bv;
            acc[NR + r] += rv * rv;
        }
    }
    block_sum<2 * NR>(acc);
    if (threadIdx.x == 0)
        for (int q = 0; q < 2 * NR; q++) P[(long)2 * NR * lb + q] = acc[q];
}

// r = b - A x for lane `lane` only (a right-hand side put into a finished
// lane), its partials (b.b, r.r) into the lane's slots of P[lb][2 NR].
template <int NR>
__global__ void __launch_bounds__(kBlock)
k_b_init_lane(OpArgs<double> o, const double *__restrict__ B, const double *__restrict__ X, double *__restrict__ R,
              float *__restrict__ R32, int lane, int chunk, double *P)
{
    const int n = o.nx * o.ny * o.nz;
    int lb, beg, end;
    chunk_range(n, chunk, lb, beg, end);
    double acc[2] = {0.0, 0.0};
    for (int s = beg + threadIdx.x; s < end; s += kBlock) {
        Nb<double> b = nb_at(o, s);
        double ax = b.e * X[(long)s * NR + lane];
        for (int f = 0; f < 6; f++) ax += b.w[f] * (X[(long)s * NR + lane] - X[(long)b.q[f] * NR + lane]);
        const double bv = B[(long)s * NR + lane];
        const double rv = bv - ax;
        R[(long)s * NR + lane] = rv;
        R32[(long)s * NR + lane] = (float)rv;
        acc[0] += bv * bv;
        acc[1] += rv * rv;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        P[(long)2 * NR * lb + lane] = acc[0];
        P[(long)2 * NR * lb + NR + lane] = acc[1];
    }
}

// Start of block CG iteration k: per right-hand side the stopping tests on
// the partial sums of the previous kernels and beta; live[r] = 0 for a
// finished one. Returns true once every right-hand side has finished. The dot
// partials have Gdot rows (G, plus one for the conduit correction).
template <int NR>
__device__ __forceinline__ bool b_iteration_scalars(int G, int Gdot, BlockScalars *sc, const double *Pinit,
                                                    const double *Prr, const double *Pdot, int &k, double (&beta)[NR],
                                                    int (&live)[NR])
{
    // a lane's own iteration count kr = k - k0: kr = 0 for a right-hand side
    // started at this iteration (its b.b and r.r are in Pinit)
    k = sc->kspmv;
    double dots[2 * NR], rrs[NR], bbs[NR];
    sum_partials<2 * NR>(Pdot, Gdot, dots);
    int fresh[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) fresh[r] = !sc->done[r] && k == sc->k0[r];
    {
        // per lane r.r (and b.b of a fresh lane) from its own partials: Pinit
        // for a fresh lane, Prr otherwise, in one pass (few registers live)
        double acc[2 * NR];
#pragma unroll
        for (int q = 0; q < 2 * NR; q++) acc[q] = 0.0;
        for (int b = threadIdx.x; b < G; b += kBlock) {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                if (fresh[r]) {
                    acc[r] += Pinit[(long)2 * NR * b + NR + r];
                    acc[NR + r] += Pinit[(long)2 * NR * b + r];
                } else if (k > 0) {
                    acc[r] += Prr[(long)NR * b + r];
                }
            }
        }
        block_sum<2 * NR>(acc);
#pragma unroll
        for (int r = 0; r < NR; r++) {
            rrs[r] = acc[r];
            bbs[r] = fresh[r] ? acc[NR + r] : sc->bb[r];
        }
    }
    int all = 1;
    const bool writer = xcd_block() == 0 && threadIdx.x == 0;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        beta[r] = 0.0;
        live[r] = 0;
        if (sc->done[r]) continue;
        const int kr = k - sc->k0[r];
        const double sigma = fabs(dots[r]);
        int stop = 0, status = HMFEM_OK;
        const double sf = kr == 0 ? (sigma == 0.0 ? 1.0 : sigma) : sc->sigma_first[r];
        if (kr == 0) {
            if (sigma == 0.0 || (sc->stop_rule == HMFEM_STOP_RESIDUAL && (bbs[r] == 0.0 || rrs[r] <= sc->rtol2 * bbs[r])))
                stop = 1;
        } else {
            beta[r] = -sc->alpha[(k - 1) & 1][r] * dots[NR + r] / sc->sigma[(k - 1) & 1][r];
            if (sc->stop_rule == HMFEM_STOP_RESIDUAL && rrs[r] <= sc->rtol2 * bbs[r]) stop = 1;
            if (sc->stop_rule == HMFEM_STOP_PRECONDITIONED && sigma <= sc->rtol2 * sf) stop = 1;
            if (!(sigma == sigma) || !(beta[r] == beta[r])) { stop = 1; status = HMFEM_ERR_BREAKDOWN; }
            else if (!stop && kr >= sc->maxit) { stop = 1; status = HMFEM_NOT_CONVERGED; }
        }
        if (writer) {
            if (kr == 0) { sc->bb[r] = bbs[r]; sc->sigma_first[r] = sf; }
            sc->rr[r] = rrs[r];
            sc->sigma[k & 1][r] = sigma;
            sc->beta[r] = beta[r];
            if (stop) { sc->done[r] = 1; sc->status[r] = status; sc->iter[r] = kr; }
        }
        if (stop) beta[r] = 0.0;
        else { live[r] = 1; all = 0; }
    }
    if (writer) {
        sc->kupd = k;
        if (all) sc->all_done = 1;
    }
    return all != 0;
}

// The stopping tests and beta of block CG iteration sc->kspmv (one block).
template <int NR>
__global__ void __launch_bounds__(kBlock)
k_b_scalars(int G, int Gdot, BlockScalars *sc, const double *Pinit, const double *Prr, const double *Pdot)
{
    if (sc->all_done) return;
    int k;
    double beta[NR];
    int live[NR];
    (void)b_iteration_scalars<NR>(G, Gdot, sc, Pinit, Prr, Pdot, k, beta, live);
}

// Block CG SpMV marching along z (as the single-RHS kernel): s = z + beta s_old
// for every live right-hand side (0 for finished ones), v = A s, partials
// <s, v>_r per logical block (Ptau[lb][NR]; blocks past the last tile write 0).
template <int NR>
__global__ void __launch_bounds__(kBlock)
k_b_spmv_z(OpArgs<double> o, const float *__restrict__ Z, double *__restrict__ S0, double *__restrict__ S1,
           double *__restrict__ V, int zlen, int G, int Gdot, BlockScalars *sc, const double *Pinit,
           const double *Prr, const double *Pdot, double *Ptau)
{
    // the iteration's scalars come from k_b_scalars (one block, just before):
    // uniform loads, so beta and the live flags sit in scalar registers (the
    // stopping tests in every block held 135 VGPRs: three waves per SIMD)
    if (sc->all_done) return;
    (void)G; (void)Gdot; (void)Pinit; (void)Prr; (void)Pdot;
    const int k = sc->kspmv;
    double beta[NR];
    int live[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        live[r] = !sc->done[r];
        beta[r] = live[r] ? sc->beta[r] : 0.0;
    }
    const double *__restrict__ sin = (k & 1) ? S1 : S0;
    double *__restrict__ sout = (k & 1) ? S0 : S1;
    __shared__ double sh[3][kGsTy + 2][kGsTx + 2][NR];
    const int nx = o.nx, ny = o.ny, nz = o.nz;
    const long nxy = (long)nx * ny;
    const int ntx = (nx + kGsTx - 1) / kGsTx, nty = (ny + kGsTy - 1) / kGsTy;
    const int ntxy = ntx * nty, nzc = (nz + zlen - 1) / zlen;
    const int lb = xcd_block();
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r] = 0.0;
    if (lb < ntxy * nzc) {
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
        // a cell's NR values of z and s_old in vector loads (all lanes: a
        // finished lane's z and s are finite, s = 0 since it finished; its
        // value is then dropped), s = z + beta s_old for the live ones
        auto cell_s = [&](long c, double (&a)[NR]) {
            float zv[NR];
            double sv[NR];
            ld_lanes<NR>(Z + c * NR, zv);
            if (k) ld_lanes<NR>(sin + c * NR, sv);
            else
#pragma unroll
                for (int r = 0; r < NR; r++) sv[r] = 0.0;
#pragma unroll
            for (int r = 0; r < NR; r++) a[r] = live[r] ? (double)zv[r] + (k ? beta[r] * sv[r] : 0.0) : 0.0;
        };
        auto fill = [&](int kk) {
            const int slot = (kk + 3) % 3;
            const bool in = kk >= 0 && kk < nz;
            double a[NR], b[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) { a[r] = 0.0; b[r] = 0.0; }
            // a lane that restarts has beta 0 (and its s_old is 0: it was
            // idle, s = 0, on the iteration before its restart)
            if (in && col) cell_s(i + (long)j * nx + (long)kk * nxy, a);
            if (in && ring_in) cell_s(ri + (long)rj * nx + (long)kk * nxy, b);
#pragma unroll
            for (int r = 0; r < NR; r++) {
                sh[slot][ty + 1][tx + 1][r] = a[r];
                if (has_ring) sh[slot][ry][rx][r] = b[r];
            }
        };
        fill(kb - 1);
        fill(kb);
        double wz_lo = col && kb > 0 ? o.wz[i + (long)j * nx + (long)(kb - 1) * nxy] : 0.0;
        for (int kk = kb; kk < ke; kk++) {
            fill(kk + 1);
            __syncthreads();
            const int c = (kk + 3) % 3, lo = (kk + 2) % 3, hi = (kk + 1) % 3;
            if (col) {
                const long q = i + (long)j * nx + (long)kk * nxy;
                const double e = o.e[q];
                const double xp = i < nx - 1 ? o.wx[q] : 0.0, xm = i > 0 ? o.wx[q - 1] : 0.0;
                const double yp = j < ny - 1 ? o.wy[q] : 0.0, ym = j > 0 ? o.wy[q - nx] : 0.0;
                const double zp = kk < nz - 1 ? o.wz[q] : 0.0;
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const double us = sh[c][ty + 1][tx + 1][r];
                    double a = e * us;
                    a += xp * (us - sh[c][ty + 1][tx + 2][r]);
                    a += xm * (us - sh[c][ty + 1][tx][r]);
                    a += yp * (us - sh[c][ty + 2][tx + 1][r]);
                    a += ym * (us - sh[c][ty][tx + 1][r]);
                    a += zp * (us - sh[hi][ty + 1][tx + 1][r]);
                    a += wz_lo * (us - sh[lo][ty + 1][tx + 1][r]);
                    sout[q * NR + r] = us;
                    V[q * NR + r] = a;
                    acc[r] += us * a;
                }
                wz_lo = zp;
            }
            __syncthreads();
        }
    }
    block_sum<NR>(acc);
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; r++) Ptau[(long)NR * lb + r] = acc[r];
}

// alpha = sigma / tau per right-hand side; x += alpha s, r -= alpha v; partials r.r.
template <int NR>
__global__ void __launch_bounds__(kBlock)
k_b_update(int n, const double *__restrict__ S0, const double *__restrict__ S1, const double *__restrict__ V,
           double *__restrict__ X, double *__restrict__ R, float *__restrict__ R32, int chunk, int Gtau,
           BlockScalars *sc, const double *Ptau, double *Prr)
{
    if (sc->all_done) return;
    const int k = sc->kupd;
    const double *__restrict__ S = (k & 1) ? S0 : S1;
    double t[NR];
    sum_partials<NR>(Ptau, Gtau, t);
    double alpha[NR];
    const bool writer = xcd_block() == 0 && threadIdx.x == 0;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        alpha[r] = 0.0;
        if (sc->done[r]) continue;
        const double tau = fabs(t[r]);
        if (tau == 0.0 || !(tau == tau)) {
            if (writer) { sc->done[r] = 1; sc->status[r] = HMFEM_ERR_BREAKDOWN; sc->iter[r] = k - sc->k0[r]; }
            continue;
        }
        alpha[r] = sc->sigma[k & 1][r] / tau;
    }
    if (writer) {
        for (int r = 0; r < NR; r++) { sc->alpha[k & 1][r] = alpha[r]; sc->tau[r] = t[r]; }
        sc->kspmv = k + 1;
    }
    int lb, beg, end;
    chunk_range(n, chunk, lb, beg, end);
    // flat over the interleaved values of the chunk's cells, one 8-byte value
    // per lane (coalesced); kBlock is a multiple of NR, so a thread always
    // serves right-hand side threadIdx.x % NR
    static_assert(kBlock % NR == 0, "a thread serves one right-hand side");
    const int mine = threadIdx.x % NR;
    double am = 0.0;
#pragma unroll
    for (int r = 0; r < NR; r++)
        if (r == mine) am = alpha[r];
    double my = 0.0;
    const long qe = (long)end * NR;
    for (long q = (long)beg * NR + threadIdx.x; q < qe; q += kBlock) {
        const double ri = R[q] - am * V[q];
        X[q] += am * S[q];
        R[q] = ri;
        R32[q] = (float)ri;
        my += ri * ri;
    }
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r] = r == mine ? my : 0.0;
    block_sum<NR>(acc);
    if (threadIdx.x == 0)
        for (int r = 0; r < NR; r++) Prr[(long)NR * lb + r] = acc[r];
}

// ---------------------------------------------------------------- V-cycle kernels

constexpr int kBCz = 4;                          // cells of a tile column per thread
// gs_sweeps 0: two sweeps (the 256^3 block workload, eight right-hand sides:
// 277 block iterations and 9.17 solves/s with two, 308 and 7.91 with three)
constexpr int kBlockSweeps = kSweepsSingle;
constexpr int kBThreads = kGsTx * kGsTy * (kBTz / kBCz);

// Block red-black Gauss-Seidel on tiles of kGsTx x kGsTy x kBTz cells for NR
// interleaved right-hand sides, as the single-RHS smoother (the ring of the
// tile frozen); a tile column is shared by kBTz / kBCz threads, so the z
// neighbour across their boundary is read from LDS. MODE 0: from zero;
// MODE 1: from u + ocf * uc[aggregate] (uo must not alias u); MODE 2: MODE 1
// plus the CG dot partials <z, f>_r, <z, v>_r of this block's tiles
// (Pdot[lb][2 NR]). The tile's coefficients are loaded once and the
// right-hand sides smoothed NP at a time: two at a time gives four waves per
// SIMD (128 VGPRs with a small spill) but fetches the interleaved lines of
// f, u and v once per pass; all four at once (two waves) was faster on the
// 256^3 block workload (9.79 against 9.57 right-hand sides/s).
template <int NR, typename FT, int MODE, int CZ = kBCz, int NTHREADS = kGsTx * kGsTy * (kBTz / CZ),
          int NP = NR>
__global__ void __launch_bounds__(NTHREADS)
k_b_gs_tile(OpArgs<float> o, int sweeps, const FT *__restrict__ f, const float *__restrict__ u,
            float *__restrict__ uo, int sx, int sy, int sz, int cx, int cy, const float *__restrict__ uc, float ocf,
            const double *__restrict__ v, int tpb, double *Pdot, const int *done)
{
    constexpr int TZ = kBTz, NTZ = kBTz / CZ;
    constexpr int RING = kGsRing;
    constexpr int RPER = (TZ * RING + NTHREADS - 1) / NTHREADS;
    static_assert(CZ % 2 == 0, "colour pairs");
    static_assert(NR % NP == 0, "whole passes");
    if (*done) return;
    __shared__ float sh[TZ][kGsTy + 2][kGsTx + 2][NP];
    const int nx = o.nx, ny = o.ny, nz = o.nz;
    const long nxy = (long)nx * ny;
    const int ntx = (nx + kGsTx - 1) / kGsTx, nty = (ny + kGsTy - 1) / kGsTy, ntz = (nz + TZ - 1) / TZ;
    const int ntiles = ntx * nty * ntz;
    const int lb = xcd_block();
    (void)tpb;   // tiles lb, lb + G, ...: neighbouring tiles are worked on at the same time (L2 reuse of rings)
    const int tx = threadIdx.x % kGsTx, ty = (threadIdx.x / kGsTx) % kGsTy, tz = threadIdx.x / (kGsTx * kGsTy);
    const int kbase = tz * CZ;   // first tile plane of this thread
    double acc[MODE == 2 ? 2 * NR : 1];
    if constexpr (MODE == 2)
        for (int q = 0; q < 2 * NR; q++) acc[q] = 0.0;

    // right-hand sides rb .. rb + NP - 1 of the cell's starting value (32-bit
    // offsets: the host keeps n * NR below 2^32, so that the loads take one
    // address register each)
    auto start = [&](int ii, int jj, int kk, int rb, float (&val)[NP]) {
#pragma unroll
        for (int r = 0; r < NP; r++) val[r] = 0.f;
        if (MODE == 0 || ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz) return;
        const unsigned s = (unsigned)ii + (unsigned)jj * (unsigned)nx + (unsigned)kk * (unsigned)nxy;
        const unsigned I = (unsigned)(ii >> sx) + (unsigned)cx * ((unsigned)(jj >> sy) + (unsigned)cy * (unsigned)(kk >> sz));
#pragma unroll
        for (int r = 0; r < NP; r++) val[r] = u[s * NR + rb + r] + ocf * uc[I * NR + rb + r];
    };

    for (int t = lb; t < ntiles; t += gridDim.x) {
        const int ti = t % ntx, tj = (t / ntx) % nty, tk = t / (ntx * nty);
        const int i0 = ti * kGsTx, j0 = tj * kGsTy, k0 = tk * TZ;
        const int i = i0 + tx, j = j0 + ty;
        const bool col = i < nx && j < ny;
        const int kc = k0 + kbase;   // grid plane of this thread's first cell
        const long s0 = i + (long)j * nx + (long)kc * nxy;

        float wxp[CZ], wxm[CZ], wyp[CZ], wym[CZ], wzp[CZ], dinv[CZ];
#pragma unroll
        for (int kk = 0; kk < CZ; kk++) {
            const bool act = col && kc + kk < nz;
            const long s = s0 + kk * nxy;
            wxp[kk] = act ? o.wx[s] : 0.f;
            wxm[kk] = act && i > 0 ? o.wx[s - 1] : 0.f;
            wyp[kk] = act ? o.wy[s] : 0.f;
            wym[kk] = act && j > 0 ? o.wy[s - nx] : 0.f;
            wzp[kk] = act ? o.wz[s] : 0.f;
            dinv[kk] = act ? o.di[s] : 0.f;
        }
        const float wzlo = col && kc > 0 && kc < nz ? o.wz[s0 - nxy] : 0.f;

#pragma unroll 1
        for (int rb = 0; rb < NR; rb += NP) {
            float ff[CZ][NP], uu[CZ][NP];
#pragma unroll
            for (int kk = 0; kk < CZ; kk++) {
                const bool act = col && kc + kk < nz;
                const unsigned s = (unsigned)(s0 + kk * nxy) * NR + rb;
                if constexpr (MODE == 0) {   // (the vector load takes it past 128 VGPRs: one workgroup per CU)
#pragma unroll
                    for (int r = 0; r < NP; r++) ff[kk][r] = act ? (float)f[s + r] : 0.f;
                } else {
                    FT fv[NP];
                    if (act) ld_lanes<NP>(f + s, fv);
#pragma unroll
                    for (int r = 0; r < NP; r++) ff[kk][r] = act ? (float)fv[r] : 0.f;
                }
            }
            // MODE 2: v for <z, v>, loaded with the pass's other values and held
            // through the sweeps (205 VGPRs, still two waves per SIMD; loaded
            // after the sweeps its latency showed: 808 -> 794 ms per 256^3 block solve)
            double vv[MODE == 2 ? CZ : 1][NP];
            if constexpr (MODE == 2)
#pragma unroll
                for (int kk = 0; kk < CZ; kk++) {
                    const bool act = v && col && kc + kk < nz;
                    const unsigned s = (unsigned)(s0 + kk * nxy) * NR + rb;
                    if (act) ld_lanes<NP>(v + s, vv[kk]);
                    else
#pragma unroll
                        for (int r = 0; r < NP; r++) vv[kk][r] = 0.0;
                }
            float zlo[NP], zhi[NP], ring[RPER][NP];
#pragma unroll
            for (int kk = 0; kk < CZ; kk++) start(i, j, kc + kk, rb, uu[kk]);
            if (tz == 0) start(i, j, k0 - 1, rb, zlo);
            if (tz == NTZ - 1) start(i, j, k0 + TZ, rb, zhi);
#pragma unroll
            for (int m = 0; m < RPER; m++) {
                const int q = threadIdx.x + m * NTHREADS;
                int hx = 0, hy = 0, hz = 0;
                if (q < TZ * RING) gs_ring_cell(q, hx, hy, hz);
                if (q < TZ * RING) start(i0 + hx - 1, j0 + hy - 1, k0 + hz, rb, ring[m]);
#pragma unroll
                for (int r = 0; r < NP; r++)
                    if (q < TZ * RING) sh[hz][hy][hx][r] = ring[m][r];
            }
#pragma unroll
            for (int kk = 0; kk < CZ; kk++)
#pragma unroll
                for (int r = 0; r < NP; r++) sh[kbase + kk][ty + 1][tx + 1][r] = uu[kk][r];
            __syncthreads();

            for (int sw = 0; sw < sweeps; sw++) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int c = MODE == 0 ? h : 1 - h;
                    const int sel = (c + i + j + kc) & 1;
#pragma unroll
                    for (int m = 0; m < CZ / 2; m++) {
                        const int ka = 2 * m, kb = 2 * m + 1;
                        const int kk = kbase + ka + sel;   // tile plane of the cell updated
                        const float cxp = sel ? wxp[kb] : wxp[ka], cxm = sel ? wxm[kb] : wxm[ka];
                        const float cyp = sel ? wyp[kb] : wyp[ka], cym = sel ? wym[kb] : wym[ka];
                        const float czp = sel ? wzp[kb] : wzp[ka];
                        const float czm = sel ? wzp[ka] : (m > 0 ? wzp[ka - 1] : wzlo);
                        const float di_ = sel ? dinv[kb] : dinv[ka];
#pragma unroll
                        for (int r = 0; r < NP; r++) {
                            float below, above;
                            if (m > 0) below = uu[ka - 1][r];
                            else below = tz == 0 ? zlo[r] : sh[kbase - 1][ty + 1][tx + 1][r];
                            if (kb + 1 < CZ) above = uu[kb + 1][r];
                            else above = tz == NTZ - 1 ? zhi[r] : sh[kbase + CZ][ty + 1][tx + 1][r];
                            const float uzm = sel ? uu[ka][r] : below;
                            const float uzp = sel ? above : uu[kb][r];
                            const float a = (sel ? ff[kb][r] : ff[ka][r]) + cxp * sh[kk][ty + 1][tx + 2][r] +
                                            cxm * sh[kk][ty + 1][tx][r] + cyp * sh[kk][ty + 2][tx + 1][r] +
                                            cym * sh[kk][ty][tx + 1][r] + czp * uzp + czm * uzm;
                            const float un = a * di_;
                            uu[ka][r] = sel ? uu[ka][r] : un;
                            uu[kb][r] = sel ? un : uu[kb][r];
                            sh[kk][ty + 1][tx + 1][r] = un;
                        }
                    }
                    __syncthreads();
                }
            }

#pragma unroll
            for (int kk = 0; kk < CZ; kk++) {
                if (col && kc + kk < nz) {
                    const unsigned s = (unsigned)(s0 + kk * nxy) * NR + rb;
#pragma unroll
                    for (int r = 0; r < NP; r++) {
                        uo[s + r] = uu[kk][r];
                        if constexpr (MODE == 2) {
                            // <z, r> from the FP32 copy of r this pass already holds: the
                            // FP64 r of the tile is gone from L2 by now, and z itself is FP32
                            acc[rb + r] += (double)uu[kk][r] * (double)ff[kk][r];
                            acc[NR + rb + r] += (double)uu[kk][r] * vv[kk][r];
                        }
                    }
                }
            }
            // the last half-sweep's barrier is the last LDS read of this pass:
            // the next pass (or tile) may fill the LDS at once
        }
    }
    if constexpr (MODE == 2) {
        block_sum_n<2 * NR, NTHREADS>(acc);
        if (threadIdx.x == 0)
            for (int q = 0; q < 2 * NR; q++) Pdot[(long)2 * NR * lb + q] = acc[q];
    }
}

// Residual f - A u summed over 2x2x2 aggregates for NR right-hand sides, one
// thread per fine cell (bricks of 32 x 4 x 2 fine cells).
template <int NR, typename FT>
__global__ void __launch_bounds__(kBlock)
k_b_restrict(OpArgs<float> o, int cx, int cy, int cz, const FT *__restrict__ f, const float *__restrict__ u,
             float *__restrict__ fc, const int *done)
{
    if (*done) return;
    __shared__ float upper[2][16][NR];
    const int nbx = (o.nx + 31) / 32, nby = (o.ny + 3) / 4;
    const int b = xcd_block();
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int t = threadIdx.x;
    const int i = bx * 32 + (t & 31), j = by * 4 + ((t >> 5) & 3), k = bz * 2 + (t >> 7);
    float res[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) res[r] = 0.f;
    if (i < o.nx && j < o.ny && k < o.nz) {
        const int s = i + j * o.nx + k * o.nx * o.ny;
        Nb<float> nb = nb_at(o, s);
        float au[NR];
        apply_block<NR, float>(nb, [&](int q, int r) { return u[(long)q * NR + r]; }, au);
#pragma unroll
        for (int r = 0; r < NR; r++) res[r] = (float)f[(long)s * NR + r] - au[r];
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        res[r] += __shfl_xor(res[r], 1);
        res[r] += __shfl_xor(res[r], 32);
    }
    const int lane = t & 63, cjl = (t >> 6) & 1, cil = (t & 31) >> 1;
    if ((t >> 7) == 1 && lane < 32 && !(lane & 1))
        for (int r = 0; r < NR; r++) upper[cjl][cil][r] = res[r];
    __syncthreads();
    if ((t >> 7) == 0 && lane < 32 && !(lane & 1)) {
        const int ci = bx * 16 + cil, cj = by * 2 + cjl, ck = bz;
        if (ci < cx && cj < cy && ck < cz) {
            const long I = ci + (long)cx * (cj + (long)cy * ck);
            for (int r = 0; r < NR; r++) fc[I * NR + r] = res[r] + upper[cjl][cil][r];
        }
    }
}

// The same marching along z (as the single-RHS k_restrict_z): a block owns a
// kGsTx x kGsTy column tile and planes [kb, kb + zlen) (both even); the
// column's NR values at k-1, k, k+1 stay in registers, the plane and its ring
// go through a double-buffered LDS plane; box sums over two planes in
// registers, then the x and y pairs by lane shuffles.
template <int NR, typename FT>
__global__ void __launch_bounds__(kBlock)
k_b_restrict_z(OpArgs<float> o, int zlen, int cx, int cy, int cz, const FT *__restrict__ f,
               const float *__restrict__ u, float *__restrict__ fc, const int *done)
{
    if (*done) return;
    __shared__ float sh[2][kGsTy + 2][kGsTx + 2][NR];
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
    float um[NR], uc[NR], up[NR], acc[NR];
    float wzm = 0.f;
#pragma unroll
    for (int r = 0; r < NR; r++) { um[r] = 0.f; uc[r] = 0.f; up[r] = 0.f; acc[r] = 0.f; }
    if (col) {
        if (kb > 0) {
            wzm = o.wz[cs + (kb - 1) * nxy];
#pragma unroll
            for (int r = 0; r < NR; r++) um[r] = u[(cs + (kb - 1) * nxy) * NR + r];
        }
#pragma unroll
        for (int r = 0; r < NR; r++) uc[r] = u[(cs + kb * nxy) * NR + r];
        if (kb + 1 < nz)
#pragma unroll
            for (int r = 0; r < NR; r++) up[r] = u[(cs + (kb + 1) * nxy) * NR + r];
    }
    for (int kk = kb; kk < ke; kk++) {
        const int bf = kk & 1;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            sh[bf][ty + 1][tx + 1][r] = uc[r];
            if (has_ring) sh[bf][ry][rx][r] = ring_in ? u[(rs + kk * nxy) * NR + r] : 0.f;
        }
        float e = 0.f, xp = 0.f, xm = 0.f, yp = 0.f, ym = 0.f, zp = 0.f, fq[NR], un[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) { fq[r] = 0.f; un[r] = 0.f; }
        if (col) {
            const long q = cs + kk * nxy;
            e = o.e[q];
            if (i < nx - 1) xp = o.wx[q];
            if (i > 0) xm = o.wx[q - 1];
            if (j < ny - 1) yp = o.wy[q];
            if (j > 0) ym = o.wy[q - nx];
            if (kk < nz - 1) zp = o.wz[q];
#pragma unroll
            for (int r = 0; r < NR; r++) fq[r] = (float)f[q * NR + r];
            if (kk + 2 < nz)
#pragma unroll
                for (int r = 0; r < NR; r++) un[r] = u[(q + 2 * nxy) * NR + r];
        }
        __syncthreads();
        if (col) {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const float us = uc[r];
                float a = e * us;
                a += xp * (us - sh[bf][ty + 1][tx + 2][r]);
                a += xm * (us - sh[bf][ty + 1][tx][r]);
                a += yp * (us - sh[bf][ty + 2][tx + 1][r]);
                a += ym * (us - sh[bf][ty][tx + 1][r]);
                a += zp * (us - up[r]);
                a += wzm * (us - um[r]);
                acc[r] += fq[r] - a;
                um[r] = uc[r]; uc[r] = up[r]; up[r] = un[r];
            }
            wzm = zp;
        }
        if (((kk - kb) & 1) || kk == ke - 1) {   // block-uniform
            const int ci = i >> 1, cj = j >> 1, ck = kk >> 1;
            const bool w = !(tx & 1) && !(ty & 1) && ci < cx && cj < cy && ck < cz;
#pragma unroll
            for (int r = 0; r < NR; r++) {
                float t = acc[r];
                t += __shfl_xor(t, 1);
                t += __shfl_xor(t, 32);
                if (w) fc[(ci + (long)cx * (cj + (long)cy * ck)) * NR + r] = t;
                acc[r] = 0.f;
            }
        }
    }
}

// Column r of a block <- a vector (plus an optional second vector).
__global__ void __launch_bounds__(kBlock)
k_b_put(long n, int nr, int r, const double *__restrict__ a, const double *__restrict__ add, double *__restrict__ Bk)
{
    long s = (long)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    double x = a ? a[s] : 0.0;
    if (add) x += add[s];
    Bk[s * nr + r] = x;
}

// A vector <- column r of a block.
__global__ void __launch_bounds__(kBlock)
k_b_get(long n, int nr, int r, const double *__restrict__ Bk, double *__restrict__ a)
{
    long s = (long)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    a[s] = Bk[s * nr + r];
}

// dst[s][0..m) <- src[s][lanes[0..m)] for m (1 or 2) lanes of a four-lane
// block (the compaction to two lanes; a lane of dst without a source is 0).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_b_pack2(long n, const T *__restrict__ src, int la, int lb, T *__restrict__ dst)
{
    long s = (long)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    dst[2 * s] = la >= 0 ? src[4 * s + la] : (T)0;
    dst[2 * s + 1] = lb >= 0 ? src[4 * s + lb] : (T)0;
}

// The dot partials of a four-lane block for two of its lanes: rows of w
// values per lane group ([row][g][4] -> [row][g][2], g < groups).
__global__ void __launch_bounds__(kBlock)
k_b_pack2_partials(int rows, int groups, const double *__restrict__ src, int la, int lb, double *__restrict__ dst)
{
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= rows * groups) return;
    const int row = q / groups, g = q % groups;
    const double *a = src + (long)row * groups * 4 + g * 4;
    double *b = dst + (long)row * groups * 2 + g * 2;
    b[0] = la >= 0 ? a[la] : 0.0;
    b[1] = lb >= 0 ? a[lb] : 0.0;
}

// Lane r of a two-lane block -> lane q of a four-lane block.
__global__ void __launch_bounds__(kBlock)
k_b_unpack2(long n, const double *__restrict__ src, int r, double *__restrict__ dst, int q)
{
    long s = (long)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    dst[4 * s + q] = src[2 * s + r];
}

// ---------------------------------------------------------------- host side

struct BlockLevelVecs {
    float *f = nullptr, *u = nullptr, *t = nullptr;
};

struct BlockState {
    int nr;
    Chunking ch;
    double *R, *S0, *S1, *V, *X;
    float *R32;                       // FP32 copy of R, the fine-level V-cycle's input
    float *Z;                         // fine-level V-cycle output (interleaved)
    BlockScalars *sc;
    double *Pinit, *Ptau, *Prr, *Pdot;
    BlockLevelVecs lv[kMaxLevels];
    int *done;                        // &sc->all_done
    const Deflation *D = nullptr;     // conduit correction (used when D->nc > 0)
};

static int lg2(int f) { return f == 2 ? 1 : 0; }

template <int NR, typename FT>
static void b_restrict(hipStream_t st, const Level &L, const Level &C, const FT *f, const float *u, float *fc,
                       const int *done)
{
    if (L.fx == 2 && L.fy == 2 && L.fz == 2) {
        const long ntxy = (long)((L.d.nx + kGsTx - 1) / kGsTx) * ((L.d.ny + kGsTy - 1) / kGsTy);
        int zl = 16;
        while (zl > 2 && ntxy * ((L.d.nz + zl - 1) / zl) < 2048) zl -= 2;
        const long nb = ntxy * ((L.d.nz + zl - 1) / zl);
        hipLaunchKernelGGL((k_b_restrict_z<NR, FT>), dim3(grid_for(nb, 1)), dim3(kBlock), 0, st, opargs<float>(L),
                           zl, C.d.nx, C.d.ny, C.d.nz, f, u, fc, done);
        return;
    }
    const int nb = ((L.d.nx + 31) / 32) * ((L.d.ny + 3) / 4) * ((L.d.nz + 1) / 2);
    hipLaunchKernelGGL((k_b_restrict<NR, FT>), dim3(grid_for(nb, 1)), dim3(kBlock), 0, st, opargs<float>(L),
                       C.d.nx, C.d.ny, C.d.nz, f, u, fc, done);
}

static int b_tiles(const Dims &d)
{
    return ((d.nx + kGsTx - 1) / kGsTx) * ((d.ny + kGsTy - 1) / kGsTy) * ((d.nz + kBTz - 1) / kBTz);
}

// V-cycle on level l >= 1 for all right-hand sides; result in S.lv[l].u.
template <int NR>
static int b_vcycle_coarse(hipStream_t st, Hierarchy &H, BlockState &S, int l, const hmfem_options &opt)
{
    Level &L = H.lev[l];
    BlockLevelVecs &X = S.lv[l];
    if (l == H.tail0) {
        for (int r = 0; r < NR; r++) {
            int rc = launch_tail_strided(st, H, opt.omega, ocf_of(opt, kOcfSingle), X.f + r, X.u + r, NR, S.done);
            if (rc) return rc;
        }
        return HMFEM_OK;
    }
    Level &C = H.lev[l + 1];
    const int GT = grid_for(b_tiles(L.d), 1);
    const int sw = opt.gs_sweeps < 1 ? kBlockSweeps : opt.gs_sweeps;
    hipLaunchKernelGGL((k_b_gs_tile<NR, float, 0>), dim3(GT), dim3(kBThreads), 0, st, opargs<float>(L), sw,
                       (const float *)X.f, (const float *)nullptr, X.u, 0, 0, 0, 0, 0, (const float *)nullptr, 0.f,
                       (const double *)nullptr, 1, (double *)nullptr, S.done);
    b_restrict<NR, float>(st, L, C, X.f, X.u, S.lv[l + 1].f, S.done);
    int rc = b_vcycle_coarse<NR>(st, H, S, l + 1, opt);
    if (rc) return rc;
    hipLaunchKernelGGL((k_b_gs_tile<NR, float, 1>), dim3(GT), dim3(kBThreads), 0, st, opargs<float>(L), sw,
                       (const float *)X.f, (const float *)X.u, X.t, lg2(L.fx), lg2(L.fy), lg2(L.fz), C.d.nx, C.d.ny,
                       (const float *)S.lv[l + 1].u, (float)ocf_of(opt, kOcfSingle), (const double *)nullptr, 1, (double *)nullptr,
                       S.done);
    float *tmp = X.u; X.u = X.t; X.t = tmp;
    return HMFEM_OK;
}

// Fine-level V-cycle: S.Z = M R with the dot partials of <z, r>, <z, v>.
template <int NR>
static int b_vcycle_fine(hipStream_t st, Hierarchy &H, BlockState &S, const double *Vp, const hmfem_options &opt)
{
    Level &L = H.lev[0], &C = H.lev[1];
    BlockLevelVecs &X = S.lv[0];
    const int nt = b_tiles(L.d);
    const int sw = opt.gs_sweeps < 1 ? kBlockSweeps : opt.gs_sweeps;
    const bool defl = S.D && S.D->nc > 0;
    if (defl) {   // per-cluster sums of r and v, the correction and its dot partials (row G of Pdot)
        int rc = deflation_sums_block(st, *S.D, NR, S.R, Vp, S.Pdot + (long)2 * NR * S.ch.G, S.done);
        if (rc) return rc;
    }
    hipLaunchKernelGGL((k_b_gs_tile<NR, float, 0>), dim3(grid_for(nt, 1)), dim3(kBThreads), 0, st, opargs<float>(L),
                       sw, (const float *)S.R32, (const float *)nullptr, X.u, 0, 0, 0, 0, 0, (const float *)nullptr,
                       0.f, (const double *)nullptr, 1, (double *)nullptr, S.done);
    b_restrict<NR, float>(st, L, C, S.R32, X.u, S.lv[1].f, S.done);
    int rc = b_vcycle_coarse<NR>(st, H, S, 1, opt);
    if (rc) return rc;
    const int tpb = (nt + S.ch.G - 1) / S.ch.G;
    hipLaunchKernelGGL((k_b_gs_tile<NR, float, 2>), dim3(S.ch.G), dim3(kBThreads), 0, st, opargs<float>(L), sw,
                       (const float *)S.R32, (const float *)X.u, S.Z, lg2(L.fx), lg2(L.fy), lg2(L.fz), C.d.nx,
                       C.d.ny, (const float *)S.lv[1].u, (float)ocf_of(opt, kOcfSingle), Vp, tpb, S.Pdot, S.done);
    if (defl) return deflation_scatter_block(st, *S.D, NR, S.Z, S.done);
    return HMFEM_OK;
}

static Chunking b_chunking(long n)
{
    Chunking c;
    long g = (n + kBlock * 4 - 1) / (kBlock * 4);
    if (g > kMaxReduceBlocks) g = kMaxReduceBlocks;
    if (g < kNumXcd) g = kNumXcd;
    g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
    long chunk = (n + g - 1) / g;
    chunk = (chunk + kBlock - 1) / kBlock * kBlock;
    c.G = (int)g;
    c.chunk = (int)chunk;
    return c;
}

size_t block_bytes(Dims fine, int nr, const hmfem_options &opt)
{
    hmfem_options o = opt;
    o.precision = HMFEM_PREC_FP32;
    const size_t n = fine.n();
    size_t b = hierarchy_bytes(fine, o);
    b += 4 * align256(n * nr * 8);                                   // R, S0, S1, V
    b += 2 * align256(n * nr * 4);                                   // Z, R32
    Dims d = fine;                                                   // f, u, t per level
    while (true) {
        b += 3 * align256((size_t)d.n() * nr * 4);
        if (d.n() <= 1 || (d.nx == 1 && d.ny == 1 && d.nz == 1)) break;
        d = Dims{d.nx > 1 ? (d.nx + 1) / 2 : 1, d.ny > 1 ? (d.ny + 1) / 2 : 1, d.nz > 1 ? (d.nz + 1) / 2 : 1};
    }
    b += align256(sizeof(BlockScalars)) + 4 * align256((size_t)(kMaxReduceBlocks + 1) * 2 * kMaxRhs * 8);
    b += align256(64 * sizeof(int)) + 64 * 256;
    return b;
}

// block_cg<4> returns this (an internal code) when the feed is empty and at
// most two lanes are still busy: the solve goes on with two lanes, whose
// kernels move half the bytes per iteration (a waiting lane of the four
// still moves its interleaved values).
constexpr int kBlockCompact = 1 << 20;

// Resume (resume = true): the state is that of a poll (S.Z and the dot
// partials current, h_sc the host copy of the scalars); lane r of this block
// is lane lane_map[r] of the caller's four-lane Xk4, into which a finished
// lane's x is copied before feed->done(lane_map[r], ...).
template <int NR>
static int block_cg(hipStream_t st, Hierarchy &H, const Stencil<double> &A, Dims d, BlockState &S,
                    const hmfem_options &opt, BlockScalars *h_sc, const double *Bk, BlockFeed *feed, int *kdone,
                    bool resume = false, const int *lane_map = nullptr, double *Xk4 = nullptr, int *busy_out = nullptr)
{
    const int n = (int)d.n();
    OpArgs<double> o;
    o.e = A.e; o.wx = A.wx; o.wy = A.wy; o.wz = A.wz; o.di = nullptr;
    o.nx = d.nx; o.ny = d.ny; o.nz = d.nz;
    const int G = S.ch.G;
    const int Gdot = S.D && S.D->nc > 0 ? G + 1 : G;
    // planes per block of the z-marching SpMV: at most kMaxReduceBlocks blocks
    const int ntxy = ((d.nx + kGsTx - 1) / kGsTx) * ((d.ny + kGsTy - 1) / kGsTy);
    if (ntxy > kMaxReduceBlocks) return HMFEM_ERR_ARG;
    int zlen = 4;
    while ((long)ntxy * ((d.nz + zlen - 1) / zlen) > kMaxReduceBlocks) zlen++;
    const int Gz = grid_for((long)ntxy * ((d.nz + zlen - 1) / zlen), 1);
    int rc = HMFEM_OK;
    if (!resume) {
        rc = b_vcycle_fine<NR>(st, H, S, nullptr, opt);
        if (rc) return rc;
        HMFEM_CHECK(hipGetLastError());
    }
    const int every = opt.check_every < 1 ? 1 : opt.check_every;
    int k = resume ? h_sc->kspmv : 0;
    // lanes refilled from the feed: a finished lane hands its result over and
    // takes the next right-hand side (its own k0, its residual from
    // k_b_init_lane), and the V-cycle is run again for every lane so that the
    // new one has z and <z, r> (the others get the same values again)
    bool busy[kMaxRhs] = {};
    for (int r = 0; r < NR; r++) busy[r] = feed != nullptr && (!resume || !h_sc->done[r]);
    bool feed_empty = resume;
    int kcap = k + opt.max_iterations + every;
    while (true) {
        for (int q = 0; q < every; q++, k++) {
            hipLaunchKernelGGL(k_b_scalars<NR>, dim3(1), dim3(kBlock), 0, st, G, Gdot, S.sc,
                               (const double *)S.Pinit, (const double *)S.Prr, (const double *)S.Pdot);
            hipLaunchKernelGGL(k_b_spmv_z<NR>, dim3(Gz), dim3(kBlock), 0, st, o, (const float *)S.Z, S.S0, S.S1,
                               S.V, zlen, G, Gdot, S.sc, S.Pinit, S.Prr, S.Pdot, S.Ptau);
            hipLaunchKernelGGL(k_b_update<NR>, dim3(G), dim3(kBlock), 0, st, n, S.S0, S.S1, S.V, S.X, S.R, S.R32,
                               S.ch.chunk, Gz, S.sc, S.Ptau, S.Prr);
            rc = b_vcycle_fine<NR>(st, H, S, S.V, opt);
            if (rc) return rc;
        }
        HMFEM_CHECK(hipGetLastError());
        HMFEM_CHECK(hipMemcpyAsync(h_sc, S.sc, sizeof(BlockScalars), hipMemcpyDeviceToHost, st));
        HMFEM_CHECK(hipStreamSynchronize(st));
        if (feed) {
            bool refilled = false;
            for (int r = 0; r < NR; r++) {
                if (!busy[r] || !h_sc->done[r]) continue;
                const double rel = h_sc->bb[r] > 0 ? sqrt(h_sc->rr[r] / h_sc->bb[r]) : 0.0;
                if (lane_map) {
                    hipLaunchKernelGGL(k_b_unpack2, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                                       (long)n, (const double *)S.X, r, Xk4, lane_map[r]);
                    feed->done(lane_map[r], h_sc->iter[r], rel, h_sc->status[r]);
                } else {
                    feed->done(r, h_sc->iter[r], rel, h_sc->status[r]);
                }
                busy[r] = false;
                if (feed_empty || !feed->load(r)) {
                    feed_empty = true;
                    continue;
                }
                hipLaunchKernelGGL(k_b_init_lane<NR>, dim3(G), dim3(kBlock), 0, st, o, Bk, (const double *)S.X, S.R,
                                   S.R32, r, S.ch.chunk, S.Pinit);
                h_sc->done[r] = 0;
                h_sc->status[r] = HMFEM_OK;
                h_sc->iter[r] = 0;
                h_sc->k0[r] = h_sc->kspmv;
                busy[r] = true;
                refilled = true;
            }
            if (refilled) {
                h_sc->all_done = 0;
                kcap = h_sc->kspmv + opt.max_iterations + every;
                HMFEM_CHECK(hipMemcpyAsync(S.sc, h_sc, sizeof(BlockScalars), hipMemcpyHostToDevice, st));
                rc = b_vcycle_fine<NR>(st, H, S, S.V, opt);
                if (rc) return rc;
                HMFEM_CHECK(hipGetLastError());
                continue;
            }
            int nbusy = 0;
            for (int r = 0; r < NR; r++) nbusy += busy[r];
            if (NR == 4 && busy_out && feed_empty && nbusy > 0 && nbusy <= 2 && !h_sc->all_done) {
                int m = 0;
                busy_out[0] = busy_out[1] = -1;
                for (int r = 0; r < NR; r++)
                    if (busy[r]) busy_out[m++] = r;
                if (kdone) *kdone = h_sc->kspmv;
                return kBlockCompact;
            }
        }
        if (h_sc->all_done || k > kcap) break;
    }
    if (kdone) *kdone = h_sc->kspmv;
    return HMFEM_OK;
}

// Continue a four-lane solve stopped by kBlockCompact with its busy lanes
// lanes[0..1] (-1: none) as a two-lane solve. The two-lane vectors take the
// space of four-lane ones whose values are not needed at a poll: R and the
// search direction into V and the spare S buffer, X, Z and R32 into R. On return
// X (Xk, four lanes) holds both lanes' results and hs the four-lane scalars.
static int block_compact(hipStream_t st, Hierarchy &H, const Stencil<double> &A, Dims d, BlockState &S,
                         const hmfem_options &opt, BlockScalars &hs, BlockFeed *feed, int *kdone, const int *lanes)
{
    const long n = d.n();
    const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
    const int la = lanes[0], lb = lanes[1];
    const int k = hs.kspmv;                                   // the next iteration
    double *sin4 = (k & 1) ? S.S1 : S.S0, *spare = (k & 1) ? S.S0 : S.S1;
    double *R2 = S.V, *V2 = S.V + 2 * n;
    double *S0_2 = spare, *S1_2 = spare + 2 * n;
    double *X2 = S.R;
    float *Z2 = (float *)(S.R + 2 * n), *R32_2 = (float *)(S.R + 3 * n);
    hipLaunchKernelGGL(k_b_pack2<double>, dim3(nb), dim3(kBlock), 0, st, n, (const double *)S.R, la, lb, R2);
    hipLaunchKernelGGL(k_b_pack2<double>, dim3(nb), dim3(kBlock), 0, st, n, (const double *)sin4, la, lb,
                       (k & 1) ? S1_2 : S0_2);
    hipLaunchKernelGGL(k_b_pack2<double>, dim3(nb), dim3(kBlock), 0, st, n, (const double *)S.X, la, lb, X2);
    hipLaunchKernelGGL(k_b_pack2<float>, dim3(nb), dim3(kBlock), 0, st, n, (const float *)S.Z, la, lb, Z2);
    hipLaunchKernelGGL(k_b_pack2<float>, dim3(nb), dim3(kBlock), 0, st, n, (const float *)S.R32, la, lb, R32_2);
    const int G = S.ch.G;
    const int Gdot = S.D && S.D->nc > 0 ? G + 1 : G;
    // partials of the next iteration's scalars: <z, r>, <z, v> rows (Pdot) and r.r rows (Prr)
    hipLaunchKernelGGL(k_b_pack2_partials, dim3((2 * Gdot + kBlock - 1) / kBlock), dim3(kBlock), 0, st, Gdot, 2,
                       (const double *)S.Pdot, la, lb, S.Ptau);
    hipLaunchKernelGGL(k_b_pack2_partials, dim3((G + kBlock - 1) / kBlock), dim3(kBlock), 0, st, G, 1,
                       (const double *)S.Prr, la, lb, S.Pinit);
    HMFEM_CHECK(hipGetLastError());
    BlockState S2 = S;
    S2.nr = 2;
    S2.R = R2; S2.V = V2; S2.S0 = S0_2; S2.S1 = S1_2; S2.X = X2; S2.Z = Z2; S2.R32 = R32_2;
    S2.Pdot = S.Ptau; S2.Ptau = S.Pdot; S2.Prr = S.Pinit; S2.Pinit = S.Prr;
    BlockScalars h2 = hs;
    const int map[2] = {la, lb};
    for (int r = 0; r < kMaxRhs; r++) {
        const int q = r < 2 ? map[r] : -1;
        if (q < 0) {
            h2.done[r] = 1; h2.status[r] = HMFEM_OK; h2.iter[r] = 0; h2.k0[r] = 0;
            h2.bb[r] = h2.rr[r] = h2.sigma_first[r] = h2.beta[r] = h2.tau[r] = 0.0;
            h2.sigma[0][r] = h2.sigma[1][r] = h2.alpha[0][r] = h2.alpha[1][r] = 0.0;
            continue;
        }
        h2.done[r] = hs.done[q]; h2.status[r] = hs.status[q]; h2.iter[r] = hs.iter[q]; h2.k0[r] = hs.k0[q];
        h2.bb[r] = hs.bb[q]; h2.rr[r] = hs.rr[q]; h2.sigma_first[r] = hs.sigma_first[q];
        h2.beta[r] = hs.beta[q]; h2.tau[r] = hs.tau[q];
        for (int p = 0; p < 2; p++) { h2.sigma[p][r] = hs.sigma[p][q]; h2.alpha[p][r] = hs.alpha[p][q]; }
    }
    h2.nr = 2;
    h2.all_done = 0;
    HMFEM_CHECK(hipMemcpyAsync(S.sc, &h2, sizeof(BlockScalars), hipMemcpyHostToDevice, st));
    int rc = block_cg<2>(st, H, A, d, S2, opt, &h2, nullptr, feed, kdone, true, map, S.X, nullptr);
    if (rc) return rc;
    // lanes still busy at the end (capped): their x back into the four-lane block
    for (int r = 0; r < 2; r++) {
        const int q = map[r];
        if (q < 0) continue;
        if (!h2.done[r])
            hipLaunchKernelGGL(k_b_unpack2, dim3(nb), dim3(kBlock), 0, st, n, (const double *)X2, r, S.X, q);
        hs.done[q] = h2.done[r]; hs.status[q] = h2.status[r]; hs.iter[q] = h2.iter[r];
        hs.bb[q] = h2.bb[r]; hs.rr[q] = h2.rr[r];
    }
    hs.kspmv = h2.kspmv;
    hs.all_done = h2.all_done;
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

// Solve A X = B for nr interleaved right-hand sides (B, X: [n][nr] FP64,
// device; X holds the initial guess). Memory comes from `alloc`.
int block_solve(hipStream_t st, Dims d, const Stencil<double> &A, int nr, const double *Bk, double *Xk,
                const hmfem_options &opt_in, char *(*alloc)(void *, size_t), void *actx, long lds_limit,
                std::vector<char> &stage, int *iterations, double *rel_residual, int *status, double *setup_ms,
                double *solve_ms, Deflation *defl, bool defl_setup, char *(*dalloc)(void *, size_t, int), void *dctx,
                BlockFeed *feed, int *block_iterations)
{
    if (nr != 2 && nr != 4) return HMFEM_ERR_ARG;
    hmfem_options opt = opt_in;
    opt.precision = HMFEM_PREC_FP32;
    if (opt.coarse_max > kCoarseMax || opt.coarse_max < 1) opt.coarse_max = kCoarseMax;
    if (opt.tail_max < kCoarseMax) opt.tail_max = kCoarseMax;
    if (opt.tail_max > 512) opt.tail_max = 512;   // the tail must run in LDS
    opt.tail_lds = 1;
    const size_t bytes = block_bytes(d, nr, opt);
    char *mem = alloc(actx, bytes);
    if (!mem) return HMFEM_ERR_NOMEM;
    Arena ar{mem, 0, bytes};
    const long n = d.n();
    BlockState S;
    S.nr = nr;
    S.ch = b_chunking(n);
    S.R = ar.take<double>(n * nr); S.S0 = ar.take<double>(n * nr); S.S1 = ar.take<double>(n * nr);
    S.V = ar.take<double>(n * nr); S.X = Xk;
    S.Z = ar.take<float>(n * nr);
    S.R32 = ar.take<float>(n * nr);
    S.sc = ar.take<BlockScalars>(1);
    S.Pinit = ar.take<double>((size_t)(kMaxReduceBlocks + 1) * 2 * kMaxRhs);
    S.Ptau = ar.take<double>((size_t)(kMaxReduceBlocks + 1) * 2 * kMaxRhs);
    S.Prr = ar.take<double>((size_t)(kMaxReduceBlocks + 1) * 2 * kMaxRhs);
    S.Pdot = ar.take<double>((size_t)(kMaxReduceBlocks + 1) * 2 * kMaxRhs);
    int *d_info = ar.take<int>(64);
    if (!S.R || !S.S0 || !S.S1 || !S.V || !S.Z || !S.R32 || !S.sc || !S.Pdot || !d_info) return HMFEM_ERR_NOMEM;
    S.done = &S.sc->all_done;

    hipEvent_t e0, e1, e2;
    HMFEM_CHECK(hipEventCreate(&e0));
    HMFEM_CHECK(hipEventCreate(&e1));
    HMFEM_CHECK(hipEventCreate(&e2));
    HMFEM_CHECK(hipEventRecord(e0, st));
    Hierarchy H;
    HMFEM_CHECK(hipMemsetAsync(d_info, 0, sizeof(int), st));
    int rc = HMFEM_OK;
    if (defl && defl_setup) {   // the conduit correction of this operator (kept for later groups)
        defl->nc = 0;
        defl->pending = false;
        defl->gen++;
        if (opt.deflate && dalloc) {
            rc = deflation_begin(st, d, A, *defl, dalloc, dctx);
            if (rc) return rc;
        }
    }
    rc = build_hierarchy(st, d, A, opt, ar, H, d_info, stage, lds_limit);
    if (rc) return rc;
    if (defl && defl->pending) {
        rc = deflation_finish(st, d, A, opt.deflate_ratio, *defl, dalloc, dctx);
        if (rc) return rc;
    }
    S.D = defl && opt.deflate ? defl : nullptr;
    if (H.nlev < 3 || H.tail0 <= 0 || !tail_in_lds(H)) return HMFEM_ERR_ARG;
    for (int l = 0; l <= H.tail0; l++) {
        const size_t nl = H.lev[l].d.n();
        if (l > 0) S.lv[l].f = ar.take<float>(nl * nr);
        S.lv[l].u = ar.take<float>(nl * nr);
        if (l < H.tail0) S.lv[l].t = ar.take<float>(nl * nr);
        if ((l > 0 && !S.lv[l].f) || !S.lv[l].u || (l < H.tail0 && !S.lv[l].t)) return HMFEM_ERR_NOMEM;
    }
    BlockScalars init{};
    init.rtol2 = opt.rtol * opt.rtol;
    init.maxit = opt.max_iterations;
    init.stop_rule = opt.stop_rule;
    init.nr = nr;
    for (int r = nr; r < kMaxRhs; r++) init.done[r] = 1;
    HMFEM_CHECK(hipMemcpyAsync(S.sc, &init, sizeof(BlockScalars), hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipEventRecord(e1, st));
    OpArgs<double> o;
    o.e = A.e; o.wx = A.wx; o.wy = A.wy; o.wz = A.wz; o.di = nullptr;
    o.nx = d.nx; o.ny = d.ny; o.nz = d.nz;
    BlockScalars hs{};
    switch (nr) {
    case 2:
        hipLaunchKernelGGL(k_b_init<2>, dim3(S.ch.G), dim3(kBlock), 0, st, o, Bk, Xk, S.R, S.R32, S.ch.chunk,
                           S.Pinit);
        rc = block_cg<2>(st, H, A, d, S, opt, &hs, Bk, feed, block_iterations);
        break;
    default: {
        hipLaunchKernelGGL(k_b_init<4>, dim3(S.ch.G), dim3(kBlock), 0, st, o, Bk, Xk, S.R, S.R32, S.ch.chunk,
                           S.Pinit);
        int lanes[2] = {-1, -1};
        rc = block_cg<4>(st, H, A, d, S, opt, &hs, Bk, feed, block_iterations, false, nullptr, nullptr,
                         opt.block_compact ? lanes : nullptr);
        if (rc == kBlockCompact) rc = block_compact(st, H, A, d, S, opt, hs, feed, block_iterations, lanes);
        break;
    }
    }
    if (rc) return rc;
    HMFEM_CHECK(hipEventRecord(e2, st));
    HMFEM_CHECK(hipEventSynchronize(e2));
    int info = 0;
    HMFEM_CHECK(hipMemcpy(&info, d_info, sizeof(int), hipMemcpyDeviceToHost));
    float t0 = 0.f, t1 = 0.f;
    (void)hipEventElapsedTime(&t0, e0, e1);
    (void)hipEventElapsedTime(&t1, e1, e2);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (setup_ms) *setup_ms = t0;
    if (solve_ms) *solve_ms = t1;
    for (int r = 0; r < nr; r++) {
        if (iterations) iterations[r] = hs.iter[r];
        if (rel_residual) rel_residual[r] = hs.bb[r] > 0 ? sqrt(hs.rr[r] / hs.bb[r]) : 0.0;
        if (status) status[r] = info ? HMFEM_ERR_BREAKDOWN : (hs.done[r] ? hs.status[r] : HMFEM_NOT_CONVERGED);
    }
    return HMFEM_OK;
}

void block_put(hipStream_t st, long n, int nr, int r, const double *a, const double *add, double *Bk)
{
    hipLaunchKernelGGL(k_b_put, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, n, nr, r, a, add,
                       Bk);
}

void block_get(hipStream_t st, long n, int nr, int r, const double *Bk, double *a)
{
    hipLaunchKernelGGL(k_b_get, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, n, nr, r, Bk, a);
}

} // namespace hmfem
