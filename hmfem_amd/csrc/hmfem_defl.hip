// Conduit deflation: an additive coarse correction on clusters of strongly
// coupled cells.
//
// Karst conduits are thin connected sets of cells whose permeability exceeds
// the matrix rock's by 1e4..1e6. Along a conduit the pressure is nearly
// constant, and the box-aggregation V-cycle resolves such long, thin, stiff
// paths slowly (tiles freeze their halos, boxes mix conduit and rock). The
// correction adds, per cluster c of "strong" cells (diagonal d_s at least
// `ratio` times the median diagonal, 7-point connected, >= kMinCluster cells),
// one piecewise-constant basis vector 1_c:
//     z = V(r) + Z E^-1 Z^T r,   Z = [1_c], E = Z^T A Z  (dense, FP64),
// which keeps the preconditioner symmetric positive definite. No counterpart
// in the reference (its AMG has no such stage, amg.h:1135-1239).
//
// Setup: a histogram of log2(d) gives the median; the flags of strong cells
// go to the host, which finds the clusters (union-find over the strong cells
// in index order, deterministic), gathers the coefficients of their cells on
// the device, builds E in a fixed order and inverts it (Cholesky). Per CG
// iteration: per-cluster sums of r and v (one block per cluster, fixed order),
// c = E^-1 (Z^T r) with the dot products c.(Z^T r), c.(Z^T v) written as one
// more partial of the fine-level dots, and z += Z c after the V-cycle.
#include "hmfem_stencil.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <numeric>
#include <thread>
#include <vector>

namespace hmfem {

constexpr int kHistBins = Deflation::kHistBins;
constexpr int kMinCluster = 8;
constexpr int kMaxClusters = 256;

__global__ void __launch_bounds__(kBlock)
k_diag_hist(OpArgs<double> o, unsigned *hist)
{
    __shared__ unsigned h[kHistBins];
    for (int q = threadIdx.x; q < kHistBins; q += kBlock) h[q] = 0;
    __syncthreads();
    const int n = o.nx * o.ny * o.nz;
    int run_bin = -1;   // runs of equal bins counted in a register, as in k_diag_hist_inst
    unsigned run = 0;
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
        Cell c = cell_at(o, s);
        const double d = diag_at<double, double>(o, c);
        int b = d > 0 ? ilogb(d) + 1100 : 0;
        b = b < 0 ? 0 : (b >= kHistBins ? kHistBins - 1 : b);
        if (b != run_bin) {
            if (run) atomicAdd(&h[run_bin], run);
            run_bin = b;
            run = 0;
        }
        run++;
    }
    if (run) atomicAdd(&h[run_bin], run);
    __syncthreads();
    for (int q = threadIdx.x; q < kHistBins; q += kBlock)
        if (h[q]) atomicAdd(&hist[q], h[q]);
}

// Grids handled by one setup: a single grid, or the instances of a batch
// concatenated in one index space. Block b of the per-cell setup kernels
// covers cells [(b - blk0) * kBlock, ...) of instance bmap[b] (bmap null: one
// instance).
struct DefInst {
    int nx, ny, nz, n;
    int off, blk0;
    double thr;   // strong-cell threshold of the instance (inf: none)
    int hblk0;    // first block of the instance in the histogram launch
};

constexpr int kHistCells = 16;   // cells per thread of the batch histogram

__device__ __forceinline__ bool inst_cell(const DefInst *__restrict__ I, const int *__restrict__ bmap,
                                          const Stencil<double> &A, DefInst &in, OpArgs<double> &o, Cell &c)
{
    const int q = bmap ? bmap[blockIdx.x] : 0;
    in = I[q];
    const int l = (blockIdx.x - in.blk0) * kBlock + threadIdx.x;
    o.e = A.e + in.off; o.wx = A.wx + in.off; o.wy = A.wy + in.off; o.wz = A.wz + in.off; o.di = nullptr;
    o.nx = in.nx; o.ny = in.ny; o.nz = in.nz;
    if (l >= in.n) return false;
    c = cell_at(o, l);
    return true;
}

// Per-instance histogram of floor(log2 d); block b covers kBlock * kHistCells
// cells of instance hmap[b].
__global__ void __launch_bounds__(kBlock)
k_diag_hist_inst(const DefInst *__restrict__ I, const int *__restrict__ hmap, Stencil<double> A, unsigned *hist)
{
    __shared__ unsigned h[kHistBins];
    for (int q = threadIdx.x; q < kHistBins; q += kBlock) h[q] = 0;
    __syncthreads();
    const int q = hmap[blockIdx.x];
    const DefInst in = I[q];
    OpArgs<double> o;
    o.e = A.e + in.off; o.wx = A.wx + in.off; o.wy = A.wy + in.off; o.wz = A.wz + in.off; o.di = nullptr;
    o.nx = in.nx; o.ny = in.ny; o.nz = in.nz;
    const int l0 = (blockIdx.x - in.hblk0) * kBlock * kHistCells + threadIdx.x;
    // runs of equal bins are counted in a register (fields are smooth or
    // piecewise constant, so most of a thread's cells share a bin): few LDS
    // atomics on the crowded bins
    int run_bin = -1;
    unsigned run = 0;
    for (int t = 0; t < kHistCells; t++) {
        const int l = l0 + t * kBlock;
        if (l >= in.n) break;
        const double d = diag_at<double, double>(o, cell_at(o, l));
        int b = d > 0 ? ilogb(d) + 1100 : 0;
        b = b < 0 ? 0 : (b >= kHistBins ? kHistBins - 1 : b);
        if (b != run_bin) {
            if (run) atomicAdd(&h[run_bin], run);
            run_bin = b;
            run = 0;
        }
        run++;
    }
    if (run) atomicAdd(&h[run_bin], run);
    __syncthreads();
    for (int t = threadIdx.x; t < kHistBins; t += kBlock)
        if (h[t]) atomicAdd(&hist[(long)q * kHistBins + t], h[t]);
}

// Strong cells (diagonal above the instance's threshold) in index order, in
// three steps: a bit mask per wave and a count per block, an exclusive scan
// of the counts, and the ordered list of global indices.
constexpr int kWavesPerBlock = kBlock / 64;

__global__ void __launch_bounds__(kBlock)
k_strong_mask(const DefInst *__restrict__ I, const int *__restrict__ bmap, Stencil<double> A,
              unsigned long long *mask, int *bcnt)
{
    __shared__ int wc[kWavesPerBlock];
    DefInst in;
    OpArgs<double> o;
    Cell c;
    bool f = false;
    if (inst_cell(I, bmap, A, in, o, c)) f = diag_at<double, double>(o, c) > in.thr;
    const unsigned long long m = __ballot(f);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        mask[(long)blockIdx.x * kWavesPerBlock + w] = m;
        wc[w] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int q = 0; q < kWavesPerBlock; q++) t += wc[q];
        bcnt[blockIdx.x] = t;
    }
}

// In place exclusive scan of cnt[0..G), total in cnt[G]: tiles of
// kScanThreads counts scanned by one block each (tile totals to tot[]), the
// totals scanned by one block, then added back.
constexpr int kScanThreads = 1024;

__device__ __forceinline__ int block_excl_scan(int v, int *sh, int &total)
{
    // inclusive scan inside each wave by shuffles, then over the wave totals
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0;
        for (int q = 0; q < kScanThreads / 64; q++) { const int t = sh[q]; sh[q] = a; a += t; }
        sh[kScanThreads / 64] = a;
    }
    __syncthreads();
    total = sh[kScanThreads / 64];
    return sh[w] + x - v;
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_tiles(int *cnt, int G, int *tot)
{
    __shared__ int sh[kScanThreads / 64 + 1];
    const int q = blockIdx.x * kScanThreads + threadIdx.x;
    const int v = q < G ? cnt[q] : 0;
    int total;
    const int e = block_excl_scan(v, sh, total);
    if (q < G) cnt[q] = e;
    if (threadIdx.x == 0) tot[blockIdx.x] = total;
}

// one block: exclusive scan of the nt <= kScanThreads tile totals; the grand
// total to cnt[G]
__global__ void __launch_bounds__(kScanThreads)
k_scan_totals(int *tot, int nt, int *cnt, int G)
{
    __shared__ int sh[kScanThreads / 64 + 1];
    const int v = threadIdx.x < nt ? tot[threadIdx.x] : 0;
    int total;
    const int e = block_excl_scan(v, sh, total);
    if (threadIdx.x < nt) tot[threadIdx.x] = e;
    if (threadIdx.x == 0) cnt[G] = total;
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_add(int *cnt, int G, const int *tot)
{
    const int q = blockIdx.x * kScanThreads + threadIdx.x;
    if (q < G) cnt[q] += tot[blockIdx.x];
}

__global__ void __launch_bounds__(kBlock)
k_strong_emit(const DefInst *__restrict__ I, const int *__restrict__ bmap, const unsigned long long *__restrict__ mask,
              const int *__restrict__ boff, int *list, int cap)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long m = mask[(long)blockIdx.x * kWavesPerBlock + w];
    if (!((m >> lane) & 1ull)) return;
    int at = boff[blockIdx.x];
    for (int q = 0; q < w; q++) at += __popcll(mask[(long)blockIdx.x * kWavesPerBlock + q]);
    at += __popcll(m & ((1ull << lane) - 1ull));
    const DefInst in = I[bmap ? bmap[blockIdx.x] : 0];
    if (at < cap) list[at] = in.off + (blockIdx.x - in.blk0) * kBlock + threadIdx.x;
}

// Coefficients of listed cells (global index, instance cinst[q]): e and the
// couplings to -x,+x,-y,+y,-z,+z (0 where none).
__global__ void __launch_bounds__(kBlock)
k_defl_gather(const DefInst *__restrict__ I, const int *__restrict__ cinst, Stencil<double> A,
              const int *__restrict__ cells, int m, double *out)
{
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= m) return;
    const DefInst in = I[cinst ? cinst[q] : 0];
    const int s = cells[q];
    OpArgs<double> o;
    o.e = A.e; o.wx = A.wx; o.wy = A.wy; o.wz = A.wz; o.di = nullptr;
    o.nx = in.nx; o.ny = in.ny; o.nz = in.nz;
    const Cell c = cell_at(o, s - in.off);
    const int nx = in.nx, nxy = in.nx * in.ny;
    double *w = out + 7L * q;
    w[0] = o.e[s];
    w[1] = c.i > 0 ? o.wx[s - 1] : 0.0;
    w[2] = c.i < in.nx - 1 ? o.wx[s] : 0.0;
    w[3] = c.j > 0 ? o.wy[s - nx] : 0.0;
    w[4] = c.j < in.ny - 1 ? o.wy[s] : 0.0;
    w[5] = c.k > 0 ? o.wz[s - nxy] : 0.0;
    w[6] = c.k < in.nz - 1 ? o.wz[s] : 0.0;
}

// The per-iteration kernels take NR interleaved right-hand sides (vectors
// [cell][NR]; NR = 1 for a single solve); per-cluster values are [cluster][NR].

// Per piece of a cluster (one block each, at most kPiece cells): sums of r
// and v over its cells, fixed order.
constexpr int kPiece = 1024;

template <int NR>
__global__ void __launch_bounds__(kBlock)
k_defl_sums(const int *__restrict__ pb, const int *__restrict__ cells, const double *__restrict__ r,
            const double *__restrict__ v, double *psr, double *psv, const int *done)
{
    if (*done) return;
    const int p = blockIdx.x;
    double acc[2 * NR];
#pragma unroll
    for (int j = 0; j < 2 * NR; j++) acc[j] = 0.0;
    for (int q = pb[p] + threadIdx.x; q < pb[p + 1]; q += kBlock) {
        const long s = cells[q];
#pragma unroll
        for (int j = 0; j < NR; j++) {
            acc[j] += r[s * NR + j];
            if (v) acc[NR + j] += v[s * NR + j];
        }
    }
    block_sum<2 * NR>(acc);
    if (threadIdx.x == 0)
        for (int j = 0; j < NR; j++) { psr[p * NR + j] = acc[j]; psv[p * NR + j] = acc[NR + j]; }
}

// Per cluster sr, sv = sums over its pieces (fixed order); c = E^-1 sr; the
// extra dot partials <Zc, r> = c.sr and <Zc, v> = c.sv (Pextra[0..NR) and
// Pextra[NR..2NR)). E^-1 is symmetric: thread a reads column a, coalesced.
template <int NR>
__global__ void __launch_bounds__(kBlock)
k_defl_apply(int nc, const int *__restrict__ cp, const double *__restrict__ einv, const double *__restrict__ psr,
             const double *__restrict__ psv, double *c, double *Pextra, const int *done)
{
    if (*done) return;
    __shared__ double srs[kMaxClusters * NR], svs[kMaxClusters * NR];
    for (int a = threadIdx.x; a < nc; a += kBlock) {
        double tr[NR], tv[NR];
#pragma unroll
        for (int j = 0; j < NR; j++) { tr[j] = 0.0; tv[j] = 0.0; }
        for (int p = cp[a]; p < cp[a + 1]; p++)
#pragma unroll
            for (int j = 0; j < NR; j++) { tr[j] += psr[p * NR + j]; tv[j] += psv[p * NR + j]; }
#pragma unroll
        for (int j = 0; j < NR; j++) { srs[a * NR + j] = tr[j]; svs[a * NR + j] = tv[j]; }
    }
    __syncthreads();
    double acc[2 * NR];
#pragma unroll
    for (int j = 0; j < 2 * NR; j++) acc[j] = 0.0;
    for (int a = threadIdx.x; a < nc; a += kBlock) {
        double t[NR];
#pragma unroll
        for (int j = 0; j < NR; j++) t[j] = 0.0;
        for (int b = 0; b < nc; b++) {
            const double e = einv[(long)b * nc + a];
#pragma unroll
            for (int j = 0; j < NR; j++) t[j] += e * srs[b * NR + j];
        }
#pragma unroll
        for (int j = 0; j < NR; j++) {
            c[a * NR + j] = t[j];
            acc[j] += t[j] * srs[a * NR + j];
            acc[NR + j] += t[j] * svs[a * NR + j];
        }
    }
    block_sum<2 * NR>(acc);
    if (threadIdx.x == 0)
        for (int j = 0; j < 2 * NR; j++) Pextra[j] = acc[j];
}

// The batch form of k_defl_apply (one right-hand side per instance): c = E^-1
// sr over all clusters (E is block diagonal over the instances), and per
// instance q the extra dot partials over its clusters [irange[q],
// irange[q+1]): Pext[2q] = c.sr, Pext[2q+1] = c.sv (0 without clusters).
__global__ void __launch_bounds__(kBlock)
k_defl_apply_inst(int nc, const int *__restrict__ cp, const double *__restrict__ einv,
                  const double *__restrict__ psr, const double *__restrict__ psv, const int *__restrict__ irange,
                  const int *__restrict__ cinst, int ninst, double *c, double *Pext)
{
    __shared__ double srs[kMaxClusters], svs[kMaxClusters], cs[kMaxClusters];
    for (int a = threadIdx.x; a < nc; a += kBlock) {
        double tr = 0.0, tv = 0.0;
        for (int p = cp[a]; p < cp[a + 1]; p++) { tr += psr[p]; tv += psv[p]; }
        srs[a] = tr;
        svs[a] = tv;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < nc; a += kBlock) {   // E^-1 is block diagonal over the instances
        const int q = cinst[a];
        double t = 0.0;
        for (int b = irange[q]; b < irange[q + 1]; b++) t += einv[(long)b * nc + a] * srs[b];
        cs[a] = t;
        c[a] = t;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < ninst; q += kBlock) {
        double er = 0.0, ev = 0.0;
        for (int a = irange[q]; a < irange[q + 1]; a++) { er += cs[a] * srs[a]; ev += cs[a] * svs[a]; }
        Pext[2 * q] = er;
        Pext[2 * q + 1] = ev;
    }
}

// z += c[cluster] on the cells of every cluster.
template <int NR, typename T>
__global__ void __launch_bounds__(kBlock)
k_defl_scatter(const int *__restrict__ cells, const int *__restrict__ comp, int m, const double *__restrict__ c,
               T *z, const int *done)
{
    if (*done) return;
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= m) return;
    const long s = cells[q];
    const int cl = comp[q];
#pragma unroll
    for (int j = 0; j < NR; j++) z[s * NR + j] = (T)((double)z[s * NR + j] + c[cl * NR + j]);
}

// ---------------------------------------------------------------- host side

// pinned host staging of at least `bytes` (grown on demand, freed with the context)
static int stage_reserve(Deflation &D, size_t bytes)
{
    if (D.h_stage_bytes >= bytes && D.h_stage) return 0;
    if (D.h_stage) (void)hipHostFree(D.h_stage);
    D.h_stage = nullptr;
    D.h_stage_bytes = 0;
    if (hipHostMalloc(&D.h_stage, bytes, hipHostMallocDefault) != hipSuccess) return 1;
    D.h_stage_bytes = bytes;
    return 0;
}

static int find_root(std::vector<int> &p, int a)
{
    while (p[a] != a) { p[a] = p[p[a]]; a = p[a]; }
    return a;
}

// Setup shared by the single grid and the batch. hist: per instance the
// host copy of its histogram; inst: the instances (thr filled here); dinst:
// their device copy (nullptr: upload to slot 0 after the histogram, at
// hist_bytes); bmap/G: the per-cell kernels' block map and grid.
static int finish_clusters(hipStream_t st, std::vector<DefInst> &inst, DefInst *dinst, const int *bmap, int G,
                           const Stencil<double> &A, const unsigned *hist, double ratio, Deflation &D,
                           char *(*alloc)(void *, size_t, int), void *actx)
{
    const int ninst = (int)inst.size();
    // per instance: median bin, threshold ~ ratio x median, count certainly
    // above it and an upper bound on the count above it
    long above = 0, ub = 0;
    for (int q = 0; q < ninst; q++) {
        const unsigned *h = hist + (long)q * kHistBins;
        DefInst &in = inst[q];
        in.thr = INFINITY;
        long acc = 0;
        int mb = 0;
        for (; mb < kHistBins; mb++) {
            acc += h[mb];
            if (2 * acc >= in.n) break;
        }
        const double thr = ratio * ldexp(1.0, mb - 1100) * 1.4142135623730951;
        const int tb = ilogb(thr) + 1100;
        long aq = 0, uq = 0;
        for (int b = tb + 1; b < kHistBins; b++) aq += h[b];
        for (int b = tb; b < kHistBins; b++) uq += h[b];
        if (aq < kMinCluster) continue;
        in.thr = thr;
        above += aq;
        ub += uq;
    }
    if (above < kMinCluster) return HMFEM_OK;
    if (ub > (1L << 22)) return HMFEM_OK;   // not conduits: a large part of the domain

    // the strong cells, compacted on the device in index order (at most ub)
    // slot 1: masks (G * kWavesPerBlock words), block offsets (G + 1), list
    // (ub; right after the total boff[G], so that one copy fetches both), scan
    // tile totals (nt)
    const int nt = (G + kScanThreads - 1) / kScanThreads;
    if (nt > kScanThreads) return HMFEM_OK;   // beyond 2^28 cells
    const size_t mb_bytes = sizeof(unsigned long long) * (size_t)G * kWavesPerBlock;
    char *sbuf = alloc(actx, mb_bytes + sizeof(int) * ((size_t)G + 1 + nt + ub), 1);
    if (!sbuf) return HMFEM_ERR_NOMEM;
    unsigned long long *mask = (unsigned long long *)sbuf;
    int *boff = (int *)(sbuf + mb_bytes);
    int *slist = boff + G + 1;
    int *tot = slist + ub;
    HMFEM_CHECK(hipMemcpyAsync(dinst, inst.data(), sizeof(DefInst) * ninst, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_strong_mask, dim3(G), dim3(kBlock), 0, st, dinst, bmap, A, mask, boff);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanThreads), 0, st, boff, G, tot);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kScanThreads), 0, st, tot, nt, boff, G);
    hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(kScanThreads), 0, st, boff, G, (const int *)tot);
    hipLaunchKernelGGL(k_strong_emit, dim3(G), dim3(kBlock), 0, st, dinst, bmap, mask, boff, slist, (int)ub);
    // the total (boff[G]) and the list in one copy to pinned staging
    const size_t cb = sizeof(int) * (1 + ub);
    if (stage_reserve(D, cb + sizeof(double) * 7 * (size_t)ub)) return HMFEM_ERR_NOMEM;
    int *hl = (int *)D.h_stage;
    HMFEM_CHECK(hipMemcpyAsync(hl, boff + G, cb, hipMemcpyDeviceToHost, st));
    HMFEM_CHECK(hipStreamSynchronize(st));
    const int ms = std::min(hl[0], (int)ub);
    if (ms < kMinCluster) return HMFEM_OK;
    DeflationHost &W = D.h;
    auto &cell = W.cell, &cin = W.cin;
    cell.assign(hl + 1, hl + 1 + ms);   // in index order: a deterministic labelling
    cin.resize(ms);                     // instance of each strong cell
    for (int a = 0, q = 0; a < ms; a++) {
        while (q + 1 < ninst && cell[a] >= inst[q + 1].off) q++;
        cin[a] = q;
    }

    // strong neighbours of each strong cell, inside its instance: ordinals of
    // the +x, +y, +z neighbours by one monotone sweep per direction (cell[]
    // is sorted and a neighbour stays inside the instance), the -x, -y, -z
    // ones by symmetry; then a union-find over the +x, +y, +z links. Both stay
    // inside an instance, so ranges of whole instances (contiguous in the
    // sorted list) run on host threads of their own. The union keeps the
    // smaller ordinal as the root, so every root is its component's smallest
    // member whatever the order of the unions.
    auto &nbr = W.nbr;
    auto &up = W.up;
    auto &par = W.par;
    auto &root = W.root, &size = W.size;
    nbr.resize(6L * ms);   // [a][-x,+x,-y,+y,-z,+z]
    up.resize(ms);         // bit f: the +f neighbour is inside the instance
    par.resize(ms);
    root.resize(ms);
    size.resize(ms);
    auto label = [&](int a0, int a1) {
        std::fill(nbr.begin() + 6L * a0, nbr.begin() + 6L * a1, -1);
        for (int a = a0; a < a1; a++) {
            const DefInst &in = inst[cin[a]];
            const int l = cell[a] - in.off, i = l % in.nx, jk = l / in.nx, j = jk % in.ny, k = jk / in.ny;
            up[a] = (unsigned char)((i + 1 < in.nx) | ((j + 1 < in.ny) << 1) | ((k + 1 < in.nz) << 2));
        }
        for (int f = 0; f < 3; f++) {
            int p = a0;
            for (int a = a0; a < a1; a++) {
                if (!((up[a] >> f) & 1)) continue;
                const DefInst &in = inst[cin[a]];
                const long t = cell[a] + (f == 0 ? 1 : f == 1 ? (long)in.nx : (long)in.nx * in.ny);
                while (p < a1 && cell[p] < t) p++;
                if (p < a1 && cell[p] == t) {
                    nbr[6L * a + 2 * f + 1] = p;
                    nbr[6L * p + 2 * f] = a;
                }
            }
        }
        std::iota(par.begin() + a0, par.begin() + a1, a0);
        for (int a = a0; a < a1; a++)
            for (int f = 1; f < 6; f += 2) {
                const int b = nbr[6L * a + f];
                if (b < 0) continue;
                int ra = find_root(par, a), rb = find_root(par, b);
                if (ra != rb) par[std::max(ra, rb)] = std::min(ra, rb);
            }
        std::fill(size.begin() + a0, size.begin() + a1, 0);
        for (int a = a0; a < a1; a++) { root[a] = find_root(par, a); size[root[a]]++; }
    };
    {
        // ranges of whole instances with about ms / nt strong cells each
        const int nt = ms < 1024 || ninst < 2
                           ? 1
                           : (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
        std::vector<int> cut(1, 0);
        for (int a = 1; a < ms && (int)cut.size() < nt; a++)
            if (cin[a] != cin[a - 1] && (long)a * nt >= (long)ms * (long)cut.size()) cut.push_back(a);
        cut.push_back(ms);
        std::vector<std::thread> th;
        for (size_t t = 1; t + 1 < cut.size(); t++) {
            try {
                th.emplace_back(label, cut[t], cut[t + 1]);
            } catch (...) {   // no thread to be had: this range on the calling thread
                label(cut[t], cut[t + 1]);
            }
        }
        label(cut[0], cut[1]);
        for (auto &x : th) x.join();
    }
    auto &roots = W.roots;
    roots.clear();
    for (int a = 0; a < ms; a++)
        if (root[a] == a && size[a] >= kMinCluster) roots.push_back(a);
    std::stable_sort(roots.begin(), roots.end(), [&](int x, int y) { return size[x] > size[y]; });
    if ((int)roots.size() > kMaxClusters) roots.resize(kMaxClusters);
    std::sort(roots.begin(), roots.end());   // cluster order: by smallest member (so by instance)
    const int nc = (int)roots.size();
    if (nc == 0) return HMFEM_OK;
    auto &cid = W.cid;
    cid.assign(ms, -1);
    for (int q = 0; q < nc; q++) cid[roots[q]] = q;
    for (int a = 0; a < ms; a++) cid[a] = cid[root[a]];
    auto &ptr = W.ptr, &lcell = W.lcell, &lcomp = W.lcomp, &lord = W.lord, &linst = W.linst;
    ptr.assign(nc + 1, 0);
    for (int a = 0; a < ms; a++)
        if (cid[a] >= 0) ptr[cid[a] + 1]++;
    for (int q = 0; q < nc; q++) ptr[q + 1] += ptr[q];
    const int m = ptr[nc];
    lcell.resize(m);
    lcomp.resize(m);
    lord.resize(m);
    linst.resize(m);
    {
        auto &fill = W.fill;
        fill.assign(ptr.begin(), ptr.end() - 1);
        for (int a = 0; a < ms; a++)
            if (cid[a] >= 0) {
                const int q = fill[cid[a]]++;
                lcell[q] = cell[a];
                lcomp[q] = cid[a];
                lord[q] = a;
                linst[q] = cin[a];
            }
    }
    // clusters of each instance: [irange[q], irange[q+1])
    auto &irange = W.irange;
    irange.assign(ninst + 1, 0);
    for (int q = 0; q < nc; q++) irange[cin[roots[q]] + 1]++;
    for (int q = 0; q < ninst; q++) irange[q + 1] += irange[q];

    // pieces of at most kPiece cells per cluster, contiguous in the list
    auto &pbh = W.pbh, &cph = W.cph;
    pbh.assign(1, 0);
    cph.assign(1, 0);
    for (int q = 0; q < nc; q++) {
        for (int b = ptr[q]; b < ptr[q + 1]; b += kPiece) pbh.push_back(std::min(b + kPiece, ptr[q + 1]));
        cph.push_back((int)pbh.size() - 1);
    }
    const int np = (int)pbh.size() - 1;
    auto &pc = W.pc;
    pc.assign(pbh.begin(), pbh.end());
    pc.insert(pc.end(), cph.begin(), cph.end());
    pc.insert(pc.end(), irange.begin(), irange.end());
    for (int q = 0; q < nc; q++) pc.push_back(cin[roots[q]]);   // instance of each cluster

    // device copies, coefficients of the listed cells
    D.cells = (int *)alloc(actx, sizeof(int) * m, 2);
    D.comp = (int *)alloc(actx, sizeof(int) * 2 * (size_t)m, 3);   // cluster, then instance, of each cell
    D.pb = (int *)alloc(actx, sizeof(int) * pc.size(), 4);
    D.einv = (double *)alloc(actx, sizeof(double) * nc * nc, 5);
    const size_t per = (size_t)Deflation::kMaxRhs;
    D.c = (double *)alloc(actx, sizeof(double) * (per * (kMaxClusters + 2 * (size_t)np) + 2 * (size_t)ninst), 6);
    double *coef = (double *)alloc(actx, sizeof(double) * 7 * m, 7);
    if (!D.cells || !D.comp || !D.pb || !D.einv || !D.c || !coef) return HMFEM_ERR_NOMEM;
    D.cp = D.pb + np + 1;
    D.irange = D.cp + nc + 1;
    D.cinst = D.irange + ninst + 1;
    D.psr = D.c + per * kMaxClusters;
    D.psv = D.psr + per * np;
    D.pext = D.psv + per * np;
    int *dlinst = D.comp + m;
    // the lists go up through the pinned staging; the coefficient copy below
    // reuses it only after these copies (same stream)
    {
        const size_t lb = sizeof(int) * (3 * (size_t)m + pc.size()), cbytes = sizeof(double) * 7 * (size_t)m;
        if (stage_reserve(D, std::max(lb, cbytes))) return HMFEM_ERR_NOMEM;   // the strong list is copied out
        int *hs = (int *)D.h_stage;
        memcpy(hs, lcell.data(), sizeof(int) * m);
        memcpy(hs + m, lcomp.data(), sizeof(int) * m);
        memcpy(hs + 2 * (size_t)m, linst.data(), sizeof(int) * m);
        memcpy(hs + 3 * (size_t)m, pc.data(), sizeof(int) * pc.size());
        HMFEM_CHECK(hipMemcpyAsync(D.cells, hs, sizeof(int) * m, hipMemcpyHostToDevice, st));
        HMFEM_CHECK(hipMemcpyAsync(D.comp, hs + m, sizeof(int) * 2 * (size_t)m, hipMemcpyHostToDevice, st));
        HMFEM_CHECK(hipMemcpyAsync(D.pb, hs + 3 * (size_t)m, sizeof(int) * pc.size(), hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_defl_gather, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, st, dinst,
                       (const int *)dlinst, A, D.cells, m, coef);
    const double *w = (const double *)D.h_stage;   // 7 m doubles: reserved above
    HMFEM_CHECK(hipMemcpyAsync(D.h_stage, coef, sizeof(double) * 7 * m, hipMemcpyDeviceToHost, st));
    HMFEM_CHECK(hipStreamSynchronize(st));

    // E = Z^T A Z in a fixed order: E_aa = sum (e + couplings leaving the
    // cluster), E_ab = - couplings between clusters a and b (block diagonal
    // over the instances); its inverse via Cholesky E = L L^T (E is
    // symmetric positive definite), block by block. An instance's clusters
    // and their cells are contiguous in the lists, so ranges of whole
    // instances run on host threads of their own (each writes only its own
    // blocks of E and of the inverse).
    auto &E = W.E;
    auto &inv = W.inv;
    E.assign((size_t)nc * nc, 0.0);
    inv.assign((size_t)nc * nc, 0.0);
    std::atomic<int> not_pd{0};
    auto build_invert = [&](int q0, int q1) {
        std::vector<double> L, y, xb;
        for (int q = q0; q < q1; q++) {
            const int lo = irange[q], nb = irange[q + 1] - lo;
            if (nb == 0) continue;
            for (int c = ptr[lo]; c < ptr[lo + nb]; c++) {
                const int a = lcomp[c];
                const double *ws = &w[7L * c];
                const int *nbp = &nbr[6L * lord[c]];
                double daa = ws[0];
                for (int f = 0; f < 6; f++) {
                    const double wf = ws[1 + f];
                    if (wf == 0.0) continue;
                    const int b = nbp[f] < 0 ? -1 : cid[nbp[f]];
                    if (b == a) continue;
                    daa += wf;
                    if (b >= 0) E[(size_t)a * nc + b] -= wf;
                }
                E[(size_t)a * nc + a] += daa;
            }
            auto Eb = [&](int i, int j) { return E[(size_t)(lo + i) * nc + lo + j]; };
            L.resize((size_t)nb * nb);
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++) L[(size_t)i * nb + j] = Eb(i, j);
            for (int j = 0; j < nb; j++) {
                double djj = L[(size_t)j * nb + j];
                for (int k = 0; k < j; k++) djj -= L[(size_t)j * nb + k] * L[(size_t)j * nb + k];
                if (!(djj > 0)) {   // not positive definite: the correction stays off
                    not_pd = 1;
                    return;
                }
                djj = sqrt(djj);
                L[(size_t)j * nb + j] = djj;
                for (int i = j + 1; i < nb; i++) {
                    double t = L[(size_t)i * nb + j];
                    for (int k = 0; k < j; k++) t -= L[(size_t)i * nb + k] * L[(size_t)j * nb + k];
                    L[(size_t)i * nb + j] = t / djj;
                }
            }
            y.resize(nb);
            xb.resize((size_t)nb * nb);
            for (int col = 0; col < nb; col++) {   // solve L L^T x = e_col
                for (int i = 0; i < nb; i++) {
                    double t = i == col ? 1.0 : 0.0;
                    for (int k = 0; k < i; k++) t -= L[(size_t)i * nb + k] * y[k];
                    y[i] = t / L[(size_t)i * nb + i];
                }
                for (int i = nb - 1; i >= 0; i--) {
                    double t = y[i];
                    for (int k = i + 1; k < nb; k++) t -= L[(size_t)k * nb + i] * xb[(size_t)k * nb + col];
                    xb[(size_t)i * nb + col] = t / L[(size_t)i * nb + i];
                }
            }
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++) inv[(size_t)(lo + i) * nc + lo + j] = xb[(size_t)i * nb + j];
        }
    };
    {
        // ranges of whole instances with about m / nt listed cells each
        const int nt = m < 1024 || ninst < 2
                           ? 1
                           : (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
        std::vector<int> cut(1, 0);
        for (int q = 1; q < ninst && (int)cut.size() < nt; q++)
            if ((long)ptr[irange[q]] * nt >= (long)m * (long)cut.size() && ptr[irange[q]] > ptr[irange[cut.back()]])
                cut.push_back(q);
        cut.push_back(ninst);
        std::vector<std::thread> th;
        for (size_t t = 1; t + 1 < cut.size(); t++) {
            try {
                th.emplace_back(build_invert, cut[t], cut[t + 1]);
            } catch (...) {   // no thread to be had: this range on the calling thread
                build_invert(cut[t], cut[t + 1]);
            }
        }
        build_invert(cut[0], cut[1]);
        for (auto &x : th) x.join();
    }
    if (not_pd) return HMFEM_OK;
    if (stage_reserve(D, sizeof(double) * nc * nc)) return HMFEM_ERR_NOMEM;   // the coefficients are used up
    memcpy(D.h_stage, inv.data(), sizeof(double) * nc * nc);
    HMFEM_CHECK(hipMemcpyAsync(D.einv, D.h_stage, sizeof(double) * nc * nc, hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipStreamSynchronize(st));
    D.nc = nc;
    D.m = m;
    D.np = np;
    D.ninst = ninst;
    return HMFEM_OK;
}

int deflation_begin(hipStream_t st, Dims d, const Stencil<double> &A, Deflation &D,
                    char *(*alloc)(void *, size_t, int), void *actx)
{
    D.nc = 0;
    D.m = 0;
    D.ninst = 0;
    D.pending = false;
    const long n = d.n();
    if (n < 64 || !D.h_hist || !D.hist_ready) return HMFEM_OK;
    OpArgs<double> o;
    o.e = A.e; o.wx = A.wx; o.wy = A.wy; o.wz = A.wz; o.di = nullptr;
    o.nx = d.nx; o.ny = d.ny; o.nz = d.nz;
    // slot 0: the histogram, then the instance descriptor
    char *h0 = alloc(actx, kHistBins * sizeof(unsigned) + 256 + sizeof(DefInst), 0);
    if (!h0) return HMFEM_ERR_NOMEM;
    unsigned *hist = (unsigned *)h0;
    HMFEM_CHECK(hipMemsetAsync(hist, 0, kHistBins * sizeof(unsigned), st));
    // a grid-stride loop over at most 512 blocks keeps the global bin updates
    // few (2048 blocks: 35 us instead of 25 at 128^3)
    const int G = std::min(grid_for(n, kBlock), 512);
    hipLaunchKernelGGL(k_diag_hist, dim3(G), dim3(kBlock), 0, st, o, hist);
    HMFEM_CHECK(hipMemcpyAsync(D.h_hist, hist, kHistBins * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HMFEM_CHECK(hipEventRecord(D.hist_ready, st));
    D.pending = true;
    return HMFEM_OK;
}

int deflation_finish(hipStream_t st, Dims d, const Stencil<double> &A, double ratio, Deflation &D,
                     char *(*alloc)(void *, size_t, int), void *actx)
{
    if (!D.pending) return HMFEM_OK;
    D.pending = false;
    if (!(ratio > 1.0)) return HMFEM_OK;
    const long n = d.n();
    DefInst *dinst = (DefInst *)(alloc(actx, kHistBins * sizeof(unsigned) + 256 + sizeof(DefInst), 0) +
                                 kHistBins * sizeof(unsigned));
    std::vector<DefInst> inst(1, DefInst{d.nx, d.ny, d.nz, (int)n, 0, 0, INFINITY, 0});
    HMFEM_CHECK(hipEventSynchronize(D.hist_ready));
    return finish_clusters(st, inst, dinst, nullptr, (int)((n + kBlock - 1) / kBlock), A, D.h_hist, ratio, D,
                           alloc, actx);
}

int deflation_setup_batch(hipStream_t st, int ninst, const int *dims, const Stencil<double> &A, double ratio,
                          Deflation &D, char *(*alloc)(void *, size_t, int), void *actx)
{
    D.nc = 0;
    D.m = 0;
    D.ninst = 0;
    D.pending = false;
    if (ninst < 1 || !(ratio > 1.0)) return HMFEM_OK;
    std::vector<DefInst> inst(ninst);
    std::vector<int> bm, hm;
    long off = 0;
    for (int q = 0; q < ninst; q++) {
        DefInst &in = inst[q];
        in.nx = dims[3 * q]; in.ny = dims[3 * q + 1]; in.nz = dims[3 * q + 2];
        in.n = in.nx * in.ny * in.nz;
        in.off = (int)off;
        in.blk0 = (int)bm.size();
        in.hblk0 = (int)hm.size();
        in.thr = INFINITY;
        bm.insert(bm.end(), (in.n + kBlock - 1) / kBlock, q);
        hm.insert(hm.end(), (in.n + kBlock * kHistCells - 1) / (kBlock * kHistCells), q);
        off += in.n;
    }
    const int G = (int)bm.size(), GH = (int)hm.size();
    // slot 0: per-instance histograms, the block maps, the instance descriptors
    auto up256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t hb = up256(sizeof(unsigned) * kHistBins * (size_t)ninst), bb = up256(sizeof(int) * (size_t)(G + GH));
    char *h0 = alloc(actx, hb + bb + sizeof(DefInst) * ninst, 0);
    if (!h0) return HMFEM_ERR_NOMEM;
    unsigned *hist = (unsigned *)h0;
    int *bmap = (int *)(h0 + hb), *hmap = bmap + G;
    DefInst *dinst = (DefInst *)(h0 + hb + bb);
    bm.insert(bm.end(), hm.begin(), hm.end());
    HMFEM_CHECK(hipMemsetAsync(hist, 0, sizeof(unsigned) * kHistBins * (size_t)ninst, st));
    HMFEM_CHECK(hipMemcpyAsync(bmap, bm.data(), sizeof(int) * (G + GH), hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipMemcpyAsync(dinst, inst.data(), sizeof(DefInst) * ninst, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_diag_hist_inst, dim3(GH), dim3(kBlock), 0, st, (const DefInst *)dinst, (const int *)hmap, A,
                       hist);
    // the histograms go through the pinned staging, which finish_clusters
    // reads before it reuses (and may regrow) it
    if (stage_reserve(D, sizeof(unsigned) * kHistBins * (size_t)ninst)) return HMFEM_ERR_NOMEM;
    HMFEM_CHECK(hipMemcpyAsync(D.h_stage, hist, sizeof(unsigned) * kHistBins * (size_t)ninst, hipMemcpyDeviceToHost,
                               st));
    HMFEM_CHECK(hipStreamSynchronize(st));
    // finish_clusters reads the histograms before it reuses the staging
    return finish_clusters(st, inst, dinst, bmap, G, A, (const unsigned *)D.h_stage, ratio, D, alloc, actx);
}

template <int NR>
static void defl_sums(hipStream_t st, const Deflation &D, const double *r, const double *v, double *Pextra,
                      const int *done)
{
    hipLaunchKernelGGL(k_defl_sums<NR>, dim3(D.np), dim3(kBlock), 0, st, D.pb, D.cells, r, v, D.psr, D.psv, done);
    hipLaunchKernelGGL(k_defl_apply<NR>, dim3(1), dim3(kBlock), 0, st, D.nc, D.cp, D.einv, D.psr, D.psv, D.c,
                       Pextra, done);
}

template <int NR, typename T>
static void defl_scatter(hipStream_t st, const Deflation &D, T *z, const int *done)
{
    hipLaunchKernelGGL((k_defl_scatter<NR, T>), dim3((D.m + kBlock - 1) / kBlock), dim3(kBlock), 0, st, D.cells,
                       D.comp, D.m, D.c, z, done);
}

void deflation_sums(hipStream_t st, const Deflation &D, const double *r, const double *v, double *Pextra,
                    const int *done)
{
    defl_sums<1>(st, D, r, v, Pextra, done);
}

void deflation_scatter(hipStream_t st, const Deflation &D, float *z, const int *done)
{
    defl_scatter<1, float>(st, D, z, done);
}

void deflation_scatter(hipStream_t st, const Deflation &D, double *z, const int *done)
{
    defl_scatter<1, double>(st, D, z, done);
}

void deflation_sums_batch(hipStream_t st, const Deflation &D, const double *r, const double *v, const int *never)
{
    hipLaunchKernelGGL(k_defl_sums<1>, dim3(D.np), dim3(kBlock), 0, st, D.pb, D.cells, r, v, D.psr, D.psv, never);
    hipLaunchKernelGGL(k_defl_apply_inst, dim3(1), dim3(kBlock), 0, st, D.nc, D.cp, D.einv, D.psr, D.psv, D.irange,
                       D.cinst, D.ninst, D.c, D.pext);
}

int deflation_sums_block(hipStream_t st, const Deflation &D, int nr, const double *r, const double *v,
                         double *Pextra, const int *done)
{
    switch (nr) {
    case 2: defl_sums<2>(st, D, r, v, Pextra, done); return HMFEM_OK;
    case 4: defl_sums<4>(st, D, r, v, Pextra, done); return HMFEM_OK;
    default: return HMFEM_ERR_ARG;
    }
}

int deflation_scatter_block(hipStream_t st, const Deflation &D, int nr, float *z, const int *done)
{
    switch (nr) {
    case 2: defl_scatter<2, float>(st, D, z, done); return HMFEM_OK;
    case 4: defl_scatter<4, float>(st, D, z, done); return HMFEM_OK;
    default: return HMFEM_ERR_ARG;
    }
}

} // namespace hmfem
