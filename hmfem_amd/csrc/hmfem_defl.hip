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
//     z = V(r) + Z E^-1 Z^T r,   Z = [1_c], E = Z^T A Z  (FP64),
// which keeps the preconditioner symmetric positive definite. The clusters
// are connected components of the strong cells, so no coupling joins two of
// them and E is diagonal. No counterpart in the reference (its AMG has no
// such stage, amg.h:1135-1239).
//
// Setup: a histogram of log2(d) gives the median (on the host, overlapped
// with the hierarchy's setup); on the device, the strong cells are compacted
// in index order and labelled by a lock-free union-find (deterministic: each
// root is its component's smallest member); the host picks the 256 largest
// components from their (root, size) table and lays out the per-cluster
// arrays; the device places every cell in its cluster's list in index order
// and sums E's diagonal in a fixed order. Per CG iteration: per-piece sums of
// r and v (fixed order), c = E^-1 (Z^T r) with the dot products c.(Z^T r),
// c.(Z^T v) written as one more partial of the fine-level dots, and
// z += Z c after the V-cycle.
#include "hmfem_stencil.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <vector>

namespace hmfem {

constexpr int kHistBins = Deflation::kHistBins;
constexpr int kMinCluster = 8;
constexpr int kMaxClusters = 256;

// The histograms also keep each cell's bin (2 bytes a cell), from which the
// strong-cell mask decides every cell outside the threshold's own bin without
// reading the stencil again.
__global__ void __launch_bounds__(kBlock)
k_diag_hist(OpArgs<double> o, unsigned *hist, unsigned short *bins)
{
    __shared__ unsigned h[kHistBins];
    for (int q = threadIdx.x; q < kHistBins; q += kBlock) h[q] = 0;
    __syncthreads();
    const int n = o.nx * o.ny * o.nz;
    int run_bin = -1;   // runs of equal bins counted in a register, as in k_diag_hist_inst
    unsigned run = 0;
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
        Cell c = cell_at(o, s);
        const int b = diag_bin(diag_at<double, double>(o, c), kHistBins);
        bins[s] = (unsigned short)b;
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
    int tb;       // diagonal bin of thr (kHistBins: none)
};

constexpr int kHistCells = 16;   // cells per thread of the batch histogram

__device__ __forceinline__ OpArgs<double> inst_op(const DefInst &in, const Stencil<double> &A)
{
    OpArgs<double> o;
    o.e = A.e + in.off; o.wx = A.wx + in.off; o.wy = A.wy + in.off; o.wz = A.wz + in.off; o.di = nullptr;
    o.nx = in.nx; o.ny = in.ny; o.nz = in.nz;
    return o;
}

// Per-instance histogram of floor(log2 d); block b covers kBlock * kHistCells
// cells of instance hmap[b].
__global__ void __launch_bounds__(kBlock)
k_diag_hist_inst(const DefInst *__restrict__ I, const int *__restrict__ hmap, Stencil<double> A, unsigned *hist,
                 unsigned short *bins)
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
        const int b = diag_bin(diag_at<double, double>(o, cell_at(o, l)), kHistBins);
        bins[in.off + l] = (unsigned short)b;
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
// of the counts, and the ordered list of global indices. A cell whose bin
// lies above the threshold's bin tb is strong (d >= 2^(tb + 1 - 1100) >
// thr), one below it is not (d < 2^(tb - 1100) <= thr); only cells in bin tb
// (which also holds the clamped ends of the table when tb is an end) compare
// their diagonal with thr. The mask and the list are kept per group of
// kBlock cells (group g: cells of the per-cell block map's block g); a
// workgroup of these two kernels takes kMaskSub groups, as one group is too
// little work for a workgroup's launch.
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int kMaskSub = 8;

__global__ void __launch_bounds__(kBlock)
k_strong_mask(const DefInst *__restrict__ I, const int *__restrict__ bmap, int G, Stencil<double> A,
              const unsigned short *__restrict__ bins, unsigned long long *mask, int *bcnt)
{
    __shared__ int wc[kMaskSub][kWavesPerBlock];
    const int w = threadIdx.x >> 6, g0 = blockIdx.x * kMaskSub;
    // the groups' loads first (instance, then bins), so that a thread has
    // kMaskSub of each in flight instead of a chain of three per group; a
    // group past the last (g >= G) reads group G - 1's and stores nothing
    int q[kMaskSub], l[kMaskSub], b[kMaskSub];
#pragma unroll
    for (int t = 0; t < kMaskSub; t++) q[t] = bmap ? bmap[min(g0 + t, G - 1)] : 0;
#pragma unroll
    for (int t = 0; t < kMaskSub; t++) {
        const int g = min(g0 + t, G - 1);
        l[t] = (g - I[q[t]].blk0) * kBlock + threadIdx.x;
        b[t] = l[t] < I[q[t]].n ? bins[I[q[t]].off + l[t]] : -1;
    }
#pragma unroll
    for (int t = 0; t < kMaskSub; t++) {
        const int g = g0 + t;
        if (g >= G) break;
        const int tb = I[q[t]].tb;
        bool f = b[t] > tb;
        if (b[t] == tb) {
            const DefInst in = I[q[t]];
            const OpArgs<double> o = inst_op(in, A);
            f = diag_at<double, double>(o, cell_at(o, l[t])) > in.thr;
        }
        const unsigned long long m = __ballot(f);
        if ((threadIdx.x & 63) == 0) {
            mask[(long)g * kWavesPerBlock + w] = m;
            wc[t][w] = __popcll(m);
        }
    }
    __syncthreads();
    if (threadIdx.x < kMaskSub && g0 + (int)threadIdx.x < G) {
        int t = 0;
        for (int q = 0; q < kWavesPerBlock; q++) t += wc[threadIdx.x][q];
        bcnt[g0 + threadIdx.x] = t;
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
k_strong_emit(const DefInst *__restrict__ I, const int *__restrict__ bmap, int G,
              const unsigned long long *__restrict__ mask, const int *__restrict__ boff, int *list, int cap)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, g0 = blockIdx.x * kMaskSub;
    unsigned long long mm[kMaskSub];   // the groups' masks first (most groups have no strong cell)
#pragma unroll
    for (int t = 0; t < kMaskSub; t++) mm[t] = mask[(long)min(g0 + t, G - 1) * kWavesPerBlock + w];
#pragma unroll
    for (int t = 0; t < kMaskSub; t++) {
        const int g = g0 + t;
        if (g >= G) break;
        const unsigned long long m = mm[t];
        if (!((m >> lane) & 1ull)) continue;
        int at = boff[g];
        for (int q = 0; q < w; q++) at += __popcll(mask[(long)g * kWavesPerBlock + q]);
        at += __popcll(m & ((1ull << lane) - 1ull));
        const DefInst in = I[bmap ? bmap[g] : 0];
        if (at < cap) list[at] = in.off + (g - in.blk0) * kBlock + threadIdx.x;
    }
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

// Per cluster sr, sv = sums over its pieces (fixed order); c = E^-1 sr (E is
// diagonal: einv[a] = 1 / E_aa); the extra dot partials <Zc, r> = c.sr and
// <Zc, v> = c.sv (Pextra[0..NR) and Pextra[NR..2NR)).
template <int NR>
__global__ void __launch_bounds__(kBlock)
k_defl_apply(int nc, const int *__restrict__ cp, const double *__restrict__ einv, const double *__restrict__ psr,
             const double *__restrict__ psv, double *c, double *Pextra, const int *done)
{
    if (*done) return;
    double acc[2 * NR];
#pragma unroll
    for (int j = 0; j < 2 * NR; j++) acc[j] = 0.0;
    for (int a = threadIdx.x; a < nc; a += kBlock) {
        double tr[NR], tv[NR];
#pragma unroll
        for (int j = 0; j < NR; j++) { tr[j] = 0.0; tv[j] = 0.0; }
        for (int p = cp[a]; p < cp[a + 1]; p++)
#pragma unroll
            for (int j = 0; j < NR; j++) { tr[j] += psr[p * NR + j]; tv[j] += psv[p * NR + j]; }
        const double ei = einv[a];
#pragma unroll
        for (int j = 0; j < NR; j++) {
            const double t = ei * tr[j];
            c[a * NR + j] = t;
            acc[j] += t * tr[j];
            acc[NR + j] += t * tv[j];
        }
    }
    block_sum<2 * NR>(acc);
    if (threadIdx.x == 0)
        for (int j = 0; j < 2 * NR; j++) Pextra[j] = acc[j];
}

// The batch form of k_defl_apply (one right-hand side per instance): c =
// E^-1 sr over all clusters, and per instance q the extra dot partials over
// its clusters [irange[q], irange[q+1]): Pext[2q] = c.sr, Pext[2q+1] = c.sv
// (0 without clusters).
__global__ void __launch_bounds__(kBlock)
k_defl_apply_inst(int nc, const int *__restrict__ cp, const double *__restrict__ einv,
                  const double *__restrict__ psr, const double *__restrict__ psv, const int *__restrict__ irange,
                  int ninst, double *c, double *Pext)
{
    __shared__ double srs[kMaxClusters], svs[kMaxClusters], cs[kMaxClusters];
    for (int a = threadIdx.x; a < nc; a += kBlock) {
        double tr = 0.0, tv = 0.0;
        for (int p = cp[a]; p < cp[a + 1]; p++) { tr += psr[p]; tv += psv[p]; }
        srs[a] = tr;
        svs[a] = tv;
        cs[a] = einv[a] * tr;
        c[a] = cs[a];
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

// ------------------------------------------------- clusters on the device
//
// The strong cells (sorted list, index order) are labelled by connected
// components on the device: a lock-free union-find in which the larger of
// two roots is hooked under the smaller by compare-and-swap (as in the
// ECL-CC method of Jaiganesh and Burtscher, 2018), so every root ends as its
// component's smallest member whatever the order of the unions: the labels
// are deterministic. A failed swap means another thread's succeeded, so the
// retry loop always ends. Parent pointers only ever point to smaller
// ordinals (no cycles); path halving writes an ancestor, which a concurrent
// writer may replace by another ancestor.

__device__ __forceinline__ int par_load(const int *p, int a)
{
    return __hip_atomic_load(p + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void par_store(int *p, int a, int v)
{
    __hip_atomic_store(p + a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ int cc_find(int *par, int a)
{
    while (true) {
        const int p = par_load(par, a);
        if (p == a) return a;
        const int g = par_load(par, p);
        if (g == p) return p;
        par_store(par, a, g);   // path halving
        a = g;
    }
}

__device__ void cc_unite(int *par, int a, int b)
{
    int ra = cc_find(par, a), rb = cc_find(par, b);
    while (ra != rb) {
        if (ra > rb) { const int t = ra; ra = rb; rb = t; }
        int seen = rb;
        if (__hip_atomic_compare_exchange_strong(par + rb, &seen, ra, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
        ra = cc_find(par, ra);
        rb = cc_find(par, seen);
    }
}

// index of key in the sorted a[lo, hi), or -1
__device__ __forceinline__ int sorted_find(const int *__restrict__ a, int lo, int hi, int key)
{
    int h = hi;
    while (lo < h) {
        const int mid = (lo + h) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else h = mid;
    }
    return lo < hi && a[lo] == key ? lo : -1;
}

// Work arrays of the device labelling (ub: bound on the strong cells).
struct CcWork {
    int *par, *root, *cin, *size, *cid, *lord, *nbr;   // [ub] each, nbr [6 ub]: -x,+x,-y,+y,-z,+z
    int *ncand, *cand;                                  // candidates (root, size, instance) x ccap
    int *ccount;                                        // per chunk of kChunk strong cells, per cluster
    int *sel;                                           // chosen roots (ascending), then the cluster offsets
};
constexpr int kChunk = 1024;
constexpr int kJumpRounds = 8;   // chains of up to 256 links flattened (longer ones are still walked)

// Per strong cell a < ms: its instance, its six strong neighbours inside the
// instance (ordinals, -1: none), parent = itself, size 0.
__global__ void __launch_bounds__(kBlock)
k_cc_links(const DefInst *__restrict__ I, int ninst, const int *__restrict__ list, const int *__restrict__ pms,
           int cap, CcWork w)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *w.ncand = 0;
    const int ms = min(*pms, cap);
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= ms) return;
    const int s = list[a];
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (I[mid].off <= s) lo = mid;
        else hi = mid - 1;
    }
    const DefInst in = I[lo];
    const int l = s - in.off, i = l % in.nx, jk = l / in.nx, j = jk % in.ny, k = jk / in.ny;
    const int sy = in.nx, sz = in.nx * in.ny;
    // a neighbour at +-step lies within step places of a in the sorted list
    int *nb = w.nbr + 6L * a;
    nb[0] = i > 0 ? sorted_find(list, max(0, a - 1), a, s - 1) : -1;
    nb[1] = i + 1 < in.nx ? sorted_find(list, a + 1, min(ms, a + 2), s + 1) : -1;
    nb[2] = j > 0 ? sorted_find(list, max(0, a - sy), a, s - sy) : -1;
    nb[3] = j + 1 < in.ny ? sorted_find(list, a + 1, min(ms, a + 1 + sy), s + sy) : -1;
    nb[4] = k > 0 ? sorted_find(list, max(0, a - sz), a, s - sz) : -1;
    nb[5] = k + 1 < in.nz ? sorted_find(list, a + 1, (int)min((long)ms, (long)a + 1 + sz), s + sz) : -1;
    w.cin[a] = lo;
    w.par[a] = a;
    w.size[a] = 0;
}

__global__ void __launch_bounds__(kBlock) k_cc_hook(const int *__restrict__ pms, int cap, CcWork w)
{
    const int ms = min(*pms, cap);
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= ms) return;
    const int *nb = w.nbr + 6L * a;
    for (int f = 1; f < 6; f += 2)
        if (nb[f] >= 0) cc_unite(w.par, a, nb[f]);
}

// Pointer jumping after the unions: par[a] = par[par[a]]. Unions of cells
// along a conduit leave chains (each root hooked under the next smaller
// one), and walking a chain is one dependent load per link (140 us on the
// 256^3 conduit field); every round halves the chains. Reads may see another
// thread's update of this round or not: either is an ancestor, and parents
// only ever point to smaller ordinals, so the roots stay what they are.
__global__ void __launch_bounds__(kBlock) k_cc_jump(const int *__restrict__ pms, int cap, CcWork w)
{
    const int ms = min(*pms, cap);
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= ms) return;
    int *par = w.par;
    const int p = par[a];
    const int g = par[p];
    if (g != p) par[a] = g;
}

// roots and component sizes (after every union has finished: the parent
// pointers no longer change, so plain loads through the caches walk them).
// The cells of one conduit are mostly neighbours in the list, so the lanes
// of a wave with the same root add their count with one atomic (one atomic
// per cell on a handful of sizes serialised: 140 us on the 256^3 conduit
// field).
__global__ void __launch_bounds__(kBlock) k_cc_roots(const int *__restrict__ pms, int cap, CcWork w)
{
    const int ms = min(*pms, cap);
    const int a = blockIdx.x * kBlock + threadIdx.x;
    int r = -1;
    if (a < ms) {
        const int *__restrict__ par = w.par;
        int p = par[a];
        r = a;
        while (p != r) {
            r = p;
            p = par[r];
        }
        w.root[a] = r;
    }
    const int lane = threadIdx.x & 63;
    unsigned long long left = __ballot(r >= 0);
    while (left) {   // one distinct root of the wave per round (wave-uniform loop)
        const int lead = __ffsll((long long)left) - 1;
        const int lr = __shfl(r, lead);
        const unsigned long long same = __ballot(r == lr) & left;
        if (lane == lead) atomicAdd(w.size + lr, __popcll(same));
        left &= ~same;
    }
}

// components of at least kMinCluster cells: (root, size, instance), in any
// order (the host sorts them)
__global__ void __launch_bounds__(kBlock) k_cc_cands(const int *__restrict__ pms, int cap, int ccap, CcWork w)
{
    const int ms = min(*pms, cap);
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= ms || w.root[a] != a || w.size[a] < kMinCluster) return;
    const int t = atomicAdd(w.ncand, 1);
    if (t < ccap) {
        w.cand[3 * t] = a;
        w.cand[3 * t + 1] = w.size[a];
        w.cand[3 * t + 2] = w.cin[a];
    }
}

// cluster of each strong cell (the rank of its root among the chosen roots, or -1)
__global__ void __launch_bounds__(kBlock) k_cc_ordinal(const int *__restrict__ pms, int cap, int nc, CcWork w)
{
    const int ms = min(*pms, cap);
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= ms) return;
    w.cid[a] = sorted_find(w.sel, 0, nc, w.root[a]);
}

// Placing the cells of every cluster, in index order, at their offset in
// the cluster-grouped list: per chunk of kChunk strong cells the count of
// each cluster, the counts turned into offsets chunk after chunk, then every
// cell at its cluster's offset for its chunk plus its rank among the chunk's
// cells of the same cluster (waves by ballots, then the waves before it).
__global__ void __launch_bounds__(kChunk) k_place_count(const int *__restrict__ pms, int cap, int nc, CcWork w)
{
    __shared__ int cnt[kMaxClusters];
    for (int c = threadIdx.x; c < nc; c += kChunk) cnt[c] = 0;
    __syncthreads();
    const int ms = min(*pms, cap);
    const int a = blockIdx.x * kChunk + threadIdx.x;
    const int c = a < ms ? w.cid[a] : -1;
    if (c >= 0) atomicAdd(cnt + c, 1);
    __syncthreads();
    for (int q = threadIdx.x; q < nc; q += kChunk) w.ccount[(long)blockIdx.x * nc + q] = cnt[q];
}

__global__ void __launch_bounds__(kMaxClusters) k_place_scan(int nc, int nchunks, CcWork w)
{
    const int c = threadIdx.x;
    if (c >= nc) return;
    int run = w.sel[kMaxClusters + c];   // the cluster's offset in the list
    // eight chunks' counts loaded before their offsets are stored (one load
    // latency per eight chunks instead of per chunk)
    int b = 0;
    for (; b + 8 <= nchunks; b += 8) {
        int v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = w.ccount[(long)(b + u) * nc + c];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            w.ccount[(long)(b + u) * nc + c] = run;
            run += v[u];
        }
    }
    for (; b < nchunks; b++) {
        const long q = (long)b * nc + c;
        const int v = w.ccount[q];
        w.ccount[q] = run;
        run += v;
    }
}

__global__ void __launch_bounds__(kChunk)
k_place_emit(const int *__restrict__ list, const int *__restrict__ pms, int cap, int nc, int m, CcWork w, int *cells,
             int *comp)
{
    constexpr int kW = kChunk / 64;
    __shared__ int wc[kW][kMaxClusters];
    for (int q = threadIdx.x; q < kW * kMaxClusters; q += kChunk) wc[q / kMaxClusters][q % kMaxClusters] = 0;
    __syncthreads();
    const int ms = min(*pms, cap);
    const int a = blockIdx.x * kChunk + threadIdx.x;
    const int c = a < ms ? w.cid[a] : -1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long left = __ballot(c >= 0), mine = 0;
    while (left) {   // one distinct cluster of the wave per round (wave-uniform loop)
        const int lead = __ffsll((long long)left) - 1;
        const int lc = __shfl(c, lead);
        const unsigned long long same = __ballot(c == lc) & left;
        if (c == lc) mine = same;
        if (lane == lead) wc[wv][lc] = __popcll(same);
        left &= ~same;
    }
    __syncthreads();
    if (c < 0) return;
    int rank = __popcll(mine & ((1ull << lane) - 1ull));
    for (int q = 0; q < wv; q++) rank += wc[q][c];
    const int pos = w.ccount[(long)blockIdx.x * nc + c] + rank;
    if (pos < m) {
        cells[pos] = list[a];
        comp[pos] = c;
        comp[m + pos] = w.cin[a];
        w.lord[pos] = a;
    }
}

// E = Z^T A Z is diagonal: a strong neighbour of a strong cell is in its
// component, so no coupling joins two clusters. Per listed cell its share of
// E_aa: e plus the couplings to cells outside its cluster (not strong).
__global__ void __launch_bounds__(kBlock)
k_defl_eshare(const DefInst *__restrict__ I, Stencil<double> A, const int *__restrict__ cells,
              const int *__restrict__ comp, int m, CcWork w, double *share)
{
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= m) return;
    const DefInst in = I[comp[m + q]];
    const int s = cells[q];
    const int l = s - in.off, i = l % in.nx, jk = l / in.nx, j = jk % in.ny, k = jk / in.ny;
    const int nx = in.nx, nxy = in.nx * in.ny;
    const int *nb = w.nbr + 6L * w.lord[q];
    double t = A.e[s];
    const double wf[6] = {i > 0 ? A.wx[s - 1] : 0.0, i < in.nx - 1 ? A.wx[s] : 0.0,
                          j > 0 ? A.wy[s - nx] : 0.0, j < in.ny - 1 ? A.wy[s] : 0.0,
                          k > 0 ? A.wz[s - nxy] : 0.0, k < in.nz - 1 ? A.wz[s] : 0.0};
    for (int f = 0; f < 6; f++)
        if (nb[f] < 0) t += wf[f];
    share[q] = t;
}

// per piece: the sum of its cells' shares (fixed order)
__global__ void __launch_bounds__(kBlock) k_defl_epiece(const int *__restrict__ pb, const double *__restrict__ share,
                                                          double *pe)
{
    const int p = blockIdx.x;
    double acc[1] = {0.0};
    for (int q = pb[p] + threadIdx.x; q < pb[p + 1]; q += kBlock) acc[0] += share[q];
    block_sum<1>(acc);
    if (threadIdx.x == 0) pe[p] = acc[0];
}

// E_aa from its pieces, and 1 / E_aa (0 for a cluster with E_aa <= 0,
// whose correction then stays off)
__global__ void __launch_bounds__(kBlock)
k_defl_einv(int nc, const int *__restrict__ cp, const double *__restrict__ pe, double *einv)
{
    for (int a = threadIdx.x; a < nc; a += kBlock) {
        double e = 0.0;
        for (int p = cp[a]; p < cp[a + 1]; p++) e += pe[p];
        einv[a] = e > 0.0 ? 1.0 / e : 0.0;
    }
}

// ------------------------------------------------- the balancing form
//
// z = Q r + (I - Q A) V (I - A Q) r with Q = Z E^-1 Z^T: the V-cycle sees the
// residual with the clusters' constant modes taken out, and its output is
// projected so that those modes are exactly Q's. With c1 = E^-1 Z^T r:
//   r' = r - A Z c1,  y = V(r'),  c2 = E^-1 Z^T A y,  z = y + Z (c1 - c2).
// A Z c1 is nonzero on the cluster cells (alpha_s c1_a, alpha_s = (A 1_a)_s)
// and on the cells just outside them (-sum of couplings x c1): the records,
// built once per setup. Z^T A y sums, per cluster, alpha_s y_s minus the
// couplings to outside cells times their y (fixed order, per piece).

__device__ __forceinline__ int face_step(int f, int nx, int nxy)
{
    const int m = f >> 1, sgn = (f & 1) ? 1 : -1;
    return sgn * (m == 0 ? 1 : m == 1 ? nx : nxy);
}

__device__ __forceinline__ bool face_in(int f, int i, int j, int k, const DefInst &in)
{
    switch (f) {
    case 0: return i > 0;
    case 1: return i + 1 < in.nx;
    case 2: return j > 0;
    case 3: return j + 1 < in.ny;
    case 4: return k > 0;
    default: return k + 1 < in.nz;
    }
}

// coupling of global cell s to its neighbour across face f (the face exists)
__device__ __forceinline__ double face_w(const Stencil<double> &A, long s, int f, int nx, int nxy)
{
    switch (f) {
    case 0: return A.wx[s - 1];
    case 1: return A.wx[s];
    case 2: return A.wy[s - nx];
    case 3: return A.wy[s];
    case 4: return A.wz[s - nxy];
    default: return A.wz[s];
    }
}

// One thread per (listed cell, face) whose neighbour t is not strong: t is
// recorded by the first of its own faces (in face order) that leads to a
// cluster cell, so every outside cell gets exactly one record (the records'
// order does not matter: each is used on its own).
__global__ void __launch_bounds__(kBlock)
k_bnn_records(const DefInst *__restrict__ I, Stencil<double> A, const int *__restrict__ cells,
              const int *__restrict__ comp, int m, const int *__restrict__ list, const int *__restrict__ pms, int cap,
              CcWork w, int *rec_cell, int *rec_cl, double *rec_w, int *rec_count, int rec_max)
{
    const long idx = (long)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= 6L * m) return;
    const int q = (int)(idx / 6), f = (int)(idx % 6);
    const int ms = min(*pms, cap);
    const DefInst in = I[comp[m + q]];
    const int s = cells[q];
    const int nx = in.nx, nxy = in.nx * in.ny;
    int l = s - in.off, i = l % nx, jk = l / nx, j = jk % in.ny, k = jk / in.ny;
    if (!face_in(f, i, j, k, in) || w.nbr[6L * w.lord[q] + f] >= 0) return;
    const int t = s + face_step(f, nx, nxy);
    l = t - in.off; i = l % nx; jk = l / nx; j = jk % in.ny; k = jk / in.ny;
    auto cluster_of = [&](int u) {
        const int b = sorted_find(list, 0, ms, u);
        return b >= 0 ? w.cid[b] : -1;
    };
    const int own = f ^ 1;   // the face of t that leads back to s
    for (int g = 0; g < own; g++)
        if (face_in(g, i, j, k, in) && cluster_of(t + face_step(g, nx, nxy)) >= 0) return;
    const int rk = atomicAdd(rec_count, 1);
    if (rk >= rec_max) return;
    rec_cell[rk] = t;
    for (int g = 0; g < 6; g++) {
        int cl = -1;
        double wg = 0.0;
        if (face_in(g, i, j, k, in)) {
            cl = cluster_of(t + face_step(g, nx, nxy));
            if (cl >= 0) wg = face_w(A, t, g, nx, nxy);
        }
        rec_cl[6L * rk + g] = cl;
        rec_w[6L * rk + g] = wg;
    }
}

// r' = r - A Z c1 in place (the overwritten values saved): listed cells
// first, then the records.
__global__ void __launch_bounds__(kBlock)
k_bnn_pre(int m, const int *__restrict__ cells, const int *__restrict__ comp, const double *__restrict__ alpha,
          const double *__restrict__ c, const int *__restrict__ rec_cell, const int *__restrict__ rec_cl,
          const double *__restrict__ rec_w, const int *__restrict__ rec_count, int rec_max, double *r,
          double *rsave, const int *done)
{
    if (*done) return;
    const long idx = (long)blockIdx.x * kBlock + threadIdx.x;
    if (idx < m) {
        const int q = (int)idx;
        const long s = cells[q];
        const double rv = r[s];
        rsave[q] = rv;
        r[s] = rv - alpha[q] * c[comp[q]];
        return;
    }
    const long k = idx - m;
    if (k >= min(*rec_count, rec_max)) return;
    const long t = rec_cell[k];
    const double rv = r[t];
    rsave[m + k] = rv;
    double a = rv;
    for (int g = 0; g < 6; g++) {
        const int cl = rec_cl[6 * k + g];
        if (cl >= 0) a += rec_w[6 * k + g] * c[cl];
    }
    r[t] = a;
}

// r restored; blocks [0, np): per piece the sum of alpha_s y_s minus the
// couplings of its cells to outside cells times their y (Z^T A y, fixed
// order); the blocks after them restore the records.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_bnn_post(const DefInst *__restrict__ I, Stencil<double> A, int np, const int *__restrict__ pb,
           const int *__restrict__ cells, const int *__restrict__ comp, int m, const double *__restrict__ alpha,
           const int *__restrict__ lord, const int *__restrict__ nbr, const T *__restrict__ y,
           const int *__restrict__ rec_cell, const int *__restrict__ rec_count, int rec_max,
           const double *__restrict__ rsave, double *r, double *pe2, const int *done)
{
    if (*done) return;
    if ((int)blockIdx.x >= np) {
        const long k = (long)(blockIdx.x - np) * kBlock + threadIdx.x;
        if (k < min(*rec_count, rec_max)) r[rec_cell[k]] = rsave[m + k];
        return;
    }
    const int p = blockIdx.x;
    double acc[1] = {0.0};
    for (int q = pb[p] + threadIdx.x; q < pb[p + 1]; q += kBlock) {
        const long s = cells[q];
        r[s] = rsave[q];
        const DefInst in = I[comp[m + q]];
        const int nx = in.nx, nxy = in.nx * in.ny;
        const int l = (int)s - in.off, i = l % nx, jk = l / nx, j = jk % in.ny, k = jk / in.ny;
        double h = alpha[q] * (double)y[s];
        const int *nb = nbr + 6L * lord[q];
        for (int f = 0; f < 6; f++)
            if (nb[f] < 0 && face_in(f, i, j, k, in))
                h -= face_w(A, s, f, nx, nxy) * (double)y[s + face_step(f, nx, nxy)];
        acc[0] += h;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) pe2[p] = acc[0];
}

// per cluster: g = Z^T A y, c2 = E^-1 g, c = c1 - c2 (for the scatter); the
// extra dot partials of z = y + Z c: <z, r> - <y, r'> = c1.g + c.(Z^T r) and
// <z, v> - <y, v> = c.(Z^T v)
__global__ void __launch_bounds__(kBlock)
k_bnn_apply(int nc, const int *__restrict__ cp, const double *__restrict__ einv, const double *__restrict__ psr,
            const double *__restrict__ psv, const double *__restrict__ pe2, double *c, double *Pextra, const int *done)
{
    if (*done) return;
    double acc[2] = {0.0, 0.0};
    for (int a = threadIdx.x; a < nc; a += kBlock) {
        double g = 0.0, sr = 0.0, sv = 0.0;
        for (int p = cp[a]; p < cp[a + 1]; p++) { g += pe2[p]; sr += psr[p]; sv += psv[p]; }
        const double c1 = c[a];
        const double d = c1 - einv[a] * g;
        c[a] = d;
        acc[0] += c1 * g + d * sr;
        acc[1] += d * sv;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) { Pextra[0] = acc[0]; Pextra[1] = acc[1]; }
}

// ---------------------------------------------------------------- host side

// pinned host staging of at least `bytes` (grown on demand, freed with the context)
static int stage_reserve(Deflation &D, size_t bytes)
{
    // the host writes the staging (or frees it) only after the copies the
    // last setup queued from it
    if (D.stage_busy && D.stage_done && hipEventSynchronize(D.stage_done) != hipSuccess) return 1;
    D.stage_busy = false;
    if (D.h_stage_bytes >= bytes && D.h_stage) return 0;
    if (D.h_stage) (void)hipHostFree(D.h_stage);
    D.h_stage = nullptr;
    D.h_stage_bytes = 0;
    if (hipHostMalloc(&D.h_stage, bytes, hipHostMallocDefault) != hipSuccess) return 1;
    D.h_stage_bytes = bytes;
    return 0;
}

// Setup shared by the single grid and the batch. hist: per instance the
// host copy of its histogram; inst: the instances (thr filled here); dinst:
// their device copy; bmap/G: the per-cell kernels' block map and grid.
// Everything per cell runs on the device; the host chooses the clusters
// from the table of components (root, size, instance) and lays out the
// per-cluster index arrays: one round trip of a few KB.
static int finish_clusters(hipStream_t st, std::vector<DefInst> &inst, DefInst *dinst, const int *bmap, int G,
                           const Stencil<double> &A, const unsigned *hist, const unsigned short *bins, double ratio,
                           Deflation &D,
                           char *(*alloc)(void *, size_t, int), void *actx, bool records)
{
    const int ninst = (int)inst.size();
    // per instance: median bin, threshold ~ ratio x median, count certainly
    // above it and an upper bound on the count above it
    long above = 0, ub = 0;
    for (int q = 0; q < ninst; q++) {
        const unsigned *h = hist + (long)q * kHistBins;
        DefInst &in = inst[q];
        in.thr = INFINITY;
        in.tb = kHistBins;
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
        in.tb = std::min(std::max(tb, 0), kHistBins - 1);
        above += aq;
        ub += uq;
    }
    if (above < kMinCluster) return HMFEM_OK;
    if (ub > (1L << 22)) return HMFEM_OK;   // not conduits: a large part of the domain

    // slot 1: masks (G * kWavesPerBlock words), block offsets (G + 1), the
    // strong list (ub), scan tile totals (nt), then the labelling's arrays
    const int nt = (G + kScanThreads - 1) / kScanThreads;
    if (nt > kScanThreads) return HMFEM_OK;   // beyond 2^28 cells
    const int cap = (int)ub, ccap = cap / kMinCluster + 1, nchunks = (cap + kChunk - 1) / kChunk;
    const size_t mb_bytes = sizeof(unsigned long long) * (size_t)G * kWavesPerBlock;
    const size_t ints = (size_t)G + 1 + nt + (size_t)cap * 13 + 1 + 3 * (size_t)ccap +
                        (size_t)nchunks * kMaxClusters + 2 * (size_t)kMaxClusters + 1;
    char *sbuf = alloc(actx, mb_bytes + sizeof(int) * ints, 1);
    if (!sbuf) return HMFEM_ERR_NOMEM;
    unsigned long long *mask = (unsigned long long *)sbuf;
    int *boff = (int *)(sbuf + mb_bytes);
    int *slist = boff + G + 1;
    int *tot = slist + cap;
    CcWork w;
    w.par = tot + nt;
    w.root = w.par + cap;
    w.cin = w.root + cap;
    w.size = w.cin + cap;
    w.cid = w.size + cap;
    w.lord = w.cid + cap;
    w.nbr = w.lord + cap;
    w.ncand = w.nbr + 6L * cap;
    w.cand = w.ncand + 1;
    w.ccount = w.cand + 3L * ccap;
    w.sel = w.ccount + (long)nchunks * kMaxClusters;
    const int *pms = boff + G;   // the number of strong cells (capped at cap by every kernel)
    HMFEM_CHECK(hipMemcpyAsync(dinst, inst.data(), sizeof(DefInst) * ninst, hipMemcpyHostToDevice, st));
    const dim3 gm((G + kMaskSub - 1) / kMaskSub);
    hipLaunchKernelGGL(k_strong_mask, gm, dim3(kBlock), 0, st, dinst, bmap, G, A, bins, mask, boff);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanThreads), 0, st, boff, G, tot);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kScanThreads), 0, st, tot, nt, boff, G);
    hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(kScanThreads), 0, st, boff, G, (const int *)tot);
    hipLaunchKernelGGL(k_strong_emit, gm, dim3(kBlock), 0, st, dinst, bmap, G, mask, boff, slist, cap);
    const dim3 gl((cap + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_cc_links, gl, dim3(kBlock), 0, st, (const DefInst *)dinst, ninst, (const int *)slist, pms,
                       cap, w);
    hipLaunchKernelGGL(k_cc_hook, gl, dim3(kBlock), 0, st, pms, cap, w);
    for (int round = 0; round < kJumpRounds; round++)
        hipLaunchKernelGGL(k_cc_jump, gl, dim3(kBlock), 0, st, pms, cap, w);
    hipLaunchKernelGGL(k_cc_roots, gl, dim3(kBlock), 0, st, pms, cap, w);
    hipLaunchKernelGGL(k_cc_cands, gl, dim3(kBlock), 0, st, pms, cap, ccap, w);

    // the components to the host: the count, then the table (a first part in
    // the same copy; the rest, if any, in a second)
    const int first = std::min(ccap, 4096);
    if (stage_reserve(D, sizeof(int) * (1 + 3 * (size_t)ccap))) return HMFEM_ERR_NOMEM;
    int *hc = (int *)D.h_stage;
    HMFEM_CHECK(hipMemcpyAsync(hc, w.ncand, sizeof(int) * (1 + 3 * (size_t)first), hipMemcpyDeviceToHost, st));
    HMFEM_CHECK(hipStreamSynchronize(st));
    const int ncand = std::min(hc[0], ccap);
    if (ncand > first) {
        HMFEM_CHECK(hipMemcpyAsync(hc + 1 + 3 * first, w.cand + 3 * first, sizeof(int) * 3 * (size_t)(ncand - first),
                                   hipMemcpyDeviceToHost, st));
        HMFEM_CHECK(hipStreamSynchronize(st));
    }
    if (ncand == 0) return HMFEM_OK;
    // the kMaxClusters largest (ties: the smaller root first), then in root
    // order, which is instance order (the list is sorted by cell index)
    DeflationHost &W = D.h;
    auto &cand = W.cand;
    cand.resize(ncand);
    for (int q = 0; q < ncand; q++) cand[q] = {hc[1 + 3 * q], hc[2 + 3 * q], hc[3 + 3 * q]};
    std::sort(cand.begin(), cand.end(), [](const std::array<int, 3> &x, const std::array<int, 3> &y) {
        return x[1] != y[1] ? x[1] > y[1] : x[0] < y[0];
    });
    if ((int)cand.size() > kMaxClusters) cand.resize(kMaxClusters);
    std::sort(cand.begin(), cand.end(),
              [](const std::array<int, 3> &x, const std::array<int, 3> &y) { return x[0] < y[0]; });
    const int nc = (int)cand.size();
    // cluster offsets in the list, pieces of at most kPiece cells, the
    // clusters of each instance, the instance of each cluster
    auto &ptr = W.ptr, &pbh = W.pbh, &cph = W.cph, &irange = W.irange, &pc = W.pc;
    ptr.assign(nc + 1, 0);
    for (int q = 0; q < nc; q++) ptr[q + 1] = ptr[q] + cand[q][1];
    const int m = ptr[nc];
    pbh.assign(1, 0);
    cph.assign(1, 0);
    for (int q = 0; q < nc; q++) {
        for (int b = ptr[q]; b < ptr[q + 1]; b += kPiece) pbh.push_back(std::min(b + kPiece, ptr[q + 1]));
        cph.push_back((int)pbh.size() - 1);
    }
    const int np = (int)pbh.size() - 1;
    irange.assign(ninst + 1, 0);
    for (int q = 0; q < nc; q++) irange[cand[q][2] + 1]++;
    for (int q = 0; q < ninst; q++) irange[q + 1] += irange[q];
    pc.assign(pbh.begin(), pbh.end());
    pc.insert(pc.end(), cph.begin(), cph.end());
    pc.insert(pc.end(), irange.begin(), irange.end());
    for (int q = 0; q < nc; q++) pc.push_back(cand[q][2]);

    // device buffers of the correction
    D.cells = (int *)alloc(actx, sizeof(int) * m, 2);
    D.comp = (int *)alloc(actx, sizeof(int) * 2 * (size_t)m, 3);   // cluster, then instance, of each cell
    // the piece arrays, then the balancing form's records (cell, clusters of six faces, count)
    const int rec_max = 6 * m;
    D.pb = (int *)alloc(actx, sizeof(int) * (pc.size() + 1 + 7 * (size_t)rec_max), 4);
    D.einv = (double *)alloc(actx, sizeof(double) * nc, 5);
    const size_t per = (size_t)Deflation::kMaxRhs;
    D.c = (double *)alloc(actx, sizeof(double) * (per * (kMaxClusters + 2 * (size_t)np) + 2 * (size_t)ninst), 6);
    // alpha (m), per-piece E (np), per-piece Z^T A y (np), saved r (m + rec_max), record couplings (6 rec_max)
    double *share = (double *)alloc(actx, sizeof(double) * ((size_t)m + 2 * (size_t)np + m + 7 * (size_t)rec_max), 7);
    if (!D.cells || !D.comp || !D.pb || !D.einv || !D.c || !share) return HMFEM_ERR_NOMEM;
    D.cp = D.pb + np + 1;
    D.irange = D.cp + nc + 1;
    D.cinst = D.irange + ninst + 1;
    D.rec_count = D.pb + pc.size();
    D.rec_cell = D.rec_count + 1;
    D.rec_cl = D.rec_cell + rec_max;
    D.psr = D.c + per * kMaxClusters;
    D.psv = D.psr + per * np;
    D.pext = D.psv + per * np;
    D.alpha = share;
    D.pe2 = share + m + np;
    D.rsave = D.pe2 + np;
    D.rec_w = D.rsave + m + rec_max;
    D.rec_max = rec_max;
    // chosen roots (kMaxClusters slots), cluster offsets, then the piece
    // arrays, up through the pinned staging (the table above is used up)
    const size_t sel_ints = 2 * (size_t)kMaxClusters + 1;
    if (stage_reserve(D, sizeof(int) * (sel_ints + pc.size()))) return HMFEM_ERR_NOMEM;
    int *hs = (int *)D.h_stage;
    for (int q = 0; q < kMaxClusters; q++) hs[q] = q < nc ? cand[q][0] : 0;
    for (int q = 0; q <= kMaxClusters; q++) hs[kMaxClusters + q] = ptr[std::min(q, nc)];
    memcpy(hs + sel_ints, pc.data(), sizeof(int) * pc.size());
    HMFEM_CHECK(hipMemcpyAsync(w.sel, hs, sizeof(int) * sel_ints, hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipMemcpyAsync(D.pb, hs + sel_ints, sizeof(int) * pc.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_cc_ordinal, gl, dim3(kBlock), 0, st, pms, cap, nc, w);
    hipLaunchKernelGGL(k_place_count, dim3(nchunks), dim3(kChunk), 0, st, pms, cap, nc, w);
    hipLaunchKernelGGL(k_place_scan, dim3(1), dim3(kMaxClusters), 0, st, nc, nchunks, w);
    hipLaunchKernelGGL(k_place_emit, dim3(nchunks), dim3(kChunk), 0, st, (const int *)slist, pms, cap, nc, m, w,
                       D.cells, D.comp);
    hipLaunchKernelGGL(k_defl_eshare, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, st, (const DefInst *)dinst,
                       A, (const int *)D.cells, (const int *)D.comp, m, w, share);
    hipLaunchKernelGGL(k_defl_epiece, dim3(np), dim3(kBlock), 0, st, (const int *)D.pb, (const double *)share,
                       share + m);
    hipLaunchKernelGGL(k_defl_einv, dim3(1), dim3(kBlock), 0, st, nc, (const int *)D.cp, (const double *)(share + m),
                       D.einv);
    // the balancing form's records: the cells just outside the clusters
    // (none unless asked for: the additive form does not read them)
    HMFEM_CHECK(hipMemsetAsync(D.rec_count, 0, sizeof(int), st));
    if (records)
        hipLaunchKernelGGL(k_bnn_records, dim3((unsigned)((6L * m + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                           (const DefInst *)dinst, A, (const int *)D.cells, (const int *)D.comp, m,
                           (const int *)slist, pms, cap, w, D.rec_cell, D.rec_cl, D.rec_w, D.rec_count, rec_max);
    D.lord = w.lord;
    D.nbr = w.nbr;
    D.inst = dinst;
    HMFEM_CHECK(hipGetLastError());
    if (D.stage_done) {
        HMFEM_CHECK(hipEventRecord(D.stage_done, st));
        D.stage_busy = true;
    } else {
        HMFEM_CHECK(hipStreamSynchronize(st));
    }
    D.nc = nc;
    D.m = m;
    D.np = np;
    D.ninst = ninst;
    return HMFEM_OK;
}

// Slot 0 of a single-grid setup: the histogram, the instance descriptor, then
// the cells' bins.
static size_t slot0_bins_offset() { return (kHistBins * sizeof(unsigned) + 256 + sizeof(DefInst) + 255) & ~(size_t)255; }

int deflation_begin(hipStream_t st, Dims d, const Stencil<double> &A, Deflation &D,
                    char *(*alloc)(void *, size_t, int), void *actx, bool external_hist)
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
    char *h0 = alloc(actx, slot0_bins_offset() + sizeof(unsigned short) * (size_t)n, 0);
    if (!h0) return HMFEM_ERR_NOMEM;
    unsigned *hist = (unsigned *)h0;
    unsigned short *bins = (unsigned short *)(h0 + slot0_bins_offset());
    HMFEM_CHECK(hipMemsetAsync(hist, 0, kHistBins * sizeof(unsigned), st));
    D.pending = true;
    D.h.bhist = hist;
    D.h.bbins = bins;
    if (external_hist) return HMFEM_OK;
    // a grid-stride loop over at most 512 blocks keeps the global bin updates
    // few (2048 blocks: 35 us instead of 25 at 128^3)
    const int G = std::min(grid_for(n, kBlock), 512);
    hipLaunchKernelGGL(k_diag_hist, dim3(G), dim3(kBlock), 0, st, o, hist, bins);
    return deflation_hist_queued(st, D);
}

int deflation_hist_queued(hipStream_t st, Deflation &D)
{
    if (!D.pending || !D.h.bhist) return HMFEM_OK;
    HMFEM_CHECK(hipMemcpyAsync(D.h_hist, D.h.bhist, kHistBins * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HMFEM_CHECK(hipEventRecord(D.hist_ready, st));
    D.h.bhist = nullptr;   // queued
    return HMFEM_OK;
}

int deflation_finish(hipStream_t st, Dims d, const Stencil<double> &A, double ratio, Deflation &D,
                     char *(*alloc)(void *, size_t, int), void *actx, bool records)
{
    if (!D.pending) return HMFEM_OK;
    D.pending = false;
    if (!(ratio > 1.0)) return HMFEM_OK;
    const long n = d.n();
    char *h0 = alloc(actx, slot0_bins_offset() + sizeof(unsigned short) * (size_t)n, 0);
    if (D.h.bhist) {   // begun with external_hist, but no pass took the histogram
        OpArgs<double> o;
        o.e = A.e; o.wx = A.wx; o.wy = A.wy; o.wz = A.wz; o.di = nullptr;
        o.nx = d.nx; o.ny = d.ny; o.nz = d.nz;
        hipLaunchKernelGGL(k_diag_hist, dim3(std::min(grid_for(n, kBlock), 512)), dim3(kBlock), 0, st, o, D.h.bhist,
                           D.h.bbins);
        int rc = deflation_hist_queued(st, D);
        if (rc) return rc;
    }
    DefInst *dinst = (DefInst *)(h0 + kHistBins * sizeof(unsigned));
    std::vector<DefInst> inst(1, DefInst{d.nx, d.ny, d.nz, (int)n, 0, 0, INFINITY, 0, kHistBins});
    HMFEM_CHECK(hipEventSynchronize(D.hist_ready));
    hipStream_t ds = D.side && D.side_done ? D.side : st;   // as in deflation_finish_batch
    if (ds != st) HMFEM_CHECK(hipStreamWaitEvent(ds, D.hist_ready, 0));
    int rc = finish_clusters(ds, inst, dinst, nullptr, (int)((n + kBlock - 1) / kBlock), A, D.h_hist,
                             (const unsigned short *)(h0 + slot0_bins_offset()), ratio, D, alloc, actx, records);
    if (ds != st) {
        HMFEM_CHECK(hipEventRecord(D.side_done, ds));
        HMFEM_CHECK(hipStreamWaitEvent(st, D.side_done, 0));
    }
    return rc;
}

int deflation_begin_batch(hipStream_t st, int ninst, const int *dims, const Stencil<double> &A, double ratio,
                          Deflation &D, char *(*alloc)(void *, size_t, int), void *actx, bool external_hist)
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
        in.tb = kHistBins;
        bm.insert(bm.end(), (in.n + kBlock - 1) / kBlock, q);
        hm.insert(hm.end(), (in.n + kBlock * kHistCells - 1) / (kBlock * kHistCells), q);
        off += in.n;
    }
    const int G = (int)bm.size(), GH = (int)hm.size();
    // slot 0: per-instance histograms, the block maps, the instance
    // descriptors, the cells' bins
    auto up256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t hb = up256(sizeof(unsigned) * kHistBins * (size_t)ninst), bb = up256(sizeof(int) * (size_t)(G + GH));
    const size_t ib = up256(sizeof(DefInst) * ninst);
    char *h0 = alloc(actx, hb + bb + ib + sizeof(unsigned short) * (size_t)off, 0);
    if (!h0) return HMFEM_ERR_NOMEM;
    unsigned *hist = (unsigned *)h0;
    int *bmap = (int *)(h0 + hb), *hmap = bmap + G;
    DefInst *dinst = (DefInst *)(h0 + hb + bb);
    unsigned short *bins = (unsigned short *)(h0 + hb + bb + ib);
    bm.insert(bm.end(), hm.begin(), hm.end());
    HMFEM_CHECK(hipMemsetAsync(hist, 0, sizeof(unsigned) * kHistBins * (size_t)ninst, st));
    HMFEM_CHECK(hipMemcpyAsync(bmap, bm.data(), sizeof(int) * (G + GH), hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipMemcpyAsync(dinst, inst.data(), sizeof(DefInst) * ninst, hipMemcpyHostToDevice, st));
    D.h.binst.assign((const char *)inst.data(), (const char *)inst.data() + sizeof(DefInst) * inst.size());
    D.h.bdinst = dinst;
    D.h.bbmap = bmap;
    D.h.bhist = hist;
    D.h.bbins = bins;
    D.h.bG = G;
    D.h.bratio = ratio;
    D.pending = true;
    if (external_hist) return HMFEM_OK;
    hipLaunchKernelGGL(k_diag_hist_inst, dim3(GH), dim3(kBlock), 0, st, (const DefInst *)dinst, (const int *)hmap, A,
                       hist, bins);
    return deflation_hist_queued_batch(st, D);
}

int deflation_hist_queued_batch(hipStream_t st, Deflation &D)
{
    if (!D.pending) return HMFEM_OK;
    // the histograms go through the pinned staging, which finish_clusters
    // reads before it reuses (and may regrow) it
    const size_t bytes = sizeof(unsigned) * kHistBins * (D.h.binst.size() / sizeof(DefInst));
    if (stage_reserve(D, bytes)) return HMFEM_ERR_NOMEM;
    HMFEM_CHECK(hipMemcpyAsync(D.h_stage, D.h.bhist, bytes, hipMemcpyDeviceToHost, st));
    HMFEM_CHECK(hipEventRecord(D.hist_ready, st));
    return HMFEM_OK;
}

int deflation_finish_batch(hipStream_t st, const Stencil<double> &A, Deflation &D,
                           char *(*alloc)(void *, size_t, int), void *actx)
{
    if (!D.pending) return HMFEM_OK;
    D.pending = false;
    std::vector<DefInst> inst(D.h.binst.size() / sizeof(DefInst));
    memcpy((void *)inst.data(), D.h.binst.data(), D.h.binst.size());
    HMFEM_CHECK(hipEventSynchronize(D.hist_ready));
    // on the side stream, if there is one, so that the cluster setup and its
    // round trip to the host overlap the hierarchy's last kernels (the dense
    // inverses use one workgroup per instance); it reads only the fine level
    // and the bins, which precede hist_ready on st, and st waits for its end
    hipStream_t ds = D.side && D.side_done ? D.side : st;
    if (ds != st) HMFEM_CHECK(hipStreamWaitEvent(ds, D.hist_ready, 0));
    // finish_clusters reads the histograms before it reuses the staging
    int rc = finish_clusters(ds, inst, (DefInst *)D.h.bdinst, D.h.bbmap, D.h.bG, A, (const unsigned *)D.h_stage,
                             D.h.bbins, D.h.bratio, D, alloc, actx, false);
    if (ds != st) {
        HMFEM_CHECK(hipEventRecord(D.side_done, ds));
        HMFEM_CHECK(hipStreamWaitEvent(st, D.side_done, 0));
    }
    return rc;
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

void deflation_pre(hipStream_t st, const Deflation &D, double *r, const int *done)
{
    const long n = (long)D.m + D.rec_max;
    hipLaunchKernelGGL(k_bnn_pre, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, D.m,
                       (const int *)D.cells, (const int *)D.comp, (const double *)D.alpha, (const double *)D.c,
                       (const int *)D.rec_cell, (const int *)D.rec_cl, (const double *)D.rec_w,
                       (const int *)D.rec_count, D.rec_max, r, D.rsave, done);
}

template <typename T>
static void bnn_post(hipStream_t st, const Deflation &D, const Stencil<double> &A, double *r, const T *y,
                     double *Pextra, const int *done)
{
    const unsigned G = (unsigned)(D.np + (D.rec_max + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((k_bnn_post<T>), dim3(G), dim3(kBlock), 0, st, (const DefInst *)D.inst, A, D.np,
                       (const int *)D.pb, (const int *)D.cells, (const int *)D.comp, D.m, (const double *)D.alpha,
                       D.lord, D.nbr, y, (const int *)D.rec_cell, (const int *)D.rec_count, D.rec_max,
                       (const double *)D.rsave, r, D.pe2, done);
    hipLaunchKernelGGL(k_bnn_apply, dim3(1), dim3(kBlock), 0, st, D.nc, (const int *)D.cp, (const double *)D.einv,
                       (const double *)D.psr, (const double *)D.psv, (const double *)D.pe2, D.c, Pextra, done);
}

void deflation_post(hipStream_t st, const Deflation &D, const Stencil<double> &A, double *r, const float *y,
                    double *Pextra, const int *done)
{
    bnn_post<float>(st, D, A, r, y, Pextra, done);
}

void deflation_post(hipStream_t st, const Deflation &D, const Stencil<double> &A, double *r, const double *y,
                    double *Pextra, const int *done)
{
    bnn_post<double>(st, D, A, r, y, Pextra, done);
}

void deflation_sums_batch(hipStream_t st, const Deflation &D, const double *r, const double *v, const int *never)
{
    hipLaunchKernelGGL(k_defl_sums<1>, dim3(D.np), dim3(kBlock), 0, st, D.pb, D.cells, r, v, D.psr, D.psv, never);
    hipLaunchKernelGGL(k_defl_apply_inst, dim3(1), dim3(kBlock), 0, st, D.nc, D.cp, D.einv, D.psr, D.psv, D.irange,
                       D.ninst, D.c, D.pext);
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
