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
#include <numeric>
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
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
        Cell c = cell_at(o, s);
        const double d = diag_at<double, double>(o, c);
        int b = d > 0 ? ilogb(d) + 1100 : 0;
        b = b < 0 ? 0 : (b >= kHistBins ? kHistBins - 1 : b);
        atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kHistBins; q += kBlock)
        if (h[q]) atomicAdd(&hist[q], h[q]);
}

// Strong cells (diagonal above thr) in index order, in three steps: a bit
// mask per wave and a count per block (logical block lb covers cells
// [lb*kBlock, lb*kBlock + kBlock)), an exclusive scan of the counts, and the
// ordered list.
constexpr int kWavesPerBlock = kBlock / 64;

__global__ void __launch_bounds__(kBlock)
k_strong_mask(OpArgs<double> o, double thr, unsigned long long *mask, int *bcnt)
{
    __shared__ int wc[kWavesPerBlock];
    const int n = o.nx * o.ny * o.nz;
    const int lb = xcd_block();
    const int s = lb * kBlock + threadIdx.x;
    bool f = false;
    if (s < n) {
        Cell c = cell_at(o, s);
        f = diag_at<double, double>(o, c) > thr;
    }
    const unsigned long long m = __ballot(f);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        mask[(long)lb * kWavesPerBlock + w] = m;
        wc[w] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int q = 0; q < kWavesPerBlock; q++) t += wc[q];
        bcnt[lb] = t;
    }
}

// In place exclusive scan of cnt[0..G), total in cnt[G]; one block.
constexpr int kScanThreads = 1024;
__global__ void __launch_bounds__(kScanThreads)
k_scan_counts(int *cnt, int G)
{
    __shared__ int part[kScanThreads];
    const int per = (G + kScanThreads - 1) / kScanThreads;
    const int beg = threadIdx.x * per, end = min(G, beg + per);
    int t = 0;
    for (int q = beg; q < end; q++) t += cnt[q];
    part[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {   // 1024 partial sums: sequential is cheap enough for a setup step
        int a = 0;
        for (int q = 0; q < kScanThreads; q++) { const int v = part[q]; part[q] = a; a += v; }
        cnt[G] = a;
    }
    __syncthreads();
    int a = part[threadIdx.x];
    for (int q = beg; q < end; q++) { const int v = cnt[q]; cnt[q] = a; a += v; }
}

__global__ void __launch_bounds__(kBlock)
k_strong_emit(const unsigned long long *__restrict__ mask, const int *__restrict__ boff, int *list, int cap)
{
    const int lb = xcd_block();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long m = mask[(long)lb * kWavesPerBlock + w];
    if (!((m >> lane) & 1ull)) return;
    int at = boff[lb];
    for (int q = 0; q < w; q++) at += __popcll(mask[(long)lb * kWavesPerBlock + q]);
    at += __popcll(m & ((1ull << lane) - 1ull));
    if (at < cap) list[at] = lb * kBlock + threadIdx.x;
}

// Coefficients of listed cells: e and the couplings to -x,+x,-y,+y,-z,+z (0 where none).
__global__ void __launch_bounds__(kBlock)
k_defl_gather(OpArgs<double> o, const int *__restrict__ cells, int m, double *out)
{
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= m) return;
    const int s = cells[q];
    Cell c = cell_at(o, s);
    const int nx = o.nx, nxy = o.nx * o.ny;
    double *w = out + 7L * q;
    w[0] = o.e[s];
    w[1] = c.i > 0 ? o.wx[s - 1] : 0.0;
    w[2] = c.i < o.nx - 1 ? o.wx[s] : 0.0;
    w[3] = c.j > 0 ? o.wy[s - nx] : 0.0;
    w[4] = c.j < o.ny - 1 ? o.wy[s] : 0.0;
    w[5] = c.k > 0 ? o.wz[s - nxy] : 0.0;
    w[6] = c.k < o.nz - 1 ? o.wz[s] : 0.0;
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

int deflation_begin(hipStream_t st, Dims d, const Stencil<double> &A, Deflation &D,
                    char *(*alloc)(void *, size_t, int), void *actx)
{
    D.nc = 0;
    D.m = 0;
    D.pending = false;
    const long n = d.n();
    if (n < 64 || !D.h_hist || !D.hist_ready) return HMFEM_OK;
    OpArgs<double> o;
    o.e = A.e; o.wx = A.wx; o.wy = A.wy; o.wz = A.wz; o.di = nullptr;
    o.nx = d.nx; o.ny = d.ny; o.nz = d.nz;
    unsigned *hist = (unsigned *)alloc(actx, kHistBins * sizeof(unsigned), 0);
    if (!hist) return HMFEM_ERR_NOMEM;
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
    const long n = d.n();
    if (!(ratio > 1.0)) return HMFEM_OK;
    OpArgs<double> o;
    o.e = A.e; o.wx = A.wx; o.wy = A.wy; o.wz = A.wz; o.di = nullptr;
    o.nx = d.nx; o.ny = d.ny; o.nz = d.nz;
    const int G = grid_for(n, kBlock);
    HMFEM_CHECK(hipEventSynchronize(D.hist_ready));
    const unsigned *h = D.h_hist;
    // median bin, threshold, and the count of cells certainly above it
    long acc = 0;
    int mb = 0;
    for (; mb < kHistBins; mb++) {
        acc += h[mb];
        if (2 * acc >= n) break;
    }
    const double thr = ratio * ldexp(1.0, mb - 1100) * 1.4142135623730951;   // ~ ratio x median
    const int tb = ilogb(thr) + 1100;
    long above = 0;
    for (int b = tb + 1; b < kHistBins; b++) above += h[b];
    if (above < kMinCluster) return HMFEM_OK;

    // the strong cells, compacted on the device (at most ub of them)
    long ub = 0;
    for (int b = tb; b < kHistBins; b++) ub += h[b];
    if (ub > (1L << 22)) return HMFEM_OK;   // not conduits: a large part of the domain
    // slot 1: masks (G * kWavesPerBlock words), block offsets (G + 1), list (ub)
    const size_t mb_bytes = sizeof(unsigned long long) * (size_t)G * kWavesPerBlock;
    char *sbuf = alloc(actx, mb_bytes + sizeof(int) * ((size_t)G + 1 + ub), 1);
    if (!sbuf) return HMFEM_ERR_NOMEM;
    unsigned long long *mask = (unsigned long long *)sbuf;
    int *boff = (int *)(sbuf + mb_bytes);
    int *slist = boff + G + 1;
    hipLaunchKernelGGL(k_strong_mask, dim3(G), dim3(kBlock), 0, st, o, thr, mask, boff);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kScanThreads), 0, st, boff, G);
    hipLaunchKernelGGL(k_strong_emit, dim3(G), dim3(kBlock), 0, st, mask, boff, slist, (int)ub);
    // the total (boff[G]) and the list in one copy to pinned staging
    const size_t cb = sizeof(int) * (1 + ub);
    if (stage_reserve(D, cb + sizeof(double) * 7 * (size_t)ub)) return HMFEM_ERR_NOMEM;
    int *hl = (int *)D.h_stage;
    HMFEM_CHECK(hipMemcpyAsync(hl, boff + G, cb, hipMemcpyDeviceToHost, st));
    HMFEM_CHECK(hipStreamSynchronize(st));
    const int ms = std::min(hl[0], (int)ub);
    if (ms < kMinCluster) return HMFEM_OK;
    std::vector<int> cell(hl + 1, hl + 1 + ms);   // in index order: a deterministic labelling

    // strong neighbours of each strong cell: ordinals of the +x, +y, +z
    // neighbours by one monotone sweep per direction (cell[] is sorted), the
    // -x, -y, -z ones by symmetry
    const long nx = d.nx, nxy = (long)d.nx * d.ny;
    std::vector<int> nbr(6L * ms, -1);   // [a][-x,+x,-y,+y,-z,+z]
    {
        const long off[3] = {1, nx, nxy};
        for (int f = 0; f < 3; f++) {
            int p = 0;
            for (int a = 0; a < ms; a++) {
                const long s = cell[a];
                const bool inside = f == 0 ? s % nx + 1 < d.nx : f == 1 ? (s / nx) % d.ny + 1 < d.ny
                                                                        : s / nxy + 1 < d.nz;
                if (!inside) continue;
                const long t = s + off[f];
                while (p < ms && cell[p] < t) p++;
                if (p < ms && cell[p] == t) {
                    nbr[6L * a + 2 * f + 1] = p;
                    nbr[6L * p + 2 * f] = a;
                }
            }
        }
    }
    // union-find over the +x, +y, +z links
    std::vector<int> par(ms);
    std::iota(par.begin(), par.end(), 0);
    for (int a = 0; a < ms; a++)
        for (int f = 1; f < 6; f += 2) {
            const int b = nbr[6L * a + f];
            if (b < 0) continue;
            int ra = find_root(par, a), rb = find_root(par, b);
            if (ra != rb) par[std::max(ra, rb)] = std::min(ra, rb);
        }
    std::vector<int> root(ms), size(ms, 0);
    for (int a = 0; a < ms; a++) { root[a] = find_root(par, a); size[root[a]]++; }
    std::vector<int> roots;
    for (int a = 0; a < ms; a++)
        if (root[a] == a && size[a] >= kMinCluster) roots.push_back(a);
    std::stable_sort(roots.begin(), roots.end(), [&](int x, int y) { return size[x] > size[y]; });
    if ((int)roots.size() > kMaxClusters) roots.resize(kMaxClusters);
    std::sort(roots.begin(), roots.end());   // cluster order: by smallest member
    const int nc = (int)roots.size();
    if (nc == 0) return HMFEM_OK;
    std::vector<int> cid(ms, -1);
    for (int q = 0; q < nc; q++) cid[roots[q]] = q;
    for (int a = 0; a < ms; a++) cid[a] = cid[root[a]];
    std::vector<int> ptr(nc + 1, 0), lcell, lcomp, lord;
    for (int a = 0; a < ms; a++)
        if (cid[a] >= 0) ptr[cid[a] + 1]++;
    for (int q = 0; q < nc; q++) ptr[q + 1] += ptr[q];
    const int m = ptr[nc];
    lcell.resize(m);
    lcomp.resize(m);
    lord.resize(m);
    {
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int a = 0; a < ms; a++)
            if (cid[a] >= 0) {
                const int q = fill[cid[a]]++;
                lcell[q] = cell[a];
                lcomp[q] = cid[a];
                lord[q] = a;
            }
    }

    // pieces of at most kPiece cells per cluster, contiguous in the list
    std::vector<int> pbh(1, 0), cph(1, 0);
    for (int q = 0; q < nc; q++) {
        for (int b = ptr[q]; b < ptr[q + 1]; b += kPiece) pbh.push_back(std::min(b + kPiece, ptr[q + 1]));
        cph.push_back((int)pbh.size() - 1);
    }
    const int np = (int)pbh.size() - 1;
    std::vector<int> pc(pbh);
    pc.insert(pc.end(), cph.begin(), cph.end());

    // device copies, coefficients of the listed cells
    D.cells = (int *)alloc(actx, sizeof(int) * m, 2);
    D.comp = (int *)alloc(actx, sizeof(int) * m, 3);
    D.pb = (int *)alloc(actx, sizeof(int) * pc.size(), 4);
    D.einv = (double *)alloc(actx, sizeof(double) * nc * nc, 5);
    const size_t per = (size_t)Deflation::kMaxRhs;
    D.c = (double *)alloc(actx, sizeof(double) * per * (kMaxClusters + 2 * (size_t)np), 6);
    double *coef = (double *)alloc(actx, sizeof(double) * 7 * m, 7);
    if (!D.cells || !D.comp || !D.pb || !D.einv || !D.c || !coef) return HMFEM_ERR_NOMEM;
    D.cp = D.pb + np + 1;
    D.psr = D.c + per * kMaxClusters;
    D.psv = D.psr + per * np;
    HMFEM_CHECK(hipMemcpyAsync(D.cells, lcell.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipMemcpyAsync(D.comp, lcomp.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipMemcpyAsync(D.pb, pc.data(), sizeof(int) * pc.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_defl_gather, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, st, o, D.cells, m, coef);
    const double *w = (const double *)D.h_stage;   // 7 m <= 7 ub doubles: reserved above
    HMFEM_CHECK(hipMemcpyAsync(D.h_stage, coef, sizeof(double) * 7 * m, hipMemcpyDeviceToHost, st));
    HMFEM_CHECK(hipStreamSynchronize(st));

    // E = Z^T A Z in a fixed order: E_aa = sum (e + couplings leaving the
    // cluster), E_ab = - couplings between clusters a and b
    std::vector<double> E((size_t)nc * nc, 0.0);
    for (int q = 0; q < m; q++) {
        const int a = lcomp[q];
        const double *ws = &w[7L * q];
        const int *nb = &nbr[6L * lord[q]];
        double daa = ws[0];
        for (int f = 0; f < 6; f++) {
            const double wf = ws[1 + f];
            if (wf == 0.0) continue;
            const int b = nb[f] < 0 ? -1 : cid[nb[f]];
            if (b == a) continue;
            daa += wf;
            if (b >= 0) E[(size_t)a * nc + b] -= wf;
        }
        E[(size_t)a * nc + a] += daa;
    }
    // inverse via Cholesky E = L L^T (E is symmetric positive definite)
    std::vector<double> L = E;
    for (int j = 0; j < nc; j++) {
        double djj = L[(size_t)j * nc + j];
        for (int k = 0; k < j; k++) djj -= L[(size_t)j * nc + k] * L[(size_t)j * nc + k];
        if (!(djj > 0)) return HMFEM_OK;   // not positive definite: leave the correction off
        djj = sqrt(djj);
        L[(size_t)j * nc + j] = djj;
        for (int i = j + 1; i < nc; i++) {
            double t = L[(size_t)i * nc + j];
            for (int k = 0; k < j; k++) t -= L[(size_t)i * nc + k] * L[(size_t)j * nc + k];
            L[(size_t)i * nc + j] = t / djj;
        }
    }
    std::vector<double> inv((size_t)nc * nc, 0.0);
    for (int col = 0; col < nc; col++) {   // solve L L^T x = e_col
        std::vector<double> y(nc, 0.0);
        for (int i = 0; i < nc; i++) {
            double t = i == col ? 1.0 : 0.0;
            for (int k = 0; k < i; k++) t -= L[(size_t)i * nc + k] * y[k];
            y[i] = t / L[(size_t)i * nc + i];
        }
        for (int i = nc - 1; i >= 0; i--) {
            double t = y[i];
            for (int k = i + 1; k < nc; k++) t -= L[(size_t)k * nc + i] * inv[(size_t)k * nc + col];
            inv[(size_t)i * nc + col] = t / L[(size_t)i * nc + i];
        }
    }
    HMFEM_CHECK(hipMemcpyAsync(D.einv, inv.data(), sizeof(double) * nc * nc, hipMemcpyHostToDevice, st));
    HMFEM_CHECK(hipStreamSynchronize(st));
    D.nc = nc;
    D.m = m;
    D.np = np;
    return HMFEM_OK;
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
