// Internal declarations shared by the HIP translation units of libhmfem_gpu.
//
// Data layout in HBM (one grid of nx*ny*nz cells, s = i + nx*(j + ny*k)):
//   a level of the hierarchy is a structured 7-point operator in flux form,
//     (A u)_s = e_s u_s + sum_faces w_f (u_s - u_nb),
//   stored as four SoA arrays: e (boundary/source "excess" of the diagonal),
//   wx, wy, wz (coupling to the +x, +y, +z neighbour; 0 on the last plane).
//   The -x coupling of cell s is wx[s-1], and so on, so each face is stored
//   once and the operator is symmetric by construction.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <array>
#include <functional>
#include <vector>

#include "../../include/hmfem_gpu.h"

#define HMFEM_CHECK(call)                                                                      \
    do {                                                                                       \
        hipError_t _e = (call);                                                                \
        if (_e != hipSuccess) {                                                                \
            fprintf(stderr, "hmfem_gpu: %s failed at %s:%d: %s\n", #call, __FILE__, __LINE__, \
                    hipGetErrorString(_e));                                                    \
            return HMFEM_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

namespace hmfem {

constexpr int kBlock = 256;        // threads per block for stencil / vector kernels
constexpr int kCellsPerThread = 4; // cells per thread in reduction kernels
constexpr int kMaxLevels = 24;
constexpr int kNumXcd = 8;         // MI355X: 8 XCDs, blocks dispatched round-robin
constexpr int kCoarseMax = 80;     // largest coarsest level inverted densely in LDS (FP64, <= 64 KB)
constexpr int kMaxReduceBlocks = 1024; // blocks of a chunked reduction launch (one partial each)

struct Dims {
    int nx, ny, nz;
    long n() const { return (long)nx * ny * nz; }
};

// Structured stencil level. T is the storage type of the coefficients.
template <typename T>
struct Stencil {
    T *e = nullptr, *wx = nullptr, *wy = nullptr, *wz = nullptr;
    T *di = nullptr;   // inverse diagonal (smoothers); computed at setup
    long eface = 0;    // 1: e is zero except on cells of the grid's six faces (no well or conduit terms)
};

struct Level {
    Dims d{0, 0, 0};
    int fx = 1, fy = 1, fz = 1;      // aggregation factors towards the next level
    Stencil<double> s64;             // FP64 coefficients (setup, FP64 V-cycle)
    Stencil<float> s32;              // FP32 coefficients (FP32 V-cycle)
    Stencil<__half> s16;             // FP16 scaled fine-level values (ablation)
    float *sc = nullptr;             // FP16 row scaling (sqrt of diagonal), fine level only
    void *f = nullptr, *u = nullptr, *t = nullptr; // V-cycle vectors (float or double)
    void *r = nullptr, *w = nullptr;               // second right-hand side and saved u (W-cycle)
};

// Device scalars of one CG solve, kept in one small allocation.
// sigma and alpha keep two slots (by parity of the CG iteration) so that a
// kernel never overwrites a value its own blocks still read.
// The iteration index lives on the device too (kspmv: next SpMV; kupd: the
// SpMV's index handed to the update), so one captured graph serves every
// iteration.
struct CgScalars {
    double bb, rr, sigma_first, tau, beta;
    double sigma[2], alpha[2];
    double rtol2;
    int done, iter, maxit, stop_rule, status, kspmv, kupd;
    // single-grid CG: the iteration whose x += alpha s is still to be applied
    // (the next SpMV applies it while it reads s; -1: none)
    int xpend;
};

// A captured CG iteration, reused while its launch parameters are unchanged.
struct GraphCache {
    hipGraphExec_t exec = nullptr;
    std::vector<char> key;
};

// The last single-solve setup of a context (hierarchy bytes and the key it
// was built for), for opt.reuse_setup.
struct SetupCache {
    bool valid = false;
    std::vector<char> key;
    std::vector<char> hier;   // the Hierarchy struct (plain data; its buffers live in the arena)
    unsigned dgen = 0;        // Deflation::gen at the build
};

// Per-axis coupling strength of a fine operator (geometric means of its
// nonzero x, y, z couplings) for the coarsening plan: an axis is coarsened
// only where its coupling is at least kSemiTheta times the strongest axis's
// (strength-based semi-coarsening, which follows anisotropy from the cell
// sizes or from K). off: 2x2x2 boxes throughout.
struct PlanHint {
    bool on = false;
    double m[3] = {0.0, 0.0, 0.0};
};
constexpr double kSemiTheta = 0.25;

// Grid launch helpers with an XCD-aware block order: logical block b of a
// launch with G blocks (G a multiple of 8) is placed so that each XCD gets a
// contiguous range of the domain and keeps its neighbours in its own L2.
__device__ __forceinline__ int xcd_block()
{
    int G = gridDim.x;
    int b = blockIdx.x;
    return (b % kNumXcd) * (G / kNumXcd) + b / kNumXcd;
}

inline int grid_for(long items, int per_block)
{
    long g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
    return (int)g;
}

} // namespace hmfem

namespace hmfem {

// launchers defined in hmfem_assemble.hip
int launch_assemble(hipStream_t st, const hmfem_grid *h, const double *Kx, const double *Ky, const double *Kz,
                    const double *KI, const double *WI, const double *pc, const double *pw,
                    const double *const bval[6], double *e, double *wx, double *wy, double *wz, double *b,
                    ELE *perm, IND *indcol);
int launch_source_from_perm(hipStream_t st, const hmfem_grid *h, const ELE *perm, const IND *indcol,
                            const double *KI, const double *WI, const double *pc, const double *pw,
                            const double *const bval[6], double *b);
int launch_perm_to_stencil(hipStream_t st, int nx, int ny, int nz, const ELE *perm, const IND *indcol, double *e,
                           double *wx, double *wy, double *wz, int *ok);
int launch_flux(hipStream_t st, const hmfem_grid *h, const double *Kx, const double *Ky, const double *Kz,
                const ELE *perm, const double *const bval[6], const double *p, ELE *v, double *partial,
                unsigned int *counter, double *out);
int launch_divergence(hipStream_t st, const hmfem_grid *h, const ELE *v, double *partial, unsigned int *counter,
                      double *out);
int launch_prod_mat_vec(hipStream_t st, long n, const ELE *perm, const IND *indcol, const double *x, double *res);

// 1-D karst conduits (hmfem_conduit.hip)
constexpr int kConduitLdsCells = 1024;   // conduits up to this length are solved in LDS (PCR)
constexpr int kRelChangeBlocks = 512;    // blocks of the relative-change reduction

struct ConduitDesc {
    int off, m, axis, bc_l, bc_r, nbrec, boff;   // cells [off, off + m); breccias [boff, boff + nbrec)
    long rs_l, rsa_l, rs_r, rsa_r;               // Robin ends: reservoir cell outside the end, end cell
    double hs, Kc, KI_c, cont_l, cont_r, ccont_l, ccont_r;
};

struct ConduitArgs {
    const ConduitDesc *desc;
    const int *res_cell;                         // reservoir cell of each conduit cell
    const int *brec_cell;                        // per breccia (conduit by conduit): cell index in its conduit
    const double *brec_beta, *brec_C;
    double *ucR, *ucL;                           // breccia velocities (carried between passes)
    const double *p;                             // reservoir pressure (exchange and Robin ends)
    const double *Kx, *Ky, *Kz;
    double hx, hy, hz, mu, rho;
    const double *pc_old;                        // previous pass's conduit pressure
    double *pc, *kl, *kr, *kc, *fbuf, *cp;       // pressure out, transmissibilities, Thomas scratch
};

int conduit_pass(hipStream_t st, const ConduitArgs &a, int count, int mmax);
int conduit_flux(hipStream_t st, const ConduitArgs &a, int count, double *ucl, double *ucr);
int conduit_scatter(hipStream_t st, int count, const int *res, const int *cell, const double *pc_old,
                    double *pc_aux);
int rel_change(hipStream_t st, long n, const double *p, double *p_old, double *partial, double *out);

// per-kernel timing collected by the solver (profile classes below)
enum {
    HMFEM_PROF_CG_SPMV = 0,
    HMFEM_PROF_CG_UPDATE,
    HMFEM_PROF_CG_DOTS,
    HMFEM_PROF_SMOOTH0,
    HMFEM_PROF_RESTRICT,
    HMFEM_PROF_PROLONG,
    HMFEM_PROF_JACOBI,
    HMFEM_PROF_COARSE,
    HMFEM_PROF_TYPES
};
struct ProfileSink {
    hipEvent_t *ev;
    int *type;
    int cap;
    double ms[HMFEM_PROF_TYPES];
    long count[HMFEM_PROF_TYPES];
    double bytes[HMFEM_PROF_TYPES];
};

// defined in hmfem_solver.hip
// The paths' own cycle settings where the options leave them at 0 (measured
// best on each: single-grid and block solves two sweeps at 2.2, batches four
// at 2.0).
constexpr double kOcfSingle = 2.2, kOcfBatch = 2.0;
constexpr int kSweepsSingle = 2, kSweepsBatch = 4;
inline double ocf_of(const hmfem_options &o, double own) { return o.ocf > 0 ? o.ocf : own; }

struct CgBuffers {
    double *r, *s0, *s1, *v;
    CgScalars *sc;
    double *partial;
};
// Additive correction on clusters of strongly coupled cells (hmfem_defl.hip).
// nc == 0: no correction for this operator.
// Host work arrays of the cluster setup, kept by the context so that repeated
// setups reuse their memory (fresh vectors of a few MB cost their page faults
// every time).
struct DeflationHost {
    std::vector<std::array<int, 3>> cand;   // components: root, size, instance
    std::vector<int> ptr, pbh, cph, irange, pc;
    // a batch setup between deflation_begin_batch and deflation_finish_batch:
    // the instance descriptors (bytes), their device copy, the block map, the
    // cells' diagonal bins
    std::vector<char> binst;
    void *bdinst = nullptr;
    const int *bbmap = nullptr;
    unsigned *bhist = nullptr;
    unsigned short *bbins = nullptr;
    int bG = 0;
    double bratio = 0.0;
};

struct Deflation {
    int nc = 0, m = 0, np = 0;               // clusters, cells listed (grouped by cluster), pieces
    int *cells = nullptr, *comp = nullptr;   // cell index and cluster of each listed cell
    int *pb = nullptr;                       // piece p covers listed cells [pb[p], pb[p+1])
    int *cp = nullptr;                       // cluster q owns pieces [cp[q], cp[q+1])
    double *einv = nullptr;                  // 1 / E_aa (E = Z^T A Z is diagonal), nc
    double *c = nullptr;                     // per-cluster correction [cluster][rhs]
    double *psr = nullptr, *psv = nullptr;   // per-piece sums of r and v [piece][rhs]
    int ninst = 0;                           // grids of the setup (batch instances; 1 for a single grid)
    int *irange = nullptr;                   // instance q owns clusters [irange[q], irange[q+1])
    int *cinst = nullptr;                    // instance of each cluster
    double *pext = nullptr;                  // batch: per-instance extra dot partials [instance][2]
    // balancing form (single grid): r' = r - A Z c1 before the V-cycle, its
    // output y projected after it (z = y + Z (c1 - E^-1 Z^T A y))
    double *alpha = nullptr;                 // (A 1_a)_s of each listed cell s (e + couplings leaving a)
    const int *lord = nullptr, *nbr = nullptr;   // strong ordinal of each listed cell; six strong neighbours
    const void *inst = nullptr;              // the instance descriptors (coordinates of listed cells)
    int *rec_cell = nullptr, *rec_cl = nullptr, *rec_count = nullptr;   // cells just outside the clusters
    double *rec_w = nullptr;                 // their couplings to cluster cells [record][face] (cluster rec_cl)
    double *rsave = nullptr;                 // r of the listed cells, then of the records, while r' is in r
    double *pe2 = nullptr;                   // per-piece sums of (A Z)^T y
    int rec_max = 0;                         // records room (6 m)
    static constexpr int kHistBins = 2048;   // bins of floor(log2 diagonal) + 1100
    static constexpr int kMaxRhs = 8;        // interleaved right-hand sides of the per-cluster buffers
    unsigned *h_hist = nullptr;              // pinned host copy of the diagonal histogram (context-owned)
    hipEvent_t hist_ready = nullptr;         // recorded after that copy (context-owned)
    void *h_stage = nullptr;                 // pinned staging of the component table and cluster arrays
    size_t h_stage_bytes = 0;
    hipEvent_t stage_done = nullptr;         // recorded after the last copy out of the staging (context-owned)
    bool stage_busy = false;                 // such a copy may still be queued
    bool pending = false;                    // deflation_begin queued a histogram
    hipStream_t side = nullptr;              // the cluster setup runs here, beside the hierarchy's (context-owned)
    hipEvent_t side_done = nullptr;          // recorded on side at the end of that setup (context-owned)
    unsigned gen = 0;                        // bumped by every setup (single, block or batch)
    DeflationHost h;
};
// deflation_begin queues the diagonal histogram and its copy to the host;
// deflation_finish (after other work has been queued) decides and builds the
// correction. D.nc == 0 afterwards if there is none. records: also find the
// cells just outside the clusters that the balancing form (opt.deflate 2)
// needs.
// With external_hist, deflation_begin only clears the histogram (D.h.bhist)
// and leaves it and the cells' bins (D.h.bbins) to the caller's pass over the
// fine level (k_to_f32_dinv_hist), after which deflation_hist_queued queues
// the copy; deflation_finish takes the histogram itself if that never came.
int deflation_begin(hipStream_t st, Dims d, const Stencil<double> &A, Deflation &D,
                    char *(*alloc)(void *, size_t, int), void *actx, bool external_hist = false);
int deflation_hist_queued(hipStream_t st, Deflation &D);
int deflation_finish(hipStream_t st, Dims d, const Stencil<double> &A, double ratio, Deflation &D,
                     char *(*alloc)(void *, size_t, int), void *actx, bool records = false);
void deflation_sums(hipStream_t st, const Deflation &D, const double *r, const double *v, double *Pextra,
                    const int *done);
// Batch: the instances (dims[3q..3q+2], concatenated in one index space) get
// their own thresholds and clusters; per CG iteration the per-instance extra
// dot partials go to D.pext. `never` points at a device int that stays 0.
// It runs in steps, so that the histograms' round trip to the host overlaps
// other queued work (the batch's hierarchy setup): begin queues the
// per-instance histograms and their copy, finish decides and builds. With
// external_hist, begin only clears the histograms (D.h.bhist, kHistBins per
// instance) and leaves them and the cells' bins (D.h.bbins, by batch index)
// to a pass of the caller's (kb_to_f32_dinv), after which
// deflation_hist_queued_batch queues the copy.
int deflation_begin_batch(hipStream_t st, int ninst, const int *dims, const Stencil<double> &A, double ratio,
                          Deflation &D, char *(*alloc)(void *, size_t, int), void *actx, bool external_hist = false);
int deflation_hist_queued_batch(hipStream_t st, Deflation &D);
int deflation_finish_batch(hipStream_t st, const Stencil<double> &A, Deflation &D,
                           char *(*alloc)(void *, size_t, int), void *actx);
void deflation_sums_batch(hipStream_t st, const Deflation &D, const double *r, const double *v, const int *never);
void deflation_scatter(hipStream_t st, const Deflation &D, float *z, const int *done);
void deflation_scatter(hipStream_t st, const Deflation &D, double *z, const int *done);
// The balancing form around the V-cycle of a single grid (after
// deflation_sums): deflation_pre puts r' = r - A Z c1 into r (saving what it
// overwrites), deflation_post restores r, turns c into c1 - E^-1 Z^T A y for
// the V-cycle's output y and rewrites the extra dot partials for
// z = y + Z c; deflation_scatter then adds Z c.
void deflation_pre(hipStream_t st, const Deflation &D, double *r, const int *done);
void deflation_post(hipStream_t st, const Deflation &D, const Stencil<double> &A, double *r, const float *y,
                    double *Pextra, const int *done);
void deflation_post(hipStream_t st, const Deflation &D, const Stencil<double> &A, double *r, const double *y,
                    double *Pextra, const int *done);
// The same for nr (2 or 4) interleaved right-hand sides ([cell][nr] vectors;
// Pextra gets c.sr for each, then c.sv for each).
int deflation_sums_block(hipStream_t st, const Deflation &D, int nr, const double *r, const double *v,
                         double *Pextra, const int *done);
int deflation_scatter_block(hipStream_t st, const Deflation &D, int nr, float *z, const int *done);

// compute() on a general symmetric count-CSR matrix (hmfem_csr.hip): host
// arrays in, x (host) in/out; device buffers from alloc(actx, bytes, slot).
int csr_solve(hipStream_t st, int n, const int *cnt, const int *col, const double *ele, const double *b, double *x,
              double rtol, int stop_rule, int maxit, char *(*alloc)(void *, size_t, int), void *actx, int *iterations,
              double *rel_residual);

size_t hierarchy_bytes(Dims fine, const hmfem_options &opt, const PlanHint *hint = nullptr);
// The coupling strength of a fine FP64 stencil (device), sampled every 16th
// cell; partial needs kAxisBlocks * 6 doubles of device memory.
constexpr int kAxisBlocks = 256;
int axis_strength(hipStream_t st, Dims d, const Stencil<double> &A, double *partial, PlanHint &hint);
// Level l of the hierarchy a cache holds (sizes, factors, FP64 coefficients to
// host memory); l < 0: only the number of levels. Returns the level count.
int hierarchy_level_copy(hipStream_t st, const SetupCache &cache, int l, int *dims, int *fac, double *e, double *wx,
                         double *wy, double *wz);
// The dense inverse of that hierarchy's coarsest level (nc x nc, host); returns nc.
int hierarchy_coarse_inverse(hipStream_t st, const SetupCache &cache, double *inv);
size_t cg_partial_count(long n);
int solve_structured(hipStream_t st, Dims d, const Stencil<double> &A, const double *b, double *x,
                     const hmfem_options &opt, char *arena, size_t arena_bytes, CgBuffers &B, CgScalars *h_sc,
                     int *d_info, hmfem_status *out, hipEvent_t ev[4], ProfileSink *sink,
                     std::vector<char> &stage, GraphCache *graph, long lds_limit, Deflation *defl = nullptr,
                     char *(*dalloc)(void *, size_t, int) = nullptr, void *dctx = nullptr,
                     SetupCache *cache = nullptr, const PlanHint *hint = nullptr);

// batched solve (hmfem_batch.hip) and batch layout / kernels (hmfem_assemble.hip)
int batch_solve(hipStream_t st, int ninst, const int *dims, const Stencil<double> &A0, const double *b, double *x,
                const hmfem_options &opt_in, char *(*alloc)(void *, size_t), void *alloc_ctx, CgScalars *h_sc,
                GraphCache *graph, std::vector<char> &stage, hipEvent_t ev[4], int *iters_out,
                double *relres_out, double *done_ms_out, int *status_out, double *setup_ms, double *solve_ms,
                Deflation *defl = nullptr, char *(*dalloc)(void *, size_t, int) = nullptr, void *dctx = nullptr,
                const int *inst_gs = nullptr, const double *inst_ocf = nullptr);
int batch_grids(hipStream_t st, int ninst, const int *dims, const double *lengths, double mu, const int bc[6],
                std::vector<char> &stage, void *d_grids, int *d_bmap, int *nblk_out);
size_t batch_grid_bytes(int ninst);

// A block solve's lanes (right-hand-side slots) refilled from a queue: when
// lane r has finished, done(r, iterations, rel_residual, status) takes its
// result out of X[.][r], then load(r) puts the next right-hand side into
// B[.][r] and X[.][r] (false: none left, the lane stays idle).
struct BlockFeed {
    std::function<void(int, int, double, int)> done;
    std::function<bool(int)> load;
};

// block solve of several right-hand sides (hmfem_block.hip): B, X interleaved [n][nr]
int block_solve(hipStream_t st, Dims d, const Stencil<double> &A, int nr, const double *Bk, double *Xk,
                const hmfem_options &opt_in, char *(*alloc)(void *, size_t), void *actx, long lds_limit,
                std::vector<char> &stage, int *iterations, double *rel_residual, int *status, double *setup_ms,
                double *solve_ms, Deflation *defl = nullptr, bool defl_setup = false,
                char *(*dalloc)(void *, size_t, int) = nullptr, void *dctx = nullptr, BlockFeed *feed = nullptr,
                int *block_iterations = nullptr);
size_t block_bytes(Dims fine, int nr, const hmfem_options &opt);
void block_put(hipStream_t st, long n, int nr, int r, const double *a, const double *add, double *Bk);
void block_get(hipStream_t st, long n, int nr, int r, const double *Bk, double *a);
int launch_batch_assemble(hipStream_t st, const void *d_grids, const int *d_bmap, int nblk, const double *Kx,
                          const double *Ky, const double *Kz, const double *const bval[6], double *e, double *wx,
                          double *wy, double *wz, double *b);
int launch_batch_flux(hipStream_t st, int ninst, const void *d_grids, const int *d_bmap, int nblk, const double *Kx,
                      const double *Ky, const double *Kz, const double *const bval[6], const double *p, ELE *v,
                      double *partial, double *out);

} // namespace hmfem
