// The AMG hierarchy shared by the single-RHS and block solvers: level
// storage, the arena it is carved from, and the host entry points of its
// setup and of the single-workgroup tail kernel.
#pragma once

#include "hmfem_stencil.h"

#include <vector>

namespace hmfem {

// Descriptor of one level run by the tail kernel.
template <typename T>
struct TailLevel {
    OpArgs<T> o;
    T *f, *u, *t, *r, *w;
    int n, fx, fy, fz;
    int lds_off;   // byte offset of this level's copy in LDS (LDS mode), -1 otherwise
};

// Block Gauss-Seidel tiles: kGsTx x kGsTy columns, one per thread.
constexpr int kGsTx = 32, kGsTy = 8, kGsTz = 8;
constexpr int kGsCz = 4;                                   // cells of a tile column per thread
constexpr int kGsThreads = kGsTx * kGsTy * (kGsTz / kGsCz); // threads of a smoother block

// The x/y ring of one tile plane: rows hy = 0 and hy = kGsTy + 1 (kGsTx + 2
// cells each) and columns hx = 0 and hx = kGsTx + 1 of rows 1..kGsTy.
constexpr int kGsRing = 2 * (kGsTx + 2) + 2 * kGsTy;

// LDS coordinates of ring cell q (0 <= q < kGsTz * kGsRing).
__device__ __forceinline__ void gs_ring_cell(int q, int &hx, int &hy, int &hz)
{
    hz = q / kGsRing;
    const int r = q - hz * kGsRing;
    if (r < kGsTx + 2) { hx = r; hy = 0; }
    else if (r < 2 * (kGsTx + 2)) { hx = r - (kGsTx + 2); hy = kGsTy + 1; }
    else { const int r2 = r - 2 * (kGsTx + 2); hy = 1 + (r2 >> 1); hx = (r2 & 1) ? kGsTx + 1 : 0; }
}

// A reduction launch: G blocks, each over a contiguous chunk of cells.
struct Chunking {
    int G, chunk;
};

constexpr int kTailThreads = 1024;
constexpr int kMaxTailLevels = 16;
constexpr int kMaxTailOps = 4096;
constexpr int kTailLdsOps = 256;   // op lists up to this length are copied into LDS

struct Hierarchy {
    int nlev = 0;
    Level lev[kMaxLevels];
    double *inv = nullptr;   // dense inverse of the coarsest level
    int nc = 0;
    int prec = HMFEM_PREC_FP32;
    int tail0 = -1;          // first level run by the tail kernel (-1: none)
    void *tail_levels = nullptr;
    int4 *tail_ops = nullptr;
    int tail_nops = 0;
    int tail_lds_bytes = 0;    // dynamic LDS of the tail kernel (0: global mode)
    int tail_lds_inv = -1;     // LDS offset of the dense inverse (-1: global mode; kTailInvGlobal: levels in
                               // LDS, the inverse read from global memory)
    // fine level stored in FP16 (HMFEM_PREC_FP16): scaled couplings and scale
    __half *f16x = nullptr, *f16y = nullptr, *f16z = nullptr;
    float *f16sq = nullptr;
};

constexpr int kTailInvGlobal = -2;

// The tail runs as the LDS kernel (its levels fit in LDS and its op list in
// the kernel's static array).
inline bool tail_in_lds(const Hierarchy &H)
{
    return (H.tail_lds_inv >= 0 || H.tail_lds_inv == kTailInvGlobal) && H.tail_nops <= kTailLdsOps;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Carve device memory out of a single allocation.
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

template <typename T>
inline OpArgs<T> opargs(const Level &L);
template <>
inline OpArgs<float> opargs<float>(const Level &L)
{
    OpArgs<float> o;
    o.e = L.s32.e; o.wx = L.s32.wx; o.wy = L.s32.wy; o.wz = L.s32.wz; o.di = L.s32.di;
    o.nx = L.d.nx; o.ny = L.d.ny; o.nz = L.d.nz;
    return o;
}
template <>
inline OpArgs<double> opargs<double>(const Level &L)
{
    OpArgs<double> o;
    o.e = L.s64.e; o.wx = L.s64.wx; o.wy = L.s64.wy; o.wz = L.s64.wz; o.di = L.s64.di;
    o.nx = L.d.nx; o.ny = L.d.ny; o.nz = L.d.nz;
    return o;
}

// Build the hierarchy from a fine FP64 stencil (device), inside `ar`.
// dh: a conduit-correction setup begun with external_hist, whose histogram
// the fine level's FP32 conversion pass then takes (and queues its copy).
struct Deflation;
int build_hierarchy(hipStream_t st, Dims fine, const Stencil<double> &s0, const hmfem_options &opt, Arena &ar,
                    Hierarchy &H, int *d_info, std::vector<char> &stage, long lds_limit,
                    const PlanHint *hint = nullptr, Deflation *dh = nullptr);

// One tail cycle of an FP32 hierarchy in LDS mode on a right-hand side read
// from ftop[s * stride] (top tail level), result to utop[s * stride].
// Returns HMFEM_ERR_ARG if this hierarchy's tail cannot run that way.
int launch_tail_strided(hipStream_t st, const Hierarchy &H, double omega, double ocf, const float *ftop,
                        float *utop, int stride, const int *done);

} // namespace hmfem
