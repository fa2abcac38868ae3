// The reference's 1-D karst conduit model on the GPU: for every conduit of a
// set, in one launch (one workgroup per conduit),
//   permeabilidade_karst  permeabilities.c:795-867   conduit transmissibilities
//   source_term1d         set_source_term.c:256-414  exchange, breccia and end terms
//   solve_linsys_1d       solve_linearsystem_AMG.c:200-291 (the reference hands
//                         the tridiagonal system to its AMG-PCG; here it is
//                         solved exactly by parallel cyclic reduction in LDS)
// and fluxo1d (flux_reconstruction.c:322-496) in a second kernel. Conduit j's
// cells are entries [off_j, off_j + m_j) of every per-cell array.
//
// The transmissibilities and right-hand sides follow the reference's FP64
// expressions with contraction off, so they match it bit for bit; the
// solve's rounding is PCR's own.
#include "hmfem_internal.h"

namespace hmfem {

__device__ __forceinline__ double hmean(double a, double b)
{
#pragma clang fp contract(off)
    return 2 * a * b / (a + b);
}

// Reservoir face between the cell s just outside a conduit end and the end
// cell sa: the transmissibility the reference reads from perm[s]
// (hm(K_s, K_sa) / (h^2 mu), permeabilities.c:216-293), the flux
// h T (p_s - p_sa) and the K-weighted mean pressure of the two cells
// (set_source_term.c:300-392, flux_reconstruction.c:368-470).
__device__ __forceinline__ void robin_face(const ConduitArgs &a, int axis, long s, long sa, double &u, double &lw)
{
#pragma clang fp contract(off)
    const double *K = axis == 1 ? a.Kx : axis == 2 ? a.Ky : a.Kz;
    const double hf = axis == 1 ? a.hx : axis == 2 ? a.hy : a.hz;
    const double den = hf * hf * a.mu;
    const double ks = K[s], ka = K[sa];
    const double T = hmean(ks, ka) / den;
    const double ps = a.p[s], pa = a.p[sa];
    u = hf * T * (ps - pa);
    lw = (ks * ps + ka * pa) / (ks + ka);
}

__device__ __forceinline__ double robin_coef(double u, double C, double cont)
{
#pragma clang fp contract(off)
    double c;
    if (u < 0) c = -C * (cont - 1);
    else c = C * (cont - 1);
    return c;
}

// Thread 0 of conduit j's workgroup: the breccia terms (state ucR/ucL
// updated, as source_term1d does) and the end terms, added to f.
__device__ void conduit_rhs_serial(const ConduitArgs &a, const ConduitDesc &D, double *f)
{
#pragma clang fp contract(off)
    const double mu = a.mu, hs = D.hs, rho = a.rho;
    const int m = D.m;
    for (int q = 0; q < D.nbrec; q++) {
        const int b = D.boff + q;
        const int s = a.brec_cell[b];
        const double cR = hmean(D.Kc, D.Kc);
        const double Cb = a.brec_C[b], beta = a.brec_beta[b];
        const double uRo = a.ucR[b], uLo = a.ucL[b];
        double R;
        if (uRo < 0) R = -Cb * beta;
        else R = Cb * beta;
        const double pcs = a.pc_old[D.off + s], pcs1 = a.pc_old[D.off + s + 1];
        const double uR = (cR / (mu * hs)) * (pcs - pcs1 - rho * uRo * uRo * R);
        const double uL = (cR / (mu * hs)) * (pcs1 - pcs + rho * uLo * uLo * R);
        a.ucR[b] = uR;
        a.ucL[b] = uL;
        f[s] = f[s] + (cR / (mu * hs)) * (rho * uR * uR * R);
        f[s + 1] = f[s + 1] - (cR / (mu * hs)) * (rho * uL * uL * R);
    }
    if (D.bc_l == 0) f[0] = f[0] + (a.kl[D.off]) * D.cont_l;
    if (D.bc_l == 1) f[0] = f[0] + D.cont_l / (hs);
    if (D.bc_l == 3) {
        double u, lw;
        robin_face(a, D.axis, D.rs_l, D.rsa_l, u, lw);
        const double c = robin_coef(u, D.ccont_l, D.cont_l);
        const double font = -lw + rho * u * u * c;
        f[0] = f[0] - 2 * D.Kc / (mu * hs * hs) * font;
    }
    if (D.bc_r == 0) f[m - 1] = f[m - 1] + a.kr[D.off + m - 1] * D.cont_r;
    if (D.bc_r == 1) f[m - 1] = f[m - 1] - D.cont_r / (hs);
    if (D.bc_r == 3) {   // (the host refuses a Robin right end along z, which the reference leaves undefined)
        double u, lw;
        robin_face(a, D.axis, D.rs_r, D.rsa_r, u, lw);
        const double c = robin_coef(u, D.ccont_r, D.cont_r);
        const double font = -lw + rho * u * u * c;
        f[m - 1] = f[m - 1] - font * 2 * D.Kc / (mu * hs * hs);
    }
}

// One pass of conduit() (karst.c:8-42) for every conduit: blockIdx.x = j.
// Dynamic LDS: 8 * cap doubles (two copies of the four tridiagonal
// diagonals); conduits longer than cap are solved by thread 0 (Thomas) in
// the global scratch a.cp.
__global__ void __launch_bounds__(kBlock) k_conduit_pass(ConduitArgs a, int cap)
{
#pragma clang fp contract(off)
    extern __shared__ double lds[];
    const ConduitDesc D = a.desc[blockIdx.x];
    const int m = D.m, o = D.off;
    const double mu = a.mu, hs = D.hs;
    const bool pcr = m <= cap;
    double *f = pcr ? lds + 3 * cap : a.fbuf + o;   // right-hand side (first copy of d)
    // permeabilidade_karst and the exchange term of source_term1d
    for (int i = threadIdx.x; i < m; i += kBlock) {
        double kl = 0.0, kr = 0.0;
        if (i < m - 1) kr = hmean(D.Kc, D.Kc) / (hs * hs * mu);
        if (i > 0) kl = hmean(D.Kc, D.Kc) / (hs * hs * mu);
        if (i == 0) {
            if (D.bc_l == 0) kl = 2 * (D.Kc) / (hs * hs * mu);
            if (D.bc_l == 1) kl = 0;
            if (D.bc_l == 3) kl = 2 * D.Kc / (mu * hs * hs);
        }
        if (i == m - 1) {
            if (D.bc_r == 0) kr = 2 * (D.Kc) / (hs * hs * mu);
            if (D.bc_r == 1) kr = 0;
            if (D.bc_r == 3) kr = 2 * D.Kc / (mu * hs * hs);
        }
        const double kc = kl + kr + D.KI_c / mu;
        a.kl[o + i] = kl;
        a.kr[o + i] = kr;
        a.kc[o + i] = kc;
        f[i] = (D.KI_c * a.p[a.res_cell[o + i]]) / mu;
    }
    __syncthreads();
    if (threadIdx.x == 0) conduit_rhs_serial(a, D, f);
    __syncthreads();
    if (!pcr) {
        // Thomas on the tridiagonal (-kl, kc, -kr); cp and f in global memory
        if (threadIdx.x == 0) {
            double *cp = a.cp + o;
            for (int i = 0; i < m; i++) {
                const double av = i > 0 ? -a.kl[o + i] : 0.0;
                const double bv = a.kc[o + i];
                const double cv = i < m - 1 ? -a.kr[o + i] : 0.0;
                const double den = i > 0 ? bv - av * cp[i - 1] : bv;
                cp[i] = cv / den;
                f[i] = (i > 0 ? f[i] - av * f[i - 1] : f[i]) / den;
            }
            for (int i = m - 1; i >= 0; i--) {
                const double x = i < m - 1 ? f[i] - cp[i] * a.pc[o + i + 1] : f[i];
                a.pc[o + i] = x;
            }
        }
        return;
    }
    // parallel cyclic reduction: after the step with stride d every row
    // couples only to rows i -/+ 2d; log2(m) steps leave diagonal rows
    double *A0 = lds, *B0 = lds + cap, *C0 = lds + 2 * cap, *D0 = lds + 3 * cap;
    double *A1 = lds + 4 * cap, *B1 = lds + 5 * cap, *C1 = lds + 6 * cap, *D1 = lds + 7 * cap;
    for (int i = threadIdx.x; i < m; i += kBlock) {
        A0[i] = i > 0 ? -a.kl[o + i] : 0.0;
        B0[i] = a.kc[o + i];
        C0[i] = i < m - 1 ? -a.kr[o + i] : 0.0;
    }
    __syncthreads();
    for (int d = 1; d < m; d *= 2) {
        for (int i = threadIdx.x; i < m; i += kBlock) {
            double an = 0.0, bn = B0[i], cn = 0.0, dn = D0[i];
            if (i - d >= 0) {
                const double k1 = A0[i] / B0[i - d];
                an = -A0[i - d] * k1;
                bn = bn - C0[i - d] * k1;
                dn = dn - D0[i - d] * k1;
            }
            if (i + d < m) {
                const double k2 = C0[i] / B0[i + d];
                cn = -C0[i + d] * k2;
                bn = bn - A0[i + d] * k2;
                dn = dn - D0[i + d] * k2;
            }
            A1[i] = an; B1[i] = bn; C1[i] = cn; D1[i] = dn;
        }
        __syncthreads();
        double *t;
        t = A0; A0 = A1; A1 = t;
        t = B0; B0 = B1; B1 = t;
        t = C0; C0 = C1; C1 = t;
        t = D0; D0 = D1; D1 = t;
    }
    for (int i = threadIdx.x; i < m; i += kBlock) a.pc[o + i] = D0[i] / B0[i];
}

// fluxo1d (flux_reconstruction.c:322-496) of x = a.pc: interior faces per
// cell, then (thread 0) the breccia faces and the ends, which overwrite.
__global__ void __launch_bounds__(kBlock) k_conduit_flux(ConduitArgs a, double *ucl, double *ucr)
{
#pragma clang fp contract(off)
    const ConduitDesc D = a.desc[blockIdx.x];
    const int m = D.m, o = D.off;
    const double mu = a.mu, hs = D.hs, rho = a.rho;
    const double *x = a.pc + o;
    for (int i = threadIdx.x; i < m; i += kBlock) {
        if (i < m - 1) ucr[o + i] = hs * a.kr[o + i] * (x[i] - x[i + 1]);
        if (i > 0) ucl[o + i] = hs * a.kl[o + i] * (x[i] - x[i - 1]);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int q = 0; q < D.nbrec; q++) {
        const int b = D.boff + q;
        const double Cb = a.brec_C[b], beta = a.brec_beta[b];
        const double uRo = a.ucR[b], uLo = a.ucL[b];
        double c;
        if (uRo < 0) c = -Cb * beta;
        else c = Cb * beta;
        const int s = a.brec_cell[b];
        const double cR = hmean(D.Kc, D.Kc);
        ucr[o + s] = (cR / (mu * hs)) * (x[s] - x[s + 1] - 0.5 * rho * uRo * uRo * c);
        ucl[o + s + 1] = (cR / (mu * hs)) * (x[s + 1] - x[s] + 0.5 * rho * uLo * uLo * c);
    }
    if (D.bc_l == 0) ucl[o] = hs * a.kl[o] * (x[0] - D.cont_l);
    if (D.bc_l == 1) ucl[o] = -D.cont_r;   // the reference reports cont_karst[1][j] here
    if (D.bc_l == 3) {
        double u, lw;
        robin_face(a, D.axis, D.rs_l, D.rsa_l, u, lw);
        const double c = robin_coef(u, D.ccont_l, D.cont_l);
        ucl[o] = (2 * D.Kc) / (mu * hs) * (x[0] - lw + rho * u * u * c);
    }
    if (D.bc_r == 0) ucr[o + m - 1] = hs * a.kr[o + m - 1] * (x[m - 1] - D.cont_r);
    if (D.bc_r == 1) ucr[o + m - 1] = D.cont_r;
    if (D.bc_r == 3) {
        double u, lw;
        robin_face(a, D.axis, D.rs_r, D.rsa_r, u, lw);
        const double c = robin_coef(u, D.ccont_r, D.cont_r);
        ucr[o + m - 1] = (2 * D.Kc) / (mu * hs) * (x[m - 1] - lw + rho * u * u * c);
    }
}

// reservoir(): pc_aux[s] = pc_old[conduit cell] on the conduits' cells
// (reservoir.c:19-45; where conduits overlap, the last one in the
// reference's loop order was chosen by the host).
__global__ void __launch_bounds__(kBlock)
k_conduit_scatter(int count, const int *__restrict__ res, const int *__restrict__ cell, const double *pc_old,
                  double *pc_aux)
{
    const int u = blockIdx.x * kBlock + threadIdx.x;
    if (u < count) pc_aux[res[u]] = pc_old[cell[u]];
}

// erro_relativo (main.c:106-119) of p against p_old, one partial pair per
// block, and p_old = p.
__global__ void __launch_bounds__(kBlock)
k_rel_change(long n, const double *__restrict__ p, double *__restrict__ p_old, double *partial)
{
    double s1 = 0.0, s2 = 0.0;
    const long G = gridDim.x;
    for (long s = (long)blockIdx.x * kBlock + threadIdx.x; s < n; s += G * kBlock) {
        const double y = p[s], d = y - p_old[s];
        s1 += d * d;
        s2 += y * y;
        p_old[s] = y;
    }
    __shared__ double w1[kBlock / 64], w2[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
    }
    if (threadIdx.x % 64 == 0) { w1[threadIdx.x / 64] = s1; w2[threadIdx.x / 64] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; q++) { s1 += w1[q]; s2 += w2[q]; }
        partial[2 * blockIdx.x] = s1;
        partial[2 * blockIdx.x + 1] = s2;
    }
}

__global__ void __launch_bounds__(kBlock) k_rel_change_final(int G, const double *partial, double *out)
{
    double s1 = 0.0, s2 = 0.0;
    for (int q = threadIdx.x; q < G; q += kBlock) { s1 += partial[2 * q]; s2 += partial[2 * q + 1]; }
    __shared__ double w1[kBlock / 64], w2[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
    }
    if (threadIdx.x % 64 == 0) { w1[threadIdx.x / 64] = s1; w2[threadIdx.x / 64] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; q++) { s1 += w1[q]; s2 += w2[q]; }
        out[0] = sqrt(s1) / sqrt(s2);
    }
}

// ------------------------------------------------------------ launchers

int conduit_pass(hipStream_t st, const ConduitArgs &a, int count, int mmax)
{
    const int cap = mmax < kConduitLdsCells ? mmax : kConduitLdsCells;
    static bool attr = false;
    if (!attr) {
        HMFEM_CHECK(hipFuncSetAttribute((const void *)k_conduit_pass, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        8 * kConduitLdsCells * (int)sizeof(double)));
        attr = true;
    }
    hipLaunchKernelGGL(k_conduit_pass, dim3(count), dim3(kBlock), 8 * cap * sizeof(double), st, a, cap);
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

int conduit_flux(hipStream_t st, const ConduitArgs &a, int count, double *ucl, double *ucr)
{
    hipLaunchKernelGGL(k_conduit_flux, dim3(count), dim3(kBlock), 0, st, a, ucl, ucr);
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

int conduit_scatter(hipStream_t st, int count, const int *res, const int *cell, const double *pc_old,
                    double *pc_aux)
{
    if (count <= 0) return HMFEM_OK;
    hipLaunchKernelGGL(k_conduit_scatter, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, st, count, res, cell,
                       pc_old, pc_aux);
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

int rel_change(hipStream_t st, long n, const double *p, double *p_old, double *partial, double *out)
{
    const int G = kRelChangeBlocks;
    hipLaunchKernelGGL(k_rel_change, dim3(G), dim3(kBlock), 0, st, n, p, p_old, partial);
    hipLaunchKernelGGL(k_rel_change_final, dim3(1), dim3(kBlock), 0, st, G, partial, out);
    HMFEM_CHECK(hipGetLastError());
    return HMFEM_OK;
}

} // namespace hmfem
