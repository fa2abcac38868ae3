/*
 * Seeded permeability-field generators for the Darcy pressure solve.
 *
 * The reference chooses its permeability in set_perm_faces()
 * (permeabilities.c:9-55); its only reachable branch is the homogeneous
 * field of perm_homogeneous() (permeabilities.c:776-789), and its readers for
 * heterogeneous files are unreachable (main.c:341 never sets the file name).
 * These generators stand in for those readers: every field is a pure function
 * of (grid, parameters, seed), written in portable C with a counter-based
 * PRNG so that the same inputs give the same bits on every x86-64 host that
 * runs this library with the same libm.
 *
 * Layout: K[s] with s = i + j*nx + k*nx*ny (x fastest), as pindex()
 * (solve_linearsystem_AMG.c:15-20).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/hmfem_fields.h"

/* SplitMix64: one 64-bit state, one output per step. */
static uint64_t sm64_next(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Uniform double in [0,1) from the top 53 bits. */
static double sm64_uniform(uint64_t *state)
{
    return (double)(sm64_next(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* Standard normal by Box-Muller; one value per call, the sine branch unused. */
static double sm64_normal(uint64_t *state)
{
    double u1 = sm64_uniform(state);
    double u2 = sm64_uniform(state);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Seed mixing so that nearby seeds give unrelated streams. */
static uint64_t stream_seed(uint64_t seed, uint64_t stream)
{
    uint64_t s = seed * 0xD1B54A32D192ED03ull + stream * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    (void)sm64_next(&s);
    return s;
}

int hmfem_field_homogeneous(int nx, int ny, int nz, double k, double *K)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !K) return -1;
    size_t n = (size_t)nx * ny * nz;
    for (size_t s = 0; s < n; s++) K[s] = k;
    return 0;
}

/*
 * Gaussian field with squared-exponential covariance exp(-r^2/l^2) (r, l in
 * cells), standardised to sample mean 0 and sample variance 1, then
 * K = exp(mean_ln + sigma * g).
 *
 * A white-noise grid padded by R cells on every side is convolved with the
 * separable kernel w(x) = exp(-2 x^2 / l^2), |x| <= R, R = ceil(2 l): the
 * autocorrelation of that kernel is proportional to exp(-x^2 / l^2) per axis,
 * and the 3-D kernel is the product of the three 1-D ones, so three 1-D
 * passes give the 3-D field. Padding removes any wrap-around.
 */
int hmfem_field_lognormal(int nx, int ny, int nz, double mean_ln, double sigma,
                          double corr_len, uint64_t seed, double *K)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !K || !(corr_len > 0.0)) return -1;
    int R = (int)ceil(2.0 * corr_len);
    int px = nx + 2 * R, py = ny + 2 * R, pz = nz + 2 * R;
    int taps = 2 * R + 1;
    double *w = (double *)malloc(sizeof(double) * taps);
    if (!w) return -2;
    for (int t = -R; t <= R; t++) w[t + R] = exp(-2.0 * (double)t * (double)t / (corr_len * corr_len));

    size_t npad = (size_t)px * py * pz;
    double *a = (double *)malloc(sizeof(double) * npad);
    if (!a) { free(w); return -2; }
    /* One PRNG stream per padded z-plane keeps the draw order fixed. */
    for (int k = 0; k < pz; k++) {
        uint64_t st = stream_seed(seed, (uint64_t)k + 1);
        double *plane = a + (size_t)k * px * py;
        for (size_t q = 0; q < (size_t)px * py; q++) plane[q] = sm64_normal(&st);
    }
    /* x pass: (px,py,pz) -> (nx,py,pz) */
    double *b = (double *)malloc(sizeof(double) * (size_t)nx * py * pz);
    if (!b) { free(w); free(a); return -2; }
    for (int k = 0; k < pz; k++)
        for (int j = 0; j < py; j++) {
            const double *row = a + ((size_t)k * py + j) * px;
            double *out = b + ((size_t)k * py + j) * nx;
            for (int i = 0; i < nx; i++) {
                double acc = 0.0;
                for (int t = 0; t < taps; t++) acc += w[t] * row[i + t];
                out[i] = acc;
            }
        }
    free(a);
    /* y pass: (nx,py,pz) -> (nx,ny,pz) */
    double *c = (double *)malloc(sizeof(double) * (size_t)nx * ny * pz);
    if (!c) { free(w); free(b); return -2; }
    for (int k = 0; k < pz; k++)
        for (int j = 0; j < ny; j++) {
            double *out = c + ((size_t)k * ny + j) * nx;
            for (int i = 0; i < nx; i++) out[i] = 0.0;
            for (int t = 0; t < taps; t++) {
                const double *row = b + ((size_t)k * py + j + t) * nx;
                double wt = w[t];
                for (int i = 0; i < nx; i++) out[i] += wt * row[i];
            }
        }
    free(b);
    /* z pass: (nx,ny,pz) -> (nx,ny,nz), written straight into K */
    size_t plane = (size_t)nx * ny;
    for (int k = 0; k < nz; k++) {
        double *out = K + (size_t)k * plane;
        for (size_t q = 0; q < plane; q++) out[q] = 0.0;
        for (int t = 0; t < taps; t++) {
            const double *src = c + (size_t)(k + t) * plane;
            double wt = w[t];
            for (size_t q = 0; q < plane; q++) out[q] += wt * src[q];
        }
    }
    free(c);
    free(w);
    /* Standardise (two passes, fixed order). */
    size_t n = plane * nz;
    double mean = 0.0;
    for (size_t s = 0; s < n; s++) mean += K[s];
    mean /= (double)n;
    double var = 0.0;
    for (size_t s = 0; s < n; s++) { double d = K[s] - mean; var += d * d; }
    var /= (double)n;
    double inv_sd = var > 0.0 ? 1.0 / sqrt(var) : 0.0;
    for (size_t s = 0; s < n; s++) K[s] = exp(mean_ln + sigma * (K[s] - mean) * inv_sd);
    return 0;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/*
 * Karst conduits: each conduit c (0-based) uses its own stream from
 * (seed, c). It is a polyline of 5..15 segments whose vertices follow a
 * random walk inside the box (step length uniform in [n/16, n/4] cells of the
 * longest axis, reflected at the walls). Each segment is traced by a voxel
 * traversal that steps across one face at a time, so a conduit is a
 * face-connected path; each traced cell is dilated by a cube of edge 1..3
 * cells, and the cells touched get the multiplier 10^U(4,6); where conduits
 * overlap the largest multiplier wins. K[s] is multiplied in place.
 */
int hmfem_field_conduits(int nx, int ny, int nz, int count, uint64_t seed, double *K)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !K || count < 0) return -1;
    size_t n = (size_t)nx * ny * nz;
    float *mult = (float *)calloc(n, sizeof(float));
    if (!mult) return -2;
    int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
    for (int c = 0; c < count; c++) {
        uint64_t st = stream_seed(seed ^ 0xA5A5A5A5u, (uint64_t)c + 1);
        int nseg = 5 + (int)(sm64_uniform(&st) * 11.0);        /* 5..15 */
        int thick = 1 + (int)(sm64_uniform(&st) * 3.0);         /* 1..3 */
        double factor = pow(10.0, 4.0 + 2.0 * sm64_uniform(&st));
        float f = (float)factor;
        double pxv = sm64_uniform(&st) * nx, pyv = sm64_uniform(&st) * ny, pzv = sm64_uniform(&st) * nz;
        int lo = -(thick - 1) / 2, hi = thick / 2;              /* cube offsets */
        for (int sgi = 0; sgi < nseg; sgi++) {
            double len = (double)nmax * (1.0 / 16.0 + (3.0 / 16.0) * sm64_uniform(&st));
            double dz = 2.0 * sm64_uniform(&st) - 1.0;
            double phi = 6.283185307179586 * sm64_uniform(&st);
            double rxy = sqrt(1.0 - dz * dz);
            double qx = pxv + len * rxy * cos(phi), qy = pyv + len * rxy * sin(phi), qz = pzv + len * dz;
            /* reflect into [0, n) */
            if (qx < 0) qx = -qx;
            if (qx >= nx) qx = 2.0 * nx - qx - 1e-9;
            if (qx < 0) qx = 0;
            if (qy < 0) qy = -qy;
            if (qy >= ny) qy = 2.0 * ny - qy - 1e-9;
            if (qy < 0) qy = 0;
            if (qz < 0) qz = -qz;
            if (qz >= nz) qz = 2.0 * nz - qz - 1e-9;
            if (qz < 0) qz = 0;
            /* Voxel traversal (Amanatides-Woo): every cell the segment
               crosses, stepping one face at a time, so the trace is a
               face-connected (7-point-stencil-connected) path. */
            double org[3] = {pxv, pyv, pzv}, dir[3] = {qx - pxv, qy - pyv, qz - pzv};
            int dim[3] = {nx, ny, nz}, cell[3], stp[3];
            double tmax[3], tdel[3];
            for (int a = 0; a < 3; a++) {
                cell[a] = clampi((int)floor(org[a]), 0, dim[a] - 1);
                if (dir[a] > 0) { stp[a] = 1; tdel[a] = 1.0 / dir[a]; tmax[a] = ((double)cell[a] + 1.0 - org[a]) / dir[a]; }
                else if (dir[a] < 0) { stp[a] = -1; tdel[a] = -1.0 / dir[a]; tmax[a] = (org[a] - (double)cell[a]) / -dir[a]; }
                else { stp[a] = 0; tdel[a] = 1e300; tmax[a] = 1e300; }
            }
            for (int guard = 0; guard < 4 * (nx + ny + nz) + 8; guard++) {
                for (int oz = lo; oz <= hi; oz++)
                    for (int oy = lo; oy <= hi; oy++)
                        for (int ox = lo; ox <= hi; ox++) {
                            int ii = cell[0] + ox, jj = cell[1] + oy, kk = cell[2] + oz;
                            if (ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz) continue;
                            size_t s = (size_t)ii + (size_t)jj * nx + (size_t)kk * nx * ny;
                            if (f > mult[s]) mult[s] = f;
                        }
                int a = tmax[0] < tmax[1] ? (tmax[0] < tmax[2] ? 0 : 2) : (tmax[1] < tmax[2] ? 1 : 2);
                if (tmax[a] > 1.0) break;
                int nxt = cell[a] + stp[a];
                if (nxt < 0 || nxt >= dim[a]) break;
                cell[a] = nxt;
                tmax[a] += tdel[a];
            }
            pxv = qx; pyv = qy; pzv = qz;
        }
    }
    for (size_t s = 0; s < n; s++)
        if (mult[s] > 0.0f) K[s] *= (double)mult[s];
    free(mult);
    return 0;
}

/* Reproducible per-instance parameter draws for a batch of instances. */
uint64_t hmfem_field_instance_seed(uint64_t seed, int instance)
{
    return stream_seed(seed, 0x100000000ull + (uint64_t)instance);
}

double hmfem_field_uniform(uint64_t *state) { return sm64_uniform(state); }
