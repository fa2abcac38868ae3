// Output writers of the reservoir step: pressure_reservoir.dat,
// flux_reservoir.dat (posprocessing_output.c:200-296) and campos.vts
// (vtk.c:8-76), byte for byte as the reference writes them, into the
// directory named by the global `dirname` (main.c:343 sets it from argv; an
// executable that defines `dirname` is read through ELF interposition, as
// `boundary`), or by hmfem_set_output_dir() in hosts without that global.
//
// The reference formats every line with fprintf in loops whose innermost
// index is k (stride nx*ny in memory). Here the outermost loop's range is cut
// into slabs that threads format with snprintf (the same conversions, so the
// same bytes) into their own buffers, written in order with one fwrite each.
#include "hmfem_gpu.h"

#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

extern "C" {
const char *dirname = nullptr;
int number_karst = 0;
}

namespace {

const char *g_outdir = nullptr;

// The output directory: hmfem_set_output_dir's, else the host's `dirname`.
// In a process without its own `dirname` variable the name binds to the C
// library's dirname() function (which precedes this library in the lookup
// order); that one is recognised by its object and ignored.
const char *out_dir()
{
    if (g_outdir) return g_outdir;
    Dl_info a{}, b{};
    if (dladdr((const void *)&dirname, &a) && dladdr((const void *)&printf, &b) && a.dli_fbase == b.dli_fbase)
        return ".";
    return dirname ? dirname : ".";
}

// Line formatting helpers: append one snprintf conversion to a string.
template <typename... A>
void add(std::string &out, const char *fmt, A... a)
{
    char line[256];
    const int k = snprintf(line, sizeof(line), fmt, a...);
    if (k <= 0) return;
    if ((size_t)k < sizeof(line)) {
        out.append(line, (size_t)k);
        return;
    }
    // a longer line (e.g. %.6lf of a huge coordinate): format it again in place
    const size_t at = out.size();
    out.resize(at + (size_t)k + 1);
    snprintf(&out[at], (size_t)k + 1, fmt, a...);
    out.resize(at + (size_t)k);
}

// Run body(lo, hi, out) over [0, count) in contiguous slabs on up to 16
// threads and return the slabs' text in order.
template <typename F>
std::vector<std::string> slabs(int count, F body)
{
    int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    nt = std::max(1, std::min(nt, count));
    std::vector<std::string> out(nt);
    std::vector<std::thread> th;
    std::atomic<bool> failed{false};
    // an exception (out of memory growing a slab's text) must not leave a
    // thread or cross the C interface: it is noted and reported after the join
    auto run = [&](int lo, int hi, std::string &o) {
        try {
            body(lo, hi, o);
        } catch (...) {
            failed = true;
        }
    };
    for (int t = 0; t < nt; t++) {
        const int lo = (int)((long)count * t / nt), hi = (int)((long)count * (t + 1) / nt);
        try {
            th.emplace_back([&, t, lo, hi] { run(lo, hi, out[t]); });
        } catch (...) {   // no thread to be had: this slab on the calling thread
            run(lo, hi, out[t]);
        }
    }
    for (auto &x : th) x.join();
    if (failed) {
        fprintf(stderr, "hmfem: out of memory formatting an output file\n");
        exit(1);
    }
    return out;
}

FILE *open_out(const char *name, char *path, size_t cap)
{
    snprintf(path, cap, "%s/%s", out_dir(), name);
    return fopen(path, "w");
}

void write_all(FILE *f, const std::vector<std::string> &parts)
{
    for (const auto &p : parts) fwrite(p.data(), 1, p.size(), f);
}

inline int pidx(int i, int j, int k, const Dim &N) { return i + j * N.x + k * N.x * N.y; }

} // namespace

extern "C" void hmfem_set_output_dir(const char *dir) { g_outdir = dir; }

extern "C" void save_pressure(double *x, Dim N, grid h)
{
    char path[2048];
    FILE *f = open_out("pressure_reservoir.dat", path, sizeof(path));
    if (!f) {
        printf("Unable to open file %s. Terminating run... \n", path);
        exit(0);
    }
    auto parts = slabs(N.x, [&](int lo, int hi, std::string &o) {
        for (int i = lo; i < hi; i++) {
            for (int j = 0; j < N.y; j++) {
                for (int k = 0; k < N.z; k++)
                    add(o, "%.6lf %.6lf %.6lf %1.9e\n", h.x * (i + 0.5), h.y * (j + 0.5), h.z * (k + 0.5),
                        x[pidx(i, j, k, N)]);
                o.append(" \n");
            }
            o.append(" \n");
        }
    });
    write_all(f, parts);
    fclose(f);
}

extern "C" void save_flux(ELE *v, Dim N, grid h)
{
    char path[2048];
    FILE *f = open_out("flux_reservoir.dat", path, sizeof(path));
    if (!f) {
        printf("Unable to open file %s. Terminating run... \n", path);
        exit(0);
    }
    auto parts = slabs(N.x, [&](int lo, int hi, std::string &o) {
        for (int i = lo; i < hi; i++)
            for (int j = 0; j < N.y; j++)
                for (int k = 0; k < N.z; k++) {
                    const ELE &c = v[pidx(i, j, k, N)];
                    add(o, "%.6lf %.6lf %.6lf %1.9e %1.9e %1.9e \n", h.x * (i + 0.5), h.y * (j + 0.5),
                        h.z * (k + 0.5), 0.5 * (c.r - c.l), 0.5 * (c.u - c.d), 0.5 * (c.t - c.b));
                }
    });
    write_all(f, parts);
    fclose(f);
}

extern "C" int vts_StructuredGrid_file(ELE *v, double *x, Dim N, grid h)
{
    char path[2048];
    FILE *f = open_out("campos.vts", path, sizeof(path));
    if (!f) {
        printf("Unable to open file %s.\n", path);
        return -1;   // the reference goes on writing to the null stream
    }
    std::string head;
    head.append("<?xml version=\"1.0\"?>\n");
    head.append("<VTKFile type=\"StructuredGrid\" version=\"0.1\">\n");
    add(head, "<StructuredGrid WholeExtent=\"0 %d 0 %d 0 %d\">\n", N.x - 1, N.y - 1, N.z - 1);
    add(head, "<Piece Extent=\"0 %d 0 %d 0 %d\">\n", N.x - 1, N.y - 1, N.z - 1);
    head.append("<PointData  Vectors=\"Velocity\" Scalars=\"pressure\">\n");
    head.append("<DataArray type=\"Float32\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    fwrite(head.data(), 1, head.size(), f);
    // k outermost, i innermost: index order
    write_all(f, slabs(N.z, [&](int lo, int hi, std::string &o) {
        for (int k = lo; k < hi; k++)
            for (int j = 0; j < N.y; j++)
                for (int i = 0; i < N.x; i++) {
                    const ELE &c = v[pidx(i, j, k, N)];
                    add(o, "%1.8e %1.8e %1.8e\n", (0.5) * (c.r - c.l), (0.5) * (c.u - c.d), (0.5) * (c.t - c.b));
                }
    }));
    std::string mid;
    mid.append("</DataArray>\n");
    mid.append("<DataArray type=\"Float32\" Name=\"pressure\" NumberOfComponents=\"1\" format=\"ascii\">\n");
    fwrite(mid.data(), 1, mid.size(), f);
    write_all(f, slabs(N.z, [&](int lo, int hi, std::string &o) {
        for (int k = lo; k < hi; k++)
            for (int j = 0; j < N.y; j++)
                for (int i = 0; i < N.x; i++) add(o, "%1.8e\n", x[pidx(i, j, k, N)]);
    }));
    std::string pts;
    pts.append("</DataArray>\n");
    pts.append("</PointData>\n");
    pts.append("<CellData>\n");
    pts.append("</CellData>\n");
    pts.append("<Points>\n");
    pts.append("<DataArray type=\"Float32\" NumberOfComponents=\"3\">\n");
    fwrite(pts.data(), 1, pts.size(), f);
    write_all(f, slabs(N.z, [&](int lo, int hi, std::string &o) {
        for (int k = lo; k < hi; k++) {
            const double z_pos = h.z * (k + 0.5);
            for (int j = 0; j < N.y; j++) {
                const double y_pos = h.y * (j + 0.5);
                for (int i = 0; i < N.x; i++) add(o, "%.4f %.4f %.4f \n", h.x * (i + 0.5), y_pos, z_pos);
            }
        }
    }));
    std::string tail;
    tail.append("</DataArray>\n");
    tail.append("</Points>\n");
    tail.append("</Piece>\n");
    tail.append("</StructuredGrid>\n");
    tail.append("</VTKFile>\n");
    fwrite(tail.data(), 1, tail.size(), f);
    fclose(f);
    return 0;
}

namespace {

// One conduit's lines (posprocessing_output.c:453-528, :571-655): the cell
// centre of step i along the conduit's direction and a value.
template <typename V>
void conduit_lines(std::string &o, const DimK &M, grid h, V value)
{
    int dx = 0, dy = 0, dz = 0;
    switch (M.pos) {
    case 1: dx = 1; break;
    case 2: dy = 1; break;
    case 3: dz = 1; break;
    case 4: dx = 1; dy = 1; break;
    case 5: dx = 1; dy = -1; break;
    case 6: dx = -1; dy = 1; break;
    case 7: dx = -1; dy = -1; break;
    default: return;   // the reference writes no line for other directions
    }
    for (int i = 0; i < M.s; i++)
        add(o, "%.6lf %.6lf %.6lf %1.9e\n", h.x * ((M.sxi + dx * i) + 0.5), h.y * ((M.syi + dy * i) + 0.5),
            h.z * ((M.szi + dz * i) + 0.5), value(i));
}

template <typename V>
void save_conduits(const char *stem, DimK *M, grid h, V value)
{
    for (int j = 0; j < number_karst; j++) {
        char name[64], path[2048];
        snprintf(name, sizeof(name), "%s_%i.dat", stem, j);
        FILE *f = open_out(name, path, sizeof(path));
        if (!f) {
            printf("Unable to open file %s. Terminating run... \n", path);
            exit(0);
        }
        std::string o;
        conduit_lines(o, M[j], h, [&](int i) { return value(i, j); });
        fwrite(o.data(), 1, o.size(), f);
        fclose(f);
    }
}

} // namespace

extern "C" void save_pressure_1d(double **x, DimK *M, double *hs, grid h)
{
    (void)hs;
    save_conduits("pressure_karst", M, h, [&](int i, int j) { return x[i][j]; });
}

extern "C" void save_flux_1d(ELEK *uc, DimK *M, double *hs, grid h)
{
    (void)hs;
    save_conduits("flux_karst", M, h, [&](int i, int j) { return 0.5 * (uc->r[i][j] - uc->l[i][j]); });
}
