// restart.cpp -- the reference's binary restart file (checkpoint/resume).
//
// Layout of State::save / State::load (reference src/State.cc:65-223),
// native endian, no padding:
//   int nx, ny, ngrid; double dx, x0, y0; int numPoints;
//   q:     for lev, for flux index: double
//   omega: for lev, i = 1..nx-1, j = 1..ny-1: double
//   f:     for point i: double fx, double fy
//   int timestep; double time
#include <cstdio>
#include <vector>

#include "../../include/ibpm_gpu.h"

namespace ibpmg {
void set_error(const char* fmt, ...);
}

extern "C" {

int ibpm_state_save(const char* path, int nx, int ny, int ngrid, double dx, double x0, double y0, int npts,
                    const double* q, const double* omega, const double* f, int timestep, double time) {
    FILE* fp = std::fopen(path, "wb");
    if (!fp) { ibpmg::set_error("cannot open %s for writing", path); return -1; }
    const long nq = (2L * nx * ny + nx + ny) * ngrid;
    const long nw = (long)(nx - 1) * (ny - 1) * ngrid;
    bool ok = true;
    ok = ok && std::fwrite(&nx, sizeof(int), 1, fp) == 1;
    ok = ok && std::fwrite(&ny, sizeof(int), 1, fp) == 1;
    ok = ok && std::fwrite(&ngrid, sizeof(int), 1, fp) == 1;
    ok = ok && std::fwrite(&dx, sizeof(double), 1, fp) == 1;
    ok = ok && std::fwrite(&x0, sizeof(double), 1, fp) == 1;
    ok = ok && std::fwrite(&y0, sizeof(double), 1, fp) == 1;
    ok = ok && std::fwrite(&npts, sizeof(int), 1, fp) == 1;
    ok = ok && std::fwrite(q, sizeof(double), nq, fp) == (size_t)nq;
    ok = ok && std::fwrite(omega, sizeof(double), nw, fp) == (size_t)nw;
    for (int i = 0; i < npts && ok; ++i) {
        ok = ok && std::fwrite(&f[i], sizeof(double), 1, fp) == 1;
        ok = ok && std::fwrite(&f[npts + i], sizeof(double), 1, fp) == 1;
    }
    ok = ok && std::fwrite(&timestep, sizeof(int), 1, fp) == 1;
    ok = ok && std::fwrite(&time, sizeof(double), 1, fp) == 1;
    std::fclose(fp);
    if (!ok) { ibpmg::set_error("short write to %s", path); return -1; }
    return 0;
}

// Reads the header (all pointers may be null) or, with buffers sized from a
// first header-only call, the whole file.  Returns 0, or -1 on a short read.
int ibpm_state_load(const char* path, int* nx, int* ny, int* ngrid, double* dx, double* x0, double* y0,
                    int* npts, double* q, double* omega, double* f, int* timestep, double* time) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) { ibpmg::set_error("cannot open %s", path); return -1; }
    int h[3];
    double d[3];
    int np = 0;
    bool ok = std::fread(h, sizeof(int), 3, fp) == 3 && std::fread(d, sizeof(double), 3, fp) == 3 &&
              std::fread(&np, sizeof(int), 1, fp) == 1;
    if (!ok) { std::fclose(fp); ibpmg::set_error("%s: truncated header", path); return -1; }
    if (nx) *nx = h[0];
    if (ny) *ny = h[1];
    if (ngrid) *ngrid = h[2];
    if (dx) *dx = d[0];
    if (x0) *x0 = d[1];
    if (y0) *y0 = d[2];
    if (npts) *npts = np;
    if (q && omega && f) {
        const long nq = (2L * h[0] * h[1] + h[0] + h[1]) * h[2];
        const long nw = (long)(h[0] - 1) * (h[1] - 1) * h[2];
        ok = std::fread(q, sizeof(double), nq, fp) == (size_t)nq;
        ok = ok && std::fread(omega, sizeof(double), nw, fp) == (size_t)nw;
        for (int i = 0; i < np && ok; ++i) {
            ok = ok && std::fread(&f[i], sizeof(double), 1, fp) == 1;
            ok = ok && std::fread(&f[np + i], sizeof(double), 1, fp) == 1;
        }
        int ts = 0;
        double tt = 0;
        ok = ok && std::fread(&ts, sizeof(int), 1, fp) == 1 && std::fread(&tt, sizeof(double), 1, fp) == 1;
        if (timestep) *timestep = ts;
        if (time) *time = tt;
    }
    std::fclose(fp);
    if (!ok) { ibpmg::set_error("%s: truncated file", path); return -1; }
    return 0;
}

}  // extern "C"
