// kf_grid.h -- a kielflow grid (three axes, grid_base.hpp:9-39) with the
// per-axis device tables the spectral kernels read.
#pragma once

#include "kf_common.h"
#include "kf_poisson.h"

namespace kf {

struct Grid {
    AxisDesc ax[3];
    int nx, ny, nz, nzc;
    long nr, nc;
    double d23_thr;
    double *k64[3] = {nullptr, nullptr, nullptr};
    float *k32[3] = {nullptr, nullptr, nullptr};
    double *q[3] = {nullptr, nullptr, nullptr};
    double *qzf = nullptr;
    int *imax = nullptr;  // GridView::imax
    // dealiased 1/S'(xi) per stretched axis (OP_partial_derivative.cpp:62-80);
    // nullptr on equidistant axes
    double *prof64[3] = {nullptr, nullptr, nullptr};
    float *prof32[3] = {nullptr, nullptr, nullptr};
    // penalization barrier cos^2(3 pi x/L) for x < -L/6, a function of the
    // x index only (rhs_standard.cpp:39-49)
    double *bar64 = nullptr;
    float *bar32 = nullptr;
    // cell volumes per axis, S'(xi_i) L / N (their product is the volume
    // element of field integrals on stretched grids)
    double *dv[3] = {nullptr, nullptr, nullptr};

    // scratch of the stretched OP_partial_derivative and of cfl_all,
    // allocated on first use and kept with the grid: three real fields, the
    // CFL spacing table and the per-block CFL maxima (device), and the
    // pinned host mirror of those maxima
    double *ws = nullptr;
    double *ws_host = nullptr;
    bool cfl_table = false;
    unsigned cfl_blocks() const { return (unsigned)((nr + 255) / 256); }
    double *workspace();

    Grid(const AxisDesc axes[3], hipStream_t s);
    ~Grid();
    Grid(const Grid &) = delete;
    Grid &operator=(const Grid &) = delete;

    bool stretched(int d) const { return ax[d].type != 1; }
    bool any_stretched() const { return stretched(0) || stretched(1) || stretched(2); }

    template <typename R>
    GridView<R> view() const
    {
        GridView<R> v;
        v.nx = nx;
        v.ny = ny;
        v.nz = nz;
        v.nzc = nzc;
        v.nr = nr;
        v.nc = nc;
        for (int d = 0; d < 3; ++d) {
            if constexpr (std::is_same<R, double>::value) v.k[d] = k64[d];
            else v.k[d] = k32[d];
            v.q[d] = q[d];
        }
        v.qzf = qzf;
        v.d23_thr = d23_thr;
        v.imax = imax;
        return v;
    }
    template <typename R>
    const R *prof(int d) const
    {
        if constexpr (std::is_same<R, double>::value) return prof64[d];
        else return prof32[d];
    }
    template <typename R>
    const R *bar() const
    {
        if constexpr (std::is_same<R, double>::value) return bar64;
        else return bar32;
    }
};

// smallest double t with sqrt(t) >= 2/3, so that for any s,
// sqrt(s) < 2./3.  <=>  s < t   (sqrt is correctly rounded)
double d23_threshold();

}  // namespace kf
