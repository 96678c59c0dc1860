// kielflow_io_compat.hpp -- kielflow's file functions (IO.hpp:69-127,
// IO.cpp:60-504, 1175-1338) over libkielflow_io (include/kielflow_io.h) and
// the field classes of kielflow_compat.hpp, with the reference's names,
// argument order and default paths.  Link with -lkielflow_io.
//
// Fields in the files are real (double[Nx][Ny][Nz], z fastest); the
// spectral save/load functions go through iFFT / FFT on the GPU, as the
// reference's go through FFTW.
#pragma once

#include <cstdio>
#include <string>

#include "../../include/kielflow_io.h"
#include "kielflow_compat.hpp"

namespace kf_compat {
inline void check_io(int code)
{
    if (code != KF_OK) throw std::runtime_error(std::string("kielflow_io: ") + kf_io_last_error());
}
}  // namespace kf_compat

// file_create (IO.cpp:60-86)
inline void file_create(const std::string &path) { kf_compat::check_io(kf_io_file_create(path.c_str(), nullptr)); }

// save_grid / load_grid (IO.cpp:173-249)
inline void save_grid(const grid &Target, const std::string &path = "./data/fields.h5")
{
    const kf_axis a[3] = {Target.x_axis->c(), Target.y_axis->c(), Target.z_axis->c()};
    kf_compat::check_io(kf_io_save_grid(path.c_str(), a));
}

inline axis *kf_make_axis(const kf_axis &a)
{
    // load_axis (IO.cpp:153-164)
    switch (a.type) {
    case KF_AXIS_EQST: return new axis_CoEqSt(a.l0, a.L, a.n);
    case KF_AXIS_HYST: return new axis_CoHySt(a.l0, a.L, a.n, (int)a.m);
    default: return new axis_CoSiSt(a.l0, a.L, a.n, a.m);
    }
}

inline grid load_grid(const std::string &path = "./data/fields.h5")
{
    kf_axis a[3];
    kf_compat::check_io(kf_io_load_grid(path.c_str(), a));
    std::unique_ptr<axis> x(kf_make_axis(a[0])), y(kf_make_axis(a[1])), z(kf_make_axis(a[2]));
    return grid(*x, *y, *z);
}

// save_time / load_time (IO.cpp:253-294)
inline void save_time(const double &Time, const std::string &path = "./data/fields.h5")
{
    kf_compat::check_io(kf_io_save_time(path.c_str(), Time));
}
inline double load_time(const std::string &path = "./data/fields.h5")
{
    double t = 0.;
    kf_compat::check_io(kf_io_load_time(path.c_str(), &t));
    return t;
}

// save_parameters / load_parameters (IO.cpp:298-360)
inline void save_parameters(const parameters &Target, const std::string &path = "./data/fields.h5")
{
    const kf_parameters p = Target.c();
    kf_compat::check_io(kf_io_save_parameters(path.c_str(), &p));
}
inline parameters load_parameters(const std::string &path = "./data/fields.h5")
{
    kf_parameters p;
    kf_compat::check_io(kf_io_load_parameters(path.c_str(), &p));
    return parameters(p.M, p.tau, p.theta, p.mu, p.beta);
}

// save_field_real (IO.cpp:365-394)
inline void save_field_real(const std::string &data_set, const field_real &Target,
                            const std::string &path = "./data/fields.h5")
{
    kf_compat::check_io(
        kf_io_save_field_real(path.c_str(), data_set.c_str(), Target.val, Target.Nx, Target.Ny, Target.Nz));
}
// load_field_real (IO.cpp:397-441): the definition takes (path, data_set),
// the reverse of its declaration at IO.hpp:103-104; callers (backend.cpp:
// 220-224) follow the definition, and so does this
inline field_real load_field_real(const std::string &path, const std::string &data_set)
{
    grid Omega = load_grid(path);
    field_real Target(Omega);
    kf_compat::check_io(
        kf_io_load_field_real(path.c_str(), data_set.c_str(), Target.val, Target.Nx, Target.Ny, Target.Nz));
    return Target;
}

// save_field_imag / load_field_imag (IO.cpp:446-504)
inline void save_field_imag(const std::string &data_set, const field_imag &Target,
                            const std::string &path = "./data/fields.h5")
{
    field_real tmp(Target.my_grid);
    iFFT(Target, tmp);
    save_field_real(data_set, tmp, path);
}
inline field_imag load_field_imag(const std::string &data_set, const std::string &path = "./data/fields.h5")
{
    field_real RF = load_field_real(path, data_set);
    field_imag IF(RF.my_grid);
    FFT(RF, IF);
    return IF;
}

// save_slice / load_slice of one field (IO.cpp:1175-1266): the z plane at
// z_axis->index_at(0), double[Nx][Ny]
inline void save_slice(const std::string &data_set, const field_imag &Target, const std::string &path)
{
    field_real tmp(Target.my_grid);
    iFFT(Target, tmp);
    const int k = tmp.my_grid.z_axis->index_at(0.);
    std::vector<double> plane((size_t)tmp.Nx * tmp.Ny);
    for (int i = 0; i < tmp.Nx; ++i)
        for (int j = 0; j < tmp.Ny; ++j) plane[(size_t)i * tmp.Ny + j] = tmp.val[tmp.index(i, j, k)];
    kf_compat::check_io(kf_io_save_plane(path.c_str(), data_set.c_str(), plane.data(), tmp.Nx, tmp.Ny));
}
inline void load_slice(const std::string &data_set, double *ptr_data, const std::string &path)
{
    grid Omega = load_grid(path);
    kf_compat::check_io(kf_io_load_plane(path.c_str(), data_set.c_str(), ptr_data, Omega.Nx, Omega.Ny));
}

// I4 (IO.cpp:1281-1287)
inline std::string I4(int number)
{
    char b[32];
    kf_compat::check_io(kf_io_i4(number, b, (int)sizeof(b)));
    return std::string(b);
}

// save_all (IO.cpp:1295-1314)
inline void save_all(const grid &Omega, const double &Time, const parameters &Params, const field_imag &FUx,
                     const field_imag &FUy, const field_imag &FUz, const field_imag &Fni, const field_imag &FPh,
                     const std::string &path = "./data/fields.h5")
{
    file_create(path);
    save_grid(Omega, path);
    save_time(Time, path);
    save_parameters(Params, path);
    save_field_imag("Ux", FUx, path);
    save_field_imag("Uy", FUy, path);
    save_field_imag("Uz", FUz, path);
    save_field_imag("ni", Fni, path);
    save_field_imag("Ph", FPh, path);
}

// save_slice of a state (IO.cpp:1317-1338): ./data/slice_<I4(step)>.h5
inline void save_slice(const int &step, const grid &Omega, const double &Time, const parameters &Params,
                       const field_imag &FUx, const field_imag &FUy, const field_imag &FUz, const field_imag &Fni,
                       const field_imag &FPh, const std::string &dir = "./data")
{
    const std::string path = dir + "/slice_" + I4(step) + ".h5";
    file_create(path);
    save_grid(Omega, path);
    save_time(Time, path);
    save_parameters(Params, path);
    save_slice("Ux", FUx, path);
    save_slice("Uy", FUy, path);
    save_slice("Uz", FUz, path);
    save_slice("ni", Fni, path);
    save_slice("Ph", FPh, path);
}
