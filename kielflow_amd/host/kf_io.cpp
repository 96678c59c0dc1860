// kf_io.cpp -- kielflow's HDF5 file layout (IO.cpp:60-504, 1175-1338) over
// the HDF5 C library.  Host code only; built into libkielflow_io.so.
//
// The reference writes through echelon (IO.hpp:40): scalar datasets for
// every number and string, one group per topic, and contiguous double
// arrays for fields.  The same names, ranks and element classes are written
// here (int for axis type and size, double for the rest, variable-length
// strings), so a file from either side loads in the other.  Reads convert to
// the requested memory type, so integer or float widths chosen by another
// writer load as well.
#include "kielflow_io.h"

#include <hdf5.h>
#include <unistd.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct IoError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// HDF5 prints its error stack by default; within an entry point the
// message returned names the failing operation instead, and the caller's
// handler is restored on every exit path
struct QuietHdf5 {
    H5E_auto2_t func = nullptr;
    void *data = nullptr;
    QuietHdf5()
    {
        H5Eget_auto2(H5E_DEFAULT, &func, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietHdf5() { H5Eset_auto2(H5E_DEFAULT, func, data); }
};

template <class F>
int guard(F &&f)
{
    try {
        QuietHdf5 q;
        return f();
    } catch (const std::exception &e) {
        g_err = e.what();
        return KF_ERR_ARG;
    }
}

void need(bool ok, const std::string &what)
{
    if (!ok) throw IoError(what);
}

// owning hid_t
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    ~Hid()
    {
        if (id >= 0 && close) close(id);
    }
    operator hid_t() const { return id; }
};

bool exists(const char *path) { return path && access(path, F_OK) == 0; }

hid_t open_file(const char *path, bool write)
{
    need(path && *path, "empty path");
    need(exists(path), std::string("no such file: ") + path);
    hid_t f = H5Fopen(path, write ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
    need(f >= 0, std::string("not an HDF5 file or not accessible: ") + path);
    return f;
}

hid_t open_group(hid_t f, const char *name)
{
    need(H5Lexists(f, name, H5P_DEFAULT) > 0, std::string("missing group /") + name);
    hid_t g = H5Gopen2(f, name, H5P_DEFAULT);
    need(g >= 0, std::string("cannot open group /") + name);
    return g;
}

void write_scalar(hid_t loc, const std::string &name, hid_t file_type, hid_t mem_type, const void *v)
{
    need(H5Lexists(loc, name.c_str(), H5P_DEFAULT) <= 0, "dataset exists: " + name);
    Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
    Hid d(H5Dcreate2(loc, name.c_str(), file_type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    need(d >= 0, "cannot create dataset " + name);
    need(H5Dwrite(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0, "cannot write dataset " + name);
}

void write_string(hid_t loc, const std::string &name, const std::string &s)
{
    Hid t(H5Tcopy(H5T_C_S1), H5Tclose);
    need(H5Tset_size(t, H5T_VARIABLE) >= 0, "string type");
    const char *p = s.c_str();
    write_scalar(loc, name, t, t, &p);
}

hid_t open_dataset(hid_t loc, const std::string &name)
{
    need(H5Lexists(loc, name.c_str(), H5P_DEFAULT) > 0, "missing dataset " + name);
    hid_t d = H5Dopen2(loc, name.c_str(), H5P_DEFAULT);
    need(d >= 0, "cannot open dataset " + name);
    return d;
}

template <class T>
T read_scalar(hid_t loc, const std::string &name, hid_t mem_type)
{
    Hid d(open_dataset(loc, name), H5Dclose);
    Hid sp(H5Dget_space(d), H5Sclose);
    need(H5Sget_simple_extent_npoints(sp) == 1, "not a scalar: " + name);
    T v{};
    need(H5Dread(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v) >= 0, "cannot read " + name);
    return v;
}

void write_array(const char *path, const char *name, const double *f, int rank, const hsize_t *dims)
{
    need(name && *name, "empty dataset name");
    need(f != nullptr, "null array");
    for (int r = 0; r < rank; ++r) need(dims[r] > 0, "non-positive extent");
    Hid file(open_file(path, true), H5Fclose);
    Hid data(open_group(file, "data"), H5Gclose);
    need(H5Lexists(data, name, H5P_DEFAULT) <= 0, std::string("dataset exists: /data/") + name);
    Hid sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    Hid d(H5Dcreate2(data, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    need(d >= 0, std::string("cannot create /data/") + name);
    need(H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, f) >= 0,
         std::string("cannot write /data/") + name);
}

void read_array(const char *path, const char *name, double *f, int rank, const hsize_t *dims)
{
    need(name && *name, "empty dataset name");
    need(f != nullptr, "null array");
    Hid file(open_file(path, false), H5Fclose);
    Hid data(open_group(file, "data"), H5Gclose);
    Hid d(open_dataset(data, name), H5Dclose);
    Hid sp(H5Dget_space(d), H5Sclose);
    const int r = H5Sget_simple_extent_ndims(sp);
    need(r == rank, std::string("/data/") + name + ": rank " + std::to_string(r) + ", expected " +
                        std::to_string(rank));
    hsize_t got[3] = {0, 0, 0};
    H5Sget_simple_extent_dims(sp, got, nullptr);
    for (int i = 0; i < rank; ++i)
        need(got[i] == dims[i], std::string("/data/") + name + ": extent " + std::to_string(got[i]) + " on axis " +
                                    std::to_string(i) + ", expected " + std::to_string(dims[i]));
    need(H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, f) >= 0,
         std::string("cannot read /data/") + name);
}

std::string self_path()
{
    // get_selfpath (IO.cpp:4-15)
    char buf[PATH_MAX];
    const ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    if (n < 0) return std::string();
    buf[n] = '\0';
    return std::string(buf);
}

const char *const kAxisName[3] = {"x", "y", "z"};
const char *const kParamName[5] = {"mach number M", "damping ratio tau", "temperature ratio theta",
                                   "velocity diffusion mu", "magnetization beta"};

void check_axis(const kf_axis &a)
{
    need(a.type == KF_AXIS_EQST || a.type == KF_AXIS_HYST || a.type == KF_AXIS_SIST,
         "axis type must be 1 (EqSt), 2 (HySt) or 3 (SiSt)");
    need(a.n > 0 && a.L > 0, "axis needs N > 0 and L > 0");
}

}  // namespace

extern "C" {

const char *kf_io_last_error(void) { return g_err.c_str(); }

int kf_io_file_create(const char *path, const char *version)
{
    return guard([&] {
        need(path && *path, "empty path");
        Hid f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
        need(f >= 0, std::string("cannot create ") + path);
        for (const char *g : {"program", "parameters", "domain", "data"}) {
            Hid h(H5Gcreate2(f, g, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            need(h >= 0, std::string("cannot create group /") + g);
        }
        write_string(f, "version", version ? version : "kielflow_amd");
        const long now = (long)time(nullptr);
        write_scalar(f, "date", H5T_NATIVE_LONG, H5T_NATIVE_LONG, &now);
        write_string(f, "path", self_path());
        return KF_OK;
    });
}

int kf_io_save_grid(const char *path, const kf_axis axes[3])
{
    return guard([&] {
        need(axes != nullptr, "null axes");
        for (int d = 0; d < 3; ++d) check_axis(axes[d]);
        if (!exists(path)) {
            const int rc = kf_io_file_create(path, nullptr);
            if (rc != KF_OK) throw IoError(g_err);
        }
        Hid f(open_file(path, true), H5Fclose);
        Hid dom(open_group(f, "domain"), H5Gclose);
        for (int d = 0; d < 3; ++d) {
            const std::string c = kAxisName[d];
            const kf_axis &a = axes[d];
            write_scalar(dom, "type_" + c, H5T_NATIVE_INT, H5T_NATIVE_INT, &a.type);
            write_scalar(dom, "m_" + c, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, &a.m);
            write_scalar(dom, "l0_" + c, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, &a.l0);
            write_scalar(dom, "N_" + c, H5T_NATIVE_INT, H5T_NATIVE_INT, &a.n);
            write_scalar(dom, "L_" + c, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, &a.L);
        }
        return KF_OK;
    });
}

int kf_io_load_grid(const char *path, kf_axis axes[3])
{
    return guard([&] {
        need(axes != nullptr, "null axes");
        Hid f(open_file(path, false), H5Fclose);
        Hid dom(open_group(f, "domain"), H5Gclose);
        for (int d = 0; d < 3; ++d) {
            const std::string c = kAxisName[d];
            kf_axis a;
            a.type = read_scalar<int>(dom, "type_" + c, H5T_NATIVE_INT);
            a.m = read_scalar<double>(dom, "m_" + c, H5T_NATIVE_DOUBLE);
            a.l0 = read_scalar<double>(dom, "l0_" + c, H5T_NATIVE_DOUBLE);
            a.n = read_scalar<int>(dom, "N_" + c, H5T_NATIVE_INT);
            a.L = read_scalar<double>(dom, "L_" + c, H5T_NATIVE_DOUBLE);
            // load_axis leaves the axis unset for an unknown type (IO.cpp:153-164);
            // here that is an error
            check_axis(a);
            axes[d] = a;
        }
        return KF_OK;
    });
}

int kf_io_save_time(const char *path, double t)
{
    return guard([&] {
        Hid f(open_file(path, true), H5Fclose);
        write_scalar(f, "time", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, &t);
        return KF_OK;
    });
}

int kf_io_load_time(const char *path, double *t)
{
    return guard([&] {
        need(t != nullptr, "null output");
        Hid f(open_file(path, false), H5Fclose);
        *t = read_scalar<double>(f, "time", H5T_NATIVE_DOUBLE);
        return KF_OK;
    });
}

int kf_io_save_parameters(const char *path, const kf_parameters *p)
{
    return guard([&] {
        need(p != nullptr, "null parameters");
        Hid f(open_file(path, true), H5Fclose);
        Hid g(open_group(f, "parameters"), H5Gclose);
        const double v[5] = {p->M, p->tau, p->theta, p->mu, p->beta};
        for (int i = 0; i < 5; ++i) write_scalar(g, kParamName[i], H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, &v[i]);
        return KF_OK;
    });
}

int kf_io_load_parameters(const char *path, kf_parameters *p)
{
    return guard([&] {
        need(p != nullptr, "null output");
        Hid f(open_file(path, false), H5Fclose);
        Hid g(open_group(f, "parameters"), H5Gclose);
        double v[5];
        for (int i = 0; i < 5; ++i) v[i] = read_scalar<double>(g, kParamName[i], H5T_NATIVE_DOUBLE);
        *p = kf_parameters{v[0], v[1], v[2], v[3], v[4]};
        return KF_OK;
    });
}

int kf_io_save_field_real(const char *path, const char *name, const double *f, int nx, int ny, int nz)
{
    return guard([&] {
        need(nx > 0 && ny > 0 && nz > 0, "non-positive extent");
        const hsize_t dims[3] = {(hsize_t)nx, (hsize_t)ny, (hsize_t)nz};
        write_array(path, name, f, 3, dims);
        return KF_OK;
    });
}

int kf_io_load_field_real(const char *path, const char *name, double *f, int nx, int ny, int nz)
{
    return guard([&] {
        need(nx > 0 && ny > 0 && nz > 0, "non-positive extent");
        const hsize_t dims[3] = {(hsize_t)nx, (hsize_t)ny, (hsize_t)nz};
        read_array(path, name, f, 3, dims);
        return KF_OK;
    });
}

int kf_io_save_plane(const char *path, const char *name, const double *f, int nx, int ny)
{
    return guard([&] {
        need(nx > 0 && ny > 0, "non-positive extent");
        const hsize_t dims[2] = {(hsize_t)nx, (hsize_t)ny};
        write_array(path, name, f, 2, dims);
        return KF_OK;
    });
}

int kf_io_load_plane(const char *path, const char *name, double *f, int nx, int ny)
{
    return guard([&] {
        need(nx > 0 && ny > 0, "non-positive extent");
        const hsize_t dims[2] = {(hsize_t)nx, (hsize_t)ny};
        read_array(path, name, f, 2, dims);
        return KF_OK;
    });
}

int kf_io_i4(int number, char *out, int len)
{
    return guard([&] {
        need(out && len > 0, "null output");
        std::string s = std::to_string(number);
        if (s.size() < 4) s = std::string(4 - s.size(), '0') + s;
        need((int)s.size() < len, "output buffer too small");
        std::memcpy(out, s.c_str(), s.size() + 1);
        return KF_OK;
    });
}

}  // extern "C"
