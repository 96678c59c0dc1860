/*
 * kielflow_io.h -- C ABI of the on-disk format of kielflow: the HDF5 files
 * that frontend.cpp writes (fields.h5), that kielflow.cpp reads and writes
 * (backup_at_<step>.h5, slice_<step>.h5) and that backend.cpp reads.
 *
 * Host code only (libkielflow_io.so, over the HDF5 C library; the reference
 * goes through the echelon C++ wrapper, IO.hpp:40).  Arrays are HOST arrays
 * in the reference's real-field order double[Nx][Ny][Nz] (z fastest).
 * The layout written (IO.cpp:60-504, 1175-1338):
 *   /version (string)  /date (long, time(2))  /path (string)  /time (double)
 *   /program/  (empty)
 *   /parameters/"mach number M", "damping ratio tau", "temperature ratio theta",
 *               "velocity diffusion mu", "magnetization beta"   (double)
 *   /domain/type_{x,y,z} (int: 1 EqSt, 2 HySt, 3 SiSt), N_{x,y,z} (int),
 *           m_, l0_, L_{x,y,z} (double)
 *   /data/<name> double[Nx][Ny][Nz] (fields) or double[Nx][Ny] (slices)
 *
 * Every entry point returns KF_OK (0) or KF_ERR_ARG with a message in
 * kf_io_last_error().  Creating a dataset that already exists is an error,
 * as it is in the reference.
 */
#ifndef KIELFLOW_IO_H_
#define KIELFLOW_IO_H_

#include "kielflow_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *kf_io_last_error(void);

/* file_create (IO.cpp:60-86): truncate, groups program/parameters/domain/data,
 * root scalars version, date, path.  version may be NULL ("kielflow_amd"). */
int kf_io_file_create(const char *path, const char *version);

/* save_grid / save_axis (IO.cpp:90-117, 173-196); creates the file first if
 * it does not exist, as save_axis does */
int kf_io_save_grid(const char *path, const kf_axis axes[3]);
/* load_grid / load_axis (IO.cpp:120-171, 199-249) */
int kf_io_load_grid(const char *path, kf_axis axes[3]);

/* save_time / load_time (IO.cpp:253-294): the dataset sits at the root */
int kf_io_save_time(const char *path, double t);
int kf_io_load_time(const char *path, double *t);

/* save_parameters / load_parameters (IO.cpp:298-360) */
int kf_io_save_parameters(const char *path, const kf_parameters *p);
int kf_io_load_parameters(const char *path, kf_parameters *p);

/* save_field_real / load_field_real (IO.cpp:365-441): /data/<name> as
 * double[nx][ny][nz]; loading checks the stored extents */
int kf_io_save_field_real(const char *path, const char *name, const double *f, int nx, int ny, int nz);
int kf_io_load_field_real(const char *path, const char *name, double *f, int nx, int ny, int nz);

/* the dataset writes of save_slice / load_slice (IO.cpp:1175-1266):
 * /data/<name> as double[nx][ny] */
int kf_io_save_plane(const char *path, const char *name, const double *f, int nx, int ny);
int kf_io_load_plane(const char *path, const char *name, double *f, int nx, int ny);

/* I4 (IO.cpp:1281-1287): the step number right-aligned in 4 characters,
 * zero-filled; writes at most len bytes including the terminator */
int kf_io_i4(int number, char *out, int len);

#ifdef __cplusplus
}
#endif

#endif /* KIELFLOW_IO_H_ */
