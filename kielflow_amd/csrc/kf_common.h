// kf_common.h -- shared types and device helpers of the kielflow_amd HIP library.
//
// Field layout is kielflow's (field_base.hpp:23-24, field_imag.cpp:15-16):
//   real     R[Nx][Ny][Nz]          (z fastest)
//   spectral C2[Nx][Ny][Nz/2+1]     (interleaved re/im)
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <type_traits>

namespace kf {

template <typename R>
struct VecT;
template <>
struct VecT<float> {
    using C2 = float2;
};
template <>
struct VecT<double> {
    using C2 = double2;
};
template <typename R>
using C2 = typename VecT<R>::C2;

template <typename C>
__host__ __device__ inline C mk(decltype(C().x) a, decltype(C().x) b)
{
    C c;
    c.x = a;
    c.y = b;
    return c;
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b mod 8).  xcd_block maps the hardware index to a logical one so that each
// XCD works through one contiguous eighth of the logical blocks: neighbouring
// tiles then share that XCD's L2.  The tail beyond a multiple of 8 keeps its
// index.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned n)
{
    const unsigned n8 = n >> 3;
    if (b >= (n8 << 3)) return b;
    return (b & 7u) * n8 + (b >> 3);
}

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C>
__device__ __forceinline__ C cmul(C a, C b)
{
    return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C>
__device__ __forceinline__ C cconj(C a) { return mk<C>(a.x, -a.y); }
template <typename C, typename S>
__device__ __forceinline__ C cscale(C a, S s) { return mk<C>(a.x * s, a.y * s); }
// multiply by +i (sign>0) or -i (sign<0)
template <int SIGN, typename C>
__device__ __forceinline__ C cmuli(C a)
{
    return SIGN > 0 ? mk<C>(-a.y, a.x) : mk<C>(a.y, -a.x);
}

// Device view of a grid: sizes plus per-axis tables (device memory).
//   k[d][i]  = k_val_at(i) of axis d in the compute precision R
//              (z axis over i = 0..Nz/2)
//   q[d][i]  = (k/k_max)^2 in FP64, k_max = k_val_at(N/2)
//   qzf[i]   = (kz/k_val_at((Nz/2+1)/2))^2 in FP64 (the Phi filter's z norm,
//              operations.hpp:23 reads Nz of the complex field)
template <typename R>
struct GridView {
    int nx, ny, nz, nzc;
    long nr, nc;
    const R *k[3];
    const double *q[3];
    const double *qzf;
    double d23_thr;   // 2/3-rule: sqrt(s) < 2/3  <=>  s < d23_thr  (s = sum of q)
    // per column (j, k): the largest |i'| (i' the signed x index) whose mode
    // passes the 2/3 rule, or -1; the kept x indices of a column are exactly
    // |i'| <= imax since q_x grows with |i'|
    const int *imax;
};

// wavenumber along axis dir at (i, j, kk): all three table values are
// loaded and one is selected, so that no run-time offset into the view is
// formed (that would force a copy of a functor holding the view into
// scratch memory)
template <typename R>
__device__ __forceinline__ R k_along(const GridView<R> &g, int dir, int i, int j, int kk)
{
    const R a = g.k[0][i], b = g.k[1][j], c = g.k[2][kk];
    return dir == 0 ? a : (dir == 1 ? b : c);
}

// D23 membership from per-axis squared ratios, evaluated as the reference
// does: ((qx + qy) + qz) then compared through the exact threshold that
// sqrt(s) < 2./3. induces (OP_dealiasing.hpp:28-29, rhs_standard.cpp:22).
__device__ __forceinline__ bool d23_keep(double qx, double qy, double qz, double thr)
{
    double s = __dadd_rn(__dadd_rn(qx, qy), qz);
    return s < thr;
}

}  // namespace kf

#define KF_HIP(call)                                                               \
    do {                                                                            \
        hipError_t _e = (call);                                                     \
        if (_e != hipSuccess)                                                       \
            throw kf::Error(std::string(#call) + ": " + hipGetErrorString(_e));      \
    } while (0)

namespace kf {
struct Error {
    std::string msg;
    explicit Error(std::string m) : msg(std::move(m)) {}
};
}  // namespace kf
