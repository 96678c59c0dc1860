// kf_poisson.hip -- nonlinear Poisson solver (Boltzmann electrons) on the GPU.
#include "kf_poisson.h"
#include "kf_prof.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace kf {

// ======================================================================== axes
double axis_k_val_at(const AxisDesc &a, int index)
{
    // axis_EqSt.cpp:57-61 (also SiSt, and the computational grid of HySt)
    double pi = acos(-1.);
    return ((index <= a.n / 2) ? (2. * pi * index / a.L) : (2. * pi * (index - a.n) / a.L));
}

double axis_dS(const AxisDesc &a, int index)
{
    if (a.type == 1) return 1.;
    if (a.type == 3) {  // axis_SiSt.cpp:129-134
        const double pi = acos(-1.);
        double x = a.l0 + (a.L / a.n) * index;
        return 1. - (2. * pi * a.m / a.L) * cos(2. * pi * x / a.L);
    }
    // out_sourced/axis_FoHySt.cpp:107-117
    double L = a.L, m = a.m;
    double x_val = a.l0 + L * (double)(index) / (double)(a.n);
    if (x_val >= L / 2.) return (m / sinh(m)) * cosh(2. * (x_val - L) * m / L);
    if (x_val < -L / 2.) return (m / sinh(m)) * cosh(2. * (x_val + L) * m / L);
    return (m / sinh(m)) * cosh(2. * (x_val)*m / L);
}

// ======================================================================== device helpers
constexpr int RED_BLOCK = 256;

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// linear index -> (i, j, k) of a [nx][ny][nz] array, 32-bit unsigned
// arithmetic (every grid here has fewer than 2^31 points)
__device__ __forceinline__ void dec3(long p, int ny, int nz, int &i, int &j, int &k)
{
    const unsigned u = (unsigned)p, z = (unsigned)nz, y = (unsigned)ny;
    const unsigned r = u / z;
    k = (int)(u - r * z);
    const unsigned q = r / y;
    j = (int)(r - q * y);
    i = (int)q;
}

// On an axis of uniform spacing the coefficients are the constants ucp, udd
// (kernel arguments, so no per-point coefficient loads); other axes gather
// them per index.  Offsets are 32-bit (every level has fewer than 2^31 points).
// value of a field at a linear index: x[q]
template <typename R>
struct Plain {
    const R *__restrict__ x;
    __device__ __forceinline__ R operator()(unsigned q) const { return x[q]; }
};
// phi[q] + x[q]: the potential after the pending Newton correction
template <typename R>
struct Corrected {
    const double *__restrict__ phi;
    const R *__restrict__ x;
    __device__ __forceinline__ double operator()(unsigned q) const { return phi[q] + (double)x[q]; }
};

template <typename R, class LD>
__device__ __forceinline__ void stencil_ld(const LevelView<R> &L, const LD &x, int i, int j, int k, R &A, R &D)
{
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const unsigned sx = (unsigned)ny * (unsigned)nz;
    const unsigned ip = wrapi(i + 1, nx), im = wrapi(i - 1, nx);
    const unsigned jp = wrapi(j + 1, ny), jm = wrapi(j - 1, ny);
    const unsigned kp = wrapi(k + 1, nz), km = wrapi(k - 1, nz);
    const unsigned ri = (unsigned)i * sx, rj = (unsigned)j * (unsigned)nz;
    const R xip = x(ip * sx + rj + (unsigned)k), xim = x(im * sx + rj + (unsigned)k);
    const R xjp = x(ri + jp * (unsigned)nz + (unsigned)k), xjm = x(ri + jm * (unsigned)nz + (unsigned)k);
    const R xkp = x(ri + rj + kp), xkm = x(ri + rj + km);
    R a, d;
    if (L.c[0].uni) {
        a = L.c[0].ucp * (xip + xim);
        d = L.c[0].udd;
    } else {
        a = L.c[0].cp[i] * xip + L.c[0].cm[i] * xim;
        d = L.c[0].dd[i];
    }
    if (L.c[1].uni) {
        a += L.c[1].ucp * (xjp + xjm);
        d += L.c[1].udd;
    } else {
        a += L.c[1].cp[j] * xjp + L.c[1].cm[j] * xjm;
        d += L.c[1].dd[j];
    }
    if (L.c[2].uni) {
        a += L.c[2].ucp * (xkp + xkm);
        d += L.c[2].udd;
    } else {
        a += L.c[2].cp[k] * xkp + L.c[2].cm[k] * xkm;
        d += L.c[2].dd[k];
    }
    A = (R)2 * a;
    D = d;
}

template <typename R>
__device__ __forceinline__ void stencil(const LevelView<R> &L, const R *__restrict__ x, int i, int j,
                                        int k, R &A, R &D)
{
    stencil_ld<R>(L, Plain<R>{x}, i, j, k, A, D);
}

// block reductions: wavefront (64 lanes) shuffles, then one LDS slot per wave
template <typename T, class Op>
__device__ __forceinline__ T block_reduce(T v, T *sh, Op op)
{
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    T r = sh[0];
    for (int w = 1; w < nw; ++w) r = op(r, sh[w]);
    __syncthreads();
    return r;
}
struct OpMax {
    template <typename T>
    __device__ T operator()(T a, T b) const { return a > b ? a : b; }
};
struct OpMin {
    template <typename T>
    __device__ T operator()(T a, T b) const { return a < b ? a : b; }
};
struct OpSum {
    template <typename T>
    __device__ T operator()(T a, T b) const { return a + b; }
};
template <typename T>
__device__ __forceinline__ T block_max(T v, T *sh) { return block_reduce(v, sh, OpMax()); }
template <typename T>
__device__ __forceinline__ T block_min(T v, T *sh) { return block_reduce(v, sh, OpMin()); }
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) { return block_reduce(v, sh, OpSum()); }

// grid used by reduction kernels: grid-stride loops over at most this many blocks
constexpr unsigned RED_MAX_BLOCKS = 2048;

// ======================================================================== Newton-MG kernels
// F = L_h phi - exp(phi) + rho (FP64); b = -F and e = exp(phi) in the sweep
// precision; per-block max|F| and max|rho|.
// Each thread takes NR_PTS points spaced RED_BLOCK apart inside its block's
// contiguous chunk, so the loads of all its points are in flight together.
constexpr int NR_PTS = 4;
template <typename RIN, typename R>
__global__ void __launch_bounds__(RED_BLOCK)
k_newton_resid(LevelView<double> L, const double *__restrict__ phi, const RIN *__restrict__ rho,
               R *__restrict__ b, R *__restrict__ e, double *__restrict__ part, const int *ctl, double theta)
{
    __shared__ double sh[RED_BLOCK / 64];
    if (ctl[0]) return;
    double fmaxv = 0., rmax = 0.;
    const int nz = L.n[2], ny = L.n[1];
    const long p0 = (long)xcd_block(blockIdx.x, gridDim.x) * (RED_BLOCK * NR_PTS) + threadIdx.x;
    double F[NR_PTS], ex[NR_PTS], rh[NR_PTS];
#pragma unroll
    for (int u = 0; u < NR_PTS; ++u) {
        const long p = p0 + (long)u * RED_BLOCK;
        F[u] = 0.;
        ex[u] = 0.;
        rh[u] = 0.;
        if (p < L.N) {
            int i, j, k;
            dec3(p, ny, nz, i, j, k);
            double A, D;
            stencil<double>(L, phi, i, j, k, A, D);
            const double x = phi[p];
            ex[u] = exp(x);
            rh[u] = (double)rho[p];
            F[u] = A - D * x - ex[u] + rh[u];
            if (theta != 0.) {
                // Boltzmann ions: F += exp(-theta x), J gains -theta exp(-theta x)
                const double ei = exp(-theta * x);
                F[u] += ei;
                ex[u] += theta * ei;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < NR_PTS; ++u) {
        const long p = p0 + (long)u * RED_BLOCK;
        if (p < L.N) {
            b[p] = (R)(-F[u]);
            e[p] = (R)ex[u];
            fmaxv = fmax(fmaxv, fabs(F[u]));
            rmax = fmax(rmax, fabs(rh[u]));
        }
    }
    fmaxv = block_max(fmaxv, sh);
    rmax = block_max(rmax, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = fmaxv;
        part[2 * blockIdx.x + 1] = rmax;
    }
}

// Newton update fused with the next residual: phi_out = phi_in + x and
// F(phi_out) as k_newton_resid (the stencil reads phi_in + x, so phi_in is
// not modified while neighbours still read it).  ctl[3] records which of the
// two potential buffers holds the newest iterate.
template <typename RIN, typename R>
__global__ void __launch_bounds__(RED_BLOCK)
k_newton_upd_resid(LevelView<double> L, const double *__restrict__ phi_in, const R *__restrict__ x,
                   double *__restrict__ phi_out, int out_id, const RIN *__restrict__ rho, R *__restrict__ b,
                   R *__restrict__ e, double *__restrict__ part, int *ctl, double theta)
{
    __shared__ double sh[RED_BLOCK / 64];
    if (ctl[0]) return;
    double fmaxv = 0., rmax = 0.;
    const int nz = L.n[2], ny = L.n[1];
    const long p0 = (long)xcd_block(blockIdx.x, gridDim.x) * (RED_BLOCK * NR_PTS) + threadIdx.x;
    const Corrected<R> cur{phi_in, x};
    double F[NR_PTS], ex[NR_PTS], rh[NR_PTS], xn[NR_PTS];
#pragma unroll
    for (int u = 0; u < NR_PTS; ++u) {
        const long p = p0 + (long)u * RED_BLOCK;
        F[u] = 0.;
        ex[u] = 0.;
        rh[u] = 0.;
        xn[u] = 0.;
        if (p < L.N) {
            int i, j, k;
            dec3(p, ny, nz, i, j, k);
            double A, D;
            stencil_ld<double>(L, cur, i, j, k, A, D);
            xn[u] = cur((unsigned)p);
            ex[u] = exp(xn[u]);
            rh[u] = (double)rho[p];
            F[u] = A - D * xn[u] - ex[u] + rh[u];
            if (theta != 0.) {
                const double ei = exp(-theta * xn[u]);
                F[u] += ei;
                ex[u] += theta * ei;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < NR_PTS; ++u) {
        const long p = p0 + (long)u * RED_BLOCK;
        if (p < L.N) {
            phi_out[p] = xn[u];
            b[p] = (R)(-F[u]);
            e[p] = (R)ex[u];
            fmaxv = fmax(fmaxv, fabs(F[u]));
            rmax = fmax(rmax, fabs(rh[u]));
        }
    }
    fmaxv = block_max(fmaxv, sh);
    rmax = block_max(rmax, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = fmaxv;
        part[2 * blockIdx.x + 1] = rmax;
        if (blockIdx.x == 0) ctl[3] = out_id;
    }
}

// Newton residual with each thread on 4 consecutive z points of one line
// (Nz % 4 == 0): one index decode, vector loads of the line and of its four
// neighbour lines, and the two z neighbours outside the quad.  With x given,
// the pending correction is applied first (phi_out = phi_in + x), as in
// k_newton_upd_resid; with x == nullptr it is k_newton_resid.
template <typename R>
__device__ __forceinline__ void load4(const double *__restrict__ phi, const R *__restrict__ x, unsigned q,
                                      double v[4])
{
    const double2 a = *reinterpret_cast<const double2 *>(phi + q);
    const double2 b = *reinterpret_cast<const double2 *>(phi + q + 2);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
    if (x) {
        if constexpr (std::is_same<R, float>::value) {
            const float4 c = *reinterpret_cast<const float4 *>(x + q);
            v[0] += (double)c.x;
            v[1] += (double)c.y;
            v[2] += (double)c.z;
            v[3] += (double)c.w;
        } else {
            const double2 c = *reinterpret_cast<const double2 *>(x + q);
            const double2 d = *reinterpret_cast<const double2 *>(x + q + 2);
            v[0] += c.x;
            v[1] += c.y;
            v[2] += d.x;
            v[3] += d.y;
        }
    }
}

template <typename R>
__device__ __forceinline__ void store4(R *__restrict__ p, unsigned q, const double v[4])
{
    if constexpr (std::is_same<R, float>::value) {
        *reinterpret_cast<float4 *>(p + q) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
    } else {
        *reinterpret_cast<double2 *>(p + q) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2 *>(p + q + 2) = make_double2(v[2], v[3]);
    }
}

// index of (i, j, k) in a colour-split field (layout described further down)
struct SplitIdx {
    unsigned ny, nzh, half;
    __device__ __forceinline__ unsigned operator()(unsigned i, unsigned j, unsigned k) const
    {
        return ((i + j + k) & 1u) * half + (i * ny + j) * nzh + (k >> 1);
    }
};

// the four values at z = k0..k0+3 (k0 % 4 == 0) of a colour-split field
// (FP32): colour (i+j) & 1 holds k0 and k0+2, the other colour k0+1 and k0+3
__device__ __forceinline__ void split4_load(const float *__restrict__ x, unsigned half, unsigned line,
                                           unsigned nzh, unsigned par0, unsigned k0, double v[4])
{
    const unsigned o = line * nzh + (k0 >> 1);
    const float2 a = *reinterpret_cast<const float2 *>(x + par0 * half + o);
    const float2 c = *reinterpret_cast<const float2 *>(x + (1u - par0) * half + o);
    v[0] += (double)a.x;
    v[1] += (double)c.x;
    v[2] += (double)a.y;
    v[3] += (double)c.y;
}

__device__ __forceinline__ void split4_store(float *__restrict__ x, unsigned half, unsigned line, unsigned nzh,
                                            unsigned par0, unsigned k0, const double v[4])
{
    const unsigned o = line * nzh + (k0 >> 1);
    *reinterpret_cast<float2 *>(x + par0 * half + o) = make_float2((float)v[0], (float)v[2]);
    *reinterpret_cast<float2 *>(x + (1u - par0) * half + o) = make_float2((float)v[1], (float)v[3]);
}

template <typename RIN, typename R, bool SPLIT = false>
__global__ void __launch_bounds__(RED_BLOCK)
k_newton_z4(LevelView<double> L, const double *__restrict__ phi_in, const R *__restrict__ x,
            double *__restrict__ phi_out, int out_id, const RIN *__restrict__ rho, R *__restrict__ b,
            R *__restrict__ e, double *__restrict__ part, int *ctl)
{
    __shared__ double sh[RED_BLOCK / 64];
    if (ctl[0]) return;
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const unsigned nq = (unsigned)(L.N >> 2);
    const unsigned t = xcd_block(blockIdx.x, gridDim.x) * RED_BLOCK + threadIdx.x;
    double fmaxv = 0., rmax = 0.;
    if (t < nq) {
        const unsigned nz4 = (unsigned)nz >> 2;
        const unsigned line = t / nz4, k0 = (t - line * nz4) << 2;
        const unsigned i = line / (unsigned)ny, j = line - i * (unsigned)ny;
        const unsigned sx = (unsigned)ny * (unsigned)nz;
        const unsigned q = line * (unsigned)nz + k0;
        const unsigned ip = wrapi((int)i + 1, nx), im = wrapi((int)i - 1, nx);
        const unsigned jp = wrapi((int)j + 1, ny), jm = wrapi((int)j - 1, ny);
        double c[4], xp[4], xm[4], yp[4], ym[4];
        const unsigned kl = wrapi((int)k0 - 1, nz), kr = wrapi((int)k0 + 4, nz);
        const unsigned rowb = line * (unsigned)nz;
        double zl = phi_in[rowb + kl], zr = phi_in[rowb + kr];
        const unsigned nzh = (unsigned)nz >> 1, half = (unsigned)(L.N >> 1);
        if constexpr (SPLIT && std::is_same<R, float>::value) {
            load4<R>(phi_in, nullptr, q, c);
            load4<R>(phi_in, nullptr, q + (ip - i) * sx, xp);
            load4<R>(phi_in, nullptr, q + (im - i) * sx, xm);
            load4<R>(phi_in, nullptr, q + (jp - j) * (unsigned)nz, yp);
            load4<R>(phi_in, nullptr, q + (jm - j) * (unsigned)nz, ym);
            if (x) {
                split4_load(x, half, line, nzh, (i + j) & 1u, k0, c);
                split4_load(x, half, ip * (unsigned)ny + j, nzh, (ip + j) & 1u, k0, xp);
                split4_load(x, half, im * (unsigned)ny + j, nzh, (im + j) & 1u, k0, xm);
                split4_load(x, half, i * (unsigned)ny + jp, nzh, (i + jp) & 1u, k0, yp);
                split4_load(x, half, i * (unsigned)ny + jm, nzh, (i + jm) & 1u, k0, ym);
                const SplitIdx S{(unsigned)ny, nzh, half};
                zl += (double)x[S(i, j, kl)];
                zr += (double)x[S(i, j, kr)];
            }
        } else {
            load4<R>(phi_in, x, q, c);
            load4<R>(phi_in, x, q + (ip - i) * sx, xp);
            load4<R>(phi_in, x, q + (im - i) * sx, xm);
            load4<R>(phi_in, x, q + (jp - j) * (unsigned)nz, yp);
            load4<R>(phi_in, x, q + (jm - j) * (unsigned)nz, ym);
            if (x) {
                zl += (double)x[rowb + kl];
                zr += (double)x[rowb + kr];
            }
        }
        // per-axis coefficients (uniform constants or gathered)
        double cpx, cmx, ddx, cpy, cmy, ddy;
        if (L.c[0].uni) {
            cpx = cmx = L.c[0].ucp;
            ddx = L.c[0].udd;
        } else {
            cpx = L.c[0].cp[i];
            cmx = L.c[0].cm[i];
            ddx = L.c[0].dd[i];
        }
        if (L.c[1].uni) {
            cpy = cmy = L.c[1].ucp;
            ddy = L.c[1].udd;
        } else {
            cpy = L.c[1].cp[j];
            cmy = L.c[1].cm[j];
            ddy = L.c[1].dd[j];
        }
        double rh[4];
        if constexpr (std::is_same<RIN, float>::value) {
            const float4 r4 = *reinterpret_cast<const float4 *>(rho + q);
            rh[0] = r4.x;
            rh[1] = r4.y;
            rh[2] = r4.z;
            rh[3] = r4.w;
        } else {
            const double2 r0 = *reinterpret_cast<const double2 *>(rho + q);
            const double2 r1 = *reinterpret_cast<const double2 *>(rho + q + 2);
            rh[0] = r0.x;
            rh[1] = r0.y;
            rh[2] = r1.x;
            rh[3] = r1.y;
        }
        double F[4], ex[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double zp = u < 3 ? c[u + 1] : zr, zm = u > 0 ? c[u - 1] : zl;
            double cpz, cmz, ddz;
            if (L.c[2].uni) {
                cpz = cmz = L.c[2].ucp;
                ddz = L.c[2].udd;
            } else {
                cpz = L.c[2].cp[k0 + u];
                cmz = L.c[2].cm[k0 + u];
                ddz = L.c[2].dd[k0 + u];
            }
            double a, d;
            if (L.c[0].uni) a = cpx * (xp[u] + xm[u]);
            else a = cpx * xp[u] + cmx * xm[u];
            if (L.c[1].uni) a += cpy * (yp[u] + ym[u]);
            else a += cpy * yp[u] + cmy * ym[u];
            if (L.c[2].uni) a += cpz * (zp + zm);
            else a += cpz * zp + cmz * zm;
            d = ddx + ddy + ddz;
            ex[u] = exp(c[u]);
            F[u] = 2. * a - d * c[u] - ex[u] + rh[u];
            fmaxv = fmax(fmaxv, fabs(F[u]));
            rmax = fmax(rmax, fabs(rh[u]));
        }
        if (x) store4<double>(phi_out, q, c);
        const double nb[4] = {-F[0], -F[1], -F[2], -F[3]};
        if constexpr (SPLIT && std::is_same<R, float>::value) {
            split4_store(b, half, line, nzh, (i + j) & 1u, k0, nb);
            split4_store(e, half, line, nzh, (i + j) & 1u, k0, ex);
        } else {
            store4<R>(b, q, nb);
            store4<R>(e, q, ex);
        }
    }
    fmaxv = block_max(fmaxv, sh);
    rmax = block_max(rmax, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = fmaxv;
        part[2 * blockIdx.x + 1] = rmax;
        if (x && blockIdx.x == 0) ctl[3] = out_id;
    }
}

// the newest iterate into the result buffer when it sits in the other one
// (ctl[3] names the buffer of the newest iterate; `other` is the value of
// ctl[3] that means the other buffer)
__global__ void k_phi_final(double *__restrict__ phi, const double *__restrict__ alt, long n, const int *ctl,
                            int other)
{
    if (ctl[3] != other) return;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) phi[p] = alt[p];
}

// ctl: [0] done, [1] iterations, [2] converged, [3] buffer of the newest
// iterate (0 caller's, 1 alternate); stat: [0] rel resid
__global__ void __launch_bounds__(RED_BLOCK)
k_newton_check(const double *__restrict__ part, int nblk, int it, int max_it, double tol, int *ctl,
               int *fail, double *stat)
{
    __shared__ double sh[RED_BLOCK / 64];
    if (ctl[0]) return;
    double fm = 0., rm = 0.;
    for (int b = threadIdx.x; b < nblk; b += RED_BLOCK) {
        fm = fmax(fm, part[2 * b]);
        rm = fmax(rm, part[2 * b + 1]);
    }
    fm = block_max(fm, sh);
    rm = block_max(rm, sh);
    if (threadIdx.x == 0) {
        double rel = fm / (rm > 0. ? rm : 1.);
        stat[0] = rel;
        ctl[1] = it;
        if (rel <= tol || !(rel == rel)) {
            ctl[0] = 1;
            ctl[2] = (rel <= tol) ? 1 : 0;
            if (!(rel <= tol) && max_it > 0) fail[0] += 1;
        } else if (it >= max_it) {
            ctl[0] = 1;
            ctl[2] = 0;
            if (max_it > 0) fail[0] += 1;
        }
    }
}

template <typename R>
__global__ void k_fill(R *__restrict__ x, long n, R v, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) x[p] = v;
}

template <typename R>
__global__ void k_update(double *__restrict__ phi, const R *__restrict__ x, long n, const int *ctl)
{
    if (ctl[0]) return;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) phi[p] += (double)x[p];
}

// the diagonal D of the stencil at (i, j, k) (the sum of the axes' dd)
template <typename R>
__device__ __forceinline__ R diag(const LevelView<R> &L, int i, int j, int k)
{
    R d = L.c[0].uni ? L.c[0].udd : L.c[0].dd[i];
    d += L.c[1].uni ? L.c[1].udd : L.c[1].dd[j];
    d += L.c[2].uni ? L.c[2].udd : L.c[2].dd[k];
    return d;
}

// red-black Gauss-Seidel half sweep of  L x - e x = b  (all dims even)
template <typename R>
__global__ void k_rb(LevelView<R> L, R *__restrict__ x, const R *__restrict__ b,
                     const R *__restrict__ e, int color, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long q = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= L.N / 2) return;
    const int nz2 = L.n[2] / 2, ny = L.n[1];
    int i, j, kk;
    dec3(q, ny, nz2, i, j, kk);
    const int k = 2 * kk + ((i + j + color) & 1);
    R A, D;
    stencil<R>(L, x, i, j, k, A, D);
    const long p = ((long)i * ny + j) * L.n[2] + k;
    x[p] = (A - b[p]) / (D + e[p]);
}

// The red update of point (i, j, k) from x (its neighbours are all black)
template <typename R>
__device__ __forceinline__ R red_update(const LevelView<R> &L, const R *__restrict__ x, const R *__restrict__ b,
                                        const R *__restrict__ e, int i, int j, int k)
{
    R A, D;
    stencil<R>(L, x, i, j, k, A, D);
    const long p = ((long)i * L.n[1] + j) * L.n[2] + k;
    return (A - b[p]) / (D + e[p]);
}

// the red values of a sweep of x, recomputed where a black point reads them
template <typename R>
struct RedOf {
    const LevelView<R> &L;
    const R *__restrict__ x;
    const R *__restrict__ b;
    const R *__restrict__ e;
    __device__ __forceinline__ R operator()(unsigned q) const
    {
        const unsigned nz = (unsigned)L.n[2], ny = (unsigned)L.n[1];
        const unsigned r = q / nz, k = q - r * nz, i = r / ny, j = r - i * ny;
        return red_update<R>(L, x, b, e, (int)i, (int)j, (int)k);
    }
};

// One whole red-black sweep of x written to y (a second buffer): red points
// as k_rb colour 0 computes them, black points from those red values, each
// recomputed from x where a black point needs it.  The same arithmetic as
// the two in-place half sweeps, in one launch; for small levels, whose
// sweeps are launch-bound.
template <typename R>
__global__ void k_rb_full(LevelView<R> L, const R *__restrict__ x, R *__restrict__ y, const R *__restrict__ b,
                          const R *__restrict__ e, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long q = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= L.N / 2) return;
    const int nz2 = L.n[2] / 2, ny = L.n[1];
    int i, j, kk;
    dec3(q, ny, nz2, i, j, kk);
    const int kr = 2 * kk + ((i + j) & 1), kb = 2 * kk + ((i + j + 1) & 1);
    const long row = ((long)i * ny + j) * L.n[2];
    y[row + kr] = red_update<R>(L, x, b, e, i, j, kr);
    R A, D;
    stencil_ld<R>(L, RedOf<R>{L, x, b, e}, i, j, kb, A, D);
    y[row + kb] = (A - b[row + kb]) / (D + e[row + kb]);
}

// the red values of the first sweep from x = 0, -b / (D + e)
template <typename R>
struct RedFromZero {
    const LevelView<R> &L;
    const R *__restrict__ b;
    const R *__restrict__ e;
    __device__ __forceinline__ R operator()(unsigned q) const
    {
        const unsigned nz = (unsigned)L.n[2], ny = (unsigned)L.n[1];
        const unsigned r = q / nz, k = q - r * nz, i = r / ny, j = r - i * ny;
        return ((R)0 - b[q]) / (diag<R>(L, (int)i, (int)j, (int)k) + e[q]);
    }
};

// One whole sweep from x = 0 (k_rb_zero, then the black half sweep) in one
// launch: it reads no x, so it writes x in place
template <typename R>
__global__ void k_rb_full_zero(LevelView<R> L, R *__restrict__ x, const R *__restrict__ b, const R *__restrict__ e,
                               const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long q = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= L.N / 2) return;
    const int nz2 = L.n[2] / 2, ny = L.n[1];
    int i, j, kk;
    dec3(q, ny, nz2, i, j, kk);
    const int kr = 2 * kk + ((i + j) & 1), kb = 2 * kk + ((i + j + 1) & 1);
    const long row = ((long)i * ny + j) * L.n[2];
    x[row + kr] = ((R)0 - b[row + kr]) / (diag<R>(L, i, j, kr) + e[row + kr]);
    R A, D;
    stencil_ld<R>(L, RedFromZero<R>{L, b, e}, i, j, kb, A, D);
    x[row + kb] = (A - b[row + kb]) / (D + e[row + kb]);
}

// the first red half sweep from x = 0: red = -b / (D + e) (the sweep's own
// arithmetic with zero neighbours), black = 0; replaces zeroing x first
template <typename R>
__global__ void k_rb_zero(LevelView<R> L, R *__restrict__ x, const R *__restrict__ b, const R *__restrict__ e,
                          const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long q = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= L.N / 2) return;
    const int nz2 = L.n[2] / 2, ny = L.n[1];
    int i, j, kk;
    dec3(q, ny, nz2, i, j, kk);
    const int par = (i + j) & 1;
    const int k = 2 * kk + par;
    const R D = diag<R>(L, i, j, k);
    const long row = ((long)i * ny + j) * L.n[2];
    x[row + k] = ((R)0 - b[row + k]) / (D + e[row + k]);
    x[row + 2 * kk + 1 - par] = (R)0;
}

// weighted Jacobi sweep (levels with an odd extent): out = x + w((A-b)/(D+e) - x)
template <typename R>
__global__ void k_jacobi(LevelView<R> L, const R *__restrict__ x, R *__restrict__ out,
                         const R *__restrict__ b, const R *__restrict__ e, R w, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long p = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= L.N) return;
    const int nz = L.n[2], ny = L.n[1];
    int i, j, k;
        dec3(p, ny, nz, i, j, k);
    R A, D;
    stencil<R>(L, x, i, j, k, A, D);
    const R xn = (A - b[p]) / (D + e[p]);
    out[p] = x[p] + w * (xn - x[p]);
}

template <typename R>
__global__ void k_copy(R *__restrict__ dst, const R *__restrict__ src, long n, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long p = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[p];
}

// r = b - (L x - e x)
template <typename R>
__global__ void k_lin_resid(LevelView<R> L, const R *__restrict__ x, const R *__restrict__ b,
                            const R *__restrict__ e, R *__restrict__ r, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long p = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= L.N) return;
    const int nz = L.n[2], ny = L.n[1];
    int i, j, k;
        dec3(p, ny, nz, i, j, k);
    R A, D;
    stencil<R>(L, x, i, j, k, A, D);
    r[p] = b[p] - (A - D * x[p] - e[p] * x[p]);
}

// full weighting (1/4, 1/2, 1/4)^3, periodic; fine dims nf, coarse nf/2
template <typename R>
__global__ void k_restrict(const R *__restrict__ f, int fx, int fy, int fz, R *__restrict__ c,
                           const int *ctl)
{
    if (ctl && ctl[0]) return;
    const int cx = fx / 2, cy = fy / 2, cz = fz / 2;
    const long p = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= (long)cx * cy * cz) return;
    int i, j, k;
        dec3(p, cy, cz, i, j, k);
    const R w[3] = {(R)0.25, (R)0.5, (R)0.25};
    R acc = 0;
    for (int a = -1; a <= 1; ++a) {
        const int fi = wrapi(2 * i + a, fx);
        for (int bb = -1; bb <= 1; ++bb) {
            const int fj = wrapi(2 * j + bb, fy);
            for (int cc = -1; cc <= 1; ++cc) {
                const int fk = wrapi(2 * k + cc, fz);
                acc += w[a + 1] * w[bb + 1] * w[cc + 1] * f[((long)fi * fy + fj) * fz + fk];
            }
        }
    }
    c[p] = acc;
}

// fine += trilinear(index space) interpolation of coarse
template <typename R>
__global__ void k_prolong_add(const R *__restrict__ c, int cx, int cy, int cz, R *__restrict__ f,
                              const int *ctl)
{
    if (ctl && ctl[0]) return;
    const int fx = 2 * cx, fy = 2 * cy, fz = 2 * cz;
    const long p = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= (long)fx * fy * fz) return;
    int i, j, k;
        dec3(p, fy, fz, i, j, k);
    const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
    const int i1 = wrapi(i0 + 1, cx), j1 = wrapi(j0 + 1, cy), k1 = wrapi(k0 + 1, cz);
    const R wi1 = (i & 1) ? (R)0.5 : (R)0, wj1 = (j & 1) ? (R)0.5 : (R)0, wk1 = (k & 1) ? (R)0.5 : (R)0;
    const R wi0 = (R)1 - wi1, wj0 = (R)1 - wj1, wk0 = (R)1 - wk1;
#define CV(a, b, d) c[((long)(a)*cy + (b)) * cz + (d)]
    R v = wi0 * (wj0 * (wk0 * CV(i0, j0, k0) + wk1 * CV(i0, j0, k1)) +
                 wj1 * (wk0 * CV(i0, j1, k0) + wk1 * CV(i0, j1, k1))) +
          wi1 * (wj0 * (wk0 * CV(i1, j0, k0) + wk1 * CV(i1, j0, k1)) +
                 wj1 * (wk0 * CV(i1, j1, k0) + wk1 * CV(i1, j1, k1)));
#undef CV
    f[p] += v;
}

// ======================================================================== colour-split finest level
// In mixed precision, when the finest level is not solved by the coarse
// kernel and its extents are even (Nz a multiple of 4), its FP32 x, b and e
// are stored colour by colour: colour c = (i+j+k) & 1 at
// c*N/2 + (i*ny + j)*nz/2 + k/2.  A red-black half sweep then reads the other
// colour's x and this colour's b and e, and writes this colour's x, as whole
// cache lines (half the bytes of the interleaved layout).  The residual r
// stays in the natural layout for the restriction.

template <typename R>
__device__ __forceinline__ void stencil_split(const LevelView<R> &L, const R *__restrict__ x, const SplitIdx &S,
                                              int i, int j, int k, R &A, R &D)
{
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const unsigned ip = wrapi(i + 1, nx), im = wrapi(i - 1, nx);
    const unsigned jp = wrapi(j + 1, ny), jm = wrapi(j - 1, ny);
    const unsigned kp = wrapi(k + 1, nz), km = wrapi(k - 1, nz);
    const R xip = x[S(ip, j, k)], xim = x[S(im, j, k)];
    const R xjp = x[S(i, jp, k)], xjm = x[S(i, jm, k)];
    const R xkp = x[S(i, j, kp)], xkm = x[S(i, j, km)];
    R a, d;
    if (L.c[0].uni) {
        a = L.c[0].ucp * (xip + xim);
        d = L.c[0].udd;
    } else {
        a = L.c[0].cp[i] * xip + L.c[0].cm[i] * xim;
        d = L.c[0].dd[i];
    }
    if (L.c[1].uni) {
        a += L.c[1].ucp * (xjp + xjm);
        d += L.c[1].udd;
    } else {
        a += L.c[1].cp[j] * xjp + L.c[1].cm[j] * xjm;
        d += L.c[1].dd[j];
    }
    if (L.c[2].uni) {
        a += L.c[2].ucp * (xkp + xkm);
        d += L.c[2].udd;
    } else {
        a += L.c[2].cp[k] * xkp + L.c[2].cm[k] * xkm;
        d += L.c[2].dd[k];
    }
    A = (R)2 * a;
    D = d;
}

// half sweep of colour `color` on the split layout: thread q -> (i, j, kk)
template <typename R>
__global__ void k_rb_split(LevelView<R> L, R *__restrict__ x, const R *__restrict__ b, const R *__restrict__ e,
                           int color, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const unsigned half = (unsigned)(L.N >> 1);
    const unsigned q = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= half) return;
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const unsigned nzh = (unsigned)nz >> 1;
    const unsigned line = q / nzh, kk = q - line * nzh;
    const unsigned i = line / (unsigned)ny, j = line - i * (unsigned)ny;
    const unsigned par = (i + j + (unsigned)color) & 1u;
    const unsigned k = 2u * kk + par;
    const R *__restrict__ o = x + (1u - (unsigned)color) * half;
    const unsigned ip = wrapi((int)i + 1, nx), im = wrapi((int)i - 1, nx);
    const unsigned jp = wrapi((int)j + 1, ny), jm = wrapi((int)j - 1, ny);
    unsigned kp = kk + par, km = kk + par + nzh - 1u;
    if (kp >= nzh) kp -= nzh;
    if (km >= nzh) km -= nzh;
    const unsigned base = line * nzh;
    const R xip = o[(ip * (unsigned)ny + j) * nzh + kk], xim = o[(im * (unsigned)ny + j) * nzh + kk];
    const R xjp = o[(i * (unsigned)ny + jp) * nzh + kk], xjm = o[(i * (unsigned)ny + jm) * nzh + kk];
    const R xkp = o[base + kp], xkm = o[base + km];
    R a, d;
    if (L.c[0].uni) {
        a = L.c[0].ucp * (xip + xim);
        d = L.c[0].udd;
    } else {
        a = L.c[0].cp[i] * xip + L.c[0].cm[i] * xim;
        d = L.c[0].dd[i];
    }
    if (L.c[1].uni) {
        a += L.c[1].ucp * (xjp + xjm);
        d += L.c[1].udd;
    } else {
        a += L.c[1].cp[j] * xjp + L.c[1].cm[j] * xjm;
        d += L.c[1].dd[j];
    }
    if (L.c[2].uni) {
        a += L.c[2].ucp * (xkp + xkm);
        d += L.c[2].udd;
    } else {
        a += L.c[2].cp[k] * xkp + L.c[2].cm[k] * xkm;
        d += L.c[2].dd[k];
    }
    const unsigned c = (unsigned)color * half + base + kk;
    x[c] = ((R)2 * a - b[c]) / (d + e[c]);
}

// k_rb_zero on the split layout: colour 0 = -b / (D + e), colour 1 = 0
template <typename R>
__global__ void k_rb_split_zero(LevelView<R> L, R *__restrict__ x, const R *__restrict__ b,
                                const R *__restrict__ e, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const unsigned half = (unsigned)(L.N >> 1);
    const unsigned q = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= half) return;
    const int ny = L.n[1], nz = L.n[2];
    const unsigned nzh = (unsigned)nz >> 1;
    const unsigned line = q / nzh, kk = q - line * nzh;
    const unsigned i = line / (unsigned)ny, j = line - i * (unsigned)ny;
    const unsigned k = 2u * kk + ((i + j) & 1u);
    const R D = diag<R>(L, (int)i, (int)j, (int)k);
    x[q] = ((R)0 - b[q]) / (D + e[q]);
    x[half + q] = (R)0;
}

// r (natural) = b - (L x - e x), x, b, e split
template <typename R>
__global__ void k_lin_resid_split(LevelView<R> L, const R *__restrict__ x, const R *__restrict__ b,
                                  const R *__restrict__ e, R *__restrict__ r, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long p = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= L.N) return;
    const int nz = L.n[2], ny = L.n[1];
    int i, j, k;
    dec3(p, ny, nz, i, j, k);
    const SplitIdx S{(unsigned)ny, (unsigned)nz >> 1, (unsigned)(L.N >> 1)};
    R A, D;
    stencil_split<R>(L, x, S, i, j, k, A, D);
    const unsigned c = S(i, j, k);
    r[p] = b[c] - (A - D * x[c] - e[c] * x[c]);
}

// coarse correction added to the split finest level
template <typename R>
__global__ void k_prolong_add_split(const R *__restrict__ c, int cx, int cy, int cz, R *__restrict__ f,
                                    const int *ctl)
{
    if (ctl && ctl[0]) return;
    const int fx = 2 * cx, fy = 2 * cy, fz = 2 * cz;
    const long p = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= (long)fx * fy * fz) return;
    int i, j, k;
    dec3(p, fy, fz, i, j, k);
    const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
    const int i1 = wrapi(i0 + 1, cx), j1 = wrapi(j0 + 1, cy), k1 = wrapi(k0 + 1, cz);
    const R wi1 = (i & 1) ? (R)0.5 : (R)0, wj1 = (j & 1) ? (R)0.5 : (R)0, wk1 = (k & 1) ? (R)0.5 : (R)0;
    const R wi0 = (R)1 - wi1, wj0 = (R)1 - wj1, wk0 = (R)1 - wk1;
#define CV(a, b, d) c[((long)(a)*cy + (b)) * cz + (d)]
    R v = wi0 * (wj0 * (wk0 * CV(i0, j0, k0) + wk1 * CV(i0, j0, k1)) +
                 wj1 * (wk0 * CV(i0, j1, k0) + wk1 * CV(i0, j1, k1))) +
          wi1 * (wj0 * (wk0 * CV(i1, j0, k0) + wk1 * CV(i1, j0, k1)) +
                 wj1 * (wk0 * CV(i1, j1, k0) + wk1 * CV(i1, j1, k1)));
#undef CV
    const SplitIdx S{(unsigned)fy, (unsigned)fz >> 1, (unsigned)((long)fx * fy * fz >> 1)};
    f[S(i, j, k)] += v;
}

// full weighting of a split fine field (the finest level's e) to the natural coarse level
template <typename R>
__global__ void k_restrict_split(const R *__restrict__ f, int fx, int fy, int fz, R *__restrict__ c,
                                 const int *ctl)
{
    if (ctl && ctl[0]) return;
    const int cx = fx / 2, cy = fy / 2, cz = fz / 2;
    const long p = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= (long)cx * cy * cz) return;
    int i, j, k;
    dec3(p, cy, cz, i, j, k);
    const SplitIdx S{(unsigned)fy, (unsigned)fz >> 1, (unsigned)((long)fx * fy * fz >> 1)};
    const R w[3] = {(R)0.25, (R)0.5, (R)0.25};
    R acc = 0;
    for (int a = -1; a <= 1; ++a) {
        const int fi = wrapi(2 * i + a, fx);
        for (int bb = -1; bb <= 1; ++bb) {
            const int fj = wrapi(2 * j + bb, fy);
            for (int cc = -1; cc <= 1; ++cc) {
                const int fk = wrapi(2 * k + cc, fz);
                acc += w[a + 1] * w[bb + 1] * w[cc + 1] * f[S(fi, fj, fk)];
            }
        }
    }
    c[p] = acc;
}

// ======================================================================== coarse levels in one workgroup
// The tail of a V-cycle (every level with at most COARSE_N points) runs in a
// single workgroup: block-stride loops separated by barriers replace the
// many tiny launches those levels would otherwise cost (their launch latency
// exceeds their work; at 32^3 one workgroup is already slower than launches).
constexpr long COARSE_N = 4096;
constexpr int COARSE_MAXL = 8;
constexpr int COARSE_BLOCK = 1024;

template <typename R>
struct CoarseSet {
    int nl;
    LevelView<R> L[COARSE_MAXL];
    R *x[COARSE_MAXL], *b[COARSE_MAXL], *e[COARSE_MAXL], *r[COARSE_MAXL];
    int nu1, nu2, nuc;
    int zero0;  // start the first level from x = 0
    int wave;   // coarsest level of at most 64 points in LDS: one wavefront smooths it
};

// wavefront-scope barrier: the LDS writes of every lane are visible to the
// wavefront's later reads
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// blk_smooth on a level of at most 64 points whose x is in LDS, run by the
// first wavefront alone (call with threadIdx.x < 64): one point per lane,
// b and e held in registers, wavefront barriers between the phases.  Same
// updates in the same colour order as blk_smooth.
template <typename R>
__device__ void wave_smooth(const LevelView<R> &L, R *x, const R *b, const R *e, int sweeps)
{
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const bool rb = !(nx & 1) && !(ny & 1) && !(nz & 1);
    const int p = threadIdx.x;
    const bool act = p < L.N;
    int i = 0, j = 0, k = 0;
    if (act) dec3(p, ny, nz, i, j, k);
    const R bp = act ? b[p] : (R)0, ep = act ? e[p] : (R)0;
    const int col = (i + j + k) & 1;
    for (int sw = 0; sw < sweeps; ++sw) {
        if (rb) {
            for (int c = 0; c < 2; ++c) {
                if (act && col == c) {
                    R A, D;
                    stencil<R>(L, x, i, j, k, A, D);
                    x[p] = (A - bp) / (D + ep);
                }
                wave_sync();
            }
        } else {
            R v = (R)0;
            if (act) {
                R A, D;
                stencil<R>(L, x, i, j, k, A, D);
                v = x[p] + (R)0.8 * ((A - bp) / (D + ep) - x[p]);
            }
            wave_sync();
            if (act) x[p] = v;
            wave_sync();
        }
    }
}

template <typename R>
__device__ void blk_smooth(const LevelView<R> &L, R *x, const R *b, const R *e, R *r, int sweeps)
{
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const bool rb = !(nx & 1) && !(ny & 1) && !(nz & 1);
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int sw = 0; sw < sweeps; ++sw) {
        if (rb) {
            const int nz2 = nz / 2;
            for (int c = 0; c < 2; ++c) {
                for (long q = tid; q < L.N / 2; q += bs) {
                    int i, j, kk;
                    dec3(q, ny, nz2, i, j, kk);
                    const int k = 2 * kk + ((i + j + c) & 1);
                    R A, D;
                    stencil<R>(L, x, i, j, k, A, D);
                    const long p = ((long)i * ny + j) * nz + k;
                    x[p] = (A - b[p]) / (D + e[p]);
                }
                __syncthreads();
            }
        } else {
            for (long p = tid; p < L.N; p += bs) {
                int i, j, k;
        dec3(p, ny, nz, i, j, k);
                R A, D;
                stencil<R>(L, x, i, j, k, A, D);
                r[p] = x[p] + (R)0.8 * ((A - b[p]) / (D + e[p]) - x[p]);
            }
            __syncthreads();
            for (long p = tid; p < L.N; p += bs) x[p] = r[p];
            __syncthreads();
        }
    }
}

// levels with at most COARSE_LDS_N points keep x and b in LDS; together
// they hold less than 8/7 of the first such level (each has 1/8 of the
// points of the one above)
constexpr long COARSE_LDS_N = 4096;
constexpr long COARSE_LDS_PTS = COARSE_LDS_N + COARSE_LDS_N / 7 + 8;

// The coarse V-cycle.  IN_LDS: x and b of the levels with at most
// COARSE_LDS_N points and their residual live in LDS, so most of the ~100
// barrier-separated phases of the cycle wait on LDS rather than on L2
// (larger levels, and e of every level, stay in global memory).
template <typename R, bool IN_LDS>
__global__ void __launch_bounds__(COARSE_BLOCK) k_coarse_vcycle(CoarseSet<R> cs, const int *ctl)
{
    __shared__ R sx[IN_LDS ? COARSE_LDS_PTS : 1], sb[IN_LDS ? COARSE_LDS_PTS : 1], sr[IN_LDS ? COARSE_LDS_N : 1];
    __shared__ long soff[COARSE_MAXL + 1];  // LDS offset of a level, -1: global memory
    if (ctl[0]) return;
    const int tid = threadIdx.x, bs = blockDim.x;
    const int nl = cs.nl;
    if (tid == 0) {
        long o = 0;
        for (int l = 0; l < nl; ++l) {
            const bool in = IN_LDS && cs.L[l].N <= COARSE_LDS_N;
            soff[l] = in ? o : -1;
            if (in) o += cs.L[l].N;
        }
    }
    __syncthreads();
#define CX(l) (soff[l] >= 0 ? sx + soff[l] : cs.x[l])
#define CB(l) (soff[l] >= 0 ? sb + soff[l] : cs.b[l])
#define CR(l) (soff[l] >= 0 ? sr : cs.r[l])
    if (soff[0] >= 0) {
        for (long p = tid; p < cs.L[0].N; p += bs) {
            sb[p] = cs.b[0][p];
            sx[p] = cs.zero0 ? (R)0 : cs.x[0][p];
        }
    } else if (cs.zero0) {
        for (long p = tid; p < cs.L[0].N; p += bs) cs.x[0][p] = (R)0;
    }
    __syncthreads();
    for (int l = 0; l < nl - 1; ++l) {
        const LevelView<R> &L = cs.L[l];
        R *x = CX(l), *b = CB(l), *r = CR(l);
        const R *e = cs.e[l];
        blk_smooth<R>(L, x, b, e, r, cs.nu1);
        const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
        for (long p = tid; p < L.N; p += bs) {
            int i, j, k;
            dec3(p, ny, nz, i, j, k);
            R A, D;
            stencil<R>(L, x, i, j, k, A, D);
            r[p] = b[p] - (A - D * x[p] - e[p] * x[p]);
        }
        __syncthreads();
        const LevelView<R> &C = cs.L[l + 1];
        R *xc = CX(l + 1), *bc = CB(l + 1);
        const int cx = C.n[0], cy = C.n[1], cz = C.n[2];
        const R w[3] = {(R)0.25, (R)0.5, (R)0.25};
        for (long p = tid; p < C.N; p += bs) {
            int i, j, k;
            dec3(p, cy, cz, i, j, k);
            R acc = 0;
            for (int a = -1; a <= 1; ++a)
                for (int bb = -1; bb <= 1; ++bb)
                    for (int cc = -1; cc <= 1; ++cc) {
                        const int fi = wrapi(2 * i + a, nx), fj = wrapi(2 * j + bb, ny), fk = wrapi(2 * k + cc, nz);
                        acc += w[a + 1] * w[bb + 1] * w[cc + 1] * r[((long)fi * ny + fj) * nz + fk];
                    }
            bc[p] = acc;
            xc[p] = (R)0;
        }
        __syncthreads();
    }
    if (cs.wave && cs.L[nl - 1].N <= 64 && soff[nl - 1] >= 0) {
        if (tid < 64) wave_smooth<R>(cs.L[nl - 1], CX(nl - 1), CB(nl - 1), cs.e[nl - 1], cs.nuc);
        __syncthreads();
    } else {
        blk_smooth<R>(cs.L[nl - 1], CX(nl - 1), CB(nl - 1), cs.e[nl - 1], CR(nl - 1), cs.nuc);
    }
    for (int l = nl - 2; l >= 0; --l) {
        const LevelView<R> &L = cs.L[l], &C = cs.L[l + 1];
        const int cx = C.n[0], cy = C.n[1], cz = C.n[2];
        const int fy = L.n[1], fz = L.n[2];
        const R *c = CX(l + 1);
        R *x = CX(l);
        for (long p = tid; p < L.N; p += bs) {
            int i, j, k;
            dec3(p, fy, fz, i, j, k);
            const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
            const int i1 = wrapi(i0 + 1, cx), j1 = wrapi(j0 + 1, cy), k1 = wrapi(k0 + 1, cz);
            const R wi1 = (i & 1) ? (R)0.5 : (R)0, wj1 = (j & 1) ? (R)0.5 : (R)0, wk1 = (k & 1) ? (R)0.5 : (R)0;
            const R wi0 = (R)1 - wi1, wj0 = (R)1 - wj1, wk0 = (R)1 - wk1;
#define CV(a, b, d) c[((long)(a)*cy + (b)) * cz + (d)]
            R v = wi0 * (wj0 * (wk0 * CV(i0, j0, k0) + wk1 * CV(i0, j0, k1)) +
                         wj1 * (wk0 * CV(i0, j1, k0) + wk1 * CV(i0, j1, k1))) +
                  wi1 * (wj0 * (wk0 * CV(i1, j0, k0) + wk1 * CV(i1, j0, k1)) +
                         wj1 * (wk0 * CV(i1, j1, k0) + wk1 * CV(i1, j1, k1)));
#undef CV
            x[p] += v;
        }
        __syncthreads();
        blk_smooth<R>(L, x, CB(l), cs.e[l], CR(l), cs.nu2);
    }
    if (soff[0] >= 0) {
        for (long p = tid; p < cs.L[0].N; p += bs) cs.x[0][p] = sx[p];
    }
#undef CX
#undef CB
#undef CR
}

// restriction of the coefficient e through the coarse levels, one workgroup
template <typename R>
__global__ void __launch_bounds__(COARSE_BLOCK) k_coarse_restrict_e(CoarseSet<R> cs, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const int tid = threadIdx.x, bs = blockDim.x;
    const R w[3] = {(R)0.25, (R)0.5, (R)0.25};
    for (int l = 0; l < cs.nl - 1; ++l) {
        const LevelView<R> &L = cs.L[l], &C = cs.L[l + 1];
        const int nx = L.n[0], ny = L.n[1], nz = L.n[2], cx = C.n[0], cy = C.n[1], cz = C.n[2];
        for (long p = tid; p < C.N; p += bs) {
            int i, j, k;
        dec3(p, cy, cz, i, j, k);
            R acc = 0;
            for (int a = -1; a <= 1; ++a)
                for (int bb = -1; bb <= 1; ++bb)
                    for (int cc = -1; cc <= 1; ++cc) {
                        const int fi = wrapi(2 * i + a, nx), fj = wrapi(2 * j + bb, ny), fk = wrapi(2 * k + cc, nz);
                        acc += w[a + 1] * w[bb + 1] * w[cc + 1] * cs.e[l][((long)fi * ny + fj) * nz + fk];
                    }
            cs.e[l + 1][p] = acc;
        }
        __syncthreads();
    }
}

// ======================================================================== BiCGStab kernels
// scalar slots: 0 rho, 1 alpha, 2 omega, 3 beta, 4 dot result, 5 (rh,v), 6 (t,s), 7 (t,t), 8 rho_new
template <typename R>
__global__ void __launch_bounds__(RED_BLOCK)
k_dot_part(const R *__restrict__ a, const R *__restrict__ b, long n, double *__restrict__ part,
           const int *ctl)
{
    __shared__ double sh[RED_BLOCK / 64];
    if (ctl[0]) return;
    double acc = 0.;
    for (long p = (long)blockIdx.x * RED_BLOCK + threadIdx.x; p < n; p += (long)gridDim.x * RED_BLOCK)
        acc += (double)a[p] * (double)b[p];
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// (a, b) and (a, c) in one pass over a: the per-block partials of each in
// the order k_dot_part forms them, (a, b) at part[0, nb), (a, c) at
// part[nb, 2 nb)
template <typename R>
__global__ void __launch_bounds__(RED_BLOCK)
k_dot2_part(const R *__restrict__ a, const R *__restrict__ b, const R *__restrict__ c, long n,
            double *__restrict__ part, const int *ctl)
{
    __shared__ double sh[RED_BLOCK / 64];
    if (ctl[0]) return;
    double ab = 0., ac = 0.;
    for (long p = (long)blockIdx.x * RED_BLOCK + threadIdx.x; p < n; p += (long)gridDim.x * RED_BLOCK) {
        const double x = (double)a[p];
        ab += x * (double)b[p];
        ac += x * (double)c[p];
    }
    ab = block_sum(ab, sh);
    ac = block_sum(ac, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ab;
        part[gridDim.x + blockIdx.x] = ac;
    }
}

__device__ __forceinline__ double safe_div(double a, double b) { return b != 0. ? a / b : 0.; }

// op 0: init (rho=alpha=omega=1); op 1: beta = (rho_new/rho)(alpha/omega), rho = rho_new;
// op 2: alpha = rho/(rh,v); op 3: omega = (t,s)/(t,t)
__device__ void scal_apply(double *scal, int op)
{
    if (op == 0) {
        scal[0] = 1.;
        scal[1] = 1.;
        scal[2] = 1.;
    } else if (op == 1) {
        scal[3] = safe_div(scal[8], scal[0]) * safe_div(scal[1], scal[2]);
        scal[0] = scal[8];
    } else if (op == 2) {
        scal[1] = safe_div(scal[0], scal[5]);
    } else {
        scal[2] = safe_div(scal[6], scal[7]);
    }
}

__global__ void k_scal(double *scal, int op, const int *ctl)
{
    if (ctl[0] || threadIdx.x != 0) return;
    scal_apply(scal, op);
}

// the sums of nd partial sets of nb (set q at part[q nb, (q+1) nb)) into
// scal[slot + q], then the scalar update op (none when op < 0)
__global__ void __launch_bounds__(RED_BLOCK)
k_dot_final(const double *__restrict__ part, int nb, double *scal, int slot, const int *ctl, int nd = 1,
            int op = -1)
{
    __shared__ double sh[RED_BLOCK / 64];
    if (ctl[0]) return;
    for (int q = 0; q < nd; ++q) {
        double acc = 0.;
        for (int b = threadIdx.x; b < nb; b += RED_BLOCK) acc += part[q * nb + b];
        acc = block_sum(acc, sh);
        if (threadIdx.x == 0) scal[slot + q] = acc;
    }
    if (threadIdx.x == 0 && op >= 0) scal_apply(scal, op);
}

// p = r + beta (p - omega v)
template <typename R>
__global__ void k_bicg_p(R *__restrict__ p, const R *__restrict__ r, const R *__restrict__ v, long n,
                         const double *scal, const int *ctl)
{
    if (ctl[0]) return;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R beta = (R)scal[3], om = (R)scal[2];
    p[i] = r[i] + beta * (p[i] - om * v[i]);
}

// s = r - alpha v
template <typename R>
__global__ void k_bicg_s(R *__restrict__ s, const R *__restrict__ r, const R *__restrict__ v, long n,
                         const double *scal, const int *ctl)
{
    if (ctl[0]) return;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s[i] = r[i] - (R)scal[1] * v[i];
}

// x += alpha ph + omega sh ; r = s - omega t
template <typename R>
__global__ void k_bicg_xr(R *__restrict__ x, R *__restrict__ r, const R *__restrict__ ph,
                          const R *__restrict__ sh, const R *__restrict__ s, const R *__restrict__ t, long n,
                          const double *scal, const int *ctl)
{
    if (ctl[0]) return;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R al = (R)scal[1], om = (R)scal[2];
    x[i] = x[i] + al * ph[i] + om * sh[i];
    r[i] = s[i] - om * t[i];
}

// y = J x = L x - e x
template <typename R>
__global__ void k_apply_J(LevelView<R> L, const R *__restrict__ x, const R *__restrict__ e, R *__restrict__ y,
                          const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L.N) return;
    const int nz = L.n[2], ny = L.n[1];
    int i, j, k;
        dec3(p, ny, nz, i, j, k);
    R A, D;
    stencil<R>(L, x, i, j, k, A, D);
    y[p] = A - D * x[p] - e[p] * x[p];
}

// y = J x, as k_apply_J, with the per-block partials of (a, y), or of (y, c)
// and (y, y) when c is set, in the partition and order of k_dot_part /
// k_dot2_part (grid-stride over the same blocks)
template <typename R>
__global__ void __launch_bounds__(RED_BLOCK)
k_apply_J_dot(LevelView<R> L, const R *__restrict__ x, const R *__restrict__ e, R *__restrict__ y,
              const R *__restrict__ a, const R *__restrict__ c, long n, double *__restrict__ part, const int *ctl)
{
    __shared__ double sh[RED_BLOCK / 64];
    if (ctl[0]) return;
    const int nz = L.n[2], ny = L.n[1];
    double s1 = 0., s2 = 0.;
    for (long p = (long)blockIdx.x * RED_BLOCK + threadIdx.x; p < n; p += (long)gridDim.x * RED_BLOCK) {
        int i, j, k;
        dec3(p, ny, nz, i, j, k);
        R A, D;
        stencil<R>(L, x, i, j, k, A, D);
        const R yv = A - D * x[p] - e[p] * x[p];
        y[p] = yv;
        if (c) {
            const double t = (double)yv;
            s1 += t * (double)c[p];
            s2 += t * (double)yv;
        } else {
            s1 += (double)a[p] * (double)yv;
        }
    }
    s1 = block_sum(s1, sh);
    if (c) s2 = block_sum(s2, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s1;
        if (c) part[gridDim.x + blockIdx.x] = s2;
    }
}

// ======================================================================== reference NLJ kernels
// newton_method::solve (newton_method.hpp:17-63), f = A - D x - e^x
// the local function of solver_poisson_jacobi_nlin.cpp:190-199; with
// Boltzmann ions (the __FRONTEND__ build, theta > 0) exp(-theta x) is added
// to f but not to its derivative (cpp:192-195)
__device__ __forceinline__ double nlj_f(double x, double A, double D, double theta)
{
    double f = A - D * x - exp(x);
    if (theta != 0.) f += exp(-theta * x);
    return f;
}

__device__ __forceinline__ double newton_ref(double x, double A, double D, double theta)
{
    const double eps = 0.05;
    const int Nmax = 20;
    double omega = 1.;
    const double g = 0.8;
    double dx = 0.;
    double dx_alt = nlj_f(x, A, D, theta) / (-D - exp(x));
    int i = 0;
    do {
        x -= omega * dx;
        dx = nlj_f(x, A, D, theta) / (-D - exp(x));
        omega = fabs(dx) > fabs(dx_alt) ? g * omega : (1. - g) * omega + g * 1.;
        dx_alt = dx;
        ++i;
    } while (fabs(dx) > eps && i < Nmax);
    return x - dx;
}

// one iteration_loop sweep (solver_poisson_jacobi_nlin.cpp:170-212); the
// stencil uses the reference's divisions, in its order
__global__ void __launch_bounds__(RED_BLOCK)
k_nlj_ref(int nx, int ny, int nz, const double *__restrict__ HX, const double *__restrict__ HY,
          const double *__restrict__ HZ, const double *__restrict__ in, double *__restrict__ out,
          const double *__restrict__ rho, double omega, double theta, double *__restrict__ part)
{
    __shared__ double sh[RED_BLOCK / 64];
    const long p = (long)blockIdx.x * RED_BLOCK + threadIdx.x;
    const long N = (long)nx * ny * nz;
    double sup = -INFINITY, inf = INFINITY, nmax = 0., nsum = 0.;
    if (p < N) {
        int i, j, k;
        dec3(p, ny, nz, i, j, k);
        const int ip = wrapi(i + 1, nx), im = wrapi(i - 1, nx), jp = wrapi(j + 1, ny),
                  jm = wrapi(j - 1, ny), kp = wrapi(k + 1, nz), km = wrapi(k - 1, nz);
#define IN(a, b, c) in[((long)(a)*ny + (b)) * nz + (c)]
        double D = 2. / (HX[i + 1] * HX[i]);
        D += 2. / (HY[j + 1] * HY[j]);
        D += 2. / (HZ[k + 1] * HZ[k]);
        double A = IN(ip, j, k) / (HX[i + 1] * (HX[i + 1] + HX[i]));
        A += IN(im, j, k) / (HX[i] * (HX[i + 1] + HX[i]));
        A += IN(i, jp, k) / (HY[j + 1] * (HY[j + 1] + HY[j]));
        A += IN(i, jm, k) / (HY[j] * (HY[j + 1] + HY[j]));
        A += IN(i, j, kp) / (HZ[k + 1] * (HZ[k + 1] + HZ[k]));
        A += IN(i, j, km) / (HZ[k] * (HZ[k + 1] + HZ[k]));
#undef IN
        A = 2. * A;
        A += rho[p];
        const double P_old = in[p];
        const double P_new = newton_ref(P_old, A, D, theta);
        const double delta = fabs(P_new - P_old);
        out[p] = (1. - omega) * P_old + omega * P_new;
        sup = P_new;
        inf = P_new;
        nmax = delta;
        nsum = delta;
    }
    sup = block_max(sup, sh);
    inf = block_min(inf, sh);
    nmax = block_max(nmax, sh);
    nsum = block_sum(nsum, sh);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 0] = sup;
        part[4 * blockIdx.x + 1] = inf;
        part[4 * blockIdx.x + 2] = nmax;
        part[4 * blockIdx.x + 3] = nsum;
    }
}

// OP_8htoh / OP_4htoh / OP_2htoh_lvl0 (field_interpolation.cpp:233-315)
__global__ void k_inject(const double *__restrict__ in, int fx, int fy, int fz, int f,
                         double *__restrict__ out)
{
    const int cx = fx / f, cy = fy / f, cz = fz / f;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)cx * cy * cz) return;
    int i, j, k;
        dec3(p, cy, cz, i, j, k);
    out[p] = in[((long)(f * i) * fy + f * j) * fz + f * k];
}

// OP_hto2h (field_interpolation.cpp:56-147), one of its four stages.
// g[] holds gamma per fine index of the stage's axis (val_at based).
__global__ void k_hto2h(int stage, const double *__restrict__ in, int cx, int cy, int cz,
                        double *__restrict__ out, const double *__restrict__ g)
{
    const int nx = 2 * cx, ny = 2 * cy, nz = 2 * cz;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)nx * ny * nz) return;
    int i, j, k;
        dec3(p, ny, nz, i, j, k);
#define O(a, b, c) out[((long)(a)*ny + (b)) * nz + (c)]
    if (stage == 0) {
        if (!(i & 1) && !(j & 1) && !(k & 1)) O(i, j, k) = in[((long)(i / 2) * cy + j / 2) * cz + k / 2];
    } else if (stage == 1) {
        if (!(i & 1) && !(j & 1) && (k & 1))
            O(i, j, k) = (1. - g[k]) * O(i, j, wrapi(k - 1, nz)) + g[k] * O(i, j, wrapi(k + 1, nz));
    } else if (stage == 2) {
        if (!(i & 1) && (j & 1))
            O(i, j, k) = (1. - g[j]) * O(i, wrapi(j - 1, ny), k) + g[j] * O(i, wrapi(j + 1, ny), k);
    } else {
        if (i & 1)
            O(i, j, k) = (1. - g[i]) * O(wrapi(i - 1, nx), j, k) + g[i] * O(wrapi(i + 1, nx), j, k);
    }
#undef O
}

// ======================================================================== host
static inline unsigned nblocks(long n, int b) { return (unsigned)((n + b - 1) / b); }

// largest level swept by k_rb_full (paired whole sweeps): up to 64^3 the
// launch saved outweighs the recomputed red values; at 128^3 a whole sweep
// takes 17 us against 2 x 5.7 us for the two half sweeps
constexpr long FULL_SWEEP_N = 64L * 64 * 64;

Poisson::Poisson(const AxisDesc ax[3], bool mixed, hipStream_t s) : mixed_(mixed), s0_(s)
{
    // mixed mode: FP32 sweeps are cheap beside the FP64 residual and update
    // of a Newton iteration, and a third post sweep saves more than it costs
    // (measured at 256^3 over 500 steps: 15.06 ms/step against 15.25 with
    // two and two or three and three; FP64 mode is fastest with two and two)
    if (mixed_) {
        nu1_ = 2;
        nu2_ = 3;
    }
    for (int d = 0; d < 3; ++d) ax_[d] = ax[d];
    // the one-workgroup tail starts at 16^3 on equidistant grids and at 8^3
    // on stretched ones, where its stencil loads per-axis coefficients
    // (measured: a 16^3 start is 5% faster on a batch of 32-96^3 EqSt boxes;
    // an 8^3 start is 2% faster at 128^3 SiSt, 3.4% on a 64^3 HySt forward
    // run and the same at 256^3 EqSt)
    coarse_n_ = ax_desc_uniform() ? COARSE_N : 512;
    if (const char *e = std::getenv("KF_COARSE_N")) coarse_n_ = std::max(8L, std::min(COARSE_N, std::atol(e)));
    coarse_wave_ = std::getenv("KF_COARSE_WAVE") == nullptr || std::atoi(std::getenv("KF_COARSE_WAVE")) != 0;
    precond_direct_ = std::getenv("KF_PRECOND_COPY") == nullptr || std::atoi(std::getenv("KF_PRECOND_COPY")) == 0;
    int n[3] = {ax[0].n, ax[1].n, ax[2].n};
    for (int l = 0; l < 16; ++l) {
        Level L;
        for (int d = 0; d < 3; ++d) L.n[d] = n[d];
        L.N = (long)n[0] * n[1] * n[2];
        std::vector<double> c(9 * (size_t)std::max(n[0], std::max(n[1], n[2])));
        std::vector<double> h64;
        std::vector<double> coef;
        coef.reserve(3 * 3 * 256);
        for (int d = 0; d < 3; ++d) {
            AxisDesc a = ax[d];
            a.n = n[d];
            L.h[d].resize(n[d] + 1);
            for (int i = 0; i < n[d] + 1; ++i) L.h[d][i] = axis_val_at(a, i) - axis_val_at(a, i - 1);
        }
        size_t off = 0;
        std::vector<double> hc(3 * (size_t)(n[0] + n[1] + n[2]));
        for (int d = 0; d < 3; ++d) {
            const std::vector<double> &H = L.h[d];
            for (int i = 0; i < n[d]; ++i) {
                hc[off + i] = 1. / (H[i + 1] * (H[i + 1] + H[i]));
                hc[off + n[d] + i] = 1. / (H[i] * (H[i + 1] + H[i]));
                hc[off + 2 * n[d] + i] = 2. / (H[i + 1] * H[i]);
            }
            off += 3 * (size_t)n[d];
        }
        KF_HIP(hipMalloc(&L.c64, sizeof(double) * hc.size()));
        KF_HIP(hipMemcpy(L.c64, hc.data(), sizeof(double) * hc.size(), hipMemcpyHostToDevice));
        std::vector<float> hf(hc.begin(), hc.end());
        KF_HIP(hipMalloc(&L.c32, sizeof(float) * hf.size()));
        KF_HIP(hipMemcpy(L.c32, hf.data(), sizeof(float) * hf.size(), hipMemcpyHostToDevice));
        if (mixed_) {
            KF_HIP(hipMalloc(&L.x32, sizeof(float) * L.N));
            KF_HIP(hipMalloc(&L.b32, sizeof(float) * L.N));
            KF_HIP(hipMalloc(&L.e32, sizeof(float) * L.N));
            KF_HIP(hipMalloc(&L.r32, sizeof(float) * L.N));
        }
        KF_HIP(hipMalloc(&L.x64, sizeof(double) * L.N));
        KF_HIP(hipMalloc(&L.b64, sizeof(double) * L.N));
        KF_HIP(hipMalloc(&L.e64, sizeof(double) * L.N));
        KF_HIP(hipMalloc(&L.r64, sizeof(double) * L.N));
        lv_.push_back(L);
        bool can = (n[0] % 2 == 0) && (n[1] % 2 == 0) && (n[2] % 2 == 0) && n[0] / 2 >= 3 &&
                   n[1] / 2 >= 3 && n[2] / 2 >= 3;
        if (!can) break;
        n[0] /= 2;
        n[1] /= 2;
        n[2] /= 2;
    }
    long nb = nblocks(lv_[0].N, RED_BLOCK);
    KF_HIP(hipMalloc(&d_red_, sizeof(double) * 4 * (nb + 1)));
    KF_HIP(hipMalloc(&d_ctl_, sizeof(int) * 8));
    KF_HIP(hipMalloc(&d_fail_, sizeof(int) * 2));
    KF_HIP(hipMemset(d_fail_, 0, sizeof(int) * 2));
    KF_HIP(hipMalloc(&d_stat_, sizeof(double) * 4));
    KF_HIP(hipHostMalloc(&h_stat_, sizeof(double) * 4, hipHostMallocDefault));
    KF_HIP(hipMalloc(&d_scal_, sizeof(double) * 16));
    KF_HIP(hipEventCreateWithFlags(&ev_[0], hipEventDisableTiming));
    KF_HIP(hipEventCreateWithFlags(&ev_[1], hipEventDisableTiming));
    krylov_ = (ax[0].type != 1) || (ax[1].type != 1) || (ax[2].type != 1);
    // colour-split finest level (FP32 arrays of the mixed solver)
    split0_ = mixed_ && !krylov_ && lv_.size() > 1 && first_coarse() > 0 && lv_[0].n[0] % 2 == 0 &&
              lv_[0].n[1] % 2 == 0 && lv_[0].n[2] % 4 == 0;
    {
        const size_t es = mixed_ ? sizeof(float) : sizeof(double);
        for (auto &v : kv_) KF_HIP(hipMalloc(&v, es * lv_[0].N));
    }
    KF_HIP(hipHostMalloc(&h_ctl_, sizeof(int) * 8, hipHostMallocDefault));
    // the second potential buffer of the Newton update (also where a caller
    // puts an initial guess, guess_buffer()); allocated here so that nothing
    // allocates inside a solve
    KF_HIP(hipMalloc(&alt64_, sizeof(double) * lv_[0].N));
    KF_HIP(hipDeviceSynchronize());
}

Poisson::~Poisson()
{
    for (Level &L : lv_) {
        (void)hipFree(L.c64);
        (void)hipFree(L.c32);
        (void)hipFree(L.x64);
        (void)hipFree(L.b64);
        (void)hipFree(L.e64);
        (void)hipFree(L.r64);
        (void)hipFree(L.x32);
        (void)hipFree(L.b32);
        (void)hipFree(L.e32);
        (void)hipFree(L.r32);
    }
    (void)hipFree(d_red_);
    (void)hipFree(d_ctl_);
    (void)hipFree(d_fail_);
    (void)hipFree(d_stat_);
    (void)hipHostFree(h_stat_);
    (void)hipHostFree(h_ctl_);
    (void)hipFree(corr64_);
    (void)hipFree(alt64_);
    (void)hipFree(d_scal_);
    (void)hipFree(lsx_);
    (void)hipFree(lsr_);
    (void)hipFree(lsb_);
    (void)hipFree(lsT_);
    for (auto &v : vol_) (void)hipFree(v);
    (void)hipEventDestroy(ev_[0]);
    (void)hipEventDestroy(ev_[1]);
    for (auto &v : kv_) (void)hipFree(v);
}

template <typename R>
LevelView<R> Poisson::view(int l) const
{
    const Level &L = lv_[l];
    LevelView<R> v;
    for (int d = 0; d < 3; ++d) v.n[d] = L.n[d];
    v.N = L.N;
    const R *base = nullptr;
    if constexpr (std::is_same<R, double>::value) base = L.c64;
    else base = L.c32;
    size_t off = 0;
    for (int d = 0; d < 3; ++d) {
        v.c[d].cp = base + off;
        v.c[d].cm = base + off + L.n[d];
        v.c[d].dd = base + off + 2 * L.n[d];
        off += 3 * (size_t)L.n[d];
        // equidistant axis: h = L/n at every level, cp = cm = 1/(2h^2), dd = 2/h^2
        v.c[d].uni = ax_[d].type == 1 ? 1 : 0;
        const double h = ax_[d].L / (double)L.n[d];
        v.c[d].ucp = (R)(1. / (2. * h * h));
        v.c[d].udd = (R)(2. / (h * h));
    }
    return v;
}

template <typename R>
void Poisson::zero_level(int l, hipStream_t s)
{
    Level &L = lv_[l];
    R *x;
    if constexpr (std::is_same<R, double>::value) x = L.x64;
    else x = L.x32;
    if (l == 0 && x0o_) x = (R *)x0o_;
    hipLaunchKernelGGL(k_fill<R>, dim3(nblocks(L.N, 256)), dim3(256), 0, s, x, L.N, (R)0, (const int *)d_ctl_);
}

template <typename R>
void Poisson::smooth(int l, int sweeps, hipStream_t s, bool zero)
{
    Level &L = lv_[l];
    LevelView<R> v = view<R>(l);
    R *x, *b, *e, *r;
    if constexpr (std::is_same<R, double>::value) { x = L.x64; b = L.b64; e = L.e64; r = L.r64; }
    else { x = L.x32; b = L.b32; e = L.e32; r = L.r32; }
    if (l == 0 && x0o_) {
        x = (R *)x0o_;
        b = (R *)b0o_;  // read only
    }
    const bool rb = (L.n[0] % 2 == 0) && (L.n[1] % 2 == 0) && (L.n[2] % 2 == 0);
    const bool sp = split_level<R>(l);
    if (zero && (!rb || sweeps < 1)) {
        zero_level<R>(l, s);
        zero = false;
    }
    // coarse levels (not the finest, whose r the Newton and Krylov loops
    // use) up to 64^3: sweeps in pairs x -> r -> x, one launch each
    const bool fuse = rb && !sp && l > 0 && L.N <= FULL_SWEEP_N;
    for (int it = 0; it < sweeps; ++it) {
        if (fuse && zero && it == 0) {
            hipLaunchKernelGGL(k_rb_full_zero<R>, dim3(nblocks(L.N / 2, 256)), dim3(256), 0, s, v, x, (const R *)b,
                               (const R *)e, (const int *)d_ctl_);
            continue;
        }
        if (fuse && it + 1 < sweeps) {
            const unsigned nb = nblocks(L.N / 2, 256);
            const int *ctl = d_ctl_;
            hipLaunchKernelGGL(k_rb_full<R>, dim3(nb), dim3(256), 0, s, v, (const R *)x, r, (const R *)b,
                               (const R *)e, ctl);
            hipLaunchKernelGGL(k_rb_full<R>, dim3(nb), dim3(256), 0, s, v, (const R *)r, x, (const R *)b,
                               (const R *)e, ctl);
            ++it;
            continue;
        }
        if (rb) {
            for (int c = 0; c < 2; ++c) {
                prof::Scope ps_(l == 0 ? prof::RB0 : 1000, s);
                const unsigned nb = nblocks(L.N / 2, 256);
                const int *ctl = d_ctl_;
                if (zero && it == 0 && c == 0) {
                    if (sp)
                        hipLaunchKernelGGL(k_rb_split_zero<R>, dim3(nb), dim3(256), 0, s, v, x, (const R *)b,
                                           (const R *)e, ctl);
                    else
                        hipLaunchKernelGGL(k_rb_zero<R>, dim3(nb), dim3(256), 0, s, v, x, (const R *)b,
                                           (const R *)e, ctl);
                } else if (sp) {
                    hipLaunchKernelGGL(k_rb_split<R>, dim3(nb), dim3(256), 0, s, v, x, (const R *)b, (const R *)e,
                                       c, ctl);
                } else {
                    hipLaunchKernelGGL(k_rb<R>, dim3(nb), dim3(256), 0, s, v, x, b, e, c, ctl);
                }
            }
        } else {
            hipLaunchKernelGGL(k_jacobi<R>, dim3(nblocks(L.N, 256)), dim3(256), 0, s, v, (const R *)x, r,
                               (const R *)b, (const R *)e, (R)0.8, (const int *)d_ctl_);
            hipLaunchKernelGGL(k_copy<R>, dim3(nblocks(L.N, 256)), dim3(256), 0, s, x, (const R *)r, L.N,
                               (const int *)d_ctl_);
        }
    }
}

template <typename R>
CoarseSet<R> Poisson::coarse_set(int l0) const
{
    CoarseSet<R> cs{};
    cs.nl = (int)lv_.size() - l0;
    if (cs.nl > COARSE_MAXL) throw Error("too many coarse levels");
    for (int i = 0; i < cs.nl; ++i) {
        const Level &L = lv_[l0 + i];
        cs.L[i] = view<R>(l0 + i);
        if constexpr (std::is_same<R, double>::value) {
            cs.x[i] = L.x64; cs.b[i] = L.b64; cs.e[i] = L.e64; cs.r[i] = L.r64;
        } else {
            cs.x[i] = L.x32; cs.b[i] = L.b32; cs.e[i] = L.e32; cs.r[i] = L.r32;
        }
    }
    cs.nu1 = nu1_;
    cs.nu2 = nu2_;
    cs.nuc = nuc_;
    cs.wave = coarse_wave_ ? 1 : 0;
    return cs;
}

int Poisson::first_coarse() const
{
    for (int l = 0; l < (int)lv_.size(); ++l)
        if (lv_[l].N <= coarse_n_ && (int)lv_.size() - l <= COARSE_MAXL) return l;
    return (int)lv_.size() - 1;
}

// exp(phi) on every level: full weighting down the hierarchy
template <typename R>
void Poisson::restrict_e(hipStream_t s, const int *ctl)
{
    const int lc = first_coarse();
    for (int l = 0; l < lc; ++l) {
        Level &A = lv_[l], &B = lv_[l + 1];
        const R *ea;
        R *eb;
        if constexpr (std::is_same<R, double>::value) { ea = A.e64; eb = B.e64; }
        else { ea = A.e32; eb = B.e32; }
        if (split_level<R>(l))
            hipLaunchKernelGGL(k_restrict_split<R>, dim3(nblocks(B.N, 256)), dim3(256), 0, s, ea, A.n[0], A.n[1],
                               A.n[2], eb, ctl);
        else
            hipLaunchKernelGGL(k_restrict<R>, dim3(nblocks(B.N, 256)), dim3(256), 0, s, ea, A.n[0], A.n[1],
                               A.n[2], eb, ctl);
    }
    if (lc < (int)lv_.size() - 1)
        hipLaunchKernelGGL(k_coarse_restrict_e<R>, dim3(1), dim3(COARSE_BLOCK), 0, s, coarse_set<R>(lc), ctl);
}

template <typename R>
void Poisson::vcycle(int l, hipStream_t s, bool zero)
{
    const int nl = (int)lv_.size();
    if (l == first_coarse() && l < nl - 1) {
        CoarseSet<R> cs = coarse_set<R>(l);
        cs.zero0 = zero ? 1 : 0;
        long pts = 0;
        for (int m = 0; m < cs.nl; ++m)
            if (cs.L[m].N <= COARSE_LDS_N) pts += cs.L[m].N;
        if (pts <= COARSE_LDS_PTS)
            hipLaunchKernelGGL((k_coarse_vcycle<R, true>), dim3(1), dim3(COARSE_BLOCK), 0, s, cs, (const int *)d_ctl_);
        else
            hipLaunchKernelGGL((k_coarse_vcycle<R, false>), dim3(1), dim3(COARSE_BLOCK), 0, s, cs, (const int *)d_ctl_);
        return;
    }
    if (l == nl - 1) {
        smooth<R>(l, nuc_, s, zero);
        return;
    }
    Level &L = lv_[l], &C = lv_[l + 1];
    R *x, *b, *e, *r, *cx, *cb;
    if constexpr (std::is_same<R, double>::value) { x = L.x64; b = L.b64; e = L.e64; r = L.r64; cx = C.x64; cb = C.b64; }
    else { x = L.x32; b = L.b32; e = L.e32; r = L.r32; cx = C.x32; cb = C.b32; }
    if (l == 0 && x0o_) {
        x = (R *)x0o_;
        b = (R *)b0o_;  // read only
    }
    smooth<R>(l, nu1_, s, zero);
    LevelView<R> v = view<R>(l);
    const bool sp = split_level<R>(l);
    if (sp)
        hipLaunchKernelGGL(k_lin_resid_split<R>, dim3(nblocks(L.N, 256)), dim3(256), 0, s, v, (const R *)x,
                           (const R *)b, (const R *)e, r, (const int *)d_ctl_);
    else
        hipLaunchKernelGGL(k_lin_resid<R>, dim3(nblocks(L.N, 256)), dim3(256), 0, s, v, (const R *)x,
                           (const R *)b, (const R *)e, r, (const int *)d_ctl_);
    hipLaunchKernelGGL(k_restrict<R>, dim3(nblocks(C.N, 256)), dim3(256), 0, s, (const R *)r, L.n[0],
                       L.n[1], L.n[2], cb, (const int *)d_ctl_);
    vcycle<R>(l + 1, s, true);
    if (sp)
        hipLaunchKernelGGL(k_prolong_add_split<R>, dim3(nblocks(L.N, 256)), dim3(256), 0, s, (const R *)cx,
                           C.n[0], C.n[1], C.n[2], x, (const int *)d_ctl_);
    else
        hipLaunchKernelGGL(k_prolong_add<R>, dim3(nblocks(L.N, 256)), dim3(256), 0, s, (const R *)cx,
                           C.n[0], C.n[1], C.n[2], x, (const int *)d_ctl_);
    smooth<R>(l, nu2_, s);
}

// Newton iterations with FP64 residual/update and R-precision linear MG.
// newton_check(it): the FP64 residual at iterate it (iterate it >= 1 is
// iterate it-1 plus the correction of iteration it-1), the right side and
// coefficient of the next correction, and the device convergence test;
// the flag is copied to the host behind an event.
template <typename R>
void Poisson::newton_check(int it)
{
    NewtonRun &r = run_;
    Level &L0 = lv_[0];
    const long N = L0.N;
    // (the quad kernel has no Boltzmann-ion term: with ions, the generic one)
    const bool z4 = (L0.n[2] % 4) == 0 && ion_theta_ == 0.;
    const unsigned nbr = z4 ? nblocks(N / 4, RED_BLOCK) : nblocks(N, RED_BLOCK * NR_PTS);
    LevelView<double> v64 = view<double>(0);
    R *x0, *b0, *e0;
    if constexpr (std::is_same<R, double>::value) { x0 = L0.x64; b0 = L0.b64; e0 = L0.e64; }
    else { x0 = L0.x32; b0 = L0.b32; e0 = L0.e32; }
    hipStream_t s = r.s;
    const double *rho64 = r.rho64;
    const float *rho32 = r.rho32;
    const int *ctl = d_ctl_;
    // odd iterations write buffer B, even ones buffer A; the guess is in A
    double *A = r.A, *B = r.B;
    if (z4) {
        const double *src = (it == 0 || (it & 1)) ? A : B;
        double *dst = (it & 1) ? B : A;
        const R *xc = it == 0 ? nullptr : (const R *)x0;
        if (split_level<R>(0)) {
            if (rho64)
                hipLaunchKernelGGL((k_newton_z4<double, R, true>), dim3(nbr), dim3(RED_BLOCK), 0, s, v64, src, xc,
                                   dst, it & 1, rho64, b0, e0, d_red_, d_ctl_);
            else
                hipLaunchKernelGGL((k_newton_z4<float, R, true>), dim3(nbr), dim3(RED_BLOCK), 0, s, v64, src, xc,
                                   dst, it & 1, rho32, b0, e0, d_red_, d_ctl_);
        } else if (rho64) {
            hipLaunchKernelGGL((k_newton_z4<double, R>), dim3(nbr), dim3(RED_BLOCK), 0, s, v64, src, xc, dst,
                               it & 1, rho64, b0, e0, d_red_, d_ctl_);
        } else {
            hipLaunchKernelGGL((k_newton_z4<float, R>), dim3(nbr), dim3(RED_BLOCK), 0, s, v64, src, xc, dst,
                               it & 1, rho32, b0, e0, d_red_, d_ctl_);
        }
    } else if (it == 0) {
        if (rho64)
            hipLaunchKernelGGL((k_newton_resid<double, R>), dim3(nbr), dim3(RED_BLOCK), 0, s, v64, (const double *)A,
                               rho64, b0, e0, d_red_, ctl, ion_theta_);
        else
            hipLaunchKernelGGL((k_newton_resid<float, R>), dim3(nbr), dim3(RED_BLOCK), 0, s, v64, (const double *)A,
                               rho32, b0, e0, d_red_, ctl, ion_theta_);
    } else {
        const double *src = (it & 1) ? A : B;
        double *dst = (it & 1) ? B : A;
        if (rho64)
            hipLaunchKernelGGL((k_newton_upd_resid<double, R>), dim3(nbr), dim3(RED_BLOCK), 0, s, v64, src,
                               (const R *)x0, dst, it & 1, rho64, b0, e0, d_red_, d_ctl_, ion_theta_);
        else
            hipLaunchKernelGGL((k_newton_upd_resid<float, R>), dim3(nbr), dim3(RED_BLOCK), 0, s, v64, src,
                               (const R *)x0, dst, it & 1, rho32, b0, e0, d_red_, d_ctl_, ion_theta_);
    }
    hipLaunchKernelGGL(k_newton_check, dim3(1), dim3(RED_BLOCK), 0, s, (const double *)d_red_, (int)nbr, it,
                       r.max_it, tol_, d_ctl_, d_fail_, d_stat_);
    if (r.graph) return;
    KF_HIP(hipMemcpyAsync(h_ctl_ + 4 * (it & 1), d_ctl_, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
    KF_HIP(hipEventRecord(ev_[it & 1], s));
}

// the correction of one Newton iteration: J x = b by one linear V-cycle
// (or BiCGStab preconditioned by V-cycles on stretched grids)
template <typename R>
void Poisson::newton_correct()
{
    hipStream_t s = run_.s;
    restrict_e<R>(s, d_ctl_);
    if (krylov_) {
        bicgstab<R>(kiters_, s);
    } else {
        for (int c = 0; c < inner_; ++c) vcycle<R>(0, s, c == 0);
    }
}

// Enqueues the first check and the corrections up to the count the last
// solve needed (all of them in graph mode).  From a guess, the guess is in
// guess_buffer() (A) and the result goes to phi (B), so that a solve that
// stops after one correction, the usual case with an extrapolated guess,
// leaves the result where it belongs without a copy; in place, A is phi.
template <typename R>
void Poisson::newton_begin(double *phi, const double *rho64, const float *rho32, hipStream_t s, bool from_guess)
{
    NewtonRun &r = run_;
    r.phi = phi;
    r.rho64 = rho64;
    r.rho32 = rho32;
    r.s = s;
    r.from_guess = from_guess;
    r.A = from_guess ? alt64_ : phi;
    r.B = from_guess ? phi : alt64_;
    r.graph = graph_it_ > 0;
    r.max_it = r.graph ? graph_it_ : max_it_;
    r.it = 0;
    r.active = true;
    r.known = known_next_ && !r.graph;
    known_next_ = false;
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(8), 0, s, d_ctl_, 8L, 0, (const int *)nullptr);
    newton_check<R>(0);
    const int upto = r.graph ? r.max_it : (r.known ? 0 : std::min(std::max(1, expect_), r.max_it));
    while (r.it < upto) {
        newton_correct<R>();
        newton_check<R>(++r.it);
    }
    if (r.graph) newton_end<R>();
}

// From the last enqueued check on, the host waits for each check before
// enqueuing another correction.  (A correction enqueued after the converged
// check still launches every kernel of a V-cycle, and on the finest levels
// an early-exiting launch costs nearly as much as a working one.)
template <typename R>
void Poisson::newton_end()
{
    NewtonRun &r = run_;
    if (!r.active) return;
    r.active = false;
    const long N = lv_[0].N;
    if (!r.graph) {
        while (true) {
            KF_HIP(hipEventSynchronize(ev_[r.it & 1]));
            if (h_ctl_[4 * (r.it & 1)] || r.it >= r.max_it) break;
            newton_correct<R>();
            newton_check<R>(++r.it);
        }
        // ctl[1]: the iteration of the check that ended the solve
        if (!r.known) expect_ = std::max(1, h_ctl_[4 * (r.it & 1) + 1]);
        n_solves_ += 1;
        n_corr_ += h_ctl_[4 * (r.it & 1) + 1];
    }
    // result in phi: in place (A = phi) copy when the newest iterate is in B
    // (ctl[3] == 1); from a guess (B = phi) copy when it is in A (ctl[3] == 0)
    hipLaunchKernelGGL(k_phi_final, dim3(nblocks(N, 256)), dim3(256), 0, r.s, r.phi, (const double *)alt64_, N,
                       (const int *)d_ctl_, r.from_guess ? 0 : 1);
}

template <typename R>
void Poisson::dot(const R *a, const R *b, int slot, hipStream_t s, int op, const R *c)
{
    const long N = lv_[0].N;
    unsigned nb = nblocks(N, RED_BLOCK);
    if (nb > 1024) nb = 1024;
    if (c)
        hipLaunchKernelGGL(k_dot2_part<R>, dim3(nb), dim3(RED_BLOCK), 0, s, a, b, c, N, d_red_,
                           (const int *)d_ctl_);
    else
        hipLaunchKernelGGL(k_dot_part<R>, dim3(nb), dim3(RED_BLOCK), 0, s, a, b, N, d_red_, (const int *)d_ctl_);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_BLOCK), 0, s, (const double *)d_red_, (int)nb, d_scal_,
                       slot, (const int *)d_ctl_, c ? 2 : 1, op);
}

// y = J x with (a, y) into scalar slot, or (y, c), (y, y) into slot, slot + 1
// when c is set; then the scalar update op
template <typename R>
void Poisson::apply_J_dot(const R *x, R *y, const R *a, const R *c, int slot, int op, hipStream_t s)
{
    const long N = lv_[0].N;
    const R *e0;
    if constexpr (std::is_same<R, double>::value) e0 = lv_[0].e64;
    else e0 = lv_[0].e32;
    unsigned nb = nblocks(N, RED_BLOCK);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_apply_J_dot<R>, dim3(nb), dim3(RED_BLOCK), 0, s, view<R>(0), x, e0, y, a, c, N, d_red_,
                       (const int *)d_ctl_);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_BLOCK), 0, s, (const double *)d_red_, (int)nb, d_scal_,
                       slot, (const int *)d_ctl_, c ? 2 : 1, op);
}

// right-preconditioned BiCGStab (van der Vorst 1992) for J x = b, x0 = 0,
// M = one V-cycle from zero; scalars stay on the device
template <typename R>
void Poisson::bicgstab(int iters, hipStream_t s)
{
    Level &L0 = lv_[0];
    const long N = L0.N;
    const unsigned nb = nblocks(N, 256);
    const int *ctl = d_ctl_;
    R *x0, *b0, *e0;
    if constexpr (std::is_same<R, double>::value) { x0 = L0.x64; b0 = L0.b64; e0 = L0.e64; }
    else { x0 = L0.x32; b0 = L0.b32; e0 = L0.e32; }
    // a finest level launched on its own (not inside the one-workgroup
    // tail, whose pointers are its levels') reads and writes the Krylov
    // vectors in place of its b and x; its own b and x then hold the
    // residual and the iterate (b is consumed, x is the result)
    const bool direct = precond_direct_ && first_coarse() > 0 && !split_level<R>(0);
    R *X = direct ? x0 : (R *)kv_[0], *r = direct ? b0 : (R *)kv_[1], *rh = (R *)kv_[2], *p = (R *)kv_[3];
    R *v = (R *)kv_[4], *sv = (R *)kv_[5], *t = (R *)kv_[6], *ph = (R *)kv_[7], *shv = (R *)kv_[8];
    LevelView<R> v0 = view<R>(0);
    // r = b, rh = b, x = 0, p = v = 0
    if (!direct) hipLaunchKernelGGL(k_copy<R>, dim3(nb), dim3(256), 0, s, r, (const R *)b0, N, ctl);
    hipLaunchKernelGGL(k_copy<R>, dim3(nb), dim3(256), 0, s, rh, (const R *)b0, N, ctl);
    hipLaunchKernelGGL(k_fill<R>, dim3(nb), dim3(256), 0, s, X, N, (R)0, ctl);
    hipLaunchKernelGGL(k_fill<R>, dim3(nb), dim3(256), 0, s, p, N, (R)0, ctl);
    hipLaunchKernelGGL(k_fill<R>, dim3(nb), dim3(256), 0, s, v, N, (R)0, ctl);
    hipLaunchKernelGGL(k_scal, dim3(1), dim3(1), 0, s, d_scal_, 0, ctl);
    auto precond = [&](const R *in, R *out) {
        if (direct) {
            x0o_ = out;
            b0o_ = in;
            vcycle<R>(0, s, true);
            x0o_ = nullptr;
            b0o_ = nullptr;
            return;
        }
        hipLaunchKernelGGL(k_copy<R>, dim3(nb), dim3(256), 0, s, b0, in, N, ctl);
        vcycle<R>(0, s, true);
        hipLaunchKernelGGL(k_copy<R>, dim3(nb), dim3(256), 0, s, out, (const R *)x0, N, ctl);
    };
    for (int k = 0; k < iters; ++k) {
        if (direct) {
            dot<R>(rh, r, 8, s, 1);
        } else {
            dot<R>(rh, r, 8, s);
            hipLaunchKernelGGL(k_scal, dim3(1), dim3(1), 0, s, d_scal_, 1, ctl);
        }
        hipLaunchKernelGGL(k_bicg_p<R>, dim3(nb), dim3(256), 0, s, p, (const R *)r, (const R *)v, N,
                           (const double *)d_scal_, ctl);
        precond(p, ph);
        if (direct) {
            // v = J ph with (rh, v), then alpha
            apply_J_dot<R>(ph, v, rh, nullptr, 5, 2, s);
        } else {
            hipLaunchKernelGGL(k_apply_J<R>, dim3(nb), dim3(256), 0, s, v0, (const R *)ph, (const R *)e0, v, ctl);
            dot<R>(rh, v, 5, s);
            hipLaunchKernelGGL(k_scal, dim3(1), dim3(1), 0, s, d_scal_, 2, ctl);
        }
        hipLaunchKernelGGL(k_bicg_s<R>, dim3(nb), dim3(256), 0, s, sv, (const R *)r, (const R *)v, N,
                           (const double *)d_scal_, ctl);
        precond(sv, shv);
        if (direct) {
            // t = J shv with (t, s) and (t, t), then omega
            apply_J_dot<R>(shv, t, nullptr, sv, 6, 3, s);
        } else {
            hipLaunchKernelGGL(k_apply_J<R>, dim3(nb), dim3(256), 0, s, v0, (const R *)shv, (const R *)e0, t, ctl);
            dot<R>(t, sv, 6, s);
            dot<R>(t, t, 7, s);
            hipLaunchKernelGGL(k_scal, dim3(1), dim3(1), 0, s, d_scal_, 3, ctl);
        }
        hipLaunchKernelGGL(k_bicg_xr<R>, dim3(nb), dim3(256), 0, s, X, r, (const R *)ph, (const R *)shv,
                           (const R *)sv, (const R *)t, N, (const double *)d_scal_, ctl);
    }
    if (!direct) hipLaunchKernelGGL(k_copy<R>, dim3(nb), dim3(256), 0, s, x0, (const R *)X, N, ctl);
}

void Poisson::set_ion_boltzmann(double theta)
{
    if (theta < 0.) throw Error("the ion temperature ratio theta must be >= 0");
    if (theta != 0. && mixed_) throw Error("Boltzmann ions are an FP64 solve (the equilibrium start)");
    ion_theta_ = theta;
}

void Poisson::solve_begin(double *phi, const double *rho64, const float *rho32, hipStream_t s, bool from_guess)
{
    if (mode_ != POISSON_NEWTON_MG) {
        solve(phi, rho64, rho32, s, false, from_guess);
        return;
    }
    if (mixed_) newton_begin<float>(phi, rho64, rho32, s, from_guess);
    else newton_begin<double>(phi, rho64, rho32, s, from_guess);
}

void Poisson::solve_end()
{
    if (mixed_) newton_end<float>();
    else newton_end<double>();
}

int Poisson::solve(double *phi, const double *rho64, const float *rho32, hipStream_t s, bool sync, bool from_guess)
{
    prof::Scope ps_(prof::POISSON, s);
    if (mode_ == POISSON_NEWTON_MG) {
        solve_begin(phi, rho64, rho32, s, from_guess);
        solve_end();
        if (!sync) return 0;
        PoissonStats st = last_stats();
        return st.converged ? 0 : 1;
    }
    if (!rho64) throw Error("reference Poisson modes need an FP64 rho");
    if (from_guess)
        KF_HIP(hipMemcpyAsync(phi, alt64_, sizeof(double) * lv_[0].N, hipMemcpyDeviceToDevice, s));
    if (mode_ == POISSON_NLJ) {
        nlj_.max_iterations = max_it_;
        nlj_.limit_max = tol_;
        int r = nlj_solve_level(0, phi, rho64, s);
        return r < 0 ? r : 0;
    }
    int r = solve_cascade(phi, rho64, s);
    return r < 0 ? r : 0;
}

PoissonStats Poisson::last_stats()
{
    PoissonStats st{};
    if (mode_ != POISSON_NEWTON_MG) {
        st.cycles = nlj_.iteration;
        st.converged = nlj_.converged;
        st.resid = h_stat_[0];
        return st;
    }
    KF_HIP(hipMemcpyAsync(h_ctl_, d_ctl_, sizeof(int) * 4, hipMemcpyDeviceToHost, s0_));
    KF_HIP(hipMemcpyAsync(h_stat_, d_stat_, sizeof(double) * 2, hipMemcpyDeviceToHost, s0_));
    KF_HIP(hipStreamSynchronize(s0_));
    st.cycles = h_ctl_[1];
    st.converged = h_ctl_[2];
    st.resid = h_stat_[0];
    return st;
}

// solver_poisson_jacobi_nlin::solve (cpp:50-108) on level l's grid
int Poisson::nlj_solve_level(int level, double *phi, const double *rho, hipStream_t s)
{
    NljState &S = nlj_;
    int n[3];
    std::vector<double> H[3];
    if (level >= 0) {
        for (int d = 0; d < 3; ++d) { n[d] = lv_[level].n[d]; H[d] = lv_[level].h[d]; }
    } else {
        // negative level: -shift below the finest grid (cascade levels)
        for (int d = 0; d < 3; ++d) {
            AxisDesc a = ax_[d];
            a.n = ax_[d].n >> (-level);
            n[d] = a.n;
            H[d].resize(a.n + 1);
            for (int i = 0; i < a.n + 1; ++i) H[d][i] = axis_val_at(a, i) - axis_val_at(a, i - 1);
        }
    }
    const long N = (long)n[0] * n[1] * n[2];
    S.supremum = 0.;
    S.infinum = 0.;
    S.iteration = 0;
    S.omega_SOR = 1.;
    S.invocation++;
    S.converged = 0;
    S.norm_sum_old = -1.;
    if (S.max_iterations < 0) S.max_iterations = 1000;
    double *dH = nullptr, *tmp = nullptr;
    KF_HIP(hipMalloc(&dH, sizeof(double) * (n[0] + n[1] + n[2] + 3)));
    KF_HIP(hipMalloc(&tmp, sizeof(double) * N));
    KF_HIP(hipMemcpyAsync(dH, H[0].data(), sizeof(double) * (n[0] + 1), hipMemcpyHostToDevice, s));
    KF_HIP(hipMemcpyAsync(dH + n[0] + 1, H[1].data(), sizeof(double) * (n[1] + 1), hipMemcpyHostToDevice, s));
    KF_HIP(hipMemcpyAsync(dH + n[0] + n[1] + 2, H[2].data(), sizeof(double) * (n[2] + 1),
                          hipMemcpyHostToDevice, s));
    const unsigned nb = nblocks(N, RED_BLOCK);
    std::vector<double> part(4 * (size_t)nb);
    int rc = 0;
    while (!S.converged && (S.iteration < S.max_iterations)) {
        S.iteration++;
        double in0;
        KF_HIP(hipMemcpyAsync(&in0, phi, sizeof(double), hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(k_nlj_ref, dim3(nb), dim3(RED_BLOCK), 0, s, n[0], n[1], n[2], (const double *)dH,
                           (const double *)(dH + n[0] + 1), (const double *)(dH + n[0] + n[1] + 2),
                           (const double *)phi, tmp, rho, S.omega_SOR, ion_theta_, d_red_);
        KF_HIP(hipMemcpyAsync(part.data(), d_red_, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, s));
        KF_HIP(hipMemcpyAsync(phi, tmp, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
        KF_HIP(hipStreamSynchronize(s));
        // iteration_loop bookkeeping (cpp:154-159, 208-259)
        S.norm_sum_old = S.norm_sum;
        S.norm_max_old = S.norm_max;
        double sup = in0, inf = in0, nmax = 0., nsum = 0.;
        for (unsigned b = 0; b < nb; ++b) {
            sup = std::max(sup, part[4 * b]);
            inf = std::min(inf, part[4 * b + 1]);
            nmax = std::max(nmax, part[4 * b + 2]);
            nsum += part[4 * b + 3];
        }
        S.supremum = sup;
        S.infinum = inf;
        S.norm_max = nmax;
        S.norm_sum = nsum / (double)N;
        if ((S.norm_max < S.norm_max_old) && (S.iteration > 1)) {
            double lambda = S.norm_sum / S.norm_sum_old;
            double w = S.omega_SOR;
            S.omega_SOR = 2. / (1. + sqrt(1. - pow(lambda + w - 1., 2.) / (lambda * pow(w, 2.))));
        } else {
            S.omega_SOR = .5;
        }
        if (S.omega_SOR != S.omega_SOR) S.omega_SOR = .5;
        if (S.norm_max < S.limit_max * (S.supremum - S.infinum)) S.converged = 1;
        if (!(S.norm_max == S.norm_max)) { rc = -2; break; }  // NaN (cpp:95-99)
    }
    (void)hipFree(dH);
    (void)hipFree(tmp);
    h_stat_[0] = S.norm_max;
    return rc < 0 ? rc : S.iteration;
}

// solver_poisson_multigrid::set_level_control (cpp:82-147): the level
// program of the cascade.  change[c] < 0 goes down by -change[c] levels
// (injection by 2^-change[c]), +1 goes up one level (OP_hto2h), 0 stays.
// The level may never rise above the finest grid (cpp:126-145) and must
// end on it (cpp:406-410).
void Poisson::set_cascade(int n, const int *change, const int *steps, const double *tol)
{
    if (n < 1) throw Error("cascade needs at least one level step");
    int lvl = 0, depth = 0;
    for (int c = 0; c < n; ++c) {
        if (change[c] < -3 || change[c] > 1) throw Error("cascade level steps are -3..1 (down3 .. up)");
        lvl -= change[c];
        if (lvl < 0) throw Error("cascade level rises above the base level");
        depth = std::max(depth, lvl);
    }
    if (lvl != 0) throw Error("cascade must end on the base level");
    for (int d = 0; d < 3; ++d)
        if ((ax_[d].n % (1 << depth)) != 0 || (ax_[d].n >> depth) < 2)
            throw Error("cascade depth does not divide the grid extents");
    casc_change_.assign(change, change + n);
    casc_steps_.assign(steps, steps + n);
    casc_tol_.assign(tol, tol + n);
}

// solver_poisson_multigrid::solve (cpp:246-418) over the level program set
// by set_cascade (default: kielflow.cpp:207-217's {down3, up, up, up} with
// steps {0, -1, -1, -1} and tolerances {1e-2, 5e-4, 1e-4, 1e-4})
int Poisson::solve_cascade(double *phi, const double *rho, hipStream_t s)
{
    const int nx = ax_[0].n, ny = ax_[1].n, nz = ax_[2].n;
    int depth = 0, lv = 0;
    for (int c : casc_change_) depth = std::max(depth, lv -= c);
    if (nx % (1 << depth) || ny % (1 << depth) || nz % (1 << depth) || (std::min(nx, std::min(ny, nz)) >> depth) < 2)
        throw Error("the cascade's depth does not divide the grid extents");
    const long N = (long)nx * ny * nz;
    double *a = nullptr, *b = nullptr, *rl = nullptr, *g = nullptr;
    KF_HIP(hipMalloc(&a, sizeof(double) * N));
    KF_HIP(hipMalloc(&b, sizeof(double) * N));
    KF_HIP(hipMalloc(&rl, sizeof(double) * N));
    KF_HIP(hipMalloc(&g, sizeof(double) * (nx + ny + nz)));
    KF_HIP(hipMemcpyAsync(a, phi, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    int lvl = 0;
    int total = 0, rc = 0;
    for (size_t c = 0; c < casc_change_.size(); ++c) {
        const int step = casc_change_[c];
        const int cx = nx >> lvl, cy = ny >> lvl, cz = nz >> lvl;
        if (step < 0) {
            // injection by 2^-step from level lvl
            const int f = 1 << (-step);
            const long Nc = (long)(cx / f) * (cy / f) * (cz / f);
            hipLaunchKernelGGL(k_inject, dim3(nblocks(Nc, 256)), dim3(256), 0, s, (const double *)a, cx, cy, cz, f,
                               b);
            std::swap(a, b);
            lvl += -step;
        } else if (step > 0) {
            // OP_hto2h from level lvl to lvl-1
            lvl -= 1;
            const int fx = 2 * cx, fy = 2 * cy, fz = 2 * cz;
            std::vector<double> gam(fx + fy + fz, 0.);
            int ff[3] = {fx, fy, fz};
            size_t off = 0;
            for (int d = 0; d < 3; ++d) {
                AxisDesc A = ax_[d];
                A.n = ff[d];
                for (int i = 1; i < ff[d]; i += 2) {
                    double x1 = axis_val_at(A, i - 1), x2 = axis_val_at(A, i + 1), x = axis_val_at(A, i);
                    gam[off + i] = (x - x1) / (x2 - x1);
                }
                off += ff[d];
            }
            KF_HIP(hipMemcpyAsync(g, gam.data(), sizeof(double) * gam.size(), hipMemcpyHostToDevice, s));
            long Nf = (long)fx * fy * fz;
            const double *gz = g + fx + fy, *gy = g + fx, *gx = g;
            hipLaunchKernelGGL(k_hto2h, dim3(nblocks(Nf, 256)), dim3(256), 0, s, 0, (const double *)a, cx, cy,
                               cz, b, (const double *)nullptr);
            hipLaunchKernelGGL(k_hto2h, dim3(nblocks(Nf, 256)), dim3(256), 0, s, 1, (const double *)a, cx, cy,
                               cz, b, gz);
            hipLaunchKernelGGL(k_hto2h, dim3(nblocks(Nf, 256)), dim3(256), 0, s, 2, (const double *)a, cx, cy,
                               cz, b, gy);
            hipLaunchKernelGGL(k_hto2h, dim3(nblocks(Nf, 256)), dim3(256), 0, s, 3, (const double *)a, cx, cy,
                               cz, b, gx);
            // the host gamma table is read by the copy above before it goes out of scope
            KF_HIP(hipStreamSynchronize(s));
            std::swap(a, b);
        }
        // rho on this level: injection from the finest grid (cpp:284-359)
        const int f = 1 << lvl;
        long Nl = N >> (3 * lvl);
        hipLaunchKernelGGL(k_inject, dim3(nblocks(Nl, 256)), dim3(256), 0, s, rho, nx, ny, nz, f, rl);
        nlj_.max_iterations = casc_steps_[c];
        nlj_.limit_max = casc_tol_[c];
        int it = nlj_solve_level(-lvl, a, rl, s);
        if (it < 0) { rc = it; break; }
        total += it;
    }
    if (rc == 0) KF_HIP(hipMemcpyAsync(phi, a, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    KF_HIP(hipStreamSynchronize(s));
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(rl);
    (void)hipFree(g);
    return rc < 0 ? rc : total;
}

__global__ void k_exp_field(const double *__restrict__ phi, double *__restrict__ e, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) e[i] = exp(phi[i]);
}

// out = in * V (mode 1) or in / V (mode 0), V = vx[i] vy[j] vz[k]
__global__ void k_vscale(const double *__restrict__ in, double *__restrict__ out, const double *__restrict__ vx,
                         const double *__restrict__ vy, const double *__restrict__ vz, int nx, int ny, int nz,
                         int mode)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)nx * ny * nz) return;
    int i, j, k;
    dec3(p, ny, nz, i, j, k);
    const double v = vx[i] * vy[j] * vz[k];
    out[p] = mode ? in[p] * v : in[p] / v;
}

__global__ void __launch_bounds__(RED_BLOCK)
k_absmax2(const double *__restrict__ a, const double *__restrict__ b, long n, double *__restrict__ part,
          const int *ctl)
{
    __shared__ double sh[RED_BLOCK / 64];
    if (ctl && ctl[0]) return;
    double ma = 0., mb = 0.;
    for (long p = (long)blockIdx.x * RED_BLOCK + threadIdx.x; p < n; p += (long)gridDim.x * RED_BLOCK) {
        ma = fmax(ma, fabs(a[p]));
        mb = fmax(mb, fabs(b[p]));
    }
    ma = block_max(ma, sh);
    mb = block_max(mb, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ma;
        part[2 * blockIdx.x + 1] = mb;
    }
}

__global__ void k_axpy64(double *__restrict__ y, const double *__restrict__ x, long n, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// the stopping test of the linear defect correction on the device: stop
// (ctl[0] = 1) when max|r| <= tol max|b| or after max_it corrections;
// ctl[1]: corrections done, stat[0]: max|r| / max|b|
__global__ void __launch_bounds__(RED_BLOCK)
k_lin_check(const double *__restrict__ part, int nblk, int it, int max_it, double tol, int *ctl, double *stat)
{
    __shared__ double sh[RED_BLOCK / 64];
    if (ctl[0]) return;
    double rm = 0., bm = 0.;
    for (int b = threadIdx.x; b < nblk; b += RED_BLOCK) {
        rm = fmax(rm, part[2 * b]);
        bm = fmax(bm, part[2 * b + 1]);
    }
    rm = block_max(rm, sh);
    bm = block_max(bm, sh);
    if (threadIdx.x == 0) {
        stat[0] = rm / (bm > 0. ? bm : 1.);
        ctl[1] = it;
        if (rm <= tol * bm || it >= max_it) ctl[0] = 1;
    }
}

void Poisson::set_linearization(const double *phi, hipStream_t s)
{
    const long N = lv_[0].N;
    hipLaunchKernelGGL(k_exp_field, dim3(nblocks(N, 256)), dim3(256), 0, s, phi, lv_[0].e64, N);
}

int Poisson::linear_solve(double *x, const double *b, double rtol, int max_it, hipStream_t s, bool guess)
{
    Level &L0 = lv_[0];
    const long N = L0.N;
    const unsigned nb = nblocks(N, 256);
    unsigned nbr = nblocks(N, RED_BLOCK);
    if (nbr > RED_MAX_BLOCKS) nbr = RED_MAX_BLOCKS;
    if (!lsx_) {
        KF_HIP(hipMalloc(&lsx_, sizeof(double) * N));
        KF_HIP(hipMalloc(&lsr_, sizeof(double) * N));
        KF_HIP(hipMalloc(&lsb_, sizeof(double) * N));
    }
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(8), 0, s, d_ctl_, 8L, 0, (const int *)nullptr);
    restrict_e<double>(s, d_ctl_);
    if (guess) KF_HIP(hipMemcpyAsync(lsx_, x, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    else hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(256), 0, s, lsx_, N, 0., (const int *)nullptr);
    LevelView<double> v = view<double>(0);
    const int *ctl = d_ctl_;
    // residual and stopping test after `it` corrections; every kernel of a
    // correction launched after the test has stopped returns at once, so
    // the corrections the last solve needed are launched without waiting
    // for the host, and the result is that of a test after each one
    auto check = [&](int it) {
        hipLaunchKernelGGL(k_lin_resid<double>, dim3(nb), dim3(256), 0, s, v, (const double *)lsx_, b,
                           (const double *)L0.e64, lsr_, ctl);
        hipLaunchKernelGGL(k_absmax2, dim3(nbr), dim3(RED_BLOCK), 0, s, (const double *)lsr_, b, N, d_red_, ctl);
        hipLaunchKernelGGL(k_lin_check, dim3(1), dim3(RED_BLOCK), 0, s, (const double *)d_red_, (int)nbr, it, max_it,
                           rtol, d_ctl_, d_stat_);
    };
    check(0);
    int it = 0, upto = std::min(std::max(1, lin_expect_), max_it);
    for (;;) {
        for (; it < upto; ++it) {
            hipLaunchKernelGGL(k_copy<double>, dim3(nb), dim3(256), 0, s, L0.b64, (const double *)lsr_, N, ctl);
            if (krylov_) bicgstab<double>(kiters_, s);
            else vcycle<double>(0, s, true);
            hipLaunchKernelGGL(k_axpy64, dim3(nb), dim3(256), 0, s, lsx_, (const double *)L0.x64, N, ctl);
            check(it + 1);
        }
        KF_HIP(hipMemcpyAsync(h_ctl_, d_ctl_, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        KF_HIP(hipMemcpyAsync(h_stat_, d_stat_, sizeof(double), hipMemcpyDeviceToHost, s));
        KF_HIP(hipStreamSynchronize(s));
        if (h_ctl_[0] || it >= max_it) break;
        upto = it + 1;
    }
    it = h_ctl_[1];
    lin_expect_ = it;
    KF_HIP(hipMemcpyAsync(x, lsx_, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    lin_it_ = it;
    return (h_stat_[0] <= rtol) ? it : -1;
}

int Poisson::linear_solve_T(double *x, const double *b, double rtol, int max_it, hipStream_t s, bool guess)
{
    const bool uniform = ax_desc_uniform();
    if (uniform) return linear_solve(x, b, rtol, max_it, s, guess);
    Level &L0 = lv_[0];
    const long N = L0.N;
    if (!vol_[0]) {
        for (int d = 0; d < 3; ++d) {
            std::vector<double> v(L0.n[d]);
            for (int i = 0; i < L0.n[d]; ++i) v[i] = 0.5 * (L0.h[d][i + 1] + L0.h[d][i]);
            KF_HIP(hipMalloc(&vol_[d], sizeof(double) * v.size()));
            KF_HIP(hipMemcpy(vol_[d], v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
        }
    }
    if (!lsT_) KF_HIP(hipMalloc(&lsT_, sizeof(double) * N));
    double *bs = lsT_;
    hipLaunchKernelGGL(k_vscale, dim3(nblocks(N, 256)), dim3(256), 0, s, b, bs, (const double *)vol_[0],
                       (const double *)vol_[1], (const double *)vol_[2], L0.n[0], L0.n[1], L0.n[2], 0);
    if (guess)  // the iterate of J y = V^{-1} b is y = V^{-1} x
        hipLaunchKernelGGL(k_vscale, dim3(nblocks(N, 256)), dim3(256), 0, s, (const double *)x, x,
                           (const double *)vol_[0], (const double *)vol_[1], (const double *)vol_[2], L0.n[0],
                           L0.n[1], L0.n[2], 0);
    int it = linear_solve(x, bs, rtol, max_it, s, guess);
    hipLaunchKernelGGL(k_vscale, dim3(nblocks(N, 256)), dim3(256), 0, s, (const double *)x, x,
                       (const double *)vol_[0], (const double *)vol_[1], (const double *)vol_[2], L0.n[0],
                       L0.n[1], L0.n[2], 1);
    return it;
}

template void Poisson::newton_begin<float>(double *, const double *, const float *, hipStream_t, bool);
template void Poisson::newton_end<float>();
template void Poisson::bicgstab<float>(int, hipStream_t);
template void Poisson::bicgstab<double>(int, hipStream_t);
template void Poisson::newton_begin<double>(double *, const double *, const float *, hipStream_t, bool);
template void Poisson::newton_end<double>();

}  // namespace kf
