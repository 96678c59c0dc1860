// kf_fft.h -- 3D R2C / C2R transforms on kielflow's field layout, built
// from three LDS-staged passes:
//   forward R2C : z (real rows -> half spectrum)  ->  y  ->  x
//   inverse C2R : x  ->  y  ->  z (half spectrum -> real rows, 1/N scale)
// The z pass packs a real row of Nz values as Nz/2 complex values and
// runs an Nz/2-point transform plus the split/merge step; the y and x passes
// are strided complex transforms over CB columns per workgroup, so every
// global access is a contiguous run of CB complex values.
//
// Pass kernels take load/store functors, which is where the RHS kernels
// fuse spectral multipliers and real-space products into the transforms.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "kf_common.h"
#include "kf_fft_engine.h"
#include "kf_prof.h"

namespace kf {
namespace fft {

// threads per transform for a length-n transform (about 8 elements/thread)
__host__ __device__ constexpr int tpt_for(int n)
{
    return n <= 8 ? 1 : (n + 7) / 8;
}

// ------------------------------------------------------------------ functors
// Strided pass addressing: element t of transform q sits at
//   (q / qdiv) * qplane + (q % qdiv) + t * tstride
struct StridedAddr {
    long qdiv, qplane, tstride;
    __device__ __forceinline__ long base(long q) const
    {
        const unsigned u = (unsigned)q, d = (unsigned)qdiv;
        const unsigned r = u / d;
        return (long)r * qplane + (long)(u - r * d);
    }
    __device__ __forceinline__ long operator()(long q, int t) const { return base(q) + (long)t * tstride; }
};

template <typename R>
struct LoadPlain {
    const C2<R> *p;
    StridedAddr a;
    __device__ __forceinline__ C2<R> operator()(long q, int t) const { return p[a(q, t)]; }
};

template <typename R>
struct StorePlain {
    C2<R> *p;
    StridedAddr a;
    __device__ __forceinline__ void operator()(long q, int t, C2<R> v) const { p[a(q, t)] = v; }
};

// ------------------------------------------------------------------ kernels
// Adaptors between the engine's src.ld(t) / dst.st(t, v) and the pass
// kernels' functors: element t of transform q, or zero / no store for a
// transform past the end.  They refer to the kernel's functor argument
// rather than copying it: a copy whose arrays are indexed at run time would
// live in scratch memory.
template <typename C, class LD>
struct SrcQ {
    const LD &f;
    long q;
    bool live;
    __device__ __forceinline__ C ld(int t, int = 0) const
    {
        using R = decltype(C().x);
        return live ? f(q, t) : mk<C>((R)0, (R)0);
    }
};

// source of the first of two transforms of the same column: loads through
// the kernel's functor and keeps each input in the thread's slot
template <typename C, class LD>
struct SrcKeep {
    const LD &f;
    long q;
    bool live;
    C *keep;
    __device__ __forceinline__ C ld(int t, int slot) const
    {
        using R = decltype(C().x);
        const C v = live ? f(q, t) : mk<C>((R)0, (R)0);
        keep[slot] = v;
        return v;
    }
};

// source of the second: the kept inputs times i k[t] (k of the pass axis)
template <typename C>
struct SrcKeptIK {
    const C *keep;
    const decltype(C().x) *k;
    __device__ __forceinline__ C ld(int t, int slot) const
    {
        const C v = keep[slot];
        const auto kv = k[t];
        return mk<C>(-kv * v.y, kv * v.x);
    }
};
template <typename C, class ST_>
struct DstQ {
    const ST_ &f;
    long q;
    bool live;
    __device__ __forceinline__ void st(int t, C v, int = 0) const
    {
        if (live) f(q, t, v);
    }
};

// Strided complex pass: CB transforms of length N per workgroup; thread
// (c = tid % CB, t0 = tid / CB) works on transform q0 + c, so every global
// access of a wave covers CB consecutive complex values of each of 64/CB
// rows.  LDS: column layout, two buffers when they fit 32 KiB.
template <typename R, int N, int CB, int DIR, class LD, class ST_>
__global__ void __launch_bounds__(tpt_for(N) * CB)
k_strided(const C2<R> *__restrict__ tw_g, long Q, LD ld, ST_ st)
{
    using C = C2<R>;
    constexpr int TPT = tpt_for(N);
    constexpr int WG = TPT * CB;
    constexpr bool PING = 2 * N * CB * (int)sizeof(C) <= 32768;
    __shared__ C lds[(PING ? 2 : 1) * N * CB];
    __shared__ C tw[N];
    const int tid = threadIdx.x;
    const long q0 = (long)xcd_block(blockIdx.x, gridDim.x) * CB;
    for (int m = tid; m < N; m += WG) tw[m] = tw_g[m];
    const int c = tid % CB, t0 = tid / CB;
    const long q = q0 + c;
    const bool live = q < Q;
    const ColBuf<C, N, CB> A{lds, c}, B{lds + (PING ? N * CB : 0), c};
    fft_io<C, N, TPT, DIR, PING, false>(tw, t0, SrcQ<C, LD>{ld, q, live}, DstQ<C, ST_>{st, q, live}, A, B);
}

// Two strided transforms of every column of CB, sharing one load: out A =
// FFT(x), out B = FFT(i k[t] x) with k the wavenumbers of the pass axis (t
// the position along it).  The inputs of the first pass stay in registers,
// which roughly doubles the registers of a single pass; with one LDS buffer
// (no ping-pong) the narrow column block is bounded by registers alone.
// FP32: at least 6 waves per SIMD (80 VGPRs, no spills; the compiler's own
// choice is 81-82, which allows 5: the 256-point passes then take 65 and
// 44.6 us against 68.9 and 48.9).  FP64 is left to the compiler (the same
// bound spills at lengths 12 and 24).
template <typename R, int N, int CB, int DIR, class LD, class STA, class STB>
__global__ void __launch_bounds__(tpt_for(N) * CB, sizeof(R) == 4 ? 6 : 1)
k_strided_ik(const C2<R> *__restrict__ tw_g, long Q, LD ld, STA sta, STB stb, const R *__restrict__ kax)
{
    using C = C2<R>;
    constexpr int TPT = tpt_for(N);
    constexpr int WG = TPT * CB;
    constexpr bool PING = false;
    __shared__ C lds[(PING ? 2 : 1) * N * CB];
    __shared__ C tw[N];
    const int tid = threadIdx.x;
    const long q0 = (long)xcd_block(blockIdx.x, gridDim.x) * CB;
    for (int m = tid; m < N; m += WG) tw[m] = tw_g[m];
    const int c = tid % CB, t0 = tid / CB;
    const long q = q0 + c;
    const bool live = q < Q;
    const ColBuf<C, N, CB> A{lds, c}, B{lds + (PING ? N * CB : 0), c};
    C keep[in_slots(N, TPT)];
    fft_io<C, N, TPT, DIR, PING, false>(tw, t0, SrcKeep<C, LD>{ld, q, live, keep}, DstQ<C, STA>{sta, q, live}, A, B);
    __syncthreads();
    fft_io<C, N, TPT, DIR, PING, false>(tw, t0, SrcKeptIK<C>{keep, kax}, DstQ<C, STB>{stb, q, live}, A, B);
}

// Real rows -> half-spectrum rows (forward, unnormalised).
// in: rows of NZ = 2*M reals; out functor gets (row, k, X_k), k = 0..M.
template <typename R>
struct RowLoadReal {
    const R *p;
    int nz;
    __device__ __forceinline__ C2<R> operator()(long row, int n) const
    {
        return reinterpret_cast<const C2<R> *>(p + row * nz)[n];
    }
};

template <typename R>
struct RowStoreC {
    C2<R> *p;
    int nzc;
    __device__ __forceinline__ void operator()(long row, int k, C2<R> v) const
    {
        p[row * nzc + k] = v;
    }
};

template <typename R, int M, int RB, class LD, class ST_>
__global__ void __launch_bounds__(tpt_for(M) * RB)
k_zr2c(const C2<R> *__restrict__ twM_g, const C2<R> *__restrict__ twN_g, long rows, LD ld,
       ST_ st)
{
    using C = C2<R>;
    using Buf = RowBuf<C, M>;
    constexpr int TPT = tpt_for(M);
    constexpr int WG = TPT * RB;
    __shared__ C lds[RB * Buf::SC];
    __shared__ C tw[M];
    __shared__ C twN[M + 1];  // exp(-2 pi i k / (2M)), k = 0..M
    const int tid = threadIdx.x;
    const long r0 = (long)xcd_block(blockIdx.x, gridDim.x) * RB;
    for (int m = tid; m < M; m += WG) tw[m] = twM_g[m];
    for (int m = tid; m <= M; m += WG) twN[m] = twN_g[m];
    // row c's Nz/2-point transform by threads tt = 0..TPT-1, result in LDS
    const int c = tid / TPT, tt = tid % TPT;
    const long row = r0 + c;
    const Buf A{lds, c};
    fft_io<C, M, TPT, -1, false, true>(tw, tt, SrcQ<C, LD>{ld, row, row < rows}, A, A, A);
    __syncthreads();
    // split: X[k] = (Z[k] + conj Z[M-k])/2 - i/2 W^k (Z[k] - conj Z[M-k])
    for (int e = tid; e < RB * (M + 1); e += WG) {
        const int cc = e / (M + 1), k = e % (M + 1);
        const long rr = r0 + cc;
        if (rr < rows) {
            const Buf Z{lds, cc};
            C zk = Z.ld(k == M ? 0 : k);
            C zm = cconj(Z.ld(k == 0 ? 0 : M - k));
            C ev = mk<C>((R)0.5 * (zk.x + zm.x), (R)0.5 * (zk.y + zm.y));
            C od = mk<C>((R)0.5 * (zk.x - zm.x), (R)0.5 * (zk.y - zm.y));
            // -i * W^k * od
            C wod = cmul(twN[k], od);
            C X = mk<C>(ev.x + wod.y, ev.y - wod.x);
            st(rr, k, X);
        }
    }
}

// Half-spectrum rows -> real rows (inverse, unnormalised; callers scale in
// the store functor).  Im X_0 and Im X_M are dropped: this is C2R of the
// Hermitian projection of the kz = 0 and kz = Nz/2 planes once the x and y
// passes have run.
template <typename R>
struct RowLoadC {
    const C2<R> *p;
    int nzc;
    __device__ __forceinline__ C2<R> operator()(long row, int k) const
    {
        return p[row * nzc + k];
    }
};

template <typename R>
struct RowStoreReal {
    R *p;
    int nz;
    R scale;
    __device__ __forceinline__ void operator()(long row, int n, C2<R> v) const
    {
        reinterpret_cast<C2<R> *>(p + row * nz)[n] = mk<C2<R>>(v.x * scale, v.y * scale);
    }
};

// C2R merge at load: Z[k] = (X[k] + conj X[M-k]) + i conj(W^k) (X[k] - conj X[M-k])
// for k = 0..M-1, with Im X_0 and Im X_M dropped
template <typename C, int M, class LD>
struct MergeLoad {
    const LD &f;
    long row;
    bool live;
    const C *__restrict__ twN;
    __device__ __forceinline__ C ld(int k, int = 0) const
    {
        using R = decltype(C().x);
        if (!live) return mk<C>((R)0, (R)0);
        C xk = f(row, k), xm = f(row, M - k);
        if (k == 0) {
            xk.y = (R)0;
            xm.y = (R)0;
        }
        xm = cconj(xm);
        const C ev = cadd(xk, xm);
        const C od = cmul(cconj(twN[k]), csub(xk, xm));
        return mk<C>(ev.x - od.y, ev.y + od.x);
    }
};

// Whether a C2R z pass of half length M can form its merged inputs in
// registers: the first pass gives each of the row's TPT lanes one butterfly
// (its inputs t = tt + TPT q), and the row's lanes are a power of two within
// a wave, so X[M - t] sits in lane (TPT - tt) % TPT, slot R1 - 1 - q
__host__ __device__ constexpr bool merge_in_regs(int m)
{
    const int tpt = tpt_for(m);
    return in_slots(m, tpt) == m / tpt && m / tpt == next_radix(m) && (tpt & (tpt - 1)) == 0 && tpt <= 64;
}

// The merged first-pass inputs Z[t] = (X[t] + conj X[M-t]) + i conj(W^t)
// (X[t] - conj X[M-t]) of the thread's slots, with one load per element:
// the partner X[M-t] comes by lane shuffle (lane 0 of a row holds its own
// partners, and loads X[M]).  Im X_0 and Im X_M are dropped.
template <typename C, int M, class LD>
__device__ __forceinline__ void merge_regs(const LD &f, long row, bool live, int tt, const C *__restrict__ twN,
                                           C *z)
{
    using R = decltype(C().x);
    constexpr int TPT = tpt_for(M);
    constexpr int R1 = M / TPT;
    C xk[R1];
#pragma unroll
    for (int q = 0; q < R1; ++q) xk[q] = live ? f(row, tt + q * TPT) : mk<C>((R)0, (R)0);
    const C xm = (live && tt == 0) ? f(row, M) : mk<C>((R)0, (R)0);
    const int lane = (int)(threadIdx.x & 63);
    const int src = (lane - tt) + ((TPT - tt) & (TPT - 1));
#pragma unroll
    for (int q = 0; q < R1; ++q) {
        C p;
        p.x = __shfl(xk[R1 - 1 - q].x, src, 64);
        p.y = __shfl(xk[R1 - 1 - q].y, src, 64);
        C a = xk[q];
        if (tt == 0) {
            if (q == 0) {
                p = xm;
                a.y = (R)0;
                p.y = (R)0;
            } else {
                p = xk[R1 - q];
            }
        }
        p = cconj(p);
        const int t = tt + q * TPT;
        const C ev = cadd(a, p);
        const C od = cmul(cconj(twN[t]), csub(a, p));
        z[q] = mk<C>(ev.x - od.y, ev.y + od.x);
    }
}

// first-pass source from registers (slot = the thread's input slot)
template <typename C>
struct SrcRegs {
    const C *z;
    __device__ __forceinline__ C ld(int, int slot) const { return z[slot]; }
};

template <typename R, int M, int RB, class LD, class ST_>
__global__ void __launch_bounds__(tpt_for(M) * RB)
k_zc2r(const C2<R> *__restrict__ twM_g, const C2<R> *__restrict__ twN_g, long rows, LD ld,
       ST_ st)
{
    using C = C2<R>;
    using Buf = RowBuf<C, M>;
    constexpr int TPT = tpt_for(M);
    constexpr int WG = TPT * RB;
    __shared__ C lds[RB * Buf::SC];
    __shared__ C tw[M];
    const int tid = threadIdx.x;
    const long r0 = (long)xcd_block(blockIdx.x, gridDim.x) * RB;
    for (int m = tid; m < M; m += WG) tw[m] = twM_g[m];
    const int c = tid / TPT, tt = tid % TPT;
    const long row = r0 + c;
    const bool live = row < rows;
    const Buf A{lds, c};
    if constexpr (merge_in_regs(M)) {
        C z[M / TPT];
        merge_regs<C, M>(ld, row, live, tt, twN_g, z);
        fft_io<C, M, TPT, +1, false, false>(tw, tt, SrcRegs<C>{z}, DstQ<C, ST_>{st, row, live}, A, A);
    } else {
        fft_io<C, M, TPT, +1, false, false>(tw, tt, MergeLoad<C, M, LD>{ld, row, live, twN_g},
                                            DstQ<C, ST_>{st, row, live}, A, A);
    }
}

// ------------------------------------------------------------------ plan
// Twiddle tables for the lengths one grid needs, in one precision.
template <typename R>
struct Plan {
    int nx, ny, nz, nzc;
    // row length (complex) of the scratch field: Nz/2+1 rounded up to 128
    // bytes, so that the y pass reads and writes whole cache lines
    int nzs;
    long nr, nc;
    // column-block variant of the x and y passes, picked by the autotuner
    // at plan creation (0: cb_for columns per workgroup, 1: half of them)
    int vx = 0, vy = 0;
    C2<R> *tw_x = nullptr, *tw_y = nullptr, *tw_m = nullptr, *tw_n = nullptr;
    C2<R> *scratch = nullptr;  // nc complex, for C2R (input is preserved)
    void *owner_pool = nullptr;
};

template <typename R>
void plan_init(Plan<R> &p, int nx, int ny, int nz, hipStream_t s);
template <typename R>
void plan_free(Plan<R> &p);

// Supported lengths: x, y in {2,3,4,6,8,12,16,24,32,48,64,96,128,192,256};
// Nz even with Nz/2 in that set.
bool length_supported(int n);
bool grid_supported(int nx, int ny, int nz);

template <typename R>
void r2c(const Plan<R> &p, const R *in, C2<R> *out, hipStream_t s);
template <typename R>
void c2r(const Plan<R> &p, const C2<R> *in, R *out, hipStream_t s);

// ------------------------------------------------------------------ launchers
// Columns per workgroup for strided passes and rows per workgroup for z
// passes: aim at 256 (FP64) / 512 (FP32) threads and 32 KiB of LDS.
template <typename R>
__host__ __device__ constexpr int cb_for(int n)
{
    int v = (sizeof(R) == 4 ? 512 : 256) / tpt_for(n);
    return v < 8 ? 8 : (v > 64 ? 64 : v);
}

#define KF_FFT_LENGTHS(X) X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)

// the two column-block variants of a strided pass (results are identical:
// only the mapping of columns to workgroups differs)
template <typename R>
__host__ __device__ constexpr int cb_variant(int n, int v)
{
    return (v == 1 && cb_for<R>(n) / 2 >= 4) ? cb_for<R>(n) / 2 : cb_for<R>(n);
}

template <typename R, int DIR, class LD, class ST_>
void launch_strided(int n, long Q, const C2<R> *tw, LD ld, ST_ st, hipStream_t s, int fam, int var = 0)
{
    prof::Scope ps_(fam, s);
    switch (n) {
#define KF_CASE(NN)                                                                       \
    case NN: {                                                                            \
        constexpr int CB0 = cb_variant<R>(NN, 0), CB1 = cb_variant<R>(NN, 1);             \
        if (var == 1 && CB1 != CB0) {                                                     \
            dim3 grid((unsigned)((Q + CB1 - 1) / CB1));                                   \
            hipLaunchKernelGGL((k_strided<R, NN, CB1, DIR, LD, ST_>), grid,               \
                               dim3(tpt_for(NN) * CB1), 0, s, tw, Q, ld, st);             \
        } else {                                                                          \
            dim3 grid((unsigned)((Q + CB0 - 1) / CB0));                                   \
            hipLaunchKernelGGL((k_strided<R, NN, CB0, DIR, LD, ST_>), grid,               \
                               dim3(tpt_for(NN) * CB0), 0, s, tw, Q, ld, st);             \
        }                                                                                 \
        break;                                                                            \
    }
        KF_FFT_LENGTHS(KF_CASE)
#undef KF_CASE
    default:
        throw Error("unsupported transform length " + std::to_string(n));
    }
}

template <typename R, int DIR, class LD, class STA, class STB>
void launch_strided_ik(int n, long Q, const C2<R> *tw, LD ld, STA sta, STB stb, const R *kax, hipStream_t s,
                       int fam, int var = 0)
{
    prof::Scope ps_(fam, s);
    switch (n) {
#define KF_CASE(NN)                                                                            \
    case NN: {                                                                                 \
        constexpr int CB0 = cb_variant<R>(NN, 0), CB1 = cb_variant<R>(NN, 1);                  \
        if (var == 1 && CB1 != CB0) {                                                          \
            dim3 grid((unsigned)((Q + CB1 - 1) / CB1));                                        \
            hipLaunchKernelGGL((k_strided_ik<R, NN, CB1, DIR, LD, STA, STB>), grid,            \
                               dim3(tpt_for(NN) * CB1), 0, s, tw, Q, ld, sta, stb, kax);       \
        } else {                                                                               \
            dim3 grid((unsigned)((Q + CB0 - 1) / CB0));                                        \
            hipLaunchKernelGGL((k_strided_ik<R, NN, CB0, DIR, LD, STA, STB>), grid,            \
                               dim3(tpt_for(NN) * CB0), 0, s, tw, Q, ld, sta, stb, kax);       \
        }                                                                                      \
        break;                                                                                 \
    }
        KF_FFT_LENGTHS(KF_CASE)
#undef KF_CASE
    default:
        throw Error("unsupported FFT length " + std::to_string(n));
    }
}

template <typename R, class LD, class ST_>
void launch_zr2c(int m, long rows, const C2<R> *twm, const C2<R> *twn, LD ld, ST_ st,
                 hipStream_t s)
{
    prof::Scope ps_(prof::FFT_ZR2C, s);
    switch (m) {
#define KF_CASE(NN)                                                                       \
    case NN: {                                                                            \
        constexpr int RB = cb_for<R>(NN);                                                 \
        dim3 grid((unsigned)((rows + RB - 1) / RB));                                      \
        hipLaunchKernelGGL((k_zr2c<R, NN, RB, LD, ST_>), grid, dim3(tpt_for(NN) * RB), 0, s, \
                           twm, twn, rows, ld, st);                                       \
        break;                                                                            \
    }
        KF_FFT_LENGTHS(KF_CASE)
#undef KF_CASE
    default:
        throw Error("unsupported z length " + std::to_string(2 * m));
    }
}

template <typename R, class LD, class ST_>
void launch_zc2r(int m, long rows, const C2<R> *twm, const C2<R> *twn, LD ld, ST_ st,
                 hipStream_t s)
{
    prof::Scope ps_(prof::FFT_ZC2R, s);
    switch (m) {
#define KF_CASE(NN)                                                                       \
    case NN: {                                                                            \
        constexpr int RB = cb_for<R>(NN);                                                 \
        dim3 grid((unsigned)((rows + RB - 1) / RB));                                      \
        hipLaunchKernelGGL((k_zc2r<R, NN, RB, LD, ST_>), grid, dim3(tpt_for(NN) * RB), 0, s, \
                           twm, twn, rows, ld, st);                                       \
        break;                                                                            \
    }
        KF_FFT_LENGTHS(KF_CASE)
#undef KF_CASE
    default:
        throw Error("unsupported z length " + std::to_string(2 * m));
    }
}

// Strided passes over the padded scratch: columns kz >= nvalid are padding
// (loaded as zero, never stored)
template <typename R>
struct LoadMasked {
    const C2<R> *p;
    StridedAddr a;
    unsigned nvalid;
    __device__ __forceinline__ C2<R> operator()(long q, int t) const
    {
        if ((unsigned)q % (unsigned)a.qdiv >= nvalid) return mk<C2<R>>((R)0, (R)0);
        return p[a(q, t)];
    }
};

template <typename R>
struct StoreMasked {
    C2<R> *p;
    StridedAddr a;
    unsigned nvalid;
    __device__ __forceinline__ void operator()(long q, int t, C2<R> v) const
    {
        if ((unsigned)q % (unsigned)a.qdiv >= nvalid) return;
        p[a(q, t)] = v;
    }
};

// scratch-side addressing: y pass over Nx*nzs columns (padding masked);
// saddr_x: forward x pass over the Ny*Nzc columns of the unpadded spectral
// layout (its stores, to that layout, stay whole lines)
template <typename R>
inline StridedAddr saddr_y(const Plan<R> &p)
{
    return StridedAddr{p.nzs, (long)p.ny * p.nzs, p.nzs};
}
template <typename R>
inline StridedAddr saddr_x(const Plan<R> &p)
{
    return StridedAddr{p.nzc, p.nzs, (long)p.ny * p.nzs};
}
template <typename R>
inline long scratch_cols_y(const Plan<R> &p)
{
    return (long)p.nx * p.nzs;
}

// Inverse x pass from the [Nx][Ny][Nzc] layout into the scratch: it walks
// the Ny*nzs scratch columns (y, kz), so that a workgroup's CB consecutive
// columns are whole 128-byte lines of the scratch rows it writes (walking
// the Ny*Nzc unpadded columns instead puts most of its stores across two
// lines).  Columns kz >= Nzc are padding: no load, no store.
template <typename R, class LD>
struct PadColLoad {
    LD f;
    unsigned nzs, nzc;
    __device__ __forceinline__ C2<R> operator()(long q, int t) const
    {
        const unsigned u = (unsigned)q, r = u / nzs, k = u - r * nzs;
        if (k >= nzc) return mk<C2<R>>((R)0, (R)0);
        return f((long)(r * nzc + k), t);
    }
};
template <typename R, class LD>
inline PadColLoad<R, LD> pad_cols(const LD &f, const Plan<R> &p)
{
    return PadColLoad<R, LD>{f, (unsigned)p.nzs, (unsigned)p.nzc};
}
template <typename R>
inline StoreMasked<R> scratch_store_x(C2<R> *dst, const Plan<R> &p)
{
    return StoreMasked<R>{dst, StridedAddr{p.nzs, p.nzs, (long)p.ny * p.nzs}, (unsigned)p.nzc};
}
template <typename R>
inline long scratch_cols_x(const Plan<R> &p)
{
    return (long)p.ny * p.nzs;
}

// address helpers for the y and x passes on the [Nx][Ny][Nzc] layout
inline StridedAddr addr_y(int nx, int ny, int nzc)
{
    return StridedAddr{nzc, (long)ny * nzc, nzc};
}
inline StridedAddr addr_x(int nx, int ny, int nzc)
{
    return StridedAddr{(long)ny * nzc, 0, (long)ny * nzc};
}

}  // namespace fft
}  // namespace kf
