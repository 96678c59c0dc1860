// kf_fft_engine.h -- LDS-resident batched 1D FFT for gfx950.
//
// A workgroup runs several transforms of length N; TPT threads cooperate on
// one transform and each owns about N/TPT elements per pass.  Passes are
// Stockham autosort steps (natural order in, natural order out) with radices
// 8, 4, 2, 3 chosen at compile time.  The first pass takes its inputs
// straight from global memory into registers and the last pass stores
// straight from registers, so only the passes in between go through LDS.
// Twiddles come from a per-length table w[m] = exp(-2 pi i m / N) held in
// LDS (loaded from a host-computed FP64 table, so FP64 transforms carry
// correctly rounded twiddles).
#pragma once

#include "kf_common.h"

namespace kf {
namespace fft {

// ---------------------------------------------------------------- radix plan
__host__ __device__ constexpr int next_radix(int n)
{
    return (n % 8 == 0) ? 8 : (n % 4 == 0) ? 4 : (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : n;
}

// ---------------------------------------------------------------- small DFTs
// DIR = -1 forward (exp(-2 pi i/r)), +1 inverse.
template <int DIR, typename C>
__device__ __forceinline__ void dft2(C *v)
{
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int DIR, typename C>
__device__ __forceinline__ void dft3(C *v)
{
    using R = decltype(C().x);
    const R s = (R)0.86602540378443864676372317075294;  // sin(2 pi/3)
    C a = v[0], b = v[1], c = v[2];
    C t1 = cadd(b, c);
    C t2 = mk<C>(a.x - (R)0.5 * t1.x, a.y - (R)0.5 * t1.y);
    C d = csub(b, c);
    // forward: X1 = t2 - i s d, X2 = t2 + i s d
    C isd = mk<C>(-s * d.y, s * d.x);  // i*s*d
    v[0] = cadd(a, t1);
    if (DIR < 0) {
        v[1] = csub(t2, isd);
        v[2] = cadd(t2, isd);
    } else {
        v[1] = cadd(t2, isd);
        v[2] = csub(t2, isd);
    }
}

template <int DIR, typename C>
__device__ __forceinline__ void dft4(C *v)
{
    C a = v[0], b = v[1], c = v[2], d = v[3];
    C s0 = cadd(a, c), d0 = csub(a, c);
    C s1 = cadd(b, d), d1 = csub(b, d);
    // forward: w = -i
    C wd1 = cmuli<DIR>(d1);
    v[0] = cadd(s0, s1);
    v[2] = csub(s0, s1);
    v[1] = cadd(d0, wd1);
    v[3] = csub(d0, wd1);
}

template <int DIR, typename C>
__device__ __forceinline__ void dft8(C *v)
{
    using R = decltype(C().x);
    const R h = (R)0.70710678118654752440084436210485;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft4<DIR>(e);
    dft4<DIR>(o);
    // w8^s, s = 0..3 ; forward w8 = (1 - i)/sqrt2
    C o1 = o[1], o2 = o[2], o3 = o[3];
    C w1o1 = (DIR < 0) ? mk<C>(h * (o1.x + o1.y), h * (o1.y - o1.x))
                       : mk<C>(h * (o1.x - o1.y), h * (o1.y + o1.x));
    C w2o2 = cmuli<DIR>(o2);
    C w3o3 = (DIR < 0) ? mk<C>(h * (o3.y - o3.x), -h * (o3.x + o3.y))
                       : mk<C>(-h * (o3.x + o3.y), h * (o3.x - o3.y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1o1);
    v[5] = csub(e[1], w1o1);
    v[2] = cadd(e[2], w2o2);
    v[6] = csub(e[2], w2o2);
    v[3] = cadd(e[3], w3o3);
    v[7] = csub(e[3], w3o3);
}

template <int R_, int DIR, typename C>
__device__ __forceinline__ void dft(C *v)
{
    if constexpr (R_ == 2) dft2<DIR>(v);
    else if constexpr (R_ == 3) dft3<DIR>(v);
    else if constexpr (R_ == 4) dft4<DIR>(v);
    else if constexpr (R_ == 8) dft8<DIR>(v);
    else static_assert(R_ == 2, "unsupported radix");
}

// ---------------------------------------------------------------- passes
// Stockham step of radix R_ at stride NS (product of earlier radices).
//   butterfly j (0 <= j < N/R_):  inputs  x[j + q*N/R_]
//   twiddle  w^(q * (j mod NS) * N/(NS*R_))
//   outputs  y[(j/NS)*NS*R_ + (j mod NS) + q*NS]
// Inputs come from src.ld(t, slot), outputs go to dst.st(t, v, slot):
// global memory (the load/store functors of the kernel) for the first and
// the last pass, an LDS buffer in between; slot numbers a thread's inputs
// (< in_slots) and outputs (< out_slots) at compile time, for sources and
// destinations that keep them in registers.  All of a thread's loads are issued before its
// arithmetic.  SYNC_MID: src and dst are the same LDS buffer, so every thread
// must finish reading before any writes.
template <typename C, int N, int R_, int NS, int TPT, int DIR, bool SYNC_MID, class SRC, class DST>
__device__ __forceinline__ void pass_io(const C *tw, int tt, const SRC &src, const DST &dst)
{
    constexpr int NB = N / R_;
    constexpr int MB = (NB + TPT - 1) / TPT;
    C v[MB][R_];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        const int j = tt + m * TPT;
        if (NB % TPT == 0 || j < NB) {
#pragma unroll
            for (int q = 0; q < R_; ++q) v[m][q] = src.ld(j + q * NB, m * R_ + q);
        }
    }
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        const int j = tt + m * TPT;
        if (NB % TPT == 0 || j < NB) {
            if constexpr (NS > 1) {
                const int kk = j % NS;
#pragma unroll
                for (int q = 1; q < R_; ++q) {
                    C w = tw[(q * kk * (N / (NS * R_))) % N];
                    if (DIR > 0) w = cconj(w);
                    v[m][q] = cmul(v[m][q], w);
                }
            }
            dft<R_, DIR>(v[m]);
        }
    }
    if constexpr (SYNC_MID) __syncthreads();
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        const int j = tt + m * TPT;
        if (NB % TPT == 0 || j < NB) {
            const int base = (j / NS) * NS * R_ + (j % NS);
#pragma unroll
            for (int q = 0; q < R_; ++q) dst.st(base + q * NS, v[m][q], m * R_ + q);
        }
    }
}

// The passes of a length-N transform.  The first reads src, the last writes
// dst; in between, data moves through LDS buffers `in`/`out` (PING: two
// buffers used alternately, one barrier per pass; else one buffer, two
// barriers per pass).  LAST_LDS: dst is the single LDS buffer itself.
// The first pass uses no twiddles, so a table the caller fills in LDS is
// read only after the first barrier.  Returns with every LDS write of the
// chain issued but not synchronised.  FIRST_LDS: src is the single LDS
// buffer itself (the first pass synchronises between its reads and writes).
template <typename C, int N, int REM, int NS, int TPT, int DIR, bool PING, bool LAST_LDS, bool FIRST_LDS,
          class SRC, class DST, class BUF>
__device__ __forceinline__ void chain_io(const C *tw, int tt, const SRC &src, const DST &dst, const BUF &in,
                                         const BUF &out)
{
    static_assert(!(PING && LAST_LDS), "an LDS destination needs the single-buffer chain");
    static_assert(!(PING && FIRST_LDS), "an LDS source needs the single-buffer chain");
    constexpr int R_ = next_radix(REM);
    constexpr bool first = NS == 1, last = REM == R_;
    if constexpr (first && last) {
        pass_io<C, N, R_, NS, TPT, DIR, FIRST_LDS>(tw, tt, src, dst);
    } else if constexpr (first) {
        pass_io<C, N, R_, NS, TPT, DIR, FIRST_LDS>(tw, tt, src, out);
        __syncthreads();
        chain_io<C, N, REM / R_, NS * R_, TPT, DIR, PING, LAST_LDS, false>(tw, tt, src, dst, out, PING ? in : out);
    } else if constexpr (last) {
        pass_io<C, N, R_, NS, TPT, DIR, LAST_LDS>(tw, tt, in, dst);
    } else {
        pass_io<C, N, R_, NS, TPT, DIR, !PING>(tw, tt, in, out);
        __syncthreads();
        chain_io<C, N, REM / R_, NS * R_, TPT, DIR, PING, LAST_LDS, false>(tw, tt, src, dst, out, PING ? in : out);
    }
}

// Transform of length N by TPT threads (this one is tt): see chain_io.
// A and B are the LDS buffers (the same buffer unless PING).
template <typename C, int N, int TPT, int DIR, bool PING, bool LAST_LDS, bool FIRST_LDS = false, class SRC,
          class DST, class BUF>
__device__ __forceinline__ void fft_io(const C *tw, int tt, const SRC &src, const DST &dst, const BUF &A,
                                       const BUF &B)
{
    chain_io<C, N, N, 1, TPT, DIR, PING, LAST_LDS, FIRST_LDS>(tw, tt, src, dst, B, A);
}

// inputs per thread of the first pass of a length-n transform by tpt threads
__host__ __device__ constexpr int in_slots(int n, int tpt)
{
    const int r = next_radix(n);
    return ((n / r + tpt - 1) / tpt) * r;
}

// outputs per thread of the last pass of a length-n transform by tpt threads
__host__ __device__ constexpr int out_slots(int n, int tpt)
{
    int rem = n, r = 1;
    while (rem > 1) {
        r = next_radix(rem);
        if (rem == r) break;
        rem /= r;
    }
    return ((n / r + tpt - 1) / tpt) * r;
}

// ---------------------------------------------------------------- LDS layouts
// Bank-conflict-avoiding permutations of the positions t of one transform:
// t ^ ((t >> A) & MASK) (a bijection since MASK < 2^A), chosen per length
// and layout with a model of the gfx950 LDS banks over every pass's reads
// and writes.
__host__ __device__ constexpr int swz(int t, int a, int mask) { return t ^ ((t >> a) & mask); }

// column layout (strided passes): element t of column c at t*CB + c;
// RBYTES = CB * sizeof(C), the bytes of one LDS row
__host__ __device__ constexpr int col_swz_a(int n, int rbytes)
{
    return (n >= 128 && rbytes == 128) || (n == 256 && rbytes == 64) ? 3 : 1;
}
__host__ __device__ constexpr int col_swz_mask(int n, int rbytes)
{
    return (n >= 128 && rbytes == 128) ? 1 : (n == 256 && rbytes == 64) ? 3 : 0;
}

template <typename C, int N, int CB>
struct ColBuf {
    C *p;
    int c;
    static constexpr int A = col_swz_a(N, CB * (int)sizeof(C)), MASK = col_swz_mask(N, CB * (int)sizeof(C));
    __device__ __forceinline__ C ld(int t, int = 0) const { return p[swz(t, A, MASK) * CB + c]; }
    __device__ __forceinline__ void st(int t, C v, int = 0) const { p[swz(t, A, MASK) * CB + c] = v; }
};

// row layout (z passes): element t of row c at c*SC + swz(t); {pad, A, MASK}
// per half-length M and complex size
struct RowLay {
    int pad, a, mask;
};
__host__ __device__ constexpr RowLay row_layout(int m, int csize)
{
    if (csize == 8) {
        switch (m) {
        case 12: return {2, 2, 1};
        case 16: return {2, 3, 1};
        case 24: return {2, 3, 1};
        case 32: return {4, 3, 3};
        case 48: return {6, 5, 1};
        case 64: return {8, 3, 7};
        case 96: return {8, 3, 7};
        case 128: return {8, 4, 7};
        case 192: return {8, 4, 7};
        default: return {1, 1, 0};
        }
    }
    switch (m) {
    case 12: return {2, 2, 1};
    case 16: return {2, 3, 1};
    case 24: return {3, 4, 1};
    case 32: return {4, 3, 3};
    case 48: return {6, 4, 1};
    case 64: return {8, 3, 7};
    case 96: return {4, 3, 7};
    case 128: return {0, 3, 7};
    case 192: return {8, 3, 7};
    default: return {1, 1, 0};
    }
}

template <typename C, int M>
struct RowBuf {
    C *p;
    int c;
    static constexpr RowLay L = row_layout(M, (int)sizeof(C));
    static constexpr int SC = M + L.pad;
    __device__ __forceinline__ C ld(int t, int = 0) const { return p[c * SC + swz(t, L.a, L.mask)]; }
    __device__ __forceinline__ void st(int t, C v, int = 0) const { p[c * SC + swz(t, L.a, L.mask)] = v; }
};

}  // namespace fft
}  // namespace kf
