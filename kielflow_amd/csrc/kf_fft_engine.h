// kf_fft_engine.h -- LDS-resident batched 1D FFT for gfx950.
//
// A workgroup holds CB transforms of length N in LDS.  TPT threads cooperate
// on one transform; each owns E = N/TPT elements per pass.  Passes are
// Stockham autosort steps (natural order in, natural order out) with radices
// 8, 4, 2, 3 chosen at compile time.  Twiddles come from a per-length table
// w[m] = exp(-2 pi i m / N) held in LDS (loaded from a host-computed FP64
// table, so FP64 transforms carry correctly rounded twiddles).
//
// Element (transform c, position t) lives at lds[c*SC + t*ST]; SC/ST are the
// layout strides the caller chose (row layout for contiguous z transforms,
// column layout for strided x/y transforms).
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

// ---------------------------------------------------------------- one pass
// Stockham step of radix R_ at stride NS (product of earlier radices).
//   butterfly j (0 <= j < N/R_):  inputs  x[j + q*N/R_]
//   twiddle  w^(q * (j mod NS) * N/(NS*R_))
//   outputs  y[(j/NS)*NS*R_ + (j mod NS) + q*NS]
template <typename C, int N, int R_, int NS, int TPT, int SC, int ST, int DIR>
__device__ __forceinline__ void pass(C *lds, const C *tw, int c, int tt)
{
    constexpr int NB = N / R_;
    constexpr int MB = (NB + TPT - 1) / TPT;
    C v[MB][R_];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        const int j = tt + m * TPT;
        if (NB % TPT == 0 || j < NB) {
#pragma unroll
            for (int q = 0; q < R_; ++q) v[m][q] = lds[c * SC + (j + q * NB) * ST];
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
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        const int j = tt + m * TPT;
        if (NB % TPT == 0 || j < NB) {
            const int base = (j / NS) * NS * R_ + (j % NS);
#pragma unroll
            for (int q = 0; q < R_; ++q) lds[c * SC + (base + q * NS) * ST] = v[m][q];
        }
    }
    __syncthreads();
}

template <typename C, int N, int REM, int NS, int TPT, int SC, int ST, int DIR>
__device__ __forceinline__ void passes(C *lds, const C *tw, int c, int tt)
{
    if constexpr (REM > 1) {
        constexpr int R_ = next_radix(REM);
        pass<C, N, R_, NS, TPT, SC, ST, DIR>(lds, tw, c, tt);
        passes<C, N, REM / R_, NS * R_, TPT, SC, ST, DIR>(lds, tw, c, tt);
    }
}

// Full length-N transform of transform c by thread tt (0..TPT-1) of its group.
// Caller must __syncthreads() after filling LDS; this returns synchronised.
template <typename C, int N, int TPT, int SC, int ST, int DIR>
__device__ __forceinline__ void fft_lds(C *lds, const C *tw, int c, int tt)
{
    passes<C, N, N, 1, TPT, SC, ST, DIR>(lds, tw, c, tt);
}

}  // namespace fft
}  // namespace kf
