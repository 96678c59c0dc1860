// kf_fft.hip -- plans and plain 3D R2C / C2R drivers (OP_FFT / OP_iFFT
// semantics: OP_FFT.cpp:49-63 unnormalised forward, OP_iFFT.cpp:49-64 inverse
// scaled by 1/N; the caller's input is preserved in both directions).
#include <cstring>

#include "kf_fft.h"

#include <algorithm>
#include <cstdlib>

#include <cmath>

namespace kf {
namespace fft {

static bool in_set(int n)
{
    switch (n) {
    case 2: case 3: case 4: case 6: case 8: case 12: case 16: case 24: case 32:
    case 48: case 64: case 96: case 128: case 192: case 256:
        return true;
    default:
        return false;
    }
}

bool length_supported(int n) { return in_set(n); }

bool grid_supported(int nx, int ny, int nz)
{
    return in_set(nx) && in_set(ny) && nz % 2 == 0 && in_set(nz / 2);
}

// w[m] = exp(-2 pi i m / n), m = 0..count-1, rounded from extended precision
template <typename R>
static C2<R> *twiddles(int n, int count, hipStream_t s)
{
    std::vector<C2<R>> h(count);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < count; ++m) {
        long double a = -2.0L * pi * (long double)m / (long double)n;
        h[m].x = (R)cosl(a);
        h[m].y = (R)sinl(a);
    }
    C2<R> *d = nullptr;
    KF_HIP(hipMalloc(&d, sizeof(C2<R>) * count));
    KF_HIP(hipMemcpyAsync(d, h.data(), sizeof(C2<R>) * count, hipMemcpyHostToDevice, s));
    KF_HIP(hipStreamSynchronize(s));
    return d;
}

// Autotuner: time the column-block variants of the x and y passes, and the
// y pass with and without the padded scratch row, on the plan's scratch
// field (median of 7 after a warm-up launch).  A choice other than the
// default (variant 0, padded rows) is kept only when it is at least 5%
// faster: the timed passes carry plain loads, and the fused loads of the RHS
// favour the wider column blocks, so near-ties go to the default.
// KF_FFT_AUTOTUNE=0 keeps variant 0 and the padded rows; KF_FFT_AUTOTUNE=alt
// takes variant 1 wherever it differs and unpadded rows (the other end of
// every choice, for tests).  Every choice computes the same numbers; only
// the speed changes.
template <typename R>
static void autotune(Plan<R> &p, hipStream_t s)
{
    p.vx = p.vy = 0;
    const char *env = std::getenv("KF_FFT_AUTOTUNE");
    if (env && env[0] == '0') return;
    if (env && std::strcmp(env, "alt") == 0) {
        p.vx = cb_variant<R>(p.nx, 1) != cb_variant<R>(p.nx, 0) ? 1 : 0;
        p.vy = cb_variant<R>(p.ny, 1) != cb_variant<R>(p.ny, 0) ? 1 : 0;
        p.nzs = p.nzc;
        return;
    }
    KF_HIP(hipMemsetAsync(p.scratch, 0, sizeof(C2<R>) * (size_t)p.nx * p.ny * p.nzs, s));
    hipEvent_t e0, e1;
    KF_HIP(hipEventCreate(&e0));
    KF_HIP(hipEventCreate(&e1));
    StridedAddr ay = saddr_y(p), ax = saddr_x(p);
    auto time_pass = [&](int pass, int var) {
        float ms[7];
        for (int r = -1; r < 7; ++r) {
            KF_HIP(hipEventRecord(e0, s));
            if (pass == 0) {
                // the x passes as the transforms run them: inverse from the
                // unpadded layout into the scratch, forward back out of it
                // (in place on the zeroed scratch: every value stays zero)
                const StridedAddr ux = addr_x(p.nx, p.ny, p.nzc);
                launch_strided<R, +1>(p.nx, scratch_cols_x(p), p.tw_x, pad_cols(LoadPlain<R>{p.scratch, ux}, p),
                                      scratch_store_x(p.scratch, p), s, 1000, var);
                launch_strided<R, -1>(p.nx, (long)p.ny * p.nzc, p.tw_x, LoadPlain<R>{p.scratch, ax},
                                      StorePlain<R>{p.scratch, ux}, s, 1000, var);
            } else
                launch_strided<R, -1>(p.ny, scratch_cols_y(p), p.tw_y, LoadMasked<R>{p.scratch, ay, (unsigned)p.nzc},
                                      StoreMasked<R>{p.scratch, ay, (unsigned)p.nzc}, s, 1000, var);
            KF_HIP(hipEventRecord(e1, s));
            KF_HIP(hipEventSynchronize(e1));
            if (r >= 0) KF_HIP(hipEventElapsedTime(&ms[r], e0, e1));
        }
        std::sort(ms, ms + 7);
        return ms[3];
    };
    // y pass: column-block variant and scratch row padding together
    const int nzs_pad = p.nzs;
    float best = 1e30f, t_default = 0.f;
    int best_nzs = nzs_pad, best_vy = 0;
    for (int nzs : {nzs_pad, p.nzc}) {
        p.nzs = nzs;
        ay = saddr_y(p);
        for (int v = 0; v < 2; ++v) {
            if (v == 1 && cb_variant<R>(p.ny, 1) == cb_variant<R>(p.ny, 0)) continue;
            const float t = time_pass(1, v);
            if (nzs == nzs_pad && v == 0) t_default = t;
            if (t < best) {
                best = t;
                best_nzs = nzs;
                best_vy = v;
            }
        }
        if (nzs_pad == p.nzc) break;
    }
    if (!(best < 0.95f * t_default)) {
        best_nzs = nzs_pad;
        best_vy = 0;
    }
    p.nzs = best_nzs;
    p.vy = best_vy;
    ax = saddr_x(p);
    if (cb_variant<R>(p.nx, 1) != cb_variant<R>(p.nx, 0)) p.vx = time_pass(0, 1) < 0.95f * time_pass(0, 0) ? 1 : 0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

template <typename R>
void plan_init(Plan<R> &p, int nx, int ny, int nz, hipStream_t s)
{
    if (!grid_supported(nx, ny, nz))
        throw Error("grid " + std::to_string(nx) + "x" + std::to_string(ny) + "x" +
                    std::to_string(nz) + " not supported by the FFT (lengths 2^a 3^b <= 256, Nz even)");
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    p.nzc = nz / 2 + 1;
    p.nr = (long)nx * ny * nz;
    p.nc = (long)nx * ny * p.nzc;
    p.tw_x = twiddles<R>(nx, nx, s);
    p.tw_y = twiddles<R>(ny, ny, s);
    p.tw_m = twiddles<R>(nz / 2, nz / 2, s);
    p.tw_n = twiddles<R>(nz, nz / 2 + 1, s);
    {
        const int al = 128 / (int)sizeof(C2<R>);
        p.nzs = (p.nzc + al - 1) / al * al;
    }
    KF_HIP(hipMalloc(&p.scratch, sizeof(C2<R>) * (size_t)nx * ny * p.nzs));
    autotune(p, s);
}

template <typename R>
void plan_free(Plan<R> &p)
{
    (void)hipFree(p.tw_x);
    (void)hipFree(p.tw_y);
    (void)hipFree(p.tw_m);
    (void)hipFree(p.tw_n);
    (void)hipFree(p.scratch);
    p.tw_x = p.tw_y = p.tw_m = p.tw_n = p.scratch = nullptr;
}

template <typename R>
void r2c(const Plan<R> &p, const R *in, C2<R> *out, hipStream_t s)
{
    const long rows = (long)p.nx * p.ny;
    launch_zr2c<R>(p.nz / 2, rows, p.tw_m, p.tw_n, RowLoadReal<R>{in, p.nz},
                   RowStoreC<R>{out, p.nzc}, s);
    StridedAddr ay = addr_y(p.nx, p.ny, p.nzc), ax = addr_x(p.nx, p.ny, p.nzc);
    launch_strided<R, -1>(p.ny, (long)p.nx * p.nzc, p.tw_y, LoadPlain<R>{out, ay},
                          StorePlain<R>{out, ay}, s, prof::FFT_Y, p.vy);
    launch_strided<R, -1>(p.nx, (long)p.ny * p.nzc, p.tw_x, LoadPlain<R>{out, ax},
                          StorePlain<R>{out, ax}, s, prof::FFT_X, p.vx);
}

template <typename R>
void c2r(const Plan<R> &p, const C2<R> *in, R *out, hipStream_t s)
{
    const StridedAddr ax = addr_x(p.nx, p.ny, p.nzc), sy = saddr_y(p);
    launch_strided<R, +1>(p.nx, scratch_cols_x(p), p.tw_x, pad_cols(LoadPlain<R>{in, ax}, p),
                          scratch_store_x(p.scratch, p), s, prof::FFT_X, p.vx);
    launch_strided<R, +1>(p.ny, scratch_cols_y(p), p.tw_y, LoadMasked<R>{p.scratch, sy, (unsigned)p.nzc},
                          StoreMasked<R>{p.scratch, sy, (unsigned)p.nzc}, s, prof::FFT_Y, p.vy);
    const long rows = (long)p.nx * p.ny;
    launch_zc2r<R>(p.nz / 2, rows, p.tw_m, p.tw_n, RowLoadC<R>{p.scratch, p.nzs},
                   RowStoreReal<R>{out, p.nz, (R)(1.0 / (double)p.nr)}, s);
}

template void plan_init<float>(Plan<float> &, int, int, int, hipStream_t);
template void plan_init<double>(Plan<double> &, int, int, int, hipStream_t);
template void plan_free<float>(Plan<float> &);
template void plan_free<double>(Plan<double> &);
template void r2c<float>(const Plan<float> &, const float *, float2 *, hipStream_t);
template void r2c<double>(const Plan<double> &, const double *, double2 *, hipStream_t);
template void c2r<float>(const Plan<float> &, const float2 *, float *, hipStream_t);
template void c2r<double>(const Plan<double> &, const double2 *, double *, hipStream_t);

}  // namespace fft
}  // namespace kf
