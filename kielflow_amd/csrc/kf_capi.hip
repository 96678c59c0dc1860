// kf_capi.hip -- the extern "C" boundary (include/kielflow_amd.h).
#include "../../include/kielflow_amd.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "kf_adjoint.h"
#include "kf_batch.h"
#include "kf_effects.h"
#include "kf_fft.h"
#include "kf_grid.h"
#include "kf_poisson.h"
#include "kf_prof.h"
#include "kf_rhs.h"

using namespace kf;

static thread_local std::string g_err;

struct kf_ctx_s {
    int device;
    hipStream_t stream;
};
struct kf_grid_s {
    kf_ctx ctx;
    Grid *g;
    fft::Plan<double> *p64;  // FP64 plan for the OP_* operators and CFL
    EffectWork *ew;          // scratch of the effect_* entries (on first use)
    EffectWork &work()
    {
        if (!ew) ew = new EffectWork(*g);
        return *ew;
    }
};
struct kf_fft_s {
    kf_grid grid;
    int precision;
    fft::Plan<double> p64;
    fft::Plan<float> p32;
};
struct kf_poisson_s {
    kf_grid grid;
    Poisson *p;
    PoissonStats last;
    double *tmp;
};
struct kf_rhs_s {
    kf_grid grid;
    int precision;
    kf_poisson solver;
    Rhs<double> *r64;
    Rhs<float> *r32;
    float2 *conv[4];
};
struct kf_adjoint_s {
    kf_rhs rhs;
    Adjoint *a;
};
struct kf_rk4_s {
    kf_rhs rhs;
    Rk4<double> *k64;
    Rk4<float> *k32;
};
struct kf_euler_s {
    kf_rhs rhs;
    Euler<double> *e64;
    Euler<float> *e32;
};

template <class F>
static int guard(F &&f)
{
    try {
        return f();
    } catch (const kf::Error &e) {
        g_err = e.msg;
        return (e.msg.find("hip") != std::string::npos) ? KF_ERR_HIP : KF_ERR_ARG;
    } catch (const std::exception &e) {
        g_err = e.what();
        return KF_ERR_ARG;
    } catch (...) {
        g_err = "unknown error";
        return KF_ERR_ARG;
    }
}

#define KF_REQUIRE(cond, msg)            \
    do {                                 \
        if (!(cond)) {                   \
            g_err = (msg);               \
            return KF_ERR_ARG;           \
        }                                \
    } while (0)

static void set_dev(kf_ctx c) { KF_HIP(hipSetDevice(c->device)); }

// ------------------------------------------------------------------ static scratch
// Device scratch of the allocation-free entry points (kf_cfl_all_noalloc):
// one array of KF_NOALLOC_SCRATCH_BYTES per GPU, held until the process
// ends.  In host code the array's name is not a device address: the
// address comes from hipGetSymbolAddress, once per device with that device
// current (kf_ctx_create), and is kept per device.  Calls that overlap on
// one GPU share the array and race.
static __device__ __align__(256) double scratch[KF_NOALLOC_SCRATCH_BYTES / sizeof(double)];
constexpr int KF_MAX_DEVICES = 64;
static std::mutex g_scratch_mtx;
static double *g_scratch_ptr[KF_MAX_DEVICES] = {};

// the current device's address of `scratch` (the device must be current)
static double *noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mtx);
    if (device < 0 || device >= KF_MAX_DEVICES) throw Error("device index beyond the scratch table");
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) throw Error(std::string("hip: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    return g_scratch_ptr[device];
}

extern "C" {

const char *kf_last_error(void) { return g_err.c_str(); }

int kf_ctx_create(int device, kf_ctx *out)
{
    return guard([&] {
        KF_REQUIRE(out, "null output");
        int n = 0;
        KF_HIP(hipGetDeviceCount(&n));
        KF_REQUIRE(device >= 0 && device < n, "no such device");
        KF_HIP(hipSetDevice(device));
        std::unique_ptr<kf_ctx_s> c(new kf_ctx_s{device, nullptr});
        KF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        // the device address of the static scratch, while the device is current
        (void)noalloc_scratch(device);
        *out = c.release();
        return KF_OK;
    });
}

int kf_ctx_destroy(kf_ctx ctx)
{
    return guard([&] {
        if (!ctx) return KF_OK;
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return KF_OK;
    });
}

int kf_ctx_synchronize(kf_ctx ctx)
{
    return guard([&] {
        KF_REQUIRE(ctx, "null context");
        KF_HIP(hipStreamSynchronize(ctx->stream));
        return KF_OK;
    });
}

void *kf_ctx_stream(kf_ctx ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int kf_malloc(kf_ctx ctx, size_t bytes, void **out)
{
    return guard([&] {
        KF_REQUIRE(ctx && out, "null argument");
        set_dev(ctx);
        KF_HIP(hipMalloc(out, bytes ? bytes : 1));
        return KF_OK;
    });
}

int kf_free(kf_ctx ctx, void *p)
{
    return guard([&] {
        KF_REQUIRE(ctx, "null context");
        KF_HIP(hipStreamSynchronize(ctx->stream));
        KF_HIP(hipFree(p));
        return KF_OK;
    });
}

int kf_memcpy_h2d(kf_ctx ctx, void *dst, const void *src, size_t bytes)
{
    return guard([&] {
        KF_REQUIRE(ctx, "null context");
        KF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        KF_HIP(hipStreamSynchronize(ctx->stream));
        return KF_OK;
    });
}

int kf_memcpy_d2h(kf_ctx ctx, void *dst, const void *src, size_t bytes)
{
    return guard([&] {
        KF_REQUIRE(ctx, "null context");
        KF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        KF_HIP(hipStreamSynchronize(ctx->stream));
        return KF_OK;
    });
}

// plain streaming copy, 16 bytes per lane and iteration (also the kernel the
// HBM byte counters are calibrated against: it reads and writes `bytes` once)
__global__ void __launch_bounds__(256) k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int kf_memcpy_d2d(kf_ctx ctx, void *dst, const void *src, size_t bytes)
{
    return guard([&] {
        KF_REQUIRE(ctx, "null context");
        if (bytes == 0) return KF_OK;
        const bool al = (((uintptr_t)dst | (uintptr_t)src | bytes) & 15) == 0;
        if (!al) {
            KF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
            return KF_OK;
        }
        const size_t n = bytes / 16;
        size_t nb = (n + 255) / 256;
        if (nb > 8192) nb = 8192;
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)nb), dim3(256), 0, ctx->stream, (const uint4 *)src, (uint4 *)dst,
                           n);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_memset(kf_ctx ctx, void *dst, int value, size_t bytes)
{
    return guard([&] {
        KF_REQUIRE(ctx, "null context");
        KF_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
        return KF_OK;
    });
}

// ------------------------------------------------------------------ grid
static AxisDesc to_desc(const kf_axis &a) { return AxisDesc{a.type, a.n, a.l0, a.L, a.m}; }

int kf_grid_create(kf_ctx ctx, const kf_axis axes[3], kf_grid *out)
{
    return guard([&] {
        KF_REQUIRE(ctx && axes && out, "null argument");
        set_dev(ctx);
        AxisDesc ad[3] = {to_desc(axes[0]), to_desc(axes[1]), to_desc(axes[2])};
        KF_REQUIRE(fft::grid_supported(ad[0].n, ad[1].n, ad[2].n),
                   "grid extents not supported (2^a 3^b <= 256; Nz even)");
        for (int d = 0; d < 3; ++d) {
            KF_REQUIRE(ad[d].type >= 1 && ad[d].type <= 3, "unknown axis type");
            KF_REQUIRE(ad[d].L > 0, "axis length must be positive");
        }
        std::unique_ptr<Grid> g(new Grid(ad, ctx->stream));
        std::unique_ptr<fft::Plan<double>> p(new fft::Plan<double>());
        fft::plan_init<double>(*p, ad[0].n, ad[1].n, ad[2].n, ctx->stream);
        *out = new kf_grid_s{ctx, g.release(), p.release(), nullptr};
        return KF_OK;
    });
}

int kf_grid_destroy(kf_grid g)
{
    return guard([&] {
        if (!g) return KF_OK;
        (void)hipStreamSynchronize(g->ctx->stream);
        fft::plan_free<double>(*g->p64);
        delete g->p64;
        delete g->ew;
        delete g->g;
        delete g;
        return KF_OK;
    });
}

int kf_grid_dims(kf_grid g, int *nx, int *ny, int *nz, size_t *nreal, size_t *ncplx)
{
    return guard([&] {
        KF_REQUIRE(g, "null grid");
        if (nx) *nx = g->g->nx;
        if (ny) *ny = g->g->ny;
        if (nz) *nz = g->g->nz;
        if (nreal) *nreal = (size_t)g->g->nr;
        if (ncplx) *ncplx = (size_t)g->g->nc;
        return KF_OK;
    });
}

int kf_axis_values(const kf_axis *a, int index, double *val, double *k, double *dS)
{
    return guard([&] {
        KF_REQUIRE(a && a->n > 0, "bad axis");
        AxisDesc d = to_desc(*a);
        if (val) *val = axis_val_at(d, index);
        if (k) *k = axis_k_val_at(d, index);
        if (dS) *dS = axis_dS(d, index);
        return KF_OK;
    });
}

int kf_axis_index_at(const kf_axis *a, double val, int *index)
{
    return guard([&] {
        KF_REQUIRE(a && index, "null argument");
        KF_REQUIRE((a->type == KF_AXIS_EQST || a->type == KF_AXIS_HYST || a->type == KF_AXIS_SIST) && a->n > 0 &&
                       a->L > 0,
                   "bad axis");
        KF_REQUIRE(a->type != KF_AXIS_HYST || a->m != 0., "a HySt axis needs m != 0");
        const int N = a->n;
        const double l0 = a->l0, L = a->L, m = a->m;
        KF_REQUIRE(std::isfinite(val), "coordinate is not finite");
        if (a->type == KF_AXIS_EQST) {
            // axis_EqSt.cpp:63-66: the conversion to int truncates
            *index = (int)(N * (val - l0) / L);
        } else if (a->type == KF_AXIS_HYST) {
            // axis_HySt.cpp:121-147: periodic shift, then the inverse map
            KF_REQUIRE(std::fabs(val - l0) < 1.e6 * L, "coordinate too far outside the box");
            double x = val;
            int sh = 0;
            while (x < l0) {
                sh -= 1;
                x += L;
            }
            while (x >= l0 + L) {
                sh += 1;
                x -= L;
            }
            x = (L / (2. * m)) * std::asinh(2. * x * std::sinh(m) / L);
            *index = (int)std::nearbyint(N * (x - l0) / L) + N * sh;
        } else {
            // axis_SiSt.cpp:98-113: wrap into [l0, l0 + L), then bisect
            // S(x) = x - m sin(2 pi x / L) for S(x) = val to 1e-6
            // (axis_SiSt.cpp:6-13), round, and N maps to 0
            KF_REQUIRE(std::fabs(val - l0) < 1.e6 * L, "coordinate too far outside the box");
            double y = val;
            while (y < l0) y += L;
            while (y >= l0 + L) y -= L;
            const double pi = std::acos(-1.);
            auto S = [&](double x) { return x - m * std::sin(2. * pi * x / L); };
            double lo = l0, hi = l0 + L, x = 0.5 * (lo + hi);
            for (int it = 0; it < 200 && !(std::fabs(y - S(x)) < 1.e-6); ++it) {
                if (S(x) < y)
                    lo = x;
                else
                    hi = x;
                x = 0.5 * (lo + hi);
            }
            int r = (int)std::nearbyint(N * (x - l0) / L);
            if (r == N) r = 0;
            *index = r;
        }
        return KF_OK;
    });
}

// ------------------------------------------------------------------ FFT
int kf_fft_create(kf_grid g, int precision, kf_fft *out)
{
    return guard([&] {
        KF_REQUIRE(g && out, "null argument");
        KF_REQUIRE(precision == KF_FP64 || precision == KF_FP32, "precision must be KF_FP64 or KF_FP32");
        set_dev(g->ctx);
        std::unique_ptr<kf_fft_s> f(new kf_fft_s());
        f->grid = g;
        f->precision = precision;
        if (precision == KF_FP64) fft::plan_init<double>(f->p64, g->g->nx, g->g->ny, g->g->nz, g->ctx->stream);
        else fft::plan_init<float>(f->p32, g->g->nx, g->g->ny, g->g->nz, g->ctx->stream);
        *out = f.release();
        return KF_OK;
    });
}

int kf_fft_destroy(kf_fft f)
{
    return guard([&] {
        if (!f) return KF_OK;
        (void)hipStreamSynchronize(f->grid->ctx->stream);
        if (f->precision == KF_FP64) fft::plan_free<double>(f->p64);
        else fft::plan_free<float>(f->p32);
        delete f;
        return KF_OK;
    });
}

int kf_fft_r2c(kf_fft f, const void *in, void *out)
{
    return guard([&] {
        KF_REQUIRE(f && in && out, "null argument");
        hipStream_t s = f->grid->ctx->stream;
        if (f->precision == KF_FP64) fft::r2c<double>(f->p64, (const double *)in, (double2 *)out, s);
        else fft::r2c<float>(f->p32, (const float *)in, (float2 *)out, s);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_fft_c2r(kf_fft f, const void *in, void *out)
{
    return guard([&] {
        KF_REQUIRE(f && in && out, "null argument");
        hipStream_t s = f->grid->ctx->stream;
        if (f->precision == KF_FP64) fft::c2r<double>(f->p64, (const double2 *)in, (double *)out, s);
        else fft::c2r<float>(f->p32, (const float2 *)in, (float *)out, s);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

// ------------------------------------------------------------------ spectral ops
int kf_partial_derivative(kf_grid g, int dir, const double *in, double *out)
{
    return guard([&] {
        KF_REQUIRE(g && in && out, "null argument");
        KF_REQUIRE(dir >= 0 && dir < 3, "direction must be 0, 1 or 2");
        op_partial_derivative(*g->g, *g->p64, dir, (const double2 *)in, (double2 *)out, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

// ---------------------------------------------------------------- effect_* interfaces
int kf_effect_advection(kf_grid g, const double *FUx, const double *FUy, const double *FUz, const double *FXX,
                        double *out)
{
    return guard([&] {
        KF_REQUIRE(g && FUx && FUy && FUz && FXX && out, "null argument");
        const double2 *FU[3] = {(const double2 *)FUx, (const double2 *)FUy, (const double2 *)FUz};
        effect_advection(*g->g, *g->p64, g->work(), FU, (const double2 *)FXX, (double2 *)out, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_effect_continuity(kf_grid g, const double *FUx, const double *FUy, const double *FUz, const double *Fni,
                         double *out)
{
    return guard([&] {
        KF_REQUIRE(g && FUx && FUy && FUz && Fni && out, "null argument");
        const double2 *FU[3] = {(const double2 *)FUx, (const double2 *)FUy, (const double2 *)FUz};
        effect_continuity(*g->g, *g->p64, g->work(), FU, (const double2 *)Fni, (double2 *)out, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_effect_force_E(kf_grid g, const double *FPhi, double *outx, double *outy, double *outz)
{
    return guard([&] {
        KF_REQUIRE(g && FPhi && outx && outy && outz, "null argument");
        double2 *o[3] = {(double2 *)outx, (double2 *)outy, (double2 *)outz};
        effect_force_E(*g->g, *g->p64, g->work(), (const double2 *)FPhi, o, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_effect_diffusion_u(kf_grid g, double theta, const double *Fni, double *outx, double *outy, double *outz)
{
    return guard([&] {
        KF_REQUIRE(g && Fni && outx && outy && outz, "null argument");
        KF_REQUIRE(theta != 0., "theta must be nonzero");
        double2 *o[3] = {(double2 *)outx, (double2 *)outy, (double2 *)outz};
        effect_diffusion_u(*g->g, *g->p64, g->work(), theta, (const double2 *)Fni, o, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_effect_spectral_viscosity(kf_grid g, const double *in, double *out)
{
    return guard([&] {
        KF_REQUIRE(g && in && out, "null argument");
        effect_spectral_viscosity(*g->g, (const double2 *)in, (double2 *)out, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_effect_translation(kf_grid g, double v, const double *in, double *out)
{
    return guard([&] {
        KF_REQUIRE(g && in && out, "null argument");
        effect_translation(*g->g, *g->p64, g->work(), v, (const double2 *)in, (double2 *)out, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_effect_force_linear(kf_grid g, double k, const double *in, double *out)
{
    return guard([&] {
        KF_REQUIRE(g && in && out, "null argument");
        effect_force_linear(*g->g, k, (const double2 *)in, (double2 *)out, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_effect_penalization(kf_grid g, const double *mask, double eta, double fix, int fix_scaled_by_mask,
                           const double *in, double *out)
{
    return guard([&] {
        KF_REQUIRE(g && mask && in && out, "null argument");
        KF_REQUIRE(eta <= 0., "penalization damping must not be positive (effect_penalization.cpp:9-15)");
        effect_penalization(*g->g, *g->p64, g->work(), mask, eta, fix, fix_scaled_by_mask ? 1 : 0,
                            (const double2 *)in, (double2 *)out, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_dealias23(kf_grid g, double *inout)
{
    return guard([&] {
        KF_REQUIRE(g && inout, "null argument");
        op_dealias23(*g->g, (double2 *)inout, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_filter_phi(kf_grid g, double *inout)
{
    return guard([&] {
        KF_REQUIRE(g && inout, "null argument");
        op_filter_phi(*g->g, (double2 *)inout, g->ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

// ------------------------------------------------------------------ Poisson
int kf_poisson_create(kf_grid g, int precision, int mode, kf_poisson *out)
{
    return guard([&] {
        KF_REQUIRE(g && out, "null argument");
        KF_REQUIRE(precision == KF_FP64 || precision == KF_MIXED, "precision must be KF_FP64 or KF_MIXED");
        KF_REQUIRE(mode >= 0 && mode <= 2, "unknown Poisson mode");
        KF_REQUIRE(!(precision == KF_MIXED && mode != KF_POISSON_NEWTON_MG),
                   "mixed precision needs the Newton-MG solver");
        set_dev(g->ctx);
        std::unique_ptr<Poisson> p(new Poisson(g->g->ax, precision == KF_MIXED, g->ctx->stream));
        p->set_mode(mode);
        *out = new kf_poisson_s{g, p.release(), PoissonStats{0, 0., 0}, nullptr};
        return KF_OK;
    });
}

int kf_poisson_destroy(kf_poisson p)
{
    return guard([&] {
        if (!p) return KF_OK;
        (void)hipStreamSynchronize(p->grid->ctx->stream);
        delete p->p;
        (void)hipFree(p->tmp);
        delete p;
        return KF_OK;
    });
}

int kf_poisson_set_tolerance(kf_poisson p, double tol)
{
    return guard([&] {
        KF_REQUIRE(p, "null solver");
        KF_REQUIRE(tol > 0, "tolerance must be positive");
        p->p->set_tolerance(tol);
        return KF_OK;
    });
}

int kf_poisson_set_max_iterations(kf_poisson p, int iters)
{
    return guard([&] {
        KF_REQUIRE(p, "null solver");
        p->p->set_max_iterations(iters);
        return KF_OK;
    });
}

int kf_poisson_set_smoothing(kf_poisson p, int pre, int post, int coarse)
{
    return guard([&] {
        KF_REQUIRE(p, "null solver");
        KF_REQUIRE(pre >= 0 && post >= 0 && coarse >= 1 && pre + post >= 1, "bad smoothing counts");
        p->p->set_smoothing(pre, post, coarse);
        return KF_OK;
    });
}

int kf_poisson_set_cascade(kf_poisson p, int n, const int *level_change, const int *steps, const double *tol)
{
    return guard([&] {
        KF_REQUIRE(p && level_change && steps && tol, "null argument");
        KF_REQUIRE(p->p->mode() == POISSON_CASCADE, "the level program belongs to the KF_POISSON_CASCADE mode");
        p->p->set_cascade(n, level_change, steps, tol);
        return KF_OK;
    });
}

int kf_poisson_set_ion_boltzmann(kf_poisson p, double theta)
{
    return guard([&] {
        KF_REQUIRE(p, "null solver");
        p->p->set_ion_boltzmann(theta);
        return KF_OK;
    });
}

int kf_frontend_equilibrium(kf_grid g, double theta, const double *nd, int mode, double tol, int max_iterations,
                            double *phi, double *ni, int *iters)
{
    return guard([&] {
        KF_REQUIRE(g && nd && phi && ni, "null argument");
        KF_REQUIRE(theta > 0., "theta must be positive");
        KF_REQUIRE(mode >= 0 && mode <= 2, "unknown Poisson mode");
        set_dev(g->ctx);
        hipStream_t s = g->ctx->stream;
        Poisson ps(g->g->ax, false, s);
        ps.set_mode(mode);
        ps.set_ion_boltzmann(theta);
        if (mode != POISSON_CASCADE) {
            ps.set_tolerance(tol);
            ps.set_max_iterations(max_iterations);
        }
        // frontend.cpp:164-190: the potential of the grain with Boltzmann
        // electrons and ions (rho = n_d), then n_i = exp(-theta phi)
        const int rc = ps.solve(phi, nd, nullptr, s, true);
        const PoissonStats st = ps.last_stats();
        if (iters) *iters = st.cycles;
        if (rc == -2) {
            g_err = "NaN in the potential";
            return KF_ERR_NAN;
        }
        const long nr = g->g->nr, nc = g->g->nc;
        boltzmann_ions(phi, theta, ni, nr, s);
        // frontend.cpp:296-303: both through dealiasing_undesignated with
        // exp(-2000 k^15)
        double2 *F = nullptr;
        KF_HIP(hipMalloc(&F, sizeof(double2) * nc));
        for (double *f : {ni, phi}) {
            fft::r2c<double>(*g->p64, f, F, s);
            op_filter_phi(*g->g, F, s);
            fft::c2r<double>(*g->p64, F, f, s);
        }
        KF_HIP(hipStreamSynchronize(s));
        (void)hipFree(F);
        if (mode == POISSON_NEWTON_MG && !st.converged) {
            g_err = "equilibrium solve did not converge within the iteration budget";
            return KF_ERR_NOT_CONVERGED;
        }
        return KF_OK;
    });
}

int kf_poisson_solve(kf_poisson p, double *phi, const double *rho, int *iters, double *resid)
{
    return guard([&] {
        KF_REQUIRE(p && phi && rho, "null argument");
        hipStream_t s = p->grid->ctx->stream;
        int rc = p->p->solve(phi, rho, nullptr, s, true);
        // Newton-MG: resid = max|F|/max|rho|; reference modes: the last
        // sweep's max update norm_max (solver_poisson_jacobi_nlin.cpp:204-210)
        PoissonStats st = p->p->last_stats();
        p->last = st;
        if (iters) *iters = st.cycles;
        if (resid) *resid = st.resid;
        if (rc == -2) {
            g_err = "NaN in the potential";
            return KF_ERR_NAN;
        }
        if (p->p->mode() == POISSON_NEWTON_MG && !st.converged) {
            g_err = "Poisson solve did not converge within the iteration budget";
            return KF_ERR_NOT_CONVERGED;
        }
        return KF_OK;
    });
}

int kf_poisson_converged(kf_poisson p) { return p ? p->last.converged : 0; }

int kf_poisson_residual(kf_poisson p, const double *phi, const double *rho, double *resid)
{
    return guard([&] {
        KF_REQUIRE(p && phi && rho && resid, "null argument");
        hipStream_t s = p->grid->ctx->stream;
        // a zero-iteration Newton solve on a copy evaluates the residual
        const long N = p->grid->g->nr;
        if (!p->tmp) KF_HIP(hipMalloc(&p->tmp, sizeof(double) * N));
        KF_HIP(hipMemcpyAsync(p->tmp, phi, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
        int mode = p->p->mode(), mi = p->p->max_iterations();
        double tol = p->p->tolerance();
        p->p->set_mode(POISSON_NEWTON_MG);
        p->p->set_max_iterations(0);
        p->p->set_tolerance(0.0);
        p->p->solve(p->tmp, rho, nullptr, s, true);
        PoissonStats st = p->p->last_stats();
        p->p->set_mode(mode);
        p->p->set_max_iterations(mi);
        p->p->set_tolerance(tol);
        *resid = st.resid;
        return KF_OK;
    });
}

// ------------------------------------------------------------------ RHS
int kf_rhs_create(kf_grid g, const kf_parameters *params, const kf_rhs_switches *sw, kf_poisson solver,
                  int precision, kf_rhs *out)
{
    return guard([&] {
        KF_REQUIRE(g && params && sw && solver && out, "null argument");
        KF_REQUIRE(precision == KF_FP64 || precision == KF_MIXED, "precision must be KF_FP64 or KF_MIXED");
        KF_REQUIRE(solver->grid->g == g->g, "solver belongs to another grid");
        KF_REQUIRE(params->theta != 0, "theta must be nonzero");
        set_dev(g->ctx);
        Params pa{params->M, params->tau, params->theta, params->mu, params->beta};
        Switches s{sw->sv_on, sw->pen_on};
        std::unique_ptr<kf_rhs_s> r(new kf_rhs_s());
        r->grid = g;
        r->precision = precision;
        r->solver = solver;
        if (precision == KF_FP64) r->r64 = new Rhs<double>(*g->g, *solver->p, pa, s, g->ctx->stream);
        else {
            r->r32 = new Rhs<float>(*g->g, *solver->p, pa, s, g->ctx->stream);
            for (int f = 0; f < 4; ++f) KF_HIP(hipMalloc(&r->conv[f], sizeof(float2) * g->g->nc));
        }
        *out = r.release();
        return KF_OK;
    });
}

int kf_rhs_destroy(kf_rhs r)
{
    return guard([&] {
        if (!r) return KF_OK;
        (void)hipStreamSynchronize(r->grid->ctx->stream);
        delete r->r64;
        delete r->r32;
        for (int f = 0; f < 4; ++f) (void)hipFree(r->conv[f]);
        delete r;
        return KF_OK;
    });
}

int kf_rhs_set_fields(kf_rhs r, double *phi, const double *nd, const double *mask)
{
    return guard([&] {
        KF_REQUIRE(r, "null rhs");
        if (r->r64) r->r64->set_fields(phi, nd, mask);
        else r->r32->set_fields(phi, nd, mask);
        return KF_OK;
    });
}

int kf_rhs_set_parameters(kf_rhs r, const kf_parameters *params)
{
    return guard([&] {
        KF_REQUIRE(r && params, "null argument");
        Params pa{params->M, params->tau, params->theta, params->mu, params->beta};
        if (r->r64) r->r64->set_params(pa);
        else r->r32->set_params(pa);
        return KF_OK;
    });
}

int kf_rhs_solve(kf_rhs r, double t, double *FUx, double *FUy, double *FUz, double *Fn)
{
    return guard([&] {
        KF_REQUIRE(r && FUx && FUy && FUz && Fn, "null argument");
        hipStream_t s = r->grid->ctx->stream;
        double2 *F[4] = {(double2 *)FUx, (double2 *)FUy, (double2 *)FUz, (double2 *)Fn};
        if (r->r64) {
            r->r64->eval(t, F);
        } else {
            const long nc = r->grid->g->nc;
            for (int f = 0; f < 4; ++f) to_single(F[f], r->conv[f], nc, s);
            r->r32->eval(t, r->conv);
            for (int f = 0; f < 4; ++f) to_double(r->conv[f], F[f], nc, s);
        }
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_rhs_poisson_failures(kf_rhs r, int *count)
{
    return guard([&] {
        KF_REQUIRE(r && count, "null argument");
        *count = r->r64 ? r->r64->poisson_failures() : r->r32->poisson_failures();
        return KF_OK;
    });
}

int kf_poisson_newton_counts(kf_poisson p, long *solves, long *corrections)
{
    return guard([&] {
        KF_REQUIRE(p && solves && corrections, "null argument");
        *solves = p->p->solves_counted();
        *corrections = p->p->corrections_counted();
        return KF_OK;
    });
}

int kf_rhs_force_general(kf_rhs r, int on)
{
    return guard([&] {
        KF_REQUIRE(r, "null rhs");
        if (r->r64) r->r64->set_force_general(on != 0);
        else r->r32->set_force_general(on != 0);
        return KF_OK;
    });
}

// ------------------------------------------------------------------ RK4
int kf_rk4_create(kf_rhs r, double dt, kf_rk4 *out)
{
    return guard([&] {
        KF_REQUIRE(r && out, "null argument");
        set_dev(r->grid->ctx);
        std::unique_ptr<kf_rk4_s> k(new kf_rk4_s());
        k->rhs = r;
        if (r->r64) k->k64 = new Rk4<double>(*r->r64, dt);
        else k->k32 = new Rk4<float>(*r->r32, dt);
        *out = k.release();
        return KF_OK;
    });
}

int kf_rk4_destroy(kf_rk4 k)
{
    return guard([&] {
        if (!k) return KF_OK;
        (void)hipStreamSynchronize(k->rhs->grid->ctx->stream);
        delete k->k64;
        delete k->k32;
        delete k;
        return KF_OK;
    });
}

int kf_rk4_steps(kf_rk4 k, int n, double *FUx, double *FUy, double *FUz, double *Fn)
{
    return guard([&] {
        KF_REQUIRE(k && FUx && FUy && FUz && Fn, "null argument");
        double2 *y[4] = {(double2 *)FUx, (double2 *)FUy, (double2 *)FUz, (double2 *)Fn};
        if (k->k64) k->k64->steps(y, n);
        else k->k32->steps(y, n);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_rk4_solve(kf_rk4 k, double *FUx, double *FUy, double *FUz, double *Fn)
{
    return kf_rk4_steps(k, 1, FUx, FUy, FUz, Fn);
}

int kf_rk4_set_dt(kf_rk4 k, double dt)
{
    return guard([&] {
        KF_REQUIRE(k, "null integrator");
        if (k->k64) k->k64->set_dt(dt);
        else k->k32->set_dt(dt);
        return KF_OK;
    });
}

int kf_rk4_time(kf_rk4 k, double *t)
{
    return guard([&] {
        KF_REQUIRE(k && t, "null argument");
        *t = k->k64 ? k->k64->t() : k->k32->t();
        return KF_OK;
    });
}

int kf_cfl_all(kf_grid g, double dt, const double *FUx, const double *FUy, const double *FUz, double *dt_new)
{
    return guard([&] {
        KF_REQUIRE(g && FUx && FUy && FUz && dt_new, "null argument");
        const double2 *F[3] = {(const double2 *)FUx, (const double2 *)FUy, (const double2 *)FUz};
        *dt_new = cfl_all(*g->g, *g->p64, dt, F, g->ctx->stream);
        return KF_OK;
    });
}

int kf_cfl_all_noalloc(kf_grid g, double dt, const double *FUx, const double *FUy, const double *FUz,
                       double *dt_new)
{
    return guard([&] {
        KF_REQUIRE(g && FUx && FUy && FUz && dt_new, "null argument");
        set_dev(g->ctx);
        const double2 *F[3] = {(const double2 *)FUx, (const double2 *)FUy, (const double2 *)FUz};
        const size_t need = cfl_all_scratch_doubles(*g->g);
        if (need > KF_NOALLOC_SCRATCH_BYTES / sizeof(double)) {
            g_err = "kf_cfl_all_noalloc: the grid needs " + std::to_string(need * sizeof(double)) +
                    " bytes of scratch, more than the " + std::to_string((size_t)KF_NOALLOC_SCRATCH_BYTES) +
                    " of the static scratch (use kf_cfl_all)";
            return KF_ERR_ARG;
        }
        *dt_new = cfl_all_into(*g->g, *g->p64, dt, F, noalloc_scratch(g->ctx->device), g->ctx->stream);
        return KF_OK;
    });
}

int kf_field_integrate(kf_grid g, const double *f, double *result)
{
    return guard([&] {
        KF_REQUIRE(g && f && result, "null argument");
        set_dev(g->ctx);
        *result = field_integrate(*g->g, f, nullptr, g->ctx->stream);
        return KF_OK;
    });
}

int kf_force_decomposition(kf_rhs r, const double *FUx, const double *FUy, const double *FUz, const double *Fn,
                           const double *phi, double *terms, double *integrals, double *drag)
{
    return guard([&] {
        KF_REQUIRE(r && FUx && FUy && FUz && Fn && phi, "null argument");
        KF_REQUIRE(r->r64, "the force decomposition takes an FP64 RHS");
        kf_grid gh = r->grid;
        set_dev(gh->ctx);
        Grid &g = *gh->g;
        fft::Plan<double> &p = *gh->p64;
        EffectWork &w = gh->work();
        hipStream_t s = gh->ctx->stream;
        const Params pa = r->r64->params();
        const Switches sw = r->r64->switches();
        const long nr = g.nr, nc = g.nc;
        const double2 *FU[3] = {(const double2 *)FUx, (const double2 *)FUy, (const double2 *)FUz};
        double2 *T[3] = {}, *Fphi = nullptr;
        double *out = nullptr;
        struct Free {
            std::vector<void *> v;
            ~Free()
            {
                for (void *q : v) (void)hipFree(q);
            }
        } fr;
        for (auto &t : T) {
            KF_HIP(hipMalloc(&t, sizeof(double2) * nc));
            fr.v.push_back(t);
        }
        KF_HIP(hipMalloc(&Fphi, sizeof(double2) * nc));
        fr.v.push_back(Fphi);
        if (!terms) {
            KF_HIP(hipMalloc(&out, sizeof(double) * nr));
            fr.v.push_back(out);
        }
        fft::r2c<double>(p, phi, Fphi, s);
        // the terms of d/dt u as rhs_standard.cpp accumulates them (each the
        // negative of its part of B), in the reference backend's names:
        // DISSIPATION, SPECTRAL_VISCOSITY, E_INT, DIFFUSION, E_EXT, ADVECTION
        for (int t = 0; t < KF_FORCE_TERMS; ++t) {
            for (auto q : T) KF_HIP(hipMemsetAsync(q, 0, sizeof(double2) * nc, s));
            switch (t) {
            case 0:
                for (int a = 0; a < 3; ++a) effect_force_linear(g, -pa.tau, FU[a], T[a], s);
                break;
            case 1:
                if (sw.sv_on)
                    for (int a = 0; a < 3; ++a) effect_spectral_viscosity(g, FU[a], T[a], s);
                break;
            case 2: effect_force_E(g, p, w, Fphi, T, s); break;
            case 3: effect_diffusion_u(g, p, w, pa.theta, (const double2 *)Fn, T, s); break;
            case 4:
                for (int a = 0; a < 3; ++a) effect_translation(g, p, w, pa.M, FU[a], T[a], s);
                break;
            default:
                for (int a = 0; a < 3; ++a) effect_advection(g, p, w, FU, FU[a], T[a], s);
            }
            for (int a = 0; a < 3; ++a) {
                double *o = terms ? terms + (size_t)(3 * t + a) * nr : out;
                fft::c2r<double>(p, T[a], o, s);
                if (integrals) integrals[3 * t + a] = field_integrate(g, o, nullptr, s);
            }
        }
        if (drag) {
            // the grain's x-drag, -integral of n_d d_x phi dV (the objective of
            // the adjoint): d_x phi in T[0], real in a scratch field
            op_partial_derivative(g, p, 0, Fphi, T[0], s);
            double *dx = out;
            if (!dx) {
                KF_HIP(hipMalloc(&dx, sizeof(double) * nr));
                fr.v.push_back(dx);
            }
            fft::c2r<double>(p, T[0], dx, s);
            *drag = -field_integrate(g, dx, r->r64->nd(), s);
        }
        KF_HIP(hipStreamSynchronize(s));
        return KF_OK;
    });
}

int kf_state_to_real(kf_grid g, const double *FUx, const double *FUy, const double *FUz, const double *Fn,
                     const double *phi, double *out)
{
    return guard([&] {
        KF_REQUIRE(g && FUx && FUy && FUz && Fn && phi && out, "null argument");
        set_dev(g->ctx);
        const double2 *F[4] = {(const double2 *)FUx, (const double2 *)FUy, (const double2 *)FUz,
                               (const double2 *)Fn};
        state_to_real(*g->g, *g->p64, F, phi, out, g->ctx->stream);
        KF_HIP(hipStreamSynchronize(g->ctx->stream));
        return KF_OK;
    });
}

int kf_state_plane_z(kf_grid g, const double *FUx, const double *FUy, const double *FUz, const double *Fn,
                     const double *phi, int k, double *out)
{
    return guard([&] {
        KF_REQUIRE(g && FUx && FUy && FUz && Fn && phi && out, "null argument");
        KF_REQUIRE(k >= 0 && k < g->g->nz, "plane index outside [0, Nz)");
        set_dev(g->ctx);
        const double2 *F[4] = {(const double2 *)FUx, (const double2 *)FUy, (const double2 *)FUz,
                               (const double2 *)Fn};
        state_plane_z(*g->g, *g->p64, F, phi, k, out, g->ctx->stream);
        KF_HIP(hipStreamSynchronize(g->ctx->stream));
        return KF_OK;
    });
}

int kf_state_from_real(kf_grid g, const double *in, double *FUx, double *FUy, double *FUz, double *Fn, double *phi)
{
    return guard([&] {
        KF_REQUIRE(g && in && FUx && FUy && FUz && Fn, "null argument");
        set_dev(g->ctx);
        double2 *F[4] = {(double2 *)FUx, (double2 *)FUy, (double2 *)FUz, (double2 *)Fn};
        KF_REQUIRE(state_from_real(*g->g, *g->p64, in, F, phi, g->ctx->stream),
                   "the ion density must be positive and finite (it is stored linear and evolved as ln n)");
        return KF_OK;
    });
}

int kf_euler_create(kf_rhs r, double dt, kf_euler *out)
{
    return guard([&] {
        KF_REQUIRE(r && out, "null argument");
        set_dev(r->grid->ctx);
        std::unique_ptr<kf_euler_s> e(new kf_euler_s());
        e->rhs = r;
        if (r->r64) e->e64 = new Euler<double>(*r->r64, dt);
        else e->e32 = new Euler<float>(*r->r32, dt);
        *out = e.release();
        return KF_OK;
    });
}

int kf_euler_destroy(kf_euler e)
{
    return guard([&] {
        if (!e) return KF_OK;
        (void)hipStreamSynchronize(e->rhs->grid->ctx->stream);
        delete e->e64;
        delete e->e32;
        delete e;
        return KF_OK;
    });
}

int kf_euler_steps(kf_euler e, int n, double *FUx, double *FUy, double *FUz, double *Fn)
{
    return guard([&] {
        KF_REQUIRE(e && FUx && FUy && FUz && Fn, "null argument");
        double2 *y[4] = {(double2 *)FUx, (double2 *)FUy, (double2 *)FUz, (double2 *)Fn};
        if (e->e64) e->e64->steps(y, n);
        else e->e32->steps(y, n);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_euler_set_dt(kf_euler e, double dt)
{
    return guard([&] {
        KF_REQUIRE(e, "null integrator");
        if (e->e64) e->e64->set_dt(dt);
        else e->e32->set_dt(dt);
        return KF_OK;
    });
}

int kf_rk4_stage_input(kf_ctx ctx, size_t n, double c, const double *y0, const double *k, double *out)
{
    return guard([&] {
        KF_REQUIRE(ctx && y0 && k && out, "null argument");
        set_dev(ctx);
        rk4_stage_input((long)n, c, y0, k, out, ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_rk4_combine(kf_ctx ctx, size_t n, double dt, const double *y0, const double *k1, const double *k2,
                   const double *k3, const double *k4, double *out)
{
    return guard([&] {
        KF_REQUIRE(ctx && y0 && k1 && k2 && k3 && k4 && out, "null argument");
        set_dev(ctx);
        const double *k[4] = {k1, k2, k3, k4};
        rk4_combine((long)n, dt, y0, k, out, ctx->stream);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

// ------------------------------------------------------------------ adjoint
int kf_adjoint_create(kf_rhs r, double dt, kf_adjoint *out)
{
    return guard([&] {
        KF_REQUIRE(r && out, "null argument");
        KF_REQUIRE(r->r64, "the adjoint runs in FP64: create the RHS with KF_FP64");
        set_dev(r->grid->ctx);
        *out = new kf_adjoint_s{r, new Adjoint(*r->r64, dt)};
        return KF_OK;
    });
}

int kf_adjoint_destroy(kf_adjoint a)
{
    return guard([&] {
        if (!a) return KF_OK;
        delete a->a;
        delete a;
        return KF_OK;
    });
}

int kf_adjoint_set_tolerance(kf_adjoint a, double tol)
{
    return guard([&] {
        KF_REQUIRE(a && tol > 0, "bad argument");
        a->a->set_tolerance(tol);
        return KF_OK;
    });
}

int kf_adjoint_reset(kf_adjoint a)
{
    return guard([&] {
        KF_REQUIRE(a, "null adjoint");
        a->a->reset_accumulators();
        return KF_OK;
    });
}

int kf_rhs_adjoint(kf_adjoint a, const double *const S[4], const double *const ob[4], double *const Fb[4])
{
    return guard([&] {
        KF_REQUIRE(a && S && ob && Fb, "null argument");
        const double2 *s4[4], *o4[4];
        double2 *f4[4];
        for (int i = 0; i < 4; ++i) {
            s4[i] = (const double2 *)S[i];
            o4[i] = (const double2 *)ob[i];
            f4[i] = (double2 *)Fb[i];
        }
        a->a->rhs_adjoint(s4, o4, f4, true);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_rk4_step_adjoint(kf_adjoint a, const double *const y0[4], double *const yb[4])
{
    return guard([&] {
        KF_REQUIRE(a && y0 && yb, "null argument");
        const double2 *y[4];
        double2 *b[4];
        for (int i = 0; i < 4; ++i) {
            y[i] = (const double2 *)y0[i];
            b[i] = (double2 *)yb[i];
        }
        a->a->step_adjoint(y, b);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_objective(kf_adjoint a, const double *const F[4], const double *dV, double *J)
{
    return guard([&] {
        KF_REQUIRE(a && F && dV && J, "null argument");
        const double2 *f[4];
        for (int i = 0; i < 4; ++i) f[i] = (const double2 *)F[i];
        *J = a->a->objective(f, dV);
        return KF_OK;
    });
}

int kf_objective_adjoint(kf_adjoint a, const double *const F[4], const double *dV, double *const Fb[4])
{
    return guard([&] {
        KF_REQUIRE(a && F && dV && Fb, "null argument");
        const double2 *f[4];
        double2 *b[4];
        for (int i = 0; i < 4; ++i) {
            f[i] = (const double2 *)F[i];
            b[i] = (double2 *)Fb[i];
        }
        a->a->objective_adjoint(f, dV, b);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_adjoint_accumulators(kf_adjoint a, double *M_bar, double *nd_bar)
{
    return guard([&] {
        KF_REQUIRE(a, "null adjoint");
        if (M_bar) *M_bar = a->a->M_bar();
        if (nd_bar) {
            Grid &g = *a->rhs->grid->g;
            KF_HIP(hipMemcpyAsync(nd_bar, a->a->nd_bar(), sizeof(double) * g.nr, hipMemcpyDeviceToDevice,
                                  a->rhs->grid->ctx->stream));
            KF_HIP(hipStreamSynchronize(a->rhs->grid->ctx->stream));
        }
        return KF_OK;
    });
}

int kf_adjoint_run(kf_adjoint a, double *const y[4], int nsteps, int k, const double *dV, const double *G, double *J,
                   double *dJdM, double *dJdQ, double *dJdFn0)
{
    return guard([&] {
        KF_REQUIRE(a && y && dV && G && J && nsteps >= 0 && k >= 1, "bad argument");
        double2 *yy[4];
        for (int i = 0; i < 4; ++i) yy[i] = (double2 *)y[i];
        *J = a->a->run(yy, nsteps, k, dV, G, dJdM, dJdQ, (double2 *)dJdFn0);
        return KF_OK;
    });
}

int kf_adjoint_set_capture_budget(kf_adjoint a, size_t bytes)
{
    return guard([&] {
        KF_REQUIRE(a, "null adjoint");
        a->a->set_capture_budget(bytes);
        return KF_OK;
    });
}

int kf_adjoint_set_keep_potentials(kf_adjoint a, int keep)
{
    return guard([&] {
        KF_REQUIRE(a, "null adjoint");
        a->a->set_keep_potentials(keep != 0);
        return KF_OK;
    });
}

int kf_adjoint_run_info(kf_adjoint a, int *mode, long *linear_iterations)
{
    return guard([&] {
        KF_REQUIRE(a, "null adjoint");
        if (mode) *mode = a->a->last_run_mode();
        if (linear_iterations) *linear_iterations = a->a->last_run_linear_iterations();
        return KF_OK;
    });
}

int kf_adjoint_step_capture_bytes(kf_adjoint a, size_t *bytes)
{
    return guard([&] {
        KF_REQUIRE(a && bytes, "null argument");
        *bytes = a->a->step_capture_bytes();
        return KF_OK;
    });
}

int kf_adjoint_spectral_to_real(kf_adjoint a, const double *dJdF, double *dJdf)
{
    return guard([&] {
        KF_REQUIRE(a && dJdF && dJdf, "null argument");
        a->a->spectral_to_real_T((const double2 *)dJdF, dJdf);
        KF_HIP(hipStreamSynchronize(a->rhs->grid->ctx->stream));
        return KF_OK;
    });
}

// ---------------------------------------------------------------- batch
struct kf_batch_s {
    kf_ctx ctx;
    Batch *b;
};

int kf_batch_create(kf_ctx ctx, int nbox, const kf_axis *axes, const kf_parameters *params, const kf_rhs_switches *sw,
                    int precision, double dt, double poisson_tol, int nstreams, kf_batch *out)
{
    return guard([&] {
        KF_REQUIRE(ctx && axes && params && sw && out && nbox > 0, "bad argument");
        KF_REQUIRE(precision == KF_FP64 || precision == KF_MIXED, "precision must be KF_FP64 or KF_MIXED");
        KF_REQUIRE(dt > 0 && poisson_tol > 0, "dt and tolerance must be positive");
        set_dev(ctx);
        std::vector<BatchBox> boxes((size_t)nbox);
        for (int b = 0; b < nbox; ++b) {
            for (int d = 0; d < 3; ++d) {
                boxes[b].ax[d] = to_desc(axes[3 * b + d]);
                KF_REQUIRE(boxes[b].ax[d].type >= 1 && boxes[b].ax[d].type <= 3, "unknown axis type");
            }
            KF_REQUIRE(fft::grid_supported(boxes[b].ax[0].n, boxes[b].ax[1].n, boxes[b].ax[2].n),
                       "grid extents not supported (2^a 3^b <= 256; Nz even)");
            boxes[b].pa = Params{params[b].M, params[b].tau, params[b].theta, params[b].mu, params[b].beta};
        }
        Switches s{sw->sv_on, sw->pen_on};
        *out = new kf_batch_s{ctx, new Batch(boxes, s, precision == KF_MIXED ? 1 : 0, dt, poisson_tol,
                                             nstreams)};
        return KF_OK;
    });
}

int kf_batch_destroy(kf_batch b)
{
    return guard([&] {
        if (!b) return KF_OK;
        delete b->b;
        delete b;
        return KF_OK;
    });
}

int kf_batch_box(kf_batch b, int box, double **FUx, double **FUy, double **FUz, double **Fn, double **phi)
{
    return guard([&] {
        KF_REQUIRE(b && box >= 0 && box < b->b->size(), "bad box index");
        double **o[4] = {FUx, FUy, FUz, Fn};
        for (int f = 0; f < 4; ++f)
            if (o[f]) *o[f] = (double *)b->b->state(box, f);
        if (phi) *phi = b->b->potential(box);
        return KF_OK;
    });
}

int kf_batch_set_fields(kf_batch b, int box, const double *nd, const double *mask)
{
    return guard([&] {
        KF_REQUIRE(b && box >= 0 && box < b->b->size(), "bad box index");
        b->b->set_fields(box, nd, mask);
        return KF_OK;
    });
}

int kf_batch_capture(kf_batch b, int poisson_iterations)
{
    return guard([&] {
        KF_REQUIRE(b && poisson_iterations > 0 && poisson_iterations <= 64, "bad argument");
        b->b->capture(poisson_iterations);
        return KF_OK;
    });
}

int kf_batch_steps(kf_batch b, int nsteps)
{
    return guard([&] {
        KF_REQUIRE(b && nsteps >= 0, "bad argument");
        b->b->steps(nsteps);
        KF_HIP(hipGetLastError());
        return KF_OK;
    });
}

int kf_batch_synchronize(kf_batch b)
{
    return guard([&] {
        KF_REQUIRE(b, "null batch");
        b->b->sync();
        return KF_OK;
    });
}

int kf_batch_poisson_failures(kf_batch b, int *count)
{
    return guard([&] {
        KF_REQUIRE(b && count, "null argument");
        *count = b->b->poisson_failures();
        return KF_OK;
    });
}

int kf_profile_enable(int family)
{
    return guard([&] {
        KF_REQUIRE(family >= -1 && family < prof::NFAM, "unknown family");
        prof::enable(family);
        return KF_OK;
    });
}

int kf_profile_read(int *launches, double *total_ms)
{
    return guard([&] {
        prof::read(launches, total_ms);
        return KF_OK;
    });
}

// ------------------------------------------------------------------ set-up (host)
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
static void philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    std::memcpy(out, c, sizeof(c));
}

int kf_setup_host(const kf_axis axes[3], unsigned long long seed, double amp, double Q, double a,
                  double *ln_n, double *nd, double *mask)
{
    return guard([&] {
        KF_REQUIRE(axes, "null axes");
        AxisDesc ad[3] = {to_desc(axes[0]), to_desc(axes[1]), to_desc(axes[2])};
        const long nx = ad[0].n, ny = ad[1].n, nz = ad[2].n, N = nx * ny * nz;
        const double pi = acos(-1.);
        if (ln_n) {
            const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
            for (long b = 0; 4 * b < N; ++b) {
                const uint32_t ctr[4] = {(uint32_t)b, (uint32_t)((uint64_t)b >> 32), 0u, 0u};
                uint32_t o[4];
                philox(ctr, key, o);
                for (int h = 0; h < 2; ++h) {
                    const double u1 = ((double)o[2 * h] + 0.5) * (1.0 / 4294967296.0);
                    const double u2 = ((double)o[2 * h + 1] + 0.5) * (1.0 / 4294967296.0);
                    const double r = sqrt(-2. * log(u1)), th = 2. * pi * u2;
                    const long i0 = 4 * b + 2 * h;
                    if (i0 < N) ln_n[i0] = log(1. + amp * (r * cos(th)));
                    if (i0 + 1 < N) ln_n[i0 + 1] = log(1. + amp * (r * sin(th)));
                }
            }
        }
        const double A = (Q / (pow(sqrt(pi), 3.) * a * a * a));
        for (long i = 0; i < nx; ++i) {
            const double x = axis_val_at(ad[0], (int)i);
            for (long j = 0; j < ny; ++j) {
                const double y = axis_val_at(ad[1], (int)j);
                for (long k = 0; k < nz; ++k) {
                    const double z = axis_val_at(ad[2], (int)k);
                    const double e = exp(-pow(x / a, 2.) - pow(y / a, 2.) - pow(z / a, 2.));
                    const long p = (i * ny + j) * nz + k;
                    if (nd) nd[p] = A * e;
                    if (mask) mask[p] = e;
                }
            }
        }
        return KF_OK;
    });
}

}  // extern "C"
