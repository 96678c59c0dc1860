// kf_effects.hip -- kielflow's effect interfaces (effect_advection.hpp:13-15,
// effect_continuityEQ.hpp:14-15, effect_force_E.hpp:14-15,
// effect_diffusion_u.hpp:15-16, effect_spectral_viscosity.hpp:20-21,
// effect_translation.hpp:15-16, effect_force_linear.hpp:14-15,
// effect_penalization.hpp:16-18) on FP64 device fields.
//
// Each effect adds its term of d/dt to a buffer, the way the reference's
// effect classes are called; the arithmetic of each term is the one that
// rhs_standard::solve inlines (rhs_standard.cpp:119-533), so a sequence of
// effects reproduces kf_rhs_solve.  The reference's library bodies differ
// from rhs_standard and are not followed: effect_diffusion_u.cpp:50-56 uses
// a linear-n pressure form and effect_spectral_viscosity.cpp:30-43 uses
// coordinates in place of k.
#include "kf_effects.h"

#include "kf_rhs_impl.h"

namespace kf {

using namespace fft;

static inline unsigned nbk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// out = [out +] c * [i k_dir] * [D23] * [filter] * in
__global__ void k_eff_spec(GridView<double> g, const double2 *in, double2 *out, double c, int d23, int dir,
                           int filt, int acc)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i, j, k;
    dec3c(idx, g.ny, g.nzc, i, j, k);
    double2 v = in[idx];
    if (d23 && !keep23(g, i, j, k)) v = mk<double2>(0., 0.);
    if (filt) v = cscale(v, phi_filter(g, i, j, k));
    if (dir >= 0) v = ik(k_along(g, dir, i, j, k), v);
    v = cscale(v, c);
    if (acc) {
        const double2 o = out[idx];
        v = mk<double2>(o.x + v.x, o.y + v.y);
    }
    out[idx] = v;
}

// out -= SV * in   (SV = 0.5 (1 - exp(-8 khat^4)), rhs_standard.cpp:51-77)
__global__ void k_eff_sv(GridView<double> g, const double2 *in, double2 *out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i, j, k;
    dec3c(idx, g.ny, g.nzc, i, j, k);
    const double sv = sv_factor(g, i, j, k);
    const double2 v = in[idx], o = out[idx];
    out[idx] = mk<double2>(o.x - sv * v.x, o.y - sv * v.y);
}

__global__ void k_eff_neg(double *a, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = -a[i];
}

// y (+)= a x  (complex)
__global__ void k_eff_axpy(double2 *y, const double2 *x, double a, int acc, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 v = x[i];
    if (acc) {
        const double2 o = y[i];
        y[i] = mk<double2>(o.x + a * v.x, o.y + a * v.y);
    } else {
        y[i] = mk<double2>(a * v.x, a * v.y);
    }
}

// minus the penalization blend of rhs_standard.cpp:483-490 in d/dt form:
//   mP = -[(1-H) F + eta H (U - fix [H]) - b U]
// F = iFFT(out), U = iFFT(in); the forward transform then takes -mP
__global__ void k_eff_pen(double *__restrict__ F, const double *__restrict__ U, const double *__restrict__ H,
                          const double *__restrict__ bar, double eta, double fix, int fix_mask, int nyz, long n)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double h = H[p], u = U[p], b = bar[(unsigned)p / (unsigned)nyz];
    const double f = fix_mask ? fix * h : fix;
    F[p] = -((1. - h) * F[p] + eta * h * (u - f) - b * u);
}

EffectWork::EffectWork(Grid &g) : g_(g)
{
    for (auto &r : r_) KF_HIP(hipMalloc(&r, sizeof(double) * g.nr));
    for (auto &c : c_) KF_HIP(hipMalloc(&c, sizeof(double2) * g.nc));
}

EffectWork::~EffectWork()
{
    for (auto &r : r_) (void)hipFree(r);
    for (auto &c : c_) (void)hipFree(c);
}

// out (+)= d_b([D23] X) with OP_partial_derivative.cpp:104-136 semantics
static void deriv(Grid &g, Plan<double> &p, const double2 *X, int d23_in, int b, double2 *out, int acc,
                  double *rtmp, hipStream_t s)
{
    if (!g.stretched(b)) {
        hipLaunchKernelGGL(k_eff_spec, dim3(nbk(g.nc)), dim3(256), 0, s, g.view<double>(), X, out, 1.0, d23_in, b,
                           0, acc);
        return;
    }
    ZStore<double> z = zs<double>(ZK_PROF, p);
    z.o0 = rtmp;
    z.prof = g.prof<double>(b);
    z.pdir = b;
    c2r_spec<double>(g, p, X, 1, b, z, s);
    XStore<double> x = xs<double>(acc ? XK_ACC : XK_WRITE, g);
    x.o0 = out;
    r2c_store<double, double>(g, p, rtmp, x, s);
}

// A = sum_b iFFT(D23 FU_b) * iFFT(D23 d_b(D23 FX))   (rhs_standard.cpp:323-343, 388-437)
static void products(Grid &g, Plan<double> &p, EffectWork &w, const double2 *FU[3], const double2 *FX, double *A,
                     hipStream_t s)
{
    for (int b = 0; b < 3; ++b) {
        ZStore<double> z = zs<double>(ZK_WRITE, p);
        z.o0 = w.r(b);
        c2r_spec<double>(g, p, FU[b], 1, -1, z, s);
    }
    for (int b = 0; b < 3; ++b) {
        ZStore<double> z = zs<double>(ZK_PROD, p);
        z.o0 = A;
        z.a0 = w.r(b);
        z.flag = (b > 0);
        if (!g.stretched(b)) {
            c2r_spec<double>(g, p, FX, 1, b, z, s);
        } else {
            deriv(g, p, FX, 1, b, w.c(0), 0, w.r(4), s);
            c2r_spec<double>(g, p, w.c(0), 1, -1, z, s);
        }
    }
}

void effect_advection(Grid &g, Plan<double> &p, EffectWork &w, const double2 *FU[3], const double2 *FX,
                      double2 *out, hipStream_t s)
{
    products(g, p, w, FU, FX, w.r(3), s);
    XStore<double> x = xs<double>(XK_ACC, g);
    x.o0 = out;
    // out -= FFT(A): accumulate FFT(-A)
    hipLaunchKernelGGL(k_eff_neg, dim3(nbk(g.nr)), dim3(256), 0, s, w.r(3), (long)g.nr);
    r2c_store<double, double>(g, p, w.r(3), x, s);
}

void effect_continuity(Grid &g, Plan<double> &p, EffectWork &w, const double2 *FU[3], const double2 *Fn,
                       double2 *out, hipStream_t s)
{
    // out = -(div(D23 FU) + FFT(u . grad(D23 Fn)))   (rhs_standard.cpp:315-345)
    products(g, p, w, FU, Fn, w.r(3), s);
    hipLaunchKernelGGL(k_eff_neg, dim3(nbk(g.nr)), dim3(256), 0, s, w.r(3), (long)g.nr);
    XStore<double> x = xs<double>(XK_WRITE, g);
    x.o0 = out;
    r2c_store<double, double>(g, p, w.r(3), x, s);
    for (int a = 0; a < 3; ++a) {
        deriv(g, p, FU[a], 1, a, w.c(1), 0, w.r(4), s);
        hipLaunchKernelGGL(k_eff_axpy, dim3(nbk(g.nc)), dim3(256), 0, s, out, (const double2 *)w.c(1), -1.0, 1,
                           (long)g.nc);
    }
}

void effect_force_E(Grid &g, Plan<double> &p, EffectWork &w, const double2 *FPhi, double2 *out[3], hipStream_t s)
{
    // out_a -= d_a(filter FFT(Phi))   (rhs_standard.cpp:219-247, filter :220)
    hipLaunchKernelGGL(k_eff_spec, dim3(nbk(g.nc)), dim3(256), 0, s, g.view<double>(), FPhi, w.c(1), 1.0, 0, -1, 1,
                       0);
    for (int a = 0; a < 3; ++a) {
        deriv(g, p, w.c(1), 0, a, w.c(0), 0, w.r(4), s);
        hipLaunchKernelGGL(k_eff_axpy, dim3(nbk(g.nc)), dim3(256), 0, s, out[a], (const double2 *)w.c(0), -1.0, 1,
                           (long)g.nc);
    }
}

void effect_diffusion_u(Grid &g, Plan<double> &p, EffectWork &w, double theta, const double2 *Fn, double2 *out[3],
                        hipStream_t s)
{
    // out_a -= d_a(Fn / theta)   (rhs_standard.cpp:219-247, log form)
    for (int a = 0; a < 3; ++a) {
        deriv(g, p, Fn, 0, a, w.c(0), 0, w.r(4), s);
        hipLaunchKernelGGL(k_eff_axpy, dim3(nbk(g.nc)), dim3(256), 0, s, out[a], (const double2 *)w.c(0),
                           -1.0 / theta, 1, (long)g.nc);
    }
}

void effect_spectral_viscosity(Grid &g, const double2 *in, double2 *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_eff_sv, dim3(nbk(g.nc)), dim3(256), 0, s, g.view<double>(), in, out);
}

void effect_translation(Grid &g, Plan<double> &p, EffectWork &w, double v, const double2 *in, double2 *out,
                        hipStream_t s)
{
    // out -= v d_x(D23 in)   (rhs_standard.cpp:279-310)
    deriv(g, p, in, 1, 0, w.c(0), 0, w.r(4), s);
    hipLaunchKernelGGL(k_eff_axpy, dim3(nbk(g.nc)), dim3(256), 0, s, out, (const double2 *)w.c(0), -v, 1, (long)g.nc);
}

void effect_force_linear(Grid &g, double k, const double2 *in, double2 *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_eff_axpy, dim3(nbk(g.nc)), dim3(256), 0, s, out, in, k, 1, (long)g.nc);
}

void effect_penalization(Grid &g, Plan<double> &p, EffectWork &w, const double *mask, double eta, double fix,
                         int fix_mask, const double2 *in, double2 *out, hipStream_t s)
{
    // out = D23 FFT((1-H) iFFT(out) + eta H (iFFT(in) - fix) - b iFFT(in))   (rhs_standard.cpp:459-533)
    ZStore<double> zu = zs<double>(ZK_WRITE, p);
    zu.o0 = w.r(0);
    c2r_spec<double>(g, p, in, 0, -1, zu, s);
    ZStore<double> zf = zs<double>(ZK_WRITE, p);
    zf.o0 = w.r(1);
    c2r_spec<double>(g, p, out, 0, -1, zf, s);
    hipLaunchKernelGGL(k_eff_pen, dim3(nbk(g.nr)), dim3(256), 0, s, w.r(1), (const double *)w.r(0), mask,
                       g.bar<double>(), eta, fix, fix_mask, g.ny * g.nz, (long)g.nr);
    XStore<double> x = xs<double>(XK_NEG_D23, g);
    x.o0 = out;
    r2c_store<double, double>(g, p, w.r(1), x, s);
}

}  // namespace kf
