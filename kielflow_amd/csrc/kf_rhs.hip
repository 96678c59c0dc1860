// kf_rhs.hip -- RHS (rhs_standard.cpp:89-566) and RK4 (RKO4.cpp:50-248).
//
// The equidistant-grid RHS fuses every spectral multiplier into a pass of an
// inverse transform and every real-space product, exp, and penalization
// blend into the last pass of the transform that produces it; 25 transforms
// per evaluation (the reference's 33 minus 8 that repeat an inverse already
// computed or undo a forward transform).  The 16 inverse transforms of the
// state fields and their derivatives share x and y passes: per field, two x
// passes (D23 F and i k_x D23 F) and three y passes serve all four z passes.
// Stretched axes
// take the general path, which materialises the stretched derivatives
// d_b X = FFT(dS_b * iFFT(D23 i k_b X)) as the reference does
// (OP_partial_derivative.cpp:104-136).
#include "kf_rhs.h"
#include "kf_rhs_impl.h"

#include <algorithm>
#include <cstdlib>

#include <cmath>
#include <vector>

namespace kf {

using namespace fft;

// functors and transform helpers: kf_rhs_impl.h

// ------------------------------------------------------------------ pointwise spectral kernels
static inline unsigned nbk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// out (+)= [i k_dir] [D23] in [* filter] ; mode: 0 write, 1 accumulate
template <typename R>
__global__ void k_spec(GridView<R> g, const C2<R> *__restrict__ in, C2<R> *out, int d23, int dir,
                       int filt, int acc)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i, j, k;
    dec3c(idx, g.ny, g.nzc, i, j, k);
    C2<R> v = in[idx];
    if (d23 && !keep23(g, i, j, k)) v = mk<C2<R>>((R)0, (R)0);
    if (filt) {
        const R f = phi_filter(g, i, j, k);
        v = mk<C2<R>>(v.x * f, v.y * f);
    }
    if (dir >= 0) v = ik(k_along(g, dir, i, j, k), v);
    if (acc) {
        C2<R> o = out[idx];
        v = mk<C2<R>>(o.x + v.x, o.y + v.y);
    }
    out[idx] = v;
}

// general path: linear terms of rhs_standard.cpp:119-163
template <typename R>
__global__ void k_lin_init(GridView<R> g, const C2<R> *F0, const C2<R> *F1, const C2<R> *F2,
                           const C2<R> *F3, C2<R> *L0, C2<R> *L1, C2<R> *L2, C2<R> *L3, R tau, int sv_on)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i, j, k;
    dec3c(idx, g.ny, g.nzc, i, j, k);
    const R sv = sv_on ? sv_factor(g, i, j, k) : (R)0;
    const R fac = sv_on ? tau + sv : tau;
    const C2<R> *F[3] = {F0, F1, F2};
    C2<R> *L[3] = {L0, L1, L2};
    for (int a = 0; a < 3; ++a) {
        C2<R> u = F[a][idx];
        L[a][idx] = mk<C2<R>>(fac * u.x, fac * u.y);
    }
    C2<R> n = F3[idx];
    L3[idx] = sv_on ? mk<C2<R>>(sv * n.x, sv * n.y) : mk<C2<R>>((R)0, (R)0);
}

// general path: lin_a += M dF_a,x ; lin_n += M dF_n,x + dF_x,x + dF_y,y + dF_z,z
template <typename R>
__global__ void k_lin_combine(long nc, C2<R> *L0, C2<R> *L1, C2<R> *L2, C2<R> *L3, const C2<R> *dx0,
                              const C2<R> *dx1, const C2<R> *dx2, const C2<R> *dnx, const C2<R> *d00,
                              const C2<R> *d11, const C2<R> *d22, R M)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nc) return;
    C2<R> *L[3] = {L0, L1, L2};
    const C2<R> *D[3] = {dx0, dx1, dx2};
    for (int a = 0; a < 3; ++a) {
        C2<R> l = L[a][idx], d = D[a][idx];
        L[a][idx] = mk<C2<R>>(l.x + M * d.x, l.y + M * d.y);
    }
    C2<R> l = L3[idx], d = dnx[idx];
    l = mk<C2<R>>(l.x + M * d.x, l.y + M * d.y);
    C2<R> a = d00[idx], b = d11[idx], c = d22[idx];
    l = mk<C2<R>>(l.x + a.x, l.y + a.y);
    l = mk<C2<R>>(l.x + b.x, l.y + b.y);
    l = mk<C2<R>>(l.x + c.x, l.y + c.y);
    L3[idx] = l;
}

template <typename R>
__global__ void k_prod_acc(R *__restrict__ o, const R *__restrict__ u, const R *__restrict__ gg, long n, int first)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (first ? (R)0 : o[i]) + u[i] * gg[i];
}

template <typename A, typename B>
__global__ void k_convert(const A *__restrict__ in, B *__restrict__ out, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (B)in[i];
}

template <typename CA, typename CB>
__global__ void k_convert2(const CA *__restrict__ in, CB *__restrict__ out, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        CA v = in[i];
        CB w;
        w.x = v.x;
        w.y = v.y;
        out[i] = w;
    }
}

void to_single(const double2 *in, float2 *out, long n, hipStream_t s)
{
    hipLaunchKernelGGL((k_convert2<double2, float2>), dim3(nbk(n)), dim3(256), 0, s, in, out, n);
}
void to_double(const float2 *in, double2 *out, long n, hipStream_t s)
{
    hipLaunchKernelGGL((k_convert2<float2, double2>), dim3(nbk(n)), dim3(256), 0, s, in, out, n);
}
void real_to_single(const double *in, float *out, long n, hipStream_t s)
{
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nbk(n)), dim3(256), 0, s, in, out, n);
}

// ------------------------------------------------------------------ Rhs
template <typename R>
static R *ralloc(long n)
{
    R *p = nullptr;
    KF_HIP(hipMalloc(&p, sizeof(R) * n));
    return p;
}

template <typename R>
Rhs<R>::Rhs(Grid &g, Poisson &p, const Params &pa, const Switches &sw, hipStream_t s)
    : g_(g), ps_(p), pa_(pa), sw_(sw), s_(s)
{
    plan_init<R>(plan_, g.nx, g.ny, g.nz, s);
    // KF_EXTRAP_ORDER=1..5 caps the order of the potential extrapolation
    if (const char *env = std::getenv("KF_EXTRAP_ORDER")) extrap_order_ = std::max(1, std::min(5, std::atoi(env)));
    pen_fused_ = std::getenv("KF_PEN_UNFUSED") == nullptr;
    prof_fused_ = std::getenv("KF_PROF_UNFUSED") == nullptr;
    xy_shared_ = prof_fused_ && std::getenv("KF_XY_UNSHARED") == nullptr && g.stretched(0) && g.stretched(1) &&
                 g.stretched(2);
    const long nr = g.nr, nc = g.nc;
    nd_ = ralloc<R>(nr);
    mask_ = ralloc<R>(nr);
    for (int b = 0; b < 3; ++b) U_[b] = ralloc<R>(nr);
    L_ = ralloc<R>(nr);
    rho_ = ralloc<R>(nr);
    rtmp_ = ralloc<R>(nr);
    for (int a = 0; a < 4; ++a) adv_[a] = ralloc<R>(nr);
    for (int a = 0; a < 4; ++a) lin_[a] = ralloc<C2<R>>(nc);
    if (!std::is_same<R, double>::value) {
        phiR_ = ralloc<R>(nr);
        // KF_RHO64=0 forms rho in R (FP32) in the mixed RHS
        const char *env = std::getenv("KF_RHO64");
        rho64_on_ = !(env && std::atoi(env) == 0);
        if (rho64_on_) {
            rho64_ = ralloc<double>(nr);
            nd64_ = ralloc<double>(nr);
            KF_HIP(hipMemsetAsync(nd64_, 0, sizeof(double) * nr, s));
            // KF_RHO_STATE=0: rho from the FP32 stage input even under RK4
            const char *e2 = std::getenv("KF_RHO_STATE");
            if (!(e2 && std::atoi(e2) == 0)) {
                plan64_ = new fft::Plan<double>();
                plan_init<double>(*plan64_, g.nx, g.ny, g.nz, s);
            }
        }
    }
    if (!g.any_stretched())
        for (auto &f : H_)
            for (auto &h : f) h = ralloc<C2<R>>((long)plan_.nx * plan_.ny * plan_.nzs);
    if (xy_shared_)
        for (auto &h : H_[0]) h = ralloc<C2<R>>((long)plan_.nx * plan_.ny * plan_.nzs);
    KF_HIP(hipMemsetAsync(nd_, 0, sizeof(R) * nr, s));
    KF_HIP(hipMemsetAsync(mask_, 0, sizeof(R) * nr, s));
    KF_HIP(hipStreamSynchronize(s));
}

template <typename R>
Rhs<R>::~Rhs()
{
    plan_free<R>(plan_);
    if (plan64_) {
        plan_free<double>(*plan64_);
        delete plan64_;
    }
    (void)hipFree(nd_);
    (void)hipFree(mask_);
    for (int b = 0; b < 3; ++b) (void)hipFree(U_[b]);
    (void)hipFree(L_);
    (void)hipFree(rho_);
    (void)hipFree(rho64_);
    (void)hipFree(nd64_);
    (void)hipFree(rtmp_);
    for (int a = 0; a < 4; ++a) {
        (void)hipFree(adv_[a]);
        (void)hipFree(lin_[a]);
    }
    (void)hipFree(F1_);
    for (auto &d : dF_) (void)hipFree(d);
    for (auto &f : H_)
        for (auto &h : f) (void)hipFree(h);
    (void)hipFree(phiR_);
    for (auto &st : ph_)
        for (auto &p : st) (void)hipFree(p);
}

template <typename R>
void Rhs<R>::set_fields(double *phi, const double *nd, const double *mask)
{
    if (phi) phi_ = phi;
    const long nr = g_.nr;
    if (nd) hipLaunchKernelGGL((k_convert<double, R>), dim3(nbk(nr)), dim3(256), 0, s_, nd, nd_, nr);
    if (nd && nd64_) KF_HIP(hipMemcpyAsync(nd64_, nd, sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
    if (mask) hipLaunchKernelGGL((k_convert<double, R>), dim3(nbk(nr)), dim3(256), 0, s_, mask, mask_, nr);
    KF_HIP(hipGetLastError());
}

template <typename R>
int Rhs<R>::poisson_failures()
{
    int f = 0;
    KF_HIP(hipMemcpyAsync(&f, ps_.device_fail_counter(), sizeof(int), hipMemcpyDeviceToHost, s_));
    KF_HIP(hipStreamSynchronize(s_));
    return f;
}

// phi = polynomial extrapolation in time (order 1..5) of the last converged
// potentials h[5] (newest), h[4], ..., h[0]: the value at the next equal
// time step of the polynomial through the newest order+1 of them
constexpr int HIST = 6;
struct Hist {
    const double *h[HIST];
};
__global__ void k_extrap(double *__restrict__ phi, Hist H, int order, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *const *h = H.h;
    double v;
    if (order >= 5) v = 6. * (h[5][i] + h[1][i]) - 15. * (h[4][i] + h[2][i]) + 20. * h[3][i] - h[0][i];
    else if (order == 4) v = 5. * (h[5][i] - h[2][i]) - 10. * (h[4][i] - h[3][i]) + h[1][i];
    else if (order == 3) v = 4. * (h[5][i] + h[3][i]) - 6. * h[4][i] - h[2][i];
    else if (order == 2) v = 3. * (h[5][i] - h[4][i]) + h[3][i];
    else v = 2. * h[5][i] - h[4][i];
    phi[i] = v;
}

// h[0] <- h[1] <- ... <- h[HIST-1] <- phi
struct HistW {
    double *h[HIST];
};
__global__ void k_shift_hist(HistW H, const double *__restrict__ phi, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int m = 0; m < HIST - 1; ++m) H.h[m][i] = H.h[m + 1][i];
    H.h[HIST - 1][i] = phi[i];
}

template <typename R>
void Rhs<R>::alloc_history()
{
    if (ph_[0][0]) return;
    for (auto &st : ph_)
        for (auto &p : st) KF_HIP(hipMalloc(&p, sizeof(double) * g_.nr));
}

template <typename R>
void Rhs<R>::set_stage(int s)
{
    stage_ = s;
    if (s >= 0) alloc_history();
}

// rhs_standard.cpp:100-113: rho = exp(iFFT(D23 Fn)) + n_d, then Poisson
template <typename R>
void Rhs<R>::poisson_step(const C2<R> *Fn, const C2<R> *staged)
{
    poisson_begin(Fn, staged);
    poisson_end();
}

// rho, the guess, and the solve's first corrections (enqueued)
template <typename R>
void Rhs<R>::poisson_begin(const C2<R> *Fn, const C2<R> *staged)
{
    if constexpr (!std::is_same<R, double>::value) {
        if (rho_src_ && plan64_) {
            // ln n and rho from the FP64 stage input (rhs_standard.cpp:100-108:
            // D23, inverse transform, exp, + n_d), ln n then rounded to FP32
            fft::Plan<double> &p = *plan64_;
            ZStore<double> z = zs<double>(ZK_RHOF, p);
            z.of = (float *)L_;
            z.od = rho64_;
            z.ad = nd64_;
            c2r_xy<double>(g_, p, rho_src_, 1, s_);
            launch_zc2r<double>(p.nz / 2, (long)p.nx * p.ny, p.tw_m, p.tw_n, RowLoadC<double>{p.scratch, p.nzs},
                                ZStoreK<double, ZK_RHOF>{z}, s_);
            staged = nullptr;
            Fn = nullptr;
        }
    }
    if (Fn || staged) {
        ZStore<R> z = zs<R>(rho64_on_ ? ZK_RHO64 : ZK_RHO, plan_);
        z.o0 = L_;
        z.o1 = rho_;
        z.a0 = nd_;
        z.od = rho64_;
        z.ad = nd64_;
        if (staged) z_staged<R>(g_, plan_, staged, false, z, s_);
        else c2r_spec<R>(g_, plan_, Fn, 1, -1, z, s_);
    }
    if (!phi_) throw Error("RHS potential field not set");
    const long nr = g_.nr;
    phi_cur_ = phi_;
    slot_ = nullptr;
    const double *rho64 = std::is_same<R, double>::value ? (const double *)rho_ : (rho64_on_ ? rho64_ : nullptr);
    const float *rho32 = std::is_same<R, double>::value || rho64_on_ ? nullptr : (const float *)rho_;
    if (src_ && stage_ >= 0 && !graph_safe_) {
        // a kept potential of this stage (a recomputation of a step whose
        // solve converged before): it passes the first convergence check
        KF_HIP(hipMemcpyAsync(ps_.guess_buffer(), src_[stage_], sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
        slot_ = ph_[stage_][0];
        ps_.next_guess_converged();
        ps_.solve_begin(slot_, rho64, rho32, s_, true);
        return;
    }
    if (stage_ >= 0 && hc_[stage_] >= 2) {
        double **h = ph_[stage_];
        const int order = std::min(hc_[stage_] - 1, extrap_order_);
        Hist H{};
        for (int m = 0; m < HIST; ++m) H.h[m] = h[m];
        if (!graph_safe_) {
            // Extrapolated guess into the solver's guess buffer; the solve
            // writes its result straight into the oldest potential of this
            // stage's history, which then becomes the newest, and the rest of
            // the evaluation reads it there.  The caller's potential receives
            // the last stage's solution once per step.
            hipLaunchKernelGGL(k_extrap, dim3(nbk(nr)), dim3(256), 0, s_, ps_.guess_buffer(), H, order, nr);
            slot_ = h[0];
            ps_.solve_begin(slot_, rho64, rho32, s_, true);
            return;
        }
        hipLaunchKernelGGL(k_extrap, dim3(nbk(nr)), dim3(256), 0, s_, phi_, H, order, nr);
    }
    ps_.solve_begin(phi_, rho64, rho32, s_, false);
}

// convergence of the solve, then the potential history
template <typename R>
void Rhs<R>::poisson_end()
{
    ps_.solve_end();
    if (stage_ < 0) return;
    const long nr = g_.nr;
    double **h = ph_[stage_];
    if (slot_) {
        for (int m = 0; m < HIST - 1; ++m) h[m] = h[m + 1];
        h[HIST - 1] = slot_;
        phi_cur_ = slot_;
        if (stage_ == 3) KF_HIP(hipMemcpyAsync(phi_, slot_, sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
    } else if (graph_safe_) {
        HistW H{};
        for (int m = 0; m < HIST; ++m) H.h[m] = h[m];
        hipLaunchKernelGGL(k_shift_hist, dim3(nbk(nr)), dim3(256), 0, s_, H, (const double *)phi_, nr);
    } else {
        double *oldest = h[0];
        for (int m = 0; m < HIST - 1; ++m) h[m] = h[m + 1];
        h[HIST - 1] = oldest;
        KF_HIP(hipMemcpyAsync(h[HIST - 1], phi_, sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
    }
    hc_[stage_]++;
    if (sink_) KF_HIP(hipMemcpyAsync(sink_[stage_], phi_cur_, sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
}

template <typename R>
void Rhs<R>::output(C2<R> *F[4])
{
    if (sw_.pen_on) {
        // rhs_standard.cpp:459-533
        for (int a = 0; a < 4; ++a) {
            ZStore<R> z = zs<R>(ZK_PEN, plan_);
            z.o0 = adv_[a];
            z.a0 = (a < 3) ? U_[a] : L_;
            z.a1 = adv_[a];
            z.a2 = mask_;
            z.a3 = g_.bar<R>();
            z.c0 = (R)2;
            z.c1 = (a == 0) ? (R)pa_.M : (a == 3 ? (R)10 : (R)0);
            z.flag = (a == 0);
            c2r_spec<R>(g_, plan_, lin_[a], 0, -1, z, s_);
            XStore<R> x = xs<R>(XK_NEG_D23, g_);
            x.o0 = F[a];
            r2c_store<R, R>(g_, plan_, adv_[a], x, s_);
        }
    } else {
        for (int a = 0; a < 4; ++a) {
            XStore<R> x = xs<R>(XK_NEG_ACC, g_);
            x.o0 = F[a];
            x.a0 = lin_[a];
            r2c_store<R, R>(g_, plan_, adv_[a], x, s_);
        }
    }
}

template <typename R>
void Rhs<R>::eval_fused(C2<R> *F[4])
{
    // x and y passes of every state field, shared by its value and its
    // three derivatives
    for (int a = 0; a < 4; ++a) stage_xy<R>(g_, plan_, F[a], H_[a][0], H_[a][1], H_[a][2], s_);
    // the Poisson solve is enqueued, then the velocities and the products,
    // which do not read the potential, and only then does the host wait for
    // the solve's convergence test
    poisson_begin(nullptr, H_[3][0]);
    // velocities (rhs_standard.cpp:258-260)
    for (int b = 0; b < 3; ++b) {
        ZStore<R> z = zs<R>(ZK_WRITE, plan_);
        z.o0 = U_[b];
        z_staged<R>(g_, plan_, H_[b][0], false, z, s_);
    }
    // advection / continuity products (rhs_standard.cpp:323-343, 403-437):
    // adv_a = sum_b U_b dF_a/db, one z pass per field (with penalization,
    // inside the output's z pass instead)
    const R *U[3] = {U_[0], U_[1], U_[2]};
    const bool pen_fused = sw_.pen_on && pen_fused_;
    if (!pen_fused)
        for (int a = 0; a < 4; ++a) zprod3<R>(g_, plan_, H_[a][1], H_[a][2], H_[a][0], U, adv_[a], s_);
    poisson_end();
    // FFT(Phi) -> filter -> + Fn/theta -> all linear terms (4 outputs)
    XStore<R> x = xs<R>(XK_LIN, g_);
    x.o0 = lin_[0];
    x.o1 = lin_[1];
    x.o2 = lin_[2];
    x.o3 = lin_[3];
    x.a0 = F[0];
    x.a1 = F[1];
    x.a2 = F[2];
    x.a3 = F[3];
    x.tau = (R)pa_.tau;
    x.M = (R)pa_.M;
    x.theta = (R)pa_.theta;
    x.sv_on = sw_.sv_on;
    r2c_store<R, double>(g_, plan_, phi_cur_, x, s_);
    if (!pen_fused) {
        output(F);
        return;
    }
    // rhs_standard.cpp:403-437 and 459-533, per field: the inverse x and y
    // passes of the linear terms, one z pass with the products, the blend and
    // the forward z transform, then the forward y and x passes with -D23
    for (int a = 0; a < 4; ++a) {
        c2r_xy<R>(g_, plan_, lin_[a], 0, s_);
        const R c1 = (a == 0) ? (R)pa_.M : (a == 3 ? (R)10 : (R)0);
        zpen<R>(g_, plan_, H_[a][1], H_[a][2], H_[a][0], U, plan_.scratch, (a < 3) ? U_[a] : L_, mask_, (R)2, c1,
                a == 0 ? 1 : 0, s_);
        XStore<R> xo = xs<R>(rk_ ? (rk_->final_ ? XK_RK_FINAL : XK_RK_STAGE) : XK_NEG_D23, g_);
        xo.o0 = F[a];
        if (rk_) {
            xo.ry = rk_->y[a];
            xo.racc = rk_->acc[a];
            xo.rc = rk_->c;
            xo.rfirst = rk_->first;
            xo.rnext = rk_->next;
            xo.r64 = (a == 3 && !rk_->final_) ? rk_->s64n : nullptr;
        }
        r2c_yx<R>(g_, plan_, xo, s_);
    }
    rk_done_ = rk_ != nullptr;
}

// d_b X of OP_partial_derivative.cpp:104-136 into out (write or accumulate)
template <typename R>
void Rhs<R>::deriv_into(const C2<R> *X, int d23_in, int b, C2<R> *out, int accumulate)
{
    if (!g_.stretched(b)) {
        hipLaunchKernelGGL(k_spec<R>, dim3(nbk(g_.nc)), dim3(256), 0, s_, g_.view<R>(), X, out, d23_in, b, 0,
                           accumulate);
        return;
    }
    XStore<R> x = xs<R>(accumulate ? XK_ACC : XK_WRITE, g_);
    x.o0 = out;
    if (prof_fused_) {
        // inverse x and y passes, then one z pass for C2R, 1/S' and R2C,
        // then the forward y and x passes
        c2r_xy<R>(g_, plan_, X, 1, s_, b);
        zprof<R>(plan_, nullptr, nullptr, g_.prof<R>(b), b, s_);
        r2c_yx<R>(g_, plan_, x, s_);
        return;
    }
    ZStore<R> z = zs<R>(ZK_PROF, plan_);
    z.o0 = rtmp_;
    z.prof = g_.prof<R>(b);
    z.pdir = b;
    c2r_spec<R>(g_, plan_, X, 1, b, z, s_);
    r2c_store<R, R>(g_, plan_, rtmp_, x, s_);
}

// d_b D23 X for b = 0, 1, 2 (OP_partial_derivative.cpp:104-136), the
// inverse x and y passes staged once (stage_xy: i k_x at the x pass, i k_y
// at the y pass, i k_z at the z load), then per axis one z pass for C2R,
// 1/S'_b and R2C, and the forward y and x passes into out[b]
template <typename R>
void Rhs<R>::derivs_shared(const C2<R> *X, C2<R> *const out[3], int accumulate)
{
    stage_xy<R>(g_, plan_, X, H_[0][0], H_[0][1], H_[0][2], s_);
    C2<R> *const src[3] = {H_[0][1], H_[0][2], H_[0][0]};
    for (int b = 0; b < 3; ++b) {
        XStore<R> x = xs<R>(accumulate ? XK_ACC : XK_WRITE, g_);
        x.o0 = out[b];
        zprof<R>(plan_, src[b], b == 2 ? g_.view<R>().k[2] : nullptr, g_.prof<R>(b), b, s_);
        r2c_yx<R>(g_, plan_, x, s_, src[b]);
    }
}

template <typename R>
void Rhs<R>::alloc_general()
{
    if (F1_) return;
    const long nc = g_.nc;
    F1_ = ralloc<C2<R>>(nc);
    for (auto &d : dF_) d = ralloc<C2<R>>(nc);
}

template <typename R>
void Rhs<R>::set_graph_safe(bool v)
{
    graph_safe_ = v;
    // a graph records buffer addresses: allocate the general path's
    // buffers now where that path will run
    if (v && (force_general_ || g_.stretched(0) || g_.stretched(1) || g_.stretched(2))) alloc_general();
}

template <typename R>
bool Rhs<R>::history_ready() const
{
    for (int h : hc_)
        if (h < extrap_order_ + 1) return false;
    return true;
}

template <typename R>
void Rhs<R>::eval_general(C2<R> *F[4])
{
    const long nc = g_.nc;
    alloc_general();
    poisson_step(F[3]);
    Capture *cap = std::is_same<R, double>::value ? active_capture() : nullptr;
    const long nr = g_.nr;
    if (cap) {
        KF_HIP(hipMemcpyAsync(cap->L, L_, sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
        KF_HIP(hipMemcpyAsync(cap->phi, phi_cur_, sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
    }
    XStore<R> x = xs<R>(XK_F1, g_);
    x.o0 = F1_;
    x.a0 = F[3];
    x.theta = (R)pa_.theta;
    r2c_store<R, double>(g_, plan_, phi_cur_, x, s_);
    hipLaunchKernelGGL(k_lin_init<R>, dim3(nbk(nc)), dim3(256), 0, s_, g_.view<R>(), (const C2<R> *)F[0],
                       (const C2<R> *)F[1], (const C2<R> *)F[2], (const C2<R> *)F[3], lin_[0], lin_[1],
                       lin_[2], lin_[3], (R)pa_.tau, sw_.sv_on);
    if (xy_shared_) {
        // the stretched derivatives apply the 2/3 rule to their input, F1
        // included (deriv_into)
        C2<R> *const o1[3] = {lin_[0], lin_[1], lin_[2]};
        derivs_shared(F1_, o1, 1);
        for (int a = 0; a < 4; ++a) {
            C2<R> *const o[3] = {dF_[3 * a], dF_[3 * a + 1], dF_[3 * a + 2]};
            derivs_shared(F[a], o, 0);
        }
    } else {
        for (int a = 0; a < 3; ++a) deriv_into(F1_, 0, a, lin_[a], 1);
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 3; ++b) deriv_into(F[a], 1, b, dF_[3 * a + b], 0);
    }
    hipLaunchKernelGGL(k_lin_combine<R>, dim3(nbk(nc)), dim3(256), 0, s_, nc, lin_[0], lin_[1], lin_[2],
                       lin_[3], (const C2<R> *)dF_[0], (const C2<R> *)dF_[3], (const C2<R> *)dF_[6],
                       (const C2<R> *)dF_[9], (const C2<R> *)dF_[0], (const C2<R> *)dF_[4],
                       (const C2<R> *)dF_[8], (R)pa_.M);
    for (int b = 0; b < 3; ++b) {
        ZStore<R> z = zs<R>(ZK_WRITE, plan_);
        z.o0 = U_[b];
        c2r_spec<R>(g_, plan_, F[b], 1, -1, z, s_);
        if (cap) KF_HIP(hipMemcpyAsync(cap->U[b], U_[b], sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
    }
    if (cap)
        for (int a = 0; a < 4; ++a)
            KF_HIP(hipMemcpyAsync(cap->dF0[a], dF_[3 * a], sizeof(double2) * nc, hipMemcpyDeviceToDevice, s_));
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 3; ++b) {
            if (cap) {
                ZStore<R> z = zs<R>(ZK_WRITE, plan_);
                z.o0 = (R *)cap->G[3 * a + b];
                c2r_spec<R>(g_, plan_, dF_[3 * a + b], 1, -1, z, s_);
                hipLaunchKernelGGL(k_prod_acc<R>, dim3(nbk(nr)), dim3(256), 0, s_, adv_[a], (const R *)U_[b],
                                   (const R *)cap->G[3 * a + b], nr, b == 0 ? 1 : 0);
                continue;
            }
            ZStore<R> z = zs<R>(ZK_PROD, plan_);
            z.o0 = adv_[a];
            z.a0 = U_[b];
            z.flag = (b > 0);
            c2r_spec<R>(g_, plan_, dF_[3 * a + b], 1, -1, z, s_);
        }
    output(F);
}

template <typename R>
void Rhs<R>::eval(double t, C2<R> *F[4])
{
    (void)t;
    if (g_.any_stretched() || force_general_ || (active_capture() && std::is_same<R, double>::value))
        eval_general(F);
    else eval_fused(F);
    KF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ RK4
template <typename R>
struct Quad {
    const double2 *y[4];
    C2<R> *S[4];
    double2 *acc[4];
};

template <typename R>
__global__ void k_rk_init(Quad<R> q, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= n) return;
    double2 v = q.y[f][i];
    q.S[f][i] = mk<C2<R>>((R)v.x, (R)v.y);
}

// after stage s: acc = (s==1 ? 1*k : acc + 2*k); S = y0 + c*k (and, where
// s64n is set, field 3's S also in FP64)
template <typename R>
__global__ void k_rk_stage(Quad<R> q, long n, double c, int first, double2 *s64n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= n) return;
    const C2<R> kk = q.S[f][i];
    const double kx = (double)kk.x, ky = (double)kk.y;
    double2 a;
    if (first) {
        a.x = 1. * kx;
        a.y = 1. * ky;
    } else {
        a = q.acc[f][i];
        a.x = a.x + 2. * kx;
        a.y = a.y + 2. * ky;
    }
    q.acc[f][i] = a;
    const double2 y0 = q.y[f][i];
    const double sx = y0.x + c * kx, sy = y0.y + c * ky;
    q.S[f][i] = mk<C2<R>>((R)sx, (R)sy);
    if (s64n && f == 3) s64n[i] = make_double2(sx, sy);
}

// y = y0 + (dt/6) (acc + 1*k4)   (RKO4.cpp:225-238); with next, also the
// next step's first stage input S = y (what k_rk_init would write)
template <typename R>
__global__ void k_rk_final(Quad<R> q, long n, double dt6, int next)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= n) return;
    double2 *Y = const_cast<double2 *>(f == 0 ? q.y[0] : f == 1 ? q.y[1] : f == 2 ? q.y[2] : q.y[3]);
    const C2<R> kk = q.S[f][i];
    const double2 a = q.acc[f][i];
    const double2 y0 = Y[i];
    double2 o;
    o.x = y0.x + dt6 * (a.x + 1. * (double)kk.x);
    o.y = y0.y + dt6 * (a.y + 1. * (double)kk.y);
    Y[i] = o;
    if (next) q.S[f][i] = mk<C2<R>>((R)o.x, (R)o.y);
}

template <typename R>
Rk4<R>::Rk4(Rhs<R> &rhs, double dt) : rhs_(rhs), dt_(dt)
{
    // the per-stage potential histories exist before any step (and so before
    // any graph capture of one)
    rhs.alloc_history();
    const long nc = rhs.grid().nc;
    for (int f = 0; f < 4; ++f) {
        S_[f] = ralloc<C2<R>>(nc);
        acc_[f] = ralloc<double2>(nc);
    }
    if (rhs.wants_rho_source()) S64n_ = ralloc<double2>(nc);
}

template <typename R>
Rk4<R>::~Rk4()
{
    for (int f = 0; f < 4; ++f) {
        (void)hipFree(S_[f]);
        (void)hipFree(acc_[f]);
    }
    (void)hipFree(S64n_);
}

// RKO4.cpp:50-248 with an FP64 running sum k1 + 2 k2 + 2 k3
template <typename R>
void Rk4<R>::step(double2 *y[4])
{
    steps(y, 1);
}

// n steps; within the call the final update of a step also writes the
// first stage input of the next one
template <typename R>
void Rk4<R>::steps(double2 *y[4], int n)
{
    hipStream_t s = rhs_.stream();
    const long nc = rhs_.grid().nc;
    Quad<R> q;
    for (int f = 0; f < 4; ++f) {
        q.y[f] = y[f];
        q.S[f] = S_[f];
        q.acc[f] = acc_[f];
    }
    dim3 grid(nbk(nc), 4);
    const double c[3] = {0.5 * dt_, 0.5 * dt_, 1.0 * dt_};
    // the update of each stage runs in the RHS's last store when the
    // evaluation path allows it, else as its own kernel
    RkEpilogue e{};
    for (int f = 0; f < 4; ++f) {
        e.y[f] = y[f];
        e.acc[f] = acc_[f];
    }
    e.s64n = S64n_;
    for (int it = 0; it < n; ++it) {
        if (it == 0) hipLaunchKernelGGL(k_rk_init<R>, grid, dim3(256), 0, s, q, nc);
        const double tt[4] = {t_, t_ + 0.5 * dt_, t_ + 0.5 * dt_, t_ + 1.0 * dt_};
        for (int st = 0; st < 4; ++st) {
            rhs_.set_stage(st);
            e.final_ = st == 3;
            e.c = st < 3 ? c[st] : dt_ / 6.;
            e.first = st == 0;
            e.next = it + 1 < n;
            rhs_.set_rk_epilogue(&e);
            // the first stage's input is the FP64 state itself
            if (S64n_) rhs_.set_rho_source(st == 0 ? (const double2 *)y[3] : (const double2 *)S64n_);
            rhs_.eval(tt[st], S_);
            rhs_.set_rho_source(nullptr);
            const bool fused = rhs_.rk_epilogue_done();
            rhs_.set_rk_epilogue(nullptr);
            if (fused) continue;
            if (st < 3)
                hipLaunchKernelGGL(k_rk_stage<R>, grid, dim3(256), 0, s, q, nc, c[st], st == 0 ? 1 : 0, S64n_);
            else
                hipLaunchKernelGGL(k_rk_final<R>, grid, dim3(256), 0, s, q, nc, dt_ / 6., it + 1 < n ? 1 : 0);
        }
        rhs_.set_stage(-1);
        KF_HIP(hipGetLastError());
        t_ += dt_;
    }
}

// euler_method.cpp:74-88: y = y + dt k
template <typename R>
__global__ void k_euler(Quad<R> q, long n, double dt)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= n) return;
    double2 *Y = const_cast<double2 *>(f == 0 ? q.y[0] : f == 1 ? q.y[1] : f == 2 ? q.y[2] : q.y[3]);
    const C2<R> kk = q.S[f][i];
    double2 v = Y[i];
    v.x = v.x + dt * (double)kk.x;
    v.y = v.y + dt * (double)kk.y;
    Y[i] = v;
}

template <typename R>
Euler<R>::Euler(Rhs<R> &rhs, double dt) : rhs_(rhs), dt_(dt)
{
    rhs.alloc_history();
    for (int f = 0; f < 4; ++f) S_[f] = ralloc<C2<R>>(rhs.grid().nc);
}

template <typename R>
Euler<R>::~Euler()
{
    for (int f = 0; f < 4; ++f) (void)hipFree(S_[f]);
}

template <typename R>
void Euler<R>::steps(double2 *y[4], int n)
{
    hipStream_t s = rhs_.stream();
    const long nc = rhs_.grid().nc;
    Quad<R> q;
    for (int f = 0; f < 4; ++f) {
        q.y[f] = y[f];
        q.S[f] = S_[f];
        q.acc[f] = nullptr;
    }
    dim3 grid(nbk(nc), 4);
    for (int it = 0; it < n; ++it) {
        hipLaunchKernelGGL(k_rk_init<R>, grid, dim3(256), 0, s, q, nc);
        rhs_.set_stage(0);
        // the stage input is the FP64 state itself
        if (rhs_.wants_rho_source()) rhs_.set_rho_source((const double2 *)y[3]);
        rhs_.eval(t_, S_);
        rhs_.set_rho_source(nullptr);
        rhs_.set_stage(-1);
        hipLaunchKernelGGL(k_euler<R>, grid, dim3(256), 0, s, q, nc, dt_);
        KF_HIP(hipGetLastError());
        t_ += dt_;
    }
}

// ------------------------------------------------------------------ OP_* on FP64 fields
void op_partial_derivative(Grid &g, Plan<double> &p, int dir, const double2 *in, double2 *out,
                           hipStream_t s)
{
    if (!g.stretched(dir)) {
        hipLaunchKernelGGL(k_spec<double>, dim3(nbk(g.nc)), dim3(256), 0, s, g.view<double>(), in, out, 0, dir,
                           0, 0);
        return;
    }
    // inverse x and y passes with D23 and i k_dir at the x load, one z pass
    // for C2R, 1/S' and R2C, the forward y and x passes
    XStore<double> x = xs<double>(XK_WRITE, g);
    x.o0 = out;
    c2r_xy<double>(g, p, in, 1, s, dir);
    zprof<double>(p, nullptr, nullptr, g.prof<double>(dir), dir, s);
    r2c_yx<double>(g, p, x, s);
}

void op_dealias23(Grid &g, double2 *inout, hipStream_t s)
{
    hipLaunchKernelGGL(k_spec<double>, dim3(nbk(g.nc)), dim3(256), 0, s, g.view<double>(),
                       (const double2 *)inout, inout, 1, -1, 0, 0);
}

void op_filter_phi(Grid &g, double2 *inout, hipStream_t s)
{
    hipLaunchKernelGGL(k_spec<double>, dim3(nbk(g.nc)), dim3(256), 0, s, g.view<double>(),
                       (const double2 *)inout, inout, 0, -1, 1, 0);
}

// CFL.cpp:43-82
__global__ void k_cfl(const double *__restrict__ Ux, const double *__restrict__ Uy,
                      const double *__restrict__ Uz, const double *__restrict__ dx, const double *__restrict__ dy,
                      const double *__restrict__ dz, int nx, int ny, int nz, double dt, double *part)
{
    __shared__ double sh[3][256];
    const long N = (long)nx * ny * nz;
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    double cx = 0., cy = 0., cz = 0.;
    if (p < N) {
        int i, j, k;
        dec3c(p, ny, nz, i, j, k);
        cx = dt * fabs(Ux[p]) / dx[i];
        cy = dt * fabs(Uy[p]) / dy[j];
        cz = dt * fabs(Uz[p]) / dz[k];
    }
    sh[0][threadIdx.x] = cx;
    sh[1][threadIdx.x] = cy;
    sh[2][threadIdx.x] = cz;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] = fmax(sh[c][threadIdx.x], sh[c][threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) part[3 * blockIdx.x + c] = sh[c][0];
}

double cfl_all(Grid &g, Plan<double> &p, double dt, const double2 *F[3], hipStream_t s)
{
    const long nr = g.nr;
    double *U = g.workspace();
    double *dd = U + 3 * nr;
    double *part = dd + (g.nx + g.ny + g.nz);
    const unsigned nb = g.cfl_blocks();
    if (!g.cfl_table) {
        std::vector<double> h(g.nx + g.ny + g.nz);
        int n3[3] = {g.nx, g.ny, g.nz};
        size_t off = 0;
        for (int d = 0; d < 3; ++d) {
            for (int i = 0; i < n3[d]; ++i) {
                // get_delta (CFL.cpp:6-12) at i and i+1, smaller of the two
                double d0 = axis_val_at(g.ax[d], i + 1) - axis_val_at(g.ax[d], i);
                double d1 = axis_val_at(g.ax[d], i + 2) - axis_val_at(g.ax[d], i + 1);
                h[off + i] = d0 < d1 ? d0 : d1;
            }
            off += n3[d];
        }
        KF_HIP(hipMemcpy(dd, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
        g.cfl_table = true;
    }
    for (int b = 0; b < 3; ++b) c2r<double>(p, F[b], U + b * nr, s);
    hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(256), 0, s, (const double *)U, (const double *)(U + nr),
                       (const double *)(U + 2 * nr), (const double *)dd, (const double *)(dd + g.nx),
                       (const double *)(dd + g.nx + g.ny), g.nx, g.ny, g.nz, dt, part);
    KF_HIP(hipMemcpyAsync(g.ws_host, part, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, s));
    KF_HIP(hipStreamSynchronize(s));
    double cmax = 0.;
    for (unsigned i = 0; i < 3 * nb; ++i) cmax = g.ws_host[i] > cmax ? g.ws_host[i] : cmax;
    if (.4 * dt / cmax >= 0.01) return 0.01;
    return .4 * dt / cmax;
}

// CFL.cpp:6-12 spacing table of the three axes, h[i] = min(x(i+1) - x(i),
// x(i+2) - x(i+1)), axis by axis
__global__ void k_cfl_spacing(AxisDesc a0, AxisDesc a1, AxisDesc a2, double *__restrict__ h)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = a0.n, n1 = a1.n, n2 = a2.n;
    if (t >= n0 + n1 + n2) return;
    const AxisDesc a = t < n0 ? a0 : (t < n0 + n1 ? a1 : a2);
    const int i = t < n0 ? t : (t < n0 + n1 ? t - n0 : t - n0 - n1);
    const double d0 = axis_val_at(a, i + 1) - axis_val_at(a, i);
    const double d1 = axis_val_at(a, i + 2) - axis_val_at(a, i + 1);
    h[t] = d0 < d1 ? d0 : d1;
}

// CFL.cpp:43-82 for one velocity component: per-block max of dt |U| / d
__global__ void k_cfl1(const double *__restrict__ U, const double *__restrict__ d, int comp, int ny, int nz,
                       long N, double dt, double *__restrict__ part)
{
    __shared__ double sh[256];
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    double c = 0.;
    if (p < N) {
        int i, j, k;
        dec3c(p, ny, nz, i, j, k);
        c = dt * fabs(U[p]) / d[comp == 0 ? i : (comp == 1 ? j : k)];
    }
    sh[threadIdx.x] = c;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// max of n partial maxima into *out (one workgroup)
__global__ void k_max_final(const double *__restrict__ part, long n, double *__restrict__ out)
{
    __shared__ double sh[256];
    double m = 0.;
    for (long q = threadIdx.x; q < n; q += 256) m = fmax(m, part[q]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

size_t cfl_all_scratch_doubles(const Grid &g)
{
    return (size_t)g.nr + (size_t)(g.nx + g.ny + g.nz) + 3 * (size_t)nbk(g.nr) + 4;
}

// cfl_all (CFL.cpp:29-82) on a scratch of cfl_all_scratch_doubles(g)
// doubles: one velocity component at a time in real space, then the
// maximum, which is independent of the order of the reductions, so the
// result is cfl_all's
double cfl_all_into(Grid &g, Plan<double> &p, double dt, const double2 *F[3], double *ws, hipStream_t s)
{
    const long nr = g.nr;
    const unsigned nb = nbk(nr);
    double *U = ws, *dd = U + nr, *part = dd + (g.nx + g.ny + g.nz), *out = part + 3 * (size_t)nb;
    const int nt = g.nx + g.ny + g.nz;
    hipLaunchKernelGGL(k_cfl_spacing, dim3((nt + 255) / 256), dim3(256), 0, s, g.ax[0], g.ax[1], g.ax[2], dd);
    const long off[3] = {0, g.nx, g.nx + g.ny};
    for (int b = 0; b < 3; ++b) {
        c2r<double>(p, F[b], U, s);
        hipLaunchKernelGGL(k_cfl1, dim3(nb), dim3(256), 0, s, (const double *)U, (const double *)(dd + off[b]), b,
                           g.ny, g.nz, nr, dt, part + (size_t)b * nb);
    }
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, s, (const double *)part, 3 * (long)nb, out);
    double cmax = 0.;
    KF_HIP(hipMemcpyAsync(&cmax, out, sizeof(double), hipMemcpyDeviceToHost, s));
    KF_HIP(hipStreamSynchronize(s));
    if (.4 * dt / cmax >= 0.01) return 0.01;
    return .4 * dt / cmax;
}

// per-block sums of f [* w] dV (dV = dvx[i] dvy[j] dvz[k])
__global__ void k_integrate(const double *__restrict__ f, const double *__restrict__ w,
                            const double *__restrict__ dvx, const double *__restrict__ dvy,
                            const double *__restrict__ dvz, int ny, int nz, long N, double *__restrict__ part)
{
    __shared__ double sh[256];
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    double v = 0.;
    if (p < N) {
        int i, j, k;
        dec3c(p, ny, nz, i, j, k);
        v = f[p] * (w ? w[p] : 1.) * (dvx[i] * dvy[j] * dvz[k]);
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

double field_integrate(Grid &g, const double *f, const double *w, hipStream_t s)
{
    const long nr = g.nr;
    const unsigned nb = nbk(nr);
    double *part = g.workspace() + 3 * nr + (g.nx + g.ny + g.nz);
    hipLaunchKernelGGL(k_integrate, dim3(nb), dim3(256), 0, s, f, w, (const double *)g.dv[0],
                       (const double *)g.dv[1], (const double *)g.dv[2], g.ny, g.nz, nr, part);
    std::vector<double> h(nb);
    KF_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    KF_HIP(hipStreamSynchronize(s));
    double sum = 0.;
    for (double v : h) sum += v;
    return sum;
}

// ------------------------------------------------ state <-> stored fields

__global__ void k_exp_inplace(double *a, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = exp(a[i]);
}

__global__ void k_fill_int(int *a, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) a[i] = 0;
}

// out = log(in); *bad |= 1 where in is not a positive finite number
__global__ void k_log_checked(const double *__restrict__ in, double *__restrict__ out, int *bad, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = in[i];
    if (!(v > 0.) || !isfinite(v)) atomicOr(bad, 1);
    out[i] = log(v);
}

// out[i][j] = f[i][j][k]
__global__ void k_plane_z(const double *__restrict__ f, double *__restrict__ out, int ny, int nz, int k, long nxy)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nxy) out[p] = f[p * nz + k];
}

void state_to_real(Grid &g, Plan<double> &p, const double2 *F[4], const double *phi, double *out, hipStream_t s)
{
    const long nr = g.nr;
    for (int a = 0; a < 4; ++a) c2r<double>(p, F[a], out + a * nr, s);
    hipLaunchKernelGGL(k_exp_inplace, dim3(nbk(nr)), dim3(256), 0, s, out + 3 * nr, nr);
    KF_HIP(hipMemcpyAsync(out + 4 * nr, phi, sizeof(double) * nr, hipMemcpyDeviceToDevice, s));
}

void state_plane_z(Grid &g, Plan<double> &p, const double2 *F[4], const double *phi, int k, double *out,
                   hipStream_t s)
{
    const long nr = g.nr, nxy = (long)g.nx * g.ny;
    double *tmp = g.workspace();
    for (int a = 0; a < 4; ++a) {
        c2r<double>(p, F[a], tmp, s);
        hipLaunchKernelGGL(k_plane_z, dim3(nbk(nxy)), dim3(256), 0, s, (const double *)tmp, out + a * nxy, g.ny,
                           g.nz, k, nxy);
    }
    hipLaunchKernelGGL(k_exp_inplace, dim3(nbk(nxy)), dim3(256), 0, s, out + 3 * nxy, nxy);
    hipLaunchKernelGGL(k_plane_z, dim3(nbk(nxy)), dim3(256), 0, s, phi, out + 4 * nxy, g.ny, g.nz, k, nxy);
    (void)nr;
}

bool state_from_real(Grid &g, Plan<double> &p, const double *in, double2 *F[4], double *phi, hipStream_t s)
{
    const long nr = g.nr;
    double *tmp = g.workspace();
    int *bad = (int *)(tmp + nr);
    for (int a = 0; a < 3; ++a) r2c<double>(p, in + a * nr, F[a], s);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, bad, 1);
    hipLaunchKernelGGL(k_log_checked, dim3(nbk(nr)), dim3(256), 0, s, in + 3 * nr, tmp, bad, nr);
    r2c<double>(p, tmp, F[3], s);
    if (phi) KF_HIP(hipMemcpyAsync(phi, in + 4 * nr, sizeof(double) * nr, hipMemcpyDeviceToDevice, s));
    int h = 0;
    KF_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    KF_HIP(hipStreamSynchronize(s));
    return h == 0;
}

// Boltzmann ions of the equilibrium start (frontend.cpp:186-189)
__global__ void k_boltzmann_ions(const double *__restrict__ phi, double theta, double *__restrict__ ni, long n)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) ni[p] = exp(-theta * phi[p]);
}

void boltzmann_ions(const double *phi, double theta, double *ni, long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_boltzmann_ions, dim3(nbk(n)), dim3(256), 0, s, phi, theta, ni, n);
}

// RKO4.cpp:140-153 (and :167-180, :194-207): out = y0 + c k, n doubles
__global__ void k_rk_axpy(long n, double c, const double *__restrict__ y0, const double *__restrict__ k,
                          double *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = y0[i] + c * k[i];
}

// RKO4.cpp:225-238: out = y0 + (dt/6)(1 k1 + 2 k2 + 2 k3 + 1 k4)
__global__ void k_rk_combine(long n, double dt6, const double *__restrict__ y0, const double *__restrict__ k1,
                             const double *__restrict__ k2, const double *__restrict__ k3,
                             const double *__restrict__ k4, double *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = y0[i] + dt6 * (1. * k1[i] + 2. * k2[i] + 2. * k3[i] + 1. * k4[i]);
}

void rk4_stage_input(long n, double c, const double *y0, const double *k, double *out, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_rk_axpy, dim3(nbk(n)), dim3(256), 0, s, n, c, y0, k, out);
}

void rk4_combine(long n, double dt, const double *y0, const double *const k[4], double *out, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_rk_combine, dim3(nbk(n)), dim3(256), 0, s, n, dt / 6., y0, k[0], k[1], k[2], k[3],
                           out);
}

template class Rhs<float>;
template class Rhs<double>;
template class Rk4<float>;
template class Rk4<double>;
template class Euler<float>;
template class Euler<double>;

}  // namespace kf
