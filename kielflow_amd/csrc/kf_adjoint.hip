// kf_adjoint.hip -- discrete adjoint of the RHS (rhs_standard.cpp:89-566),
// of the RK4 step (RKO4.cpp:50-248) and of the grain drag objective.
#include "kf_adjoint.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "kf_rhs_impl.h"

namespace kf {

using namespace fft;

static inline unsigned nbk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }
constexpr int AB = 256;
constexpr unsigned AMAXB = 1024;

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ double wsum(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double r = 0.;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
    __syncthreads();
    return r;
}

// partial sums of a . b (n reals)
__global__ void __launch_bounds__(AB) k_dotp(const double *__restrict__ a, const double *__restrict__ b, long n,
                                             double *__restrict__ part)
{
    __shared__ double sh[AB / 64];
    double acc = 0.;
    for (long p = (long)blockIdx.x * AB + threadIdx.x; p < n; p += (long)gridDim.x * AB) acc += a[p] * b[p];
    acc = wsum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// *target += scale * sum(part)
__global__ void __launch_bounds__(AB) k_addsum(const double *__restrict__ part, int nb, double scale, double *target)
{
    __shared__ double sh[AB / 64];
    double acc = 0.;
    for (int b = threadIdx.x; b < nb; b += AB) acc += part[b];
    acc = wsum(acc, sh);
    if (threadIdx.x == 0) *target += scale * acc;
}

// penalization blend transposed (rhs_standard.cpp:483-490):
//   adv_bar = (1-H) Pb;  f_bar += Pb (eps H + b);  M_bar partial += Pb eps H H
__global__ void __launch_bounds__(AB)
k_pen_adj(const double *__restrict__ Pb, const double *__restrict__ H, const double *__restrict__ bar, int nx_, int ny,
          int nz, double *__restrict__ advb, double *__restrict__ fb, double eps, int mflag,
          double *__restrict__ part)
{
    __shared__ double sh[AB / 64];
    const long N = (long)nx_ * ny * nz;
    double acc = 0.;
    for (long p = (long)blockIdx.x * AB + threadIdx.x; p < N; p += (long)gridDim.x * AB) {
        const double h = H[p], pb = Pb[p];
        const double b = bar[(unsigned)p / (unsigned)(ny * nz)];
        advb[p] = (1. - h) * pb;
        fb[p] += pb * (eps * h + b);
        if (mflag) acc += pb * eps * h * h;
    }
    if (mflag) {
        acc = wsum(acc, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = acc;
    }
}

__global__ void k_neg_copy(const double2 *__restrict__ a, double2 *__restrict__ b, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double2 v = a[i];
        v.x = -v.x;
        v.y = -v.y;
        b[i] = v;
    }
}

__global__ void k_zero(double *__restrict__ a, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.;
}

// Ub += sum_a advb_a G_a
__global__ void k_ub_acc(double *__restrict__ ub, const double *__restrict__ a0, const double *__restrict__ a1,
                         const double *__restrict__ a2, const double *__restrict__ a3, const double *__restrict__ g0,
                         const double *__restrict__ g1, const double *__restrict__ g2, const double *__restrict__ g3,
                         long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ub[i] += a0[i] * g0[i] + a1[i] * g1[i] + a2[i] * g2[i] + a3[i] * g3[i];
}

// out (+)= coef [D23] (-i k_dir) in   (pointwise d_dir^T on an equidistant axis)
__global__ void k_specT(GridView<double> g, const double2 *in, double2 *out, int dir, double coef,
                        int d23, int acc)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i, j, k;
    dec3c(idx, g.ny, g.nzc, i, j, k);
    double2 v = in[idx];
    const double kv = -k_along(g, dir, i, j, k) * coef;
    double2 w = mk<double2>(-kv * v.y, kv * v.x);
    if (d23 && !keep23(g, i, j, k)) w = mk<double2>(0., 0.);
    if (acc) {
        double2 o = out[idx];
        w.x += o.x;
        w.y += o.y;
    }
    out[idx] = w;
}

// Fb_a += (tau + SV) linb_a (a < 3); Fb_n += SV linb_n (sv on)
__global__ void k_lin_adj(GridView<double> g, double2 *F0, double2 *F1, double2 *F2, double2 *F3, const double2 *l0,
                          const double2 *l1, const double2 *l2, const double2 *l3, double tau, int sv_on)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i, j, k;
    dec3c(idx, g.ny, g.nzc, i, j, k);
    const double sv = sv_on ? sv_factor(g, i, j, k) : 0.;
    const double fac = sv_on ? tau + sv : tau;
    double2 *F[3] = {F0, F1, F2};
    const double2 *Lb[3] = {l0, l1, l2};
    for (int a = 0; a < 3; ++a) {
        double2 o = F[a][idx], l = Lb[a][idx];
        o.x += fac * l.x;
        o.y += fac * l.y;
        F[a][idx] = o;
    }
    if (sv_on) {
        double2 o = F3[idx], l = l3[idx];
        o.x += sv * l.x;
        o.y += sv * l.y;
        F3[idx] = o;
    }
}

__global__ void k_axpy2(double2 *__restrict__ y, const double2 *__restrict__ x, double a, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double2 v = y[i], w = x[i];
        v.x += a * w.x;
        v.y += a * w.y;
        y[i] = v;
    }
}

__global__ void k_scale2(double2 *__restrict__ y, const double2 *__restrict__ x, double a, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double2 w = x[i];
        y[i] = mk<double2>(a * w.x, a * w.y);
    }
}

// y = x0 + a x1
__global__ void k_lincomb2(double2 *__restrict__ y, const double2 *__restrict__ x0, const double2 *__restrict__ x1,
                           double a, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double2 u = x0[i], w = x1[i];
        y[i] = mk<double2>(u.x + a * w.x, u.y + a * w.y);
    }
}

// Lb -= lam exp(L); ndb -= lam      (rho_bar = -lam)
__global__ void k_L_adj(double *__restrict__ Lb, double *__restrict__ ndb, const double *__restrict__ lam,
                        const double *__restrict__ L, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double r = -lam[i];
        Lb[i] += r * exp(L[i]);
        ndb[i] += r;
    }
}

__global__ void k_axpy64r(double *__restrict__ y, const double *__restrict__ x, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// out = a * b * c
__global__ void k_mul3(double *__restrict__ out, const double *__restrict__ a, const double *__restrict__ b,
                       const double *__restrict__ c, double s, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = s * a[i] * b[i] * (c ? c[i] : 1.);
}

// x = 3 (h2 - h1) + h0 (order 2) or 2 h2 - h1 (order 1)
// the polynomial of degree `order` (1..4) through the newest order + 1 of
// the equally spaced solutions h[4] (newest), h[3], ..., at the next point:
// binomial weights with alternating signs
struct LamHist {
    const double *h[5];
};
__global__ void k_lam_extrap(double *__restrict__ x, LamHist hs, int order, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = hs.h[4][i], b = hs.h[3][i];
    switch (order) {
    case 1: x[i] = 2. * a - b; break;
    case 2: x[i] = 3. * (a - b) + hs.h[2][i]; break;
    case 3: x[i] = 4. * (a + hs.h[2][i]) - 6. * b - hs.h[1][i]; break;
    default: x[i] = 5. * (a - hs.h[1][i]) - 10. * (b - hs.h[2][i]) + hs.h[0][i]; break;
    }
}

// ------------------------------------------------------------------ class
template <typename T>
static T *dalloc(long n)
{
    T *p = nullptr;
    KF_HIP(hipMalloc(&p, sizeof(T) * n));
    KF_HIP(hipMemset(p, 0, sizeof(T) * n));
    return p;
}

Adjoint::Adjoint(Rhs<double> &rhs, double dt) : rhs_(rhs), g_(rhs.grid()), p_(rhs.plan()), s_(rhs.stream()), dt_(dt)
{
    const long nr = g_.nr, nc = g_.nc;
    prof_fused_ = std::getenv("KF_PROF_UNFUSED") == nullptr;
    if (const char *e = std::getenv("KF_ADJ_EXTRAP")) lam_order_ = std::max(0, std::min(4, std::atoi(e)));
    for (auto &c : cap_) {
        for (auto &u : c.U) u = dalloc<double>(nr);
        c.L = dalloc<double>(nr);
        c.phi = dalloc<double>(nr);
        for (auto &gg : c.G) gg = dalloc<double>(nr);
        for (auto &d : c.dF0) d = dalloc<double2>(nc);
    }
    for (auto &st : kb_)
        for (auto &f : st) f = dalloc<double2>(nc);
    for (auto &l : lam_) l = dalloc<double>(nr);
    for (auto &st : lamh_)
        for (auto &l : st) l = dalloc<double>(nr);
    rk_ = new Rk4<double>(rhs_, dt_);
    for (int a = 0; a < 4; ++a) {
        Fb_[a] = dalloc<double2>(nc);
        linb_[a] = dalloc<double2>(nc);
        advb_[a] = dalloc<double>(nr);
        yscratch_[a] = dalloc<double2>(nc);
    }
    for (auto &u : Ub_) u = dalloc<double>(nr);
    F1b_ = dalloc<double2>(nc);
    Yt_ = dalloc<double2>(nc);
    Lb_ = dalloc<double>(nr);
    rt_ = dalloc<double>(nr);
    Pb_ = dalloc<double>(nr);
    phib_ = dalloc<double>(nr);
    rhob_ = dalloc<double>(nr);
    ndb_ = dalloc<double>(nr);
    phiJ_ = dalloc<double>(nr);
    E_ = dalloc<double>(nr);
    Mb_ = dalloc<double>(4);
    part_ = dalloc<double>(AMAXB + 8);
}

static void free_capture(Capture &c)
{
    for (auto &u : c.U) (void)hipFree(u);
    (void)hipFree(c.L);
    (void)hipFree(c.phi);
    for (auto &gg : c.G) (void)hipFree(gg);
    for (auto &d : c.dF0) (void)hipFree(d);
}

Adjoint::~Adjoint()
{
    (void)hipStreamSynchronize(s_);
    for (auto &c : cap_) free_capture(c);
    for (auto &st : kb_)
        for (auto &f : st) (void)hipFree(f);
    for (auto &l : lam_) (void)hipFree(l);
    for (auto &st : lamh_)
        for (auto &l : st) (void)hipFree(l);
    for (void *p : slabs_) (void)hipFree(p);
    delete rk_;
    for (int a = 0; a < 4; ++a) {
        (void)hipFree(Fb_[a]);
        (void)hipFree(linb_[a]);
        (void)hipFree(advb_[a]);
        (void)hipFree(yscratch_[a]);
    }
    for (auto &u : Ub_) (void)hipFree(u);
    for (double *p : {(double *)F1b_, (double *)Yt_, Lb_, rt_, Pb_, phib_, rhob_, ndb_, phiJ_, E_, Mb_, part_})
        (void)hipFree(p);
}

void Adjoint::reset_accumulators()
{
    lam_valid_ = false;
    for (int &c : lam_count_) c = 0;
    KF_HIP(hipMemsetAsync(Mb_, 0, sizeof(double) * 4, s_));
    KF_HIP(hipMemsetAsync(ndb_, 0, sizeof(double) * g_.nr, s_));
}

double Adjoint::M_bar()
{
    double v = 0.;
    KF_HIP(hipMemcpyAsync(&v, Mb_, sizeof(double), hipMemcpyDeviceToHost, s_));
    KF_HIP(hipStreamSynchronize(s_));
    return v;
}

// out (+)= coef [D23] d_b^T Y ; stretched: d_b^T Y = -ik_b D W' R2C(dS_b C2R(W Y))
void Adjoint::deriv_T(const double2 *Y, int b, double2 *out, int acc, double coef, int d23_out)
{
    if (!g_.stretched(b)) {
        hipLaunchKernelGGL(k_specT, dim3(nbk(g_.nc)), dim3(256), 0, s_, g_.view<double>(), Y, out, b, coef, d23_out,
                           acc);
        return;
    }
    XLoadSpec<double> ld = xload<double>(Y, g_.view<double>(), 0, -1);
    ld.w = 1;
    const StridedAddr ay = saddr_y(p_);
    launch_strided<double, +1>(p_.nx, scratch_cols_x(p_), p_.tw_x, pad_cols(ld, p_), scratch_store_x(p_.scratch, p_),
                               s_, prof::FFT_X, p_.vx);
    launch_strided<double, +1>(p_.ny, scratch_cols_y(p_), p_.tw_y, LoadMasked<double>{p_.scratch, ay, (unsigned)p_.nzc},
                               StoreMasked<double>{p_.scratch, ay, (unsigned)p_.nzc}, s_, prof::FFT_Y, p_.vy);
    XStore<double> x = xs<double>(XK_ADJ, g_);
    x.o0 = out;
    x.d23 = 1;
    x.dir = b;
    x.acc = acc;
    x.coef = coef * (double)g_.nr;
    if (prof_fused_) {
        zprof<double>(p_, nullptr, nullptr, g_.prof<double>(b), b, s_);
        r2c_yx<double>(g_, p_, x, s_);
        return;
    }
    ZStore<double> z = zs<double>(ZK_PROF, p_);
    z.o0 = rt_;
    z.prof = g_.prof<double>(b);
    z.pdir = b;
    launch_zc2r<double>(p_.nz / 2, (long)p_.nx * p_.ny, p_.tw_m, p_.tw_n, RowLoadC<double>{p_.scratch, p_.nzs}, z, s_);
    r2c_store<double, double>(g_, p_, rt_, x, s_);
}

// inverse transform with the W weight, optional D23 / Phi filter, scale
static void c2r_w(Grid &g, Plan<double> &p, const double2 *src, int d23, int filt, double scale, double *out,
                  hipStream_t s)
{
    XLoadSpec<double> ld = xload<double>(src, g.view<double>(), d23, -1);
    ld.w = 1;
    ld.filt = filt;
    ZStore<double> z = zs<double>(ZK_WRITE, p);
    z.o0 = out;
    z.scale = scale;
    const StridedAddr ay = saddr_y(p);
    launch_strided<double, +1>(p.nx, scratch_cols_x(p), p.tw_x, pad_cols(ld, p), scratch_store_x(p.scratch, p), s,
                               prof::FFT_X, p.vx);
    launch_strided<double, +1>(p.ny, scratch_cols_y(p), p.tw_y, LoadMasked<double>{p.scratch, ay, (unsigned)p.nzc},
                               StoreMasked<double>{p.scratch, ay, (unsigned)p.nzc}, s, prof::FFT_Y, p.vy);
    launch_zc2r<double>(p.nz / 2, (long)p.nx * p.ny, p.tw_m, p.tw_n, RowLoadC<double>{p.scratch, p.nzs}, z, s);
}

// o (+)= coef [D23] [-ik_dir] W'/N R2C(a0 [* a1])
static void r2c_adj(Grid &g, Plan<double> &p, const double *a0, const double *a1, double2 *o, int d23, int dir,
                    int acc, double coef, hipStream_t s)
{
    XStore<double> x = xs<double>(XK_ADJ, g);
    x.o0 = o;
    x.d23 = d23;
    x.dir = dir;
    x.acc = acc;
    x.coef = coef;
    const StridedAddr ax = saddr_x(p), ay = saddr_y(p);
    if (a1)
        launch_zr2c<double>(p.nz / 2, (long)p.nx * p.ny, p.tw_m, p.tw_n, RowLoadProd<double>{a0, a1, p.nz},
                            RowStoreC<double>{p.scratch, p.nzs}, s);
    else
        launch_zr2c<double>(p.nz / 2, (long)p.nx * p.ny, p.tw_m, p.tw_n, RowLoadConv<double, double>{a0, p.nz},
                            RowStoreC<double>{p.scratch, p.nzs}, s);
    launch_strided<double, -1>(p.ny, scratch_cols_y(p), p.tw_y, LoadMasked<double>{p.scratch, ay, (unsigned)p.nzc},
                               StoreMasked<double>{p.scratch, ay, (unsigned)p.nzc}, s, prof::FFT_Y, p.vy);
    launch_strided<double, -1>(p.nx, (long)p.ny * p.nzc, p.tw_x, LoadPlain<double>{p.scratch, ax}, x, s, prof::FFT_X, p.vx);
}

// one RHS evaluation transposed, at the forward intermediates in c
static void rhs_adjoint_core(Adjoint *A, Rhs<double> &rhs, Grid &g, Plan<double> &p, hipStream_t s, const Capture &c,
                             const double2 *ob[4], double2 *Fb[4], double2 **linb, double **advb, double **Ub,
                             double *Lb, double2 *F1b, double2 *Yt, double *Pb, double *phib, double *rhob,
                             double *ndb, double *Mb, double *part, double ltol, bool guess,
                             void (Adjoint::*derivT)(const double2 *, int, double2 *, int, double, int))
{
    const long nr = g.nr, nc = g.nc;
    const Params &pa = rhs.params();
    const Switches &sw = rhs.switches();
    const unsigned nbr = nbk(nr) > AMAXB ? AMAXB : nbk(nr);
    for (int a = 0; a < 4; ++a) KF_HIP(hipMemsetAsync(Fb[a], 0, sizeof(double2) * nc, s));
    for (int b = 0; b < 3; ++b) KF_HIP(hipMemsetAsync(Ub[b], 0, sizeof(double) * nr, s));
    KF_HIP(hipMemsetAsync(Lb, 0, sizeof(double) * nr, s));
    // output stage: out = -B
    if (sw.pen_on) {
        for (int a = 0; a < 4; ++a) {
            // P_bar = R2C^T(D23 (-ob)) = -N C2R(W D23 ob)
            c2r_w(g, p, ob[a], 1, 0, -1.0, Pb, s);
            double *fb = (a < 3) ? Ub[a] : Lb;
            hipLaunchKernelGGL(k_pen_adj, dim3(nbr), dim3(AB), 0, s, (const double *)Pb, rhs.mask(), g.bar<double>(),
                               g.nx, g.ny, g.nz, advb[a], fb, 2.0, a == 0 ? 1 : 0, part);
            if (a == 0) hipLaunchKernelGGL(k_addsum, dim3(1), dim3(AB), 0, s, (const double *)part, (int)nbr, 1.0, Mb);
            // lin_bar = C2R^T((1-H) P_bar) = W'/N R2C(adv_bar)
            r2c_adj(g, p, advb[a], nullptr, linb[a], 0, -1, 0, 1.0, s);
        }
    } else {
        for (int a = 0; a < 4; ++a) {
            hipLaunchKernelGGL(k_neg_copy, dim3(nbk(nc)), dim3(256), 0, s, ob[a], linb[a], nc);
            c2r_w(g, p, ob[a], 0, 0, -1.0, advb[a], s);
        }
    }
    // advection / continuity products adv_a = sum_b U_b G_ab
    for (int b = 0; b < 3; ++b)
        hipLaunchKernelGGL(k_ub_acc, dim3(nbk(nr)), dim3(256), 0, s, Ub[b], (const double *)advb[0],
                           (const double *)advb[1], (const double *)advb[2], (const double *)advb[3],
                           (const double *)c.G[b], (const double *)c.G[3 + b], (const double *)c.G[6 + b],
                           (const double *)c.G[9 + b], nr);
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 3; ++b) {
            if (!g.stretched(b)) {
                r2c_adj(g, p, advb[a], c.U[b], Fb[a], 1, b, 1, 1.0, s);
            } else {
                r2c_adj(g, p, advb[a], c.U[b], Yt, 1, -1, 0, 1.0, s);
                (A->*derivT)(Yt, b, Fb[a], 1, 1.0, 1);
            }
        }
    // U_b = C2R(D23 F_b)
    for (int b = 0; b < 3; ++b) r2c_adj(g, p, Ub[b], nullptr, Fb[b], 1, -1, 1, 1.0, s);
    // linear terms
    hipLaunchKernelGGL(k_lin_adj, dim3(nbk(nc)), dim3(256), 0, s, g.view<double>(), Fb[0], Fb[1], Fb[2], Fb[3],
                       (const double2 *)linb[0], (const double2 *)linb[1], (const double2 *)linb[2],
                       (const double2 *)linb[3], pa.tau, sw.sv_on);
    {
        const unsigned nbc = nbk(2 * nc) > AMAXB ? AMAXB : nbk(2 * nc);
        for (int a = 0; a < 4; ++a) {
            hipLaunchKernelGGL(k_dotp, dim3(nbc), dim3(AB), 0, s, (const double *)linb[a], (const double *)c.dF0[a],
                               2 * nc, part);
            hipLaunchKernelGGL(k_addsum, dim3(1), dim3(AB), 0, s, (const double *)part, (int)nbc, 1.0, Mb);
        }
    }
    for (int a = 0; a < 4; ++a) (A->*derivT)(linb[a], 0, Fb[a], 1, pa.M, 1);          // M d_x(D23 F_a)
    for (int a = 0; a < 3; ++a) (A->*derivT)(linb[3], a, Fb[a], 1, 1.0, 1);           // div u
    for (int a = 0; a < 3; ++a) (A->*derivT)(linb[a], a, F1b, a > 0 ? 1 : 0, 1.0, 0); // grad F1
    // F1 = filter * FFT(phi) + Fn/theta
    hipLaunchKernelGGL(k_axpy2, dim3(nbk(nc)), dim3(256), 0, s, Fb[3], (const double2 *)F1b, 1.0 / pa.theta, nc);
    c2r_w(g, p, F1b, 0, 1, 1.0, phib, s);  // phi_bar = N C2R(W f F1_bar)
    // Poisson: rho_bar = -J^{-T} phi_bar
    Poisson &ps = rhs.poisson();
    ps.set_linearization(c.phi, s);
    int it = ps.linear_solve_T(rhob, phib, ltol, 200, s, guess);
    if (it < 0) throw Error("adjoint Poisson solve did not converge");
    hipLaunchKernelGGL(k_L_adj, dim3(nbk(nr)), dim3(256), 0, s, Lb, ndb, (const double *)rhob,
                       (const double *)c.L, nr);
    // L = C2R(D23 Fn)
    r2c_adj(g, p, Lb, nullptr, Fb[3], 1, -1, 1, 1.0, s);
}

void Adjoint::rhs_adjoint(const double2 *S[4], const double2 *ob[4], double2 *Fb[4], bool recompute)
{
    if (recompute) {
        for (int a = 0; a < 4; ++a)
            KF_HIP(hipMemcpyAsync(yscratch_[a], S[a], sizeof(double2) * g_.nc, hipMemcpyDeviceToDevice, s_));
        rhs_.set_stage(-1);
        rhs_.set_capture(&cap_[0]);
        rhs_.eval(0., yscratch_);
        rhs_.set_capture(nullptr);
    }
    rhs_adjoint_core(this, rhs_, g_, p_, s_, cap_[0], ob, Fb, linb_, advb_, Ub_, Lb_, F1b_, Yt_, Pb_, phib_, rhob_,
                     ndb_, Mb_, part_, ltol_, false, &Adjoint::deriv_T);
}

void Adjoint::step_adjoint(const double2 *y0[4], double2 *yb[4])
{
    // recompute the step from y0 through the forward RK4, keeping the
    // intermediates of its four stages
    for (int a = 0; a < 4; ++a)
        KF_HIP(hipMemcpyAsync(yscratch_[a], y0[a], sizeof(double2) * g_.nc, hipMemcpyDeviceToDevice, s_));
    rhs_.reset_history();
    rhs_.set_capture_stages(cap_);
    rk_->step(yscratch_);
    rhs_.set_capture_stages(nullptr);
    reverse_step(cap_, yb);
}

// y1 = y0 + dt/6 (k1 + 2 k2 + 2 k3 + k4), k_s = f(S_s), S_1 = y0,
// S_{s+1} = y0 + c_s k_s (c = dt/2, dt/2, dt) run backwards: ybar_0 = ybar_1 +
// sum_s Sbar_s, kbar_s = w_s ybar_1 + c_s Sbar_{s+1}, Sbar_s = f'(S_s)^T kbar_s
void Adjoint::reverse_step(const Capture caps[4], double2 *yb[4])
{
    const long nc = g_.nc;
    const unsigned nb = nbk(nc);
    const double wk[4] = {dt_ / 6., dt_ / 3., dt_ / 3., dt_ / 6.};
    for (int st = 0; st < 4; ++st)
        for (int a = 0; a < 4; ++a)
            hipLaunchKernelGGL(k_scale2, dim3(nb), dim3(256), 0, s_, kb_[st][a], (const double2 *)yb[a], wk[st], nc);
    const double back[4] = {0., 0.5 * dt_, 0.5 * dt_, 1.0 * dt_};  // S_st = y0 + back[st] k_{st-1}
    for (int st = 3; st >= 0; --st) {
        const double2 *ob[4] = {kb_[st][0], kb_[st][1], kb_[st][2], kb_[st][3]};
        // warm start: polynomial extrapolation (in the reverse time direction)
        // of this stage's last transposed Poisson solutions
        const long nr = g_.nr;
        double **h = lamh_[st];
        const int order = std::min(lam_order_, lam_count_[st] - 1);
        if (order >= 1) {
            LamHist hs;
            for (int q = 0; q < LAM_H; ++q) hs.h[q] = h[q];
            hipLaunchKernelGGL(k_lam_extrap, dim3(nbk(nr)), dim3(256), 0, s_, lam_[st], hs, order, nr);
        }
        rhs_adjoint_core(this, rhs_, g_, p_, s_, caps[st], ob, Fb_, linb_, advb_, Ub_, Lb_, F1b_, Yt_, Pb_, phib_,
                         lam_[st], ndb_, Mb_, part_, ltol_, lam_valid_, &Adjoint::deriv_T);
        lin_its_ += rhs_.poisson().last_linear_iterations();
        // the solution into the oldest slot, which becomes the newest
        double *old = h[0];
        for (int q = 0; q + 1 < LAM_H; ++q) h[q] = h[q + 1];
        h[LAM_H - 1] = old;
        KF_HIP(hipMemcpyAsync(old, lam_[st], sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
        if (lam_count_[st] < LAM_H) ++lam_count_[st];
        for (int a = 0; a < 4; ++a) {
            hipLaunchKernelGGL(k_axpy2, dim3(nb), dim3(256), 0, s_, yb[a], (const double2 *)Fb_[a], 1.0, nc);
            if (st > 0)
                hipLaunchKernelGGL(k_axpy2, dim3(nb), dim3(256), 0, s_, kb_[st - 1][a], (const double2 *)Fb_[a],
                                   back[st], nc);
        }
    }
    lam_valid_ = true;
    KF_HIP(hipGetLastError());
}

// J = -sum n_d dV d_x phi
double Adjoint::objective(const double2 *F[4], const double *dV)
{
    const long nr = g_.nr;
    ZStore<double> z = zs<double>(ZK_RHO, p_);
    z.o0 = E_;     // L of the final state (kept for the adjoint in rt_)
    z.o1 = rhob_;  // rho
    z.a0 = rhs_.nd();
    c2r_spec<double>(g_, p_, F[3], 1, -1, z, s_);
    KF_HIP(hipMemcpyAsync(Lb_, E_, sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
    KF_HIP(hipMemcpyAsync(phiJ_, rhs_.phi(), sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
    rhs_.poisson().solve(phiJ_, rhob_, nullptr, s_, true);
    // Phi_hat, d_x, back to real space
    XStore<double> xw = xs<double>(XK_WRITE, g_);
    xw.o0 = Yt_;
    r2c_store<double, double>(g_, p_, phiJ_, xw, s_);
    if (!g_.stretched(0)) {
        GridView<double> gv = g_.view<double>();
        hipLaunchKernelGGL(k_specT, dim3(nbk(g_.nc)), dim3(256), 0, s_, gv, (const double2 *)Yt_, Yt_, 0, -1.0, 0, 0);
    } else {
        ZStore<double> zp = zs<double>(ZK_PROF, p_);
        zp.o0 = rt_;
        zp.prof = g_.prof<double>(0);
        zp.pdir = 0;
        c2r_spec<double>(g_, p_, Yt_, 1, 0, zp, s_);
        r2c_store<double, double>(g_, p_, rt_, xw, s_);
    }
    ZStore<double> ze = zs<double>(ZK_WRITE, p_);
    ze.o0 = E_;
    c2r_spec<double>(g_, p_, Yt_, 0, -1, ze, s_);
    // weight n_d dV
    hipLaunchKernelGGL(k_mul3, dim3(nbk(nr)), dim3(256), 0, s_, rt_, rhs_.nd(), dV, (const double *)nullptr, 1.0, nr);
    const unsigned nbr = nbk(nr) > AMAXB ? AMAXB : nbk(nr);
    hipLaunchKernelGGL(k_dotp, dim3(nbr), dim3(AB), 0, s_, (const double *)rt_, (const double *)E_, nr, part_);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, s_, Mb_ + 2, 1L);
    hipLaunchKernelGGL(k_addsum, dim3(1), dim3(AB), 0, s_, (const double *)part_, (int)nbr, -1.0, Mb_ + 2);
    double J = 0.;
    KF_HIP(hipMemcpyAsync(&J, Mb_ + 2, sizeof(double), hipMemcpyDeviceToHost, s_));
    KF_HIP(hipStreamSynchronize(s_));
    return J;
}

// after objective(): Fb = dJ/dF, nd_bar += dJ/dn_d
void Adjoint::objective_adjoint(const double2 *F[4], const double *dV, double2 *Fb[4])
{
    (void)F;
    const long nr = g_.nr, nc = g_.nc;
    for (int a = 0; a < 4; ++a) KF_HIP(hipMemsetAsync(Fb[a], 0, sizeof(double2) * nc, s_));
    // E_bar = -n_d dV ; explicit nd_bar += -dV E
    hipLaunchKernelGGL(k_mul3, dim3(nbk(nr)), dim3(256), 0, s_, rt_, rhs_.nd(), dV, (const double *)nullptr, -1.0,
                       nr);
    hipLaunchKernelGGL(k_mul3, dim3(nbk(nr)), dim3(256), 0, s_, Pb_, dV, (const double *)E_, (const double *)nullptr,
                       -1.0, nr);
    hipLaunchKernelGGL(k_axpy64r, dim3(nbk(nr)), dim3(256), 0, s_, ndb_, (const double *)Pb_, nr);
    // (FFT phi)_bar = d_x^T C2R^T(E_bar)
    r2c_adj(g_, p_, rt_, nullptr, F1b_, 0, -1, 0, 1.0, s_);
    deriv_T(F1b_, 0, Yt_, 0, 1.0, 0);
    // phi_bar = R2C^T(.)
    c2r_w(g_, p_, Yt_, 0, 0, 1.0, phib_, s_);
    Poisson &ps = rhs_.poisson();
    ps.set_linearization(phiJ_, s_);
    int it = ps.linear_solve_T(rhob_, phib_, ltol_, 200, s_);
    if (it < 0) throw Error("adjoint Poisson solve did not converge");
    KF_HIP(hipMemsetAsync(E_, 0, sizeof(double) * nr, s_));
    hipLaunchKernelGGL(k_L_adj, dim3(nbk(nr)), dim3(256), 0, s_, E_, ndb_, (const double *)rhob_, (const double *)Lb_,
                       nr);
    r2c_adj(g_, p_, E_, nullptr, Fb[3], 1, -1, 0, 1.0, s_);
}

void Adjoint::spectral_to_real_T(const double2 *Y, double *out) { c2r_w(g_, p_, Y, 0, 0, 1.0, out, s_); }

// one stage's intermediates are carved from a slab in pieces rounded up to
// 256 bytes, so that every field starts on an aligned address
static size_t piece(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

static size_t stage_capture_bytes(const Grid &g)
{
    return 17 * piece(sizeof(double) * (size_t)g.nr) + 4 * piece(sizeof(double2) * (size_t)g.nc);
}

size_t Adjoint::step_capture_bytes() const { return 4 * stage_capture_bytes(g_); }

// grow the pool to the intermediates of `steps` steps, within the capture
// budget and half of the free HBM; one allocation per growth
bool Adjoint::ensure_pool(int steps)
{
    const size_t per = step_capture_bytes();
    if ((size_t)steps * per > capture_budget_) return false;
    if ((int)pool_.size() >= 4 * steps) return true;
    const int add = steps - (int)pool_.size() / 4;
    const size_t need = (size_t)add * per;
    size_t fr = 0, tot = 0;
    KF_HIP(hipMemGetInfo(&fr, &tot));
    if (need > fr / 2) return false;
    char *slab = nullptr;
    KF_HIP(hipMalloc(&slab, need));
    slabs_.push_back(slab);
    const size_t br = piece(sizeof(double) * (size_t)g_.nr), bc = piece(sizeof(double2) * (size_t)g_.nc);
    for (int q = 0; q < 4 * add; ++q) {
        Capture c{};
        for (auto &u : c.U) u = (double *)slab, slab += br;
        c.L = (double *)slab, slab += br;
        c.phi = (double *)slab, slab += br;
        for (auto &gg : c.G) gg = (double *)slab, slab += br;
        for (auto &d : c.dF0) d = (double2 *)slab, slab += bc;
        pool_.push_back(c);
    }
    return true;
}

double Adjoint::run(double2 *y[4], int nsteps, int k, const double *dV, const double *G, double *dM, double *dQ,
                    double2 *dFn0)
{
    if (k < 1) k = 1;
    const long nc = g_.nc, nr = g_.nr;
    const int nck = (nsteps + k - 1) / k;
    // checkpoints: state and converged potential every k steps
    struct Owned {
        std::vector<void *> v;
        ~Owned()
        {
            for (auto p : v) (void)hipFree(p);
        }
    } own;
    auto alloc = [&](size_t bytes) {
        void *p = nullptr;
        KF_HIP(hipMalloc(&p, bytes));
        own.v.push_back(p);
        return p;
    };
    double2 *yb[4];
    for (auto &p : yb) p = (double2 *)alloc(sizeof(double2) * nc);
    // Where HBM allows, the forward sweep keeps every stage's intermediates
    // and nothing is recomputed.  Otherwise states and potentials are stored
    // every k steps and each segment is recomputed once with its
    // intermediates kept; when even one segment does not fit, every step is
    // recomputed from the segment's stored states.
    const bool whole = ensure_pool(nsteps);
    const bool keep = whole || ensure_pool(std::min(k, nsteps));
    std::vector<double2 *> ck, seg;
    std::vector<double *> ckphi;
    if (!whole) {
        ck.resize((size_t)nck * 4);
        ckphi.resize(nck);
        for (auto &p : ck) p = (double2 *)alloc(sizeof(double2) * nc);
        for (auto &p : ckphi) p = (double *)alloc(sizeof(double) * nr);
    }
    if (!keep) {
        seg.resize((size_t)k * 4);
        for (auto &p : seg) p = (double2 *)alloc(sizeof(double2) * nc);
    }
    // With segments recomputed, the forward sweep also keeps every stage's
    // converged potential (one real field per stage, about 1/21 of a stage's
    // intermediates) where half the free HBM holds them: the recomputation
    // then starts each solve from it, and the solve ends at its first check.
    std::vector<double *> phis;
    if (keep && !whole) {
        size_t fr = 0, tot = 0;
        KF_HIP(hipMemGetInfo(&fr, &tot));
        const size_t need = (size_t)nsteps * 4 * sizeof(double) * nr;
        if (keep_potentials_ && need <= fr / 2) {
            double *slab = (double *)alloc(need);
            phis.resize((size_t)nsteps * 4);
            for (size_t q = 0; q < phis.size(); ++q) phis[q] = slab + q * nr;
        }
    }
    rk_->set_dt(dt_);
    rhs_.reset_history();
    for (int n = 0; n < nsteps; ++n) {
        if (whole) {
            rhs_.set_capture_stages(&pool_[(size_t)n * 4]);
        } else if (n % k == 0) {
            const int c = n / k;
            for (int a = 0; a < 4; ++a)
                KF_HIP(hipMemcpyAsync(ck[(size_t)c * 4 + a], y[a], sizeof(double2) * nc, hipMemcpyDeviceToDevice, s_));
            KF_HIP(hipMemcpyAsync(ckphi[c], rhs_.phi(), sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
        }
        if (!phis.empty()) rhs_.set_phi_sink(&phis[(size_t)n * 4]);
        rk_->step(y);
    }
    rhs_.set_phi_sink(nullptr);
    rhs_.set_capture_stages(nullptr);
    const double2 *yc[4] = {y[0], y[1], y[2], y[3]};
    const double J = objective(yc, dV);
    reset_accumulators();
    lam_valid_ = false;
    for (int &c : lam_count_) c = 0;
    lin_its_ = 0;
    objective_adjoint(yc, dV, yb);
    if (whole) {
        for (int n = nsteps - 1; n >= 0; --n) reverse_step(&pool_[(size_t)n * 4], yb);
    }
    for (int c = whole ? -1 : nck - 1; c >= 0; --c) {
        const int n0 = c * k, n1 = std::min(nsteps, n0 + k);
        KF_HIP(hipMemcpyAsync(rhs_.phi(), ckphi[c], sizeof(double) * nr, hipMemcpyDeviceToDevice, s_));
        rhs_.reset_history();
        if (keep) {
            for (int a = 0; a < 4; ++a)
                KF_HIP(hipMemcpyAsync(yscratch_[a], ck[(size_t)c * 4 + a], sizeof(double2) * nc,
                                      hipMemcpyDeviceToDevice, s_));
            for (int n = n0; n < n1; ++n) {
                rhs_.set_capture_stages(&pool_[(size_t)(n - n0) * 4]);
                if (!phis.empty()) rhs_.set_phi_source(&phis[(size_t)n * 4]);
                rk_->step(yscratch_);
            }
            rhs_.set_phi_source(nullptr);
            rhs_.set_capture_stages(nullptr);
            for (int n = n1 - 1; n >= n0; --n) reverse_step(&pool_[(size_t)(n - n0) * 4], yb);
        } else {
            for (int a = 0; a < 4; ++a)
                KF_HIP(hipMemcpyAsync(seg[a], ck[(size_t)c * 4 + a], sizeof(double2) * nc, hipMemcpyDeviceToDevice,
                                      s_));
            for (int n = n0 + 1; n < n1; ++n) {
                double2 *cur[4];
                for (int a = 0; a < 4; ++a) {
                    cur[a] = seg[(size_t)(n - n0) * 4 + a];
                    KF_HIP(hipMemcpyAsync(cur[a], seg[(size_t)(n - n0 - 1) * 4 + a], sizeof(double2) * nc,
                                          hipMemcpyDeviceToDevice, s_));
                }
                rk_->step(cur);
            }
            for (int n = n1 - 1; n >= n0; --n) {
                const double2 *st[4];
                for (int a = 0; a < 4; ++a) st[a] = seg[(size_t)(n - n0) * 4 + a];
                step_adjoint(st, yb);
            }
        }
    }
    if (dM) *dM = M_bar();
    if (dQ) {
        // n_d = Q G: the dynamics and the weight of J both carry Q
        const unsigned nbr = nbk(nr) > AMAXB ? AMAXB : nbk(nr);
        hipLaunchKernelGGL(k_dotp, dim3(nbr), dim3(AB), 0, s_, (const double *)ndb_, G, nr, part_);
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, s_, Mb_ + 3, 1L);
        hipLaunchKernelGGL(k_addsum, dim3(1), dim3(AB), 0, s_, (const double *)part_, (int)nbr, 1.0, Mb_ + 3);
        KF_HIP(hipMemcpyAsync(dQ, Mb_ + 3, sizeof(double), hipMemcpyDeviceToHost, s_));
    }
    if (dFn0) KF_HIP(hipMemcpyAsync(dFn0, yb[3], sizeof(double2) * nc, hipMemcpyDeviceToDevice, s_));
    KF_HIP(hipStreamSynchronize(s_));
    run_mode_ = whole ? RUN_TRAJECTORY : (keep ? RUN_SEGMENTS : RUN_RECOMPUTE);
    return J;
}

}  // namespace kf
