// kf_rhs_impl.h -- load/store functors fused into the FFT passes and the
// transform helpers built from them (shared by the RHS and its adjoint).
#pragma once

#include <type_traits>

#include "kf_fft.h"
#include "kf_grid.h"

namespace kf {

using namespace fft;

enum { ZK_WRITE = 0, ZK_PROF, ZK_RHO, ZK_PROD, ZK_PEN, ZK_RHO64, ZK_RHOF };
enum { XK_WRITE = 0, XK_ACC, XK_NEG_D23, XK_NEG_ACC, XK_F1, XK_LIN, XK_ADJ, XK_RK_STAGE, XK_RK_FINAL };

// ------------------------------------------------------------------ spectral factors
// 2/3-rule membership of mode (i, j, k): the exact per-element test
// d23_keep(q_x, q_y, q_z) reduced to a bound per column (GridView::imax)
template <typename R>
__device__ __forceinline__ bool keep23(const GridView<R> &g, int i, int j, int k)
{
    const int ip = i <= (g.nx >> 1) ? i : g.nx - i;
    return ip <= g.imax[j * g.nzc + k];
}

// spectral viscosity 0.5 (1 - exp(-8 sqrt(knorm)^8)) (rhs_standard.cpp:68-71)
template <typename R>
__device__ __forceinline__ R sv_factor(const GridView<R> &g, int i, int j, int k)
{
    const R s = (R)((g.q[0][i] + g.q[1][j]) + g.q[2][k]);
    return (R)0.5 * ((R)1 - exp(-(R)8 * (s * s)));
}

// exp(-2000 kabs^15) with the z norm of operations.hpp:23 (rhs_standard.cpp:220).
// kabs^15 = s^7 sqrt(s) for s = kabs^2, by multiplication: a few ulps from a
// general pow, at a fraction of its instructions (it runs for every element
// of the passes that filter Phi)
template <typename R>
__device__ __forceinline__ R phi_filter(const GridView<R> &g, int i, int j, int k)
{
    const R s = (R)((g.q[0][i] + g.q[1][j]) + g.qzf[k]);
    const R kab = sqrt(s);
    const R s2 = s * s;
    const R k15 = ((s2 * s2) * (s2 * s)) * kab;
    return exp((R)-2000 * k15);
}

template <typename R>
__device__ __forceinline__ C2<R> ik(R kv, C2<R> v)
{
    return mk<C2<R>>(-kv * v.y, kv * v.x);
}

template <typename R>
__device__ __forceinline__ void decode_x(const GridView<R> &g, long q, int t, int &i, int &j, int &k,
                                         long &idx)
{
    const unsigned u = (unsigned)q, z = (unsigned)g.nzc;
    const unsigned r = u / z;
    i = t;
    j = (int)r;
    k = (int)(u - r * z);
    idx = (long)t * g.ny * g.nzc + q;
}

__device__ __forceinline__ void dec3c(long p, int ny, int nz, int &i, int &j, int &k)
{
    const unsigned u = (unsigned)p, z = (unsigned)nz, y = (unsigned)ny;
    const unsigned r = u / z;
    k = (int)(u - r * z);
    const unsigned q = r / y;
    j = (int)(r - q * y);
    i = (int)q;
}

// ------------------------------------------------------------------ functors
// first pass of an inverse transform: v = [D23] [i k_dir] src
// half-spectrum weight of the transposed transforms: interior kz planes
// count twice in a C2R (W = 1/2 undoes that, W' = 2 applies it)
template <typename R>
__device__ __forceinline__ R wplane(const GridView<R> &g, int k)
{
    return (k == 0 || 2 * k == g.nz) ? (R)1 : (R)0.5;
}

template <typename R>
struct XLoadSpec {
    const C2<R> *src;
    GridView<R> g;
    int d23, dir;
    int w = 0;     // multiply by W (adjoint of R2C)
    int filt = 0;  // multiply by the Phi filter
    // the wavenumber table of axis dir (set by xload): one load per element
    const R *kd = nullptr;
    __device__ __forceinline__ C2<R> operator()(long q, int t) const
    {
        int i, j, k;
        long idx;
        decode_x(g, q, t, i, j, k, idx);
        C2<R> v = src[idx];
        if (d23 && !keep23(g, i, j, k)) return mk<C2<R>>((R)0, (R)0);
        if (dir >= 0) v = ik(kd[dir == 0 ? i : (dir == 1 ? j : k)], v);
        if (w) v = cscale(v, wplane(g, k));
        if (filt) v = cscale(v, phi_filter(g, i, j, k));
        return v;
    }
};

// XLoadSpec of src with [D23] and [i k_dir] (dir = -1: none)
template <typename R>
inline XLoadSpec<R> xload(const C2<R> *src, const GridView<R> &gv, int d23, int dir)
{
    XLoadSpec<R> l{src, gv, d23, dir};
    l.kd = dir >= 0 ? gv.k[dir] : nullptr;
    return l;
}

// real rows of a product a0*a1 (forward transform of a product)
template <typename R>
struct RowLoadProd {
    const R *a0, *a1;
    int nz;
    __device__ __forceinline__ C2<R> operator()(long row, int n) const
    {
        const C2<R> u = reinterpret_cast<const C2<R> *>(a0 + row * nz)[n];
        const C2<R> v = reinterpret_cast<const C2<R> *>(a1 + row * nz)[n];
        return mk<C2<R>>(u.x * v.x, u.y * v.y);
    }
};

// last pass of an inverse transform (real rows, pairs 2n, 2n+1)
template <typename R>
struct ZStore {
    int kind, nz, ny;
    R scale;
    R *o0, *o1;
    const R *a0, *a1, *a2, *a3;
    R c0, c1;
    int flag;
    const R *prof;
    int pdir;
    double *od;         // ZK_RHO64, ZK_RHOF: rho in FP64
    const double *ad;   // ZK_RHO64, ZK_RHOF: n_d in FP64
    float *of;          // ZK_RHOF: ln n rounded to FP32
    __device__ __forceinline__ void operator()(long row, int n, C2<R> v) const { apply<-1>(row, n, v); }
    // K >= 0: the kind fixed at compile time (ZStoreK), so only its code is built
    template <int K>
    __device__ __forceinline__ void apply(long row, int n, C2<R> v) const
    {
        using C = C2<R>;
        const R x0 = v.x * scale, x1 = v.y * scale;
        const long base = row * nz + 2 * n;
        switch (K >= 0 ? K : kind) {
        case ZK_WRITE:
            *reinterpret_cast<C *>(o0 + base) = mk<C>(x0, x1);
            break;
        case ZK_PROF: {
            R p0, p1;
            if (pdir == 2) {
                p0 = prof[2 * n];
                p1 = prof[2 * n + 1];
            } else {
                const unsigned rr = (unsigned)row, yy = (unsigned)ny, qq = rr / yy;
                const int id = (pdir == 1) ? (int)(rr - qq * yy) : (int)qq;
                p0 = p1 = prof[id];
            }
            *reinterpret_cast<C *>(o0 + base) = mk<C>(x0 * p0, x1 * p1);
            break;
        }
        case ZK_RHO: {
            const C d = *reinterpret_cast<const C *>(a0 + base);
            *reinterpret_cast<C *>(o0 + base) = mk<C>(x0, x1);
            *reinterpret_cast<C *>(o1 + base) = mk<C>(exp(x0) + d.x, exp(x1) + d.y);
            break;
        }
        case ZK_RHO64: {
            // ln n in R; rho = exp(ln n) + n_d formed in FP64, so that the
            // FP64 Newton residual sees no rounding of rho ~ 1 to R
            const double2 d = *reinterpret_cast<const double2 *>(ad + base);
            *reinterpret_cast<C *>(o0 + base) = mk<C>(x0, x1);
            *reinterpret_cast<double2 *>(od + base) = make_double2(exp((double)x0) + d.x, exp((double)x1) + d.y);
            break;
        }
        case ZK_RHOF: {
            // an FP64 transform of the FP64 stage input: ln n to FP32 for the
            // rest of the mixed RHS, rho = exp(ln n) + n_d in FP64 from the
            // unrounded ln n
            const double2 d = *reinterpret_cast<const double2 *>(ad + base);
            *reinterpret_cast<float2 *>(of + base) = make_float2((float)x0, (float)x1);
            *reinterpret_cast<double2 *>(od + base) = make_double2(exp((double)x0) + d.x, exp((double)x1) + d.y);
            break;
        }
        case ZK_PROD: {
            const C u = *reinterpret_cast<const C *>(a0 + base);
            C prev = flag ? *reinterpret_cast<const C *>(o0 + base) : mk<C>((R)0, (R)0);
            *reinterpret_cast<C *>(o0 + base) = mk<C>(prev.x + u.x * x0, prev.y + u.y * x1);
            break;
        }
        case ZK_PEN: {
            // rhs_standard.cpp:483-490: eps1 H (f + c H) + (1-H) B + b f
            const C f = *reinterpret_cast<const C *>(a0 + base);
            const C ad = *reinterpret_cast<const C *>(a1 + base);
            const C H = *reinterpret_cast<const C *>(a2 + base);
            const R b = a3[(unsigned)row / (unsigned)ny];
            const R t0 = flag ? c1 * H.x : c1, t1 = flag ? c1 * H.y : c1;
            R p0 = c0 * H.x * (f.x + t0);
            R p1 = c0 * H.y * (f.y + t1);
            p0 += ((R)1 - H.x) * (x0 + ad.x);
            p1 += ((R)1 - H.y) * (x1 + ad.y);
            p0 += b * f.x;
            p1 += b * f.y;
            *reinterpret_cast<C *>(o0 + base) = mk<C>(p0, p1);
            break;
        }
        }
    }
};

// last pass of a forward transform
template <typename R>
struct XStore {
    int kind;
    GridView<R> g;
    C2<R> *o0, *o1, *o2, *o3;
    const C2<R> *a0, *a1, *a2, *a3;
    R tau, M, theta;
    int sv_on;
    // XK_ADJ: o0 (+)= coef * [D23] * [-i k_dir] * W'/N * X   (adjoint of C2R)
    int d23, dir, acc;
    R coef, invN;
    // XK_RK_STAGE / XK_RK_FINAL: the RHS k = -D23 X of one field (in R, as
    // XK_NEG_D23 stores it) goes straight into the RK4 update of that field
    // (the arithmetic of k_rk_stage / k_rk_final): ry the FP64 state,
    // racc the FP64 accumulator, o0 the next stage input
    double2 *ry, *racc;
    double rc;
    int rfirst, rnext;
    // XK_RK_STAGE: the next stage input also in FP64 (before its rounding
    // to R), where set
    double2 *r64;
    // linear terms of velocity component a (rhs_standard.cpp:119-140,
    // 219-247, 279-310) from its spectrum u; adds i k_a [D23] u to ln
    __device__ __forceinline__ C2<R> lin_u(C2<R> u, R fac, C2<R> F1, bool kp, R kx, R ka, C2<R> &ln) const
    {
        using C = C2<R>;
        C l = mk<C>(fac * u.x, fac * u.y);
        const C dF1 = ik(ka, F1);
        l = mk<C>(l.x + dF1.x, l.y + dF1.y);
        const C ud = kp ? u : mk<C>((R)0, (R)0);
        const C dx = ik(kx, ud);
        l = mk<C>(l.x + M * dx.x, l.y + M * dx.y);
        const C da = ik(ka, ud);
        ln = mk<C>(ln.x + da.x, ln.y + da.y);
        return l;
    }
    __device__ __forceinline__ void operator()(long q, int t, C2<R> X) const { apply<-1>(q, t, X); }
    // K >= 0: the kind fixed at compile time (XStoreK), so only its code is built
    template <int K>
    __device__ __forceinline__ void apply(long q, int t, C2<R> X) const
    {
        using C = C2<R>;
        int i, j, k;
        long idx;
        decode_x(g, q, t, i, j, k, idx);
        switch (K >= 0 ? K : kind) {
        case XK_ADJ: {
            C v = mk<C>((R)0, (R)0);
            if (!d23 || keep23(g, i, j, k)) {
                const R f = coef * invN / wplane(g, k);
                v = cscale(X, f);
                if (dir >= 0) v = ik(-k_along(g, dir, i, j, k), v);
            }
            if (acc) {
                C o = o0[idx];
                v = mk<C>(o.x + v.x, o.y + v.y);
            }
            o0[idx] = v;
            break;
        }
        case XK_WRITE:
            o0[idx] = X;
            break;
        case XK_ACC: {
            C o = o0[idx];
            o0[idx] = mk<C>(o.x + X.x, o.y + X.y);
            break;
        }
        case XK_NEG_D23:
            o0[idx] = keep23(g, i, j, k) ? mk<C>(-X.x, -X.y) : mk<C>((R)0, (R)0);
            break;
        case XK_RK_STAGE: {
            // after stage s: acc = (s == 1 ? 1 k : acc + 2 k); S = y0 + c k
            const C kk = keep23(g, i, j, k) ? mk<C>(-X.x, -X.y) : mk<C>((R)0, (R)0);
            const double kx = (double)kk.x, ky = (double)kk.y;
            double2 a;
            if (rfirst) {
                a.x = 1. * kx;
                a.y = 1. * ky;
            } else {
                a = racc[idx];
                a.x = a.x + 2. * kx;
                a.y = a.y + 2. * ky;
            }
            racc[idx] = a;
            const double2 y0 = ry[idx];
            const double sx = y0.x + rc * kx, sy = y0.y + rc * ky;
            o0[idx] = mk<C>((R)sx, (R)sy);
            if (r64) r64[idx] = make_double2(sx, sy);
            break;
        }
        case XK_RK_FINAL: {
            // y = y0 + (dt/6) (acc + 1 k4)   (RKO4.cpp:225-238); with rnext,
            // also the next step's first stage input S = y
            const C kk = keep23(g, i, j, k) ? mk<C>(-X.x, -X.y) : mk<C>((R)0, (R)0);
            const double2 a = racc[idx];
            const double2 y0 = ry[idx];
            double2 o;
            o.x = y0.x + rc * (a.x + 1. * (double)kk.x);
            o.y = y0.y + rc * (a.y + 1. * (double)kk.y);
            ry[idx] = o;
            if (rnext) o0[idx] = mk<C>((R)o.x, (R)o.y);
            break;
        }
        case XK_NEG_ACC: {
            C l = a0[idx];
            o0[idx] = mk<C>(-(l.x + X.x), -(l.y + X.y));
            break;
        }
        case XK_F1: {
            const R f = phi_filter(g, i, j, k), rth = (R)1 / theta;
            C n = a0[idx];
            o0[idx] = mk<C>(f * X.x + n.x * rth, f * X.y + n.y * rth);
            break;
        }
        case XK_LIN: {
            // rhs_standard.cpp:119-140, 219-247, 279-319 on an equidistant grid
            const R f = phi_filter(g, i, j, k), rth = (R)1 / theta;
            const C fn = a3[idx];
            const C F1 = mk<C>(f * X.x + fn.x * rth, f * X.y + fn.y * rth);
            const bool kp = keep23(g, i, j, k);
            const R sv = sv_on ? sv_factor(g, i, j, k) : (R)0;
            const R kx = g.k[0][i], ky = g.k[1][j], kz = g.k[2][k];
            C dnx;
            {
                const C fd = kp ? fn : mk<C>((R)0, (R)0);
                dnx = ik(kx, fd);
            }
            C ln = sv_on ? mk<C>(sv * fn.x, sv * fn.y) : mk<C>((R)0, (R)0);
            ln = mk<C>(ln.x + M * dnx.x, ln.y + M * dnx.y);
            const R fac = sv_on ? (tau + sv) : tau;
            o0[idx] = lin_u(a0[idx], fac, F1, kp, kx, kx, ln);
            o1[idx] = lin_u(a1[idx], fac, F1, kp, kx, ky, ln);
            o2[idx] = lin_u(a2[idx], fac, F1, kp, kx, kz, ln);
            o3[idx] = ln;
            break;
        }
        }
    }
};

// Store functors with the kind fixed at compile time: a kernel then holds
// only that kind's code (the run-time switch builds every kind into every
// pass kernel)
template <typename R, int K>
struct ZStoreK : ZStore<R> {
    __device__ __forceinline__ void operator()(long row, int n, C2<R> v) const { this->template apply<K>(row, n, v); }
};
template <typename R, int K>
struct XStoreK : XStore<R> {
    __device__ __forceinline__ void operator()(long q, int t, C2<R> X) const { this->template apply<K>(q, t, X); }
};

// call f with the store as its compile-time-kind wrapper for the kinds of
// the fused RHS, else as is
template <typename R, class F>
inline void with_zstore(const ZStore<R> &z, F &&f)
{
    switch (z.kind) {
    case ZK_WRITE: f(ZStoreK<R, ZK_WRITE>{z}); break;
    case ZK_RHO: f(ZStoreK<R, ZK_RHO>{z}); break;
    case ZK_RHO64: f(ZStoreK<R, ZK_RHO64>{z}); break;
    case ZK_PEN: f(ZStoreK<R, ZK_PEN>{z}); break;
    default: f(z);
    }
}
template <typename R, class F>
inline void with_xstore(const XStore<R> &x, F &&f)
{
    switch (x.kind) {
    case XK_NEG_D23: f(XStoreK<R, XK_NEG_D23>{x}); break;
    case XK_RK_STAGE: f(XStoreK<R, XK_RK_STAGE>{x}); break;
    case XK_RK_FINAL: f(XStoreK<R, XK_RK_FINAL>{x}); break;
    case XK_NEG_ACC: f(XStoreK<R, XK_NEG_ACC>{x}); break;
    case XK_LIN: f(XStoreK<R, XK_LIN>{x}); break;
    default: f(x);
    }
}

// first load of an inverse transform with at most the 2/3 rule (no
// derivative, weight or filter): XLoadSpec's common case, built alone
template <typename R, int D23>
struct XLoadMask {
    const C2<R> *src;
    GridView<R> g;
    __device__ __forceinline__ C2<R> operator()(long q, int t) const
    {
        int i, j, k;
        long idx;
        decode_x(g, q, t, i, j, k, idx);
        const C2<R> v = src[idx];
        if (D23 && !keep23(g, i, j, k)) return mk<C2<R>>((R)0, (R)0);
        return v;
    }
};

template <typename R, class F>
inline void with_xload(const XLoadSpec<R> &l, F &&f)
{
    if (l.dir < 0 && !l.w && !l.filt) {
        if (l.d23) f(XLoadMask<R, 1>{l.src, l.g});
        else f(XLoadMask<R, 0>{l.src, l.g});
    } else {
        f(l);
    }
}

template <typename R, typename RS>
struct RowLoadConv {
    const RS *p;
    int nz;
    __device__ __forceinline__ C2<R> operator()(long row, int n) const
    {
        const C2<RS> v = reinterpret_cast<const C2<RS> *>(p + row * nz)[n];
        return mk<C2<R>>((R)v.x, (R)v.y);
    }
};

// ------------------------------------------------------------------ transforms with functors
template <typename R>
inline void c2r_spec(Grid &g, Plan<R> &p, const C2<R> *src, int d23, int dir, const ZStore<R> &st,
                     hipStream_t s)
{
    GridView<R> gv = g.view<R>();
    const StridedAddr ay = saddr_y(p);
    with_xload<R>(xload<R>(src, gv, d23, dir), [&](const auto &ld) {
        launch_strided<R, +1>(p.nx, scratch_cols_x(p), p.tw_x, pad_cols(ld, p), scratch_store_x(p.scratch, p), s,
                              prof::FFT_X, p.vx);
    });
    launch_strided<R, +1>(p.ny, scratch_cols_y(p), p.tw_y, LoadMasked<R>{p.scratch, ay, (unsigned)p.nzc},
                          StoreMasked<R>{p.scratch, ay, (unsigned)p.nzc}, s, prof::FFT_Y, p.vy);
    with_zstore<R>(st, [&](const auto &z) {
        launch_zc2r<R>(p.nz / 2, (long)p.nx * p.ny, p.tw_m, p.tw_n, RowLoadC<R>{p.scratch, p.nzs}, z, s);
    });
}

// y pass load of the scratch layout with i k_y (the y index is t) applied
template <typename R>
struct LoadMaskedK {
    const C2<R> *p;
    StridedAddr a;
    unsigned nvalid;
    const R *k;
    __device__ __forceinline__ C2<R> operator()(long q, int t) const
    {
        if ((unsigned)q % (unsigned)a.qdiv >= nvalid) return mk<C2<R>>((R)0, (R)0);
        return ik(k[t], p[a(q, t)]);
    }
};

// z pass load of half-spectrum rows of the scratch layout with i k_z applied
template <typename R>
struct RowLoadCK {
    const C2<R> *p;
    int nzs;
    const R *k;
    __device__ __forceinline__ C2<R> operator()(long row, int kk) const { return ik(k[kk], p[row * nzs + kk]); }
};

// The x and y passes of the inverse transforms of a state field F and its
// x and y derivatives, in the scratch layout:
//   H  = FFT_y FFT_x (D23 F)
//   Hx = FFT_y FFT_x (i k_x D23 F)
//   Hy = FFT_y (i k_y FFT_x (D23 F))      (i k_y commutes with FFT_x)
// A z pass of H gives F in real space, of H with i k_z at its load dF/dz,
// of Hx dF/dx and of Hy dF/dy (OP_partial_derivative.cpp:104-116 on an
// equidistant grid, dealiased first as at rhs_standard.cpp:253-256).
template <typename R>
inline void stage_xy(Grid &g, Plan<R> &p, const C2<R> *F, C2<R> *H, C2<R> *Hx, C2<R> *Hy, hipStream_t s)
{
    GridView<R> gv = g.view<R>();
    const StridedAddr ay = saddr_y(p);
    const long qx = scratch_cols_x(p), qy = scratch_cols_y(p);
    const unsigned nv = (unsigned)p.nzc;
    // one load of F: D23 F -> H and i k_x D23 F -> Hx
    launch_strided_ik<R, +1>(p.nx, qx, p.tw_x, pad_cols(XLoadMask<R, 1>{F, gv}, p), scratch_store_x(H, p),
                             scratch_store_x(Hx, p), gv.k[0], s, prof::FFT_X, p.vx);
    // one load of H: H in place and i k_y H -> Hy
    launch_strided_ik<R, +1>(p.ny, qy, p.tw_y, LoadMasked<R>{H, ay, nv}, StoreMasked<R>{H, ay, nv},
                             StoreMasked<R>{Hy, ay, nv}, gv.k[1], s, prof::FFT_Y, p.vy);
    launch_strided<R, +1>(p.ny, qy, p.tw_y, LoadMasked<R>{Hx, ay, nv}, StoreMasked<R>{Hx, ay, nv}, s,
                          prof::FFT_Y, p.vy);
}

// adv = sum_b U_b dF/db for one field (rhs_standard.cpp:323-343, 403-437)
// in one z pass: the C2R transforms of the staged Hx (dF/dx), Hy (dF/dy) and
// H with i k_z (dF/dz) run one after another in the workgroup, each product
// is accumulated in registers in the order b = 0, 1, 2, and adv is written
// once.  Same arithmetic as three accumulating passes (ZK_PROD), a third of
// the U reads and adv traffic.
template <typename R>
struct ProdAcc {
    C2<R> *acc;   // the thread's accumulator slots
    int *pos;     // their output positions (pairs n of the real row)
    const R *u;   // U_b
    long row;
    int nz;
    R scale;
    bool first, live;
    __device__ __forceinline__ void st(int t, C2<R> v, int slot) const
    {
        if (!live) return;
        const R x0 = v.x * scale, x1 = v.y * scale;
        const C2<R> uu = reinterpret_cast<const C2<R> *>(u + row * nz)[t];
        const C2<R> prev = first ? mk<C2<R>>((R)0, (R)0) : acc[slot];
        acc[slot] = mk<C2<R>>(prev.x + uu.x * x0, prev.y + uu.y * x1);
        pos[slot] = t;
    }
};

template <typename R, int M, int RB>
__global__ void __launch_bounds__(tpt_for(M) * RB)
k_zprod3(const C2<R> *__restrict__ twM_g, const C2<R> *__restrict__ twN_g, long rows, int nz, int nzs,
         const C2<R> *Hx, const C2<R> *Hy, const C2<R> *H, const R *kz, const R *U0, const R *U1, const R *U2,
         R *adv, R scale)
{
    using C = C2<R>;
    using Buf = RowBuf<C, M>;
    constexpr int TPT = tpt_for(M);
    constexpr int WG = TPT * RB;
    constexpr int E = out_slots(M, TPT);
    __shared__ C lds[RB * Buf::SC];
    __shared__ C tw[M];
    const int tid = threadIdx.x;
    const long r0 = (long)xcd_block(blockIdx.x, gridDim.x) * RB;
    for (int m = tid; m < M; m += WG) tw[m] = twM_g[m];
    const int c = tid / TPT, tt = tid % TPT;
    const long row = r0 + c;
    const bool live = row < rows;
    const Buf A{lds, c};
    C acc[E];
    int pos[E];
#pragma unroll
    for (int e = 0; e < E; ++e) pos[e] = -1;
    const RowLoadC<R> lx{Hx, nzs}, ly{Hy, nzs};
    const RowLoadCK<R> lz{H, nzs, kz};
    if constexpr (merge_in_regs(M)) {
        C z[M / TPT];
        merge_regs<C, M>(lx, row, live, tt, twN_g, z);
        fft_io<C, M, TPT, +1, false, false>(tw, tt, SrcRegs<C>{z}, ProdAcc<R>{acc, pos, U0, row, nz, scale, true, live},
                                            A, A);
        __syncthreads();
        merge_regs<C, M>(ly, row, live, tt, twN_g, z);
        fft_io<C, M, TPT, +1, false, false>(tw, tt, SrcRegs<C>{z},
                                            ProdAcc<R>{acc, pos, U1, row, nz, scale, false, live}, A, A);
        __syncthreads();
        merge_regs<C, M>(lz, row, live, tt, twN_g, z);
        fft_io<C, M, TPT, +1, false, false>(tw, tt, SrcRegs<C>{z},
                                            ProdAcc<R>{acc, pos, U2, row, nz, scale, false, live}, A, A);
    } else {
        fft_io<C, M, TPT, +1, false, false>(tw, tt, MergeLoad<C, M, RowLoadC<R>>{lx, row, live, twN_g},
                                            ProdAcc<R>{acc, pos, U0, row, nz, scale, true, live}, A, A);
        __syncthreads();
        fft_io<C, M, TPT, +1, false, false>(tw, tt, MergeLoad<C, M, RowLoadC<R>>{ly, row, live, twN_g},
                                            ProdAcc<R>{acc, pos, U1, row, nz, scale, false, live}, A, A);
        __syncthreads();
        fft_io<C, M, TPT, +1, false, false>(tw, tt, MergeLoad<C, M, RowLoadCK<R>>{lz, row, live, twN_g},
                                            ProdAcc<R>{acc, pos, U2, row, nz, scale, false, live}, A, A);
    }
    if (live) {
        C *o = reinterpret_cast<C *>(adv + row * nz);
#pragma unroll
        for (int e = 0; e < E; ++e)
            if (pos[e] >= 0) o[pos[e]] = acc[e];
    }
}

template <typename R>
inline void zprod3(Grid &g, Plan<R> &p, const C2<R> *Hx, const C2<R> *Hy, const C2<R> *H, const R *const U[3],
                   R *adv, hipStream_t s)
{
    prof::Scope ps_(prof::FFT_ZC2R, s);
    const long rows = (long)p.nx * p.ny;
    const R *kz = g.view<R>().k[2];
    const R scale = (R)(1.0 / (double)p.nr);
    // FP32: half the usual rows per workgroup.  The register count (112)
    // allows four waves per SIMD either way; with four-wave workgroups a
    // barrier holds fewer waves (111 us against 114-116 at 256^3).
    switch (p.nz / 2) {
#define KF_CASE(NN)                                                                                            \
    case NN: {                                                                                                 \
        constexpr int RB = cb_for<R>(NN) / (sizeof(R) == 4 ? 2 : 1);                                          \
        hipLaunchKernelGGL((k_zprod3<R, NN, RB>), dim3((unsigned)((rows + RB - 1) / RB)), dim3(tpt_for(NN) * RB), \
                           0, s, p.tw_m, p.tw_n, rows, p.nz, p.nzs, Hx, Hy, H, kz, U[0], U[1], U[2], adv, scale); \
        break;                                                                                                 \
    }
        KF_FFT_LENGTHS(KF_CASE)
#undef KF_CASE
    default:
        throw Error("unsupported z length " + std::to_string(p.nz));
    }
}

// The penalization blend of rhs_standard.cpp:483-490, eps1 H (f + c H) +
// (1-H) B + b f, applied to the C2R output of the linear terms: the thread's
// slots hold the products sum_b U_b dF_a/db (the real-space part of B) on
// entry and the blend on exit.  The last pass of every C2R of one length
// gives a thread the same output positions, so slot e sits at pos[e] in
// each of the four transforms.
template <typename R>
struct PenBlend {
    C2<R> *acc;
    const R *f, *mask;
    R b;
    long row;
    int nz;
    R scale, c0, c1;
    int flag;
    bool live;
    __device__ __forceinline__ void st(int t, C2<R> v, int slot) const
    {
        if (!live) return;
        using C = C2<R>;
        const R x0 = v.x * scale, x1 = v.y * scale;
        const C ad = acc[slot];
        const C ff = reinterpret_cast<const C *>(f + row * nz)[t];
        const C H = reinterpret_cast<const C *>(mask + row * nz)[t];
        const R t0 = flag ? c1 * H.x : c1, t1 = flag ? c1 * H.y : c1;
        R p0 = c0 * H.x * (ff.x + t0);
        R p1 = c0 * H.y * (ff.y + t1);
        p0 += ((R)1 - H.x) * (x0 + ad.x);
        p1 += ((R)1 - H.y) * (x1 + ad.y);
        p0 += b * ff.x;
        p1 += b * ff.y;
        acc[slot] = mk<C>(p0, p1);
    }
};

// One field's products, penalization blend and forward z transform in one
// z pass (rhs_standard.cpp:403-437 and 459-533).  Per row: the three C2R
// transforms of the staged Hx, Hy and i k_z H accumulate sum_b U_b dF_a/db
// in registers (as k_zprod3); the C2R of the linear terms' x/y-transformed
// spectrum (lin, rows of nzs) is blended with them, with f (U_a or ln n),
// the mask and the barrier; the blend goes to LDS and its R2C overwrites
// the row of lin in place.  Same arithmetic as k_zprod3, the ZK_PEN z pass
// and k_zr2c one after another, without the two real fields in between.
template <typename R, int M, int RB>
__global__ void __launch_bounds__(tpt_for(M) * RB)
k_zpen(const C2<R> *__restrict__ twM_g, const C2<R> *__restrict__ twN_g, long rows, int nz, int nzs, int ny,
       const C2<R> *Hx, const C2<R> *Hy, const C2<R> *H, const R *kz, const R *U0, const R *U1, const R *U2,
       C2<R> *lin, const R *f, const R *mask, const R *bar, R c0, R c1, int flag, R scale)
{
    using C = C2<R>;
    using Buf = RowBuf<C, M>;
    constexpr int TPT = tpt_for(M);
    constexpr int WG = TPT * RB;
    constexpr int E = out_slots(M, TPT);
    __shared__ C lds[RB * Buf::SC];
    __shared__ C tw[M];
    __shared__ C twN[M + 1];
    const int tid = threadIdx.x;
    const long r0 = (long)xcd_block(blockIdx.x, gridDim.x) * RB;
    for (int m = tid; m < M; m += WG) tw[m] = twM_g[m];
    for (int m = tid; m <= M; m += WG) twN[m] = twN_g[m];
    __syncthreads();
    const int c = tid / TPT, tt = tid % TPT;
    const long row = r0 + c;
    const bool live = row < rows;
    const Buf A{lds, c};
    C acc[E];
    int pos[E];
#pragma unroll
    for (int e = 0; e < E; ++e) pos[e] = -1;
    const RowLoadC<R> lx{Hx, nzs}, ly{Hy, nzs}, ll{lin, nzs};
    const RowLoadCK<R> lz{H, nzs, kz};
    const R b = live ? bar[(unsigned)row / (unsigned)ny] : (R)0;
    const PenBlend<R> pen{acc, f, mask, b, row, nz, scale, c0, c1, flag, live};
    if constexpr (merge_in_regs(M)) {
        C z[M / TPT];
        merge_regs<C, M>(lx, row, live, tt, twN, z);
        fft_io<C, M, TPT, +1, false, false>(tw, tt, SrcRegs<C>{z}, ProdAcc<R>{acc, pos, U0, row, nz, scale, true, live},
                                            A, A);
        __syncthreads();
        merge_regs<C, M>(ly, row, live, tt, twN, z);
        fft_io<C, M, TPT, +1, false, false>(tw, tt, SrcRegs<C>{z},
                                            ProdAcc<R>{acc, pos, U1, row, nz, scale, false, live}, A, A);
        __syncthreads();
        merge_regs<C, M>(lz, row, live, tt, twN, z);
        fft_io<C, M, TPT, +1, false, false>(tw, tt, SrcRegs<C>{z},
                                            ProdAcc<R>{acc, pos, U2, row, nz, scale, false, live}, A, A);
        __syncthreads();
        merge_regs<C, M>(ll, row, live, tt, twN, z);
        fft_io<C, M, TPT, +1, false, false>(tw, tt, SrcRegs<C>{z}, pen, A, A);
    } else {
        fft_io<C, M, TPT, +1, false, false>(tw, tt, MergeLoad<C, M, RowLoadC<R>>{lx, row, live, twN},
                                            ProdAcc<R>{acc, pos, U0, row, nz, scale, true, live}, A, A);
        __syncthreads();
        fft_io<C, M, TPT, +1, false, false>(tw, tt, MergeLoad<C, M, RowLoadC<R>>{ly, row, live, twN},
                                            ProdAcc<R>{acc, pos, U1, row, nz, scale, false, live}, A, A);
        __syncthreads();
        fft_io<C, M, TPT, +1, false, false>(tw, tt, MergeLoad<C, M, RowLoadCK<R>>{lz, row, live, twN},
                                            ProdAcc<R>{acc, pos, U2, row, nz, scale, false, live}, A, A);
        __syncthreads();
        fft_io<C, M, TPT, +1, false, false>(tw, tt, MergeLoad<C, M, RowLoadC<R>>{ll, row, live, twN}, pen, A, A);
    }
    // the blended row into LDS (pairs 2t, 2t+1 at t), then its forward
    // Nz/2-point transform in place
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e)
        if (pos[e] >= 0) A.st(pos[e], live ? acc[e] : mk<C>((R)0, (R)0));
    __syncthreads();
    fft_io<C, M, TPT, -1, false, true, true>(tw, tt, A, A, A, A);
    __syncthreads();
    // split: X[k] = (Z[k] + conj Z[M-k])/2 - i/2 W^k (Z[k] - conj Z[M-k]), as k_zr2c
    for (int e = tid; e < RB * (M + 1); e += WG) {
        const int cc = e / (M + 1), k = e % (M + 1);
        const long rr = r0 + cc;
        if (rr < rows) {
            const Buf Z{lds, cc};
            C zk = Z.ld(k == M ? 0 : k);
            C zm = cconj(Z.ld(k == 0 ? 0 : M - k));
            C ev = mk<C>((R)0.5 * (zk.x + zm.x), (R)0.5 * (zk.y + zm.y));
            C od = mk<C>((R)0.5 * (zk.x - zm.x), (R)0.5 * (zk.y - zm.y));
            C wod = cmul(twN[k], od);
            lin[rr * nzs + k] = mk<C>(ev.x + wod.y, ev.y - wod.x);
        }
    }
}

// first-pass store of k_zprof: the real pair t of the row, times the
// stretched axis' 1/S' exactly as the ZK_PROF z store multiplies it, kept in
// the thread's slots
template <typename R>
struct ProfMul {
    C2<R> *acc;
    int *pos;
    const R *prof;
    int pdir, ny;
    long row;
    R scale;
    bool live;
    __device__ __forceinline__ void st(int t, C2<R> v, int slot) const
    {
        if (!live) return;
        const R x0 = v.x * scale, x1 = v.y * scale;
        R p0, p1;
        if (pdir == 2) {
            p0 = prof[2 * t];
            p1 = prof[2 * t + 1];
        } else {
            const unsigned rr = (unsigned)row, yy = (unsigned)ny, qq = rr / yy;
            const int id = (pdir == 1) ? (int)(rr - qq * yy) : (int)qq;
            p0 = p1 = prof[id];
        }
        acc[slot] = mk<C2<R>>(x0 * p0, x1 * p1);
        pos[slot] = t;
    }
};

// The metric multiply of a derivative along a stretched axis in one z pass
// (OP_partial_derivative.cpp:104-136): per row, the C2R of the x/y-
// transformed spectrum (the scratch, rows of nzs), times 1/S' of the axis,
// into LDS, and its R2C back over the row in place.  Same arithmetic as the
// ZK_PROF z pass and k_zr2c one after another, without the real field in
// between.
template <typename R, int M, int RB, typename L>
__global__ void __launch_bounds__(tpt_for(M) * RB)
k_zprof(const C2<R> *__restrict__ twM_g, const C2<R> *__restrict__ twN_g, long rows, int ny, int nzs,
        C2<R> *rowsbuf, L ld, const R *prof, int pdir, R scale)
{
    using C = C2<R>;
    using Buf = RowBuf<C, M>;
    constexpr int TPT = tpt_for(M);
    constexpr int WG = TPT * RB;
    constexpr int E = out_slots(M, TPT);
    __shared__ C lds[RB * Buf::SC];
    __shared__ C tw[M];
    __shared__ C twN[M + 1];
    const int tid = threadIdx.x;
    const long r0 = (long)xcd_block(blockIdx.x, gridDim.x) * RB;
    for (int m = tid; m < M; m += WG) tw[m] = twM_g[m];
    for (int m = tid; m <= M; m += WG) twN[m] = twN_g[m];
    __syncthreads();
    const int c = tid / TPT, tt = tid % TPT;
    const long row = r0 + c;
    const bool live = row < rows;
    const Buf A{lds, c};
    C acc[E];
    int pos[E];
#pragma unroll
    for (int e = 0; e < E; ++e) pos[e] = -1;
    const ProfMul<R> pm{acc, pos, prof, pdir, ny, row, scale, live};
    if constexpr (merge_in_regs(M)) {
        C z[M / TPT];
        merge_regs<C, M>(ld, row, live, tt, twN, z);
        fft_io<C, M, TPT, +1, false, false>(tw, tt, SrcRegs<C>{z}, pm, A, A);
    } else {
        fft_io<C, M, TPT, +1, false, false>(tw, tt, MergeLoad<C, M, L>{ld, row, live, twN}, pm, A, A);
    }
    // the product row into LDS (pairs 2t, 2t+1 at t), then its forward
    // Nz/2-point transform in place
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e)
        if (pos[e] >= 0) A.st(pos[e], live ? acc[e] : mk<C>((R)0, (R)0));
    __syncthreads();
    fft_io<C, M, TPT, -1, false, true, true>(tw, tt, A, A, A, A);
    __syncthreads();
    // split, as k_zr2c
    for (int e = tid; e < RB * (M + 1); e += WG) {
        const int cc = e / (M + 1), k = e % (M + 1);
        const long rr = r0 + cc;
        if (rr < rows) {
            const Buf Z{lds, cc};
            C zk = Z.ld(k == M ? 0 : k);
            C zm = cconj(Z.ld(k == 0 ? 0 : M - k));
            C ev = mk<C>((R)0.5 * (zk.x + zm.x), (R)0.5 * (zk.y + zm.y));
            C od = mk<C>((R)0.5 * (zk.x - zm.x), (R)0.5 * (zk.y - zm.y));
            C wod = cmul(twN[k], od);
            rowsbuf[rr * nzs + k] = mk<C>(ev.x + wod.y, ev.y - wod.x);
        }
    }
}

// k_zprof over all rows of buf (scratch layout; the plan's scratch when
// null), with i k_z applied at the load when kz is set
template <typename R>
inline void zprof(Plan<R> &p, C2<R> *buf, const R *kz, const R *prof, int pdir, hipStream_t s)
{
    prof::Scope ps_(prof::FFT_ZC2R, s);
    const long rows = (long)p.nx * p.ny;
    const R scale = (R)(1.0 / (double)p.nr);
    C2<R> *b = buf ? buf : p.scratch;
    auto go = [&](const auto &ld) {
        using L = std::decay_t<decltype(ld)>;
        switch (p.nz / 2) {
#define KF_CASE(NN)                                                                                              \
    case NN: {                                                                                                   \
        constexpr int RB = cb_for<R>(NN);                                                                        \
        hipLaunchKernelGGL((k_zprof<R, NN, RB, L>), dim3((unsigned)((rows + RB - 1) / RB)),                      \
                           dim3(tpt_for(NN) * RB), 0, s, p.tw_m, p.tw_n, rows, p.ny, p.nzs, b, ld, prof, pdir,  \
                           scale);                                                                               \
        break;                                                                                                   \
    }
            KF_FFT_LENGTHS(KF_CASE)
#undef KF_CASE
        default:
            throw Error("unsupported z length " + std::to_string(p.nz));
        }
    };
    if (kz) go(RowLoadCK<R>{b, p.nzs, kz});
    else go(RowLoadC<R>{b, p.nzs});
}

// k_zpen over all rows: lin (scratch layout, after the inverse x and y
// passes of the linear terms) is replaced by the forward z transform of the
// blend
template <typename R>
inline void zpen(Grid &g, Plan<R> &p, const C2<R> *Hx, const C2<R> *Hy, const C2<R> *H, const R *const U[3],
                 C2<R> *lin, const R *f, const R *mask, R c0, R c1, int flag, hipStream_t s)
{
    prof::Scope ps_(prof::FFT_ZC2R, s);
    const long rows = (long)p.nx * p.ny;
    const R *kz = g.view<R>().k[2];
    const R scale = (R)(1.0 / (double)p.nr);
    switch (p.nz / 2) {
#define KF_CASE(NN)                                                                                            \
    case NN: {                                                                                                 \
        constexpr int RB = cb_for<R>(NN) / (sizeof(R) == 4 ? 2 : 1);                                          \
        hipLaunchKernelGGL((k_zpen<R, NN, RB>), dim3((unsigned)((rows + RB - 1) / RB)), dim3(tpt_for(NN) * RB), \
                           0, s, p.tw_m, p.tw_n, rows, p.nz, p.nzs, p.ny, Hx, Hy, H, kz, U[0], U[1], U[2], lin, f, \
                           mask, g.bar<R>(), c0, c1, flag, scale);                                             \
        break;                                                                                                 \
    }
        KF_FFT_LENGTHS(KF_CASE)
#undef KF_CASE
    default:
        throw Error("unsupported z length " + std::to_string(p.nz));
    }
}

// the x and y passes of an inverse transform of src (2/3 rule at the load
// when d23) into the plan's scratch, for a z pass of the caller's own
template <typename R>
inline void c2r_xy(Grid &g, Plan<R> &p, const C2<R> *src, int d23, hipStream_t s, int dir = -1)
{
    GridView<R> gv = g.view<R>();
    const StridedAddr ay = saddr_y(p);
    with_xload<R>(xload<R>(src, gv, d23, dir), [&](const auto &ld) {
        launch_strided<R, +1>(p.nx, scratch_cols_x(p), p.tw_x, pad_cols(ld, p), scratch_store_x(p.scratch, p), s,
                              prof::FFT_X, p.vx);
    });
    launch_strided<R, +1>(p.ny, scratch_cols_y(p), p.tw_y, LoadMasked<R>{p.scratch, ay, (unsigned)p.nzc},
                          StoreMasked<R>{p.scratch, ay, (unsigned)p.nzc}, s, prof::FFT_Y, p.vy);
}

// the y and x passes of a forward transform from buf (scratch layout; the
// plan's scratch when null), after a z pass of the caller's own, into an
// XStore; the y pass runs in place on buf
template <typename R>
inline void r2c_yx(Grid &g, Plan<R> &p, const XStore<R> &st, hipStream_t s, C2<R> *buf = nullptr)
{
    const StridedAddr ax = saddr_x(p), ay = saddr_y(p);
    C2<R> *b = buf ? buf : p.scratch;
    launch_strided<R, -1>(p.ny, scratch_cols_y(p), p.tw_y, LoadMasked<R>{b, ay, (unsigned)p.nzc},
                          StoreMasked<R>{b, ay, (unsigned)p.nzc}, s, prof::FFT_Y, p.vy);
    with_xstore<R>(st, [&](const auto &x) {
        launch_strided<R, -1>(p.nx, (long)p.ny * p.nzc, p.tw_x, LoadPlain<R>{b, ax}, x, s, prof::FFT_X, p.vx);
    });
}

// z pass of a staged field into a ZStore; with i k_z when dz
template <typename R>
inline void z_staged(Grid &g, Plan<R> &p, const C2<R> *H, bool dz, const ZStore<R> &st, hipStream_t s)
{
    const long rows = (long)p.nx * p.ny;
    with_zstore<R>(st, [&](const auto &z) {
        if (dz)
            launch_zc2r<R>(p.nz / 2, rows, p.tw_m, p.tw_n, RowLoadCK<R>{H, p.nzs, g.view<R>().k[2]}, z, s);
        else
            launch_zc2r<R>(p.nz / 2, rows, p.tw_m, p.tw_n, RowLoadC<R>{H, p.nzs}, z, s);
    });
}

template <typename R, typename RS>
inline void r2c_store(Grid &g, Plan<R> &p, const RS *src, const XStore<R> &st, hipStream_t s)
{
    const StridedAddr ax = saddr_x(p), ay = saddr_y(p);
    launch_zr2c<R>(p.nz / 2, (long)p.nx * p.ny, p.tw_m, p.tw_n, RowLoadConv<R, RS>{src, p.nz},
                   RowStoreC<R>{p.scratch, p.nzs}, s);
    launch_strided<R, -1>(p.ny, scratch_cols_y(p), p.tw_y, LoadMasked<R>{p.scratch, ay, (unsigned)p.nzc},
                          StoreMasked<R>{p.scratch, ay, (unsigned)p.nzc}, s, prof::FFT_Y, p.vy);
    with_xstore<R>(st, [&](const auto &x) {
        launch_strided<R, -1>(p.nx, (long)p.ny * p.nzc, p.tw_x, LoadPlain<R>{p.scratch, ax}, x, s, prof::FFT_X, p.vx);
    });
}

template <typename R>
inline ZStore<R> zs(int kind, const Plan<R> &p)
{
    ZStore<R> z{};
    z.kind = kind;
    z.nz = p.nz;
    z.ny = p.ny;
    z.scale = (R)(1.0 / (double)p.nr);
    return z;
}

template <typename R>
inline XStore<R> xs(int kind, Grid &g)
{
    XStore<R> x{};
    x.kind = kind;
    x.g = g.view<R>();
    x.dir = -1;
    x.coef = (R)1;
    x.invN = (R)(1.0 / (double)g.nr);
    return x;
}


}  // namespace kf
