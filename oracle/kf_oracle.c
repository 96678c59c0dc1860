/*
 * kf_oracle.c -- CPU FP64 restatement of kielflow's RHS + RK4 + nonlinear
 * Poisson path.  TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Each function names the reference file:line it follows.  The FFT is this
 * file's own mixed-radix DFT; FFTW (the reference's unvendored, unpinned
 * dependency, makefile:46) is absent here, so the transform semantics are
 * restated from its documented definition: unnormalised forward R2C,
 * and C2R scaled by 1/N as OP_iFFT.cpp:59-61 does.  C2R of a spectrum whose
 * kz=0 / kz=Nz/2 planes are not Hermitian is defined as C2R of their
 * Hermitian projection (F + conj F(-k))/2.
 */
#include "kf_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define OR_PI (acos(-1.))

/* ======================================================================== */
/* axes                                                                     */
/* ======================================================================== */

/* axis_SiSt.cpp:123-127 */
static double sist_S(const or_axis *a, double x)
{
    const double pi = acos(-1.);
    return x - a->m * sin(2. * pi * x / a->L);
}

double or_axis_val_at(const or_axis *a, int index)
{
    if (a->type == OR_EQST) {
        /* axis_EqSt.cpp:52-55 */
        return a->l0 + index * (a->L / a->N);
    }
    if (a->type == OR_SIST) {
        /* axis_SiSt.cpp:64-89 */
        int N = a->N, shift = 0;
        if (index < 0) {
            do { shift += 1; } while (index + shift * N < 0);
        }
        if (index >= N) {
            do { shift -= 1; } while (index + shift * N >= N - 1);
        }
        double x = a->l0 + a->L * (double)(index + shift * N) / (N);
        return sist_S(a, x) - shift * a->L;
    }
    /* OR_HYST: axis_HySt.cpp:49-80 */
    {
        double L = a->L;
        double x_val = a->l0 + L * (double)(index) / (a->N);
        int shift = 0;
        if (x_val < -L / 2.) {
            do { x_val += L; shift -= 1; } while (x_val < -L / 2.);
        }
        if (x_val >= L / 2.) {
            do { x_val -= L; shift += 1; } while (x_val >= a->l0 + L);
        }
        return shift * L + (L / 2.) * sinh(2 * (x_val) * a->m / L) / sinh(a->m);
    }
}

/* axis_EqSt.cpp:57-61, axis_SiSt.cpp:92-96; HySt uses the same
 * computational-grid wavenumber (its own k_val_at returns a coordinate,
 * axis_HySt.cpp:82-113, which the derivative cannot use). */
double or_axis_k_val_at(const or_axis *a, int index)
{
    double pi = acos(-1.);
    int N = a->N;
    double L = a->L;
    return ((index <= N / 2) ? (2. * pi * index / L) : (2. * pi * (index - N) / L));
}

/* axis_EqSt.cpp:81-84, axis_SiSt.cpp:129-134, out_sourced/axis_FoHySt.cpp:107-117 */
double or_axis_dS(const or_axis *a, int index)
{
    if (a->type == OR_EQST) return 1.;
    if (a->type == OR_SIST) {
        const double pi = acos(-1.);
        double x = a->l0 + (a->L / a->N) * index;
        return 1. - (2. * pi * a->m / a->L) * cos(2. * pi * x / a->L);
    }
    {
        double L = a->L, m = a->m;
        double x_val = a->l0 + L * (double)(index) / (double)(a->N);
        double y_val;
        if (x_val >= L / 2.) y_val = (m / sinh(m)) * cosh(2. * (x_val - L) * m / L);
        else if (x_val < -L / 2.) y_val = (m / sinh(m)) * cosh(2. * (x_val + L) * m / L);
        else y_val = (m / sinh(m)) * cosh(2. * (x_val) * m / L);
        return y_val;
    }
}

/* ======================================================================== */
/* 1D mixed-radix DFT                                                       */
/* ======================================================================== */

typedef struct {
    int n;
    cplx *w; /* w[t] = exp(-2 pi i t / n) */
} fftplan;

static fftplan *plan_create(int n)
{
    fftplan *p = (fftplan *)malloc(sizeof(fftplan));
    p->n = n;
    p->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    const double pi = acos(-1.);
    for (int t = 0; t < n; ++t) {
        double ang = -2. * pi * (double)t / (double)n;
        p->w[t] = cos(ang) + I * sin(ang);
    }
    return p;
}

static void plan_destroy(fftplan *p)
{
    if (!p) return;
    free(p->w);
    free(p);
}

static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int r = 3; r * r <= n; r += 2)
        if (n % r == 0) return r;
    return n;
}

/* Decimation-in-time recursion: out[0..n) = DFT of in[0], in[s], ... */
static void fft_rec(const cplx *in, cplx *out, int n, int istride, int sign,
                    const fftplan *p, int wstride)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int r = smallest_factor(n);
    int m = n / r;
    for (int q = 0; q < r; ++q)
        fft_rec(in + (size_t)q * istride, out + (size_t)q * m, m, istride * r, sign, p,
                wstride * r);
    const int N = p->n;
    cplx t[64];
    cplx *tt = (r <= 64) ? t : (cplx *)malloc(sizeof(cplx) * (size_t)r);
    cplx *yy = (cplx *)malloc(sizeof(cplx) * (size_t)r);
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < r; ++q) {
            long idx = ((long)q * k * wstride) % N;
            cplx w = p->w[idx];
            if (sign > 0) w = conj(w);
            tt[q] = out[k + (size_t)m * q] * w;
        }
        for (int s = 0; s < r; ++s) {
            cplx acc = 0;
            for (int q = 0; q < r; ++q) {
                long idx = ((long)q * s * (N / r)) % N;
                cplx w = p->w[idx];
                if (sign > 0) w = conj(w);
                acc += tt[q] * w;
            }
            yy[s] = acc;
        }
        for (int s = 0; s < r; ++s) out[k + (size_t)m * s] = yy[s];
    }
    if (tt != t) free(tt);
    free(yy);
}

static void fft_line(const fftplan *p, cplx *buf, cplx *tmp, int sign)
{
    fft_rec(buf, tmp, p->n, 1, sign, p, 1);
    memcpy(buf, tmp, sizeof(cplx) * (size_t)p->n);
}

void or_fft1d(double *x, int n, int sign)
{
    fftplan *p = plan_create(n);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int i = 0; i < n; ++i) buf[i] = x[2 * i] + I * x[2 * i + 1];
    fft_line(p, buf, tmp, sign);
    for (int i = 0; i < n; ++i) {
        x[2 * i] = creal(buf[i]);
        x[2 * i + 1] = cimag(buf[i]);
    }
    free(buf);
    free(tmp);
    plan_destroy(p);
}

/* ======================================================================== */
/* grid                                                                     */
/* ======================================================================== */

struct or_grid_s {
    or_axis ax[3];
    int n[3];
    int nzc;
    size_t nr, nc;
    double *k[3];   /* k_val_at per axis index (z: 0..nzc-1) */
    fftplan *pl[3];
    cplx *full;     /* nx*ny*nz complex work */
    cplx *line, *line2;
    double *dSval[3];  /* dealiased 1/S' (OP_partial_derivative.cpp:62-80), lazily */
    double *d23;       /* dealias table */
    double *ws_c;      /* complex scratch nc */
    double *ws_r;      /* real scratch nr */
};

or_grid *or_grid_create(const or_axis *ax, const or_axis *ay, const or_axis *az)
{
    or_grid *g = (or_grid *)calloc(1, sizeof(or_grid));
    g->ax[0] = *ax;
    g->ax[1] = *ay;
    g->ax[2] = *az;
    for (int d = 0; d < 3; ++d) g->n[d] = g->ax[d].N;
    g->nzc = g->n[2] / 2 + 1;
    g->nr = (size_t)g->n[0] * g->n[1] * g->n[2];
    g->nc = (size_t)g->n[0] * g->n[1] * g->nzc;
    for (int d = 0; d < 3; ++d) {
        int len = (d == 2) ? g->nzc : g->n[d];
        g->k[d] = (double *)malloc(sizeof(double) * (size_t)len);
        for (int i = 0; i < len; ++i) g->k[d][i] = or_axis_k_val_at(&g->ax[d], i);
        g->pl[d] = plan_create(g->n[d]);
    }
    g->full = (cplx *)malloc(sizeof(cplx) * g->nr);
    int mx = g->n[0];
    if (g->n[1] > mx) mx = g->n[1];
    if (g->n[2] > mx) mx = g->n[2];
    g->line = (cplx *)malloc(sizeof(cplx) * (size_t)mx);
    g->line2 = (cplx *)malloc(sizeof(cplx) * (size_t)mx);
    g->d23 = (double *)malloc(sizeof(double) * g->nc);
    or_dealias23_table(g, g->d23);
    g->ws_c = (double *)malloc(sizeof(double) * 2 * g->nc);
    g->ws_r = (double *)malloc(sizeof(double) * g->nr);
    return g;
}

void or_grid_destroy(or_grid *g)
{
    if (!g) return;
    for (int d = 0; d < 3; ++d) {
        free(g->k[d]);
        plan_destroy(g->pl[d]);
        free(g->dSval[d]);
    }
    free(g->full);
    free(g->line);
    free(g->line2);
    free(g->d23);
    free(g->ws_c);
    free(g->ws_r);
    free(g);
}

size_t or_grid_nreal(const or_grid *g) { return g->nr; }
size_t or_grid_ncplx(const or_grid *g) { return g->nc; }

/* 3D complex DFT of g->full along all axes (sign -1 forward) */
static void fft3_full(or_grid *g, int sign)
{
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const int mx = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
    cplx *a = g->full;
    /* independent lines: threads share nothing but the plans (read only) */
#pragma omp parallel
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)mx);
        cplx *line2 = (cplx *)malloc(sizeof(cplx) * (size_t)mx);
#pragma omp for collapse(2)
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                cplx *row = a + ((size_t)i * ny + j) * nz;
                fft_line(g->pl[2], row, line2, sign);
            }
#pragma omp for collapse(2)
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) {
                for (int j = 0; j < ny; ++j) line[j] = a[((size_t)i * ny + j) * nz + k];
                fft_line(g->pl[1], line, line2, sign);
                for (int j = 0; j < ny; ++j) a[((size_t)i * ny + j) * nz + k] = line[j];
            }
#pragma omp for collapse(2)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                for (int i = 0; i < nx; ++i) line[i] = a[((size_t)i * ny + j) * nz + k];
                fft_line(g->pl[0], line, line2, sign);
                for (int i = 0; i < nx; ++i) a[((size_t)i * ny + j) * nz + k] = line[i];
            }
        free(line);
        free(line2);
    }
}

/* host threads of the oracle (1 = the reference's single thread) */
#ifdef _OPENMP
#include <omp.h>
void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int or_get_threads(void) { return omp_get_max_threads(); }
#else
void or_set_threads(int n) { (void)n; }
int or_get_threads(void) { return 1; }
#endif

/* OP_FFT::operator() (OP_FFT.cpp:49-63): unnormalised R2C */
void or_fft_r2c(const or_grid *gc, const double *in, double *out)
{
    or_grid *g = (or_grid *)gc;
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2], nzc = g->nzc;
    for (size_t p = 0; p < g->nr; ++p) g->full[p] = in[p];
    fft3_full(g, -1);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nzc; ++k) {
                cplx v = g->full[((size_t)i * ny + j) * nz + k];
                size_t q = ((size_t)i * ny + j) * nzc + k;
                out[2 * q] = creal(v);
                out[2 * q + 1] = cimag(v);
            }
}

/* OP_iFFT::operator() (OP_iFFT.cpp:49-64): C2R scaled by 1/N */
void or_fft_c2r(const or_grid *gc, const double *in, double *out)
{
    or_grid *g = (or_grid *)gc;
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2], nzc = g->nzc;
    cplx *a = g->full;
#define CIN(i, j, k) (in[2 * (((size_t)(i) * ny + (j)) * nzc + (k))] + \
                      I * in[2 * (((size_t)(i) * ny + (j)) * nzc + (k)) + 1])
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nzc; ++k) {
                cplx v = CIN(i, j, k);
                int self_conj_plane = (k == 0) || (2 * k == nz);
                if (self_conj_plane) {
                    int im = (nx - i) % nx, jm = (ny - j) % ny;
                    v = 0.5 * (v + conj(CIN(im, jm, k)));
                }
                a[((size_t)i * ny + j) * nz + k] = v;
            }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = nzc; k < nz; ++k) {
                int im = (nx - i) % nx, jm = (ny - j) % ny;
                a[((size_t)i * ny + j) * nz + k] = conj(a[((size_t)im * ny + jm) * nz + (nz - k)]);
            }
#undef CIN
    fft3_full(g, +1);
    double norm = 1. / ((double)g->nr);
    for (size_t p = 0; p < g->nr; ++p) out[p] = creal(a[p]) * norm;
}

/* ======================================================================== */
/* spectral operators                                                       */
/* ======================================================================== */

/* OP_dealiasing.hpp:12-32 with the 2/3 lambda of rhs_standard.cpp:22 */
void or_dealias23_table(const or_grid *g, double *out)
{
    const int nx = g->n[0], ny = g->n[1], nzc = g->nzc;
    double kx_max = or_axis_k_val_at(&g->ax[0], g->n[0] / 2);
    double ky_max = or_axis_k_val_at(&g->ax[1], g->n[1] / 2);
    double kz_max = or_axis_k_val_at(&g->ax[2], g->n[2] / 2);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nzc; k++) {
                size_t ijk = k + (size_t)nzc * (j + (size_t)i * ny);
                double kx = or_axis_k_val_at(&g->ax[0], i);
                double ky = or_axis_k_val_at(&g->ax[1], j);
                double kz = or_axis_k_val_at(&g->ax[2], k);
                double kabs = sqrt(pow(kx / kx_max, 2.) + pow(ky / ky_max, 2.) + pow(kz / kz_max, 2.));
                out[ijk] = kabs < (2. / 3.) ? 1. : 0.;
            }
}

/* OP_dealiasing::operator() (OP_dealiasing.cpp:5-13) */
void or_dealias23(const or_grid *g, double *inout)
{
    for (size_t ijk = 0; ijk < g->nc; ijk++) {
        inout[2 * ijk] *= g->d23[ijk];
        inout[2 * ijk + 1] *= g->d23[ijk];
    }
}

/* dealiasing_undesignated (operations.hpp:20-46) with the filter of
 * rhs_standard.cpp:220.  The field's Nz is the complex extent Nz/2+1, so
 * kz_max is k_val_at((Nz/2+1)/2), as written in operations.hpp:23. */
void or_filter_phi(const or_grid *g, double *inout)
{
    const int nx = g->n[0], ny = g->n[1], nzc = g->nzc;
    double kx_max = or_axis_k_val_at(&g->ax[0], nx / 2);
    double ky_max = or_axis_k_val_at(&g->ax[1], ny / 2);
    double kz_max = or_axis_k_val_at(&g->ax[2], nzc / 2);
    for (int i = 0; i < nx; ++i) {
        double kx = or_axis_k_val_at(&g->ax[0], i);
        for (int j = 0; j < ny; ++j) {
            double ky = or_axis_k_val_at(&g->ax[1], j);
            for (int k = 0; k < nzc; ++k) {
                double kz = or_axis_k_val_at(&g->ax[2], k);
                size_t ijk = k + (size_t)nzc * (j + (size_t)i * ny);
                double kabs = sqrt(pow(kx / kx_max, 2.) + pow(ky / ky_max, 2.) + pow(kz / kz_max, 2.));
                double f = exp(-2000. * pow(kabs, 15.));
                inout[2 * ijk] *= f;
                inout[2 * ijk + 1] *= f;
            }
        }
    }
}

/* rhs_standard.cpp:51-77 */
void or_sv_table(const or_grid *g, double *out)
{
    const int nx = g->n[0], ny = g->n[1], nzc = g->nzc;
    double kx_max = or_axis_k_val_at(&g->ax[0], g->n[0] / 2);
    double ky_max = or_axis_k_val_at(&g->ax[1], g->n[1] / 2);
    double kz_max = or_axis_k_val_at(&g->ax[2], g->n[2] / 2);
    double eps = .5;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nzc; ++k) {
                size_t ijk = (k + (size_t)nzc * (j + (size_t)i * ny));
                double kx = or_axis_k_val_at(&g->ax[0], i);
                double ky = or_axis_k_val_at(&g->ax[1], j);
                double kz = or_axis_k_val_at(&g->ax[2], k);
                double knorm = sqrt(pow(kx / kx_max, 2.) + pow(ky / ky_max, 2.) + pow(kz / kz_max, 2.));
                double filter = 1. - exp(-8. * pow(sqrt(knorm), 8.));
                out[ijk] = eps * filter;
            }
}

/* rhs_standard.cpp:39-49 */
void or_barrier_table(const or_grid *g, double *out)
{
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                size_t ijk = (k + (size_t)nz * (j + (size_t)i * ny));
                double x = or_axis_val_at(&g->ax[0], i);
                double L = g->ax[0].L;
                const double pi = acos(-1);
                out[ijk] = (x < -0.5 * L / 3.) ? pow(cos(2 * (1.5) * pi / L * x), 2.) : 0.;
            }
}

/* OP_partial_derivative ctor (OP_partial_derivative.cpp:62-80): dealiased
 * 1/S' sampled on the full grid */
static const double *dS_val(or_grid *g, int dir)
{
    if (g->dSval[dir]) return g->dSval[dir];
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    double *v = (double *)malloc(sizeof(double) * g->nr);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                int idx = dir == 0 ? i : (dir == 1 ? j : k);
                v[(k + (size_t)nz * (j + (size_t)i * ny))] = 1. / or_axis_dS(&g->ax[dir], idx);
            }
    double *F = (double *)malloc(sizeof(double) * 2 * g->nc);
    or_fft_r2c(g, v, F);
    or_dealias23(g, F);
    or_fft_c2r(g, F, v);
    free(F);
    g->dSval[dir] = v;
    return v;
}

/* OP_partial_derivative::operator() (OP_partial_derivative.cpp:104-136) */
void or_deriv(const or_grid *gc, int dir, const double *in, double *out)
{
    or_grid *g = (or_grid *)gc;
    const int nx = g->n[0], ny = g->n[1], nzc = g->nzc;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nzc; ++k) {
                size_t ijk = k + (size_t)nzc * (j + (size_t)i * ny);
                double kv = g->k[dir][dir == 0 ? i : (dir == 1 ? j : k)];
                double tmp = in[2 * ijk];
                out[2 * ijk] = -kv * in[2 * ijk + 1];
                out[2 * ijk + 1] = kv * tmp;
            }
    if (g->ax[dir].type == OR_EQST) return; /* IsLinear (cpp:35-44) */
    const double *ds = dS_val(g, dir);
    or_dealias23(g, out);
    double *tmp = (double *)malloc(sizeof(double) * g->nr);
    or_fft_c2r(g, out, tmp);
    for (size_t p = 0; p < g->nr; ++p) tmp[p] = tmp[p] * ds[p];
    or_fft_r2c(g, tmp, out);
    free(tmp);
}

/* ======================================================================== */
/* Poisson: reference NLJ                                                   */
/* ======================================================================== */

static double omax(double a, double b) { return a > b ? a : b; }
static double omin(double a, double b) { return a < b ? a : b; }

typedef struct {
    int n[3];
    double *H[3];  /* H[d][i] = val_at(i) - val_at(i-1), i = 0..N (cpp:112-126) */
} spacing;

static void spacing_make(spacing *s, const or_axis *ax)
{
    for (int d = 0; d < 3; ++d) {
        s->n[d] = ax[d].N;
        s->H[d] = (double *)malloc(sizeof(double) * (size_t)(ax[d].N + 1));
        for (int i = 0; i < ax[d].N + 1; ++i)
            s->H[d][i] = or_axis_val_at(&ax[d], i) - or_axis_val_at(&ax[d], i - 1);
    }
}

static void spacing_free(spacing *s)
{
    for (int d = 0; d < 3; ++d) free(s->H[d]);
}

static inline int wrap(int i, int n)
{
    while (i < 0) i += n;
    while (i > n - 1) i -= n;
    return i;
}

#define PIDX(i, j, k) ((size_t)(k) + (size_t)nz * ((size_t)(j) + (size_t)(i) * ny))

/* newton_method::solve (newton_method.hpp:17-63) for f = A - D x - e^x */
/* the local function of solver_poisson_jacobi_nlin.cpp:190-199; with the
 * __FRONTEND__ build (theta > 0) the Boltzmann ions add exp(-theta x) to f
 * but not to its derivative (cpp:192-195) */
static double nlj_f(double x, double A, double D, double theta)
{
    double f = A - D * x - exp(x);
    if (theta != 0.) f += exp(-theta * x);
    return f;
}

static double newton_ref(double x, double A, double D, double theta)
{
    const double eps = 0.05;
    const int Nmax = 20;
    double omega = 1.;
    double gg = 0.8;
    double dx = 0.;
    double dx_alt = nlj_f(x, A, D, theta) / (-D - exp(x));
    int i = 0;
    do {
        x -= omega * dx;
        dx = nlj_f(x, A, D, theta) / (-D - exp(x));
        omega = fabs(dx) > fabs(dx_alt) ? gg * omega : (1. - gg) * omega + gg * 1.;
        dx_alt = dx;
        ++i;
    } while (fabs(dx) > eps && i < Nmax);
    return x - dx;
}

void or_nlj_init(or_nlj_state *s)
{
    /* solver_poisson_jacobi_nlin.cpp:8-44 */
    memset(s, 0, sizeof(*s));
    s->norm_sum_old = -1.;
    s->norm_max_old = -1.;
    s->limit_max = 1.45e-06;
    s->max_iterations = 50;
}

/* iteration_loop (solver_poisson_jacobi_nlin.cpp:140-265) */
static void nlj_iteration(or_nlj_state *s, const spacing *sp, const double *in, double *out,
                          const double *rho)
{
    const int nx = sp->n[0], ny = sp->n[1], nz = sp->n[2];
    const double *HX = sp->H[0], *HY = sp->H[1], *HZ = sp->H[2];
    s->norm_sum_old = s->norm_sum;
    s->norm_max_old = s->norm_max;
    s->norm_max = 0.;
    s->norm_sum = 0.;
    s->supremum = in[0];
    s->infinum = in[0];
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                double D = 2. / (HX[i + 1] * HX[i]);
                D += 2. / (HY[j + 1] * HY[j]);
                D += 2. / (HZ[k + 1] * HZ[k]);
                double A = in[PIDX(wrap(i + 1, nx), j, k)] / (HX[i + 1] * (HX[i + 1] + HX[i]));
                A += in[PIDX(wrap(i - 1, nx), j, k)] / (HX[i] * (HX[i + 1] + HX[i]));
                A += in[PIDX(i, wrap(j + 1, ny), k)] / (HY[j + 1] * (HY[j + 1] + HY[j]));
                A += in[PIDX(i, wrap(j - 1, ny), k)] / (HY[j] * (HY[j + 1] + HY[j]));
                A += in[PIDX(i, j, wrap(k + 1, nz))] / (HZ[k + 1] * (HZ[k + 1] + HZ[k]));
                A += in[PIDX(i, j, wrap(k - 1, nz))] / (HZ[k] * (HZ[k + 1] + HZ[k]));
                A = 2. * A;
                A += rho[PIDX(i, j, k)];
                double P_old = in[PIDX(i, j, k)];
                double P_new = newton_ref(P_old, A, D, s->ion_theta);
                double delta = fabs(P_new - P_old);
                out[PIDX(i, j, k)] = (1. - s->omega_SOR) * P_old + s->omega_SOR * P_new;
                s->supremum = omax(s->supremum, P_new);
                s->infinum = omin(s->infinum, P_new);
                s->norm_max = omax(s->norm_max, delta);
                s->norm_sum += delta;
            }
    s->norm_sum = s->norm_sum / ((double)nx * ny * nz);
    if ((s->norm_max < s->norm_max_old) && (s->iteration > 1)) {
        double lambda = s->norm_sum / s->norm_sum_old;
        double w = s->omega_SOR;
        s->omega_SOR = 2. / (1. + sqrt(1. - pow(lambda + w - 1., 2.) / (lambda * pow(w, 2.))));
    } else {
        s->omega_SOR = .5;
    }
    if (s->omega_SOR != s->omega_SOR) s->omega_SOR = .5;
    if (s->norm_max < s->limit_max * (s->supremum - s->infinum)) s->converged = 1;
}

static int has_nan(const double *v, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (v[i] != v[i]) return 1;
    return 0;
}

static int nlj_solve_axes(or_nlj_state *s, const or_axis *ax, double *phi, const double *rho)
{
    /* solver_poisson_jacobi_nlin::solve (cpp:50-108) */
    spacing sp;
    size_t N = (size_t)ax[0].N * ax[1].N * ax[2].N;
    s->supremum = 0.;
    s->infinum = 0.;
    s->iteration = 0;
    s->omega_SOR = 1.;
    s->invocation++;
    s->converged = 0;
    s->norm_sum_old = -1.;
    if (s->max_iterations < 0) s->max_iterations = 1000;
    if (has_nan(phi, N)) return -2;
    spacing_make(&sp, ax);
    double *phi_n = (double *)malloc(sizeof(double) * N);
    while (!s->converged && (s->iteration < s->max_iterations)) {
        s->iteration++;
        nlj_iteration(s, &sp, phi, phi_n, rho);
        memcpy(phi, phi_n, sizeof(double) * N);
    }
    free(phi_n);
    spacing_free(&sp);
    if (has_nan(phi, N)) return -2;
    return s->iteration;
}

int or_nlj_solve(or_nlj_state *s, const or_grid *g, double *phi, const double *rho)
{
    return nlj_solve_axes(s, g->ax, phi, rho);
}

/* ---- reference MG cascade ---------------------------------------------- */

static void axes_level(const or_axis *fine, int shift, or_axis *out)
{
    for (int d = 0; d < 3; ++d) {
        out[d] = fine[d];
        out[d].N = fine[d].N >> shift;
    }
}

/* OP_8htoh / OP_4htoh / OP_2htoh_lvl0 (field_interpolation.cpp:233-315):
 * injection by a factor f in every direction */
static void inject(const double *in, const int *nin, double *out, int f)
{
    int nx = nin[0] / f, ny = nin[1] / f, nz = nin[2] / f;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
                out[(size_t)k + (size_t)nz * ((size_t)j + (size_t)i * ny)] =
                    in[(size_t)(f * k) + (size_t)nin[2] * ((size_t)(f * j) + (size_t)(f * i) * nin[1])];
}

/* OP_hto2h (field_interpolation.cpp:56-147); out_ax are the fine axes */
static void prolong_hto2h(const double *in, const int *nin, double *out, const or_axis *out_ax)
{
    const int nx = 2 * nin[0], ny = 2 * nin[1], nz = 2 * nin[2];
    for (int i = 0; i < nin[0]; i++)
        for (int j = 0; j < nin[1]; j++)
            for (int k = 0; k < nin[2]; k++)
                out[PIDX(2 * i, 2 * j, 2 * k)] =
                    in[(size_t)k + (size_t)nin[2] * ((size_t)j + (size_t)i * nin[1])];
    for (int i = 0; i < nx - 1; i += 2)
        for (int j = 0; j < ny; j += 2)
            for (int k = 1; k < nz; k += 2) {
                double z1 = or_axis_val_at(&out_ax[2], k - 1);
                double z2 = or_axis_val_at(&out_ax[2], k + 1);
                double z = or_axis_val_at(&out_ax[2], k);
                double gz = (z - z1) / (z2 - z1);
                out[PIDX(i, j, k)] = (1. - gz) * out[PIDX(i, j, wrap(k - 1, nz))] +
                                     gz * out[PIDX(i, j, wrap(k + 1, nz))];
            }
    for (int i = 0; i < nx; i += 2)
        for (int j = 1; j < ny; j += 2) {
            double y1 = or_axis_val_at(&out_ax[1], j - 1);
            double y2 = or_axis_val_at(&out_ax[1], j + 1);
            double y = or_axis_val_at(&out_ax[1], j);
            double gy = (y - y1) / (y2 - y1);
            for (int k = 0; k < nz; k += 1)
                out[PIDX(i, j, k)] = (1. - gy) * out[PIDX(i, wrap(j - 1, ny), k)] +
                                     gy * out[PIDX(i, wrap(j + 1, ny), k)];
        }
    for (int i = 1; i < nx; i += 2) {
        double x1 = or_axis_val_at(&out_ax[0], i - 1);
        double x2 = or_axis_val_at(&out_ax[0], i + 1);
        double x = or_axis_val_at(&out_ax[0], i);
        double gx = (x - x1) / (x2 - x1);
        for (int j = 0; j < ny; j += 1)
            for (int k = 0; k < nz; k += 1)
                out[PIDX(i, j, k)] = (1. - gx) * out[PIDX(wrap(i - 1, nx), j, k)] +
                                     gx * out[PIDX(wrap(i + 1, nx), j, k)];
    }
}

/* solver_poisson_multigrid::solve (cpp:246-418).  lvl_change[c] is the
 * level step of cascade c: -3 (lvl_down3), -2, -1, 0 (lvl_keep), +1 (lvl_up). */
int or_mg_cascade_solve(or_nlj_state *s, const or_grid *g, double *phi, const double *rho,
                        int ncascade, const int *lvl_change, const int *steps, const double *tol)
{
    int lvl = 0;
    size_t N0 = g->nr;
    double *cur = (double *)malloc(sizeof(double) * N0);
    double *nxt = (double *)malloc(sizeof(double) * N0);
    double *rho_l = (double *)malloc(sizeof(double) * N0);
    memcpy(cur, phi, sizeof(double) * N0);
    int total = 0;
    for (int c = 0; c < ncascade; ++c) {
        or_axis ax_cur[3], ax_new[3];
        axes_level(g->ax, lvl, ax_cur);
        int ncur[3] = {ax_cur[0].N, ax_cur[1].N, ax_cur[2].N};
        int step = lvl_change[c];
        if (step < 0) {
            inject(cur, ncur, nxt, 1 << (-step));
            lvl += -step;
        } else if (step > 0) {
            axes_level(g->ax, lvl - 1, ax_new);
            prolong_hto2h(cur, ncur, nxt, ax_new);
            lvl -= 1;
        } else {
            memcpy(nxt, cur, sizeof(double) * ncur[0] * ncur[1] * ncur[2]);
        }
        axes_level(g->ax, lvl, ax_new);
        /* rho restricted from the finest grid by injection (cpp:284-359) */
        inject(rho, g->n, rho_l, 1 << lvl);
        s->max_iterations = steps[c];
        s->limit_max = tol[c];
        int it = nlj_solve_axes(s, ax_new, nxt, rho_l);
        if (it < 0) {
            free(cur); free(nxt); free(rho_l);
            return it;
        }
        total += it;
        double *t = cur; cur = nxt; nxt = t;
    }
    if (lvl != 0) {
        free(cur); free(nxt); free(rho_l);
        return -3; /* cpp:406-410 */
    }
    memcpy(phi, cur, sizeof(double) * N0);
    free(cur); free(nxt); free(rho_l);
    return total;
}

/* ======================================================================== */
/* Poisson: FAS V-cycle on the same discrete equation                      */
/* ======================================================================== */

typedef struct {
    int n[3];
    spacing sp;
    double *phi, *f, *tmp, *old;
} fas_level;

#define FAS_MAXLV 12

/* L_h phi at (i,j,k) split into off-diagonal sum A and diagonal D
 * (solver_poisson_jacobi_nlin.cpp:177-188) */
static inline void stencil(const fas_level *lv, const double *phi, int i, int j, int k,
                           double *A_out, double *D_out)
{
    const int nx = lv->n[0], ny = lv->n[1], nz = lv->n[2];
    const double *HX = lv->sp.H[0], *HY = lv->sp.H[1], *HZ = lv->sp.H[2];
    double D = 2. / (HX[i + 1] * HX[i]);
    D += 2. / (HY[j + 1] * HY[j]);
    D += 2. / (HZ[k + 1] * HZ[k]);
    double A = phi[PIDX(wrap(i + 1, nx), j, k)] / (HX[i + 1] * (HX[i + 1] + HX[i]));
    A += phi[PIDX(wrap(i - 1, nx), j, k)] / (HX[i] * (HX[i + 1] + HX[i]));
    A += phi[PIDX(i, wrap(j + 1, ny), k)] / (HY[j + 1] * (HY[j + 1] + HY[j]));
    A += phi[PIDX(i, wrap(j - 1, ny), k)] / (HY[j] * (HY[j + 1] + HY[j]));
    A += phi[PIDX(i, j, wrap(k + 1, nz))] / (HZ[k + 1] * (HZ[k + 1] + HZ[k]));
    A += phi[PIDX(i, j, wrap(k - 1, nz))] / (HZ[k] * (HZ[k + 1] + HZ[k]));
    *A_out = 2. * A;
    *D_out = D;
}

/* red-black nonlinear Gauss-Seidel: local equation A - f - D x - e^x = 0 */
static void fas_smooth(fas_level *lv, int sweeps)
{
    const int nx = lv->n[0], ny = lv->n[1], nz = lv->n[2];
    for (int s = 0; s < sweeps; ++s)
        for (int color = 0; color < 2; ++color)
#pragma omp parallel for collapse(2)
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j)
                    for (int k = (i + j + color) & 1; k < nz; k += 2) {
                        double A, D;
                        stencil(lv, lv->phi, i, j, k, &A, &D);
                        A -= lv->f[PIDX(i, j, k)];
                        double x = lv->phi[PIDX(i, j, k)];
                        for (int it = 0; it < 2; ++it) {
                            double ex = exp(x);
                            x -= (A - D * x - ex) / (-D - ex);
                        }
                        lv->phi[PIDX(i, j, k)] = x;
                    }
}

/* r = f - N(phi), N(phi) = L_h phi - e^phi */
static void fas_residual(const fas_level *lv, double *r)
{
    const int nx = lv->n[0], ny = lv->n[1], nz = lv->n[2];
#pragma omp parallel for
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                double A, D;
                double x = lv->phi[PIDX(i, j, k)];
                stencil(lv, lv->phi, i, j, k, &A, &D);
                r[PIDX(i, j, k)] = lv->f[PIDX(i, j, k)] - (A - D * x - exp(x));
            }
}

static void fas_apply_N(const fas_level *lv, const double *phi, double *out)
{
    const int nx = lv->n[0], ny = lv->n[1], nz = lv->n[2];
#pragma omp parallel for
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                double A, D;
                double x = phi[PIDX(i, j, k)];
                stencil(lv, phi, i, j, k, &A, &D);
                out[PIDX(i, j, k)] = A - D * x - exp(x);
            }
}

/* full weighting (1/4,1/2,1/4)^3 restriction, periodic */
static void restrict_fw(const double *in, const int *nf, double *out)
{
    const int nx = nf[0] / 2, ny = nf[1] / 2, nz = nf[2] / 2;
    static const double w[3] = {0.25, 0.5, 0.25};
#pragma omp parallel for
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                double acc = 0.;
                for (int a = -1; a <= 1; ++a)
                    for (int b = -1; b <= 1; ++b)
                        for (int c = -1; c <= 1; ++c) {
                            int fi = wrap(2 * i + a, nf[0]), fj = wrap(2 * j + b, nf[1]),
                                fk = wrap(2 * k + c, nf[2]);
                            acc += w[a + 1] * w[b + 1] * w[c + 1] *
                                   in[(size_t)fk + (size_t)nf[2] * ((size_t)fj + (size_t)fi * nf[1])];
                        }
                out[(size_t)k + (size_t)nz * ((size_t)j + (size_t)i * ny)] = acc;
            }
}

/* trilinear (index-space) prolongation of a coarse correction, added to fine:
 * even fine index -> coarse i/2 (weight 1); odd -> coarse (i-1)/2 and
 * (i+1)/2 (periodic), weight 1/2 each; tensor product over the axes */
static void prolong_add(const double *ec, const int *nc, double *fine)
{
    const int nx = 2 * nc[0], ny = 2 * nc[1], nz = 2 * nc[2];
#pragma omp parallel for
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                int ci[2] = {i >> 1, wrap((i >> 1) + 1, nc[0])};
                int cj[2] = {j >> 1, wrap((j >> 1) + 1, nc[1])};
                int ck[2] = {k >> 1, wrap((k >> 1) + 1, nc[2])};
                double wi[2] = {(i & 1) ? 0.5 : 1.0, (i & 1) ? 0.5 : 0.0};
                double wj[2] = {(j & 1) ? 0.5 : 1.0, (j & 1) ? 0.5 : 0.0};
                double wk[2] = {(k & 1) ? 0.5 : 1.0, (k & 1) ? 0.5 : 0.0};
                double v = 0.;
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int c = 0; c < 2; ++c) {
                            double ww = wi[a] * wj[b] * wk[c];
                            if (ww != 0.)
                                v += ww * ec[(size_t)ck[c] +
                                             (size_t)nc[2] * ((size_t)cj[b] + (size_t)ci[a] * nc[1])];
                        }
                fine[PIDX(i, j, k)] += v;
            }
}

static int fas_levels(const or_axis *ax, fas_level *lv)
{
    int L = 0;
    int n[3] = {ax[0].N, ax[1].N, ax[2].N};
    for (;;) {
        or_axis a[3];
        axes_level(ax, L, a);
        lv[L].n[0] = n[0]; lv[L].n[1] = n[1]; lv[L].n[2] = n[2];
        spacing_make(&lv[L].sp, a);
        size_t N = (size_t)n[0] * n[1] * n[2];
        lv[L].phi = (double *)malloc(sizeof(double) * N);
        lv[L].f = (double *)malloc(sizeof(double) * N);
        lv[L].tmp = (double *)malloc(sizeof(double) * N);
        lv[L].old = (double *)malloc(sizeof(double) * N);
        L++;
        int can = (n[0] % 2 == 0) && (n[1] % 2 == 0) && (n[2] % 2 == 0) &&
                  (n[0] / 2 >= 3) && (n[1] / 2 >= 3) && (n[2] / 2 >= 3) && (L < FAS_MAXLV);
        if (!can) break;
        n[0] /= 2; n[1] /= 2; n[2] /= 2;
    }
    return L;
}

static void fas_vcycle(fas_level *lv, int l, int nl)
{
    if (l == nl - 1) {
        fas_smooth(&lv[l], 20);
        return;
    }
    fas_smooth(&lv[l], 2);
    fas_residual(&lv[l], lv[l].tmp);
    fas_level *c = &lv[l + 1];
    size_t Nc = (size_t)c->n[0] * c->n[1] * c->n[2];
    restrict_fw(lv[l].tmp, lv[l].n, c->f);          /* r_c */
    inject(lv[l].phi, lv[l].n, c->phi, 2);            /* phi_c = I phi */
    memcpy(c->old, c->phi, sizeof(double) * Nc);
    fas_apply_N(c, c->phi, c->tmp);
    for (size_t p = 0; p < Nc; ++p) c->f[p] += c->tmp[p]; /* f_c = N_c(I phi) + R r */
    fas_vcycle(lv, l + 1, nl);
    for (size_t p = 0; p < Nc; ++p) c->tmp[p] = c->phi[p] - c->old[p];
    prolong_add(c->tmp, c->n, lv[l].phi);
    fas_smooth(&lv[l], 2);
}

double or_poisson_residual(const or_grid *g, const double *phi, const double *rho, double *r_out)
{
    fas_level lv;
    lv.n[0] = g->n[0]; lv.n[1] = g->n[1]; lv.n[2] = g->n[2];
    spacing_make(&lv.sp, g->ax);
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    double rmax = 0., qmax = 0.;
#pragma omp parallel for reduction(max : rmax, qmax)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                double A, D;
                double x = phi[PIDX(i, j, k)];
                stencil(&lv, phi, i, j, k, &A, &D);
                double r = A - D * x - exp(x) + rho[PIDX(i, j, k)];
                if (r_out) r_out[PIDX(i, j, k)] = r;
                rmax = omax(rmax, fabs(r));
                qmax = omax(qmax, fabs(rho[PIDX(i, j, k)]));
            }
    spacing_free(&lv.sp);
    return rmax / (qmax > 0. ? qmax : 1.);
}

int or_poisson_fas(const or_grid *g, double *phi, const double *rho, double tol, int max_cycles,
                   double *resid_out)
{
    fas_level lv[FAS_MAXLV];
    int nl = fas_levels(g->ax, lv);
    size_t N = g->nr;
    memcpy(lv[0].phi, phi, sizeof(double) * N);
    for (size_t p = 0; p < N; ++p) lv[0].f[p] = -rho[p];
    int cyc = 0;
    double res = or_poisson_residual(g, lv[0].phi, rho, NULL);
    while (res > tol && cyc < max_cycles) {
        fas_vcycle(lv, 0, nl);
        cyc++;
        res = or_poisson_residual(g, lv[0].phi, rho, NULL);
    }
    memcpy(phi, lv[0].phi, sizeof(double) * N);
    for (int l = 0; l < nl; ++l) {
        spacing_free(&lv[l].sp);
        free(lv[l].phi); free(lv[l].f); free(lv[l].tmp); free(lv[l].old);
    }
    if (resid_out) *resid_out = res;
    return res <= tol ? cyc : -1;
}

/* ======================================================================== */
/* RHS                                                                      */
/* ======================================================================== */

struct or_rhs_s {
    or_grid *g;
    or_params p;
    double *phi, *nd, *mask, *sv, *barrier;
    double *Ux, *Uy, *Uz, *B1, *B2;
    double *FBux, *FBuy, *FBuz, *FBn, *FB1, *FB2;
    double *dF[3][3]; /* dF[a][b] = d_b FU_a */
    or_nlj_state nlj;
    double last_resid;
    int last_iters;
    int status;
};

or_rhs *or_rhs_create(const or_grid *g, const or_params *p)
{
    or_rhs *r = (or_rhs *)calloc(1, sizeof(or_rhs));
    r->g = (or_grid *)g;
    r->p = *p;
    size_t nr = g->nr, nc = g->nc;
    r->phi = (double *)calloc(nr, sizeof(double));
    r->nd = (double *)calloc(nr, sizeof(double));
    r->mask = (double *)calloc(nr, sizeof(double));
    r->sv = (double *)malloc(sizeof(double) * nc);
    or_sv_table(g, r->sv);
    r->barrier = (double *)malloc(sizeof(double) * nr);
    or_barrier_table(g, r->barrier);
    r->Ux = (double *)malloc(sizeof(double) * nr);
    r->Uy = (double *)malloc(sizeof(double) * nr);
    r->Uz = (double *)malloc(sizeof(double) * nr);
    r->B1 = (double *)malloc(sizeof(double) * nr);
    r->B2 = (double *)malloc(sizeof(double) * nr);
    r->FBux = (double *)malloc(sizeof(double) * 2 * nc);
    r->FBuy = (double *)malloc(sizeof(double) * 2 * nc);
    r->FBuz = (double *)malloc(sizeof(double) * 2 * nc);
    r->FBn = (double *)malloc(sizeof(double) * 2 * nc);
    r->FB1 = (double *)malloc(sizeof(double) * 2 * nc);
    r->FB2 = (double *)malloc(sizeof(double) * 2 * nc);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) r->dF[a][b] = (double *)malloc(sizeof(double) * 2 * nc);
    or_nlj_init(&r->nlj);
    return r;
}

void or_rhs_destroy(or_rhs *r)
{
    if (!r) return;
    free(r->phi); free(r->nd); free(r->mask); free(r->sv); free(r->barrier);
    free(r->Ux); free(r->Uy); free(r->Uz); free(r->B1); free(r->B2);
    free(r->FBux); free(r->FBuy); free(r->FBuz); free(r->FBn); free(r->FB1); free(r->FB2);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) free(r->dF[a][b]);
    free(r);
}

void or_rhs_set_fields(or_rhs *r, const double *phi, const double *nd, const double *mask)
{
    size_t nr = r->g->nr;
    if (phi) memcpy(r->phi, phi, sizeof(double) * nr);
    if (nd) memcpy(r->nd, nd, sizeof(double) * nr);
    if (mask) memcpy(r->mask, mask, sizeof(double) * nr);
}

void or_rhs_get_phi(const or_rhs *r, double *phi) { memcpy(phi, r->phi, sizeof(double) * r->g->nr); }

int or_rhs_poisson_status(const or_rhs *r, double *last_resid, int *last_iters)
{
    if (last_resid) *last_resid = r->last_resid;
    if (last_iters) *last_iters = r->last_iters;
    return r->status;
}

static void cadd(double *a, const double *b, size_t nc)
{
    for (size_t i = 0; i < nc; ++i) {
        a[2 * i] += b[2 * i];
        a[2 * i + 1] += b[2 * i + 1];
    }
}

static int poisson_call(or_rhs *r, double *phi, const double *rho)
{
    if (r->p.poisson_mode == OR_POISSON_FAS) {
        double res;
        int it = or_poisson_fas(r->g, phi, rho, r->p.poisson_tol, r->p.poisson_max, &res);
        r->last_resid = res;
        r->last_iters = it;
        return it < 0 ? -1 : 0;
    }
    if (r->p.poisson_mode == OR_POISSON_NLJ) {
        r->nlj.max_iterations = r->p.poisson_max;
        r->nlj.limit_max = r->p.poisson_tol;
        int it = or_nlj_solve(&r->nlj, r->g, phi, rho);
        r->last_iters = it;
        r->last_resid = or_poisson_residual(r->g, phi, rho, NULL);
        return it < 0 ? it : 0;
    }
    /* kielflow.cpp:207-217 */
    {
        int lvl[4] = {-3, 1, 1, 1};
        int steps[4] = {0, -1, -1, -1};
        double tol[4] = {0.01, 0.0005, 0.0001, 0.0001};
        int it = or_mg_cascade_solve(&r->nlj, r->g, phi, rho, 4, lvl, steps, tol);
        r->last_iters = it;
        r->last_resid = or_poisson_residual(r->g, phi, rho, NULL);
        return it < 0 ? it : 0;
    }
}

/* rhs_standard::solve (rhs_standard.cpp:89-566); switches as
 * rhs_standard.hpp:22-30 with _RHS_SVISCOSITY / _RHS_PENALIZATION runtime */
int or_rhs_eval(or_rhs *r, double t, double *FUx, double *FUy, double *FUz, double *Fni)
{
    (void)t;
    or_grid *g = r->g;
    const size_t nc = g->nc, nr = g->nr;
    const or_params *P = &r->p;
    double *FB1 = r->FB1, *FB2 = r->FB2, *B1 = r->B1, *B2 = r->B2;
    double *FU[3] = {FUx, FUy, FUz};
    double *FBu[3] = {r->FBux, r->FBuy, r->FBuz};
    double *U[3] = {r->Ux, r->Uy, r->Uz};

    /* POISSON (cpp:100-113) */
    memcpy(FB1, Fni, sizeof(double) * 2 * nc);
    or_dealias23(g, FB1);
    or_fft_c2r(g, FB1, B1);
    for (size_t ijk = 0; ijk < nr; ++ijk) B1[ijk] = exp(B1[ijk]) + r->nd[ijk];
    r->status = poisson_call(r, r->phi, B1);

    /* dissipation & spectral viscosity (cpp:119-163) */
    for (int a = 0; a < 3; ++a)
        for (size_t ijk = 0; ijk < nc; ++ijk) {
            double f = P->sv_on ? (P->tau + r->sv[ijk]) : P->tau;
            FBu[a][2 * ijk] = f * FU[a][2 * ijk];
            FBu[a][2 * ijk + 1] = f * FU[a][2 * ijk + 1];
        }
    for (size_t ijk = 0; ijk < nc; ++ijk) {
        if (P->sv_on) {
            r->FBn[2 * ijk] = r->sv[ijk] * Fni[2 * ijk];
            r->FBn[2 * ijk + 1] = r->sv[ijk] * Fni[2 * ijk + 1];
        } else {
            r->FBn[2 * ijk] = 0.;
            r->FBn[2 * ijk + 1] = 0.;
        }
    }

    /* internal E-forces + diffusion (cpp:218-248) */
    or_fft_r2c(g, r->phi, FB1);
    or_filter_phi(g, FB1);
    for (size_t ijk = 0; ijk < nc; ijk++) {
        FB1[2 * ijk] += Fni[2 * ijk] / P->theta;
        FB1[2 * ijk + 1] += Fni[2 * ijk + 1] / P->theta;
    }
    for (int a = 0; a < 3; ++a) {
        or_deriv(g, a, FB1, FB2);
        cadd(FBu[a], FB2, nc);
    }

    /* dealias inputs (cpp:253-256) */
    for (int a = 0; a < 3; ++a) or_dealias23(g, FU[a]);
    or_dealias23(g, Fni);
    for (int a = 0; a < 3; ++a) or_fft_c2r(g, FU[a], U[a]);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) or_deriv(g, b, FU[a], r->dF[a][b]);

    /* grain-frame translation (cpp:279-295, 304-311) */
    for (int a = 0; a < 3; ++a)
        for (size_t i = 0; i < nc; ++i) {
            FBu[a][2 * i] += P->M * r->dF[a][0][2 * i];
            FBu[a][2 * i + 1] += P->M * r->dF[a][0][2 * i + 1];
        }
    or_deriv(g, 0, Fni, FB1);
    for (size_t i = 0; i < nc; ++i) {
        r->FBn[2 * i] += P->M * FB1[2 * i];
        r->FBn[2 * i + 1] += P->M * FB1[2 * i + 1];
    }

    /* continuity (cpp:315-345) */
    cadd(r->FBn, r->dF[0][0], nc);
    cadd(r->FBn, r->dF[1][1], nc);
    cadd(r->FBn, r->dF[2][2], nc);
    or_dealias23(g, FB1);
    or_fft_c2r(g, FB1, B1);
    for (size_t i = 0; i < nr; ++i) B2[i] = B1[i] * r->Ux[i];
    or_deriv(g, 1, Fni, FB1);
    or_dealias23(g, FB1);
    or_fft_c2r(g, FB1, B1);
    for (size_t i = 0; i < nr; ++i) B2[i] += B1[i] * r->Uy[i];
    or_deriv(g, 2, Fni, FB1);
    or_dealias23(g, FB1);
    or_fft_c2r(g, FB1, B1);
    for (size_t i = 0; i < nr; ++i) B2[i] += B1[i] * r->Uz[i];
    or_fft_r2c(g, B2, FB1);
    cadd(r->FBn, FB1, nc);

    /* advection (cpp:388-439) */
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) or_dealias23(g, r->dF[a][b]);
    for (int a = 0; a < 3; ++a) {
        or_fft_c2r(g, r->dF[a][0], B1);
        for (size_t i = 0; i < nr; ++i) B2[i] = r->Ux[i] * B1[i];
        or_fft_c2r(g, r->dF[a][1], B1);
        for (size_t i = 0; i < nr; ++i) B2[i] += r->Uy[i] * B1[i];
        or_fft_c2r(g, r->dF[a][2], B1);
        for (size_t i = 0; i < nr; ++i) B2[i] += r->Uz[i] * B1[i];
        or_fft_r2c(g, B2, FB1);
        cadd(FBu[a], FB1, nc);
    }

    /* penalization (cpp:459-533) */
    if (P->pen_on) {
        double eps1 = 2.;
        double *Fs[4] = {FUx, FUy, FUz, Fni};
        double *Bs[4] = {r->FBux, r->FBuy, r->FBuz, r->FBn};
        double target[4] = {P->M, 0., 0., 10.};
        for (int a = 0; a < 4; ++a) {
            or_fft_c2r(g, Fs[a], B1);
            or_fft_c2r(g, Bs[a], B2);
            for (size_t ijk = 0; ijk < nr; ijk++) {
                double H = r->mask[ijk];
                double tmp;
                if (a == 0)
                    tmp = eps1 * H * (B1[ijk] + target[a] * H);
                else
                    tmp = eps1 * H * (B1[ijk] + target[a]);
                tmp += (1. - H) * B2[ijk];
                tmp += (r->barrier[ijk]) * B1[ijk];
                B1[ijk] = tmp;
            }
            or_fft_r2c(g, B1, Bs[a]);
            or_dealias23(g, Bs[a]);
        }
    }

    /* output (cpp:538-560) */
    for (int a = 0; a < 3; ++a)
        for (size_t ijk = 0; ijk < nc; ++ijk) {
            FU[a][2 * ijk] = -FBu[a][2 * ijk];
            FU[a][2 * ijk + 1] = -FBu[a][2 * ijk + 1];
        }
    for (size_t ijk = 0; ijk < nc; ++ijk) {
        Fni[2 * ijk] = -r->FBn[2 * ijk];
        Fni[2 * ijk + 1] = -r->FBn[2 * ijk + 1];
    }
    return r->status;
}

/* ======================================================================== */
/* RK4 (RKO4.cpp:50-248) and CFL (CFL.cpp:6-82)                             */
/* ======================================================================== */

int or_rk4_step(or_rhs *r, double dt, double *t_io, double *FUx, double *FUy, double *FUz,
                double *Fn)
{
    size_t n2 = 2 * r->g->nc;
    double *y[4] = {FUx, FUy, FUz, Fn};
    double *RK[5][4];
    for (int s = 0; s < 5; ++s)
        for (int f = 0; f < 4; ++f) RK[s][f] = (double *)malloc(sizeof(double) * n2);
    int st = 0;
    double t = *t_io;
    for (int f = 0; f < 4; ++f) {
        memcpy(RK[0][f], y[f], sizeof(double) * n2);
        memcpy(RK[1][f], y[f], sizeof(double) * n2);
    }
    st |= or_rhs_eval(r, t, RK[1][0], RK[1][1], RK[1][2], RK[1][3]);
    for (int f = 0; f < 4; ++f)
        for (size_t i = 0; i < n2; ++i) RK[2][f][i] = RK[0][f][i] + 0.5 * dt * RK[1][f][i];
    st |= or_rhs_eval(r, t + 0.5 * dt, RK[2][0], RK[2][1], RK[2][2], RK[2][3]);
    for (int f = 0; f < 4; ++f)
        for (size_t i = 0; i < n2; ++i) RK[3][f][i] = RK[0][f][i] + 0.5 * dt * RK[2][f][i];
    st |= or_rhs_eval(r, t + 0.5 * dt, RK[3][0], RK[3][1], RK[3][2], RK[3][3]);
    for (int f = 0; f < 4; ++f)
        for (size_t i = 0; i < n2; ++i) RK[4][f][i] = RK[0][f][i] + 1.0 * dt * RK[3][f][i];
    st |= or_rhs_eval(r, t + 1.0 * dt, RK[4][0], RK[4][1], RK[4][2], RK[4][3]);
    for (int f = 0; f < 4; ++f)
        for (size_t i = 0; i < n2; ++i)
            y[f][i] = RK[0][f][i] + (dt / 6.) * (1. * RK[1][f][i] + 2. * RK[2][f][i] +
                                                2. * RK[3][f][i] + 1. * RK[4][f][i]);
    *t_io = t + dt;
    for (int s = 0; s < 5; ++s)
        for (int f = 0; f < 4; ++f) free(RK[s][f]);
    return st;
}

/* euler_method::solve (euler_method.cpp:13-91) */
int or_euler_step(or_rhs *r, double dt, double *t_io, double *FUx, double *FUy, double *FUz, double *Fn)
{
    size_t n2 = 2 * r->g->nc;
    double *y[4] = {FUx, FUy, FUz, Fn};
    double *k[4];
    for (int f = 0; f < 4; ++f) {
        k[f] = (double *)malloc(sizeof(double) * n2);
        memcpy(k[f], y[f], sizeof(double) * n2);
    }
    int st = or_rhs_eval(r, *t_io, k[0], k[1], k[2], k[3]);
    for (int f = 0; f < 4; ++f)
        for (size_t i = 0; i < n2; ++i) y[f][i] = y[f][i] + dt * k[f][i];
    *t_io += dt;
    for (int f = 0; f < 4; ++f) free(k[f]);
    return st;
}

static double get_delta(const or_axis *A, int i)
{
    /* CFL.cpp:6-12 (the edge branch can never be taken) */
    if ((i == A->N - 1) && (i == -1)) return or_axis_val_at(A, 1) - or_axis_val_at(A, 0);
    return or_axis_val_at(A, i + 1) - or_axis_val_at(A, i);
}

double or_cfl(const or_grid *g, double dt, const double *FUx, const double *FUy, const double *FUz)
{
    size_t nr = g->nr;
    double *U[3];
    const double *F[3] = {FUx, FUy, FUz};
    for (int a = 0; a < 3; ++a) {
        U[a] = (double *)malloc(sizeof(double) * nr);
        or_fft_c2r(g, F[a], U[a]);
    }
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    double CFLx = 0., CFLy = 0., CFLz = 0., CFLmax = 0.;
    for (int i = 0; i < nx; ++i) {
        double dx = omin(get_delta(&g->ax[0], i), get_delta(&g->ax[0], i + 1));
        for (int j = 0; j < ny; ++j) {
            double dy = omin(get_delta(&g->ax[1], j), get_delta(&g->ax[1], j + 1));
            for (int k = 0; k < nz; ++k) {
                double dz = omin(get_delta(&g->ax[2], k), get_delta(&g->ax[2], k + 1));
                size_t idx = PIDX(i, j, k);
                CFLx = omax(CFLx, dt * fabs(U[0][idx]) / dx);
                CFLy = omax(CFLy, dt * fabs(U[1][idx]) / dy);
                CFLz = omax(CFLz, dt * fabs(U[2][idx]) / dz);
                CFLmax = omax(CFLx, omax(CFLy, CFLz));
            }
        }
    }
    for (int a = 0; a < 3; ++a) free(U[a]);
    if (.4 * dt / CFLmax >= 0.01) return 0.01;
    return .4 * dt / CFLmax;
}

/* ======================================================================== */
/* set-up                                                                   */
/* ======================================================================== */

/* Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as
 * 1, 2, 3"); multipliers and Weyl constants from that paper. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* standard normals: block b -> Philox(ctr={b_lo,b_hi,0,0}, key={seed_lo,seed_hi})
 * -> two Box-Muller pairs -> normals 4b..4b+3 */
void or_normal_fill(uint64_t seed, size_t n, double *out)
{
    const double pi = acos(-1.);
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (size_t b = 0; 4 * b < n; ++b) {
        uint32_t ctr[4] = {(uint32_t)b, (uint32_t)((uint64_t)b >> 32), 0u, 0u};
        uint32_t o[4];
        or_philox4x32_10(ctr, key, o);
        for (int h = 0; h < 2; ++h) {
            double u1 = ((double)o[2 * h] + 0.5) * (1.0 / 4294967296.0);
            double u2 = ((double)o[2 * h + 1] + 0.5) * (1.0 / 4294967296.0);
            double rr = sqrt(-2. * log(u1));
            double th = 2. * pi * u2;
            size_t i0 = 4 * b + 2 * h;
            if (i0 < n) out[i0] = rr * cos(th);
            if (i0 + 1 < n) out[i0 + 1] = rr * sin(th);
        }
    }
}

/* fkt3d_Gauss (masks.hpp:351-366): A/(pi^{3/2} a^3) exp(-(x/a)^2-(y/a)^2-(z/a)^2),
 * so the amplitude is the integral over space */
void or_gauss_fill(const or_grid *g, double amplitude, double a, double *out)
{
    const double pi = acos(-1.);
    double A = (amplitude / (pow(sqrt(pi), 3.) * a * a * a));
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    for (int i = 0; i < nx; ++i) {
        double x = or_axis_val_at(&g->ax[0], i);
        for (int j = 0; j < ny; ++j) {
            double y = or_axis_val_at(&g->ax[1], j);
            for (int k = 0; k < nz; ++k) {
                double z = or_axis_val_at(&g->ax[2], k);
                out[PIDX(i, j, k)] = A * exp(-pow(x / a, 2.) - pow(y / a, 2.) - pow(z / a, 2.));
            }
        }
    }
}

/* solid mask of kielflow.cpp:228-229 with radius a: exp(-(x/a)^2-(y/a)^2-(z/a)^2) */
void or_mask_fill(const or_grid *g, double a, double *out)
{
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    for (int i = 0; i < nx; ++i) {
        double x = or_axis_val_at(&g->ax[0], i);
        for (int j = 0; j < ny; ++j) {
            double y = or_axis_val_at(&g->ax[1], j);
            for (int k = 0; k < nz; ++k) {
                double z = or_axis_val_at(&g->ax[2], k);
                out[PIDX(i, j, k)] = exp(-pow(x / a, 2.) - pow(y / a, 2.) - pow(z / a, 2.));
            }
        }
    }
}

/* kielflow.cpp:172-182: state is FFT(ln n) with n = 1 + amp*N(0,1) */
void or_init_density(const or_grid *g, uint64_t seed, double amp, double *Fn)
{
    size_t nr = g->nr;
    double *v = (double *)malloc(sizeof(double) * nr);
    or_normal_fill(seed, nr, v);
    for (size_t i = 0; i < nr; ++i) v[i] = log(1. + amp * v[i]);
    or_fft_r2c(g, v, Fn);
    free(v);
}

/* ======================================================================== */
/* Tangent-linear model of the RHS, the RK4 step and the grain drag         */
/* objective (forward-mode derivative, FP64).  It linearises each line of  */
/* or_rhs_eval above; it checks the GPU adjoint, which is its transpose.   */
/* ======================================================================== */

/* (e^phi - L_h) x = b: Jacobi-preconditioned CG on the system multiplied by
 * the cell volumes V = prod_d (H[i+1] + H[i]) / 2, which makes it symmetric
 * positive definite (V L_h is symmetric for the stencil above). */
static int tl_poisson(const or_grid *g, const double *phi, const double *b, double *x, double tol)
{
    fas_level lv;
    lv.n[0] = g->n[0]; lv.n[1] = g->n[1]; lv.n[2] = g->n[2];
    spacing_make(&lv.sp, g->ax);
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const size_t N = g->nr;
    double *V = (double *)malloc(sizeof(double) * N), *E = (double *)malloc(sizeof(double) * N);
    double *dg = (double *)malloc(sizeof(double) * N), *r = (double *)malloc(sizeof(double) * N);
    double *z = (double *)malloc(sizeof(double) * N), *p = (double *)malloc(sizeof(double) * N);
    double *Ap = (double *)malloc(sizeof(double) * N);
    double bmax = 0.;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t q = PIDX(i, j, k);
                double A, D;
                stencil(&lv, phi, i, j, k, &A, &D);
                V[q] = 0.125 * (lv.sp.H[0][i + 1] + lv.sp.H[0][i]) * (lv.sp.H[1][j + 1] + lv.sp.H[1][j]) *
                       (lv.sp.H[2][k + 1] + lv.sp.H[2][k]);
                E[q] = exp(phi[q]);
                dg[q] = V[q] * (E[q] + D);
                x[q] = 0.;
                r[q] = V[q] * b[q];
                bmax = omax(bmax, fabs(r[q]));
            }
    double rz = 0.;
    for (size_t q = 0; q < N; ++q) {
        z[q] = r[q] / dg[q];
        p[q] = z[q];
        rz += r[q] * z[q];
    }
    int it = 0, ok = (bmax == 0.);
    for (; !ok && it < 20000; ++it) {
        double pAp = 0.;
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    const size_t q = PIDX(i, j, k);
                    double A, D;
                    stencil(&lv, p, i, j, k, &A, &D);
                    Ap[q] = V[q] * (E[q] * p[q] - (A - D * p[q]));
                    pAp += p[q] * Ap[q];
                }
        const double al = rz / pAp;
        double rmax = 0., rz2 = 0.;
        for (size_t q = 0; q < N; ++q) {
            x[q] += al * p[q];
            r[q] -= al * Ap[q];
            rmax = omax(rmax, fabs(r[q]));
            z[q] = r[q] / dg[q];
            rz2 += r[q] * z[q];
        }
        if (rmax <= tol * bmax) ok = 1;
        const double be = rz2 / rz;
        rz = rz2;
        for (size_t q = 0; q < N; ++q) p[q] = z[q] + be * p[q];
    }
    spacing_free(&lv.sp);
    free(V); free(E); free(dg); free(r); free(z); free(p); free(Ap);
    return ok ? it : -1;
}

static double *tl_alloc(size_t n) { return (double *)calloc(n, sizeof(double)); }

/* out = C2R(D23 d_b X) (real) */
static void tl_dd_c2r(or_grid *g, int b, const double *X, double *tmp, double *out)
{
    or_deriv(g, b, X, tmp);
    or_dealias23(g, tmp);
    or_fft_c2r(g, tmp, out);
}

int or_rhs_eval_tl(or_rhs *r, double *F[4], double *T[4], double dM, const double *dnd)
{
    or_grid *g = r->g;
    const size_t nc = g->nc, nr = g->nr, n2 = 2 * nc;
    const or_params *P = &r->p;
    double *FB[4], *TB[4], *U[3], *dU[3], *dF[3][3], *dT[3][3];
    for (int a = 0; a < 4; ++a) { FB[a] = tl_alloc(n2); TB[a] = tl_alloc(n2); }
    for (int a = 0; a < 3; ++a) {
        U[a] = tl_alloc(nr); dU[a] = tl_alloc(nr);
        for (int b = 0; b < 3; ++b) { dF[a][b] = tl_alloc(n2); dT[a][b] = tl_alloc(n2); }
    }
    double *c1 = tl_alloc(n2), *c2 = tl_alloc(n2), *c3 = tl_alloc(n2);
    double *L = tl_alloc(nr), *dL = tl_alloc(nr), *rho = tl_alloc(nr), *drho = tl_alloc(nr), *dphi = tl_alloc(nr);
    double *B1 = tl_alloc(nr), *B2 = tl_alloc(nr), *B3 = tl_alloc(nr), *TB2 = tl_alloc(nr);

    /* Poisson and its linearisation: (e^phi - L_h) dphi = e^L dL + dnd */
    memcpy(c1, F[3], sizeof(double) * n2);
    or_dealias23(g, c1);
    or_fft_c2r(g, c1, L);
    memcpy(c1, T[3], sizeof(double) * n2);
    or_dealias23(g, c1);
    or_fft_c2r(g, c1, dL);
    for (size_t q = 0; q < nr; ++q) {
        rho[q] = exp(L[q]) + r->nd[q];
        drho[q] = exp(L[q]) * dL[q] + (dnd ? dnd[q] : 0.);
    }
    r->status = poisson_call(r, r->phi, rho);
    int st = tl_poisson(g, r->phi, drho, dphi, 1e-15);
    if (st < 0) r->status = -7;

    /* dissipation and spectral viscosity */
    for (size_t i = 0; i < nc; ++i) {
        const double f = P->sv_on ? P->tau + r->sv[i] : P->tau;
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 2; ++c) {
                FB[a][2 * i + c] = f * F[a][2 * i + c];
                TB[a][2 * i + c] = f * T[a][2 * i + c];
            }
        for (int c = 0; c < 2; ++c) {
            FB[3][2 * i + c] = P->sv_on ? r->sv[i] * F[3][2 * i + c] : 0.;
            TB[3][2 * i + c] = P->sv_on ? r->sv[i] * T[3][2 * i + c] : 0.;
        }
    }
    /* E force and pressure: grad(filter FFT(phi) + Fn / theta) */
    or_fft_r2c(g, r->phi, c1);
    or_filter_phi(g, c1);
    or_fft_r2c(g, dphi, c2);
    or_filter_phi(g, c2);
    for (size_t i = 0; i < n2; ++i) {
        c1[i] += F[3][i] / P->theta;
        c2[i] += T[3][i] / P->theta;
    }
    for (int a = 0; a < 3; ++a) {
        or_deriv(g, a, c1, c3);
        cadd(FB[a], c3, nc);
        or_deriv(g, a, c2, c3);
        cadd(TB[a], c3, nc);
    }
    /* dealiased inputs, velocities and their gradients */
    for (int a = 0; a < 4; ++a) {
        or_dealias23(g, F[a]);
        or_dealias23(g, T[a]);
    }
    for (int a = 0; a < 3; ++a) {
        or_fft_c2r(g, F[a], U[a]);
        or_fft_c2r(g, T[a], dU[a]);
        for (int b = 0; b < 3; ++b) {
            or_deriv(g, b, F[a], dF[a][b]);
            or_deriv(g, b, T[a], dT[a][b]);
        }
    }
    /* translation M d_x */
    for (int a = 0; a < 3; ++a)
        for (size_t i = 0; i < n2; ++i) {
            FB[a][i] += P->M * dF[a][0][i];
            TB[a][i] += P->M * dT[a][0][i] + dM * dF[a][0][i];
        }
    or_deriv(g, 0, F[3], c1);
    or_deriv(g, 0, T[3], c2);
    for (size_t i = 0; i < n2; ++i) {
        FB[3][i] += P->M * c1[i];
        TB[3][i] += P->M * c2[i] + dM * c1[i];
    }
    /* continuity: div u + FFT(u . grad ln n) */
    for (int a = 0; a < 3; ++a) {
        cadd(FB[3], dF[a][a], nc);
        cadd(TB[3], dT[a][a], nc);
    }
    for (size_t q = 0; q < nr; ++q) { B2[q] = 0.; TB2[q] = 0.; }
    for (int b = 0; b < 3; ++b) {
        tl_dd_c2r(g, b, F[3], c3, B1);
        tl_dd_c2r(g, b, T[3], c3, B3);
        for (size_t q = 0; q < nr; ++q) {
            B2[q] += B1[q] * U[b][q];
            TB2[q] += B3[q] * U[b][q] + B1[q] * dU[b][q];
        }
    }
    or_fft_r2c(g, B2, c1);
    cadd(FB[3], c1, nc);
    or_fft_r2c(g, TB2, c1);
    cadd(TB[3], c1, nc);
    /* advection FFT(u . grad u_a) */
    for (int a = 0; a < 3; ++a) {
        for (size_t q = 0; q < nr; ++q) { B2[q] = 0.; TB2[q] = 0.; }
        for (int b = 0; b < 3; ++b) {
            or_dealias23(g, dF[a][b]);
            or_dealias23(g, dT[a][b]);
            or_fft_c2r(g, dF[a][b], B1);
            or_fft_c2r(g, dT[a][b], B3);
            for (size_t q = 0; q < nr; ++q) {
                B2[q] += U[b][q] * B1[q];
                TB2[q] += dU[b][q] * B1[q] + U[b][q] * B3[q];
            }
        }
        or_fft_r2c(g, B2, c1);
        cadd(FB[a], c1, nc);
        or_fft_r2c(g, TB2, c1);
        cadd(TB[a], c1, nc);
    }
    /* penalization blend */
    if (P->pen_on) {
        const double eps1 = 2., target[4] = {P->M, 0., 0., 10.};
        double *TB1r = dL, *TB2r = drho; /* reuse */
        for (int a = 0; a < 4; ++a) {
            or_fft_c2r(g, F[a], B1);
            or_fft_c2r(g, FB[a], B2);
            or_fft_c2r(g, T[a], TB1r);
            or_fft_c2r(g, TB[a], TB2r);
            for (size_t q = 0; q < nr; ++q) {
                const double H = r->mask[q], bar = r->barrier[q];
                double p = (a == 0) ? eps1 * H * (B1[q] + target[a] * H) : eps1 * H * (B1[q] + target[a]);
                p += (1. - H) * B2[q] + bar * B1[q];
                double tp = eps1 * H * TB1r[q] + (1. - H) * TB2r[q] + bar * TB1r[q];
                if (a == 0) tp += eps1 * H * H * dM;
                B1[q] = p;
                B3[q] = tp;
            }
            or_fft_r2c(g, B1, FB[a]);
            or_dealias23(g, FB[a]);
            or_fft_r2c(g, B3, TB[a]);
            or_dealias23(g, TB[a]);
        }
    }
    for (int a = 0; a < 4; ++a)
        for (size_t i = 0; i < n2; ++i) {
            F[a][i] = -FB[a][i];
            T[a][i] = -TB[a][i];
        }
    for (int a = 0; a < 4; ++a) { free(FB[a]); free(TB[a]); }
    for (int a = 0; a < 3; ++a) {
        free(U[a]); free(dU[a]);
        for (int b = 0; b < 3; ++b) { free(dF[a][b]); free(dT[a][b]); }
    }
    free(c1); free(c2); free(c3); free(L); free(dL); free(rho); free(drho); free(dphi);
    free(B1); free(B2); free(B3); free(TB2);
    return r->status;
}

/* one RK4 step of the state y and of its tangent ty (RKO4.cpp:50-248) */
int or_rk4_step_tl(or_rhs *r, double dt, double *y[4], double *ty[4], double dM, const double *dnd)
{
    const size_t n2 = 2 * r->g->nc;
    double *S[4], *TS[4], *acc[4], *tacc[4];
    for (int f = 0; f < 4; ++f) {
        S[f] = tl_alloc(n2); TS[f] = tl_alloc(n2); acc[f] = tl_alloc(n2); tacc[f] = tl_alloc(n2);
        memcpy(S[f], y[f], sizeof(double) * n2);
        memcpy(TS[f], ty[f], sizeof(double) * n2);
    }
    const double c[4] = {0.5, 0.5, 1.0, 0.}, w[4] = {1., 2., 2., 1.};
    int st = 0;
    for (int s = 0; s < 4; ++s) {
        st |= or_rhs_eval_tl(r, S, TS, dM, dnd);
        for (int f = 0; f < 4; ++f)
            for (size_t i = 0; i < n2; ++i) {
                acc[f][i] += w[s] * S[f][i];
                tacc[f][i] += w[s] * TS[f][i];
                if (s < 3) {
                    S[f][i] = y[f][i] + c[s] * dt * S[f][i];
                    TS[f][i] = ty[f][i] + c[s] * dt * TS[f][i];
                }
            }
    }
    for (int f = 0; f < 4; ++f) {
        for (size_t i = 0; i < n2; ++i) {
            y[f][i] += (dt / 6.) * acc[f][i];
            ty[f][i] += (dt / 6.) * tacc[f][i];
        }
        free(S[f]); free(TS[f]); free(acc[f]); free(tacc[f]);
    }
    return st;
}

/* J = -sum n_d dV E_x with E_x = C2R(d_x FFT(phi)), phi the Poisson solution
 * of rho = exp(C2R(D23 Fn)) + n_d (warm start: the RHS's phi), and its
 * tangent dJ along (dFn, dnd).  Returns J. */
double or_objective_tl(or_rhs *r, const double *Fn, const double *Tn, const double *dV, const double *dnd,
                       double *dJ)
{
    or_grid *g = r->g;
    const size_t nc = g->nc, nr = g->nr, n2 = 2 * nc;
    double *c1 = tl_alloc(n2), *c2 = tl_alloc(n2);
    double *L = tl_alloc(nr), *dL = tl_alloc(nr), *rho = tl_alloc(nr), *drho = tl_alloc(nr);
    double *phi = tl_alloc(nr), *dphi = tl_alloc(nr), *E = tl_alloc(nr), *dE = tl_alloc(nr);
    memcpy(c1, Fn, sizeof(double) * n2);
    or_dealias23(g, c1);
    or_fft_c2r(g, c1, L);
    if (Tn) {
        memcpy(c1, Tn, sizeof(double) * n2);
        or_dealias23(g, c1);
        or_fft_c2r(g, c1, dL);
    }
    for (size_t q = 0; q < nr; ++q) {
        rho[q] = exp(L[q]) + r->nd[q];
        drho[q] = exp(L[q]) * dL[q] + (dnd ? dnd[q] : 0.);
    }
    memcpy(phi, r->phi, sizeof(double) * nr);
    double res;
    if (or_poisson_fas(g, phi, rho, r->p.poisson_tol, r->p.poisson_max, &res) < 0) r->status = -1;
    if (tl_poisson(g, phi, drho, dphi, 1e-15) < 0) r->status = -7;
    or_fft_r2c(g, phi, c1);
    or_deriv(g, 0, c1, c2);
    or_fft_c2r(g, c2, E);
    or_fft_r2c(g, dphi, c1);
    or_deriv(g, 0, c1, c2);
    or_fft_c2r(g, c2, dE);
    double J = 0., t = 0.;
    for (size_t q = 0; q < nr; ++q) {
        J -= r->nd[q] * dV[q] * E[q];
        t -= r->nd[q] * dV[q] * dE[q] + (dnd ? dnd[q] : 0.) * dV[q] * E[q];
    }
    if (dJ) *dJ = t;
    free(c1); free(c2); free(L); free(dL); free(rho); free(drho); free(phi); free(dphi); free(E); free(dE);
    return J;
}

/* create_input_from_MGsolver (frontend.cpp:133-198) and the smoothing of
 * frontend.cpp:291-310: the potential of the grain's charge n_d with
 * Boltzmann electrons and ions, solver_poisson_jacobi_nlin's __FRONTEND__
 * equation L_h phi = exp(phi) - exp(-theta phi) - n_d (cpp:192-195), by the
 * reference cascade (the level program of frontend.cpp:170-183) or, with
 * use_cascade = 0, by that NLJ run to its update test at limit `tol` with
 * at most max_it iterations; then n_i = exp(-theta phi), and n_i and phi
 * both through exp(-2000 k^15) (dealiasing_undesignated, operations.hpp:
 * 20-46).  phi: in/out (the initial guess); returns the iterations or < 0. */
int or_frontend_init(const or_grid *g, double theta, const double *nd, int use_cascade, double tol, int max_it,
                     double *phi, double *ni)
{
    or_nlj_state s;
    or_nlj_init(&s);
    s.ion_theta = theta;
    int it;
    if (use_cascade) {
        const int lvl[4] = {-3, 1, 1, 1}, steps[4] = {0, -1, -1, -1};
        const double tl[4] = {.01, 0.0005, 0.0001, 0.0001};
        it = or_mg_cascade_solve(&s, g, phi, nd, 4, lvl, steps, tl);
    } else {
        s.max_iterations = max_it;
        s.limit_max = tol;
        it = or_nlj_solve(&s, g, phi, nd);
    }
    if (it < 0) return it;
    const size_t nr = g->nr, n2 = 2 * g->nc;
    double *F = (double *)malloc(sizeof(double) * n2);
    for (size_t q = 0; q < nr; ++q) ni[q] = exp(-theta * phi[q]);
    or_fft_r2c(g, ni, F);
    or_filter_phi(g, F);
    or_fft_c2r(g, F, ni);
    or_fft_r2c(g, phi, F);
    or_filter_phi(g, F);
    or_fft_c2r(g, F, phi);
    free(F);
    return it;
}

/* the cascade's transfer operators, exported for their known-answer tests */
void or_mg_inject(const double *in, const int nin[3], double *out, int f) { inject(in, nin, out, f); }
void or_mg_prolong(const double *in, const int nin[3], double *out, const or_axis out_ax[3])
{
    prolong_hto2h(in, nin, out, out_ax);
}
