// kf_grid.hip -- host construction of the per-axis device tables.
#include "kf_grid.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace kf {

double d23_threshold()
{
    const double c = 2. / 3.;
    double t = c * c;
    while (sqrt(t) >= c) t = nextafter(t, 0.);
    while (!(sqrt(t) >= c)) t = nextafter(t, 1.);
    return t;
}

template <typename T>
static T *upload(const std::vector<T> &h, hipStream_t s)
{
    T *d = nullptr;
    KF_HIP(hipMalloc(&d, sizeof(T) * (h.empty() ? 1 : h.size())));
    // synchronous: callers pass temporaries
    if (!h.empty()) KF_HIP(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    (void)s;
    return d;
}

Grid::Grid(const AxisDesc axes[3], hipStream_t s)
{
    for (int d = 0; d < 3; ++d) {
        ax[d] = axes[d];
        if (ax[d].type != 1 && ax[d].type != 2 && ax[d].type != 3) throw Error("unknown axis type");
        if (ax[d].n < 2) throw Error("axis too short");
    }
    nx = ax[0].n;
    ny = ax[1].n;
    nz = ax[2].n;
    nzc = nz / 2 + 1;
    nr = (long)nx * ny * nz;
    nc = (long)nx * ny * nzc;
    d23_thr = d23_threshold();
    for (int d = 0; d < 3; ++d) {
        const int len = (d == 2) ? nzc : ax[d].n;
        std::vector<double> kk(len), qq(len);
        // k_max = k_val_at(N/2) (OP_dealiasing.hpp:17-19, rhs_standard.cpp:52-54)
        const double kmax = axis_k_val_at(ax[d], ax[d].n / 2);
        for (int i = 0; i < len; ++i) {
            kk[i] = axis_k_val_at(ax[d], i);
            qq[i] = pow(kk[i] / kmax, 2.);
        }
        k64[d] = upload(kk, s);
        k32[d] = upload(std::vector<float>(kk.begin(), kk.end()), s);
        q[d] = upload(qq, s);
        if (d == 2) {
            // operations.hpp:23: kz_max = k_val_at(Nz_complex/2)
            const double kzf = axis_k_val_at(ax[2], nzc / 2);
            std::vector<double> qf(len);
            for (int i = 0; i < len; ++i) qf[i] = pow(kk[i] / kzf, 2.);
            qzf = upload(qf, s);
        }
        if (ax[d].type != 1) {
            // dealiased 1/S': DFT of the 1D profile, keep the modes whose
            // 3D 2/3-rule test passes on the k_d axis line, inverse DFT
            const int n = ax[d].n;
            std::vector<long double> f(n), re(n), im(n);
            for (int i = 0; i < n; ++i) f[i] = 1.0L / (long double)axis_dS(ax[d], i);
            const long double pi = 3.141592653589793238462643383279502884L;
            for (int m = 0; m < n; ++m) {
                long double a = 0, b = 0;
                for (int i = 0; i < n; ++i) {
                    long double th = -2.0L * pi * (long double)(((long)m * i) % n) / n;
                    a += f[i] * cosl(th);
                    b += f[i] * sinl(th);
                }
                re[m] = a;
                im[m] = b;
            }
            std::vector<double> pr(n);
            for (int i = 0; i < n; ++i) {
                long double acc = 0;
                for (int m = 0; m < n; ++m) {
                    // mode m <-> k index m (negative half included); the
                    // squared ratio of the other two axes is 0 on this line
                    double qm = pow(axis_k_val_at(ax[d], m) / kmax, 2.);
                    double ssum = (d == 0) ? ((qm + 0.) + 0.) : (d == 1 ? ((0. + qm) + 0.) : ((0. + 0.) + qm));
                    if (!(ssum < d23_thr)) continue;
                    long double th = 2.0L * pi * (long double)(((long)m * i) % n) / n;
                    acc += re[m] * cosl(th) - im[m] * sinl(th);
                }
                pr[i] = (double)(acc / n);
            }
            prof64[d] = upload(pr, s);
            prof32[d] = upload(std::vector<float>(pr.begin(), pr.end()), s);
        }
    }
    {
        // 2/3-rule membership per column as a bound on |i'|, from the same
        // FP64 expression the kernels evaluated per element, ((q_x + q_y) + q_z)
        // < d23_thr; checked to be exactly a bound (q_x grows with |i'|)
        std::vector<double> qh[3];
        for (int d = 0; d < 3; ++d) {
            const int len = (d == 2) ? nzc : ax[d].n;
            const double kmax = axis_k_val_at(ax[d], ax[d].n / 2);
            qh[d].resize(len);
            for (int i = 0; i < len; ++i) qh[d][i] = pow(axis_k_val_at(ax[d], i) / kmax, 2.);
        }
        std::vector<int> im((size_t)ny * nzc);
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nzc; ++k) {
                int m = -1;
                for (int i = 0; i < nx; ++i) {
                    const int ip = i <= nx / 2 ? i : nx - i;
                    if ((qh[0][i] + qh[1][j]) + qh[2][k] < d23_thr) m = std::max(m, ip);
                }
                for (int i = 0; i < nx; ++i) {
                    const int ip = i <= nx / 2 ? i : nx - i;
                    const bool keep = (qh[0][i] + qh[1][j]) + qh[2][k] < d23_thr;
                    if (keep != (ip <= m)) throw Error("2/3-rule set of a column is not a bound on |i'|");
                }
                im[(size_t)j * nzc + k] = m;
            }
        imax = upload(im, s);
    }
    {
        std::vector<double> b(nx);
        const double L = ax[0].L;
        const double pi = acos(-1);
        for (int i = 0; i < nx; ++i) {
            double x = axis_val_at(ax[0], i);
            b[i] = (x < -0.5 * L / 3.) ? pow(cos(2 * (1.5) * pi / L * x), 2.) : 0.;
        }
        bar64 = upload(b, s);
        bar32 = upload(std::vector<float>(b.begin(), b.end()), s);
    }
    for (int d = 0; d < 3; ++d) {
        std::vector<double> v(ax[d].n);
        for (int i = 0; i < ax[d].n; ++i) v[i] = axis_dS(ax[d], i) * ax[d].L / ax[d].n;
        dv[d] = upload(v, s);
    }
    KF_HIP(hipStreamSynchronize(s));
}

double *Grid::workspace()
{
    if (!ws) {
        const size_t n = 3 * (size_t)nr + (size_t)(nx + ny + nz) + 3 * (size_t)cfl_blocks();
        KF_HIP(hipMalloc(&ws, sizeof(double) * n));
        KF_HIP(hipHostMalloc(&ws_host, sizeof(double) * 3 * cfl_blocks(), hipHostMallocDefault));
    }
    return ws;
}

Grid::~Grid()
{
    (void)hipFree(ws);
    (void)hipHostFree(ws_host);
    for (int d = 0; d < 3; ++d) {
        (void)hipFree(k64[d]);
        (void)hipFree(k32[d]);
        (void)hipFree(q[d]);
        (void)hipFree(prof64[d]);
        (void)hipFree(prof32[d]);
    }
    (void)hipFree(qzf);
    (void)hipFree(imax);
    (void)hipFree(bar64);
    (void)hipFree(bar32);
    for (int d = 0; d < 3; ++d) (void)hipFree(dv[d]);
}

}  // namespace kf
