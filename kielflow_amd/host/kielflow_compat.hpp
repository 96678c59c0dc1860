// kielflow_compat.hpp -- kielflow's class interfaces over the C ABI of
// libkielflow_amd (include/kielflow_amd.h), so that a driver written against
// the reference's headers compiles unchanged and runs its pseudo-spectral
// step on the GPU.
//
// Names, signatures and argument meaning follow the reference headers:
//   parameters            parameters.hpp:7-21
//   axis, axis_Co*St      axis.hpp:20-41, axis_EqSt.hpp, axis_SiSt.hpp:15, axis_HySt.hpp:15
//   grid                  grid_base.hpp:10-29
//   field, field_real,    field_base.hpp:10-29, field_real.hpp, field_imag.hpp
//   field_imag
//   OP_FFT, OP_iFFT,      OP_FFT.hpp:13-26, OP_iFFT.hpp:13-26, operations.hpp:48-49
//   FFT, iFFT
//   OP_partial_derivative OP_partial_derivative.cpp:10,104
//   OP_dealiasing_23rd    OP_dealiasing_23rd.hpp
//   interface_*           interface_rhs.hpp, interface_ODE_solver.hpp,
//                         interface_poisson_solver.hpp
//   solver_poisson_jacobi_nlin   solver_poisson_jacobi_nlin.hpp:48-56
//   solver_poisson_multigrid     solver_poisson_multigrid.hpp:38-44
//   rhs_standard          rhs_standard.hpp:35-37
//   Runge_kutta_O4        RKO4.hpp:29
//   cfl_all               CFL.hpp:25
//   effect_*              effect_advection.hpp:13-15 ... effect_penalization.hpp:16-18
//   fkt3d_Gauss, fkt3d_const   masks.hpp:351-366
//
// Fields keep the reference's host layout (`val`: double[N] real,
// double[N][2] complex, z fastest) and own a device mirror.  Every operator
// uploads its inputs, runs on the GPU and downloads its outputs, so host
// code may read and write `val` between calls exactly as with the
// reference.  Errors are thrown as std::runtime_error with the library's
// message (the reference throws strings).
//
// One addition: solver_poisson_newton_mg, the GPU's default Poisson solver
// (residual-based convergence), usable wherever an
// interface_relaxation_solver is.
#pragma once

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kielflow_amd.h"

#ifndef FFTW3_H
typedef double fftw_complex[2];
#endif

namespace kf_compat {
inline void check(int code)
{
    if (code != KF_OK) throw std::runtime_error(std::string("kielflow_amd: ") + kf_last_error());
}
// one context (device 0, one stream) per process: the reference is
// single-threaded and not re-entrant
inline kf_ctx context()
{
    static kf_ctx c = [] {
        kf_ctx h = nullptr;
        check(kf_ctx_create(0, &h));
        return h;
    }();
    return c;
}
struct DeviceBuffer {
    void *p = nullptr;
    size_t bytes = 0;
    explicit DeviceBuffer(size_t b) : bytes(b) { check(kf_malloc(context(), b, &p)); }
    ~DeviceBuffer() { kf_free(context(), p); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
};
}  // namespace kf_compat

// ------------------------------------------------------------------ parameters
class parameters {
  public:
    parameters(double M_, double tau_, double theta_, double mu_, double beta_)
        : pi(std::acos(-1.)), M(M_), tau(tau_), theta(theta_), mu(mu_), beta(beta_)
    {
    }
    parameters() : parameters(0., 0., 0., 0., 0.) {}
    double pi, M, tau, theta, mu, beta;
    kf_parameters c() const { return kf_parameters{M, tau, theta, mu, beta}; }
};

// ------------------------------------------------------------------ axes
typedef enum { e_x, e_y, e_z, e_r, e_phi, e_theta } direction;

class axis {
  public:
    virtual ~axis() {}
    unsigned int type_id;
    direction e_i = e_x;
    int N;
    double l0, L, m;
    virtual axis *clone() const = 0;
    virtual void resize(const int &N_) { N = N_; }
    double val_at(const int &index) const { return values(index, 0); }
    double k_val_at(const int &index) const { return values(index, 1); }
    virtual double dS(const int &index) const { return values(index, 2); }
    int index_at(const double &val) const
    {
        const kf_axis a = c();
        int i = 0;
        kf_compat::check(kf_axis_index_at(&a, val, &i));
        return i;
    }
    kf_axis c() const { return kf_axis{(int)type_id, N, l0, L, m}; }

  protected:
    axis(unsigned t, double l0_, double L_, int N_, double m_) : type_id(t), N(N_), l0(l0_), L(L_), m(m_) {}
    double values(int index, int which) const
    {
        const kf_axis a = c();
        double v[3];
        kf_compat::check(kf_axis_values(&a, index, &v[0], &v[1], &v[2]));
        return v[which];
    }
};

class axis_CoEqSt : public axis {
  public:
    axis_CoEqSt(const double &l0_, const double &L_, const int &N_) : axis(KF_AXIS_EQST, l0_, L_, N_, 0.) {}
    axis *clone() const override { return new axis_CoEqSt(*this); }
};

class axis_CoSiSt : public axis {
  public:
    axis_CoSiSt(const double &l0_, const double &L_, const int &N_, const double &m_ = 0.)
        : axis(KF_AXIS_SIST, l0_, L_, N_, m_)
    {
    }
    axis *clone() const override { return new axis_CoSiSt(*this); }
};

class axis_CoHySt : public axis {
  public:
    axis_CoHySt(const double &l0_, const double &L_, const int &N_, const int &m_ = 1)
        : axis(KF_AXIS_HYST, l0_, L_, N_, (double)m_)
    {
    }
    axis *clone() const override { return new axis_CoHySt(*this); }
};

// ------------------------------------------------------------------ grid
class grid {
  public:
    axis *x_axis, *y_axis, *z_axis;
    grid(const axis &Ax, const axis &Ay, const axis &Az)
        : x_axis(Ax.clone()), y_axis(Ay.clone()), z_axis(Az.clone()), Nx(x_axis->N), Ny(y_axis->N), Nz(z_axis->N)
    {
    }
    // copies share one device grid (created here if it does not exist yet)
    grid(const grid &that) : grid(*that.x_axis, *that.y_axis, *that.z_axis)
    {
        that.handle();
        h_ = that.h_;
    }
    ~grid()
    {
        delete x_axis;
        delete y_axis;
        delete z_axis;
    }
    grid &operator=(const grid &) = delete;
    const int &Nx;
    const int &Ny;
    const int &Nz;
    // the device grid (created on first use, shared by copies)
    kf_grid handle() const
    {
        if (!h_) {
            const kf_axis a[3] = {x_axis->c(), y_axis->c(), z_axis->c()};
            kf_grid g = nullptr;
            kf_compat::check(kf_grid_create(kf_compat::context(), a, &g));
            h_ = std::shared_ptr<kf_grid_s>(g, [](kf_grid p) { kf_grid_destroy(p); });
        }
        return h_.get();
    }

  private:
    mutable std::shared_ptr<kf_grid_s> h_;
};

// ------------------------------------------------------------------ fields
class field {
  public:
    int Nx, Ny, Nz, N;
    grid my_grid;
    field(const grid &domain, int nx, int ny, int nz) : Nx(nx), Ny(ny), Nz(nz), N(nx * ny * nz), my_grid(domain) {}
    inline int index(const int &i, const int &j, const int &k) const { return k + Nz * (j + i * Ny); }
};

class field_real : public field {
  public:
    explicit field_real(const grid &Omega)
        : field(Omega, Omega.Nx, Omega.Ny, Omega.Nz), buf_((size_t)N), val(buf_.data())
    {
    }
    field_real(const field_real &that) : field(that.my_grid, that.Nx, that.Ny, that.Nz), buf_(that.buf_), val(buf_.data())
    {
    }
    field_real &operator=(const field_real &rhs)
    {
        std::memcpy(val, rhs.val, sizeof(double) * (size_t)N);
        return *this;
    }
    double &operator()(const int &i, const int &j, const int &k) { return val[index(i, j, k)]; }
    double const &operator()(const int &i, const int &j, const int &k) const { return val[index(i, j, k)]; }
    double val_at(int ix, int iy, int iz) const
    {
        ix = ((ix % Nx) + Nx) % Nx;
        iy = ((iy % Ny) + Ny) % Ny;
        iz = ((iz % Nz) + Nz) % Nz;
        return val[index(ix, iy, iz)];
    }
    template <typename Func>
    void fill2(Func fkt)
    {
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < Nz; ++k)
                    val[index(i, j, k)] = fkt(my_grid.x_axis->val_at(i), my_grid.y_axis->val_at(j),
                                              my_grid.z_axis->val_at(k));
    }
    template <class F3>
    void fill(const F3 &f)
    {
        fill2([&](double x, double y, double z) { return f(x, y, z); });
    }
    // device mirror
    void *dev() const
    {
        if (!d_) d_ = std::make_shared<kf_compat::DeviceBuffer>(sizeof(double) * (size_t)N);
        return d_->p;
    }
    void *upload() const
    {
        kf_compat::check(kf_memcpy_h2d(kf_compat::context(), dev(), val, sizeof(double) * (size_t)N));
        return dev();
    }
    void download()
    {
        kf_compat::check(kf_memcpy_d2h(kf_compat::context(), val, dev(), sizeof(double) * (size_t)N));
    }

  private:
    std::vector<double> buf_;
    mutable std::shared_ptr<kf_compat::DeviceBuffer> d_;

  public:
    double *val;
};

class field_imag : public field {
  public:
    explicit field_imag(const grid &Omega)
        : field(Omega, Omega.Nx, Omega.Ny, Omega.Nz / 2 + 1), buf_(2 * (size_t)N),
          val(reinterpret_cast<fftw_complex *>(buf_.data()))
    {
    }
    field_imag(const field_imag &that)
        : field(that.my_grid, that.Nx, that.Ny, that.Nz), buf_(that.buf_),
          val(reinterpret_cast<fftw_complex *>(buf_.data()))
    {
    }
    field_imag &operator=(const field_imag &rhs)
    {
        std::memcpy(val, rhs.val, sizeof(fftw_complex) * (size_t)N);
        return *this;
    }
    field_imag &operator+=(const field_imag &rhs)
    {
        for (int i = 0; i < N; ++i) {
            val[i][0] += rhs.val[i][0];
            val[i][1] += rhs.val[i][1];
        }
        return *this;
    }
    fftw_complex &operator()(int ix, int iy, int iz) { return val[index(ix, iy, iz)]; }
    fftw_complex &operator()(const int &n) { return val[n]; }
    double &operator()(int ix, int iy, int iz, int re_im) { return val[index(ix, iy, iz)][re_im]; }
    double &operator()(const int &n, int re_im) { return val[n][re_im]; }
    void *dev() const
    {
        if (!d_) d_ = std::make_shared<kf_compat::DeviceBuffer>(sizeof(fftw_complex) * (size_t)N);
        return d_->p;
    }
    void *upload() const
    {
        kf_compat::check(kf_memcpy_h2d(kf_compat::context(), dev(), val, sizeof(fftw_complex) * (size_t)N));
        return dev();
    }
    void download()
    {
        kf_compat::check(kf_memcpy_d2h(kf_compat::context(), val, dev(), sizeof(fftw_complex) * (size_t)N));
    }

  private:
    std::vector<double> buf_;
    mutable std::shared_ptr<kf_compat::DeviceBuffer> d_;

  public:
    fftw_complex *val;
};

// ------------------------------------------------------------------ 3D functions (masks.hpp)
class interface_3d_fkt {
  public:
    virtual ~interface_3d_fkt() {}
    virtual double operator()(const double &x, const double &y, const double &z) const = 0;
};

class fkt3d_const : public interface_3d_fkt {
  public:
    explicit fkt3d_const(const double &c) : c_(c) {}
    double operator()(const double &, const double &, const double &) const override { return c_; }

  private:
    double c_;
};

// A / (pi^(3/2) a_x a_y a_z) exp(-(x/a_x)^2 - (y/a_y)^2 - (z/a_z)^2)  (masks.hpp:351-366)
class fkt3d_Gauss : public interface_3d_fkt {
  public:
    fkt3d_Gauss(const double &A, const double &ax, const double &ay, const double &az)
        : A_(A), ax_(ax), ay_(ay), az_(az)
    {
    }
    double operator()(const double &x, const double &y, const double &z) const override
    {
        const double pi = std::acos(-1.);
        return A_ / (std::pow(pi, 1.5) * ax_ * ay_ * az_) *
               std::exp(-(x / ax_) * (x / ax_) - (y / ay_) * (y / ay_) - (z / az_) * (z / az_));
    }

  private:
    double A_, ax_, ay_, az_;
};

// ------------------------------------------------------------------ operators
class OP_FFT {
  public:
    explicit OP_FFT(const grid &domain) : g_(domain)
    {
        kf_compat::check(kf_fft_create(g_.handle(), KF_FP64, &f_));
    }
    ~OP_FFT() { kf_fft_destroy(f_); }
    OP_FFT(const OP_FFT &) = delete;
    void operator()(const field_real &in, field_imag &out)
    {
        kf_compat::check(kf_fft_r2c(f_, in.upload(), out.dev()));
        out.download();
    }

  private:
    grid g_;
    kf_fft f_ = nullptr;
};

class OP_iFFT {
  public:
    explicit OP_iFFT(const grid &domain) : g_(domain)
    {
        kf_compat::check(kf_fft_create(g_.handle(), KF_FP64, &f_));
    }
    ~OP_iFFT() { kf_fft_destroy(f_); }
    OP_iFFT(const OP_iFFT &) = delete;
    void operator()(const field_imag &in, field_real &out)
    {
        kf_compat::check(kf_fft_c2r(f_, in.upload(), out.dev()));
        out.download();
    }

  private:
    grid g_;
    kf_fft f_ = nullptr;
};

inline void FFT(const field_real &in, field_imag &out)
{
    OP_FFT op(in.my_grid);
    op(in, out);
}

inline void iFFT(const field_imag &in, field_real &out)
{
    OP_iFFT op(in.my_grid);
    op(in, out);
}

class OP_partial_derivative {
  public:
    OP_partial_derivative(const grid &domain, const direction &e_i) : g_(domain), dir_((int)e_i)
    {
        if (dir_ > 2) throw std::runtime_error("OP_partial_derivative: cartesian directions only");
    }
    OP_partial_derivative(const field_real &f, const direction &e_i) : OP_partial_derivative(f.my_grid, e_i) {}
    OP_partial_derivative(const field_imag &f, const direction &e_i) : OP_partial_derivative(f.my_grid, e_i) {}
    void operator()(const field_imag &in, field_imag &out)
    {
        kf_compat::check(kf_partial_derivative(g_.handle(), dir_, (const double *)in.upload(), (double *)out.dev()));
        out.download();
    }

  private:
    grid g_;
    int dir_;
};

class OP_dealiasing_23rd {
  public:
    explicit OP_dealiasing_23rd(const grid &domain) : g_(domain) {}
    void operator()(field_imag &target) const
    {
        kf_compat::check(kf_dealias23(g_.handle(), (double *)target.upload()));
        target.download();
    }

  private:
    grid g_;
};

// dealiasing_undesignated(F, exp(-2000 k^15)) of rhs_standard.cpp:220
inline void dealiasing_phi_filter(field_imag &F)
{
    kf_compat::check(kf_filter_phi(F.my_grid.handle(), (double *)F.upload()));
    F.download();
}

inline double cfl_all(const double &dt, const field_imag &FUx, const field_imag &FUy, const field_imag &FUz)
{
    double out = 0.;
    kf_compat::check(kf_cfl_all(FUx.my_grid.handle(), dt, (const double *)FUx.upload(), (const double *)FUy.upload(),
                                (const double *)FUz.upload(), &out));
    return out;
}

// ------------------------------------------------------------------ effects
class effect {};

class effect_advection : public effect {
  public:
    effect_advection(OP_partial_derivative &, OP_partial_derivative &, OP_partial_derivative &) {}
    void execute(const field_imag &FUx, const field_imag &FUy, const field_imag &FUz, const field_imag FXX,
                 field_imag &return_Buffer)
    {
        kf_compat::check(kf_effect_advection(FUx.my_grid.handle(), (const double *)FUx.upload(),
                                             (const double *)FUy.upload(), (const double *)FUz.upload(),
                                             (const double *)FXX.upload(), (double *)return_Buffer.upload()));
        return_Buffer.download();
    }
};

class effect_continuityEQ : public effect {
  public:
    effect_continuityEQ() {}
    void execute(field_imag &FUx, field_imag &FUy, field_imag &FUz, field_imag &Fni, field_imag &Buffer_Fni)
    {
        kf_compat::check(kf_effect_continuity(FUx.my_grid.handle(), (const double *)FUx.upload(),
                                              (const double *)FUy.upload(), (const double *)FUz.upload(),
                                              (const double *)Fni.upload(), (double *)Buffer_Fni.dev()));
        Buffer_Fni.download();
    }
};

class effect_force_E : public effect {
  public:
    effect_force_E() {}
    void execute(const field_imag &FPhi, field_imag &Bx, field_imag &By, field_imag &Bz)
    {
        kf_compat::check(kf_effect_force_E(FPhi.my_grid.handle(), (const double *)FPhi.upload(),
                                           (double *)Bx.upload(), (double *)By.upload(), (double *)Bz.upload()));
        Bx.download();
        By.download();
        Bz.download();
    }
};

class effect_diffusion_u : public effect {
  public:
    explicit effect_diffusion_u(double temperature_ratio) : theta(temperature_ratio) {}
    void execute(const field_imag &Fni, field_imag &Bx, field_imag &By, field_imag &Bz)
    {
        kf_compat::check(kf_effect_diffusion_u(Fni.my_grid.handle(), theta, (const double *)Fni.upload(),
                                               (double *)Bx.upload(), (double *)By.upload(), (double *)Bz.upload()));
        Bx.download();
        By.download();
        Bz.download();
    }

  private:
    const double theta;
};

class effect_spectral_viscosity : public effect {
  public:
    effect_spectral_viscosity(const double &, const grid &) {}
    void execute(const field_imag &in, field_imag &out) const
    {
        kf_compat::check(kf_effect_spectral_viscosity(in.my_grid.handle(), (const double *)in.upload(),
                                                      (double *)out.upload()));
        out.download();
    }
};

class effect_translation : public effect {
  public:
    effect_translation(const double &v, OP_partial_derivative &) : my_velocity(v) {}
    void execute(const field_imag &in, field_imag &out)
    {
        kf_compat::check(kf_effect_translation(in.my_grid.handle(), my_velocity, (const double *)in.upload(),
                                               (double *)out.upload()));
        out.download();
    }

  private:
    double my_velocity;
};

class effect_force_linear : public effect {
  public:
    explicit effect_force_linear(const double &k) : hook_constant(k) {}
    void execute(const field_imag &in, field_imag &out)
    {
        kf_compat::check(kf_effect_force_linear(in.my_grid.handle(), hook_constant, (const double *)in.upload(),
                                                (double *)out.upload()));
        out.download();
    }

  private:
    double hook_constant;
};

class effect_penalization : public effect {
  public:
    effect_penalization(const double &eta, field_real &H, const double &boundary_value, bool scale_by_mask = false)
        : damping(eta), fix_value(boundary_value), scaled(scale_by_mask), mask(H)
    {
        if (damping > 0) throw std::runtime_error("error in penalization method due to damping");
    }
    void execute(const field_imag &in, field_imag &out) const
    {
        kf_compat::check(kf_effect_penalization(in.my_grid.handle(), (const double *)mask.upload(), damping,
                                                fix_value, scaled ? 1 : 0, (const double *)in.upload(),
                                                (double *)out.upload()));
        out.download();
    }

  private:
    double damping, fix_value;
    bool scaled;
    field_real &mask;
};

// ------------------------------------------------------------------ interfaces
class interface_poisson_solver {
  public:
    virtual ~interface_poisson_solver() {}
    virtual void solve(field_real &Phi_IO, field_real &rho) = 0;

  protected:
    interface_poisson_solver() {}
};

class interface_relaxation_solver : public interface_poisson_solver {
  public:
    virtual ~interface_relaxation_solver() {}
    virtual void set_max_iterations(const int &iter) = 0;
    virtual void set_tolerance(const double &tolerance) = 0;
    virtual bool converged(void) const = 0;
};

class interface_rhs {
  public:
    virtual ~interface_rhs() {}
    virtual void solve(const double &t, field_imag &FUx, field_imag &FUy, field_imag &FUz, field_imag &Fni) = 0;

  protected:
    interface_rhs() {}
};

class interface_ode_solver {
  public:
    virtual ~interface_ode_solver() {}
    virtual void solve(field_imag &FUx, field_imag &FUy, field_imag &FUz, field_imag &Fni) = 0;

  protected:
    interface_ode_solver() {}
};

// ------------------------------------------------------------------ Poisson solvers
// Common part: a device solver of a given mode, created for the grid of the
// first field it sees.
class kf_poisson_adaptor : public interface_relaxation_solver {
  public:
    ~kf_poisson_adaptor() override
    {
        if (p_) kf_poisson_destroy(p_);
    }
    void solve(field_real &Phi_IO, field_real &rho) override
    {
        kf_poisson p = device_solver(Phi_IO.my_grid);
        int it = 0;
        double res = 0.;
        const int code = kf_poisson_solve(p, (double *)Phi_IO.upload(), (const double *)rho.upload(), &it, &res);
        converged_ = (code == KF_OK);
        if (code != KF_OK && code != KF_ERR_NOT_CONVERGED) kf_compat::check(code);
        Phi_IO.download();
    }
    void set_max_iterations(const int &iter) override
    {
        max_it_ = iter;
        if (p_) kf_compat::check(kf_poisson_set_max_iterations(p_, iter));
    }
    void set_tolerance(const double &tolerance) override
    {
        tol_ = tolerance;
        if (p_) kf_compat::check(kf_poisson_set_tolerance(p_, tolerance));
    }
    bool converged(void) const override { return converged_; }
    // the device solver (used by rhs_standard)
    kf_poisson device_solver(const grid &g)
    {
        if (!p_) {
            g_.reset(new grid(g));
            kf_compat::check(kf_poisson_create(g_->handle(), KF_FP64, mode_, &p_));
            kf_compat::check(kf_poisson_set_tolerance(p_, tol_));
            kf_compat::check(kf_poisson_set_max_iterations(p_, max_it_));
            on_create();
        }
        return p_;
    }

  protected:
    explicit kf_poisson_adaptor(int mode, double tol, int max_it) : mode_(mode), tol_(tol), max_it_(max_it) {}
    // settings held until the device solver exists
    virtual void on_create() {}
    int mode_;
    double tol_;
    int max_it_;
    bool converged_ = false;
    kf_poisson p_ = nullptr;
    std::unique_ptr<grid> g_;
};

// the reference's nonlinear Jacobi (pure Jacobi, Carre omega, update stop)
class solver_poisson_jacobi_nlin : public kf_poisson_adaptor {
  public:
    solver_poisson_jacobi_nlin(interface_3d_fkt &, interface_3d_fkt &, const double & = 1.)
        : kf_poisson_adaptor(KF_POISSON_NLJ, 1e-10, 100)
    {
    }
    void set_limit_NormMax(const double &val) { set_tolerance(val); }
    void set_limit_NormSum(const double &) {}
};

typedef enum {
    lvl_down, lvl_keep, lvl_up, lvl_down2, lvl_down3, lvl_down_x, lvl_up_x, lvl_down_y, lvl_up_y, lvl_down_z,
    lvl_up_z
} MG_lvl_control;

// the reference's coarse-to-fine cascade around a relaxation method
class solver_poisson_multigrid : public kf_poisson_adaptor {
  public:
    explicit solver_poisson_multigrid(interface_relaxation_solver &method)
        : kf_poisson_adaptor(KF_POISSON_CASCADE, 1e-4, 80), method_(&method)
    {
    }
    void set_method(interface_relaxation_solver &method) { method_ = &method; }
    // the cascade's level program, steps and tolerances
    // (solver_poisson_multigrid.cpp:82-147); the isotropic level controls
    // are implemented, the per-axis ones (lvl_down_x ... lvl_up_z) throw
    void set_level_control(const int &N, int *lvl_steps, MG_lvl_control *lvl_C, double *lvl_tol)
    {
        change_.assign(N, 0);
        steps_.assign(lvl_steps, lvl_steps + N);
        tol_lv_.assign(lvl_tol, lvl_tol + N);
        for (int i = 0; i < N; ++i) {
            switch (lvl_C[i]) {
            case lvl_keep: change_[i] = 0; break;
            case lvl_up: change_[i] = 1; break;
            case lvl_down: change_[i] = -1; break;
            case lvl_down2: change_[i] = -2; break;
            case lvl_down3: change_[i] = -3; break;
            default:
                throw std::runtime_error("solver_poisson_multigrid: per-axis level controls are not implemented");
            }
        }
        if (p_) apply_levels();
    }

  protected:
    void on_create() override
    {
        if (!change_.empty()) apply_levels();
    }

  private:
    void apply_levels()
    {
        kf_compat::check(kf_poisson_set_cascade(p_, (int)change_.size(), change_.data(), steps_.data(),
                                                tol_lv_.data()));
    }
    interface_relaxation_solver *method_;
    std::vector<int> change_, steps_;
    std::vector<double> tol_lv_;
};

// Newton-multigrid with a residual-based stop (the GPU default)
class solver_poisson_newton_mg : public kf_poisson_adaptor {
  public:
    explicit solver_poisson_newton_mg(double tolerance = 1e-10, int max_iterations = 100)
        : kf_poisson_adaptor(KF_POISSON_NEWTON_MG, tolerance, max_iterations)
    {
    }
};

// ------------------------------------------------------------------ RHS and RK4
class rhs_standard : public interface_rhs {
  public:
    // the term switches as rhs_standard.hpp:22-30 ships them: _RHS_SVISCOSITY
    // defined, _RHS_PENALIZATION commented out
    static constexpr bool default_spectral_viscosity = true;
    static constexpr bool default_penalization = false;
    rhs_standard(parameters &Params, interface_poisson_solver &Solver, field_real &potential,
                 field_real &density_dust, field_real &solid_mask, const grid &domain,
                 bool spectral_viscosity = default_spectral_viscosity, bool penalization = default_penalization)
        : my_params(Params), Phi(potential), nd(density_dust), my_solids(solid_mask), g_(domain)
    {
        auto *ad = dynamic_cast<kf_poisson_adaptor *>(&Solver);
        if (!ad) throw std::runtime_error("rhs_standard: the Poisson solver must be one of kielflow_amd's");
        const kf_parameters pa = Params.c();
        const kf_rhs_switches sw{spectral_viscosity ? 1 : 0, penalization ? 1 : 0};
        kf_compat::check(kf_rhs_create(g_.handle(), &pa, &sw, ad->device_solver(domain), KF_FP64, &r_));
        kf_compat::check(kf_rhs_set_fields(r_, (double *)Phi.upload(), (const double *)nd.upload(),
                                           (const double *)my_solids.upload()));
    }
    ~rhs_standard() override { kf_rhs_destroy(r_); }
    void solve(const double &t, field_imag &FUx, field_imag &FUy, field_imag &FUz, field_imag &Fni) override
    {
        sync_parameters();
        kf_compat::check(kf_rhs_solve(r_, t, (double *)FUx.upload(), (double *)FUy.upload(), (double *)FUz.upload(),
                                      (double *)Fni.upload()));
        FUx.download();
        FUy.download();
        FUz.download();
        Fni.download();
        Phi.download();
    }
    kf_rhs handle() { return r_; }
    void sync_parameters()
    {
        const kf_parameters pa = my_params.c();
        kf_compat::check(kf_rhs_set_parameters(r_, &pa));
    }
    field_real &potential() { return Phi; }

  protected:
    parameters &my_params;
    field_real &Phi, &nd, &my_solids;
    grid g_;
    kf_rhs r_ = nullptr;
};

// euler_method (euler_method.hpp:27-39): y = y + dt k per step.  With
// kielflow_amd's rhs_standard the step runs in the library (kf_euler_*);
// any other interface_rhs is evaluated through its solve() and the update
// runs on the device (kf_rk4_stage_input with c = dt).
class euler_method : public interface_ode_solver {
  public:
    explicit euler_method(interface_rhs &rhs, double dt = 0.1) : dt_(dt), my_rhs(rhs)
    {
        if (auto *r = dynamic_cast<rhs_standard *>(&rhs)) kf_compat::check(kf_euler_create(r->handle(), dt, &e_));
    }
    ~euler_method() override
    {
        if (e_) kf_euler_destroy(e_);
    }
    void solve(field_imag &FUx, field_imag &FUy, field_imag &FUz, field_imag &Fni) override
    {
        field_imag *y[4] = {&FUx, &FUy, &FUz, &Fni};
        if (e_) {
            auto &r = dynamic_cast<rhs_standard &>(my_rhs);
            r.sync_parameters();
            kf_compat::check(kf_euler_steps(e_, 1, (double *)FUx.upload(), (double *)FUy.upload(),
                                            (double *)FUz.upload(), (double *)Fni.upload()));
            for (field_imag *f : y) f->download();
            r.potential().download();
        } else {
            std::vector<field_imag> k;
            k.reserve(4);
            for (int f = 0; f < 4; ++f) k.emplace_back(*y[f]);
            my_rhs.solve(t_, k[0], k[1], k[2], k[3]);
            for (int f = 0; f < 4; ++f) {
                // y <- y + dt k, into y's device mirror (a separate buffer)
                field_imag y0(*y[f]);
                kf_compat::check(kf_rk4_stage_input(kf_compat::context(), 2 * (size_t)y[f]->N, dt_,
                                                    (const double *)y0.upload(), (const double *)k[f].upload(),
                                                    (double *)y[f]->dev()));
                y[f]->download();
            }
        }
        t_ += dt_;
    }

  private:
    double dt_;
    double t_ = 0.;
    interface_rhs &my_rhs;
    kf_euler e_ = nullptr;
};

class Runge_kutta_O4 : public interface_ode_solver {
  public:
    // With kielflow_amd's rhs_standard the whole step runs in the library
    // (kf_rk4_*).  Any other interface_rhs is evaluated four times per step
    // through its own solve(), as RKO4.cpp:105-212 does, and the stage
    // inputs and the update run on the device (kf_rk4_stage_input,
    // kf_rk4_combine) with the reference's arithmetic.
    explicit Runge_kutta_O4(interface_rhs &rhs, double dt = 0.01) : dt_(dt), my_rhs(rhs)
    {
        if (auto *r = dynamic_cast<rhs_standard *>(&rhs)) kf_compat::check(kf_rk4_create(r->handle(), dt, &k_));
    }
    ~Runge_kutta_O4() override
    {
        if (k_) kf_rk4_destroy(k_);
    }
    void solve(field_imag &FUx, field_imag &FUy, field_imag &FUz, field_imag &Fni) override
    {
        if (!k_) {
            solve_any(FUx, FUy, FUz, Fni);
            return;
        }
        auto &r = dynamic_cast<rhs_standard &>(my_rhs);
        r.sync_parameters();
        kf_compat::check(kf_rk4_solve(k_, (double *)FUx.upload(), (double *)FUy.upload(), (double *)FUz.upload(),
                                      (double *)Fni.upload()));
        FUx.download();
        FUy.download();
        FUz.download();
        Fni.download();
        r.potential().download();
        t_ += dt_;
    }

  private:
    // RKO4.cpp:50-248 around an arbitrary interface_rhs
    void solve_any(field_imag &FUx, field_imag &FUy, field_imag &FUz, field_imag &Fni)
    {
        field_imag *y[4] = {&FUx, &FUy, &FUz, &Fni};
        const size_t n = 2 * (size_t)FUx.N;
        const kf_ctx ctx = kf_compat::context();
        // RK0 (the step's start) and the four coefficients, host copies
        // (reserved: a copy of a field does not carry its device mirror)
        std::vector<field_imag> y0, k[4];
        y0.reserve(4);
        for (int f = 0; f < 4; ++f) y0.emplace_back(*y[f]);
        for (int f = 0; f < 4; ++f) y0[f].upload();
        for (int s = 0; s < 4; ++s) {
            k[s].reserve(4);
            for (int f = 0; f < 4; ++f) k[s].emplace_back(*y[f]);
        }
        const double c[3] = {0.5 * dt_, 0.5 * dt_, 1.0 * dt_};
        const double tt[4] = {t_, t_ + 0.5 * dt_, t_ + 0.5 * dt_, t_ + 1.0 * dt_};
        for (int s = 0; s < 4; ++s) {
            if (s > 0) {
                // stage input RK_{s+1} = RK0 + c_s RK_s (the coefficient just evaluated)
                for (int f = 0; f < 4; ++f) {
                    kf_compat::check(kf_rk4_stage_input(ctx, n, c[s - 1], (const double *)y0[f].dev(),
                                                        (const double *)k[s - 1][f].upload(),
                                                        (double *)k[s][f].dev()));
                    k[s][f].download();
                }
            }
            my_rhs.solve(tt[s], k[s][0], k[s][1], k[s][2], k[s][3]);
        }
        for (int f = 0; f < 4; ++f) {
            kf_compat::check(kf_rk4_combine(ctx, n, dt_, (const double *)y0[f].dev(),
                                            (const double *)k[0][f].upload(), (const double *)k[1][f].upload(),
                                            (const double *)k[2][f].upload(), (const double *)k[3][f].upload(),
                                            (double *)y[f]->dev()));
            y[f]->download();
        }
        t_ += dt_;
    }

    double dt_;
    double t_ = 0.;
    interface_rhs &my_rhs;
    kf_rk4 k_ = nullptr;
};
