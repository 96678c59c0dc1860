// kf_driver.cpp -- a kielflow-style driver (kielflow.cpp:172-331) written
// against the reference's class interfaces (kielflow_compat.hpp): grid and
// parameters, log ion density, grain charge and mask, a Poisson solver,
// rhs_standard, Runge_kutta_O4 and the time loop.  Inputs are synthetic
// (seeded noise) instead of an HDF5 file.
//
//   kf_driver N STEPS OUT [eqst|sist|hyst] [newton|nlj|cascade] [M] [Q] [direct|wrapped]
//
// "wrapped" hands Runge_kutta_O4 an interface_rhs of the driver's own that
// forwards to rhs_standard, so the stepper takes its path for right-hand
// sides that are not the library's (four solve() calls per step).
// writes FUx, FUy, FUz, Fni (complex, [N][N][N/2+1]) and Phi (real) to OUT
// as raw little-endian doubles and prints one summary line.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "kielflow_compat.hpp"

// an interface_rhs that is not rhs_standard (RKO4.cpp:105-212 calls it)
class forwarding_rhs : public interface_rhs {
  public:
    explicit forwarding_rhs(interface_rhs &inner) : inner_(inner) {}
    void solve(const double &t, field_imag &FUx, field_imag &FUy, field_imag &FUz, field_imag &Fni) override
    {
        inner_.solve(t, FUx, FUy, FUz, Fni);
    }

  private:
    interface_rhs &inner_;
};

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s N STEPS OUT [eqst|sist|hyst] [newton|nlj|cascade] [M] [Q] [direct|wrapped]\n",
                     argv[0]);
        return 2;
    }
    const int N = std::atoi(argv[1]), steps = std::atoi(argv[2]);
    const std::string out = argv[3], kind = argc > 4 ? argv[4] : "eqst", solver = argc > 5 ? argv[5] : "newton";
    const double M = argc > 6 ? std::atof(argv[6]) : 1.0, Q = argc > 7 ? std::atof(argv[7]) : -0.1;
    const bool wrapped = argc > 8 && std::string(argv[8]) == "wrapped";
    const double L = 16.;
    try {
        axis *ax;
        if (kind == "sist") ax = new axis_CoSiSt(-L / 2, L, N, L * (1. - 1. / 3.) / (2. * std::acos(-1.)));
        else if (kind == "hyst") ax = new axis_CoHySt(-L / 2, L, N, 3);
        else ax = new axis_CoEqSt(-L / 2, L, N);
        grid Omega(*ax, *ax, *ax);
        delete ax;
        parameters Params(M, 0.1, 100., 0., 0.);

        // ln n, grain charge and mask (kielflow.cpp:172-193, 227-233): the
        // noise, the Gaussians of radius a = 2 local cells
        field_real ln_n(Omega), nd(Omega), mask(Omega);
        const double a = 2. * (Omega.x_axis->val_at(N / 2 + 1) - Omega.x_axis->val_at(N / 2));
        const kf_axis axes[3] = {Omega.x_axis->c(), Omega.y_axis->c(), Omega.z_axis->c()};
        kf_compat::check(kf_setup_host(axes, 0, 1e-3, Q, a, ln_n.val, nd.val, mask.val));
        field_imag FUx(Omega), FUy(Omega), FUz(Omega), Fni(Omega);
        std::memset(FUx.val, 0, sizeof(fftw_complex) * FUx.N);
        std::memset(FUy.val, 0, sizeof(fftw_complex) * FUy.N);
        std::memset(FUz.val, 0, sizeof(fftw_complex) * FUz.N);
        FFT(ln_n, Fni);

        // the Poisson solver
        fkt3d_const boundary_shape(0.), boundary_value(0.);
        solver_poisson_jacobi_nlin NLJ(boundary_shape, boundary_value, .8);
        NLJ.set_max_iterations(2000);
        NLJ.set_limit_NormMax(1e-10);
        solver_poisson_multigrid MG(NLJ);
        int MG_steps[4] = {0, -1, -1, -1};
        MG_lvl_control shape[4] = {lvl_down3, lvl_up, lvl_up, lvl_up};
        double MG_err[4] = {0.01, 0.0005, 0.0001, 0.0001};
        MG.set_level_control(4, MG_steps, shape, MG_err);
        solver_poisson_newton_mg NMG(1e-10, 100);
        interface_poisson_solver *ps = &NMG;
        if (solver == "nlj") ps = &NLJ;
        else if (solver == "cascade") ps = &MG;

        field_real Ph(Omega);
        std::memset(Ph.val, 0, sizeof(double) * Ph.N);
        rhs_standard rhs(Params, *ps, Ph, nd, mask, Omega, false, true);
        forwarding_rhs fwd(rhs);
        Runge_kutta_O4 time_integrator(wrapped ? static_cast<interface_rhs &>(fwd) : rhs, 1e-3);
        const auto t0 = std::chrono::steady_clock::now();
        for (int step = 0; step < steps; ++step) time_integrator.solve(FUx, FUy, FUz, Fni);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

        FILE *f = std::fopen(out.c_str(), "wb");
        if (!f) throw std::runtime_error("cannot open " + out);
        for (field_imag *F : {&FUx, &FUy, &FUz, &Fni}) std::fwrite(F->val, sizeof(fftw_complex), F->N, f);
        std::fwrite(Ph.val, sizeof(double), Ph.N, f);
        std::fclose(f);
        std::printf("{\"n\": %d, \"steps\": %d, \"seconds\": %.6f, \"axis\": \"%s\", \"poisson\": \"%s\"}\n", N, steps,
                    secs, kind.c_str(), solver.c_str());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "kf_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
