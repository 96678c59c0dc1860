// kf_driver.cpp -- a kielflow-style driver (kielflow.cpp:172-331) written
// against the reference's class interfaces (kielflow_compat.hpp): grid and
// parameters, log ion density, grain charge and mask, a Poisson solver,
// rhs_standard, Runge_kutta_O4 and the time loop.  Inputs are synthetic
// (seeded noise) instead of an HDF5 file.
//
//   kf_driver N STEPS OUT [eqst|sist|hyst] [newton|nlj|cascade] [M] [Q] [direct|wrapped] [rk4|euler]
//
// "wrapped" hands Runge_kutta_O4 an interface_rhs of the driver's own that
// forwards to rhs_standard, so the stepper takes its path for right-hand
// sides that are not the library's (four solve() calls per step).
// writes FUx, FUy, FUz, Fni (complex, [N][N][N/2+1]) and Phi (real) to OUT
// as raw little-endian doubles and prints one summary line.
//
//   kf_driver file IN.h5 STEPS DATA_DIR [newton|nlj|cascade] [SLICE_EVERY] [BACKUP_EVERY] [DT]
//
// runs kielflow's main program from an HDF5 state (kielflow.cpp:147-331):
// grid, time, parameters and fields from IN.h5, n to ln n, the grain charge
// and barrier of the main program, rhs_standard with its shipped switches,
// RK4, and the loop with slice_<I4(step)>.h5 every SLICE_EVERY steps and
// backup_at_<step>.h5 every BACKUP_EVERY (defaults 10, 50, dt 0.01).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "kielflow_io_compat.hpp"

// counter (counter.hpp:6-19)
class counter {
  public:
    explicit counter(int s_max) : step_max(s_max), step(0) {}
    void reset() { step = 0; }
    bool good() const { return step < step_max; }
    void operator++() { step++; }

  private:
    int step_max, step;
};

static int run_file(int argc, char **argv)
{
    if (argc < 5) {
        std::fprintf(stderr,
                     "usage: %s file IN.h5 STEPS DATA_DIR [newton|nlj|cascade] [SLICE_EVERY] [BACKUP_EVERY] [DT]\n",
                     argv[0]);
        return 2;
    }
    const std::string IO_file = argv[2], data_dir = argv[4], solver = argc > 5 ? argv[5] : "cascade";
    const int iterations = std::atoi(argv[3]);
    const int every_slice = argc > 6 ? std::atoi(argv[6]) : 10, every_backup = argc > 7 ? std::atoi(argv[7]) : 50;
    const double t_delta = argc > 8 ? std::atof(argv[8]) : 0.01;
    try {
        counter i_output(every_slice), i_backup(every_backup);
        const double charge_Q = -11498.5, radius_a = 0.15;  // kielflow.cpp:72-73
        grid Omega = load_grid(IO_file);
        double t_total = load_time(IO_file);
        parameters Params = load_parameters(IO_file);
        field_imag FUx = load_field_imag("Ux", IO_file);
        field_imag FUy = load_field_imag("Uy", IO_file);
        field_imag FUz = load_field_imag("Uz", IO_file);
        field_imag Fni = load_field_imag("ni", IO_file);
        field_imag FPh = load_field_imag("Ph", IO_file);

        // ion density to log scale (kielflow.cpp:172-182)
        field_real ni_tmp(Omega);
        field_imag Fni_tmp(Omega);
        iFFT(Fni, ni_tmp);
        for (int i = 0; i < ni_tmp.N; ++i) ni_tmp.val[i] = std::log(ni_tmp.val[i]);
        FFT(ni_tmp, Fni);

        // dust charge and barrier (kielflow.cpp:186-189, 227-229)
        const double scale_Q = 8.1720e-06;
        field_real nd(Omega);
        fkt3d_Gauss dust_3d_fkt(charge_Q * scale_Q, radius_a, radius_a, radius_a);
        nd.fill(dust_3d_fkt);
        field_real Barrier(Omega);
        Barrier.fill2([](double x, double y, double z) {
            return std::exp(-std::pow(x / 0.15, 2.) - std::pow(y / 0.15, 2.) - std::pow(z / 0.15, 2.));
        });

        // the Poisson solvers (kielflow.cpp:197-217)
        fkt3d_const boundary_shape(0.), boundary_value(0.);
        solver_poisson_jacobi_nlin NLJ(boundary_shape, boundary_value, .8);
        NLJ.set_max_iterations(80);
        NLJ.set_limit_NormMax(.5e-5);
        NLJ.set_limit_NormSum(.5e-7);
        solver_poisson_multigrid MG(NLJ);
        int MG_steps[4] = {0, -1, -1, -1};
        MG_lvl_control shape[4] = {lvl_down3, lvl_up, lvl_up, lvl_up};
        double MG_err[4] = {0.01, 0.0005, 0.0001, 0.0001};
        MG.set_level_control(4, MG_steps, shape, MG_err);
        solver_poisson_newton_mg NMG(1e-10, 100);
        interface_poisson_solver *ps = &MG;
        if (solver == "nlj") ps = &NLJ;
        else if (solver == "newton") ps = &NMG;

        field_real Ph(FPh.my_grid);
        iFFT(FPh, Ph);
        rhs_standard rhs(Params, *ps, Ph, nd, Barrier, Omega);
        Runge_kutta_O4 time_integrator(rhs, t_delta);

        const auto t0 = std::chrono::steady_clock::now();
        int slices = 0, backups = 0;
        for (int step = 0; step < iterations; step++, ++i_output, ++i_backup) {
            time_integrator.solve(FUx, FUy, FUz, Fni);
            if (!i_output.good()) {
                FFT(Ph, FPh);
                iFFT(Fni, ni_tmp);
                for (int i = 0; i < ni_tmp.N; ++i) ni_tmp.val[i] = std::exp(ni_tmp.val[i]);
                FFT(ni_tmp, Fni_tmp);
                save_slice(step, Omega, t_total, Params, FUx, FUy, FUz, Fni_tmp, FPh, data_dir);
                ++slices;
                i_output.reset();
            }
            if (!i_backup.good()) {
                iFFT(Fni, ni_tmp);
                for (int i = 0; i < ni_tmp.N; ++i) ni_tmp.val[i] = std::exp(ni_tmp.val[i]);
                FFT(ni_tmp, Fni_tmp);
                save_all(Omega, t_total, Params, FUx, FUy, FUz, Fni_tmp, FPh,
                         data_dir + "/backup_at_" + std::to_string(step) + ".h5");
                ++backups;
                i_backup.reset();
            }
            t_total += t_delta;
        }
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("{\"steps\": %d, \"seconds\": %.6f, \"t\": %.17g, \"slices\": %d, \"backups\": %d, "
                    "\"poisson\": \"%s\"}\n",
                    iterations, secs, t_total, slices, backups, solver.c_str());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "kf_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}

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
    if (argc > 1 && std::string(argv[1]) == "file") return run_file(argc, argv);
    if (argc < 4) {
        std::fprintf(stderr,
                     "usage: %s N STEPS OUT [eqst|sist|hyst] [newton|nlj|cascade] [M] [Q] [direct|wrapped] [rk4|euler]\n",
                     argv[0]);
        return 2;
    }
    const int N = std::atoi(argv[1]), steps = std::atoi(argv[2]);
    const std::string out = argv[3], kind = argc > 4 ? argv[4] : "eqst", solver = argc > 5 ? argv[5] : "newton";
    const double M = argc > 6 ? std::atof(argv[6]) : 1.0, Q = argc > 7 ? std::atof(argv[7]) : -0.1;
    const bool wrapped = argc > 8 && std::string(argv[8]) == "wrapped";
    const bool euler = argc > 9 && std::string(argv[9]) == "euler";
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
        interface_rhs &the_rhs = wrapped ? static_cast<interface_rhs &>(fwd) : rhs;
        std::unique_ptr<interface_ode_solver> integrator;
        if (euler) integrator.reset(new euler_method(the_rhs, 1e-3));
        else integrator.reset(new Runge_kutta_O4(the_rhs, 1e-3));
        interface_ode_solver &time_integrator = *integrator;
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
