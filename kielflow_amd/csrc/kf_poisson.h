// kf_poisson.h -- nonlinear Poisson solver for Boltzmann electrons,
//   L_h phi - exp(phi) + rho = 0,
// on kielflow's non-uniform 7-point stencil (solver_poisson_jacobi_nlin.cpp:
// 177-199): L_h phi = 2 sum_nb phi_nb / (h_nb (h_+ + h_-)) - D phi with
// D = sum_axes 2/(h_+ h_-), h from axis val_at differences (cpp:112-126).
//
// Solvers:
//  * Newton-multigrid (default): outer Newton iterations on F(phi) with the
//    residual, convergence test  max|F| <= tol * max|rho|  and update in
//    FP64; the Newton correction J d = -F, J = L_h - diag(exp(phi)), is
//    solved by linear V-cycles with red-black Gauss-Seidel smoothing in the
//    sweep precision (FP64, or FP32 in mixed mode: FP64 defect correction).
//    Kernels check a device-side done flag, so a solve can be enqueued
//    without host round trips; an exhausted budget is reported as an error.
//  * the reference iteration: point-Jacobi with damped Newton
//    (newton_method.hpp:17-63), SOR blend, Carre omega and the update-based
//    stop test of solver_poisson_jacobi_nlin.cpp:140-265.
//  * the reference coarse-to-fine cascade (solver_poisson_multigrid.cpp:
//    246-418) built on that iteration.
#pragma once

#include <type_traits>

#include <vector>

#include "kf_common.h"

namespace kf {

struct AxisDesc {
    int type;  // 1 EqSt, 2 HySt, 3 SiSt (IO.cpp:153-164)
    int n;
    double l0, L, m;
};

// axis functions restated (axis_EqSt.cpp:52-84, axis_SiSt.cpp:64-134,
// axis_HySt.cpp:49-80, axis_FoHySt.cpp:107-117); val_at also on the device
__host__ __device__ inline double axis_val_at(const AxisDesc &a, int index)
{
    if (a.type == 1) return a.l0 + index * (a.L / a.n);  // axis_EqSt.cpp:52-55
    if (a.type == 3) {                                  // axis_SiSt.cpp:64-89
        int N = a.n, shift = 0;
        if (index < 0) {
            do { shift += 1; } while (index + shift * N < 0);
        }
        if (index >= N) {
            do { shift -= 1; } while (index + shift * N >= N - 1);
        }
        const double pi = acos(-1.);
        double x = a.l0 + a.L * (double)(index + shift * N) / (N);
        return (x - a.m * sin(2. * pi * x / a.L)) - shift * a.L;
    }
    // axis_HySt.cpp:49-80
    double L = a.L;
    double x_val = a.l0 + L * (double)(index) / (a.n);
    int shift = 0;
    if (x_val < -L / 2.) {
        do { x_val += L; shift -= 1; } while (x_val < -L / 2.);
    }
    if (x_val >= L / 2.) {
        do { x_val -= L; shift += 1; } while (x_val >= a.l0 + L);
    }
    return shift * L + (L / 2.) * sinh(2 * (x_val) * a.m / L) / sinh(a.m);
}
double axis_k_val_at(const AxisDesc &a, int index);
double axis_dS(const AxisDesc &a, int index);

// stencil coefficients of one axis at one level, index i = 0..n-1
template <typename R>
struct AxisCoef {
    const R *cp, *cm, *dd;  // 1/(h+ (h+ + h-)), 1/(h- (h+ + h-)), 2/(h+ h-)
    int uni;                // uniform spacing: the constants below replace the arrays
    R ucp, udd;
};

template <typename R>
struct LevelView {
    int n[3];
    long N;
    AxisCoef<R> c[3];
};

template <typename R>
struct CoarseSet;

struct PoissonStats {
    int cycles;      // Newton iterations, or sweeps of the reference modes
    double resid;    // max|F| / max|rho|
    int converged;
};

enum PoissonMode { POISSON_NEWTON_MG = 0, POISSON_NLJ = 1, POISSON_CASCADE = 2 };

// One solver instance per grid and precision of the sweeps.  FP64 mode
// smooths in FP64.  Mixed mode smooths in FP32 inside a defect-correction
// loop whose residual and correction are FP64.
class Poisson {
  public:
    Poisson(const AxisDesc ax[3], bool mixed, hipStream_t s);
    ~Poisson();
    Poisson(const Poisson &) = delete;
    Poisson &operator=(const Poisson &) = delete;

    void set_mode(int mode) { mode_ = mode; }
    void set_tolerance(double t) { tol_ = t; }
    void set_max_iterations(int m) { max_it_ = m; }
    void set_smoothing(int pre, int post, int coarse) { nu1_ = pre; nu2_ = post; nuc_ = coarse; }
    void set_krylov(bool on, int iters) { krylov_ = on; kiters_ = iters; }
    // k > 0: every Newton-MG solve enqueues exactly k iterations with no
    // host synchronisation, so a stream capture can record it; iterations
    // after convergence return at once on the device, and a solve still
    // unconverged after k counts as a failure.  k = 0: host-driven loop.
    void set_graph_iterations(int k) { graph_it_ = k; }
    int mode() const { return mode_; }
    double tolerance() const { return tol_; }
    int max_iterations() const { return max_it_; }

    // J = L_h - diag(exp(phi)) at the given phi (FP64 operator of the adjoint)
    void set_linearization(const double *phi, hipStream_t s);
    // J x = b to max|b - J x| <= rtol max|b| (FP64; x overwritten; defect
    // correction with one V-cycle, or two BiCGStab steps on stretched grids,
    // per iteration).  Returns iterations, or -1 when max_it is reached.
    // With guess, x holds the initial iterate on entry.
    int linear_solve(double *x, const double *b, double rtol, int max_it, hipStream_t s, bool guess = false);
    // J^T x = b through  V J V^{-1} = J^T  (V = cell volumes, V L_h symmetric)
    int linear_solve_T(double *x, const double *b, double rtol, int max_it, hipStream_t s, bool guess = false);
    // iterations of the last linear solve
    int last_linear_iterations() const { return lin_it_; }

    // phi: FP64 in/out (warm start), rho: FP64 or FP32 real field.
    // Enqueues the solve on stream s.  When sync is true the call waits and
    // fills stats (and returns nonzero if the solve did not converge);
    // otherwise stats are available from last_stats() after the stream ends.
    // With from_guess the initial guess is read from guess_buffer() (which
    // the solve then uses as scratch) and the result is written to phi.
    int solve(double *phi, const double *rho64, const float *rho32, hipStream_t s, bool sync,
              bool from_guess = false);
    double *guess_buffer() { return alt64_; }
    // the next Newton-MG solve starts from a potential that converged for the
    // same right side: enqueue no correction ahead of its first check (which
    // ends it), and keep the expected count of the other solves
    void next_guess_converged() { known_next_ = true; }
    // Newton-MG solves finished with the host loop since creation, and the
    // corrections they took (captured solves are not counted)
    long solves_counted() const { return n_solves_; }
    long corrections_counted() const { return n_corr_; }
    // solve() in two halves, so that work that does not read phi can be
    // enqueued between them (Newton-MG; other modes solve in solve_begin):
    // solve_begin enqueues the corrections the last solve needed, solve_end
    // waits for convergence, enqueuing more corrections as needed.
    void solve_begin(double *phi, const double *rho64, const float *rho32, hipStream_t s, bool from_guess);
    void solve_end();
    PoissonStats last_stats();          // synchronises the stream
    int *device_fail_counter() { return d_fail_; }

    // reference NLJ state (solver_poisson_jacobi_nlin.hpp:66-98)
    struct NljState {
        double norm_max = 0, norm_sum = 0, supremum = 0, infinum = 0;
        double norm_max_old = -1, norm_sum_old = -1;
        double limit_max = 1.45e-06, omega_SOR = 1.;
        int iteration = 0, invocation = 0, max_iterations = 50, converged = 0;
    };
    NljState &nlj() { return nlj_; }
    int nlj_solve_level(int level, double *phi, const double *rho, hipStream_t s);
    // Boltzmann ions exp(-theta phi) in the equation (the reference
    // frontend's build of the solver, solver_poisson_jacobi_nlin.cpp:
    // 192-195): L_h phi = exp(phi) - exp(-theta phi) - rho; 0 turns them off.
    // FP64 solvers only.
    void set_ion_boltzmann(double theta);
    // the cascade's level program (solver_poisson_multigrid::set_level_control)
    void set_cascade(int n, const int *change, const int *steps, const double *tol);

    int levels() const { return (int)lv_.size(); }
    bool ax_desc_uniform() const { return ax_[0].type == 1 && ax_[1].type == 1 && ax_[2].type == 1; }

  private:
    struct Level {
        int n[3];
        long N;
        std::vector<double> h[3];  // H[d][i], i=0..n  (host, for the reference mode)
        double *c64 = nullptr;     // device: [3 axes][cp, cm, dd][n]
        float *c32 = nullptr;
        // linear MG fields: x (correction), b (right side), e (exp(phi) coefficient), r
        double *x64 = nullptr, *b64 = nullptr, *e64 = nullptr, *r64 = nullptr;
        float *x32 = nullptr, *b32 = nullptr, *e32 = nullptr, *r32 = nullptr;
    };
    template <typename R>
    LevelView<R> view(int l) const;
    // zero: x of level l starts at 0 (nothing needs to clear it first)
    template <typename R>
    void vcycle(int l, hipStream_t s, bool zero);
    template <typename R>
    void smooth(int l, int sweeps, hipStream_t s, bool zero = false);
    template <typename R>
    void zero_level(int l, hipStream_t s);
    template <typename R>
    void newton_begin(double *phi, const double *rho64, const float *rho32, hipStream_t s, bool from_guess);
    template <typename R>
    void newton_end();
    template <typename R>
    void newton_check(int it);
    template <typename R>
    void newton_correct();
    // the Newton solve between newton_begin and newton_end
    struct NewtonRun {
        double *phi = nullptr, *A = nullptr, *B = nullptr;
        const double *rho64 = nullptr;
        const float *rho32 = nullptr;
        hipStream_t s = nullptr;
        bool from_guess = false, graph = false, active = false, known = false;
        int it = 0, max_it = 0;
    } run_;
    // BiCGStab on J x = b (level-0 b holds the right side on entry), right
    // preconditioned by one V-cycle; result in level-0 x
    template <typename R>
    void bicgstab(int iters, hipStream_t s);
    template <typename R>
    // (a, b) into scalar slot (and (a, c) into slot + 1 when c is set), then
    // the BiCGStab scalar update op when op >= 0
    void dot(const R *a, const R *b, int slot, hipStream_t s, int op = -1, const R *c = nullptr);
    template <typename R>
    void apply_J_dot(const R *x, R *y, const R *a, const R *c, int slot, int op, hipStream_t s);
    template <typename R>
    CoarseSet<R> coarse_set(int l0) const;
    int first_coarse() const;
    template <typename R>
    void restrict_e(hipStream_t s, const int *ctl);
    int solve_cascade(double *phi, const double *rho, hipStream_t s);
    // level program of the reference cascade (kielflow.cpp:207-217 by default)
    std::vector<int> casc_change_{-3, 1, 1, 1}, casc_steps_{0, -1, -1, -1};
    std::vector<double> casc_tol_{0.01, 0.0005, 0.0001, 0.0001};

    AxisDesc ax_[3];
    bool mixed_;
    int mode_ = POISSON_NEWTON_MG;
    double tol_ = 1e-10;
    int max_it_ = 100;
    // sweeps on the coarsest level: 8 (measured: 24 gave the same Newton
    // counts and cost 4-6% of a stretched-grid step; 4 cost Newton iterations)
    int nu1_ = 2, nu2_ = 2, nuc_ = 8, inner_ = 1;
    long coarse_n_ = 4096;        // levels of at most this many points run in the one-workgroup tail (KF_COARSE_N)
    bool coarse_wave_ = true;     // one wavefront smooths the coarsest level of the tail (KF_COARSE_WAVE=0: all)
    int graph_it_ = 0;
    // the finest level's FP32 x, b, e are stored colour by colour
    bool split0_ = false;
    template <typename R>
    bool split_level(int l) const
    {
        return l == 0 && split0_ && std::is_same<R, float>::value;
    }
    std::vector<Level> lv_;
    // device control block: [0] done flag, [1] cycles, [2] fail counter
    int *d_ctl_ = nullptr;
    int *d_fail_ = nullptr;
    double *d_red_ = nullptr;     // per-block partial reductions
    double *d_stat_ = nullptr;    // [0] resid, [1] rhomax, ...
    double *h_stat_ = nullptr;    // pinned mirror
    int *h_ctl_ = nullptr;
    double *corr64_ = nullptr;    // FP64 scratch (mixed defect correction)
    double *alt64_ = nullptr;     // second potential buffer of the fused Newton update
    int expect_ = 1;              // corrections the last solve needed (host sync point of the next)
    bool known_next_ = false;     // the next solve's guess is a converged potential of the same rho
    long n_solves_ = 0, n_corr_ = 0;
    double ion_theta_ = 0.;       // Boltzmann-ion theta (0: none)
    bool krylov_ = false;         // BiCGStab acceleration (stretched grids)
    int kiters_ = 2;
    // the preconditioning V-cycle reads its right side from, and writes its
    // solution to, the Krylov vectors themselves (finest level launched on
    // its own); KF_PRECOND_COPY=1 copies them through the level's x and b
    bool precond_direct_ = true;
    void *x0o_ = nullptr;         // finest-level x and b of the current V-cycle, when set
    const void *b0o_ = nullptr;
    void *kv_[10] = {};           // BiCGStab vectors (precision of the sweeps)
    double *d_scal_ = nullptr;    // device scalars of BiCGStab
    hipEvent_t ev_[2] = {};       // behind the convergence-flag copies of the last two checks
    double *lsx_ = nullptr, *lsr_ = nullptr, *lsb_ = nullptr;  // linear solve work
    double *lsT_ = nullptr;  // right-hand side of the transposed solve
    int lin_it_ = 0;
    int lin_expect_ = 1;          // corrections the last linear solve needed
    double *vol_[3] = {};         // per-axis cell volumes (H[i+1]+H[i])/2
    NljState nlj_;
    hipStream_t s0_;
};

}  // namespace kf
