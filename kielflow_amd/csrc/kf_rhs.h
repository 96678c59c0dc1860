// kf_rhs.h -- kielflow's right-hand side (rhs_standard::solve,
// rhs_standard.cpp:89-566) and classic RK4 (Runge_kutta_O4::solve,
// RKO4.cpp:50-248) on the GPU.
//
// State: FUx, FUy, FUz (velocity perturbation) and Fn (spectral ln n), each
// complex [Nx][Ny][Nz/2+1].  The RHS overwrites its input with d/dt of the
// state (interface_rhs.hpp:11-12).
#pragma once

#include "kf_common.h"
#include "kf_fft.h"
#include "kf_grid.h"
#include "kf_poisson.h"

namespace kf {

struct Params {
    double M, tau, theta, mu, beta;  // parameters.hpp:13-18
};

struct Switches {
    int sv_on;   // _RHS_SVISCOSITY  (rhs_standard.hpp:30)
    int pen_on;  // _RHS_PENALIZATION (rhs_standard.hpp:29)
};

// Forward intermediates of one evaluation kept for the adjoint (FP64):
// real U_b, L = iFFT(D23 Fn), phi, G_ab = iFFT(D23 d_b(D23 F_a)) and the
// spectral d_x(D23 F_a) of a = x, y, z, n.
struct Capture {
    double *U[3];
    double *L;
    double *phi;
    double *G[12];
    double2 *dF0[4];
};

// The RK4 update of the fields, run by the last store of an RHS evaluation
// when its path allows (the fused equidistant path with penalization):
// after stage s < 3, acc = (s == 0 ? k : acc + 2 k) and S = y0 + c k; after
// the last stage, y = y0 + c (acc + k) and, with next, S = y.
struct RkEpilogue {
    double2 *y[4];
    double2 *acc[4];
    double c;
    int final_, first, next;
    double2 *s64n;  // the next stage input of ln n in FP64 (or null)
};

// RHS in compute precision R (double: FP64 mode; float: mixed mode).
template <typename R>
class Rhs {
  public:
    Rhs(Grid &g, Poisson &p, const Params &pa, const Switches &sw, hipStream_t s);
    ~Rhs();
    Rhs(const Rhs &) = delete;
    Rhs &operator=(const Rhs &) = delete;

    // phi: caller-owned persistent FP64 potential (warm start, in/out);
    // nd, mask: real FP64 fields copied (and converted) into the RHS.
    void set_fields(double *phi, const double *nd, const double *mask);
    // F[0..3]: stage input, overwritten with the RHS (same precision R)
    void eval(double t, C2<R> *F[4]);
    void set_force_general(bool v) { force_general_ = v; }
    // RK stage (0..3) of the next evaluations, or -1.  Within an RK run the
    // Poisson solve of stage s starts from the quartic extrapolation in time
    // of the last five converged potentials of stage s (of lower order while
    // fewer exist).
    void set_stage(int s);
    // the per-stage potential histories (six FP64 fields per stage)
    void alloc_history();
    // the next evaluation runs *e as its last store when it can; done tells
    // whether it did (else the caller runs the RK4 update itself)
    void set_rk_epilogue(const RkEpilogue *e)
    {
        rk_ = e;
        rk_done_ = false;
    }
    bool rk_epilogue_done() const { return rk_done_; }
    // evaluations store their intermediates into *c (general path) while set
    void set_capture(Capture *c) { cap_ = c; }
    // evaluations at RK stage s store their intermediates into c4[s] while set
    void set_capture_stages(Capture *c4) { cap_stages_ = c4; }
    // graph-safe history: rotate the per-stage potentials by copying
    // (fixed buffers) rather than by swapping pointers on the host
    void set_graph_safe(bool v);
    // every stage's history holds enough potentials for the full order of
    // the extrapolation (a captured step keeps the order it was recorded with)
    bool history_ready() const;
    static constexpr int history_steps() { return 5; }
    // Mixed precision: the FP64 spectrum of ln n that the next evaluation's
    // rho = exp(ln n) + n_d is formed from (an FP64 inverse transform), in
    // place of its FP32 stage input (null: off).  The FP32 stage input's
    // rounding otherwise reaches rho and, through the potential history, the
    // Newton residual of later solves.  Set per stage by the RK4 stepper.
    void set_rho_source(const double2 *Fn64) { rho_src_ = Fn64; }
    bool wants_rho_source() const { return plan64_ != nullptr; }
    // keep the converged potential of RK stage s in p4[s] after each solve
    // (null: off); phi_cur_ is that potential, in phi_ or a history slot
    void set_phi_sink(double *const *p4) { sink_ = p4; }
    // start the solve of RK stage s from p4[s] instead of the extrapolated
    // history (null: off).  A recomputation of a step from the same state
    // with the potentials its first run kept converges at the first check and
    // reproduces that run's numbers.
    void set_phi_source(const double *const *p4) { src_ = p4; }
    // forget the per-stage potential history (the run jumped in time)
    void reset_history()
    {
        for (int &h : hc_) h = 0;
    }
    void set_params(const Params &p) { pa_ = p; }
    const Params &params() const { return pa_; }
    const Switches &switches() const { return sw_; }
    const R *mask() const { return mask_; }
    const R *nd() const { return nd_; }
    Poisson &poisson() { return ps_; }
    fft::Plan<R> &plan() { return plan_; }
    Grid &grid() { return g_; }
    hipStream_t stream() const { return s_; }
    int poisson_failures();   // synchronises
    // real-space fields of the last evaluation (device)
    const R *U(int b) const { return U_[b]; }
    const R *lnn() const { return L_; }
    double *phi() { return phi_; }

  private:
    void eval_fused(C2<R> *F[4]);
    void eval_general(C2<R> *F[4]);
    void alloc_general();
    // rho from Fn (or from its staged x/y passes), then the Poisson solve
    void poisson_step(const C2<R> *Fn, const C2<R> *staged = nullptr);
    // poisson_step in two halves (work that does not read phi may go between)
    void poisson_begin(const C2<R> *Fn, const C2<R> *staged = nullptr);
    void poisson_end();
    void output(C2<R> *F[4]);
    void deriv_into(const C2<R> *X, int d23_in, int b, C2<R> *out, int accumulate);
    // the three derivatives of D23 X on a grid stretched along every axis,
    // from one staging of the inverse x and y passes
    void derivs_shared(const C2<R> *X, C2<R> *const out[3], int accumulate);

    Grid &g_;
    Poisson &ps_;
    Params pa_;
    Switches sw_;
    hipStream_t s_;
    fft::Plan<R> plan_;
    bool force_general_ = false;
    // penalization fused into one z pass per field with the products (the
    // environment variable KF_PEN_UNFUSED selects the separate passes, which
    // compute the same numbers)
    bool pen_fused_ = true;
    // the metric multiply of a stretched-axis derivative fused into one z
    // pass (KF_PROF_UNFUSED selects the separate C2R and R2C z passes, which
    // compute the same numbers)
    bool prof_fused_ = true;
    // on grids stretched along every axis, the three derivatives of a field
    // share its inverse x and y passes (KF_XY_UNSHARED: one x/y/z round
    // trip per derivative; the two orders round differently)
    bool xy_shared_ = true;
    const RkEpilogue *rk_ = nullptr;
    bool rk_done_ = false;
    double *phi_ = nullptr;     // caller-owned
    double *phi_cur_ = nullptr; // the potential of the current evaluation (phi_ or a history slot)
    double *slot_ = nullptr;    // history slot the current solve writes (or null)
    R *phiR_ = nullptr;         // FP32 copy (mixed) for the forward transform
    R *nd_ = nullptr, *mask_ = nullptr;
    R *U_[3] = {nullptr, nullptr, nullptr};
    R *L_ = nullptr, *rho_ = nullptr;
    double *rho64_ = nullptr, *nd64_ = nullptr;  // mixed: rho and n_d in FP64 (rho64_on_)
    bool rho64_on_ = false;
    R *adv_[4] = {nullptr, nullptr, nullptr, nullptr};
    R *rtmp_ = nullptr;
    C2<R> *lin_[4] = {nullptr, nullptr, nullptr, nullptr};
    C2<R> *F1_ = nullptr;
    C2<R> *dF_[12] = {};
    C2<R> *H_[4][3] = {};       // staged x/y passes of the state fields (equidistant grids)
    int stage_ = -1;
    Capture *cap_ = nullptr;
    Capture *cap_stages_ = nullptr;
    bool graph_safe_ = false;
    Capture *active_capture() const
    {
        if (cap_) return cap_;
        return (cap_stages_ && stage_ >= 0) ? &cap_stages_[stage_] : nullptr;
    }
    double *ph_[4][6] = {};  // per RK stage: the last six converged potentials, newest last
    int hc_[4] = {0, 0, 0, 0};
    int extrap_order_ = 4;   // highest order of the extrapolation in time
    double *const *sink_ = nullptr;
    const double *const *src_ = nullptr;
    const double2 *rho_src_ = nullptr;
    fft::Plan<double> *plan64_ = nullptr;  // mixed, with rho in FP64: the FP64 transform of rho_src_
};

// RK4 with an FP64 state and accumulator; stages evaluated in precision R.
template <typename R>
class Rk4 {
  public:
    Rk4(Rhs<R> &rhs, double dt);
    ~Rk4();
    Rk4(const Rk4 &) = delete;
    Rk4 &operator=(const Rk4 &) = delete;
    void step(double2 *y[4]);          // in place
    void steps(double2 *y[4], int n);  // n steps in place
    double dt() const { return dt_; }
    void set_dt(double dt) { dt_ = dt; }
    double t() const { return t_; }
    void set_t(double t) { t_ = t; }
    Rhs<R> &rhs() { return rhs_; }

  private:
    Rhs<R> &rhs_;
    double dt_, t_ = 0.;
    C2<R> *S_[4] = {};
    double2 *acc_[4] = {};
    double2 *S64n_ = nullptr;  // FP64 stage input of ln n, where the RHS takes it
};

// euler_method::solve (euler_method.cpp:13-91): y = y + dt k with an FP64
// state, the RHS evaluated in precision R (its Poisson solves warm-started
// from the extrapolation of the previous steps' potentials)
template <typename R>
class Euler {
  public:
    Euler(Rhs<R> &rhs, double dt);
    ~Euler();
    Euler(const Euler &) = delete;
    Euler &operator=(const Euler &) = delete;
    void steps(double2 *y[4], int n);
    double dt() const { return dt_; }
    void set_dt(double dt) { dt_ = dt; }
    double t() const { return t_; }

  private:
    Rhs<R> &rhs_;
    double dt_, t_ = 0.;
    C2<R> *S_[4] = {};
};

// conversions and small field kernels shared with the C ABI
void to_single(const double2 *in, float2 *out, long n, hipStream_t s);
void to_double(const float2 *in, double2 *out, long n, hipStream_t s);
void real_to_single(const double *in, float *out, long n, hipStream_t s);

// spectral operators on FP64 fields (OP_partial_derivative, OP_dealiasing,
// dealiasing_undesignated with the Phi filter)
void op_partial_derivative(Grid &g, fft::Plan<double> &p, int dir, const double2 *in, double2 *out,
                           hipStream_t s);
void op_dealias23(Grid &g, double2 *inout, hipStream_t s);
void op_filter_phi(Grid &g, double2 *inout, hipStream_t s);
// CFL.cpp:29-82 (perturbation velocity, 0.01 cap)
double cfl_all(Grid &g, fft::Plan<double> &p, double dt, const double2 *F[3], hipStream_t s);
// the same with a caller-provided device scratch of at least
// cfl_all_scratch_doubles(g) doubles (nothing is allocated)
size_t cfl_all_scratch_doubles(const Grid &g);
double cfl_all_into(Grid &g, fft::Plan<double> &p, double dt, const double2 *F[3], double *ws, hipStream_t s);
// the volume integral of f [* w] over the box, dV = prod_b S'_b L_b / N_b
// (field_integrate: used at frontend.cpp:157-159, its header is missing
// from the reference); blocks of the grid's workspace hold the partial sums
double field_integrate(Grid &g, const double *f, const double *w, hipStream_t s);
// The stored form of a state (the fields of kielflow's HDF5 files,
// kielflow.cpp:290-327): out = [Ux, Uy, Uz, n, Ph], 5 real fields, with
// U = iFFT(FU), n = exp(iFFT(F ln n)) (linear density) and Ph = phi.
void state_to_real(Grid &g, fft::Plan<double> &p, const double2 *F[4], const double *phi, double *out,
                   hipStream_t s);
// the same 5 fields on the z plane k only: out = double[5][nx][ny]
void state_plane_z(Grid &g, fft::Plan<double> &p, const double2 *F[4], const double *phi, int k, double *out,
                   hipStream_t s);
// the inverse (kielflow.cpp:156-182): F = FFT(U), F ln n = FFT(log n),
// phi = Ph (if phi is not null); false if some n is not positive and finite
bool state_from_real(Grid &g, fft::Plan<double> &p, const double *in, double2 *F[4], double *phi, hipStream_t s);
// n_i = exp(-theta phi) (frontend.cpp:186-189)
void boltzmann_ions(const double *phi, double theta, double *ni, long n, hipStream_t s);
// the RK4 arithmetic of RKO4.cpp for a right-hand side evaluated elsewhere:
// out = y0 + c k, and out = y0 + (dt/6)(k1 + 2 k2 + 2 k3 + k4), over n doubles
void rk4_stage_input(long n, double c, const double *y0, const double *k, double *out, hipStream_t s);
void rk4_combine(long n, double dt, const double *y0, const double *const k[4], double *out, hipStream_t s);

}  // namespace kf
