// kf_adjoint.h -- discrete adjoint of the FP64 RHS, of one RK4 step and of
// the grain drag objective, plus a checkpointed reverse sweep over a run.
//
// Inner products are the Euclidean ones of the stored numbers (a spectral
// field is 2*Nx*Ny*(Nz/2+1) reals).  With that convention
//   R2C^T = N * C2R(W .)       W = 1/2 on interior kz planes, 1 on kz = 0, Nz/2
//   C2R^T = (1/N) W' R2C(.)    W' = 2 on interior kz planes, 1 on kz = 0, Nz/2
//   (i k)^T = -i k,  D23^T = D23,  (Poisson)^T: rho_bar = -J^{-T} phi_bar
// and every transform of the forward RHS (rhs_standard.cpp:89-566) is
// mirrored by one of the reverse sweep.
#pragma once

#include <vector>

#include "kf_rhs.h"

namespace kf {

class Adjoint {
  public:
    Adjoint(Rhs<double> &rhs, double dt);
    ~Adjoint();
    Adjoint(const Adjoint &) = delete;
    Adjoint &operator=(const Adjoint &) = delete;

    // cotangent of one RHS evaluation at stage input S: ob -> Fb (written),
    // and accumulation into M_bar (device scalar) and nd_bar (real field).
    // Runs the forward evaluation with capture first when `recompute`.
    void rhs_adjoint(const double2 *S[4], const double2 *ob[4], double2 *Fb[4], bool recompute = true);
    // adjoint of one RK4 step from state y0: yb holds ybar_{n+1} on entry and
    // ybar_n on return (M_bar / nd_bar accumulate)
    void step_adjoint(const double2 *y0[4], double2 *yb[4]);

    // J = -sum_p W(p) d_x phi(p), W = n_d dV, phi = Poisson(exp(iFFT(D23 Fn)) + n_d)
    double objective(const double2 *F[4], const double *W);
    // dJ/dF (Fn component only, others zero) and nd_bar accumulation
    void objective_adjoint(const double2 *F[4], const double *W, double2 *Fb[4]);

    // forward run of nsteps from y (in place) with checkpoints every k steps,
    // objective at the end, reverse sweep; returns J and the gradients
    // dJ/dM, dJ/dQ (through n_d = Q G) and dJ/dFn0 (spectral, device).
    double run(double2 *y[4], int nsteps, int k, const double *W, const double *G, double *dM, double *dQ,
               double2 *dFn0);

    // out = R2C^T Y (the real-space cotangent of a field whose spectrum is Y)
    void spectral_to_real_T(const double2 *Y, double *out);

    void reset_accumulators();
    double M_bar();              // synchronises
    double *nd_bar() { return ndb_; }
    void set_tolerance(double t) { ltol_ = t; }

    // capture budget (bytes) of the stage intermediates kept in HBM
    void set_capture_budget(size_t b) { capture_budget_ = b; }
    // segments mode: keep every stage's converged potential in the forward
    // sweep and start the recomputation's solves from them (default on)
    void set_keep_potentials(bool v) { keep_potentials_ = v; }
    // bytes of the intermediates of one RK4 step (four stages)
    size_t step_capture_bytes() const;
    // how the last run got its intermediates: RUN_RECOMPUTE (every step
    // recomputed from the segment's stored states), RUN_SEGMENTS (each
    // segment recomputed once and kept) or RUN_TRAJECTORY (the whole
    // trajectory kept by the forward sweep, nothing recomputed); and the
    // iterations of its transposed Poisson solves
    enum { RUN_RECOMPUTE = 0, RUN_SEGMENTS = 1, RUN_TRAJECTORY = 2 };
    int last_run_mode() const { return run_mode_; }
    long last_run_linear_iterations() const { return lin_its_; }

  private:
    void deriv_T(const double2 *Y, int b, double2 *out, int acc, double coef, int d23_out);
    void reverse_step(const Capture caps[4], double2 *yb[4]);
    bool ensure_pool(int steps);
    Rhs<double> &rhs_;
    Grid &g_;
    fft::Plan<double> &p_;
    hipStream_t s_;
    double dt_;
    double ltol_ = 1e-12;
    Capture cap_[4];
    Rk4<double> *rk_ = nullptr;      // forward steps of runs and recomputations
    std::vector<Capture> pool_;      // per-stage intermediates of kept steps
    std::vector<void *> slabs_;      // the pool's allocations
    size_t capture_budget_ = (size_t)64 << 30;
    bool keep_potentials_ = true;
    // the C2R, dS_b multiply and R2C of a stretched-axis transposed
    // derivative in one z pass (KF_PROF_UNFUSED: separate z passes, the same
    // numbers)
    bool prof_fused_ = true;
    int run_mode_ = RUN_RECOMPUTE;
    double *lam_[4] = {};            // transposed Poisson solutions per stage (warm start)
    bool lam_valid_ = false;
    static constexpr int LAM_H = 5;  // transposed solutions kept per stage for the warm start
    double *lamh_[4][LAM_H] = {};    // per stage: the last LAM_H solutions, newest last
    int lam_count_[4] = {0, 0, 0, 0};
    int lam_order_ = 3;              // extrapolation order of the warm start (KF_ADJ_EXTRAP, 0..4)
    long lin_its_ = 0;
    double2 *Fb_[4] = {}, *linb_[4] = {}, *F1b_ = nullptr, *Yt_ = nullptr, *kb_[4][4] = {};
    double *advb_[4] = {}, *Ub_[3] = {}, *Lb_ = nullptr, *rt_ = nullptr, *Pb_ = nullptr;
    double *phib_ = nullptr, *rhob_ = nullptr, *ndb_ = nullptr, *phiJ_ = nullptr, *E_ = nullptr;
    double *Mb_ = nullptr, *part_ = nullptr;
    double2 *yscratch_[4] = {};
};

}  // namespace kf
