// kf_batch.h -- independent boxes of uneven sizes advanced together, one
// captured HIP graph per box and RK4 step, boxes spread over a few streams.
#pragma once

#include <vector>

#include "kf_poisson.h"
#include "kf_rhs.h"

namespace kf {

struct BatchBox {
    AxisDesc ax[3];
    Params pa;
};

class Batch {
  public:
    // precision: 0 FP64, 1 mixed
    Batch(const std::vector<BatchBox> &boxes, const Switches &sw, int precision, double dt, double poisson_tol,
          int nstreams);
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    int size() const { return (int)box_.size(); }
    // FP64 device state of box b (FUx, FUy, FUz, Fn) and its potential
    double2 *state(int b, int f) { return box_.at(b)->F[f]; }
    double *potential(int b) { return box_.at(b)->phi; }
    Grid &grid(int b) { return *box_.at(b)->g; }
    // n_d and mask of box b (device, FP64; copied)
    void set_fields(int b, const double *nd, const double *mask);
    // record one RK4 step of every box as a graph; later steps replay it.
    // The Poisson solves of the graph run exactly poisson_iterations
    // Newton iterations (early exit on the device once converged).
    void capture(int poisson_iterations);
    bool captured() const { return captured_; }
    // n steps of every box (enqueued; call sync() to wait)
    void steps(int n);
    void sync();
    int poisson_failures();  // synchronises

  private:
    struct Box {
        AxisDesc ad[3];
        hipStream_t s = nullptr;
        Grid *g = nullptr;
        Poisson *ps = nullptr;
        Rhs<float> *r32 = nullptr;
        Rhs<double> *r64 = nullptr;
        Rk4<float> *k32 = nullptr;
        Rk4<double> *k64 = nullptr;
        double2 *F[4] = {};
        double *phi = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    void step_box(Box &x);
    int precision_;
    double dt_;
    bool captured_ = false;
    std::vector<hipStream_t> streams_;
    std::vector<Box *> box_;
};

}  // namespace kf
