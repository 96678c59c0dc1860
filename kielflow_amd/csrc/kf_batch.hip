// kf_batch.hip -- many independent boxes of uneven sizes advanced together
// (parameter sweeps).  Every box is a complete solver (grid, FFT plan,
// Newton-MG Poisson, RHS, RK4) on one of a few HIP streams.  After a few
// eager steps fill the per-stage potential histories, one RK4 step of every
// box is captured into a HIP graph; a batch step is then one graph launch
// per box, round-robin over the streams, so the small boxes' kernels overlap
// across streams and no host launch loop remains.  Nothing is shared between
// boxes, so a box's result does not depend on the batch it runs in.
#include "kf_batch.h"

#include <algorithm>

namespace kf {

Batch::Batch(const std::vector<BatchBox> &boxes, const Switches &sw, int precision, double dt, double poisson_tol,
             int nstreams)
    : precision_(precision), dt_(dt)
{
    if (boxes.empty()) throw Error("empty batch");
    nstreams = std::max(1, std::min(nstreams, 16));
    streams_.resize(nstreams);
    for (auto &s : streams_) KF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (size_t b = 0; b < boxes.size(); ++b) {
        Box *x = new Box();
        x->s = streams_[b % streams_.size()];
        for (int d = 0; d < 3; ++d) x->ad[d] = boxes[b].ax[d];
        x->g = new Grid(x->ad, x->s);
        x->ps = new Poisson(x->ad, precision == 1, x->s);
        x->ps->set_tolerance(poisson_tol);
        x->ps->set_max_iterations(100);
        if (precision == 1) {
            x->r32 = new Rhs<float>(*x->g, *x->ps, boxes[b].pa, sw, x->s);
            x->r32->set_graph_safe(true);
            x->k32 = new Rk4<float>(*x->r32, dt);
        } else {
            x->r64 = new Rhs<double>(*x->g, *x->ps, boxes[b].pa, sw, x->s);
            x->r64->set_graph_safe(true);
            x->k64 = new Rk4<double>(*x->r64, dt);
        }
        const long nc = x->g->nc, nr = x->g->nr;
        for (auto &f : x->F) {
            KF_HIP(hipMalloc(&f, sizeof(double2) * nc));
            KF_HIP(hipMemsetAsync(f, 0, sizeof(double2) * nc, x->s));
        }
        KF_HIP(hipMalloc(&x->phi, sizeof(double) * nr));
        KF_HIP(hipMemsetAsync(x->phi, 0, sizeof(double) * nr, x->s));
        box_.push_back(x);
    }
    sync();
}

Batch::~Batch()
{
    sync();
    for (Box *x : box_) {
        if (x->exec) (void)hipGraphExecDestroy(x->exec);
        delete x->k32;
        delete x->k64;
        delete x->r32;
        delete x->r64;
        delete x->ps;
        delete x->g;
        for (auto f : x->F) (void)hipFree(f);
        (void)hipFree(x->phi);
        delete x;
    }
    for (auto s : streams_) (void)hipStreamDestroy(s);
}

void Batch::set_fields(int b, const double *nd, const double *mask)
{
    Box &x = *box_.at(b);
    if (x.r32) x.r32->set_fields(x.phi, nd, mask);
    else x.r64->set_fields(x.phi, nd, mask);
    KF_HIP(hipStreamSynchronize(x.s));
}

void Batch::step_box(Box &x)
{
    if (x.exec) {
        KF_HIP(hipGraphLaunch(x.exec, x.s));
        return;
    }
    if (x.k32) x.k32->step(x.F);
    else x.k64->step(x.F);
}

void Batch::steps(int n)
{
    for (int i = 0; i < n; ++i)
        for (Box *x : box_) step_box(*x);
}

void Batch::capture(int poisson_iterations)
{
    sync();
    // a graph keeps the extrapolation order and the buffers it was recorded
    // with: the histories must be full (all lazy allocations done) first
    for (Box *x : box_) {
        const bool ready = x->r32 ? x->r32->history_ready() : x->r64->history_ready();
        if (!x->exec && !ready)
            throw Error("batch capture needs " + std::to_string(Rhs<double>::history_steps()) +
                        " eager steps first (the potential histories are not full)");
    }
    for (Box *x : box_) {
        if (x->exec) continue;
        x->ps->set_graph_iterations(poisson_iterations);
        // a captured solve runs a fixed number of Newton iterations, so
        // extra sweeps per iteration save nothing: two pre and two post
        x->ps->set_smoothing(2, 2, 24);
        hipGraph_t graph = nullptr;
        KF_HIP(hipStreamBeginCapture(x->s, hipStreamCaptureModeThreadLocal));
        try {
            if (x->k32) x->k32->step(x->F);
            else x->k64->step(x->F);
        } catch (...) {
            // leave the stream out of capture mode whatever failed
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(x->s, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            throw;
        }
        KF_HIP(hipStreamEndCapture(x->s, &graph));
        hipError_t e = hipGraphInstantiate(&x->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            x->exec = nullptr;
            KF_HIP(e);
        }
    }
    // the capture only recorded the step: nothing advanced
    captured_ = true;
}

void Batch::sync()
{
    for (auto s : streams_) KF_HIP(hipStreamSynchronize(s));
}

int Batch::poisson_failures()
{
    int f = 0;
    for (Box *x : box_) f += x->r32 ? x->r32->poisson_failures() : x->r64->poisson_failures();
    return f;
}

}  // namespace kf
