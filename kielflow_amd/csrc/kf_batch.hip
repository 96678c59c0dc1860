// kf_batch.hip -- many independent boxes of uneven sizes advanced together
// (parameter sweeps).  Every box is a complete solver (grid, FFT plan,
// Newton-MG Poisson, RHS, RK4) on one of a few HIP streams.  After a few
// eager steps fill the per-stage potential histories, one RK4 step of every
// box is captured into ONE HIP graph: the first stream forks to the others
// through an event, each box's step is recorded on its stream (a branch of
// the graph), and the other streams join back before the capture ends.  A
// batch step is then one graph launch; the small boxes' kernels overlap
// across the branches and no host launch loop remains.  Nothing is shared
// between boxes, so a box's result does not depend on the batch it runs in.
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
    events_.resize(nstreams + 1);
    for (auto &e : events_) KF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
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
    if (exec_) (void)hipGraphExecDestroy(exec_);
    for (Box *x : box_) {
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
    for (auto e : events_) (void)hipEventDestroy(e);
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
    if (x.k32) x.k32->step(x.F);
    else x.k64->step(x.F);
}

void Batch::steps(int n)
{
    for (int i = 0; i < n; ++i) {
        if (exec_) {
            KF_HIP(hipGraphLaunch(exec_, streams_[0]));
            continue;
        }
        for (Box *x : box_) step_box(*x);
    }
}

void Batch::capture(int poisson_iterations)
{
    if (exec_) return;
    sync();
    // a graph keeps the extrapolation order and the buffers it was recorded
    // with: the histories must be full (all lazy allocations done) first
    for (Box *x : box_) {
        const bool ready = x->r32 ? x->r32->history_ready() : x->r64->history_ready();
        if (!ready)
            throw Error("batch capture needs " + std::to_string(Rhs<double>::history_steps()) +
                        " eager steps first (the potential histories are not full)");
    }
    for (Box *x : box_) {
        x->ps->set_graph_iterations(poisson_iterations);
        // a captured solve runs a fixed number of Newton iterations, so
        // extra sweeps per iteration save nothing: two pre and two post
        x->ps->set_smoothing(2, 2, 24);
    }
    const int ns = (int)streams_.size();
    hipStream_t origin = streams_[0];
    // join every other stream back into the origin (also on failure, so
    // that the capture can end and no stream stays in capture mode)
    auto join = [&]() {
        for (int k = 1; k < ns; ++k) {
            KF_HIP(hipEventRecord(events_[k], streams_[k]));
            KF_HIP(hipStreamWaitEvent(origin, events_[k], 0));
        }
    };
    hipGraph_t graph = nullptr;
    KF_HIP(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal));
    try {
        KF_HIP(hipEventRecord(events_[0], origin));
        for (int k = 1; k < ns; ++k) KF_HIP(hipStreamWaitEvent(streams_[k], events_[0], 0));
        for (Box *x : box_) step_box(*x);
        join();
    } catch (...) {
        try {
            join();
        } catch (...) {
        }
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(origin, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
    }
    KF_HIP(hipStreamEndCapture(origin, &graph));
    hipError_t e = hipGraphInstantiate(&exec_, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        exec_ = nullptr;
        KF_HIP(e);
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
