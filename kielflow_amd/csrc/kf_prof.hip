// kf_prof.hip -- event bookkeeping of kf_prof.h.
#include "kf_prof.h"

#include <mutex>
#include <vector>

namespace kf {
namespace prof {

int g_active = -1;
static std::mutex g_mu;
static std::vector<hipEvent_t> g_ev;   // start/end pairs in launch order
static std::vector<hipEvent_t> g_pool;

static hipEvent_t get_event()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

void begin(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t e = get_event();
    (void)hipEventRecord(e, s);
    g_ev.push_back(e);
}

void end(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t e = get_event();
    (void)hipEventRecord(e, s);
    g_ev.push_back(e);
}

int enable(int fam)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (hipEvent_t e : g_ev) {
        (void)hipEventSynchronize(e);
        g_pool.push_back(e);
    }
    g_ev.clear();
    g_active = (fam >= 0 && fam < NFAM) ? fam : -1;
    return 0;
}

int read(int *launches, double *total_ms)
{
    std::lock_guard<std::mutex> lk(g_mu);
    double tot = 0.;
    int n = 0;
    for (size_t i = 0; i + 1 < g_ev.size(); i += 2) {
        (void)hipEventSynchronize(g_ev[i + 1]);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, g_ev[i], g_ev[i + 1]) == hipSuccess) {
            tot += ms;
            ++n;
        }
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = tot;
    return 0;
}

}  // namespace prof
}  // namespace kf
