// kf_prof.h -- optional per-family launch timing with HIP events, for the
// roofline figures.  When a family is enabled, each launch of that family is
// bracketed by two events on the launch stream; the sum of the elapsed times
// is read back after the stream has drained.
#pragma once

#include <hip/hip_runtime.h>

namespace kf {
namespace prof {

enum Family {
    FFT_X = 0,      // strided pass along x (first of C2R, last of R2C)
    FFT_Y = 1,      // strided pass along y
    FFT_ZR2C = 2,   // z pass, real rows -> half spectrum
    FFT_ZC2R = 3,   // z pass, half spectrum -> real rows
    POISSON = 4,    // all Poisson kernels
    RB0 = 5,        // red-black half sweeps on the finest multigrid level
    NFAM = 6
};

extern int g_active;  // enabled family or -1
void begin(hipStream_t s);
void end(hipStream_t s);
inline bool on(int fam) { return g_active == fam; }

struct Scope {
    bool a;
    hipStream_t s;
    Scope(int fam, hipStream_t st) : a(on(fam)), s(st)
    {
        if (a) begin(s);
    }
    ~Scope()
    {
        if (a) end(s);
    }
};

int enable(int fam);                                 // resets the accumulator
int read(int *launches, double *total_ms);           // synchronises recorded events

}  // namespace prof
}  // namespace kf
