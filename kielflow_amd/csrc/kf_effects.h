// kf_effects.h -- kielflow's effect_* interfaces on FP64 device fields.
#pragma once

#include "kf_fft.h"
#include "kf_grid.h"

namespace kf {

// scratch of the effect operators: 5 real and 2 spectral fields
class EffectWork {
  public:
    explicit EffectWork(Grid &g);
    ~EffectWork();
    EffectWork(const EffectWork &) = delete;
    EffectWork &operator=(const EffectWork &) = delete;
    double *r(int i) { return r_[i]; }
    double2 *c(int i) { return c_[i]; }

  private:
    Grid &g_;
    double *r_[5] = {};
    double2 *c_[2] = {};
};

// out -= FFT(u . grad(D23 FX)), u = iFFT(D23 FU)         effect_advection::execute
void effect_advection(Grid &g, fft::Plan<double> &p, EffectWork &w, const double2 *FU[3], const double2 *FX,
                      double2 *out, hipStream_t s);
// out = -(div(D23 FU) + FFT(u . grad(D23 Fn)))            effect_continuityEQ::execute
void effect_continuity(Grid &g, fft::Plan<double> &p, EffectWork &w, const double2 *FU[3], const double2 *Fn,
                       double2 *out, hipStream_t s);
// out_a -= d_a(Phi-filter FPhi)                            effect_force_E::execute
void effect_force_E(Grid &g, fft::Plan<double> &p, EffectWork &w, const double2 *FPhi, double2 *out[3],
                    hipStream_t s);
// out_a -= d_a(Fn) / theta                                 effect_diffusion_u::execute
void effect_diffusion_u(Grid &g, fft::Plan<double> &p, EffectWork &w, double theta, const double2 *Fn,
                        double2 *out[3], hipStream_t s);
// out -= SV in                                              effect_spectral_viscosity::execute
void effect_spectral_viscosity(Grid &g, const double2 *in, double2 *out, hipStream_t s);
// out -= v d_x(D23 in)                                      effect_translation::execute
void effect_translation(Grid &g, fft::Plan<double> &p, EffectWork &w, double v, const double2 *in, double2 *out,
                        hipStream_t s);
// out += k in                                               effect_force_linear::execute
void effect_force_linear(Grid &g, double k, const double2 *in, double2 *out, hipStream_t s);
// out = D23 FFT((1-H) iFFT(out) + eta H (iFFT(in) - fix [H]) - b iFFT(in))   effect_penalization::execute
void effect_penalization(Grid &g, fft::Plan<double> &p, EffectWork &w, const double *mask, double eta, double fix,
                         int fix_mask, const double2 *in, double2 *out, hipStream_t s);

}  // namespace kf
