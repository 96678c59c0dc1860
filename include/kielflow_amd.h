/*
 * kielflow_amd.h -- C ABI of the MI355X (gfx950) implementation of kielflow's
 * pseudo-spectral right-hand side, RK4 time step and nonlinear Poisson solve.
 *
 * Every entry point returns KF_OK (0) or a KF_ERR_* code; kf_last_error()
 * gives the message of the last failure on the calling thread.  No C++
 * exception crosses this boundary.  All field pointers are DEVICE pointers
 * (use kf_malloc / kf_memcpy_*), laid out as the reference lays out its
 * fields (field_base.hpp:23-24, field_imag.cpp:15-16):
 *   real      double[Nx][Ny][Nz]            z fastest
 *   spectral  double[Nx][Ny][Nz/2+1][2]     interleaved re/im (fftw_complex)
 * Work is enqueued on the context's stream; entry points that return
 * results to the host synchronise that stream.
 *
 * Reference citations name the kielflow file:line of the interface each
 * entry point replaces.
 */
#ifndef KIELFLOW_AMD_H_
#define KIELFLOW_AMD_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    KF_OK = 0,
    KF_ERR_ARG = 1,             /* invalid argument or unsupported size      */
    KF_ERR_HIP = 2,             /* HIP runtime error                          */
    KF_ERR_NOT_CONVERGED = 3,   /* Poisson solve hit its iteration budget     */
    KF_ERR_NAN = 4              /* NaN in the potential (solver_poisson_jacobi_nlin.cpp:78-99) */
};

enum { KF_FP64 = 0, KF_MIXED = 1, KF_FP32 = 2 };

/* axis types as stored by the reference's HDF5 files (IO.cpp:153-164) */
enum { KF_AXIS_EQST = 1, KF_AXIS_HYST = 2, KF_AXIS_SIST = 3 };

/* Poisson algorithms */
enum {
    KF_POISSON_NEWTON_MG = 0,  /* default: Newton + red-black multigrid, FP64 residual test */
    KF_POISSON_NLJ = 1,        /* the reference iteration (solver_poisson_jacobi_nlin.cpp:140-265) */
    KF_POISSON_CASCADE = 2     /* the reference MG cascade (solver_poisson_multigrid.cpp:246-418,
                                  cycle of kielflow.cpp:207-217) */
};

typedef struct kf_ctx_s *kf_ctx;
typedef struct kf_grid_s *kf_grid;
typedef struct kf_fft_s *kf_fft;
typedef struct kf_poisson_s *kf_poisson;
typedef struct kf_rhs_s *kf_rhs;
typedef struct kf_rk4_s *kf_rk4;
typedef struct kf_euler_s *kf_euler;

/* axis_CoEqSt(l0, L, N) (axis_EqSt.hpp:15), axis_CoSiSt(l0, L, N, m)
 * (axis_SiSt.hpp), axis_CoHySt(l0, L, N, m) (axis_HySt.hpp) */
typedef struct {
    int type;
    int n;
    double l0;
    double L;
    double m;
} kf_axis;

/* parameters (parameters.hpp:7-21) */
typedef struct {
    double M, tau, theta, mu, beta;
} kf_parameters;

/* compile-time term switches of rhs_standard.hpp:22-30 made runtime:
 * _RHS_SVISCOSITY and _RHS_PENALIZATION (the others are always on, as shipped) */
typedef struct {
    int sv_on;
    int pen_on;
} kf_rhs_switches;

/* ---------------------------------------------------------------- context */
const char *kf_last_error(void);
int kf_ctx_create(int device, kf_ctx *out);
int kf_ctx_destroy(kf_ctx ctx);
int kf_ctx_synchronize(kf_ctx ctx);
void *kf_ctx_stream(kf_ctx ctx);                 /* hipStream_t */
int kf_malloc(kf_ctx ctx, size_t bytes, void **out);
int kf_free(kf_ctx ctx, void *p);
int kf_memcpy_h2d(kf_ctx ctx, void *dst, const void *src, size_t bytes);
int kf_memcpy_d2h(kf_ctx ctx, void *dst, const void *src, size_t bytes);
int kf_memcpy_d2d(kf_ctx ctx, void *dst, const void *src, size_t bytes);
int kf_memset(kf_ctx ctx, void *dst, int value, size_t bytes);

/* ---------------------------------------------------------------- grid
 * grid(const axis&, const axis&, const axis&) (grid_base.hpp:9-15).
 * Supported extents: 2^a 3^b up to 256 in x and y; Nz even with Nz/2 of
 * that form. */
int kf_grid_create(kf_ctx ctx, const kf_axis axes[3], kf_grid *out);
int kf_grid_destroy(kf_grid g);
int kf_grid_dims(kf_grid g, int *nx, int *ny, int *nz, size_t *nreal, size_t *ncplx);
/* axis::val_at / k_val_at / dS (axis_interface.hpp:10-12, axis.hpp:44) */
int kf_axis_values(const kf_axis *a, int index, double *val, double *k, double *dS);
/* axis::index_at (axis_EqSt.cpp:63-66, axis_HySt.cpp:121-147,
 * axis_SiSt.cpp:98-113): the grid index of coordinate val (EqSt truncates,
 * HySt and SiSt round; SiSt inverts its map by bisection to 1e-6, as the
 * reference does); save_slice takes the z plane at index_at(0) */
int kf_axis_index_at(const kf_axis *a, double val, int *index);

/* ---------------------------------------------------------------- FFT
 * OP_FFT(grid) / OP_iFFT(grid) with operator()(in, out)
 * (OP_FFT.hpp:13-26, OP_iFFT.hpp:13-26; free FFT/iFFT operations.hpp:48-49).
 * r2c is unnormalised; c2r scales by 1/(Nx Ny Nz) (OP_iFFT.cpp:59-61) and
 * transforms the Hermitian projection of the kz=0 and kz=Nz/2 planes.  The
 * input is preserved.  precision KF_FP64 (double fields) or KF_FP32 (float). */
int kf_fft_create(kf_grid g, int precision, kf_fft *out);
int kf_fft_destroy(kf_fft f);
int kf_fft_r2c(kf_fft f, const void *in_real, void *out_cplx);
int kf_fft_c2r(kf_fft f, const void *in_cplx, void *out_real);

/* ---------------------------------------------------------------- spectral operators (FP64)
 * OP_partial_derivative(grid, dir)::operator()(in, out), dir 0/1/2 = e_x/e_y/e_z
 * (OP_partial_derivative.cpp:10-136); in == out allowed. */
int kf_partial_derivative(kf_grid g, int dir, const double *in, double *out);
/* OP_dealiasing(grid, k<2/3)::operator()(field_imag&) (OP_dealiasing.hpp:12-32) */
int kf_dealias23(kf_grid g, double *inout);
/* dealiasing_undesignated(F, exp(-2000 k^15)) (operations.hpp:20-46, rhs_standard.cpp:220) */
int kf_filter_phi(kf_grid g, double *inout);

/* ---------------------------------------------------------------- effects
 * kielflow's effect_* interfaces on FP64 spectral device fields.  Each adds
 * its term of d/dt to the output buffer(s), as the reference's effect
 * classes are used; the arithmetic is the term rhs_standard::solve inlines,
 * so the sequence below reproduces kf_rhs_solve (rhs_standard.cpp:89-566):
 *   out_a  = force_linear(-tau) FU_a [+ spectral_viscosity FU_a]
 *   out_n  = [spectral_viscosity Fn]
 *   force_E(FFT(Phi)); diffusion_u(theta, Fn)       (Fn not yet dealiased)
 *   dealias FU, Fn; translation(M) on FU_a and Fn
 *   out_n += continuity(FU, Fn); advection(FU, FU_a) into out_a
 *   penalization(mask, -2, fix, FU_a or Fn) on each buffer
 *     (fix = -M scaled by the mask for u_x, 0 for u_y, u_z, -10 for ln n) */
/* out -= FFT(u . grad(D23 FXX)), u = iFFT(D23 FU)
 * (effect_advection::execute, effect_advection.hpp:13-15; rhs_standard.cpp:388-437) */
int kf_effect_advection(kf_grid g, const double *FUx, const double *FUy, const double *FUz, const double *FXX,
                        double *out);
/* out = -(div(D23 FU) + FFT(u . grad(D23 Fni)))
 * (effect_continuityEQ::execute, effect_continuityEQ.hpp:14-15; rhs_standard.cpp:315-345) */
int kf_effect_continuity(kf_grid g, const double *FUx, const double *FUy, const double *FUz, const double *Fni,
                         double *out);
/* out_a -= d_a(filter(FPhi)), filter = exp(-2000 khat^15)
 * (effect_force_E::execute, effect_force_E.hpp:14-15; rhs_standard.cpp:219-247) */
int kf_effect_force_E(kf_grid g, const double *FPhi, double *outx, double *outy, double *outz);
/* out_a -= d_a(Fni) / theta, log form
 * (effect_diffusion_u::execute, effect_diffusion_u.hpp:15-16; rhs_standard.cpp:219-247) */
int kf_effect_diffusion_u(kf_grid g, double theta, const double *Fni, double *outx, double *outy, double *outz);
/* out -= SV in, SV = 0.5 (1 - exp(-8 khat^4))
 * (effect_spectral_viscosity::execute, effect_spectral_viscosity.hpp:20-21; rhs_standard.cpp:51-77,119-140) */
int kf_effect_spectral_viscosity(kf_grid g, const double *in, double *out);
/* out -= v d_x(D23 in)
 * (effect_translation::execute, effect_translation.hpp:15-16; rhs_standard.cpp:279-310) */
int kf_effect_translation(kf_grid g, double v, const double *in, double *out);
/* out += k in   (effect_force_linear::execute, effect_force_linear.hpp:14-15; k = -tau in rhs_standard.cpp:119-140) */
int kf_effect_force_linear(kf_grid g, double k, const double *in, double *out);
/* out = D23 FFT((1-H) iFFT(out) + eta H (iFFT(in) - fix [H]) - b iFFT(in)), b the upstream barrier
 * (effect_penalization::execute, effect_penalization.hpp:16-18; rhs_standard.cpp:459-533); eta <= 0 */
int kf_effect_penalization(kf_grid g, const double *mask, double eta, double fix, int fix_scaled_by_mask,
                           const double *in, double *out);

/* ---------------------------------------------------------------- Poisson
 * interface_relaxation_solver (interface_poisson_solver.hpp:14-34) and
 * solver_poisson_jacobi_nlin / solver_poisson_multigrid (their .hpp:38-56). */
int kf_poisson_create(kf_grid g, int precision, int mode, kf_poisson *out);
int kf_poisson_destroy(kf_poisson p);
int kf_poisson_set_tolerance(kf_poisson p, double tol);
int kf_poisson_set_max_iterations(kf_poisson p, int iters);
/* multigrid smoothing of the Newton-MG solver: pre/post red-black sweeps per
 * level and sweeps on the coarsest level (defaults 2, 2, 8 in FP64, 2, 3,
 * 8 in mixed precision) */
int kf_poisson_set_smoothing(kf_poisson p, int pre, int post, int coarse);
/* solver_poisson_multigrid::set_level_control (solver_poisson_multigrid.cpp:
 * 82-147), for the KF_POISSON_CASCADE mode: n level steps, each going down by
 * -level_change[c] levels (1..3: lvl_down, lvl_down2, lvl_down3; injection),
 * up one level (+1: lvl_up; OP_hto2h) or staying (0: lvl_keep), followed by
 * the reference NLJ with steps[c] iterations (-1: up to 1000) and update
 * tolerance tol[c].  The level may not rise above the base grid and must end
 * on it.  Default: kielflow.cpp:207-217's {-3, 1, 1, 1}, {0, -1, -1, -1},
 * {1e-2, 5e-4, 1e-4, 1e-4}. */
int kf_poisson_set_cascade(kf_poisson p, int n, const int *level_change, const int *steps, const double *tol);
/* Newton-MG solves this solver has finished under the host's loop since its
 * creation (those of kf_rhs_solve / kf_rk4_steps included, captured ones
 * not), and the Newton corrections they took: a diagnostic with no
 * counterpart in the reference, whose NLJ reports nothing. */
int kf_poisson_newton_counts(kf_poisson p, long *solves, long *corrections);
/* Boltzmann ions in the equation, as the reference frontend builds the
 * solver (solver_poisson_jacobi_nlin.cpp:192-195, __FRONTEND__):
 * L_h phi = exp(phi) - exp(-theta phi) - rho; theta = 0 turns them off.
 * FP64 solvers only.  The reference modes keep the reference's local Newton
 * (the ion term is in f, not in its derivative); Newton-MG uses the exact
 * Jacobian. */
int kf_poisson_set_ion_boltzmann(kf_poisson p, double theta);
/* The equilibrium start of the reference frontend (create_input_from_MGsolver,
 * frontend.cpp:133-198, and the smoothing of frontend.cpp:291-310): with the
 * grain's charge density nd (FP64 real, from fkt3d_Gauss), solve the
 * equation above with rho = nd by the given mode (KF_POISSON_CASCADE: the
 * reference cascade with frontend.cpp:170-183's level program, tol and
 * max_iterations unused; KF_POISSON_NLJ: its update test at tol;
 * KF_POISSON_NEWTON_MG: the residual test at tol), set ni = exp(-theta phi),
 * and pass ni and phi through exp(-2000 k^15) (dealiasing_undesignated).
 * phi: in/out (the initial guess); ni: out.  The frontend's velocities are
 * zero. */
int kf_frontend_equilibrium(kf_grid g, double theta, const double *nd, int mode, double tol, int max_iterations,
                            double *phi, double *ni, int *iters);
/* solve(Phi_IO, rho): phi FP64 in/out (warm start), rho FP64.  Returns
 * KF_ERR_NOT_CONVERGED when the budget is spent; iters/resid may be NULL. */
int kf_poisson_solve(kf_poisson p, double *phi, const double *rho, int *iters, double *resid);
int kf_poisson_converged(kf_poisson p);         /* converged() of the last solve */
/* residual max|L_h phi - exp(phi) + rho| / max|rho| of the discrete equation */
int kf_poisson_residual(kf_poisson p, const double *phi, const double *rho, double *resid);

/* ---------------------------------------------------------------- RHS
 * rhs_standard(Params, Solver, Phi, n_d, mask, grid) (rhs_standard.hpp:35-37)
 * with interface_rhs::solve(t, FUx, FUy, FUz, Fn) (interface_rhs.hpp:11-12).
 * precision KF_FP64 or KF_MIXED (FFT/RHS/sweeps in FP32). */
int kf_rhs_create(kf_grid g, const kf_parameters *params, const kf_rhs_switches *sw,
                  kf_poisson solver, int precision, kf_rhs *out);
int kf_rhs_destroy(kf_rhs r);
/* phi: persistent FP64 potential the RHS keeps using (held by reference, as
 * rhs_standard::Phi); nd, mask: FP64 real fields (copied). NULL keeps the old. */
int kf_rhs_set_fields(kf_rhs r, double *phi, const double *nd, const double *mask);
int kf_rhs_set_parameters(kf_rhs r, const kf_parameters *params);
/* in place: the four FP64 spectral fields become d/dt of the state */
int kf_rhs_solve(kf_rhs r, double t, double *FUx, double *FUy, double *FUz, double *Fn);
/* number of Poisson solves that did not converge since creation */
int kf_rhs_poisson_failures(kf_rhs r, int *count);
/* diagnostic: evaluate equidistant grids through the general (stretched-axis)
 * path, which materialises every derivative as OP_partial_derivative does */
int kf_rhs_force_general(kf_rhs r, int on);

/* ---------------------------------------------------------------- RK4
 * Runge_kutta_O4(rhs, dt) (RKO4.hpp:29) and interface_ode_solver::solve
 * (interface_ODE_solver.hpp:11-12); FP64 state, FP64 accumulation. */
int kf_rk4_create(kf_rhs r, double dt, kf_rk4 *out);
int kf_rk4_destroy(kf_rk4 k);
int kf_rk4_solve(kf_rk4 k, double *FUx, double *FUy, double *FUz, double *Fn);
/* n steps back to back (no host synchronisation between steps) */
int kf_rk4_steps(kf_rk4 k, int n, double *FUx, double *FUy, double *FUz, double *Fn);
int kf_rk4_set_dt(kf_rk4 k, double dt);
int kf_rk4_time(kf_rk4 k, double *t);
/* The RK4 arithmetic of Runge_kutta_O4::solve for a right-hand side that is
 * not kielflow_amd's (any interface_rhs): a stage input
 * out = y0 + c k (RKO4.cpp:140-153, 167-180, 194-207; c = dt/2, dt/2, dt)
 * and the update out = y0 + (dt/6)(1 k1 + 2 k2 + 2 k3 + 1 k4)
 * (RKO4.cpp:225-238), over n doubles of device memory; out must not overlap
 * the inputs. */
int kf_rk4_stage_input(kf_ctx ctx, size_t n, double c, const double *y0, const double *k, double *out);
int kf_rk4_combine(kf_ctx ctx, size_t n, double dt, const double *y0, const double *k1, const double *k2,
                   const double *k3, const double *k4, double *out);

/* field_integrate::execute: the volume integral of a real field, with
 * dV = prod_b S'_b(xi) L_b / N_b on stretched axes (used at frontend.cpp:
 * 157-159; its header is missing from the reference). */
int kf_field_integrate(kf_grid g, const double *f, double *result);

/* The force decomposition of the reference backend (backend.cpp:271-654):
 * the terms of d/dt u that rhs_standard.cpp accumulates, each as three real
 * fields (x, y, z) in terms[(3 t + a) * Nreal] (may be NULL) with its volume
 * integral in integrals[3 t + a] (may be NULL), t in the backend's order
 *   0 DISSIPATION  -tau u            1 SPECTRAL_VISCOSITY  -SV u
 *   2 E_INT  -d_a filt(phi)          3 DIFFUSION  -d_a ln n / theta
 *   4 E_EXT  -M d_x (D23 u)          5 ADVECTION  -u . grad u
 * from the spectral state and the real potential phi, with the parameters
 * of the FP64 RHS r; and the grain's x-drag -integral n_d d_x phi dV into
 * *drag (may be NULL).  Without penalization the six terms sum to the RHS.
 * The terms follow rhs_standard's arithmetic, not the effect_* library
 * bodies the backend calls (which differ from it). */
#define KF_FORCE_TERMS 6
int kf_force_decomposition(kf_rhs r, const double *FUx, const double *FUy, const double *FUz, const double *Fn,
                           const double *phi, double *terms, double *integrals, double *drag);

/* The stored form of a state, as kielflow writes and reads its HDF5 files
 * (kielflow.cpp:156-182 on load, 290-327 before save_slice / save_all, and
 * the iFFT of save_field_imag, IO.cpp:446-481): out = device double[5][Nreal]
 * holding Ux, Uy, Uz = iFFT(FU), ni = exp(iFFT(F ln n)) and Ph = phi.
 * The reference's FFT(exp(...)) then iFFT round trip is the identity and is
 * not repeated.  Host files go through kielflow_io.h. */
int kf_state_to_real(kf_grid g, const double *FUx, const double *FUy, const double *FUz, const double *Fn,
                     const double *phi, double *out);
/* the same five fields on z plane k only (save_slice, IO.cpp:1175-1219, takes
 * k = z_axis->index_at(0)): out = device double[5][Nx][Ny] */
int kf_state_plane_z(kf_grid g, const double *FUx, const double *FUy, const double *FUz, const double *Fn,
                     const double *phi, int k, double *out);
/* the inverse (load_field_imag, IO.cpp:484-504, and the conversion to
 * ln n, kielflow.cpp:172-182): in = device double[5][Nreal] as above; phi
 * may be NULL.  KF_ERR_ARG when some ni is not positive and finite. */
int kf_state_from_real(kf_grid g, const double *in, double *FUx, double *FUy, double *FUz, double *Fn, double *phi);

/* euler_method(interface_rhs&, dt) and euler_method::solve (euler_method.hpp:
 * 27-31, euler_method.cpp:13-91): y = y + dt k per step, FP64 state */
int kf_euler_create(kf_rhs r, double dt, kf_euler *out);
int kf_euler_destroy(kf_euler e);
int kf_euler_steps(kf_euler e, int n, double *FUx, double *FUy, double *FUz, double *Fn);
int kf_euler_set_dt(kf_euler e, double dt);

/* cfl_all(dt, FUx, FUy, FUz) (CFL.cpp:29-82) */
int kf_cfl_all(kf_grid g, double dt, const double *FUx, const double *FUy, const double *FUz,
               double *dt_new);

/* Allocation-free variant of kf_cfl_all; it has no counterpart in the
 * reference.  It is for callers that cannot allocate device or pinned host
 * memory: code under graph capture or in a tight loop, and harnesses that
 * forbid hipMalloc.  Its scratch (one real field, the spacing table and the
 * per-block maxima) lives in a static device array of KF_NOALLOC_SCRATCH_BYTES
 * per GPU, held until the process ends, so grids of up to about 4.1 million
 * points (e.g. 128^3) fit; larger grids return KF_ERR_ARG.  Every call on one
 * GPU shares that array: calls that overlap on one GPU race.  The result is
 * kf_cfl_all's to rounding (the spacing table is formed on the device, where
 * its multiply-adds may be fused). */
#define KF_NOALLOC_SCRATCH_BYTES (32u << 20)
int kf_cfl_all_noalloc(kf_grid g, double dt, const double *FUx, const double *FUy, const double *FUz,
                       double *dt_new);

/* ---------------------------------------------------------------- batched boxes
 * Parameter sweeps over many independent boxes of uneven sizes (no
 * counterpart in the reference, which runs one box per process).  Each box
 * is a complete solver; after kf_batch_capture one RK4 step of every box is
 * a HIP graph, and kf_batch_steps replays them round-robin over nstreams
 * streams.  precision: KF_FP64 or KF_MIXED. */
typedef struct kf_batch_s *kf_batch;
int kf_batch_create(kf_ctx ctx, int nbox, const kf_axis *axes /* 3 per box */, const kf_parameters *params,
                    const kf_rhs_switches *sw, int precision, double dt, double poisson_tol, int nstreams,
                    kf_batch *out);
int kf_batch_destroy(kf_batch b);
/* device pointers of a box's state (spectral FP64) and potential (real FP64) */
int kf_batch_box(kf_batch b, int box, double **FUx, double **FUy, double **FUz, double **Fn, double **phi);
/* n_d and solid mask of a box (device FP64 real fields, copied) */
int kf_batch_set_fields(kf_batch b, int box, const double *nd, const double *mask);
/* record the graphs; Poisson solves inside run exactly poisson_iterations
 * Newton iterations (unconverged solves are counted as failures) with two
 * pre and two post sweeps per level.  The order of the potential warm start
 * (extrapolation in time) is recorded too, so capture needs at least 5
 * kf_batch_steps first, when it is steady (quartic): KF_ERR_ARG otherwise.
 * A failure during a capture leaves no stream in capture mode. */
int kf_batch_capture(kf_batch b, int poisson_iterations);
int kf_batch_steps(kf_batch b, int nsteps);
int kf_batch_synchronize(kf_batch b);
int kf_batch_poisson_failures(kf_batch b, int *count);

/* ---------------------------------------------------------------- adjoint (FP64)
 * Discrete adjoint of rhs_standard::solve, of Runge_kutta_O4::solve and of
 * the grain drag J = -sum n_d dV d_x phi (phi of the final state), with
 * checkpointed reverse sweeps.  No counterpart in the reference (it has no
 * differentiation); the forward operators are the ones of the entries
 * above.  Cotangents use the Euclidean inner product of the stored numbers
 * (spectral fields as 2 Nx Ny (Nz/2+1) reals). */
typedef struct kf_adjoint_s *kf_adjoint;
int kf_adjoint_create(kf_rhs r, double dt, kf_adjoint *out);
int kf_adjoint_destroy(kf_adjoint a);
/* tolerance of the transposed Poisson solves (default 1e-12) */
int kf_adjoint_set_tolerance(kf_adjoint a, double tol);
int kf_adjoint_reset(kf_adjoint a);
/* S: stage input (4 spectral fields), ob: cotangent of the RHS output;
 * Fb receives the cotangent of S; M_bar and n_d_bar accumulate */
int kf_rhs_adjoint(kf_adjoint a, const double *const S[4], const double *const ob[4], double *const Fb[4]);
/* y0: state before the step; yb: cotangent after the step on entry, before
 * the step on return */
int kf_rk4_step_adjoint(kf_adjoint a, const double *const y0[4], double *const yb[4]);
/* J of a state (dV: cell volumes of the quadrature, real field) */
int kf_objective(kf_adjoint a, const double *const F[4], const double *dV, double *J);
/* dJ/dF of the state of the last kf_objective call; n_d_bar accumulates */
int kf_objective_adjoint(kf_adjoint a, const double *const F[4], const double *dV, double *const Fb[4]);
/* accumulated M_bar (scalar) and n_d_bar (device real field, copied to nd_bar) */
int kf_adjoint_accumulators(kf_adjoint a, double *M_bar, double *nd_bar);
/* forward nsteps from y (in place), checkpoints every k steps, J at the end,
 * reverse sweep with stage recomputation: J, dJ/dM, dJ/dQ (n_d = Q G, G the
 * unit-charge grain shape) and dJ/dFn0 (spectral, device) */
int kf_adjoint_run(kf_adjoint a, double *const y[4], int nsteps, int k, const double *dV, const double *G,
                   double *J, double *dJdM, double *dJdQ, double *dJdFn0);
/* HBM budget (bytes) for kept stage intermediates (default 64 GiB).  When
 * the whole run's intermediates fit (and half the free HBM holds them), the
 * forward sweep keeps them and nothing is recomputed; else when one
 * checkpoint segment's fit, each segment is recomputed once and kept; else
 * every step is recomputed from the segment's stored states */
int kf_adjoint_set_capture_budget(kf_adjoint a, size_t bytes);
/* in the segments mode, the forward sweep also keeps every stage's converged
 * potential (one real field per stage, where half the free HBM holds them)
 * and the recomputation starts each Poisson solve from it: the solve ends at
 * its first check and the recomputed segment reproduces the forward sweep's
 * numbers.  On by default; keep = 0 recomputes the solves from the
 * checkpointed potential. */
int kf_adjoint_set_keep_potentials(kf_adjoint a, int keep);
/* bytes of the kept intermediates of one RK4 step (four stages) */
int kf_adjoint_step_capture_bytes(kf_adjoint a, size_t *bytes);
/* of the last kf_adjoint_run: how it got its intermediates (KF_ADJ_*), and
 * the iterations of its transposed Poisson solves */
enum { KF_ADJ_RECOMPUTE = 0, KF_ADJ_SEGMENTS = 1, KF_ADJ_TRAJECTORY = 2 };
int kf_adjoint_run_info(kf_adjoint a, int *mode, long *linear_iterations);
/* real-space cotangent of a spectral one through Fn = OP_FFT(ln n):
 * dJ/d(ln n)(x) = (R2C^T dJ/dFn)(x) (device fields; dJ/dn0 = that / n0) */
int kf_adjoint_spectral_to_real(kf_adjoint a, const double *dJdF, double *dJdf);

/* ---------------------------------------------------------------- launch timing
 * Diagnostic for roofline figures (no counterpart in the reference, which
 * has no profiler hooks, logger.hpp:13-77 aside).  family: 0 FFT x passes,
 * 1 FFT y passes, 2 FFT z R2C passes, 3 FFT z C2R passes, 4 whole Poisson
 * solves, 5 red-black half sweeps on the finest multigrid level; -1 disables.  Each launch of the enabled family is bracketed by
 * HIP events on its stream; read returns the count and summed milliseconds
 * since the last enable. */
int kf_profile_enable(int family);
int kf_profile_read(int *launches, double *total_ms);

/* ---------------------------------------------------------------- set-up
 * Initial state of the kielflow driver (kielflow.cpp:172-193, 227-233) on
 * the host: seeded ion density n = 1 + amp N(0,1) (Philox4x32-10 +
 * Box-Muller, i-major / z-fastest), state ln n; grain charge density
 * fkt3d_Gauss(Q, a) (masks.hpp:351-366); solid mask exp(-(r/a)^2). Host
 * arrays: ln_n (nreal), nd (nreal), mask (nreal). */
int kf_setup_host(const kf_axis axes[3], unsigned long long seed, double amp, double Q, double a,
                  double *ln_n, double *nd, double *mask);

#ifdef __cplusplus
}
#endif
#endif
