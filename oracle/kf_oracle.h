/*
 * kf_oracle.h -- CPU FP64 restatement of kielflow's pseudo-spectral RHS,
 * RK4 step and nonlinear Poisson path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in kielflow_amd/.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  The product never
 * links or calls it.
 *
 * Every function cites the reference file:line it restates.  Reference
 * paths are relative to the kielflow source tree.
 *
 * Layout (field_base.hpp:23-24, field_imag.cpp:15-16):
 *   real    fields  double[Nx][Ny][Nz]        (k fastest)
 *   complex fields  double[Nx][Ny][Nz/2+1][2] (interleaved re/im)
 */
#ifndef KF_ORACLE_H_
#define KF_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* axis type ids as in axis_EqSt.cpp:15, axis_SiSt.cpp:29, IO.cpp:153-164 */
enum { OR_EQST = 1, OR_HYST = 2, OR_SIST = 3 };

typedef struct {
    int type;
    int N;
    double l0, L, m;
} or_axis;

typedef struct or_grid_s or_grid;

/* host threads (OpenMP) of the oracle; 1 is the reference's single thread */
void or_set_threads(int n);
int or_get_threads(void);

/* ---- axes ------------------------------------------------------------- */
double or_axis_val_at(const or_axis *a, int index);
double or_axis_k_val_at(const or_axis *a, int index);
double or_axis_dS(const or_axis *a, int index);

/* ---- grid ------------------------------------------------------------- */
or_grid *or_grid_create(const or_axis *ax, const or_axis *ay, const or_axis *az);
void or_grid_destroy(or_grid *g);
size_t or_grid_nreal(const or_grid *g);
size_t or_grid_ncplx(const or_grid *g);

/* ---- FFT (OP_FFT.cpp:49-63, OP_iFFT.cpp:49-64) ------------------------ */
void or_fft_r2c(const or_grid *g, const double *in, double *out);
void or_fft_c2r(const or_grid *g, const double *in, double *out);
/* 1D complex DFT of length n (sign -1 forward), exposed for pinning. */
void or_fft1d(double *x, int n, int sign);

/* ---- spectral operators ---------------------------------------------- */
void or_dealias23(const or_grid *g, double *inout);                 /* OP_dealiasing */
void or_filter_phi(const or_grid *g, double *inout);                /* operations.hpp:20-46 */
void or_deriv(const or_grid *g, int dir, const double *in, double *out); /* OP_partial_derivative */
void or_sv_table(const or_grid *g, double *out_nc);                 /* rhs_standard.cpp:51-77 */
void or_barrier_table(const or_grid *g, double *out_nr);            /* rhs_standard.cpp:39-49 */
void or_dealias23_table(const or_grid *g, double *out_nc);

/* ---- Poisson ---------------------------------------------------------- */
typedef struct {
    /* reference NLJ state (solver_poisson_jacobi_nlin.hpp:66-98) */
    double norm_max, norm_sum, supremum, infinum;
    double norm_max_old, norm_sum_old;
    double limit_max;
    double omega_SOR;
    int iteration, invocation, max_iterations, converged;
    /* theta of the __FRONTEND__ build's Boltzmann ions (cpp:192-195), 0: off */
    double ion_theta;
} or_nlj_state;

void or_nlj_init(or_nlj_state *s);
/* one solve of solver_poisson_jacobi_nlin::solve (cpp:50-108) */
int or_nlj_solve(or_nlj_state *s, const or_grid *g, double *phi, const double *rho);

/* reference MG cascade (solver_poisson_multigrid.cpp:246-418) with the
 * cycle of kielflow.cpp:207-217.  Works on the given grid; levels are the
 * grid resized by powers of two (grid_base.hpp resize). */
int or_mg_cascade_solve(or_nlj_state *s, const or_grid *g, double *phi, const double *rho,
                        int ncascade, const int *lvl_change, const int *steps, const double *tol);

/* FAS V-cycle solver of the same discrete equation, run to a residual
 * tolerance: max|L_h phi - exp(phi) + rho| <= tol * max|rho|.
 * Returns cycles used (>=0) or -1 when max_cycles is reached. */
int or_poisson_fas(const or_grid *g, double *phi, const double *rho, double tol,
                   int max_cycles, double *resid_out);
/* residual of the discrete equation (solver_poisson_jacobi_nlin.cpp:177-199) */
double or_poisson_residual(const or_grid *g, const double *phi, const double *rho, double *r_out);

/* equilibrium start of the reference frontend (frontend.cpp:133-198,
 * 291-310); see kf_oracle.c */
int or_frontend_init(const or_grid *g, double theta, const double *nd, int use_cascade, double tol, int max_it,
                     double *phi, double *ni);

/* ---- RHS (rhs_standard.cpp:89-566) ----------------------------------- */
enum { OR_POISSON_FAS = 0, OR_POISSON_NLJ = 1, OR_POISSON_CASCADE = 2 };

typedef struct {
    double M, tau, theta, mu, beta;   /* parameters.hpp:13-18 */
    int sv_on;                        /* _RHS_SVISCOSITY  (rhs_standard.hpp:30) */
    int pen_on;                       /* _RHS_PENALIZATION (rhs_standard.hpp:29) */
    int poisson_mode;
    double poisson_tol;
    int poisson_max;                  /* cycles (FAS) or iterations (NLJ) */
} or_params;

typedef struct or_rhs_s or_rhs;

or_rhs *or_rhs_create(const or_grid *g, const or_params *p);
void or_rhs_destroy(or_rhs *r);
/* phi (warm start, in/out), dust density and solid mask, real fields */
void or_rhs_set_fields(or_rhs *r, const double *phi, const double *nd, const double *mask);
void or_rhs_get_phi(const or_rhs *r, double *phi);
int  or_rhs_eval(or_rhs *r, double t, double *FUx, double *FUy, double *FUz, double *Fn);
int  or_rhs_poisson_status(const or_rhs *r, double *last_resid, int *last_iters);

/* ---- Euler (euler_method.cpp:13-91) ------------------------------------ */
int or_euler_step(or_rhs *r, double dt, double *t_io, double *FUx, double *FUy, double *FUz, double *Fn);

/* ---- RK4 (RKO4.cpp:50-248) -------------------------------------------- */
int or_rk4_step(or_rhs *r, double dt, double *t_io,
                double *FUx, double *FUy, double *FUz, double *Fn);

/* OP_8htoh-style injection by f and OP_hto2h prolongation (field_interpolation.cpp:56-147,233-315) */
void or_mg_inject(const double *in, const int nin[3], double *out, int f);
void or_mg_prolong(const double *in, const int nin[3], double *out, const or_axis out_ax[3]);

/* tangent-linear model (forward-mode derivative): F is the state (in place,
 * becomes the RHS as or_rhs_eval), T the tangent direction (becomes the
 * tangent of the RHS) for parameter directions dM and dnd (may be NULL) */
int or_rhs_eval_tl(or_rhs *r, double *F[4], double *T[4], double dM, const double *dnd);
int or_rk4_step_tl(or_rhs *r, double dt, double *y[4], double *ty[4], double dM, const double *dnd);
double or_objective_tl(or_rhs *r, const double *Fn, const double *Tn, const double *dV, const double *dnd,
                       double *dJ);
/* CFL.cpp:29-82 */
double or_cfl(const or_grid *g, double dt, const double *FUx, const double *FUy, const double *FUz);

/* ---- set-up ------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_normal_fill(uint64_t seed, size_t n, double *out);
/* fkt3d_Gauss (masks.hpp:351-366) sampled at the grid points */
void or_gauss_fill(const or_grid *g, double amplitude, double a, double *out);
/* solid mask exp(-(r/a)^2) with peak 1 (kielflow.cpp:228-229) */
void or_mask_fill(const or_grid *g, double a, double *out);
/* ln(1 + amp*N(0,1)) -> spectral (kielflow.cpp:172-182) */
void or_init_density(const or_grid *g, uint64_t seed, double amp, double *Fn);

#ifdef __cplusplus
}
#endif
#endif
