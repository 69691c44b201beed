/*
 * mfc_rhs.h -- C ABI of the MI355X (gfx950) implementation of MicroFC's
 * simulation right-hand side and TVD Runge-Kutta time stepper.
 *
 * Every entry point returns 0 on success and a negative code on failure
 * (mfc_last_error() then holds a message); nothing throws across the ABI.
 * Field arguments are host arrays of sys_size DEVICE pointers, one per
 * variable, each pointing at the first element of that variable's array,
 * i.e. c_loc(q_cons_vf(i)%sf(-b, -b, -b)) on the Fortran side:
 *   ghosted fields (q_cons_vf, q_prim_vf, q_cons_ts): (m+2b+1) x (n+2b+1) x
 *     (p+2b+1) doubles, x fastest (a collapsed direction has extent 1);
 *   rhs_vf: interior only, (m+1) x (n+1) x (p+1) doubles, x fastest.
 * b = buff_size as the reference sizes it (src/simulation/m_global_parameters.fpp:
 * 355-359): polyn + 2, or 2*polyn + 2 when any fluid_pp(i)%Re is positive,
 * with polyn = (weno_order-1)/2; mfc_sizes reports it.  All work is queued on the library's
 * stream (mfc_stream()); calls return before the device has finished unless
 * they return a host value.  Call mfc_sync() before reading results on the host.
 */
#ifndef MFC_RHS_H
#define MFC_RHS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFC_NUM_FLUIDS_MAX 10

/* Arithmetic of the WENO/HLLC bulk. */
#define MFC_PREC_FP64 0      /* plain path: FP64 throughout, reference evaluation order */
#define MFC_PREC_FP32_BULK 1 /* WENO and HLLC in FP32; cons->prim, flux divergence and RK in FP64 */
#define MFC_PREC_FP32_LEARNED 2 /* FP32 bulk with learned WENO5 weights (needs mfc_mlp_init) */
#define MFC_PREC_FP64_FAST 3 /* FP64 throughout with FMA contraction and reciprocal-multiply divisions
                                (WENO weights normalised by one reciprocal of their sum); within the
                                reference's 1e-12 relative bar of the plain form, not bit-equal */

/* Learned WENO5 weights: a 5 -> 32 -> 32 -> 3 MLP (ReLU) whose logits z
 * modulate the plain weights, w_k ~ w_plain_k exp(g z_k), with the gate
 * g = min(1, max(0, sum_k |w_plain_k - d_k| - 0.005) / 0.03) and
 * d = (0.3, 0.6, 0.1) the uniform-grid ideal weights of the oriented face
 * (w = w_plain where g = 0).  Parameters are one
 * FP32 block of MFC_MLP_NPARAM floats: W1[16][32] (input rows >= 5 zero),
 * b1[32], W2[32][32], b2[32], W3[32][16] (output columns >= 3 zero), b3[16].
 * A training sample is 12 floats (structure of arrays, 12 planes of n): the
 * normalised stencil x[5], the normalised candidate polynomials p[3], the
 * plain weights w_plain[3] and the normalised fine-grid face value t; the
 * loss is mean((sum_k w_k p_k - t)^2).  None of this has a counterpart in the
 * reference; it replaces the nonlinear weights of src/simulation/m_weno.fpp:508-528. */
#define MFC_MLP_NPARAM 2128
#define MFC_MLP_SAMPLE_FIELDS 12

/* Case parameters; names and meaning follow the namelist &user_inputs of
 * src/simulation/m_start_up.fpp:73-83 (p and bc_z are the 3D extension). */
typedef struct mfc_params {
    int m, n, p;
    int num_fluids;
    int weno_order;
    int time_stepper;
    int t_step_stop;
    int bc_x_beg, bc_x_end, bc_y_beg, bc_y_end, bc_z_beg, bc_z_end;
    double weno_eps;
    double gammas[MFC_NUM_FLUIDS_MAX];  /* fluid_pp(i)%gamma  */
    double pi_infs[MFC_NUM_FLUIDS_MAX]; /* fluid_pp(i)%pi_inf */
    int precision;                      /* MFC_PREC_* */
    /* fluid_pp(i)%Re(1:2): shear and volume Reynolds numbers.  An entry <= 0
     * (the reference's default dflt_real = -1e6, or a zeroed struct) leaves
     * fluid i out of that viscous term (m_global_parameters.fpp:314-338); any
     * positive entry turns the viscous terms on (m_viscous.fpp,
     * m_riemann_solvers.fpp:683-902), in 1D and 2D and in MFC_PREC_FP64. */
    double Re[MFC_NUM_FLUIDS_MAX][2];
    /* weno_Re_flux (m_global_parameters.fpp:103): nonzero takes the viscous
     * velocity gradients from WENO face values by the divergence theorem and
     * reconstructs them to the faces with WENO (m_viscous.fpp:200-214,
     * :372-591, m_rhs.fpp:514-528); zero uses finite differences
     * (m_viscous.fpp:216-355).  Ignored by inviscid cases. */
    int weno_Re_flux;
} mfc_params;

/* Replaces the module set-up chain s_initialize_global_parameters_module,
 * s_initialize_variables_conversion_module, s_initialize_riemann_solvers_module,
 * s_initialize_rhs_module and s_initialize_time_steppers_module
 * (src/simulation/p_main.fpp:125-183) and the input checks of
 * s_check_input_file (src/simulation/m_start_up.fpp:117-290). */
int mfc_rhs_init(const mfc_params *prm);

/* Replaces s_initialize_weno_module / s_compute_weno_coefficients
 * (src/simulation/m_weno.fpp:103-363) for direction dir (0 = x, 1 = y, 2 = z):
 * s_cb is the ghosted host array x_cb(-1-b : s+b) and ds the ghosted host
 * array dx(-b : s+b) as left by s_populate_grid_variables_buffers
 * (src/simulation/m_start_up.fpp:517-655). */
/* In the forms other than MFC_PREC_FP64, a direction whose cells' WENO
 * coefficients all lie within 5e-13 (relative) of the first reconstructed
 * cell's (equal widths up to the rounding of the cell boundaries) is run
 * with the first cell's coefficients; MFC_PREC_FP64 does so only when they
 * are equal bit for bit. */
int mfc_rhs_set_grid(int dir, const double *s_cb, const double *ds);

/* Replaces s_compute_rhs(q_cons_vf, q_prim_vf, rhs_vf, t_step)
 * (src/simulation/m_rhs.fpp:405-679).  The ghost cells of q_cons_vf are
 * filled in place (the reference fills a private copy; the ghost layers are
 * scratch in both).  q_prim_vf may be NULL; when given it receives the ghosted
 * primitive field (the reference copies it out when run_time_info is set).
 * When t_step == t_step_stop only the ghost fill happens, as in the reference's
 * early return at m_rhs.fpp:452. */
int mfc_compute_rhs(double *const *q_cons_vf, double *const *q_prim_vf, double *const *rhs_vf,
                    int t_step);

/* Replaces s_{1st,2nd,3rd}_order_tvd_rk(t_step, time_avg)
 * (src/simulation/m_time_steppers.fpp:129-362) for the configured
 * time_stepper: advances q_cons_ts1 in place by one step of size dt.
 * q_cons_ts2 is the second state of RK2/RK3; rhs_vf is interior scratch.
 * q_prim_vf (may be NULL) receives the primitive field of the first stage,
 * which is what the reference's run-time information reads.  The RK stage
 * update is fused into the last direction's sweep; RK3 uses one more ghosted
 * state owned by the library.  Returns 1 without advancing when
 * t_step == t_step_stop (m_time_steppers.fpp:296). */
int mfc_tvd_rk_step(int t_step, double dt, double *const *q_cons_ts1, double *const *q_cons_ts2,
                    double *const *q_prim_vf, double *const *rhs_vf);

/* Replaces the conversion of s_convert_conservative_to_primitive_variables
 * (src/common/m_variables_conversion.fpp:313-375) over the ghosted range. */
int mfc_cons_to_prim(double *const *q_cons_vf, double *const *q_prim_vf);

/* Maximum ICFL over the interior, from a primitive field, as computed by
 * s_write_run_time_information (src/simulation/m_data_output.fpp:197-251).
 * Blocks until the value is on the host.  A cell whose ICFL is NaN (a NaN
 * density, pressure or volume fraction) makes the maximum NaN. */
int mfc_icfl_max(double *const *q_prim_vf, double dt, double *icfl_max);

/* The stability criteria of s_write_run_time_information
 * (src/simulation/m_data_output.fpp:197-263) over the interior, from a
 * primitive field: out[0] = maximum ICFL (as mfc_icfl_max), and in viscous
 * cases out[1] = maximum VCFL and out[2] = minimum cell Reynolds number Rc,
 * with the mixture Reynolds numbers of
 * src/common/m_variables_conversion.fpp:210-223 (0 and 0 when inviscid).
 * Blocks until the values are on the host.  A cell whose criterion is NaN
 * makes that extremum NaN (min and max drop a NaN operand, as written in the
 * reference, so e.g. a NaN velocity alone does not reach ICFL). */
int mfc_stability_criteria(double *const *q_prim_vf, double dt, double *out);

/* CFL time step: dt = cfl / max over cells of the ICFL at dt = 1, evaluated
 * from a conservative field (cons->prim fused).  The reference has no adaptive
 * dt (its dt is an input); this is the same max-reduction inverted. */
int mfc_cfl_dt(double *const *q_cons_vf, double cfl, double *dt_out);

/* mfc_cfl_dt without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight time loops, harnesses that forbid hipMalloc).  Same result as
 * mfc_cfl_dt, bit for bit; the block partials live in a fixed 64 KiB device
 * array (8192 doubles, one copy per GPU, kept until the process ends) in
 * place of the buffer mfc_rhs_init allocates.  The reduction is a
 * grid-stride pass over at most 8192 blocks, so every grid size fits.  Calls
 * that overlap on one GPU (several host threads, or several streams set with
 * mfc_set_stream) share that array and race.  Needs mfc_rhs_init and
 * mfc_rhs_set_grid first, like mfc_cfl_dt; blocks until dt is on the host. */
int mfc_cfl_dt_noalloc(double *const *q_cons_vf, double cfl, double *dt_out);

/* Sizes implied by the initialised case. */
int mfc_sizes(int *sys_size, int *buff_size, long *ghost_len, long *interior_len);

/* Host copy of the WENO coefficient table of direction dir: 27 planes of
 * (s+2b+1) doubles, cell j at offset j+b (plane layout in mfc_rhs.hip). */
int mfc_weno_coefficients(int dir, double *out);

/* Device-memory plumbing for hosts without their own HIP bindings (Fortran). */
int mfc_device_alloc(size_t bytes, void **ptr);
int mfc_device_free(void *ptr);
int mfc_memcpy_h2d(void *dst_device, const void *src_host, size_t bytes);
int mfc_memcpy_d2h(void *dst_host, const void *src_device, size_t bytes);
int mfc_sync(void);
void *mfc_stream(void);
/* Queue all further work on a caller's HIP stream (NULL: the library's own). */
int mfc_set_stream(void *stream);

/* Loads parameters (host, MFC_MLP_NPARAM floats), resets the Adam state and
 * packs the BF16 operand copies used by training and by the learned sweeps. */
int mfc_mlp_init(const float *params);
int mfc_mlp_get(float *params);
/* Gradient (host, MFC_MLP_NPARAM floats) and loss of n device samples. */
int mfc_mlp_gradient(const float *samples, long n, float *grad, float *loss);
/* One Adam step (beta 0.9/0.999) on n device samples; loss before the step. */
int mfc_mlp_train_step(const float *samples, long n, float lr, float *loss);
/* Inference on n device samples: out (device) = 4 planes: w0, w1, w2, value. */
int mfc_mlp_infer(const float *samples, long n, float *out);
/* Training pairs from the plain solver: coarse cell averages of a 2x finer
 * field (one variable, ghosted layouts with b ghost layers; n = 0 for 1D). */
int mfc_ml_restrict(int m, int n, int b, const double *fine, double *coarse);
/* Samples of direction dir (0 x, 1 y) for every coarse interior cell and
 * both of its faces: out (device) has 13 planes of 2(m+1)(n+1) floats, the
 * 12 sample fields and a validity flag (the stencil's largest
 * neighbour difference exceeds 1e-6 of its largest value).  coef27: one cell's
 * WENO5 coefficients of the uniform coarse grid (mfc_weno_coefficients layout). */
int mfc_ml_samples(int dir, int m, int n, int b, const double *coarse, const double *fine, const double *coef27,
                   double eps, float *out);
/* The same for 1D, 2D and 3D fields (p = 0 below 3D): the 3D restriction
 * averages 2x2x2 fine cells, and a 3D sample's target averages the 4 fine
 * faces that tile the coarse face; dir 2 is z.  out has 13 planes of
 * 2(m+1)(n+1)(p+1) floats.  mfc_ml_restrict and mfc_ml_samples are these
 * with p = 0. */
int mfc_ml_restrict3(int m, int n, int p, int b, const double *fine, double *coarse);
int mfc_ml_samples3(int dir, int m, int n, int p, int b, const double *coarse, const double *fine,
                    const double *coef27, double eps, float *out);
const char *mfc_mlp_last_error(void);

/* Per-kernel timing with HIP events recorded on the library's stream around
 * every launch.  Kernel ids: 0 bc_fill, 1 cons_to_prim, 2 sweep_x, 3 sweep_y,
 * 4 sweep_z, 5 icfl, 6 rk1_update, 7 viscous (one direction of a viscous
 * run: face fluxes and finish), 8 sweep_xy (the fused x and y sweep of the
 * contracted FP64 form on uniform grids).  The reference's only timing is the mean
 * CPU time per step (src/simulation/m_time_steppers.fpp:281,352-358). */
int mfc_timing_enable(int on);
int mfc_timing_reset(void);
int mfc_timing_read(int kernel, const char **name, double *total_ms, long *count);

/* Replaces s_finalize_*_module (src/simulation/p_main.fpp:329-342). */
int mfc_rhs_finalize(void);

const char *mfc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
