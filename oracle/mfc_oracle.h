/*
 * mfc_oracle.h -- CPU FP64 restatement of MicroFC's simulation right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product (microfc_amd/) never links it.
 *
 * It restates, in plain C compiled with -ffp-contract=off, the arithmetic of
 *   src/simulation/m_rhs.fpp            (s_compute_rhs, ghost-cell fill, flux divergence)
 *   src/simulation/m_weno.fpp           (WENO coefficients, WENO1/3/5)
 *   src/simulation/m_riemann_solvers.fpp(HLLC, Riemann-state buffers)
 *   src/common/m_variables_conversion.fpp (cons->prim, mixture EOS)
 *   src/simulation/m_time_steppers.fpp  (TVD RK1/2/3)
 *   src/simulation/m_start_up.fpp       (grid ghost buffers)
 *   src/simulation/m_data_output.fpp    (ICFL)
 *   src/simulation/m_viscous.fpp        (velocity gradients, weno_Re_flux = F and T)
 * keeping the reference's evaluation order operation by operation.  The 3D
 * case is the dimension-split generalisation (the reference is 1D/2D only).
 *
 * Field layout (shared with the HIP library): variable v of a ghosted field is
 * the contiguous block  f + v*ng  with ng = nxg*nyg*nzg and element (i,j,k),
 * i in [-b, m+b] (j, k likewise when n, p > 0, else 0) at
 *   (i+b) + nxg*((j+bj) + nyg*(k+bk)).
 * x is fastest, matching the Fortran column-major sf(-b:m+b, -b:n+b[, -b:p+b]).
 * The rhs field is interior only: (m+1)(n+1)(p+1) per variable.
 */
#ifndef MFC_ORACLE_H
#define MFC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NUM_FLUIDS_MAX 10

typedef struct {
    int m, n, p;                 /* last interior index per direction; n = p = 0 collapse */
    int num_fluids;              /* 1..10 */
    int weno_order;              /* 1, 3 or 5 */
    int time_stepper;            /* 1, 2 or 3 */
    int t_step_stop;             /* s_compute_rhs returns early when t_step == t_step_stop */
    double weno_eps;
    double gammas[ORC_NUM_FLUIDS_MAX];   /* fluid_pp(i)%gamma  = 1/(gamma_i - 1) */
    double pi_infs[ORC_NUM_FLUIDS_MAX];  /* fluid_pp(i)%pi_inf = gamma_i*pi_inf_i/(gamma_i - 1) */
    int bc_beg[3], bc_end[3];    /* bc_x%beg, bc_x%end, bc_y..., bc_z... */
    /* fluid_pp(i)%Re(1:2): shear and volume Reynolds numbers; a value <= 0
     * leaves fluid i out of that viscous term (m_global_parameters.fpp:314-338).
     * Any positive entry makes the case viscous (1D and 2D only). */
    double Re[ORC_NUM_FLUIDS_MAX][2];
    /* weno_Re_flux: velocity gradients from WENO-reconstructed face values
     * by the divergence theorem, then WENO-reconstructed to the faces
     * (nonzero), or finite differences (0). */
    int weno_Re_flux;
} orc_case;

/* Host threads of the pencil and update loops (OpenMP builds; the results do
 * not depend on the count) and the current count. */
void orc_set_threads(int n);
int orc_max_threads(void);

/* Derived sizes (m_global_parameters.fpp:285-381). */
int orc_num_dims(const orc_case *c);
int orc_sys_size(const orc_case *c);
int orc_buff_size(const orc_case *c);
/* Number of doubles per variable of a ghosted field. */
long orc_ghost_len(const orc_case *c);

/* Grid ghost buffers (m_start_up.fpp:517-655).  cb has s+2b+2 entries,
 * cb[0] = s_cb(-1-b); on entry cb(-1:s) (entries b..s+b+1) hold the interior
 * cell boundaries.  On return cb, ds (s+2b+1 entries, ds[0] = ds(-b)) and
 * cc (same) are filled including ghosts. */
int orc_grid_buffers(int s, int b, int bc_beg, int bc_end, double *cb, double *ds, double *cc);

/* WENO coefficient tables for one direction (m_weno.fpp:168-363).  cb as
 * above.  Output has 27 planes of ncell = s+2b+1 doubles (cell index
 * j = -b..s+b at offset j+b); only j in [-b+polyn, s+b-polyn] are set, the
 * rest are zero.  Plane layout:
 *   0..5   poly_coef_cbL(j,r,q)  at 0 + 2r + q   (r = 0..2, q = 0..1)
 *   6..11  poly_coef_cbR(j,r,q)  at 6 + 2r + q
 *   12..14 d_cbL(r,j), 15..17 d_cbR(r,j)
 *   18..26 beta_coef(j,r,q)      at 18 + 3r + q  (q = 0..2)
 * For weno_order 3 only r <= 1, q = 0 entries are used. */
int orc_weno_coefficients(int s, int b, int weno_order, const double *cb, double *out);

/* Conservative -> primitive over the whole ghosted range
 * (m_variables_conversion.fpp:313-375). */
int orc_cons_to_prim(const orc_case *c, const double *q_cons, double *q_prim);

/* Ghost-cell fill of a conservative field (m_rhs.fpp:686-908 plus the z
 * analogue).  bc codes <= -3 extrapolate, -2 reflect, -1 periodic. */
int orc_bc_fill(const orc_case *c, double *q);

/* s_compute_rhs (m_rhs.fpp:405-679): copies q_cons into a private buffer,
 * fills its ghosts, converts to primitive, and (unless t_step == t_step_stop)
 * accumulates the dimension-split flux divergence and alpha*div(u) source
 * into rhs.  q_prim (may be NULL) receives the ghosted primitive field.
 * cb[d], ds[d]: ghosted grid arrays of direction d (NULL beyond num_dims). */
int orc_compute_rhs(const orc_case *c, const double *const cb[3], const double *const ds[3],
                    const double *q_cons, double *q_prim, double *rhs, int t_step);

/* One time step of s_{1st,2nd,3rd}_order_tvd_rk (m_time_steppers.fpp:129-362).
 * q1 is advanced in place; q2 is the stage buffer (unused for RK1); rhs is
 * interior scratch.  Returns 1 (and leaves q1 unchanged) when t_step ==
 * t_step_stop, like the reference's early return. */
int orc_tvd_rk_step(const orc_case *c, const double *const cb[3], const double *const ds[3],
                    double dt, double *q1, double *q2, double *q_prim, double *rhs, int t_step);

/* Maximum ICFL over interior cells (m_data_output.fpp:197-251) from a ghosted
 * primitive field.  NaN in any cell gives NaN. */
int orc_icfl_max(const orc_case *c, const double *const ds[3], const double *q_prim, double dt,
                 double *icfl_max);
int orc_stability(const orc_case *c, const double *const ds[3], const double *q_prim, double dt, double out[3]);

/* Single-face kernels exposed for unit tests. */
/* WENO5/3/1 of one variable at one cell from the 5-point stencil v[0..4]
 * (cells j-2..j+2) with that cell's 27 coefficients; writes vL (face j-1/2)
 * and vR (face j+1/2) (m_weno.fpp:391-540). */
void orc_weno_cell(int weno_order, double weno_eps, const double *coef27, const double *v,
                   double *vL, double *vR);
/* Learned WENO5 weights (no reference counterpart; they replace the
 * nonlinear weights of m_weno.fpp:508-510,526-528).  params: the 2128 FP32
 * parameters of the 5 -> 32 -> 32 -> 3 MLP in the layout of
 * include/mfc_rhs.h, or NULL to return to the plain weights.  While set,
 * orc_compute_rhs (and the RK step) use them for WENO5, in FP64 arithmetic
 * with BF16 round-to-nearest-even applied where the GPU feeds its matrix
 * cores: features, weights and both hidden activations.  The features are
 * evaluated in FP32 on the FP32-rounded stencil, as the model defines them.
 * Not thread safe. */
void orc_set_learned(const float *params);
/* The learned reconstruction of one variable at one WENO5 cell.  kind: bit 0
 * = the variable is >= 0 where its plain value is (partial densities, volume
 * fractions), bit 1 = <= 1 where the plain value is (volume fractions), bit 2
 * = > 0 where the plain value is (pressure).  Writes vL, vR as
 * orc_weno_cell. */
void orc_weno5_learned_cell(double weno_eps, const double *coef27, const double *v, int kind, double *vL,
                            double *vR);
/* HLLC flux at one face (m_riemann_solvers.fpp:134-325).  qL, qR: primitive
 * states (sys_size each) left and right of the face.  dir: 0, 1, 2.  flux
 * receives sys_size values, *ustar the contact velocity. */
void orc_hllc_face(const orc_case *c, int dir, const double *qL, const double *qR, double *flux,
                   double *ustar);

#ifdef __cplusplus
}
#endif
#endif
