/*
 * amr_oracle.h -- CPU oracle for the finite-volume Euler hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU algorithm (ackirby88/amr-demo, src/solver + src/physics).  It exists to
 * check the HIP implementation; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * or calls it.
 *
 * Layout is the reference one: one patch per quadrant, AoS within the patch,
 * soln[c*4*n*n + f + 4*(i + n*j)] (Fortran Q(4,n,n)), i,j 0-based here.
 * cell_info[3*c + {type, level, soln offset}], face_info[11*f + ...] and
 * geom[12*c] (4 corners x xyz, z-order) as built by
 * src/amr/src/amr_external.c:337-508.
 *
 * Pinning (what checks each function; tests/test_oracle_*.py,
 * tests/test_walls.py; golden vectors come from the reference's own
 * hot-path sources compiled into oracle/_ref by tests/golden/make_golden.py):
 *   orc_flux_roe_2d ........ golden vectors, consistency F(U,U)=F_phys(U)n,
 *                            supersonic upwinding, x<->y mirror symmetry
 *   orc_flux_bc_2d, orc_bc_2d, orc_flux_native_2d
 *                            golden vectors (bc_2d + flux_bc_2d, whose physical
 *                            flux is flux_native_2d); slip wall pure-pressure
 *   orc_residual_*_2d, orc_bilinear_node_2d, orc_rhs_structured(_range)
 *                            golden vectors of physics_rhs_structured on
 *                            uniform, hanging and boundary forests (n = 2..16);
 *                            free-stream preservation; discrete conservation
 *   orc_residual_bc_ext .... wall types 3-5, project additions with no
 *                            reference counterpart: closed forms (a uniform
 *                            prescribed state is steady; over gas at rest only
 *                            the cells behind the shock line / before x0 change)
 *   orc_time_step_2d, orc_time_step
 *                            golden physics_time_step minima (bit for bit);
 *                            closed form on uniform states
 *   orc_rk_step, orc_rk_generic, orc_run_steps(_threaded)
 *                            closed-form stability polynomials on L(Q)=lambda Q,
 *                            convergence orders on a rotation system, LSERK45
 *                            local order 5 (the reference rk_* needs mpi.h and
 *                            is not built)
 *   orc_initial_condition_2d golden vectors from initial_condition_2d
 *   orc_refine_2d, orc_coarsen_2d, orc_tag_feature
 *                            golden vectors of solver_refine_operator_2d,
 *                            solver_coarsen_operator_2d, solver_tag_feature
 *   orc_bf16 ... orc_mlp_grad (tagger_oracle.c) have no reference counterpart:
 *                            pinned by the mathematics in tests/test_tagger_oracle.py
 */
#ifndef AMR_ORACLE_H
#define AMR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Physics settings the reference keeps in header globals
 * (src/physics/src/physics_var_physics.h:21-49,
 *  src/physics/src/physics_var_simulation.h:20-37,
 *  src/physics/src/physics_var_interface.h:20-39). */
typedef struct {
    int    nelem_subgrid;        /* n */
    int    time_scheme;          /* 1..5 */
    int    bc[6];                /* per side: 0 copy, 1 slip, 2 no-slip */
    double cfl;                  /* effective CFL (already 0.25*cfl*scheme) */
    double dt;                   /* input dt cap */
    double coarse_grid_h;        /* h_base = mesh_scale */
    /* wall types 3-5 (project additions, no reference counterpart; see
     * include/amr_hip.h): boundary states, shock line, time of the residual */
    double bstate[2][4];
    double shock_x0, shock_angle, shock_speed;
    double time;
} orc_params;

/* a1: src/solver/src/flux_2d_module.F90:63-182 */
void orc_flux_roe_2d(const double ql[4], const double qr[4], const double norm[2], double flux[4]);
/* src/solver/src/flux_2d_module.F90:32-60 ; f is Fortran flux(4,2) */
void orc_flux_native_2d(const double q[4], double f[8]);
/* src/solver/src/flux_2d_module.F90:13-29 */
void orc_flux_bc_2d(const double q[4], const double normal[2], double flux[4]);
/* a6: src/solver/src/bc_2d_module.F90:14-75 */
void orc_bc_2d(int bc_type, const double normal[2], const double q[4], double qbc[4]);
/* a7: src/solver/src/bilinear_node_2d_module.F90:14-36 */
void orc_bilinear_node_2d(const double geom[12], const double nat[2], double out[2]);

/* a2: src/solver/src/solver_residual_subgrid_2d_module.F90:32-164 */
void orc_residual_subgrid_2d(int nfields, int n, int type, int level,
                             const double geom[12], const double *Q, double *res);
/* a3: src/solver/src/solver_residual_face_full_structured_2d_module.F90:35-109.
 * Contributions go only to indices in [lo, hi) -- the reference filter is
 * ind < dof, i.e. lo = 0, hi = dof. */
void orc_residual_face_full_2d(int nfields, int n, int side, int level,
                               int ind_l, int ind_r, int lo, int hi, double h_base,
                               const double *ql, const double *qr, double *res);
/* a4: src/solver/src/solver_residual_face_full_hang_structured_2d.F90:34-128 */
void orc_residual_face_full_hang_2d(int nfields, int n, int side, int level_l, int level_r,
                                    int ind_l, int ind_r1, int ind_r2, int lo, int hi,
                                    double h_base, const double *ql, const double *qr1,
                                    const double *qr2, double *res);
/* a5: src/solver/src/solver_residual_face_hang_full_structured_2d.F90:34-128 */
void orc_residual_face_hang_full_2d(int nfields, int n, int side, int level_l, int level_r,
                                    int ind_l1, int ind_l2, int ind_r, int lo, int hi,
                                    double h_base, const double *ql1, const double *ql2,
                                    const double *qr, double *res);
/* a6: src/solver/src/solver_residual_bc_structured_2d_module.F90:31-88 */
void orc_residual_bc_2d(int nfields, int n, int side, int type, int level, int bc,
                        double h_base, const double *soln, double *res);
/* Wall types 3-5 on side `side` of a quadrant with corners geom (project
 * additions: prescribed inflow state, post-shock / slip bottom wall, moving
 * shock along the top; face centres from the quadrant's lower-left corner).
 * Boundary flux = flux_bc_2d of the boundary state, as the reference's walls. */
void orc_residual_bc_ext(const orc_params *p, int n, int side, int level, const double geom[12],
                         const double *soln, double *res);
/* a11: src/solver/src/solver_time_step_2d_module.F90:26-122 */
void orc_time_step_2d(int nfields, int n, int type, int level, const double geom[12],
                      const double *Q, double cfl, double *dt);

/* a8: src/physics/src/physics_rhs_structured.c:19-283 (dim = 2, nfields = 4).
 * Writes rhs[0:dof). */
void orc_rhs_structured(const orc_params *p, int ncell, const int *cell_info,
                        const double *geom, int nface, const int *face_info, int dof,
                        const double *soln, double *rhs);
/* Same operator restricted to the cells [c0, c1) a worker owns, the way an MPI
 * rank of the reference evaluates its part (faces to cells outside the range
 * are evaluated, their contributions dropped, as the ind < dof filter does). */
void orc_rhs_structured_range(const orc_params *p, int ncell, const int *cell_info,
                              const double *geom, int nface, const int *face_info,
                              const double *soln, double *rhs, int c0, int c1);
/* a11 driver: src/physics/src/physics_time_step.c:16-49 (running min into *dt) */
void orc_time_step(const orc_params *p, int ncell, const int *cell_info,
                   const double *geom, const double *soln, double *dt);
/* a10: src/physics/src/physics_norm_residual.c:15-56 (sum of squares, no sqrt) */
double orc_norm_L2_residual(int n, int ncell, const int *cell_info, const double *res);

/* a9: one step of scheme p->time_scheme (src/physics/src/rk_time_integrators.c:22-540)
 * in the reference's own k-form.  work must hold 5*dof doubles.  Returns the
 * sqrt of the last stage's residual sum of squares (the reference resnorm). */
double orc_rk_step(const orc_params *p, int ncell, const int *cell_info, const double *geom,
                   int nface, const int *face_info, int dof, double dt, double *Q,
                   double *work);
/* The same stage arithmetic applied to a caller-supplied operator
 * k = L(Q) (used to pin the schemes on L(Q) = lambda Q in closed form). */
double orc_rk_generic(int scheme, int dof, double dt, double *Q, double *work,
                      void (*rhs)(void *ctx, double *Q, double *k), void *ctx);
/* a12: the step loop of src/physics/src/physics_runge_kutta_interface.c:185-239:
 * per step dt = min(input dt, CFL) (skipped for RK2, :142-144), then one RK step.
 * Runs nsteps steps; returns last resnorm; *dt_last gets the last dt. */
double orc_run_steps(const orc_params *p, int ncell, const int *cell_info, const double *geom,
                     int nface, const int *face_info, int dof, int nsteps, double *Q,
                     double *dt_last);
/* The same step loop run by nthreads workers, each owning a contiguous range
 * of quadrants (weight-1 partition), synchronised per stage: the reference's
 * MPI decomposition restated with shared memory.  Returns last resnorm. */
double orc_run_steps_threaded(const orc_params *p, int ncell, const int *cell_info,
                              const double *geom, int nface, const int *face_info, int dof,
                              int nsteps, double *Q, double *dt_last, int nthreads);

/* src/solver/src/initial_condition_2d_module.F90:12-131 with the vortex
 * strength beta_L and centre (xo, yo) as parameters (reference: 3, origin). */
void orc_initial_condition_2d(int ic, int n, const double geom[12], double *Q,
                              double mach, double alpha, double beta, double gamma,
                              double density, double pressure,
                              double beta_L, double xo, double yo);

/* src/solver/src/solver_refine_operator_2d.F90:28-55 and
 * src/solver/src/solver_coarsen_operator_2d.F90:29-53 */
void orc_refine_2d(int n, const double *parent, double *c1, double *c2, double *c3, double *c4);
void orc_coarsen_2d(int n, const double *c1, const double *c2, const double *c3,
                    const double *c4, double *parent);
/* src/solver/src/tagging_methods_module.F90:14-71 via solver_tag_feature.F90:9-52 */
int orc_tag_feature(int ntags, const int *methods, const double *tol, int n, const double *Q);

/* ---- learned tagger (no reference counterpart; tagger_oracle.c) --------- */
/* The reference has no learned tagger.  These restate the library's own
 * definitions (include/amr_tagger.h) on the CPU so the GPU kernels have a
 * checker; the MLP backward pass is pinned by finite differences and by an
 * independent autograd implementation in the tests. */

/* FP32 -> BF16 bits, round to nearest even. */
unsigned short orc_bf16(float x);
float orc_bf16_to_float(unsigned short h);
/* Features of every quadrant of a mesh given in the reference's arrays
 * (face_info as hpc_amr_evolve_solver receives it, soln Q(4,8,8) per
 * quadrant); X[ncell][384] BF16 bits. */
int orc_tagger_features(int n, int ncell, int nface, const int *face_info, const double *soln,
                        unsigned short *X);
/* Logits [n][3] of the BF16 network (exact = 0) or of the same network with
 * no rounding of weights and activations (exact = 1). */
void orc_mlp_forward(const float *P, const unsigned short *X, int n, int exact, double *logits);
/* Class-weighted mean cross-entropy and its gradient (AMR_MLP_NPARAM
 * doubles).  exact = 0 rounds as the GPU pass does; exact = 1 is the smooth
 * function of the FP32 parameters. */
double orc_mlp_grad(const float *P, const unsigned short *X, const int *y, int n, const float cw[3],
                    int exact, double *grad);

#ifdef __cplusplus
}
#endif
#endif
