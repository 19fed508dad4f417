/*
 * amr_hip.h -- C ABI of the MI355X implementation of the demo's finite-volume
 * Euler hot path (patch residual, RK stage update, CFL time step).
 *
 * Two layers:
 *
 * 1. Drop-in physics entry points: the GPU seam physics_go_gpu that the
 *    reference's physics_evolve_interface calls under _GPU_ (its header,
 *    physics_go_gpu.h, is in this directory), and GPU counterparts of the
 *    reference's residual and time-step functions.  Same argument lists and
 *    meaning; the only deviation is the communicator: one GPU owns the whole
 *    forest, so the MPI_Comm argument is an opaque pointer that is ignored.
 *    Nothing here re-declares a function the reference's physics module
 *    defines, so this header can sit beside the reference's headers.  The
 *    library also exports replacements of hpc_amr_initialize_solver and
 *    hpc_amr_evolve_solver for drivers that link it instead of the
 *    reference's physics module (amr_hip_dropin.h).
 *
 * 2. A device-resident context (amrgpu_*), with no reference counterpart,
 *    that keeps the state in HBM across steps, runs whole evolve intervals
 *    (stages launched directly, or replayed from captured HIP graphs) and
 *    exposes the timing hooks the benchmark needs.
 *    It also carries the batched mode: several independent forests
 *    ("instances") advanced by the same launches.
 *
 * All arrays are host pointers unless stated otherwise.  Return codes: 0 ok,
 * negative on argument/HIP/mesh errors, AMRGPU_ERR_NAN (3, the reference's
 * MPI_Abort code) when the residual norm is NaN.
 */
#ifndef AMR_HIP_H
#define AMR_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define AMRGPU_OK        0
#define AMRGPU_ERR_ARG  -1
#define AMRGPU_ERR_HIP  -2
#define AMRGPU_ERR_MESH -3
#define AMRGPU_ERR_NAN   3

/* Physics settings: the reference keeps these in header globals filled from
 * input.physics (src/physics/src/physics_initialize.c:61-197, 312-316). */
typedef struct {
    int    dim, nfields, nelem_subgrid;
    int    extern_cell_size, extern_face_size;
    int    time_scheme;            /* 1 RK1, 2 SSP-RK2, 3 SSP-RK3, 4 RK4, 5 LSERK45 */
    int    time_steps;
    double dt;                     /* input dt (cap) */
    int    checkpoint_interval, visualization_interval, regrid_interval;
    int    bc[6];
    double cfl;                    /* effective CFL = 0.25*input*time_scheme */
    double coarse_grid_h;          /* mesh_scale */
    double t_end;                  /* physical-time stop (project addition; <= 0: off) */
    int    verbose;
} physics_gpu_params_t;

/* Programmatic access to the same settings (no reference counterpart; the
 * reference writes the globals directly). */
void physics_gpu_set_params(const physics_gpu_params_t *p);
void physics_gpu_get_params(physics_gpu_params_t *p);
/* Physical time accumulated by physics_go_gpu since the last set_params. */
double physics_gpu_time(void);

/* The GPU seam: physics_go_gpu has the argument list of physics_go
 * (src/physics/src/physics_go.h:59-84), as physics_evolve_interface calls it
 * under _GPU_ (src/physics/src/physics_evolve_interface.c:42-46).  Uploads the
 * interval's arrays, runs steps until time_steps or an interval flag
 * (src/physics/src/physics_runge_kutta_interface.c:195-230) on the GPU, and
 * writes soln back in place.  Returns AMRGPU_ERR_NAN where the reference calls
 * MPI_Abort(comm, 3). */
int physics_go_gpu(int *dim, int *unstructured, int *nfields, int *nelem_subgrid,
                   int *ncell, int *nghost, int *nface, int *dof,
                   int *cell_info, int *face_info, int *nsend, int *nrecv,
                   int *send, int *recv, int *sizes, double *geom, double *soln,
                   int *istep, int *regrid, int *visualize, int *checkpoint,
                   int *evolve_solution, void *mpicomm_ext,
                   double *total_time, double *compute_time, double *residual_time);

/* Replaces physics_rhs_structured (src/physics/src/physics_rhs_structured.c:19-22):
 * rhs[0:dof) = -R/V of soln, evaluated on the GPU (dim 2, nfields 4). */
int physics_rhs_structured_gpu(int *dim, int *nfields, int *nelem_subgrid, int *ncell,
                               int *cell_info, double *geom, int *nface, int *face_info,
                               int *dof, double *soln, double *rhs);

/* Replaces physics_time_step (src/physics/src/physics_time_step.c:16-19):
 * *dt = min(*dt, min over cells of cfl*sqrt(V)/(|u|+c)), on the GPU. */
int physics_time_step_gpu(int *dim, int *nfields, int *nelem_subgrid, int *ncell,
                          int *cell_info, double *geom, double *soln, double *dt);

/* ------------------------------------------------------------------------ */
/* Device-resident context                                                  */
/* ------------------------------------------------------------------------ */
typedef struct amrgpu_ctx amrgpu_ctx;

/* Per-instance settings (one forest = one instance). */
/* Wall types per domain side.  0-2 are the reference's (bc_2d_module.F90:
 * 44-72); 3-5 are project additions for the double-Mach reflection, whose
 * boundary states depend on position and time (boundary flux = physical flux
 * of the boundary state with the outward normal, as flux_bc_2d):
 *   3 inflow      the prescribed state bstate[0]
 *   4 shock wall  bstate[0] for x < shock_x0, slip wall beyond
 *   5 shock top   bstate[0] behind the moving shock line
 *                 x_s(y, t) = shock_x0 + y / tan(angle) + speed / sin(angle) t,
 *                 bstate[1] ahead of it
 * Types 4 and 5 need the quadrant coordinates (amrgpu_setup's geom, or
 * amrgpu_set_geometry after a remesh); t is the instance time at the start
 * of the step. */
enum { AMRGPU_BC_COPY = 0, AMRGPU_BC_SLIP = 1, AMRGPU_BC_NOSLIP = 2, AMRGPU_BC_INFLOW = 3,
       AMRGPU_BC_SHOCK_WALL = 4, AMRGPU_BC_SHOCK_TOP = 5 };

typedef struct {
    double h_base;      /* coarse_grid_h of this forest */
    int    bc[4];       /* xlo, xhi, ylo, yhi */
    double cfl;         /* effective CFL */
    double dt_cap;      /* input dt */
    double t_end;       /* physical-time stop, <= 0: none */
    double bstate[2][4];/* boundary states (conservative) of types 3-5 */
    double shock_x0, shock_angle, shock_speed;   /* types 4-5: degrees, speed */
} amrgpu_inst_t;

amrgpu_ctx *amrgpu_create(void);
void amrgpu_destroy(amrgpu_ctx *c);
const char *amrgpu_last_error(void);

/* Upload a mesh + state.  cell_info/face_info/soln use the reference layout
 * (cell ids and soln offsets global over all instances); cell_inst[ncell]
 * gives each quadrant's instance (NULL: all instance 0).  geom is accepted for
 * interface symmetry: the structured path derives sub-cell sizes from the
 * level, as the reference face kernels do. */
int amrgpu_setup(amrgpu_ctx *c, int nelem_subgrid, int time_scheme,
                 int ninst, const amrgpu_inst_t *inst,
                 int ncell, const int *cell_info, const int *cell_inst,
                 int nface, const int *face_info, const double *geom, const double *soln);
int amrgpu_upload_soln(amrgpu_ctx *c, const double *soln);
int amrgpu_download_soln(amrgpu_ctx *c, double *soln);
/* The state of quadrants [first, first + count) only, in the same layout
 * (soln holds 4 n^2 count doubles). */
int amrgpu_download_quads(amrgpu_ctx *c, int first, int count, double *soln);
/* Project-defined density-gradient indicator of the current state, one
 * value per quadrant (host out); tagging is ind > tolerance. */
int amrgpu_indicator(amrgpu_ctx *c, double *ind);
/* The reference's feature tags of the current state, one flag per quadrant
 * (host out): solver_tag_feature (src/solver/src/solver_tag_feature.F90:9-52)
 * over ntags methods, tag_density_below (1: some sub-cell density <=
 * tols[k]) and tag_density_above (2: some sub-cell density >= tols[k]) of
 * src/solver/src/tagging_methods_module.F90:14-71; tags[q] = 1 if any method
 * fires, else 0 (other method numbers tag nothing, as in the reference).
 * 0 <= ntags <= AMRGPU_MAX_FEATURE_TAGS. */
#define AMRGPU_MAX_FEATURE_TAGS 8
int amrgpu_feature_tags(amrgpu_ctx *c, int ntags, const int *methods, const double *tols, int *tags);
/* Replace the mesh after a regrid without leaving the device: src[ncell]
 * from amr_forest_track_src (copy of an old quadrant, or mean of an old
 * family); cell_info/face_info/cell_inst describe the new mesh.  Instance
 * settings and time are kept; an instance the new mesh leaves without
 * quadrants stops advancing (its time stays where it is). */
int amrgpu_remesh(amrgpu_ctx *c, int ncell, const int *cell_info, const int *cell_inst, int nface,
                  const int *face_info, const int *src);
/* The same remesh given the device mesh table directly (levels, instances
 * and nbr as amr_forest_build_nbr makes them) instead of the reference's
 * cell_info / face_info; every neighbour id is checked to lie in the new
 * mesh.  No counterpart in the reference: it spares a driver that owns the
 * forest the reference-format arrays on every regrid. */
int amrgpu_remesh_nbr(amrgpu_ctx *c, int ncell, const int *level, const int *cell_inst, const int *nbr,
                      const int *src);
/* Quadrant coordinates of the current mesh (geom as amrgpu_setup) for the
 * position-dependent wall types; a remesh drops them until this is called. */
int amrgpu_set_geometry(amrgpu_ctx *c, const double *geom);
/* The same from the quadrants' lower-left corners (org[2k], org[2k + 1]). */
int amrgpu_set_origins(amrgpu_ctx *c, const double *org);
/* The device mesh table that amrgpu_setup / amrgpu_remesh derive from the
 * reference's face_info (nbr[8 ncell], format of amr_forest_build_nbr);
 * host only, no device needed.  Rejects inconsistent face tables. */
int amrgpu_nbr_from_faces(int ncell, int nface, const int *face_info, int *nbr);
/* Residual rhs = -R/V of the current state (reference layout, host out). */
int amrgpu_rhs(amrgpu_ctx *c, double *rhs);
/* CFL minimum of the current state per instance (no dt cap). */
int amrgpu_cfl(amrgpu_ctx *c, double *dt_inst);
/* Advance nsteps steps.  use_graph: replay one captured step when graph
 * replay is enabled (amrgpu_set_graphs), else launch the stages directly.
 * On return resnorm (if non-NULL) holds sqrt(sum rhs^2) of the last stage. */
int amrgpu_run(amrgpu_ctx *c, int nsteps, int use_graph, double *resnorm);
/* Step launch mode: 0 direct stream launches (the default: on MI355X a
 * dependent kernel costs less launched from the stream than replayed from a
 * captured graph), 1 replay of one captured HIP graph per step.  The
 * environment variable AMRGPU_GRAPHS=0/1 sets the initial mode;
 * amrgpu_autotune times both for the chosen tile and keeps the faster. */
int amrgpu_set_graphs(amrgpu_ctx *c, int enable);
int amrgpu_get_graphs(const amrgpu_ctx *c);
/* Per-instance physical time and last dt. */
int amrgpu_get_time(amrgpu_ctx *c, double *t_inst, double *dt_inst);
int amrgpu_set_time(amrgpu_ctx *c, const double *t_inst);

/* Benchmark hooks: run nsteps steps in the current launch mode, timing the
 * whole region with HIP events -> *ms_total; then nsteps more steps launched
 * directly with events around every stage kernel -> avg_stage_ms[s] for
 * s < nstages. */
int amrgpu_bench(amrgpu_ctx *c, int nsteps, double *ms_total, double *avg_stage_ms,
                 int *nstages);
int amrgpu_num_cells(const amrgpu_ctx *c);   /* sub-grid cells */
/* Tile shape of the stage kernel: patches per workgroup, the minimum waves
 * per SIMD reserved by register allocation, and the face-evaluation design
 * (0: face task list, 1: paired +x/+y faces per cell).  amrgpu_autotune
 * times every compiled variant for the mesh's patch size over nsteps steps
 * with direct launches (state and time are restored afterwards), keeps the
 * fastest and reports each variant's ms per step, then sets the launch mode
 * (amrgpu_set_graphs) that is faster for that tile. */
int amrgpu_set_tile(amrgpu_ctx *c, int ppb, int minw, int design);
/* The compiled variants for patch size n (returns their number). */
int amrgpu_tile_variants(int n, int *ppb, int *minw, int *design, int max);
int amrgpu_get_tile(amrgpu_ctx *c, int *ppb, int *minw, int *design);
int amrgpu_autotune(amrgpu_ctx *c, int nsteps, int *nvar, int *ppb, int *minw, int *design,
                    double *ms_per_step);
int amrgpu_num_quads(const amrgpu_ctx *c);
/* Read-only view of the device-resident mesh and state for kernels outside
 * the library's stage path (the learned tagger): state U[quad][field][j][i]
 * (FP64, n x n patches), neighbour table nbr[quad][side][2] (side -x,+x,-y,+y;
 * a < 0 boundary, b < 0 one index-aligned neighbour, else two finer ones in
 * increasing coordinate), and the stream that orders work on them. */
int amrgpu_device_state(amrgpu_ctx *c, const double **U, const int **nbr, int *nquad, int *n,
                        void **stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the per-instance CFL minimum of amrgpu_cfl, for a state
 * the caller keeps on the device, with no allocation of any kind.  Inputs:
 * d_U the state in the context layout U[quad][field][j][i] (FP64, n x n
 * patches, nquad quadrants), d_level / d_inst each quadrant's level and
 * instance (int32, device), h_base[ninst] / cfl[ninst] per instance (host,
 * passed by value); output d_dt[ninst] (device): min over the instance's
 * cells of cfl sqrt(V) / (|u| + c), bit for bit amrgpu_cfl's value.  The
 * work goes on `stream` (NULL: the default stream) and is not waited for.
 * Scratch is one static device array of the library (instance table + CFL
 * sub-slots, about 130 KB, held by each GPU until the process ends); it
 * keeps the last call's contents and is re-initialised by every call, so
 * calls that overlap on one GPU race.  1 <= ninst <= AMRGPU_NOALLOC_MAX_INST,
 * else AMRGPU_ERR_ARG. */
#define AMRGPU_NOALLOC_MAX_INST 64
int amrgpu_cfl_noalloc(int nelem_subgrid, int nquad, int ninst, const double *h_base, const double *cfl,
                       const int *d_level, const int *d_inst, const double *d_U, double *d_dt, void *stream);

/* Device-side helpers for the HBM roofline probe: copy nbytes between two
 * device buffers allocated by the library, timed with events. */
int amrgpu_copy_probe(long long nbytes, int reps, double *gbps);

#ifdef __cplusplus
}
#endif
#endif
