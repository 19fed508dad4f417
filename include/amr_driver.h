/*
 * amr_driver.h -- host AMR driver pieces around the GPU physics path:
 * initial conditions, the density-gradient indicator and the regrid
 * sequence of src/amr/src/amr_regrid.c:15-81 on the host forest.
 *
 * Initial conditions: free stream and the isentropic vortex follow
 * src/solver/src/initial_condition_2d_module.F90:12-131 (vortex strength
 * and centre as parameters); the four-quadrant Riemann problem and the
 * double-Mach reflection are project additions the configs need.
 *
 * The density-gradient indicator is project-defined (the reference tags by
 * density thresholds only, src/solver/src/tagging_methods_module.F90:14-71):
 * a quadrant is tagged when max |rho_a - rho_b| / (0.5 (rho_a + rho_b)) over
 * adjacent sub-cell pairs of its patch and across its four faces (the same
 * index-aligned pairing the residual uses) exceeds the tolerance.
 */
#ifndef AMR_DRIVER_H
#define AMR_DRIVER_H

#include "amr_forest.h"
#include "amr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { AMR_IC_FREESTREAM = 0, AMR_IC_VORTEX = 1, AMR_IC_RIEMANN4 = 2, AMR_IC_DMR = 3 };

typedef struct {
    int kind;
    /* free stream / vortex (reference inputs) */
    double mach, alpha, beta, gamma, density, pressure;
    double vortex_strength, xo, yo;
    /* four-quadrant Riemann: split point and primitive states (rho,u,v,p)
     * of quadrants 0:(x<xs,y<ys) 1:(x>xs,y<ys) 2:(x<xs,y>ys) 3:(x>xs,y>ys) */
    double xs, ys;
    double state[4][4];
    /* double-Mach reflection: shock through (x0, 0) at angle (deg) to the
     * x axis, post-shock state left of it, pre-shock (rho_pre, p_pre at rest)
     * right of it, shock Mach number */
    double dmr_x0, dmr_angle, dmr_mach, dmr_rho, dmr_p;
} amr_ic_t;

/* Fill one patch Q(4,n,n) from the corner geometry (bilinear centres). */
void amr_ic_patch(const amr_ic_t *ic, int n, const double geom[12], double *Q);
/* Fill every quadrant of the forest. */
void amr_ic_apply(amr_forest_t *f, const amr_ic_t *ic);
/* Exact vortex density at (x, y, t) on a periodic box [x0,x0+Lx] x [y0,y0+Ly]
 * (nearest periodic image of the advected centre). */
double amr_ic_vortex_density(const amr_ic_t *ic, double x, double y, double t,
                             double x0, double Lx, double y0, double Ly);

/* Project-defined density-gradient indicator per quadrant (see above);
 * ind[k] receives the indicator value.  On a forest that is not 2:1
 * balanced (between the passes of amr_regrid_initial) the sides facing a
 * neighbour two or more levels away contribute no pairs. */
void amr_indicator_gradient(amr_forest_t *f, double *ind);

/* Regrid in the order of regrid_solution (src/amr/src/amr_regrid.c:15-81):
 * reset tags, tag by indicator > tol and refine tagged quadrants one level
 * (below max_level), coarsen untagged families (unless initial), 2:1 balance.
 * Returns the new number of quadrants, or -1 on failure (as
 * amr_regrid_tags). */
int amr_regrid(amr_forest_t *f, double tol, int allow_coarsening);
/* The same regrid driven by given per-quadrant tags: 1 refine, 0 may be
 * coarsened with its family, any other value keeps the quadrant (no refine,
 * its family is not coarsened).  Returns the new number of quadrants, or -1
 * if an allocation failed (the forest is then refined / coarsened but not
 * balanced). */
int amr_regrid_tags(amr_forest_t *f, const int *tags, int allow_coarsening);
/* Initial refinement (regrid_feature with initial = 1): repeatedly tag and
 * refine, re-evaluating the initial condition on new quadrants
 * (initialize_quadrant_data), then balance.  Returns the number of
 * quadrants, or -1 if an allocation failed. */
int amr_regrid_initial(amr_forest_t *f, const amr_ic_t *ic, double tol);

/* Output of the evolve loops (as src/amr/src/amr_evolve.c:45-66): when a
 * directory is set, the visualization and checkpoint flags of the interval
 * settings (and the end of the run) write <dir>/soln_<step>.vtu
 * (amr_forest_write_vtk, per quadrant as the reference) and
 * <dir>/checkpoint/amr_soln_<step>.bin (amr_forest_save).  NULL or "" (the
 * default) writes nothing. */
void amr_set_output_dir(const char *dir);

/* Statistics of an adaptive run. */
typedef struct {
    int    steps;            /* time steps taken */
    int    regrids;          /* regrid calls */
    int    peak_quads;       /* largest quadrant count seen */
    double cell_stages;      /* sum over steps of sub-cells x RK stages */
    double t;                /* physical time reached */
    double wall_gpu;         /* seconds inside physics_go_gpu / time stepping */
    double wall_regrid;      /* seconds in regrids (the three parts below) + setup */
    double wall_tags;        /* of which: indicator / tagger and tag download */
    double wall_forest;      /* of which: host forest regrid + boundary arrays */
    double wall_remesh;      /* of which: device face table + state remap */
} amr_evolve_stats_t;

/* The evolve loop of src/amr/src/amr_evolve.c:15-73 on the host forest: per
 * interval build the boundary arrays (external_setup), call the GPU physics
 * module through its seam, physics_go_gpu (where the reference calls
 * hpc_amr_evolve_solver -> physics_evolve_interface), and regrid when it raises
 * the flag.  The physics settings come from physics_gpu_set_params /
 * hpc_amr_initialize_solver; tol is the density-gradient tagging tolerance.
 * Returns 0, or the physics module's error code. */
int amr_evolve(amr_forest_t *f, double tol, int allow_coarsening, amr_evolve_stats_t *stats);

struct amrgpu_ctx;
/* Tag source for amr_evolve_device / amr_evolve_batch: fills tags[q] for
 * every quadrant q of ctx (amr_regrid_tags values) from its device state;
 * f is the forest (NULL for a batch context); returns 0 or an error code. */
typedef int (*amr_tagger_fn)(void *user, struct amrgpu_ctx *ctx, amr_forest_t *f, int *tags);

typedef struct {
    double tol;                  /* density-gradient tolerance */
    int    allow_coarsening;
    int    regrid_interval;      /* steps between regrids (<= 0: never) */
    amr_tagger_fn tagger;        /* NULL: device density-gradient indicator */
    void  *tagger_ctx;
    int    tile_ppb, tile_minw, tile_design;   /* stage-kernel tile (ppb == 0: default; ppb < 0:
                                                  amrgpu_autotune on the run's mesh at the start
                                                  and whenever the mesh has doubled since) */
    const amrgpu_inst_t *inst;   /* NULL: instance from the physics parameters; else these
                                    settings (e.g. the double-Mach walls and states) */
    int    regrid_lag;           /* 0: regrid at the tagging step (the reference's order);
                                    k in (0, regrid_interval): the host regrids the topology on
                                    a thread while the GPU advances k more steps on the old mesh,
                                    then the state k steps later is remapped (tags k steps old) */
    int    groups;               /* amr_evolve_batch only.  0 or 1: one device context for every
                                    instance.  G > 1 (at most AMR_MAX_GROUPS): the instances are
                                    split into G groups of about equal estimated work, each
                                    advanced by its own context, stream and host thread, so that
                                    one group's host regrid overlaps the others' time steps; every
                                    instance's steps, regrids and results are unchanged */
} amr_evolve_opts_t;

#define AMR_MAX_GROUPS 8

/* The reference's own tag source, as a tagger for amr_evolve_device /
 * amr_evolve_batch: solver_tag_feature's density thresholds
 * (src/solver/src/solver_tag_feature.F90:9-52, methods 1 "density below"
 * and 2 "density above" of src/solver/src/tagging_methods_module.F90:14-71)
 * evaluated on the device by amrgpu_feature_tags.  Tag 1 refines; untagged
 * families are coarsened (src/amr/src/amr_tag.c:202-220).  Pass the address
 * of an amr_feature_tagger_t as tagger_ctx. */
typedef struct {
    int    ntags;                               /* <= AMRGPU_MAX_FEATURE_TAGS */
    int    methods[AMRGPU_MAX_FEATURE_TAGS];
    double tols[AMRGPU_MAX_FEATURE_TAGS];
} amr_feature_tagger_t;
int amr_feature_tagger(void *user, struct amrgpu_ctx *ctx, amr_forest_t *f, int *tags);

/* The evolve loop with the state resident in HBM across regrids: tags are
 * computed on the device, the forest topology is regridded on the host in
 * tracking mode, the state is remapped on the device (amrgpu_remesh).  Time
 * stepping settings come from the physics parameters; the final state is
 * downloaded into the forest. */
int amr_evolve_device(amr_forest_t *f, const amr_evolve_opts_t *opts, amr_evolve_stats_t *stats);

/* Many independent problems in one device context: the quadrants of every
 * forest are concatenated (instance k owns a contiguous range, cell_inst = k)
 * and each RK stage is one launch over all of them, with per-instance CFL
 * minimum, dt, physical time and stop (inst[k]).  Every `regrid_interval`
 * steps each instance that has not reached its stop is regridded on its own
 * forest by the device indicator (tracking mode) and the combined state is
 * remapped on the device.  An instance past its stop keeps its state (dt =
 * 0) and leaves the device mesh at the next regrid (its final state is
 * written to its forest then), so each instance evolves exactly as it would
 * alone under amr_evolve_device with the same settings and later launches
 * carry only the cells of instances still advancing.  All forests must have the same
 * patch size; time scheme and step limit come from the physics parameters.
 * per_inst[ninst] (may be NULL) receives each instance's statistics
 * (cell_stages counted while it has not reached its stop at the start of a
 * poll of at most 16 steps); *total the whole run's (cell_stages of every
 * launched cell; with opts->groups > 1 the sums over the groups, steps and
 * the wall_* times those of the longest group).  Final states are downloaded
 * into the forests. */
int amr_evolve_batch(amr_forest_t **forests, const amrgpu_inst_t *inst, int ninst,
                     const amr_evolve_opts_t *opts, amr_evolve_stats_t *per_inst, amr_evolve_stats_t *total);

#ifdef __cplusplus
}
#endif
#endif
