/*
 * amr_forest.h -- host quadtree forest that stands in for p4est on the AMR
 * side of the boundary (p4est is not available in this image).
 *
 * It owns the mesh (brick of trees, linear Morton-ordered quadrants, 2:1
 * face+corner balance, refine/coarsen with the reference's data operators)
 * and produces the exact arrays the reference AMR module hands to the physics
 * module through hpc_amr_evolve_solver (src/amr/src/amr_external.c:211-508):
 *
 *   cell_info[3*c + {0,1,2}] = {type, level, soln offset = 4*n*n*c}
 *   face_info[11*f + k]      = {facetype (1 bc, 2 full, 3 hanging 2D), side of the
 *                               full quad, its cell id, neighbour ids [3..6],
 *                               neighbour sides [7..10]}, unused entries -1
 *   geom[12*c]               = 4 corners x (x,y,z), z-order, as
 *                               p4est_utilities_quad_coordinates
 *                               (src/amr/src/amr_p4est_utilities.c:15-81)
 *   soln                     = per quadrant patch Q(4,n,n), patch-contiguous
 *
 * One process owns the whole forest: no ghosts, nsend = nrecv = 0.
 */
#ifndef AMR_FOREST_H
#define AMR_FOREST_H

#ifdef __cplusplus
extern "C" {
#endif

#define AMR_MAXLEVEL 30                 /* P4EST_MAXLEVEL in 2D */
#define AMR_ROOT_LEN (1 << AMR_MAXLEVEL)

typedef struct {
    int tree;       /* brick tree id */
    int level;
    int x, y;       /* integer coordinates in [0, AMR_ROOT_LEN) */
} amr_quad_t;

typedef struct amr_forest amr_forest_t;

/* Brick of ntx x nty trees of edge `scale`, lower corner (xlo, ylo), periodic
 * flags per direction; n x n sub-grid patch per quadrant; uniform refinement
 * to min_level.  Mirrors initialize_grid + p4est_new_ext
 * (src/amr/src/amr_initialize.c:213-314, 199-207). */
amr_forest_t *amr_forest_new(int ntx, int nty, int periodic_x, int periodic_y,
                             double xlo, double ylo, double scale, int n,
                             int min_level, int max_level);
void amr_forest_destroy(amr_forest_t *f);

int amr_forest_num_quads(const amr_forest_t *f);
int amr_forest_nelem_subgrid(const amr_forest_t *f);
/* Whether the calling thread's per-quadrant forest loops run on OpenMP
 * threads: on by default; a batch regrid turns it off on its own per-instance
 * workers only, and leaves the caller's setting as it found it. */
int amr_forest_host_parallel_enabled(void);
const amr_quad_t *amr_forest_quads(const amr_forest_t *f);
double *amr_forest_soln(amr_forest_t *f);     /* nquad * 4*n*n, reference layout */
int *amr_forest_tags(amr_forest_t *f);        /* nquad */
double amr_forest_scale(const amr_forest_t *f);

/* Corner coordinates of quadrant k (12 doubles), reference formula. */
void amr_forest_quad_geom(const amr_forest_t *f, int k, double xyz[12]);

/* Build the boundary arrays (memory owned by the forest, valid until the next
 * call or destroy).  Returns 0 on success. */
int amr_forest_build_external(amr_forest_t *f, int *ncell, int *nghost, int *nface, int *dof,
                              int **cell_info, int **face_info, double **geom, double **soln);
/* The device mesh table of the current forest, without the reference's
 * boundary arrays (what amrgpu_remesh_nbr takes): per quadrant k and side s
 * (-x, +x, -y, +y) nbr[8k + 2s] = a, nbr[8k + 2s + 1] = b with a < 0 a
 * physical boundary, b < 0 one neighbour a (same size, or the coarser one),
 * else the two finer neighbours a, b in increasing coordinate along the side;
 * level[k]; org[2k], org[2k + 1] the lower-left corner.  The same table as
 * amrgpu_nbr_from_faces makes of amr_forest_build_external's face_info.
 * Arrays of 8, 1 and 2 entries per quadrant.  Returns 0, -1 if inconsistent. */
int amr_forest_build_nbr(amr_forest_t *f, int *nbr, int *level, double *org);

/* Refine every quadrant with flags[k] != 0 (and level < max_level) by one
 * level; children copy the parent patch (solver_refine_operator_2d) and
 * inherit its tag.  Returns the number of quadrants refined. */
int amr_forest_refine(amr_forest_t *f, const int *flags);
/* Coarsen every complete family whose four members have tag == 0 and level >
 * min_level (tag_coarsen, src/amr/src/amr_tag.c:202-220); the parent patch is
 * the element-wise mean (solver_coarsen_operator_2d).  Returns families
 * coarsened. */
int amr_forest_coarsen(amr_forest_t *f);
/* 2:1 balance across faces and corners (P4EST_CONNECT_FULL); refined
 * quadrants copy their parent patch.  Returns quadrants refined. */
int amr_forest_balance(amr_forest_t *f);
/* The same closure after a regrid of a balanced forest, examining only the
 * quadrants created by refine / coarsen since amr_forest_clear_marks (their
 * neighbours are the only possible violations).  Returns quadrants refined,
 * or -1 with the forest unchanged if an allocation failed. */
int amr_forest_balance_marked(amr_forest_t *f);
void amr_forest_clear_marks(amr_forest_t *f);
/* 1 if the forest is 2:1 balanced across faces and corners. */
int amr_forest_is_balanced(amr_forest_t *f);

/* Regrid tracking.  Between track_begin and track_end, refine / coarsen /
 * balance record where each quadrant's patch comes from; with keep_data = 0
 * the host patches are not moved at all (the state lives on the device and
 * is remapped there).  amr_forest_track_src() gives, per current quadrant,
 * src >= 0: copy of old quadrant src (refine copies the parent, as
 * solver_refine_operator_2d); src < 0: element-wise mean of the old family
 * -src-1 .. -src+2 (solver_coarsen_operator_2d).  track_end returns 1 when
 * every quadrant has such a source. */
void amr_forest_track_begin(amr_forest_t *f, int keep_data);
const int *amr_forest_track_src(const amr_forest_t *f);
int amr_forest_track_end(amr_forest_t *f);

/* Leaf index containing the point (tree, x, y) or -1. */
int amr_forest_find_leaf(const amr_forest_t *f, int tree, int x, int y);

/* Construction parameters of the forest (any pointer may be NULL). */
void amr_forest_brick(const amr_forest_t *f, int *ntx, int *nty, int *periodic_x, int *periodic_y,
                      double *xlo, double *ylo, double *scale, int *min_level, int *max_level);
/* Forest with the given leaves (tree-major Morton order, each tree tiled
 * exactly once, levels within [0, max_level]); patches zero, tags 0.  NULL if
 * the quadrants do not form such a forest. */
amr_forest_t *amr_forest_from_quads(int ntx, int nty, int periodic_x, int periodic_y, double xlo,
                                    double ylo, double scale, int n, int min_level, int max_level,
                                    int nquad, const amr_quad_t *quads);

/* ---- output in this project's formats (amr_io.c) ----------------------
 * The reference checkpoints with p4est_save and visualizes through p4est's
 * VTK writer (src/amr/src/amr_evolve.c:55-64, src/amr/src/amr_vtk.c:23-106);
 * without p4est both are self-contained files. */

/* Binary checkpoint (header with brick, patch size, levels, step and time;
 * quadrants; patches in the reference layout).  0 on success. */
int amr_forest_save(amr_forest_t *f, const char *path, int istep, double t);
/* Forest restored from amr_forest_save's file (NULL on a bad file). */
amr_forest_t *amr_forest_load(const char *path, int *istep, double *t);
/* XML VTK unstructured grid (.vtu, raw appended binary).  subcells = 0: one
 * cell per quadrant with tree and level, and the patch-averaged rho, rhou,
 * rhov, rhoe at its corners (the reference's fields); subcells = 1: one cell
 * per sub-grid cell with its state.  0 on success. */
int amr_forest_write_vtk(amr_forest_t *f, const char *path, int subcells);

#ifdef __cplusplus
}
#endif
#endif
