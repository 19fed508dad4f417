/*
 * physics_gpu.c -- host side of the GPU physics module: settings, the
 * physics_go_gpu step driver and the drop-in residual / time-step entry
 * points.  Compute runs in the HIP kernels of amr_gpu.hip through the
 * amrgpu_* context; nothing here evaluates the scheme on the CPU.
 */
#define _POSIX_C_SOURCE 200809L
#include "amr_hip_dropin.h"
#include "amr_internal.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* defaults of physics_initialize_inputs (src/physics/src/physics_initialize.c:65-85) */
static physics_gpu_params_t g_p = {
    .dim = 2, .nfields = 4, .nelem_subgrid = 2,
    .extern_cell_size = 3, .extern_face_size = 11,
    .time_scheme = 1, .time_steps = 1, .dt = 0.001,
    .checkpoint_interval = INT_MAX, .visualization_interval = INT_MAX,
    .regrid_interval = INT_MAX,
    .bc = {1, 1, 1, 1, 1, 1},
    .cfl = 1.0, .coarse_grid_h = 1.0, .t_end = 0.0, .verbose = 0,
};
static double g_time = 0.0;          /* physical time accumulated across intervals */
static amrgpu_ctx *g_ctx = NULL;

void physics_gpu_set_params(const physics_gpu_params_t *p)
{
    if (p) g_p = *p;
    g_time = 0.0;
}

void physics_gpu_get_params(physics_gpu_params_t *p)
{
    if (p) *p = g_p;
}

double physics_gpu_time(void) { return g_time; }

/* keyword lookup with the semantics of utilities_find_keyword_*
 * (src/physics/src/physics_utilities.c:135-178): the first line containing
 * the keyword wins and the value is read at offset strlen(keyword) of that
 * line.  Returns 0 when found. */
static int find_keyword(const char *file, const char *key, char *val, size_t vlen)
{
    FILE *fp = fopen(file, "r");
    char buff[256];
    size_t len = strlen(key);
    if (!fp) return 1;
    while (fgets(buff, sizeof(buff), fp)) {
        if (strstr(buff, key)) {
            snprintf(val, vlen, "%s", strlen(buff) > len ? &buff[len] : "");
            fclose(fp);
            return 0;
        }
    }
    fclose(fp);
    return 1;
}
static void kw_int(const char *file, const char *key, int *v)
{
    char s[256];
    if (!find_keyword(file, key, s, sizeof(s))) *v = atoi(s);
}
static void kw_dbl(const char *file, const char *key, double *v)
{
    char s[256];
    if (!find_keyword(file, key, s, sizeof(s))) *v = atof(s);
}

void hpc_amr_initialize_solver(int *rank, int *dim, double *mesh_scale, int *nlevels,
                               int *nfields, int *nfringe, int *nelem_subgrid,
                               int *extern_cell_size, int *extern_face_size,
                               int *extern_edge_size, int *extern_node_size,
                               int *extern_mpi_comm_size)
{
    const char *file = "input.physics";
    (void)nlevels; (void)nfringe; (void)extern_edge_size; (void)extern_node_size;
    (void)extern_mpi_comm_size;
    physics_gpu_params_t p = g_p;
    p.dim = *dim; p.nfields = *nfields; p.nelem_subgrid = *nelem_subgrid;
    p.extern_cell_size = *extern_cell_size; p.extern_face_size = *extern_face_size;
    p.coarse_grid_h = *mesh_scale;
    p.checkpoint_interval = 0; p.visualization_interval = 0; p.regrid_interval = 0;
    if (access(file, F_OK) == 0) {
        kw_int(file, "checkpoint_interval:", &p.checkpoint_interval);
        kw_int(file, "visualization_interval:", &p.visualization_interval);
        kw_int(file, "regrid_interval:", &p.regrid_interval);
        kw_int(file, "time_scheme:", &p.time_scheme);
        kw_int(file, "time_steps:", &p.time_steps);
        kw_dbl(file, "dt:", &p.dt);
        kw_int(file, "bc_xlo:", &p.bc[0]);
        kw_int(file, "bc_xhi:", &p.bc[1]);
        kw_int(file, "bc_ylo:", &p.bc[2]);
        kw_int(file, "bc_yhi:", &p.bc[3]);
        kw_int(file, "bc_zlo:", &p.bc[4]);
        kw_int(file, "bc_zhi:", &p.bc[5]);
        kw_dbl(file, "cfl:", &p.cfl);
        kw_dbl(file, "t_end:", &p.t_end);
        /* physics_initialize.c:193 */
        p.cfl = 0.25 * p.cfl * p.time_scheme;
    } else if (*rank == 0) {
        printf("[physics] The physics input file does not exist!\n");
    }
    /* physics_initialize.c:313-316 */
    if (p.checkpoint_interval <= 0) p.checkpoint_interval = INT_MAX;
    if (p.visualization_interval <= 0) p.visualization_interval = INT_MAX;
    if (p.regrid_interval <= 0) p.regrid_interval = INT_MAX;
    g_p = p;
    g_time = 0.0;
}

static double wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int ensure_ctx(void)
{
    if (!g_ctx) g_ctx = amrgpu_create();
    if (!g_ctx) {
        fprintf(stderr, "[physics] GPU context creation failed: %s\n", amrgpu_last_error());
        return AMRGPU_ERR_HIP;
    }
    return AMRGPU_OK;
}

static int check_structured(const int *dim, const int *nfields)
{
    if (*dim != 2 || *nfields != 4) {
        fprintf(stderr, "[physics] GPU path supports dim=2, nfields=4 only (got %d, %d)\n",
                *dim, *nfields);
        return AMRGPU_ERR_ARG;
    }
    return AMRGPU_OK;
}

static amrgpu_inst_t inst_from_params(void)
{
    amrgpu_inst_t ip;
    memset(&ip, 0, sizeof(ip));
    ip.h_base = g_p.coarse_grid_h;
    for (int s = 0; s < 4; ++s) ip.bc[s] = g_p.bc[s];
    ip.cfl = g_p.cfl;
    ip.dt_cap = g_p.dt;
    ip.t_end = g_p.t_end;
    return ip;
}

/* steps until the next multiple of interval (the reference stops when
 * istep % interval == 0 after incrementing) */
static long long to_next(int istep, int interval)
{
    if (interval <= 0 || interval == INT_MAX) return LLONG_MAX;
    return (long long)interval - (istep % interval);
}

int physics_go_gpu(int *dim, int *unstructured, int *nfields, int *nelem_subgrid,
                   int *ncell, int *nghost, int *nface, int *dof,
                   int *cell_info, int *face_info, int *nsend, int *nrecv,
                   int *send, int *recv, int *sizes, double *geom, double *soln,
                   int *istep, int *regrid, int *visualize, int *checkpoint,
                   int *evolve_solution, void *mpicomm_ext,
                   double *total_time, double *compute_time, double *residual_time)
{
    (void)nghost; (void)dof; (void)nsend; (void)nrecv; (void)send; (void)recv;
    (void)sizes; (void)mpicomm_ext;
    int rc;
    if (*unstructured) {
        fprintf(stderr, "[physics] unstructured grids are not supported on the GPU path\n");
        return AMRGPU_ERR_ARG;
    }
    if ((rc = check_structured(dim, nfields))) return rc;
    /* physics_go.c:66-69 */
    if (*istep >= g_p.time_steps) {
        *evolve_solution = 0;
        return AMRGPU_OK;
    }
    /* one GPU owns the whole forest, so an interval always has quadrants
     * (the reference's empty MPI ranks have no counterpart here) */
    if (*ncell <= 0 || *nface < 0) {
        amr_set_error("physics_go_gpu: empty mesh (ncell %d, nface %d)", *ncell, *nface);
        return AMRGPU_ERR_ARG;
    }
    if ((rc = ensure_ctx())) return rc;

    double t1 = wall();
    amrgpu_inst_t ip = inst_from_params();
    rc = amrgpu_setup(g_ctx, *nelem_subgrid, g_p.time_scheme, 1, &ip, *ncell, cell_info, NULL,
                      *nface, face_info, geom, soln);
    if (rc) {
        fprintf(stderr, "[physics] GPU setup failed: %s\n", amrgpu_last_error());
        return rc;
    }
    if ((rc = amrgpu_set_time(g_ctx, &g_time))) return rc;

    /* steps this interval: physics_runge_kutta_interface.c:213-229 */
    long long n = (long long)g_p.time_steps - *istep;
    long long r = to_next(*istep, g_p.regrid_interval);
    long long k = to_next(*istep, g_p.checkpoint_interval);
    long long v = to_next(*istep, g_p.visualization_interval);
    long long nsteps = n;
    if (r < nsteps) nsteps = r;
    if (k < nsteps) nsteps = k;
    if (v < nsteps) nsteps = v;

    double resnorm = -1.0;
    long long done = 0;
    int stop_time = 0;
    double tc0 = wall();
    while (done < nsteps) {
        long long chunk = nsteps - done;
        if (g_p.t_end > 0.0 && chunk > 16) chunk = 16;   /* poll the physical time */
        rc = amrgpu_run(g_ctx, (int)chunk, 1, &resnorm);
        done += chunk;
        if (rc == AMRGPU_ERR_NAN) {
            printf("[solve] nan detected in residual aborting\n");
            *istep += (int)done;
            *evolve_solution = 0;
            return rc;
        }
        if (rc) {
            fprintf(stderr, "[physics] GPU run failed: %s\n", amrgpu_last_error());
            return rc;
        }
        if (g_p.t_end > 0.0) {
            amrgpu_get_time(g_ctx, &g_time, NULL);
            if (g_time >= g_p.t_end) { stop_time = 1; break; }
        }
    }
    double tc1 = wall();
    if (g_p.t_end <= 0.0) amrgpu_get_time(g_ctx, &g_time, NULL);
    *istep += (int)done;
    if (*istep >= g_p.time_steps || stop_time) *evolve_solution = 0;
    if (!stop_time) {
        if (*istep % g_p.regrid_interval == 0) *regrid = 1;
        if (*istep % g_p.checkpoint_interval == 0) *checkpoint = 1;
        if (*istep % g_p.visualization_interval == 0) *visualize = 1;
    } else {
        *visualize = 1;
        *checkpoint = 1;
    }
    if ((rc = amrgpu_download_soln(g_ctx, soln))) return rc;
    double t2 = wall();
    *total_time += t2 - t1;
    *compute_time += tc1 - tc0;
    *residual_time += tc1 - tc0;
    if (g_p.verbose) {
        int total_dof = *ncell * (*nelem_subgrid) * (*nelem_subgrid);
        printf("[ amr ] istep: %d\n", *istep);
        printf("[physics] iterations: %lld, gpu time: %e, residual: %1.15e, t = %.6f\n", done,
               tc1 - tc0, resnorm / pow((double)total_dof, 1.0 / (double)*dim), g_time);
    }
    return AMRGPU_OK;
}

int hpc_amr_evolve_solver(int *dim, int *unstructured, int *nfields, int *nelem_subgrid,
                          int *ncell, int *nghost, int *nface, int *dof,
                          int *cell_info, int *face_info, int *nsend, int *nrecv,
                          int *send, int *recv, int *sizes, double *geom, double *soln,
                          int *istep, int *regrid, int *visualize, int *checkpoint,
                          int *evolve_solution, void *mpicomm_ext,
                          double *total_time, double *compute_time, double *residual_time)
{
    return physics_go_gpu(dim, unstructured, nfields, nelem_subgrid, ncell, nghost, nface, dof,
                          cell_info, face_info, nsend, nrecv, send, recv, sizes, geom, soln,
                          istep, regrid, visualize, checkpoint, evolve_solution, mpicomm_ext,
                          total_time, compute_time, residual_time);
}

int physics_rhs_structured_gpu(int *dim, int *nfields, int *nelem_subgrid, int *ncell,
                               int *cell_info, double *geom, int *nface, int *face_info,
                               int *dof, double *soln, double *rhs)
{
    int rc;
    (void)dof;
    if ((rc = check_structured(dim, nfields))) return rc;
    if ((rc = ensure_ctx())) return rc;
    amrgpu_inst_t ip = inst_from_params();
    rc = amrgpu_setup(g_ctx, *nelem_subgrid, g_p.time_scheme, 1, &ip, *ncell, cell_info, NULL,
                      *nface, face_info, geom, soln);
    if (rc) return rc;
    return amrgpu_rhs(g_ctx, rhs);
}

int physics_time_step_gpu(int *dim, int *nfields, int *nelem_subgrid, int *ncell,
                          int *cell_info, double *geom, double *soln, double *dt)
{
    int rc;
    if ((rc = check_structured(dim, nfields))) return rc;
    if ((rc = ensure_ctx())) return rc;
    /* the time step needs no faces: build a boundary-only table */
    int *fi = (int *)malloc(sizeof(int) * 11 * 4 * (size_t)(*ncell));
    if (!fi) return AMRGPU_ERR_ARG;
    for (int c = 0; c < *ncell; ++c)
        for (int s = 0; s < 4; ++s) {
            int *f = &fi[11 * (4 * c + s)];
            for (int e = 0; e < 11; ++e) f[e] = -1;
            f[0] = 1; f[1] = s; f[2] = c;
        }
    amrgpu_inst_t ip = inst_from_params();
    rc = amrgpu_setup(g_ctx, *nelem_subgrid, g_p.time_scheme, 1, &ip, *ncell, cell_info, NULL,
                      4 * (*ncell), fi, geom, soln);
    free(fi);
    if (rc) return rc;
    double d;
    if ((rc = amrgpu_cfl(g_ctx, &d))) return rc;
    if (d < *dt) *dt = d;
    return AMRGPU_OK;
}
