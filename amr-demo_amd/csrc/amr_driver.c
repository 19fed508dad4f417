/*
 * amr_driver.c -- initial conditions, density-gradient indicator and the
 * regrid sequence on the host forest.  See include/amr_driver.h.
 */
#define _POSIX_C_SOURCE 200809L
#include "amr_driver.h"
#include "amr_hip.h"
#include "amr_internal.h"

#include <limits.h>
#include <pthread.h>
#include <stdio.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

static void bilinear(const double g[12], double xi, double eta, double out[2])
{
    /* src/solver/src/bilinear_node_2d_module.F90:14-36 */
    double oneMxi = 0.5 * (1.0 - xi), onePxi = 0.5 * (1.0 + xi);
    double oneMeta = 0.5 * (1.0 - eta), onePeta = 0.5 * (1.0 + eta);
    for (int d = 0; d < 2; ++d)
        out[d] = g[0 + d] * oneMxi * oneMeta + g[3 + d] * onePxi * oneMeta +
                 g[6 + d] * oneMxi * onePeta + g[9 + d] * onePxi * onePeta;
}

static void centre(const double g[12], int n, int i, int j, double out[2])
{
    /* sub-cell centres as initial_condition_2d_module.F90:100-111 */
    double sf = 2.0 / (double)n;
    double v1 = sf * (double)j - 1.0, v2 = sf * (double)(j + 1) - 1.0;
    double w1 = sf * (double)i - 1.0, w2 = sf * (double)(i + 1) - 1.0;
    bilinear(g, 0.5 * (w1 + w2), 0.5 * (v1 + v2), out);
}

static void prim_to_cons(double gamma, double rho, double u, double v, double p, double *q)
{
    q[0] = rho;
    q[1] = rho * u;
    q[2] = rho * v;
    q[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v);
}

void amr_ic_patch(const amr_ic_t *ic, int n, const double geom[12], double *Q)
{
    const double pi = 3.141592653589793238462643383279502884197;
    const double twopi = 6.283185307179586476925286766559005768394;
    const double gamma = ic->gamma;
    const double gm1 = gamma - 1.0, oneOgm1 = 1.0 / gm1;
    /* free-stream quantities (initial_condition_2d_module.F90:51-67) */
    double RT0 = ic->pressure / ic->density;
    double V_0 = ic->mach * sqrt(gamma * RT0);
    double u_in = V_0 * cos(ic->alpha * pi / 180.0) * cos(ic->beta * pi / 180.0);
    double v_in = V_0 * sin(ic->alpha * pi / 180.0) * cos(ic->beta * pi / 180.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double *q = &Q[4 * (i + n * j)];
            double x[2];
            centre(geom, n, i, j, x);
            switch (ic->kind) {
            case AMR_IC_VORTEX: {
                /* initial_condition_2d_module.F90:83-129, strength/centre as inputs */
                double a = 0.5, two_a = 2.0 * a, bl = ic->vortex_strength;
                double dT1 = -gm1 * bl * bl / (16.0 * a * gamma * pi * pi);
                double r2 = (x[0] - ic->xo) * (x[0] - ic->xo) + (x[1] - ic->yo) * (x[1] - ic->yo);
                double dT = dT1 * exp(two_a * (1.0 - r2));
                double rho = pow(1.0 + dT, oneOgm1);
                double ff = bl / twopi * exp(a * (1.0 - r2));
                double u = u_in - (x[1] - ic->yo) * ff;
                double v = v_in + (x[0] - ic->xo) * ff;
                double p = ic->pressure + pow(rho, gamma) - 1.0;
                q[0] = rho;
                q[1] = rho * u;
                q[2] = rho * v;
                q[3] = p * oneOgm1 + 0.5 * rho * (u * u + v * v);
                break;
            }
            case AMR_IC_RIEMANN4: {
                int k = (x[0] > ic->xs ? 1 : 0) + (x[1] > ic->ys ? 2 : 0);
                const double *s = ic->state[k];
                prim_to_cons(gamma, s[0], s[1], s[2], s[3], q);
                break;
            }
            case AMR_IC_DMR: {
                /* shock line through (x0, 0) making angle th with the x axis;
                 * post-shock state behind it (x < x0 + y / tan th) */
                double th = ic->dmr_angle * pi / 180.0;
                double xs = ic->dmr_x0 + x[1] / tan(th);
                double rho1 = ic->dmr_rho, p1 = ic->dmr_p;
                double M = ic->dmr_mach;
                double c1 = sqrt(gamma * p1 / rho1);
                double w = M * c1;                          /* shock speed */
                double rho2 = rho1 * (gamma + 1.0) * M * M / ((gamma - 1.0) * M * M + 2.0);
                double p2 = p1 * (2.0 * gamma * M * M - (gamma - 1.0)) / (gamma + 1.0);
                double un = w * (1.0 - rho1 / rho2);        /* post-shock normal speed */
                if (x[0] < xs)
                    prim_to_cons(gamma, rho2, un * sin(th), -un * cos(th), p2, q);
                else
                    prim_to_cons(gamma, rho1, 0.0, 0.0, p1, q);
                break;
            }
            default:
                q[0] = ic->density;
                q[1] = ic->density * u_in;
                q[2] = ic->density * v_in;
                q[3] = (ic->pressure * oneOgm1) + 0.5 * ic->density * (u_in * u_in + v_in * v_in);
                break;
            }
        }
}

void amr_ic_apply(amr_forest_t *f, const amr_ic_t *ic)
{
    int nq = amr_forest_num_quads(f), n = amr_forest_nelem_subgrid(f);
    double *soln = amr_forest_soln(f);
    double g[12];
    for (int k = 0; k < nq; ++k) {
        amr_forest_quad_geom(f, k, g);
        amr_ic_patch(ic, n, g, &soln[(size_t)4 * n * n * k]);
    }
}

double amr_ic_vortex_density(const amr_ic_t *ic, double x, double y, double t,
                             double x0, double Lx, double y0, double Ly)
{
    const double pi = 3.141592653589793238462643383279502884197;
    const double gamma = ic->gamma, gm1 = gamma - 1.0;
    double RT0 = ic->pressure / ic->density;
    double V_0 = ic->mach * sqrt(gamma * RT0);
    double u_in = V_0 * cos(ic->alpha * pi / 180.0) * cos(ic->beta * pi / 180.0);
    double v_in = V_0 * sin(ic->alpha * pi / 180.0) * cos(ic->beta * pi / 180.0);
    double cx = ic->xo + u_in * t, cy = ic->yo + v_in * t;
    double dx = x - cx, dy = y - cy;
    dx -= Lx * floor((dx - x0 + x0) / Lx + 0.5);
    dy -= Ly * floor((dy - y0 + y0) / Ly + 0.5);
    double a = 0.5, bl = ic->vortex_strength;
    double dT1 = -gm1 * bl * bl / (16.0 * a * gamma * pi * pi);
    double dT = dT1 * exp(2.0 * a * (1.0 - (dx * dx + dy * dy)));
    return pow(1.0 + dT, 1.0 / gm1);
}

static inline double jump(double a, double b)
{
    return fabs(a - b) / (0.5 * (a + b));
}

/* sub-cell (i, j) at position t along side s (-x, +x, -y, +y) of a patch */
static inline void side_cell(int n, int s, int t, int *i, int *j)
{
    switch (s) {
    case 0: *i = 0; *j = t; break;
    case 1: *i = n - 1; *j = t; break;
    case 2: *i = t; *j = 0; break;
    default: *i = t; *j = n - 1; break;
    }
}

void amr_indicator_gradient(amr_forest_t *f, double *ind)
{
    int ncell, nghost, nface, dof;
    int *cell_info, *face_info;
    double *geom, *soln;
    /* between the passes of the initial refinement the forest is not yet
     * 2:1 balanced: the face list then omits the sides that face a
     * neighbour two or more levels away (the builder reports it), and those
     * sides contribute no pairs */
    (void)amr_forest_build_external(f, &ncell, &nghost, &nface, &dof, &cell_info, &face_info, &geom, &soln);
    const int n = amr_forest_nelem_subgrid(f);
    /* interior pairs */
    for (int k = 0; k < ncell; ++k) {
        const double *Q = &soln[(size_t)4 * n * n * k];
        double m = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double r = Q[4 * (i + n * j)];
                if (i + 1 < n) { double v = jump(r, Q[4 * (i + 1 + n * j)]); if (v > m) m = v; }
                if (j + 1 < n) { double v = jump(r, Q[4 * (i + n * (j + 1))]); if (v > m) m = v; }
            }
        ind[k] = m;
    }
    /* across faces: both sides see the same pairs (the neighbour's edge
     * cells lie on its opposite side s ^ 1, index-aligned) */
    for (int fc = 0; fc < nface; ++fc) {
        const int *fi = &face_info[11 * fc];
        if (fi[0] == 1) continue;
        const int s = fi[1], a = fi[2];
        const double *Qa = &soln[(size_t)4 * n * n * a];
        const int nbs[2] = {fi[3], fi[0] == 3 ? fi[4] : -1};
        for (int h = 0; h < 2; ++h) {
            const int b = nbs[h];
            if (b < 0) continue;
            const double *Qb = &soln[(size_t)4 * n * n * b];
            double m = 0.0;
            for (int t = 0; t < n; ++t) {
                int i, j, ib, jb;
                side_cell(n, s, t, &i, &j);
                side_cell(n, s ^ 1, t, &ib, &jb);
                double v = jump(Qa[4 * (i + n * j)], Qb[4 * (ib + n * jb)]);
                if (v > m) m = v;
            }
            if (m > ind[a]) ind[a] = m;
            if (m > ind[b]) ind[b] = m;
        }
    }
}

int amr_regrid_tags(amr_forest_t *f, const int *tags, int allow_coarsening)
{
    int nq = amr_forest_num_quads(f);
    int *tag = amr_forest_tags(f);
    /* tag_reset + regrid_feature; a "keep" tag (2) refines nothing but, being
     * non-zero, stops tag_coarsen from merging the family */
    int *flags = (int *)malloc(sizeof(int) * (size_t)(nq > 0 ? nq : 1));
    if (!flags) return -1;
    for (int k = 0; k < nq; ++k) {
        tag[k] = tags[k] == 1 ? 1 : (tags[k] ? 2 : 0);
        flags[k] = tags[k] == 1;
    }
    /* the forest is balanced on entry (every regrid leaves it so): balance
     * only around what this regrid refines and coarsens */
    amr_forest_clear_marks(f);
    amr_forest_refine(f, flags);
    free(flags);
    if (allow_coarsening) amr_forest_coarsen(f);
    if (amr_forest_balance_marked(f) < 0) return -1;
    return amr_forest_num_quads(f);
}

int amr_regrid(amr_forest_t *f, double tol, int allow_coarsening)
{
    int nq = amr_forest_num_quads(f);
    double *ind = (double *)malloc(sizeof(double) * (size_t)(nq > 0 ? nq : 1));
    int *tags = (int *)malloc(sizeof(int) * (size_t)(nq > 0 ? nq : 1));
    int r = -1;
    if (ind && tags) {
        amr_indicator_gradient(f, ind);
        for (int k = 0; k < nq; ++k) tags[k] = ind[k] > tol;
        r = amr_regrid_tags(f, tags, allow_coarsening);
    }
    free(tags);
    free(ind);
    return r;
}

int amr_regrid_initial(amr_forest_t *f, const amr_ic_t *ic, double tol)
{
    for (int pass = 0; pass < AMR_MAXLEVEL; ++pass) {
        int nq = amr_forest_num_quads(f);
        double *ind = (double *)malloc(sizeof(double) * (size_t)(nq > 0 ? nq : 1));
        int *flags = (int *)malloc(sizeof(int) * (size_t)(nq > 0 ? nq : 1));
        if (!ind || !flags) {
            free(ind);
            free(flags);
            return -1;
        }
        amr_indicator_gradient(f, ind);
        for (int k = 0; k < nq; ++k) flags[k] = ind[k] > tol;
        int r = amr_forest_refine(f, flags);
        free(ind);
        free(flags);
        if (!r) break;
        /* initialize_quadrant_data on every quadrant (new children get the IC) */
        amr_ic_apply(f, ic);
    }
    amr_forest_balance(f);
    amr_ic_apply(f, ic);
    int nq = amr_forest_num_quads(f);
    int *tag = amr_forest_tags(f);
    for (int k = 0; k < nq; ++k) tag[k] = 0;
    return nq;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static char g_outdir[1024];

void amr_set_output_dir(const char *dir)
{
    g_outdir[0] = '\0';
    if (dir && dir[0]) snprintf(g_outdir, sizeof(g_outdir), "%s", dir);
}

/* vtk_write_solution / p4est_save of the reference evolve loop */
static int write_output(amr_forest_t *f, int istep, double t, int visualize, int checkpoint)
{
    if (!g_outdir[0]) return 0;
    char path[1200];
    int rc = 0;
    if (visualize) {
        snprintf(path, sizeof(path), "%s/soln_%07d.vtu", g_outdir, istep);
        if (amr_forest_write_vtk(f, path, 0)) rc = AMRGPU_ERR_ARG;
    }
    if (checkpoint) {
        snprintf(path, sizeof(path), "%s/checkpoint", g_outdir);
        mkdir(path, 0777);
        snprintf(path, sizeof(path), "%s/checkpoint/amr_soln_%07d.bin", g_outdir, istep);
        if (amr_forest_save(f, path, istep, t)) rc = AMRGPU_ERR_ARG;
    }
    return rc;
}

int amr_evolve(amr_forest_t *f, double tol, int allow_coarsening, amr_evolve_stats_t *st)
{
    physics_gpu_params_t p;
    physics_gpu_get_params(&p);
    static const int stages[6] = {0, 1, 2, 3, 4, 5};
    int istep = 0, evolve = 1, rc = 0;
    double total = 0.0, comp = 0.0, res = 0.0;
    amr_evolve_stats_t s;
    memset(&s, 0, sizeof(s));
    s.peak_quads = amr_forest_num_quads(f);
    while (evolve) {
        int regrid = 0, visualize = 0, checkpoint = 0;
        int ncell, nghost, nface, dof, dim = 2, uns = 0, nfields = 4, n = amr_forest_nelem_subgrid(f);
        int nsend = 0, nrecv = 0, dummy = 0;
        int *cell_info, *face_info;
        double *geom, *soln;
        double t0 = now_s();
        if (amr_forest_build_external(f, &ncell, &nghost, &nface, &dof, &cell_info, &face_info,
                                      &geom, &soln))
            return AMRGPU_ERR_MESH;
        double t1 = now_s();
        int before = istep;
        /* soln is the forest's own patch array: updated in place
         * (external_p4est_to_soln / external_soln_to_p4est) */
        rc = physics_go_gpu(&dim, &uns, &nfields, &n, &ncell, &nghost, &nface, &dof,
                            cell_info, face_info, &nsend, &nrecv, &dummy, &dummy, &dummy,
                            geom, soln, &istep, &regrid, &visualize, &checkpoint, &evolve,
                            NULL, &total, &comp, &res);
        double t2 = now_s();
        s.wall_gpu += t2 - t1;
        s.wall_regrid += t1 - t0;
        if (rc) break;
        int nst = istep - before;
        s.steps += nst;
        s.cell_stages += (double)nst * ncell * n * n * stages[p.time_scheme];
        if (!evolve) visualize = checkpoint = 1;     /* end of the run */
        rc = write_output(f, istep, physics_gpu_time(), visualize, checkpoint);
        if (rc) break;
        if (regrid && evolve) {
            double t3 = now_s();
            if (amr_regrid(f, tol, allow_coarsening) < 0) { rc = AMRGPU_ERR_MESH; break; }
            s.wall_regrid += now_s() - t3;
            s.regrids++;
            if (amr_forest_num_quads(f) > s.peak_quads) s.peak_quads = amr_forest_num_quads(f);
        }
    }
    physics_gpu_get_params(&p);
    s.t = physics_gpu_time();
    if (st) *st = s;
    return rc;
}

/* interval length of the reference step loop for the regrid flag */
static long long steps_to(int istep, int interval)
{
    if (interval <= 0 || interval == 2147483647) return 1LL << 40;
    return (long long)interval - (istep % interval);
}

int amr_device_setup(amr_forest_t *f, const amrgpu_inst_t *override, amrgpu_ctx **out)
{
    physics_gpu_params_t p;
    physics_gpu_get_params(&p);
    int ncell, nghost, nface, dof;
    int *cell_info, *face_info;
    double *geom, *soln;
    *out = NULL;
    if (amr_forest_build_external(f, &ncell, &nghost, &nface, &dof, &cell_info, &face_info, &geom, &soln))
        return AMRGPU_ERR_MESH;
    amrgpu_ctx *g = amrgpu_create();
    if (!g) return AMRGPU_ERR_HIP;
    amrgpu_inst_t ip;
    memset(&ip, 0, sizeof(ip));
    ip.h_base = p.coarse_grid_h;
    for (int k = 0; k < 4; ++k) ip.bc[k] = p.bc[k];
    ip.cfl = p.cfl;
    ip.dt_cap = p.dt;
    ip.t_end = p.t_end;
    if (override) ip = *override;
    int rc = amrgpu_setup(g, amr_forest_nelem_subgrid(f), p.time_scheme, 1, &ip, ncell, cell_info, NULL, nface,
                          face_info, geom, soln);
    if (rc) {
        amrgpu_destroy(g);
        return rc;
    }
    *out = g;
    return AMRGPU_OK;
}

typedef struct {
    int ncell, nface, rc;
    int *cell_info, *face_info;
    double *geom, *soln;
} ext_t;

static int ext_of(amr_forest_t *f, ext_t *e)
{
    int nghost, dof;
    e->rc = amr_forest_build_external(f, &e->ncell, &nghost, &e->nface, &dof, &e->cell_info, &e->face_info,
                                      &e->geom, &e->soln)
                ? AMRGPU_ERR_MESH
                : AMRGPU_OK;
    return e->rc;
}

/* the device mesh table of a forest (amr_forest_build_nbr), for remeshes
 * that skip the reference-format arrays */
typedef struct {
    int nq, rc;
    int *nbr, *level;
    double *org;
} tab_t;

static int tab_of(amr_forest_t *f, tab_t *t)
{
    const int nq = amr_forest_num_quads(f);
    t->nq = nq;
    t->nbr = (int *)malloc(sizeof(int) * 8 * (size_t)(nq > 0 ? nq : 1));
    t->level = (int *)malloc(sizeof(int) * (size_t)(nq > 0 ? nq : 1));
    t->org = (double *)malloc(sizeof(double) * 2 * (size_t)(nq > 0 ? nq : 1));
    t->rc = (!t->nbr || !t->level || !t->org) ? -1 : amr_forest_build_nbr(f, t->nbr, t->level, t->org);
    return t->rc;
}

static void tab_free(tab_t *t)
{
    free(t->nbr); free(t->level); free(t->org);
    t->nbr = t->level = NULL;
    t->org = NULL;
}

/* Host half of a lagged regrid: topology and mesh table of one forest on a
 * thread of its own while the GPU keeps stepping the old mesh. */
typedef struct {
    amr_forest_t *f;
    const int *tags;
    int allow, ok;
    tab_t tab;
    double wall;
} lag_job_t;

static void *lag_regrid_thread(void *arg)
{
    lag_job_t *j = (lag_job_t *)arg;
    double t0 = now_s();
    amr_forest_track_begin(j->f, 0);
    const int r = amr_regrid_tags(j->f, j->tags, j->allow);
    j->ok = amr_forest_track_end(j->f) && r >= 0;
    tab_of(j->f, &j->tab);
    j->wall = now_s() - t0;
    return NULL;
}

double amr_now(void) { return now_s(); }

int amr_device_regrid(amr_forest_t *f, amrgpu_ctx *g, const int *tags, int allow_coarsening, double *t_forest,
                      double *t_remesh)
{
    double t0 = now_s();
    amr_forest_track_begin(f, 0);
    const int r = amr_regrid_tags(f, tags, allow_coarsening);
    if (!amr_forest_track_end(f) || r < 0) return AMRGPU_ERR_MESH;
    /* the device takes its mesh table straight from the forest: the
     * reference-format arrays are not needed between device steps */
    tab_t tb;
    if (tab_of(f, &tb)) { tab_free(&tb); return AMRGPU_ERR_MESH; }
    double t1 = now_s();
    int rc = amrgpu_remesh_nbr(g, tb.nq, tb.level, NULL, tb.nbr, amr_forest_track_src(f));
    if (!rc) rc = amrgpu_set_origins(g, tb.org);
    tab_free(&tb);
    if (t_forest) *t_forest += t1 - t0;
    if (t_remesh) *t_remesh += now_s() - t1;
    return rc;
}

/* tile_ppb < 0 in the options: the stage-kernel tile and launch mode are
 * chosen by amrgpu_autotune on the loop's own mesh, at the start and again
 * whenever the mesh has doubled since the last choice (the trial steps are
 * undone, so the run is the one it would be with that tile set by hand).
 * *tuned_at holds the cell count of the last choice (0: none yet). */
static int autotune_tile(amrgpu_ctx *g, const amr_evolve_opts_t *o, int nquad, int *tuned_at)
{
    if (!o || o->tile_ppb >= 0) return AMRGPU_OK;
    if (*tuned_at > 0 && nquad < 2 * *tuned_at) return AMRGPU_OK;
    *tuned_at = nquad;
    return amrgpu_autotune(g, 5, NULL, NULL, NULL, NULL, NULL);
}

int amr_feature_tagger(void *user, amrgpu_ctx *ctx, amr_forest_t *f, int *tags)
{
    (void)f;
    const amr_feature_tagger_t *u = (const amr_feature_tagger_t *)user;
    if (!u) {
        amr_set_error("amr_feature_tagger: no settings");
        return AMRGPU_ERR_ARG;
    }
    return amrgpu_feature_tags(ctx, u->ntags, u->methods, u->tols, tags);
}

int amr_evolve_device(amr_forest_t *f, const amr_evolve_opts_t *o, amr_evolve_stats_t *st)
{
    physics_gpu_params_t p;
    physics_gpu_get_params(&p);
    static const int stages[6] = {0, 1, 2, 3, 4, 5};
    amr_evolve_stats_t s;
    memset(&s, 0, sizeof(s));
    s.peak_quads = amr_forest_num_quads(f);
    const int n = amr_forest_nelem_subgrid(f);
    int ncell = amr_forest_num_quads(f), rc = 0;
    double t0 = now_s();
    amrgpu_ctx *g = NULL;
    rc = amr_device_setup(f, o ? o->inst : NULL, &g);
    if (rc) return rc;
    if (o && o->tile_ppb > 0) rc = amrgpu_set_tile(g, o->tile_ppb, o->tile_minw, o->tile_design);
    int tuned_at = 0;
    if (!rc) rc = autotune_tile(g, o, ncell, &tuned_at);
    double *ind = NULL;
    int *tags = NULL;
    int cap = 0, istep = 0;
    double t = 0.0, t_print = now_s();
    s.wall_regrid += now_s() - t0;
    while (!rc) {
        long long nsteps = (long long)p.time_steps - istep;
        long long r = steps_to(istep, o ? o->regrid_interval : p.regrid_interval);
        if (r < nsteps) nsteps = r;
        if (g_outdir[0]) {        /* stop at the output intervals too */
            long long v = steps_to(istep, p.visualization_interval), c = steps_to(istep, p.checkpoint_interval);
            if (v < nsteps) nsteps = v;
            if (c < nsteps) nsteps = c;
        }
        if (nsteps <= 0) break;
        if (p.t_end > 0.0 && nsteps > 16) nsteps = 16;       /* poll the physical time */
        double t1 = now_s();
        double resnorm;
        rc = amrgpu_run(g, (int)nsteps, 1, &resnorm);
        if (rc) break;
        istep += (int)nsteps;
        s.steps += (int)nsteps;
        s.cell_stages += (double)nsteps * ncell * n * n * stages[p.time_scheme];
        if (p.t_end > 0.0 || g_outdir[0]) amrgpu_get_time(g, &t, NULL);
        s.wall_gpu += now_s() - t1;
        if (p.verbose > 0 && now_s() - t_print > 10.0) {       /* long runs: a progress line every 10 s */
            fprintf(stderr, "amr_evolve_device: step %d, t = %.6e, %d quadrants, %d regrids\n", istep, t, ncell,
                    s.regrids);
            t_print = now_s();
        }
        const int done = istep >= p.time_steps || (p.t_end > 0.0 && t >= p.t_end);
        if (g_outdir[0] && !done) {
            const int vis = p.visualization_interval > 0 && istep % p.visualization_interval == 0;
            const int ckp = p.checkpoint_interval > 0 && istep % p.checkpoint_interval == 0;
            if (vis || ckp) {
                rc = amrgpu_download_soln(g, amr_forest_soln(f));
                if (!rc) rc = write_output(f, istep, t, vis, ckp);
                if (rc) break;
            }
        }
        if (done) break;
        int interval = o ? o->regrid_interval : p.regrid_interval;
        if (interval <= 0 || istep % interval != 0) continue;
        /* regrid: tags on the device, topology on the host, data on the device */
        double t2 = now_s();
        int nq = amr_forest_num_quads(f);
        if (nq > cap) {
            cap = nq + nq / 2;
            free(ind); free(tags);
            ind = (double *)malloc(sizeof(double) * (size_t)cap);
            tags = (int *)malloc(sizeof(int) * (size_t)cap);
            if (!ind || !tags) { rc = AMRGPU_ERR_ARG; break; }
        }
        if (o && o->tagger) {
            rc = o->tagger(o->tagger_ctx, g, f, tags);
        } else {
            rc = amrgpu_indicator(g, ind);
            for (int k = 0; k < nq; ++k) tags[k] = ind[k] > (o ? o->tol : 0.1);
        }
        s.wall_tags += now_s() - t2;
        if (rc) break;
        const int lag = o ? o->regrid_lag : 0;
        if (lag > 0 && lag < interval) {
            lag_job_t job;
            memset(&job, 0, sizeof(job));
            job.f = f; job.tags = tags; job.allow = o->allow_coarsening;
            pthread_t th;
            const int threaded = pthread_create(&th, NULL, lag_regrid_thread, &job) == 0;
            if (!threaded) lag_regrid_thread(&job);
            /* the GPU advances the old mesh meanwhile */
            int nl = lag;
            if (istep + nl > p.time_steps) nl = p.time_steps - istep;
            double t4 = now_s();
            if (nl > 0) rc = amrgpu_run(g, nl, 1, &resnorm);
            if (!rc && nl > 0) {
                istep += nl;
                s.steps += nl;
                s.cell_stages += (double)nl * ncell * n * n * stages[p.time_scheme];
                if (p.t_end > 0.0 || g_outdir[0]) amrgpu_get_time(g, &t, NULL);
            }
            const double t5 = now_s();
            s.wall_gpu += t5 - t4;
            if (threaded) pthread_join(th, NULL);
            s.wall_forest += job.wall;
            if (!rc && (!job.ok || job.tab.rc)) rc = AMRGPU_ERR_MESH;
            const double t6 = now_s();
            if (!rc) rc = amrgpu_remesh_nbr(g, job.tab.nq, job.tab.level, NULL, job.tab.nbr,
                                            amr_forest_track_src(f));
            if (!rc) rc = amrgpu_set_origins(g, job.tab.org);
            tab_free(&job.tab);
            s.wall_remesh += now_s() - t6;
            /* the regrid wall counts what the host made the GPU wait for:
             * tagging, the join after the overlapped steps, the remesh */
            t2 += t5 - t4;
        } else {
            rc = amr_device_regrid(f, g, tags, o ? o->allow_coarsening : 1, &s.wall_forest, &s.wall_remesh);
        }
        ncell = amr_forest_num_quads(f);
        s.regrids++;
        if (ncell > s.peak_quads) s.peak_quads = ncell;
        if (!rc) rc = autotune_tile(g, o, ncell, &tuned_at);
        s.wall_regrid += now_s() - t2;
        if (istep >= p.time_steps || (p.t_end > 0.0 && t >= p.t_end)) break;   /* stop reached while lagging */
    }
    if (!rc) {
        amrgpu_get_time(g, &t, NULL);
        rc = amrgpu_download_soln(g, amr_forest_soln(f));
    }
    if (!rc) rc = write_output(f, istep, t, 1, 1);      /* end of the run */
    s.t = t;
    amrgpu_destroy(g);
    free(ind);
    free(tags);
    if (st) *st = s;
    return rc;
}

/* ---- many problems in one context -------------------------------------- */

typedef struct {
    int *ci, *fi, *cinst, *src;
    double *geom, *soln;
    size_t cap_c, cap_f;
    int *nbr, *lev, *tinst, *tsrc;     /* concatenated mesh tables (regrids) */
    double *org;
    size_t cap_t;
} batch_bufs_t;

static void batch_free(batch_bufs_t *b)
{
    free(b->ci); free(b->fi); free(b->cinst); free(b->src); free(b->geom); free(b->soln);
    free(b->nbr); free(b->lev); free(b->tinst); free(b->tsrc); free(b->org);
    memset(b, 0, sizeof(*b));
}

static int batch_grow(batch_bufs_t *b, size_t nc, size_t nf, int n2)
{
    if (nc > b->cap_c) {
        size_t c = nc + nc / 2;
        int *ci = (int *)realloc(b->ci, sizeof(int) * 3 * c);
        if (ci) b->ci = ci;
        int *cinst = (int *)realloc(b->cinst, sizeof(int) * c);
        if (cinst) b->cinst = cinst;
        int *src = (int *)realloc(b->src, sizeof(int) * c);
        if (src) b->src = src;
        double *geom = (double *)realloc(b->geom, sizeof(double) * 12 * c);
        if (geom) b->geom = geom;
        double *soln = (double *)realloc(b->soln, sizeof(double) * 4 * (size_t)n2 * c);
        if (soln) b->soln = soln;
        if (!ci || !cinst || !src || !geom || !soln) return -1;
        b->cap_c = c;
    }
    if (nf > b->cap_f || !b->fi) {
        size_t c = nf + nf / 2 + 1;
        int *fi = (int *)realloc(b->fi, sizeof(int) * 11 * c);
        if (!fi) return -1;
        b->fi = fi;
        b->cap_f = c;
    }
    return 0;
}

/* Concatenated boundary arrays and states of all forests for the batch's
 * setup (quadrants of instance k at base[k] .. base[k+1]-1; face cell ids
 * shifted by the same offsets) from each forest's own arrays ext[k]. */
static int batch_build(const ext_t *ext, int ninst, int n, int *base, batch_bufs_t *b, int *ncell_out,
                       int *nface_out)
{
    const int n2 = n * n;
    size_t nc = 0, nf = 0;
    for (int k = 0; k < ninst; ++k) {
        if (ext[k].rc) return ext[k].rc;
        const int ncell = ext[k].ncell, nface = ext[k].nface;
        const int *cell_info = ext[k].cell_info, *face_info = ext[k].face_info;
        const double *geom = ext[k].geom, *soln = ext[k].soln;
        base[k] = (int)nc;
        if (batch_grow(b, nc + (size_t)ncell, nf + (size_t)nface, n2)) return AMRGPU_ERR_ARG;
        for (int q = 0; q < ncell; ++q) {
            const size_t g = nc + (size_t)q;
            b->ci[3 * g] = cell_info[3 * q];
            b->ci[3 * g + 1] = cell_info[3 * q + 1];
            b->ci[3 * g + 2] = 4 * n2 * (int)g;
            b->cinst[g] = k;
        }
        memcpy(b->geom + 12 * nc, geom, sizeof(double) * 12 * (size_t)ncell);
        memcpy(b->soln + 4 * (size_t)n2 * nc, soln, sizeof(double) * 4 * (size_t)n2 * ncell);
        for (int e = 0; e < nface; ++e) {
            const int *src = face_info + 11 * e;
            int *dst = b->fi + 11 * (nf + (size_t)e);
            for (int j = 0; j < 11; ++j) dst[j] = src[j];
            dst[2] = src[2] + (int)nc;
            for (int j = 3; j < 7; ++j) dst[j] = src[j] >= 0 ? src[j] + (int)nc : src[j];
        }
        nc += (size_t)ncell;
        nf += (size_t)nface;
    }
    base[ninst] = (int)nc;
    *ncell_out = (int)nc;
    *nface_out = (int)nf;
    return AMRGPU_OK;
}

/* Concatenated mesh tables of all forests from each forest's own table
 * tab[k] (quadrants of instance k at base[k] .. base[k+1]-1, neighbour ids
 * shifted by the same offsets), with the instance of every quadrant; a
 * dropped instance contributes no quadrants. */
static int batch_build_tab(const tab_t *tab, const int *dropped, int ninst, int *base, batch_bufs_t *b,
                           int *ncell_out)
{
    size_t nc = 0;
    for (int k = 0; k < ninst; ++k) {
        if (dropped[k]) continue;
        if (tab[k].rc) return AMRGPU_ERR_MESH;
        nc += (size_t)tab[k].nq;
    }
    if (nc > (size_t)INT_MAX / 8) return AMRGPU_ERR_ARG;
    if (nc > b->cap_t || !b->nbr) {
        const size_t c = nc + nc / 2 + 1;
        int *nbr = (int *)realloc(b->nbr, sizeof(int) * 8 * c);
        if (nbr) b->nbr = nbr;
        int *lev = (int *)realloc(b->lev, sizeof(int) * c);
        if (lev) b->lev = lev;
        double *org = (double *)realloc(b->org, sizeof(double) * 2 * c);
        if (org) b->org = org;
        int *tinst = (int *)realloc(b->tinst, sizeof(int) * c);
        if (tinst) b->tinst = tinst;
        int *tsrc = (int *)realloc(b->tsrc, sizeof(int) * c);
        if (tsrc) b->tsrc = tsrc;
        if (!nbr || !lev || !org || !tinst || !tsrc) return AMRGPU_ERR_ARG;
        b->cap_t = c;
    }
    nc = 0;
    for (int k = 0; k < ninst; ++k) {
        const int nq = dropped[k] ? 0 : tab[k].nq, off = (int)nc;
        base[k] = off;
        for (size_t e = 0; e < 8 * (size_t)nq; ++e) {
            const int v = tab[k].nbr[e];
            b->nbr[8 * nc + e] = v >= 0 ? v + off : v;
        }
        memcpy(b->lev + nc, tab[k].level, sizeof(int) * (size_t)nq);
        memcpy(b->org + 2 * nc, tab[k].org, sizeof(double) * 2 * (size_t)nq);
        for (int q = 0; q < nq; ++q) b->tinst[nc + (size_t)q] = k;
        nc += (size_t)nq;
    }
    base[ninst] = (int)nc;
    *ncell_out = (int)nc;
    return AMRGPU_OK;
}

/* Per-instance regrid work of one batch regrid, shared by the host threads:
 * instance k (taken in turn from a counter, largest first) is regridded by
 * its tags and its mesh table rebuilt when it has not reached its stop; one
 * that has leaves the device mesh at this regrid and needs neither. */
typedef struct {
    amr_forest_t **f;
    const int *done, *tags, *obase;
    int allow, ninst;
    tab_t *tab;
    int *ok;
    int *order;      /* instances by decreasing quadrant count */
    int next;
    pthread_mutex_t lock;
} regrid_job_t;

static void *regrid_worker(void *arg)
{
    regrid_job_t *j = (regrid_job_t *)arg;
    /* the instances are the parallelism here; the caller's thread also runs
     * this worker, so its own setting is restored on the way out */
    const int was_parallel = amr_forest_host_parallel(0);
    for (;;) {
        pthread_mutex_lock(&j->lock);
        const int i = j->next++;
        pthread_mutex_unlock(&j->lock);
        if (i >= j->ninst) break;
        const int k = j->order ? j->order[i] : i;
        j->ok[k] = 1;
        if (j->done[k]) continue;      /* leaves the device mesh at this regrid */
        amr_forest_track_begin(j->f[k], 0);
        const int r = amr_regrid_tags(j->f[k], j->tags + j->obase[k], j->allow);
        j->ok[k] = amr_forest_track_end(j->f[k]) && r >= 0;
        tab_free(&j->tab[k]);
        tab_of(j->f[k], &j->tab[k]);
    }
    amr_forest_host_parallel(was_parallel);
    return NULL;
}

static int host_threads(int work)
{
    const char *e = getenv("AMR_HOST_THREADS");
    long n = e ? strtol(e, NULL, 10) : sysconf(_SC_NPROCESSORS_ONLN);
    if (n > 16) n = 16;
    if (n > work) n = work;
    return n < 1 ? 1 : (int)n;
}

static void *regrid_all_thread(void *arg);

static void regrid_all(regrid_job_t *j)
{
    const int nt = host_threads(j->ninst);
    j->next = 0;
    /* the largest instances first: the threads' last instances are then
     * small ones, and the regrid ends close to the total work / threads
     * (an instance's regrid is serial, and one large instance taken last
     * would hold the others) */
    j->order = (int *)malloc(sizeof(int) * (size_t)j->ninst);
    if (j->order) {     /* insertion sort by old quadrant count, stable */
        for (int k = 0; k < j->ninst; ++k) {
            const int sk = j->obase[k + 1] - j->obase[k];
            int i = k;
            while (i > 0 && j->obase[j->order[i - 1] + 1] - j->obase[j->order[i - 1]] < sk) {
                j->order[i] = j->order[i - 1];
                --i;
            }
            j->order[i] = k;
        }
    }
    pthread_mutex_init(&j->lock, NULL);
    pthread_t th[16];
    int started = 0;
    for (int t = 1; t < nt; ++t)
        if (pthread_create(&th[started], NULL, regrid_worker, j) == 0) ++started;
    regrid_worker(j);
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&j->lock);
    free(j->order);
    j->order = NULL;
}

static void *regrid_all_thread(void *arg)
{
    regrid_all((regrid_job_t *)arg);
    return NULL;
}

/* ---- instance groups: one device context and host thread per group ----- */

typedef struct {               /* a tag source shared by the groups, one call at a time */
    amr_tagger_fn fn;
    void *ctx;
    pthread_mutex_t lock;
} shared_tagger_t;

static int shared_tagger(void *user, struct amrgpu_ctx *ctx, amr_forest_t *f, int *tags)
{
    shared_tagger_t *s = (shared_tagger_t *)user;
    pthread_mutex_lock(&s->lock);
    const int rc = s->fn(s->ctx, ctx, f, tags);
    pthread_mutex_unlock(&s->lock);
    return rc;
}

typedef struct {
    amr_forest_t **f;
    amrgpu_inst_t *inst;
    int n;
    amr_evolve_opts_t o;
    amr_evolve_stats_t *per, tot;
    int device;                /* the caller's HIP device */
    int rc;
} batch_group_t;

static void *batch_group_thread(void *arg)
{
    batch_group_t *g = (batch_group_t *)arg;
    amrgpu_use_device(g->device);
    g->rc = amr_evolve_batch(g->f, g->inst, g->n, &g->o, g->per, &g->tot);
    return NULL;
}

/* Estimated work of an instance: quadrants x 2^(finest level), the steps to a
 * given time growing with the finest level's 1/h. */
static double inst_work(const amr_forest_t *f)
{
    const amr_quad_t *q = amr_forest_quads(f);
    const int nq = amr_forest_num_quads(f);
    int lmax = 0;
    for (int i = 0; i < nq; ++i)
        if (q[i].level > lmax) lmax = q[i].level;
    return (double)nq * ldexp(1.0, lmax);
}

static int evolve_groups(amr_forest_t **f, const amrgpu_inst_t *inst, int ninst, const amr_evolve_opts_t *o,
                         amr_evolve_stats_t *per_inst, amr_evolve_stats_t *total)
{
    int ng = o->groups < ninst ? o->groups : ninst;
    if (ng > AMR_MAX_GROUPS) ng = AMR_MAX_GROUPS;
    batch_group_t grp[AMR_MAX_GROUPS];
    memset(grp, 0, sizeof(grp));
    int *order = (int *)malloc(sizeof(int) * (size_t)ninst);
    int *of = (int *)malloc(sizeof(int) * (size_t)ninst);           /* group of each instance */
    double *w = (double *)malloc(sizeof(double) * (size_t)ninst);
    amr_forest_t **fs = (amr_forest_t **)malloc(sizeof(amr_forest_t *) * (size_t)ninst);
    amrgpu_inst_t *is = (amrgpu_inst_t *)malloc(sizeof(amrgpu_inst_t) * (size_t)ninst);
    amr_evolve_stats_t *ps = (amr_evolve_stats_t *)calloc((size_t)ninst, sizeof(amr_evolve_stats_t));
    int rc = 0;
    if (!order || !of || !w || !fs || !is || !ps) rc = AMRGPU_ERR_ARG;
    shared_tagger_t tg;
    tg.fn = o->tagger;
    tg.ctx = o->tagger_ctx;
    pthread_mutex_init(&tg.lock, NULL);
    const int dev = amrgpu_current_device();
    if (!rc) {
        /* largest estimated work first, each to the group with the least so far */
        double load[AMR_MAX_GROUPS] = {0};
        for (int k = 0; k < ninst; ++k) {
            w[k] = inst_work(f[k]);
            int i = k;
            while (i > 0 && w[order[i - 1]] < w[k]) {
                order[i] = order[i - 1];
                --i;
            }
            order[i] = k;
        }
        for (int i = 0; i < ninst; ++i) {
            int best = 0;
            for (int gi = 1; gi < ng; ++gi)
                if (load[gi] < load[best]) best = gi;
            of[order[i]] = best;
            load[best] += w[order[i]];
        }
        /* each group's instances in their input order, contiguous in fs / is / ps */
        int pos = 0;
        for (int gi = 0; gi < ng; ++gi) {
            batch_group_t *g = &grp[gi];
            g->f = fs + pos;
            g->inst = is + pos;
            g->per = ps + pos;
            for (int k = 0; k < ninst; ++k)
                if (of[k] == gi) {
                    fs[pos] = f[k];
                    is[pos] = inst[k];
                    ++pos;
                    ++g->n;
                }
            g->o = *o;
            g->o.groups = 1;
            g->device = dev;
            if (o->tagger) {
                g->o.tagger = shared_tagger;
                g->o.tagger_ctx = &tg;
            }
        }
        pthread_t th[AMR_MAX_GROUPS];
        int started[AMR_MAX_GROUPS] = {0};
        for (int gi = 1; gi < ng; ++gi)
            started[gi] = pthread_create(&th[gi], NULL, batch_group_thread, &grp[gi]) == 0;
        for (int gi = 1; gi < ng; ++gi)      /* a thread that could not start runs here */
            if (!started[gi]) batch_group_thread(&grp[gi]);
        batch_group_thread(&grp[0]);
        for (int gi = 1; gi < ng; ++gi)
            if (started[gi]) pthread_join(th[gi], NULL);
        for (int gi = 0; gi < ng && !rc; ++gi) rc = grp[gi].rc;
    }
    if (order && of && w && fs && is && ps) {
        amr_evolve_stats_t tot;
        memset(&tot, 0, sizeof(tot));
        for (int gi = 0; gi < ng; ++gi) {
            const amr_evolve_stats_t *t = &grp[gi].tot;
            if (t->steps > tot.steps) tot.steps = t->steps;
            tot.regrids += t->regrids;
            tot.peak_quads += t->peak_quads;
            tot.cell_stages += t->cell_stages;
            if (t->t > tot.t) tot.t = t->t;
            /* the groups overlap: each phase's time is the longest group's */
            if (t->wall_gpu > tot.wall_gpu) tot.wall_gpu = t->wall_gpu;
            if (t->wall_regrid > tot.wall_regrid) tot.wall_regrid = t->wall_regrid;
            if (t->wall_tags > tot.wall_tags) tot.wall_tags = t->wall_tags;
            if (t->wall_forest > tot.wall_forest) tot.wall_forest = t->wall_forest;
            if (t->wall_remesh > tot.wall_remesh) tot.wall_remesh = t->wall_remesh;
        }
        if (total) *total = tot;
        if (per_inst) {
            int pos = 0;
            for (int gi = 0; gi < ng; ++gi)
                for (int k = 0; k < ninst; ++k)
                    if (of[k] == gi) per_inst[k] = ps[pos++];
        }
    }
    pthread_mutex_destroy(&tg.lock);
    free(order); free(of); free(w); free(fs); free(is); free(ps);
    return rc;
}

int amr_evolve_batch(amr_forest_t **f, const amrgpu_inst_t *inst, int ninst, const amr_evolve_opts_t *o,
                     amr_evolve_stats_t *per_inst, amr_evolve_stats_t *total)
{
    if (!f || !inst || ninst <= 0) return AMRGPU_ERR_ARG;
    if (o && o->groups > 1 && ninst > 1) return evolve_groups(f, inst, ninst, o, per_inst, total);
    const int n = amr_forest_nelem_subgrid(f[0]);
    for (int k = 1; k < ninst; ++k)
        if (amr_forest_nelem_subgrid(f[k]) != n) return AMRGPU_ERR_ARG;
    physics_gpu_params_t p;
    physics_gpu_get_params(&p);
    static const int stages[6] = {0, 1, 2, 3, 4, 5};
    const int n2 = n * n, interval = o ? o->regrid_interval : p.regrid_interval;
    int poll = 0;
    for (int k = 0; k < ninst; ++k)
        if (inst[k].t_end > 0.0) poll = 1;
    amr_evolve_stats_t tot;
    memset(&tot, 0, sizeof(tot));
    amr_evolve_stats_t *st = (amr_evolve_stats_t *)calloc((size_t)ninst, sizeof(amr_evolve_stats_t));
    int *base = (int *)malloc(sizeof(int) * (size_t)(ninst + 1));
    int *done = (int *)calloc((size_t)ninst, sizeof(int));
    double *t = (double *)calloc((size_t)ninst, sizeof(double));
    ext_t *ext = (ext_t *)calloc((size_t)ninst, sizeof(ext_t));
    tab_t *tab = (tab_t *)calloc((size_t)ninst, sizeof(tab_t));
    int *ok = (int *)calloc((size_t)ninst, sizeof(int));
    int *obase = (int *)malloc(sizeof(int) * (size_t)(ninst + 1));
    int *dropped = (int *)calloc((size_t)ninst, sizeof(int));   /* left the device mesh */
    batch_bufs_t b;
    memset(&b, 0, sizeof(b));
    double *ind = NULL;
    int *tags = NULL;
    int ncell = 0, nface = 0, rc = 0, istep = 0, cap = 0;
    amrgpu_ctx *g = NULL;
    double t0 = now_s();
    if (!st || !base || !done || !t || !ext || !tab || !ok || !obase || !dropped) rc = AMRGPU_ERR_ARG;
    for (int k = 0; k < ninst && !rc; ++k) rc = ext_of(f[k], &ext[k]);
    if (!rc) rc = batch_build(ext, ninst, n, base, &b, &ncell, &nface);
    if (!rc) {
        g = amrgpu_create();
        if (!g) rc = AMRGPU_ERR_HIP;
    }
    if (!rc)
        rc = amrgpu_setup(g, n, p.time_scheme, ninst, inst, ncell, b.ci, b.cinst, nface, b.fi, b.geom, b.soln);
    {   /* the reference-format arrays serve the setup only; regrids use the mesh tables */
        free(b.ci); free(b.fi); free(b.cinst); free(b.src); free(b.geom); free(b.soln);
        b.ci = b.fi = b.cinst = b.src = NULL;
        b.geom = b.soln = NULL;
        b.cap_c = b.cap_f = 0;
    }
    if (!rc && o && o->tile_ppb > 0) rc = amrgpu_set_tile(g, o->tile_ppb, o->tile_minw, o->tile_design);
    int tuned_at = 0;
    if (!rc) rc = autotune_tile(g, o, ncell, &tuned_at);
    for (int k = 0; k < ninst && !rc; ++k) st[k].peak_quads = base[k + 1] - base[k];
    tot.peak_quads = ncell;
    tot.wall_regrid += now_s() - t0;
    while (!rc) {
        long long nsteps = (long long)p.time_steps - istep;
        long long r = steps_to(istep, interval);
        if (r < nsteps) nsteps = r;
        if (nsteps <= 0) break;
        if (poll && nsteps > 16) nsteps = 16;
        double t1 = now_s();
        double resnorm;
        rc = amrgpu_run(g, (int)nsteps, 1, &resnorm);
        if (rc) break;
        for (int k = 0; k < ninst; ++k) {
            if (done[k]) continue;
            st[k].steps += (int)nsteps;
            st[k].cell_stages += (double)nsteps * (base[k + 1] - base[k]) * n2 * stages[p.time_scheme];
        }
        istep += (int)nsteps;
        tot.steps += (int)nsteps;
        tot.cell_stages += (double)nsteps * ncell * n2 * stages[p.time_scheme];
        rc = amrgpu_get_time(g, t, NULL);
        tot.wall_gpu += now_s() - t1;
        if (rc) break;
        int all = 1;
        for (int k = 0; k < ninst; ++k) {
            if (inst[k].t_end > 0.0 && t[k] >= inst[k].t_end) done[k] = 1;
            if (!done[k]) all = 0;
        }
        if (istep >= p.time_steps || all) break;
        if (interval <= 0 || istep % interval != 0) continue;
        /* regrid every unfinished instance on its own forest */
        double t2 = now_s();
        if (ncell > cap) {
            cap = ncell + ncell / 2;
            free(ind); free(tags);
            ind = (double *)malloc(sizeof(double) * (size_t)cap);
            tags = (int *)malloc(sizeof(int) * (size_t)cap);
            if (!ind || !tags) { rc = AMRGPU_ERR_ARG; break; }
        }
        if (o && o->tagger) {
            rc = o->tagger(o->tagger_ctx, g, NULL, tags);
        } else {
            rc = amrgpu_indicator(g, ind);
            for (int q = 0; q < ncell; ++q) tags[q] = ind[q] > (o ? o->tol : 0.1);
        }
        double t3 = now_s();
        tot.wall_tags += t3 - t2;
        if (rc) break;
        /* old quadrant ranges, the instances regridded on the host threads,
         * then the tracked sources shifted into the old ranges */
        memcpy(obase, base, sizeof(int) * (size_t)(ninst + 1));
        regrid_job_t job;
        memset(&job, 0, sizeof(job));
        job.f = f; job.done = done; job.tags = tags; job.obase = obase;
        job.allow = o ? o->allow_coarsening : 1;
        job.ninst = ninst; job.tab = tab; job.ok = ok;
        const int lag = o ? o->regrid_lag : 0;
        double t_gpu_lag = 0.0;
        if (lag > 0 && lag < interval) {
            /* the host regrids every instance while the GPU advances the old
             * meshes lag more steps; instances that reached their stop before
             * this regrid stay frozen and unregridded */
            pthread_t th;
            const int threaded = pthread_create(&th, NULL, regrid_all_thread, &job) == 0;
            if (!threaded) regrid_all(&job);
            int nl = lag;
            if (istep + nl > p.time_steps) nl = p.time_steps - istep;
            const double t4 = now_s();
            if (nl > 0) rc = amrgpu_run(g, nl, 1, &resnorm);
            if (!rc && nl > 0) {
                for (int k = 0; k < ninst; ++k) {
                    if (done[k]) continue;
                    st[k].steps += nl;
                    st[k].cell_stages += (double)nl * (base[k + 1] - base[k]) * n2 * stages[p.time_scheme];
                }
                istep += nl;
                tot.steps += nl;
                tot.cell_stages += (double)nl * ncell * n2 * stages[p.time_scheme];
                rc = amrgpu_get_time(g, t, NULL);
            }
            t_gpu_lag = now_s() - t4;
            tot.wall_gpu += t_gpu_lag;
            if (threaded) pthread_join(th, NULL);
        } else {
            regrid_all(&job);
        }
        for (int k = 0; k < ninst; ++k) {
            if (!ok[k]) rc = AMRGPU_ERR_MESH;
            if (!done[k]) st[k].regrids++;
        }
        /* instances that reached their stop leave the device mesh: their
         * final state goes to their own forest now, and no later launch
         * carries their cells */
        for (int k = 0; k < ninst && !rc; ++k) {
            if (!done[k] || dropped[k]) continue;
            rc = amrgpu_download_quads(g, obase[k], obase[k + 1] - obase[k], amr_forest_soln(f[k]));
            st[k].t = t[k];
            dropped[k] = 1;
        }
        if (!rc) rc = batch_build_tab(tab, dropped, ninst, base, &b, &ncell);
        for (int k = 0; k < ninst && !rc; ++k) {
            if (dropped[k]) continue;
            const int nk = base[k + 1] - base[k];
            const int *s = amr_forest_track_src(f[k]);
            for (int q = 0; q < nk; ++q)
                b.tsrc[base[k] + q] = s[q] >= 0 ? obase[k] + s[q] : -(obase[k] + (-s[q] - 1)) - 1;
            if (nk > st[k].peak_quads) st[k].peak_quads = nk;
        }
        double t4 = now_s();
        tot.wall_forest += t4 - t3 - t_gpu_lag;
        if (!rc) rc = amrgpu_remesh_nbr(g, ncell, b.lev, b.tinst, b.nbr, b.tsrc);
        if (!rc) rc = amrgpu_set_origins(g, b.org);
        tot.wall_remesh += now_s() - t4;
        if (!rc) rc = autotune_tile(g, o, ncell, &tuned_at);
        tot.regrids++;
        if (ncell > tot.peak_quads) tot.peak_quads = ncell;
        tot.wall_regrid += now_s() - t2 - t_gpu_lag;
        if (t_gpu_lag > 0.0) {       /* stops reached during the overlapped steps */
            int all_done = 1;
            for (int k = 0; k < ninst; ++k) {
                if (inst[k].t_end > 0.0 && t[k] >= inst[k].t_end) done[k] = 1;
                if (!done[k]) all_done = 0;
            }
            if (istep >= p.time_steps || all_done) break;
        }
    }
    if (!rc) {
        rc = amrgpu_get_time(g, t, NULL);
        double *all = (double *)malloc(sizeof(double) * 4 * (size_t)n2 * (size_t)(ncell > 0 ? ncell : 1));
        if (!all) rc = AMRGPU_ERR_ARG;
        if (!rc) rc = amrgpu_download_soln(g, all);
        for (int k = 0; k < ninst && !rc; ++k) {
            if (dropped[k]) continue;          /* written when it left the mesh */
            memcpy(amr_forest_soln(f[k]), all + 4 * (size_t)n2 * base[k],
                   sizeof(double) * 4 * (size_t)n2 * (size_t)(base[k + 1] - base[k]));
            st[k].t = t[k];
        }
        free(all);
    }
    if (g) amrgpu_destroy(g);
    if (per_inst && st) memcpy(per_inst, st, sizeof(amr_evolve_stats_t) * (size_t)ninst);
    if (total) *total = tot;
    batch_free(&b);
    if (tab)
        for (int k = 0; k < ninst; ++k) tab_free(&tab[k]);
    free(st); free(base); free(done); free(t); free(ind); free(tags); free(ext); free(tab); free(ok); free(obase);
    free(dropped);
    return rc;
}
