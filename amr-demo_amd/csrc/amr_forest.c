/*
 * amr_forest.c -- host quadtree forest (p4est stand-in) and the AMR->physics
 * boundary builder.  See include/amr_forest.h.
 *
 * Quadrants live in one array sorted by (tree, Morton key of (x, y)); a
 * complete linear quadtree per tree, like p4est's local quadrant arrays.
 * Neighbour queries are binary searches for the leaf containing a point.
 */
#include "amr_forest.h"
#include "amr_internal.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct amr_forest {
    int ntx, nty, periodic[2];
    double xlo, ylo, scale;
    int n, min_level, max_level;
    int nquad, cap;
    int soln_cap;         /* patches the soln array holds (>= nquad when read) */
    amr_quad_t *q;
    uint64_t *key;        /* Morton key per quadrant (lower-left corner) */
    double *soln;         /* nquad * 4*n*n */
    int *tag, *type;
    /* 1: created by the last refine, 2: by the last coarsen (worklist of the
     * regrid's balance) */
    unsigned char *mark;
    /* boundary arrays (rebuilt by amr_forest_build_external) */
    int *cell_info, *face_info;
    double *geom;
    int nface, ext_cap;
    int *ext_rec;         /* per quadrant up to 4 face records (parallel build) */
    int ext_rec_cap;
    /* regrid tracking: source of every quadrant's patch (see track_begin) */
    int *src;
    int track, track_data, track_ok;
    /* hash index (tree, level, Morton key) -> quadrant, rebuilt lazily after
     * any change of the quadrant array */
    struct hentry *hash;
    int hcap, hvalid;
};

struct hentry {
    uint64_t key;      /* Morton key of the lower-left corner */
    int meta;          /* tree * 32 + level, -1 = empty */
    int idx;
};

static inline int qlen(int level) { return 1 << (AMR_MAXLEVEL - level); }

/* Host threads for the per-quadrant loops (OpenMP), off inside the batch's
 * own per-instance threads. */
static _Thread_local int t_parallel = 1;
#define PAR_MIN 1024

int amr_forest_host_parallel(int on)
{
    const int was = t_parallel;
    t_parallel = on;
    return was;
}

int amr_forest_host_parallel_enabled(void) { return t_parallel; }

static inline uint64_t spread(uint32_t v)
{
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}
static inline uint64_t morton(int x, int y) { return spread((uint32_t)x) | (spread((uint32_t)y) << 1); }

static int patch_size(const amr_forest_t *f) { return 4 * f->n * f->n; }

/* Allocation failure ends the process with a message, as p4est's allocator
 * does: a half-rebuilt forest has no state to return to.  (The balance
 * after a regrid is the exception: it reports failure and leaves the forest
 * as it was.) */
static void *xrealloc(void *p, size_t bytes)
{
    void *r = realloc(p, bytes ? bytes : 1);
    if (!r) {
        fprintf(stderr, "amr_forest: out of memory (%zu bytes)\n", bytes);
        abort();
    }
    return r;
}
static void *xmalloc(size_t bytes) { return xrealloc(NULL, bytes); }

static inline uint64_t hmix(uint64_t key, int meta)
{
    uint64_t z = key ^ ((uint64_t)(uint32_t)meta * 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void hash_build(amr_forest_t *f)
{
    int cap = 16;
    while (cap < 2 * f->nquad) cap <<= 1;
    if (cap != f->hcap) {
        free(f->hash);
        f->hash = (struct hentry *)xmalloc(sizeof(struct hentry) * (size_t)cap);
        f->hcap = cap;
    }
    for (int i = 0; i < cap; ++i) f->hash[i].meta = -1;
    const uint64_t mask = (uint64_t)cap - 1;
    for (int k = 0; k < f->nquad; ++k) {
        const int meta = f->q[k].tree * 32 + f->q[k].level;
        uint64_t h = hmix(f->key[k], meta) & mask;
        while (f->hash[h].meta >= 0) h = (h + 1) & mask;
        f->hash[h].key = f->key[k];
        f->hash[h].meta = meta;
        f->hash[h].idx = k;
    }
    f->hvalid = 1;
}

/* index of the quadrant (tree, level, x, y) with (x, y) its lower-left
 * corner, or -1 */
static int hash_find(amr_forest_t *f, int tree, int level, int x, int y)
{
    if (!f->hvalid) hash_build(f);
    const uint64_t key = morton(x, y);
    const int meta = tree * 32 + level;
    const uint64_t mask = (uint64_t)f->hcap - 1;
    for (uint64_t h = hmix(key, meta) & mask;; h = (h + 1) & mask) {
        const struct hentry *e = &f->hash[h];
        if (e->meta < 0) return -1;
        if (e->meta == meta && e->key == key) return e->idx;
    }
}

static void reserve(amr_forest_t *f, int cap)
{
    if (cap <= f->cap) return;
    int ps = patch_size(f);
    f->q = (amr_quad_t *)xrealloc(f->q, sizeof(amr_quad_t) * (size_t)cap);
    f->key = (uint64_t *)xrealloc(f->key, sizeof(uint64_t) * (size_t)cap);
    f->soln = (double *)xrealloc(f->soln, sizeof(double) * (size_t)ps * (size_t)cap);
    f->tag = (int *)xrealloc(f->tag, sizeof(int) * (size_t)cap);
    f->type = (int *)xrealloc(f->type, sizeof(int) * (size_t)cap);
    f->mark = (unsigned char *)xrealloc(f->mark, (size_t)cap);
    f->cap = cap;
    f->soln_cap = cap;
}

amr_forest_t *amr_forest_new(int ntx, int nty, int periodic_x, int periodic_y,
                             double xlo, double ylo, double scale, int n,
                             int min_level, int max_level)
{
    if (ntx < 1 || nty < 1 || n < 1 || min_level < 0 || max_level < min_level ||
        max_level > AMR_MAXLEVEL - 1)
        return NULL;
    amr_forest_t *f = (amr_forest_t *)calloc(1, sizeof(amr_forest_t));
    if (!f) return NULL;
    f->ntx = ntx; f->nty = nty;
    f->periodic[0] = periodic_x; f->periodic[1] = periodic_y;
    f->xlo = xlo; f->ylo = ylo; f->scale = scale;
    f->n = n; f->min_level = min_level; f->max_level = max_level;
    long long per_tree = 1LL << (2 * min_level);
    long long total = per_tree * ntx * nty;
    if (total > (1LL << 30)) { free(f); return NULL; }
    reserve(f, (int)total);
    int k = 0, m = 1 << min_level, len = qlen(min_level);
    for (int t = 0; t < ntx * nty; ++t) {
        /* Morton order within the tree */
        for (long long z = 0; z < per_tree; ++z) {
            int x = 0, y = 0;
            for (int b = 0; b < min_level; ++b) {
                x |= (int)((z >> (2 * b)) & 1) << b;
                y |= (int)((z >> (2 * b + 1)) & 1) << b;
            }
            (void)m;
            f->q[k].tree = t; f->q[k].level = min_level;
            f->q[k].x = x * len; f->q[k].y = y * len;
            f->key[k] = morton(f->q[k].x, f->q[k].y);
            f->tag[k] = 0; f->type[k] = 0; f->mark[k] = 0;
            ++k;
        }
    }
    f->nquad = k;
    memset(f->soln, 0, sizeof(double) * (size_t)patch_size(f) * (size_t)k);
    return f;
}

void amr_forest_destroy(amr_forest_t *f)
{
    if (!f) return;
    free(f->q); free(f->key); free(f->soln); free(f->tag); free(f->type); free(f->mark);
    free(f->cell_info); free(f->face_info); free(f->geom); free(f->src); free(f->hash); free(f->ext_rec);
    free(f);
}

int amr_forest_num_quads(const amr_forest_t *f) { return f->nquad; }
int amr_forest_nelem_subgrid(const amr_forest_t *f) { return f->n; }
const amr_quad_t *amr_forest_quads(const amr_forest_t *f) { return f->q; }
double *amr_forest_soln(amr_forest_t *f)
{
    /* regrids that leave the data on the device do not move patches; the
     * array only grows here, contents of new entries undefined */
    if (f->soln_cap < f->nquad) {
        f->soln = (double *)xrealloc(f->soln, sizeof(double) * (size_t)patch_size(f) * (size_t)f->nquad);
        f->soln_cap = f->nquad;
    }
    return f->soln;
}
int *amr_forest_tags(amr_forest_t *f) { return f->tag; }
double amr_forest_scale(const amr_forest_t *f) { return f->scale; }

/* p4est_utilities_quad_coordinates (src/amr/src/amr_p4est_utilities.c:15-81)
 * on brick vertices scaled as initialize_grid does
 * (src/amr/src/amr_initialize.c:299-303). */
void amr_forest_quad_geom(const amr_forest_t *f, int k, double xyz[12])
{
    const amr_quad_t *q = &f->q[k];
    int tx = q->tree % f->ntx, ty = q->tree / f->ntx;
    double v[4][3];
    for (int c = 0; c < 4; ++c) {
        int ix = tx + (c & 1), iy = ty + (c >> 1);
        v[c][0] = (double)ix * f->scale + f->xlo;
        v[c][1] = (double)iy * f->scale + f->ylo;
        v[c][2] = 0.0 * f->scale + 0.0;
    }
    const double intsize = 1.0 / (double)AMR_ROOT_LEN;
    double h2 = 0.5 * intsize * (double)qlen(q->level);
    double eta_z = 0.0;
    int kk = 0;
    for (int yi = 0; yi < 2; ++yi) {
        double eta_y = intsize * (double)q->y + h2 * (1.0 + (2 * yi - 1));
        for (int xi = 0; xi < 2; ++xi) {
            double eta_x = intsize * (double)q->x + h2 * (1.0 + (2 * xi - 1));
            for (int j = 0; j < 3; ++j) {
                xyz[3 * kk + j] =
                    ((1. - eta_z) * ((1. - eta_y) * ((1. - eta_x) * v[0][j] + eta_x * v[1][j]) +
                                     eta_y * ((1. - eta_x) * v[2][j] + eta_x * v[3][j])));
            }
            ++kk;
        }
    }
}

/* The lower-left corner of quadrant k: corner 0 of amr_forest_quad_geom,
 * operation for operation (so bit for bit), without the other corners. */
static void quad_origin(const amr_forest_t *f, int k, double *ox, double *oy)
{
    const amr_quad_t *q = &f->q[k];
    int tx = q->tree % f->ntx, ty = q->tree / f->ntx;
    double v[4][2];
    for (int c = 0; c < 4; ++c) {
        int ix = tx + (c & 1), iy = ty + (c >> 1);
        v[c][0] = (double)ix * f->scale + f->xlo;
        v[c][1] = (double)iy * f->scale + f->ylo;
    }
    const double intsize = 1.0 / (double)AMR_ROOT_LEN;
    double h2 = 0.5 * intsize * (double)qlen(q->level);
    double eta_z = 0.0;
    double eta_y = intsize * (double)q->y + h2 * (1.0 + (2 * 0 - 1));
    double eta_x = intsize * (double)q->x + h2 * (1.0 + (2 * 0 - 1));
    double o[2];
    for (int j = 0; j < 2; ++j)
        o[j] = ((1. - eta_z) * ((1. - eta_y) * ((1. - eta_x) * v[0][j] + eta_x * v[1][j]) +
                                eta_y * ((1. - eta_x) * v[2][j] + eta_x * v[3][j])));
    *ox = o[0];
    *oy = o[1];
}

static int cmp_pos(const amr_forest_t *f, int k, int tree, uint64_t key)
{
    if (f->q[k].tree != tree) return f->q[k].tree < tree ? -1 : 1;
    if (f->key[k] != key) return f->key[k] < key ? -1 : 1;
    return 0;
}

int amr_forest_find_leaf(const amr_forest_t *f, int tree, int x, int y)
{
    uint64_t key = morton(x, y);
    int lo = 0, hi = f->nquad - 1, best = -1;
    while (lo <= hi) {
        int mid = lo + (hi - lo) / 2;
        if (cmp_pos(f, mid, tree, key) <= 0) { best = mid; lo = mid + 1; }
        else hi = mid - 1;
    }
    if (best < 0) return -1;
    const amr_quad_t *q = &f->q[best];
    int len = qlen(q->level);
    if (q->tree != tree || x < q->x || y < q->y || (long long)x >= (long long)q->x + len ||
        (long long)y >= (long long)q->y + len)
        return -1;
    return best;
}

/* Leaf containing the point (tree, x, y), looked up first at levels
 * level, level - 1 and level + 1 through the hash index (the only
 * possibilities across a face or corner of a 2:1 balanced forest), then by
 * binary search. */
static int leaf_near(amr_forest_t *f, int tree, int x, int y, int level)
{
    static const int order[3] = {0, -1, 1};
    for (int o = 0; o < 3; ++o) {
        const int l = level + order[o];
        if (l < 0 || l >= AMR_MAXLEVEL) continue;
        const int len = qlen(l);
        const int k = hash_find(f, tree, l, x & ~(len - 1), y & ~(len - 1));
        if (k >= 0) return k;
    }
    return amr_forest_find_leaf(f, tree, x, y);
}

/* Same-size neighbour region of quadrant q in direction (dx, dy); returns 0
 * if it lies outside a non-periodic domain. */
static int neighbour_region(const amr_forest_t *f, const amr_quad_t *q, int dx, int dy,
                            int *tree, int *nx, int *ny)
{
    long long len = qlen(q->level);
    long long x = q->x + dx * len, y = q->y + dy * len;
    int tx = q->tree % f->ntx, ty = q->tree / f->ntx;
    if (x < 0) { x += AMR_ROOT_LEN; tx -= 1; }
    else if (x >= AMR_ROOT_LEN) { x -= AMR_ROOT_LEN; tx += 1; }
    if (y < 0) { y += AMR_ROOT_LEN; ty -= 1; }
    else if (y >= AMR_ROOT_LEN) { y -= AMR_ROOT_LEN; ty += 1; }
    if (tx < 0 || tx >= f->ntx) {
        if (!f->periodic[0]) return 0;
        tx = (tx + f->ntx) % f->ntx;
    }
    if (ty < 0 || ty >= f->nty) {
        if (!f->periodic[1]) return 0;
        ty = (ty + f->nty) % f->nty;
    }
    *tree = ty * f->ntx + tx;
    *nx = (int)x; *ny = (int)y;
    return 1;
}

static const int SIDE_DX[4] = {-1, 1, 0, 0};
static const int SIDE_DY[4] = {0, 0, -1, 1};

/* Classify side s of quadrant k: returns 0 boundary, 1 same-size neighbour
 * (nb[0]), 2 coarser neighbour (nb[0]), 3 two finer neighbours (nb[0], nb[1])
 * in order of increasing coordinate along the face, -1 on an inconsistency. */
static int classify_side(amr_forest_t *f, int k, int s, int nb[2])
{
    const amr_quad_t *q = &f->q[k];
    int t, nx, ny;
    if (!neighbour_region(f, q, SIDE_DX[s], SIDE_DY[s], &t, &nx, &ny)) return 0;
    int L = leaf_near(f, t, nx, ny, q->level);
    if (L < 0) return -1;
    if (f->q[L].level == q->level) { nb[0] = L; return 1; }
    if (f->q[L].level < q->level) { nb[0] = L; return 2; }
    int h = qlen(q->level) / 2;
    int ax = nx, ay = ny, bx = nx, by = ny;
    switch (s) {
    case 0: ax = nx + h; bx = nx + h; by = ny + h; break;  /* neighbour's xhi children */
    case 1: by = ny + h; break;                             /* neighbour's xlo children */
    case 2: ay = ny + h; by = ny + h; bx = nx + h; break;   /* neighbour's yhi children */
    case 3: bx = nx + h; break;                             /* neighbour's ylo children */
    }
    nb[0] = leaf_near(f, t, ax, ay, q->level + 1);
    nb[1] = leaf_near(f, t, bx, by, q->level + 1);
    if (nb[0] < 0 || nb[1] < 0) return -1;
    if (f->q[nb[0]].level != q->level + 1 || f->q[nb[1]].level != q->level + 1) return -1;
    return 3;
}

int amr_forest_build_external(amr_forest_t *f, int *ncell, int *nghost, int *nface, int *dof,
                              int **cell_info, int **face_info, double **geom, double **soln)
{
    int nq = f->nquad, ps = patch_size(f);
    if (nq > f->ext_cap) {
        const int c = nq + nq / 4 + 16;
        free(f->cell_info); free(f->face_info); free(f->geom);
        f->cell_info = (int *)xmalloc(sizeof(int) * 3 * (size_t)c);
        f->face_info = (int *)xmalloc(sizeof(int) * 11 * 4 * (size_t)c);
        f->geom = (double *)xmalloc(sizeof(double) * 12 * (size_t)c);
        f->ext_cap = c;
    }
    if (nq + 1 > f->ext_rec_cap) {
        const int c = nq + nq / 4 + 16;
        free(f->ext_rec);
        f->ext_rec = (int *)xmalloc(sizeof(int) * 45 * (size_t)c);    /* 4 records of 11 + count */
        f->ext_rec_cap = c;
    }
    if (!f->hvalid) hash_build(f);               /* shared read-only below */
    int err = 0;
    /* each quadrant's sides, in side order, into its own slots */
#pragma omp parallel for schedule(static) reduction(| : err) if (nq >= PAR_MIN && t_parallel)
    for (int k = 0; k < nq; ++k) {
        f->cell_info[3 * k + 0] = f->type[k];
        f->cell_info[3 * k + 1] = f->q[k].level;
        f->cell_info[3 * k + 2] = ps * k;
        amr_forest_quad_geom(f, k, &f->geom[12 * k]);
        int *rec = &f->ext_rec[45 * k];
        int cnt = 0;
        for (int s = 0; s < 4; ++s) {
            int nb[2] = {-1, -1};
            int c = classify_side(f, k, s, nb);
            if (c < 0) { err = 1; continue; }
            if (c == 2) continue;                      /* emitted by the coarse side */
            if (c == 1 && (s % 2) == 0) continue;      /* emitted from the odd side */
            int *fi = &rec[1 + 11 * cnt];
            for (int e = 0; e < 11; ++e) fi[e] = -1;
            fi[1] = s;
            fi[2] = k;
            if (c == 0) {
                fi[0] = 1;
            } else if (c == 1) {
                fi[0] = 2;
                fi[3] = nb[0]; fi[7] = s ^ 1;
            } else {
                fi[0] = 3;
                fi[3] = nb[0]; fi[4] = nb[1];
                fi[7] = s ^ 1; fi[8] = s ^ 1;
            }
            ++cnt;
        }
        rec[0] = cnt;
    }
    /* quadrant order, as the serial loop */
    int nf = 0;
    for (int k = 0; k < nq; ++k) {
        const int cnt = f->ext_rec[45 * k];
        f->ext_rec[45 * k] = nf;                      /* now: first face of quadrant k */
        nf += cnt;
    }
#pragma omp parallel for schedule(static) if (nq >= PAR_MIN && t_parallel)
    for (int k = 0; k < nq; ++k) {
        const int first = f->ext_rec[45 * k];
        const int cnt = (k + 1 < nq ? f->ext_rec[45 * (k + 1)] : nf) - first;
        memcpy(&f->face_info[11 * (size_t)first], &f->ext_rec[45 * k + 1], sizeof(int) * 11 * (size_t)cnt);
    }
    f->nface = nf;
    *ncell = nq; *nghost = 0; *nface = nf; *dof = ps * nq;
    *cell_info = f->cell_info; *face_info = f->face_info; *geom = f->geom; *soln = amr_forest_soln(f);
    return err ? -1 : 0;
}

int amr_forest_build_nbr(amr_forest_t *f, int *nbr, int *level, double *org)
{
    const int nq = f->nquad;
    if (!f->hvalid) hash_build(f);               /* shared read-only below */
    int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err) if (nq >= PAR_MIN && t_parallel)
    for (int k = 0; k < nq; ++k) {
        level[k] = f->q[k].level;
        quad_origin(f, k, &org[2 * (size_t)k], &org[2 * (size_t)k + 1]);
        for (int s = 0; s < 4; ++s) {
            int nb[2] = {-1, -1};
            const int c = classify_side(f, k, s, nb);
            int *e = &nbr[8 * (size_t)k + 2 * s];
            if (c < 0) { err = 1; e[0] = e[1] = -1; continue; }
            e[0] = c == 0 ? -1 : nb[0];                 /* boundary; same size or coarser; finer */
            e[1] = c == 3 ? nb[1] : -1;
        }
    }
    return err ? -1 : 0;
}

/* Replace the quadrant array by applying refine flags (one level). */
int amr_forest_refine(amr_forest_t *f, const int *flags)
{
    if (!f->track || f->track_data) (void)amr_forest_soln(f);    /* patches are copied: full array */
    int nq = f->nquad, ps = patch_size(f), cnt = 0;
    for (int k = 0; k < nq; ++k) if (flags[k] && f->q[k].level < f->max_level) ++cnt;
    if (!cnt) return 0;
    int nn = nq + 3 * cnt;
    amr_quad_t *q = (amr_quad_t *)xmalloc(sizeof(amr_quad_t) * (size_t)nn);
    uint64_t *key = (uint64_t *)xmalloc(sizeof(uint64_t) * (size_t)nn);
    const int move = !f->track || f->track_data;
    double *soln = move ? (double *)xmalloc(sizeof(double) * (size_t)ps * (size_t)nn) : NULL;
    int *tag = (int *)xmalloc(sizeof(int) * (size_t)nn);
    int *type = (int *)xmalloc(sizeof(int) * (size_t)nn);
    unsigned char *mark = (unsigned char *)xmalloc((size_t)nn);
    int *src = f->track ? (int *)xmalloc(sizeof(int) * (size_t)nn) : NULL;
    int j = 0;
    for (int k = 0; k < nq; ++k) {
        if (flags[k] && f->q[k].level < f->max_level) {
            int h = qlen(f->q[k].level + 1);
            for (int c = 0; c < 4; ++c) {
                q[j].tree = f->q[k].tree;
                q[j].level = f->q[k].level + 1;
                q[j].x = f->q[k].x + (c & 1) * h;
                q[j].y = f->q[k].y + (c >> 1) * h;
                key[j] = morton(q[j].x, q[j].y);
                /* solver_refine_operator_2d.F90:49-52: children = parent */
                if (move) memcpy(&soln[(size_t)ps * j], &f->soln[(size_t)ps * k], sizeof(double) * ps);
                if (src) src[j] = f->src[k];
                tag[j] = f->tag[k];
                type[j] = f->type[k];
                mark[j] = 1;
                ++j;
            }
        } else {
            q[j] = f->q[k]; key[j] = f->key[k];
            if (move) memcpy(&soln[(size_t)ps * j], &f->soln[(size_t)ps * k], sizeof(double) * ps);
            if (src) src[j] = f->src[k];
            tag[j] = f->tag[k]; type[j] = f->type[k];
            mark[j] = f->mark[k];
            ++j;
        }
    }
    free(f->q); free(f->key); free(f->tag); free(f->type); free(f->mark);
    f->q = q; f->key = key; f->tag = tag; f->type = type; f->mark = mark;
    if (move) {
        free(f->soln);
        f->soln = soln;
        f->soln_cap = nn;
    }
    if (src) { free(f->src); f->src = src; }
    f->nquad = nn; f->cap = nn;
    f->hvalid = 0;
    return cnt;
}

static int is_family(const amr_forest_t *f, int k)
{
    if (k + 3 >= f->nquad) return 0;
    const amr_quad_t *a = &f->q[k];
    if (a->level == 0) return 0;
    int h = qlen(a->level);
    if ((a->x % (2 * h)) || (a->y % (2 * h))) return 0;
    for (int c = 1; c < 4; ++c) {
        const amr_quad_t *b = &f->q[k + c];
        if (b->tree != a->tree || b->level != a->level) return 0;
        if (b->x != a->x + (c & 1) * h || b->y != a->y + (c >> 1) * h) return 0;
    }
    return 1;
}

int amr_forest_coarsen(amr_forest_t *f)
{
    if (!f->track || f->track_data) (void)amr_forest_soln(f);    /* patches are moved: full array */
    int nq = f->nquad, ps = patch_size(f), j = 0, cnt = 0;
    for (int k = 0; k < nq;) {
        int fam = is_family(f, k) && f->q[k].level > f->min_level;
        if (fam) {
            for (int c = 0; c < 4; ++c) if (f->tag[k + c] > 0) fam = 0;
        }
        if (fam) {
            amr_quad_t p = f->q[k];
            int h = qlen(p.level - 1);
            p.level -= 1;
            p.x = (p.x / h) * h; p.y = (p.y / h) * h;
            if (!f->track || f->track_data) {
                double *dst = &f->soln[(size_t)ps * j];
                const double *c1 = &f->soln[(size_t)ps * k], *c2 = c1 + ps, *c3 = c2 + ps, *c4 = c3 + ps;
                /* solver_coarsen_operator_2d.F90:49-50 */
                for (int e = 0; e < ps; ++e) {
                    double s = c1[e] + c2[e] + c3[e] + c4[e];
                    dst[e] = 0.25 * s;
                }
            }
            if (f->track) {
                /* a family is always four consecutive copies of an old family
                 * (children refined in this regrid inherit their parent's tag
                 * and are never coarsened in it) */
                int s0 = f->src[k];
                if (s0 < 0 || f->src[k + 1] != s0 + 1 || f->src[k + 2] != s0 + 2 ||
                    f->src[k + 3] != s0 + 3)
                    f->track_ok = 0;
                f->src[j] = -(s0 + 1);
            }
            f->q[j] = p; f->key[j] = morton(p.x, p.y);
            f->tag[j] = 0;
            f->type[j] = f->type[k + 3];  /* last child's type (amr_regrid.c:264-268) */
            f->mark[j] = 2;
            ++j; k += 4; ++cnt;
        } else {
            if (j != k) {
                f->q[j] = f->q[k]; f->key[j] = f->key[k];
                if (!f->track || f->track_data)
                    memmove(&f->soln[(size_t)ps * j], &f->soln[(size_t)ps * k], sizeof(double) * ps);
                if (f->track) f->src[j] = f->src[k];
                f->tag[j] = f->tag[k]; f->type[j] = f->type[k]; f->mark[j] = f->mark[k];
            }
            ++j; ++k;
        }
    }
    f->nquad = j;
    f->hvalid = 0;
    return cnt;
}

static int balance_flags(amr_forest_t *f, int *flags)
{
    int any = 0;
    memset(flags, 0, sizeof(int) * (size_t)f->nquad);
    if (!f->hvalid) hash_build(f);
#pragma omp parallel for schedule(static) reduction(| : any) if (f->nquad >= PAR_MIN && t_parallel)
    for (int k = 0; k < f->nquad; ++k) {
        const amr_quad_t *q = &f->q[k];
        if (q->level < 2) continue;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dx && !dy) continue;
                int t, nx, ny;
                if (!neighbour_region(f, q, dx, dy, &t, &nx, &ny)) continue;
                int L = leaf_near(f, t, nx, ny, q->level);
                if (L >= 0 && f->q[L].level < q->level - 1) { __atomic_store_n(&flags[L], 1, __ATOMIC_RELAXED); any = 1; }
            }
    }
    return any;
}

int amr_forest_balance(amr_forest_t *f)
{
    int total = 0;
    for (;;) {
        int *flags = (int *)xmalloc(sizeof(int) * (size_t)(f->nquad > 0 ? f->nquad : 1));
        int any = balance_flags(f, flags);
        if (any) {
            /* balance may refine past max_level only if the forest already
             * violated it; refine ignores max_level here */
            int save = f->max_level;
            f->max_level = AMR_MAXLEVEL - 1;
            total += amr_forest_refine(f, flags);
            f->max_level = save;
        }
        free(flags);
        if (!any) break;
    }
    return total;
}

int amr_forest_is_balanced(amr_forest_t *f)
{
    int *flags = (int *)xmalloc(sizeof(int) * (size_t)(f->nquad > 0 ? f->nquad : 1));
    int any = balance_flags(f, flags);
    free(flags);
    return !any;
}

/* ---- balance after a regrid, as a worklist over a dynamic node set ------
 * Only quadrants this regrid created (marks: 1 refined child, 2 coarsened
 * parent) and what balance refines can take part in a violation of a forest
 * that was balanced before.  The tree nodes -- leaves and their ancestors --
 * live in an open-addressing set keyed by (tree, level, corner) with a leaf
 * flag, so each neighbour question is one lookup: the leaf containing a point
 * is coarser than level l exactly when no node of level l contains it, and
 * finer than l exactly when the node of level l containing it is internal.
 * Refining a leaf flips its flag and inserts its four children, so the ripple
 * needs no array rebuilds.  The Morton-ordered arrays are emitted once at the
 * end by expanding every refined original quadrant into its leaves (children
 * in z order).  The result is the same unique balanced closure as
 * amr_forest_balance's full passes. */
typedef struct {
    uint64_t key;
    int meta;          /* tree * 32 + level; -1 empty */
    int leaf;
} nodeent_t;

typedef struct {
    nodeent_t *e;
    uint64_t mask;
    int used;
    int oom;           /* an allocation failed: the set is no longer complete */
} nodeset_t;

static nodeent_t *ns_find(const nodeset_t *S, int meta, uint64_t key)
{
    for (uint64_t h = hmix(key, meta) & S->mask;; h = (h + 1) & S->mask) {
        nodeent_t *e = &S->e[h];
        if (e->meta == -1) return NULL;
        if (e->meta == meta && e->key == key) return e;
    }
}

static int ns_alloc(nodeset_t *S, uint64_t cap)
{
    S->e = (nodeent_t *)malloc(sizeof(nodeent_t) * (size_t)cap);
    S->mask = cap - 1;
    S->used = 0;
    S->oom = S->e == NULL;
    if (S->oom) return -1;
    for (uint64_t i = 0; i < cap; ++i) S->e[i].meta = -1;
    return 0;
}

static void ns_put(nodeset_t *S, int meta, uint64_t key, int leaf)
{
    uint64_t h = hmix(key, meta) & S->mask;
    while (S->e[h].meta != -1) h = (h + 1) & S->mask;
    S->e[h].key = key;
    S->e[h].meta = meta;
    S->e[h].leaf = leaf;
    S->used++;
}

/* insert a node known to be absent, keeping the load at most 1/2 */
static void ns_insert(nodeset_t *S, int meta, uint64_t key, int leaf)
{
    if (S->oom) return;
    if (2 * (uint64_t)(S->used + 1) > S->mask + 1) {
        nodeset_t T;
        if (ns_alloc(&T, 2 * (S->mask + 1))) { S->oom = 1; return; }
        for (uint64_t i = 0; i <= S->mask; ++i)
            if (S->e[i].meta != -1) ns_put(&T, S->e[i].meta, S->e[i].key, S->e[i].leaf);
        free(S->e);
        *S = T;
    }
    ns_put(S, meta, key, leaf);
}

/* the node of level l containing (tree, x, y), or NULL */
static inline nodeent_t *ns_at(const nodeset_t *S, int tree, int l, int x, int y)
{
    const int len = qlen(l);
    return ns_find(S, tree * 32 + l, morton(x & ~(len - 1), y & ~(len - 1)));
}

typedef struct {
    amr_quad_t *q;
    unsigned char *kind;   /* 1: leaf, check for too-coarse neighbours; 2: also too-fine;
                              3: a refined quadrant, check for neighbours too coarse
                              for its children */
    int n, cap, oom;
} qqueue_t;

static void qq_push(qqueue_t *Q, amr_quad_t q, unsigned char kind)
{
    if (Q->oom) return;
    if (Q->n == Q->cap) {
        const int cap = Q->cap ? 2 * Q->cap : 256;
        amr_quad_t *nq = (amr_quad_t *)realloc(Q->q, sizeof(amr_quad_t) * (size_t)cap);
        if (nq) Q->q = nq;
        unsigned char *nk = nq ? (unsigned char *)realloc(Q->kind, (size_t)cap) : NULL;
        if (nk) Q->kind = nk;
        if (!nq || !nk) { Q->oom = 1; return; }
        Q->cap = cap;
    }
    Q->q[Q->n] = q;
    Q->kind[Q->n] = kind;
    Q->n++;
}

/* neighbour_region with q's tree coordinates (tx, ty) already known */
static inline int near_region(const amr_forest_t *f, const amr_quad_t *q, int tx, int ty, int dx, int dy,
                              int *tree, int *nx, int *ny)
{
    const int len = qlen(q->level);
    int x = q->x + dx * len, y = q->y + dy * len;
    if (x < 0) { x += AMR_ROOT_LEN; tx -= 1; }
    else if (x >= AMR_ROOT_LEN) { x -= AMR_ROOT_LEN; tx += 1; }
    if (y < 0) { y += AMR_ROOT_LEN; ty -= 1; }
    else if (y >= AMR_ROOT_LEN) { y -= AMR_ROOT_LEN; ty += 1; }
    if (tx < 0 || tx >= f->ntx) {
        if (!f->periodic[0]) return 0;
        tx = tx < 0 ? tx + f->ntx : tx - f->ntx;
    }
    if (ty < 0 || ty >= f->nty) {
        if (!f->periodic[1]) return 0;
        ty = ty < 0 ? ty + f->nty : ty - f->nty;
    }
    *tree = ty * f->ntx + tx;
    *nx = x; *ny = y;
    return 1;
}

static void refine_leaf(nodeset_t *S, qqueue_t *Q, nodeent_t *e, amr_quad_t q, int *count)
{
    e->leaf = 0;     /* before the inserts, which may move the table */
    const int h = qlen(q.level + 1);
    for (int c = 0; c < 4; ++c) {
        amr_quad_t ch = q;
        ch.level = q.level + 1;
        ch.x = q.x + (c & 1) * h;
        ch.y = q.y + (c >> 1) * h;
        ns_insert(S, ch.tree * 32 + ch.level, morton(ch.x, ch.y), 1);
    }
    qq_push(Q, q, 3);    /* the four children share the parent's neighbourhood */
    (*count)++;
}

/* emission of the balanced arrays */
typedef struct {
    amr_quad_t *q;
    uint64_t *key;
    int *tag, *type, *src;
    unsigned char *mark;
    double *soln;
    int n, cap, move, ps;
} emit_t;

static void emit_leaves(const amr_forest_t *f, const nodeset_t *S, emit_t *E, amr_quad_t q, int k)
{
    if (!ns_find(S, q.tree * 32 + q.level, morton(q.x, q.y))->leaf) {
        const int h = qlen(q.level + 1);
        for (int c = 0; c < 4; ++c) {
            amr_quad_t ch = q;
            ch.level = q.level + 1;
            ch.x = q.x + (c & 1) * h;
            ch.y = q.y + (c >> 1) * h;
            emit_leaves(f, S, E, ch, k);
        }
        return;
    }
    if (E->n >= E->cap) return;       /* cannot happen: the capacity is the exact leaf count */
    const int j = E->n++;
    E->q[j] = q;
    E->key[j] = morton(q.x, q.y);
    E->tag[j] = f->tag[k];            /* balance-refined children inherit, as refine */
    E->type[j] = f->type[k];
    E->mark[j] = 0;
    if (E->src) E->src[j] = f->src[k];
    if (E->move) memcpy(&E->soln[(size_t)E->ps * j], &f->soln[(size_t)E->ps * k], sizeof(double) * (size_t)E->ps);
}

int amr_forest_balance_marked(amr_forest_t *f)
{
    const int nq = f->nquad;
    nodeset_t S;
    uint64_t cap = 16;
    while (cap < 4 * (uint64_t)nq) cap <<= 1;
    if (ns_alloc(&S, cap)) return -1;
    qqueue_t Q = {NULL, NULL, 0, 0, 0};
    for (int k = 0; k < nq; ++k) {
        const amr_quad_t *q = &f->q[k];
        ns_insert(&S, q->tree * 32 + q->level, f->key[k], 1);
        /* ancestors up to the first one already present; in Morton order
         * the parent of a child other than the first is present already */
        const int own = q->level > 0 ? qlen(q->level) : 0;
        for (int l = (q->x | q->y) & own ? -1 : q->level - 1; l >= 0; --l) {
            const int len = qlen(l);
            const int meta = q->tree * 32 + l;
            const uint64_t key = morton(q->x & ~(len - 1), q->y & ~(len - 1));
            if (ns_find(&S, meta, key)) break;
            ns_insert(&S, meta, key, 0);
        }
        if (f->mark[k] == 2) {
            qq_push(&Q, *q, 2);
        } else if (f->mark[k] && !((q->x | q->y) & qlen(q->level))) {
            /* the first child of a quadrant refined by this regrid: one
             * check of the parent's neighbourhood stands for all four */
            const int pl = qlen(q->level - 1);
            amr_quad_t par = {q->tree, q->level - 1, q->x & ~(pl - 1), q->y & ~(pl - 1)};
            qq_push(&Q, par, 3);
        }
    }
    int refined = 0;
    for (int it = 0; it < Q.n && !S.oom && !Q.oom; ++it) {
        const amr_quad_t q = Q.q[it];
        const unsigned char kind = Q.kind[it];
        const nodeent_t *self = ns_find(&S, q.tree * 32 + q.level, morton(q.x, q.y));
        if (kind != 3 && !self->leaf) continue;                                /* refined since */
        /* lc: level of the leaves whose neighbours are checked (q's children
         * for kind 3); a neighbour leaf is too coarse below lc - 1 */
        const int lc = kind == 3 ? q.level + 1 : q.level;
        int recheck = 0, too_coarse_self = 0;
        const int len = qlen(q.level);
        /* q's parent is not a leaf, so every leaf inside it is at q's level
         * or finer: the directions towards the siblings are never too coarse */
        const int sx_dir = (q.x & len) ? -1 : 1, sy_dir = (q.y & len) ? -1 : 1;
        const int qtx = q.tree % f->ntx, qty = q.tree / f->ntx;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if ((!dx && !dy) || S.oom || Q.oom) continue;
                const int sibling = q.level > 0 && (!dx || dx == sx_dir) && (!dy || dy == sy_dir);
                if (sibling && kind != 2) continue;
                int t, nx, ny;
                if (!near_region(f, &q, qtx, qty, dx, dy, &t, &nx, &ny)) continue;
                if (!sibling && lc >= 2 && !ns_at(&S, t, lc - 1, nx, ny)) {
                    /* the neighbour leaf containing (nx, ny) is coarser than
                     * lc - 1: find it and refine it */
                    for (int l = lc - 2; l >= 0; --l) {
                        nodeent_t *e = ns_at(&S, t, l, nx, ny);
                        if (e && e->leaf) {
                            const int ln = qlen(l);
                            amr_quad_t L = {t, l, nx & ~(ln - 1), ny & ~(ln - 1)};
                            refine_leaf(&S, &Q, e, L, &refined);
                            recheck = 1;
                            break;
                        }
                    }
                }
                if (kind == 2 && !too_coarse_self) {
                    /* a neighbour two or more levels finer touches q when one
                     * of the level + 1 nodes along the side (two; one at a
                     * corner) is internal: one point in each */
                    const int px = dx > 0 ? nx : nx + len - 1, py = dy > 0 ? ny : ny + len - 1;
                    const int ns = (dx && dy) ? 1 : 2;
                    for (int i = 0; i < ns && !too_coarse_self; ++i) {
                        const int off = i * (len / 2);
                        const int sx = dx ? px : nx + off, sy = dy ? py : ny + off;
                        const nodeent_t *e = ns_at(&S, t, q.level + 1, sx, sy);
                        if (e && !e->leaf) too_coarse_self = 1;
                    }
                }
            }
        if (too_coarse_self) {
            nodeent_t *e = ns_find(&S, q.tree * 32 + q.level, morton(q.x, q.y));
            refine_leaf(&S, &Q, e, q, &refined);
        } else if (recheck) {
            qq_push(&Q, q, kind == 3 ? 3 : 1);
        }
    }
    if (S.oom || Q.oom) refined = -1;      /* the forest is left as refine/coarsen made it */
    if (refined > 0) {
        emit_t E;
        memset(&E, 0, sizeof(E));
        E.cap = nq + 3 * refined;           /* each balance refinement adds three leaves */
        E.q = (amr_quad_t *)malloc(sizeof(amr_quad_t) * (size_t)E.cap);
        E.key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)E.cap);
        E.tag = (int *)malloc(sizeof(int) * (size_t)E.cap);
        E.type = (int *)malloc(sizeof(int) * (size_t)E.cap);
        E.mark = (unsigned char *)malloc((size_t)E.cap);
        E.src = f->track ? (int *)malloc(sizeof(int) * (size_t)E.cap) : NULL;
        E.move = !f->track || f->track_data;
        E.ps = patch_size(f);
        if (E.move) {
            (void)amr_forest_soln(f);
            E.soln = (double *)malloc(sizeof(double) * (size_t)E.ps * (size_t)E.cap);
        }
        if (!E.q || !E.key || !E.tag || !E.type || !E.mark || (f->track && !E.src) || (E.move && !E.soln)) {
            free(E.q); free(E.key); free(E.tag); free(E.type); free(E.mark); free(E.src); free(E.soln);
            free(S.e);
            free(Q.q);
            free(Q.kind);
            return -1;
        }
        for (int k = 0; k < nq; ++k) emit_leaves(f, &S, &E, f->q[k], k);
        free(f->q); free(f->key); free(f->tag); free(f->type); free(f->mark);
        f->q = E.q; f->key = E.key; f->tag = E.tag; f->type = E.type; f->mark = E.mark;
        if (E.src) { free(f->src); f->src = E.src; }
        if (E.move) { free(f->soln); f->soln = E.soln; f->soln_cap = E.cap; }
        f->nquad = E.n;
        f->cap = E.cap;
        f->hvalid = 0;
    } else if (refined == 0) {
        memset(f->mark, 0, (size_t)(nq > 0 ? nq : 1));
    }
    free(S.e);
    free(Q.q);
    free(Q.kind);
    return refined;
}

void amr_forest_clear_marks(amr_forest_t *f)
{
    memset(f->mark, 0, (size_t)(f->nquad > 0 ? f->nquad : 1));
}

void amr_forest_track_begin(amr_forest_t *f, int keep_data)
{
    free(f->src);
    f->src = (int *)xmalloc(sizeof(int) * (size_t)(f->nquad > 0 ? f->nquad : 1));
    for (int k = 0; k < f->nquad; ++k) f->src[k] = k;
    f->track = 1;
    f->track_data = keep_data;
    f->track_ok = 1;
}

const int *amr_forest_track_src(const amr_forest_t *f) { return f->src; }

int amr_forest_track_end(amr_forest_t *f)
{
    f->track = 0;
    f->track_data = 1;
    return f->track_ok;
}

void amr_forest_brick(const amr_forest_t *f, int *ntx, int *nty, int *periodic_x, int *periodic_y,
                      double *xlo, double *ylo, double *scale, int *min_level, int *max_level)
{
    if (ntx) *ntx = f->ntx;
    if (nty) *nty = f->nty;
    if (periodic_x) *periodic_x = f->periodic[0];
    if (periodic_y) *periodic_y = f->periodic[1];
    if (xlo) *xlo = f->xlo;
    if (ylo) *ylo = f->ylo;
    if (scale) *scale = f->scale;
    if (min_level) *min_level = f->min_level;
    if (max_level) *max_level = f->max_level;
}

amr_forest_t *amr_forest_from_quads(int ntx, int nty, int periodic_x, int periodic_y, double xlo,
                                    double ylo, double scale, int n, int min_level, int max_level,
                                    int nquad, const amr_quad_t *quads)
{
    if (nquad < ntx * nty || !quads) return NULL;
    /* validate: order, alignment, and each tree tiled exactly once (the
     * leaves of a complete quadtree in Morton order cover consecutive
     * Morton ranges: each starts where the previous ended) */
    const uint64_t full = (uint64_t)1 << (2 * AMR_MAXLEVEL);
    int tree = 0;
    uint64_t next = 0;
    for (int k = 0; k < nquad; ++k) {
        const amr_quad_t *q = &quads[k];
        if (q->level < 0 || q->level > max_level || q->tree < 0 || q->tree >= ntx * nty) return NULL;
        const int len = qlen(q->level);
        if (q->x < 0 || q->y < 0 || q->x % len || q->y % len || q->x >= AMR_ROOT_LEN || q->y >= AMR_ROOT_LEN)
            return NULL;
        if (q->tree != tree) {
            if (q->tree != tree + 1 || next != full) return NULL;
            tree = q->tree;
            next = 0;
        }
        if (morton(q->x, q->y) != next) return NULL;
        next += (uint64_t)len * (uint64_t)len;
    }
    if (tree != ntx * nty - 1 || next != full) return NULL;
    amr_forest_t *f = amr_forest_new(ntx, nty, periodic_x, periodic_y, xlo, ylo, scale, n, 0, max_level);
    if (!f) return NULL;
    f->min_level = min_level;
    reserve(f, nquad);
    for (int k = 0; k < nquad; ++k) {
        f->q[k] = quads[k];
        f->key[k] = morton(quads[k].x, quads[k].y);
        f->tag[k] = 0;
        f->type[k] = 0;
        f->mark[k] = 0;
    }
    f->nquad = nquad;
    f->hvalid = 0;
    memset(f->soln, 0, sizeof(double) * (size_t)patch_size(f) * (size_t)nquad);
    return f;
}
