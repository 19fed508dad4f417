/*
 * amr_tagger.c -- host side of the learned refinement tagger: collection of
 * labelled patch snapshots from solver runs, and the tagger that
 * amr_evolve_device calls at each regrid.  See include/amr_tagger.h.
 *
 * The reference has no counterpart; the regrid sequence it plugs into is the
 * reference's tag -> refine -> coarsen -> balance order
 * (src/amr/src/amr_regrid.c:15-81), with the keep tag described in
 * amr_regrid_tags.
 */
#define _POSIX_C_SOURCE 200809L
#include "amr_tagger.h"
#include "amr_driver.h"
#include "amr_forest.h"
#include "amr_hip.h"
#include "amr_internal.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int cap;
    double *ind, *indt, *indmax;
    int *y, *tags;
    uint16_t *X;
} host_bufs_t;

static void bufs_free(host_bufs_t *b)
{
    free(b->ind); free(b->indt); free(b->indmax); free(b->y); free(b->tags); free(b->X);
    memset(b, 0, sizeof(*b));
}

static int bufs_grow(host_bufs_t *b, int need)
{
    if (need <= b->cap) return 0;
    int c = need + need / 2;
    bufs_free(b);
    b->ind = (double *)malloc(sizeof(double) * (size_t)c);
    b->indt = (double *)malloc(sizeof(double) * (size_t)c);
    b->indmax = (double *)malloc(sizeof(double) * (size_t)c);
    b->y = (int *)malloc(sizeof(int) * (size_t)c);
    b->tags = (int *)malloc(sizeof(int) * (size_t)c);
    b->X = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)c * AMR_MLP_NIN);
    if (!b->ind || !b->indt || !b->indmax || !b->y || !b->tags || !b->X) return -1;
    b->cap = c;
    return 0;
}

static int finished(const physics_gpu_params_t *p, int istep, double t)
{
    return istep >= p->time_steps || (p->t_end > 0.0 && t >= p->t_end);
}

int amr_tagger_collect(amr_forest_t *f, const amr_collect_opts_t *o, amr_tagset_t *set, int *nrows)
{
    if (!f || !o || !set || o->window <= 0 || o->label_stride <= 0 || o->window % o->label_stride) {
        amr_set_error("amr_tagger_collect: window must be a positive multiple of label_stride");
        return AMRGPU_ERR_ARG;
    }
    if (amr_forest_nelem_subgrid(f) != 8 && amr_forest_nelem_subgrid(f) != 16) {
        amr_set_error("amr_tagger_collect: 8x8 or 16x16 patches required");
        return AMRGPU_ERR_ARG;
    }
    physics_gpu_params_t p;
    physics_gpu_get_params(&p);
    amrgpu_ctx *g = NULL;
    int rc = amr_device_setup(f, o->inst, &g);
    if (rc) return rc;
    host_bufs_t b;
    memset(&b, 0, sizeof(b));
    int istep = 0, rows = 0;
    double t = 0.0;
    const double ctol = o->coarsen_frac * o->tol;
    while (!rc && !finished(&p, istep, t)) {
        const int nq = amr_forest_num_quads(f);
        if (bufs_grow(&b, nq)) {
            amr_set_error("amr_tagger_collect: out of host memory");
            rc = AMRGPU_ERR_ARG;
            break;
        }
        /* snapshot: features and the indicator now */
        rc = amrgpu_tagger_features_host(g, b.X);
        if (!rc) rc = amrgpu_indicator(g, b.ind);
        if (rc) break;
        for (int k = 0; k < nq; ++k) b.indmax[k] = 0.0;
        /* the indicator at every label time of the window, on the same mesh */
        int nlabel = 0;
        for (int s = o->label_stride; s <= o->window && !finished(&p, istep, t); s += o->label_stride) {
            int nst = o->label_stride;
            if (istep + nst > p.time_steps) nst = p.time_steps - istep;
            double resnorm;
            rc = amrgpu_run(g, nst, 1, &resnorm);
            if (!rc) rc = amrgpu_indicator(g, b.indt);
            if (rc) break;
            istep += nst;
            amrgpu_get_time(g, &t, NULL);
            for (int k = 0; k < nq; ++k)
                if (b.indt[k] > b.indmax[k]) b.indmax[k] = b.indt[k];
            nlabel++;
        }
        if (rc || !nlabel) break;
        for (int k = 0; k < nq; ++k)
            b.y[k] = b.indmax[k] > o->tol ? AMR_TAG_REFINE : (b.indmax[k] < ctol ? AMR_TAG_COARSEN : AMR_TAG_KEEP);
        rc = amr_tagset_append(set, b.X, b.y, b.ind, nq);
        if (rc) break;
        rows += nq;
        if (finished(&p, istep, t)) break;
        /* regrid with the plain indicator of the current state */
        for (int k = 0; k < nq; ++k) b.tags[k] = b.indt[k] > o->tol;
        rc = amr_device_regrid(f, g, b.tags, o->allow_coarsening, NULL, NULL);
    }
    if (!rc) rc = amrgpu_download_soln(g, amr_forest_soln(f));
    amrgpu_destroy(g);
    bufs_free(&b);
    if (nrows) *nrows = rows;
    return rc;
}

int amr_mlp_tagger(void *user, amrgpu_ctx *ctx, amr_forest_t *f, int *tags)
{
    amr_mlp_tagger_t *u = (amr_mlp_tagger_t *)user;
    if (!u || !u->mlp || !ctx || !tags) {
        amr_set_error("amr_mlp_tagger: bad argument");
        return AMRGPU_ERR_ARG;
    }
    /* every quadrant of the context (f is NULL when the context holds many
     * forests, amr_evolve_batch) */
    const int nq = amrgpu_num_quads(ctx);
    (void)f;
    int *cls = (int *)malloc(sizeof(int) * (size_t)(nq > 0 ? nq : 1));
    double *ind = u->union_plain ? (double *)malloc(sizeof(double) * (size_t)(nq > 0 ? nq : 1)) : NULL;
    if (!cls || (u->union_plain && !ind)) {
        free(cls); free(ind);
        amr_set_error("amr_mlp_tagger: out of host memory");
        return AMRGPU_ERR_ARG;
    }
    int rc = amr_mlp_predict_ctx(u->mlp, ctx, cls);
    if (!rc && ind) rc = amrgpu_indicator(ctx, ind);
    if (!rc) {
        for (int k = 0; k < nq; ++k) {
            const int c = cls[k];
            u->predicted[c]++;
            int tg = c == AMR_TAG_REFINE ? 1 : (c == AMR_TAG_KEEP ? 2 : 0);
            if (ind && ind[k] > u->tol) {
                u->plain_refine++;
                if (tg != 1) u->added_by_plain++;
                tg = 1;
            }
            tags[k] = tg;
        }
    }
    free(cls);
    free(ind);
    return rc;
}
