/*
 * amr_tagger.h -- learned refinement tagger.
 *
 * The reference refines by density thresholds over a patch (binary flags,
 * src/solver/src/tagging_methods_module.F90:14-71) and coarsens a family when
 * none of its members is tagged and its level is above the minimum
 * (src/amr/src/amr_tag.c:202-220).  It has no learned tagger: everything here
 * is a project addition with no counterpart in the reference.
 *
 * A multilayer perceptron 384 -> 64 -> 64 -> 3 (ReLU, BF16 weights and
 * activations, FP32 accumulation on the MFMA units) reads one 8x8 patch and
 * predicts what the plain gradient indicator (amr_indicator_gradient) will
 * ask for over the next regrid interval:
 *
 *   class 0 "coarsen": indicator below coarsen_frac * tol at every label time
 *   class 1 "keep"   : otherwise
 *   class 2 "refine" : indicator above tol at some label time
 *
 * Features (384 per quadrant, BF16, channel-major c*64 + j*8 + i):
 *   c = 0..3  U_c(i,j) / max over the patch of |U_c|       (rho, rho u, rho v, rho E)
 *   c = 4     (rho(i+1,j) - rho(i-1,j)) / max rho         (x difference)
 *   c = 5     (rho(i,j+1) - rho(i,j-1)) / max rho         (y difference)
 * Past a patch edge the neighbour value is the face neighbour's
 * index-aligned edge cell (the mean of the two when the side has two finer
 * neighbours), or the cell itself on a domain boundary.  Each feature is
 * computed in FP64 (IEEE divide), rounded to FP32, then to BF16 (nearest even).
 * A 16x16 patch is read coarsened to 8x8 (each coarse cell the mean of its
 * 2x2 block, ((a + b) + c) + d times 0.25 with a..d in z-order), its
 * neighbours coarsened alike: the same 384 features of the coarsened state.
 *
 * Parameters (FP32 master copy, AMR_MLP_NPARAM floats):
 *   W1[64][384], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]
 * Training minimises the class-weighted cross-entropy
 *   L = sum_i w[y_i] CE_i / sum_i w[y_i]
 * with Adam; gradients follow the BF16 forward pass exactly (activations and
 * the backward pre-activation gradients are rounded to BF16 where they feed
 * an MFMA; bias gradients sum those rounded values in FP32).
 */
#ifndef AMR_TAGGER_H
#define AMR_TAGGER_H

#include <stdint.h>

#include "amr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

struct amrgpu_ctx;
struct amr_forest;

#define AMR_MLP_NIN 384
#define AMR_MLP_NH 64
#define AMR_MLP_NOUT 3
#define AMR_MLP_NPARAM (64 * 384 + 64 + 64 * 64 + 64 + 3 * 64 + 3)

enum { AMR_TAG_COARSEN = 0, AMR_TAG_KEEP = 1, AMR_TAG_REFINE = 2 };

typedef struct amr_mlp amr_mlp_t;
typedef struct amr_tagset amr_tagset_t;

/* Network with He-uniform weights drawn from `seed` (zero biases). */
amr_mlp_t *amr_mlp_create(unsigned long long seed);
void amr_mlp_destroy(amr_mlp_t *m);
int amr_mlp_get_params(amr_mlp_t *m, float *params);            /* AMR_MLP_NPARAM */
int amr_mlp_set_params(amr_mlp_t *m, const float *params);      /* resets Adam state */

/* Forward pass over n feature rows (host BF16 bits, n x 384); logits n x 3. */
int amr_mlp_forward(amr_mlp_t *m, const uint16_t *X, int n, float *logits);
/* Loss and its gradient w.r.t. every parameter (AMR_MLP_NPARAM floats) over
 * n labelled rows with class weights cw[3]; no parameter update. */
int amr_mlp_grad(amr_mlp_t *m, const uint16_t *X, const int *y, int n, const float cw[3],
                 float *grad, float *loss);

/* Training set of (features, label, indicator now) rows held in device
 * memory. */
amr_tagset_t *amr_tagset_create(void);
void amr_tagset_destroy(amr_tagset_t *s);
int amr_tagset_size(const amr_tagset_t *s);
int amr_tagset_append(amr_tagset_t *s, const uint16_t *X, const int *y, const double *ind_now, int n);
/* Copies rows [first, first+n) back to the host (any pointer may be NULL). */
int amr_tagset_get(const amr_tagset_t *s, int first, int n, uint16_t *X, int *y, double *ind_now);
/* Rows [first, first+n) of `src` appended to `dst` (device to device). */
int amr_tagset_append_from(amr_tagset_t *dst, const amr_tagset_t *src, int first, int n);

/* Features of every quadrant of the context's current state (8x8 or 16x16
 * patches) into X (device pointer, nquad x 384 BF16 bits), on the context's
 * stream. */
int amrgpu_tagger_features(struct amrgpu_ctx *c, uint16_t *d_X);
/* Same, copied to the host (nquad x 384). */
int amrgpu_tagger_features_host(struct amrgpu_ctx *c, uint16_t *X);

/* Adam training over the set: `epochs` passes over a seeded shuffle in
 * minibatches of `batch` rows (a multiple of 64; the remainder of each
 * shuffled pass is skipped).  cw = class weights (NULL: inverse class
 * frequency, clipped to [0.2, 20]).  loss[epochs] receives the mean
 * minibatch loss of each pass (may be NULL). */
int amr_mlp_train(amr_mlp_t *m, const amr_tagset_t *s, int epochs, int batch, float lr,
                  const float *cw, unsigned long long seed, float *loss);
/* Decision rule of the predictions below: class = argmax(z0, z1, z2 + bias)
 * (bias > 0 trades precision for recall of "refine"; default 0). */
int amr_mlp_set_refine_bias(amr_mlp_t *m, float bias);
/* Predicted class of every row of the set. */
int amr_mlp_predict_set(amr_mlp_t *m, const amr_tagset_t *s, int *cls);
/* Predicted class of every quadrant of the context's current state. */
int amr_mlp_predict_ctx(amr_mlp_t *m, struct amrgpu_ctx *c, int *cls);

/* ---- data collection and deployment (host driver, amr_tagger.c) ---- */

typedef struct {
    double tol;              /* plain indicator threshold (refine above)        */
    double coarsen_frac;     /* coarsen label below coarsen_frac * tol (0.5)     */
    int window;              /* steps between regrids = label horizon (40)       */
    int label_stride;        /* label times every label_stride steps (5)         */
    int allow_coarsening;
    const amrgpu_inst_t *inst;   /* NULL: instance from the physics parameters */
} amr_collect_opts_t;

/* Advances the forest (initial condition already applied) to the physical
 * stop of the current physics settings on the GPU, regridding with the plain
 * indicator every `window` steps.  At each regrid the features of every
 * quadrant are taken; the indicator is then evaluated every label_stride
 * steps up to the window on the same mesh and turned into the row's label.
 * Rows are appended to `set`.  Returns 0 or an error code. */
int amr_tagger_collect(struct amr_forest *f, const amr_collect_opts_t *o, amr_tagset_t *set,
                       int *nrows);

/* Tagger for amr_evolve_device (amr_tagger_fn): `user` points to this. */
typedef struct {
    amr_mlp_t *mlp;
    double tol;              /* plain indicator threshold for the fallback      */
    int union_plain;         /* 1: also refine what the plain indicator tags now */
    long long predicted[3];  /* running counts of predicted classes             */
    long long plain_refine;  /* quadrants the plain indicator tagged            */
    long long added_by_plain;/* of those, refined only because of the fallback  */
} amr_mlp_tagger_t;

int amr_mlp_tagger(void *user, struct amrgpu_ctx *ctx, struct amr_forest *f, int *tags);

#ifdef __cplusplus
}
#endif
#endif
