/*
 * tagger_oracle.c -- CPU restatement of the learned tagger's features and
 * network (test infrastructure only; see amr_oracle.h).
 *
 * The reference has no learned tagger; what is restated here is the
 * definition in include/amr_tagger.h: features from the reference's own
 * boundary arrays (face_info, src/amr/src/amr_external.c:211-508 format) and
 * the 384-64-64-3 network with its BF16 roundings.  Accumulation is in
 * double (the GPU accumulates in FP32), so GPU and oracle agree to FP32
 * rounding, not bitwise, for the network; the features agree bitwise.
 */
#define _POSIX_C_SOURCE 200809L
#include "amr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NIN 384
#define NH 64
#define NOUT 3
#define OW1 0
#define OB1 (NH * NIN)
#define OW2 (OB1 + NH)
#define OB2 (OW2 + NH * NH)
#define OW3 (OB2 + NH)
#define OB3 (OW3 + NOUT * NH)
#define NP (OB3 + NOUT)

unsigned short orc_bf16(float x)
{
    unsigned int u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

float orc_bf16_to_float(unsigned short h)
{
    unsigned int u = (unsigned int)h << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

static double round_bf(double v) { return (double)orc_bf16_to_float(orc_bf16((float)v)); }

/* ------------------------------------------------------------------------- */
/* features                                                                  */
/* ------------------------------------------------------------------------- */
/* density of the neighbour's facing edge cell ec across side s of quadrant q:
 * one neighbour, the mean of two finer ones, or the own value r at a wall */
static double edge_rho(const int *nb, const double *soln, int q, int s, int ec, double r)
{
    const int a = nb[8 * q + 2 * s], b = nb[8 * q + 2 * s + 1];
    if (a < 0) return r;
    const double va = soln[(size_t)256 * a + 4 * ec];
    if (b < 0) return va;
    return 0.5 * (va + soln[(size_t)256 * b + 4 * ec]);
}

int orc_tagger_features(int n, int ncell, int nface, const int *face_info, const double *soln,
                        unsigned short *X)
{
    if (n == 16) {
        /* 16x16 patches: the features of the patch coarsened to 8x8, every
         * coarse cell the mean of its 2x2 block, ((a + b) + c) + d times 0.25
         * with a..d in z-order (the summation of the reference's coarsen
         * operator, solver_coarsen_operator_2d.F90:49-50, within one patch);
         * the neighbours coarsened alike, so the 8x8 definition applies */
        double *P = (double *)malloc(sizeof(double) * 256 * (size_t)(ncell > 0 ? ncell : 1));
        if (!P) return -1;
        for (int q = 0; q < ncell; ++q) {
            const double *Q = soln + (size_t)1024 * q;
            for (int J = 0; J < 8; ++J)
                for (int I = 0; I < 8; ++I)
                    for (int f = 0; f < 4; ++f) {
                        const double a = Q[f + 4 * (2 * I + 16 * (2 * J))];
                        const double b = Q[f + 4 * (2 * I + 1 + 16 * (2 * J))];
                        const double c = Q[f + 4 * (2 * I + 16 * (2 * J + 1))];
                        const double d = Q[f + 4 * (2 * I + 1 + 16 * (2 * J + 1))];
                        P[(size_t)256 * q + f + 4 * (I + 8 * J)] = 0.25 * (((a + b) + c) + d);
                    }
        }
        const int rc = orc_tagger_features(8, ncell, nface, face_info, P, X);
        free(P);
        return rc;
    }
    if (n != 8) return -1;
    /* neighbours per side from the face list: full faces pair the two sides,
     * a hanging face gives the coarse side both fine quadrants (in face order)
     * and each fine quadrant the coarse one */
    int *nb = (int *)malloc(sizeof(int) * 8 * (size_t)(ncell > 0 ? ncell : 1));
    if (!nb) return -1;
    for (int k = 0; k < 8 * ncell; ++k) nb[k] = -1;
    for (int fc = 0; fc < nface; ++fc) {
        const int *fi = face_info + 11 * fc;
        const int type = fi[0], side = fi[1], c = fi[2];
        if (type == 2) {
            nb[8 * c + 2 * side] = fi[3];
            nb[8 * fi[3] + 2 * fi[7]] = c;
        } else if (type == 3) {
            nb[8 * c + 2 * side] = fi[3];
            nb[8 * c + 2 * side + 1] = fi[4];
            nb[8 * fi[3] + 2 * fi[7]] = c;
            nb[8 * fi[4] + 2 * fi[8]] = c;
        }
    }
    const int n2 = 64;
    for (int q = 0; q < ncell; ++q) {
        const double *Q = soln + (size_t)4 * n2 * q;
        double smax[4] = {0.0, 0.0, 0.0, 0.0};
        for (int c = 0; c < n2; ++c)
            for (int f = 0; f < 4; ++f) {
                double a = fabs(Q[f + 4 * c]);
                if (a > smax[f]) smax[f] = a;
            }
        for (int j = 0; j < 8; ++j)
            for (int i = 0; i < 8; ++i) {
                const int c = i + 8 * j;
                double feat[6];
                for (int f = 0; f < 4; ++f) feat[f] = smax[f] > 0.0 ? Q[f + 4 * c] / smax[f] : 0.0;
                const double r = Q[4 * c];
                const double re = i < 7 ? Q[4 * (c + 1)] : edge_rho(nb, soln, q, 1, 8 * j, r);
                const double rw = i > 0 ? Q[4 * (c - 1)] : edge_rho(nb, soln, q, 0, 8 * j + 7, r);
                const double rn = j < 7 ? Q[4 * (c + 8)] : edge_rho(nb, soln, q, 3, i, r);
                const double rs = j > 0 ? Q[4 * (c - 8)] : edge_rho(nb, soln, q, 2, 56 + i, r);
                feat[4] = smax[0] > 0.0 ? (re - rw) / smax[0] : 0.0;
                feat[5] = smax[0] > 0.0 ? (rn - rs) / smax[0] : 0.0;
                for (int f = 0; f < 6; ++f) X[(size_t)NIN * q + 64 * f + c] = orc_bf16((float)feat[f]);
            }
    }
    free(nb);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* network                                                                   */
/* ------------------------------------------------------------------------- */
typedef struct {
    double *W1, *W2, *W3;     /* weights as used by the products */
    double *h1, *h2;          /* activations [n][64] as used downstream */
    double *a1, *a2;          /* pre-activations (exact mode masks) */
    double *z;                /* logits [n][3] */
} fwd_t;

static void fwd_free(fwd_t *f)
{
    free(f->W1); free(f->W2); free(f->W3); free(f->h1); free(f->h2); free(f->a1); free(f->a2); free(f->z);
}

static int forward(const float *P, const unsigned short *X, int n, int exact, fwd_t *f)
{
    memset(f, 0, sizeof(*f));
    f->W1 = (double *)malloc(sizeof(double) * NH * NIN);
    f->W2 = (double *)malloc(sizeof(double) * NH * NH);
    f->W3 = (double *)malloc(sizeof(double) * NOUT * NH);
    f->h1 = (double *)malloc(sizeof(double) * (size_t)n * NH);
    f->h2 = (double *)malloc(sizeof(double) * (size_t)n * NH);
    f->a1 = (double *)malloc(sizeof(double) * (size_t)n * NH);
    f->a2 = (double *)malloc(sizeof(double) * (size_t)n * NH);
    f->z = (double *)malloc(sizeof(double) * (size_t)n * NOUT);
    if (!f->W1 || !f->W2 || !f->W3 || !f->h1 || !f->h2 || !f->a1 || !f->a2 || !f->z) return -1;
    for (int i = 0; i < NH * NIN; ++i) f->W1[i] = exact ? (double)P[OW1 + i] : round_bf(P[OW1 + i]);
    for (int i = 0; i < NH * NH; ++i) f->W2[i] = exact ? (double)P[OW2 + i] : round_bf(P[OW2 + i]);
    for (int i = 0; i < NOUT * NH; ++i) f->W3[i] = exact ? (double)P[OW3 + i] : round_bf(P[OW3 + i]);
    for (int s = 0; s < n; ++s) {
        const unsigned short *x = X + (size_t)NIN * s;
        for (int h = 0; h < NH; ++h) {
            double acc = 0.0;
            for (int k = 0; k < NIN; ++k) acc += (double)orc_bf16_to_float(x[k]) * f->W1[h * NIN + k];
            double a = exact ? acc + (double)P[OB1 + h] : (double)((float)acc + P[OB1 + h]);
            f->a1[s * NH + h] = a;
            f->h1[s * NH + h] = exact ? (a > 0.0 ? a : 0.0) : round_bf(a > 0.0 ? a : 0.0);
        }
        for (int o = 0; o < NH; ++o) {
            double acc = 0.0;
            for (int k = 0; k < NH; ++k) acc += f->h1[s * NH + k] * f->W2[o * NH + k];
            double a = exact ? acc + (double)P[OB2 + o] : (double)((float)acc + P[OB2 + o]);
            f->a2[s * NH + o] = a;
            f->h2[s * NH + o] = exact ? (a > 0.0 ? a : 0.0) : round_bf(a > 0.0 ? a : 0.0);
        }
        for (int c = 0; c < NOUT; ++c) {
            double acc = 0.0;
            for (int k = 0; k < NH; ++k) acc += f->h2[s * NH + k] * f->W3[c * NH + k];
            f->z[s * NOUT + c] = exact ? acc + (double)P[OB3 + c] : (double)((float)acc + P[OB3 + c]);
        }
    }
    return 0;
}

void orc_mlp_forward(const float *P, const unsigned short *X, int n, int exact, double *logits)
{
    fwd_t f;
    if (forward(P, X, n, exact, &f) == 0) memcpy(logits, f.z, sizeof(double) * (size_t)n * NOUT);
    fwd_free(&f);
}

double orc_mlp_grad(const float *P, const unsigned short *X, const int *y, int n, const float cw[3],
                    int exact, double *grad)
{
    fwd_t f;
    double loss = 0.0;
    double *dZ = (double *)malloc(sizeof(double) * (size_t)n * NOUT);
    double *dA2 = (double *)malloc(sizeof(double) * (size_t)n * NH);
    double *dA1 = (double *)malloc(sizeof(double) * (size_t)n * NH);
    if (!dZ || !dA2 || !dA1 || forward(P, X, n, exact, &f)) {
        free(dZ); free(dA2); free(dA1);
        fwd_free(&f);
        return NAN;
    }
    double sw = 0.0;
    for (int s = 0; s < n; ++s) sw += cw[y[s]];
    memset(grad, 0, sizeof(double) * NP);
    for (int s = 0; s < n; ++s) {
        const double *z = f.z + NOUT * s;
        double m = z[0] > z[1] ? z[0] : z[1];
        if (z[2] > m) m = z[2];
        double e[3], S = 0.0;
        for (int c = 0; c < 3; ++c) { e[c] = exp(z[c] - m); S += e[c]; }
        const double g = cw[y[s]] / sw;
        loss += g * (log(S) - (z[y[s]] - m));
        for (int c = 0; c < 3; ++c) {
            double d = g * (e[c] / S - (c == y[s] ? 1.0 : 0.0));
            dZ[s * NOUT + c] = exact ? d : round_bf(d);
        }
    }
    /* layer 3 */
    for (int s = 0; s < n; ++s)
        for (int c = 0; c < NOUT; ++c) {
            const double d = dZ[s * NOUT + c];
            grad[OB3 + c] += d;
            for (int j = 0; j < NH; ++j) grad[OW3 + c * NH + j] += d * f.h2[s * NH + j];
        }
    for (int s = 0; s < n; ++s)
        for (int j = 0; j < NH; ++j) {
            double acc = 0.0;
            for (int c = 0; c < NOUT; ++c) acc += dZ[s * NOUT + c] * f.W3[c * NH + j];
            const int on = exact ? f.a2[s * NH + j] > 0.0 : f.h2[s * NH + j] > 0.0;
            dA2[s * NH + j] = on ? (exact ? acc : round_bf(acc)) : 0.0;
        }
    /* layer 2 */
    for (int s = 0; s < n; ++s)
        for (int o = 0; o < NH; ++o) {
            const double d = dA2[s * NH + o];
            grad[OB2 + o] += d;
            for (int i = 0; i < NH; ++i) grad[OW2 + o * NH + i] += d * f.h1[s * NH + i];
        }
    for (int s = 0; s < n; ++s)
        for (int i = 0; i < NH; ++i) {
            double acc = 0.0;
            for (int o = 0; o < NH; ++o) acc += dA2[s * NH + o] * f.W2[o * NH + i];
            const int on = exact ? f.a1[s * NH + i] > 0.0 : f.h1[s * NH + i] > 0.0;
            dA1[s * NH + i] = on ? (exact ? acc : round_bf(acc)) : 0.0;
        }
    /* layer 1 */
    for (int s = 0; s < n; ++s) {
        const unsigned short *x = X + (size_t)NIN * s;
        for (int h = 0; h < NH; ++h) {
            const double d = dA1[s * NH + h];
            grad[OB1 + h] += d;
            if (d == 0.0) continue;
            for (int k = 0; k < NIN; ++k) grad[OW1 + h * NIN + k] += d * (double)orc_bf16_to_float(x[k]);
        }
    }
    free(dZ); free(dA2); free(dA1);
    fwd_free(&f);
    return loss;
}
