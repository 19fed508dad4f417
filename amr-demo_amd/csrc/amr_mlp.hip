// amr_mlp.hip -- learned refinement tagger on gfx950: patch features, and
// the forward / backward passes of a 384 -> 64 -> 64 -> 3 ReLU perceptron on
// the MFMA units (gfx950's v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
// accumulate, K = 32 per instruction; the one K = 16 product, by the padded
// logit gradients, uses v_mfma_f32_16x16x16_bf16),
// with a deterministic two-pass gradient reduction and Adam.
//
// The reference has no learned tagger (it refines on density thresholds,
// src/solver/src/tagging_methods_module.F90:14-71); features, labels and the
// network are defined in include/amr_tagger.h.
//
// Work decomposition.  A workgroup of four waves takes 64 samples at a time
// ("chunk"); wave w owns samples 16w..16w+15 through the forward pass and the
// backward pass down to the pre-activation gradients, while the weight
// gradients (a reduction over samples) are split by output tile across the
// waves and stay in registers across the workgroup's chunks.  W1 (48 KB BF16)
// is staged in LDS once per workgroup; one chunk of features (48 KB) is staged
// per chunk.  Each workgroup writes one partial gradient; a second kernel sums
// the partials in a fixed order (bitwise reproducible) and applies Adam.
//
// MFMA operand layouts (lane l, r = l % 16, g = l / 16):
//   16x16x32: A[i][k]: lane holds A[r][8g + t], B[k][j]: lane holds
//             B[8g + t][r], t = 0..7 (eight consecutive k per lane)
//   16x16x16: the same with 4g + t, t = 0..3
//   D[i][j] (both): lane holds D[4g + t][r], t = 0..3

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "amr_hip.h"
#include "amr_internal.h"
#include "amr_tagger.h"

namespace {

constexpr int NIN = AMR_MLP_NIN, NH = AMR_MLP_NH, NOUT = AMR_MLP_NOUT;
constexpr int OW1 = 0, OB1 = NH * NIN, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NOUT * NH, NP = OB3 + NOUT;
static_assert(NP == AMR_MLP_NPARAM, "parameter layout");
constexpr int NPP = (NP + 1 + 63) / 64 * 64;     // partial row: gradient + loss, padded
constexpr int SX = NIN + 8;                      // LDS row strides (elements)
constexpr int SH = NH + 8;
constexpr int SZ = 16 + 8;
constexpr int ROWS = 64;                         // samples per chunk
constexpr int BLK = 256;

typedef short s4v __attribute__((ext_vector_type(4)));
typedef short s8v __attribute__((ext_vector_type(8)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);     // round to nearest even (no NaN here)
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f4v mfma(s4v a, s4v b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma32(s8v a, s8v b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8v, a), __builtin_bit_cast(b8v, b), c, 0,
                                                   0, 0);
}

__device__ __forceinline__ unsigned long long d2bits(double v) { return (unsigned long long)__double_as_longlong(v); }
__device__ __forceinline__ double bits2d(unsigned long long b) { return __longlong_as_double((long long)b); }

// ---------------------------------------------------------------------------
// features: one thread per cell of an 8x8 patch, four patches per workgroup
// ---------------------------------------------------------------------------
__device__ __forceinline__ double edge_value(const double *U, const int *nbr, int quad, int side, int ec,
                                             double own)
{
    const int na = nbr[quad * 8 + 2 * side], nb = nbr[quad * 8 + 2 * side + 1];
    if (na < 0) return own;
    const double a = U[(size_t)na * 256 + ec];
    if (nb < 0) return a;
    return 0.5 * (a + U[(size_t)nb * 256 + ec]);
}

__global__ __launch_bounds__(256) void features_kernel(const double *__restrict__ U,
                                                       const int *__restrict__ nbr, int nquad,
                                                       uint16_t *__restrict__ X)
{
    __shared__ unsigned long long smax[4][4];
    const int tid = threadIdx.x, p = tid >> 6, c = tid & 63, ci = c & 7, cj = c >> 3;
    const int quad = blockIdx.x * 4 + p;
    const bool active = quad < nquad;
    if (tid < 16) smax[tid >> 2][tid & 3] = 0ull;
    __syncthreads();
    double u[4] = {0.0, 0.0, 0.0, 0.0};
    const double *Q = U + (size_t)quad * 256;
    if (active) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            u[f] = Q[f * 64 + c];
            atomicMax(&smax[p][f], d2bits(fabs(u[f])));   // non-negative doubles order as integers
        }
    }
    __syncthreads();
    if (!active) return;
    double feat[6];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const double s = bits2d(smax[p][f]);
        feat[f] = s > 0.0 ? u[f] / s : 0.0;
    }
    const double r = u[0], s0 = bits2d(smax[p][0]);
    const double re = ci < 7 ? Q[c + 1] : edge_value(U, nbr, quad, 1, cj * 8, r);
    const double rw = ci > 0 ? Q[c - 1] : edge_value(U, nbr, quad, 0, cj * 8 + 7, r);
    const double rn = cj < 7 ? Q[c + 8] : edge_value(U, nbr, quad, 3, ci, r);
    const double rs = cj > 0 ? Q[c - 8] : edge_value(U, nbr, quad, 2, 56 + ci, r);
    feat[4] = s0 > 0.0 ? (re - rw) / s0 : 0.0;
    feat[5] = s0 > 0.0 ? (rn - rs) / s0 : 0.0;
    uint16_t *o = X + (size_t)quad * NIN + c;
#pragma unroll
    for (int f = 0; f < 6; ++f) o[f * 64] = f2bf((float)feat[f]);
}

// 16x16 patches: the tagger reads each patch coarsened to 8x8, every coarse
// cell the mean of its 2x2 block, ((a + b) + c) + d times 0.25 with a..d in
// z-order (the summation of solver_coarsen_operator_2d.F90:49-50 applied
// within a patch).  The neighbours are coarsened alike and keep their ids, so
// features_kernel then runs unchanged on the coarsened state.
__global__ __launch_bounds__(256) void pool16_kernel(const double *__restrict__ U, double *__restrict__ P,
                                                     int nquad)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;     // one coarse cell
    if (t >= (long long)nquad * 64) return;
    const long long q = t >> 6;
    const int c = (int)(t & 63), I = c & 7, J = c >> 3;
    const double *s = U + q * 1024 + (2 * J) * 16 + 2 * I;
    double *d = P + q * 256 + c;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const double *sf = s + f * 256;
        d[f * 64] = 0.25 * (((sf[0] + sf[1]) + sf[16]) + sf[17]);
    }
}

// ---------------------------------------------------------------------------
// shared pieces of the network kernels
// ---------------------------------------------------------------------------
struct __attribute__((aligned(16))) NetLds {     // rows read as 16-byte vectors
    uint16_t W1[NH * SX];
    uint16_t X[ROWS * SX];
    uint16_t W2[NH * SH];
    uint16_t W3[16 * SH];       // rows 3..15 zero
    uint16_t H1[ROWS * SH];
    uint16_t H2[ROWS * SH];
    float b[2 * NH + 4];
    float Z[ROWS * 4];
};

struct __attribute__((aligned(16))) TrainLds {
    NetLds net;
    uint16_t dA2[ROWS * SH];
    uint16_t dA1[ROWS * SH];
    uint16_t dZ[ROWS * SZ];     // classes 3..15 zero
    float w[ROWS];
    int y[ROWS];
    float bg[2 * NH + 4];       // bias gradients, loss in bg[2*NH+3]
    float loss[ROWS];
};

struct Weights {
    const uint16_t *W1, *W2, *W3;   // BF16 copies: [64][384], [64][64], [16][64]
    const float *P;                 // FP32 master (biases read from here)
};

__device__ void stage_weights(NetLds &s, const Weights &wt)
{
    const int tid = threadIdx.x;
    for (int i = tid; i < NH * NIN / 8; i += BLK) {
        const int row = i / (NIN / 8), c8 = i % (NIN / 8);
        *(uint4 *)&s.W1[row * SX + c8 * 8] = ((const uint4 *)wt.W1)[i];
    }
    for (int i = tid; i < NH * NH / 8; i += BLK) {
        const int row = i / (NH / 8), c8 = i % (NH / 8);
        *(uint4 *)&s.W2[row * SH + c8 * 8] = ((const uint4 *)wt.W2)[i];
    }
    for (int i = tid; i < 16 * NH / 8; i += BLK) {
        const int row = i / (NH / 8), c8 = i % (NH / 8);
        *(uint4 *)&s.W3[row * SH + c8 * 8] = ((const uint4 *)wt.W3)[i];
    }
    if (tid < NH) {
        s.b[tid] = wt.P[OB1 + tid];
        s.b[NH + tid] = wt.P[OB2 + tid];
    }
    if (tid < 4) s.b[2 * NH + tid] = tid < NOUT ? wt.P[OB3 + tid] : 0.0f;
}

// Feature rows of one chunk into LDS; rows past `count` are zero.
__device__ void stage_rows(NetLds &s, const uint16_t *__restrict__ X, const int *__restrict__ perm,
                           long long first, int local0, int count)
{
    for (int i = threadIdx.x; i < ROWS * NIN / 8; i += BLK) {
        const int row = i / (NIN / 8), c8 = i % (NIN / 8);
        const int loc = local0 + row;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (loc < count) {
            const long long g = first + loc;
            const long long src = perm ? (long long)perm[g] : g;
            v = ((const uint4 *)(X + (size_t)src * NIN))[c8];
        }
        *(uint4 *)&s.X[row * SX + c8 * 8] = v;
    }
}

// Forward pass of wave w's 16 samples: H1, H2 (BF16, ReLU) and logits Z.
__device__ void forward_rows(NetLds &s)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int row = 16 * w + r;
    {
        f4v acc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
        for (int kk = 0; kk < NIN / 32; ++kk) {
            const s8v a = *(const s8v *)&s.X[row * SX + 32 * kk + 8 * g];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const s8v bv = *(const s8v *)&s.W1[(16 * c + r) * SX + 32 * kk + 8 * g];
                acc[c] = mfma32(a, bv, acc[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int smp = 16 * w + 4 * g + t, h = 16 * c + r;
                s.H1[smp * SH + h] = f2bf(fmaxf(acc[c][t] + s.b[h], 0.0f));
            }
    }
    __syncthreads();
    {
        f4v acc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < NH / 32; ++kk) {
            const s8v a = *(const s8v *)&s.H1[row * SH + 32 * kk + 8 * g];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const s8v bv = *(const s8v *)&s.W2[(16 * c + r) * SH + 32 * kk + 8 * g];
                acc[c] = mfma32(a, bv, acc[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int smp = 16 * w + 4 * g + t, h = 16 * c + r;
                s.H2[smp * SH + h] = f2bf(fmaxf(acc[c][t] + s.b[NH + h], 0.0f));
            }
    }
    __syncthreads();
    {
        f4v acc = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < NH / 32; ++kk) {
            const s8v a = *(const s8v *)&s.H2[row * SH + 32 * kk + 8 * g];
            const s8v bv = *(const s8v *)&s.W3[r * SH + 32 * kk + 8 * g];
            acc = mfma32(a, bv, acc);
        }
        if (r < NOUT)
#pragma unroll
            for (int t = 0; t < 4; ++t) s.Z[(16 * w + 4 * g + t) * 4 + r] = acc[t] + s.b[2 * NH + r];
    }
    __syncthreads();
}

// Inference: logits (optional) and argmax class (optional) of `count` rows.
__global__ __launch_bounds__(BLK) void infer_kernel(Weights wt, const uint16_t *__restrict__ X, int count,
                                                    float refine_bias, float *__restrict__ logits,
                                                    int *__restrict__ cls)
{
    __shared__ NetLds s;
    stage_weights(s, wt);
    for (int chunk = blockIdx.x; chunk * ROWS < count; chunk += gridDim.x) {
        stage_rows(s, X, nullptr, 0, chunk * ROWS, count);
        __syncthreads();
        forward_rows(s);
        const int t = threadIdx.x;
        const int row = chunk * ROWS + t;
        if (t < ROWS && row < count) {
            const float z0 = s.Z[t * 4], z1 = s.Z[t * 4 + 1], z2 = s.Z[t * 4 + 2] + refine_bias;
            if (logits) {
                logits[(size_t)row * 3] = z0;
                logits[(size_t)row * 3 + 1] = z1;
                logits[(size_t)row * 3 + 2] = s.Z[t * 4 + 2];
            }
            if (cls) cls[row] = (z0 >= z1 && z0 >= z2) ? 0 : (z1 >= z2 ? 1 : 2);
        }
        __syncthreads();
    }
}

struct GradArgs {
    const uint16_t *X;
    const int *y;
    const int *perm;       // nullptr: rows first.. in order
    long long first;
    int count;             // rows of this launch
    int chunks;            // chunks per workgroup
    float cw[3];           // class weights
    float scale;           // 1 / sum of row weights
    float *partial;        // [gridDim.x][NPP]
};

// Loss and parameter gradient partials of `chunks` chunks per workgroup.
__global__ __launch_bounds__(BLK) void grad_kernel(Weights wt, GradArgs ga)
{
    __shared__ TrainLds s;
    NetLds &n = s.net;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, r = lane & 15, g = lane >> 4;
    stage_weights(n, wt);
    for (int i = tid; i < ROWS * SZ; i += BLK) s.dZ[i] = 0;
    if (tid < 2 * NH + 4) s.bg[tid] = 0.0f;

    f4v g1[4][6], g2[4], g3;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 6; ++b) g1[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
        g2[a] = f4v{0.f, 0.f, 0.f, 0.f};
    }
    g3 = f4v{0.f, 0.f, 0.f, 0.f};

    for (int k = 0; k < ga.chunks; ++k) {
        const int local0 = (blockIdx.x * ga.chunks + k) * ROWS;
        stage_rows(n, ga.X, ga.perm, ga.first, local0, ga.count);
        if (tid < ROWS) {
            const int loc = local0 + tid;
            int yy = 0;
            float wv = 0.0f;
            if (loc < ga.count) {
                const long long gi = ga.first + loc;
                yy = ga.y[ga.perm ? (long long)ga.perm[gi] : gi];
                wv = ga.cw[yy];
            }
            s.y[tid] = yy;
            s.w[tid] = wv;
        }
        __syncthreads();
        forward_rows(n);

        // softmax cross-entropy, one thread per sample
        if (tid < ROWS) {
            const float z[3] = {n.Z[tid * 4], n.Z[tid * 4 + 1], n.Z[tid * 4 + 2]};
            const float m = fmaxf(z[0], fmaxf(z[1], z[2]));
            const float e0 = expf(z[0] - m), e1 = expf(z[1] - m), e2 = expf(z[2] - m);
            const float S = e0 + e1 + e2;
            const int yy = s.y[tid];
            const float wv = s.w[tid], gs = wv * ga.scale;
            const float p[3] = {e0 / S, e1 / S, e2 / S};
#pragma unroll
            for (int c = 0; c < 3; ++c) s.dZ[tid * SZ + c] = f2bf(gs * (p[c] - (c == yy ? 1.0f : 0.0f)));
            s.loss[tid] = wv > 0.0f ? gs * (logf(S) - (z[yy] - m)) : 0.0f;
        }
        __syncthreads();

        // dW3 += dZ^T H2 (wave w: hidden tile w)
#pragma unroll
        for (int kk = 0; kk < ROWS / 32; ++kk) {
            s8v a, bv;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int smp = 32 * kk + 8 * g + t;
                a[t] = (short)s.dZ[smp * SZ + r];
                bv[t] = (short)n.H2[smp * SH + 16 * w + r];
            }
            g3 = mfma32(a, bv, g3);
        }
        // dA2 = (dZ W3) * [H2 > 0] for wave w's samples
        {
            const s4v a = *(const s4v *)&s.dZ[(16 * w + r) * SZ + 4 * g];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                s4v bv;
#pragma unroll
                for (int t = 0; t < 4; ++t) bv[t] = (short)n.W3[(4 * g + t) * SH + 16 * c + r];
                const f4v d = mfma(a, bv, f4v{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int idx = (16 * w + 4 * g + t) * SH + 16 * c + r;
                    s.dA2[idx] = f2bf(bf2f(n.H2[idx]) > 0.0f ? d[t] : 0.0f);
                }
            }
        }
        __syncthreads();
        // dA1 = (dA2 W2) * [H1 > 0] for wave w's samples
        {
            f4v acc[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < NH / 32; ++kk) {
                const s8v a = *(const s8v *)&s.dA2[(16 * w + r) * SH + 32 * kk + 8 * g];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    s8v bv;
#pragma unroll
                    for (int t = 0; t < 8; ++t) bv[t] = (short)n.W2[(32 * kk + 8 * g + t) * SH + 16 * c + r];
                    acc[c] = mfma32(a, bv, acc[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int idx = (16 * w + 4 * g + t) * SH + 16 * c + r;
                    s.dA1[idx] = f2bf(bf2f(n.H1[idx]) > 0.0f ? acc[c][t] : 0.0f);
                }
        }
        // dW2 += dA2^T H1 (wave w: output-row tile w)
#pragma unroll
        for (int kk = 0; kk < ROWS / 32; ++kk) {
            s8v a;
#pragma unroll
            for (int t = 0; t < 8; ++t) a[t] = (short)s.dA2[(32 * kk + 8 * g + t) * SH + 16 * w + r];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                s8v bv;
#pragma unroll
                for (int t = 0; t < 8; ++t) bv[t] = (short)n.H1[(32 * kk + 8 * g + t) * SH + 16 * c + r];
                g2[c] = mfma32(a, bv, g2[c]);
            }
        }
        __syncthreads();
        // dW1 += dA1^T X (wave w: feature tiles 6w..6w+5)
#pragma unroll
        for (int kk = 0; kk < ROWS / 32; ++kk) {
            s8v a[4];
#pragma unroll
            for (int hc = 0; hc < 4; ++hc)
#pragma unroll
                for (int t = 0; t < 8; ++t) a[hc][t] = (short)s.dA1[(32 * kk + 8 * g + t) * SH + 16 * hc + r];
#pragma unroll
            for (int fi = 0; fi < 6; ++fi) {
                s8v bv;
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    bv[t] = (short)n.X[(32 * kk + 8 * g + t) * SX + 16 * (6 * w + fi) + r];
#pragma unroll
                for (int hc = 0; hc < 4; ++hc) g1[hc][fi] = mfma32(a[hc], bv, g1[hc][fi]);
            }
        }
        // bias gradients and loss: fixed summation order over the chunk's samples
        if (tid < NH) {
            float acc = 0.0f;
            for (int q = 0; q < ROWS; ++q) acc += bf2f(s.dA1[q * SH + tid]);
            s.bg[tid] += acc;
        } else if (tid < 2 * NH) {
            float acc = 0.0f;
            for (int q = 0; q < ROWS; ++q) acc += bf2f(s.dA2[q * SH + tid - NH]);
            s.bg[tid] += acc;
        } else if (tid < 2 * NH + NOUT) {
            float acc = 0.0f;
            for (int q = 0; q < ROWS; ++q) acc += bf2f(s.dZ[q * SZ + tid - 2 * NH]);
            s.bg[tid] += acc;
        } else if (tid == 2 * NH + 3) {
            float acc = 0.0f;
            for (int q = 0; q < ROWS; ++q) acc += s.loss[q];
            s.bg[tid] += acc;
        }
        __syncthreads();
    }

    float *part = ga.partial + (size_t)blockIdx.x * NPP;
#pragma unroll
    for (int hc = 0; hc < 4; ++hc)
#pragma unroll
        for (int fi = 0; fi < 6; ++fi)
#pragma unroll
            for (int t = 0; t < 4; ++t)
                part[OW1 + (16 * hc + 4 * g + t) * NIN + 16 * (6 * w + fi) + r] = g1[hc][fi][t];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int t = 0; t < 4; ++t) part[OW2 + (16 * w + 4 * g + t) * NH + 16 * c + r] = g2[c][t];
    if (g == 0)
#pragma unroll
        for (int t = 0; t < NOUT; ++t) part[OW3 + t * NH + 16 * w + r] = g3[t];
    if (tid < NH) {
        part[OB1 + tid] = s.bg[tid];
        part[OB2 + tid] = s.bg[NH + tid];
    }
    if (tid < NOUT) part[OB3 + tid] = s.bg[2 * NH + tid];
    if (tid == 0) part[NP] = s.bg[2 * NH + 3];
}

struct AdamArgs {
    int on;
    float lr, beta1, beta2, eps, bc1, bc2;   // bc = 1 / (1 - beta^t)
};

// Sum of the partials in workgroup order; optional Adam update and BF16 copy.
__global__ __launch_bounds__(256) void reduce_kernel(const float *__restrict__ partial, int nparts,
                                                     float *__restrict__ grad, float *__restrict__ loss,
                                                     float *P, float *M, float *V, uint16_t *W1b,
                                                     uint16_t *W2b, uint16_t *W3b, AdamArgs ad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NP) return;
    float gsum = 0.0f;
    for (int b = 0; b < nparts; ++b) gsum += partial[(size_t)b * NPP + p];
    if (p == NP) {
        if (loss) *loss = gsum;
        return;
    }
    if (grad) grad[p] = gsum;
    if (!ad.on) return;
    const float m = ad.beta1 * M[p] + (1.0f - ad.beta1) * gsum;
    const float v = ad.beta2 * V[p] + (1.0f - ad.beta2) * gsum * gsum;
    M[p] = m;
    V[p] = v;
    const float x = P[p] - ad.lr * (m * ad.bc1) / (sqrtf(v * ad.bc2) + ad.eps);
    P[p] = x;
    if (p < OB1) W1b[p] = f2bf(x);
    else if (p >= OW2 && p < OB2) W2b[p - OW2] = f2bf(x);
    else if (p >= OW3 && p < OB3) W3b[p - OW3] = f2bf(x);
}

__global__ void bf16_copy_kernel(const float *P, uint16_t *W1b, uint16_t *W2b, uint16_t *W3b)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < OB1) W1b[p] = f2bf(P[p]);
    else if (p >= OW2 && p < OB2) W2b[p - OW2] = f2bf(P[p]);
    else if (p >= OW3 && p < OB3) W3b[p - OW3] = f2bf(P[p]);
}

#define MCHK(x)                                                                               \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) {                                                               \
            amr_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));     \
            return AMRGPU_ERR_HIP;                                                            \
        }                                                                                     \
    } while (0)

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

template <class T>
int grow(T **p, size_t *cap, size_t need, bool keep, size_t used)
{
    if (need <= *cap) return AMRGPU_OK;
    size_t nc = std::max(need, *cap + *cap / 2);
    T *q = nullptr;
    MCHK(hipMalloc((void **)&q, sizeof(T) * nc));
    if (keep && *p && used) MCHK(hipMemcpy(q, *p, sizeof(T) * used, hipMemcpyDeviceToDevice));
    if (*p) (void)hipFree(*p);
    *p = q;
    *cap = nc;
    return AMRGPU_OK;
}

}  // namespace

struct amr_mlp {
    hipStream_t stream = nullptr;
    float *P = nullptr, *M = nullptr, *V = nullptr, *grad = nullptr, *loss = nullptr;
    uint16_t *W1b = nullptr, *W2b = nullptr, *W3b = nullptr;
    float *partial = nullptr;
    size_t cap_partial = 0;
    uint16_t *X = nullptr;      // scratch rows (host uploads, context features)
    size_t cap_X = 0;
    int *y = nullptr, *perm = nullptr, *cls = nullptr;
    size_t cap_y = 0, cap_perm = 0, cap_cls = 0;
    float *logits = nullptr;
    size_t cap_logits = 0;
    long long step = 0;
    float refine_bias = 0.0f;
};

struct amr_tagset {
    uint16_t *X = nullptr;
    int *y = nullptr;
    size_t cap = 0, capy = 0;
    int n = 0;
    std::vector<int> h_y;
    std::vector<double> h_ind;
};

namespace {

Weights weights_of(const amr_mlp *m) { return Weights{m->W1b, m->W2b, m->W3b, m->P}; }

int launch_infer(amr_mlp *m, const uint16_t *dX, int count, float *dlogits, int *dcls, hipStream_t st)
{
    if (count <= 0) return AMRGPU_OK;
    int chunks = (count + ROWS - 1) / ROWS;
    int grid = std::min(chunks, 1024);
    hipLaunchKernelGGL(infer_kernel, dim3(grid), dim3(BLK), 0, st, weights_of(m), dX, count, m->refine_bias,
                       dlogits, dcls);
    MCHK(hipGetLastError());
    return AMRGPU_OK;
}

// One gradient evaluation over `count` rows; partial sums reduced into
// m->grad / m->loss, with an Adam step when ad.on.
int launch_grad(amr_mlp *m, const uint16_t *dX, const int *dy, const int *dperm, long long first, int count,
                const float cw[3], float scale, int chunks, const AdamArgs &ad, float *dloss)
{
    int nblk = (count + ROWS * chunks - 1) / (ROWS * chunks);
    int rc = grow(&m->partial, &m->cap_partial, (size_t)nblk * NPP, false, 0);
    if (rc) return rc;
    GradArgs ga;
    ga.X = dX; ga.y = dy; ga.perm = dperm; ga.first = first; ga.count = count; ga.chunks = chunks;
    ga.cw[0] = cw[0]; ga.cw[1] = cw[1]; ga.cw[2] = cw[2];
    ga.scale = scale;
    ga.partial = m->partial;
    hipLaunchKernelGGL(grad_kernel, dim3(nblk), dim3(BLK), 0, m->stream, weights_of(m), ga);
    MCHK(hipGetLastError());
    hipLaunchKernelGGL(reduce_kernel, dim3((NP + 1 + 255) / 256), dim3(256), 0, m->stream, m->partial, nblk,
                       m->grad, dloss, m->P, m->M, m->V, m->W1b, m->W2b, m->W3b, ad);
    MCHK(hipGetLastError());
    return AMRGPU_OK;
}

void default_weights(const std::vector<int> &y, int n, float cw[3])
{
    long long cnt[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) cnt[y[i]]++;
    for (int c = 0; c < 3; ++c) {
        double v = cnt[c] ? (double)n / (3.0 * (double)cnt[c]) : 1.0;
        cw[c] = (float)std::min(20.0, std::max(0.2, v));
    }
}

}  // namespace

extern "C" {

amr_mlp_t *amr_mlp_create(unsigned long long seed)
{
    amr_mlp *m = new amr_mlp();
    std::vector<float> P(NP, 0.0f);
    uint64_t s = seed * 0x2545F4914F6CDD1Dull + 1;
    const double a1 = std::sqrt(6.0 / NIN), a2 = std::sqrt(6.0 / NH);
    for (int i = 0; i < NH * NIN; ++i) P[OW1 + i] = (float)((2.0 * uniform01(s) - 1.0) * a1);
    for (int i = 0; i < NH * NH; ++i) P[OW2 + i] = (float)((2.0 * uniform01(s) - 1.0) * a2);
    for (int i = 0; i < NOUT * NH; ++i) P[OW3 + i] = (float)((2.0 * uniform01(s) - 1.0) * a2);
    bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->P, sizeof(float) * NP) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->M, sizeof(float) * NP) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->V, sizeof(float) * NP) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->grad, sizeof(float) * NP) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->loss, sizeof(float) * 4) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->W1b, sizeof(uint16_t) * NH * NIN) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->W2b, sizeof(uint16_t) * NH * NH) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->W3b, sizeof(uint16_t) * 16 * NH) == hipSuccess;
    ok = ok && hipMemset(m->W3b, 0, sizeof(uint16_t) * 16 * NH) == hipSuccess;
    if (!ok) {
        amr_set_error("amr_mlp_create: device allocation failed");
        amr_mlp_destroy(m);
        return nullptr;
    }
    if (amr_mlp_set_params(m, P.data())) {
        amr_mlp_destroy(m);
        return nullptr;
    }
    return m;
}

void amr_mlp_destroy(amr_mlp_t *m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void *bufs[] = {m->P, m->M, m->V, m->grad, m->loss, m->W1b, m->W2b, m->W3b, m->partial,
                    m->X, m->y, m->perm, m->cls, m->logits};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int amr_mlp_get_params(amr_mlp_t *m, float *params)
{
    if (!m || !params) { amr_set_error("amr_mlp_get_params: null argument"); return AMRGPU_ERR_ARG; }
    MCHK(hipStreamSynchronize(m->stream));
    MCHK(hipMemcpy(params, m->P, sizeof(float) * NP, hipMemcpyDeviceToHost));
    return AMRGPU_OK;
}

int amr_mlp_set_params(amr_mlp_t *m, const float *params)
{
    if (!m || !params) { amr_set_error("amr_mlp_set_params: null argument"); return AMRGPU_ERR_ARG; }
    MCHK(hipStreamSynchronize(m->stream));
    MCHK(hipMemcpy(m->P, params, sizeof(float) * NP, hipMemcpyHostToDevice));
    MCHK(hipMemset(m->M, 0, sizeof(float) * NP));
    MCHK(hipMemset(m->V, 0, sizeof(float) * NP));
    hipLaunchKernelGGL(bf16_copy_kernel, dim3((NP + 255) / 256), dim3(256), 0, m->stream, m->P, m->W1b, m->W2b,
                       m->W3b);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(m->stream));
    m->step = 0;
    return AMRGPU_OK;
}

int amr_mlp_set_refine_bias(amr_mlp_t *m, float bias)
{
    if (!m || !std::isfinite(bias)) { amr_set_error("amr_mlp_set_refine_bias: bad argument"); return AMRGPU_ERR_ARG; }
    m->refine_bias = bias;
    return AMRGPU_OK;
}

int amr_mlp_forward(amr_mlp_t *m, const uint16_t *X, int n, float *logits)
{
    if (!m || n < 0 || (n && (!X || !logits))) { amr_set_error("amr_mlp_forward: bad argument"); return AMRGPU_ERR_ARG; }
    if (!n) return AMRGPU_OK;
    int rc = grow(&m->X, &m->cap_X, (size_t)n * NIN, false, 0);
    if (!rc) rc = grow(&m->logits, &m->cap_logits, (size_t)n * 3, false, 0);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(m->X, X, sizeof(uint16_t) * n * NIN, hipMemcpyHostToDevice, m->stream));
    rc = launch_infer(m, m->X, n, m->logits, nullptr, m->stream);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(logits, m->logits, sizeof(float) * n * 3, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return AMRGPU_OK;
}

int amr_mlp_grad(amr_mlp_t *m, const uint16_t *X, const int *y, int n, const float cw[3], float *grad,
                 float *loss)
{
    if (!m || !X || !y || n <= 0 || !cw || !grad) { amr_set_error("amr_mlp_grad: bad argument"); return AMRGPU_ERR_ARG; }
    double sw = 0.0;
    for (int i = 0; i < n; ++i) {
        if (y[i] < 0 || y[i] > 2) { amr_set_error("amr_mlp_grad: label %d of row %d", y[i], i); return AMRGPU_ERR_ARG; }
        sw += cw[y[i]];
    }
    if (!(sw > 0.0)) { amr_set_error("amr_mlp_grad: zero total weight"); return AMRGPU_ERR_ARG; }
    int rc = grow(&m->X, &m->cap_X, (size_t)n * NIN, false, 0);
    if (!rc) rc = grow(&m->y, &m->cap_y, (size_t)n, false, 0);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(m->X, X, sizeof(uint16_t) * n * NIN, hipMemcpyHostToDevice, m->stream));
    MCHK(hipMemcpyAsync(m->y, y, sizeof(int) * n, hipMemcpyHostToDevice, m->stream));
    AdamArgs ad;
    memset(&ad, 0, sizeof(ad));
    rc = launch_grad(m, m->X, m->y, nullptr, 0, n, cw, (float)(1.0 / sw), 1, ad, m->loss);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(grad, m->grad, sizeof(float) * NP, hipMemcpyDeviceToHost, m->stream));
    float l = 0.0f;
    MCHK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    if (loss) *loss = l;
    return AMRGPU_OK;
}

amr_tagset_t *amr_tagset_create(void) { return new amr_tagset(); }

void amr_tagset_destroy(amr_tagset_t *s)
{
    if (!s) return;
    if (s->X) (void)hipFree(s->X);
    if (s->y) (void)hipFree(s->y);
    delete s;
}

int amr_tagset_size(const amr_tagset_t *s) { return s ? s->n : 0; }

int amr_tagset_append(amr_tagset_t *s, const uint16_t *X, const int *y, const double *ind_now, int n)
{
    if (!s || n < 0 || (n && (!X || !y))) { amr_set_error("amr_tagset_append: bad argument"); return AMRGPU_ERR_ARG; }
    for (int i = 0; i < n; ++i)
        if (y[i] < 0 || y[i] > 2) { amr_set_error("amr_tagset_append: label %d of row %d", y[i], i); return AMRGPU_ERR_ARG; }
    size_t need = (size_t)s->n + n;
    int rc = grow(&s->X, &s->cap, need * NIN, true, (size_t)s->n * NIN);
    if (!rc) rc = grow(&s->y, &s->capy, need, true, (size_t)s->n);
    if (rc) return rc;
    MCHK(hipMemcpy(s->X + (size_t)s->n * NIN, X, sizeof(uint16_t) * n * NIN, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(s->y + s->n, y, sizeof(int) * n, hipMemcpyHostToDevice));
    s->h_y.insert(s->h_y.end(), y, y + n);
    for (int i = 0; i < n; ++i) s->h_ind.push_back(ind_now ? ind_now[i] : 0.0);
    s->n += n;
    return AMRGPU_OK;
}

int amr_tagset_get(const amr_tagset_t *s, int first, int n, uint16_t *X, int *y, double *ind_now)
{
    if (!s || first < 0 || n < 0 || first + n > s->n) { amr_set_error("amr_tagset_get: rows out of range"); return AMRGPU_ERR_ARG; }
    if (!n) return AMRGPU_OK;
    if (X) MCHK(hipMemcpy(X, s->X + (size_t)first * NIN, sizeof(uint16_t) * n * NIN, hipMemcpyDeviceToHost));
    if (y) memcpy(y, s->h_y.data() + first, sizeof(int) * n);
    if (ind_now) memcpy(ind_now, s->h_ind.data() + first, sizeof(double) * n);
    return AMRGPU_OK;
}

int amr_tagset_append_from(amr_tagset_t *dst, const amr_tagset_t *src, int first, int n)
{
    if (!dst || !src || dst == src || first < 0 || n < 0 || first + n > src->n) {
        amr_set_error("amr_tagset_append_from: rows out of range");
        return AMRGPU_ERR_ARG;
    }
    size_t need = (size_t)dst->n + n;
    int rc = grow(&dst->X, &dst->cap, need * NIN, true, (size_t)dst->n * NIN);
    if (!rc) rc = grow(&dst->y, &dst->capy, need, true, (size_t)dst->n);
    if (rc) return rc;
    if (n) {
        MCHK(hipMemcpy(dst->X + (size_t)dst->n * NIN, src->X + (size_t)first * NIN, sizeof(uint16_t) * n * NIN,
                       hipMemcpyDeviceToDevice));
        MCHK(hipMemcpy(dst->y + dst->n, src->y + first, sizeof(int) * n, hipMemcpyDeviceToDevice));
    }
    dst->h_y.insert(dst->h_y.end(), src->h_y.begin() + first, src->h_y.begin() + first + n);
    dst->h_ind.insert(dst->h_ind.end(), src->h_ind.begin() + first, src->h_ind.begin() + first + n);
    dst->n += n;
    return AMRGPU_OK;
}

int amrgpu_tagger_features(amrgpu_ctx *c, uint16_t *d_X)
{
    const double *U;
    const int *nbr;
    int nquad, n;
    void *st;
    int rc = amrgpu_device_state(c, &U, &nbr, &nquad, &n, &st);
    if (rc) return rc;
    if (n != 8 && n != 16) {
        amr_set_error("tagger features need 8x8 or 16x16 patches (got %d)", n);
        return AMRGPU_ERR_ARG;
    }
    if (!d_X) { amr_set_error("amrgpu_tagger_features: null output"); return AMRGPU_ERR_ARG; }
    if (n == 16) {
        // the features of the patches coarsened to 8x8, in the context's
        // staging buffer (a quarter of one state vector)
        double *P;
        size_t cap;
        rc = amrgpu_staging(c, &P, &cap);
        if (rc) return rc;
        if (cap < (size_t)nquad * 256) { amr_set_error("amrgpu_tagger_features: staging too small"); return AMRGPU_ERR_ARG; }
        hipLaunchKernelGGL(pool16_kernel, dim3((unsigned)(((long long)nquad * 64 + 255) / 256)), dim3(256), 0,
                           (hipStream_t)st, U, P, nquad);
        MCHK(hipGetLastError());
        U = P;
    }
    hipLaunchKernelGGL(features_kernel, dim3((nquad + 3) / 4), dim3(256), 0, (hipStream_t)st, U, nbr, nquad, d_X);
    MCHK(hipGetLastError());
    return AMRGPU_OK;
}

int amrgpu_tagger_features_host(amrgpu_ctx *c, uint16_t *X)
{
    int nquad = amrgpu_num_quads(c);
    void *st;
    int rc = amrgpu_device_state(c, nullptr, nullptr, nullptr, nullptr, &st);
    if (rc) return rc;
    if (!X) { amr_set_error("amrgpu_tagger_features_host: null output"); return AMRGPU_ERR_ARG; }
    uint16_t *d = nullptr;
    MCHK(hipMalloc((void **)&d, sizeof(uint16_t) * (size_t)nquad * NIN));
    rc = amrgpu_tagger_features(c, d);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(X, d, sizeof(uint16_t) * (size_t)nquad * NIN, hipMemcpyDeviceToHost,
                                      (hipStream_t)st);
        if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)st);
        if (e != hipSuccess) {
            amr_set_error("amrgpu_tagger_features_host: %s", hipGetErrorString(e));
            rc = AMRGPU_ERR_HIP;
        }
    }
    (void)hipFree(d);
    return rc;
}

int amr_mlp_train(amr_mlp_t *m, const amr_tagset_t *s, int epochs, int batch, float lr, const float *cw_in,
                  unsigned long long seed, float *loss)
{
    if (!m || !s || epochs < 0 || batch <= 0 || batch % ROWS || s->n < batch || !(lr > 0.0f)) {
        amr_set_error("amr_mlp_train: bad argument (batch %d, rows %d)", batch, s ? s->n : 0);
        return AMRGPU_ERR_ARG;
    }
    const int N = s->n, nb = N / batch;
    float cw[3];
    if (cw_in) memcpy(cw, cw_in, sizeof(cw));
    else default_weights(s->h_y, N, cw);
    const int chunks = std::max(1, batch / (ROWS * 32));
    int rc = grow(&m->perm, &m->cap_perm, (size_t)N, false, 0);
    if (!rc) rc = grow(&m->logits, &m->cap_logits, (size_t)nb, false, 0);   // per-minibatch losses
    if (rc) return rc;
    std::vector<int> perm(N);
    std::vector<float> lb(nb);
    for (int i = 0; i < N; ++i) perm[i] = i;
    AdamArgs ad;
    ad.on = 1;
    ad.lr = lr;
    ad.beta1 = 0.9f;
    ad.beta2 = 0.999f;
    ad.eps = 1e-8f;
    for (int e = 0; e < epochs; ++e) {
        uint64_t st = seed * 0x9E3779B97F4A7C15ull + (uint64_t)e + 1;
        for (int i = N - 1; i > 0; --i) {
            int j = (int)(splitmix64(st) % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        MCHK(hipMemcpyAsync(m->perm, perm.data(), sizeof(int) * N, hipMemcpyHostToDevice, m->stream));
        for (int b = 0; b < nb; ++b) {
            double sw = 0.0;
            for (int i = 0; i < batch; ++i) sw += cw[s->h_y[perm[(size_t)b * batch + i]]];
            m->step++;
            ad.bc1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)m->step)));
            ad.bc2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)m->step)));
            rc = launch_grad(m, s->X, s->y, m->perm, (long long)b * batch, batch, cw, (float)(1.0 / sw), chunks,
                             ad, m->logits + b);
            if (rc) return rc;
        }
        MCHK(hipMemcpyAsync(lb.data(), m->logits, sizeof(float) * nb, hipMemcpyDeviceToHost, m->stream));
        MCHK(hipStreamSynchronize(m->stream));
        if (loss) {
            double a = 0.0;
            for (int b = 0; b < nb; ++b) a += lb[b];
            loss[e] = (float)(a / nb);
        }
    }
    return AMRGPU_OK;
}

int amr_mlp_predict_set(amr_mlp_t *m, const amr_tagset_t *s, int *cls)
{
    if (!m || !s || (s->n && !cls)) { amr_set_error("amr_mlp_predict_set: bad argument"); return AMRGPU_ERR_ARG; }
    if (!s->n) return AMRGPU_OK;
    int rc = grow(&m->cls, &m->cap_cls, (size_t)s->n, false, 0);
    if (!rc) rc = launch_infer(m, s->X, s->n, nullptr, m->cls, m->stream);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(cls, m->cls, sizeof(int) * s->n, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return AMRGPU_OK;
}

int amr_mlp_predict_ctx(amr_mlp_t *m, amrgpu_ctx *c, int *cls)
{
    if (!m || !c || !cls) { amr_set_error("amr_mlp_predict_ctx: bad argument"); return AMRGPU_ERR_ARG; }
    void *st;
    int nquad;
    int rc = amrgpu_device_state(c, nullptr, nullptr, &nquad, nullptr, &st);
    if (!rc) rc = grow(&m->X, &m->cap_X, (size_t)nquad * NIN, false, 0);
    if (!rc) rc = grow(&m->cls, &m->cap_cls, (size_t)nquad, false, 0);
    if (!rc) rc = amrgpu_tagger_features(c, m->X);
    if (!rc) rc = launch_infer(m, m->X, nquad, nullptr, m->cls, (hipStream_t)st);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(cls, m->cls, sizeof(int) * nquad, hipMemcpyDeviceToHost, (hipStream_t)st));
    MCHK(hipStreamSynchronize((hipStream_t)st));
    return AMRGPU_OK;
}

}  // extern "C"
