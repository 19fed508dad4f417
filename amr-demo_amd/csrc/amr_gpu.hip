// amr_gpu.hip -- CDNA4 (gfx950) kernels and the device-resident context for the
// finite-volume Euler hot path of amr-demo: per-quadrant sub-grid patch
// residual (first-order faces, HLLE-type "Roe" flux, conforming and 2:1
// hanging faces, wall/copy boundaries), the RK stage update and the CFL
// time-step reduction, fused into one kernel per RK stage.
//
// Reference arithmetic: src/solver/src/flux_2d_module.F90:63-182 (flux),
// bc_2d_module.F90:14-75 + flux_2d_module.F90:13-60 (boundary flux),
// solver_residual_subgrid_2d_module.F90:32-164 (interior faces, /V),
// solver_residual_face_{full_structured_2d_module,full_hang_structured_2d,
// hang_full_structured_2d}.F90 (patch faces), solver_time_step_2d_module.F90
// :26-122 (CFL), src/physics/src/rk_time_integrators.c (RK schemes).
//
// HBM layout: state U[quad][field][j][i] (FP64 struct-of-arrays per patch,
// 4*n*n doubles per quadrant, patches in Morton order); per-quadrant face
// table nbr[quad][side][2]; per-quadrant level and instance id.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "amr_hip.h"
#include "amr_internal.h"

// No multiply-add is fused unless the source says fma(): every kernel
// instantiation (tile variants, the paired and the task rounds) then rounds
// alike, so the two evaluations of a face that two patches share agree bit
// for bit and a run's results do not depend on the tile it ran with.
#pragma clang fp contract(off)

namespace {

constexpr int NFLD = 4;
constexpr unsigned long long DBL_INF_BITS = 0x7ff0000000000000ull;
// The CFL minimum of an instance is spread over NSLOT sub-slots (workgroup b
// folds into sub-slot b % NSLOT): atomics on one address serialise, and on a
// 16.8M-cell mesh with one patch per workgroup 64 sub-slots still left the
// last stage ~17 % slower than the others (4096 atomics per address); 256
// recover most of it, and the one wave per instance of the next step's
// dt_prep_kernel reduces them with four loads per lane.
constexpr int NSLOT = 256;

struct InstDev {
    double h_base;
    int bc[4];
    double cfl;
    double dt_cap;
    double t_end;
    double bstate[2][4];              // boundary states of wall types 3-5
    double shock_x0, shock_cot, shock_vx;   // x_s(y,t) = x0 + y cot(a) + vx t
    // with h0 = h_base / n, the sub-cell size at level 0: -1 / (h0 h0) and
    // cfl sqrt(h0 h0).  Sizes of other levels are h0 2^-level, so -1 / (h h)
    // and cfl sqrt(h h) are these scaled by 4^level and 2^-level exactly
    // (one ldexp per patch instead of a divide or a square root)
    double ninv_vol0, cfl_h0;
};

// Per-quadrant face table entry for side s (0 xlo, 1 xhi, 2 ylo, 3 yhi):
//   a <  0          physical boundary (bc type from the instance)
//   a >= 0, b < 0   one neighbour a whose opposite edge couples index-aligned
//                   (same level, or a coarser quadrant across a hanging face)
//   a >= 0, b >= 0  two finer neighbours (a, b) in hanging order
struct MeshDev {
    int nquad;
    int ninst;
    const int *nbr;      // nquad*8
    const int *level;    // nquad
    const int *inst;     // nquad
    const InstDev *ip;   // ninst
    const double *org;   // nquad*2 lower-left corner (position-dependent walls)
    const double *t_step;// ninst: instance time at the start of the step
};

enum { MODE_SO = 0, MODE_RK4_ACC = 1, MODE_RK4_FINAL = 2, MODE_LSERK = 3, MODE_RHS = 4 };
enum { ROLE_CFL = 4 };

struct StageArgs {
    MeshDev m;
    const double *Us;    // state the residual is evaluated on
    const double *U0;    // base state of the step (may alias Uout pointwise)
    double *Uout;
    double *Acc;         // RK4 accumulator / LSERK register
    double c0, c1, c2;
    int mode;
    int first;           // first stage of an accumulating scheme
    int role;            // ROLE_CFL: the stage folds the next step's CFL minimum
    const double *dt_cur;// ninst: the step's dt (dt_prep_kernel)
    unsigned long long *slot;
    double *norm_part;
    int per_block;       // persistent design: patches per workgroup (a contiguous range)
};

__device__ __forceinline__ double bits2d(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long d2bits(double d) { return (unsigned long long)__double_as_longlong(d); }
// A device pointer taken as a global-memory one: loads through it are
// global_load (vector memory counter only) instead of flat loads, which
// also count against the LDS counter and so make every later LDS wait wait
// for them too.
__device__ __forceinline__ const double __attribute__((address_space(1))) *gptr(const double *p)
{
    return (const double __attribute__((address_space(1))) *)p;
}

// Logical block index with each XCD receiving a contiguous range, so that
// Morton-adjacent patches (which share halo cells) are processed by the same
// XCD and its L2.  Workgroups are dispatched round-robin over the 8 XCDs.
__device__ __forceinline__ int xcd_block(int b, int nblk)
{
    const int NX = 8;
    int q = nblk / NX, r = nblk % NX;
    int x = b % NX, k = b / NX;
    return x * q + (x < r ? x : r) + k;
}

struct Prim {
    double r, ru, rv, rE, u, v, p, H, sr;
};

// 1/x from the hardware estimate (v_rcp_f64) refined by two Newton steps in
// FMA form: within ~1 ulp of the correctly rounded quotient the reference
// computes, at half the instructions of an IEEE divide.
__device__ __forceinline__ double frcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// sqrt(x) and 1/sqrt(x) together from the hardware estimate (v_rsq_f64)
// refined by two coupled Newton (Goldschmidt) steps and a final residual
// correction of the root; both within ~1 ulp.  Arguments are positive
// normal numbers (densities, squared sound speeds).
__device__ __forceinline__ void fsqrt_rsq(double x, double &s, double &rs)
{
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    s = fma(d, h, g);
    rs = 2.0 * h;
}

// A face's primitive record oriented along the face normal: n is the normal
// direction (x for an x face), t the tangential one.
struct OPrim {
    double r, rn, rt, rE, un, ut, p, H, sr;
};

__device__ __forceinline__ OPrim orient(const Prim &P, bool xdir)
{
    OPrim o;
    o.r = P.r; o.rE = P.rE; o.p = P.p; o.H = P.H; o.sr = P.sr;
    o.rn = xdir ? P.ru : P.rv;
    o.rt = xdir ? P.rv : P.ru;
    o.un = xdir ? P.u : P.v;
    o.ut = xdir ? P.v : P.u;
    return o;
}

// HLLE-type flux of flux_2d_module.F90:63-182 on an axis-aligned face of
// length h, in the face's own frame (unit normal +n): F[0] mass, F[1] normal
// and F[2] tangential momentum, F[3] energy.  x and y faces are the same
// arithmetic on oriented records, so both sides of a face and both
// directions round alike.  sqrt(rho_l*rho_r) is taken as
// sqrt(rho_l)*sqrt(rho_r), abar and 1/abar come from one rsq, and the
// divisions become products with refined reciprocals: all within a few ulp
// of the reference's correctly rounded operations.
__device__ __forceinline__ void flux_normal(const OPrim &L, const OPrim &R, double h, double F[4])
{
    double inv_rf = frcp(L.sr + R.sr);
    double rhobar = L.sr * R.sr;
    // the low side's products stand alone: a cell's two paired faces share them
    double unbar = fma(R.un, R.sr, L.un * L.sr) * inv_rf;
    double utbar = fma(R.ut, R.sr, L.ut * L.sr) * inv_rf;
    double Hbar = fma(R.H, R.sr, L.H * L.sr) * inv_rf;
    double abar, inva;
    fsqrt_rsq(0.4 * fma(-0.5, fma(unbar, unbar, utbar * utbar), Hbar), abar, inva);
    double Mn = unbar * inva;
    double eigp = unbar + abar, eigm = unbar - abar;
    double eigp_max = R.un + abar, eigm_min = L.un - abar;
    // b1 = max(eig+, eig+_max) if either is positive, else 0 (b2 alike):
    // the same values as the reference's branches, NaN operands included
    double b1 = fmax(fmax(eigp, eigp_max), 0.0);
    double b2 = fmin(fmin(eigm, eigm_min), 0.0);
    // 1/(b1 - b2) and 1/(1 + |Mn|) from one refined reciprocal of their
    // product (b1 - b2 > 0 whenever the sound speed is positive)
    const double d1 = b1 - b2, d2 = 1.0 + fabs(Mn);
    const double rr = frcp(d1 * d2);
    double invb = rr * d2;
    double Mc = rr * d1;
    double bcoeff = b1 * b2 * invb;
    double delta1 = fma(-(R.p - L.p), inva * inva, R.r - L.r);
    double dH = R.H - L.H;
    // the face length folded into the coefficients
    const double invbh = invb * h;
    const double c1 = b1 * invbh, c2 = b2 * invbh;
    const double bch = bcoeff * h, bMh = bcoeff * Mc * h;
    double fl1 = fma(L.rn, L.un, L.p), fr1 = fma(R.rn, R.un, R.p);
    double fl2 = L.rn * L.ut, fr2 = R.rn * R.ut;
    double dn = delta1 * unbar;                           // the shear term vanishes along n
    double dt = fma(delta1, utbar, rhobar * (R.ut - L.ut));
    double fl3 = (L.rE + L.p) * L.un, fr3 = (R.rE + R.p) * R.un;
    F[0] = fma(-bMh, delta1, fma(bch, R.r - L.r, fma(-c2, R.rn, L.rn * c1)));
    F[1] = fma(-bMh, dn, fma(bch, R.rn - L.rn, fma(-c2, fr1, fl1 * c1)));
    F[2] = fma(-bMh, dt, fma(bch, R.rt - L.rt, fma(-c2, fr2, fl2 * c1)));
    F[3] = fma(-bMh, fma(delta1, Hbar, rhobar * dH),
               fma(bch, fma(R.r, R.H, -(L.r * L.H)), fma(-c2, fr3, fl3 * c1)));
}

// Flux through an x face (XDIR) or y face in the (x, y) frame.
template <bool XDIR>
__device__ __forceinline__ void flux_face(const Prim &L, const Prim &R, double h, double F[4])
{
    double G[4];
    flux_normal(orient(L, XDIR), orient(R, XDIR), h, G);
    F[0] = G[0];
    F[1] = XDIR ? G[1] : G[2];
    F[2] = XDIR ? G[2] : G[1];
    F[3] = G[3];
}

// The same with the direction known only at run time (lanes of one wave
// may differ): one flux evaluation on selected operands instead of a
// divergent pair.
__device__ __forceinline__ void flux_face_dir(const Prim &L, const Prim &R, double h, bool xdir, double F[4])
{
    double G[4];
    flux_normal(orient(L, xdir), orient(R, xdir), h, G);
    F[0] = G[0];
    F[1] = xdir ? G[1] : G[2];
    F[2] = xdir ? G[2] : G[1];
    F[3] = G[3];
}

// Physical boundary flux: bc_2d (bc_2d_module.F90:14-75) then flux_bc_2d
// (flux_2d_module.F90:13-60) with outward normal nrm = (+-h, 0) or (0, +-h),
// operation for operation.  Types 3-5 (project additions) take a prescribed
// state that depends on the face centre (org + (fx, fy) h) and time t.
__device__ __forceinline__ void flux_boundary(const InstDev &ip, int side, const double nrm[2], const double q[4],
                                              const double *org, double ofx, double ofy, double h, double t,
                                              double F[4])
{
    int bc = ip.bc[side];
    if (bc >= 3) {
        const double *st = ip.bstate[0];
        if (bc >= 4) {
            const double xf = org[0] + ofx * h, yf = org[1] + ofy * h;
            if (bc == 4) {
                if (!(xf < ip.shock_x0)) bc = 1;             // slip wall beyond x0
            } else if (!(xf < ip.shock_x0 + yf * ip.shock_cot + ip.shock_vx * t)) {
                st = ip.bstate[1];                           // ahead of the shock
            }
        }
        if (bc != 1) {
            const double r = st[0];
            const double oneOr = 1.0 / r;
            const double ub = st[1] * oneOr, vb = st[2] * oneOr;
            const double p = (1.4 - 1.0) * (st[3] - 0.5 * r * (ub * ub + vb * vb));
            const double fx4[4] = {st[1], st[1] * ub + p, st[1] * vb, ub * (st[3] + p)};
            const double fy4[4] = {st[2], st[2] * ub, st[2] * vb + p, vb * (st[3] + p)};
#pragma unroll
            for (int k = 0; k < 4; ++k) F[k] = fx4[k] * nrm[0] + fy4[k] * nrm[1];
            return;
        }
    }
    double rho = q[0];
    double oneOrho = 1.0 / rho;
    double u = q[1] * oneOrho;
    double v = q[2] * oneOrho;
    double qb[4];
    if (bc == 1) {
        double mag = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1]);
        double nx = nrm[0] / mag, ny = nrm[1] / mag;
        double Udn = u * nx + v * ny;
        double ub = u - Udn * nx, vb = v - Udn * ny;
        qb[0] = q[0]; qb[1] = q[0] * ub; qb[2] = q[0] * vb; qb[3] = q[3];
    } else if (bc == 2) {
        qb[0] = q[0]; qb[1] = 0.0; qb[2] = 0.0;
        qb[3] = q[3] - 0.5 * rho * (u * u + v * v);
    } else {
        qb[0] = q[0]; qb[1] = q[1]; qb[2] = q[2]; qb[3] = q[3];
    }
    double r = qb[0];
    double oneOr = 1.0 / r;
    double ub = qb[1] * oneOr, vb = qb[2] * oneOr;
    double p = (1.4 - 1.0) * (qb[3] - 0.5 * r * (ub * ub + vb * vb));
    double fx[4] = {qb[1], qb[1] * ub + p, qb[1] * vb, ub * (qb[3] + p)};
    double fy[4] = {qb[2], qb[2] * ub, qb[2] * vb + p, vb * (qb[3] + p)};
#pragma unroll
    for (int k = 0; k < 4; ++k) F[k] = fx[k] * nrm[0] + fy[k] * nrm[1];
}

// Per-cell CFL candidate, solver_time_step_2d_module.F90:107-116, evaluated
// without FMA contraction so that it reproduces the reference bit for bit;
// cfl_h = cfl * sqrt(vol), the patch's numerator.
__device__ __forceinline__ double cfl_candidate_h(const double q[4], double cfl_h)
{
#pragma clang fp contract(off)
    double oneOrho = 1.0 / q[0];
    double u = q[1] * oneOrho;
    double v = q[2] * oneOrho;
    double V2 = u * u + v * v;
    double pOverRho = 0.4 * (q[3] * oneOrho - 0.5 * V2);
    double c = sqrt(1.4 * pOverRho);
    double eig = sqrt(V2) + c;
    return cfl_h / eig;
}

__device__ __forceinline__ double cfl_candidate(const double q[4], double cfl, double vol)
{
#pragma clang fp contract(off)
    return cfl_candidate_h(q, cfl * sqrt(vol));
}

// v from the lane `CTRL` (a DPP row-shift control) selects, or v itself
// where the shift has no source lane
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    int lo = (int)b, hi = (int)(b >> 32);
    lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// A wave-uniform double moved to scalar registers (its value in lane 0):
// it then holds no vector registers while it waits for its uses.
__device__ __forceinline__ double uniform_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_f64(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Minimum over the 64 lanes of a wave (all lanes active), returned to every
// lane.  Row shifts by 1, 2, 4, 8 leave each 16-lane row's minimum in its
// last lane (DPP moves: no LDS round trips); the four row minima are read
// as scalars.  A minimum is exact, so the order does not change the value.
__device__ __forceinline__ double wave_min(double v)
{
    v = fmin(v, dpp_f64<0x111>(v));    // row_shr:1
    v = fmin(v, dpp_f64<0x112>(v));    // row_shr:2
    v = fmin(v, dpp_f64<0x114>(v));    // row_shr:4
    v = fmin(v, dpp_f64<0x118>(v));    // row_shr:8
    return fmin(fmin(lane_f64(v, 15), lane_f64(v, 31)), fmin(lane_f64(v, 47), lane_f64(v, 63)));
}
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide min of the CFL candidates, then one 64-bit atomicMin per
// instance (positive doubles order like their bit patterns).  uinst is the
// block's common instance, or -1 when the block straddles two instances
// (then every thread issues its own atomic).
template <int BLK>
__device__ void cfl_reduce(double cand, int inst, bool active, int uinst,
                           unsigned long long *slot, double *red)
{
    const int tid = threadIdx.x;
    if (uinst < 0) {
        if (active) atomicMin(&slot[inst * NSLOT + (blockIdx.x & (NSLOT - 1))], d2bits(cand));
        return;
    }
    double v = active ? cand : INFINITY;
    v = wave_min(v);
    constexpr int NW = (BLK + 63) / 64;
    if (NW == 1) {
        if (tid == 0) atomicMin(&slot[uinst * NSLOT + (blockIdx.x & (NSLOT - 1))], d2bits(v));
        return;
    }
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double m = red[0];
        for (int w = 1; w < NW; ++w) m = fmin(m, red[w]);
        atomicMin(&slot[uinst * NSLOT + (blockIdx.x & (NSLOT - 1))], d2bits(m));
    }
}

template <int BLK>
__device__ void sum_reduce_store(double v, double *out, double *red)
{
    const int tid = threadIdx.x;
    v = wave_sum(v);
    constexpr int NW = (BLK + 63) / 64;
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double s = red[0];
        for (int w = 1; w < NW; ++w) s += red[w];
        *out = s;
    }
}

// CFL minimum of instance inst over its sub-slots; called by a whole wave
// (every lane returns the result).
__device__ __forceinline__ double slot_min(const unsigned long long *slot, int inst)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long *s = slot + (size_t)inst * NSLOT;
    double m = INFINITY;
#pragma unroll
    for (int k = 0; k < (NSLOT + 63) / 64; ++k)
        if (lane + 64 * k < NSLOT) m = fmin(m, bits2d(s[lane + 64 * k]));
    return wave_min(m);
}

// dt of the step for instance inst: min(input dt, CFL minimum, time left);
// called by a whole wave.
__device__ __forceinline__ double step_dt(const InstDev &ip, const unsigned long long *slot,
                                          const double *t_inst, int inst, int use_cfl)
{
    double dt = ip.dt_cap;
    if (use_cfl) {
        double c = slot_min(slot, inst);
        if (c < dt) dt = c;
    }
    if (ip.t_end > 0.0) {
        double rem = ip.t_end - t_inst[inst];
        if (rem < 0.0) rem = 0.0;
        if (rem < dt) dt = rem;
    }
    return dt;
}

// ---------------------------------------------------------------------------
// Fused stage kernel: one workgroup = PPB patches of N x N cells, one thread
// per cell.  LDS holds the patches' primitive variables with a one-cell ghost
// ring gathered from the neighbours (two rings' worth on coarse sides of
// hanging faces); every face flux of the patch is evaluated once per patch
// into registers, staged to LDS and gathered per cell; the epilogue divides
// by the volume, applies the RK combination, writes the new state
// (coalesced) and, on the last stage, folds the CFL minimum and the residual
// sum of squares into block reductions.
// ---------------------------------------------------------------------------
template <int N, int PPB>
struct Tile {
    static constexpr int N2 = N * N;
    static constexpr int BLK = N2 * PPB;
    // LDS record per tile position: r, ru, rv, rE, p, 1/r, sqrt(r), one
    // array per quantity (array stride NPOS).  Positions [0, N2) are the
    // patch cells row by row (pitch N: the lanes of a wave read consecutive
    // addresses, free of bank conflicts), then GHOST + s N + k the ghost
    // cell of each side s = 0..3 (-x, +x, -y, +y) and SECOND + s N + k the
    // second fine neighbour of each coarse side of a hanging face.
    // 8x8 patches (SPLIT): the halo's 8N positions equal the N2 cells, so
    // the halo gets a block of its own after the cells' seven arrays, at the
    // same stride N2.  The face and second-neighbour arrays written after
    // the flux rounds then overlay everything but the cells' r, ru, rv, rE,
    // which the epilogue reads back instead of keeping its own copy: 7.5 KB
    // of LDS per patch instead of 9, which is what lets a CU hold 20
    // one-patch workgroups (five waves per SIMD).
    static constexpr int NP = 7;
    static constexpr bool SPLIT = (8 * N == N2);
    static constexpr int NPOS = SPLIT ? N2 : N2 + 8 * N;
    static constexpr int GHOST = SPLIT ? NP * N2 : N2;
    static constexpr int SECOND = GHOST + 4 * N;
    // tile position of cell (i, j), -1 <= i, j <= N, not a corner
    static __device__ __forceinline__ int pos(int i, int j)
    {
        if (i < 0) return GHOST + 0 * N + j;
        if (i >= N) return GHOST + 1 * N + j;
        if (j < 0) return GHOST + 2 * N + i;
        if (j >= N) return GHOST + 3 * N + i;
        return j * N + i;
    }
    static constexpr int PRIM_PER_PATCH = NP * NPOS;
    static constexpr int FLUX_PER_PATCH = NFLD * 2 * N * (N + 1);
    // the second-neighbour fluxes of each patch side (4 sides x
    // 4 fields x N) beside the face arrays, and the rounds of extra tasks
    // (at most 2N edge faces + 4N second fluxes per patch)
    static constexpr int F2_PER_PATCH = 4 * NFLD * N;
    // SPLIT: one block per patch -- cells' r, ru, rv, rE (4 N2), then the
    // faces and second-neighbour fluxes over the rest of the records
    static constexpr int PB = 4 * N2 + FLUX_PER_PATCH + F2_PER_PATCH;
    static constexpr int LDS_DOUBLES =
        SPLIT ? PPB * PB : PPB * (PRIM_PER_PATCH > FLUX_PER_PATCH ? PRIM_PER_PATCH : FLUX_PER_PATCH);
    static constexpr int F2BASE = PPB * FLUX_PER_PATCH;     // (not SPLIT)
    // where patch p's records, face arrays and second-neighbour arrays start
    static __device__ __forceinline__ constexpr int rec(int p) { return p * (SPLIT ? PB : PRIM_PER_PATCH); }
    static __device__ __forceinline__ constexpr int face(int p)
    {
        return SPLIT ? p * PB + 4 * N2 : p * FLUX_PER_PATCH;
    }
    static __device__ __forceinline__ constexpr int f2(int p)
    {
        return SPLIT ? p * PB + 4 * N2 + FLUX_PER_PATCH : F2BASE + p * F2_PER_PATCH;
    }
    static constexpr int XROUNDS = (PPB * 6 * N + BLK - 1) / BLK;
    static constexpr int NINT = 2 * N * (N - 1);           // interior faces per patch
    static constexpr int NEDGE = 4 * N;                    // patch-edge faces per patch
    static constexpr int NTASK = PPB * (NINT + NEDGE);
    static constexpr int ROUNDS = (NTASK + BLK - 1) / BLK;
    static_assert(SPLIT || F2BASE + PPB * F2_PER_PATCH <= LDS_DOUBLES, "second-neighbour fluxes fit beside the faces");
    static_assert(!SPLIT || PB >= 2 * NP * N2, "a patch's block holds its records");
};

template <int N, int PPB>
__device__ __forceinline__ void store_rec(double *lds, int p, int t, double r, double ru, double rv,
                                          double rE)
{
    using TL = Tile<N, PPB>;
    // the primitive variables of flux_2d_module.F90:93-99, with sqrt(rho)
    // and 1/rho from one rsq; u, v and H are formed again by load_prim
    double sr, rsr;
    fsqrt_rsq(r, sr, rsr);
    const double ir = rsr * rsr;
    const double u = ru * ir, v = rv * ir;
    // flux_2d_module.F90:98 pressure
    const double pr = (1.4 - 1.0) * fma(-(0.5 * r), fma(u, u, v * v), rE);
    double *b = lds + TL::rec(p) + t;
    b[0 * TL::NPOS] = r;  b[1 * TL::NPOS] = ru; b[2 * TL::NPOS] = rv; b[3 * TL::NPOS] = rE;
    b[4 * TL::NPOS] = pr; b[5 * TL::NPOS] = ir; b[6 * TL::NPOS] = sr;
}

template <int N, int PPB>
__device__ __forceinline__ void load_prim(const double *lds, int p, int t, Prim &P)
{
    using TL = Tile<N, PPB>;
    const double *b = lds + TL::rec(p) + t;
    P.r = b[0 * TL::NPOS]; P.ru = b[1 * TL::NPOS]; P.rv = b[2 * TL::NPOS]; P.rE = b[3 * TL::NPOS];
    P.p = b[4 * TL::NPOS];
    const double ir = b[5 * TL::NPOS];
    P.sr = b[6 * TL::NPOS];
    P.u = P.ru * ir;
    P.v = P.rv * ir;
    P.H = (P.rE + P.p) * ir;     // E + p/rho
}

// The same record oriented along a face normal chosen per lane: offn / offt
// are the LDS array offsets of the normal and tangential momenta.
template <int N, int PPB>
__device__ __forceinline__ void load_oprim(const double *lds, int p, int t, int offn, int offt, OPrim &P)
{
    using TL = Tile<N, PPB>;
    const double *b = lds + TL::rec(p) + t;
    P.r = b[0]; P.rn = b[offn]; P.rt = b[offt]; P.rE = b[3 * TL::NPOS];
    P.p = b[4 * TL::NPOS];
    const double ir = b[5 * TL::NPOS];
    P.sr = b[6 * TL::NPOS];
    P.un = P.rn * ir;
    P.ut = P.rt * ir;
    P.H = (P.rE + P.p) * ir;
}

// The stage update of one cell from its residual k4 = R(U) / V (the schemes
// of rk_time_integrators.c: Shu-Osher stages for SSP-RK2/3, the RK4
// accumulator, the 2N-storage LSERK45 register), shared by the stage
// kernels; B is the step's base state where the mode reads one.
__device__ __forceinline__ void stage_update(const StageArgs &a, size_t base, int N2, double dt, const double U[4],
                                             const double B[4], const double k4[4], double out[4])
{
    switch (a.mode) {
    case MODE_SO:
        if (a.U0) {
#pragma unroll
            for (int f = 0; f < 4; ++f) out[f] = fma(a.c2 * dt, k4[f], fma(a.c0, B[f], a.c1 * U[f]));
        } else {
#pragma unroll
            for (int f = 0; f < 4; ++f) out[f] = fma(a.c2 * dt, k4[f], a.c1 * U[f]);
        }
        break;
    case MODE_RK4_ACC:
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double acc = a.first ? k4[f] : fma(2.0, k4[f], a.Acc[base + f * N2]);
            a.Acc[base + f * N2] = acc;
            out[f] = fma(a.c2 * dt, k4[f], B[f]);
        }
        break;
    case MODE_RK4_FINAL:
#pragma unroll
        for (int f = 0; f < 4; ++f)
            out[f] = fma(dt / 6.0, a.Acc[base + f * N2] + k4[f], B[f]);
        break;
    case MODE_LSERK:
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double dq = a.first ? dt * k4[f] : fma(a.Acc[base + f * N2], a.c1, dt * k4[f]);
            a.Acc[base + f * N2] = dq;
            out[f] = fma(a.c2, dq, U[f]);
        }
        break;
    default:   // MODE_RHS
#pragma unroll
        for (int f = 0; f < 4; ++f) out[f] = k4[f];
        break;
    }
}

template <int N, int PPB, int MINW, int DESIGN>
__global__ __launch_bounds__((Tile<N, PPB>::BLK), (MINW)) void stage_kernel(StageArgs a)
{
    using TL = Tile<N, PPB>;
    constexpr int N2 = TL::N2;
    __shared__ double lds[TL::LDS_DOUBLES];
    // own state, kept out of registers (SPLIT: read back from the records)
    __shared__ double ldsU[TL::SPLIT ? 1 : PPB * NFLD * N2];
    __shared__ double red[(TL::BLK + 63) / 64];
    __shared__ int pquad[PPB], plevel[PPB], pinst[PPB];
    __shared__ int pnbr[PPB][8];
    __shared__ double pdt[PPB];
    __shared__ double ptime[PPB];
    __shared__ double ph[PPB];     // cell size of each patch

    const int tid = threadIdx.x;
    const int nblk = gridDim.x;
    const int lb = xcd_block(blockIdx.x, nblk);
    const int p = tid / N2, c = tid % N2;
    const int ci = c % N, cj = c / N;
    const int quad = lb * PPB + p;
    // one patch per workgroup: the grid has exactly one workgroup per patch
    const bool active = PPB == 1 || quad < a.m.nquad;
    // threads past the last patch read patch 0 (nquad >= 1 when launched)
    // and store nothing, so that the loads need no branch
    const int quadc = active ? quad : 0;
    const size_t base = (size_t)quadc * NFLD * N2 + c;
    // field f of the cell's own state, from LDS
    auto own_u = [=](int f) {
        return TL::SPLIT ? lds[TL::rec(p) + f * TL::NPOS + c] : ldsU[(p * NFLD + f) * N2 + c];
    };

    // ---- prologue -----------------------------------------------------------
    // The own-state loads and the halo loads go out first; the waits that
    // the dependent chains need (patch parameters behind the instance id,
    // the step's dt behind the CFL slots) come after them, so that no wait
    // holds back an independent load.  With one patch per wave (N2 a
    // multiple of 64) the patch's ids and parameters are wave-uniform and
    // are read through the scalar cache (constant address space): they wait
    // on their own counter, not behind the vector loads.
    constexpr bool UNI = (N2 % 64 == 0);
    constexpr bool ONE_GHOST = (4 * N <= N2);     // at most one halo cell per thread
    // Per-thread values below are left unset where only the paths that
    // assign them read them (an inactive thread stores and reduces nothing):
    // initialising them costs a move per register on every path.
    double U[4];
    double B[4];     // base state U0 of the step
    const bool needB = (a.mode == MODE_SO && a.U0) || a.mode == MODE_RK4_ACC || a.mode == MODE_RK4_FINAL;
#pragma unroll
    for (int f = 0; f < 4; ++f) U[f] = a.Us[base + f * N2];
    int nbw[8];
    int levw = 0, inw = 0;
    bool actw = false;
    // one paired-face patch per workgroup: every reader of the patch's ids
    // and parameters takes them from the wave's scalars, and the LDS copies
    // the other tiles share between waves are not written
    constexpr bool SMETA = (PPB == 1 && UNI && DESIGN == 1);
    double tpw = 0.0;     // the step's start time of the patch's instance (wave-uniform)
    if constexpr (UNI) {
        typedef const int __attribute__((address_space(4))) cint4;
        const int qw = __builtin_amdgcn_readfirstlane(quad);
        actw = PPB == 1 || qw < a.m.nquad;
        const int qc = actw ? qw : 0;      // a valid index (nquad >= 1 when launched); discarded if inactive
        const cint4 *nbc = (const cint4 *)a.m.nbr + (size_t)qc * 8;
#pragma unroll
        for (int k = 0; k < 8; ++k) nbw[k] = nbc[k];
        levw = ((const cint4 *)a.m.level)[qc];
        inw = ((const cint4 *)a.m.inst)[qc];
        if (!actw) {
#pragma unroll
            for (int k = 0; k < 8; ++k) nbw[k] = -1;
            levw = 0;
            inw = 0;
        }
    }
    // whether the thread's patch exists, as a wave-uniform (scalar) value
    // when a wave holds one patch: branches on it are scalar, and values
    // assigned under it need no definition for threads that skip it
    const bool actu = UNI ? actw : active;
    // halo: the ghost cell of side s = hh / N, position k = hh % N (and the
    // second fine neighbour of a coarse side), loaded here and stored as
    // records after the parameter chains below when there is at most one
    // per thread
    double g0, g1, g2, g3, s0, s1, s2, s3;
    bool gsec = false;
    // with at most one halo cell per thread, threads without one load side
    // 0's first (unused): one pass, no branch
    const int hh0 = ONE_GHOST ? (c < 4 * N ? c : 0) : c;
    for (int hh = hh0, it = 0; ONE_GHOST ? it < 1 : hh < 4 * N; hh += N2, ++it) {
        if (!ONE_GHOST && !active) break;
        const int sd = hh / N, k = hh % N;
        const bool sx = sd < 2, shi = sd & 1;    // an x side; the high (+) side
        int na, nb;
        if constexpr (UNI) {
            // selects of wave-uniform ids by two bits of the side: no branches
            const int na_x = shi ? nbw[2] : nbw[0], na_y = shi ? nbw[6] : nbw[4];
            const int nb_x = shi ? nbw[3] : nbw[1], nb_y = shi ? nbw[7] : nbw[5];
            na = sx ? na_x : na_y;
            nb = sx ? nb_x : nb_y;
        } else {
            na = a.m.nbr[quadc * 8 + 2 * sd];
            nb = a.m.nbr[quadc * 8 + 2 * sd + 1];
        }
        // neighbour's edge cell ec and own edge cell oc: row / column k of
        // the side, at the far (N - 1) or near (0) end of the normal
        const int kb = sx ? k * N : k, far = sx ? N - 1 : (N - 1) * N;
        const int ec = kb + (shi ? 0 : far), oc = kb + (shi ? far : 0);
        // physical boundary: the ghost holds the own edge cell, so that the
        // generic face evaluation stays finite before the boundary flux
        // replaces it
        const double *sa = na < 0 ? a.Us + (size_t)quadc * NFLD * N2 + oc
                                  : a.Us + (size_t)na * NFLD * N2 + ec;
        g0 = sa[0]; g1 = sa[N2]; g2 = sa[2 * N2]; g3 = sa[3 * N2];
        gsec = nb >= 0;
        if (gsec) {
            const double *sb = a.Us + (size_t)nb * NFLD * N2 + ec;
            s0 = sb[0]; s1 = sb[N2]; s2 = sb[2 * N2]; s3 = sb[3 * N2];
        }
        if constexpr (!ONE_GHOST) {
            if (gsec) store_rec<N, PPB>(lds, p, TL::SECOND + hh, s0, s1, s2, s3);
            store_rec<N, PPB>(lds, p, TL::GHOST + hh, g0, g1, g2, g3);
        }
    }
    // patch parameters that depend on the instance
    double hw = 0.0, dtw = 0.0;
    if constexpr (UNI) {
        typedef const InstDev __attribute__((address_space(4))) cinst4;
        typedef const double __attribute__((address_space(4))) cdbl4;
        const cinst4 *ipc = (const cinst4 *)a.m.ip;
        hw = actw ? ldexp(ipc[inw].h_base, -levw) / (double)N : 0.0;
        // the step's dt and start time, written by the step's prep kernel
        dtw = actw && a.mode != MODE_RHS ? ((const cdbl4 *)a.dt_cur)[inw] : 0.0;
        const double tw = actw ? ((const cdbl4 *)a.m.t_step)[inw] : 0.0;
        tpw = tw;
        if (!SMETA && c == 0) {
            pquad[p] = actw ? lb * PPB + p : -1;
            plevel[p] = levw;
            pinst[p] = inw;
#pragma unroll
            for (int k = 0; k < 8; ++k) pnbr[p][k] = nbw[k];
            ph[p] = hw;
            pdt[p] = dtw;
            ptime[p] = tw;
        }
    } else if (tid < PPB) {
        int qq = lb * PPB + tid;
        bool act = qq < a.m.nquad;
        const int lev = act ? a.m.level[qq] : 0, in = act ? a.m.inst[qq] : 0;
        pquad[tid] = act ? qq : -1;
        plevel[tid] = lev;
        pinst[tid] = in;
#pragma unroll
        for (int k = 0; k < 8; ++k) pnbr[tid][k] = act ? a.m.nbr[qq * 8 + k] : -1;
        ph[tid] = act ? ldexp(a.m.ip[in].h_base, -lev) / (double)N : 0.0;
        pdt[tid] = act && a.mode != MODE_RHS ? a.dt_cur[in] : 0.0;
        // the step's start time, stable while the time advances
        ptime[tid] = act ? a.m.t_step[in] : 0.0;
    }
    if constexpr (ONE_GHOST) {
        if (active && c < 4 * N) {
            if (gsec) store_rec<N, PPB>(lds, p, TL::SECOND + c, s0, s1, s2, s3);
            store_rec<N, PPB>(lds, p, TL::GHOST + c, g0, g1, g2, g3);
        }
    }
    if (actu) {
        store_rec<N, PPB>(lds, p, c, U[0], U[1], U[2], U[3]);
#pragma unroll
        for (int f = 0; f < 4; ++f)
            if constexpr (!TL::SPLIT) ldsU[(p * NFLD + f) * N2 + c] = U[f];
    }
    __syncthreads();

    // ---- face fluxes into registers --------------------------------------
    if constexpr (DESIGN == 1) {
        // paired faces: every cell evaluates its +x and +y faces from one
        // load of its own record (two independent flux chains per thread).
        // Every other face evaluation of the workgroup -- the -x and -y
        // patch edges, and the second fine neighbour of each side with two
        // -- is a task of one more round, packed onto as few waves as they
        // fill (a patch's extra fluxes would otherwise cost its wave a round
        // per kind); the epilogue adds each second-neighbour flux to its
        // first, in the same order as a single thread would.
        double FA[4], FB[4];
        if (actu) {
            const double h = UNI ? hw : ph[p];
            const int na1 = UNI ? nbw[2] : pnbr[p][2], nb1 = UNI ? nbw[3] : pnbr[p][3];
            const int na3 = UNI ? nbw[6] : pnbr[p][6], nb3 = UNI ? nbw[7] : pnbr[p][7];
            const double tp_ = UNI ? tpw : ptime[p];
            const bool ex = (ci == N - 1), ey = (cj == N - 1);
            // wall parameters: wave-uniform (scalar) when a wave holds one patch
            const int wq = UNI ? __builtin_amdgcn_readfirstlane(quad) : quad;
            const int wi = UNI ? inw : pinst[p];
            const double hx = (ex && nb1 >= 0) ? 0.5 * h : h;
            const double hy = (ey && nb3 >= 0) ? 0.5 * h : h;
            Prim L, RX, RY;
            load_prim<N, PPB>(lds, p, c, L);
            load_prim<N, PPB>(lds, p, ex ? TL::GHOST + 1 * N + cj : c + 1, RX);
            load_prim<N, PPB>(lds, p, ey ? TL::GHOST + 3 * N + ci : c + N, RY);
            flux_face<true>(L, RX, hx, FA);
            flux_face<false>(L, RY, hy, FB);
            // both fluxes complete here, before the wall branches: a wall
            // replaces the value, and the compiler would otherwise sink the
            // second flux past the first branch and lose the two chains'
            // interleaving (an empty statement, no instruction)
            asm volatile("" : "+v"(FA[0]), "+v"(FA[1]), "+v"(FA[2]), "+v"(FA[3]), "+v"(FB[0]), "+v"(FB[1]),
                         "+v"(FB[2]), "+v"(FB[3]));
            if (ex && na1 < 0) {
                const double nrm[2] = {h, 0.0};
                const double q[4] = {own_u(0), own_u(1), own_u(2), own_u(3)};
                flux_boundary(a.m.ip[wi], 1, nrm, q, a.m.org + 2 * (size_t)wq, (double)N,
                              cj + 0.5, h, tp_, FA);
            }
            if (ey && na3 < 0) {
                const double nrm[2] = {0.0, h};
                const double q[4] = {own_u(0), own_u(1), own_u(2), own_u(3)};
                flux_boundary(a.m.ip[wi], 3, nrm, q, a.m.org + 2 * (size_t)wq, ci + 0.5,
                              (double)N, h, tp_, FB);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        // the tasks start on a wave that rotates with the block index,
        // spreading the extra round over the SIMDs; when the halo is wave
        // 0's alone (one patch of several waves), on one of the others
        constexpr int NWV = (TL::BLK + 63) / 64;
        constexpr bool HALO0 = (PPB == 1 && NWV > 1 && 4 * N <= 64);
        const int ew = HALO0 ? 1 + lb % (NWV - 1) : lb % NWV;
        const int t0 = (tid - ew * 64 + TL::BLK) % TL::BLK;
        // one patch per workgroup held by whole waves: its neighbour ids are
        // the scalars nbw, and the task list's shape (length, the order of
        // the sides with two neighbours) is worked out once, in scalar code
        constexpr bool SCALAR_TASKS = (PPB == 1 && UNI);
        int T1 = 0, hs0 = 0, hs1 = 0, hs2 = 0, hs3 = 0;
        if constexpr (SCALAR_TASKS) {
            int m = 0;
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                const int so = o == 0 ? 1 : o == 1 ? 3 : o == 2 ? 0 : 2;     // +x, +y, -x, -y
                if (nbw[2 * so + 1] >= 0) {
                    if (m == 0) hs0 = so; else if (m == 1) hs1 = so; else if (m == 2) hs2 = so; else hs3 = so;
                    ++m;
                }
            }
            T1 = actw ? 2 * N + N * m : 0;
        }
        double FC[TL::XROUNDS][4];
        int dC[TL::XROUNDS], fnC[TL::XROUNDS];
        bool sC[TL::XROUNDS];      // a second neighbour's flux (side arrays, field stride N)
#pragma unroll
        for (int r = 0; r < TL::XROUNDS; ++r) {
            dC[r] = -1;
            fnC[r] = 1;
            int t = t0 + r * TL::BLK;
            // patch of task t: 2N edge faces, then N per side with two
            // neighbours, in the order +x, +y, -x, -y
            int tp = -1;
            if constexpr (SCALAR_TASKS) {
                if (t < T1) tp = 0;
            } else {
                for (int q = 0; q < PPB; ++q) {
                    if (pquad[q] < 0) break;
                    const int m = (pnbr[q][1] >= 0) + (pnbr[q][3] >= 0) + (pnbr[q][5] >= 0) + (pnbr[q][7] >= 0);
                    const int Tq = 2 * N + N * m;
                    if (t < Tq) { tp = q; break; }
                    t -= Tq;
                }
            }
            if (tp < 0) continue;
            const int kind = t / N, k = t % N;
            const double h = SCALAR_TASKS ? hw : ph[tp];
            bool xs, second, wall = false;
            bool two;     // the face's side has two (fine) neighbours: fine normal length
            int side, tl, tr;
            if (kind < 2) {                          // patch edge, first (or only) neighbour
                side = kind == 0 ? 0 : 2;
                xs = kind == 0;
                second = false;
                tl = TL::GHOST + side * N + k;
                tr = xs ? k * N : k;
                if constexpr (SCALAR_TASKS) {
                    wall = kind == 0 ? nbw[0] < 0 : nbw[4] < 0;
                    two = kind == 0 ? nbw[1] >= 0 : nbw[5] >= 0;
                } else {
                    wall = pnbr[tp][2 * side] < 0;
                    two = pnbr[tp][2 * side + 1] >= 0;
                }
            } else {                                 // second fine neighbour of the (kind-2)-th side with two
                two = true;
                if constexpr (SCALAR_TASKS) {
                    const int cnt = kind - 2;
                    side = cnt == 0 ? hs0 : cnt == 1 ? hs1 : cnt == 2 ? hs2 : hs3;
                } else {
                    side = -1;
                    int cnt = kind - 2;
#pragma unroll
                    for (int o = 0; o < 4; ++o) {
                        const int so = o == 0 ? 1 : o == 1 ? 3 : o == 2 ? 0 : 2;     // +x, +y, -x, -y
                        if (pnbr[tp][2 * so + 1] >= 0) {
                            if (cnt == 0 && side < 0) side = so;
                            --cnt;
                        }
                    }
                }
                xs = side < 2;
                second = true;
                const int own = side == 0 ? k * N : side == 1 ? k * N + (N - 1) : side == 2 ? k : (N - 1) * N + k;
                const int sec = TL::SECOND + side * N + k;
                const bool low = (side == 0 || side == 2);    // the neighbour is on the low side
                tl = low ? sec : own;
                tr = low ? own : sec;
            }
            const double hf = two ? 0.5 * h : h;
            const int offn = xs ? TL::NPOS : 2 * TL::NPOS, offt = 3 * TL::NPOS - offn;
            OPrim L, R;
            load_oprim<N, PPB>(lds, tp, tl, offn, offt, L);
            load_oprim<N, PPB>(lds, tp, tr, offn, offt, R);
            flux_normal(L, R, hf, FC[r]);
            if (wall) {
                const double nrm[2] = {xs ? -h : 0.0, xs ? 0.0 : -h};
                const double q[4] = {R.r, xs ? R.rn : R.rt, xs ? R.rt : R.rn, R.rE};
                double F[4];
                const int wi = SCALAR_TASKS ? inw : pinst[tp];
                const int wq = SCALAR_TASKS ? __builtin_amdgcn_readfirstlane(quad) : pquad[tp];
                flux_boundary(a.m.ip[wi], side, nrm, q, a.m.org + 2 * (size_t)wq,
                              xs ? 0.0 : k + 0.5, xs ? k + 0.5 : 0.0, h, SCALAR_TASKS ? tpw : ptime[tp], F);
                FC[r][0] = -F[0]; FC[r][1] = -(xs ? F[1] : F[2]); FC[r][2] = -(xs ? F[2] : F[1]);
                FC[r][3] = -F[3];
            }
            dC[r] = second ? TL::f2(tp) + (side * NFLD) * N + k
                           : TL::face(tp) + (xs ? k * (N + 1) : NFLD * N * (N + 1) + k);
            fnC[r] = xs ? 1 : 2;
            sC[r] = second;
        }
        if (actu && needB) {   // base state: loaded late, its latency overlaps the flux gather
#pragma unroll
            for (int f = 0; f < 4; ++f) B[f] = gptr(a.U0)[base + f * N2];
        }
        __syncthreads();   // all reads of the primitive tile are done
        const int stride = N * (N + 1);
        if (actu) {
            const int dA = TL::face(p) + cj * (N + 1) + ci + 1;
            const int dB = TL::face(p) + NFLD * N * (N + 1) + (cj + 1) * N + ci;
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                lds[dA + f * stride] = FA[f];
                lds[dB + f * stride] = FB[f];
            }
        }
#pragma unroll
        for (int r = 0; r < TL::XROUNDS; ++r) {
            if (dC[r] < 0) continue;
            // first fluxes in the face arrays (field stride N (N + 1)), second
            // ones in the side arrays (field stride N)
            const int fs = sC[r] ? N : stride;
            lds[dC[r]] = FC[r][0];
            lds[dC[r] + fnC[r] * fs] = FC[r][1];
            lds[dC[r] + (3 - fnC[r]) * fs] = FC[r][2];
            lds[dC[r] + 3 * fs] = FC[r][3];
        }
        __syncthreads();
    } else {
        double Fr[TL::ROUNDS][4];
        int Fdst[TL::ROUNDS];     // LDS destination of the face value
    #pragma unroll
        for (int r = 0; r < TL::ROUNDS; ++r) {
            Fdst[r] = -1;
            int task = tid + r * TL::BLK;
            if (task >= TL::NTASK) continue;
            int tp, isx, fi, fj;   // patch; x-face (fi in 0..N, row fj) or y-face (col fi, fj in 0..N)
            bool edge;
            int side = -1;
            if (task < PPB * TL::NINT) {
                tp = task / TL::NINT;
                int t = task % TL::NINT;
                edge = false;
                if (t < N * (N - 1)) { isx = 1; fi = 1 + t % (N - 1); fj = t / (N - 1); }
                else { t -= N * (N - 1); isx = 0; fi = t % N; fj = 1 + t / N; }
            } else {
                int t = task - PPB * TL::NINT;
                tp = t / TL::NEDGE;
                int e = t % TL::NEDGE;
                edge = true;
                side = e / N;
                int k = e % N;
                switch (side) {
                case 0: isx = 1; fi = 0; fj = k; break;
                case 1: isx = 1; fi = N; fj = k; break;
                case 2: isx = 0; fi = k; fj = 0; break;
                default: isx = 0; fi = k; fj = N; break;
                }
            }
            if (pquad[tp] < 0) continue;
            const double h = ph[tp];
            double F[4];
            // tile positions of the low (L) and high (R) cell of the face
            int tl, tr;
            if (isx) { tl = TL::pos(fi - 1, fj); tr = TL::pos(fi, fj); }
            else     { tl = TL::pos(fi, fj - 1); tr = TL::pos(fi, fj); }
            if (!edge) {
                Prim L, R;
                load_prim<N, PPB>(lds, tp, tl, L);
                load_prim<N, PPB>(lds, tp, tr, R);
                flux_face_dir(L, R, h, isx, F);
            } else {
                int na = pnbr[tp][2 * side], nb = pnbr[tp][2 * side + 1];
                bool lowside = (side == 0 || side == 2);
                int town = lowside ? tr : tl;
                if (na < 0) {
                    // physical boundary: outward normal; the face value is stored
                    // with the +x/+y orientation of the interior fluxes
                    double nrm[2] = {0.0, 0.0};
                    if (side == 0) nrm[0] = -h; else if (side == 1) nrm[0] = h;
                    else if (side == 2) nrm[1] = -h; else nrm[1] = h;
                    const double *b = lds + TL::rec(tp) + town;
                    double q[4] = {b[0], b[TL::NPOS], b[2 * TL::NPOS], b[3 * TL::NPOS]};
                    const double fxo = side == 0 ? 0.0 : (side == 1 ? (double)N : fi + 0.5);
                    const double fyo = side == 2 ? 0.0 : (side == 3 ? (double)N : fj + 0.5);
                    flux_boundary(a.m.ip[pinst[tp]], side, nrm, q, a.m.org + 2 * (size_t)pquad[tp], fxo, fyo, h,
                                  ptime[tp], F);
                    if (lowside) { F[0] = -F[0]; F[1] = -F[1]; F[2] = -F[2]; F[3] = -F[3]; }
                } else if (nb < 0) {
                    Prim L, R;
                    load_prim<N, PPB>(lds, tp, tl, L);
                    load_prim<N, PPB>(lds, tp, tr, R);
                    flux_face_dir(L, R, h, isx, F);
                } else {
                    // coarse side of a hanging face: sub-face k couples to sub-face
                    // k of both fine neighbours, normal length of the fine level
                    // (solver_residual_face_full_hang_structured_2d.F90:86-122)
                    const double hf = 0.5 * h;
                    int k = isx ? fj : fi;
                    Prim own;
                    load_prim<N, PPB>(lds, tp, town, own);
                    F[0] = F[1] = F[2] = F[3] = 0.0;
                    for (int q = 0; q < 2; ++q) {
                        // one fine neighbour at a time keeps a single flux live
                        Prim fn;
                        load_prim<N, PPB>(lds, tp, q == 0 ? (lowside ? tl : tr) : TL::SECOND + side * N + k, fn);
                        double G[4];
                        if (lowside) {
                            flux_face_dir(fn, own, hf, isx, G);
                        } else {
                            flux_face_dir(own, fn, hf, isx, G);
                        }
                        F[0] += G[0]; F[1] += G[1]; F[2] += G[2]; F[3] += G[3];
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
            }
    #pragma unroll
            for (int f = 0; f < 4; ++f) Fr[r][f] = F[f];
            // FX[p][f][fj][fi] (fi in 0..N) ; FY[p][f][fj][fi] (fj in 0..N)
            Fdst[r] = TL::face(tp) + (isx ? (fj * (N + 1) + fi)
                                                      : (NFLD * N * (N + 1) + fj * N + fi));
            // keep the rounds apart: interleaving three flux evaluations buys
            // little ILP and costs the registers that bound occupancy
            __builtin_amdgcn_sched_barrier(0);
        }
        if (actu && needB) {   // base state: loaded late, its latency overlaps the flux gather
#pragma unroll
            for (int f = 0; f < 4; ++f) B[f] = gptr(a.U0)[base + f * N2];
        }
        __syncthreads();   // all reads of the primitive tile are done
    #pragma unroll
        for (int r = 0; r < TL::ROUNDS; ++r) {
            if (Fdst[r] < 0) continue;
            const int stride = N * (N + 1);
    #pragma unroll
            for (int f = 0; f < 4; ++f) lds[Fdst[r] + f * stride] = Fr[r][f];
        }
        __syncthreads();
    }

    // ---- per-cell residual and stage update ------------------------------
    if (actu) {
#pragma unroll
        for (int f = 0; f < 4; ++f) U[f] = own_u(f);
    }
    double k4[4];
    double out[4];
    double cand = INFINITY;
    double sq = 0.0;
    double cflh = 0.0;      // cfl sqrt(h h) of the patch
    if (actu) {
        // -1 / (h h) and cfl sqrt(h h) from the instance's level-0 values
        double ninvV;
        if constexpr (UNI) {
            typedef const InstDev __attribute__((address_space(4))) cinst4;
            const cinst4 *ipc = (const cinst4 *)a.m.ip;
            ninvV = ldexp(ipc[inw].ninv_vol0, 2 * levw);
            cflh = ldexp(ipc[inw].cfl_h0, -levw);
        } else {
            const InstDev &ip = a.m.ip[pinst[p]];
            ninvV = ldexp(ip.ninv_vol0, 2 * plevel[p]);
            cflh = ldexp(ip.cfl_h0, -plevel[p]);
        }
        const double *FX = lds + TL::face(p);
        const double *FY = FX + NFLD * N * (N + 1);
        const int stride = N * (N + 1);
        // second fine neighbours' fluxes on the patch edges
        const double *F2 = lds + TL::f2(p);
        const int n1 = UNI ? nbw[1] : pnbr[p][1], n3 = UNI ? nbw[3] : pnbr[p][3];
        const int n5 = UNI ? nbw[5] : pnbr[p][5], n7 = UNI ? nbw[7] : pnbr[p][7];
        const bool s0 = DESIGN == 1 && ci == 0 && n1 >= 0, s1 = DESIGN == 1 && ci == N - 1 && n3 >= 0;
        const bool s2 = DESIGN == 1 && cj == 0 && n5 >= 0, s3 = DESIGN == 1 && cj == N - 1 && n7 >= 0;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double fxr = FX[f * stride + cj * (N + 1) + ci + 1], fxl = FX[f * stride + cj * (N + 1) + ci];
            double fyt = FY[f * stride + (cj + 1) * N + ci], fyb = FY[f * stride + cj * N + ci];
            if (s1) fxr += F2[(1 * NFLD + f) * N + cj];
            if (s0) fxl += F2[(0 * NFLD + f) * N + cj];
            if (s3) fyt += F2[(3 * NFLD + f) * N + ci];
            if (s2) fyb += F2[(2 * NFLD + f) * N + ci];
            double Rf = (fxr - fxl) + (fyt - fyb);
            k4[f] = Rf * ninvV;
        }
        if (a.norm_part) {        // the residual norm, only on the launches that report it
#pragma unroll
            for (int f = 0; f < 4; ++f) sq = fma(k4[f], k4[f], sq);
        }
        double dt = 0.0;
        if (a.mode != MODE_RHS) dt = UNI ? dtw : pdt[p];
        stage_update(a, base, N2, dt, U, B, k4, out);
        // an instance past its physical stop (dt = 0) keeps its state bit for
        // bit: every stage state equals U^n then, while a convex combination
        // such as 0.75 U + 0.25 U can round away from U
        // (a wave-uniform test when a wave holds one patch: a scalar branch
        // around the rare case instead of selects on every path)
        const bool frozen = a.mode != MODE_RHS && dt == 0.0;
        if (UNI ? (bool)__builtin_amdgcn_readfirstlane((int)frozen) : frozen) {
#pragma unroll
            for (int f = 0; f < 4; ++f) out[f] = U[f];
        }
#pragma unroll
        for (int f = 0; f < 4; ++f)   // streamed out: no other wave of this launch reads it
            __builtin_nontemporal_store(out[f], &a.Uout[base + f * N2]);
        if (a.role & ROLE_CFL) cand = cfl_candidate_h(out, cflh);
    }
    if (a.role & ROLE_CFL) {
        int uinst = SMETA ? inw : pinst[0];
#pragma unroll
        for (int k = 1; k < PPB; ++k)
            if (pquad[k] >= 0 && pinst[k] != uinst) uinst = -1;
        cfl_reduce<TL::BLK>(cand, active ? (SMETA ? inw : pinst[p]) : 0, active, uinst, a.slot, red);
    }
    if (a.norm_part) sum_reduce_store<TL::BLK>(sq, &a.norm_part[blockIdx.x], red);
}



// ---------------------------------------------------------------------------
// Persistent, prefetching stage kernel (DESIGN 2, 16x16 patches): a grid of
// about as many workgroups as the GPU holds at once, each walking a
// contiguous Morton range of patches (per_block of them; the XCD-aware block
// order keeps each XCD's ranges contiguous).  While one patch is evaluated,
// the next patch's own state and halo cells travel into the other half of a
// double-buffered LDS area by direct global-to-LDS loads (no registers), so
// a patch's memory latency is paid behind the arithmetic of the patch before
// it.  Per patch the arithmetic is stage_kernel's paired-face design,
// operation for operation (the results are bitwise those of DESIGN 1).  The
// CFL minimum is kept per wave across patches of one instance and folded
// into the instance's sub-slot once per instance run; the residual sum of
// squares once per workgroup.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dma16(const double *src, double *lds)    // 16 B per lane, lane-contiguous
{
    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)src,
                                     (void __attribute__((address_space(3))) *)lds, 16, 0, 0);
}
__device__ __forceinline__ void dma4(const unsigned *src, double *lds)   // 4 B per lane, lane-contiguous
{
    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)src,
                                     (void __attribute__((address_space(3))) *)lds, 4, 0, 0);
}

// edge cell of the neighbour across side sd at position k (ec) and the own
// edge cell (oc), cell index j * N + i
template <int N>
__device__ __forceinline__ void edge_cells(int sd, int k, int &ec, int &oc)
{
    switch (sd) {
    case 0: ec = k * N + (N - 1); oc = k * N;           break;
    case 1: ec = k * N + 0;       oc = k * N + (N - 1); break;
    case 2: ec = (N - 1) * N + k; oc = k;               break;
    default: ec = 0 * N + k;      oc = (N - 1) * N + k; break;
    }
}

// Issue the loads of patch q into one LDS buffer: its own state (4 N^2
// doubles, contiguous: 1 KB per wave instruction) and its halo, the index-
// aligned edge cells of the face neighbours (own edge cells on a wall; both
// fine neighbours on a coarse side), gathered as 32-bit halves (a lane pair
// per double).  Every wave issues its share; the caller waits for them
// (vmcnt) and synchronises before the buffer is read.
template <int N>
__device__ __forceinline__ void prefetch_patch(const double *Us, int q, const int nb[8], bool hang, double *raw,
                                               double *gh, int tid)
{
    constexpr int N2 = N * N, NG = 4 * N, NW = N2 / 64;
    const int lane = tid & 63, w = tid >> 6;
    const double *own = Us + (size_t)q * NFLD * N2;
    constexpr int CH = NFLD * N2 / 128;                 // 1 KB chunks
#pragma unroll
    for (int j = w; j < CH; j += NW) dma16(own + j * 128 + lane * 2, raw + j * 128);
    constexpr int HCH = NFLD * NG / 32;                 // 32 doubles per instruction
#pragma unroll
    for (int i = w; i < HCH; i += NW) {
        const int d = 32 * i + (lane >> 1);
        const int f = d / NG, hh = d % NG, sd = hh / N, k = hh % N;
        const int na = sd == 0 ? nb[0] : sd == 1 ? nb[2] : sd == 2 ? nb[4] : nb[6];
        int ec, oc;
        edge_cells<N>(sd, k, ec, oc);
        const double *sa = na < 0 ? own + f * N2 + oc : Us + ((size_t)na * NFLD + f) * N2 + ec;
        dma4((const unsigned *)sa + (lane & 1), gh + 32 * i);
        if (hang) {
            const int nbb = sd == 0 ? nb[1] : sd == 1 ? nb[3] : sd == 2 ? nb[5] : nb[7];
            if (nbb >= 0)
                dma4((const unsigned *)(Us + ((size_t)nbb * NFLD + f) * N2 + ec) + (lane & 1),
                     gh + NFLD * NG + 32 * i);
        }
    }
}

template <int N, int MINW>
__global__ __launch_bounds__(N * N, MINW) void stage_kernel_pp(StageArgs a)
{
    using TL = Tile<N, 1>;
    constexpr int N2 = N * N, NG = 4 * N, NW = N2 / 64;
    static_assert(N2 % 64 == 0 && NG <= N2, "one patch of whole waves, one halo cell per thread at most");
    __shared__ double lds[TL::LDS_DOUBLES];
    __shared__ __attribute__((aligned(16))) double raw[2][NFLD * N2];
    __shared__ __attribute__((aligned(16))) double gh[2][2 * NFLD * NG];
    __shared__ double red[NW];
    typedef const int __attribute__((address_space(4))) cint4;
    typedef const InstDev __attribute__((address_space(4))) cinst4;
    typedef const double __attribute__((address_space(4))) cdbl4;

    const int tid = threadIdx.x, c = tid;
    const int ci = c % N, cj = c / N;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int q0 = lb * a.per_block;
    const int q1 = min(q0 + a.per_block, a.m.nquad);
    const bool needB = (a.mode == MODE_SO && a.U0) || a.mode == MODE_RK4_ACC || a.mode == MODE_RK4_FINAL;
    const cint4 *nbrc = (const cint4 *)a.m.nbr;
    const cinst4 *ipc = (const cinst4 *)a.m.ip;

    int nbw[8];
    bool hangw = false;
    if (q0 < q1) {
#pragma unroll
        for (int k = 0; k < 8; ++k) nbw[k] = nbrc[(size_t)q0 * 8 + k];
        hangw = nbw[1] >= 0 || nbw[3] >= 0 || nbw[5] >= 0 || nbw[7] >= 0;
        prefetch_patch<N>(a.Us, q0, nbw, hangw, raw[0], gh[0], tid);
    }
    double sq = 0.0;
    double cmin = INFINITY;          // CFL minimum of the current instance run (per wave)
    int cinst = -1;
    int buf = 0;
    for (int q = q0; q < q1; ++q, buf ^= 1) {
        // this patch's halo and state are in buffer buf once every wave's
        // loads are complete
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const int levw = ((const cint4 *)a.m.level)[q];
        const int inw = ((const cint4 *)a.m.inst)[q];
        double B[4] = {0.0, 0.0, 0.0, 0.0};
        const size_t base = (size_t)q * NFLD * N2 + c;
        if (needB) {
#pragma unroll
            for (int f = 0; f < 4; ++f) B[f] = gptr(a.U0)[base + f * N2];
        }
        // the next patch's neighbour ids (scalar loads, used below)
        int nbn[8];
        const bool more = q + 1 < q1;
        const int qn = more ? q + 1 : q;         // a valid index either way
#pragma unroll
        for (int k = 0; k < 8; ++k) nbn[k] = nbrc[(size_t)qn * 8 + k];
        // records of the own cells and of the halo
        const double *rw = raw[buf];
        const double *gw = gh[buf];
        store_rec<N, 1>(lds, 0, c, rw[c], rw[N2 + c], rw[2 * N2 + c], rw[3 * N2 + c]);
        if (c < NG) {
            const int sd = c / N;
            const int nbb = sd == 0 ? nbw[1] : sd == 1 ? nbw[3] : sd == 2 ? nbw[5] : nbw[7];
            if (nbb >= 0)
                store_rec<N, 1>(lds, 0, TL::SECOND + c, gw[NFLD * NG + c], gw[NFLD * NG + NG + c],
                                gw[NFLD * NG + 2 * NG + c], gw[NFLD * NG + 3 * NG + c]);
            store_rec<N, 1>(lds, 0, TL::GHOST + c, gw[c], gw[NG + c], gw[2 * NG + c], gw[3 * NG + c]);
        }
        // patch parameters behind the instance id
        const double hw = ldexp(ipc[inw].h_base, -levw) / (double)N;
        // the step's dt and start time, written by the step's prep kernel
        const double dtw = a.mode != MODE_RHS ? ((const cdbl4 *)a.dt_cur)[inw] : 0.0;
        const double tw = ((const cdbl4 *)a.m.t_step)[inw];
        // the next patch's loads go out now, into the other buffer (read
        // last by the patch before this one, which every wave has finished)
        if (more) {
            const bool hangn = nbn[1] >= 0 || nbn[3] >= 0 || nbn[5] >= 0 || nbn[7] >= 0;
            prefetch_patch<N>(a.Us, q + 1, nbn, hangn, raw[buf ^ 1], gh[buf ^ 1], tid);
        }
        __syncthreads();

        // ---- face fluxes (stage_kernel DESIGN 1, one patch) ----------------
        double FA[4], FB[4], FC[4] = {0, 0, 0, 0};
        int dC = -1, fnC = 1;
        {
            const double h = hw;
            const int na1 = nbw[2], nb1 = nbw[3], na3 = nbw[6], nb3 = nbw[7];
            const bool ex = (ci == N - 1), ey = (cj == N - 1);
            const double hx = (ex && nb1 >= 0) ? 0.5 * h : h;
            const double hy = (ey && nb3 >= 0) ? 0.5 * h : h;
            Prim L, RX, RY;
            load_prim<N, 1>(lds, 0, c, L);
            load_prim<N, 1>(lds, 0, ex ? TL::GHOST + 1 * N + cj : c + 1, RX);
            load_prim<N, 1>(lds, 0, ey ? TL::GHOST + 3 * N + ci : c + N, RY);
            flux_face<true>(L, RX, hx, FA);
            flux_face<false>(L, RY, hy, FB);
            if (ex && (na1 < 0 || nb1 >= 0)) {
                if (na1 < 0) {
                    const double nrm[2] = {h, 0.0};
                    const double qq[4] = {L.r, L.ru, L.rv, L.rE};
                    flux_boundary(a.m.ip[inw], 1, nrm, qq, a.m.org + 2 * (size_t)q, (double)N, cj + 0.5, h, tw, FA);
                } else {
                    Prim F2;
                    load_prim<N, 1>(lds, 0, TL::SECOND + 1 * N + cj, F2);
                    double G[4];
                    flux_face<true>(L, F2, hx, G);
                    FA[0] += G[0]; FA[1] += G[1]; FA[2] += G[2]; FA[3] += G[3];
                }
            }
            if (ey && (na3 < 0 || nb3 >= 0)) {
                if (na3 < 0) {
                    const double nrm[2] = {0.0, h};
                    const double qq[4] = {L.r, L.ru, L.rv, L.rE};
                    flux_boundary(a.m.ip[inw], 3, nrm, qq, a.m.org + 2 * (size_t)q, ci + 0.5, (double)N, h, tw, FB);
                } else {
                    Prim F2;
                    load_prim<N, 1>(lds, 0, TL::SECOND + 3 * N + ci, F2);
                    double G[4];
                    flux_face<false>(L, F2, hy, G);
                    FB[0] += G[0]; FB[1] += G[1]; FB[2] += G[2]; FB[3] += G[3];
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        {
            // the patch-edge round on one wave other than the halo's wave 0,
            // rotating with the patch
            const int et = tid - (2 * N <= 64 ? (1 + (lb + q) % (NW - 1)) * 64 : 0);
            if (et >= 0 && et < 2 * N) {
                const double h = hw;
                const bool xs = et < N;
                const int k = xs ? et : et - N, side = xs ? 0 : 2;
                const int na = xs ? nbw[0] : nbw[4], nb = xs ? nbw[1] : nbw[5];
                const double hf = nb >= 0 ? 0.5 * h : h;
                const int tl = TL::GHOST + side * N + k;
                const int tr = xs ? k * N : k;
                const int offn = xs ? TL::NPOS : 2 * TL::NPOS, offt = 3 * TL::NPOS - offn;
                OPrim L, R;
                load_oprim<N, 1>(lds, 0, tl, offn, offt, L);
                load_oprim<N, 1>(lds, 0, tr, offn, offt, R);
                flux_normal(L, R, hf, FC);
                if (na < 0) {
                    const double nrm[2] = {xs ? -h : 0.0, xs ? 0.0 : -h};
                    const double qq[4] = {R.r, xs ? R.rn : R.rt, xs ? R.rt : R.rn, R.rE};
                    double F[4];
                    flux_boundary(a.m.ip[inw], side, nrm, qq, a.m.org + 2 * (size_t)q, xs ? 0.0 : k + 0.5,
                                  xs ? k + 0.5 : 0.0, h, tw, F);
                    FC[0] = -F[0]; FC[1] = -(xs ? F[1] : F[2]); FC[2] = -(xs ? F[2] : F[1]); FC[3] = -F[3];
                } else if (nb >= 0) {
                    OPrim F2;
                    load_oprim<N, 1>(lds, 0, TL::SECOND + side * N + k, offn, offt, F2);
                    double G[4];
                    flux_normal(F2, R, hf, G);
                    FC[0] += G[0]; FC[1] += G[1]; FC[2] += G[2]; FC[3] += G[3];
                }
                dC = xs ? k * (N + 1) : NFLD * N * (N + 1) + k;
                fnC = xs ? 1 : 2;
            }
        }
        __syncthreads();   // all reads of the record tile are done
        const int stride = N * (N + 1);
        {
            const int dA = cj * (N + 1) + ci + 1;
            const int dB = NFLD * N * (N + 1) + (cj + 1) * N + ci;
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                lds[dA + f * stride] = FA[f];
                lds[dB + f * stride] = FB[f];
            }
        }
        if (dC >= 0) {
            lds[dC] = FC[0];
            lds[dC + fnC * stride] = FC[1];
            lds[dC + (3 - fnC) * stride] = FC[2];
            lds[dC + 3 * stride] = FC[3];
        }
        __syncthreads();

        // ---- per-cell residual and stage update ---------------------------
        double U[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) U[f] = rw[f * N2 + c];
        double k4[4], out[4];
        const double ninvV = ldexp(ipc[inw].ninv_vol0, 2 * levw);
        const double cflh = ldexp(ipc[inw].cfl_h0, -levw);
        {
            const double *FX = lds;
            const double *FY = FX + NFLD * N * (N + 1);
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                double Rf = (FX[f * stride + cj * (N + 1) + ci + 1] - FX[f * stride + cj * (N + 1) + ci]) +
                            (FY[f * stride + (cj + 1) * N + ci] - FY[f * stride + cj * N + ci]);
                k4[f] = Rf * ninvV;
                sq = fma(k4[f], k4[f], sq);
            }
        }
        const double dt = a.mode != MODE_RHS ? dtw : 0.0;
        stage_update(a, base, N2, dt, U, B, k4, out);
        if (a.mode != MODE_RHS && dt == 0.0) {
#pragma unroll
            for (int f = 0; f < 4; ++f) out[f] = U[f];
        }
#pragma unroll
        for (int f = 0; f < 4; ++f) __builtin_nontemporal_store(out[f], &a.Uout[base + f * N2]);
        if (a.role & ROLE_CFL) {
            // the running minimum of this wave, folded into the instance's
            // sub-slot when the instance changes (uniform across the block)
            if (inw != cinst) {
                if (cinst >= 0) cfl_reduce<N2>(cmin, cinst, true, cinst, a.slot, red);
                cinst = inw;
                cmin = INFINITY;
            }
            cmin = fmin(cmin, wave_min(cfl_candidate_h(out, cflh)));
        }
        // the next patch's ids
#pragma unroll
        for (int k = 0; k < 8; ++k) nbw[k] = nbn[k];
    }
    if ((a.role & ROLE_CFL) && cinst >= 0) cfl_reduce<N2>(cmin, cinst, true, cinst, a.slot, red);
    if (a.norm_part) sum_reduce_store<N2>(sq, &a.norm_part[blockIdx.x], red);
}

// Standalone CFL minimum (physics_time_step): one thread per cell.
template <int N, int PPB>
__global__ __launch_bounds__((Tile<N, PPB>::BLK)) void cfl_kernel(MeshDev m, const double *U,
                                                                unsigned long long *slot)
{
    using TL = Tile<N, PPB>;
    constexpr int N2 = TL::N2;
    __shared__ double red[(TL::BLK + 63) / 64];
    __shared__ int pinst[PPB];
    const int tid = threadIdx.x;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int p = tid / N2, c = tid % N2;
    const int quad = lb * PPB + p;
    // a quadrant whose instance id is outside [0, ninst) is skipped (ids
    // from a caller's device array cannot be checked on the host)
    const bool active = quad < m.nquad && (unsigned)m.inst[quad] < (unsigned)m.ninst;
    if (tid < PPB) {
        int qq = lb * PPB + tid;
        const int in = qq < m.nquad ? m.inst[qq] : -1;
        pinst[tid] = (unsigned)in < (unsigned)m.ninst ? in : -1;
    }
    __syncthreads();
    int uinst = pinst[0];
#pragma unroll
    for (int k = 1; k < PPB; ++k)
        if (pinst[k] >= 0 && pinst[k] != uinst) uinst = -1;
    double cand = INFINITY;
    int inst = 0;
    if (active) {
        inst = m.inst[quad];
        const double h = ldexp(m.ip[inst].h_base, -m.level[quad]) / (double)N;
        const double *src = U + (size_t)quad * NFLD * N2 + c;
        double q[4] = {src[0], src[N2], src[2 * N2], src[3 * N2]};
        cand = cfl_candidate(q, m.ip[inst].cfl, h * h);
    }
    cfl_reduce<TL::BLK>(cand, inst, active, uinst, slot, red);
}

__global__ void slot_reset_kernel(unsigned long long *slot, int ninst)
{
    for (int i = threadIdx.x; i < ninst * NSLOT; i += blockDim.x) slot[i] = DBL_INF_BITS;
}

// The step's dt and time update, ahead of its first stage: one wave per
// instance forms min(input dt, CFL minimum over the instance's sub-slots,
// time left), publishes it with the step's start time, advances the time and
// re-arms the sub-slots that the step's last stage fills.  The stages then
// read the published dt instead of every wave reducing the sub-slots.
__global__ __launch_bounds__(64) void dt_prep_kernel(MeshDev m, int use_cfl, double *dt_cur, double *t_inst,
                                                     double *t_step, unsigned long long *slot)
{
    const int i = blockIdx.x;
    if (i >= m.ninst) return;
    const double dt = step_dt(m.ip[i], slot, t_inst, i, use_cfl);
    if (threadIdx.x == 0) {
        dt_cur[i] = dt;
        t_step[i] = t_inst[i];
        t_inst[i] += dt;
    }
    for (int k = threadIdx.x; k < NSLOT; k += 64) slot[(size_t)i * NSLOT + k] = DBL_INF_BITS;
}

// reference patch layout Q(4,n,n) (AoS) <-> U[quad][field][cell] (SoA)
__global__ void aos_to_soa_kernel(const double *aos, double *soa, long long ncells, int n2)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells) return;
    long long q = t / n2, c = t % n2;
    const double *s = aos + t * NFLD;
#pragma unroll
    for (int f = 0; f < 4; ++f) soa[(q * NFLD + f) * n2 + c] = s[f];
}
__global__ void soa_to_aos_kernel(const double *soa, double *aos, long long ncells, int n2)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells) return;
    long long q = t / n2, c = t % n2;
    double *d = aos + t * NFLD;
#pragma unroll
    for (int f = 0; f < 4; ++f) d[f] = soa[(q * NFLD + f) * n2 + c];
}

// Project-defined density-gradient indicator, one thread per sub-cell: the
// largest |rho_a - rho_b| / (0.5 (rho_a + rho_b)) over the cell's +x and +y
// neighbours in its patch and, on patch edges, over the index-aligned edge
// cells of the face neighbours (both fine ones on a coarse side); reduced to
// one value per quadrant.  Max is order-free, so the result equals the host
// indicator bit for bit.
__device__ __forceinline__ double rel_jump(double a, double b) { return fabs(a - b) / (0.5 * (a + b)); }

template <int N, int PPB>
__global__ __launch_bounds__((Tile<N, PPB>::BLK)) void indicator_kernel(MeshDev m, const double *U,
                                                                        double *ind)
{
    constexpr int N2 = N * N;
    __shared__ unsigned long long pmax[PPB];
    const int tid = threadIdx.x;
    const int p = tid / N2, c = tid % N2, ci = c % N, cj = c / N;
    const int quad = blockIdx.x * PPB + p;
    const bool active = quad < m.nquad;
    if (tid < PPB) pmax[tid] = 0ull;
    __syncthreads();
    if (active) {
        const double *Q = U + (size_t)quad * NFLD * N2;      // field 0 plane
        const double r = Q[c];
        double v = 0.0;
        if (ci + 1 < N) v = fmax(v, rel_jump(r, Q[c + 1]));
        if (cj + 1 < N) v = fmax(v, rel_jump(r, Q[c + N]));
        const bool on[4] = {ci == 0, ci == N - 1, cj == 0, cj == N - 1};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (!on[s]) continue;
            const int k = (s < 2) ? cj : ci;
            int ec;
            switch (s) {
            case 0: ec = k * N + (N - 1); break;
            case 1: ec = k * N; break;
            case 2: ec = (N - 1) * N + k; break;
            default: ec = k; break;
            }
            const int na = m.nbr[quad * 8 + 2 * s], nb = m.nbr[quad * 8 + 2 * s + 1];
            if (na >= 0) v = fmax(v, rel_jump(r, U[(size_t)na * NFLD * N2 + ec]));
            if (nb >= 0) v = fmax(v, rel_jump(r, U[(size_t)nb * NFLD * N2 + ec]));
        }
        atomicMax(&pmax[p], d2bits(v));    // non-negative doubles order as integers
    }
    __syncthreads();
    if (tid < PPB && (int)blockIdx.x * PPB + tid < m.nquad) ind[blockIdx.x * PPB + tid] = bits2d(pmax[tid]);
}

// The reference's feature tags (solver_tag_feature.F90:9-52 with
// tag_density_below / tag_density_above, tagging_methods_module.F90:14-71):
// a quadrant is tagged when, for some method k, one of its sub-cell
// densities is <= tol[k] (method 1) or >= tol[k] (method 2).  "Some density
// <= tol" is "the patch minimum <= tol" (fmin skips NaN, and a NaN density
// satisfies neither comparison), so one min / max reduction per patch gives
// the reference's flag exactly.  Methods 3 and 4 (vorticity, Q criterion) are
// commented out in the reference and tag nothing, as any other value.
struct FeatureTagArgs {
    int ntags;
    int method[AMRGPU_MAX_FEATURE_TAGS];
    double tol[AMRGPU_MAX_FEATURE_TAGS];
};

template <int N>
__global__ __launch_bounds__(256) void feature_tag_kernel(int nquad, const double *U, FeatureTagArgs t,
                                                          int *tags)
{
    constexpr int N2 = N * N;
    constexpr int PPB = N2 >= 256 ? 1 : 256 / N2;
    constexpr int W = N2 < 64 ? N2 : 64;            // lanes of one patch within a wave
    __shared__ double smin[4], smax[4];
    const int tid = threadIdx.x;
    const int p = tid / N2, c = tid % N2;
    const int quad = blockIdx.x * PPB + p;
    const bool active = p < PPB && quad < nquad;
    const double r = active ? U[(size_t)quad * NFLD * N2 + c] : NAN;
    double mn = r, mx = r;
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) {           // segments of W aligned lanes
        mn = fmin(mn, __shfl_xor(mn, o, 64));
        mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    if constexpr (N2 > 64) {
        static_assert(N2 == 256, "one patch of four waves");
        if ((tid & 63) == 0) { smin[tid >> 6] = mn; smax[tid >> 6] = mx; }
        __syncthreads();
        mn = fmin(fmin(smin[0], smin[1]), fmin(smin[2], smin[3]));
        mx = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
    }
    if (active && c == 0) {
        int tag = 0;
        for (int k = 0; k < t.ntags; ++k) {
            if (t.method[k] == 1 && mn <= t.tol[k]) tag = 1;
            if (t.method[k] == 2 && mx >= t.tol[k]) tag = 1;
        }
        tags[quad] = tag;
    }
}

// New state after a regrid, one thread per new sub-cell: src >= 0 copies
// old quadrant src (solver_refine_operator_2d.F90:49-52 copies the parent);
// src < 0 averages the old family -src-1.. (solver_coarsen_operator_2d.F90:
// 49-50: sum of the four children, then times 0.25).
__global__ void remap_kernel(const double *Uold, double *Unew, const int *src, long long ncells, int n2)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncells) return;
    const long long q = t / n2, c = t % n2;
    const int s = src[q];
    double *d = Unew + (size_t)q * NFLD * n2 + c;
    if (s >= 0) {
        const double *o = Uold + (size_t)s * NFLD * n2 + c;
#pragma unroll
        for (int f = 0; f < 4; ++f) d[f * n2] = o[f * n2];
    } else {
        const long long k = -(long long)s - 1;
        const double *o = Uold + (size_t)k * NFLD * n2 + c;
        const size_t ps = (size_t)NFLD * n2;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double sum = o[f * n2] + o[ps + f * n2];
            sum = sum + o[2 * ps + f * n2];
            sum = sum + o[3 * ps + f * n2];
            d[f * n2] = 0.25 * sum;
        }
    }
}

__global__ void copy_kernel(const double4 *__restrict__ src, double4 *__restrict__ dst, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
thread_local char g_err[512];

void set_err(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

#define HIPCHK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            set_err("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            return AMRGPU_ERR_HIP;                                                  \
        }                                                                           \
    } while (0)

struct StageLaunch {
    StageArgs a;
};

}  // namespace

extern "C" void amr_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

struct amrgpu_ctx {
    int n = 0, nquad = 0, ninst = 0, scheme = 3;
    hipStream_t stream = nullptr;
    int *d_nbr = nullptr, *d_level = nullptr, *d_inst = nullptr;
    InstDev *d_ip = nullptr;
    double *U[3] = {nullptr, nullptr, nullptr};
    double *Acc = nullptr;
    double *d_dt_cur = nullptr, *d_t = nullptr;
    unsigned long long *d_slot = nullptr;
    double *d_norm = nullptr;
    double *d_stage = nullptr;      // AoS staging for up/download
    int *d_map = nullptr;           // regrid source map
    double *d_ind = nullptr;        // per-quadrant indicator
    double *d_org = nullptr;        // per-quadrant lower-left corner
    double *d_tstep = nullptr;      // per-instance time at the step start
    bool org_ok = false, need_org = false;
    int nblocks = 0;
    int variant = -1;
    int cap_quad = 0, cap_inst = 0, cap_blocks = 0;
    size_t cap_state = 0;
    int cur = 0;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    // Step launch mode: direct stream launches by default (measured on
    // MI355X, ROCm 7: back-to-back dependent kernels cost ~2.3 us each from a
    // stream and ~4.7-6.7 us each replayed from a captured graph); graph
    // replay on request (AMRGPU_GRAPHS=1, amrgpu_set_graphs) or when the
    // autotuner measures it faster.
    bool graphs_enabled = false;
    std::vector<InstDev> h_ip;
    int *h_pin = nullptr;      // pinned staging of the mesh uploads (ints)
    size_t h_pin_cap = 0;
    int cu_count = 256;        // compute units of the context's device
};

namespace {

MeshDev mesh_of(const amrgpu_ctx *c)
{
    MeshDev m;
    m.nquad = c->nquad; m.ninst = c->ninst;
    m.nbr = c->d_nbr; m.level = c->d_level; m.inst = c->d_inst; m.ip = c->d_ip;
    m.org = c->d_org; m.t_step = c->d_tstep;
    return m;
}

// Tile variants of the stage kernel: patches per workgroup (PPB) and the
// minimum waves per SIMD the register allocator must leave room for (MINW).
// The autotuner times every variant of the mesh's patch size.
struct TileVariant {
    int n, ppb, minw, design;
};
const TileVariant kVariants[] = {
    {2, 16, 1, 0}, {2, 16, 1, 1},
    {4, 8, 1, 0}, {4, 8, 1, 1},
    {8, 2, 1, 0}, {8, 4, 1, 0}, {8, 4, 4, 0}, {8, 2, 1, 1}, {8, 4, 1, 1}, {8, 4, 4, 1},
    {16, 1, 1, 0}, {16, 1, 4, 0}, {16, 1, 1, 1}, {16, 1, 4, 1},
    {8, 1, 1, 1}, {8, 1, 4, 1}, {8, 2, 4, 1}, {8, 1, 5, 1},
    {16, 1, 3, 2},
    {16, 1, 5, 1},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

// Persistent design: the workgroups resident at once (LDS bounds a 16x16
// workgroup to 3 per CU, registers to MINW), each given a contiguous range
// of patches.
constexpr int PERSIST_WG_PER_CU = 3;

void persistent_grid(const amrgpu_ctx *c, int nquad, int minw, int *nblk, int *per)
{
    const int gmax = std::max(8, c->cu_count * std::min(minw, PERSIST_WG_PER_CU));
    const int p = std::max(1, (nquad + gmax - 1) / gmax);
    *per = p;
    *nblk = (nquad + p - 1) / p;
}

// Workgroups of one stage launch for tile variant v (the residual-norm
// partials are one per workgroup).
int grid_of(const amrgpu_ctx *c, int v, int nquad)
{
    if (kVariants[v].design == 2) {
        int nblk, per;
        persistent_grid(c, nquad, kVariants[v].minw, &nblk, &per);
        return nblk;
    }
    return (nquad + kVariants[v].ppb - 1) / kVariants[v].ppb;
}


// Stage-kernel tile used until amrgpu_set_tile / amrgpu_autotune picks one
// (every tile computes the same bits): for n = 8 and 16 the paired-faces
// tile with one patch per workgroup, four and five waves per SIMD, the
// fastest or within 2 % of it on the meshes measured (207k to 16.8M cells).
int default_variant(int n)
{
    switch (n) {
    case 2: return 0;
    case 4: return 2;
    case 8: return 15;
    case 16: return 19;
    default: return -1;
    }
}

// Patches per workgroup of the auxiliary per-patch kernels (CFL, indicator),
// fixed by their template instantiations below; 0 for an unsupported n.
int ppb_of(int n)
{
    switch (n) {
    case 2: return 16;
    case 4: return 8;
    case 8: return 4;
    case 16: return 1;
    default: return 0;
    }
}

template <int N, int MINW>
int launch_pp(amrgpu_ctx *c, const StageArgs &a0)
{
    StageArgs a = a0;
    int nblk, per;
    persistent_grid(c, c->nquad, MINW, &nblk, &per);
    a.per_block = per;
    hipLaunchKernelGGL((stage_kernel_pp<N, MINW>), dim3(nblk), dim3(N * N), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return AMRGPU_OK;
}

template <int N, int PPB, int MINW, int D>
int launch_v(amrgpu_ctx *c, const StageArgs &a)
{
    int nblk = (c->nquad + PPB - 1) / PPB;
    hipLaunchKernelGGL((stage_kernel<N, PPB, MINW, D>), dim3(nblk), dim3(Tile<N, PPB>::BLK), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return AMRGPU_OK;
}

int launch_stage(amrgpu_ctx *c, const StageArgs &a)
{
    switch (c->variant) {
    case 0: return launch_v<2, 16, 1, 0>(c, a);
    case 1: return launch_v<2, 16, 1, 1>(c, a);
    case 2: return launch_v<4, 8, 1, 0>(c, a);
    case 3: return launch_v<4, 8, 1, 1>(c, a);
    case 4: return launch_v<8, 2, 1, 0>(c, a);
    case 5: return launch_v<8, 4, 1, 0>(c, a);
    case 6: return launch_v<8, 4, 4, 0>(c, a);
    case 7: return launch_v<8, 2, 1, 1>(c, a);
    case 8: return launch_v<8, 4, 1, 1>(c, a);
    case 9: return launch_v<8, 4, 4, 1>(c, a);
    case 10: return launch_v<16, 1, 1, 0>(c, a);
    case 11: return launch_v<16, 1, 4, 0>(c, a);
    case 12: return launch_v<16, 1, 1, 1>(c, a);
    case 13: return launch_v<16, 1, 4, 1>(c, a);
    case 14: return launch_v<8, 1, 1, 1>(c, a);
    case 15: return launch_v<8, 1, 4, 1>(c, a);
    case 16: return launch_v<8, 2, 4, 1>(c, a);
    case 17: return launch_v<8, 1, 5, 1>(c, a);
    case 18: return launch_pp<16, 3>(c, a);
    case 19: return launch_v<16, 1, 5, 1>(c, a);
    default: set_err("bad tile variant %d", c->variant); return AMRGPU_ERR_ARG;
    }
}

int launch_cfl(amrgpu_ctx *c, const double *U)
{
    int ppb = ppb_of(c->n);
    int nblk = (c->nquad + ppb - 1) / ppb;
    MeshDev m = mesh_of(c);
    switch (c->n) {
    case 2: hipLaunchKernelGGL((cfl_kernel<2, 16>), dim3(nblk), dim3(64), 0, c->stream, m, U, c->d_slot); break;
    case 4: hipLaunchKernelGGL((cfl_kernel<4, 8>), dim3(nblk), dim3(128), 0, c->stream, m, U, c->d_slot); break;
    case 8: hipLaunchKernelGGL((cfl_kernel<8, 4>), dim3(nblk), dim3(256), 0, c->stream, m, U, c->d_slot); break;
    case 16: hipLaunchKernelGGL((cfl_kernel<16, 1>), dim3(nblk), dim3(256), 0, c->stream, m, U, c->d_slot); break;
    default: set_err("unsupported nelem_subgrid %d", c->n); return AMRGPU_ERR_ARG;
    }
    HIPCHK(hipGetLastError());
    return AMRGPU_OK;
}

int launch_indicator(amrgpu_ctx *c, double *d_ind)
{
    int ppb = ppb_of(c->n);
    int nblk = (c->nquad + ppb - 1) / ppb;
    MeshDev m = mesh_of(c);
    const double *U = c->U[c->cur];
    switch (c->n) {
    case 2: hipLaunchKernelGGL((indicator_kernel<2, 16>), dim3(nblk), dim3(64), 0, c->stream, m, U, d_ind); break;
    case 4: hipLaunchKernelGGL((indicator_kernel<4, 8>), dim3(nblk), dim3(128), 0, c->stream, m, U, d_ind); break;
    case 8: hipLaunchKernelGGL((indicator_kernel<8, 4>), dim3(nblk), dim3(256), 0, c->stream, m, U, d_ind); break;
    case 16: hipLaunchKernelGGL((indicator_kernel<16, 1>), dim3(nblk), dim3(256), 0, c->stream, m, U, d_ind); break;
    default: set_err("unsupported nelem_subgrid %d", c->n); return AMRGPU_ERR_ARG;
    }
    HIPCHK(hipGetLastError());
    return AMRGPU_OK;
}

StageArgs base_args(amrgpu_ctx *c)
{
    StageArgs a;
    memset(&a, 0, sizeof(a));
    a.m = mesh_of(c);
    a.dt_cur = c->d_dt_cur;
    a.slot = c->d_slot;
    return a;
}

// Stage launches of one step starting from state buffer `cur`; returns the
// buffer holding the new state.  Shu-Osher forms need the stage state and
// the step's base state only (2 reads + 1 write per cell-stage).
int build_step(amrgpu_ctx *c, int cur, std::vector<StageArgs> &st, bool with_norm)
{
    st.clear();
    double *A = c->U[cur], *B = c->U[(cur + 1) % 3], *C = c->U[(cur + 2) % 3];
    StageArgs a = base_args(c);
    int cfl = (c->scheme != 2) ? ROLE_CFL : 0;
    switch (c->scheme) {
    case 1:
        a.Us = A; a.U0 = nullptr; a.Uout = B; a.mode = MODE_SO; a.c0 = 0; a.c1 = 1; a.c2 = 1;
        a.role = cfl;
        if (with_norm) a.norm_part = c->d_norm;
        st.push_back(a);
        return (cur + 1) % 3;
    case 2:
        a.Us = A; a.U0 = nullptr; a.Uout = B; a.mode = MODE_SO; a.c0 = 0; a.c1 = 1; a.c2 = 1;
        a.role = 0; st.push_back(a);
        a.Us = B; a.U0 = A; a.Uout = A; a.c0 = 0.5; a.c1 = 0.5; a.c2 = 0.5;
        a.role = 0; if (with_norm) a.norm_part = c->d_norm; st.push_back(a);
        return cur;
    case 3:
        a.Us = A; a.U0 = nullptr; a.Uout = B; a.mode = MODE_SO; a.c0 = 0; a.c1 = 1; a.c2 = 1;
        a.role = 0; st.push_back(a);
        a.Us = B; a.U0 = A; a.Uout = C; a.c0 = 0.75; a.c1 = 0.25; a.c2 = 0.25;
        a.role = 0; st.push_back(a);
        a.Us = C; a.U0 = A; a.Uout = A; a.c0 = 1.0 / 3.0; a.c1 = 2.0 / 3.0; a.c2 = 2.0 / 3.0;
        a.role = cfl; if (with_norm) a.norm_part = c->d_norm; st.push_back(a);
        return cur;
    case 4:
        a.Acc = c->Acc;
        a.Us = A; a.U0 = A; a.Uout = B; a.mode = MODE_RK4_ACC; a.first = 1; a.c2 = 0.5;
        a.role = 0; st.push_back(a);
        a.Us = B; a.Uout = C; a.first = 0; a.c2 = 0.5; a.role = 0; st.push_back(a);
        a.Us = C; a.Uout = B; a.c2 = 1.0; a.role = 0; st.push_back(a);
        a.Us = B; a.Uout = A; a.mode = MODE_RK4_FINAL; a.role = cfl;
        if (with_norm) a.norm_part = c->d_norm; st.push_back(a);
        return cur;
    case 5: {
        static const double ra[5] = {0.0,
            -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
            -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
        static const double rb[5] = {1432997174477.0 / 9575080441755.0,
            5161836677717.0 / 13612068292357.0, 1720146321549.0 / 2090206949498.0,
            3134564353537.0 / 4481467310338.0, 2277821191437.0 / 14882151754819.0};
        a.Acc = c->Acc;
        a.mode = MODE_LSERK;
        double *X = A, *Y = B;
        for (int s = 0; s < 5; ++s) {
            a.Us = X; a.U0 = nullptr; a.Uout = Y; a.first = (s == 0); a.c1 = ra[s]; a.c2 = rb[s];
            a.role = s == 4 ? cfl : 0;
            a.norm_part = (s == 4 && with_norm) ? c->d_norm : nullptr;
            st.push_back(a);
            double *tmp = X; X = Y; Y = tmp;
        }
        return (cur + 1) % 3;   // five ping-pong stages end in B
    }
    default:
        return -1;
    }
}

int enqueue_step(amrgpu_ctx *c, int cur, bool with_norm, int *next)
{
    std::vector<StageArgs> st;
    int nx = build_step(c, cur, st, with_norm);
    if (nx < 0) { set_err("bad time scheme %d", c->scheme); return AMRGPU_ERR_ARG; }
    hipLaunchKernelGGL(dt_prep_kernel, dim3(std::max(1, c->ninst)), dim3(64), 0, c->stream, mesh_of(c),
                       c->scheme != 2, c->d_dt_cur, c->d_t, c->d_tstep, c->d_slot);
    HIPCHK(hipGetLastError());
    for (auto &a : st) {
        int rc = launch_stage(c, a);
        if (rc) return rc;
    }
    *next = nx;
    return AMRGPU_OK;
}

void free_graphs(amrgpu_ctx *c)
{
    for (int i = 0; i < 3; ++i)
        if (c->graph[i]) { (void)hipGraphExecDestroy(c->graph[i]); c->graph[i] = nullptr; }
}

void free_mesh(amrgpu_ctx *c)
{
    free_graphs(c);
    (void)hipFree(c->d_nbr); (void)hipFree(c->d_level); (void)hipFree(c->d_inst); (void)hipFree(c->d_ip);
    for (int i = 0; i < 3; ++i) (void)hipFree(c->U[i]);
    (void)hipFree(c->Acc); (void)hipFree(c->d_dt_cur); (void)hipFree(c->d_t); (void)hipFree(c->d_slot);
    (void)hipFree(c->d_norm); (void)hipFree(c->d_stage);
    (void)hipFree(c->d_map); (void)hipFree(c->d_ind); (void)hipFree(c->d_org); (void)hipFree(c->d_tstep);
    c->d_map = nullptr; c->d_ind = nullptr; c->d_org = nullptr; c->d_tstep = nullptr;
    c->org_ok = false;
    c->d_nbr = c->d_level = c->d_inst = nullptr; c->d_ip = nullptr;
    c->U[0] = c->U[1] = c->U[2] = nullptr; c->Acc = nullptr;
    c->d_dt_cur = c->d_t = nullptr; c->d_slot = nullptr; c->d_norm = nullptr;
    c->d_stage = nullptr;
    c->nquad = 0;
    c->cap_quad = c->cap_inst = c->cap_blocks = 0;
    c->cap_state = 0;
}

// face_info (reference format) -> per-quadrant face table
int build_nbr(int ncell, int nface, const int *face_info, std::vector<int> &nbr)
{
    nbr.assign((size_t)ncell * 8, -2);
    auto setside = [&](int q, int s, int a, int b) -> int {
        if (q < 0 || q >= ncell || s < 0 || s > 3) return -1;
        int *e = &nbr[(size_t)q * 8 + 2 * s];
        if (e[0] != -2) return -1;   // side already described
        e[0] = a; e[1] = b;
        return 0;
    };
    for (int f = 0; f < nface; ++f) {
        const int *fi = &face_info[11 * f];
        int type = fi[0], side = fi[1], cell = fi[2];
        int rc = 0;
        if (type == 1) {
            rc |= setside(cell, side, -1, -1);
        } else if (type == 2) {
            int o = fi[3];
            if (o < 0 || o >= ncell) { set_err("face %d: bad neighbour %d", f, o); return AMRGPU_ERR_MESH; }
            rc |= setside(cell, side, o, -1);
            if (o != cell || (side ^ 1) != side) rc |= setside(o, side ^ 1, cell, -1);
        } else if (type == 3) {
            int f1 = fi[3], f2 = fi[4];
            if (f1 < 0 || f1 >= ncell || f2 < 0 || f2 >= ncell) {
                set_err("face %d: bad hanging neighbours %d %d", f, f1, f2);
                return AMRGPU_ERR_MESH;
            }
            rc |= setside(cell, side, f1, f2);
            rc |= setside(f1, side ^ 1, cell, -1);
            rc |= setside(f2, side ^ 1, cell, -1);
        } else {
            set_err("face %d: unsupported face type %d", f, type);
            return AMRGPU_ERR_MESH;
        }
        if (rc) { set_err("face %d: inconsistent face table", f); return AMRGPU_ERR_MESH; }
    }
    for (int q = 0; q < ncell; ++q)
        for (int s = 0; s < 4; ++s)
            if (nbr[(size_t)q * 8 + 2 * s] == -2) {
                set_err("quadrant %d side %d has no face", q, s);
                return AMRGPU_ERR_MESH;
            }
    return AMRGPU_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *amrgpu_last_error(void) { return g_err; }

int amrgpu_current_device(void)
{
    int dev = 0;
    return hipGetDevice(&dev) == hipSuccess ? dev : 0;
}

void amrgpu_use_device(int dev)
{
    (void)hipSetDevice(dev);
}

amrgpu_ctx *amrgpu_create(void)
{
    amrgpu_ctx *c = new amrgpu_ctx();
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
        c->cu_count = ncu;
    const char *ge = getenv("AMRGPU_GRAPHS");
    if (ge) c->graphs_enabled = ge[0] == '1';
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_err("hipStreamCreate failed");
        delete c;
        return nullptr;
    }
    return c;
}

void amrgpu_destroy(amrgpu_ctx *c)
{
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    free_mesh(c);
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int amrgpu_staging(amrgpu_ctx *c, double **buf, size_t *ndoubles)
{
    if (!c || !c->d_stage) { set_err("amrgpu_staging: no mesh"); return AMRGPU_ERR_ARG; }
    *buf = c->d_stage;
    *ndoubles = c->cap_state;
    return AMRGPU_OK;
}

int amrgpu_num_cells(const amrgpu_ctx *c) { return c ? c->nquad * c->n * c->n : 0; }
int amrgpu_num_quads(const amrgpu_ctx *c) { return c ? c->nquad : 0; }

int amrgpu_device_state(amrgpu_ctx *c, const double **U, const int **nbr, int *nquad, int *n,
                        void **stream)
{
    if (!c || !c->nquad) { set_err("amrgpu_device_state: no mesh"); return AMRGPU_ERR_ARG; }
    if (U) *U = c->U[c->cur];
    if (nbr) *nbr = c->d_nbr;
    if (nquad) *nquad = c->nquad;
    if (n) *n = c->n;
    if (stream) *stream = (void *)c->stream;
    return AMRGPU_OK;
}

static int validate_cells(int n2, int ninst, int ncell, const int *cell_info, const int *cell_inst,
                          const char *who)
{
    for (int q = 0; q < ncell; ++q) {
        if (cell_info[3 * q + 2] != 4 * n2 * q) {
            set_err("%s: cell %d soln offset %d is not patch-contiguous", who, q, cell_info[3 * q + 2]);
            return AMRGPU_ERR_MESH;
        }
        int lv = cell_info[3 * q + 1];
        if (lv < 0 || lv > 29) { set_err("%s: cell %d level %d", who, q, lv); return AMRGPU_ERR_MESH; }
        if (cell_inst && (cell_inst[q] < 0 || cell_inst[q] >= ninst)) {
            set_err("%s: cell %d instance %d", who, q, cell_inst[q]);
            return AMRGPU_ERR_ARG;
        }
    }
    return AMRGPU_OK;
}

// Buffers are kept across setups and remeshes (an adaptive run changes the
// mesh once per regrid interval) and grown with 25% headroom when the mesh
// outgrows them.  *fresh tells whether they were reallocated.
static int ensure_capacity(amrgpu_ctx *c, int n, int ncell, int ninst, bool need_acc, bool *fresh)
{
    const int n2 = n * n;
    const size_t need_state = (size_t)ncell * 4 * n2;
    const bool reuse = c->U[0] && c->n == n && ncell <= c->cap_quad && ninst <= c->cap_inst &&
                       need_state <= c->cap_state && (!need_acc || c->Acc);
    *fresh = !reuse;
    if (reuse) return AMRGPU_OK;
    free_mesh(c);
    const int capq = ncell + ncell / 4 + 16;
    c->cap_quad = capq;
    c->cap_inst = ninst;
    c->cap_state = (size_t)capq * 4 * n2;
    c->cap_blocks = capq + 64;   // enough for one patch per workgroup
    const size_t cq = (size_t)c->cap_quad, cs = c->cap_state;
    HIPCHK(hipMalloc(&c->d_nbr, sizeof(int) * 8 * cq));
    HIPCHK(hipMalloc(&c->d_level, sizeof(int) * cq));
    HIPCHK(hipMalloc(&c->d_inst, sizeof(int) * cq));
    HIPCHK(hipMalloc(&c->d_ip, sizeof(InstDev) * c->cap_inst));
    for (int i = 0; i < 3; ++i) HIPCHK(hipMalloc(&c->U[i], sizeof(double) * cs));
    if (need_acc) HIPCHK(hipMalloc(&c->Acc, sizeof(double) * cs));
    HIPCHK(hipMalloc(&c->d_dt_cur, sizeof(double) * c->cap_inst));
    HIPCHK(hipMalloc(&c->d_t, sizeof(double) * c->cap_inst));
    HIPCHK(hipMalloc(&c->d_slot, sizeof(unsigned long long) * c->cap_inst * NSLOT));
    HIPCHK(hipMalloc(&c->d_norm, sizeof(double) * c->cap_blocks));
    HIPCHK(hipMalloc(&c->d_stage, sizeof(double) * cs));
    HIPCHK(hipMalloc(&c->d_map, sizeof(int) * cq));
    HIPCHK(hipMalloc(&c->d_ind, sizeof(double) * cq));
    HIPCHK(hipMalloc(&c->d_org, sizeof(double) * 2 * cq));
    HIPCHK(hipMalloc(&c->d_tstep, sizeof(double) * c->cap_inst));
    return AMRGPU_OK;
}

// Levels of the reference's cell_info (already validated).
static std::vector<int> levels_of(int ncell, const int *cell_info)
{
    std::vector<int> level(ncell);
    for (int q = 0; q < ncell; ++q) level[q] = cell_info[3 * q + 1];
    return level;
}

// Upload the face table, levels and instances of a mesh already validated.
// Pinned host staging for the uploads of setup, remesh and origins: copies
// from it are plain asynchronous DMA, where pageable ones go through the
// runtime's own staging call by call.  Every user synchronises the stream
// before it returns, so no copy from the buffer is pending whenever it is
// written; growing it synchronises first.
static int pin_staging(amrgpu_ctx *c, size_t nint)
{
    if (nint <= c->h_pin_cap) return AMRGPU_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    c->h_pin = nullptr;
    c->h_pin_cap = 0;
    const size_t cap = nint + nint / 2 + 64;
    HIPCHK(hipHostMalloc((void **)&c->h_pin, sizeof(int) * cap, hipHostMallocDefault));
    c->h_pin_cap = cap;
    return AMRGPU_OK;
}

static int upload_mesh(amrgpu_ctx *c, int ncell, const int *level, const int *cell_inst, const int *nbr)
{
    int rc = pin_staging(c, 10 * (size_t)ncell);     // nbr, level, instance
    if (rc) return rc;
    int *pn = c->h_pin, *pl = pn + 8 * (size_t)ncell, *pq = pl + ncell;
    memcpy(pn, nbr, sizeof(int) * 8 * (size_t)ncell);
    memcpy(pl, level, sizeof(int) * (size_t)ncell);
    if (cell_inst) memcpy(pq, cell_inst, sizeof(int) * (size_t)ncell);
    else memset(pq, 0, sizeof(int) * (size_t)ncell);
    c->nquad = ncell;
    if (c->variant < 0 || kVariants[c->variant].n != c->n) c->variant = default_variant(c->n);
    c->nblocks = grid_of(c, c->variant, ncell);
    if (c->nblocks > c->cap_blocks || (size_t)ncell * 4 * c->n * c->n > c->cap_state) {
        set_err("internal capacity error");
        return AMRGPU_ERR_ARG;
    }
    HIPCHK(hipMemcpyAsync(c->d_nbr, pn, sizeof(int) * 8 * (size_t)ncell, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_level, pl, sizeof(int) * (size_t)ncell, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_inst, pq, sizeof(int) * (size_t)ncell, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMRGPU_OK;
}

int amrgpu_setup(amrgpu_ctx *c, int nelem_subgrid, int time_scheme, int ninst,
                 const amrgpu_inst_t *inst, int ncell, const int *cell_info,
                 const int *cell_inst, int nface, const int *face_info, const double *geom,
                 const double *soln)
{
    if (!c || !inst || !cell_info || !face_info || ncell <= 0 || ninst <= 0 || nface < 0) {
        set_err("amrgpu_setup: bad arguments");
        return AMRGPU_ERR_ARG;
    }
    if (ppb_of(nelem_subgrid) == 0) {
        set_err("amrgpu_setup: nelem_subgrid %d not in {2,4,8,16}", nelem_subgrid);
        return AMRGPU_ERR_ARG;
    }
    if (time_scheme < 1 || time_scheme > 5) {
        set_err("amrgpu_setup: time_scheme %d", time_scheme);
        return AMRGPU_ERR_ARG;
    }
    for (int i = 0; i < ninst; ++i)
        for (int s = 0; s < 4; ++s)
            if (inst[i].bc[s] < 0 || inst[i].bc[s] > AMRGPU_BC_SHOCK_TOP) {
                set_err("amrgpu_setup: instance %d side %d wall type %d", i, s, inst[i].bc[s]);
                return AMRGPU_ERR_ARG;
            }
    const int n2 = nelem_subgrid * nelem_subgrid;
    int rc = validate_cells(n2, ninst, ncell, cell_info, cell_inst, "amrgpu_setup");
    if (rc) return rc;
    std::vector<int> nbr;
    rc = build_nbr(ncell, nface, face_info, nbr);
    if (rc) return rc;

    HIPCHK(hipStreamSynchronize(c->stream));
    free_graphs(c);
    bool fresh;
    rc = ensure_capacity(c, nelem_subgrid, ncell, ninst, time_scheme == 4 || time_scheme == 5, &fresh);
    if (rc) return rc;
    c->n = nelem_subgrid; c->ninst = ninst; c->scheme = time_scheme;
    c->cur = 0;
    c->h_ip.resize(ninst);
    c->need_org = false;
    for (int i = 0; i < ninst; ++i) {
        InstDev &d = c->h_ip[i];
        memset(&d, 0, sizeof(d));
        d.h_base = inst[i].h_base;
        for (int s = 0; s < 4; ++s) {
            d.bc[s] = inst[i].bc[s];
            if (d.bc[s] >= AMRGPU_BC_SHOCK_WALL) c->need_org = true;
        }
        d.cfl = inst[i].cfl;
        d.dt_cap = inst[i].dt_cap;
        d.t_end = inst[i].t_end;
        memcpy(d.bstate, inst[i].bstate, sizeof(d.bstate));
        const double a = inst[i].shock_angle * 3.141592653589793238462643383279502884197 / 180.0;
        d.shock_x0 = inst[i].shock_x0;
        d.shock_cot = a != 0.0 ? cos(a) / sin(a) : 0.0;
        d.shock_vx = a != 0.0 ? inst[i].shock_speed / sin(a) : 0.0;
        const double h0 = inst[i].h_base / nelem_subgrid, v0 = h0 * h0;
        d.ninv_vol0 = -1.0 / v0;
        d.cfl_h0 = inst[i].cfl * sqrt(v0);
    }
    HIPCHK(hipMemcpyAsync(c->d_ip, c->h_ip.data(), sizeof(InstDev) * ninst, hipMemcpyHostToDevice, c->stream));
    std::vector<double> zeros(ninst, 0.0);
    HIPCHK(hipMemcpyAsync(c->d_t, zeros.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_tstep, zeros.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_dt_cur, zeros.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, c->stream));
    rc = upload_mesh(c, ncell, levels_of(ncell, cell_info).data(), cell_inst, nbr.data());
    if (rc) return rc;
    c->org_ok = false;
    if (geom) {
        rc = amrgpu_set_geometry(c, geom);
        if (rc) return rc;
    }
    if (soln) return amrgpu_upload_soln(c, soln);
    return AMRGPU_OK;
}

int amrgpu_set_geometry(amrgpu_ctx *c, const double *geom)
{
    if (!c || !c->nquad || !geom) { set_err("amrgpu_set_geometry: no mesh"); return AMRGPU_ERR_ARG; }
    std::vector<double> org(2 * (size_t)c->nquad);
    for (int q = 0; q < c->nquad; ++q) {
        org[2 * (size_t)q] = geom[12 * (size_t)q];          // corner 0: (x, y) of the lower-left node
        org[2 * (size_t)q + 1] = geom[12 * (size_t)q + 1];
    }
    HIPCHK(hipMemcpyAsync(c->d_org, org.data(), sizeof(double) * org.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->org_ok = true;
    return AMRGPU_OK;
}

int amrgpu_indicator(amrgpu_ctx *c, double *ind)
{
    if (!c || !ind || !c->nquad) { set_err("amrgpu_indicator: no mesh"); return AMRGPU_ERR_ARG; }
    int rc = launch_indicator(c, c->d_ind);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ind, c->d_ind, sizeof(double) * c->nquad, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMRGPU_OK;
}

int amrgpu_feature_tags(amrgpu_ctx *c, int ntags, const int *methods, const double *tols, int *tags)
{
    if (!c || !tags || !c->nquad || ntags < 0 || ntags > AMRGPU_MAX_FEATURE_TAGS ||
        (ntags > 0 && (!methods || !tols))) {
        set_err("amrgpu_feature_tags: bad arguments");
        return AMRGPU_ERR_ARG;
    }
    FeatureTagArgs t;
    memset(&t, 0, sizeof(t));
    t.ntags = ntags;
    for (int k = 0; k < ntags; ++k) { t.method[k] = methods[k]; t.tol[k] = tols[k]; }
    const int n2 = c->n * c->n, ppb = n2 >= 256 ? 1 : 256 / n2;
    const int nblk = (c->nquad + ppb - 1) / ppb;
    int *d_tags = c->d_map;      // regrid source map: free between remeshes
    const double *U = c->U[c->cur];
    switch (c->n) {
    case 2: hipLaunchKernelGGL((feature_tag_kernel<2>), dim3(nblk), dim3(256), 0, c->stream, c->nquad, U, t, d_tags); break;
    case 4: hipLaunchKernelGGL((feature_tag_kernel<4>), dim3(nblk), dim3(256), 0, c->stream, c->nquad, U, t, d_tags); break;
    case 8: hipLaunchKernelGGL((feature_tag_kernel<8>), dim3(nblk), dim3(256), 0, c->stream, c->nquad, U, t, d_tags); break;
    case 16: hipLaunchKernelGGL((feature_tag_kernel<16>), dim3(nblk), dim3(256), 0, c->stream, c->nquad, U, t, d_tags); break;
    default: set_err("unsupported nelem_subgrid %d", c->n); return AMRGPU_ERR_ARG;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(tags, d_tags, sizeof(int) * c->nquad, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMRGPU_OK;
}

// The source map of a remesh refers to the old mesh only.
static int check_src(int ncell, const int *src, int nold, const char *who)
{
    for (int q = 0; q < ncell; ++q) {
        const int s = src[q];
        if ((s >= 0 && s >= nold) || (s < 0 && (-s - 1) + 3 >= nold)) {
            set_err("%s: quadrant %d source %d outside the old mesh (%d)", who, q, s, nold);
            return AMRGPU_ERR_ARG;
        }
    }
    return AMRGPU_OK;
}

// Remap the state onto a validated new mesh and upload its table.
static int remesh_to(amrgpu_ctx *c, int ncell, const int *level, const int *cell_inst, const int *nbr,
                     const int *src)
{
    const int n2 = c->n * c->n, nold = c->nquad;
    int rc = AMRGPU_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    free_graphs(c);
    const long long ncells = (long long)ncell * n2;
    const size_t old_state = (size_t)nold * 4 * n2;
    const bool fits = ncell <= c->cap_quad && (size_t)ncell * 4 * n2 <= c->cap_state;
    if (fits) {
        // old state in U[cur], new state into the next buffer; the source
        // map goes through the staging after the table's part of it
        const int nx = (c->cur + 1) % 3;
        rc = pin_staging(c, 11 * (size_t)ncell);
        if (rc) return rc;
        int *ps = c->h_pin + 10 * (size_t)ncell;
        memcpy(ps, src, sizeof(int) * (size_t)ncell);
        HIPCHK(hipMemcpyAsync(c->d_map, ps, sizeof(int) * ncell, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(remap_kernel, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, c->stream,
                           c->U[c->cur], c->U[nx], c->d_map, ncells, n2);
        HIPCHK(hipGetLastError());
        c->cur = nx;
    } else {
        // keep the old state, time, dt and step start time in scratch
        // allocations, regrow, remap (the residual of a remeshed state reads
        // the step start time before any step rewrites it)
        double *keep = nullptr;
        int *map = nullptr;
        std::vector<double> t(c->ninst), dt(c->ninst), ts(c->ninst);
        HIPCHK(hipMemcpyAsync(t.data(), c->d_t, sizeof(double) * c->ninst, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(dt.data(), c->d_dt_cur, sizeof(double) * c->ninst, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(ts.data(), c->d_tstep, sizeof(double) * c->ninst, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMalloc(&keep, sizeof(double) * old_state));
        HIPCHK(hipMalloc(&map, sizeof(int) * ncell));
        HIPCHK(hipMemcpyAsync(keep, c->U[c->cur], sizeof(double) * old_state, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(map, src, sizeof(int) * ncell, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const int ninst = c->ninst;
        bool fresh;
        rc = ensure_capacity(c, c->n, ncell + ncell / 2, ninst, c->scheme == 4 || c->scheme == 5, &fresh);
        if (rc) { (void)hipFree(keep); (void)hipFree(map); return rc; }
        HIPCHK(hipMemcpyAsync(c->d_ip, c->h_ip.data(), sizeof(InstDev) * ninst, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_t, t.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_dt_cur, dt.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_tstep, ts.data(), sizeof(double) * ninst, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(remap_kernel, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, c->stream,
                           keep, c->U[0], map, ncells, n2);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        (void)hipFree(keep);
        (void)hipFree(map);
        c->cur = 0;
    }
    rc = upload_mesh(c, ncell, level, cell_inst, nbr);
    if (rc) return rc;
    // an instance left without quadrants (a batch drops those that reached
    // their stop) no longer advances: no wave publishes its dt any more, so
    // its last dt is cleared here and the time bookkeeping of the stages
    // keeps its time as it is
    if (cell_inst && c->ninst > 1) {
        std::vector<char> has(c->ninst, 0);
        for (int q = 0; q < ncell; ++q) has[cell_inst[q]] = 1;
        bool empty = false;
        for (int i = 0; i < c->ninst; ++i) empty = empty || !has[i];
        if (empty) {
            std::vector<double> dt(c->ninst);
            HIPCHK(hipMemcpyAsync(dt.data(), c->d_dt_cur, sizeof(double) * c->ninst, hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int i = 0; i < c->ninst; ++i)
                if (!has[i]) dt[i] = 0.0;
            HIPCHK(hipMemcpyAsync(c->d_dt_cur, dt.data(), sizeof(double) * c->ninst, hipMemcpyHostToDevice,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    }
    return AMRGPU_OK;
}

int amrgpu_remesh(amrgpu_ctx *c, int ncell, const int *cell_info, const int *cell_inst, int nface,
                  const int *face_info, const int *src)
{
    if (!c || !c->nquad || !cell_info || !face_info || !src || ncell <= 0 || nface < 0) {
        set_err("amrgpu_remesh: bad arguments");
        return AMRGPU_ERR_ARG;
    }
    int rc = check_src(ncell, src, c->nquad, "amrgpu_remesh");
    if (!rc) rc = validate_cells(c->n * c->n, c->ninst, ncell, cell_info, cell_inst, "amrgpu_remesh");
    if (rc) return rc;
    std::vector<int> nbr;
    rc = build_nbr(ncell, nface, face_info, nbr);
    if (rc) return rc;                 // a refused call leaves the context as it was
    c->org_ok = false;                 // the new quadrants' coordinates come with amrgpu_set_geometry
    return remesh_to(c, ncell, levels_of(ncell, cell_info).data(), cell_inst, nbr.data(), src);
}

int amrgpu_remesh_nbr(amrgpu_ctx *c, int ncell, const int *level, const int *cell_inst, const int *nbr,
                      const int *src)
{
    if (!c || !c->nquad || !level || !nbr || !src || ncell <= 0) {
        set_err("amrgpu_remesh_nbr: bad arguments");
        return AMRGPU_ERR_ARG;
    }
    int rc = check_src(ncell, src, c->nquad, "amrgpu_remesh_nbr");
    if (rc) return rc;
    // every id the kernels will index with must lie in the new mesh
    for (int q = 0; q < ncell; ++q) {
        if (level[q] < 0 || level[q] > 29) {
            set_err("amrgpu_remesh_nbr: quadrant %d level %d", q, level[q]);
            return AMRGPU_ERR_MESH;
        }
        if (cell_inst && (cell_inst[q] < 0 || cell_inst[q] >= c->ninst)) {
            set_err("amrgpu_remesh_nbr: quadrant %d instance %d", q, cell_inst[q]);
            return AMRGPU_ERR_ARG;
        }
        for (int sd = 0; sd < 4; ++sd) {
            const int na = nbr[8 * (size_t)q + 2 * sd], nb = nbr[8 * (size_t)q + 2 * sd + 1];
            if (na < -1 || na >= ncell || nb < -1 || nb >= ncell || (nb >= 0 && na < 0)) {
                set_err("amrgpu_remesh_nbr: quadrant %d side %d neighbours %d %d", q, sd, na, nb);
                return AMRGPU_ERR_MESH;
            }
        }
    }
    c->org_ok = false;                 // the new quadrants' coordinates come with amrgpu_set_origins
    return remesh_to(c, ncell, level, cell_inst, nbr, src);
}

int amrgpu_set_origins(amrgpu_ctx *c, const double *org)
{
    if (!c || !c->nquad || !org) { set_err("amrgpu_set_origins: no mesh"); return AMRGPU_ERR_ARG; }
    // no copy from the staging is pending (its users synchronise first)
    int rc = pin_staging(c, 4 * (size_t)c->nquad);
    if (rc) return rc;
    memcpy(c->h_pin, org, sizeof(double) * 2 * (size_t)c->nquad);
    HIPCHK(hipMemcpyAsync(c->d_org, c->h_pin, sizeof(double) * 2 * (size_t)c->nquad, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->org_ok = true;
    return AMRGPU_OK;
}

int amrgpu_nbr_from_faces(int ncell, int nface, const int *face_info, int *nbr)
{
    if (ncell <= 0 || nface < 0 || !face_info || !nbr) {
        set_err("amrgpu_nbr_from_faces: bad arguments");
        return AMRGPU_ERR_ARG;
    }
    std::vector<int> t;
    int rc = build_nbr(ncell, nface, face_info, t);
    if (rc) return rc;
    memcpy(nbr, t.data(), sizeof(int) * t.size());
    return AMRGPU_OK;
}

int amrgpu_upload_soln(amrgpu_ctx *c, const double *soln)
{
    if (!c || !soln || !c->nquad) { set_err("amrgpu_upload_soln: no mesh"); return AMRGPU_ERR_ARG; }
    const int n2 = c->n * c->n;
    const long long ncells = (long long)c->nquad * n2;
    HIPCHK(hipMemcpyAsync(c->d_stage, soln, sizeof(double) * 4 * ncells, hipMemcpyHostToDevice, c->stream));
    c->cur = 0;
    hipLaunchKernelGGL(aos_to_soa_kernel, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, c->stream,
                       c->d_stage, c->U[0], ncells, n2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMRGPU_OK;
}

int amrgpu_download_soln(amrgpu_ctx *c, double *soln)
{
    if (!c || !soln || !c->nquad) { set_err("amrgpu_download_soln: no mesh"); return AMRGPU_ERR_ARG; }
    const int n2 = c->n * c->n;
    const long long ncells = (long long)c->nquad * n2;
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, c->stream,
                       c->U[c->cur], c->d_stage, ncells, n2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(soln, c->d_stage, sizeof(double) * 4 * ncells, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMRGPU_OK;
}

int amrgpu_download_quads(amrgpu_ctx *c, int first, int count, double *soln)
{
    if (!c || !soln || first < 0 || count < 0 || first + (long long)count > c->nquad) {
        set_err("amrgpu_download_quads: range [%d, %d + %d) outside the %d quadrants", first, first, count,
                c ? c->nquad : 0);
        return AMRGPU_ERR_ARG;
    }
    if (count == 0) return AMRGPU_OK;
    const int n2 = c->n * c->n;
    const long long ncells = (long long)count * n2;
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, c->stream,
                       c->U[c->cur] + (size_t)first * 4 * n2, c->d_stage, ncells, n2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(soln, c->d_stage, sizeof(double) * 4 * ncells, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMRGPU_OK;
}

// position-dependent walls need the quadrant coordinates of the current mesh
static int geometry_ready(amrgpu_ctx *c)
{
    if (c->need_org && !c->org_ok) {
        set_err("wall types 4-5 need amrgpu_set_geometry after a remesh");
        return AMRGPU_ERR_MESH;
    }
    return AMRGPU_OK;
}

int amrgpu_rhs(amrgpu_ctx *c, double *rhs)
{
    if (!c || !rhs || !c->nquad) { set_err("amrgpu_rhs: no mesh"); return AMRGPU_ERR_ARG; }
    if (int rc = geometry_ready(c)) return rc;
    StageArgs a = base_args(c);
    int other = (c->cur + 1) % 3;
    a.Us = c->U[c->cur]; a.Uout = c->U[other]; a.mode = MODE_RHS; a.role = 0;
    int rc = launch_stage(c, a);
    if (rc) return rc;
    const int n2 = c->n * c->n;
    const long long ncells = (long long)c->nquad * n2;
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, c->stream,
                       c->U[other], c->d_stage, ncells, n2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(rhs, c->d_stage, sizeof(double) * 4 * ncells, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMRGPU_OK;
}

int amrgpu_cfl(amrgpu_ctx *c, double *dt_inst)
{
    if (!c || !dt_inst || !c->nquad) { set_err("amrgpu_cfl: no mesh"); return AMRGPU_ERR_ARG; }
    hipLaunchKernelGGL(slot_reset_kernel, dim3(1), dim3(256), 0, c->stream, c->d_slot, c->ninst);
    HIPCHK(hipGetLastError());
    int rc = launch_cfl(c, c->U[c->cur]);
    if (rc) return rc;
    std::vector<unsigned long long> b((size_t)c->ninst * NSLOT);
    HIPCHK(hipMemcpyAsync(b.data(), c->d_slot, sizeof(unsigned long long) * b.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->ninst; ++i) {
        unsigned long long mn = DBL_INF_BITS;
        for (int k = 0; k < NSLOT; ++k) mn = std::min(mn, b[(size_t)i * NSLOT + k]);
        memcpy(&dt_inst[i], &mn, sizeof(double));
    }
    return AMRGPU_OK;
}

static int prime_run(amrgpu_ctx *c)
{
    if (int rc = geometry_ready(c)) return rc;
    // CFL slot of the current state for the first step's dt
    hipLaunchKernelGGL(slot_reset_kernel, dim3(1), dim3(256), 0, c->stream, c->d_slot, c->ninst);
    HIPCHK(hipGetLastError());
    if (c->scheme != 2) return launch_cfl(c, c->U[c->cur]);
    return AMRGPU_OK;
}

static int graph_for(amrgpu_ctx *c, int cur, hipGraphExec_t *out)
{
    if (c->graph[cur]) { *out = c->graph[cur]; return AMRGPU_OK; }
    hipGraph_t g;
    int nx;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(c, cur, false, &nx);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) { set_err("graph capture: %s", hipGetErrorString(e)); return AMRGPU_ERR_HIP; }
    HIPCHK(hipGraphInstantiate(&c->graph[cur], g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    *out = c->graph[cur];
    return AMRGPU_OK;
}

static int step_next(amrgpu_ctx *c, int cur)
{
    std::vector<StageArgs> st;
    return build_step(c, cur, st, false);
}

static int finish_norm(amrgpu_ctx *c, double *resnorm)
{
    std::vector<double> part(c->nblocks);
    HIPCHK(hipMemcpyAsync(part.data(), c->d_norm, sizeof(double) * c->nblocks, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (double v : part) s += v;
    double rn = sqrt(s);
    if (resnorm) *resnorm = rn;
    if (std::isnan(rn)) { set_err("nan detected in residual"); return AMRGPU_ERR_NAN; }
    return AMRGPU_OK;
}

// One step: a replay of the captured step graph, or direct launches.
static int one_step(amrgpu_ctx *c, bool graph, bool with_norm)
{
    int nx, rc;
    if (graph && c->graphs_enabled && !with_norm) {
        hipGraphExec_t g;
        rc = graph_for(c, c->cur, &g);
        if (rc) return rc;
        HIPCHK(hipGraphLaunch(g, c->stream));
        nx = step_next(c, c->cur);
    } else {
        rc = enqueue_step(c, c->cur, with_norm, &nx);
        if (rc) return rc;
    }
    c->cur = nx;
    return AMRGPU_OK;
}

int amrgpu_run(amrgpu_ctx *c, int nsteps, int use_graph, double *resnorm)
{
    if (!c || !c->nquad || nsteps < 0) { set_err("amrgpu_run: no mesh"); return AMRGPU_ERR_ARG; }
    if (nsteps == 0) return AMRGPU_OK;
    int rc = prime_run(c);
    if (rc) return rc;
    // the residual norm (NaN check, reported norm) is taken on the last stage
    // of the last step, the only value the reference keeps
    for (int s = 0; s < nsteps; ++s) {
        rc = one_step(c, use_graph != 0, s == nsteps - 1);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return finish_norm(c, resnorm);
}

int amrgpu_get_time(amrgpu_ctx *c, double *t_inst, double *dt_inst)
{
    if (!c || !c->nquad) { set_err("amrgpu_get_time: no mesh"); return AMRGPU_ERR_ARG; }
    if (t_inst) HIPCHK(hipMemcpyAsync(t_inst, c->d_t, sizeof(double) * c->ninst, hipMemcpyDeviceToHost, c->stream));
    if (dt_inst) HIPCHK(hipMemcpyAsync(dt_inst, c->d_dt_cur, sizeof(double) * c->ninst, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMRGPU_OK;
}

int amrgpu_set_time(amrgpu_ctx *c, const double *t_inst)
{
    if (!c || !c->nquad || !t_inst) { set_err("amrgpu_set_time: no mesh"); return AMRGPU_ERR_ARG; }
    HIPCHK(hipMemcpyAsync(c->d_t, t_inst, sizeof(double) * c->ninst, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_tstep, t_inst, sizeof(double) * c->ninst, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMRGPU_OK;
}

int amrgpu_bench(amrgpu_ctx *c, int nsteps, double *ms_total, double *avg_stage_ms, int *nstages)
{
    if (!c || !c->nquad || nsteps <= 0) { set_err("amrgpu_bench: no mesh"); return AMRGPU_ERR_ARG; }
    int rc = prime_run(c);
    if (rc) return rc;
    // in graph mode, capture the graphs of every starting buffer before timing
    if (c->graphs_enabled) {
        int cur = c->cur;
        for (int s = 0; s < 2; ++s) {
            hipGraphExec_t g;
            rc = graph_for(c, cur, &g);
            if (rc) return rc;
            cur = step_next(c, cur);
        }
    }
    // every event is destroyed on every exit path
    struct Events {
        std::vector<hipEvent_t> e;
        ~Events() { for (auto &x : e) if (x) (void)hipEventDestroy(x); }
    } evs;
    std::vector<StageArgs> st;
    build_step(c, c->cur, st, false);
    const int ns = (int)st.size();
    evs.e.assign(2 + (size_t)nsteps * (ns + 1), nullptr);
    for (auto &e : evs.e) HIPCHK(hipEventCreate(&e));
    hipEvent_t e0 = evs.e[0], e1 = evs.e[1];
    hipEvent_t *ev = evs.e.data() + 2;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int s = 0; s < nsteps; ++s) {
        rc = one_step(c, true, false);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (ms_total) *ms_total = ms;

    // per-stage kernel durations: direct launches bracketed by events
    for (int s = 0; s < nsteps; ++s) {
        int nx = build_step(c, c->cur, st, false);
        hipLaunchKernelGGL(dt_prep_kernel, dim3(std::max(1, c->ninst)), dim3(64), 0, c->stream, mesh_of(c),
                           c->scheme != 2, c->d_dt_cur, c->d_t, c->d_tstep, c->d_slot);
        HIPCHK(hipEventRecord(ev[(size_t)s * (ns + 1)], c->stream));
        for (int k = 0; k < ns; ++k) {
            rc = launch_stage(c, st[k]);
            if (rc) return rc;
            HIPCHK(hipEventRecord(ev[(size_t)s * (ns + 1) + k + 1], c->stream));
        }
        c->cur = nx;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < ns; ++k) {
        double tot = 0.0;
        for (int s = 0; s < nsteps; ++s) {
            float m = 0.f;
            HIPCHK(hipEventElapsedTime(&m, ev[(size_t)s * (ns + 1) + k], ev[(size_t)s * (ns + 1) + k + 1]));
            tot += m;
        }
        if (avg_stage_ms) avg_stage_ms[k] = tot / nsteps;
    }
    if (nstages) *nstages = ns;
    // one more step carrying the residual norm: NaN check of the result
    rc = one_step(c, false, true);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return finish_norm(c, nullptr);
}

int amrgpu_set_tile(amrgpu_ctx *c, int ppb, int minw, int design)
{
    if (!c || !c->nquad) { set_err("amrgpu_set_tile: no mesh"); return AMRGPU_ERR_ARG; }
    for (int v = 0; v < kNumVariants; ++v) {
        if (kVariants[v].n == c->n && kVariants[v].ppb == ppb && kVariants[v].minw == minw &&
            kVariants[v].design == design) {
            HIPCHK(hipStreamSynchronize(c->stream));
            free_graphs(c);
            c->variant = v;
            c->nblocks = grid_of(c, v, c->nquad);
            return AMRGPU_OK;
        }
    }
    set_err("amrgpu_set_tile: no variant n=%d ppb=%d minw=%d design=%d", c->n, ppb, minw, design);
    return AMRGPU_ERR_ARG;
}

int amrgpu_tile_variants(int n, int *ppb, int *minw, int *design, int max)
{
    int k = 0;
    for (int v = 0; v < kNumVariants; ++v) {
        if (kVariants[v].n != n) continue;
        if (k < max) {
            if (ppb) ppb[k] = kVariants[v].ppb;
            if (minw) minw[k] = kVariants[v].minw;
            if (design) design[k] = kVariants[v].design;
        }
        ++k;
    }
    return k;
}

int amrgpu_get_tile(amrgpu_ctx *c, int *ppb, int *minw, int *design)
{
    if (!c || c->variant < 0) { set_err("amrgpu_get_tile: no mesh"); return AMRGPU_ERR_ARG; }
    if (ppb) *ppb = kVariants[c->variant].ppb;
    if (minw) *minw = kVariants[c->variant].minw;
    if (design) *design = kVariants[c->variant].design;
    return AMRGPU_OK;
}

// The autotuner's trials, each followed by a restore of the state copy in
// d_stage and the per-instance time t0; the caller restores both once more
// on every exit, so an error in a trial leaves the context as it was.
static int restore_trial(amrgpu_ctx *c, int cur0, size_t nstate, const std::vector<double> &t0)
{
    HIPCHK(hipMemcpyAsync(c->U[cur0], c->d_stage, sizeof(double) * nstate, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_t, t0.data(), sizeof(double) * c->ninst, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->cur = cur0;
    return AMRGPU_OK;
}

static int timed_trial(amrgpu_ctx *c, int nsteps, hipEvent_t e0, hipEvent_t e1, double *per)
{
    int rc = prime_run(c);
    if (rc) return rc;
    for (int s = 0; s < 2; ++s) if ((rc = one_step(c, true, false))) return rc;   // warm + capture
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int s = 0; s < nsteps; ++s) if ((rc = one_step(c, true, false))) return rc;
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *per = ms / nsteps;
    return AMRGPU_OK;
}

static int autotune_trials(amrgpu_ctx *c, int nsteps, int cur0, size_t nstate, const std::vector<double> &t0,
                           hipEvent_t e0, hipEvent_t e1, int *nvar, int *ppb, int *minw, int *design,
                           double *ms_per_step)
{
    int best = -1, k = 0, rc;
    double best_ms = 1e300;
    for (int v = 0; v < kNumVariants; ++v) {
        if (kVariants[v].n != c->n) continue;
        if ((rc = amrgpu_set_tile(c, kVariants[v].ppb, kVariants[v].minw, kVariants[v].design))) return rc;
        double per = 0.0;
        if ((rc = timed_trial(c, nsteps, e0, e1, &per))) return rc;
        if (ppb) ppb[k] = kVariants[v].ppb;
        if (minw) minw[k] = kVariants[v].minw;
        if (design) design[k] = kVariants[v].design;
        if (ms_per_step) ms_per_step[k] = per;
        ++k;
        if (per < best_ms) { best_ms = per; best = v; }
        if ((rc = restore_trial(c, cur0, nstate, t0))) return rc;
    }
    if (nvar) *nvar = k;
    if (best < 0) { set_err("amrgpu_autotune: no tile variant for n = %d", c->n); return AMRGPU_ERR_ARG; }
    if ((rc = amrgpu_set_tile(c, kVariants[best].ppb, kVariants[best].minw, kVariants[best].design))) return rc;
    // launch mode for the chosen tile: direct launches against graph replay
    double per_mode[2] = {0.0, 0.0};
    for (int m = 0; m < 2; ++m) {
        c->graphs_enabled = (m == 1);
        if ((rc = timed_trial(c, nsteps, e0, e1, &per_mode[m]))) return rc;
        if ((rc = restore_trial(c, cur0, nstate, t0))) return rc;
    }
    c->graphs_enabled = per_mode[1] < per_mode[0];
    if (!c->graphs_enabled) free_graphs(c);
    return AMRGPU_OK;
}

int amrgpu_autotune(amrgpu_ctx *c, int nsteps, int *nvar, int *ppb, int *minw, int *design,
                    double *ms_per_step)
{
    if (!c || !c->nquad || nsteps <= 0) { set_err("amrgpu_autotune: no mesh"); return AMRGPU_ERR_ARG; }
    // the trial steps advance the state: keep a device copy of it and the
    // per-instance time, and restore both afterwards, on every exit path
    const size_t nstate = (size_t)c->nquad * 4 * c->n * c->n;
    std::vector<double> t0(c->ninst);
    HIPCHK(hipMemcpyAsync(t0.data(), c->d_t, sizeof(double) * c->ninst, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_stage, c->U[c->cur], sizeof(double) * nstate, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int cur0 = c->cur;
    const bool graphs0 = c->graphs_enabled;
    const int variant0 = c->variant, nblocks0 = c->nblocks;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = AMRGPU_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_err("amrgpu_autotune: hipEventCreate failed");
        rc = AMRGPU_ERR_HIP;
    } else {
        rc = autotune_trials(c, nsteps, cur0, nstate, t0, e0, e1, nvar, ppb, minw, design, ms_per_step);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) {
        // a failed trial: the tile and launch mode from before, the state
        // and time from before (the error message is kept)
        char msg[sizeof(g_err)];
        memcpy(msg, g_err, sizeof(msg));
        c->variant = variant0;
        c->nblocks = nblocks0;
        c->graphs_enabled = graphs0;
        free_graphs(c);
        (void)hipStreamSynchronize(c->stream);
        (void)restore_trial(c, cur0, nstate, t0);
        memcpy(g_err, msg, sizeof(msg));
    }
    return rc;
}

int amrgpu_set_graphs(amrgpu_ctx *c, int enable)
{
    if (!c) { set_err("amrgpu_set_graphs: no context"); return AMRGPU_ERR_ARG; }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->graphs_enabled = enable != 0;
    if (!c->graphs_enabled) free_graphs(c);
    return AMRGPU_OK;
}

int amrgpu_get_graphs(const amrgpu_ctx *c) { return c && c->graphs_enabled ? 1 : 0; }

// ---------------------------------------------------------------------------
// amrgpu_cfl_noalloc: the per-instance CFL minimum of amrgpu_cfl for callers
// that cannot allocate.  Its device memory beyond the caller's inputs and
// outputs is the file-scope array below (instance table + CFL sub-slots for
// at most AMRGPU_NOALLOC_MAX_INST instances, 130 KB), which each GPU the
// library runs on holds until the process ends.  Host code reaches it only
// through the address hipGetSymbolAddress returns on that device.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t NOALLOC_IP_BYTES = ((sizeof(InstDev) * AMRGPU_NOALLOC_MAX_INST + 255) / 256) * 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_IP_BYTES + sizeof(unsigned long long) * AMRGPU_NOALLOC_MAX_INST * NSLOT;
static __device__ __align__(256) unsigned long long scratch[NOALLOC_BYTES / sizeof(unsigned long long)];

constexpr int NOALLOC_MAX_DEVICES = 64;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES];     // the array's address on each device
std::mutex g_scratch_lock;

struct NoallocInst {
    double h_base[AMRGPU_NOALLOC_MAX_INST];
    double cfl[AMRGPU_NOALLOC_MAX_INST];
};

// instance table (the fields cfl_kernel reads) and every sub-slot reset, in
// one kernel: no memset node if the caller captures the launches
__global__ void noalloc_prep_kernel(InstDev *ip, unsigned long long *slot, int ninst, NoallocInst p)
{
    for (int i = threadIdx.x; i < ninst; i += blockDim.x) {
        InstDev d;
        memset(&d, 0, sizeof(d));
        d.h_base = p.h_base[i];
        d.cfl = p.cfl[i];
        ip[i] = d;
    }
    for (int i = threadIdx.x; i < ninst * NSLOT; i += blockDim.x) slot[i] = DBL_INF_BITS;
}

__global__ void noalloc_finish_kernel(const unsigned long long *slot, int ninst, double *dt)
{
    // one wave per instance
    const int i = blockIdx.x;
    if (i < ninst) {
        const double m = slot_min(slot, i);
        if (threadIdx.x == 0) dt[i] = m;
    }
}

int noalloc_scratch(void **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) { set_err("amrgpu_cfl_noalloc: device %d", dev); return AMRGPU_ERR_ARG; }
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        // the array's own name is a host-side stand-in: only this address,
        // taken once per device with that device current, is valid in
        // kernels and copies
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return AMRGPU_OK;
}
}  // namespace

int amrgpu_cfl_noalloc(int nelem_subgrid, int nquad, int ninst, const double *h_base, const double *cfl,
                       const int *d_level, const int *d_inst, const double *d_U, double *d_dt, void *stream)
{
    if (ninst < 1 || ninst > AMRGPU_NOALLOC_MAX_INST) {
        set_err("amrgpu_cfl_noalloc: %d instances (1 to %d)", ninst, AMRGPU_NOALLOC_MAX_INST);
        return AMRGPU_ERR_ARG;
    }
    if (nquad < 1 || !h_base || !cfl || !d_level || !d_inst || !d_U || !d_dt || ppb_of(nelem_subgrid) == 0) {
        set_err("amrgpu_cfl_noalloc: bad arguments");
        return AMRGPU_ERR_ARG;
    }
    void *base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    InstDev *ip = (InstDev *)base;
    unsigned long long *slot = (unsigned long long *)((char *)base + NOALLOC_IP_BYTES);
    NoallocInst p;
    memset(&p, 0, sizeof(p));
    for (int i = 0; i < ninst; ++i) { p.h_base[i] = h_base[i]; p.cfl[i] = cfl[i]; }
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(noalloc_prep_kernel, dim3(1), dim3(256), 0, st, ip, slot, ninst, p);
    HIPCHK(hipGetLastError());
    MeshDev m;
    memset(&m, 0, sizeof(m));
    m.nquad = nquad;
    m.ninst = ninst;
    m.level = d_level;
    m.inst = d_inst;
    m.ip = ip;
    const int ppb = ppb_of(nelem_subgrid), nblk = (nquad + ppb - 1) / ppb;
    switch (nelem_subgrid) {
    case 2: hipLaunchKernelGGL((cfl_kernel<2, 16>), dim3(nblk), dim3(64), 0, st, m, d_U, slot); break;
    case 4: hipLaunchKernelGGL((cfl_kernel<4, 8>), dim3(nblk), dim3(128), 0, st, m, d_U, slot); break;
    case 8: hipLaunchKernelGGL((cfl_kernel<8, 4>), dim3(nblk), dim3(256), 0, st, m, d_U, slot); break;
    default: hipLaunchKernelGGL((cfl_kernel<16, 1>), dim3(nblk), dim3(256), 0, st, m, d_U, slot); break;
    }
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(noalloc_finish_kernel, dim3(ninst), dim3(64), 0, st, slot, ninst, d_dt);
    HIPCHK(hipGetLastError());
    return AMRGPU_OK;
}

int amrgpu_copy_probe(long long nbytes, int reps, double *gbps)
{
    long long n4 = nbytes / 32;
    if (n4 <= 0 || reps <= 0) { set_err("amrgpu_copy_probe: bad size"); return AMRGPU_ERR_ARG; }
    double4 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, sizeof(double4) * n4));
    HIPCHK(hipMalloc(&b, sizeof(double4) * n4));
    HIPCHK(hipMemset(a, 0, sizeof(double4) * n4));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int blocks = 256 * 8 * 4;
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n4);
    HIPCHK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n4);
    HIPCHK(hipEventRecord(e1, 0));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (gbps) *gbps = 2.0 * (double)n4 * 32.0 * reps / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return AMRGPU_OK;
}

}  // extern "C"
