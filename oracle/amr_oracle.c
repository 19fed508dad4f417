/*
 * amr_oracle.c -- CPU oracle (test infrastructure only; see amr_oracle.h).
 *
 * Every function restates the reference arithmetic operation by operation,
 * in the same order, so that on the same inputs it reproduces the reference's
 * own Fortran/C results (compiled from the reference sources into
 * oracle/_ref by oracle/Makefile) to the last bit where no FMA contraction is
 * involved.  Citations are to files of ackirby88/amr-demo.
 */
#define _POSIX_C_SOURCE 200809L
#include "amr_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define NF 4
/* Fortran Q(4,n,n): element (f,i,j), 0-based here */
#define QIDX(f, i, j, n) ((f) + NF * ((i) + (n) * (j)))

/* ------------------------------------------------------------------------- */
/* a1  src/solver/src/flux_2d_module.F90:63-182                              */
/* ------------------------------------------------------------------------- */
void orc_flux_roe_2d(const double ql[4], const double qr[4], const double norm[2], double flux[4])
{
    const double gamma = 1.4;   /* :82 */
    const double gm1 = 0.4;     /* :83 -- literal, not gamma-1 */
    const double half = 0.5;
    double mag, nx, ny;
    double rhol, oneOrhol, ul, vl, El, Hl, pl, unl;
    double rhor, oneOrhor, ur, vr, Er, Hr, pr, unr;
    double roe_fact, oneOroe_fact, rhobar, ubar, vbar, Hbar, abar, unbar, Mnbar;
    double eigp, eigm, eigp_max, eigm_min, b1, b2, bcoeff, delta1, Mcoeff;
    double fluxl, fluxr, disp;

    mag = sqrt(norm[0] * norm[0] + norm[1] * norm[1]);          /* :88 */
    nx = norm[0] / mag;
    ny = norm[1] / mag;

    rhol = ql[0];                                               /* :93-100 */
    oneOrhol = 1.0 / rhol;
    ul = ql[1] * oneOrhol;
    vl = ql[2] * oneOrhol;
    El = ql[3] * oneOrhol;
    pl = (gamma - 1.0) * (ql[3] - half * rhol * (ul * ul + vl * vl));
    Hl = El + pl * oneOrhol;
    unl = ul * nx + vl * ny;

    rhor = qr[0];                                               /* :103-110 */
    oneOrhor = 1.0 / rhor;
    ur = qr[1] * oneOrhor;
    vr = qr[2] * oneOrhor;
    Er = qr[3] * oneOrhor;
    pr = (gamma - 1.0) * (qr[3] - half * rhor * (ur * ur + vr * vr));
    Hr = Er + pr * oneOrhor;
    unr = ur * nx + vr * ny;

    roe_fact = sqrt(rhol) + sqrt(rhor);                         /* :113-121 */
    oneOroe_fact = 1.0 / roe_fact;
    rhobar = sqrt(rhol * rhor);
    ubar = (ul * sqrt(rhol) + ur * sqrt(rhor)) * oneOroe_fact;
    vbar = (vl * sqrt(rhol) + vr * sqrt(rhor)) * oneOroe_fact;
    Hbar = (Hl * sqrt(rhol) + Hr * sqrt(rhor)) * oneOroe_fact;
    abar = sqrt(gm1 * (Hbar - half * (ubar * ubar + vbar * vbar)));
    unbar = ubar * nx + vbar * ny;
    Mnbar = unbar / abar;

    eigp = unbar + abar;                                        /* :124-141 */
    eigm = unbar - abar;
    eigp_max = unr + abar;
    eigm_min = unl - abar;
    if (eigp > 0.0 || eigp_max > 0.0) b1 = (eigp > eigp_max) ? eigp : eigp_max;
    else b1 = 0.0;
    if (eigm < 0.0 || eigm_min < 0.0) b2 = (eigm < eigm_min) ? eigm : eigm_min;
    else b2 = 0.0;

    bcoeff = b1 * b2 / (b1 - b2);                               /* :144-146 */
    delta1 = (rhor - rhol) - (pr - pl) / (abar * abar);
    Mcoeff = 1.0 / (1.0 + fabs(Mnbar));

    fluxl = ql[1] * nx + ql[2] * ny;                            /* :150-154 */
    fluxr = qr[1] * nx + qr[2] * ny;
    disp = bcoeff * (rhor - rhol) - bcoeff * Mcoeff * delta1;
    flux[0] = (b1 * fluxl - b2 * fluxr) / (b1 - b2) + disp;

    fluxl = (ql[1] * ul + pl) * nx + (ql[2] * ul) * ny;         /* :157-162 */
    fluxr = (qr[1] * ur + pr) * nx + (qr[2] * ur) * ny;
    disp = bcoeff * (qr[1] - ql[1])
         - bcoeff * Mcoeff * (delta1 * ubar + rhobar * (ur - ul - nx * (unr - unl)));
    flux[1] = (b1 * fluxl - b2 * fluxr) / (b1 - b2) + disp;

    fluxl = (ql[1] * vl) * nx + (ql[2] * vl + pl) * ny;         /* :165-170 */
    fluxr = (qr[1] * vr) * nx + (qr[2] * vr + pr) * ny;
    disp = bcoeff * (qr[2] - ql[2])
         - bcoeff * Mcoeff * (delta1 * vbar + rhobar * (vr - vl - ny * (unr - unl)));
    flux[2] = (b1 * fluxl - b2 * fluxr) / (b1 - b2) + disp;

    fluxl = (ql[3] + pl) * unl;                                 /* :173-178 */
    fluxr = (qr[3] + pr) * unr;
    disp = bcoeff * (rhor * Hr - rhol * Hl)
         - bcoeff * Mcoeff * (delta1 * Hbar + rhobar * (Hr - Hl));
    flux[3] = (b1 * fluxl - b2 * fluxr) / (b1 - b2) + disp;

    flux[0] *= mag;                                             /* :180 */
    flux[1] *= mag;
    flux[2] *= mag;
    flux[3] *= mag;
}

/* src/solver/src/flux_2d_module.F90:32-60 */
void orc_flux_native_2d(const double q[4], double f[8])
{
    const double gamma = 1.4;
    double rho = q[0];
    double oneOrho = 1.0 / rho;
    double u = q[1] * oneOrho;
    double v = q[2] * oneOrho;
    double p = (gamma - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
    f[0] = q[1];
    f[1] = q[1] * u + p;
    f[2] = q[1] * v;
    f[3] = u * (q[3] + p);
    f[4] = q[2];
    f[5] = q[2] * u;
    f[6] = q[2] * v + p;
    f[7] = v * (q[3] + p);
}

/* src/solver/src/flux_2d_module.F90:13-29 */
void orc_flux_bc_2d(const double q[4], const double normal[2], double flux[4])
{
    double f[8];
    orc_flux_native_2d(q, f);
    for (int k = 0; k < 4; ++k) flux[k] = f[k] * normal[0] + f[4 + k] * normal[1];
}

/* src/solver/src/bc_2d_module.F90:14-75 */
void orc_bc_2d(int bc_type, const double normal[2], const double q[4], double qbc[4])
{
    const double gamma = 1.4;
    const double gm1 = gamma - 1.0;   /* :308 runtime, 0.3999... */
    double rho = q[0];
    double oneOrho = 1.0 / rho;
    double u = q[1] * oneOrho;
    double v = q[2] * oneOrho;
    (void)gm1;
    switch (bc_type) {
    case 0:                                                   /* :321-323 */
        qbc[0] = q[0]; qbc[1] = q[1]; qbc[2] = q[2]; qbc[3] = q[3];
        break;
    case 1: {                                                 /* :325-339 */
        double mag = sqrt(normal[0] * normal[0] + normal[1] * normal[1]);
        double nx = normal[0] / mag;
        double ny = normal[1] / mag;
        double U_dot_n = u * nx + v * ny;
        double ub = u - U_dot_n * nx;
        double vb = v - U_dot_n * ny;
        qbc[0] = q[0];
        qbc[1] = q[0] * ub;
        qbc[2] = q[0] * vb;
        qbc[3] = q[3];
        break;
    }
    case 2:                                                   /* :341-345 */
        qbc[0] = q[0];
        qbc[1] = 0.0;
        qbc[2] = 0.0;
        qbc[3] = q[3] - 0.5 * rho * (u * u + v * v);
        break;
    default:
        /* Fortran select case with no match leaves Q_bc undefined; keep copy */
        qbc[0] = q[0]; qbc[1] = q[1]; qbc[2] = q[2]; qbc[3] = q[3];
        break;
    }
}

/* src/solver/src/bilinear_node_2d_module.F90:14-36 */
void orc_bilinear_node_2d(const double g[12], const double nat[2], double out[2])
{
    double oneMxi = 0.5 * (1.0 - nat[0]);
    double onePxi = 0.5 * (1.0 + nat[0]);
    double oneMeta = 0.5 * (1.0 - nat[1]);
    double onePeta = 0.5 * (1.0 + nat[1]);
    for (int d = 0; d < 2; ++d) {
        out[d] = g[0 + d] * oneMxi * oneMeta + g[3 + d] * onePxi * oneMeta
               + g[6 + d] * oneMxi * onePeta + g[9 + d] * onePxi * onePeta;
    }
}

static inline void add4(double *r, const double *f) { r[0] += f[0]; r[1] += f[1]; r[2] += f[2]; r[3] += f[3]; }
static inline void sub4(double *r, const double *f) { r[0] -= f[0]; r[1] -= f[1]; r[2] -= f[2]; r[3] -= f[3]; }

/* ------------------------------------------------------------------------- */
/* a2  src/solver/src/solver_residual_subgrid_2d_module.F90:32-164           */
/* ------------------------------------------------------------------------- */
void orc_residual_subgrid_2d(int nfields, int n, int type, int level,
                             const double geom[12], const double *Q, double *res)
{
    double scale_factor = 2.0 / (double)n;                      /* :66 */
    double n1[2], n2[2], n3[2], n4[2], a1[2], a2[2], a3[2], a4[2];
    double normvec[2], flux[4];
    (void)nfields; (void)type; (void)level;

    /* x faces (:82-102), Fortran j=1..n, i=1..n-1 */
    for (int j = 1; j <= n; ++j) {
        a1[1] = scale_factor * (double)(j - 1) - 1.0;
        a2[1] = scale_factor * (double)j - 1.0;
        for (int i = 1; i <= n - 1; ++i) {
            a1[0] = scale_factor * (double)i - 1.0;
            a2[0] = scale_factor * (double)i - 1.0;
            orc_bilinear_node_2d(geom, a1, n1);
            orc_bilinear_node_2d(geom, a2, n2);
            normvec[0] = n2[1] - n1[1];
            normvec[1] = n1[0] - n2[0];
            orc_flux_roe_2d(&Q[QIDX(0, i - 1, j - 1, n)], &Q[QIDX(0, i, j - 1, n)], normvec, flux);
            add4(&res[QIDX(0, i - 1, j - 1, n)], flux);
            sub4(&res[QIDX(0, i, j - 1, n)], flux);
        }
    }
    /* y faces (:105-126), j=1..n-1, i=1..n */
    for (int j = 1; j <= n - 1; ++j) {
        a1[1] = scale_factor * (double)j - 1.0;
        a2[1] = scale_factor * (double)j - 1.0;
        for (int i = 1; i <= n; ++i) {
            a1[0] = scale_factor * (double)i - 1.0;
            a2[0] = scale_factor * (double)(i - 1) - 1.0;
            orc_bilinear_node_2d(geom, a1, n1);
            orc_bilinear_node_2d(geom, a2, n2);
            normvec[0] = n2[1] - n1[1];
            normvec[1] = n1[0] - n2[0];
            orc_flux_roe_2d(&Q[QIDX(0, i - 1, j - 1, n)], &Q[QIDX(0, i - 1, j, n)], normvec, flux);
            add4(&res[QIDX(0, i - 1, j - 1, n)], flux);
            sub4(&res[QIDX(0, i - 1, j, n)], flux);
        }
    }
    /* divide by volume (:129-163) */
    for (int j = 1; j <= n; ++j) {
        double v1 = scale_factor * (double)(j - 1) - 1.0;
        double v2 = scale_factor * (double)j - 1.0;
        a1[1] = v1; a2[1] = v1; a3[1] = v2; a4[1] = v2;
        for (int i = 1; i <= n; ++i) {
            double w1 = scale_factor * (double)(i - 1) - 1.0;
            double w2 = scale_factor * (double)i - 1.0;
            double vec1[2], vec2[2], cr, volume;
            a1[0] = w1; a2[0] = w2; a3[0] = w1; a4[0] = w2;
            orc_bilinear_node_2d(geom, a1, n1);
            orc_bilinear_node_2d(geom, a2, n2);
            orc_bilinear_node_2d(geom, a3, n3);
            orc_bilinear_node_2d(geom, a4, n4);
            vec1[0] = n4[0] - n1[0]; vec1[1] = n4[1] - n1[1];
            vec2[0] = n3[0] - n2[0]; vec2[1] = n3[1] - n2[1];
            cr = vec1[0] * vec2[1] - vec1[1] * vec2[0];
            volume = 0.5 * sqrt(cr * cr);
            for (int f = 0; f < 4; ++f) res[QIDX(f, i - 1, j - 1, n)] = res[QIDX(f, i - 1, j - 1, n)] / volume;
        }
    }
}

static inline int owned(int ind, int lo, int hi) { return ind >= lo && ind < hi; }

static double cell_h_of(double h_base, int level, int n)
{
    /* h_base * (1/2**level) / real(n)  (e.g. ..._face_full_structured_2d_module.F90:75) */
    return h_base * (1.0 / pow(2.0, (double)level)) / (double)n;
}

/* a3 */
void orc_residual_face_full_2d(int nfields, int n, int side, int level,
                               int ind_l, int ind_r, int lo, int hi, double h_base,
                               const double *ql, const double *qr, double *res)
{
    double normal[2], flux[4];
    double cell_h = cell_h_of(h_base, level, n);
    (void)nfields;
    if (side == 1) {                                            /* :80-91 */
        normal[0] = cell_h; normal[1] = 0.0;
        for (int j = 0; j < n; ++j) {
            orc_flux_roe_2d(&ql[QIDX(0, n - 1, j, n)], &qr[QIDX(0, 0, j, n)], normal, flux);
            if (owned(ind_l, lo, hi)) add4(&res[ind_l + QIDX(0, n - 1, j, n)], flux);
            if (owned(ind_r, lo, hi)) sub4(&res[ind_r + QIDX(0, 0, j, n)], flux);
        }
    } else if (side == 3) {                                     /* :93-104 */
        normal[0] = 0.0; normal[1] = cell_h;
        for (int i = 0; i < n; ++i) {
            orc_flux_roe_2d(&ql[QIDX(0, i, n - 1, n)], &qr[QIDX(0, i, 0, n)], normal, flux);
            if (owned(ind_l, lo, hi)) add4(&res[ind_l + QIDX(0, i, n - 1, n)], flux);
            if (owned(ind_r, lo, hi)) sub4(&res[ind_r + QIDX(0, i, 0, n)], flux);
        }
    }
}

/* a4 */
void orc_residual_face_full_hang_2d(int nfields, int n, int side, int level_l, int level_r,
                                    int ind_l, int ind_r1, int ind_r2, int lo, int hi,
                                    double h_base, const double *ql, const double *qr1,
                                    const double *qr2, double *res)
{
    double normal[2], flux[4];
    double cell_h = cell_h_of(h_base, level_r, n);              /* :81 fine level */
    (void)nfields; (void)level_l;
    if (side == 1) {
        normal[0] = cell_h; normal[1] = 0.0;
        for (int j = 0; j < n; ++j) {
            orc_flux_roe_2d(&ql[QIDX(0, n - 1, j, n)], &qr1[QIDX(0, 0, j, n)], normal, flux);
            if (owned(ind_l, lo, hi)) add4(&res[ind_l + QIDX(0, n - 1, j, n)], flux);
            if (owned(ind_r1, lo, hi)) sub4(&res[ind_r1 + QIDX(0, 0, j, n)], flux);
            orc_flux_roe_2d(&ql[QIDX(0, n - 1, j, n)], &qr2[QIDX(0, 0, j, n)], normal, flux);
            if (owned(ind_l, lo, hi)) add4(&res[ind_l + QIDX(0, n - 1, j, n)], flux);
            if (owned(ind_r2, lo, hi)) sub4(&res[ind_r2 + QIDX(0, 0, j, n)], flux);
        }
    } else if (side == 3) {
        normal[0] = 0.0; normal[1] = cell_h;
        for (int i = 0; i < n; ++i) {
            orc_flux_roe_2d(&ql[QIDX(0, i, n - 1, n)], &qr1[QIDX(0, i, 0, n)], normal, flux);
            if (owned(ind_l, lo, hi)) add4(&res[ind_l + QIDX(0, i, n - 1, n)], flux);
            if (owned(ind_r1, lo, hi)) sub4(&res[ind_r1 + QIDX(0, i, 0, n)], flux);
            orc_flux_roe_2d(&ql[QIDX(0, i, n - 1, n)], &qr2[QIDX(0, i, 0, n)], normal, flux);
            if (owned(ind_l, lo, hi)) add4(&res[ind_l + QIDX(0, i, n - 1, n)], flux);
            if (owned(ind_r2, lo, hi)) sub4(&res[ind_r2 + QIDX(0, i, 0, n)], flux);
        }
    }
}

/* a5 */
void orc_residual_face_hang_full_2d(int nfields, int n, int side, int level_l, int level_r,
                                    int ind_l1, int ind_l2, int ind_r, int lo, int hi,
                                    double h_base, const double *ql1, const double *ql2,
                                    const double *qr, double *res)
{
    double normal[2], flux[4];
    double cell_h = cell_h_of(h_base, level_l, n);              /* :81 fine level */
    (void)nfields; (void)level_r;
    if (side == 1) {
        normal[0] = cell_h; normal[1] = 0.0;
        for (int j = 0; j < n; ++j) {
            orc_flux_roe_2d(&ql1[QIDX(0, n - 1, j, n)], &qr[QIDX(0, 0, j, n)], normal, flux);
            if (owned(ind_l1, lo, hi)) add4(&res[ind_l1 + QIDX(0, n - 1, j, n)], flux);
            if (owned(ind_r, lo, hi)) sub4(&res[ind_r + QIDX(0, 0, j, n)], flux);
            orc_flux_roe_2d(&ql2[QIDX(0, n - 1, j, n)], &qr[QIDX(0, 0, j, n)], normal, flux);
            if (owned(ind_l2, lo, hi)) add4(&res[ind_l2 + QIDX(0, n - 1, j, n)], flux);
            if (owned(ind_r, lo, hi)) sub4(&res[ind_r + QIDX(0, 0, j, n)], flux);
        }
    } else if (side == 3) {
        normal[0] = 0.0; normal[1] = cell_h;
        for (int i = 0; i < n; ++i) {
            orc_flux_roe_2d(&ql1[QIDX(0, i, n - 1, n)], &qr[QIDX(0, i, 0, n)], normal, flux);
            if (owned(ind_l1, lo, hi)) add4(&res[ind_l1 + QIDX(0, i, n - 1, n)], flux);
            if (owned(ind_r, lo, hi)) sub4(&res[ind_r + QIDX(0, i, 0, n)], flux);
            orc_flux_roe_2d(&ql2[QIDX(0, i, n - 1, n)], &qr[QIDX(0, i, 0, n)], normal, flux);
            if (owned(ind_l2, lo, hi)) add4(&res[ind_l2 + QIDX(0, i, n - 1, n)], flux);
            if (owned(ind_r, lo, hi)) sub4(&res[ind_r + QIDX(0, i, 0, n)], flux);
        }
    }
}

/* a6  src/solver/src/solver_residual_bc_structured_2d_module.F90:31-88,
 *     index map src/solver/src/subgrid_index_2d_module.F90:14-46 */
void orc_residual_bc_2d(int nfields, int n, int side, int type, int level, int bc,
                        double h_base, const double *soln, double *res)
{
    double normal[2] = {0.0, 0.0}, qbc[4], flux[4];
    double cell_h = cell_h_of(h_base, level, n);
    (void)nfields; (void)type;
    switch (side) {
    case 0: normal[0] = -cell_h; break;
    case 1: normal[0] = cell_h; break;
    case 2: normal[1] = -cell_h; break;
    case 3: normal[1] = cell_h; break;
    }
    for (int k = 0; k < n; ++k) {
        int i = 0, j = 0;
        switch (side) {
        case 0: i = 0; j = k; break;
        case 1: i = n - 1; j = k; break;
        case 2: i = k; j = 0; break;
        case 3: i = k; j = n - 1; break;
        }
        orc_bc_2d(bc, normal, &soln[QIDX(0, i, j, n)], qbc);
        orc_flux_bc_2d(qbc, normal, flux);
        add4(&res[QIDX(0, i, j, n)], flux);
    }
}

/* project wall types 3-5 (no reference counterpart) */
void orc_residual_bc_ext(const orc_params *p, int n, int side, int level, const double geom[12],
                         const double *soln, double *res)
{
    const double pi = 3.141592653589793238462643383279502884197;
    double normal[2] = {0.0, 0.0}, qbc[4], flux[4];
    const double h = cell_h_of(p->coarse_grid_h, level, n);
    const double a = p->shock_angle * pi / 180.0;
    const double cot = a != 0.0 ? cos(a) / sin(a) : 0.0;
    const double vx = a != 0.0 ? p->shock_speed / sin(a) : 0.0;
    const double x0 = geom[0], y0 = geom[1];
    switch (side) {
    case 0: normal[0] = -h; break;
    case 1: normal[0] = h; break;
    case 2: normal[1] = -h; break;
    case 3: normal[1] = h; break;
    }
    for (int k = 0; k < n; ++k) {
        int i = 0, j = 0;
        double xf = x0, yf = y0;
        switch (side) {
        case 0: i = 0; j = k; yf = y0 + (k + 0.5) * h; break;
        case 1: i = n - 1; j = k; xf = x0 + (double)n * h; yf = y0 + (k + 0.5) * h; break;
        case 2: i = k; j = 0; xf = x0 + (k + 0.5) * h; break;
        case 3: i = k; j = n - 1; xf = x0 + (k + 0.5) * h; yf = y0 + (double)n * h; break;
        }
        const double *q = &soln[QIDX(0, i, j, n)];
        const int bc = p->bc[side];
        const double *st = p->bstate[0];
        if (bc == 4 && !(xf < p->shock_x0)) {
            orc_bc_2d(1, normal, q, qbc);                      /* slip wall beyond x0 */
        } else {
            if (bc == 5 && !(xf < p->shock_x0 + yf * cot + vx * p->time)) st = p->bstate[1];
            for (int f = 0; f < 4; ++f) qbc[f] = st[f];
        }
        orc_flux_bc_2d(qbc, normal, flux);
        add4(&res[QIDX(0, i, j, n)], flux);
    }
}

/* a11  src/solver/src/solver_time_step_2d_module.F90:26-122 */
void orc_time_step_2d(int nfields, int n, int type, int level, const double geom[12],
                      const double *Q, double cfl, double *dt)
{
    const double gamma = 1.4, gm1 = 0.4, half = 0.5;
    double scale_factor = 2.0 / (double)n;
    double a1[2], a2[2], a3[2], a4[2], n1[2], n2[2], n3[2], n4[2];
    (void)nfields; (void)type; (void)level;
    for (int j = 1; j <= n; ++j) {
        double v1 = scale_factor * (double)(j - 1) - 1.0;
        double v2 = scale_factor * (double)j - 1.0;
        a1[1] = v1; a2[1] = v1; a3[1] = v2; a4[1] = v2;
        for (int i = 1; i <= n; ++i) {
            double w1 = scale_factor * (double)(i - 1) - 1.0;
            double w2 = scale_factor * (double)i - 1.0;
            double vec1[2], vec2[2], cr, volume, oneOrho, u, v, V2, pOverRho, c, eig, cand;
            const double *q = &Q[QIDX(0, i - 1, j - 1, n)];
            a1[0] = w1; a2[0] = w2; a3[0] = w1; a4[0] = w2;
            orc_bilinear_node_2d(geom, a1, n1);
            orc_bilinear_node_2d(geom, a2, n2);
            orc_bilinear_node_2d(geom, a3, n3);
            orc_bilinear_node_2d(geom, a4, n4);
            vec1[0] = n4[0] - n1[0]; vec1[1] = n4[1] - n1[1];
            vec2[0] = n3[0] - n2[0]; vec2[1] = n3[1] - n2[1];
            cr = vec1[0] * vec2[1] - vec1[1] * vec2[0];
            volume = 0.5 * sqrt(cr * cr);
            oneOrho = 1.0 / q[0];
            u = q[1] * oneOrho;
            v = q[2] * oneOrho;
            V2 = u * u + v * v;
            pOverRho = gm1 * (q[3] * oneOrho - half * V2);
            c = sqrt(gamma * pOverRho);
            eig = sqrt(V2) + c;
            cand = cfl * sqrt(volume) / eig;
            /* Fortran min(dt, cand) */
            if (cand < *dt) *dt = cand;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* a8  src/physics/src/physics_rhs_structured.c:19-283                       */
/* ------------------------------------------------------------------------- */
static void face_residual(const orc_params *p, const int *cell_info, const double *geom,
                          const int *fi, const double *soln, double *rhs, int lo, int hi)
{
    const int n = p->nelem_subgrid;
    int facetype = fi[0], side = fi[1], cell = fi[2];
    int level = cell_info[3 * cell + 1];
    int ind = cell_info[3 * cell + 2];
    int type = cell_info[3 * cell + 0];
    int ql[4], qr[4], ind_l[4], ind_r[4];
    (void)geom;

    if (side % 2 == 0) {                                        /* :87-110 */
        for (int k = 0; k < 4; ++k) {
            ql[k] = fi[3 + k];
            ind_l[k] = (ql[k] >= 0) ? cell_info[3 * ql[k] + 2] : -1;
        }
        qr[0] = cell; ind_r[0] = cell_info[3 * cell + 2];
    } else {                                                    /* :111-135 */
        for (int k = 0; k < 4; ++k) {
            qr[k] = fi[3 + k];
            ind_r[k] = (qr[k] >= 0) ? cell_info[3 * qr[k] + 2] : -1;
        }
        ql[0] = cell; ind_l[0] = cell_info[3 * cell + 2];
    }
    /* :137-138 -- level of ql[0] / qr[0] (a boundary face has qr/ql = -1;
     * its level is not used) */
    int level_l = (ql[0] >= 0) ? cell_info[3 * ql[0] + 1] : 0;
    int level_r = (qr[0] >= 0) ? cell_info[3 * qr[0] + 1] : 0;

    switch (facetype) {
    case 1:                                                     /* :142-155 */
        if (owned(ind, lo, hi)) {
            if (p->bc[side] >= 3)
                orc_residual_bc_ext(p, n, side, level, &geom[12 * cell], &soln[ind], &rhs[ind]);
            else
                orc_residual_bc_2d(4, n, side, type, level, p->bc[side], p->coarse_grid_h,
                                   &soln[ind], &rhs[ind]);
        }
        break;
    case 2: {                                                   /* :157-176 */
        int side_plus = (side % 2 == 0) ? side + 1 : side;
        orc_residual_face_full_2d(4, n, side_plus, level_l, ind_l[0], ind_r[0], lo, hi,
                                  p->coarse_grid_h, &soln[ind_l[0]], &soln[ind_r[0]], rhs);
        break;
    }
    case 3:                                                     /* :178-218 */
        if (side % 2 == 0) {
            orc_residual_face_hang_full_2d(4, n, side + 1, level_l, level_r,
                                           ind_l[0], ind_l[1], ind_r[0], lo, hi,
                                           p->coarse_grid_h, &soln[ind_l[0]], &soln[ind_l[1]],
                                           &soln[ind_r[0]], rhs);
        } else {
            orc_residual_face_full_hang_2d(4, n, side, level_l, level_r,
                                           ind_l[0], ind_r[0], ind_r[1], lo, hi,
                                           p->coarse_grid_h, &soln[ind_l[0]], &soln[ind_r[0]],
                                           &soln[ind_r[1]], rhs);
        }
        break;
    default:
        break;
    }
}

static void rhs_impl(const orc_params *p, int ncell, const int *cell_info, const double *geom,
                     int nface, const int *face_info, const int *face_ids, int nfid,
                     const double *soln, double *rhs, int c0, int c1)
{
    const int n = p->nelem_subgrid;
    const int nftm = NF * n * n;
    int lo = c0 * nftm, hi = c1 * nftm;
    (void)ncell;
    /* zero the rhs (:67-69) */
    for (int i = lo; i < hi; ++i) rhs[i] = 0.0;
    /* face loop (:73-262) */
    if (face_ids) {
        for (int k = 0; k < nfid; ++k)
            face_residual(p, cell_info, geom, &face_info[11 * face_ids[k]], soln, rhs, lo, hi);
    } else {
        for (int f = 0; f < nface; ++f)
            face_residual(p, cell_info, geom, &face_info[11 * f], soln, rhs, lo, hi);
    }
    /* element loop (:266-276) */
    for (int c = c0; c < c1; ++c) {
        int type = cell_info[3 * c + 0], level = cell_info[3 * c + 1], ind = cell_info[3 * c + 2];
        orc_residual_subgrid_2d(4, n, type, level, &geom[12 * c], &soln[ind], &rhs[ind]);
    }
    /* rhs = -R (:279-281) */
    for (int i = lo; i < hi; ++i) rhs[i] *= -1.0;
}

void orc_rhs_structured(const orc_params *p, int ncell, const int *cell_info,
                        const double *geom, int nface, const int *face_info, int dof,
                        const double *soln, double *rhs)
{
    (void)dof;
    rhs_impl(p, ncell, cell_info, geom, nface, face_info, NULL, 0, soln, rhs, 0, ncell);
}

void orc_rhs_structured_range(const orc_params *p, int ncell, const int *cell_info,
                              const double *geom, int nface, const int *face_info,
                              const double *soln, double *rhs, int c0, int c1)
{
    rhs_impl(p, ncell, cell_info, geom, nface, face_info, NULL, 0, soln, rhs, c0, c1);
}

/* a11 driver  src/physics/src/physics_time_step.c:16-49 */
void orc_time_step(const orc_params *p, int ncell, const int *cell_info,
                   const double *geom, const double *soln, double *dt)
{
    for (int c = 0; c < ncell; ++c)
        orc_time_step_2d(4, p->nelem_subgrid, cell_info[3 * c], cell_info[3 * c + 1],
                         &geom[12 * c], &soln[cell_info[3 * c + 2]], p->cfl, dt);
}

/* a10  src/physics/src/physics_norm_residual.c:15-56 */
double orc_norm_L2_residual(int n, int ncell, const int *cell_info, const double *res)
{
    double s = 0.0;
    int nftm = NF * n * n;
    for (int c = 0; c < ncell; ++c) {
        const double *r = &res[cell_info[3 * c + 2]];
        for (int j = 0; j < nftm; ++j) s += r[j] * r[j];
    }
    return s;
}

/* ------------------------------------------------------------------------- */
/* a9  src/physics/src/rk_time_integrators.c                                 */
/* Stage-update arithmetic restated per scheme; the update runs over [i0,i1) */
/* ------------------------------------------------------------------------- */
typedef void (*orc_rhs_cb)(void *ctx, double *Q, double *k);

typedef struct {
    const orc_params *p;
    int ncell, nface, dof;
    const int *cell_info, *face_info;
    const double *geom;
    const int *face_ids; int nfid;   /* faces this worker evaluates */
    int c0, c1;                      /* owned quadrants */
    int scheme;
    int i0, i1;                      /* owned dof range */
    orc_rhs_cb cb; void *cb_ctx;     /* generic operator (closed-form tests) */
} rk_ctx;

typedef void (*sync_fn)(void *);

static void rk_rhs(const rk_ctx *c, double *Q, double *k)
{
    if (c->cb) { c->cb(c->cb_ctx, Q, k); return; }
    rhs_impl(c->p, c->ncell, c->cell_info, c->geom, c->nface, c->face_info,
             c->face_ids, c->nfid, Q, k, c->c0, c->c1);
}

static double rk_norm_part(const rk_ctx *c, const double *k)
{
    double s = 0.0;
    for (int i = c->i0; i < c->i1; ++i) s += k[i] * k[i];
    return s;
}

/* One step of the scheme; sync() is a barrier across workers (NULL when
 * single-threaded).  Returns this worker's part of the last stage's sum of
 * squares. */
static double rk_step_part(const rk_ctx *c, double dt, double *Q, double *work,
                           sync_fn sync, void *sarg)
{
    const int dof = c->dof;
    const int i0 = c->i0, i1 = c->i1;
    double *k1 = work, *k2 = work + dof, *k3 = work + 2 * dof, *k4 = work + 3 * dof;
    double *Qs = work + 4 * dof;
    double rn = 0.0;
#define SYNC() do { if (sync) sync(sarg); } while (0)
    switch (c->scheme) {
    case 1: {                                                  /* rk_1 :22-77 */
        double beta1 = 1.0 * dt;
        rk_rhs(c, Q, k1); rn = rk_norm_part(c, k1); SYNC();
        for (int i = i0; i < i1; ++i) Q[i] = Q[i] + beta1 * k1[i];
        SYNC();
        break;
    }
    case 2: {                                                  /* rk_2 :80-165 */
        double alpha1 = 1.0 * dt, beta1 = 0.5 * dt, beta2 = 0.5 * dt;
        for (int i = i0; i < i1; ++i) Qs[i] = Q[i];
        rk_rhs(c, Q, k1); rn = rk_norm_part(c, k1); SYNC();
        for (int i = i0; i < i1; ++i) Q[i] = Qs[i] + alpha1 * k1[i];
        SYNC();
        rk_rhs(c, Q, k2); rn = rk_norm_part(c, k2); SYNC();
        for (int i = i0; i < i1; ++i) Q[i] = Qs[i] + beta1 * k1[i] + beta2 * k2[i];
        SYNC();
        break;
    }
    case 3: {                                                  /* rk_3 :169-277 */
        double alpha1 = 1.00 * dt, alpha21 = 0.25 * dt, alpha22 = 0.25 * dt;
        double beta1 = 1.0 / 6.0 * dt, beta2 = 1.0 / 6.0 * dt, beta3 = 2.0 / 3.0 * dt;
        for (int i = i0; i < i1; ++i) Qs[i] = Q[i];
        rk_rhs(c, Q, k1); rn = rk_norm_part(c, k1); SYNC();
        for (int i = i0; i < i1; ++i) Q[i] = Qs[i] + alpha1 * k1[i];
        SYNC();
        rk_rhs(c, Q, k2); rn = rk_norm_part(c, k2); SYNC();
        for (int i = i0; i < i1; ++i) Q[i] = Qs[i] + alpha21 * k1[i] + alpha22 * k2[i];
        SYNC();
        rk_rhs(c, Q, k3); rn = rk_norm_part(c, k3); SYNC();
        for (int i = i0; i < i1; ++i) Q[i] = Qs[i] + beta1 * k1[i] + beta2 * k2[i] + beta3 * k3[i];
        SYNC();
        break;
    }
    case 4: {                                                  /* rk_4 :280-400 */
        double alpha1 = 0.5 * dt, alpha2 = 0.5 * dt, alpha3 = 1.0 * dt, beta1 = 1.0 / 6.0 * dt;
        for (int i = i0; i < i1; ++i) Qs[i] = Q[i];
        rk_rhs(c, Q, k1); rn = rk_norm_part(c, k1); SYNC();
        for (int i = i0; i < i1; ++i) Q[i] = Qs[i] + alpha1 * k1[i];
        SYNC();
        rk_rhs(c, Q, k2); rn = rk_norm_part(c, k2); SYNC();
        for (int i = i0; i < i1; ++i) Q[i] = Qs[i] + alpha2 * k2[i];
        SYNC();
        rk_rhs(c, Q, k3); rn = rk_norm_part(c, k3); SYNC();
        for (int i = i0; i < i1; ++i) Q[i] = Qs[i] + alpha3 * k3[i];
        SYNC();
        rk_rhs(c, Q, k4); rn = rk_norm_part(c, k4); SYNC();
        for (int i = i0; i < i1; ++i) Q[i] = Qs[i] + beta1 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
        SYNC();
        break;
    }
    case 5: {                                                  /* lserk_45 :403-540 */
        static const double a[5] = {0.0,
            -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
            -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
        static const double b[5] = {1432997174477.0 / 9575080441755.0,
            5161836677717.0 / 13612068292357.0, 1720146321549.0 / 2090206949498.0,
            3134564353537.0 / 4481467310338.0, 2277821191437.0 / 14882151754819.0};
        for (int s = 0; s < 5; ++s) {
            rk_rhs(c, Q, k1); rn = rk_norm_part(c, k1); SYNC();
            if (s == 0) {
                for (int i = i0; i < i1; ++i) { Qs[i] = dt * k1[i]; Q[i] = Q[i] + b[0] * Qs[i]; }
            } else {
                for (int i = i0; i < i1; ++i) { Qs[i] = Qs[i] * a[s] + dt * k1[i]; Q[i] = Q[i] + b[s] * Qs[i]; }
            }
            SYNC();
        }
        break;
    }
    default:
        break;
    }
#undef SYNC
    return rn;
}

double orc_rk_step(const orc_params *p, int ncell, const int *cell_info, const double *geom,
                   int nface, const int *face_info, int dof, double dt, double *Q,
                   double *work)
{
    const int nftm = NF * p->nelem_subgrid * p->nelem_subgrid;
    rk_ctx c = {p, ncell, nface, dof, cell_info, face_info, geom, NULL, 0, 0, ncell,
                p->time_scheme, 0, ncell * nftm, NULL, NULL};
    return sqrt(rk_step_part(&c, dt, Q, work, NULL, NULL));
}

double orc_rk_generic(int scheme, int dof, double dt, double *Q, double *work,
                      orc_rhs_cb cb, void *ctx)
{
    rk_ctx c;
    memset(&c, 0, sizeof(c));
    c.dof = dof; c.scheme = scheme; c.i0 = 0; c.i1 = dof; c.cb = cb; c.cb_ctx = ctx;
    return sqrt(rk_step_part(&c, dt, Q, work, NULL, NULL));
}

double orc_run_steps(const orc_params *p, int ncell, const int *cell_info, const double *geom,
                     int nface, const int *face_info, int dof, int nsteps, double *Q,
                     double *dt_last)
{
    double *work = (double *)malloc(sizeof(double) * 5 * (size_t)dof);
    double rn = -1.0, dt = p->dt;
    for (int s = 0; s < nsteps; ++s) {
        dt = p->dt;                                /* physics_runge_kutta_interface.c:198 */
        if (p->time_scheme != 2)                   /* RK2 skips the CFL (:142-144) */
            orc_time_step(p, ncell, cell_info, geom, Q, &dt);
        rn = orc_rk_step(p, ncell, cell_info, geom, nface, face_info, dof, dt, Q, work);
    }
    free(work);
    if (dt_last) *dt_last = dt;
    return rn;
}

/* ------------------------------------------------------------------------- */
/* Threaded step loop (CPU baseline): one worker per core, each owning a     */
/* contiguous quadrant range.                                                */
/* ------------------------------------------------------------------------- */
typedef struct {
    rk_ctx c;
    pthread_barrier_t *bar;
    double *Q, *work;
    int nsteps, tid, nthreads;
    double *dt_part, *rn_part;
    double dt_out, rn_out;
} worker_arg;

static void bar_wait(void *b) { pthread_barrier_wait((pthread_barrier_t *)b); }

static void *worker_main(void *vp)
{
    worker_arg *w = (worker_arg *)vp;
    const orc_params *p = w->c.p;
    double dt = p->dt, rn = 0.0;
    for (int s = 0; s < w->nsteps; ++s) {
        double dloc = p->dt;
        if (p->time_scheme != 2) {
            for (int c = w->c.c0; c < w->c.c1; ++c)
                orc_time_step_2d(4, p->nelem_subgrid, w->c.cell_info[3 * c],
                                 w->c.cell_info[3 * c + 1], &w->c.geom[12 * c],
                                 &w->Q[w->c.cell_info[3 * c + 2]], p->cfl, &dloc);
        }
        w->dt_part[w->tid] = dloc;
        pthread_barrier_wait(w->bar);
        dt = w->dt_part[0];                         /* MPI_Allreduce(MIN) */
        for (int t = 1; t < w->nthreads; ++t) if (w->dt_part[t] < dt) dt = w->dt_part[t];
        pthread_barrier_wait(w->bar);
        w->rn_part[w->tid] = rk_step_part(&w->c, dt, w->Q, w->work, bar_wait, w->bar);
        pthread_barrier_wait(w->bar);
        rn = 0.0;                                   /* MPI_Allreduce(SUM) */
        for (int t = 0; t < w->nthreads; ++t) rn += w->rn_part[t];
    }
    w->dt_out = dt;
    w->rn_out = sqrt(rn);
    return NULL;
}

double orc_run_steps_threaded(const orc_params *p, int ncell, const int *cell_info,
                              const double *geom, int nface, const int *face_info, int dof,
                              int nsteps, double *Q, double *dt_last, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > ncell) nthreads = ncell > 0 ? ncell : 1;
    double *work = (double *)malloc(sizeof(double) * 5 * (size_t)dof);
    double *dt_part = (double *)calloc((size_t)nthreads, sizeof(double));
    double *rn_part = (double *)calloc((size_t)nthreads, sizeof(double));
    worker_arg *wa = (worker_arg *)calloc((size_t)nthreads, sizeof(worker_arg));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    int **fids = (int **)calloc((size_t)nthreads, sizeof(int *));
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, (unsigned)nthreads);

    /* weight-1 partition into contiguous quadrant ranges; each worker keeps the
     * faces that touch its range (its local + ghost-adjacent faces) */
    for (int t = 0; t < nthreads; ++t) {
        int c0 = (int)((long long)ncell * t / nthreads);
        int c1 = (int)((long long)ncell * (t + 1) / nthreads);
        int cnt = 0;
        fids[t] = (int *)malloc(sizeof(int) * (size_t)(nface > 0 ? nface : 1));
        for (int f = 0; f < nface; ++f) {
            const int *fi = &face_info[11 * f];
            int touch = (fi[2] >= c0 && fi[2] < c1);
            for (int k = 3; k < 7; ++k) if (fi[k] >= c0 && fi[k] < c1) touch = 1;
            if (touch) fids[t][cnt++] = f;
        }
        const int nftm = NF * p->nelem_subgrid * p->nelem_subgrid;
        rk_ctx c = {p, ncell, nface, dof, cell_info, face_info, geom, fids[t], cnt, c0, c1,
                    p->time_scheme, c0 * nftm, c1 * nftm, NULL, NULL};
        wa[t].c = c;
        wa[t].bar = &bar;
        wa[t].Q = Q;
        wa[t].work = work;
        wa[t].nsteps = nsteps;
        wa[t].tid = t;
        wa[t].nthreads = nthreads;
        wa[t].dt_part = dt_part;
        wa[t].rn_part = rn_part;
    }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, worker_main, &wa[t]);
    worker_main(&wa[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    double rn = wa[0].rn_out;
    if (dt_last) *dt_last = wa[0].dt_out;
    pthread_barrier_destroy(&bar);
    for (int t = 0; t < nthreads; ++t) free(fids[t]);
    free(fids); free(th); free(wa); free(rn_part); free(dt_part); free(work);
    return rn;
}

/* ------------------------------------------------------------------------- */
/* IC / refine / coarsen / tags                                              */
/* ------------------------------------------------------------------------- */
void orc_initial_condition_2d(int ic, int n, const double geom[12], double *Q,
                              double mach, double alpha, double beta, double gamma,
                              double density, double pressure,
                              double beta_L, double xo, double yo)
{
    /* src/solver/src/initial_condition_2d_module.F90:51-129 */
    const double pi = 3.141592653589793238462643383279502884197;
    const double twopi = 6.283185307179586476925286766559005768394;
    const double half = 0.5;
    double gm1 = gamma - 1.0;
    double oneOgm1 = 1.0 / gm1;
    double RT0 = pressure / density;
    double C_0 = sqrt(gamma * RT0);
    double V_0 = mach * C_0;
    double u_in = V_0 * cos(alpha * pi / 180.0) * cos(beta * pi / 180.0);
    double v_in = V_0 * sin(alpha * pi / 180.0) * cos(beta * pi / 180.0);
    double rhou = density * u_in;
    double rhov = density * v_in;
    double rhoe = (pressure * oneOgm1) + half * density * (u_in * u_in + v_in * v_in);
    if (ic == 0) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double *q = &Q[QIDX(0, i, j, n)];
                q[0] = density; q[1] = rhou; q[2] = rhov; q[3] = rhoe;
            }
        return;
    }
    /* isentropic vortex; time = 0 so the centre is (xo, yo) */
    double a = 0.5, two_a = 2.0 * a;
    double dT1 = -gm1 * beta_L * beta_L / (16.0 * a * gamma * pi * pi);
    double scale_factor = 2.0 / (double)n;
    double nat[2], node[2];
    for (int j = 1; j <= n; ++j) {
        double v1 = scale_factor * (double)(j - 1) - 1.0;
        double v2 = scale_factor * (double)j - 1.0;
        nat[1] = 0.5 * (v1 + v2);
        for (int i = 1; i <= n; ++i) {
            double w1 = scale_factor * (double)(i - 1) - 1.0;
            double w2 = scale_factor * (double)i - 1.0;
            double u, v, radius2, dT, rho, ff, p;
            nat[0] = 0.5 * (w1 + w2);
            orc_bilinear_node_2d(geom, nat, node);
            u = u_in; v = v_in;
            radius2 = (node[0] - xo) * (node[0] - xo) + (node[1] - yo) * (node[1] - yo);
            dT = dT1 * exp(two_a * (1.0 - radius2));
            rho = pow(1.0 + dT, oneOgm1);
            ff = beta_L / (twopi) * exp(a * (1.0 - radius2));
            u = u - (node[1] - yo) * ff;
            v = v + (node[0] - xo) * ff;
            p = pressure + pow(rho, gamma) - 1.0;
            double *q = &Q[QIDX(0, i - 1, j - 1, n)];
            q[0] = rho;
            q[1] = rho * u;
            q[2] = rho * v;
            q[3] = p * oneOgm1 + half * rho * (u * u + v * v);
        }
    }
}

void orc_refine_2d(int n, const double *parent, double *c1, double *c2, double *c3, double *c4)
{
    size_t m = (size_t)NF * n * n;
    memcpy(c1, parent, m * sizeof(double));
    memcpy(c2, parent, m * sizeof(double));
    memcpy(c3, parent, m * sizeof(double));
    memcpy(c4, parent, m * sizeof(double));
}

void orc_coarsen_2d(int n, const double *c1, const double *c2, const double *c3,
                    const double *c4, double *parent)
{
    int m = NF * n * n;
    for (int k = 0; k < m; ++k) parent[k] = c1[k] + c2[k] + c3[k] + c4[k];
    for (int k = 0; k < m; ++k) parent[k] = 0.25 * parent[k];
}

int orc_tag_feature(int ntags, const int *methods, const double *tol, int n, const double *Q)
{
    int tag = 0;
    for (int t = 0; t < ntags; ++t) {
        if (methods[t] == 1) {
            for (int c = 0; c < n * n; ++c) if (Q[NF * c] <= tol[t]) { tag = 1; break; }
        } else if (methods[t] == 2) {
            for (int c = 0; c < n * n; ++c) if (Q[NF * c] >= tol[t]) { tag = 1; break; }
        }
    }
    return tag;
}
