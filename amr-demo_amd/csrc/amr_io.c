/*
 * amr_io.c -- checkpoint and visualization output of the host forest in this
 * project's formats.  See include/amr_forest.h.
 *
 * The reference checkpoints with p4est_save (src/amr/src/amr_evolve.c:60-64)
 * and visualizes through p4est's VTK writer with, per quadrant, the tree,
 * level and rank as cell data and the patch-averaged rho, rhou, rhov, rhoe at
 * the four corners as point data (src/amr/src/amr_vtk.c:23-106, 110-200).
 * p4est is absent here, so both are restated in self-contained formats:
 *
 *   checkpoint: little-endian binary, header (magic "AMRCKPT1", format
 *   version, brick, patch size, levels, quadrant count, step, time), the
 *   quadrant array (tree, level, x, y as int32) and the patches
 *   (nquad x 4 n^2 doubles, reference layout Q(4,n,n)).
 *
 *   VTK: one XML UnstructuredGrid (.vtu) file with raw appended binary
 *   data; subcells = 0 writes one pixel per quadrant with the reference's
 *   fields, subcells = 1 one pixel per sub-grid cell with its own state.
 */
#define _POSIX_C_SOURCE 200809L
#include "amr_forest.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const char MAGIC[8] = {'A', 'M', 'R', 'C', 'K', 'P', 'T', '1'};
#define CKPT_VERSION 1

typedef struct {
    char magic[8];
    int32_t version;
    int32_t ntx, nty, periodic_x, periodic_y;
    int32_t n, min_level, max_level;
    int32_t nquad;
    int32_t istep;
    double xlo, ylo, scale;
    double time;
} ckpt_header_t;

int amr_forest_save(amr_forest_t *f, const char *path, int istep, double t)
{
    if (!f || !path) return -1;
    ckpt_header_t h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, MAGIC, 8);
    h.version = CKPT_VERSION;
    amr_forest_brick(f, &h.ntx, &h.nty, &h.periodic_x, &h.periodic_y, &h.xlo, &h.ylo, &h.scale, &h.min_level,
                     &h.max_level);
    h.n = amr_forest_nelem_subgrid(f);
    h.nquad = amr_forest_num_quads(f);
    h.istep = istep;
    h.time = t;
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    const size_t ps = (size_t)4 * h.n * h.n;
    int ok = fwrite(&h, sizeof(h), 1, fp) == 1;
    ok = ok && fwrite(amr_forest_quads(f), sizeof(amr_quad_t), (size_t)h.nquad, fp) == (size_t)h.nquad;
    ok = ok && fwrite(amr_forest_soln(f), sizeof(double) * ps, (size_t)h.nquad, fp) == (size_t)h.nquad;
    ok = (fclose(fp) == 0) && ok;
    return ok ? 0 : -1;
}

amr_forest_t *amr_forest_load(const char *path, int *istep, double *t)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return NULL;
    ckpt_header_t h;
    amr_forest_t *f = NULL;
    amr_quad_t *q = NULL;
    if (fread(&h, sizeof(h), 1, fp) != 1 || memcmp(h.magic, MAGIC, 8) || h.version != CKPT_VERSION ||
        h.nquad < 0 || h.n < 1)
        goto done;
    q = (amr_quad_t *)malloc(sizeof(amr_quad_t) * (size_t)(h.nquad > 0 ? h.nquad : 1));
    if (!q || fread(q, sizeof(amr_quad_t), (size_t)h.nquad, fp) != (size_t)h.nquad) goto done;
    f = amr_forest_from_quads(h.ntx, h.nty, h.periodic_x, h.periodic_y, h.xlo, h.ylo, h.scale, h.n, h.min_level,
                              h.max_level, h.nquad, q);
    if (!f) goto done;
    if (fread(amr_forest_soln(f), sizeof(double) * 4 * (size_t)h.n * h.n, (size_t)h.nquad, fp) !=
        (size_t)h.nquad) {
        amr_forest_destroy(f);
        f = NULL;
        goto done;
    }
    if (istep) *istep = h.istep;
    if (t) *t = h.time;
done:
    free(q);
    fclose(fp);
    return f;
}

/* raw appended block: uint32 byte count, then the bytes */
static int put_block(FILE *fp, const void *data, size_t bytes)
{
    uint32_t nb = (uint32_t)bytes;
    return fwrite(&nb, 4, 1, fp) == 1 && (bytes == 0 || fwrite(data, 1, bytes, fp) == bytes);
}

int amr_forest_write_vtk(amr_forest_t *f, const char *path, int subcells)
{
    if (!f || !path) return -1;
    const int nq = amr_forest_num_quads(f), n = amr_forest_nelem_subgrid(f);
    const int per = subcells ? n * n : 1;                 /* cells per quadrant */
    const size_t ncell = (size_t)nq * per, npt = 4 * ncell;
    const double *soln = amr_forest_soln(f);
    const amr_quad_t *quads = amr_forest_quads(f);
    double *pts = (double *)malloc(sizeof(double) * 3 * npt);
    int32_t *conn = (int32_t *)malloc(sizeof(int32_t) * npt);
    int32_t *offs = (int32_t *)malloc(sizeof(int32_t) * ncell);
    uint8_t *types = (uint8_t *)malloc(ncell);
    int32_t *tree = (int32_t *)malloc(sizeof(int32_t) * ncell);
    int32_t *level = (int32_t *)malloc(sizeof(int32_t) * ncell);
    double *fld = (double *)malloc(sizeof(double) * 4 * (subcells ? ncell : npt));
    int rc = -1;
    FILE *fp = NULL;
    if (!pts || !conn || !offs || !types || !tree || !level || !fld) goto done;
    for (int k = 0; k < nq; ++k) {
        double g[12];
        amr_forest_quad_geom(f, k, g);
        const double x0 = g[0], y0 = g[1], hx = (g[3] - g[0]) / (subcells ? n : 1);
        const double hy = (g[7] - g[1]) / (subcells ? n : 1);
        const double *Q = soln + (size_t)4 * n * n * k;
        double avg[4] = {0.0, 0.0, 0.0, 0.0};
        if (!subcells) {
            /* patch average as interpolate_solution (amr_vtk.c:141-170) */
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < 4; ++v) avg[v] += Q[4 * (i + n * j) + v];
            const double inv = 1.0 / (double)(n * n);
            for (int v = 0; v < 4; ++v) avg[v] *= inv;
        }
        for (int c = 0; c < per; ++c) {
            const int i = subcells ? c % n : 0, j = subcells ? c / n : 0;
            const size_t e = (size_t)k * per + (size_t)c;
            for (int corner = 0; corner < 4; ++corner) {   /* VTK_PIXEL: z-order corners */
                const size_t pidx = 4 * e + (size_t)corner;
                pts[3 * pidx + 0] = x0 + hx * (i + (corner & 1));
                pts[3 * pidx + 1] = y0 + hy * (j + (corner >> 1));
                pts[3 * pidx + 2] = 0.0;
                conn[pidx] = (int32_t)pidx;
                if (!subcells)
                    for (int v = 0; v < 4; ++v) fld[(size_t)v * npt + pidx] = avg[v];
            }
            if (subcells)
                for (int v = 0; v < 4; ++v) fld[(size_t)v * ncell + e] = Q[4 * c + v];
            offs[e] = (int32_t)(4 * (e + 1));
            types[e] = 8;   /* VTK_PIXEL */
            tree[e] = quads[k].tree;
            level[e] = quads[k].level;
        }
    }
    fp = fopen(path, "wb");
    if (!fp) goto done;
    {
        static const char *names[4] = {"rho", "rhou", "rhov", "rhoe"};
        size_t off = 0;
        const size_t b_pts = sizeof(double) * 3 * npt, b_conn = sizeof(int32_t) * npt;
        const size_t b_offs = sizeof(int32_t) * ncell, b_types = ncell, b_int = sizeof(int32_t) * ncell;
        const size_t b_fld = sizeof(double) * (subcells ? ncell : npt);
        fprintf(fp, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
                    "byte_order=\"LittleEndian\" header_type=\"UInt32\">\n<UnstructuredGrid>\n");
        fprintf(fp, "<Piece NumberOfPoints=\"%zu\" NumberOfCells=\"%zu\">\n", npt, ncell);
        fprintf(fp, "<Points><DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" "
                    "offset=\"%zu\"/></Points>\n", off);
        off += 4 + b_pts;
        fprintf(fp, "<Cells>\n<DataArray type=\"Int32\" Name=\"connectivity\" format=\"appended\" offset=\"%zu\"/>\n",
                off);
        off += 4 + b_conn;
        fprintf(fp, "<DataArray type=\"Int32\" Name=\"offsets\" format=\"appended\" offset=\"%zu\"/>\n", off);
        off += 4 + b_offs;
        fprintf(fp, "<DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"%zu\"/>\n</Cells>\n", off);
        off += 4 + b_types;
        fprintf(fp, "<CellData Scalars=\"level\">\n");
        fprintf(fp, "<DataArray type=\"Int32\" Name=\"treeid\" format=\"appended\" offset=\"%zu\"/>\n", off);
        off += 4 + b_int;
        fprintf(fp, "<DataArray type=\"Int32\" Name=\"level\" format=\"appended\" offset=\"%zu\"/>\n", off);
        off += 4 + b_int;
        if (subcells) {
            for (int v = 0; v < 4; ++v) {
                fprintf(fp, "<DataArray type=\"Float64\" Name=\"%s\" format=\"appended\" offset=\"%zu\"/>\n",
                        names[v], off);
                off += 4 + b_fld;
            }
        }
        fprintf(fp, "</CellData>\n");
        if (!subcells) {
            fprintf(fp, "<PointData>\n");
            for (int v = 0; v < 4; ++v) {
                fprintf(fp, "<DataArray type=\"Float64\" Name=\"%s\" format=\"appended\" offset=\"%zu\"/>\n",
                        names[v], off);
                off += 4 + b_fld;
            }
            fprintf(fp, "</PointData>\n");
        }
        fprintf(fp, "</Piece>\n</UnstructuredGrid>\n<AppendedData encoding=\"raw\">\n_");
        int ok = put_block(fp, pts, b_pts) && put_block(fp, conn, b_conn) && put_block(fp, offs, b_offs) &&
                 put_block(fp, types, b_types) && put_block(fp, tree, b_int) && put_block(fp, level, b_int);
        for (int v = 0; v < 4 && ok; ++v) ok = put_block(fp, fld + (size_t)v * (subcells ? ncell : npt), b_fld);
        fprintf(fp, "\n</AppendedData>\n</VTKFile>\n");
        rc = ok ? 0 : -1;
    }
done:
    if (fp && fclose(fp) != 0) rc = -1;
    free(pts); free(conn); free(offs); free(types); free(tree); free(level); free(fld);
    return rc;
}
