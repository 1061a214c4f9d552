/*
 * psi_adjoint.c -- TEST INFRASTRUCTURE ONLY (part of the oracle).
 *
 * FP64 CPU reference for the mass adjoint of PSI.voxels(mode='density') and
 * the FP64 central finite differences that pin it.  Only tests/ use it.
 *
 * The reference has no adjoint; the quantity is defined by the forward map
 * restated in psi_oracle.c: the mass a tet T deposits in cell C is
 * mass_T * F(T, C), F = vol(T n C) / vol(T) (the m0 of the fragment,
 * src/geometry.c:725-728 with m0 = 6 vol in q-space).  For a loss L with
 * w_C = dL/dm_C:
 *
 *   dL/dx_{v,a} = sum_T mass_T / vol(T) * sum over grid planes x_a = c
 *                 crossing T, over cells (j, k) of that plane:
 *                 (w_above - w_below) * int_{T n plane n cell(j,k)} lambda_v dA
 *
 * where lambda_v is the barycentric coordinate of tet vertex v (moving v
 * moves the cell-plane faces of phi^-1(C) in the tet's mass coordinates;
 * the flux through a face is the face integral of lambda_v).  Cells outside
 * the grid window have w = 0, so the window planes enter like any other
 * plane.  Computed here independently of the GPU method: the cross-section
 * polygon of each tet copy with each plane, clipped to each 2D cell with a
 * Sutherland-Hodgman clip, shoelace area and centroid, barycentric solve.
 *
 * Mesh handling (wrap, window reject, periodic ghosts) is the forward's:
 * o_wrap_and_test, o_ghosts in psi_oracle.c (src/cpsi.c:309-373).  Only
 * refine_max_lvl = 0 (tet vertices are mesh vertices).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "psi_oracle.h"

/* shared with psi_oracle.c */
int orc_internal_wrap(double (*p)[3], int nv, double lo[3], double hi[3], const orc_grid* g, const orc_mesh* m);
int orc_internal_ghosts(double (*gp)[3], double (*gb)[3], const orc_grid* g, const orc_mesh* m);
int64_t orc_internal_tet_fragments(const orc_mesh* mesh, const orc_grid* g, const double (*x)[3], const double (*rb)[3],
                                   double mass, int64_t* gi, double* mm, int64_t cap);

/* the 6 tets of a trilinear element on the 3x3x3 lattice (src/refine.c:39-41) */
static const int A_TET27[6][4] = {{2, 0, 6, 18}, {8, 2, 6, 18}, {8, 20, 2, 18},
                                  {8, 24, 20, 18}, {8, 6, 24, 18}, {8, 26, 20, 24}};

/* element-local node of tet vertex v of tet t (refine level 0) */
static int a_tet_node(int vpe, int t, int v) {
    if (vpe == 4) return v;
    int L = A_TET27[t][v];
    if (vpe == 27) return L;
    int i = L / 9, j = (L / 3) % 3, k = L % 3;
    return 4 * (i / 2) + 2 * (j / 2) + (k / 2);
}

static double a_det3(const double (*M)[3]) {
    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

/* barycentric coordinates of point X in tet x[4] */
static void a_bary(const double (*x)[3], const double* X, double* lam) {
    double D[3][3], r[3];
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) D[a][b] = x[b + 1][a] - x[0][a];
        r[a] = X[a] - x[0][a];
    }
    double det = a_det3((const double (*)[3])D);
    double l[3];
    for (int c = 0; c < 3; ++c) { /* Cramer */
        double Dc[3][3];
        memcpy(Dc, D, sizeof D);
        for (int a = 0; a < 3; ++a) Dc[a][c] = r[a];
        l[c] = a_det3((const double (*)[3])Dc) / det;
    }
    lam[0] = 1.0 - l[0] - l[1] - l[2];
    lam[1] = l[0];
    lam[2] = l[1];
    lam[3] = l[2];
}

/* clip a 2D polygon by u_k >= c (sgn = +1) or u_k <= c (sgn = -1) */
static int a_clip2(const double (*in)[2], int n, double (*out)[2], int k, double c, double sgn) {
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const double* p = in[i];
        const double* q = in[(i + 1) % n];
        double sp = sgn * (p[k] - c), sq = sgn * (q[k] - c);
        if (sp >= 0) {
            out[m][0] = p[0];
            out[m][1] = p[1];
            m++;
        }
        if ((sp >= 0) != (sq >= 0)) {
            double t = sp / (sp - sq);
            out[m][0] = p[0] + t * (q[0] - p[0]);
            out[m][1] = p[1] + t * (q[1] - p[1]);
            out[m][k] = c;
            m++;
        }
    }
    return m;
}

/* gradient contribution of one tet copy x[4] (positions of the copy; the
 * shift is a constant) into g[4][3] */
static void a_tet_copy(const orc_grid* gr, const double* w, const double (*x)[3], double scale, double (*g)[3]) {
    for (int a = 0; a < 3; ++a) {
        const int b = (a + 1) % 3, c2 = (a + 2) % 3;
        double lo = x[0][a], hi = x[0][a];
        for (int v = 1; v < 4; ++v) {
            lo = fmin(lo, x[v][a]);
            hi = fmax(hi, x[v][a]);
        }
        int p0 = (int)ceil((lo - gr->win[0][a]) / gr->d[a]);
        int p1 = (int)floor((hi - gr->win[0][a]) / gr->d[a]);
        if (p0 < 0) p0 = 0;
        if (p1 > gr->n[a]) p1 = gr->n[a];
        for (int p = p0; p <= p1; ++p) {
            const double c = (p == gr->n[a]) ? gr->win[1][a] : gr->win[0][a] + gr->d[a] * p;
            /* cross-section: crossing points of the 6 edges */
            double poly[8][2];
            int np = 0;
            for (int u = 0; u < 4; ++u)
                for (int v = u + 1; v < 4; ++v) {
                    double su = x[u][a] - c, sv = x[v][a] - c;
                    if ((su < 0) == (sv < 0)) continue;
                    double t = su / (su - sv);
                    poly[np][0] = x[u][b] + t * (x[v][b] - x[u][b]);
                    poly[np][1] = x[u][c2] + t * (x[v][c2] - x[u][c2]);
                    np++;
                }
            if (np < 3) continue;
            /* order by angle about the mean */
            double mc[2] = {0, 0};
            for (int i = 0; i < np; ++i) {
                mc[0] += poly[i][0] / np;
                mc[1] += poly[i][1] / np;
            }
            double ang[8];
            for (int i = 0; i < np; ++i) ang[i] = atan2(poly[i][1] - mc[1], poly[i][0] - mc[0]);
            for (int i = 1; i < np; ++i)
                for (int j = i; j > 0 && ang[j] < ang[j - 1]; --j) {
                    double ta = ang[j];
                    ang[j] = ang[j - 1];
                    ang[j - 1] = ta;
                    double t0 = poly[j][0], t1 = poly[j][1];
                    poly[j][0] = poly[j - 1][0];
                    poly[j][1] = poly[j - 1][1];
                    poly[j - 1][0] = t0;
                    poly[j - 1][1] = t1;
                }
            double bl = poly[0][0], bh = poly[0][0], cl = poly[0][1], ch = poly[0][1];
            for (int i = 1; i < np; ++i) {
                bl = fmin(bl, poly[i][0]);
                bh = fmax(bh, poly[i][0]);
                cl = fmin(cl, poly[i][1]);
                ch = fmax(ch, poly[i][1]);
            }
            int j0 = (int)floor((bl - gr->win[0][b]) / gr->d[b]), j1 = (int)floor((bh - gr->win[0][b]) / gr->d[b]);
            int k0 = (int)floor((cl - gr->win[0][c2]) / gr->d[c2]), k1 = (int)floor((ch - gr->win[0][c2]) / gr->d[c2]);
            if (j0 < 0) j0 = 0;
            if (k0 < 0) k0 = 0;
            if (j1 > gr->n[b] - 1) j1 = gr->n[b] - 1;
            if (k1 > gr->n[c2] - 1) k1 = gr->n[c2] - 1;
            for (int j = j0; j <= j1; ++j)
                for (int k = k0; k <= k1; ++k) {
                    const double bj0 = gr->win[0][b] + gr->d[b] * j;
                    const double bj1 = (j + 1 == gr->n[b]) ? gr->win[1][b] : gr->win[0][b] + gr->d[b] * (j + 1);
                    const double ck0 = gr->win[0][c2] + gr->d[c2] * k;
                    const double ck1 = (k + 1 == gr->n[c2]) ? gr->win[1][c2] : gr->win[0][c2] + gr->d[c2] * (k + 1);
                    double t1[16][2], t2[16][2];
                    int m = a_clip2((const double (*)[2])poly, np, t1, 0, bj0, 1.0);
                    m = a_clip2((const double (*)[2])t1, m, t2, 0, bj1, -1.0);
                    m = a_clip2((const double (*)[2])t2, m, t1, 1, ck0, 1.0);
                    m = a_clip2((const double (*)[2])t1, m, t2, 1, ck1, -1.0);
                    if (m < 3) continue;
                    /* shoelace about the cell corner: absolute coordinates
                     * would cancel catastrophically on sliver pieces */
                    double A = 0, Mb = 0, Mc = 0;
                    for (int i = 0; i < m; ++i) {
                        const double P0 = t2[i][0] - bj0, P1 = t2[i][1] - ck0;
                        const double Q0 = t2[(i + 1) % m][0] - bj0, Q1 = t2[(i + 1) % m][1] - ck0;
                        double cr = P0 * Q1 - Q0 * P1;
                        A += cr;
                        Mb += (P0 + Q0) * cr;
                        Mc += (P1 + Q1) * cr;
                    }
                    if (A == 0.0) continue;
                    double X[3];
                    X[a] = c;
                    X[b] = bj0 + Mb / (3.0 * A);
                    X[c2] = ck0 + Mc / (3.0 * A);
                    A = 0.5 * fabs(A);
                    int64_t idx[3];
                    idx[b] = j;
                    idx[c2] = k;
                    double wa = 0.0, wb = 0.0;
                    if (p < gr->n[a]) {
                        idx[a] = p;
                        wa = w[(int64_t)gr->n[1] * gr->n[2] * idx[0] + (int64_t)gr->n[2] * idx[1] + idx[2]];
                    }
                    if (p > 0) {
                        idx[a] = p - 1;
                        wb = w[(int64_t)gr->n[1] * gr->n[2] * idx[0] + (int64_t)gr->n[2] * idx[1] + idx[2]];
                    }
                    double lam[4];
                    a_bary(x, X, lam);
                    for (int v = 0; v < 4; ++v) g[v][a] += scale * (wa - wb) * A * lam[v];
                }
        }
    }
}

/* dL/dpos (npart x 3, overwritten) from w = dL/dm (grid cells) for
 * refine_max_lvl = 0 */
int orc_mass_adjoint(const orc_mesh* mesh, const orc_grid* gr, const double* w, double* dpos) {
    memset(dpos, 0, sizeof(double) * 3 * (size_t)mesh->npart);
    const int vpe = mesh->elemtype;
    const int ntet = vpe == 4 ? 1 : 6;
    double p[27][3], lo[3], hi[3];
    for (int64_t e = 0; e < mesh->nelem; ++e) {
        for (int v = 0; v < vpe; ++v) {
            int64_t id = mesh->conn[(int64_t)e * vpe + v];
            for (int a = 0; a < 3; ++a) p[v][a] = mesh->pos[3 * id + a];
        }
        if (!orc_internal_wrap(p, vpe, lo, hi, gr, mesh)) continue;
        for (int t = 0; t < ntet; ++t) {
            double gp[32][3], gb[16][3];
            int node[4];
            for (int v = 0; v < 4; ++v) {
                node[v] = a_tet_node(vpe, t, v);
                for (int a = 0; a < 3; ++a) gp[v][a] = p[node[v]][a];
            }
            double D[3][3];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) D[a][b] = gp[b + 1][a] - gp[0][a];
            const double vol = fabs(a_det3((const double (*)[3])D)) / 6.0;
            if (!(vol > 0.0)) continue;
            const double mass = mesh->mass[e] / ntet;
            double bl[3], bh[3];
            for (int a = 0; a < 3; ++a) {
                bl[a] = bh[a] = gp[0][a];
                for (int v = 1; v < 4; ++v) {
                    bl[a] = fmin(bl[a], gp[v][a]);
                    bh[a] = fmax(bh[a], gp[v][a]);
                }
                gb[0][a] = bl[a];
                gb[1][a] = bh[a];
            }
            int ng = orc_internal_ghosts(gp, gb, gr, mesh);
            double g[4][3];
            memset(g, 0, sizeof g);
            for (int h = 0; h < ng; ++h) a_tet_copy(gr, w, (const double (*)[3])&gp[4 * h], mass / vol, g);
            for (int v = 0; v < 4; ++v) {
                int64_t id = mesh->conn[(int64_t)e * vpe + node[v]];
                for (int a = 0; a < 3; ++a) dpos[3 * id + a] += g[v][a];
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* central finite differences of L = sum_C (m_C / V_C - rho_t)^2        */
/* ------------------------------------------------------------------ */

typedef struct {
    int64_t cap;
    int64_t* key;
    double* val[3];
} a_map;

static double* a_slot(a_map* M, int64_t gi, int s) {
    uint64_t h = (uint64_t)gi * 0x9E3779B97F4A7C15ull;
    int64_t i = (int64_t)(h >> 20) & (M->cap - 1);
    while (M->key[i] != -1 && M->key[i] != gi) i = (i + 1) & (M->cap - 1);
    if (M->key[i] == -1) M->key[i] = gi;
    return &M->val[s][i];
}

/* mass of the listed elements into slot s of the map (same wrap, window,
 * ghosts and fragments as the forward, src/cpsi.c:41-175) */
static int a_local_mass(const orc_mesh* mesh, const orc_grid* gr, const int64_t* elems, int nel, a_map* M, int s) {
    const int vpe = mesh->elemtype;
    const int ntet = vpe == 4 ? 1 : 6;
    double p[27][3], lo[3], hi[3];
    int64_t gi[4096];
    double mm[4096];
    for (int q = 0; q < nel; ++q) {
        const int64_t e = elems[q];
        for (int v = 0; v < vpe; ++v) {
            int64_t id = mesh->conn[(int64_t)e * vpe + v];
            for (int a = 0; a < 3; ++a) p[v][a] = mesh->pos[3 * id + a];
        }
        if (!orc_internal_wrap(p, vpe, lo, hi, gr, mesh)) continue;
        for (int t = 0; t < ntet; ++t) {
            double gp[32][3], gb[16][3];
            for (int v = 0; v < 4; ++v)
                for (int a = 0; a < 3; ++a) gp[v][a] = p[a_tet_node(vpe, t, v)][a];
            for (int a = 0; a < 3; ++a) {
                gb[0][a] = gb[1][a] = gp[0][a];
                for (int v = 1; v < 4; ++v) {
                    gb[0][a] = fmin(gb[0][a], gp[v][a]);
                    gb[1][a] = fmax(gb[1][a], gp[v][a]);
                }
            }
            int ng = orc_internal_ghosts(gp, gb, gr, mesh);
            for (int h = 0; h < ng; ++h) {
                int64_t nf = orc_internal_tet_fragments(mesh, gr, (const double (*)[3])&gp[4 * h],
                                                        (const double (*)[3])&gb[2 * h], mesh->mass[e] / ntet, gi, mm,
                                                        4096);
                if (nf > 4096) return -1;
                for (int64_t f = 0; f < nf; ++f) *a_slot(M, gi[f], s) += mm[f];
            }
        }
    }
    return 0;
}

/* (L(x + h e) - L(x - h e)) / 2h for particle p, axis a, where the density
 * residual r = m / V - rho_t of the unperturbed mesh is given on the grid
 * and elems lists the elements that use particle p.  pos is modified and
 * restored. */
int orc_fd_mass_loss(orc_mesh* mesh, const orc_grid* gr, const double* r, const int64_t* elems, int nel, int64_t p,
                     int a, double h, double* out) {
    a_map M;
    M.cap = 1 << 14;
    M.key = (int64_t*)malloc(sizeof(int64_t) * M.cap);
    for (int s = 0; s < 3; ++s) M.val[s] = (double*)calloc((size_t)M.cap, sizeof(double));
    for (int64_t i = 0; i < M.cap; ++i) M.key[i] = -1;
    double* pos = (double*)mesh->pos;
    const double x0 = pos[3 * p + a];
    int rc = 0;
    pos[3 * p + a] = x0 + h;
    rc |= a_local_mass(mesh, gr, elems, nel, &M, 0);
    pos[3 * p + a] = x0;
    rc |= a_local_mass(mesh, gr, elems, nel, &M, 1);
    pos[3 * p + a] = x0 - h;
    rc |= a_local_mass(mesh, gr, elems, nel, &M, 2);
    pos[3 * p + a] = x0;
    const double V = gr->d[0] * gr->d[1] * gr->d[2];
    double dL = 0.0;
    for (int64_t i = 0; i < M.cap; ++i) {
        if (M.key[i] < 0) continue;
        const double dp = (M.val[0][i] - M.val[1][i]) / V, dm = (M.val[2][i] - M.val[1][i]) / V;
        const double rc0 = r[M.key[i]];
        dL += (dp - dm) * (2.0 * rc0 + dp + dm);
    }
    *out = dL / (2.0 * h);
    free(M.key);
    for (int s = 0; s < 3; ++s) free(M.val[s]);
    return rc;
}
