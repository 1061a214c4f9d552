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

/* ------------------------------------------------------------------ */
/* the adjoint of the fields m, v, vv                                   */
/* ------------------------------------------------------------------ */

/*
 * The fields of cell C are finalized from additive sums over the fragments
 * of every tet T in C: M = sum mass m0, Sv = sum mass m0 E[v] and Svv = sum
 * mass m0 E[v v^T] (m = M, v = Sv / M, vv = Svv / M - v v^T; E: the mean
 * over the fragment in mass coordinates, v the tet's linear velocity
 * interpolation; src/geometry.c:106-186 up to the order of summation).  W
 * holds per cell the weights dL/d(M, Sv, Svv): 13 values, W_M, W_Sv[3],
 * W_Svv[3][3].  Then
 *   dL/dx_{v,a} = sum_T mass_T / vol T * sum over planes x_a = c and cells
 *                 of the plane of int_{T n plane n cell} lambda_v (g_above
 *                 - g_below) dA,   g_C = W_M + W_Sv . v + v^T W_Svv v,
 * (as the mass adjoint above, the integrand now cubic: each piece is fanned
 * into triangles integrated with the 7-point rule exact to degree 5), and
 *   dL/du_{i,b} = sum over fragments of mass m0 (W_Sv,b E[lambda_i]
 *                 + sum_c (W_Svv,bc + W_Svv,cb) E[lambda_i v_c]),
 * E[lambda_i v_c] from the fragment's q-moments (o_reduce order 2).
 */
int64_t orc_internal_tet_moments(const orc_mesh* mesh, const orc_grid* g, const double (*x)[3], const double (*rb)[3],
                                 int64_t* gi, double* m10, int64_t cap);

static double a_g(const double* W, const double* v) {
    double s = W[0] + W[1] * v[0] + W[2] * v[1] + W[3] * v[2];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) s += W[4 + 3 * a + b] * v[a] * v[b];
    return s;
}

/* Dunavant's 7-point rule on a triangle, exact to degree 5: barycentric
 * points and weights (summing to 1) */
static const double A_D5[7][4] = {
    {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 0.225},
    {0.059715871789770, 0.470142064105115, 0.470142064105115, 0.132394152788506},
    {0.470142064105115, 0.059715871789770, 0.470142064105115, 0.132394152788506},
    {0.470142064105115, 0.470142064105115, 0.059715871789770, 0.132394152788506},
    {0.797426985353087, 0.101286507323456, 0.101286507323456, 0.125939180544827},
    {0.101286507323456, 0.797426985353087, 0.101286507323456, 0.125939180544827},
    {0.101286507323456, 0.101286507323456, 0.797426985353087, 0.125939180544827}};

/* position gradient of one tet copy (positions x, velocities u) into g */
static void a_tet_copy_full(const orc_grid* gr, const double* W, const double (*x)[3], const double (*u)[3],
                            double scale, double (*g)[3]) {
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
            double poly[8][2];
            int np = 0;
            for (int i = 0; i < 4; ++i)
                for (int j = i + 1; j < 4; ++j) {
                    double si = x[i][a] - c, sj = x[j][a] - c;
                    if ((si < 0) == (sj < 0)) continue;
                    double t = si / (si - sj);
                    poly[np][0] = x[i][b] + t * (x[j][b] - x[i][b]);
                    poly[np][1] = x[i][c2] + t * (x[j][c2] - x[i][c2]);
                    np++;
                }
            if (np < 3) continue;
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
                    int64_t idx[3];
                    idx[b] = j;
                    idx[c2] = k;
                    const double *Wa = NULL, *Wb = NULL;
                    if (p < gr->n[a]) {
                        idx[a] = p;
                        Wa = W + 13 * ((int64_t)gr->n[1] * gr->n[2] * idx[0] + (int64_t)gr->n[2] * idx[1] + idx[2]);
                    }
                    if (p > 0) {
                        idx[a] = p - 1;
                        Wb = W + 13 * ((int64_t)gr->n[1] * gr->n[2] * idx[0] + (int64_t)gr->n[2] * idx[1] + idx[2]);
                    }
                    /* fan from the first vertex, coordinates about the cell corner */
                    for (int i = 1; i + 1 < m; ++i) {
                        const double P0[2] = {t2[0][0] - bj0, t2[0][1] - ck0};
                        const double P1[2] = {t2[i][0] - bj0, t2[i][1] - ck0};
                        const double P2[2] = {t2[i + 1][0] - bj0, t2[i + 1][1] - ck0};
                        const double A = 0.5 * fabs((P1[0] - P0[0]) * (P2[1] - P0[1]) - (P2[0] - P0[0]) * (P1[1] - P0[1]));
                        if (A == 0.0) continue;
                        for (int q = 0; q < 7; ++q) {
                            double X[3], lam[4], v[3];
                            X[a] = c;
                            X[b] = bj0 + A_D5[q][0] * P0[0] + A_D5[q][1] * P1[0] + A_D5[q][2] * P2[0];
                            X[c2] = ck0 + A_D5[q][0] * P0[1] + A_D5[q][1] * P1[1] + A_D5[q][2] * P2[1];
                            a_bary(x, X, lam);
                            for (int e = 0; e < 3; ++e)
                                v[e] = lam[0] * u[0][e] + lam[1] * u[1][e] + lam[2] * u[2][e] + lam[3] * u[3][e];
                            const double f = A * A_D5[q][3] * ((Wa ? a_g(Wa, v) : 0.0) - (Wb ? a_g(Wb, v) : 0.0));
                            for (int vv = 0; vv < 4; ++vv) g[vv][a] += scale * f * lam[vv];
                        }
                    }
                }
        }
    }
}

/* velocity gradient of one tet copy from its fragments' q-moments */
static int a_tet_vel(const orc_mesh* mesh, const orc_grid* gr, const double* W, const double (*x)[3],
                     const double (*rb)[3], const double (*u)[3], double mass, double (*gu)[3]) {
    static int64_t gi[4096];
    static double mm[40960];
    int64_t nf = orc_internal_tet_moments(mesh, gr, x, rb, gi, mm, 4096);
    if (nf < 0 || nf > 4096) return -1;
    double DV[3][3];
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) DV[c][k] = u[k + 1][c] - u[0][c];
    for (int64_t f = 0; f < nf; ++f) {
        const double* m = mm + 10 * f;
        const double* Wc = W + 13 * gi[f];
        const double m0 = m[0];
        const double Eq[3] = {m[1] / m0, m[2] / m0, m[3] / m0};
        const double E2[3][3] = {{m[4] / m0, m[5] / m0, m[6] / m0},
                                 {m[5] / m0, m[7] / m0, m[8] / m0},
                                 {m[6] / m0, m[8] / m0, m[9] / m0}};
        double El[4], Elq[4][3];
        El[0] = 1.0 - Eq[0] - Eq[1] - Eq[2];
        for (int c = 0; c < 3; ++c) {
            El[c + 1] = Eq[c];
            Elq[0][c] = Eq[c] - E2[0][c] - E2[1][c] - E2[2][c];
            for (int l = 0; l < 3; ++l) Elq[l + 1][c] = E2[l][c];
        }
        for (int i = 0; i < 4; ++i) {
            double Ev[3];
            for (int c = 0; c < 3; ++c) {
                Ev[c] = u[0][c] * El[i];
                for (int k = 0; k < 3; ++k) Ev[c] += DV[c][k] * Elq[i][k];
            }
            for (int b = 0; b < 3; ++b) {
                double t = Wc[1 + b] * El[i];
                for (int c = 0; c < 3; ++c) t += (Wc[4 + 3 * b + c] + Wc[4 + 3 * c + b]) * Ev[c];
                gu[i][b] += mass * m0 * t;
            }
        }
    }
    return 0;
}

/* dpos, dvel (npart x 3 each, overwritten; either may be NULL) from the
 * cells' weights W (13 per cell), refine_max_lvl = 0 */
int orc_adjoint(const orc_mesh* mesh, const orc_grid* gr, const double* W, double* dpos, double* dvel) {
    if (dpos) memset(dpos, 0, sizeof(double) * 3 * (size_t)mesh->npart);
    if (dvel) memset(dvel, 0, sizeof(double) * 3 * (size_t)mesh->npart);
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
            double gp[32][3], gb[16][3], u[4][3];
            int node[4];
            for (int v = 0; v < 4; ++v) {
                node[v] = a_tet_node(vpe, t, v);
                const int64_t id = mesh->conn[(int64_t)e * vpe + node[v]];
                for (int a = 0; a < 3; ++a) {
                    gp[v][a] = p[node[v]][a];
                    u[v][a] = mesh->vel[3 * id + a];
                }
            }
            double D[3][3];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) D[a][b] = gp[b + 1][a] - gp[0][a];
            const double vol = fabs(a_det3((const double (*)[3])D)) / 6.0;
            const double mass = mesh->mass[e] / ntet;
            for (int a = 0; a < 3; ++a) {
                gb[0][a] = gb[1][a] = gp[0][a];
                for (int v = 1; v < 4; ++v) {
                    gb[0][a] = fmin(gb[0][a], gp[v][a]);
                    gb[1][a] = fmax(gb[1][a], gp[v][a]);
                }
            }
            int ng = orc_internal_ghosts(gp, gb, gr, mesh);
            double g[4][3], gu[4][3];
            memset(g, 0, sizeof g);
            memset(gu, 0, sizeof gu);
            for (int h = 0; h < ng; ++h) {
                if (dpos && vol > 0.0)
                    a_tet_copy_full(gr, W, (const double (*)[3])&gp[4 * h], (const double (*)[3])u, mass / vol, g);
                if (dvel &&
                    a_tet_vel(mesh, gr, W, (const double (*)[3])&gp[4 * h], (const double (*)[3])&gb[2 * h],
                              (const double (*)[3])u, mass, gu))
                    return -1;
            }
            for (int v = 0; v < 4; ++v) {
                int64_t id = mesh->conn[(int64_t)e * vpe + node[v]];
                for (int a = 0; a < 3; ++a) {
                    if (dpos) dpos[3 * id + a] += g[v][a];
                    if (dvel) dvel[3 * id + a] += gu[v][a];
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* central finite differences of a loss linear in the fields m, v, vv  */
/* ------------------------------------------------------------------ */

/* raw sums (M, Sv[3], Svv: xx xy xz yy yz zz) of the listed elements per
 * cell into slot s of the map (the forward's wrap, window, ghosts and
 * fragments) */
typedef struct {
    int64_t cap;
    int64_t* key;
    double* val; /* cap x 3 slots x 10 */
} a_map10;

static double* a_slot10(a_map10* M, int64_t gi, int s) {
    uint64_t h = (uint64_t)gi * 0x9E3779B97F4A7C15ull;
    int64_t i = (int64_t)(h >> 20) & (M->cap - 1);
    while (M->key[i] != -1 && M->key[i] != gi) i = (i + 1) & (M->cap - 1);
    if (M->key[i] == -1) M->key[i] = gi;
    return &M->val[(i * 3 + s) * 10];
}

static int a_local_raw(const orc_mesh* mesh, const orc_grid* gr, const int64_t* elems, int nel, a_map10* M, int s) {
    const int vpe = mesh->elemtype;
    const int ntet = vpe == 4 ? 1 : 6;
    double p[27][3], lo[3], hi[3];
    static int64_t gi[4096];
    static double mm[40960];
    for (int q = 0; q < nel; ++q) {
        const int64_t e = elems[q];
        for (int v = 0; v < vpe; ++v) {
            int64_t id = mesh->conn[(int64_t)e * vpe + v];
            for (int a = 0; a < 3; ++a) p[v][a] = mesh->pos[3 * id + a];
        }
        if (!orc_internal_wrap(p, vpe, lo, hi, gr, mesh)) continue;
        for (int t = 0; t < ntet; ++t) {
            double gp[32][3], gb[16][3], u[4][3];
            for (int v = 0; v < 4; ++v) {
                const int nd = a_tet_node(vpe, t, v);
                const int64_t id = mesh->conn[(int64_t)e * vpe + nd];
                for (int a = 0; a < 3; ++a) {
                    gp[v][a] = p[nd][a];
                    u[v][a] = mesh->vel[3 * id + a];
                }
            }
            for (int a = 0; a < 3; ++a) {
                gb[0][a] = gb[1][a] = gp[0][a];
                for (int v = 1; v < 4; ++v) {
                    gb[0][a] = fmin(gb[0][a], gp[v][a]);
                    gb[1][a] = fmax(gb[1][a], gp[v][a]);
                }
            }
            double DV[3][3];
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < 3; ++k) DV[c][k] = u[k + 1][c] - u[0][c];
            const double mass = mesh->mass[e] / ntet;
            int ng = orc_internal_ghosts(gp, gb, gr, mesh);
            for (int h = 0; h < ng; ++h) {
                int64_t nf = orc_internal_tet_moments(mesh, gr, (const double (*)[3])&gp[4 * h],
                                                      (const double (*)[3])&gb[2 * h], gi, mm, 4096);
                if (nf < 0 || nf > 4096) return -1;
                for (int64_t f = 0; f < nf; ++f) {
                    const double* m = mm + 10 * f;
                    /* first and second velocity moments of the fragment:
                     * v = u0 + DV q, int v = m0 u0 + DV m1, int v v^T */
                    double s1[3], M2[3][3] = {{m[4], m[5], m[6]}, {m[5], m[7], m[8]}, {m[6], m[8], m[9]}};
                    for (int c = 0; c < 3; ++c) s1[c] = DV[c][0] * m[1] + DV[c][1] * m[2] + DV[c][2] * m[3];
                    double* r = a_slot10(M, gi[f], s);
                    r[0] += mass * m[0];
                    for (int c = 0; c < 3; ++c) r[1 + c] += mass * (m[0] * u[0][c] + s1[c]);
                    const int SI[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
                    for (int z = 0; z < 6; ++z) {
                        const int c = SI[z][0], d = SI[z][1];
                        double t = m[0] * u[0][c] * u[0][d] + u[0][c] * s1[d] + s1[c] * u[0][d];
                        for (int k = 0; k < 3; ++k)
                            for (int l = 0; l < 3; ++l) t += DV[c][k] * M2[k][l] * DV[d][l];
                        r[4 + z] += mass * t;
                    }
                }
            }
        }
    }
    return 0;
}

/* the loss of one cell from its raw sums r[10] and coefficients G[13]
 * (G_m, G_v[3], G_vv[3][3]): G_m m + G_v . v + G_vv : vv */
static double a_cell_loss(const double* r, const double* G) {
    const double M = r[0];
    double L = G[0] * M;
    if (M == 0.0) return L;
    double v[3], S2[3][3];
    for (int c = 0; c < 3; ++c) v[c] = r[1 + c] / M;
    const int SI[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) S2[c][d] = r[4 + SI[c][d]] / M - v[c] * v[d];
    for (int c = 0; c < 3; ++c) {
        L += G[1 + c] * v[c];
        for (int d = 0; d < 3; ++d) L += G[4 + 3 * c + d] * S2[c][d];
    }
    return L;
}

/* (L(+h) - L(-h)) / 2h for particle p, axis a of its position (which = 0)
 * or velocity (which = 1); base: the unperturbed mesh's raw sums per cell
 * (10), G the loss coefficients per cell (13); elems lists the elements
 * that use particle p.  pos / vel are modified and restored. */
int orc_fd_fields_loss(orc_mesh* mesh, const orc_grid* gr, const double* base, const double* G, const int64_t* elems,
                       int nel, int64_t p, int a, int which, double h, double* out) {
    a_map10 M;
    M.cap = 1 << 14;
    M.key = (int64_t*)malloc(sizeof(int64_t) * M.cap);
    M.val = (double*)calloc((size_t)M.cap * 30, sizeof(double));
    for (int64_t i = 0; i < M.cap; ++i) M.key[i] = -1;
    double* arr = which ? (double*)mesh->vel : (double*)mesh->pos;
    const double x0 = arr[3 * p + a];
    int rc = 0;
    arr[3 * p + a] = x0 + h;
    rc |= a_local_raw(mesh, gr, elems, nel, &M, 0);
    arr[3 * p + a] = x0;
    rc |= a_local_raw(mesh, gr, elems, nel, &M, 1);
    arr[3 * p + a] = x0 - h;
    rc |= a_local_raw(mesh, gr, elems, nel, &M, 2);
    arr[3 * p + a] = x0;
    double dL = 0.0;
    for (int64_t i = 0; i < M.cap; ++i) {
        if (M.key[i] < 0) continue;
        const int64_t c = M.key[i];
        double rp[10], rm[10];
        for (int z = 0; z < 10; ++z) {
            const double b0 = base[10 * c + z] - M.val[(i * 3 + 1) * 10 + z];
            rp[z] = b0 + M.val[(i * 3 + 0) * 10 + z];
            rm[z] = b0 + M.val[(i * 3 + 2) * 10 + z];
        }
        dL += a_cell_loss(rp, G + 13 * c) - a_cell_loss(rm, G + 13 * c);
    }
    *out = dL / (2.0 * h);
    free(M.key);
    free(M.val);
    return rc;
}
