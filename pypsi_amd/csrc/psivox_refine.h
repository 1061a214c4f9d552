/*
 * psivox_refine.h -- element -> tets on the device (and on the host for
 * tests): periodic wrap of the element, then the reference's adaptive
 * refinement (src/refine.c:58-368), written as an explicit per-level DFS
 * so one thread handles one element without recursion.
 *
 * Tet record (TREC doubles): pos[4][3], vel[4][3], mass, valid, target, pad.
 * target is the exact (FP64) fraction of the tet's mass coordinates inside
 * the grid window summed over its periodic copies: 1 when the window holds
 * every copy, or when the window is the whole periodic box.
 */
#ifndef PSIVOX_REFINE_H
#define PSIVOX_REFINE_H

#include "psivox_geom.h"

namespace pvx {

constexpr int TREC = 28;
constexpr int MAXLVL = 6;

/* corners of the 6-tet split of the 3x3x3 lattice, index 9i+3j+k
 * (src/refine.c:39-41) */
PVX_HD int tet27(int t, int v) {
    const int T[6][4] = {{2, 0, 6, 18}, {8, 2, 6, 18}, {8, 20, 2, 18}, {8, 24, 20, 18}, {8, 6, 24, 18}, {8, 26, 20, 24}};
    return T[t][v];
}

/* 2x2x2 corners (4i+2j+k) to the 3x3x3 lattice by midpoints along x, then y,
 * then z (src/refine.c:292-319) */
PVX_HD void lin27(const double (*in)[3], double (*out)[3]) {
#pragma clang fp contract(off)
    for (int a = 0; a < 3; ++a) {
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) out[18 * i + 6 * j + 2 * k][a] = in[4 * i + 2 * j + k][a];
        for (int j = 0; j < 3; j += 2)
            for (int k = 0; k < 3; k += 2) out[9 + 3 * j + k][a] = 0.5 * (out[3 * j + k][a] + out[18 + 3 * j + k][a]);
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; k += 2)
                out[9 * i + 3 + k][a] = 0.5 * (out[9 * i + k][a] + out[9 * i + 6 + k][a]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                out[9 * i + 3 * j + 1][a] = 0.5 * (out[9 * i + 3 * j][a] + out[9 * i + 3 * j + 2][a]);
    }
}

/* 27-node quadratic element to the 5x5x5 lattice with the quarter-point
 * Lagrange weights (3, 6, -1)/8, separably along x, y, z
 * (src/refine.c:321-368) */
PVX_HD void quad125(const double (*in)[3], double (*out)[3]) {
#pragma clang fp contract(off)
#define PVX_I5(i, j, k) (25 * (i) + 5 * (j) + (k))
    for (int a = 0; a < 3; ++a) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) out[PVX_I5(2 * i, 2 * j, 2 * k)][a] = in[9 * i + 3 * j + k][a];
        for (int j = 0; j < 5; j += 2)
            for (int k = 0; k < 5; k += 2) {
                out[PVX_I5(1, j, k)][a] = 0.125 * (3.0 * out[PVX_I5(0, j, k)][a] + 6.0 * out[PVX_I5(2, j, k)][a] -
                                                   out[PVX_I5(4, j, k)][a]);
                out[PVX_I5(3, j, k)][a] = 0.125 * (3.0 * out[PVX_I5(4, j, k)][a] + 6.0 * out[PVX_I5(2, j, k)][a] -
                                                   out[PVX_I5(0, j, k)][a]);
            }
        for (int i = 0; i < 5; ++i)
            for (int k = 0; k < 5; k += 2) {
                out[PVX_I5(i, 1, k)][a] = 0.125 * (3.0 * out[PVX_I5(i, 0, k)][a] + 6.0 * out[PVX_I5(i, 2, k)][a] -
                                                   out[PVX_I5(i, 4, k)][a]);
                out[PVX_I5(i, 3, k)][a] = 0.125 * (3.0 * out[PVX_I5(i, 4, k)][a] + 6.0 * out[PVX_I5(i, 2, k)][a] -
                                                   out[PVX_I5(i, 0, k)][a]);
            }
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) {
                out[PVX_I5(i, j, 1)][a] = 0.125 * (3.0 * out[PVX_I5(i, j, 0)][a] + 6.0 * out[PVX_I5(i, j, 2)][a] -
                                                   out[PVX_I5(i, j, 4)][a]);
                out[PVX_I5(i, j, 3)][a] = 0.125 * (3.0 * out[PVX_I5(i, j, 4)][a] + 6.0 * out[PVX_I5(i, j, 2)][a] -
                                                   out[PVX_I5(i, j, 0)][a]);
            }
    }
#undef PVX_I5
}

/* does this lattice node need refining? (src/refine.c:245-254, 176-185) */
PVX_HD bool needs_refine(const double (*p)[3], bool quadratic, double tol2) {
#pragma clang fp contract(off)
    const int EL[7][3] = {{2, 4, 6}, {2, 10, 18}, {6, 12, 18}, {8, 13, 18}, {8, 14, 20}, {8, 16, 24}, {20, 22, 24}};
    const int EQ[19][3] = {{0, 1, 2},   {0, 3, 6},   {0, 9, 18},  {2, 4, 6},   {2, 5, 8},   {2, 10, 18}, {2, 11, 20},
                           {6, 7, 8},   {6, 12, 18}, {6, 15, 24}, {8, 13, 18}, {8, 14, 20}, {8, 16, 24}, {8, 17, 26},
                           {18, 19, 20}, {18, 21, 24}, {20, 22, 24}, {20, 23, 26}, {24, 25, 26}};
    const int ne = quadratic ? 19 : 7;
    for (int e = 0; e < ne; ++e) {
        const int* E = quadratic ? EQ[e] : EL[e];
        double err2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            double mt = 0.5 * (p[E[0]][a] + p[E[2]][a]);
            double ml = p[E[1]][a];
            err2 += (mt - ml) * (mt - ml);
        }
        if (err2 > tol2) return true;
    }
    return false;
}

/* the 8 corners (4i+2j+k) of child c (bits 4i+2j+k) of a linear node:
 * points of the node's 27-point lattice */
PVX_HD void child_corners(const double (*p)[3], const double (*u)[3], int c, double (*c8p)[3], double (*c8u)[3]) {
    const int ci = (c >> 2) & 1, cj = (c >> 1) & 1, ck = c & 1;
    for (int ii = 0; ii < 2; ++ii)
        for (int jj = 0; jj < 2; ++jj)
            for (int kk = 0; kk < 2; ++kk) {
                int src = 9 * (ii + ci) + 3 * (jj + cj) + (kk + ck);
                for (int a = 0; a < 3; ++a) {
                    c8p[4 * ii + 2 * jj + kk][a] = p[src][a];
                    c8u[4 * ii + 2 * jj + kk][a] = u[src][a];
                }
            }
}

/* child c (bits 4i+2j+k) of a lattice node: 27-point lattice of the child
 * (positions; velocities too when vel) */
PVX_HD void child_node(const double (*p)[3], const double (*u)[3], int c, bool quadratic, double (*cp)[3],
                       double (*cu)[3], double (*tmp)[3], bool vel = true) {
    const int ci = (c >> 2) & 1, cj = (c >> 1) & 1, ck = c & 1;
    if (!quadratic) {
        double c8p[8][3], c8u[8][3];
        child_corners(p, u, c, c8p, c8u);
        lin27((const double (*)[3])c8p, cp);
        if (vel) lin27((const double (*)[3])c8u, cu);
    } else {
        quad125(p, tmp);
        for (int ii = 0; ii < 3; ++ii)
            for (int jj = 0; jj < 3; ++jj)
                for (int kk = 0; kk < 3; ++kk)
                    for (int a = 0; a < 3; ++a)
                        cp[9 * ii + 3 * jj + kk][a] = tmp[25 * (ii + 2 * ci) + 5 * (jj + 2 * cj) + (kk + 2 * ck)][a];
        if (!vel) return;
        quad125(u, tmp);
        for (int ii = 0; ii < 3; ++ii)
            for (int jj = 0; jj < 3; ++jj)
                for (int kk = 0; kk < 3; ++kk)
                    for (int a = 0; a < 3; ++a)
                        cu[9 * ii + 3 * jj + kk][a] = tmp[25 * (ii + 2 * ci) + 5 * (jj + 2 * cj) + (kk + 2 * ck)][a];
    }
}

PVX_HD void emit6(double* out, const double (*p)[3], const double (*u)[3], double mass) {
    for (int t = 0; t < 6; ++t) {
        double* r = out + TREC * t;
        for (int v = 0; v < 4; ++v)
            for (int a = 0; a < 3; ++a) {
                r[3 * v + a] = p[tet27(t, v)][a];
                r[12 + 3 * v + a] = u[tet27(t, v)][a];
            }
        r[24] = mass * (1.0 / 6);
        r[25] = 1.0;
        r[26] = 1.0;
        r[27] = 0.0;
    }
}

/* emit6 straight from the 8 corners (4i+2j+k) of a linear element: the six
 * tets use only the 3x3x3 lattice's corners, which lin27 copies unchanged,
 * so this writes the same records without the 27-point arrays */
PVX_HD void emit6_corners(double* out, const double (*p)[3], const double (*u)[3], double mass) {
    for (int t = 0; t < 6; ++t) {
        double* r = out + TREC * t;
        for (int v = 0; v < 4; ++v) {
            /* lattice point l = 9i + 3j + k with i, j, k in {0, 2} */
            const int l = tet27(t, v), c = 4 * ((l / 9) / 2) + 2 * (((l / 3) % 3) / 2) + (l % 3) / 2;
            for (int a = 0; a < 3; ++a) {
                r[3 * v + a] = p[c][a];
                r[12 + 3 * v + a] = u[c][a];
            }
        }
        r[24] = mass * (1.0 / 6);
        r[25] = 1.0;
        r[26] = 1.0;
        r[27] = 0.0;
    }
}

/* FP64 q-volume fraction of the tet record inside the window, summed over
 * its periodic copies (src/cpsi.c:343-373 copies, src/geometry.c:816-843
 * window planes, src/geometry.c:692-767 volume) */
PVX_HD double window_target(const double* rec, const GridInfo& g, const MeshBox& mb, Poly64* pool) {
    double x0[4][3];
    for (int v = 0; v < 4; ++v)
        for (int a = 0; a < 3; ++a) x0[v][a] = rec[3 * v + a];
    double lo[3], hi[3];
    aabb((const double (*)[3])x0, 4, lo, hi);
    int sh[8];
    const int ng = ghost_shifts(lo, hi, g, mb, sh);
    if (mb.periodic && mb.full_window) return 1.0;
    double tot = 0.0;
    for (int h = 0; h < ng; ++h) {
        double x[4][3];
        bool inside = true;
        for (int a = 0; a < 3; ++a) {
            double s = mb.periodic ? (mb.hi[a] - mb.lo[a]) * (double)(1 & (sh[h] >> a)) : 0.0;
            for (int v = 0; v < 4; ++v) x[v][a] = x0[v][a] + s;
            if (lo[a] + s < g.w0[a] || hi[a] + s > g.w1[a]) inside = false;
        }
        if (inside) {
            tot += 1.0;
            continue;
        }
        Poly64* P = &pool[0];
        Poly64* T = &pool[1];
        unsigned st = 0;
        init_tet(P, (const double (*)[3])x);
        for (int a = 0; a < 3; ++a) {
            if (g.w0[a] > lo[a] + (mb.periodic ? (mb.hi[a] - mb.lo[a]) * (double)(1 & (sh[h] >> a)) : 0.0))
                clip_lo64(P, T, g.w0[a], a, st);
            if (g.w1[a] < hi[a] + (mb.periodic ? (mb.hi[a] - mb.lo[a]) * (double)(1 & (sh[h] >> a)) : 0.0))
                clip_hi64(P, T, g.w1[a], a, st);
        }
        if (P->nv > 0) {
            double m[10];
            reduce64(P, m, 0);
            tot += m[0];
        }
    }
    return tot;
}

/* Per-level DFS state of one element's refinement.  Scratch: L+1 lattice
 * nodes (27 pos + 27 vel each) and one 125-point temporary. */
struct RefineScratch {
    double p[MAXLVL + 1][27][3];
    double u[MAXLVL + 1][27][3];
    double tmp[125][3];
    int child[MAXLVL + 1];
    int org[MAXLVL + 1][3]; /* each level's node corner, in units of 1/PARAM_UNIT of the element */
};

/*
 * Where a refined tet's vertices sit in its element: every lattice point of
 * the refinement is the element's own map (trilinear over its 8 corners,
 * triquadratic over its 27 nodes) at a dyadic parameter point, since the
 * midpoints of lin27 and the quarter-point weights of quad125 are that map
 * at those points.  A tet's parameter record is 3 words: the element index,
 * then the 4 vertices' parameters (x | y << 8 | z << 16, each in units of
 * 1/PARAM_UNIT, two vertices per word).  The adjoint maps a vertex's
 * gradient onto the element's nodes with the map's weights there.
 */
constexpr int PARAM_UNIT = 2 << MAXLVL; /* 128: the lattice of the deepest level */
constexpr int PREC = 3;

PVX_HD uint32_t pack_param(int x, int y, int z) { return (uint32_t)x | ((uint32_t)y << 8) | ((uint32_t)z << 16); }

/* the parameter record of tet t of the 6-tet split of a 27-point lattice
 * with corner o and point spacing st (lattice point l = 9i + 3j + k) */
PVX_HD void emit6_params(uint64_t* prm, int64_t e, const int* o, int st) {
    for (int t = 0; t < 6; ++t) {
        uint32_t xi[4];
        for (int v = 0; v < 4; ++v) {
            const int l = tet27(t, v);
            xi[v] = pack_param(o[0] + st * (l / 9), o[1] + st * ((l / 3) % 3), o[2] + st * (l % 3));
        }
        prm[PREC * t] = (uint64_t)e;
        prm[PREC * t + 1] = (uint64_t)xi[0] | ((uint64_t)xi[1] << 32);
        prm[PREC * t + 2] = (uint64_t)xi[2] | ((uint64_t)xi[3] << 32);
    }
}

/* vertex v's parameters of a parameter record */
PVX_HD void unpack_param(const uint64_t* prm, int v, double* xi) {
    const uint32_t w = (uint32_t)(prm[1 + (v >> 1)] >> (32 * (v & 1)));
    xi[0] = (double)(w & 255u) / PARAM_UNIT;
    xi[1] = (double)((w >> 8) & 255u) / PARAM_UNIT;
    xi[2] = (double)((w >> 16) & 255u) / PARAM_UNIT;
}

/* the element map's one-axis factors at xi: node (i, j, k) weighs
 * F[0][i] F[1][j] F[2][k], linear (vpe 8: nodes at 0, 1) or quadratic (vpe
 * 27: nodes at 0, 1/2, 1); returns the nodes per axis */
PVX_HD int node_factors(int vpe, const double* xi, double (*F)[3]) {
    for (int a = 0; a < 3; ++a) {
        const double t = xi[a];
        if (vpe == 8) {
            F[a][0] = 1.0 - t;
            F[a][1] = t;
            F[a][2] = 0.0;
        } else {
            F[a][0] = 2.0 * (t - 0.5) * (t - 1.0);
            F[a][1] = 4.0 * t * (1.0 - t);
            F[a][2] = 2.0 * t * (t - 0.5);
        }
    }
    return vpe == 8 ? 2 : 3;
}

/* factor i of one axis's F (i in 0..2), without indexing at run time */
PVX_HD double pick_factor(const double* F, int i) { return i == 0 ? F[0] : (i == 1 ? F[1] : F[2]); }

/* the element map's weights at xi: trilinear over the corners 4i+2j+k
 * (vpe 8) or triquadratic over the nodes 9i+3j+k at 0, 1/2, 1 (vpe 27) */
PVX_HD int node_weights(int vpe, const double* xi, double* w) {
    double F[3][3];
    const int q = node_factors(vpe, xi, F);
    for (int n = 0; n < q * q * q; ++n) {
        const int i = n / (q * q), j = (n / q) % q, k = n % q;
        w[n] = pick_factor(F[0], i) * pick_factor(F[1], j) * pick_factor(F[2], k);
    }
    return q * q * q;
}

/*
 * Expand element e into tets.  If out is null only counts.  Returns the tet
 * count (0 if the element misses the window).  Corners in the mesh's
 * connectivity order; vpe in {4, 8, 27}.  prm (optional, with out): each
 * tet's parameter record (element index and its vertices' parameters).
 */
PVX_HD int64_t expand_element(const double* pos, const double* vel, const double* mass, const int32_t* conn,
                              int vpe, int64_t e, const GridInfo& g, const MeshBox& mb, int maxlvl, double tol2,
                              RefineScratch& S, double* out, uint64_t* prm = nullptr) {
    double P[27][3], U[27][3];
    for (int v = 0; v < vpe; ++v) {
        int64_t id = conn[e * vpe + v];
        for (int a = 0; a < 3; ++a) {
            P[v][a] = pos[3 * id + a];
            U[v][a] = vel ? vel[3 * id + a] : 0.0;
        }
    }
    if (!wrap_and_test(P, vpe, g, mb)) return 0;
    const double m = mass[e];
    if (vpe == 4) {
        if (out) {
            for (int v = 0; v < 4; ++v)
                for (int a = 0; a < 3; ++a) {
                    out[3 * v + a] = P[v][a];
                    out[12 + 3 * v + a] = U[v][a];
                }
            out[24] = m;
            out[25] = 1.0;
            out[26] = 1.0;
            out[27] = 0.0;
            if (prm) {
                /* a tet element is its own tet: vertex v is node v */
                prm[0] = (uint64_t)e;
                prm[1] = prm[2] = 0;
            }
        }
        return 1;
    }
    const bool quad = (vpe == 27);
    /* counting needs neither velocities nor the lattices of the last
     * level's nodes (their tets are counted, not placed) */
    const bool counting = out == nullptr;
    if (quad) {
        for (int v = 0; v < 27; ++v)
            for (int a = 0; a < 3; ++a) {
                S.p[0][v][a] = P[v][a];
                S.u[0][v][a] = U[v][a];
            }
    } else {
        lin27((const double (*)[3])P, S.p[0]);
        if (!counting) lin27((const double (*)[3])U, S.u[0]);
    }
    if (maxlvl > MAXLVL) maxlvl = MAXLVL;
    int64_t count = 0;
    int d = 0;
    double dm = m;
    S.org[0][0] = S.org[0][1] = S.org[0][2] = 0;
    /* visit node d */
    for (;;) {
        bool refine = d < maxlvl && needs_refine((const double (*)[3])S.p[d], quad, tol2);
        /* a refined node one level above the last: its 8 children are
         * leaves, counted without their lattices, and (linear elements)
         * emitted from their corners, which are points of this node's
         * lattice (lin27 copies corners unchanged: the same records) */
        const bool last = refine && d == maxlvl - 1 && (counting || !quad);
        if (refine && !last) {
            S.child[d] = 0;
        } else {
            if (out) {
                /* a node of level d spans PARAM_UNIT >> d; its lattice points
                 * are half that apart */
                const int half = (PARAM_UNIT / 2) >> d;
                if (last) {
                    for (int c = 0; c < 8; ++c) {
                        double c8p[8][3], c8u[8][3];
                        child_corners((const double (*)[3])S.p[d], (const double (*)[3])S.u[d], c, c8p, c8u);
                        emit6_corners(out + TREC * (count + 6 * c), (const double (*)[3])c8p,
                                      (const double (*)[3])c8u, dm * 0.125);
                        if (prm) {
                            const int oc[3] = {S.org[d][0] + half * ((c >> 2) & 1), S.org[d][1] + half * ((c >> 1) & 1),
                                               S.org[d][2] + half * (c & 1)};
                            emit6_params(prm + PREC * (count + 6 * c), e, oc, half / 2);
                        }
                    }
                } else {
                    emit6(out + TREC * count, (const double (*)[3])S.p[d], (const double (*)[3])S.u[d], dm);
                    if (prm) emit6_params(prm + PREC * count, e, S.org[d], half);
                }
            }
            count += last ? 48 : 6;
            /* ascend to the next unvisited sibling */
            for (;;) {
                if (d == 0) return count;
                d--;
                dm *= 8.0;
                if (S.child[d] < 8) break;
            }
        }
        /* descend into child S.child[d] */
        child_node((const double (*)[3])S.p[d], (const double (*)[3])S.u[d], S.child[d], quad, S.p[d + 1],
                   S.u[d + 1], S.tmp, !counting);
        {
            const int half = (PARAM_UNIT / 2) >> d, c = S.child[d];
            S.org[d + 1][0] = S.org[d][0] + half * ((c >> 2) & 1);
            S.org[d + 1][1] = S.org[d][1] + half * ((c >> 1) & 1);
            S.org[d + 1][2] = S.org[d][2] + half * (c & 1);
        }
        S.child[d]++;
        d++;
        dm *= 0.125;
    }
}

}  // namespace pvx

#endif
