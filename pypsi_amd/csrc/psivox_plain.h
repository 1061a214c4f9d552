/*
 * psivox_plain.h -- FP64 voxelization of one (tet, periodic copy): window
 * clip, then a slab -> column -> voxel sweep of the clipped tet, reducing each
 * voxel fragment to its q-moments and depositing additive raw moments.
 *
 * Follows src/geometry.c:60-188 (deformation matrices, moment transform)
 * and :806-896 (window planes, index box); the bisection DFS of
 * src/geometry.c:852-896 is replaced by the sweep, which produces the same
 * fragments (tet intersected with each cell, points on a plane going to the
 * lower cell) through a different sequence of cuts.
 *
 * Raw moments per cell, with c the cell centre (see psivox.hip, finalize):
 *   M += w,  Sx += w (cm - c),  Sv += w cv,
 *   Sxx += w (sxx + (cm-c)(cm-c)^T),  Sxv += w (sxv + (cm-c) cv^T),
 *   Svv += w (svv + cv cv^T),   w = mass * m0,
 * where cm, cv, sxx, sxv, svv are the reference's per-fragment mean and
 * covariances (src/geometry.c:106-161).  These sums merge order-free; the
 * reference's online update (src/geometry.c:164-186) is the same merge
 * performed one fragment at a time.
 */
#ifndef PSIVOX_PLAIN_H
#define PSIVOX_PLAIN_H

#include "psivox_geom.h"

namespace pvx {

/* per-fragment transform + raw deposit; Sink::add(index, value) */
template <class Sink>
PVX_HD void deposit64(Sink& sink, const GridInfo& g, int i, int j, int k, const double* m, const double (*x)[3],
                      const double (*u)[3], const double DX[3][3], const double DV[3][3], double mass) {
    const int QI[3][3] = {{4, 5, 6}, {5, 7, 8}, {6, 8, 9}};
    const int fl = g.flags;
    double cm[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    if (fl & (F_X | F_XX | F_XV))
        for (int a = 0; a < 3; ++a) {
            cm[a] = m[0] * x[0][a];
            for (int b = 0; b < 3; ++b) cm[a] += m[b + 1] * DX[a][b];
            cm[a] /= m[0];
        }
    if (fl & (F_V | F_VV | F_XV))
        for (int a = 0; a < 3; ++a) {
            cv[a] = m[0] * u[0][a];
            for (int b = 0; b < 3; ++b) cv[a] += m[b + 1] * DV[a][b];
            cv[a] /= m[0];
        }
    const double w = mass * m[0];
    const double c[3] = {g.w0[0] + g.d[0] * (i + 0.5), g.w0[1] + g.d[1] * (j + 0.5), g.w0[2] + g.d[2] * (k + 0.5)};
    double dx[3];
    for (int a = 0; a < 3; ++a) dx[a] = cm[a] - c[a];
    const int64_t gi = ((int64_t)g.n[1] * g.n[2] * i + (int64_t)g.n[2] * j + k) * g.nacc;
    sink.add(gi, w);
    if (fl & F_X)
        for (int a = 0; a < 3; ++a) sink.add(gi + g.ox + a, w * dx[a]);
    if (fl & F_V)
        for (int a = 0; a < 3; ++a) sink.add(gi + g.ov + a, w * cv[a]);
    if (fl & F_XX) {
        int s = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b, ++s) {
                double t = m[0] * (x[0][a] - cm[a]) * (x[0][b] - cm[b]);
                for (int r = 0; r < 3; ++r) {
                    t += DX[a][r] * m[1 + r] * (x[0][b] - cm[b]);
                    t += DX[b][r] * m[1 + r] * (x[0][a] - cm[a]);
                }
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) t += DX[a][r] * DX[b][q] * m[QI[r][q]];
                t /= m[0];
                sink.add(gi + g.oxx + s, w * (t + dx[a] * dx[b]));
            }
    }
    if (fl & F_XV) {
        const bool fix = (fl & F_XVFIX) != 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double t = m[0] * (x[0][a] - cm[a]) * (u[0][b] - cv[b]);
                for (int r = 0; r < 3; ++r) {
                    if (fix) {
                        t += DX[a][r] * m[1 + r] * (u[0][b] - cv[b]);
                        t += DV[b][r] * m[1 + r] * (x[0][a] - cm[a]);
                    } else { /* reference formula (src/geometry.c:139-140) */
                        t += DX[a][r] * m[1 + r] * (x[0][b] - cm[b]);
                        t += DV[b][r] * m[1 + r] * (u[0][a] - cv[a]);
                    }
                }
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) t += DX[a][r] * DV[b][q] * m[QI[r][q]];
                t /= m[0];
                sink.add(gi + g.oxv + 3 * a + b, w * (t + dx[a] * cv[b]));
            }
    }
    if (fl & F_VV) {
        int s = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b, ++s) {
                double t = m[0] * (u[0][a] - cv[a]) * (u[0][b] - cv[b]);
                for (int r = 0; r < 3; ++r) {
                    t += DV[a][r] * m[1 + r] * (u[0][b] - cv[b]);
                    t += DV[b][r] * m[1 + r] * (u[0][a] - cv[a]);
                }
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) t += DV[a][r] * DV[b][q] * m[QI[r][q]];
                t /= m[0];
                sink.add(gi + g.ovv + s, w * (t + cv[a] * cv[b]));
            }
    }
}

struct PlainScratch {
    Poly64 pool[4];
};

/*
 * The window planes and index box of a polyhedron whose bounding box is
 * [lo, hi] (src/geometry.c:816-843): P is clipped to the window where the
 * box crosses it; [ib0, ib1) are its cells.  False when nothing is left.
 * (voxelize_copy64 below does the same steps inline.)
 */
PVX_HD bool window_prologue64(Poly64*& P, Poly64*& tmp, const GridInfo& g, const double* lo, const double* hi,
                              int* ib0, int* ib1, unsigned& status) {
    for (int a = 0; a < 3; ++a) {
        double cmin = lo[a], cmax = hi[a];
        bool clo = false, chi = false;
        if (g.w0[a] > cmin) {
            cmin = g.w0[a];
            clo = true;
        }
        if (g.w1[a] < cmax) {
            cmax = g.w1[a];
            chi = true;
        }
        if (clo) clip_lo64(P, tmp, cmin, a, status);
        if (chi) clip_hi64(P, tmp, cmax, a, status);
        index_range(g, a, cmin, cmax, ib0[a], ib1[a]);
        /* zero extent: one cell, as the reference's leaf rule */
        if (ib1[a] <= ib0[a]) ib1[a] = ib0[a] + 1;
        if (ib0[a] >= g.n[a]) return false;
    }
    return P->nv > 0;
}

/*
 * Sweep P through its cells [ib0, ib1): x slabs, then y columns, then z
 * cells, each cut off the remainder in turn; leaf(i, j, k, V) gets every
 * non-empty cell piece and returns 1 when it deposited a fragment.  The
 * pieces are the reference's DFS leaves (src/geometry.c:852-896).  Returns
 * the fragment count.
 */
template <class Leaf>
PVX_HD int64_t sweep_cells64(Poly64* P, Poly64* S, Poly64* C, Poly64* V, const GridInfo& g, const int* ib0,
                             const int* ib1, unsigned& status, Leaf&& leaf) {
    int64_t nfrag = 0;
    for (int i = ib0[0]; i < ib1[0]; ++i) {
        if (i < ib1[0] - 1) {
            split64(P, S, grid_plane(g, 0, i + 1), 0, status);
        } else {
            swap_ptr(P, S);
            P->nv = 0;
        }
        if (S->nv > 0) {
            int j0, j1;
            poly_range64(S, g, 1, ib0[1], ib1[1], j0, j1);
            for (int j = j0; j < j1; ++j) {
                if (j < j1 - 1) {
                    split64(S, C, grid_plane(g, 1, j + 1), 1, status);
                } else {
                    swap_ptr(S, C);
                    S->nv = 0;
                }
                if (C->nv > 0) {
                    int k0, k1;
                    poly_range64(C, g, 2, ib0[2], ib1[2], k0, k1);
                    for (int k = k0; k < k1; ++k) {
                        if (k < k1 - 1) {
                            split64(C, V, grid_plane(g, 2, k + 1), 2, status);
                        } else {
                            swap_ptr(C, V);
                            C->nv = 0;
                        }
                        if (V->nv > 0) nfrag += leaf(i, j, k, V);
                        if (C->nv <= 0) break;
                    }
                }
                if (S->nv <= 0) break;
            }
        }
        if (P->nv <= 0) break;
    }
    return nfrag;
}

/*
 * Voxelize one periodic copy: x = its 4 (shifted) positions, u velocities,
 * [lo, hi] its bounding box.  Returns the number of fragments deposited.
 */
template <class Sink>
PVX_HD int64_t voxelize_copy64(Sink& sink, const GridInfo& g, const double (*x)[3], const double (*u)[3],
                               double mass, const double* lo, const double* hi, PlainScratch& W, unsigned& status) {
    Poly64* P = &W.pool[0];
    Poly64* S = &W.pool[1];
    Poly64* C = &W.pool[2];
    Poly64* V = &W.pool[3];
    init_tet(P, x);
    /* window planes and index box (src/geometry.c:816-843) */
    int ib0[3], ib1[3];
    for (int a = 0; a < 3; ++a) {
        double cmin = lo[a], cmax = hi[a];
        bool clo = false, chi = false;
        if (g.w0[a] > cmin) {
            cmin = g.w0[a];
            clo = true;
        }
        if (g.w1[a] < cmax) {
            cmax = g.w1[a];
            chi = true;
        }
        if (clo) clip_lo64(P, V, cmin, a, status);
        if (chi) clip_hi64(P, V, cmax, a, status);
        index_range(g, a, cmin, cmax, ib0[a], ib1[a]);
        /* zero extent: one cell, as the reference's leaf rule */
        if (ib1[a] <= ib0[a]) ib1[a] = ib0[a] + 1;
        if (ib0[a] >= g.n[a]) return 0;
    }
    if (P->nv <= 0) return 0;

    double DX[3][3], DV[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            DX[a][b] = x[b + 1][a] - x[0][a];
            DV[a][b] = u[b + 1][a] - u[0][a];
        }
    const int order = g.order;
    double m[10];
    int64_t nfrag = 0;
    for (int i = ib0[0]; i < ib1[0]; ++i) {
        if (i < ib1[0] - 1) {
            split64(P, S, grid_plane(g, 0, i + 1), 0, status);
        } else {
            swap_ptr(P, S);
            P->nv = 0;
        }
        if (S->nv > 0) {
            int j0, j1;
            poly_range64(S, g, 1, ib0[1], ib1[1], j0, j1);
            for (int j = j0; j < j1; ++j) {
                if (j < j1 - 1) {
                    split64(S, C, grid_plane(g, 1, j + 1), 1, status);
                } else {
                    swap_ptr(S, C);
                    S->nv = 0;
                }
                if (C->nv > 0) {
                    int k0, k1;
                    poly_range64(C, g, 2, ib0[2], ib1[2], k0, k1);
                    for (int k = k0; k < k1; ++k) {
                        if (k < k1 - 1) {
                            split64(C, V, grid_plane(g, 2, k + 1), 2, status);
                        } else {
                            swap_ptr(C, V);
                            C->nv = 0;
                        }
                        if (V->nv > 0) {
                            reduce64(V, m, order);
                            if (m[0] > 0.0) {
                                deposit64(sink, g, i, j, k, m, x, u, DX, DV, mass);
                                nfrag++;
                            }
                        }
                        if (C->nv <= 0) break;
                    }
                }
                if (S->nv <= 0) break;
            }
        }
        if (P->nv <= 0) break;
    }
    return nfrag;
}

/* all periodic copies of one tet record (pos[12], vel[12], mass) */
template <class Sink>
PVX_HD int64_t voxelize_tet64(Sink& sink, const GridInfo& g, const MeshBox& mb, const double* rec, PlainScratch& W,
                              unsigned& status, int64_t* ncopies) {
    double x0[4][3], u[4][3];
    bool finite = true;
    for (int v = 0; v < 4; ++v)
        for (int a = 0; a < 3; ++a) {
            x0[v][a] = rec[3 * v + a];
            u[v][a] = rec[12 + 3 * v + a];
            finite = finite && (x0[v][a] - x0[v][a] == 0.0);
        }
    if (!finite) {
        status |= ST_BADGEOM;
        return 0;
    }
    double lo[3], hi[3];
    aabb((const double (*)[3])x0, 4, lo, hi);
    int sh[8];
    int ng = ghost_shifts(lo, hi, g, mb, sh);
    if (ncopies) *ncopies += ng;
    int64_t nf = 0;
    for (int h = 0; h < ng; ++h) {
        double x[4][3], glo[3], ghi[3];
        for (int a = 0; a < 3; ++a) {
            double s = (mb.hi[a] - mb.lo[a]) * (double)(1 & (sh[h] >> a));
            if (!mb.periodic) s = 0.0;
            for (int v = 0; v < 4; ++v) x[v][a] = x0[v][a] + s;
            glo[a] = lo[a] + s;
            ghi[a] = hi[a] + s;
        }
        nf += voxelize_copy64(sink, g, (const double (*)[3])x, (const double (*)[3])u, rec[24], glo, ghi, W, status);
    }
    return nf;
}

}  // namespace pvx

#endif
