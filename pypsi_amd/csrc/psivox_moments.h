/*
 * psivox_moments.h -- per-cell conversion between the reference's output
 * layout (mass, mass-weighted means, mass-weighted covariances;
 * src/geometry.c:164-186, src/PSI.c:119-186) and additive raw moments about
 * the cell centre c:
 *   A0 = M,  Ax = M (X - c),  Av = M V,
 *   Axx = M (XX + (X-c)(X-c)^T)  [6: xx xy xz yy yz zz],
 *   Axv = M (XV + (X-c) V^T)     [9],
 *   Avv = M (VV + V V^T)         [6].
 * The covariance part is the parallel merge identity the reference applies
 * one fragment at a time; centring x on the cell keeps the subtraction in
 * finalize well conditioned (fragments lie inside the cell).
 */
#ifndef PSIVOX_MOMENTS_H
#define PSIVOX_MOMENTS_H

#include "psivox_geom.h"

namespace pvx {

struct FieldPtrs {
    double* f[6]; /* m, x, v, xx, xv, vv (reference layout), null if absent */
};

PVX_HD void cell_centre(const GridInfo& g, int64_t c, double* cc) {
    const int64_t n12 = (int64_t)g.n[1] * g.n[2];
    const int i = (int)(c / n12), j = (int)((c / g.n[2]) % g.n[1]), k = (int)(c % g.n[2]);
    cc[0] = g.w0[0] + g.d[0] * (i + 0.5);
    cc[1] = g.w0[1] + g.d[1] * (j + 0.5);
    cc[2] = g.w0[2] + g.d[2] * (k + 0.5);
}

/* the raw moments of cell c's present contents into acc (ADD: added to
 * what acc holds for the cell) */
template <bool ADD>
PVX_HD void cell_raw(const GridInfo& g, const FieldPtrs& F, double* acc, int64_t c) {
    double* A = acc + c * g.nacc;
    auto put = [A](int t, double x) { A[t] = ADD ? A[t] + x : x; };
    const double M = F.f[0][c];
    if (M == 0.0) {
        if (!ADD)
            for (int t = 0; t < g.nacc; ++t) A[t] = 0.0;
        return;
    }
    double cc[3];
    cell_centre(g, c, cc);
    double dx[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    put(0, M);
    if (F.f[1])
        for (int a = 0; a < 3; ++a) {
            dx[a] = F.f[1][3 * c + a] - cc[a];
            put(g.ox + a, M * dx[a]);
        }
    if (F.f[2])
        for (int a = 0; a < 3; ++a) {
            v[a] = F.f[2][3 * c + a];
            put(g.ov + a, M * v[a]);
        }
    if (F.f[3]) {
        int s = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b, ++s) put(g.oxx + s, M * (F.f[3][9 * c + 3 * a + b] + dx[a] * dx[b]));
    }
    if (F.f[4])
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) put(g.oxv + 3 * a + b, M * (F.f[4][9 * c + 3 * a + b] + dx[a] * v[b]));
    if (F.f[5]) {
        int s = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b, ++s) put(g.ovv + s, M * (F.f[5][9 * c + 3 * a + b] + v[a] * v[b]));
    }
}

PVX_HD void cell_to_raw(const GridInfo& g, const FieldPtrs& F, double* acc, int64_t c) { cell_raw<false>(g, F, acc, c); }

PVX_HD void cell_finalize(const GridInfo& g, const double* acc, const FieldPtrs& F, int64_t c) {
    const double* A = acc + c * g.nacc;
    const double M = A[0];
    if (M == 0.0) return;
    F.f[0][c] = M;
    const double inv = 1.0 / M;
    double cc[3];
    cell_centre(g, c, cc);
    double dx[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    if (F.f[1])
        for (int a = 0; a < 3; ++a) {
            dx[a] = A[g.ox + a] * inv;
            F.f[1][3 * c + a] = cc[a] + dx[a];
        }
    if (F.f[2])
        for (int a = 0; a < 3; ++a) {
            v[a] = A[g.ov + a] * inv;
            F.f[2][3 * c + a] = v[a];
        }
    if (F.f[3]) {
        int s = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b, ++s) {
                double t = A[g.oxx + s] * inv - dx[a] * dx[b];
                F.f[3][9 * c + 3 * a + b] = t;
                F.f[3][9 * c + 3 * b + a] = t;
            }
    }
    if (F.f[4])
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) F.f[4][9 * c + 3 * a + b] = A[g.oxv + 3 * a + b] * inv - dx[a] * v[b];
    if (F.f[5]) {
        int s = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b, ++s) {
                double t = A[g.ovv + s] * inv - v[a] * v[b];
                F.f[5][9 * c + 3 * a + b] = t;
                F.f[5][9 * c + 3 * b + a] = t;
            }
    }
}

/* The merge done at the end instead of the start: acc holds only the new
 * deposits; the cell's present contents are added to them as raw moments
 * and the sum finalized.  The same sums as cell_to_raw before the deposit
 * and cell_finalize after it, in another order.  A cell that received no
 * mass keeps its contents bit for bit. */
PVX_HD void cell_merge_finalize(const GridInfo& g, double* acc, const FieldPtrs& F, int64_t c) {
    /* both loads first: one memory round trip decides the common cases
     * (nothing new; new mass into an empty cell) */
    const double a0 = acc[c * g.nacc], m_old = F.f[0][c];
    if (a0 == 0.0) return;
    if (m_old != 0.0) cell_raw<true>(g, F, acc, c);
    cell_finalize(g, acc, F, c);
}

}  // namespace pvx

#endif
