/*
 * psivox_annihilate.h -- PSI.voxels(mode='annihilation') for one pair of
 * tets, FP64, for host and device (src/cpsi.c:115-158,
 * src/geometry.c:190-252, 503-558).
 *
 * For tets t0 and t1 whose boxes overlap with positive volume, the mass
 * field gains rho0 rho1 vol(t0 ∩ t1 ∩ cell) in every cell, rho = mass / volume.
 * Summed over all ordered pairs of overlapping tets this is the integral of
 * rho^2 over each cell.
 */
#pragma once
#include "psivox_plain.h"

namespace pvx {

/* Signed volume of the tet p; a negative one gets vertices 0 and 1 swapped
 * and its volume returned positive (psi_orient_tet, src/geometry.c:190-217). */
PVX_HD double orient_tet(double (*p)[3]) {
    double e[3][3];
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) e[k][a] = p[k + 1][a] - p[0][a];
    const double det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                       e[1][0] * (e[0][1] * e[2][2] - e[0][2] * e[2][1]) +
                       e[2][0] * (e[0][1] * e[1][2] - e[0][2] * e[1][1]);
    double vol = det / 6.0;
    if (vol < 0.0) {
        for (int a = 0; a < 3; ++a) {
            const double t = p[0][a];
            p[0][a] = p[1][a];
            p[1][a] = t;
        }
        vol = -vol;
    }
    return vol;
}

/* The inward face planes n . x + d >= 0 of a positively oriented tet
 * (psi_tet_faces_from_verts, src/geometry.c:503-558): face f is opposite
 * vertex f, n = (c - a) x (b - a) over its vertices (a, b, c) taken in the
 * order that makes n point at vertex f, d = -n . (a + b + c) / 3. */
PVX_HD void tet_faces(const double (*p)[3], double (*n)[3], double* d) {
    for (int f = 0; f < 4; ++f) {
        /* (a, b, c) = (1,2,3), (0,3,2), (0,1,3), (0,2,1) */
        const int ia = f == 0 ? 1 : 0;
        const int ib = f == 0 ? 2 : (f == 1 ? 3 : (f == 2 ? 1 : 2));
        const int ic = f == 0 ? 3 : (f == 1 ? 2 : (f == 2 ? 3 : 1));
        double u[3], w[3], c[3];
        for (int a = 0; a < 3; ++a) {
            u[a] = p[ic][a] - p[ia][a];
            w[a] = p[ib][a] - p[ia][a];
            c[a] = (p[ia][a] + p[ib][a] + p[ic][a]) / 3.0;
        }
        n[f][0] = u[1] * w[2] - u[2] * w[1];
        n[f][1] = u[2] * w[0] - u[0] * w[2];
        n[f][2] = u[0] * w[1] - u[1] * w[0];
        d[f] = -(n[f][0] * c[0] + n[f][1] * c[1] + n[f][2] * c[2]);
    }
}

/* Keep the part of P with d + n . x >= 0 (psi_clip, src/geometry.c:391-501):
 * nothing changes when no vertex is below the plane, P empties when none is
 * above it. */
PVX_HD void clip_plane64(Poly64*& P, Poly64*& tmp, const double* n, double d, unsigned& status) {
    if (P->nv <= 0) return;
    double s[NV64], smax = -HUGE_VAL;
    for (int v = 0; v < P->nv; ++v) {
        s[v] = d + P->x[v][0] * n[0] + P->x[v][1] * n[1] + P->x[v][2] * n[2];
        smax = s[v] > smax ? s[v] : smax;
    }
    if (smax <= 0.0) {
        P->nv = 0;
        return;
    }
    cut64(P, tmp, s, status); /* P: s < 0 (dropped), tmp: kept */
    swap_ptr(P, tmp);
}

/* the strict overlap [lo, hi] of the two tets' boxes (psi_aabb_ixn,
 * src/cpsi.c:293-307): false when it has no volume */
PVX_HD bool tet_box_overlap(const double* ra, const double* rb, double* lo, double* hi) {
    double la[3], ha[3], lb[3], hb[3];
    aabb((const double (*)[3])ra, 4, la, ha);
    aabb((const double (*)[3])rb, 4, lb, hb);
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
        lo[a] = la[a] > lb[a] ? la[a] : lb[a];
        hi[a] = ha[a] < hb[a] ? ha[a] : hb[a];
        ok = ok && hi[a] > lo[a];
    }
    return ok;
}

/*
 * One pair: tet records ra, rb (TREC layout: positions 0..11, mass 24).
 * Adds rho_a rho_b vol(a ∩ b ∩ cell) to the mass field through
 * sink.add(cell index, value) for every cell the overlap reaches, within
 * the boxes' common box [lo, hi].  Returns the fragment count.
 */
template <class Sink>
PVX_HD int64_t annihilate_pair64(Sink& sink, const GridInfo& g, const double* ra, const double* rb, const double* lo,
                                 const double* hi, PlainScratch& W, unsigned& status) {
    double a[4][3], b[4][3];
    for (int v = 0; v < 4; ++v)
        for (int k = 0; k < 3; ++k) {
            a[v][k] = ra[3 * v + k];
            b[v][k] = rb[3 * v + k];
        }
    const double vol0 = orient_tet(a), vol1 = orient_tet(b);
    const double w = ra[24] / vol0 * (rb[24] / vol1) * vol0;
    Poly64* P = &W.pool[0];
    Poly64* V = &W.pool[3];
    init_tet(P, (const double (*)[3])a);
    double fn[4][3], fd[4];
    tet_faces((const double (*)[3])b, fn, fd);
    for (int f = 0; f < 4; ++f) clip_plane64(P, V, fn[f], fd[f], status);
    if (P->nv <= 0) return 0;
    int ib0[3], ib1[3];
    if (!window_prologue64(P, V, g, lo, hi, ib0, ib1, status)) return 0;
    return sweep_cells64(P, &W.pool[1], &W.pool[2], V, g, ib0, ib1, status,
                         [&](int i, int j, int k, const Poly64* piece) -> int {
                             double m[1];
                             reduce64(piece, m, 0);
                             if (!(m[0] > 0.0)) return 0;
                             sink.add(((int64_t)i * g.n[1] + j) * g.n[2] + k, w * m[0]);
                             return 1;
                         });
}

}  // namespace pvx
