/*
 * psivox_adjoint.h -- the mass adjoint: dL/dpos of the mass field that
 * PSI.voxels(mode='density') deposits (the reference has no derivatives; the
 * forward is psi_voxels, src/cpsi.c:41-175, density branch :100-112).
 *
 * A tet T puts mass_T * vol(T n C) / vol(T) into cell C (the fragment's m0,
 * src/geometry.c:725-728).  Moving tet vertex v along axis a moves only the
 * faces of T n C that lie on the cell's a-planes, with normal speed
 * lambda_v (v's barycentric coordinate) there, so for w_C = dL/dm_C
 *
 *   dL/dx_{v,a} = mass_T / vol(T) * sum over grid planes x_a = c crossing T,
 *                 over the cells (j, k) of the plane:
 *                 (w_above - w_below) * int_{T n plane n cell(j,k)} lambda_v dA
 *
 * A cell outside the grid window has w = 0, so window planes enter like any
 * other.  The formula is exact: in the tet's mass coordinates q the tet is
 * the fixed unit simplex and only the cell planes x_a(q) = c move, the mass
 * fraction is 6 vol_q(T n C), its derivative is 6 times the signed integral
 * of lambda_v over the fragment's a-plane faces divided by |grad_q x_a|,
 * and dA_q / |grad_q x_a| = dA_x / (6 vol(T)).
 *
 * One task = (tet, periodic copy, axis, plane).  The section of the tet by
 * the plane is a triangle or a quadrilateral taken in FP64 straight from the
 * tet record; it is swept row by row through the plane's cells, each piece
 * clipped to its cell and integrated (area, first moment) on the fly in
 * cell-local coordinates; lambda comes from an FP64 inverse of the tet's edge
 * matrix.  Everything is FP64: the per-plane terms are large and cancel in
 * the sum, so their rounding has to be far below the gradient's.
 */
#ifndef PSIVOX_ADJOINT_H
#define PSIVOX_ADJOINT_H

#include "psivox_fast.h"
#include "psivox_geom.h"
#include "psivox_refine.h"

namespace pvx {

/* adjoint task word: tet (32 bits), copy (3), axis (2), plane (17) */
constexpr int ADJ_MAXPLANE = 1 << 17;

PVX_HD uint64_t pack_adj(int64_t t, int h, int ax, int p) {
    return (uint64_t)t | ((uint64_t)h << 32) | ((uint64_t)ax << 35) | ((uint64_t)p << 37);
}
PVX_HD void unpack_adj(uint64_t w, int64_t& t, int& h, int& ax, int& p) {
    t = (int64_t)(w & 0xffffffffull);
    h = (int)((w >> 32) & 7u);
    ax = (int)((w >> 35) & 3u);
    p = (int)((w >> 37) & 0x1ffffu);
}

/* planes p (index 0..n, plane p between cells p-1 and p) crossing the copy
 * with offset s along axis a: p0..p1 (empty when p1 < p0) */
PVX_HD void adj_planes(const GridInfo& g, int a, double lo, double hi, double s, int& p0, int& p1) {
    const double w0 = pick3(g.w0, a), d = pick3(g.d, a);
    const int n = pick3(g.n, a);
    const double fl = ceil((lo + s - w0) / d), fh = floor((hi + s - w0) / d);
    p0 = fl < 0.0 ? 0 : (fl > (double)n ? n + 1 : (int)fl);
    p1 = fh > (double)n ? n : (fh < 0.0 ? -1 : (int)fh);
}

/* the copy offsets of a tet (periodic ghosts, src/cpsi.c:343-373) */
PVX_HD int adj_copies(const double* rec, const GridInfo& g, const MeshBox& mb, double* lo, double* hi, uint32_t& shp) {
    aabb((const double (*)[3])rec, 4, lo, hi);
    return ghost_shifts_packed(lo, hi, g, mb, shp);
}

PVX_HD double copy_shift(const MeshBox& mb, uint32_t shp, int h, int a) {
    return mb.periodic ? (pick3(mb.hi, a) - pick3(mb.lo, a)) * (double)(1 & (shp >> (3 * h + a))) : 0.0;
}

/* Enumerate the adjoint tasks of one tet record (writes them when out is
 * given); returns the count, -1 for a record the adjoint cannot take. */
PVX_HD int64_t adj_tasks(const double* rec, int64_t t, const GridInfo& g, const MeshBox& mb, uint64_t* out) {
    bool fin = true;
    for (int q = 0; q < 12; ++q) fin = fin && (rec[q] - rec[q] == 0.0);
    if (!fin) return -1;
    double lo[3], hi[3];
    uint32_t shp;
    const int ng = adj_copies(rec, g, mb, lo, hi, shp);
    int64_t n = 0;
    for (int h = 0; h < ng; ++h)
        for (int a = 0; a < 3; ++a) {
            int p0, p1;
            adj_planes(g, a, pick3(lo, a), pick3(hi, a), copy_shift(mb, shp, h, a), p0, p1);
            for (int p = p0; p <= p1; ++p) {
                if (out) out[n] = pack_adj(t, h, a, p);
                n++;
            }
        }
    return n;
}

/* The cells a plane task's section spans: the rows of its index box in the
 * plane times their cells (adj_plane_task visits at most these), from the
 * points where the tet's edges cross the plane.  Only the processing order
 * depends on it (the sort key of the plane tasks). */
PVX_HD int adj_task_cells(const GridInfo& g, const MeshBox& mb, const double* rec, uint64_t task) {
    int64_t t;
    int h, a, p;
    unpack_adj(task, t, h, a, p);
    double lo[3], hi[3];
    uint32_t shp;
    adj_copies(rec, g, mb, lo, hi, shp);
    const int b = a == 2 ? 0 : a + 1, c = a == 0 ? 2 : a - 1;
    const double sa = copy_shift(mb, shp, h, a), sb = copy_shift(mb, shp, h, b), sc = copy_shift(mb, shp, h, c);
    const int na = pick3(g.n, a);
    const double xp = (p == na ? pick3(g.w1, a) : pick3(g.w0, a) + pick3(g.d, a) * (double)p) - sa;
    const double w0b = pick3(g.w0, b), w0c = pick3(g.w0, c), ib = pick3(g.dinv, b), ic = pick3(g.dinv, c);
    /* FP32 from here, with the hardware reciprocal: a cell more or less
     * only moves the task to a neighbouring class */
    float umin = 3e38f, umax = -3e38f, vmin = 3e38f, vmax = -3e38f;
    for (int u = 0; u < 3; ++u)
        for (int v = u + 1; v < 4; ++v) {
            const double* X = rec + 3 * u;
            const double* Y = rec + 3 * v;
            if ((X[a] < xp) == (Y[a] < xp)) continue;
            const float f = (float)(xp - X[a]) * frcp((float)(Y[a] - X[a]));
            const float su = (float)((X[b] + sb - w0b) * ib) + f * (float)((Y[b] - X[b]) * ib);
            const float sv = (float)((X[c] + sc - w0c) * ic) + f * (float)((Y[c] - X[c]) * ic);
            umin = fminf(umin, su);
            umax = fmaxf(umax, su);
            vmin = fminf(vmin, sv);
            vmax = fmaxf(vmax, sv);
        }
    if (umin > umax) return 0;
    const int nb = pick3(g.n, b), nc = pick3(g.n, c);
    const int j0 = umin > 0.f ? (int)floorf(umin) : 0, j1 = umax < (float)(nb - 1) ? (int)floorf(umax) : nb - 1;
    const int k0 = vmin > 0.f ? (int)floorf(vmin) : 0, k1 = vmax < (float)(nc - 1) ? (int)floorf(vmax) : nc - 1;
    return j1 < j0 || k1 < k0 ? 0 : (j1 - j0 + 1) * (k1 - k0 + 1);
}

/* the tet's FP64 inverse edge matrix: lambda_{1..3} = B (x - x0) */
struct TetInv64 {
    double x0[3], B[3][3], vol;
};

PVX_HD bool tet_inv64(const double* rec, TetInv64& T) {
    /* no fused multiply-adds here: a tet with a repeated vertex has two equal
     * columns of D, whose cofactor products cancel exactly only when each
     * product is rounded alike; a fused a*b - c*d leaves the rounding error
     * of c*d, a determinant of order 1e-20 instead of 0, and the tet would
     * not be dropped as the oracle drops it */
#pragma clang fp contract(off)
    double D[3][3];
    for (int a = 0; a < 3; ++a) {
        T.x0[a] = rec[a];
        for (int b = 0; b < 3; ++b) D[a][b] = rec[3 * (b + 1) + a] - rec[a];
    }
    const double c00 = D[1][1] * D[2][2] - D[1][2] * D[2][1], c01 = D[1][2] * D[2][0] - D[1][0] * D[2][2],
                 c02 = D[1][0] * D[2][1] - D[1][1] * D[2][0];
    const double det = D[0][0] * c00 + D[0][1] * c01 + D[0][2] * c02;
    T.vol = fabs(det) / 6.0;
    if (!(T.vol > 0.0)) return false;
    const double id = 1.0 / det;
    T.B[0][0] = c00 * id;
    T.B[1][0] = c01 * id;
    T.B[2][0] = c02 * id;
    T.B[0][1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) * id;
    T.B[1][1] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * id;
    T.B[2][1] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * id;
    T.B[0][2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) * id;
    T.B[1][2] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * id;
    T.B[2][2] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * id;
    return true;
}

/* a lane's row polygon (<= 8 vertex slots) in LDS (or host memory), stride FB */
template <int FB>
struct Poly2 {
    double* u;
    double* v;
    PVX_HD void set(int i, double a, double b) {
        u[i * FB] = a;
        v[i * FB] = b;
    }
};

/* integrates emitted polygon points: twice the area and twice the first
 * moments (shoelace), coordinates relative to the cell corner */
struct Shoelace {
    double a2 = 0.0, mu = 0.0, mv = 0.0;
    double fu = 0.0, fv = 0.0, pu = 0.0, pv = 0.0;
    int n = 0;
    PVX_HD void emit(double u, double v) {
        if (n == 0) {
            fu = u;
            fv = v;
        } else {
            const double cr = pu * v - u * pv;
            a2 += cr;
            mu += (pu + u) * cr;
            mv += (pv + v) * cr;
        }
        pu = u;
        pv = v;
        n++;
    }
    PVX_HD void close() {
        if (n > 2) emit(fu, fv);
    }
};

/* The moments int u^p v^q du dv (p + q <= 3) of a polygon from its
 * vertices in order, by Green's theorem edge by edge (each edge (u0, v0) ->
 * (u1, v1) contributes cr = u0 v1 - u1 v0 times a symmetric polynomial of
 * its ends), about the polygon's first vertex (u, v measured from it: a
 * small polygon far from the coordinates' origin would lose its area to
 * cancellation between the cross products).  m[0] M00, m[1] M10, m[2] M01,
 * m[3] M20, m[4] M11, m[5] M02, m[6] M30, m[7] M21, m[8] M12, m[9] M03. */
struct PolyMom3 {
    double s[10] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double fu = 0.0, fv = 0.0, pu = 0.0, pv = 0.0;
    int n = 0;
    PVX_HD void edge(double u0, double v0, double u1, double v1) {
        const double cr = u0 * v1 - u1 * v0;
        const double uu = u0 * u0 + u0 * u1 + u1 * u1, vv = v0 * v0 + v0 * v1 + v1 * v1;
        s[0] += cr;
        s[1] += cr * (u0 + u1);
        s[2] += cr * (v0 + v1);
        s[3] += cr * uu;
        s[4] += cr * (2.0 * u0 * v0 + u0 * v1 + u1 * v0 + 2.0 * u1 * v1);
        s[5] += cr * vv;
        s[6] += cr * (u0 + u1) * (u0 * u0 + u1 * u1);
        s[7] += cr * (3.0 * u0 * u0 * v0 + 2.0 * u0 * u1 * v0 + u1 * u1 * v0 + u0 * u0 * v1 + 2.0 * u0 * u1 * v1 +
                      3.0 * u1 * u1 * v1);
        s[8] += cr * (3.0 * v0 * v0 * u0 + 2.0 * v0 * v1 * u0 + v1 * v1 * u0 + v0 * v0 * u1 + 2.0 * v0 * v1 * u1 +
                      3.0 * v1 * v1 * u1);
        s[9] += cr * (v0 + v1) * (v0 * v0 + v1 * v1);
    }
    PVX_HD void emit(double u, double v) {
        if (n == 0) {
            fu = u;
            fv = v;
        } else {
            edge(pu, pv, u - fu, v - fv);
        }
        pu = u - fu;
        pv = v - fv;
        n++;
    }
    /* close the polygon and scale the sums to the moments (about the first
     * vertex, fu, fv) */
    PVX_HD void close(double* m) {
        if (n > 2) edge(pu, pv, 0.0, 0.0);
        const double f[10] = {1.0 / 2, 1.0 / 6, 1.0 / 6, 1.0 / 12, 1.0 / 24, 1.0 / 12, 1.0 / 20, 1.0 / 60, 1.0 / 60, 1.0 / 20};
        for (int i = 0; i < 10; ++i) m[i] = n > 2 ? s[i] * f[i] : 0.0;
    }
};

/* points of the edge P -> Q inside lo <= x_k <= hi, in order: P if inside,
 * then the crossings of the two lines met along the open edge */
template <class Emit>
PVX_HD void slab_edge(double pu, double pv, double qu, double qv, int k, double lo, double hi, Emit&& emit) {
    const double pk = k == 0 ? pu : pv, qk = k == 0 ? qu : qv;
    if (pk >= lo && pk <= hi) emit(pu, pv);
    double t0 = 2.0, t1 = 2.0;
    if ((pk < lo && qk > lo) || (pk > lo && qk < lo)) t0 = (lo - pk) / (qk - pk);
    if ((pk < hi && qk > hi) || (pk > hi && qk < hi)) t1 = (hi - pk) / (qk - pk);
    const double ta = t0 < t1 ? t0 : t1, tb = t0 < t1 ? t1 : t0;
    if (ta <= 1.0) {
        const double cu = pu + ta * (qu - pu), cv = pv + ta * (qv - pv);
        emit(k == 0 ? (ta == t0 ? lo : hi) : cu, k == 0 ? cv : (ta == t0 ? lo : hi));
    }
    if (tb <= 1.0) {
        const double cu = pu + tb * (qu - pu), cv = pv + tb * (qv - pv);
        emit(k == 0 ? (tb == t0 ? lo : hi) : cu, k == 0 ? cv : (tb == t0 ? lo : hi));
    }
}

/*
 * Per-cell weights of the adjoint: dL/d(raw sums) of the cell, w[ADJ_NW * c
 * ...]: M (the mass), then with velocity fields Sv (3) and Svv (3 x 3),
 * the mass-weighted sums of v and v v^T that the fields are finalized from
 * (m = M, v = Sv / M, vv = Svv / M - v v^T).  A tet's deposit into cell C
 * is (mass / vol) times the integral over T n C of
 *   g_C(x) = W_M + W_Sv . v(x) + v(x)^T W_Svv v(x),
 * v(x) the tet's linear velocity interpolation.
 */
constexpr int ADJ_NW = 13;

/* g_C at a point with velocity v (W: the cell's ADJ_NW weights) */
PVX_HD double adj_g(const double* W, const double* v) {
    double s = W[0] + W[1] * v[0] + W[2] * v[1] + W[3] * v[2];
    for (int a = 0; a < 3; ++a) s += v[a] * (W[4 + 3 * a] * v[0] + W[5 + 3 * a] * v[1] + W[6 + 3 * a] * v[2]);
    return s;
}

/* The point X (record frame) of a section piece: barycentric coordinates l
 * in the tet and the velocity there */
PVX_HD void adj_point(const TetInv64& T, const double* rec, const double* X, double* l, double* v) {
    double q[3];
    for (int r = 0; r < 3; ++r)
        q[r] = T.B[r][0] * (X[0] - T.x0[0]) + T.B[r][1] * (X[1] - T.x0[1]) + T.B[r][2] * (X[2] - T.x0[2]);
    l[0] = 1.0 - q[0] - q[1] - q[2];
    l[1] = q[0];
    l[2] = q[1];
    l[3] = q[2];
    for (int a = 0; a < 3; ++a)
        v[a] = l[0] * rec[12 + a] + l[1] * rec[15 + a] + l[2] * rec[18 + a] + l[3] * rec[21 + a];
}

/*
 * One adjoint task.  w: the cells' weights (grid layout n0 x n1 x n2; FULL:
 * ADJ_NW per cell, else dL/dm alone); R: a lane's row polygon buffer.
 * out[v] gets dL/dx_{v, axis} of this tet for the task's plane: the flux
 * of lambda_v (g_above - g_below) through each cell's piece of the section.
 * Without velocity fields g is constant per cell and the piece's centroid
 * gives the integral; with them the integrand lambda_v (g_above - g_below)
 * is a cubic polynomial on the plane (lambda and the velocity are affine
 * there), integrated exactly from the piece's polygon moments up to third
 * order.  Tets of zero volume contribute nothing.
 */
template <int FB, bool FULL = false>
PVX_HD void adj_plane_task(const GridInfo& g, const MeshBox& mb, const double* rec, uint64_t task,
                           const double* __restrict__ w, Poly2<FB> R, double* out) {
    int64_t t;
    int h, a, p;
    unpack_adj(task, t, h, a, p);
    out[0] = out[1] = out[2] = out[3] = 0.0;
    TetInv64 T;
    if (!tet_inv64(rec, T)) return;
    double lo[3], hi[3];
    uint32_t shp;
    adj_copies(rec, g, mb, lo, hi, shp);
    const int b = a == 2 ? 0 : a + 1, c = a == 0 ? 2 : a - 1;
    const double sa = copy_shift(mb, shp, h, a), sb = copy_shift(mb, shp, h, b), sc = copy_shift(mb, shp, h, c);
    const int na = pick3(g.n, a), nb = pick3(g.n, b), nc = pick3(g.n, c);
    const double w0b = pick3(g.w0, b), w0c = pick3(g.w0, c), db = pick3(g.d, b), dc = pick3(g.d, c);
    /* the plane in the record's frame */
    const double xp = (p == na ? pick3(g.w1, a) : pick3(g.w0, a) + pick3(g.d, a) * (double)p) - sa;
    /* section: crossing points of the edges between the two sides, in
     * cyclic order (1-3 split: the three edges at the lone vertex; 2-2
     * split: b1a1, b1a2, b2a2, b2a1) */
    /* vertices below the plane as a bit mask; the k-th set (clear) bit is
     * the k-th vertex below (above) in index order */
    unsigned bm = 0;
    for (int v = 0; v < 4; ++v) bm |= (rec[3 * v + a] < xp ? 1u : 0u) << v;
    const int nbl = __builtin_popcount(bm), nab = 4 - nbl;
    if (nbl == 0 || nab == 0) return;
    const unsigned am = ~bm & 15u;
    const int b0 = __builtin_ctz(bm), b1 = __builtin_ctz(bm & (bm - 1u) | 16u);
    const int a0 = __builtin_ctz(am), a1 = __builtin_ctz(am & (am - 1u) | 16u);
    int eu[4], ev[4], ne;
    if (nbl == 1 || nab == 1) {
        const int lone = nbl == 1 ? b0 : a0;
        /* the other three vertices in index order */
        const unsigned om = (nbl == 1 ? am : bm) & 15u, om1 = om & (om - 1u), om2 = om1 & (om1 - 1u);
        const int o0 = __builtin_ctz(om), o1 = __builtin_ctz(om1), o2 = __builtin_ctz(om2);
        eu[0] = eu[1] = eu[2] = eu[3] = lone;
        ev[0] = o0;
        ev[1] = o1;
        ev[2] = o2;
        ev[3] = o2;
        ne = 3;
    } else {
        eu[0] = b0, ev[0] = a0;
        eu[1] = b0, ev[1] = a1;
        eu[2] = b1, ev[2] = a1;
        eu[3] = b1, ev[3] = a0;
        ne = 4;
    }
    double su[4], sv[4];
    double umin = 1e300, umax = -1e300;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double* X = rec + 3 * eu[q];
        const double* Y = rec + 3 * ev[q];
        const double f = (xp - X[a]) / (Y[a] - X[a]);
        /* cell-local in-plane coordinates of the copy */
        su[q] = (X[b] + f * (Y[b] - X[b]) + sb - w0b) / db;
        sv[q] = (X[c] + f * (Y[c] - X[c]) + sc - w0c) / dc;
        if (q < ne) {
            umin = fmin(umin, su[q]);
            umax = fmax(umax, su[q]);
        }
    }
    int j0 = (int)floor(umin), j1 = (int)floor(umax);
    if (j0 < 0) j0 = 0;
    if (j1 > nb - 1) j1 = nb - 1;
    const int64_t Sa = a == 0 ? (int64_t)g.n[1] * g.n[2] : (a == 1 ? (int64_t)g.n[2] : 1);
    const int64_t Sb = b == 0 ? (int64_t)g.n[1] * g.n[2] : (b == 1 ? (int64_t)g.n[2] : 1);
    const int64_t Sc = c == 0 ? (int64_t)g.n[1] * g.n[2] : (c == 1 ? (int64_t)g.n[2] : 1);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int j = j0; j <= j1; ++j) {
        /* the row u in [j, j+1], stored relative to the row */
        int m = 0;
        double vmin = 1e300, vmax = -1e300;
        auto put = [&](double uu, double vv) {
            if (m < 8) R.set(m, uu - (double)j, vv);
            m++;
            vmin = fmin(vmin, vv);
            vmax = fmax(vmax, vv);
        };
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (q >= ne) break;
            /* the next crossing point, wrapping to the first after the last */
            const bool wrap = q + 1 == ne;
            const double nu = wrap ? su[0] : su[q < 3 ? q + 1 : 0], nv = wrap ? sv[0] : sv[q < 3 ? q + 1 : 0];
            slab_edge(su[q], sv[q], nu, nv, 0, (double)j, (double)(j + 1), put);
        }
        if (m < 3 || m > 8) continue; /* a quad cut by a slab has at most 6 */
        int k0 = (int)floor(vmin), k1 = (int)floor(vmax);
        if (k0 < 0) k0 = 0;
        if (k1 > nc - 1) k1 = nc - 1;
#pragma unroll 1
        for (int k = k0; k <= k1; ++k) {
            const int64_t gi = (int64_t)j * Sb + (int64_t)k * Sc;
            if (FULL) {
                /* G = g_above - g_below, the cells' weights differenced (none
                 * outside the grid) */
                const double* Wa = p < na ? w + ADJ_NW * (gi + (int64_t)p * Sa) : nullptr;
                const double* Wb = p > 0 ? w + ADJ_NW * (gi + (int64_t)(p - 1) * Sa) : nullptr;
                double Wd[ADJ_NW];
                bool any = false;
                for (int i = 0; i < ADJ_NW; ++i) {
                    Wd[i] = (Wa ? Wa[i] : 0.0) - (Wb ? Wb[i] : 0.0);
                    any = any || Wd[i] != 0.0;
                }
                if (!any) continue;
                /* the piece's moments in cell-local (u, v) */
                PolyMom3 PM;
                auto emitm = [&](double uu, double vv) { PM.emit(uu, vv - (double)k); };
                for (int q = 0; q < m; ++q) {
                    const int r = q + 1 == m ? 0 : q + 1;
                    slab_edge(R.u[q * FB], R.v[q * FB], R.u[r * FB], R.v[r * FB], 1, (double)k, (double)(k + 1),
                              emitm);
                }
                double M[10];
                PM.close(M);
                if (M[0] == 0.0) continue;
                if (M[0] < 0.0) /* the piece's orientation: moments of the area itself */
                    for (int i = 0; i < 10; ++i) M[i] = -M[i];
                /* lambda and the velocity as affine functions of (u, v) on the
                 * plane: value at the piece's first vertex (the moments'
                 * origin), rate along u and v */
                const double xb = w0b + db * ((double)j + PM.fu) - sb, xc = w0c + dc * ((double)k + PM.fv) - sc;
                const double X[3] = {a == 0 ? xp : (b == 0 ? xb : xc), a == 1 ? xp : (b == 1 ? xb : xc),
                                     a == 2 ? xp : (b == 2 ? xb : xc)};
                double l0[4], lu[4], lv[4];
                for (int r = 0; r < 3; ++r) {
                    l0[r + 1] = T.B[r][0] * (X[0] - T.x0[0]) + T.B[r][1] * (X[1] - T.x0[1]) +
                                T.B[r][2] * (X[2] - T.x0[2]);
                    lu[r + 1] = pick3(T.B[r], b) * db;
                    lv[r + 1] = pick3(T.B[r], c) * dc;
                }
                l0[0] = 1.0 - l0[1] - l0[2] - l0[3];
                lu[0] = -(lu[1] + lu[2] + lu[3]);
                lv[0] = -(lv[1] + lv[2] + lv[3]);
                double v0[3], vu[3], vw[3];
                for (int e = 0; e < 3; ++e) {
                    v0[e] = vu[e] = vw[e] = 0.0;
                    for (int i = 0; i < 4; ++i) {
                        const double ui = rec[12 + 3 * i + e];
                        v0[e] += l0[i] * ui;
                        vu[e] += lu[i] * ui;
                        vw[e] += lv[i] * ui;
                    }
                }
                /* G(u, v) = c00 + c10 u + c01 v + c20 u^2 + c11 u v + c02 v^2
                 * with vel = v0 + vu u + vw v: G = Wd0 + Wd1 . vel + vel^T Wd2 vel */
                double S0[3], Su[3], Sw[3]; /* Wd2 times v0, vu, vw, and its transpose times them */
                double T0[3], Tu[3];
                for (int e = 0; e < 3; ++e) {
                    S0[e] = Wd[4 + 3 * e] * v0[0] + Wd[5 + 3 * e] * v0[1] + Wd[6 + 3 * e] * v0[2];
                    Su[e] = Wd[4 + 3 * e] * vu[0] + Wd[5 + 3 * e] * vu[1] + Wd[6 + 3 * e] * vu[2];
                    Sw[e] = Wd[4 + 3 * e] * vw[0] + Wd[5 + 3 * e] * vw[1] + Wd[6 + 3 * e] * vw[2];
                    T0[e] = Wd[4 + e] * v0[0] + Wd[7 + e] * v0[1] + Wd[10 + e] * v0[2];
                    Tu[e] = Wd[4 + e] * vu[0] + Wd[7 + e] * vu[1] + Wd[10 + e] * vu[2];
                }
                const double c00 = Wd[0] + Wd[1] * v0[0] + Wd[2] * v0[1] + Wd[3] * v0[2] +
                                   v0[0] * S0[0] + v0[1] * S0[1] + v0[2] * S0[2];
                const double c10 = Wd[1] * vu[0] + Wd[2] * vu[1] + Wd[3] * vu[2] + vu[0] * S0[0] + vu[1] * S0[1] +
                                   vu[2] * S0[2] + vu[0] * T0[0] + vu[1] * T0[1] + vu[2] * T0[2];
                const double c01 = Wd[1] * vw[0] + Wd[2] * vw[1] + Wd[3] * vw[2] + vw[0] * S0[0] + vw[1] * S0[1] +
                                   vw[2] * S0[2] + vw[0] * T0[0] + vw[1] * T0[1] + vw[2] * T0[2];
                const double c20 = vu[0] * Su[0] + vu[1] * Su[1] + vu[2] * Su[2];
                const double c11 = vw[0] * Su[0] + vw[1] * Su[1] + vw[2] * Su[2] + vw[0] * Tu[0] + vw[1] * Tu[1] +
                                   vw[2] * Tu[2];
                const double c02 = vw[0] * Sw[0] + vw[1] * Sw[1] + vw[2] * Sw[2];
                /* int G, int u G, int v G over the piece */
                const double I0 = c00 * M[0] + c10 * M[1] + c01 * M[2] + c20 * M[3] + c11 * M[4] + c02 * M[5];
                const double Iu = c00 * M[1] + c10 * M[3] + c01 * M[4] + c20 * M[6] + c11 * M[7] + c02 * M[8];
                const double Iw = c00 * M[2] + c10 * M[4] + c01 * M[5] + c20 * M[7] + c11 * M[8] + c02 * M[9];
                const double area = db * dc;
                for (int r = 0; r < 4; ++r) acc[r] += area * (l0[r] * I0 + lu[r] * Iu + lv[r] * Iw);
                continue;
            }
            const double wa = p < na ? w[gi + (int64_t)p * Sa] : 0.0;
            const double wb = p > 0 ? w[gi + (int64_t)(p - 1) * Sa] : 0.0;
            const double J = wa - wb;
            if (J == 0.0) continue;
            Shoelace S;
            auto emit = [&](double uu, double vv) { S.emit(uu, vv - (double)k); };
            for (int q = 0; q < m; ++q) {
                const int r = q + 1 == m ? 0 : q + 1;
                slab_edge(R.u[q * FB], R.v[q * FB], R.u[r * FB], R.v[r * FB], 1, (double)k, (double)(k + 1), emit);
            }
            S.close();
            if (S.a2 == 0.0) continue;
            const double cu = S.mu / (3.0 * S.a2), cv = S.mv / (3.0 * S.a2);
            /* the section's centroid in the record's frame (axes a, b, c
             * are a cyclic permutation of x, y, z) */
            const double xb = w0b + db * ((double)j + cu) - sb, xc = w0c + dc * ((double)k + cv) - sc;
            const double X0 = a == 0 ? xp : (b == 0 ? xb : xc), X1 = a == 1 ? xp : (b == 1 ? xb : xc),
                         X2 = a == 2 ? xp : (b == 2 ? xb : xc);
            double l[3];
            for (int r = 0; r < 3; ++r)
                l[r] = T.B[r][0] * (X0 - T.x0[0]) + T.B[r][1] * (X1 - T.x0[1]) + T.B[r][2] * (X2 - T.x0[2]);
            const double f = J * 0.5 * fabs(S.a2) * db * dc;
            acc[0] += f * (1.0 - l[0] - l[1] - l[2]);
            acc[1] += f * l[0];
            acc[2] += f * l[1];
            acc[3] += f * l[2];
        }
    }
    const double s = rec[24] / T.vol;
    for (int v = 0; v < 4; ++v) out[v] = acc[v] * s;
}

/*
 * The velocity part of the adjoint for one column task (copy, column) of a
 * tet record: out[i][b] += dL/du_{i,b}, u_i the velocity of tet vertex i.
 * The fields Sv and Svv of cell C are sums over fragments of mass m0 E[v]
 * and mass m0 E[v v^T] (E: the mean over the fragment in mass coordinates,
 * v = sum_i lambda_i u_i), so
 *   dL/du_{i,b} = mass m0 (W_Sv,b E[lambda_i]
 *                          + sum_c (W_Svv,bc + W_Svv,cb) E[lambda_i v_c])
 * with E[lambda_i v_c] from the fragment's first and second q-moments.  The
 * moments come from the column path in precision R.  W: ADJ_NW weights per
 * cell.  False when the column overflows its vertex budget.
 */
template <class R>
PVX_HD bool adj_vel_task(const GridInfo& g, const MeshBox& mb, const double* rec, uint64_t task,
                         const double* __restrict__ W, LPoly<R>& P, double (*out)[3]) {
    ColumnState<R> S;
    const int st = column_begin(g, mb, rec, task, P, S);
    if (st < 0) return false;
    if (st == 0) return true;
    double u0[3], DV[3][3];
    for (int a = 0; a < 3; ++a) {
        u0[a] = rec[12 + a];
        for (int k = 0; k < 3; ++k) DV[a][k] = rec[12 + 3 * (k + 1) + a] - rec[12 + a];
    }
    const double mass = rec[24];
    /* one cell's fragment (moments M about O) into out */
    auto cell = [&](int k, const Mom<R>& M, const R* O) {
        if (!(M.m0 > R(0))) return;
        const double* Wc = W + ADJ_NW * (((int64_t)g.n[1] * S.i + S.j) * g.n[2] + k);
        /* fragment mean q and E[q q^T] (about q = 0) */
        const double m0 = (double)M.m0, inv = 1.0 / m0;
        double qb[3], E2[3][3];
        for (int a = 0; a < 3; ++a) qb[a] = (double)O[a] + 0.25 * (double)M.s1[a] * inv;
        const int SI[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) {
                /* second moment about O, shifted to q = 0 */
                const double mu_a = qb[a] - (double)O[a], mu_c = qb[c] - (double)O[c];
                const double cov = 0.05 * (double)M.s2[SI[a][c]] * inv - mu_a * mu_c;
                E2[a][c] = cov + qb[a] * qb[c];
            }
        /* E[lambda_i], E[lambda_i q_c] with lambda = (1 - sum q, q0, q1, q2) */
        double El[4], Elq[4][3];
        El[0] = 1.0 - qb[0] - qb[1] - qb[2];
        for (int c = 0; c < 3; ++c) {
            El[c + 1] = qb[c];
            Elq[0][c] = qb[c] - E2[0][c] - E2[1][c] - E2[2][c];
            for (int l = 0; l < 3; ++l) Elq[l + 1][c] = E2[l][c];
        }
        /* symmetric part of W_Svv, doubled */
        double Ws[3][3];
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) Ws[b][c] = Wc[4 + 3 * b + c] + Wc[4 + 3 * c + b];
        const double f = mass * m0;
        /* unrolled: the vertex index must be a constant, or El, Elq and out
         * are indexed at run time and live in scratch memory */
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            /* E[lambda_i v_c] */
            double Ev[3];
#pragma unroll
            for (int c = 0; c < 3; ++c)
                Ev[c] = u0[c] * El[i] + DV[c][0] * Elq[i][0] + DV[c][1] * Elq[i][1] + DV[c][2] * Elq[i][2];
#pragma unroll
            for (int b = 0; b < 3; ++b)
                out[i][b] += f * (Wc[1 + b] * El[i] + Ws[b][0] * Ev[0] + Ws[b][1] * Ev[1] + Ws[b][2] * Ev[2]);
        }
    };
    /* a two-cell column takes both cells in its first step (as the forward
     * does) and keeps the upper one's moments for the second; one call site
     * of the cell's body, so the kernel stays small */
    const bool two = two_cell_column(S);
    PlaneCarry<R> pc{false, false, {R(0), R(0), R(0)}};
    Mom<R> Mu;
    R O[3];
#pragma unroll 1
    for (int k = S.k0; k < S.k1; ++k) {
        Mom<R> M;
        if (two) {
            if (k == S.k0)
                slab_moments_2<2>(P, S.C.G, 2, first_plane(S), M, Mu, O);
            else
                M = Mu;
        } else {
            bool hl, hh;
            double lc, hc;
            cell_planes(S, k, hl, lc, hh, hc);
            slab_moments_u<2>(P, S.C.G, 2, hl, lc, hh, hc, M, O, &pc);
        }
        cell(k, M, O);
    }
    return true;
}

/* element-local node of vertex v of tet t at refinement level 0
 * (the 6-tet table on the 3x3x3 lattice, src/refine.c:39-41) */
PVX_HD int tet_node(int vpe, int t, int v) {
    if (vpe == 4) return v;
    const int L = tet27(t, v);
    if (vpe == 27) return L;
    const int i = L / 9, j = (L / 3) % 3, k = L % 3;
    return 4 * (i / 2) + 2 * (j / 2) + (k / 2);
}

}  // namespace pvx

#endif
