/*
 * psivox_fast.h -- the column voxelization path: clipping in cell-local
 * coordinates, FP64 deposit, one lane per column task (a tet's periodic copy
 * restricted to one (x, y) column of its index box).
 *
 * A task clips the tet by the column's four planes into a polyhedron held in
 * the lane's slice of LDS (vertex graph with 3 neighbours per vertex as in
 * the reference, include/geometry.h:118-129, but storing only the mass
 * coordinate q: positions are x(q) = x0 + DX q in grid units relative to the
 * copy's index box).  Each cell of the column is then integrated without
 * building its fragment: the column's faces are fanned once as the
 * reference's reduce fans them (src/geometry.c:692-767), and every fan
 * triangle's piece inside the cell's z-slab is integrated by a fixed set of
 * weighted terms (no branches on the geometry, so the lanes of a wavefront
 * stay together).
 *
 * Everything geometric is templated on the real type R.  R = float is the
 * fast path; R = double is the same method in FP64, used for the few tets
 * wide enough (tet_wide) that FP32 cell-local coordinates would lose the
 * accuracy of their smallest fragments: an FP32 cut point carries an error
 * of about FP32 epsilon times the tet's extent, so a sliver fragment of a
 * tet spanning many cells gets a large relative error.
 *
 * Differences from the reference arithmetic (the method is the same):
 *   - cuts and moments in R on cell-local coordinates, plane offsets
 *     formed in FP64; in FP32 a vertex within 1e-5 cells of a plane is
 *     classified with an FP64 evaluation of its signed distance;
 *   - each column is clipped from the tet rather than from its neighbour
 *     slab, so shared faces of neighbouring columns agree only to R
 *     rounding; the per-tet mass residual (FP64 exact target minus FP64
 *     sum of the fragment masses) goes to its largest fragment when
 *     small, otherwise it is spread over the largest fragment of each of
 *     its columns; total mass is conserved to FP64 rounding;
 *   - deposits are the same additive FP64 raw moments as the FP64 kernel.
 */
#ifndef PSIVOX_FAST_H
#define PSIVOX_FAST_H

#include "psivox_geom.h"
#include <string.h>

#ifndef PVX_DBG
#define PVX_DBG(...)
#endif

namespace pvx {

constexpr int NVF = 16;   /* vertex slots of the column polyhedron (tet and 4 planes: <= 12) */
constexpr int FBLOCK = 64; /* lanes per workgroup (one wavefront) */

/* Words of one vertex record in LDS: q0, q1, q2 (as R), then the link word
 * (padded to 8 words for R = double).  The 64 lanes' copies of one slot are
 * consecutive, so one wide read fetches a vertex and neighbouring lanes
 * cover all 32 banks. */
template <class R>
struct VRec {
    static constexpr int W = sizeof(R) == 4 ? 4 : 8;       /* words per vertex */
    static constexpr int LINK = sizeof(R) == 4 ? 3 : 6;    /* word of the link */
    static constexpr int STRIDE = W * FBLOCK;              /* words between a lane's slots */
    static constexpr int WORDS = NVF * STRIDE;             /* words of a wavefront's polyhedra */
};
/* FP32 layout: words between a lane's consecutive vertex slots */
constexpr int VSTRIDE = VRec<float>::STRIDE;

/* Tets whose index box spans more than WIDE_CELLS cells along an axis go
 * through the FP64 column path under the automatic kernel choice: measured
 * on the host build, FP32 second moments stay within 1e-3 of the FP64
 * oracle on every heavy cell for tets up to about this extent. */
constexpr double WIDE_CELLS = 4.0;

/* Link word layout: neighbour k in bits [5k, 5k+5); the fan apex of the
 * face to the left of the directed edge (v -> neighbour k) in bits
 * [15+5k, 20+5k): one vertex of that face, the same for all its directed
 * edges (set by tet_start, kept by every clip).  31 marks an unset link. */
constexpr uint32_t NBUNSET = 31u;

PVX_HD int nbf(uint32_t w, int k) { return (int)((w >> (5 * k)) & 31u); }
PVX_HD int apf(uint32_t w, int k) { return (int)((w >> (15 + 5 * k)) & 31u); }
PVX_HD uint32_t with_nb(uint32_t w, int k, int v) { return (w & ~(31u << (5 * k))) | ((uint32_t)v << (5 * k)); }
PVX_HD uint32_t with_apex(uint32_t w, int k, int v) {
    return (w & ~(31u << (15 + 5 * k))) | ((uint32_t)v << (15 + 5 * k));
}
PVX_HD uint32_t pack3(int a, int b, int c) { return (uint32_t)a | ((uint32_t)b << 5) | ((uint32_t)c << 10); }
PVX_HD float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
PVX_HD uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* 1 / x: in FP32 the hardware reciprocal on the GPU (about 1 ulp, where a
 * correctly rounded division costs some ten instructions; the FP32 path's
 * tolerance is far coarser), an exact division on the host; in FP64 the
 * division */
PVX_HD float frcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
PVX_HD double frcp(double x) { return 1.0 / x; }

/* min, max, |x|, floor, ceil in the argument's precision */
PVX_HD float rmin(float a, float b) { return fminf(a, b); }
PVX_HD double rmin(double a, double b) { return fmin(a, b); }
PVX_HD float rmax(float a, float b) { return fmaxf(a, b); }
PVX_HD double rmax(double a, double b) { return fmax(a, b); }
PVX_HD float rabs(float a) { return fabsf(a); }
PVX_HD double rabs(double a) { return fabs(a); }
PVX_HD float rfloor(float a) { return floorf(a); }
PVX_HD double rfloor(double a) { return floor(a); }
PVX_HD float rceil(float a) { return ceilf(a); }
PVX_HD double rceil(double a) { return ceil(a); }

/* Keeps a value in a register from here on.  Values derived from the tet
 * record are otherwise re-read from memory where the compiler likes (the
 * record is read-only), also inside the cell loop, and vector memory
 * operations retire in order: such a load waits for every deposit atomic
 * issued before it. */
PVX_HD void pin(float& x) {
#ifdef __HIP_DEVICE_COMPILE__
    __asm__ volatile("" : "+v"(x));
#else
    (void)x;
#endif
}
PVX_HD void pin(double& x) {
#ifdef __HIP_DEVICE_COMPILE__
    __asm__ volatile("" : "+v"(x));
#else
    (void)x;
#endif
}

/* A lane's polyhedron: vertex v at w[v * STRIDE .. + W) */
template <class R>
struct LPoly {
    using L = VRec<R>;
    uint32_t* w;
    uint32_t live; /* bit v: slot v holds a vertex */
    int nv;        /* vertices */
    uint64_t ord;  /* their slots in order, 4 bits each (rank r in bits [4r, 4r+4)) */
    PVX_HD int at(int r) const { return (int)((ord >> (4 * r)) & 15u); }
    PVX_HD R qa(int v, int a) const {
        if (sizeof(R) == 4) return (R)u2f(w[v * L::STRIDE + a]);
        R x;
        memcpy(&x, &w[v * L::STRIDE + 2 * a], sizeof(R));
        return x;
    }
    PVX_HD void setq(int v, R a, R b, R c) {
        if (sizeof(R) == 4) {
            w[v * L::STRIDE] = f2u((float)a);
            w[v * L::STRIDE + 1] = f2u((float)b);
            w[v * L::STRIDE + 2] = f2u((float)c);
        } else {
            memcpy(&w[v * L::STRIDE], &a, sizeof(R));
            memcpy(&w[v * L::STRIDE + 2], &b, sizeof(R));
            memcpy(&w[v * L::STRIDE + 4], &c, sizeof(R));
        }
    }
    PVX_HD int nbr(int v, int k) const { return nbf(w[v * L::STRIDE + L::LINK], k); }
    PVX_HD uint32_t nbw(int v) const { return w[v * L::STRIDE + L::LINK]; }
    PVX_HD void setnbw(int v, uint32_t x) { w[v * L::STRIDE + L::LINK] = x; }
    PVX_HD void setnbr(int v, int k, int val) {
        uint32_t x = w[v * L::STRIDE + L::LINK];
        x = (x & ~(31u << (5 * k))) | ((uint32_t)val << (5 * k));
        w[v * L::STRIDE + L::LINK] = x;
    }
};

/* The copy being processed: geometry in grid units relative to its index
 * box corner in R; the FP64 tie-break re-reads the tet record. */
template <class R>
struct CopyGeo {
    R x0[3];
    double x0d[3]; /* x0 in FP64: plane offsets c - x0 are formed in FP64 */
    R DX[3][3];    /* DX[a][b] = x_{b+1,a} - x_{0,a} */
    const double* rec;
    double shift[3];
    int ib0[3];
};

/* one axis of the copy's affine map, x_ax(q) = o + d0 q0 + d1 q1 + d2 q2,
 * plus what the FP64 tie-break needs */
template <class R>
struct AxisMap {
    R o, d0, d1, d2;
    double od;
    int ax;
    double shift, w0, dinv;
    int ib0;
};

template <class R>
PVX_HD AxisMap<R> axis_map(const CopyGeo<R>& G, const GridInfo& g, int ax) {
    AxisMap<R> m;
    m.ax = ax;
    m.o = pick3(G.x0, ax);
    m.od = pick3(G.x0d, ax);
    m.d0 = ax == 0 ? G.DX[0][0] : (ax == 1 ? G.DX[1][0] : G.DX[2][0]);
    m.d1 = ax == 0 ? G.DX[0][1] : (ax == 1 ? G.DX[1][1] : G.DX[2][1]);
    m.d2 = ax == 0 ? G.DX[0][2] : (ax == 1 ? G.DX[1][2] : G.DX[2][2]);
    m.shift = pick3(G.shift, ax);
    m.w0 = pick3(g.w0, ax);
    m.dinv = pick3(g.dinv, ax);
    m.ib0 = pick3(G.ib0, ax);
    return m;
}

/* FP64 c - x_ax(q) from the record: positions shifted by the copy's
 * periodic offset, in grid units relative to the index box corner */
template <class R>
PVX_HD double sdist64(const double* r, const AxisMap<R>& m, double c, R q0, R q1, R q2) {
    const double ra = m.ax == 0 ? r[0] : (m.ax == 1 ? r[1] : r[2]);
    const double rb = m.ax == 0 ? r[3] : (m.ax == 1 ? r[4] : r[5]);
    const double rc = m.ax == 0 ? r[6] : (m.ax == 1 ? r[7] : r[8]);
    const double rd = m.ax == 0 ? r[9] : (m.ax == 1 ? r[10] : r[11]);
    const double xa = (ra + m.shift) + (rb - ra) * (double)q0 + (rc - ra) * (double)q1 + (rd - ra) * (double)q2;
    return c - ((xa - m.w0) * m.dinv - (double)m.ib0);
}

/* signed distance c - x_ax(q) in cells as off - DX_ax q with off = c - x0
 * formed in FP64 (the plane crosses the tet, so |off| is at most the tet's
 * extent and its rounding scales with the tet, not with the cell index);
 * in FP32 re-evaluated in FP64 from the record within 1e-5 cells of the
 * plane */
template <class R>
PVX_HD R sdist(const double* rec, const AxisMap<R>& m, double c, R off, R q0, R q1, R q2) {
    R s = off - (m.d0 * q0 + m.d1 * q1 + m.d2 * q2);
    if (sizeof(R) == 4 && rabs(s) < R(1e-5)) {
        double sd = sdist64(rec, m, c, q0, q1, q2);
        s = (sd < 0.0) ? rmin(R(-1e-30), (R)sd) : rmax(R(0), (R)sd);
    }
    return s;
}

/*
 * The graph part of a clip: drop the vertices in gone (a non-empty proper
 * subset of P's), put a new vertex into a free slot on each edge from a
 * survivor vc to a dropped vn (place(vx, vc, vn) sets its q), link the new
 * vertices around their faces, and keep the order: survivors in their order,
 * then the new vertices as they are made (the order a compacted graph has).
 * The graph is the reference's (src/geometry.c:391-501); nothing is
 * renumbered, the slots of dropped vertices are freed.
 * Fan apexes: the faces the clip leaves alone keep theirs; each cut face is
 * walked once from a new vertex (the walk that links it to the next new
 * vertex around the face) and fanned from that new vertex; the new face on
 * the plane is fanned from the first new vertex.  Any vertex of a convex
 * face fans it, so the moments are those of the reference's fans up to
 * rounding.  False on slot overflow or a broken graph.
 */
template <class R, class Place>
PVX_HD bool clip_graph(LPoly<R>& P, uint32_t gone, Place&& place) {
    const uint32_t live = P.live;
    uint64_t nord = 0;
    int nk = 0;
    for (int r = 0; r < P.nv; ++r) {
        const int v = P.at(r);
        if (!((gone >> v) & 1u)) nord |= (uint64_t)v << (4 * nk++);
    }
    int nt = nk;
    uint32_t avail = ~live & ((1u << NVF) - 1u), fresh = 0u;
    for (int r = 0; r < nk; ++r) {
        const int vc = (int)((nord >> (4 * r)) & 15u);
        for (int k = 0; k < 3; ++k) {
            const int vn = P.nbr(vc, k);
            if (vn >= NVF || !((live >> vn) & 1u)) {
                PVX_DBG("clip: bad nbr %d of %d\n", vn, vc);
                return false;
            }
            if (!((gone >> vn) & 1u)) continue;
            if (avail == 0u) {
                PVX_DBG("clip: overflow\n");
                return false;
            }
            const int vx = __builtin_ctz(avail);
            avail &= avail - 1u;
            place(vx, vc, vn);
            P.setnbw(vx, pack3(vc, NBUNSET, NBUNSET));
            P.setnbr(vc, k, vx);
            fresh |= 1u << vx;
            nord |= (uint64_t)vx << (4 * nt++);
        }
    }
    const uint32_t after = (live & ~gone) | fresh;
    /* new vertices take the lowest free slots in turn: the first one made */
    const int cap = fresh ? __builtin_ctz(fresh) : 0;
    /* link each new vertex vs to the next one around the face to the left
     * of vs -> vc, and fan that face from vs */
    for (uint32_t fm = fresh; fm; fm &= fm - 1u) {
        const int vs = __builtin_ctz(fm);
        int vcur = vs, vnext = P.nbr(vs, 0);
        int guard = 0;
        for (;;) {
            if (vnext >= NVF || !((after >> vnext) & 1u) || ++guard > 2 * NVF) {
                PVX_DBG("clip: walk lost vnext %d guard %d\n", vnext, guard);
                return false;
            }
            const uint32_t w = P.nbw(vnext);
            int k = 0;
            while (k < 3 && nbf(w, k) != vcur) k++;
            if (k == 3) {
                PVX_DBG("clip: walk no back link %d -> %d\n", vcur, vnext);
                return false;
            }
            vcur = vnext;
            if ((fresh >> vcur) & 1u) break;
            const int kn = k == 2 ? 0 : k + 1;
            P.setnbw(vcur, with_apex(w, kn, vs));
            vnext = nbf(w, kn);
        }
        /* vs -> vc lies on this face, vs -> vcur on the new one */
        P.setnbw(vs, with_apex(with_apex(with_nb(P.nbw(vs), 2, vcur), 0, vs), 2, cap));
        /* vcur -> vs closes this face */
        P.setnbw(vcur, with_apex(with_nb(P.nbw(vcur), 1, vs), 1, vs));
    }
    P.live = after;
    P.nv = nt;
    P.ord = nord;
    return true;
}

/* the point of the edge vc -> vn on the plane, from the two vertices'
 * signed distances sc, sn (of opposite signs) */
template <class R>
PVX_HD void cut_point(LPoly<R>& P, int vx, int vc, int vn, R sc, R sn) {
    const R c0 = P.qa(vc, 0), c1 = P.qa(vc, 1), c2 = P.qa(vc, 2);
    const R n0 = P.qa(vn, 0), n1 = P.qa(vn, 1), n2 = P.qa(vn, 2);
    const R inv = frcp(sc - sn);
    P.setq(vx, inv * (n0 * sc - c0 * sn), inv * (n1 * sc - c1 * sn), inv * (n2 * sc - c2 * sn));
}

/* Clip P in place to one side of the plane x_ax = c.  keep_upper: keep
 * x > c (drop x <= c); otherwise keep x <= c.  Only registers and the
 * lane's LDS slice are used.  False on slot overflow or a broken graph. */
template <class R>
PVX_HD bool clip_f(LPoly<R>& P, const CopyGeo<R>& G, const GridInfo& g, double c, int ax, bool keep_upper) {
    const uint32_t live = P.live;
    if (live == 0u) return true;
    const AxisMap<R> m = axis_map(G, g, ax);
    const R off = (R)(c - m.od);
    uint32_t gone = 0;
    for (uint32_t lm = live; lm; lm &= lm - 1u) {
        const int v = __builtin_ctz(lm);
        const R t = sdist(G.rec, m, c, off, P.qa(v, 0), P.qa(v, 1), P.qa(v, 2));
        const bool drop = keep_upper ? (t >= R(0)) : (t < R(0));
        if (drop) gone |= 1u << v;
    }
    if (gone == 0u) return true;
    if (gone == live) {
        P.live = 0u;
        P.nv = 0;
        return true;
    }
    return clip_graph(P, gone, [&](int vx, int vc, int vn) {
        const R sc = sdist(G.rec, m, c, off, P.qa(vc, 0), P.qa(vc, 1), P.qa(vc, 2));
        const R sn = sdist(G.rec, m, c, off, P.qa(vn, 0), P.qa(vn, 1), P.qa(vn, 2));
        cut_point(P, vx, vc, vn, sc, sn);
    });
}

/* q-moments accumulator about an origin: m0 = 6 vol, m1 = sum d S / 4,
 * M2 = sum d (AA^T + BB^T + CC^T + SS^T) / 20 with S = A+B+C, over signed
 * tets (origin, A, B, C); same integrals as src/geometry.c:725-740 */
template <class R>
struct Mom {
    R m0, s1[3], s2[6];
    PVX_HD void zero() {
        m0 = R(0);
        for (int i = 0; i < 3; ++i) s1[i] = R(0);
        for (int i = 0; i < 6; ++i) s2[i] = R(0);
    }
};

template <int ORDER, class R>
PVX_HD void tri_acc(Mom<R>& M, const R* A, const R* B, const R* C) {
    R d = A[0] * (B[1] * C[2] - B[2] * C[1]) - A[1] * (B[0] * C[2] - B[2] * C[0]) + A[2] * (B[0] * C[1] - B[1] * C[0]);
    M.m0 += d;
    if (ORDER > 0) {
        R S0 = A[0] + B[0] + C[0], S1 = A[1] + B[1] + C[1], S2 = A[2] + B[2] + C[2];
        M.s1[0] += d * S0;
        M.s1[1] += d * S1;
        M.s1[2] += d * S2;
        if (ORDER > 1) {
            M.s2[0] += d * (A[0] * A[0] + B[0] * B[0] + C[0] * C[0] + S0 * S0);
            M.s2[1] += d * (A[0] * A[1] + B[0] * B[1] + C[0] * C[1] + S0 * S1);
            M.s2[2] += d * (A[0] * A[2] + B[0] * B[2] + C[0] * C[2] + S0 * S2);
            M.s2[3] += d * (A[1] * A[1] + B[1] * B[1] + C[1] * C[1] + S1 * S1);
            M.s2[4] += d * (A[1] * A[2] + B[1] * B[2] + C[1] * C[2] + S1 * S2);
            M.s2[5] += d * (A[2] * A[2] + B[2] * B[2] + C[2] * C[2] + S2 * S2);
        }
    }
}

/* tri_acc with the signed volume scaled by w (0 or 1: a term that is
 * absent contributes nothing without a branch) */
template <int ORDER, class R>
PVX_HD void tri_acc_w(Mom<R>& M, const R* A, const R* B, const R* C, R w) {
    R d = A[0] * (B[1] * C[2] - B[2] * C[1]) - A[1] * (B[0] * C[2] - B[2] * C[0]) + A[2] * (B[0] * C[1] - B[1] * C[0]);
    d *= w;
    M.m0 += d;
    if (ORDER > 0) {
        const R S0 = A[0] + B[0] + C[0], S1 = A[1] + B[1] + C[1], S2 = A[2] + B[2] + C[2];
        M.s1[0] += d * S0;
        M.s1[1] += d * S1;
        M.s1[2] += d * S2;
        if (ORDER > 1) {
            M.s2[0] += d * (A[0] * A[0] + B[0] * B[0] + C[0] * C[0] + S0 * S0);
            M.s2[1] += d * (A[0] * A[1] + B[0] * B[1] + C[0] * C[1] + S0 * S1);
            M.s2[2] += d * (A[0] * A[2] + B[0] * B[2] + C[0] * C[2] + S0 * S2);
            M.s2[3] += d * (A[1] * A[1] + B[1] * B[1] + C[1] * C[1] + S1 * S1);
            M.s2[4] += d * (A[1] * A[2] + B[1] * B[2] + C[1] * C[2] + S1 * S2);
            M.s2[5] += d * (A[2] * A[2] + B[2] * B[2] + C[2] * C[2] + S2 * S2);
        }
    }
}

/* A point of P on the plane z = p (z = zd . q): on the first edge, in rank
 * order then neighbour order, from a vertex at or below p to one above it;
 * le has bit v set when z(v) <= p, so only that edge's vertices are read.
 * False when no edge crosses (P lies on one side). */
template <class R>
PVX_HD bool plane_point_le(const LPoly<R>& P, R zd0, R zd1, R zd2, R p, uint32_t le, R* out) {
    for (int r = 0; r < P.nv; ++r) {
        const int v = P.at(r);
        if (!((le >> v) & 1u)) continue;
        const uint32_t w = P.nbw(v);
        for (int k = 0; k < 3; ++k) {
            const int b = nbf(w, k);
            if ((le >> b) & 1u) continue;
            const R q0 = P.qa(v, 0), q1 = P.qa(v, 1), q2 = P.qa(v, 2);
            const R zv = zd0 * q0 + zd1 * q1 + zd2 * q2;
            const R r0 = P.qa(b, 0), r1 = P.qa(b, 1), r2 = P.qa(b, 2);
            const R zb = zd0 * r0 + zd1 * r1 + zd2 * r2;
            const R t = (p - zv) * frcp(zb - zv);
            out[0] = q0 + t * (r0 - q0);
            out[1] = q1 + t * (r1 - q1);
            out[2] = q2 + t * (r2 - q2);
            return true;
        }
    }
    return false;
}

/* exit and entry points of the triangle (c, a, b), traversed c -> a -> b,
 * on the plane z = p; inside(z) is the kept side.  X: where the boundary
 * leaves the kept side, N: where it re-enters; valid when both exist.
 * Invalid segments come back as (c, c). */
template <bool UPPER_KEPT_IS_LOWER, class R>
PVX_HD bool tri_segment(const R* qc, const R* qa, const R* qb, R zc, R za, R zb, R p, R rca, R rab, R rbc, R* X,
                        R* N) {
    /* kept: z > p (lower plane of a slab) or z <= p (upper plane) */
    const bool ic = UPPER_KEPT_IS_LOWER ? (zc <= p) : (zc > p);
    const bool ia = UPPER_KEPT_IS_LOWER ? (za <= p) : (za > p);
    const bool ib = UPPER_KEPT_IS_LOWER ? (zb <= p) : (zb > p);
    const bool valid = (ic != ia) || (ia != ib);
    /* crossing parameters of the three edges (only used where they cross) */
    const R tca = (za != zc) ? (p - zc) * rca : R(0);
    const R tab = (zb != za) ? (p - za) * rab : R(0);
    const R tbc = (zc != zb) ? (p - zb) * rbc : R(0);
    R Pca[3], Pab[3], Pbc[3];
    for (int i = 0; i < 3; ++i) {
        Pca[i] = qc[i] + tca * (qa[i] - qc[i]);
        Pab[i] = qa[i] + tab * (qb[i] - qa[i]);
        Pbc[i] = qb[i] + tbc * (qc[i] - qb[i]);
    }
    const bool xca = ic && !ia, xab = ia && !ib;
    const bool nca = !ic && ia, nab = !ia && ib;
    for (int i = 0; i < 3; ++i) {
        const R x = xca ? Pca[i] : (xab ? Pab[i] : Pbc[i]);
        const R n = nca ? Pca[i] : (nab ? Pab[i] : Pbc[i]);
        X[i] = valid ? x : qc[i];
        N[i] = valid ? n : qc[i];
    }
    return valid;
}

/* the piece of the edge u -> v inside lo < z <= hi as parameters t0 < t1
 * (valid) of u + t (v - u); inv = 1 / (zv - zu) (read only when zv != zu) */
template <class R>
PVX_HD bool edge_piece(R zu, R zv, R inv, R lo, R hi, R& t0, R& t1) {
    const R dz = zv - zu;
    if (dz != R(0)) {
        const R tl = (lo - zu) * inv, th = (hi - zu) * inv;
        t0 = rmax(R(0), rmin(tl, th));
        t1 = rmin(R(1), rmax(tl, th));
    } else {
        const bool in = zu > lo && zu <= hi;
        t0 = in ? R(0) : R(1);
        t1 = in ? R(1) : R(0);
    }
    return t0 < t1;
}

/* One cell's slab of a column: the planes as offsets from the tet's vertex
 * 0 along the sweep axis, the moment origin O (on the lower plane, on the
 * upper one when there is no lower plane), the upper cap's anchor
 * relative to O, and the fan triangles reaching into the slab. */
template <class R>
struct SlabCell {
    R lo, hi; /* -3e38 / +3e38 when the side is open */
    R O[3], gh[3];
    R zd[3]; /* x_ax(q) - x_ax(q = 0) = zd . q */
    bool has_lo, has_hi;
    uint32_t le; /* with one plane: the vertices at or below it (by slot) */
    uint64_t mask;   /* bit 3r + k: the fan triangle of directed edge a -> k, a the vertex of rank r,
                        reaching into the slab and cut by one of its planes */
    uint64_t inside; /* the same for fan triangles wholly inside the slab */
};

/* The point plane_point_le found on a cell's upper plane, kept for the
 * next cell of the column, whose lower plane is the same plane (the same
 * offset and the same vertex sides, so the search would find the same
 * point): each interior plane is searched once. */
template <class R>
struct PlaneCarry {
    bool valid; /* the last cell had an upper plane */
    bool found; /* and an edge crossed it at p */
    R p[3];
    /* that plane's sides: the live vertices at or below it, and the fan
     * triangles (bit 3r + k, as in SlabCell) wholly at or below it and
     * crossing it; all: every fan triangle of the column */
    uint32_t below;
    uint64_t bm, xm, all;
};

/* One pass over P's fan triangles against the vertex set le (the vertices
 * at or below a plane): all fan triangles, those wholly in le, and those
 * with vertices on both sides.  Bit 3r + k: the triangle of the directed
 * edge from the vertex of rank r to its neighbour k (an edge whose fan apex
 * is one of its ends has no triangle). */
template <class R>
PVX_HD void tri_sides(const LPoly<R>& P, uint32_t le, uint64_t& all, uint64_t& bm, uint64_t& xm) {
    all = bm = xm = 0ull;
    for (int ra = 0; ra < P.nv; ++ra) {
        const int a = P.at(ra);
        const uint32_t wa = P.nbw(a);
        /* the rank's three directed edges as a 3-bit group each, without branches */
        uint32_t ga = 0u, gb = 0u, gx = 0u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int b = nbf(wa, k), c = apf(wa, k);
            const uint32_t tri = (1u << a) | (1u << b) | (1u << c);
            const bool valid = c != a && c != b;
            const uint32_t in = le & tri;
            ga |= (valid ? 1u : 0u) << k;
            gb |= (valid && in == tri ? 1u : 0u) << k;
            gx |= (valid && in != 0u && in != tri ? 1u : 0u) << k;
        }
        all |= (uint64_t)ga << (3 * ra);
        bm |= (uint64_t)gb << (3 * ra);
        xm |= (uint64_t)gx << (3 * ra);
    }
}

/* the live vertices of P at or below the plane z = p (z = zd . q) */
template <class R>
PVX_HD uint32_t side_le(const LPoly<R>& P, R zd0, R zd1, R zd2, R p) {
    uint32_t le = 0u;
    for (uint32_t lm = P.live; lm; lm &= lm - 1u) {
        const int v = __builtin_ctz(lm);
        const R z = zd0 * P.qa(v, 0) + zd1 * P.qa(v, 1) + zd2 * P.qa(v, 2);
        le |= (z <= p ? 1u : 0u) << v;
    }
    return le;
}

/*
 * The slab lo < x_ax <= hi of P (either side may be open): its planes, the
 * origin and cap anchor on them, and the masks of fan triangles reaching
 * into it (z_max > lo and z_min <= hi: a triangle lying on the upper plane
 * belongs to the piece, one on the lower plane not): cut by one of its
 * planes (crossing lo or hi) or wholly inside (wholly above lo and wholly
 * at or below hi).  pc (optional) brings the lower plane's point and sides
 * from the column's previous cell, whose upper plane it is, and takes the
 * upper plane's for the next: each plane's vertex sides and triangle pass
 * are computed once, and the column's last cell needs neither.
 */
template <class R>
PVX_HD void slab_setup(const LPoly<R>& P, const CopyGeo<R>& G, int ax, bool has_lo, double lo_c, bool has_hi,
                       double hi_c, SlabCell<R>& S, PlaneCarry<R>* pc = nullptr) {
    const double x0d = pick3(G.x0d, ax);
    S.has_lo = has_lo;
    S.has_hi = has_hi;
    S.lo = has_lo ? (R)(lo_c - x0d) : R(-3.0e38);
    S.hi = has_hi ? (R)(hi_c - x0d) : R(3.0e38);
    const R zd0 = ax == 0 ? G.DX[0][0] : (ax == 1 ? G.DX[1][0] : G.DX[2][0]),
            zd1 = ax == 0 ? G.DX[0][1] : (ax == 1 ? G.DX[1][1] : G.DX[2][1]),
            zd2 = ax == 0 ? G.DX[0][2] : (ax == 1 ? G.DX[1][2] : G.DX[2][2]);
    S.zd[0] = zd0;
    S.zd[1] = zd1;
    S.zd[2] = zd2;
    const bool carried = has_lo && pc && pc->valid;
    /* the upper plane's sides (all of P below an open upper side) */
    const uint32_t below = has_hi ? side_le(P, zd0, zd1, zd2, S.hi) : P.live;
    uint64_t all, bhi, xhi;
    if (has_hi || !carried) {
        tri_sides(P, below, all, bhi, xhi);
    } else {
        all = pc->all;
        bhi = all;
        xhi = 0ull;
    }
    /* the lower plane's: carried, computed, or none */
    uint32_t above;
    uint64_t alo, xlo;
    S.le = below;
    if (carried) {
        above = P.live & ~pc->below;
        xlo = pc->xm;
        alo = all & ~(pc->bm | pc->xm);
        if (!has_hi) S.le = pc->below;
    } else if (has_lo) {
        const uint32_t le = side_le(P, zd0, zd1, zd2, S.lo);
        uint64_t a2, blo;
        tri_sides(P, le, a2, blo, xlo);
        above = P.live & ~le;
        alo = all & ~(blo | xlo);
        if (!has_hi) S.le = le;
    } else {
        above = P.live;
        xlo = 0ull;
        alo = all;
    }
    S.mask = xlo | xhi;
    S.inside = alo & bhi;
    const int v0 = P.at(0);
    R O[3] = {P.qa(v0, 0), P.qa(v0, 1), P.qa(v0, 2)};
    R Gh[3] = {O[0], O[1], O[2]};
    bool fh = false; /* a point on the upper plane was found */
    if (has_lo) {
        if (carried) {
            if (pc->found) O[0] = pc->p[0], O[1] = pc->p[1], O[2] = pc->p[2];
        } else {
            plane_point_le(P, zd0, zd1, zd2, S.lo, ~above, O);
        }
        if (has_hi) {
            fh = plane_point_le(P, zd0, zd1, zd2, S.hi, below, Gh);
            if (pc) pc->p[0] = Gh[0], pc->p[1] = Gh[1], pc->p[2] = Gh[2];
            if (!fh) Gh[0] = O[0], Gh[1] = O[1], Gh[2] = O[2];
        }
    } else if (has_hi) {
        fh = plane_point_le(P, zd0, zd1, zd2, S.hi, below, O);
        if (pc) pc->p[0] = O[0], pc->p[1] = O[1], pc->p[2] = O[2];
    }
    if (pc) {
        pc->valid = has_hi;
        pc->found = fh;
        pc->below = below;
        pc->bm = bhi;
        pc->xm = xhi;
        pc->all = all;
    }
    for (int i = 0; i < 3; ++i) {
        S.O[i] = O[i];
        S.gh[i] = Gh[i] - O[i];
    }
}

/*
 * The moments (about S.O) of one fan triangle T = (c, a, b) of P (directed
 * edge e = 3a + k) cut to the slab, added to M, with no branches on the
 * geometry.  The cone over T n slab from O is fanned from G, the start of
 * the piece of edge a -> b inside the slab (else of another edge piece:
 * always a vertex of T n slab, so every term is of the fragment's size):
 * one term per boundary piece -- the clipped edges c -> a and b -> c
 * (a -> b passes through G and adds nothing) and T's segments on the two
 * planes.  The cap on the upper plane is fanned from the anchor through the
 * same segments; the cap on the lower plane adds nothing (O lies on it).
 * Five fixed terms, each weighted 0 or 1.
 */
template <int ORDER, class R>
PVX_HD void pair_terms(const LPoly<R>& P, int e, const SlabCell<R>& S, Mom<R>& M) {
    const int ra = e / 3, k = e - 3 * ra, a = P.at(ra);
    const R* O = S.O;
    const R zd0 = S.zd[0], zd1 = S.zd[1], zd2 = S.zd[2];
    const R lo = S.lo, hi = S.hi;
    const uint32_t wa = P.nbw(a);
    const R pa0 = P.qa(a, 0), pa1 = P.qa(a, 1), pa2 = P.qa(a, 2);
    const R qa[3] = {pa0 - O[0], pa1 - O[1], pa2 - O[2]};
    const R za = zd0 * pa0 + zd1 * pa1 + zd2 * pa2;
    const int b = nbf(wa, k), c = apf(wa, k);
    const R pb0 = P.qa(b, 0), pb1 = P.qa(b, 1), pb2 = P.qa(b, 2);
    const R pc0 = P.qa(c, 0), pc1 = P.qa(c, 1), pc2 = P.qa(c, 2);
    const R qb[3] = {pb0 - O[0], pb1 - O[1], pb2 - O[2]};
    const R qc[3] = {pc0 - O[0], pc1 - O[1], pc2 - O[2]};
    const R zb = zd0 * pb0 + zd1 * pb1 + zd2 * pb2;
    const R zc = zd0 * pc0 + zd1 * pc1 + zd2 * pc2;
    /* the edges' reciprocal z-extents, shared by the edge pieces and the
     * plane segments */
    const R rca = za != zc ? frcp(za - zc) : R(0), rab = zb != za ? frcp(zb - za) : R(0),
            rbc = zc != zb ? frcp(zc - zb) : R(0);
    /* the edge pieces inside the slab, in traversal order */
    R s0, s1, u0, u1, v0, v1;
    const bool eca = edge_piece(zc, za, rca, lo, hi, s0, s1);
    const bool eab = edge_piece(za, zb, rab, lo, hi, u0, u1);
    const bool ebc = edge_piece(zb, zc, rbc, lo, hi, v0, v1);
    (void)u1;
    R A1[3], B1[3], A3[3], B3[3], Gp[3];
    for (int i = 0; i < 3; ++i) {
        A1[i] = qc[i] + s0 * (qa[i] - qc[i]);
        B1[i] = qc[i] + s1 * (qa[i] - qc[i]);
        const R A2 = qa[i] + u0 * (qb[i] - qa[i]);
        A3[i] = qb[i] + v0 * (qc[i] - qb[i]);
        B3[i] = qb[i] + v1 * (qc[i] - qb[i]);
        Gp[i] = eab ? A2 : (eca ? A1[i] : (ebc ? A3[i] : qc[i]));
    }
    tri_acc_w<ORDER>(M, Gp, B1, A1, eca ? R(1) : R(0));
    tri_acc_w<ORDER>(M, Gp, B3, A3, ebc ? R(1) : R(0));
    if (S.has_lo) {
        R X[3], N[3];
        const bool v = tri_segment<false>(qc, qa, qb, zc, za, zb, lo, rca, rab, rbc, X, N);
        tri_acc_w<ORDER>(M, Gp, N, X, v ? R(1) : R(0));
    }
    if (S.has_hi) {
        R X[3], N[3];
        const bool v = tri_segment<true>(qc, qa, qb, zc, za, zb, hi, rca, rab, rbc, X, N);
        tri_acc_w<ORDER>(M, Gp, N, X, v ? R(1) : R(0));
        if (S.has_lo) tri_acc_w<ORDER>(M, S.gh, X, N, v ? R(1) : R(0));
    }
}

/*
 * The crossing of a fan triangle T = (c, a, b) of P (directed edge e) with
 * the plane z = p, whose side mask le (vertices at or below p, by slot) T
 * crosses: the vertex s alone on its side, r1 and r2 after it in traversal
 * order (c -> a -> b), all relative to O, and the points P1 on s -> r1 and
 * P2 on s -> r2 where the plane cuts T, both measured from s (so the cells
 * on the two sides of p get the same points).  Returns whether s lies at or
 * below p.
 */
template <class R>
PVX_HD bool tri_cross(const LPoly<R>& P, int e, uint32_t le, const R* O, R zd0, R zd1, R zd2, R p, R* qs, R* q1,
                      R* q2, R* P1, R* P2) {
    const int ra = e / 3, k = e - 3 * ra, a = P.at(ra);
    const uint32_t wa = P.nbw(a);
    const int b = nbf(wa, k), c = apf(wa, k);
    const bool lc = (le >> c) & 1u, la = (le >> a) & 1u, lb = (le >> b) & 1u;
    const int s = lc == la ? b : (lc == lb ? a : c);
    const int r1 = lc == la ? c : (lc == lb ? b : a);
    const int r2 = lc == la ? a : (lc == lb ? c : b);
    const R s0 = P.qa(s, 0), s1 = P.qa(s, 1), s2 = P.qa(s, 2);
    const R u0 = P.qa(r1, 0), u1 = P.qa(r1, 1), u2 = P.qa(r1, 2);
    const R v0 = P.qa(r2, 0), v1 = P.qa(r2, 1), v2 = P.qa(r2, 2);
    const R zs = zd0 * s0 + zd1 * s1 + zd2 * s2;
    const R zu = zd0 * u0 + zd1 * u1 + zd2 * u2;
    const R zv = zd0 * v0 + zd1 * v1 + zd2 * v2;
    qs[0] = s0 - O[0], qs[1] = s1 - O[1], qs[2] = s2 - O[2];
    q1[0] = u0 - O[0], q1[1] = u1 - O[1], q1[2] = u2 - O[2];
    q2[0] = v0 - O[0], q2[1] = v1 - O[1], q2[2] = v2 - O[2];
    /* the sides come from le; a z that rounds onto s's own z (a vertex on
     * p, classified apart from s elsewhere) gives that vertex's own point */
    const R t1 = zu != zs ? (p - zs) * frcp(zu - zs) : R(0), t2 = zv != zs ? (p - zs) * frcp(zv - zs) : R(0);
    for (int i = 0; i < 3; ++i) {
        P1[i] = qs[i] + t1 * (q1[i] - qs[i]);
        P2[i] = qs[i] + t2 * (q2[i] - qs[i]);
    }
    return (le >> s) & 1u;
}

/*
 * pair_terms for a slab with one plane p (the first or last cell of a
 * column): S.O lies on p, so the kept piece K of T has no cap term, and
 * fanned from a point of K on p every boundary piece through that point
 * adds nothing either.  With s, r1, r2, P1, P2 as in tri_cross: when s is
 * kept, K is the triangle (P2, s, P1) (traversal order) and its cone is the
 * term (P1, s, P2); otherwise K is the quadrilateral (P1, r1, r2, P2),
 * fanned from P2: (P2, r2, r1) and (P2, r1, P1).  Two fixed terms, the second
 * weighted 0 or 1; the same integrals as pair_terms (the reference's fan
 * moments, src/geometry.c:692-767).
 */
template <int ORDER, class R>
PVX_HD void pair_terms_1(const LPoly<R>& P, int e, const SlabCell<R>& S, Mom<R>& M) {
    const bool upper = S.has_hi; /* kept side: z <= p, or z > p */
    const R p = upper ? S.hi : S.lo;
    R qs[3], q1[3], q2[3], P1[3], P2[3];
    const bool sb = tri_cross(P, e, S.le, S.O, S.zd[0], S.zd[1], S.zd[2], p, qs, q1, q2, P1, P2);
    const bool skept = sb == upper;
    R A[3], B[3], C[3];
    for (int i = 0; i < 3; ++i) {
        A[i] = skept ? P1[i] : P2[i];
        B[i] = skept ? qs[i] : q2[i];
        C[i] = skept ? P2[i] : q1[i];
    }
    tri_acc<ORDER>(M, A, B, C);
    tri_acc_w<ORDER>(M, P2, q1, P1, skept ? R(0) : R(1));
}

/* The moments (about S.O) of a fan triangle T = (c, a, b) wholly inside the
 * slab: the cone over T from O, the one term of pair_terms that does not
 * vanish there (its edge pieces are the whole edges, it has no segment on
 * either plane). */
template <int ORDER, class R>
PVX_HD void inside_terms(const LPoly<R>& P, int e, const SlabCell<R>& S, Mom<R>& M) {
    const int ra = e / 3, k = e - 3 * ra, a = P.at(ra);
    const uint32_t wa = P.nbw(a);
    const int b = nbf(wa, k), c = apf(wa, k);
    const R* O = S.O;
    const R qa[3] = {P.qa(a, 0) - O[0], P.qa(a, 1) - O[1], P.qa(a, 2) - O[2]};
    const R qb[3] = {P.qa(b, 0) - O[0], P.qa(b, 1) - O[1], P.qa(b, 2) - O[2]};
    const R qc[3] = {P.qa(c, 0) - O[0], P.qa(c, 1) - O[1], P.qa(c, 2) - O[2]};
    tri_acc<ORDER>(M, qa, qc, qb);
}

/* moments of P's piece in the slab, one lane doing every fan triangle */
template <int ORDER, class R>
PVX_HD void slab_moments_u(const LPoly<R>& P, const CopyGeo<R>& G, int ax, bool has_lo, double lo_c, bool has_hi,
                           double hi_c, Mom<R>& M, R* O, PlaneCarry<R>* pc = nullptr) {
    M.zero();
    SlabCell<R> S;
    slab_setup(P, G, ax, has_lo, lo_c, has_hi, hi_c, S, pc);
    O[0] = S.O[0];
    O[1] = S.O[1];
    O[2] = S.O[2];
    uint64_t in = S.inside;
    while (in) {
        const int e = __builtin_ctzll(in);
        in &= in - 1;
        inside_terms<ORDER>(P, e, S, M);
    }
    uint64_t mask = S.mask;
    if (S.has_lo && S.has_hi) {
        while (mask) {
            const int e = __builtin_ctzll(mask);
            mask &= mask - 1;
            pair_terms<ORDER>(P, e, S, M);
        }
    } else {
        while (mask) {
            const int e = __builtin_ctzll(mask);
            mask &= mask - 1;
            pair_terms_1<ORDER>(P, e, S, M);
        }
    }
}

/* tri_acc of the cone over (A, B, C) into Ml with weight wl and into Mu
 * with weight 1 - wl (wl 0 or 1): the integrals once, two accumulations */
template <int ORDER, class R>
PVX_HD void tri_acc2(Mom<R>& Ml, Mom<R>& Mu, const R* A, const R* B, const R* C, bool lower) {
    const R d = A[0] * (B[1] * C[2] - B[2] * C[1]) - A[1] * (B[0] * C[2] - B[2] * C[0]) +
                A[2] * (B[0] * C[1] - B[1] * C[0]);
    const R dl = lower ? d : R(0), du = lower ? R(0) : d;
    Ml.m0 += dl;
    Mu.m0 += du;
    if (ORDER > 0) {
        const R S0 = A[0] + B[0] + C[0], S1 = A[1] + B[1] + C[1], S2 = A[2] + B[2] + C[2];
        const R s1[3] = {S0, S1, S2};
        for (int i = 0; i < 3; ++i) {
            Ml.s1[i] += dl * s1[i];
            Mu.s1[i] += du * s1[i];
        }
        if (ORDER > 1) {
            const R t[6] = {A[0] * A[0] + B[0] * B[0] + C[0] * C[0] + S0 * S0,
                            A[0] * A[1] + B[0] * B[1] + C[0] * C[1] + S0 * S1,
                            A[0] * A[2] + B[0] * B[2] + C[0] * C[2] + S0 * S2,
                            A[1] * A[1] + B[1] * B[1] + C[1] * C[1] + S1 * S1,
                            A[1] * A[2] + B[1] * B[2] + C[1] * C[2] + S1 * S2,
                            A[2] * A[2] + B[2] * B[2] + C[2] * C[2] + S2 * S2};
            for (int i = 0; i < 6; ++i) {
                Ml.s2[i] += dl * t[i];
                Mu.s2[i] += du * t[i];
            }
        }
    }
}

/*
 * The two pieces of a fan triangle T of P (directed edge e) crossing the
 * plane z = p (side mask le), for the cell below p (Ml) and the cell above
 * (Mu), about O on p: with s, r1, r2, P1, P2 as in tri_cross, the piece on
 * s's side is the triangle (P2, s, P1), the other the quadrilateral (P1, r1,
 * r2, P2): the same three cone terms pair_terms_1 takes for the two cells,
 * from the same points, where two single-plane passes take four.
 */
template <int ORDER, class R>
PVX_HD void pair_terms_2(const LPoly<R>& P, int e, uint32_t le, const R* O, R zd0, R zd1, R zd2, R p, Mom<R>& Ml,
                         Mom<R>& Mu) {
    R qs[3], q1[3], q2[3], P1[3], P2[3];
    const bool sb = tri_cross(P, e, le, O, zd0, zd1, zd2, p, qs, q1, q2, P1, P2);
    tri_acc2<ORDER>(Ml, Mu, P1, qs, P2, sb);
    tri_acc2<ORDER>(Ml, Mu, P2, q2, q1, !sb);
    tri_acc2<ORDER>(Ml, Mu, P2, q1, P1, !sb);
}

/*
 * The moments of both cells of a column with exactly two cells and one
 * plane between them (hi_c along ax; no window plane in either cell), in
 * one pass over the fan triangles: cell by cell (slab_moments_u) each cell
 * walks them and cuts each crossing triangle, here each is cut once, into
 * both cells' pieces.  Ml, Mu: the lower and upper cell, both about O (on
 * the plane, the point the lower cell's pass takes).
 */
template <int ORDER, class R>
PVX_HD void slab_moments_2(const LPoly<R>& P, const CopyGeo<R>& G, int ax, double hi_c, Mom<R>& Ml, Mom<R>& Mu,
                           R* O) {
    Ml.zero();
    Mu.zero();
    const double x0d = pick3(G.x0d, ax);
    const R p = (R)(hi_c - x0d);
    const R zd0 = ax == 0 ? G.DX[0][0] : (ax == 1 ? G.DX[1][0] : G.DX[2][0]),
            zd1 = ax == 0 ? G.DX[0][1] : (ax == 1 ? G.DX[1][1] : G.DX[2][1]),
            zd2 = ax == 0 ? G.DX[0][2] : (ax == 1 ? G.DX[1][2] : G.DX[2][2]);
    const uint32_t below = side_le(P, zd0, zd1, zd2, p);
    uint64_t all, bm, xm;
    tri_sides(P, below, all, bm, xm);
    const int v0 = P.at(0);
    O[0] = P.qa(v0, 0);
    O[1] = P.qa(v0, 1);
    O[2] = P.qa(v0, 2);
    plane_point_le(P, zd0, zd1, zd2, p, below, O);
    /* triangles wholly on one side: the cone over each, to its side's cell */
    uint64_t in = all & ~xm;
    while (in) {
        const int e = __builtin_ctzll(in);
        in &= in - 1;
        const int ra = e / 3, k = e - 3 * ra, a = P.at(ra);
        const uint32_t wa = P.nbw(a);
        const int b = nbf(wa, k), c = apf(wa, k);
        const R qa[3] = {P.qa(a, 0) - O[0], P.qa(a, 1) - O[1], P.qa(a, 2) - O[2]};
        const R qb[3] = {P.qa(b, 0) - O[0], P.qa(b, 1) - O[1], P.qa(b, 2) - O[2]};
        const R qc[3] = {P.qa(c, 0) - O[0], P.qa(c, 1) - O[1], P.qa(c, 2) - O[2]};
        tri_acc2<ORDER>(Ml, Mu, qa, qc, qb, ((bm >> e) & 1ull) != 0ull);
    }
    uint64_t mask = xm;
    while (mask) {
        const int e = __builtin_ctzll(mask);
        mask &= mask - 1;
        pair_terms_2<ORDER>(P, e, below, O, zd0, zd1, zd2, p, Ml, Mu);
    }
}

/* index range of the x_ax-extent of a polyhedron inside [plo, phi), with
 * the zero-extent rule of the FP64 path (one cell) */
template <class R>
PVX_HD void lpoly_range(const LPoly<R>& P, const CopyGeo<R>& G, int ax, int ib0, int plo, int phi, int n, int& lo,
                        int& hi) {
    R zmin = R(3.0e38), zmax = R(-3.0e38);
    for (uint32_t lm = P.live; lm; lm &= lm - 1u) {
        const int v = __builtin_ctz(lm);
        R z = G.x0[ax] + G.DX[ax][0] * P.qa(v, 0) + G.DX[ax][1] * P.qa(v, 1) + G.DX[ax][2] * P.qa(v, 2);
        zmin = rmin(zmin, z);
        zmax = rmax(zmax, z);
    }
    /* local coordinates are relative to the copy's index box corner ib0 */
    R fl = rfloor(zmin), fh = rceil(zmax);
    lo = (fl < (R)(plo - ib0)) ? plo : ib0 + (int)fl;
    hi = (fh > (R)(phi - ib0)) ? phi : ib0 + (int)fh;
    if (lo < plo) lo = plo;
    if (hi > phi) hi = phi;
    if (hi <= lo) hi = lo + 1;
    if (hi > n) hi = n;
}

/* y-extent of the tet restricted to the x-slab lo < X <= hi (local units):
 * tet vertices inside plus edge crossings of the two planes */
template <class R>
PVX_HD void slab_yrange(const R (*X)[3], bool has_lo, R lo, bool has_hi, R hi, R& ymin, R& ymax) {
    ymin = R(3.0e38);
    ymax = R(-3.0e38);
    /* the closed slab: a vertex on a plane bounds the piece next to it */
    for (int v = 0; v < 4; ++v) {
        bool in = (!has_lo || X[v][0] >= lo) && (!has_hi || X[v][0] <= hi);
        if (in) {
            ymin = rmin(ymin, X[v][1]);
            ymax = rmax(ymax, X[v][1]);
        }
    }
    for (int a = 0; a < 4; ++a)
        for (int b = a + 1; b < 4; ++b) {
            R xa = X[a][0], xb = X[b][0];
            for (int pl = 0; pl < 2; ++pl) {
                if (pl == 0 && !has_lo) continue;
                if (pl == 1 && !has_hi) continue;
                R c = pl == 0 ? lo : hi;
                if ((xa - c) * (xb - c) < R(0)) {
                    R t = (c - xa) / (xb - xa);
                    R y = X[a][1] + t * (X[b][1] - X[a][1]);
                    ymin = rmin(ymin, y);
                    ymax = rmax(ymax, y);
                }
            }
        }
}

/* per-fragment deposit from q-moments about origin O (FP64 sums); FMASK:
 * the field flags a caller can have (the others compile away) */
template <int ORDER, int FMASK = -1, class R, class Sink>
PVX_HD void deposit_f(Sink& sink, const GridInfo& g, const CopyGeo<R>& G, const int* ib0, int i, int j, int k,
                      const Mom<R>& Mm, const R* O, const double* v0, const R (*DV)[3], double mass, double wm0) {
    /* weight: mass times the fragment's mass fraction wm0 (its m0, or m0
     * plus the tet's residual for the fragment that absorbs it) */
    const double w = mass * wm0;
    const int64_t gi = ((int64_t)g.n[1] * g.n[2] * i + (int64_t)g.n[2] * j + k) * g.nacc;
    sink.add(gi, w);
    if (ORDER == 0) return;
    const R inv = frcp(Mm.m0);
    /* mean q (origin shifted back) and q covariance about the mean */
    const R mu[3] = {R(0.25) * Mm.s1[0] * inv, R(0.25) * Mm.s1[1] * inv, R(0.25) * Mm.s1[2] * inv};
    const R qb[3] = {O[0] + mu[0], O[1] + mu[1], O[2] + mu[2]};
    R cov[3][3];
    if (ORDER > 1) {
        const R c20 = R(0.05) * inv;
        cov[0][0] = c20 * Mm.s2[0] - mu[0] * mu[0];
        cov[0][1] = cov[1][0] = c20 * Mm.s2[1] - mu[0] * mu[1];
        cov[0][2] = cov[2][0] = c20 * Mm.s2[2] - mu[0] * mu[2];
        cov[1][1] = c20 * Mm.s2[3] - mu[1] * mu[1];
        cov[1][2] = cov[2][1] = c20 * Mm.s2[4] - mu[1] * mu[2];
        cov[2][2] = c20 * Mm.s2[5] - mu[2] * mu[2];
    }
    const int fl = g.flags & FMASK;
    const int idx[3] = {i, j, k};
    double dx[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    R b_[3] = {R(0), R(0), R(0)}; /* DV qbar */
    if (fl & (F_X | F_XX | F_XV))
        for (int a = 0; a < 3; ++a) {
            /* centroid minus cell centre: the offset from the cell centre is
             * accurate to R of a cell size */
            const R dq = G.DX[a][0] * qb[0] + G.DX[a][1] * qb[1] + G.DX[a][2] * qb[2];
            const R cl = (R)(G.x0d[a] - ((double)(idx[a] - ib0[a]) + 0.5)) + dq;
            dx[a] = (double)cl * g.d[a];
        }
    if (fl & (F_V | F_VV | F_XV))
        for (int a = 0; a < 3; ++a) {
            b_[a] = DV[a][0] * qb[0] + DV[a][1] * qb[1] + DV[a][2] * qb[2];
            cv[a] = v0[a] + (double)b_[a];
        }
    if (fl & F_X)
        for (int a = 0; a < 3; ++a) sink.add(gi + g.ox + a, w * dx[a]);
    if (fl & F_V)
        for (int a = 0; a < 3; ++a) sink.add(gi + g.ov + a, w * cv[a]);
    if (ORDER < 2) return;
    /* covariances of x and v inside the fragment: J cov J^T */
    R DXa[3][3];
    if (fl & (F_XX | F_XV))
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) DXa[a][b] = G.DX[a][b] * (R)g.d[a];
    if (fl & F_XX) {
        R T[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) T[a][b] = DXa[a][0] * cov[0][b] + DXa[a][1] * cov[1][b] + DXa[a][2] * cov[2][b];
        int s = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b, ++s) {
                R sxx = T[a][0] * DXa[b][0] + T[a][1] * DXa[b][1] + T[a][2] * DXa[b][2];
                sink.add(gi + g.oxx + s, w * ((double)sxx + dx[a] * dx[b]));
            }
    }
    if (fl & F_VV) {
        R T[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) T[a][b] = DV[a][0] * cov[0][b] + DV[a][1] * cov[1][b] + DV[a][2] * cov[2][b];
        int s = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b, ++s) {
                R svv = T[a][0] * DV[b][0] + T[a][1] * DV[b][1] + T[a][2] * DV[b][2];
                sink.add(gi + g.ovv + s, w * ((double)svv + cv[a] * cv[b]));
            }
    }
    if (fl & F_XV) {
        R T[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) T[a][b] = DXa[a][0] * cov[0][b] + DXa[a][1] * cov[1][b] + DXa[a][2] * cov[2][b];
        const bool fix = (fl & F_XVFIX) != 0;
        /* the reference formula (src/geometry.c:134-147) adds terms in the
         * mean offsets from vertex 0, A = DX qbar and Bv = DV qbar, which are
         * tet-sized: form them in FP64 from the record */
        double A[3] = {0, 0, 0}, Bv[3] = {0, 0, 0};
        if (!fix) {
            const double* r = G.rec;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    A[a] += (r[3 * (b + 1) + a] - r[a]) * (double)qb[b];
                    Bv[a] += (r[12 + 3 * (b + 1) + a] - r[12 + a]) * (double)qb[b];
                }
        }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double sxv = (double)(T[a][0] * DV[b][0] + T[a][1] * DV[b][1] + T[a][2] * DV[b][2]);
                if (!fix) sxv += 2.0 * A[a] * Bv[b] - A[a] * A[b] - Bv[a] * Bv[b];
                sink.add(gi + g.oxv + 3 * a + b, w * (sxv + dx[a] * cv[b]));
            }
    }
}

struct FastCounters {
    int64_t nfrag, ncopies;
};

/* ------------------------------------------------------------------ */
/* column tasks                                                        */
/* ------------------------------------------------------------------ */

/* A copy of a tet prepared for column work: its cell-local geometry, index
 * box, window-plane mask and window planes in local units. */
template <class R>
struct CopySetup {
    CopyGeo<R> G;
    int ib1[3];
    unsigned wmask; /* window planes crossed: bit 2a lower, 2a+1 upper */
    double wlo[3], whi[3]; /* window planes, local units */
    R X[4][3];
};

/* copy h of the tet record (shift pattern from ghost_shifts_packed);
 * returns false when the copy's index box starts beyond the grid */
template <class R>
PVX_HD bool copy_setup(const double* rec, const double* lo, const double* hi, uint32_t shp, int h, const GridInfo& g,
                       const MeshBox& mb, CopySetup<R>& C) {
    C.G.rec = rec;
    C.wmask = 0;
    for (int a = 0; a < 3; ++a) {
        const double s = mb.periodic ? (mb.hi[a] - mb.lo[a]) * (double)(1 & (shp >> (3 * h + a))) : 0.0;
        C.G.shift[a] = s;
        /* window planes and index box, as src/geometry.c:816-841 */
        double cmin = lo[a] + s, cmax = hi[a] + s;
        if (g.w0[a] > cmin) {
            cmin = g.w0[a];
            C.wmask |= 1u << (2 * a);
        }
        if (g.w1[a] < cmax) {
            cmax = g.w1[a];
            C.wmask |= 2u << (2 * a);
        }
        index_range(g, a, cmin, cmax, C.G.ib0[a], C.ib1[a]);
        if (C.ib1[a] <= C.G.ib0[a]) C.ib1[a] = C.G.ib0[a] + 1;
    }
    if (C.G.ib0[0] >= g.n[0] || C.G.ib0[1] >= g.n[1] || C.G.ib0[2] >= g.n[2]) return false;
    for (int a = 0; a < 3; ++a) {
        const double inv = g.dinv[a];
        for (int v = 0; v < 4; ++v) C.X[v][a] = (R)((rec[3 * v + a] + C.G.shift[a] - g.w0[a]) * inv - C.G.ib0[a]);
        C.G.x0[a] = C.X[0][a];
        C.G.x0d[a] = (rec[a] + C.G.shift[a] - g.w0[a]) * inv - C.G.ib0[a];
        for (int b = 0; b < 3; ++b) C.G.DX[a][b] = (R)((rec[3 * (b + 1) + a] - rec[a]) * inv);
        C.wlo[a] = (double)(-C.G.ib0[a]);
        C.whi[a] = g.span[a] - C.G.ib0[a];
    }
    return true;
}

/* y-column range of slab i of a prepared copy */
template <class R>
PVX_HD bool slab_columns(const CopySetup<R>& C, int i, const GridInfo& g, int& j0, int& j1) {
    const int il = i - C.G.ib0[0];
    const bool xlo = i > C.G.ib0[0], xhi = i < C.ib1[0] - 1;
    const bool sxlo = xlo || (C.wmask & 1u), sxhi = xhi || (C.wmask & 2u);
    const R cxlo = xlo ? (R)il : (R)C.wlo[0], cxhi = xhi ? (R)(il + 1) : (R)C.whi[0];
    R ymin, ymax;
    slab_yrange(C.X, sxlo, cxlo, sxhi, cxhi, ymin, ymax);
    if (ymin > ymax) return false;
    j0 = C.G.ib0[1] + (int)rfloor(ymin);
    j1 = C.G.ib0[1] + (int)rceil(ymax);
    if (j0 < C.G.ib0[1]) j0 = C.G.ib0[1];
    if (j1 > C.ib1[1]) j1 = C.ib1[1];
    if (j1 <= j0) j1 = j0 + 1;
    if (j1 > g.n[1]) j1 = g.n[1];
    return j1 > j0;
}

/* column task: tet index, copy, the column's cell (i, j) relative to the
 * copy's index box, and whether the column is the first and the last of its
 * x-slab's columns (its y planes are then the copy's window planes or open;
 * the emit pass knows this, the column kernel need not recompute the
 * slab's y range); 32 + 3 + 14 + 13 + 1 + 1 bits */
PVX_HD uint64_t pack_task(int64_t t, int h, int il, int jl, bool first, bool last) {
    return (uint64_t)t | ((uint64_t)h << 32) | ((uint64_t)il << 35) | ((uint64_t)jl << 49) |
           ((uint64_t)(first ? 1 : 0) << 62) | ((uint64_t)(last ? 1 : 0) << 63);
}
PVX_HD void unpack_task(uint64_t w, int64_t& t, int& h, int& il, int& jl) {
    t = (int64_t)(w & 0xffffffffull);
    h = (int)((w >> 32) & 7u);
    il = (int)((w >> 35) & 0x3fffu);
    jl = (int)((w >> 49) & 0x1fffu);
}
PVX_HD bool task_first(uint64_t w) { return ((w >> 62) & 1u) != 0u; }
PVX_HD bool task_last(uint64_t w) { return ((w >> 63) & 1u) != 0u; }

PVX_HD bool finite_rec(const double* rec) {
    bool f = true;
    for (int t = 0; t < 12; ++t) f = f && (rec[t] - rec[t] == 0.0);
    return f;
}

/* whether a tet record spans more than WIDE_CELLS cells along an axis (its
 * columns then run in FP64 under the automatic kernel choice) */
PVX_HD bool tet_wide(const double* rec, const GridInfo& g) {
    double lo[3], hi[3];
    aabb((const double (*)[3])rec, 4, lo, hi);
    bool w = false;
    for (int a = 0; a < 3; ++a) w = w || (hi[a] - lo[a]) * g.dinv[a] > WIDE_CELLS;
    return w;
}

/* limits of the task encoding: column offsets within a copy's index box,
 * cells along z, and tasks per tet (the largest-fragment key) */
constexpr int TASK_MAXI = 1 << 14;
constexpr int TASK_MAXJ = 1 << 13;
constexpr int TASK_MAXK = 1 << 10;
constexpr int64_t TASK_MAXPERTET = 1 << 22;

/* Enumerate the column tasks of one tet record; writes them to out (when
 * given) and returns the count, or -1 when the tet cannot take the column
 * path (non-finite vertices, or an index box beyond the task encoding): the
 * caller then runs the FP64 kernel.  ncopies gets the number of copies.
 * R: the precision of the column kernel that will take the tasks. */
template <class R>
PVX_HDI int64_t tet_columns(const double* rec, int64_t t, const GridInfo& g, const MeshBox& mb, uint64_t* out,
                            int* ncopies) {
    if (ncopies) *ncopies = 0;
    if (!finite_rec(rec)) return -1;
    double lo[3], hi[3];
    aabb((const double (*)[3])rec, 4, lo, hi);
    uint32_t shp;
    const int ng = ghost_shifts_packed(lo, hi, g, mb, shp);
    if (ncopies) *ncopies = ng;
    int64_t n = 0;
    for (int h = 0; h < ng; ++h) {
        CopySetup<R> C;
        if (!copy_setup(rec, lo, hi, shp, h, g, mb, C)) continue;
        if (C.ib1[0] - C.G.ib0[0] >= TASK_MAXI || C.ib1[1] - C.G.ib0[1] >= TASK_MAXJ ||
            C.ib1[2] - C.G.ib0[2] >= TASK_MAXK)
            return -1;
        for (int i = C.G.ib0[0]; i < C.ib1[0]; ++i) {
            int j0, j1;
            if (!slab_columns(C, i, g, j0, j1)) continue;
            for (int j = j0; j < j1; ++j) {
                if (out) out[n] = pack_task(t, h, i - C.G.ib0[0], j - C.G.ib0[1], j == j0, j == j1 - 1);
                n++;
            }
        }
    }
    return n < TASK_MAXPERTET ? n : -1;
}

/*
 * z-extent of the tet X (cell-local vertices) inside the column xl <= x <=
 * xh, yl <= y <= yh, without building it: a linear function over the
 * convex intersection takes its extremes at vertices of it -- tet vertices
 * inside the column, tet edges crossing a side plane of the column, and
 * the column's vertical edges where they pass through the tet (the interval
 * on which all four barycentric coordinates stay non-negative).  False when
 * the intersection is empty.  nvert (optional) gets the number of those
 * vertices (a point on several planes counts more than once).  FP32: only
 * the processing order depends on it.
 */
PVX_HD bool column_zext(const float (*X)[3], float xl, float xh, float yl, float yh, float& zmin, float& zmax,
                        int* nvert = nullptr) {
    float lo = 3.0e38f, hi = -3.0e38f;
    int nvx = 0;
    auto take = [&](float x, float y, float z) {
        if (x >= xl && x <= xh && y >= yl && y <= yh) {
            lo = fminf(lo, z);
            hi = fmaxf(hi, z);
            nvx++;
        }
    };
    for (int v = 0; v < 4; ++v) take(X[v][0], X[v][1], X[v][2]);
    for (int u = 0; u < 3; ++u)
        for (int v = u + 1; v < 4; ++v) {
            /* the edge's reciprocal extent along x and y, shared by the two
             * planes of each (read only where the edge crosses one) */
            const float rx = frcp(X[u][0] - X[v][0]), ry = frcp(X[u][1] - X[v][1]);
            for (int s = 0; s < 4; ++s) {
                const int ax = s >> 1;
                const float c = s == 0 ? xl : (s == 1 ? xh : (s == 2 ? yl : yh));
                const float du = X[u][ax] - c, dv = X[v][ax] - c;
                if ((du < 0.f) == (dv < 0.f)) continue;
                const float t = du * (ax == 0 ? rx : ry);
                const float px = ax == 0 ? c : X[u][0] + t * (X[v][0] - X[u][0]);
                const float py = ax == 1 ? c : X[u][1] + t * (X[v][1] - X[u][1]);
                take(px, py, X[u][2] + t * (X[v][2] - X[u][2]));
            }
        }
    /* barycentrics l_1..3 = Inv (p - X0), l_0 = 1 - sum; Inv = adj(D) / det */
    float D[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) D[a][b] = X[b + 1][a] - X[0][a];
    const float c00 = D[1][1] * D[2][2] - D[1][2] * D[2][1], c01 = D[0][2] * D[2][1] - D[0][1] * D[2][2],
                c02 = D[0][1] * D[1][2] - D[0][2] * D[1][1];
    const float det = D[0][0] * c00 + D[1][0] * c01 + D[2][0] * c02;
    if (det != 0.f) {
        const float id = frcp(det);
        float inv[3][3];
        inv[0][0] = c00 * id;
        inv[0][1] = c01 * id;
        inv[0][2] = c02 * id;
        inv[1][0] = (D[1][2] * D[2][0] - D[1][0] * D[2][2]) * id;
        inv[1][1] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) * id;
        inv[1][2] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) * id;
        inv[2][0] = (D[1][0] * D[2][1] - D[1][1] * D[2][0]) * id;
        inv[2][1] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) * id;
        inv[2][2] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) * id;
        const float b1 = inv[0][2], b2 = inv[1][2], b3 = inv[2][2], b0 = -(b1 + b2 + b3);
        /* reciprocals of the four rates, shared by the four vertical edges */
        const float rb[4] = {b0 != 0.f ? frcp(b0) : 0.f, b1 != 0.f ? frcp(b1) : 0.f, b2 != 0.f ? frcp(b2) : 0.f,
                             b3 != 0.f ? frcp(b3) : 0.f};
        for (int q = 0; q < 4; ++q) {
            const float rx = ((q & 1) ? xh : xl) - X[0][0], ry = ((q & 2) ? yh : yl) - X[0][1], rz = -X[0][2];
            const float a1 = inv[0][0] * rx + inv[0][1] * ry + b1 * rz, a2 = inv[1][0] * rx + inv[1][1] * ry + b2 * rz,
                        a3 = inv[2][0] * rx + inv[2][1] * ry + b3 * rz, a0 = 1.f - (a1 + a2 + a3);
            const float av[4] = {a0, a1, a2, a3}, bv[4] = {b0, b1, b2, b3};
            float zl = -3.0e38f, zh = 3.0e38f;
            for (int i = 0; i < 4; ++i) {
                if (bv[i] > 0.f)
                    zl = fmaxf(zl, -av[i] * rb[i]);
                else if (bv[i] < 0.f)
                    zh = fminf(zh, -av[i] * rb[i]);
                else if (av[i] < 0.f)
                    zl = 3.0e38f;
            }
            if (zl <= zh) {
                lo = fminf(lo, zl);
                hi = fmaxf(hi, zh);
                nvx += 2;
            }
        }
    }
    zmin = lo;
    zmax = hi;
    if (nvert) *nvert = nvx;
    return lo <= hi;
}

/* Number of z cells of a column task (copy C, column (il, jl) relative to
 * the copy's index box) and (nvert) the column polyhedron's vertex count,
 * from column_zext: the sort key that groups tasks of equal work into the
 * same wavefront.  Only the processing order depends on it, so column_zext
 * divides with the hardware reciprocal. */
PVX_HD int column_cells(const CopySetup<float>& C, int il, int jl, int* nvert = nullptr) {
    float zmin, zmax;
    if (!column_zext(C.X, (float)il, (float)(il + 1), (float)jl, (float)(jl + 1), zmin, zmax, nvert)) return 0;
    const int nk = C.ib1[2] - C.G.ib0[2];
    int k0 = (int)floorf(zmin), k1 = (int)ceilf(zmax);
    if (k0 < 0) k0 = 0;
    if (k1 > nk) k1 = nk;
    return k1 > k0 ? k1 - k0 : 1;
}

/*
 * The x-slab step of a column, tabulated.  Clipping the tet by its upper x
 * plane and then its lower one gives a graph that depends only on where
 * each vertex lies: state 0 dropped by the lower plane, 1 kept, 2 dropped by
 * the upper plane; 81 patterns, pattern = sum of state_v 3^v.  Each entry is
 * the result of clip_graph on the tet for that pattern: per output vertex
 * (rank order) its slot and its link word, and the new vertices in the order
 * they are made (kind 1: cut of the upper plane on the edge from tet vertex
 * vc to tet vertex vn; kind 2: cut of the lower plane on the edge from slot
 * vc to tet vertex vn; each into slot vx).  The device evaluates the cuts
 * with the generic clip's own formula in the generic clip's sequence, so
 * the result is the same polyhedron.
 */
struct alignas(16) SlabStep {
    uint32_t link[8];
    uint8_t slot[8];
    uint16_t cut[8]; /* kind << 12 | vx << 8 | vc << 4 | vn */
    uint8_t nv, ncut;
    uint8_t pad[2];
    int32_t y2base; /* the y-slab table's first entry for this pattern */
};
static_assert(sizeof(SlabStep) == 64, "four 16-byte words per entry");

/* A table entry copied into registers with 16-byte loads: the lane then
 * decodes it without further memory round trips (a field read straight
 * from the table is one dependent load each). */
template <class T>
PVX_HD T load_entry(const T& src) {
    static_assert(sizeof(T) % 16 == 0 && alignof(T) >= 16, "whole aligned 16-byte words");
#ifdef __HIP_DEVICE_COMPILE__
    /* every word is loaded here, all loads in flight together: the empty
     * asm statements need the words in registers, so the compiler can
     * neither split a load by field nor defer one past a branch */
    constexpr int NW = (int)(sizeof(T) / 16);
    const uint4* p = reinterpret_cast<const uint4*>(&src);
    uint4 w[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = p[i];
#pragma unroll
    for (int i = 0; i < NW; ++i) __asm__ volatile("" : "+v"(w[i].x), "+v"(w[i].y), "+v"(w[i].z), "+v"(w[i].w));
    T v;
    memcpy(&v, w, sizeof(T));
    return v;
#else
    return src;
#endif
}

/* component i of a 4-vector without dynamic indexing */
template <class R>
PVX_HD R pick4(const R* v, int i) {
    return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3]));
}

template <class R>
PVX_HD void tet_start(LPoly<R>& P) {
    P.live = 0xfu;
    P.nv = 4;
    P.ord = 0x3210ull;
    P.setq(0, R(0), R(0), R(0));
    P.setq(1, R(1), R(0), R(0));
    P.setq(2, R(0), R(1), R(0));
    P.setq(3, R(0), R(0), R(1));
    /* neighbours, and the fan apex of each face: its first vertex */
    P.setnbw(0, pack3(1, 3, 2));
    P.setnbw(1, with_apex(pack3(2, 3, 0), 1, 1));
    P.setnbw(2, with_apex(pack3(0, 3, 1), 2, 1));
    P.setnbw(3, with_apex(pack3(1, 2, 0), 1, 1));
}

/* one table entry, from the graph steps alone (no coordinates) */
inline bool make_slab_step(int pattern, SlabStep& E) {
    static uint32_t mem[VRec<float>::WORDS];
    LPoly<float> P{mem, 0};
    tet_start(P);
    uint16_t cut[2 * NVF];
    int nc = 0;
    int st[4];
    for (int v = 0, x = pattern; v < 4; ++v, x /= 3) st[v] = x % 3;
    E = SlabStep{};
    uint32_t gone = 0;
    for (int v = 0; v < 4; ++v) gone |= (st[v] == 2 ? 1u : 0u) << v;
    if (gone == P.live) return true;
    if (gone && !clip_graph(P, gone, [&](int vx, int vc, int vn) {
            if (nc < 2 * NVF) cut[nc++] = (uint16_t)((1 << 12) | (vx << 8) | (vc << 4) | vn);
        }))
        return false;
    gone = 0;
    for (int v = 0; v < 4; ++v) gone |= (st[v] == 0 && ((P.live >> v) & 1u) ? 1u : 0u) << v;
    if (gone == P.live) return true;
    if (gone && !clip_graph(P, gone, [&](int vx, int vc, int vn) {
            if (nc < 2 * NVF) cut[nc++] = (uint16_t)((2 << 12) | (vx << 8) | (vc << 4) | vn);
        }))
        return false;
    if (P.nv > 8 || nc > 8) return false;
    E.nv = (uint8_t)P.nv;
    E.ncut = (uint8_t)nc;
    for (int r = 0; r < P.nv; ++r) {
        const int sl = P.at(r);
        E.slot[r] = (uint8_t)sl;
        E.link[r] = P.nbw(sl);
    }
    for (int i = 0; i < nc; ++i) E.cut[i] = cut[i];
    return true;
}

/* the 81 entries, built once on the host, each with the index of its first
 * y-slab table entry (the patterns' 3^nv state sets one after another) */
inline const SlabStep* host_slab_steps() {
    static SlabStep T[81];
    static bool made = false;
    if (!made) {
        int o = 0;
        for (int p = 0; p < 81; ++p) {
            make_slab_step(p, T[p]);
            T[p].y2base = o;
            int n = 1;
            for (int r = 0; r < T[p].nv; ++r) n *= 3;
            o += T[p].nv ? n : 0;
        }
        made = true;
    }
    return T;
}

#ifdef __HIPCC__
__constant__ SlabStep slab_steps[81];
#endif

PVX_HD const SlabStep& slab_step(int p) {
#ifdef __HIP_DEVICE_COMPILE__
    return slab_steps[p];
#else
    return host_slab_steps()[p];
#endif
}

/* P = the tet clipped by the upper x plane chi (when has_hi) and then the
 * lower one clo (when has_lo), from the table.  False when the table does
 * not apply (a cut of the upper plane on the dropped side of the lower one:
 * a window slab thinner than the precision resolves); P then holds the tet. */
template <class R>
PVX_HD bool slab_from_table(LPoly<R>& P, const CopyGeo<R>& G, const GridInfo& g, bool has_hi, double chi, bool has_lo,
                            double clo, int* y2p = nullptr) {
    tet_start(P);
    const AxisMap<R> m = axis_map(G, g, 0);
    const R offh = (R)(chi - m.od), offl = (R)(clo - m.od);
    R th[4], tl[4];
    int pat = 0;
    for (int v = 3; v >= 0; --v) {
        const R q0 = v == 1 ? R(1) : R(0), q1 = v == 2 ? R(1) : R(0), q2 = v == 3 ? R(1) : R(0);
        th[v] = has_hi ? sdist(G.rec, m, chi, offh, q0, q1, q2) : R(1);
        tl[v] = has_lo ? sdist(G.rec, m, clo, offl, q0, q1, q2) : R(-1);
        pat = 3 * pat + (th[v] < R(0) ? 2 : (tl[v] >= R(0) ? 0 : 1));
    }
    const SlabStep E = load_entry(slab_step(pat));
    if (y2p) *y2p = E.nv ? E.y2base : -1;
    if (E.nv == 0) {
        P.live = 0u;
        P.nv = 0;
        return true;
    }
    const int nv = E.nv, nc = E.ncut;
    /* the cut points in the order the generic clips make them (upper cuts,
     * between tet vertices, then lower cuts, from a survivor to a tet
     * vertex); a lane writes no lower cut once one of its upper cuts lies
     * on the dropped side of the lower plane (the table does not apply) */
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i >= nc) continue;
        const int cw = E.cut[i];
        const bool up = (cw >> 12) == 1;
        if (!up && !ok) continue;
        const int vx = (cw >> 8) & 15, vc = (cw >> 4) & 15, vn = cw & 3;
        const R sn = up ? pick4(th, vn) : pick4(tl, vn);
        const R sc = up ? pick4(th, vc) : sdist(G.rec, m, clo, offl, P.qa(vc, 0), P.qa(vc, 1), P.qa(vc, 2));
        cut_point(P, vx, vc, vn, sc, sn);
        if (up && has_lo && sdist(G.rec, m, clo, offl, P.qa(vx, 0), P.qa(vx, 1), P.qa(vx, 2)) >= R(0)) ok = false;
    }
    if (!ok) {
        tet_start(P);
        return false;
    }
    uint32_t live = 0;
    uint64_t ord = 0;
    for (int r = 0; r < 8; ++r) {
        if (r >= nv) continue;
        const int sl = E.slot[r];
        P.setnbw(sl, E.link[r]);
        live |= 1u << sl;
        ord |= (uint64_t)sl << (4 * r);
    }
    P.live = live;
    P.nv = nv;
    P.ord = ord;
    return true;
}

/*
 * The y-slab step of a column, tabulated like the x-slab: clipping the
 * x-slab polyhedron of table pattern pat by the column's upper y plane and
 * then its lower one gives a graph that depends only on where each of its
 * vertices lies (state 0 dropped by the lower plane, 1 kept, 2 dropped by
 * the upper plane; state = sum of state_r 3^r over the vertex ranks r).
 * Each entry is the result of clip_graph on entry pat's graph for those
 * states: per output vertex (rank order) its slot and its link word (fan
 * apexes included), and the new vertices in the order they are made (kind
 * 1: cut of the upper plane on the edge from slot vc to slot vn; kind 2:
 * cut of the lower plane on the edge from slot vc to slot vn; each into
 * slot vx).  The lane evaluates the cuts with clip_f's own formula in clip_f's
 * sequence, so the result is the two generic clips', bit for bit.  Entries
 * are one 128-byte line each; nv = Y2NONE marks a state set that no convex
 * polyhedron produces (the generic clips then run).
 */
constexpr int Y2NV = 14;
constexpr uint8_t Y2NONE = 0xffu;

struct alignas(16) Y2Step {
    uint32_t link[Y2NV];
    uint8_t slot[Y2NV];
    uint16_t cut[Y2NV]; /* the new vertices in the order they are made (the upper plane's first):
                           kind << 12 | slot << 8 | vc << 4 | vn */
    uint8_t nv, ncut;
    uint8_t pad[128 - 4 * Y2NV - Y2NV - 2 * Y2NV - 2];
};
static_assert(sizeof(Y2Step) == 128, "one line per entry");

/* entries: sum over the 81 patterns of 3^nv */
constexpr int Y2STEPS = 296703;

/* the entry of pattern pat for the vertex states st, from the graph steps alone */
inline void make_y2_step(int pat, int st, Y2Step& Y) {
    Y = Y2Step{};
    Y.nv = Y2NONE;
    const SlabStep& E = host_slab_steps()[pat];
    const int nv = E.nv;
    static uint32_t mem[VRec<float>::WORDS];
    LPoly<float> P{mem, 0u, 0, 0ull};
    int sr[8];
    for (int r = 0, x = st; r < nv; ++r, x /= 3) sr[r] = x % 3;
    uint16_t cut[2 * NVF];
    int nc = 0;
    uint32_t gone = 0u;
    for (int r = 0; r < nv; ++r) {
        const int sl = E.slot[r];
        P.setnbw(sl, E.link[r]);
        P.live |= 1u << sl;
        P.ord |= (uint64_t)sl << (4 * r);
        if (sr[r] == 2) gone |= 1u << sl;
    }
    P.nv = nv;
    if (gone == P.live) {
        Y.nv = 0;
        return;
    }
    if (gone && !clip_graph(P, gone, [&](int vx, int vc, int vn) {
            if (nc < 2 * NVF) cut[nc++] = (uint16_t)((1 << 12) | (vx << 8) | (vc << 4) | vn);
        }))
        return;
    gone = 0u;
    for (int r = 0; r < nv; ++r)
        if (sr[r] == 0 && ((P.live >> E.slot[r]) & 1u)) gone |= 1u << E.slot[r];
    if (gone == P.live) {
        Y.nv = 0;
        return;
    }
    if (gone && !clip_graph(P, gone, [&](int vx, int vc, int vn) {
            if (nc < 2 * NVF) cut[nc++] = (uint16_t)((2 << 12) | (vx << 8) | (vc << 4) | vn);
        }))
        return;
    if (P.nv > Y2NV || nc > Y2NV) return;
    /* an upper cut whose vertex the lower clip drops again has no place in
     * the list (the device takes the upper cuts as kept): no entry */
    for (int i = 0; i < nc; ++i)
        if ((cut[i] >> 12) == 1 && !((P.live >> ((cut[i] >> 8) & 15)) & 1u)) return;
    Y.nv = (uint8_t)P.nv;
    Y.ncut = (uint8_t)nc;
    for (int r = 0; r < P.nv; ++r) {
        const int sl = P.at(r);
        Y.link[r] = P.nbw(sl);
        Y.slot[r] = (uint8_t)sl;
    }
    for (int i = 0; i < nc; ++i) Y.cut[i] = cut[i];
}

/* the number of y-slab entries the x-slab table asks for */
inline int host_y2_count() {
    const SlabStep& E = host_slab_steps()[80];
    int n = 1;
    for (int r = 0; r < E.nv; ++r) n *= 3;
    return E.y2base + (E.nv ? n : 0);
}

/* the Y2STEPS entries, built once on the host */
inline const Y2Step* host_y2_steps() {
    static Y2Step* T = nullptr;
    if (!T) {
        const SlabStep* X = host_slab_steps();
        Y2Step* t = new Y2Step[Y2STEPS];
        for (int p = 0; p < 81; ++p) {
            int n = 1;
            for (int r = 0; r < X[p].nv; ++r) n *= 3;
            for (int q = 0; X[p].nv && q < n && X[p].y2base + q < Y2STEPS; ++q) make_y2_step(p, q, t[X[p].y2base + q]);
        }
        T = t;
    }
    return T;
}

#ifdef __HIPCC__
__device__ Y2Step y2_steps[Y2STEPS];
#endif

PVX_HD const Y2Step& y2_step(int i) {
#ifdef __HIP_DEVICE_COMPILE__
    return y2_steps[i];
#else
    return host_y2_steps()[i];
#endif
}

/* Clip P, the x-slab polyhedron as slab_from_table left it (y2first: its
 * pattern's first y-slab entry), by the upper y plane chi (when has_hi) and
 * then the lower one clo (when has_lo) through the table.  False when the table does not apply (no
 * entry, or a cut of the upper plane on the dropped side of the lower
 * one); P's graph is then unchanged and the generic clips run. */
template <class R>
PVX_HD bool yslab_from_table(LPoly<R>& P, int y2first, const CopyGeo<R>& G, const GridInfo& g, bool has_hi,
                             double chi, bool has_lo, double clo) {
    if (P.live == 0u) return true;
    const AxisMap<R> m = axis_map(G, g, 1);
    const R offh = (R)(chi - m.od), offl = (R)(clo - m.od);
    const int nv = P.nv;
    int st = 0;
    for (int r = 7; r >= 0; --r) {
        if (r >= nv) continue;
        const int v = P.at(r);
        const R q0 = P.qa(v, 0), q1 = P.qa(v, 1), q2 = P.qa(v, 2);
        const R th = has_hi ? sdist(G.rec, m, chi, offh, q0, q1, q2) : R(1);
        const R tl = has_lo ? sdist(G.rec, m, clo, offl, q0, q1, q2) : R(-1);
        st = 3 * st + (th < R(0) ? 2 : (tl >= R(0) ? 0 : 1));
    }
    const Y2Step Y = load_entry(y2_step(y2first + st));
    const int ny = Y.nv;
    if (ny == (int)Y2NONE) return false;
    if (ny == 0) {
        P.live = 0u;
        P.nv = 0;
        return true;
    }
    /* the cut points in the order the generic clips make them (each lane
     * walks its own list: upper cuts into free slots, then lower cuts), with
     * the generic clips' formula; a lane writes no lower cut once one of its
     * upper cuts lies on the dropped side of the lower plane (a slab thinner
     * than the precision resolves: its polyhedron is left for the generic
     * clips) */
    const int nc = Y.ncut;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < Y2NV; ++i) {
        if (i >= nc) continue;
        const int cw = Y.cut[i];
        const bool up = (cw >> 12) == 1;
        if (!up && !ok) continue;
        const int vx = (cw >> 8) & 15, vc = (cw >> 4) & 15, vn = cw & 15;
        const double cp = up ? chi : clo;
        const R off = up ? offh : offl;
        const R sc = sdist(G.rec, m, cp, off, P.qa(vc, 0), P.qa(vc, 1), P.qa(vc, 2));
        const R sn = sdist(G.rec, m, cp, off, P.qa(vn, 0), P.qa(vn, 1), P.qa(vn, 2));
        cut_point(P, vx, vc, vn, sc, sn);
        if (up && has_lo && sdist(G.rec, m, clo, offl, P.qa(vx, 0), P.qa(vx, 1), P.qa(vx, 2)) >= R(0)) ok = false;
    }
    if (!ok) return false;
    uint32_t live = 0u;
    uint64_t ord = 0ull;
    for (int r = 0; r < Y2NV; ++r) {
        if (r >= ny) continue;
        const int sl = Y.slot[r];
        P.setnbw(sl, Y.link[r]);
        live |= 1u << sl;
        ord |= (uint64_t)sl << (4 * r);
    }
    P.live = live;
    P.nv = ny;
    P.ord = ord;
    return true;
}

/* Build the column polyhedron of task (copy C, cell i, j) in P: the tet
 * clipped by the column's four planes -- interior grid planes or, on the
 * copy's first/last slab or column, the window planes it crosses
 * (src/geometry.c:816-843).  ylo, yhi: the column has a slab column below
 * and above it (interior y planes).  False on a vertex-budget overflow. */
template <class R>
PVX_HD bool build_column(LPoly<R>& P, const CopySetup<R>& C, const GridInfo& g, int i, int j, bool ylo, bool yhi) {
    const CopyGeo<R>& G = C.G;
    const int il = i - G.ib0[0], jl = j - G.ib0[1];
    const bool xlo = i > G.ib0[0], xhi = i < C.ib1[0] - 1;
    /* the x-slab: from the table, or (when it does not apply) by the two
     * generic clips */
    const bool has_hi = xhi || (C.wmask & 2u), has_lo = xlo || (C.wmask & 1u);
    const double chi = xhi ? (double)(il + 1) : C.whi[0], clo = xlo ? (double)il : C.wlo[0];
    bool ok = true;
    int y2first = -1; /* the y-slab table's entries for P's x-slab pattern, or -1 */
    if (!slab_from_table(P, G, g, has_hi, chi, has_lo, clo, &y2first)) {
        y2first = -1;
        if (has_hi) ok = clip_f(P, G, g, chi, 0, false);
        if (ok && has_lo) ok = clip_f(P, G, g, clo, 0, true);
    }
    /* the y planes: both through the y-slab table, or (when it does not
     * apply) by the two generic clips */
    const bool yh = yhi || (j == C.ib1[1] - 1 && (C.wmask & 8u));
    const bool yl = ylo || (j == G.ib0[1] && (C.wmask & 4u));
    const double cyh = yhi ? (double)(jl + 1) : C.whi[1], cyl = ylo ? (double)jl : C.wlo[1];
    if (ok && (yh || yl) && (y2first < 0 || !yslab_from_table(P, y2first, G, g, yh, cyh, yl, cyl))) {
        if (yh) ok = clip_f(P, G, g, cyh, 1, false);
        if (ok && yl) ok = clip_f(P, G, g, cyl, 1, true);
    }
    return ok;
}

/* a column task after its column is built: the copy, the column's cell
 * (i, j) and its z-cell range [k0, k1) */
template <class R>
struct ColumnState {
    CopySetup<R> C;
    int i, j, k0, k1;
};

/* the planes of cell k of a built column: interior planes between the
 * column's cells, window planes where the copy crosses them, open ends
 * elsewhere (src/geometry.c:816-843) */
template <class R>
PVX_HD void cell_planes(const ColumnState<R>& S, int k, bool& has_lo, double& lo_c, bool& has_hi, double& hi_c) {
    const CopySetup<R>& C = S.C;
    const int kl = k - C.G.ib0[2];
    has_lo = k > S.k0 || (k == C.G.ib0[2] && (C.wmask & 16u));
    has_hi = k < S.k1 - 1 || (k == C.ib1[2] - 1 && (C.wmask & 32u));
    lo_c = k > S.k0 ? (double)kl : C.wlo[2];
    hi_c = k < S.k1 - 1 ? (double)(kl + 1) : C.whi[2];
}

/* a column of exactly two cells whose one plane is the one between them
 * (no window plane in either): slab_moments_2 takes both cells at once */
template <class R>
PVX_HD bool two_cell_column(const ColumnState<R>& S) {
    const CopySetup<R>& C = S.C;
    const bool wlo = S.k0 == C.G.ib0[2] && (C.wmask & 16u);
    const bool whi = S.k1 == C.ib1[2] && (C.wmask & 32u);
    return S.k1 - S.k0 == 2 && !wlo && !whi;
}

/* the plane between a column's first two cells, in the copy's local units */
template <class R>
PVX_HD double first_plane(const ColumnState<R>& S) {
    return (double)(S.k0 - S.C.G.ib0[2] + 1);
}

/* Build the column of a task: 1 with cells [k0, k1), 0 when the column is
 * empty, -1 on a vertex-budget overflow (the FP64 kernel takes over). */
template <class R>
PVX_HD int column_begin(const GridInfo& g, const MeshBox& mb, const double* rec, uint64_t task, LPoly<R>& P,
                        ColumnState<R>& S) {
    int64_t t;
    int h, il, jl;
    unpack_task(task, t, h, il, jl);
    double lo[3], hi[3];
    aabb((const double (*)[3])rec, 4, lo, hi);
    uint32_t shp;
    ghost_shifts_packed(lo, hi, g, mb, shp);
    if (!copy_setup(rec, lo, hi, shp, h, g, mb, S.C)) return 0;
    S.i = S.C.G.ib0[0] + il;
    S.j = S.C.G.ib0[1] + jl;
    if (!build_column(P, S.C, g, S.i, S.j, !task_first(task), !task_last(task))) return -1;
    if (P.live == 0u) return 0;
    lpoly_range(P, S.C.G, 2, S.C.G.ib0[2], S.C.G.ib0[2], S.C.ib1[2], g.n[2], S.k0, S.k1);
    return S.k1 > S.k0 ? 1 : 0;
}

/* per-tet mass and velocity data of a record */
template <class R>
struct TetVel {
    double u0[3];
    R DV[3][3];
    double mass;
};

template <class R>
PVX_HD void tet_vel(const double* rec, TetVel<R>& V) {
    for (int a = 0; a < 3; ++a) {
        V.u0[a] = rec[12 + a];
        for (int b = 0; b < 3; ++b) V.DV[a][b] = (R)(rec[12 + 3 * (b + 1) + a] - rec[12 + a]);
    }
}

/* the largest fragment of a column task: its moments, origin and cell (in
 * FP32: the residual that is deposited with them is small) */
struct BestFrag {
    Mom<float> M;
    float O[3];
    int kl; /* z cell relative to the copy's index box; -1: none */
};

/* key of a task's largest fragment for the per-tet max: the positive FP32
 * m0 orders as its bit pattern */
PVX_HD uint64_t best_key(float m0, int kl) {
    uint32_t mb;
    memcpy(&mb, &m0, 4);
    return ((uint64_t)mb << 32) | (uint64_t)(uint32_t)kl;
}

/* A task's record of its cells: fragment count, FP64 mass sum and its
 * largest fragment (for the tet's residual pass). */
struct TaskTally {
    int64_t nfrag;
    double sum_m0;
    BestFrag B;
    PVX_HD void reset() {
        nfrag = 0;
        sum_m0 = 0.0;
        B.kl = -1;
        B.M.m0 = 0.0f;
    }
};

/* deposit cell k of a task from its slab moments Mm about O */
template <int ORDER, class R, class Sink>
PVX_HD void cell_finish(Sink& sink, const GridInfo& g, const ColumnState<R>& S, const TetVel<R>& V, int k,
                        const Mom<R>& Mm, const R* O, TaskTally& T) {
    if (!(Mm.m0 > R(0))) return;
    deposit_f<ORDER>(sink, g, S.C.G, S.C.G.ib0, S.i, S.j, k, Mm, O, V.u0, (const R (*)[3])V.DV, V.mass,
                     (double)Mm.m0);
    T.nfrag++;
    T.sum_m0 += (double)Mm.m0;
    if ((float)Mm.m0 > T.B.M.m0) {
        T.B.M.m0 = (float)Mm.m0;
        for (int i = 0; i < 3; ++i) T.B.M.s1[i] = (float)Mm.s1[i];
        for (int i = 0; i < 6; ++i) T.B.M.s2[i] = (float)Mm.s2[i];
        T.B.O[0] = (float)O[0];
        T.B.O[1] = (float)O[1];
        T.B.O[2] = (float)O[2];
        T.B.kl = k - S.C.G.ib0[2];
    }
}

/*
 * One column task done by one lane: build the column, integrate and
 * deposit its cells.  Returns false on overflow (the caller falls back to
 * the FP64 kernel).  T gets the task's tally for the residual pass.
 */
template <int ORDER, class R, class Sink>
PVX_HD bool column_task(Sink& sink, const GridInfo& g, const MeshBox& mb, const double* rec, uint64_t task,
                        LPoly<R>& P, TaskTally& T, bool carry = true, bool pair = true) {
    T.reset();
    ColumnState<R> S;
    const int st = column_begin(g, mb, rec, task, P, S);
    if (st < 0) return false;
    if (st == 0) return true;
    TetVel<R> V;
    if (ORDER > 0) tet_vel(rec, V);
    V.mass = rec[24];
    pin(V.mass);
    if (pair && two_cell_column(S)) {
        Mom<R> Ml, Mu;
        R O[3];
        slab_moments_2<ORDER>(P, S.C.G, 2, first_plane(S), Ml, Mu, O);
        cell_finish<ORDER>(sink, g, S, V, S.k0, Ml, O, T);
        cell_finish<ORDER>(sink, g, S, V, S.k0 + 1, Mu, O, T);
        return true;
    }
    PlaneCarry<R> pc{false, false, {R(0), R(0), R(0)}};
#pragma unroll 1
    for (int k = S.k0; k < S.k1; ++k) {
        bool hl, hh;
        double lc, hc;
        cell_planes(S, k, hl, lc, hh, hc);
        Mom<R> Mm;
        R O[3];
        slab_moments_u<ORDER>(P, S.C.G, 2, hl, lc, hh, hc, Mm, O, carry ? &pc : nullptr);
        cell_finish<ORDER>(sink, g, S, V, k, Mm, O, T);
    }
    return true;
}

/*
 * The residual of one tet, r = (exact in-window mass fraction, record slot
 * 26) - (FP64 sum of its fragment masses), is spread over its column tasks:
 * task c's largest fragment B takes r * sum_c / sum (sum_c: the task's
 * fragment-mass sum), deposited with that fragment's own means and
 * covariances so every moment stays consistent.  Spread this way each
 * correction is about the size of the fragments' own rounding error (a
 * single receiving fragment would take the whole tet's).  This deposits one
 * task's share: weight = r * sum_c / sum (or r for the one receiver of a
 * tet with a small residual, RESID_WHOLE).
 */
template <int ORDER, class Sink>
PVX_HD void task_residual(Sink& sink, const GridInfo& g, const MeshBox& mb, const double* rec, uint64_t task,
                          const BestFrag& B, double weight) {
    if (B.kl < 0 || !(B.M.m0 > 0.0f) || weight == 0.0) return;
    int64_t t;
    int h, il, jl;
    unpack_task(task, t, h, il, jl);
    double lo[3], hi[3];
    aabb((const double (*)[3])rec, 4, lo, hi);
    uint32_t shp;
    ghost_shifts_packed(lo, hi, g, mb, shp);
    CopySetup<float> C;
    if (!copy_setup(rec, lo, hi, shp, h, g, mb, C)) return;
    TetVel<float> V;
    if (ORDER > 0) tet_vel(rec, V);
    deposit_f<ORDER>(sink, g, C.G, C.G.ib0, C.G.ib0[0] + il, C.G.ib0[1] + jl, C.G.ib0[2] + B.kl, B.M, B.O, V.u0,
                     (const float (*)[3])V.DV, rec[24], weight);
}

/* the per-tet factor r / sum of the residual spread */
PVX_HD double residual_factor(const double* rec, double sum) { return sum > 0.0 ? (rec[26] - sum) / sum : 0.0; }

/* A tet whose residual is at most RESID_WHOLE times its largest fragment's
 * mass gives the whole residual to that fragment (one deposit; its
 * relative change stays below RESID_WHOLE); larger residuals are spread. */
constexpr double RESID_WHOLE = 2e-6;

}  // namespace pvx

#endif
