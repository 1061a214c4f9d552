/*
 * psivox_geom.h -- polyhedron geometry of the voxelizer (device code; the
 * same source also compiles for the host so tests can run the kernel
 * algorithm on the CPU).
 *
 * A polyhedron is a vertex graph with exactly three neighbours per vertex,
 * listed counter-clockwise seen from outside (r3d-style, as in the reference:
 * include/geometry.h:118-129).  Every vertex carries its position x and its
 * mass coordinate q (q = 0, e_x, e_y, e_z at the four tet corners,
 * src/geometry.c:769-804); cuts interpolate both, so fragment moments are
 * exact integrals in q-space whatever the orientation or flatness of the tet.
 *
 * Cutting rule (src/geometry.c:560-690): with s = c - x[ax], vertices with
 * s < 0 go to the upper piece, the rest (including vertices on the plane) to
 * the lower piece; each cut edge yields one new vertex per piece at the same
 * interpolated point, so the two pieces tile the parent.
 *
 * Unlike the reference, which recursively bisects a polyhedron stack of
 * 64 x 3616 B per tet, the sweep here keeps four live polyhedra: the part of
 * the tet above the current x plane, the part of the current x slab above the
 * current y plane, the part of the current column above the current z plane,
 * and the current voxel fragment.
 */
#ifndef PSIVOX_GEOM_H
#define PSIVOX_GEOM_H

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define PVX_HD __host__ __device__ inline
#else
#include <math.h>
#define PVX_HD inline
#endif
/* inlined even where a call site sits in a loop (a call keeps its frame in
 * scratch memory) */
#define PVX_HDI PVX_HD __attribute__((always_inline))

namespace pvx {

/* vertex budget of one polyhedron in the FP64 path; a tet cut by the six
 * faces of a box has at most 16 vertices in general position */
constexpr int NV64 = 32;
/* new vertices one cut may create (cap polygon size) */
constexpr int NCUT = 16;

/* pick component ax of a 3-vector without dynamic indexing (keeps the
 * values in registers) */
template <typename T>
PVX_HD T pick3(const T* v, int ax) {
    return ax == 0 ? v[0] : (ax == 1 ? v[1] : v[2]);
}

enum { ST_OVERFLOW = 1, ST_BADGEOM = 2, ST_FALLBACK = 4 };

struct Poly64 {
    int nv;
    double x[NV64][3];
    double q[NV64][3];
    uint8_t nb[NV64][3];
};

template <typename T>
PVX_HD void swap_ptr(T*& a, T*& b) {
    T* t = a;
    a = b;
    b = t;
}

PVX_HD void init_tet(Poly64* P, const double (*x)[3]) {
    const uint8_t NB[4][3] = {{1, 3, 2}, {2, 3, 0}, {0, 3, 1}, {1, 2, 0}};
    P->nv = 4;
    for (int v = 0; v < 4; ++v)
        for (int a = 0; a < 3; ++a) {
            P->x[v][a] = x[v][a];
            P->q[v][a] = (v == a + 1) ? 1.0 : 0.0;
            P->nb[v][a] = NB[v][a];
        }
}

/*
 * Core cut.  R holds the polyhedron and the signed distances s[v]; on return
 * R holds the part with s < 0 ("upper") and L the part with s >= 0 ("lower").
 * The pointers may be swapped when one side is empty.  Vertex order of each
 * piece: its original vertices in order, then its new vertices in creation
 * order, as the reference produces them.
 */
PVX_HD void cut64(Poly64*& R, Poly64*& L, const double* s, unsigned& status) {
    const int onv = R->nv;
    uint32_t up = 0;
    int nup = 0;
    for (int v = 0; v < onv; ++v)
        if (s[v] < 0.0) {
            up |= 1u << v;
            nup++;
        }
    if (nup == 0) { /* all lower */
        swap_ptr(R, L);
        R->nv = 0;
        return;
    }
    if (nup == onv) { /* all upper */
        L->nv = 0;
        return;
    }
    /* new vertices: cut k has combined indices onv+2k (lower copy) and
     * onv+2k+1 (upper copy), sharing the point (nx[k], nq[k]) */
    double nx[NCUT][3], nq[NCUT][3];
    uint8_t nnb[2 * NCUT][3];
    for (int t = 0; t < 2 * NCUT; ++t) nnb[t][0] = nnb[t][1] = nnb[t][2] = 0xff;
    int ncut = 0;
    for (int vc = 0; vc < onv; ++vc) {
        if (up & (1u << vc)) continue;
        for (int k = 0; k < 3; ++k) {
            int vn = R->nb[vc][k];
            if (vn >= onv) {
                status |= ST_BADGEOM;
                R->nv = 0;
                L->nv = 0;
                return;
            }
            if (!(up & (1u << vn))) continue;
            if (ncut >= NCUT) {
                status |= ST_OVERFLOW;
                continue;
            }
            double inv = 1.0 / (s[vc] - s[vn]);
            for (int a = 0; a < 3; ++a) {
                nx[ncut][a] = inv * (R->x[vn][a] * s[vc] - R->x[vc][a] * s[vn]);
                nq[ncut][a] = inv * (R->q[vn][a] * s[vc] - R->q[vc][a] * s[vn]);
            }
            const int lo = onv + 2 * ncut, hi = lo + 1;
            nnb[2 * ncut][0] = (uint8_t)vc;
            R->nb[vc][k] = (uint8_t)lo;
            nnb[2 * ncut + 1][0] = (uint8_t)vn;
            int kk = 0;
            while (kk < 3 && R->nb[vn][kk] != vc) kk++;
            if (kk < 3) R->nb[vn][kk] = (uint8_t)hi;
            ncut++;
        }
    }
    /* walk the face right of each new vertex to the next new vertex
     * (src/geometry.c:646-669) */
    const int tot = onv + 2 * ncut;
    for (int vs = onv; vs < tot; ++vs) {
        int vc = vs, vn = nnb[vs - onv][0];
        int guard = 0;
        for (;;) {
            if (vn >= tot) {
                status |= ST_BADGEOM;
                break;
            }
            const uint8_t* nbn = (vn < onv) ? R->nb[vn] : nnb[vn - onv];
            int k = 0;
            while (k < 3 && nbn[k] != vc) k++;
            if (k == 3 || ++guard > 2 * NV64) {
                status |= ST_BADGEOM;
                break;
            }
            vc = vn;
            if (vc >= onv) break;
            vn = R->nb[vc][(k + 1) % 3];
        }
        nnb[vs - onv][2] = (uint8_t)vc;
        if (vc >= onv && vc < tot) nnb[vc - onv][1] = (uint8_t)vs;
    }
    if (status & ST_BADGEOM) {
        R->nv = 0;
        L->nv = 0;
        return;
    }
    /* remap combined indices to piece-local ones */
    uint8_t remap[NV64 + 2 * NCUT];
    const int nlo0 = onv - nup;
    if (nlo0 + ncut > NV64 || nup + ncut > NV64) {
        status |= ST_OVERFLOW;
        L->nv = 0;
        R->nv = 0;
        return;
    }
    {
        int il = 0, iu = 0;
        for (int v = 0; v < onv; ++v) remap[v] = (up & (1u << v)) ? (uint8_t)(iu++) : (uint8_t)(il++);
        for (int k = 0; k < ncut; ++k) {
            remap[onv + 2 * k] = (uint8_t)(nlo0 + k);
            remap[onv + 2 * k + 1] = (uint8_t)(nup + k);
        }
    }
    /* lower piece: originals then new lower copies; upper piece compacted
     * in place in R (targets never pass the source index) */
    int il = 0, iu = 0;
    for (int v = 0; v < onv; ++v) {
        if (up & (1u << v)) {
            for (int a = 0; a < 3; ++a) {
                R->x[iu][a] = R->x[v][a];
                R->q[iu][a] = R->q[v][a];
                R->nb[iu][a] = remap[R->nb[v][a]];
            }
            iu++;
        } else {
            for (int a = 0; a < 3; ++a) {
                L->x[il][a] = R->x[v][a];
                L->q[il][a] = R->q[v][a];
                L->nb[il][a] = remap[R->nb[v][a]];
            }
            il++;
        }
    }
    for (int k = 0; k < ncut; ++k) {
        for (int a = 0; a < 3; ++a) {
            L->x[il + k][a] = nx[k][a];
            L->q[il + k][a] = nq[k][a];
            L->nb[il + k][a] = remap[nnb[2 * k][a]];
            R->x[iu + k][a] = nx[k][a];
            R->q[iu + k][a] = nq[k][a];
            R->nb[iu + k][a] = remap[nnb[2 * k + 1][a]];
        }
    }
    L->nv = il + ncut;
    R->nv = iu + ncut;
}

/* split at x[ax] = c: R keeps x > c, L gets x <= c */
PVX_HD void split64(Poly64*& R, Poly64*& L, double c, int ax, unsigned& status) {
    double s[NV64];
    for (int v = 0; v < R->nv; ++v) s[v] = c - R->x[v][ax];
    cut64(R, L, s, status);
}

/* keep x[ax] >= c (the reference's window plane n = +e_ax, d = -c) */
PVX_HD void clip_lo64(Poly64*& P, Poly64*& tmp, double c, int ax, unsigned& status) {
    double s[NV64];
    for (int v = 0; v < P->nv; ++v) s[v] = -c + P->x[v][ax];
    cut64(P, tmp, s, status); /* P: s < 0 (dropped), tmp: kept */
    swap_ptr(P, tmp);
}

/* keep x[ax] <= c (the reference's window plane n = -e_ax, d = c) */
PVX_HD void clip_hi64(Poly64*& P, Poly64*& tmp, double c, int ax, unsigned& status) {
    double s[NV64];
    for (int v = 0; v < P->nv; ++v) s[v] = c - P->x[v][ax];
    cut64(P, tmp, s, status);
    swap_ptr(P, tmp);
}

/*
 * q-space moments by fan-triangulating every face from its first vertex and
 * summing signed tets with the origin (src/geometry.c:692-767, same
 * expressions): m[0] = 6 vol_q, m[1..3] = first moments, m[4..9] = second
 * moments xx, xy, xz, yy, yz, zz.
 */
PVX_HD void reduce64(const Poly64* P, double* m, int order) {
    const int nmom = order == 0 ? 1 : (order == 1 ? 4 : 10);
    for (int i = 0; i < nmom; ++i) m[i] = 0.0;
    uint8_t mark[NV64];
    for (int v = 0; v < P->nv; ++v) mark[v] = 0;
    for (int vs = 0; vs < P->nv; ++vs)
        for (int ks = 0; ks < 3; ++ks) {
            if (mark[vs] & (1 << ks)) continue;
            int vc = vs, kn = ks;
            mark[vc] |= (uint8_t)(1 << kn);
            int vn = P->nb[vc][kn];
            if (vn >= P->nv) return;
            const double* q0 = P->q[vc];
            int k = 0;
            while (k < 2 && P->nb[vn][k] != vc) k++;
            vc = vn;
            kn = (k + 1) % 3;
            mark[vc] |= (uint8_t)(1 << kn);
            vn = P->nb[vc][kn];
            int guard = 0;
            while (vn != vs && vn < P->nv && ++guard <= NV64) {
                const double* q1 = P->q[vn];
                const double* q2 = P->q[vc];
                double w = (-q2[0] * q1[1] * q0[2] + q1[0] * q2[1] * q0[2] + q2[0] * q0[1] * q1[2] -
                            q0[0] * q2[1] * q1[2] - q1[0] * q0[1] * q2[2] + q0[0] * q1[1] * q2[2]);
                m[0] += w;
                if (order > 0) {
                    m[1] += 0.25 * w * (q0[0] + q1[0] + q2[0]);
                    m[2] += 0.25 * w * (q0[1] + q1[1] + q2[1]);
                    m[3] += 0.25 * w * (q0[2] + q1[2] + q2[2]);
                }
                if (order > 1) {
                    m[4] += 0.1 * w *
                            (q0[0] * q0[0] + q1[0] * q1[0] + q2[0] * q2[0] + q1[0] * q2[0] + q0[0] * (q1[0] + q2[0]));
                    m[5] += 0.05 * w *
                            (q2[0] * q0[1] + q2[0] * q1[1] + 2 * q2[0] * q2[1] + q0[0] * (2 * q0[1] + q1[1] + q2[1]) +
                             q1[0] * (q0[1] + 2 * q1[1] + q2[1]));
                    m[6] += 0.05 * w *
                            (q2[0] * q0[2] + q2[0] * q1[2] + 2 * q2[0] * q2[2] + q0[0] * (2 * q0[2] + q1[2] + q2[2]) +
                             q1[0] * (q0[2] + 2 * q1[2] + q2[2]));
                    m[7] += 0.1 * w *
                            (q0[1] * q0[1] + q1[1] * q1[1] + q2[1] * q2[1] + q1[1] * q2[1] + q0[1] * (q1[1] + q2[1]));
                    m[8] += 0.05 * w *
                            (q2[1] * q0[2] + q2[1] * q1[2] + 2 * q2[1] * q2[2] + q0[1] * (2 * q0[2] + q1[2] + q2[2]) +
                             q1[1] * (q0[2] + 2 * q1[2] + q2[2]));
                    m[9] += 0.1 * w *
                            (q0[2] * q0[2] + q1[2] * q1[2] + q2[2] * q2[2] + q1[2] * q2[2] + q0[2] * (q1[2] + q2[2]));
                }
                k = 0;
                while (k < 2 && P->nb[vn][k] != vc) k++;
                vc = vn;
                kn = (k + 1) % 3;
                mark[vc] |= (uint8_t)(1 << kn);
                vn = P->nb[vc][kn];
            }
        }
}

/* ------------------------------------------------------------------ */
/* grid, accumulator layout                                             */
/* ------------------------------------------------------------------ */

enum { F_X = 1, F_V = 2, F_XX = 4, F_XV = 8, F_VV = 16, F_XVFIX = 32 };

struct GridInfo {
    int n[3];
    int flags;
    double w0[3], w1[3], d[3];
    /* 1 / d and the window's extent in cells, (w1 - w0) / d, formed once
     * (set_layout) with the same FP64 divisions a kernel would make */
    double dinv[3], span[3];
    /* raw accumulator layout per cell */
    int nacc, ox, ov, oxx, oxv, ovv;
    int order;
};

/* raw accumulator offsets and moment order from g.flags */
PVX_HD void set_layout(GridInfo& g) {
    for (int a = 0; a < 3; ++a) {
        g.dinv[a] = 1.0 / g.d[a];
        g.span[a] = (g.w1[a] - g.w0[a]) / g.d[a];
    }
    int o = 1;
    g.ox = g.ov = g.oxx = g.oxv = g.ovv = 0;
    if (g.flags & F_X) {
        g.ox = o;
        o += 3;
    }
    if (g.flags & F_V) {
        g.ov = o;
        o += 3;
    }
    if (g.flags & F_XX) {
        g.oxx = o;
        o += 6;
    }
    if (g.flags & F_XV) {
        g.oxv = o;
        o += 9;
    }
    if (g.flags & F_VV) {
        g.ovv = o;
        o += 6;
    }
    g.nacc = o;
    g.order = (g.flags & (F_XX | F_XV | F_VV)) ? 2 : ((g.flags & (F_X | F_V)) ? 1 : 0);
}

/* floor/ceil index range of [lo, hi] along ax, clamped to [0, n]
 * (src/geometry.c:838-841); clamped in floating point first so huge or
 * non-finite coordinates cannot overflow the integer conversion */
PVX_HD int clamp_index(double t, int n) {
    if (!(t >= 0.0)) return 0; /* also NaN */
    if (t > (double)n) return n;
    return (int)t;
}

/* the grid plane x_ax = w0 + d idx, rounded as the reference rounds it (a
 * multiply, then an add: src/geometry.c:887-888), not as one fused
 * multiply-add: every sweep of a polyhedron through its cells cuts at the
 * same planes as the reference's */
PVX_HD double grid_plane(const GridInfo& g, int ax, int idx) {
#pragma clang fp contract(off)
    return g.d[ax] * (double)idx + g.w0[ax];
}

PVX_HD void index_range(const GridInfo& g, int ax, double lo, double hi, int& ilo, int& ihi) {
    ilo = clamp_index(floor((lo - g.w0[ax]) / g.d[ax]), g.n[ax]);
    ihi = clamp_index(ceil((hi - g.w0[ax]) / g.d[ax]), g.n[ax]);
}

/* tight index range of a polyhedron along ax inside [plo, phi): a flat piece
 * lying on a plane keeps the cell above it, as the reference's leaf rule
 * does for a zero-extent index box (src/geometry.c:865-884) */
PVX_HD void poly_range64(const Poly64* P, const GridInfo& g, int ax, int plo, int phi, int& ilo, int& ihi) {
    double lo = 1.0e300, hi = -1.0e300;
    for (int v = 0; v < P->nv; ++v) {
        double t = P->x[v][ax];
        lo = t < lo ? t : lo;
        hi = t > hi ? t : hi;
    }
    index_range(g, ax, lo, hi, ilo, ihi);
    if (ilo < plo) ilo = plo;
    if (ihi > phi) ihi = phi;
    if (ihi <= ilo) ihi = ilo + 1;
    if (ihi > g.n[ax]) ihi = g.n[ax];
}

/* ------------------------------------------------------------------ */
/* periodic wrap and ghosts (src/cpsi.c:277-373)                        */
/* ------------------------------------------------------------------ */

struct MeshBox {
    int periodic;
    int full_window; /* the grid window is exactly the periodic box */
    double lo[3], hi[3];
};

PVX_HD void aabb(const double (*p)[3], int n, double* lo, double* hi) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = 1.0e30;
        hi[a] = -1.0e30;
    }
    for (int v = 0; v < n; ++v)
        for (int a = 0; a < 3; ++a) {
            if (p[v][a] < lo[a]) lo[a] = p[v][a];
            if (p[v][a] > hi[a]) hi[a] = p[v][a];
        }
}

/* wrap an element spanning more than half the box, reject it if it misses
 * the window; returns 0 to skip (src/cpsi.c:309-341) */
PVX_HD int wrap_and_test(double (*p)[3], int nv, const GridInfo& g, const MeshBox& mb) {
    double lo[3], hi[3];
    aabb((const double (*)[3])p, nv, lo, hi);
    for (int a = 0; a < 3; ++a) {
        if (mb.periodic) {
            double span = mb.hi[a] - mb.lo[a];
            if (hi[a] - lo[a] > 0.5 * span) {
                lo[a] = 1.0e30;
                hi[a] = -1.0e30;
                for (int v = 0; v < nv; ++v) {
                    if (p[v][a] > mb.lo[a] + 0.5 * span) p[v][a] -= span;
                    if (p[v][a] < lo[a]) lo[a] = p[v][a];
                    if (p[v][a] > hi[a]) hi[a] = p[v][a];
                }
            }
            if (lo[a] > g.w1[a]) return 0;
            if (hi[a] < g.w0[a] && lo[a] + span > g.w1[a]) return 0;
        } else {
            if (lo[a] > g.w1[a]) return 0;
            if (hi[a] < g.w0[a]) return 0;
        }
    }
    return 1;
}

/* ghost shift patterns packed 3 bits per copy (copy 0 = no shift); returns
 * the count.  Same rules as ghost_shifts, without an indexed array. */
PVX_HD int ghost_shifts_packed(const double* lo, const double* hi, const GridInfo& g, const MeshBox& mb,
                               uint32_t& packed) {
    packed = 0;
    if (!mb.periodic) return 1;
    int flags = 0, num = 1;
    for (int a = 0; a < 3; ++a)
        if (lo[a] < mb.lo[a]) flags |= 1 << a;
    for (int p = 1; p < 8; ++p) {
        if ((flags & p) != p) continue;
        int ok = 1;
        for (int a = 0; a < 3; ++a) {
            double sh = (mb.hi[a] - mb.lo[a]) * (double)(1 & (p >> a));
            if (lo[a] + sh > g.w1[a] || hi[a] + sh < g.w0[a]) {
                ok = 0;
                break;
            }
        }
        if (ok) packed |= (uint32_t)p << (3 * num++);
    }
    return num;
}

/* ghost shifts of a tet with box [lo, hi]: bit a of the returned mask p's
 * entry says shift +span along a.  Slot 0 is always the tet itself
 * (src/cpsi.c:343-373).  Returns the count; shifts[] gets the patterns. */
PVX_HD int ghost_shifts(const double* lo, const double* hi, const GridInfo& g, const MeshBox& mb, int* shifts) {
    shifts[0] = 0;
    int num = 1;
    if (!mb.periodic) return 1;
    int flags = 0;
    for (int a = 0; a < 3; ++a)
        if (lo[a] < mb.lo[a]) flags |= 1 << a;
    for (int p = 1; p < 8; ++p) {
        if ((flags & p) != p) continue;
        int ok = 1;
        for (int a = 0; a < 3; ++a) {
            double sh = (mb.hi[a] - mb.lo[a]) * (double)(1 & (p >> a));
            if (lo[a] + sh > g.w1[a] || hi[a] + sh < g.w0[a]) {
                ok = 0;
                break;
            }
        }
        if (ok) shifts[num++] = p;
    }
    return num;
}

}  // namespace pvx

#endif
