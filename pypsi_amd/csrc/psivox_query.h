/*
 * psivox_query.h -- point queries on a mesh: element boxes and the
 * multistream sample of PSI.VDF (src/cpsi.c:177-275), for host and device.
 */
#pragma once
#include "psivox_geom.h"

namespace pvx {

/* axis-aligned query box, closed on both sides */
struct QueryBox {
    double lo[3], hi[3];
};

/* The closed boxes meet (the overlap test of src/rtree.c:106-110). */
PVX_HD bool boxes_meet(const double* blo, const double* bhi, const QueryBox& q) {
    for (int a = 0; a < 3; ++a)
        if (q.hi[a] < blo[a] || q.lo[a] > bhi[a]) return false;
    return true;
}

/* det[a, b, c] of three 3-vectors */
PVX_HD double det3(const double* a, const double* b, const double* c) {
    return a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}

/*
 * Barycentric coordinates of x in the tet p (4 x 3) and the tet's volume.
 * Each coordinate is the signed volume of the tet with that vertex moved to
 * x, divided by the sum of the four (which is the tet's own signed volume);
 * the volume returned is |sum| / 6.  Same quantities as psi_barycentric
 * (src/geometry.c:335-389), which gets each coordinate from its face plane:
 * zero volume leaves the coordinates undefined and the caller skips the tet.
 */
PVX_HD double barycentric(const double* p, const double* x, double* b) {
    double e[4][3], r[4][3];
    for (int a = 0; a < 3; ++a) {
        for (int v = 0; v < 4; ++v) r[v][a] = p[3 * v + a] - x[a]; /* vertex v relative to x */
        for (int v = 1; v < 4; ++v) e[v][a] = p[3 * v + a] - p[a]; /* edge 0 -> v */
        e[0][a] = x[a] - p[a];                                    /* 0 -> x */
    }
    double s[4];
    s[0] = det3(r[1], r[2], r[3]);
    s[1] = det3(e[0], e[2], e[3]);
    s[2] = det3(e[1], e[0], e[3]);
    s[3] = det3(e[1], e[2], e[0]);
    const double tot = ((s[0] + s[1]) + s[2]) + s[3];
    for (int v = 0; v < 4; ++v) b[v] = s[v] / tot;
    return fabs(tot / 6.0);
}

/*
 * One stream of the sample at x from a tet record (TREC layout: pos 0..11,
 * vel 12..23, mass 24): 1 and (rho, v) when x lies in the tet (all
 * coordinates >= 0, nonzero volume), else 0.  rho = mass / volume, v the
 * barycentric interpolation of the vertex velocities (src/cpsi.c:238-263).
 * As there, a tet with non-finite vertices yields a NaN stream.
 */
PVX_HD int sample_tet(const double* rec, const double* x, double* rho, double* v) {
    double b[4];
    const double vol = barycentric(rec, x, b);
    if (vol <= 0.0) return 0;
    for (int k = 0; k < 4; ++k)
        if (b[k] < 0.0) return 0;
    for (int a = 0; a < 3; ++a) {
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += b[k] * rec[12 + 3 * k + a];
        v[a] = s;
    }
    *rho = rec[24] / vol;
    return 1;
}

}  // namespace pvx
