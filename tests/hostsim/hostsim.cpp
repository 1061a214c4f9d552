/*
 * hostsim.cpp -- TEST HARNESS: the device algorithm of libpsivox (element
 * expansion, FP64 sweep, raw deposit, finalize; pypsi_amd/csrc/*.h) compiled
 * for the host with plain += deposits, so the CPU test suite can check the
 * kernel logic against the oracle without a GPU.  Not used by the product.
 */
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../pypsi_amd/csrc/psivox_geom.h"
#include "../../pypsi_amd/csrc/psivox_moments.h"
#include "../../pypsi_amd/csrc/psivox_plain.h"
#include "../../pypsi_amd/csrc/psivox_refine.h"
#include "../../pypsi_amd/csrc/psivox_fast.h"
#include "../../pypsi_amd/csrc/psivox_adjoint.h"
#include "../../pypsi_amd/csrc/psivox_query.h"
#include "../../pypsi_amd/csrc/psivox_annihilate.h"

using namespace pvx;

struct HostSink {
    double* acc;
    void add(int64_t i, double v) { acc[i] += v; }
};

static void fill_grid(GridInfo& g, MeshBox& mb, const int* n, const double* win6, const double* d, double** fields,
                      int xvfix, int periodic, const double* box6) {
    for (int a = 0; a < 3; ++a) {
        g.n[a] = n[a];
        g.w0[a] = win6[a];
        g.w1[a] = win6[3 + a];
        g.d[a] = d[a];
    }
    g.flags = (fields[1] ? F_X : 0) | (fields[2] ? F_V : 0) | (fields[3] ? F_XX : 0) | (fields[4] ? F_XV : 0) |
              (fields[5] ? F_VV : 0) | (xvfix ? F_XVFIX : 0);
    set_layout(g);
    mb.periodic = periodic;
    mb.full_window = periodic;
    for (int a = 0; a < 3; ++a) {
        mb.lo[a] = box6[a];
        mb.hi[a] = box6[3 + a];
        if (mb.lo[a] != g.w0[a] || mb.hi[a] != g.w1[a]) mb.full_window = 0;
    }
}

/* whether the column path carries each interior plane's point to the next
 * cell (as the device does) or searches every plane afresh */
static bool g_carry = true;
extern "C" void hs_set_plane_carry(int on) { g_carry = on != 0; }

/* whether a column of two cells takes both in one pass over its fan
 * triangles (slab_moments_2, as the device does) or one pass per cell */
static bool g_pair = true;
extern "C" void hs_set_pair_cells(int on) { g_pair = on != 0; }

/* 1: wide tets (tet_wide) take the FP64 column path, as the device's
 * automatic kernel choice does; 0: every tet in FP32 (kernel 'fast') */
static int g_route = 0;
extern "C" void hs_set_route_wide(int on) { g_route = on; }

/* the column tasks of one tet record in precision R, deposited through
 * sink, with the per-tet residual; returns false on a task failure */
template <class R>
static bool tet_by_columns(HostSink& sink, const GridInfo& g, const MeshBox& mb, const double* r, int64_t t,
                           LPoly<R>& P, std::vector<uint64_t>& tasks, int64_t& nfrag, int64_t& ncopies,
                           int64_t& ntasks, int64_t& fails) {
    int nc = 0;
    const int64_t ncol = tet_columns<R>(r, t, g, mb, nullptr, &nc);
    ncopies += nc;
    if (ncol < 0) {
        fails++;
        return false;
    }
    tasks.resize((size_t)ncol);
    tet_columns<R>(r, t, g, mb, tasks.data(), nullptr);
    ntasks += ncol;
    std::vector<double> tsum((size_t)ncol);
    std::vector<BestFrag> tbest((size_t)ncol);
    double sum = 0.0;
    bool all = true;
    for (int64_t c = 0; c < ncol; ++c) {
        bool ok;
        TaskTally T;
        if (g.order == 0)
            ok = column_task<0>(sink, g, mb, r, tasks[c], P, T, g_carry, g_pair);
        else if (g.order == 1)
            ok = column_task<1>(sink, g, mb, r, tasks[c], P, T, g_carry, g_pair);
        else
            ok = column_task<2>(sink, g, mb, r, tasks[c], P, T, g_carry, g_pair);
        if (!ok) {
            fails++;
            all = false;
        }
        nfrag += T.nfrag;
        tsum[c] = T.sum_m0;
        tbest[c] = T.B;
        sum += tsum[c];
    }
    /* the largest fragment (first task on ties, as the GPU's key max) */
    int64_t cb = -1;
    uint64_t kb = 0;
    for (int64_t c = 0; c < ncol; ++c) {
        const uint64_t k = tbest[c].kl >= 0 ? best_key(tbest[c].M.m0, tbest[c].kl) : 0ull;
        if (k > kb) {
            kb = k;
            cb = c;
        }
    }
    const double resid = r[26] - sum;
    const bool whole = cb >= 0 && fabs(resid) <= RESID_WHOLE * (double)tbest[cb].M.m0;
    const double f = residual_factor(r, sum);
    for (int64_t c = 0; c < ncol; ++c) {
        const double wgt = whole ? (c == cb ? resid : 0.0) : f * tsum[c];
        if (g.order == 0)
            task_residual<0>(sink, g, mb, r, tasks[c], tbest[c], wgt);
        else if (g.order == 1)
            task_residual<1>(sink, g, mb, r, tasks[c], tbest[c], wgt);
        else
            task_residual<2>(sink, g, mb, r, tasks[c], tbest[c], wgt);
    }
    return all;
}

/* the FP32 fast path on the host: kernel == 2 */
extern "C" int hs_voxels_fast(int vpe, int64_t nelem, const double* pos, const double* vel, const double* mass,
                              const int32_t* conn, int periodic, const double* box6, const int* n, const double* win6,
                              const double* d, double** fields, int xvfix, double reftol, int maxlvl,
                              int64_t* out_stats) {
    GridInfo g;
    MeshBox mb;
    fill_grid(g, mb, n, win6, d, fields, xvfix, periodic, box6);
    const int64_t nvox = (int64_t)n[0] * n[1] * n[2];
    std::vector<double> acc((size_t)nvox * g.nacc, 0.0);
    FieldPtrs F;
    for (int f = 0; f < 6; ++f) F.f[f] = fields[f];
    /* the deposit starts from zero; the fields' contents are merged at the end, as on the device */
    double dxmin = d[0] < d[1] ? d[0] : d[1];
    dxmin = dxmin < d[2] ? dxmin : d[2];
    double tol2 = (reftol * dxmin) * (reftol * dxmin);
    if (vpe == 4) maxlvl = 0;
    RefineScratch* RS = new RefineScratch;
    Poly64* pool = new Poly64[2];
    std::vector<uint32_t> lv(VRec<float>::WORDS);
    LPoly<float> P{&lv[0], 0, 0, 0};
    std::vector<double> lv64((size_t)VRec<double>::WORDS / 2);
    LPoly<double> P64{(uint32_t*)lv64.data(), 0, 0, 0};
    HostSink sink{acc.data()};
    std::vector<double> tets;
    std::vector<uint64_t> tasks;
    int64_t ntets = 0, fails = 0, nfrag = 0, ncopies = 0, ntasks = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        int64_t nt = expand_element(pos, vel, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, nullptr);
        if (nt == 0) continue;
        tets.resize((size_t)nt * TREC);
        expand_element(pos, vel, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, tets.data());
        for (int64_t t = 0; t < nt; ++t) {
            double* r = &tets[(size_t)t * TREC];
            if (r[25] == 0.0) continue;
            if (!(mb.periodic && mb.full_window)) r[26] = window_target(r, g, mb, pool);
            if (g_route && tet_wide(r, g))
                tet_by_columns<double>(sink, g, mb, r, t, P64, tasks, nfrag, ncopies, ntasks, fails);
            else
                tet_by_columns<float>(sink, g, mb, r, t, P, tasks, nfrag, ncopies, ntasks, fails);
        }
        ntets += nt;
    }
    for (int64_t c = 0; c < nvox; ++c) cell_merge_finalize(g, acc.data(), F, c);
    delete RS;
    delete[] pool;
    out_stats[0] = nfrag;
    out_stats[1] = ncopies;
    out_stats[2] = ntets;
    out_stats[3] = fails;
    out_stats[4] = ntasks;
    return 0;
}

extern "C" int hs_voxels_plain(int vpe, int64_t npart, int64_t nelem, const double* pos, const double* vel,
                               const double* mass, const int32_t* conn, int periodic, const double* box6,
                               const int* n, const double* win6, const double* d, double** fields, int xvfix,
                               double reftol, int maxlvl, int64_t* out_stats) {
    (void)npart;
    GridInfo g;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = n[a];
        g.w0[a] = win6[a];
        g.w1[a] = win6[3 + a];
        g.d[a] = d[a];
    }
    g.flags = (fields[1] ? F_X : 0) | (fields[2] ? F_V : 0) | (fields[3] ? F_XX : 0) | (fields[4] ? F_XV : 0) |
              (fields[5] ? F_VV : 0) | (xvfix ? F_XVFIX : 0);
    set_layout(g);
    MeshBox mb;
    mb.periodic = periodic;
    for (int a = 0; a < 3; ++a) {
        mb.lo[a] = box6[a];
        mb.hi[a] = box6[3 + a];
    }
    const int64_t nvox = (int64_t)n[0] * n[1] * n[2];
    std::vector<double> acc((size_t)nvox * g.nacc, 0.0);
    FieldPtrs F;
    for (int f = 0; f < 6; ++f) F.f[f] = fields[f];
    /* the deposit starts from zero; the fields' contents are merged at the end, as on the device */
    double dxmin = d[0] < d[1] ? d[0] : d[1];
    dxmin = dxmin < d[2] ? dxmin : d[2];
    double tol2 = (reftol * dxmin) * (reftol * dxmin);
    if (vpe == 4) maxlvl = 0;
    RefineScratch* RS = new RefineScratch;
    PlainScratch* W = new PlainScratch;
    HostSink sink{acc.data()};
    std::vector<double> tets;
    int64_t nfrag = 0, ncopies = 0, ntets = 0;
    unsigned status = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        int64_t nt = expand_element(pos, vel, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, nullptr);
        if (nt == 0) continue;
        tets.resize((size_t)nt * TREC);
        expand_element(pos, vel, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, tets.data());
        for (int64_t t = 0; t < nt; ++t) {
            nfrag += voxelize_tet64(sink, g, mb, &tets[(size_t)t * TREC], *W, status, &ncopies);
        }
        ntets += nt;
    }
    for (int64_t c = 0; c < nvox; ++c) cell_merge_finalize(g, acc.data(), F, c);
    delete RS;
    delete W;
    out_stats[0] = nfrag;
    out_stats[1] = ncopies;
    out_stats[2] = ntets;
    out_stats[3] = status;
    return 0;
}

/* the mass adjoint of the GPU kernels on the host (refine level 0): dpos
 * (npart x 3) overwritten with dL/dpos for w = dL/dm */
/* every refined tet's parameter record checked against its vertices: the
 * element's map (wrapped nodes for positions) at each vertex's parameters
 * reproduces the vertex.  out: max position error, max velocity error, tet
 * count */
extern "C" int hs_expand_params(int vpe, int64_t nelem, const double* pos, const double* vel, const double* mass,
                                const int32_t* conn, int periodic, const double* box6, const int* n,
                                const double* win6, const double* d, double reftol, int maxlvl, double* out) {
    double* fields[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    GridInfo g;
    MeshBox mb;
    fill_grid(g, mb, n, win6, d, fields, 0, periodic, box6);
    const double dxmin = std::fmin(g.d[0], std::fmin(g.d[1], g.d[2]));
    const double tol2 = (reftol * dxmin) * (reftol * dxmin);
    RefineScratch* RS = new RefineScratch;
    std::vector<double> tets;
    std::vector<uint64_t> prm;
    double ep = 0.0, eu = 0.0, ntot = 0.0;
    for (int64_t e = 0; e < nelem; ++e) {
        const int64_t nt = expand_element(pos, vel, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, nullptr);
        if (nt == 0) continue;
        tets.resize((size_t)nt * TREC);
        prm.resize((size_t)nt * PREC);
        expand_element(pos, vel, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, tets.data(), prm.data());
        double P[27][3];
        for (int v = 0; v < vpe; ++v)
            for (int a = 0; a < 3; ++a) P[v][a] = pos[3 * (int64_t)conn[e * vpe + v] + a];
        wrap_and_test(P, vpe, g, mb);
        for (int64_t t = 0; t < nt; ++t) {
            const uint64_t* r = &prm[(size_t)t * PREC];
            if ((int64_t)r[0] != e) return -1;
            for (int v = 0; v < 4; ++v) {
                double w[27];
                int nn = 0;
                if (vpe == 4) {
                    for (int q = 0; q < 4; ++q) w[q] = q == v ? 1.0 : 0.0;
                    nn = 4;
                } else {
                    double xi[3];
                    unpack_param(r, v, xi);
                    nn = node_weights(vpe, xi, w);
                }
                for (int a = 0; a < 3; ++a) {
                    double x = 0.0, u = 0.0;
                    for (int q = 0; q < nn; ++q) {
                        x += w[q] * P[q][a];
                        u += w[q] * vel[3 * (int64_t)conn[e * vpe + q] + a];
                    }
                    ep = std::fmax(ep, std::fabs(x - tets[(size_t)t * TREC + 3 * v + a]));
                    eu = std::fmax(eu, std::fabs(u - tets[(size_t)t * TREC + 12 + 3 * v + a]));
                }
            }
        }
        ntot += (double)nt;
    }
    delete RS;
    out[0] = ep;
    out[1] = eu;
    out[2] = ntot;
    return 0;
}

/* the refined expansion of elements [0, nelem): tet records (TREC doubles
 * each) and parameter records (PREC words each) into caller buffers of cap
 * tets; returns the tet count (or -1 when cap is too small) */
extern "C" int64_t hs_expand_records(int vpe, int64_t nelem, const double* pos, const double* vel, const double* mass,
                                     const int32_t* conn, int periodic, const double* box6, const int* n,
                                     const double* win6, const double* d, double reftol, int maxlvl, int64_t cap,
                                     double* tets, uint64_t* prm) {
    double* fields[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    GridInfo g;
    MeshBox mb;
    fill_grid(g, mb, n, win6, d, fields, 0, periodic, box6);
    const double dxmin = std::fmin(g.d[0], std::fmin(g.d[1], g.d[2]));
    const double tol2 = (reftol * dxmin) * (reftol * dxmin);
    RefineScratch* RS = new RefineScratch;
    int64_t tot = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        const int64_t nt = expand_element(pos, vel, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, nullptr);
        if (tot + nt > cap) {
            delete RS;
            return -1;
        }
        if (nt) expand_element(pos, vel, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, tets + TREC * tot,
                               prm + PREC * tot);
        tot += nt;
    }
    delete RS;
    return tot;
}

/* every column task of the mesh built twice: by build_column (the y-slab table
 * where it applies) and by the generic clips alone (slab_from_table, then
 * clip_f for each plane); the two polyhedra compared bit for bit (live
 * slots, order, each live slot's q words and link word).  out: tasks,
 * tasks whose y planes cut the slab polyhedron, mismatches */
template <class R>
static void ycheck_tet(const double* r, int64_t t, const GridInfo& g, const MeshBox& mb, LPoly<R>& A, LPoly<R>& B,
                       int64_t* out) {
    std::vector<uint64_t> tasks(4096);
    const int64_t k = tet_columns<R>(r, t, g, mb, tasks.data(), nullptr);
    for (int64_t c = 0; c < k; ++c) {
        int64_t tt;
        int h, il, jl;
        unpack_task(tasks[c], tt, h, il, jl);
        double lo[3], hi[3];
        aabb((const double (*)[3])r, 4, lo, hi);
        uint32_t shp;
        ghost_shifts_packed(lo, hi, g, mb, shp);
        CopySetup<R> C;
        if (!copy_setup(r, lo, hi, shp, h, g, mb, C)) continue;
        const int i = C.G.ib0[0] + il, j = C.G.ib0[1] + jl;
        int j0, j1;
        if (!slab_columns(C, i, g, j0, j1)) continue;
        out[0]++;
        const bool okA = build_column(A, C, g, i, j, j > j0, j < j1 - 1);
        /* the generic sequence */
        const CopyGeo<R>& G = C.G;
        const bool xlo = i > G.ib0[0], xhi = i < C.ib1[0] - 1, ylo = j > j0, yhi = j < j1 - 1;
        const bool has_hi = xhi || (C.wmask & 2u), has_lo = xlo || (C.wmask & 1u);
        const double chi = xhi ? (double)(il + 1) : C.whi[0], clo = xlo ? (double)il : C.wlo[0];
        bool okB = true;
        const bool table = slab_from_table(B, G, g, has_hi, chi, has_lo, clo);
        if (!table) {
            if (has_hi) okB = clip_f(B, G, g, chi, 0, false);
            if (okB && has_lo) okB = clip_f(B, G, g, clo, 0, true);
        }
        const uint32_t slab_live = B.live;
        if (okB && yhi)
            okB = clip_f(B, G, g, (double)(jl + 1), 1, false);
        else if (okB && j == C.ib1[1] - 1 && (C.wmask & 8u))
            okB = clip_f(B, G, g, C.whi[1], 1, false);
        if (okB && ylo)
            okB = clip_f(B, G, g, (double)jl, 1, true);
        else if (okB && j == G.ib0[1] && (C.wmask & 4u))
            okB = clip_f(B, G, g, C.wlo[1], 1, true);
        if (table && B.live != slab_live) out[1]++;
        bool same = okA == okB && A.live == B.live && A.nv == B.nv && A.ord == B.ord;
        for (uint32_t lm = A.live; same && lm; lm &= lm - 1u) {
            const int v = __builtin_ctz(lm);
            for (int w = 0; w < VRec<R>::W; ++w)
                if (w == VRec<R>::LINK || w < (sizeof(R) == 4 ? 3 : 6))
                    same = same && A.w[v * VRec<R>::STRIDE + w] == B.w[v * VRec<R>::STRIDE + w];
        }
        if (!same) out[2]++;
    }
}

extern "C" int hs_ytable_check(int vpe, int64_t nelem, const double* pos, const double* vel, const double* mass,
                               const int32_t* conn, int periodic, const double* box6, const int* n, const double* win6,
                               const double* d, double reftol, int maxlvl, int fp64, int64_t* out) {
    double* fields[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    GridInfo g;
    MeshBox mb;
    fill_grid(g, mb, n, win6, d, fields, 0, periodic, box6);
    const double dxmin = std::fmin(g.d[0], std::fmin(g.d[1], g.d[2]));
    const double tol2 = (reftol * dxmin) * (reftol * dxmin);
    RefineScratch* RS = new RefineScratch;
    std::vector<uint32_t> la(VRec<double>::WORDS), lb(VRec<double>::WORDS);
    LPoly<float> A32{la.data(), 0, 0, 0}, B32{lb.data(), 0, 0, 0};
    LPoly<double> A64{la.data(), 0, 0, 0}, B64{lb.data(), 0, 0, 0};
    std::vector<double> tets;
    out[0] = out[1] = out[2] = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        const int64_t nt = expand_element(pos, vel, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, nullptr);
        if (nt == 0) continue;
        tets.resize((size_t)nt * TREC);
        expand_element(pos, vel, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, tets.data());
        for (int64_t t = 0; t < nt; ++t) {
            const double* r = &tets[(size_t)t * TREC];
            if (r[25] == 0.0) continue;
            if (fp64)
                ycheck_tet<double>(r, t, g, mb, A64, B64, out);
            else
                ycheck_tet<float>(r, t, g, mb, A32, B32, out);
        }
    }
    delete RS;
    return 0;
}

/* The ordering's sort key against the built columns: for every FP32
 * column task with a non-empty column, column_cells (the key's cell count,
 * from column_zext without building) against the built column's k1 - k0.
 * out = {tasks, equal, key below the built count, key above it}. */
extern "C" int hs_zext_check(int vpe, int64_t nelem, const double* pos, const double* vel, const double* mass,
                             const int32_t* conn, int periodic, const double* box6, const int* n, const double* win6,
                             const double* d, int64_t* out) {
    double* fields[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    GridInfo g;
    MeshBox mb;
    fill_grid(g, mb, n, win6, d, fields, 0, periodic, box6);
    RefineScratch* RS = new RefineScratch;
    std::vector<uint32_t> la(VRec<float>::WORDS);
    LPoly<float> P{la.data(), 0, 0, 0};
    std::vector<double> tets;
    std::vector<uint64_t> tasks;
    out[0] = out[1] = out[2] = out[3] = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        const int64_t nt = expand_element(pos, vel, mass, conn, vpe, e, g, mb, 0, 0.0, *RS, nullptr);
        if (nt == 0) continue;
        tets.resize((size_t)nt * TREC);
        expand_element(pos, vel, mass, conn, vpe, e, g, mb, 0, 0.0, *RS, tets.data());
        for (int64_t t = 0; t < nt; ++t) {
            const double* r = &tets[(size_t)t * TREC];
            if (r[25] == 0.0 || tet_wide(r, g)) continue;
            const int64_t k = tet_columns<float>(r, t, g, mb, nullptr, nullptr);
            if (k <= 0) continue;
            tasks.resize((size_t)k);
            tet_columns<float>(r, t, g, mb, tasks.data(), nullptr);
            for (int64_t c = 0; c < k; ++c) {
                ColumnState<float> S;
                if (column_begin(g, mb, r, tasks[c], P, S) <= 0) continue;
                int64_t tt;
                int h, il, jl;
                unpack_task(tasks[c], tt, h, il, jl);
                const int key = column_cells(S.C, il, jl);
                const int built = S.k1 - S.k0;
                out[0]++;
                out[key == built ? 1 : (key < built ? 2 : 3)]++;
            }
        }
    }
    delete RS;
    return 0;
}

extern "C" int hs_mass_adjoint(int vpe, int64_t npart, int64_t nelem, const double* pos, const double* mass,
                               const int32_t* conn, int periodic, const double* box6, const int* n, const double* win6,
                               const double* d, const double* w, double* dpos, int64_t* out_stats) {
    double* fields[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    GridInfo g;
    MeshBox mb;
    fill_grid(g, mb, n, win6, d, fields, 0, periodic, box6);
    for (int64_t q = 0; q < 3 * npart; ++q) dpos[q] = 0.0;
    std::vector<double> zv(3 * (size_t)(npart > 0 ? npart : 1), 0.0);
    RefineScratch* RS = new RefineScratch;
    std::vector<double> ru(8), rv(8);
    Poly2<1> R{ru.data(), rv.data()};
    std::vector<double> tets;
    std::vector<uint64_t> tasks;
    int64_t ntasks = 0, fails = 0;
    const int per = vpe == 4 ? 1 : 6;
    for (int64_t e = 0; e < nelem; ++e) {
        int64_t nt = expand_element(pos, zv.data(), mass, conn, vpe, e, g, mb, 0, 0.0, *RS, nullptr);
        if (nt == 0) continue;
        tets.resize((size_t)nt * TREC);
        expand_element(pos, zv.data(), mass, conn, vpe, e, g, mb, 0, 0.0, *RS, tets.data());
        for (int64_t t = 0; t < nt; ++t) {
            const double* r = &tets[(size_t)t * TREC];
            if (r[25] == 0.0) continue;
            const int64_t nk = adj_tasks(r, t, g, mb, nullptr);
            if (nk < 0) {
                fails++;
                continue;
            }
            tasks.resize((size_t)nk);
            adj_tasks(r, t, g, mb, tasks.data());
            ntasks += nk;
            for (int64_t c = 0; c < nk; ++c) {
                double o4[4];
                adj_plane_task<1>(g, mb, r, tasks[c], w, R, o4);
                const int ax = (int)((tasks[c] >> 35) & 3u);
                for (int v = 0; v < 4; ++v) {
                    const int64_t id = conn[e * vpe + tet_node(vpe, (int)(t % per), v)];
                    dpos[3 * id + ax] += o4[v];
                }
            }
        }
    }
    delete RS;
    out_stats[0] = ntasks;
    out_stats[1] = fails;
    return 0;
}

/* the adjoint of the fields m, v, vv with the device's task functions: W
 * holds ADJ_NW weights per cell; plane tasks for dpos, column tasks (FP64
 * moments) for dvel */
extern "C" int hs_adjoint(int vpe, int64_t npart, int64_t nelem, const double* pos, const double* vel,
                          const double* mass, const int32_t* conn, int periodic, const double* box6, const int* n,
                          const double* win6, const double* d, const double* W, double* dpos, double* dvel) {
    double* fields[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    GridInfo g;
    MeshBox mb;
    fill_grid(g, mb, n, win6, d, fields, 0, periodic, box6);
    for (int64_t q = 0; q < 3 * npart; ++q) dpos[q] = dvel[q] = 0.0;
    RefineScratch* RS = new RefineScratch;
    std::vector<double> ru(8), rv(8);
    Poly2<1> R{ru.data(), rv.data()};
    std::vector<double> lv((size_t)VRec<double>::WORDS / 2);
    LPoly<double> P{(uint32_t*)lv.data(), 0, 0, 0};
    std::vector<double> tets;
    std::vector<uint64_t> tasks;
    int fails = 0;
    const int per = vpe == 4 ? 1 : 6;
    for (int64_t e = 0; e < nelem; ++e) {
        int64_t nt = expand_element(pos, vel, mass, conn, vpe, e, g, mb, 0, 0.0, *RS, nullptr);
        if (nt == 0) continue;
        tets.resize((size_t)nt * TREC);
        expand_element(pos, vel, mass, conn, vpe, e, g, mb, 0, 0.0, *RS, tets.data());
        for (int64_t t = 0; t < nt; ++t) {
            const double* r = &tets[(size_t)t * TREC];
            if (r[25] == 0.0) continue;
            int64_t ids[4];
            for (int v = 0; v < 4; ++v) ids[v] = conn[e * vpe + tet_node(vpe, (int)(t % per), v)];
            const int64_t nk = adj_tasks(r, t, g, mb, nullptr);
            if (nk < 0) {
                fails++;
                continue;
            }
            tasks.resize((size_t)nk);
            adj_tasks(r, t, g, mb, tasks.data());
            for (int64_t c = 0; c < nk; ++c) {
                double o4[4];
                adj_plane_task<1, true>(g, mb, r, tasks[c], W, R, o4);
                const int ax = (int)((tasks[c] >> 35) & 3u);
                for (int v = 0; v < 4; ++v) dpos[3 * ids[v] + ax] += o4[v];
            }
            const int64_t ncol = tet_columns<double>(r, t, g, mb, nullptr, nullptr);
            if (ncol < 0) {
                fails++;
                continue;
            }
            tasks.resize((size_t)ncol);
            tet_columns<double>(r, t, g, mb, tasks.data(), nullptr);
            for (int64_t c = 0; c < ncol; ++c) {
                double o[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                if (!adj_vel_task<double>(g, mb, r, tasks[c], W, P, o)) fails++;
                for (int v = 0; v < 4; ++v)
                    for (int b = 0; b < 3; ++b) dvel[3 * ids[v] + b] += o[v][b];
            }
        }
    }
    delete RS;
    return fails;
}

/* the device query path on the host: element boxes as k_elem_boxes, the
 * box test of k_box_meets, candidates expanded as k_vdf_count/emit and
 * sampled as k_vdf_sample; returns the stream count, writes up to cap */
extern "C" int64_t hs_sample_vdf(int vpe, int64_t nelem, const double* pos, const double* vel, const double* mass,
                                 const int32_t* conn, int periodic, const double* box6, const double* win6,
                                 const double* x, double reftol, int maxlvl, double* rho, double* velout,
                                 int64_t cap) {
    GridInfo g;
    memset(&g, 0, sizeof g);
    MeshBox mb;
    mb.periodic = periodic;
    mb.full_window = 0;
    for (int a = 0; a < 3; ++a) {
        g.w0[a] = win6[a];
        g.w1[a] = win6[3 + a];
        mb.lo[a] = box6[a];
        mb.hi[a] = box6[3 + a];
    }
    QueryBox q;
    for (int a = 0; a < 3; ++a) q.lo[a] = q.hi[a] = x[a];
    GridInfo gu;
    memset(&gu, 0, sizeof gu);
    MeshBox mu;
    mu.periodic = 0;
    mu.full_window = 0;
    for (int a = 0; a < 3; ++a) {
        gu.w0[a] = -HUGE_VAL;
        gu.w1[a] = HUGE_VAL;
        mu.lo[a] = mu.hi[a] = 0.0;
    }
    if (maxlvl > MAXLVL) maxlvl = MAXLVL;
    RefineScratch* RS = new RefineScratch;
    std::vector<double> tets;
    int64_t n = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        double P[27][3], lo[3], hi[3];
        for (int v = 0; v < vpe; ++v)
            for (int a = 0; a < 3; ++a) P[v][a] = pos[3 * (int64_t)conn[e * vpe + v] + a];
        const int ok = wrap_and_test(P, vpe, g, mb);
        aabb((const double (*)[3])P, vpe, lo, hi);
        if (!ok || !boxes_meet(lo, hi, q)) continue;
        const int64_t nt = expand_element(pos, vel, mass, conn, vpe, e, gu, mu, maxlvl, reftol * reftol, *RS, nullptr);
        tets.resize((size_t)nt * TREC);
        expand_element(pos, vel, mass, conn, vpe, e, gu, mu, maxlvl, reftol * reftol, *RS, tets.data());
        for (int64_t t = 0; t < nt; ++t) {
            double r, v[3];
            if (!sample_tet(&tets[(size_t)t * TREC], x, &r, v)) continue;
            if (n < cap) {
                rho[n] = r;
                for (int a = 0; a < 3; ++a) velout[3 * n + a] = v[a];
            }
            n++;
        }
    }
    delete RS;
    return n;
}

/* annihilation mode with the device's pair code (annihilate_pair64) and
 * expansion (expand_element): element boxes as k_elem_boxes, every ordered
 * pair of elements whose boxes meet, tets of the first wrapped and of the
 * second raw, tet pairs whose boxes overlap.  Adds into m; returns the
 * fragment count, out_pairs = tet pairs. */
extern "C" int64_t hs_annihilation(int vpe, int64_t nelem, const double* pos, const double* mass,
                                   const int32_t* conn, int periodic, const double* box6, const int* n,
                                   const double* win6, const double* d, double* m, double reftol, int maxlvl,
                                   int64_t* out_pairs) {
    GridInfo g;
    MeshBox mb;
    double* fields[6] = {m, nullptr, nullptr, nullptr, nullptr, nullptr};
    fill_grid(g, mb, n, win6, d, fields, 0, periodic, box6);
    mb.full_window = 0;
    GridInfo gu = g;
    MeshBox mu;
    mu.periodic = 0;
    mu.full_window = 0;
    for (int a = 0; a < 3; ++a) {
        gu.w0[a] = -HUGE_VAL;
        gu.w1[a] = HUGE_VAL;
        mu.lo[a] = mu.hi[a] = 0.0;
    }
    double dxmin = d[0] < d[1] ? d[0] : d[1];
    dxmin = dxmin < d[2] ? dxmin : d[2];
    const double tol2 = (reftol * dxmin) * (reftol * dxmin);
    if (vpe == 4) maxlvl = 0;
    std::vector<double> blo((size_t)nelem * 3), bhi((size_t)nelem * 3);
    std::vector<int> valid((size_t)nelem);
    for (int64_t e = 0; e < nelem; ++e) {
        double P[27][3];
        for (int v = 0; v < vpe; ++v)
            for (int a = 0; a < 3; ++a) P[v][a] = pos[3 * (int64_t)conn[e * vpe + v] + a];
        valid[(size_t)e] = wrap_and_test(P, vpe, g, mb);
        aabb((const double (*)[3])P, vpe, &blo[(size_t)e * 3], &bhi[(size_t)e * 3]);
    }
    RefineScratch* RS = new RefineScratch;
    PlainScratch* W = new PlainScratch;
    struct MSink {
        double* m;
        void add(int64_t i, double v) { m[i] += v; }
    } sink{m};
    std::vector<double> ta, tb;
    int64_t nfrag = 0, npairs = 0;
    unsigned status = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        if (!valid[(size_t)e]) continue;
        const int64_t na = expand_element(pos, nullptr, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, nullptr);
        ta.resize((size_t)na * TREC);
        expand_element(pos, nullptr, mass, conn, vpe, e, g, mb, maxlvl, tol2, *RS, ta.data());
        QueryBox q;
        for (int a = 0; a < 3; ++a) {
            q.lo[a] = blo[(size_t)e * 3 + a];
            q.hi[a] = bhi[(size_t)e * 3 + a];
        }
        for (int64_t e1 = 0; e1 < nelem; ++e1) {
            if (!valid[(size_t)e1] || !boxes_meet(&blo[(size_t)e1 * 3], &bhi[(size_t)e1 * 3], q)) continue;
            const int64_t nb = expand_element(pos, nullptr, mass, conn, vpe, e1, gu, mu, maxlvl, tol2, *RS, nullptr);
            tb.resize((size_t)nb * TREC);
            expand_element(pos, nullptr, mass, conn, vpe, e1, gu, mu, maxlvl, tol2, *RS, tb.data());
            for (int64_t t = 0; t < na; ++t)
                for (int64_t t1 = 0; t1 < nb; ++t1) {
                    double lo[3], hi[3];
                    if (!tet_box_overlap(&ta[(size_t)t * TREC], &tb[(size_t)t1 * TREC], lo, hi)) continue;
                    npairs++;
                    nfrag += annihilate_pair64(sink, g, &ta[(size_t)t * TREC], &tb[(size_t)t1 * TREC], lo, hi, *W,
                                               status);
                }
        }
    }
    delete RS;
    delete W;
    *out_pairs = npairs;
    return status ? -(int64_t)status : nfrag;
}

/* the grid-side conversions on the host: fields -> raw moments (acc), and
 * the end-of-call merge of new raw moments (acc) into the fields */
extern "C" void hs_cells_to_raw(const int* n, const double* win6, const double* d, double** fields, int xvfix,
                                double* acc) {
    GridInfo g;
    MeshBox mb;
    const double box6[6] = {0, 0, 0, 0, 0, 0};
    fill_grid(g, mb, n, win6, d, fields, xvfix, 0, box6);
    FieldPtrs F;
    for (int f = 0; f < 6; ++f) F.f[f] = fields[f];
    const int64_t nvox = (int64_t)n[0] * n[1] * n[2];
    for (int64_t c = 0; c < nvox; ++c) cell_to_raw(g, F, acc, c);
}

extern "C" void hs_cells_merge_finalize(const int* n, const double* win6, const double* d, double** fields,
                                        int xvfix, double* acc) {
    GridInfo g;
    MeshBox mb;
    const double box6[6] = {0, 0, 0, 0, 0, 0};
    fill_grid(g, mb, n, win6, d, fields, xvfix, 0, box6);
    FieldPtrs F;
    for (int f = 0; f < 6; ++f) F.f[f] = fields[f];
    const int64_t nvox = (int64_t)n[0] * n[1] * n[2];
    for (int64_t c = 0; c < nvox; ++c) cell_merge_finalize(g, acc, F, c);
}
