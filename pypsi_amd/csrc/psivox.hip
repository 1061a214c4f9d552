/*
 * psivox.hip -- MI355X (gfx950) voxelizer behind PSI.voxels() and its C ABI
 * (include/psivox.h).
 *
 * Device pipeline for one call (on one HIP stream):
 *   1. a zeroed accumulator of additive raw moments per cell
 *   2. k_expand     element -> wrapped, refined tets (count/scan/emit for
 *                   refine_max_lvl > 0, fixed 6-slot output otherwise)
 *   3. k_vox_plain  one thread per tet: FP64 clip + sweep + atomic deposit
 *      k_vox_cols   FP32 column tasks, FP64 deposit (psivox_fast.h)
 *   4. k_merge_raw, k_finalize  the fields' present contents (reference
 *                   layout) plus the new raw moments -> mass, means, covariances
 * Fields travel host<->device only when the caller passes host pointers;
 * large ones go up on a second stream while steps 1-3 run.
 */
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <new>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/psivox.h"
#include "psivox_adjoint.h"
#include "psivox_fast.h"
#include "psivox_geom.h"
#include "psivox_plain.h"
#include "psivox_refine.h"
#include "psivox_moments.h"
#include "psivox_query.h"
#include "psivox_annihilate.h"

using namespace pvx;

/* ------------------------------------------------------------------ */
/* errors                                                              */
/* ------------------------------------------------------------------ */

static thread_local std::string g_err;

static int set_err(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return set_err(e_ == hipErrorOutOfMemory ? PSIVOX_E_NOMEM : PSIVOX_E_HIP, "%s failed: %s", \
                           #call, hipGetErrorString(e_));                                              \
    } while (0)

/* ------------------------------------------------------------------ */
/* device sinks                                                        */
/* ------------------------------------------------------------------ */

struct AtomicSink {
    double* acc;
    __device__ void add(int64_t i, double v) { unsafeAtomicAdd(acc + i, v); }
};

/* ------------------------------------------------------------------ */
/* kernels                                                             */
/* ------------------------------------------------------------------ */

/* The end-of-call merge (cell_merge_finalize) in two passes: the fields'
 * present contents join the new deposits of a cell as raw moments, then
 * the sums are finalized.  Two kernels, so that the second reads acc as
 * read-only memory and issues all of a cell's loads at once (one kernel
 * that may write acc took twice as long). */
__global__ void k_merge_raw(GridInfo g, double* __restrict__ acc, FieldPtrs f, int64_t nvox) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nvox) return;
    if (acc[c * g.nacc] != 0.0 && f.f[0][c] != 0.0) cell_raw<true>(g, f, acc, c);
}

__global__ void k_finalize(GridInfo g, const double* __restrict__ acc, FieldPtrs f, int64_t nvox) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nvox) return;
    cell_finalize(g, acc, f, c);
}

static void launch_merge_finalize(const GridInfo& g, double* acc, const FieldPtrs& f, int64_t nvox, hipStream_t s) {
    const unsigned nb = (unsigned)((nvox + 255) / 256);
    k_merge_raw<<<nb, 256, 0, s>>>(g, acc, f, nvox);
    k_finalize<<<nb, 256, 0, s>>>(g, acc, f, nvox);
}

struct MeshDev {
    const double* pos;
    const double* vel;
    const double* mass;
    const int32_t* conn;
    int vpe;
};

/* p[0, n) = v: clears flags and counters with a kernel (a memset node
 * recorded into a captured graph is not used) */
__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* *flag |= 1 when an entry of conn[0, n) lies outside [0, npart): every
 * kernel reads particle arrays through conn, so a bad entry must stop the
 * call before any of them runs */
__global__ __launch_bounds__(256) void k_conn_check(const int32_t* __restrict__ conn, int64_t n, int64_t npart,
                                                    unsigned* __restrict__ flag) {
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = conn[i];
        bad = bad || v < 0 || v >= npart;
    }
    if (__any(bad) && (threadIdx.x & (warpSize - 1)) == 0) atomicOr(flag, 1u);
}

/* refine_max_lvl == 0: fixed slots (6 per element, 1 for simplices);
 * elements outside the window leave invalid records */
/* tag: written to every record's slot 27 (the mesh's index in a batched
 * call, 0 otherwise) */
/* Expansion at refinement level 0 for simplex and trilinear elements: one
 * lane per element, corners in registers.  The wavefront's records go out
 * through LDS one tet index at a time (record s of its 64 elements, at a
 * 1344-byte stride in the output for trilinear elements), so that
 * neighbouring lanes store neighbouring doubles instead of records 1344
 * bytes apart. */
__global__ __launch_bounds__(64) void k_expand_fixed(MeshDev m, GridInfo g, MeshBox mb, int64_t e0, int64_t ne,
                                                     double* __restrict__ tets, double tag = 0.0) {
    __shared__ double st[64 * TREC];
    const int lane = threadIdx.x;
    const int64_t w0 = (int64_t)blockIdx.x * 64;
    const int64_t t = w0 + lane;
    const int nel = (int)(ne - w0 < 64 ? ne - w0 : 64);
    const bool simplex = m.vpe == 4;
    const int per = simplex ? 1 : 6;
    /* simplex and linear elements: corners only, all in registers */
    double P[8][3], U[8][3];
    bool ok = false;
    double mass = 0.0;
    if (t < ne) {
        const int64_t e = e0 + t;
        const int nv = simplex ? 4 : 8;
        for (int v = 0; v < 8; ++v) {
            const int64_t id = m.conn[e * m.vpe + (v < nv ? v : 0)];
            for (int a = 0; a < 3; ++a) {
                P[v][a] = m.pos[3 * id + a];
                U[v][a] = m.vel ? m.vel[3 * id + a] : 0.0;
            }
        }
        ok = simplex ? wrap_and_test(P, 4, g, mb) : wrap_and_test(P, 8, g, mb);
        mass = m.mass[e];
    }
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        if (s >= per) break;
        /* this lane's record s (the 6-tet table on the corners, or the tet) */
        double* r = st + TREC * lane;
        if (t < ne) {
            for (int v = 0; v < 4; ++v) {
                const int l = tet27(s, v), c = simplex ? v : 4 * ((l / 9) / 2) + 2 * (((l / 3) % 3) / 2) + (l % 3) / 2;
                for (int a = 0; a < 3; ++a) {
                    r[3 * v + a] = P[c][a];
                    r[12 + 3 * v + a] = U[c][a];
                }
            }
            r[24] = simplex ? mass : mass * (1.0 / 6);
            r[25] = ok ? 1.0 : 0.0;
            r[26] = 1.0;
            r[27] = tag;
        }
        __syncthreads();
        /* the wavefront stores the records, neighbouring lanes neighbouring
         * doubles (each record's 28 contiguous) */
        for (int q = lane; q < nel * TREC; q += 64) {
            const int l = q / TREC, wq = q - TREC * l;
            tets[(int64_t)TREC * (per * (w0 + l) + s) + wq] = st[q];
        }
        __syncthreads(); /* the stage is read before the next record is written */
    }
}

/* the same for triquadratic elements: the 27 nodes are the lattice */
__global__ __launch_bounds__(128) void k_expand_fixed27(MeshDev m, GridInfo g, MeshBox mb, int64_t e0, int64_t ne,
                                 double* __restrict__ tets, double tag = 0.0) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne) return;
    const int64_t e = e0 + t;
    double* out = tets + (int64_t)TREC * 6 * t;
    double P[27][3], U[27][3];
    for (int v = 0; v < 27; ++v) {
        const int64_t id = m.conn[e * 27 + v];
        for (int a = 0; a < 3; ++a) {
            P[v][a] = m.pos[3 * id + a];
            U[v][a] = m.vel ? m.vel[3 * id + a] : 0.0;
        }
    }
    if (!wrap_and_test(P, 27, g, mb)) {
        for (int s = 0; s < 6; ++s) {
            out[TREC * s + 25] = 0.0;
            out[TREC * s + 27] = tag;
        }
        return;
    }
    emit6(out, (const double (*)[3])P, (const double (*)[3])U, m.mass[e]);
    for (int s = 0; s < 6; ++s) out[TREC * s + 27] = tag;
}

/* fixed expansion of elements [e0, e0 + ne) */
static void launch_expand_fixed(const MeshDev& md, const GridInfo& g, const MeshBox& mb, int64_t e0, int64_t ne,
                                double* tets, double tag, hipStream_t s) {
    const unsigned nb = (unsigned)((ne + 127) / 128);
    if (md.vpe == 27)
        k_expand_fixed27<<<nb, 128, 0, s>>>(md, g, mb, e0, ne, tets, tag);
    else
        k_expand_fixed<<<(unsigned)((ne + 63) / 64), 64, 0, s>>>(md, g, mb, e0, ne, tets, tag);
}

/* refine_max_lvl > 0: count pass and emit pass of the per-element DFS */
__global__ void k_expand_count(MeshDev m, GridInfo g, MeshBox mb, int64_t ne, int maxlvl, double tol2,
                               int64_t* __restrict__ counts) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    RefineScratch S;
    counts[e] = expand_element(m.pos, m.vel, m.mass, m.conn, m.vpe, e, g, mb, maxlvl, tol2, S, nullptr);
}

__global__ void k_expand_emit(MeshDev m, GridInfo g, MeshBox mb, int64_t e0, int64_t e1, int maxlvl, double tol2,
                              const int64_t* __restrict__ offsets, int64_t base, double* __restrict__ tets,
                              uint64_t* __restrict__ prm = nullptr) {
    int64_t e = e0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    RefineScratch S;
    expand_element(m.pos, m.vel, m.mass, m.conn, m.vpe, e, g, mb, maxlvl, tol2, S,
                   tets + (int64_t)TREC * (offsets[e] - base), prm ? prm + PREC * (offsets[e] - base) : nullptr);
}

/* ------------------------------------------------------------------ */
/* element boxes and point queries (RStarTree, PSI.VDF)                */
/* ------------------------------------------------------------------ */

/* element e's box: its vertex box after the periodic wrap of
 * psi_aabb_periodic (src/cpsi.c:309-341); valid[e] = 0 when that test
 * rejects it for the window (src/rtree.c:62-70 leaves it out) */
__global__ void k_elem_boxes(MeshDev m, GridInfo g, MeshBox mb, int64_t ne, double* __restrict__ box,
                             int8_t* __restrict__ valid) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double P[27][3], lo[3], hi[3];
    for (int v = 0; v < m.vpe; ++v) {
        const int64_t id = m.conn[e * m.vpe + v];
        for (int a = 0; a < 3; ++a) P[v][a] = m.pos[3 * id + a];
    }
    valid[e] = (int8_t)wrap_and_test(P, m.vpe, g, mb);
    aabb((const double (*)[3])P, m.vpe, lo, hi);
    for (int a = 0; a < 3; ++a) {
        box[6 * e + a] = lo[a];
        box[6 * e + 3 + a] = hi[a];
    }
}

/* flags[e] = element e is in the index and its box meets q */
__global__ void k_box_meets(const double* __restrict__ box, const int8_t* __restrict__ valid, int64_t ne, QueryBox q,
                            uint8_t* __restrict__ flags) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    flags[e] = (uint8_t)(valid[e] && boxes_meet(box + 6 * e, box + 6 * e + 3, q));
}

/* tets of the candidate elements, refined as PSI.voxels refines them but on
 * the raw vertex positions (src/cpsi.c:217-224: no periodic wrap); g's
 * window is unbounded so every candidate expands */
__global__ void k_vdf_count(MeshDev m, GridInfo g, MeshBox mb, const int32_t* __restrict__ cand, int64_t nc,
                            int maxlvl, double tol2, int64_t* __restrict__ counts) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    RefineScratch S;
    counts[c] = expand_element(m.pos, m.vel, m.mass, m.conn, m.vpe, cand[c], g, mb, maxlvl, tol2, S, nullptr);
}

__global__ void k_vdf_emit(MeshDev m, GridInfo g, MeshBox mb, const int32_t* __restrict__ cand, int64_t nc,
                           int maxlvl, double tol2, const int64_t* __restrict__ offs, double* __restrict__ tets) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    RefineScratch S;
    expand_element(m.pos, m.vel, m.mass, m.conn, m.vpe, cand[c], g, mb, maxlvl, tol2, S, tets + TREC * offs[c]);
}

/* per tet: out[5t] = 1 when x lies in it, then rho and v */
__global__ void k_vdf_sample(const double* __restrict__ tets, int64_t nt, QueryBox x, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double rho = 0.0, v[3] = {0.0, 0.0, 0.0};
    const int in = sample_tet(tets + TREC * t, x.lo, &rho, v);
    out[5 * t] = in ? 1.0 : 0.0;
    out[5 * t + 1] = rho;
    for (int a = 0; a < 3; ++a) out[5 * t + 2 + a] = v[a];
}

/* per-block reduction of two counters into global ones */
__device__ void block_count(int64_t a, int64_t b, unsigned long long* out) {
    __shared__ unsigned long long sa[32], sb[32];
    const int lane = threadIdx.x & (warpSize - 1), wid = threadIdx.x / warpSize;
    unsigned long long va = (unsigned long long)a, vb = (unsigned long long)b;
    for (int o = warpSize / 2; o > 0; o >>= 1) {
        va += __shfl_down(va, o);
        vb += __shfl_down(vb, o);
    }
    if (lane == 0) {
        sa[wid] = va;
        sb[wid] = vb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long ta = 0, tb = 0;
        const int nw = (blockDim.x + warpSize - 1) / warpSize;
        for (int w = 0; w < nw; ++w) {
            ta += sa[w];
            tb += sb[w];
        }
        /* every block adds to the same two words, so the adds serialize:
         * leave out the zeros (most blocks have no wide tet, and the count
         * pass has no fragments) */
        if (ta) atomicAdd(out, ta);
        if (tb) atomicAdd(out + 1, tb);
    }
    /* thread 0 has read sa, sb before a following call writes them again */
    __syncthreads();
}

/* one thread per tet record, FP64 throughout.  The polyhedra live in
 * scratch, so the kernel waits on memory: capped at 128 VGPRs it holds 4
 * wavefronts per SIMD instead of 2 (2.43 -> 1.70 s on the 128^3 -> 256^3
 * m, v, vv workload; 8 wavefronts spill too much, 3.18 s) */
__global__ __launch_bounds__(256, 4) void k_vox_plain(const double* __restrict__ tets, int64_t ntet, GridInfo g,
                                                   MeshBox mb, double* __restrict__ acc, unsigned* status,
                                                   unsigned long long* counters) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nf = 0, nc = 0;
    if (t < ntet) {
        const double* r = tets + (int64_t)TREC * t;
        if (r[25] != 0.0) {
            PlainScratch W;
            AtomicSink sink{acc};
            unsigned st = 0;
            nf = voxelize_tet64(sink, g, mb, r, W, st, &nc);
            if (st) atomicOr(status, st);
        }
    }
    block_count(nf, nc, counters);
}

/* ------------------------------------------------------------------ */
/* mode='annihilation'                                                 */
/* ------------------------------------------------------------------ */

/* sort keys of the broad phase: the box's low x, +inf for elements outside
 * the window (they sort last and end every sweep) */
__global__ void k_sweep_keys(const double* __restrict__ box, const int8_t* __restrict__ valid, int64_t ne,
                             double* __restrict__ key, int32_t* __restrict__ id) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    key[e] = valid[e] ? box[6 * e] : HUGE_VAL;
    id[e] = (int32_t)e;
}

/* Broad phase by sort and sweep along x: element id[i] meets every later
 * id[j] whose low x is at most its high x and whose y, z intervals meet its
 * own (closed, as the reference's tree query, src/rtree.c:106-110).  Each
 * such pair is two ordered pairs; every element also pairs with itself.
 * emit == nullptr: count into cnt[i]; else write the pairs at off[i]. */
__global__ void k_sweep_pairs(const double* __restrict__ key, const int32_t* __restrict__ id,
                              const double* __restrict__ box, int64_t ne, int64_t* __restrict__ cnt,
                              const int64_t* __restrict__ off, int2* __restrict__ emit) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    if (!(key[i] < HUGE_VAL)) {
        if (!emit) cnt[i] = 0;
        return;
    }
    const int32_t e = id[i];
    const double* b = box + 6 * (int64_t)e;
    const double hx = b[3];
    int64_t n = 0, o = emit ? off[i] : 0;
    if (emit) emit[o++] = make_int2(e, e);
    n++;
    for (int64_t j = i + 1; j < ne && key[j] <= hx; ++j) {
        const int32_t f = id[j];
        const double* c = box + 6 * (int64_t)f;
        if (c[4] < b[1] || c[1] > b[4] || c[5] < b[2] || c[2] > b[5]) continue;
        if (emit) {
            emit[o++] = make_int2(e, f);
            emit[o++] = make_int2(f, e);
        }
        n += 2;
    }
    if (!emit) cnt[i] = n;
}

/* Tet pairs of element pair p = (e, e1): tets [offA[e], offA[e+1]) of the
 * wrapped expansion against [offB[e1], offB[e1+1]) of the raw one, kept
 * when their boxes overlap with positive volume (src/cpsi.c:138-152).
 * emit == nullptr: count; else write (tet a, tet b) at toff[p] - base. */
__global__ void k_tet_pairs(const int2* __restrict__ pairs, int64_t p0, int64_t p1, const int64_t* __restrict__ offA,
                            const int64_t* __restrict__ offB, const double* __restrict__ ta,
                            const double* __restrict__ tb, int64_t* __restrict__ cnt,
                            const int64_t* __restrict__ toff, int64_t base, uint2* __restrict__ emit) {
    const int64_t p = p0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= p1) return;
    const int2 pr = pairs[p];
    int64_t n = 0, o = emit ? toff[p] - base : 0;
    for (int64_t t = offA[pr.x]; t < offA[pr.x + 1]; ++t)
        for (int64_t u = offB[pr.y]; u < offB[pr.y + 1]; ++u) {
            double lo[3], hi[3];
            if (!tet_box_overlap(ta + TREC * t, tb + TREC * u, lo, hi)) continue;
            if (emit) emit[o++] = make_uint2((unsigned)t, (unsigned)u);
            n++;
        }
    if (!emit) cnt[p] = n;
}

struct MassSink {
    double* m;
    __device__ void add(int64_t i, double v) { unsafeAtomicAdd(m + i, v); }
};

/* one thread per tet pair, FP64: the pair's overlap swept through the cells */
__global__ __launch_bounds__(128) void k_annihilate(const double* __restrict__ ta, const double* __restrict__ tb,
                                                    const uint2* __restrict__ tasks, int64_t n, GridInfo g,
                                                    double* __restrict__ m, unsigned* status,
                                                    unsigned long long* counters) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nf = 0;
    if (t < n) {
        const uint2 k = tasks[t];
        const double* ra = ta + (int64_t)TREC * k.x;
        const double* rb = tb + (int64_t)TREC * k.y;
        double lo[3], hi[3];
        if (tet_box_overlap(ra, rb, lo, hi)) {
            PlainScratch W;
            MassSink sink{m};
            unsigned st = 0;
            nf = annihilate_pair64(sink, g, ra, rb, lo, hi, W, st);
            if (st) atomicOr(status, st);
        }
    }
    block_count(nf, 0, counters);
}

/* exact in-window mass fraction of each tet (only when the window is not
 * the whole periodic box) */
__global__ __launch_bounds__(64) void k_targets(double* __restrict__ tets, int64_t ntet, GridInfo g, MeshBox mb) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntet) return;
    double* r = tets + (int64_t)TREC * t;
    if (r[25] == 0.0) return;
    Poly64 pool[2];
    r[26] = window_target(r, g, mb, pool);
}

/* One entry per mesh of a batched call: its grid, box and accumulator.
 * Tet records carry their entry's index in slot 27. */
struct BatchGeo {
    GridInfo g;
    MeshBox mb;
    double* acc;
};

/* f(g, mb, acc, id) with the geometry of record r's mesh.  Batched: a
 * waterfall over the distinct meshes among the wave's active lanes, each
 * table entry read at a wave-uniform index (scalar loads). */
template <bool BATCH, class F>
__device__ __forceinline__ void with_mesh(const double* r, const GridInfo& g0, const MeshBox& mb0, double* acc0,
                                          const BatchGeo* table, F&& f) {
    if (!BATCH) {
        f(g0, mb0, acc0, 0);
        return;
    }
    const int my = (int)r[27];
    for (;;) {
        const int id = __builtin_amdgcn_readfirstlane(my);
        if (my == id) {
            const BatchGeo* e = table + id;
            f(e->g, e->mb, e->acc, id);
            break;
        }
    }
}

/* per-mesh counters of a batched call: fragments, copies, wide tets; each
 * in MSTRIPE stripes one 128-byte line apart (stripe: the workgroup's
 * index), summed on the host: the wavefronts of one mesh would otherwise
 * all add to one word, and such adds serialize */
constexpr int MCNT = 4;
constexpr int MSTRIPE = 16;
constexpr int MLINE = 16; /* words per 128-byte line */

PVX_HD int64_t mcnt_words(int nm) { return (int64_t)MCNT * nm * MSTRIPE * MLINE; }
PVX_HD int64_t mcnt_word(int id, int slot, int stripe) {
    return (((int64_t)MCNT * id + slot) * MSTRIPE + stripe) * MLINE;
}

/* per-mesh counter add, aggregated over the wave when its lanes share a
 * mesh; every lane of the wave calls it (id < 0: nothing) */
__device__ void mesh_count(unsigned long long* mcnt, int id, int64_t v, int slot) {
    const int id0 = __builtin_amdgcn_readfirstlane(id);
    const int stripe = (int)(blockIdx.x & (MSTRIPE - 1));
    if (__all(id == id0)) {
        unsigned long long sum = (unsigned long long)v;
        for (int o = warpSize / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if ((threadIdx.x & (warpSize - 1)) == 0 && id0 >= 0 && sum) atomicAdd(mcnt + mcnt_word(id0, slot, stripe), sum);
    } else if (id >= 0 && v) {
        atomicAdd(mcnt + mcnt_word(id, slot, stripe), (unsigned long long)v);
    }
}

/* Column passes.  PASS_ALL: every tet in FP32 (kernel FAST).  Under the
 * automatic choice, PASS_NARROW takes the tets that are not tet_wide in
 * FP32 and counts the wide ones; PASS_WIDE takes only those, in FP64.
 * PASS_ALL64: every tet in FP64 (kernel FAST64, the velocity adjoint). */
enum { PASS_ALL = 0, PASS_NARROW = 1, PASS_WIDE = 2, PASS_ALL64 = 3 };

/* whether a pass's columns are built in FP64 */
PVX_HD bool pass_fp64(int pass) { return pass == PASS_WIDE || pass == PASS_ALL64; }

/* column tasks per tet record for one pass (0 for invalid records and for
 * tets of the other pass); a tet the column path cannot take flags the
 * FP64 fallback.  counters[2] gets the wide tets PASS_NARROW leaves out. */
/* words before the wide-tet list in its buffer: its counter, alone in a
 * 64-byte line */
constexpr int WLIST_HEAD = 16;

template <bool BATCH, int PASS>
__global__ __launch_bounds__(256) void k_col_count(const double* __restrict__ tets, int64_t ntet, GridInfo g0,
                                                   MeshBox mb0, const BatchGeo* table, int64_t* __restrict__ cnt,
                                                   unsigned* status, unsigned long long* counters,
                                                   unsigned long long* mcnt, uint32_t* __restrict__ wlist) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int nc = 0, id = -1, nwide = 0;
    if (t < ntet) {
        const double* r = tets + (int64_t)TREC * t;
        int64_t n = 0;
        if (r[25] != 0.0) {
            with_mesh<BATCH>(r, g0, mb0, nullptr, table, [&](const GridInfo& g, const MeshBox& mb, double*, int m) {
                const bool wide = (PASS == PASS_NARROW || PASS == PASS_WIDE) && tet_wide(r, g);
                if (PASS == PASS_NARROW && wide)
                    nwide = 1;
                else if (PASS != PASS_WIDE || wide)
                    n = pass_fp64(PASS) ? tet_columns<double>(r, t, g, mb, nullptr, &nc)
                                        : tet_columns<float>(r, t, g, mb, nullptr, &nc);
                id = m;
            });
            if (n < 0) {
                atomicOr(status, ST_FALLBACK);
                n = 0;
            }
        }
        cnt[t] = n;
    }
    block_count(0, nc, counters);
    if (PASS == PASS_NARROW) block_count(nwide, 0, counters + 2);
    /* PASS_NARROW with a list: the wide tets' indices after its 64-byte
     * counter (a line of its own: the tallies' line takes every block's
     * adds), a workgroup's in thread order at a place taken with one add */
    if (PASS == PASS_NARROW && wlist) {
        __shared__ unsigned wsum[4];
        __shared__ unsigned long long wbase;
        const uint64_t m = __ballot(nwide != 0);
        const int wv = (int)(threadIdx.x / warpSize), lane = (int)(threadIdx.x & (warpSize - 1));
        if (lane == 0) wsum[wv] = (unsigned)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
            wbase = tot ? atomicAdd((unsigned long long*)wlist, (unsigned long long)tot) : 0ull;
        }
        __syncthreads();
        if (nwide) {
            unsigned off = 0;
            for (int k = 0; k < wv; ++k) off += wsum[k];
            wlist[WLIST_HEAD + wbase + off + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)t;
        }
    }
    if (BATCH) {
        mesh_count(mcnt, id, nc, 1);
        if (PASS == PASS_NARROW) mesh_count(mcnt, id, nwide, 2);
    }
}

/* write the column tasks of tets [t0, t1) at colstart[t] - base (the tets
 * with tasks in this pass; WIDE: enumerated for the FP64 kernel) */
template <bool BATCH, bool WIDE>
__global__ __launch_bounds__(256) void k_col_emit(const double* __restrict__ tets, int64_t t0, int64_t t1, GridInfo g0,
                                                  MeshBox mb0, const BatchGeo* table,
                                                  const int64_t* __restrict__ colstart, int64_t base,
                                                  uint64_t* __restrict__ tasks) {
    const int64_t t = t0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1) return;
    const double* r = tets + (int64_t)TREC * t;
    if (r[25] == 0.0 || colstart[t + 1] == colstart[t]) return;
    with_mesh<BATCH>(r, g0, mb0, nullptr, table, [&](const GridInfo& g, const MeshBox& mb, double*, int) {
        if (WIDE)
            tet_columns<double>(r, t, g, mb, tasks + (colstart[t] - base), nullptr);
        else
            tet_columns<float>(r, t, g, mb, tasks + (colstart[t] - base), nullptr);
    });
}

/* wide[TREC i + q] = record wlist[i], slot q, for the n listed tets */
__global__ __launch_bounds__(256) void k_wide_copy(const double* __restrict__ tets, const uint32_t* __restrict__ wlist,
                                                   int64_t n, double* __restrict__ wide) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * TREC) return;
    const int64_t k = i / TREC;
    wide[i] = tets[(int64_t)TREC * wlist[k] + (i - k * TREC)];
}

/* per-task outputs for the residual pass: FP64 fragment-mass sum, the
 * largest fragment's key and (ORDER > 0) its moments and origin */
struct TaskOut {
    double* sum;
    uint64_t* key;
    float* mom; /* 16 floats per task: m0, s1[3], s2[6], O[3], unused */
};

/* Tasks per ordering window and key classes.  A lane's work grows with its
 * column's cell count and vertex count, and a wavefront runs as long as its
 * longest lane; the column kernel takes its tasks through perm, which orders
 * each window of consecutive tasks by (cells, vertices), largest first.
 * Both counts are clamped to 15: 256 classes.  Wider windows fill
 * wavefronts with more nearly equal lanes: on the Zel'dovich 128^3 ->
 * 256^3 workload the column pass takes 56.7 ms with windows of 1024 tasks,
 * 53.0 ms with 4096, 51.7 ms with 16384 and 51.4 ms with 65536 (whose
 * ordering costs more than that saves); the 64-mesh batch, 530 -> 488 ms
 * (a wavefront spans two meshes only where a window crosses a mesh
 * boundary). */
constexpr int ORDW = 16384;
constexpr int ORDK = 256;
static_assert(ORDW % 256 == 0, "whole threads per window");

/* The column kernel's wavefront counters, one per XCD (an MI355X has 8,
 * each with its own L2; the hardware places workgroup b on XCD b mod 8),
 * 128 bytes apart.  Counter x hands out the ordering windows x, x + 8,
 * x + 16, ... wavefront by wavefront, so the tet records of a window are
 * read into one XCD's L2 instead of all eight; a workgroup whose own
 * windows are done takes from the other counters in turn. */
constexpr int NXCD = 8;
constexpr int XSTRIDE = 32;
constexpr int64_t XCH = ORDW / FBLOCK; /* wavefronts per window */

/* the next wavefront (index into the ordered tasks, in units of FBLOCK) for
 * a workgroup on XCD home that has exhausted the counters home .. home +
 * cur - 1; nwaves when all are exhausted */
__device__ int64_t claim_wave(unsigned* next, int home, int& cur, int64_t nwaves) {
    while (cur < NXCD) {
        const int x = home + cur < NXCD ? home + cur : home + cur - NXCD;
        const unsigned k = atomicAdd(next + x * XSTRIDE, 1u);
        const int64_t w = ((int64_t)(k / XCH) * NXCD + x) * XCH + (int64_t)(k % XCH);
        if (w < nwaves) return w;
        ++cur;
    }
    return nwaves;
}

/* One window of WIN consecutive tasks (the workgroup's, NT >= 256
 * threads), counting-sorted by class in LDS: key(e) gives task w0 + e's
 * class (0 is processed first), out(e, pos) places task w0 + e at w0 + pos.
 * Order within a class is arbitrary. */
template <int WIN, int NT, class Key, class Out>
__device__ void window_sort(int64_t ntask, Key&& key, Out&& out) {
    static_assert(NT >= ORDK, "one thread per class in the scan");
    __shared__ int cnt[ORDK];
    __shared__ uint8_t skey[WIN]; /* each task's class, kept here between the two passes */
    const int64_t w0 = (int64_t)blockIdx.x * WIN;
    const int nw = (int)(ntask - w0 < WIN ? ntask - w0 : WIN);
    const int tid = (int)threadIdx.x;
    if (tid < ORDK) cnt[tid] = 0;
    __syncthreads();
    for (int e = tid; e < nw; e += NT) {
        const int k = key(e);
        skey[e] = (uint8_t)k;
        atomicAdd(&cnt[k], 1);
    }
    __syncthreads();
    /* exclusive scan of the 256 class counts (Hillis-Steele) */
    int v = tid < ORDK ? cnt[tid] : 0;
    for (int d = 1; d < ORDK; d <<= 1) {
        __syncthreads();
        const int u = tid < ORDK && tid >= d ? cnt[tid - d] : 0;
        __syncthreads();
        v += u;
        if (tid < ORDK) cnt[tid] = v;
    }
    __syncthreads();
    const int ex = tid > 0 && tid < ORDK ? cnt[tid - 1] : 0;
    __syncthreads();
    if (tid < ORDK) cnt[tid] = ex;
    __syncthreads();
    for (int e = tid; e < nw; e += NT) out(e, atomicAdd(&cnt[skey[e]], 1));
}

/* threads per window of the column ordering: the whole Zel'dovich 128^3 ->
 * 256^3 call takes 40.49 ms with 256, 40.30 ms with 512, 41.17 ms with 1024 */
constexpr int ORDNT = 512;

/* perm[w0 .. w0 + ORDW) of each window: the window's task indices, largest
 * class first, and sorted[] the task words in the same order (the column
 * kernel then reads both coalesced and independently).  Also clears the
 * column kernel's wavefront counters.  ORDNT threads a workgroup. */
template <bool BATCH>
__global__ __launch_bounds__(ORDNT) void k_col_order(const double* __restrict__ tets, const uint64_t* __restrict__ tasks,
                                                   int64_t ntask, GridInfo g0, MeshBox mb0, const BatchGeo* table,
                                                   uint32_t* __restrict__ perm, uint64_t* __restrict__ sorted,
                                                   uint16_t* __restrict__ mid, unsigned* __restrict__ next) {
    const int64_t w0 = (int64_t)blockIdx.x * ORDW;
    if (blockIdx.x == 0 && threadIdx.x < NXCD) next[threadIdx.x * XSTRIDE] = 0u;
    window_sort<ORDW, ORDNT>(
        ntask,
        [&](int e) {
            const uint64_t w = tasks[w0 + e];
            int64_t t;
            int h, il, jl;
            unpack_task(w, t, h, il, jl);
            const double* r = tets + (int64_t)TREC * t;
            int nc = 0, nv = 0;
            with_mesh<BATCH>(r, g0, mb0, nullptr, table, [&](const GridInfo& g, const MeshBox& mb, double*, int) {
                double lo[3], hi[3];
                aabb((const double (*)[3])r, 4, lo, hi);
                uint32_t shp;
                ghost_shifts_packed(lo, hi, g, mb, shp);
                CopySetup<float> C;
                if (copy_setup(r, lo, hi, shp, h, g, mb, C)) nc = column_cells(C, il, jl, &nv);
            });
            nc = nc < 15 ? nc : 15;
            nv = nv < 15 ? nv : 15;
            return ORDK - 1 - (16 * nc + nv);
        },
        [&](int e, int pos) {
            const uint64_t w = tasks[w0 + e];
            perm[w0 + pos] = (uint32_t)(w0 + e);
            sorted[w0 + pos] = w;
            /* batched: the task's mesh, in processing order (the column
             * kernel then need not read it from the scattered tet record
             * first) */
            if (BATCH) mid[w0 + pos] = (uint16_t)tets[(int64_t)TREC * (int64_t)(w & 0xffffffffull) + 27];
        });
}

/* The adjoint's plane tasks of each window sorted by the cells their
 * section spans, most first (adj_task_cells): a lane's work grows with
 * that count and a wavefront runs as long as its longest lane (in emit
 * order the lanes did 37% of the longest lane's work on the Zel'dovich
 * 128^3 -> 256^3 workload, 98% sorted; host model of the lanes).  1024
 * threads a window: a call's few million tasks make only hundreds of
 * windows (windows of 2048 tasks made the plane-task kernel itself 1.8x
 * slower). */
__global__ __launch_bounds__(1024) void k_adj_order(const double* __restrict__ tets, const uint64_t* __restrict__ tasks,
                                                    int64_t ntask, GridInfo g, MeshBox mb,
                                                    uint64_t* __restrict__ sorted) {
    const int64_t w0 = (int64_t)blockIdx.x * ORDW;
    window_sort<ORDW, 1024>(
        ntask,
        [&](int e) {
            const uint64_t w = tasks[w0 + e];
            const int n = adj_task_cells(g, mb, tets + (int64_t)TREC * (int64_t)(w & 0xffffffffull), w);
            return ORDK - 1 - (n < ORDK - 1 ? n : ORDK - 1);
        },
        [&](int e, int pos) { sorted[w0 + pos] = tasks[w0 + e]; });
}

/* A fragment's deposit staged in LDS: add(i, v) puts v at row[i - base]
 * (deposit_f adds each of the cell's nacc slots exactly once). */
struct StageSink {
    double* row;
    int64_t base;
    __device__ void add(int64_t i, double v) { row[i - base] = v; }
};

/* bytes of a deposit stage of nf fragments: nacc values each, then their
 * cell offsets */
static size_t stage_bytes(const GridInfo& g, int nf) { return (size_t)nf * (g.nacc + 1) * sizeof(double); }

/* fragments the column kernel stages at once: half a wavefront, so that the
 * stage (5.6 KB for m, v, vv) costs no occupancy next to the polyhedra */
constexpr int SFRAG = FBLOCK / 2;

/* field flags of grids without position moments (x, xx, xv): the column
 * kernel built for them keeps no position data through its cell loop */
constexpr int FMASK_NOX = ~(F_X | F_XX | F_XV);

/*
 * The work of one wavefront on the tasks of its lanes with act set, all of
 * one mesh (geometry g, mb, accumulator acc): build each lane's column in
 * its LDS slice, step the cells in lockstep, stage each step's fragments
 * in LDS and issue their FP64 atomics slot by slot, and write each task's
 * tally (fragment-mass sum, largest fragment) for the residual pass.
 * Every lane of the wavefront calls it.
 */
template <int ORDER, int FMASK, class R>
__device__ __forceinline__ int64_t wave_columns(bool act, int64_t p, uint64_t w, const double* r, const GridInfo& g,
                                                const MeshBox& mb, double* __restrict__ acc, LPoly<R>& P,
                                                double* stage, TaskOut out, unsigned* status) {
    const int na = g.nacc;
    int64_t* sgi = (int64_t*)(stage + SFRAG * na);
    const int lane = threadIdx.x;
    /* a lane's first staged value in a round of the transposed deposit, as
     * (fragment, slot), and the step to its next */
    const int f0 = lane / na, j0 = lane - f0 * na, fstep = FBLOCK / na, jstep = FBLOCK - fstep * na;
    /* the task's tally: fragments, FP64 mass sum and its largest fragment
     * (moments written to out.mom as they improve) */
    int tfrag = 0;
    double tsum = 0.0;
    float bm0 = 0.0f;
    int bkl = -1;
    ColumnState<R> S;
    TetVel<R> V;
    int ncell = 0;
    if (act) {
        const int st = column_begin(g, mb, r, w, P, S);
        if (st < 0) atomicOr(status, ST_FALLBACK);
        if (st > 0) {
            ncell = S.k1 - S.k0;
            if (ORDER > 0) tet_vel(r, V);
            V.mass = r[24];
            pin(V.mass);
        }
    }
    int nmax = ncell;
    for (int o = FBLOCK / 2; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
    PlaneCarry<R> pc{false, false, {R(0), R(0), R(0)}};
    /* a column of two cells takes both in its first step (one pass over
     * its fan triangles); its second step deposits the upper cell, Mu */
    const bool pair = ncell == 2 && two_cell_column(S);
    Mom<R> Mu;
    R O[3];
    /* cell steps in lockstep: every lane takes part in each step's
     * deposit, which issues the wavefront's fragments slot by slot so
     * that neighbouring lanes add into the same cell's consecutive
     * slots (the memory-side atomic units then take several lanes per
     * request; one lane per cell would give one request per lane) */
    for (int it = 0; it < nmax; ++it) {
        Mom<R> Mm;
        int k = 0;
        bool has = false;
        if (it < ncell) {
            k = S.k0 + it;
            if (pair) {
                if (it == 0)
                    slab_moments_2<ORDER>(P, S.C.G, 2, first_plane(S), Mm, Mu, O);
                else
                    Mm = Mu;
            } else {
                bool hl, hh;
                double lc, hc;
                cell_planes(S, k, hl, lc, hh, hc);
                slab_moments_u<ORDER>(P, S.C.G, 2, hl, lc, hh, hc, Mm, O, &pc);
            }
            has = Mm.m0 > R(0);
        }
        const uint64_t mask = __ballot(has);
        if (mask == 0ull) continue;
        const int nf = __popcll(mask);
        const int rank = __popcll(mask & ((1ull << lane) - 1ull));
        /* the step's fragments in groups of SFRAG: stage, then issue */
        for (int r0 = 0; r0 < nf; r0 += SFRAG) {
            if (has && rank >= r0 && rank < r0 + SFRAG) {
                const int64_t gi = ((int64_t)g.n[1] * g.n[2] * S.i + (int64_t)g.n[2] * S.j + k) * na;
                StageSink ss{stage + (rank - r0) * na, gi};
                deposit_f<ORDER, FMASK>(ss, g, S.C.G, S.C.G.ib0, S.i, S.j, k, Mm, O, V.u0, (const R (*)[3])V.DV,
                                        V.mass, (double)Mm.m0);
                sgi[rank - r0] = gi;
            }
            __syncthreads();
            const int np = (nf - r0 < SFRAG ? nf - r0 : SFRAG) * na;
            int f = f0, j = j0;
            for (int q = lane; q < np; q += FBLOCK) {
                unsafeAtomicAdd(acc + sgi[f] + j, stage[q]);
                f += fstep;
                j += jstep;
                if (j >= na) {
                    j -= na;
                    ++f;
                }
            }
            __syncthreads(); /* the stage is read before it is written again */
        }
        if (has) {
            tfrag++;
            tsum += (double)Mm.m0;
            if ((float)Mm.m0 > bm0) {
                bm0 = (float)Mm.m0;
                bkl = k - S.C.G.ib0[2];
                if (ORDER > 0) {
                    float4* m = (float4*)(out.mom + 16 * p);
                    m[0] = make_float4((float)Mm.m0, (float)Mm.s1[0], (float)Mm.s1[1], (float)Mm.s1[2]);
                    m[1] = make_float4((float)Mm.s2[0], (float)Mm.s2[1], (float)Mm.s2[2], (float)Mm.s2[3]);
                    m[2] = make_float4((float)Mm.s2[4], (float)Mm.s2[5], (float)O[0], (float)O[1]);
                    m[3] = make_float4((float)O[2], 0.f, 0.f, 0.f);
                }
            }
        }
    }
    if (act) {
        out.sum[p] = tsum;
        out.key[p] = bkl >= 0 ? best_key(bm0, bkl) : 0ull;
    }
    return tfrag;
}

/*
 * One lane per column task: column clip in this lane's LDS slice (R =
 * float, or double for the wide tets of the automatic choice), cell
 * moments, FP64 deposit.  The grid is as many workgroups as the chip holds
 * at once; each takes wavefronts of tasks (perm order) from the counter
 * until it runs past the end.  The hardware places workgroups on XCDs and
 * CUs in a fixed rotation, and the XCDs do not run at quite the same speed:
 * with a fixed mapping the slowest XCD would set the time.
 * BATCH: the tasks of many meshes in one launch.  A task's mesh is its tet
 * record's slot 27; a wavefront works through the meshes among its lanes
 * one at a time (tasks are in tet order and tets in mesh order, so only a
 * wavefront at a boundary between meshes has more than one), with that
 * mesh's table entry in scalar registers.
 */
template <int ORDER, int FMASK, class R, bool BATCH>
__global__ __launch_bounds__(FBLOCK, ORDER == 0 && sizeof(R) == 4 ? 3 : 2) void k_vox_cols(const double* __restrict__ tets,
                                                        const uint64_t* __restrict__ sorted,
                                                        const uint32_t* __restrict__ perm,
                                                        const uint16_t* __restrict__ mid, unsigned* __restrict__ next,
                                                        int64_t ntask, GridInfo g0, MeshBox mb0,
                                                        double* __restrict__ acc0,
                                                        const BatchGeo* __restrict__ table, TaskOut out,
                                                        unsigned* status, unsigned long long* counters,
                                                        unsigned long long* mcnt) {
    __shared__ __align__(32) uint32_t sv[VRec<R>::WORDS];
    __shared__ unsigned swave;
    /* the deposit stage: up to SFRAG fragments of one cell step, in rank
     * order, nacc values each, then their accumulator offsets */
    extern __shared__ double stage[];
    const int lane = threadIdx.x;
    LPoly<R> P{(uint32_t*)__builtin_assume_aligned(&sv[VRec<R>::W * lane], 16), 0, 0, 0};
    int64_t nfrag = 0;
    /* each wavefront of tasks is claimed one round ahead, so the counter's
     * round trip overlaps the work on the one before */
    const int home = (int)(blockIdx.x % NXCD);
    const int64_t nwaves = (ntask + FBLOCK - 1) / FBLOCK;
    int cur = 0;
    int64_t claim = 0;
    if (lane == 0) claim = claim_wave(next, home, cur, nwaves);
    for (;;) {
        if (lane == 0) swave = (unsigned)claim;
        __syncthreads();
        const int64_t c0 = (int64_t)swave * FBLOCK;
        __syncthreads(); /* everyone has read swave before lane 0 writes it again */
        if (c0 >= ntask) break;
        if (lane == 0) claim = claim_wave(next, home, cur, nwaves);
        const int64_t c = c0 + lane;
        int64_t p = 0;
        uint64_t w = 0;
        const double* r = tets;
        int my = -1;
        if (c < ntask) {
            p = perm[c];
            w = sorted[c];
            r = tets + (int64_t)TREC * (int64_t)(w & 0xffffffffull);
            my = BATCH ? (int)mid[c] : 0;
        }
        if (!BATCH) {
            nfrag += wave_columns<ORDER, FMASK, R>(my >= 0, p, w, r, g0, mb0, acc0, P, stage, out, status);
            continue;
        }
        uint64_t pending = __ballot(my >= 0);
        while (pending) {
            const int id = __builtin_amdgcn_readlane(my, (int)__builtin_ctzll(pending));
            const bool act = my == id;
            pending &= ~__ballot(act);
            /* the entry in place at a wave-uniform address (scalar loads); a
             * local copy would live in scratch, as some of its fields are
             * indexed at run time */
            const BatchGeo* e = table + id;
            /* inactive lanes count no fragments */
            const int64_t nf = wave_columns<ORDER, FMASK, R>(act, p, w, r, e->g, e->mb, e->acc, P, stage, out, status);
            mesh_count(mcnt, id, nf, 0);
            nfrag += nf;
        }
    }
    block_count(nfrag, 0, counters);
}

/* one thread per tet of [t0, t1): how its residual is deposited.  A small
 * residual goes whole to the largest fragment: tetf[t] = the residual and
 * tetb[t] = that task; otherwise it is spread: tetf[t] = (target - sum) /
 * sum and tetb[t] = -1 (sums: the tasks' FP64 fragment-mass sums), and a
 * nonzero spread lists t - t0 in slist (count in slist[0], entries from
 * slist[WLIST_HEAD]; the count starts at 0) */
__global__ __launch_bounds__(256) void k_resid_factor(const double* __restrict__ tets, int64_t t0, int64_t t1,
                                                      const int64_t* __restrict__ colstart, int64_t base,
                                                      const double* __restrict__ tsum,
                                                      const uint64_t* __restrict__ tkey, double* __restrict__ tetf,
                                                      int64_t* __restrict__ tetb, uint32_t* __restrict__ slist) {
    const int64_t t = t0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1) return;
    const int64_t c0 = colstart[t] - base, c1 = colstart[t + 1] - base;
    /* RS interleaved streams (task c0 + RS q + u in stream u), so that a tet
     * with many tasks has RS loads of each array in flight at a time; the
     * largest key wins, the lowest task among equal keys */
    constexpr int RS = 8;
    double ps[RS];
    uint64_t pk[RS];
    int64_t pb[RS];
#pragma unroll
    for (int u = 0; u < RS; ++u) {
        ps[u] = 0.0;
        pk[u] = 0ull;
        pb[u] = -1;
    }
    for (int64_t cq = c0; cq < c1; cq += RS) {
        double v[RS];
        uint64_t k[RS];
#pragma unroll
        for (int u = 0; u < RS; ++u) {
            const bool in = cq + u < c1;
            v[u] = in ? tsum[cq + u] : 0.0;
            k[u] = in ? tkey[cq + u] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < RS; ++u) {
            ps[u] += v[u];
            if (k[u] > pk[u]) {
                pk[u] = k[u];
                pb[u] = cq + u;
            }
        }
    }
    double sum = 0.0;
    uint64_t kb = 0;
    int64_t cb = -1;
#pragma unroll
    for (int u = 0; u < RS; ++u) {
        sum += ps[u];
        if (pk[u] > kb || (pk[u] == kb && pk[u] != 0ull && pb[u] < cb)) {
            kb = pk[u];
            cb = pb[u];
        }
    }
    const double* r = tets + (int64_t)TREC * t;
    const double resid = r[26] - sum;
    float m0b = 0.f;
    const uint32_t mbits = (uint32_t)(kb >> 32);
    memcpy(&m0b, &mbits, 4);
    bool spread = false;
    if (cb >= 0 && fabs(resid) <= RESID_WHOLE * (double)m0b) {
        tetf[t] = resid;
        tetb[t] = cb + base;
    } else {
        const double f = residual_factor(r, sum);
        tetf[t] = f;
        tetb[t] = -1;
        spread = f != 0.0;
    }
    /* one atomic per wavefront for its spread tets' list slots */
    const uint64_t bal = __ballot(spread);
    if (bal) {
        const int lane = (int)(threadIdx.x & 63u), lead = __ffsll((unsigned long long)bal) - 1;
        int b0 = 0;
        if (lane == lead) b0 = (int)atomicAdd(slist, (unsigned)__popcll(bal));
        b0 = __shfl(b0, lead);
        if (spread) slist[WLIST_HEAD + b0 + __popcll(bal & ((1ull << lane) - 1ull))] = (uint32_t)(t - t0);
    }
}

/* the largest fragment of task c from its key (cell offset, FP32 mass) and,
 * for ORDER > 0, its moments and origin */
template <int ORDER>
__device__ BestFrag best_frag(uint64_t kb, const float* __restrict__ mom, int64_t c) {
    BestFrag B;
    B.kl = (int)(kb & 0xffffffffull);
    const uint32_t mbits = (uint32_t)(kb >> 32);
    memcpy(&B.M.m0, &mbits, 4);
    if (ORDER > 0) {
        const float4* m = (const float4*)(mom + 16 * c);
        const float4 a = m[0], b = m[1], cc = m[2], d = m[3];
        B.M.s1[0] = a.y;
        B.M.s1[1] = a.z;
        B.M.s1[2] = a.w;
        B.M.s2[0] = b.x;
        B.M.s2[1] = b.y;
        B.M.s2[2] = b.z;
        B.M.s2[3] = b.w;
        B.M.s2[4] = cc.x;
        B.M.s2[5] = cc.y;
        B.O[0] = cc.z;
        B.O[1] = cc.w;
        B.O[2] = d.x;
    }
    return B;
}

/* A deposit staged in LDS when its base is not known beforehand: the
 * first add (deposit_f's mass slot, offset 0) sets it. */
struct RowSink {
    double* row;
    int64_t base;
    __device__ void add(int64_t i, double v) {
        if (base < 0) base = i;
        row[i - base] = v;
    }
};

/* The whole residuals, staged: one thread per tet of [t0, t1) whose
 * residual goes whole to its largest fragment (tetb >= 0), for that
 * fragment's task.  The lanes that deposit stage their rows (nacc values)
 * in LDS in rank order; the wavefront then issues the staged rows slot by
 * slot, so that neighbouring lanes add into the same cell's consecutive
 * slots, and the issue loop covers only the rows staged. */
template <int ORDER, bool BATCH>
__global__ __launch_bounds__(256) void k_resid_whole(const double* __restrict__ tets, const uint64_t* __restrict__ tasks,
                                                     int64_t base, int64_t t0, int64_t t1, GridInfo g0, MeshBox mb0,
                                                     double* __restrict__ acc0, const BatchGeo* table, TaskOut out,
                                                     const double* __restrict__ tetf,
                                                     const int64_t* __restrict__ tetb) {
    extern __shared__ double rstage[]; /* per wavefront: FBLOCK rows of nacc values, FBLOCK targets */
    const int na = g0.nacc;
    const int lane = threadIdx.x & (FBLOCK - 1), wv = threadIdx.x / FBLOCK;
    double* rows = rstage + (size_t)wv * FBLOCK * (na + 1);
    double** dst = (double**)(rows + FBLOCK * na);
    double* target = nullptr;
    const int64_t t = t0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool will = false;
    int64_t c = 0;
    uint64_t kb = 0, w = 0;
    double weight = 0.0;
    const int64_t cb = t < t1 ? tetb[t] : -1;
    if (cb >= 0) {
        c = cb - base;
        kb = out.key[c];
        w = tasks[c];
        weight = tetf[t];
        will = kb != 0ull && weight != 0.0;
    }
    const uint64_t active = __ballot(will);
    const int nact = __popcll(active);
    const int rank = __popcll(active & ((1ull << lane) - 1ull));
    if (will) {
        const BestFrag B = best_frag<ORDER>(kb, out.mom, c);
        const double* r = tets + (int64_t)TREC * t;
        with_mesh<BATCH>(r, g0, mb0, acc0, table, [&](const GridInfo& g, const MeshBox& mb, double* acc, int) {
            RowSink rs{rows + rank * na, -1};
            task_residual<ORDER>(rs, g, mb, r, w, B, weight);
            if (rs.base >= 0) target = acc + rs.base;
        });
        dst[rank] = target;
    }
    __syncthreads();
    /* value q = lane + FBLOCK s of the wavefront's staged rows: row q / na, slot q % na */
    const int fstep = FBLOCK / na, jstep = FBLOCK - fstep * na;
    int f = lane / na, j = lane - (lane / na) * na;
    for (int q = lane; q < nact * na; q += FBLOCK) {
        double* d = dst[f];
        if (d) unsafeAtomicAdd(d + j, rows[q]);
        f += fstep;
        j += jstep;
        if (j >= na) {
            j -= na;
            ++f;
        }
    }
}

/* The spread residuals: the tets k_resid_factor listed, one wavefront per
 * listed tet (grid-stride), its lanes over the tet's tasks, each lane
 * depositing its task's share (weight = factor x the task's mass sum) with
 * direct atomics.  Few tets spread their residual (one in 238,000 on the
 * Zel'dovich 128^3 -> 256^3 workload, host model of the lanes), and the
 * list spares reading every task's key to find them. */
template <int ORDER, bool BATCH>
__global__ __launch_bounds__(256) void k_resid_spread(const double* __restrict__ tets, const uint64_t* __restrict__ tasks,
                                                      int64_t base, int64_t t0, const int64_t* __restrict__ colstart,
                                                      GridInfo g0, MeshBox mb0, double* __restrict__ acc0,
                                                      const BatchGeo* table, TaskOut out,
                                                      const double* __restrict__ tetf,
                                                      const uint32_t* __restrict__ slist) {
    const int64_t n = (int64_t)slist[0];
    const int lane = (int)(threadIdx.x & (FBLOCK - 1));
    const int64_t nwv = (int64_t)gridDim.x * (blockDim.x / FBLOCK);
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / FBLOCK; i < n; i += nwv) {
        const int64_t t = t0 + (int64_t)slist[WLIST_HEAD + i];
        const int64_t c0 = colstart[t] - base, c1 = colstart[t + 1] - base;
        const double f = tetf[t];
        const double* r = tets + (int64_t)TREC * t;
        for (int64_t c = c0 + lane; c < c1; c += FBLOCK) {
            const uint64_t kb = out.key[c];
            const double weight = f * out.sum[c];
            if (kb == 0ull || weight == 0.0) continue;
            const BestFrag B = best_frag<ORDER>(kb, out.mom, c);
            const uint64_t w = tasks[c];
            with_mesh<BATCH>(r, g0, mb0, acc0, table, [&](const GridInfo& g, const MeshBox& mb, double* acc, int) {
                AtomicSink as{acc};
                task_residual<ORDER>(as, g, mb, r, w, B, weight);
            });
        }
    }
}

/* the geometry of one column pass: one mesh, or a table of meshes */
struct PassGeo {
    GridInfo g;
    MeshBox mb;
    double* acc;
    const BatchGeo* table; /* non-null: batched */
    unsigned long long* mcnt;
};

typedef void (*KColsFn)(const double*, const uint64_t*, const uint32_t*, const uint16_t*, unsigned*, int64_t, GridInfo,
                        MeshBox, double*, const BatchGeo*, TaskOut, unsigned*, unsigned long long*, unsigned long long*);

/* workgroups of a persistent one-wavefront kernel (the column kernel, the
 * adjoint's velocity tasks) the current device holds at once, with sb
 * bytes of dynamic LDS */
template <class K>
static int64_t resident_blocks(K kern, size_t sb) {
    static thread_local int cached_dev = -1;
    static thread_local int ncu = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1 << 20;
    if (dev != cached_dev) {
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
            ncu = 256;
        cached_dev = dev;
    }
    int per = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, FBLOCK, sb);
    return (int64_t)ncu * (per > 0 ? per : 8);
}

/* the column kernel built for a grid's moment order and field set, in FP32
 * or (wide tets) FP64, for one mesh or a batch */
template <class R, bool BATCH>
static KColsFn cols_kernel_t(const GridInfo& g) {
    const bool nox = !(g.flags & (F_X | F_XX | F_XV));
    if (g.order == 0) return k_vox_cols<0, -1, R, BATCH>;
    /* the one-mesh FP32 kernel, the hot one, is also built without the
     * position-moment code */
    if constexpr (sizeof(R) == 4 && !BATCH) {
        if (nox) return g.order == 1 ? k_vox_cols<1, FMASK_NOX, R, BATCH> : k_vox_cols<2, FMASK_NOX, R, BATCH>;
    }
    (void)nox;
    return g.order == 1 ? k_vox_cols<1, -1, R, BATCH> : k_vox_cols<2, -1, R, BATCH>;
}

static KColsFn cols_kernel(const GridInfo& g, bool fp64, bool batch) {
    if (fp64) return batch ? cols_kernel_t<double, true>(g) : cols_kernel_t<double, false>(g);
    return batch ? cols_kernel_t<float, true>(g) : cols_kernel_t<float, false>(g);
}

/* The ordering scratch of ntask tasks: perm (4 bytes a task) and the
 * column kernel's counters, then the sorted task words (8 bytes a task) and
 * their mesh ids (2 bytes a task, batched calls). */
static size_t order_head(int64_t ntask) {
    return ((size_t)ntask * 4 + 128 + NXCD * XSTRIDE * 4 + 255) & ~(size_t)255;
}
static size_t order_bytes(int64_t ntask) { return order_head(ntask) + (size_t)ntask * 8 + (size_t)ntask * 2; }
static uint32_t* order_perm(void* b) { return (uint32_t*)b; }
/* the column kernel's NXCD wavefront counters, 128 bytes apart, after perm */
static unsigned* perm_counter(void* b, int64_t nsub) { return (unsigned*)b + ((nsub + 31) & ~(int64_t)31); }
static uint64_t* order_sorted(void* b, int64_t nsub) { return (uint64_t*)((char*)b + order_head(nsub)); }
static uint16_t* order_mid(void* b, int64_t nsub) { return (uint16_t*)(order_sorted(b, nsub) + nsub); }

/* the column kernel over tasks [0, ntask) of the pass geometry G (one mesh,
 * or a batch's table), FP64 for wide tets (ord: order_bytes(ntask) of
 * scratch for the task order and the counter); mark, when given, is
 * recorded between the ordering and the column kernel */
static int launch_cols(const double* tets, const uint64_t* tasks, int64_t ntask, const PassGeo& G, bool fp64,
                       TaskOut out, void* ord, unsigned* status, unsigned long long* counters,
                       hipStream_t s, hipEvent_t mark = nullptr) {
    uint32_t* perm = order_perm(ord);
    uint64_t* sorted = order_sorted(ord, ntask);
    uint16_t* mid = order_mid(ord, ntask);
    unsigned* next = perm_counter(ord, ntask);
    const unsigned nord = (unsigned)((ntask + ORDW - 1) / ORDW);
    if (G.table)
        k_col_order<true><<<nord, ORDNT, 0, s>>>(tets, tasks, ntask, G.g, G.mb, G.table, perm, sorted, mid, next);
    else
        k_col_order<false><<<nord, ORDNT, 0, s>>>(tets, tasks, ntask, G.g, G.mb, nullptr, perm, sorted, mid, next);
    if (mark) (void)hipEventRecord(mark, s);
    const int64_t nwave = (ntask + FBLOCK - 1) / FBLOCK;
    const size_t sb = stage_bytes(G.g, SFRAG);
    const KColsFn kern = cols_kernel(G.g, fp64, G.table != nullptr);
    const int64_t hold = resident_blocks(kern, sb);
    const unsigned nb = (unsigned)(nwave < hold ? nwave : hold);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(FBLOCK), sb, s, tets, sorted, perm, mid, next, ntask, G.g, G.mb, G.acc,
                       G.table, out, status, counters, G.mcnt);
    return (int)hipGetLastError();
}

template <int ORDER>
static void launch_resid_dep_o(const double* tets, const uint64_t* tasks, int64_t base, int64_t t0, int64_t t1,
                               const PassGeo& G, TaskOut out, const int64_t* colstart, const double* tetf,
                               const int64_t* tetb, const uint32_t* slist, hipStream_t s) {
    const unsigned nbt = (unsigned)((t1 - t0 + 255) / 256);
    const size_t sb = (size_t)(256 / FBLOCK) * stage_bytes(G.g, FBLOCK);
    if (nbt == 0u) return;
    if (G.table) {
        k_resid_whole<ORDER, true><<<nbt, 256, sb, s>>>(tets, tasks, base, t0, t1, G.g, G.mb, G.acc, G.table, out,
                                                         tetf, tetb);
        k_resid_spread<ORDER, true><<<nbt, 256, 0, s>>>(tets, tasks, base, t0, colstart, G.g, G.mb, G.acc, G.table,
                                                         out, tetf, slist);
    } else {
        k_resid_whole<ORDER, false><<<nbt, 256, sb, s>>>(tets, tasks, base, t0, t1, G.g, G.mb, G.acc, nullptr, out,
                                                          tetf, tetb);
        k_resid_spread<ORDER, false><<<nbt, 256, 0, s>>>(tets, tasks, base, t0, colstart, G.g, G.mb, G.acc, nullptr,
                                                          out, tetf, slist);
    }
}

/* the residual pass over the tets [t0, t1) and their ntask tasks; tetf,
 * tetb: one double and one int64 per tet of the batch; slist: WLIST_HEAD +
 * t1 - t0 words (the spread tets' list) */
static int launch_residual(const double* tets, int64_t t0, int64_t t1, const PassGeo& G, const int64_t* colstart,
                           int64_t base, const uint64_t* tasks, TaskOut out, double* tetf, int64_t* tetb,
                           uint32_t* slist, hipStream_t s) {
    k_fill_u32<<<1, 64, 0, s>>>(slist, 1, 0u);
    k_resid_factor<<<(unsigned)((t1 - t0 + 255) / 256), 256, 0, s>>>(tets, t0, t1, colstart, base, out.sum, out.key,
                                                                     tetf, tetb, slist);
    if (G.g.order == 0)
        launch_resid_dep_o<0>(tets, tasks, base, t0, t1, G, out, colstart, tetf, tetb, slist, s);
    else if (G.g.order == 1)
        launch_resid_dep_o<1>(tets, tasks, base, t0, t1, G, out, colstart, tetf, tetb, slist, s);
    else
        launch_resid_dep_o<2>(tets, tasks, base, t0, t1, G, out, colstart, tetf, tetb, slist, s);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------ */
/* mass adjoint                                                        */
/* ------------------------------------------------------------------ */

/* adjoint tasks per tet record (0 for invalid records) */
__global__ __launch_bounds__(256) void k_adj_count(const double* __restrict__ tets, int64_t ntet, GridInfo g,
                                                   MeshBox mb, int64_t* __restrict__ cnt, unsigned* status) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntet) return;
    const double* r = tets + (int64_t)TREC * t;
    int64_t n = 0;
    if (r[25] != 0.0) {
        n = adj_tasks(r, t, g, mb, nullptr);
        if (n < 0) {
            atomicOr(status, ST_BADGEOM);
            n = 0;
        }
    }
    cnt[t] = n;
}

__global__ __launch_bounds__(256) void k_adj_emit(const double* __restrict__ tets, int64_t t0, int64_t t1, GridInfo g,
                                                  MeshBox mb, const int64_t* __restrict__ start, int64_t base,
                                                  uint64_t* __restrict__ tasks) {
    const int64_t t = t0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1) return;
    const double* r = tets + (int64_t)TREC * t;
    if (r[25] == 0.0 || start[t + 1] == start[t]) return;
    adj_tasks(r, t, g, mb, tasks + (start[t] - base));
}

struct AdjArgs {
    const double* w; /* per-cell weights: dL/dm, or ADJ_NW per cell */
    MeshDev md;
    int64_t e0;      /* first element of the batch */
    int per;         /* tets per element */
    double* dpos;    /* dL/dpos, npart x 3 */
    double* dvel;    /* dL/dvel, npart x 3 */
    const uint64_t* prm; /* refined meshes: the tets' parameter records (else null) */
};

/* add val to component b of particle-indexed out for vertex v of tet t:
 * the vertex is a node of the element (level 0: the fixed 6-tet split), or
 * (refined) the element's map at the tet's parameter point, whose weights
 * spread the value over the element's nodes */
__device__ __forceinline__ void adj_scatter(const AdjArgs& A, int64_t t, int v, int b, double val, double* out) {
    const MeshDev& md = A.md;
    if (!A.prm) {
        const int64_t e = A.e0 + t / A.per;
        const int64_t id = md.conn[e * md.vpe + tet_node(md.vpe, (int)(t % A.per), v)];
        unsafeAtomicAdd(out + 3 * id + b, val);
        return;
    }
    const uint64_t* r = A.prm + PREC * t;
    const int64_t e = (int64_t)r[0];
    if (md.vpe == 4) {
        unsafeAtomicAdd(out + 3 * (int64_t)md.conn[4 * e + v] + b, val);
        return;
    }
    /* the node weights from their one-axis factors, node by node (an array
     * of the 27 weights would live in scratch memory) */
    double xi[3], F[3][3];
    unpack_param(r, v, xi);
    const int q = node_factors(md.vpe, xi, F);
    for (int n = 0; n < q * q * q; ++n) {
        const int i = n / (q * q), j = (n / q) % q, k = n % q;
        const double wn = pick_factor(F[0], i) * pick_factor(F[1], j) * pick_factor(F[2], k);
        if (wn != 0.0) unsafeAtomicAdd(out + 3 * (int64_t)md.conn[e * md.vpe + n] + b, val * wn);
    }
}

/* Per cell: the adjoint weights dL/d(M, Sv, Svv) (ADJ_NW values) from the
 * forward's fields m, v, vv and the upstream gradients gm, gv, gvv (any may
 * be null: zero).  With m = M, v = Sv / M, vv = Svv / M - v v^T:
 *   W_Svv = Gvv / M,  W_Sv = Gv / M - (Gvv + Gvv^T) v / M,
 *   W_M = Gm - Gv . v / M + Gvv : (v v^T - vv) / M;
 * a cell without mass has only W_M = Gm (its means are not defined). */
__global__ void k_adj_weights(int64_t nvox, const double* __restrict__ m, const double* __restrict__ v,
                              const double* __restrict__ vv, const double* __restrict__ gm,
                              const double* __restrict__ gv, const double* __restrict__ gvv, double* __restrict__ W) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nvox) return;
    double w[ADJ_NW];
    for (int q = 0; q < ADJ_NW; ++q) w[q] = 0.0;
    w[0] = gm ? gm[c] : 0.0;
    const double M = m[c];
    if (M != 0.0) {
        const double iM = 1.0 / M;
        const double V[3] = {v[3 * c], v[3 * c + 1], v[3 * c + 2]};
        if (gv)
            for (int a = 0; a < 3; ++a) {
                w[1 + a] += gv[3 * c + a] * iM;
                w[0] -= gv[3 * c + a] * V[a] * iM;
            }
        if (gvv)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double G = gvv[9 * c + 3 * a + b];
                    w[4 + 3 * a + b] = G * iM;
                    w[1 + a] -= G * V[b] * iM;
                    w[1 + b] -= G * V[a] * iM;
                    w[0] += G * (V[a] * V[b] - vv[9 * c + 3 * a + b]) * iM;
                }
    }
    for (int q = 0; q < ADJ_NW; ++q) W[ADJ_NW * c + q] = w[q];
}

/* one lane per (tet, copy, axis, plane): FP64 section integrals scattered
 * to the tet's particles; FULL: weights of the velocity fields too */
template <bool FULL>
__global__ __launch_bounds__(FBLOCK) void k_adj(const double* __restrict__ tets, const uint64_t* __restrict__ tasks,
                                                int64_t ntask, GridInfo g, MeshBox mb, AdjArgs A) {
    __shared__ double ru[8][FBLOCK];
    __shared__ double rv[8][FBLOCK];
    const int lane = threadIdx.x;
    const int64_t c = (int64_t)blockIdx.x * FBLOCK + lane;
    if (c >= ntask) return;
    const uint64_t w = tasks[c];
    int64_t t;
    int h, ax, p;
    unpack_adj(w, t, h, ax, p);
    double o[4];
    Poly2<FBLOCK> R{&ru[0][lane], &rv[0][lane]};
    adj_plane_task<FBLOCK, FULL>(g, mb, tets + (int64_t)TREC * t, w, A.w, R, o);
    for (int v = 0; v < 4; ++v) {
        const double val = v == 0 ? o[0] : (v == 1 ? o[1] : (v == 2 ? o[2] : o[3]));
        if (val != 0.0) adj_scatter(A, t, v, ax, val, A.dpos);
    }
}

/* one lane per column task: the velocity part of the adjoint (FP64 column
 * moments: FP32 ones left 8e-5 of the finite differences' gradient on
 * small components) scattered to the tet's particles.  The tasks come in
 * the column kernel's cost order (sorted: the task words in that order),
 * and the grid is persistent: each workgroup takes the next wavefront of
 * tasks from the counter until it runs past the end. */
__global__ __launch_bounds__(FBLOCK) void k_adj_vel(const double* __restrict__ tets, const uint64_t* __restrict__ sorted,
                                                    unsigned* __restrict__ next, int64_t ntask, GridInfo g, MeshBox mb,
                                                    AdjArgs A, unsigned* status) {
    __shared__ __align__(32) uint32_t sv[VRec<double>::WORDS];
    __shared__ unsigned swave;
    const int lane = threadIdx.x;
    LPoly<double> P{(uint32_t*)__builtin_assume_aligned(&sv[VRec<double>::W * lane], 16), 0, 0, 0};
    /* wavefronts claimed one round ahead from the per-XCD counters, as the
     * column kernel does */
    const int home = (int)(blockIdx.x % NXCD);
    const int64_t nwaves = (ntask + FBLOCK - 1) / FBLOCK;
    int cur = 0;
    int64_t claim = 0;
    if (lane == 0) claim = claim_wave(next, home, cur, nwaves);
    for (;;) {
        if (lane == 0) swave = (unsigned)claim;
        __syncthreads();
        const int64_t c0 = (int64_t)swave * FBLOCK;
        __syncthreads();
        if (c0 >= ntask) break;
        if (lane == 0) claim = claim_wave(next, home, cur, nwaves);
        const int64_t c = c0 + lane;
        if (c >= ntask) continue;
        const uint64_t w = sorted[c];
        const int64_t t = (int64_t)(w & 0xffffffffull);
        double o[4][3];
        for (int i = 0; i < 4; ++i) o[i][0] = o[i][1] = o[i][2] = 0.0;
        if (!adj_vel_task<double>(g, mb, tets + (int64_t)TREC * t, w, A.w, P, o)) {
            atomicOr(status, ST_OVERFLOW);
            continue;
        }
        /* one scatter site for the 12 values, each picked with selects (a
         * run-time index would put o in scratch memory; unrolled sites make
         * the kernel outgrow the instruction cache) */
        for (int q = 0; q < 12; ++q) {
            double val = 0.0;
#pragma unroll
            for (int u = 0; u < 12; ++u) val = q == u ? o[u / 3][u % 3] : val;
            if (val != 0.0) adj_scatter(A, t, q / 3, q % 3, val, A.dvel);
        }
    }
}

/* ------------------------------------------------------------------ */
/* host side                                                           */
/* ------------------------------------------------------------------ */

namespace {

/* A fixed region of device memory handed out in order (the allocation-free
 * entry point's scratch): while one is active on a thread, DevBuf::get
 * carves from it instead of calling hipMalloc. */
struct Arena {
    char* base = nullptr;
    size_t size = 0, off = 0;
};
thread_local Arena* t_arena = nullptr;

/* grow-only device buffers reused across calls on this thread's device */
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    int dev = -1;
    int get(size_t bytes, int device) {
        if (p && n >= bytes && dev == device) return 0;
        if (t_arena) {
            /* buffers are carved once per call at their largest size: a
             * second, larger request would leave the first in use */
            if (p) return set_err(PSIVOX_E_NOMEM, "scratch buffer of %zu bytes asked to grow to %zu", n, bytes);
            const size_t need = (std::max<size_t>(bytes, 256) + 255) & ~(size_t)255;
            if (t_arena->off + need > t_arena->size)
                return set_err(PSIVOX_E_NOMEM, "the problem needs more than the %zu bytes of scratch (%zu in use, %zu more asked)",
                               t_arena->size, t_arena->off, need);
            p = t_arena->base + t_arena->off;
            t_arena->off += need;
            n = need;
            dev = device;
            return 0;
        }
        if (p) {
            int cur = 0;
            (void)hipGetDevice(&cur);
            (void)hipSetDevice(dev);
            (void)hipFree(p);
            (void)hipSetDevice(cur);
            p = nullptr;
            n = 0;
        }
        if (bytes == 0) bytes = 256;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return set_err(PSIVOX_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        }
        n = bytes;
        dev = device;
        return 0;
    }
};

struct Workspace {
    DevBuf connchk, pos, vel, mass, conn, fields[6], acc, tets, counts, offs, scan_tmp, counters, colcnt, colstart, tasks,
        colsum, colkey, colmom, perm, cut, tetb, ptets, adj_w, adj_g, bmesh, bfield, bacc, btable, bmcnt,
        an_box, an_valid, an_key, an_key2, an_id, an_id2, an_pcnt, an_poff, an_pairs, an_cntA, an_offA, an_cntB,
        an_offB, an_ta, an_tb, an_tcnt, an_toff, an_tasks, adj_in[7], adj_v, adj_prm;
    hipStream_t stream = nullptr;
    int stream_dev = -1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf wtets;                           /* the wide tets' records for the FP64 column pass */
    DevBuf wlist;                           /* the wide tets' indices, from the count pass */
    DevBuf slist;                           /* the tets whose residual is spread, from k_resid_factor */
    DevBuf asorted;                         /* the adjoint's plane tasks in processing order */
    bool compact_wide = true;               /* the FP64 pass on a copy of the wide tets' records */
    bool timing = true;                     /* kernel_ms from events (off: no events, kernel_ms = 0) */
    int64_t task_cap = (int64_t)1 << 28;    /* column tasks per chunk */
    hipStream_t side = nullptr; /* the field upload beside the kernels */
    int side_dev = -1;
    hipEvent_t side_done = nullptr;
};

std::mutex g_ws_mutex;
Workspace g_ws;

const int FIELD_WIDTH[6] = {1, 3, 3, 9, 9, 9};

int fill_grid_info(const psivox_grid* grid, int xvfix, GridInfo& g) {
    for (int a = 0; a < 3; ++a) {
        if (grid->n[a] <= 0) return set_err(PSIVOX_E_ARG, "grid n[%d] = %d must be positive", a, grid->n[a]);
        g.n[a] = grid->n[a];
        g.w0[a] = grid->window[0][a];
        g.w1[a] = grid->window[1][a];
        g.d[a] = grid->d[a];
        if (!(g.d[a] > 0.0)) return set_err(PSIVOX_E_ARG, "grid d[%d] must be positive", a);
    }
    if (!grid->fields[PSIVOX_M]) return set_err(PSIVOX_E_ARG, "grid field m is required");
    const double* const* f = grid->fields;
    if (f[PSIVOX_XX] && !f[PSIVOX_X]) return set_err(PSIVOX_E_ARG, "Field 'xx' requires you to also provide field 'x'.");
    if (f[PSIVOX_XV] && !(f[PSIVOX_X] && f[PSIVOX_V]))
        return set_err(PSIVOX_E_ARG, "Field 'xv' requires you to also provide fields 'x' and 'v'.");
    if (f[PSIVOX_VV] && !f[PSIVOX_V]) return set_err(PSIVOX_E_ARG, "Field 'vv' requires you to also provide field 'v'.");
    g.flags = (f[PSIVOX_X] ? F_X : 0) | (f[PSIVOX_V] ? F_V : 0) | (f[PSIVOX_XX] ? F_XX : 0) |
              (f[PSIVOX_XV] ? F_XV : 0) | (f[PSIVOX_VV] ? F_VV : 0) | (xvfix ? F_XVFIX : 0);
    set_layout(g);
    return 0;
}

int check_mesh(const psivox_mesh* m) {
    if (!m) return set_err(PSIVOX_E_ARG, "mesh is NULL");
    if (m->elemtype != 4 && m->elemtype != 8 && m->elemtype != 27)
        return set_err(PSIVOX_E_ARG, "elemtype %d must be 4, 8 or 27", m->elemtype);
    if (m->nelem < 0 || m->npart < 0) return set_err(PSIVOX_E_ARG, "negative mesh sizes");
    if (m->nelem > 0 && (!m->pos || !m->mass || !m->conn)) return set_err(PSIVOX_E_ARG, "mesh arrays missing");
    return 0;
}

/* Host <-> device copies of large caller arrays (pageable memory) go
 * through two pinned buffers in chunks: a multi-threaded host memcpy fills
 * (or drains) one while the DMA engine moves the other.  Pageable copies
 * run at 29 GB/s up and 17 GB/s down on the MI355X host; pinned DMA at 57.
 * Callers hold g_ws_mutex. */
constexpr size_t STAGE_CHUNK = (size_t)64 << 20;
constexpr size_t STAGE_MIN = (size_t)16 << 20; /* smaller copies go direct */

struct Staging {
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
};
Staging g_stage;
Staging g_stage_side; /* for the field upload that runs beside the kernels */

void par_memcpy(void* dst, const void* src, size_t n) {
    unsigned nth = std::thread::hardware_concurrency();
    nth = nth < 1 ? 1 : (nth > 8 ? 8 : nth);
    if (n < ((size_t)4 << 20) || nth == 1) {
        memcpy(dst, src, n);
        return;
    }
    std::vector<std::thread> ts;
    const size_t ch = (n + nth - 1) / nth;
    for (unsigned t = 1; t < nth; ++t) {
        const size_t o = t * ch;
        if (o >= n) break;
        const size_t m = std::min(ch, n - o);
        ts.emplace_back([=] { memcpy((char*)dst + o, (const char*)src + o, m); });
    }
    memcpy(dst, src, std::min(ch, n));
    for (auto& t : ts) t.join();
}

int stage_ready(Staging& S) {
    for (int b = 0; b < 2; ++b) {
        if (!S.buf[b]) HIPCHK(hipHostMalloc(&S.buf[b], STAGE_CHUNK, hipHostMallocDefault));
        if (!S.ev[b]) HIPCHK(hipEventCreateWithFlags(&S.ev[b], hipEventDisableTiming));
    }
    return 0;
}

/* wait until the DMA that last used stage buffer b has finished */
int stage_wait(Staging& S, int b) {
    if (S.pending[b]) HIPCHK(hipEventSynchronize(S.ev[b]));
    S.pending[b] = false;
    return 0;
}

int copy_h2d(void* d, const void* h, size_t bytes, hipStream_t s, Staging& S = g_stage) {
    if (bytes < STAGE_MIN) {
        HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
        return 0;
    }
    int rc = stage_ready(S);
    if (rc) return rc;
    for (size_t off = 0, i = 0; off < bytes; off += STAGE_CHUNK, ++i) {
        const int b = (int)(i & 1);
        const size_t m = std::min(STAGE_CHUNK, bytes - off);
        if ((rc = stage_wait(S, b))) return rc;
        par_memcpy(S.buf[b], (const char*)h + off, m);
        HIPCHK(hipMemcpyAsync((char*)d + off, S.buf[b], m, hipMemcpyHostToDevice, s));
        HIPCHK(hipEventRecord(S.ev[b], s));
        S.pending[b] = true;
    }
    return 0;
}

/* returns when the host copy is complete (work queued on s before it
 * included) */
int copy_d2h(void* h, const void* d, size_t bytes, hipStream_t s) {
    if (bytes < STAGE_MIN) {
        HIPCHK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    }
    Staging& S = g_stage;
    int rc = stage_ready(S);
    if (rc) return rc;
    const size_t nch = (bytes + STAGE_CHUNK - 1) / STAGE_CHUNK;
    for (size_t i = 0; i <= nch; ++i) {
        if (i < nch) {
            const int b = (int)(i & 1);
            const size_t off = i * STAGE_CHUNK, m = std::min(STAGE_CHUNK, bytes - off);
            if ((rc = stage_wait(S, b))) return rc;
            HIPCHK(hipMemcpyAsync(S.buf[b], (const char*)d + off, m, hipMemcpyDeviceToHost, s));
            HIPCHK(hipEventRecord(S.ev[b], s));
            S.pending[b] = true;
        }
        if (i >= 1) {
            const int b = (int)((i - 1) & 1);
            const size_t off = (i - 1) * STAGE_CHUNK, m = std::min(STAGE_CHUNK, bytes - off);
            if ((rc = stage_wait(S, b))) return rc;
            par_memcpy((char*)h + off, S.buf[b], m);
        }
    }
    return 0;
}

int upload(DevBuf& b, const void* src, size_t bytes, int dev, hipStream_t s, const void** out) {
    int rc = b.get(bytes, dev);
    if (rc) return rc;
    if ((rc = copy_h2d(b.p, src, bytes, s))) return rc;
    *out = b.p;
    return 0;
}

/* A host-to-device upload of the grid's fields run by a thread of its own
 * on the workspace's side stream, so that the PCIe transfer overlaps the
 * deposit (which starts from a zeroed accumulator; the merge at the end adds
 * the fields once W.side_done has passed).  The destructor joins, so every
 * return path of the caller waits for the thread. */
struct SideUpload {
    std::thread t;
    int rc = 0;
    std::string err;
    hipStream_t used = nullptr;
    ~SideUpload() {
        if (t.joinable()) t.join();
        /* an early return of the caller leaves no copy in flight into the
         * workspace's field buffers */
        if (used) (void)hipStreamSynchronize(used);
    }
    int start(Workspace& W, int dev, double* const* fd, const psivox_grid* grid, int64_t nvox) {
        if (!W.side || W.side_dev != dev) {
            HIPCHK(hipStreamCreateWithFlags(&W.side, hipStreamNonBlocking));
            W.side_dev = dev;
        }
        if (!W.side_done) HIPCHK(hipEventCreateWithFlags(&W.side_done, hipEventDisableTiming));
        double* d[6];
        for (int f = 0; f < 6; ++f) d[f] = fd[f];
        used = W.side;
        t = std::thread([this, &W, dev, d, grid, nvox] {
            rc = side_body(W, dev, d, grid, nvox);
            if (rc) err = g_err;
        });
        return 0;
    }
    static int side_body(Workspace& W, int dev, double* const* d, const psivox_grid* grid, int64_t nvox) {
        HIPCHK(hipSetDevice(dev));
        for (int f = 0; f < 6; ++f) {
            if (!grid->fields[f]) continue;
            const int rc = copy_h2d(d[f], grid->fields[f], (size_t)nvox * FIELD_WIDTH[f] * sizeof(double), W.side,
                                    g_stage_side);
            if (rc) return rc;
        }
        HIPCHK(hipEventRecord(W.side_done, W.side));
        return 0;
    }
    /* wait for the thread; s then waits for the copies it queued */
    int finish(Workspace& W, hipStream_t s) {
        if (t.joinable()) t.join();
        if (rc) {
            g_err = err;
            return rc;
        }
        HIPCHK(hipStreamWaitEvent(s, W.side_done, 0));
        return 0;
    }
};

inline unsigned grid_blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

/* PSIVOX_E_ARG unless every entry of the device connectivity conn (nelem x
 * vpe) lies in [0, npart); waits for the check */
int check_conn(Workspace& W, const int32_t* conn, int64_t nelem, int vpe, int64_t npart, int dev, hipStream_t s) {
    if (nelem <= 0) return 0;
    int rc = W.connchk.get(sizeof(unsigned), dev);
    if (rc) return rc;
    unsigned* flag = (unsigned*)W.connchk.p;
    k_fill_u32<<<1, 64, 0, s>>>(flag, 1, 0u);
    const int64_t n = nelem * vpe;
    k_conn_check<<<std::min(grid_blocks(n, 256), 4096u), 256, 0, s>>>(conn, n, npart, flag);
    HIPCHK(hipGetLastError());
    unsigned h = 0;
    HIPCHK(hipMemcpyAsync(&h, flag, sizeof h, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (h)
        return set_err(PSIVOX_E_ARG, "connectivity entries must lie in [0, npart) with npart = %lld",
                       (long long)npart);
    return 0;
}

/* the mesh arrays as device pointers (uploaded when they are on the host),
 * connectivity checked against npart */
int mesh_to_device(Workspace& W, const psivox_mesh* mesh, int dev, hipStream_t s, MeshDev& md) {
    int rc;
    md = MeshDev{};
    md.vpe = mesh->elemtype;
    if (mesh->nelem <= 0) return 0;
    if (mesh->on_device) {
        md.pos = mesh->pos;
        md.vel = mesh->vel;
        md.mass = mesh->mass;
        md.conn = mesh->conn;
        return check_conn(W, md.conn, mesh->nelem, md.vpe, mesh->npart, dev, s);
    }
    const void* p;
    if ((rc = upload(W.pos, mesh->pos, (size_t)mesh->npart * 3 * sizeof(double), dev, s, &p))) return rc;
    md.pos = (const double*)p;
    if (mesh->vel) {
        if ((rc = upload(W.vel, mesh->vel, (size_t)mesh->npart * 3 * sizeof(double), dev, s, &p))) return rc;
        md.vel = (const double*)p;
    }
    if ((rc = upload(W.mass, mesh->mass, (size_t)mesh->nelem * sizeof(double), dev, s, &p))) return rc;
    md.mass = (const double*)p;
    if ((rc = upload(W.conn, mesh->conn, (size_t)mesh->nelem * md.vpe * sizeof(int32_t), dev, s, &p))) return rc;
    md.conn = (const int32_t*)p;
    return check_conn(W, md.conn, mesh->nelem, md.vpe, mesh->npart, dev, s);
}

/* The end a1 of the task chunk that starts at tet a0: the largest a1 <= nt
 * with colstart[a1] - colstart[a0] <= cap (a single tet has fewer than
 * TASK_MAXPERTET < cap tasks, so a1 > a0).  out = {a1, colstart[a0],
 * colstart[a1]}.  One thread: a binary search over the nondecreasing
 * colstart. */
__global__ void k_chunk_cut(const int64_t* __restrict__ colstart, int64_t nt, int64_t a0, int64_t cap,
                            int64_t* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int64_t lim = colstart[a0] + cap;
    int64_t lo = a0 + 1, hi = nt;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo + 1) / 2;
        if (colstart[mid] <= lim)
            lo = mid;
        else
            hi = mid - 1;
    }
    out[0] = lo;
    out[1] = colstart[a0];
    out[2] = colstart[lo];
}

/* the column kernel's wavefront counter: the word after a chunk's nsub
 * task-order entries */
/* k_col_count for a pass, one mesh or a batch */
template <bool BATCH>
static void launch_col_count(int pass, const double* tets, int64_t nt, const PassGeo& G, int64_t* cnt, unsigned* status,
                             unsigned long long* counters, uint32_t* wlist, hipStream_t s) {
    const unsigned nb = grid_blocks(nt, 256);
    if (pass == PASS_NARROW)
        k_col_count<BATCH, PASS_NARROW>
            <<<nb, 256, 0, s>>>(tets, nt, G.g, G.mb, G.table, cnt, status, counters, G.mcnt, wlist);
    else if (pass == PASS_WIDE)
        k_col_count<BATCH, PASS_WIDE>
            <<<nb, 256, 0, s>>>(tets, nt, G.g, G.mb, G.table, cnt, status, counters, G.mcnt, nullptr);
    else if (pass == PASS_ALL64)
        k_col_count<BATCH, PASS_ALL64>
            <<<nb, 256, 0, s>>>(tets, nt, G.g, G.mb, G.table, cnt, status, counters, G.mcnt, nullptr);
    else
        k_col_count<BATCH, PASS_ALL>
            <<<nb, 256, 0, s>>>(tets, nt, G.g, G.mb, G.table, cnt, status, counters, G.mcnt, nullptr);
}

template <bool BATCH>
static void launch_col_emit(bool wide, const double* tets, int64_t t0, int64_t t1, const PassGeo& G,
                            const int64_t* colstart, int64_t base, uint64_t* tasks, hipStream_t s) {
    const unsigned nb = grid_blocks(t1 - t0, 256);
    if (wide)
        k_col_emit<BATCH, true><<<nb, 256, 0, s>>>(tets, t0, t1, G.g, G.mb, G.table, colstart, base, tasks);
    else
        k_col_emit<BATCH, false><<<nb, 256, 0, s>>>(tets, t0, t1, G.g, G.mb, G.table, colstart, base, tasks);
}

/*
 * Column tasks of tet records [0, nt) for one pass (PASS_*): count, scan,
 * and for each chunk of tets whose tasks fit the task buffer emit them and
 * call fn(a0, a1, colstart, base, tasks, ntask_chunk).  Returns PSIVOX_E_*
 * or 0; *fell is set (and nothing run) when a tet cannot take the column
 * path.  nwide (PASS_NARROW): the wide tets left for PASS_WIDE.
 */
template <class Fn>
int for_task_chunks(Workspace& W, const double* tets, int64_t nt, const PassGeo& G, int pass, unsigned* status,
                    unsigned long long* counters, hipStream_t s, int dev, bool* fell, int64_t* nwide, Fn fn) {
    int rc;
    *fell = false;
    if (nwide) *nwide = 0;
    if ((rc = W.colcnt.get((size_t)(nt + 1) * sizeof(int64_t), dev))) return rc;
    if ((rc = W.colstart.get((size_t)(nt + 1) * sizeof(int64_t), dev))) return rc;
    int64_t* colcnt = (int64_t*)W.colcnt.p;
    int64_t* colstart = (int64_t*)W.colstart.p;
    /* colcnt[nt] = 0, the wide-tet counter counters[2] = 0 and, when the
     * wide tets are copied for their own pass, their list's counter = 0 */
    uint32_t* wlist = nullptr;
    if (pass == PASS_NARROW && W.compact_wide) {
        if ((rc = W.wlist.get((size_t)(WLIST_HEAD + nt) * sizeof(uint32_t), dev))) return rc;
        wlist = (uint32_t*)W.wlist.p;
        k_fill_u32<<<1, 64, 0, s>>>(wlist, 2, 0u);
    }
    k_fill_u32<<<1, 64, 0, s>>>((uint32_t*)(colcnt + nt), 2, 0u);
    k_fill_u32<<<1, 64, 0, s>>>((uint32_t*)(counters + 2), 2, 0u);
    if (G.table)
        launch_col_count<true>(pass, tets, nt, G, colcnt, status, counters, wlist, s);
    else
        launch_col_count<false>(pass, tets, nt, G, colcnt, status, counters, wlist, s);
    HIPCHK(hipGetLastError());
    size_t tb = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, colcnt, colstart, nt + 1, s));
    if ((rc = W.scan_tmp.get(tb, dev))) return rc;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(W.scan_tmp.p, tb, colcnt, colstart, nt + 1, s));
    int64_t ntask = 0;
    unsigned hs = 0;
    unsigned long long hw = 0;
    HIPCHK(hipMemcpyAsync(&ntask, colstart + nt, sizeof ntask, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&hs, status, sizeof hs, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&hw, counters + 2, sizeof hw, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (nwide) *nwide = (int64_t)hw;
    if (hs & ST_FALLBACK) {
        *fell = true;
        return 0;
    }
    if (ntask == 0) return 0;
    /* tet sub-ranges whose tasks fit the task buffers, cut on the device
     * (colstart stays there: for a large batch it is hundreds of MB) */
    const int64_t cap = W.task_cap;
    const int64_t tcap = std::min(ntask, cap);
    if ((rc = W.tasks.get((size_t)tcap * sizeof(uint64_t), dev))) return rc;
    if ((rc = W.cut.get(4 * sizeof(int64_t), dev))) return rc;
    uint64_t* tasks = (uint64_t*)W.tasks.p;
    int64_t a1 = 0;
    for (int64_t a0 = 0; a0 < nt; a0 = a1) {
        int64_t base = 0, nsub = ntask;
        a1 = nt;
        if (ntask > cap) {
            int64_t hc[3];
            k_chunk_cut<<<1, 64, 0, s>>>(colstart, nt, a0, cap, (int64_t*)W.cut.p);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(hc, W.cut.p, sizeof hc, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            a1 = hc[0];
            base = hc[1];
            nsub = hc[2] - hc[1];
            /* a single tet with more tasks than a chunk holds (only with a
             * small scratch) */
            if (nsub > cap)
                return set_err(PSIVOX_E_NOMEM, "a tet has %lld column tasks, more than the %lld a chunk holds",
                               (long long)nsub, (long long)cap);
        }
        if (nsub == 0) continue;
        if (G.table)
            launch_col_emit<true>(pass_fp64(pass), tets, a0, a1, G, colstart, base, tasks, s);
        else
            launch_col_emit<false>(pass_fp64(pass), tets, a0, a1, G, colstart, base, tasks, s);
        HIPCHK(hipGetLastError());
        if ((rc = fn(a0, a1, (const int64_t*)colstart, base, (const uint64_t*)tasks, nsub))) return rc;
    }
    return 0;
}

/* The tet records of a column pass: for PASS_WIDE (when W compacts them)
 * a copy of the nwide wide tets among tets[0, nt), in *pt, *pn, from the
 * list the PASS_NARROW count made (W.wlist); otherwise tets and nt
 * themselves. */
int pass_tets(Workspace& W, int pass, const double* tets, int64_t nt, const PassGeo& G, int64_t nwide,
              unsigned long long* counters, hipStream_t s, int dev, const double** pt, int64_t* pn) {
    (void)nt;
    (void)G;
    (void)counters;
    *pt = tets;
    *pn = nt;
    if (pass != PASS_WIDE || !W.compact_wide) return 0;
    int rc = W.wtets.get((size_t)nwide * TREC * sizeof(double), dev);
    if (rc) return rc;
    k_wide_copy<<<grid_blocks(nwide * TREC, 256), 256, 0, s>>>(tets, (const uint32_t*)W.wlist.p + WLIST_HEAD, nwide,
                                                               (double*)W.wtets.p);
    HIPCHK(hipGetLastError());
    *pt = (const double*)W.wtets.p;
    *pn = nwide;
    return 0;
}

/* One (mesh, grid) voxelization on stream s with workspace W. */
/* exclusive sum of cnt[0..n) into off[0..n] (cnt[n] is set to 0 first);
 * returns off[n] on the host */
int exclusive_sum(Workspace& W, int64_t* cnt, int64_t* off, int64_t n, hipStream_t s, int dev, int64_t* total) {
    HIPCHK(hipMemsetAsync(cnt + n, 0, sizeof(int64_t), s));
    size_t tb = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, off, n + 1, s));
    int rc = W.scan_tmp.get(tb, dev);
    if (rc) return rc;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(W.scan_tmp.p, tb, cnt, off, n + 1, s));
    HIPCHK(hipMemcpyAsync(total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

/* all tets of the elements, expanded with (g, mb): counts, offsets
 * (offs[e] .. offs[e+1]) and records */
int expand_all(Workspace& W, DevBuf& cntb, DevBuf& offb, DevBuf& tetb, const MeshDev& md, const GridInfo& g,
               const MeshBox& mb, int64_t ne, int maxlvl, double tol2, hipStream_t s, int dev, int64_t* nt) {
    int rc;
    if ((rc = cntb.get((size_t)(ne + 1) * sizeof(int64_t), dev)) || (rc = offb.get((size_t)(ne + 1) * sizeof(int64_t), dev)))
        return rc;
    int64_t* cnt = (int64_t*)cntb.p;
    int64_t* off = (int64_t*)offb.p;
    k_expand_count<<<grid_blocks(ne, 128), 128, 0, s>>>(md, g, mb, ne, maxlvl, tol2, cnt);
    HIPCHK(hipGetLastError());
    if ((rc = exclusive_sum(W, cnt, off, ne, s, dev, nt))) return rc;
    if (*nt >= ((int64_t)1 << 32)) return set_err(PSIVOX_E_ARG, "annihilation: more than 2^32 tets");
    if ((rc = tetb.get((size_t)(*nt > 0 ? *nt : 1) * TREC * sizeof(double), dev))) return rc;
    k_expand_emit<<<grid_blocks(ne, 128), 128, 0, s>>>(md, g, mb, 0, ne, maxlvl, tol2, off, 0, (double*)tetb.p);
    HIPCHK(hipGetLastError());
    return 0;
}

/*
 * mode='annihilation' (src/cpsi.c:66-77, 115-158): for every element e in
 * the window and every element e1 whose wrapped box meets e's, every pair of
 * their tets (e's wrapped, e1's raw, as the reference refines them) whose
 * boxes overlap adds rho0 rho1 vol(overlap ∩ cell) into the mass field.
 * The other fields are untouched.
 */
int run_annihilation(Workspace& W, const psivox_mesh* mesh, psivox_grid* grid, const GridInfo& g,
                     const psivox_opts* o, psivox_stats* st, hipStream_t s, int dev) {
    auto t_start = std::chrono::steady_clock::now();
    int rc;
    MeshBox mb;
    mb.periodic = mesh->periodic ? 1 : 0;
    mb.full_window = 0;
    for (int a = 0; a < 3; ++a) {
        mb.lo[a] = mesh->box[0][a];
        mb.hi[a] = mesh->box[1][a];
    }
    GridInfo gu = g; /* unbounded window, no wrap: the raw expansion */
    MeshBox mu;
    mu.periodic = 0;
    mu.full_window = 0;
    for (int a = 0; a < 3; ++a) {
        gu.w0[a] = -HUGE_VAL;
        gu.w1[a] = HUGE_VAL;
        mu.lo[a] = mu.hi[a] = 0.0;
    }
    const int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int64_t ne = mesh->nelem;
    int maxlvl = o->refine_max_lvl < 0 ? 0 : (o->refine_max_lvl > MAXLVL ? MAXLVL : o->refine_max_lvl);
    if (mesh->elemtype == 4) maxlvl = 0;
    const double dxmin = std::min(g.d[0], std::min(g.d[1], g.d[2]));
    const double tol = o->refine_tolerance * dxmin;
    const double tol2 = tol * tol;
    if (ne > INT32_MAX) return set_err(PSIVOX_E_ARG, "annihilation: more than 2^31-1 elements");

    MeshDev md{};
    if ((rc = mesh_to_device(W, mesh, dev, s, md))) return rc;
    double* mf;
    if (grid->on_device) {
        mf = grid->fields[PSIVOX_M];
    } else {
        const void* p;
        if ((rc = upload(W.fields[0], grid->fields[PSIVOX_M], (size_t)nvox * sizeof(double), dev, s, &p))) return rc;
        mf = (double*)p;
    }
    if ((rc = W.counters.get(64, dev))) return rc;
    unsigned long long* counters = (unsigned long long*)W.counters.p;
    unsigned* status = (unsigned*)(counters + 4);
    HIPCHK(hipMemsetAsync(counters, 0, 64, s));
    int64_t nused = 0, ntask_total = 0;
    double kms = 0.0;
    if (ne > 0) {
        /* element boxes and the broad phase */
        if ((rc = W.an_box.get((size_t)ne * 6 * sizeof(double), dev)) || (rc = W.an_valid.get((size_t)ne, dev)) ||
            (rc = W.an_key.get((size_t)ne * sizeof(double), dev)) ||
            (rc = W.an_key2.get((size_t)ne * sizeof(double), dev)) ||
            (rc = W.an_id.get((size_t)ne * sizeof(int32_t), dev)) ||
            (rc = W.an_id2.get((size_t)ne * sizeof(int32_t), dev)) ||
            (rc = W.an_pcnt.get((size_t)(ne + 1) * sizeof(int64_t), dev)) ||
            (rc = W.an_poff.get((size_t)(ne + 1) * sizeof(int64_t), dev)))
            return rc;
        double* box = (double*)W.an_box.p;
        k_elem_boxes<<<grid_blocks(ne, 128), 128, 0, s>>>(md, g, mb, ne, box, (int8_t*)W.an_valid.p);
        HIPCHK(hipGetLastError());
        k_sweep_keys<<<grid_blocks(ne, 256), 256, 0, s>>>(box, (const int8_t*)W.an_valid.p, ne, (double*)W.an_key.p,
                                                          (int32_t*)W.an_id.p);
        HIPCHK(hipGetLastError());
        size_t tb = 0;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const double*)W.an_key.p, (double*)W.an_key2.p,
                                                  (const int32_t*)W.an_id.p, (int32_t*)W.an_id2.p, (int)ne, 0, 64, s));
        if ((rc = W.scan_tmp.get(tb, dev))) return rc;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(W.scan_tmp.p, tb, (const double*)W.an_key.p, (double*)W.an_key2.p,
                                                  (const int32_t*)W.an_id.p, (int32_t*)W.an_id2.p, (int)ne, 0, 64, s));
        const double* key = (const double*)W.an_key2.p;
        const int32_t* id = (const int32_t*)W.an_id2.p;
        int64_t* pcnt = (int64_t*)W.an_pcnt.p;
        int64_t* poff = (int64_t*)W.an_poff.p;
        k_sweep_pairs<<<grid_blocks(ne, 128), 128, 0, s>>>(key, id, box, ne, pcnt, nullptr, nullptr);
        HIPCHK(hipGetLastError());
        int64_t npairs = 0;
        if ((rc = exclusive_sum(W, pcnt, poff, ne, s, dev, &npairs))) return rc;
        if ((rc = W.an_pairs.get((size_t)(npairs > 0 ? npairs : 1) * sizeof(int2), dev))) return rc;
        int2* pairs = (int2*)W.an_pairs.p;
        k_sweep_pairs<<<grid_blocks(ne, 128), 128, 0, s>>>(key, id, box, ne, nullptr, poff, pairs);
        HIPCHK(hipGetLastError());

        /* the two expansions: wrapped and window-tested (the outer element),
         * raw (the tree's element) */
        int64_t nA = 0, nB = 0;
        if ((rc = expand_all(W, W.an_cntA, W.an_offA, W.an_ta, md, g, mb, ne, maxlvl, tol2, s, dev, &nA))) return rc;
        if ((rc = expand_all(W, W.an_cntB, W.an_offB, W.an_tb, md, gu, mu, ne, maxlvl, tol2, s, dev, &nB))) return rc;
        const int64_t* offA = (const int64_t*)W.an_offA.p;
        const int64_t* offB = (const int64_t*)W.an_offB.p;
        const double* ta = (const double*)W.an_ta.p;
        const double* tbp = (const double*)W.an_tb.p;

        /* tet-pair tasks, in chunks of at most cap */
        if ((rc = W.an_tcnt.get((size_t)(npairs + 1) * sizeof(int64_t), dev)) ||
            (rc = W.an_toff.get((size_t)(npairs + 1) * sizeof(int64_t), dev)))
            return rc;
        int64_t* tcnt = (int64_t*)W.an_tcnt.p;
        int64_t* toff = (int64_t*)W.an_toff.p;
        if (npairs > 0) {
            k_tet_pairs<<<grid_blocks(npairs, 128), 128, 0, s>>>(pairs, 0, npairs, offA, offB, ta, tbp, tcnt, nullptr,
                                                                 0, nullptr);
            HIPCHK(hipGetLastError());
        }
        if ((rc = exclusive_sum(W, tcnt, toff, npairs, s, dev, &ntask_total))) return rc;
        const int64_t cap = std::max<int64_t>(1 << 20, std::min<int64_t>(ntask_total, (int64_t)1 << 26));
        if ((rc = W.an_tasks.get((size_t)cap * sizeof(uint2), dev)) || (rc = W.cut.get(4 * sizeof(int64_t), dev)))
            return rc;
        uint2* tasks = (uint2*)W.an_tasks.p;
        int64_t p0 = 0;
        while (p0 < npairs) {
            /* the largest p1 with toff[p1] - toff[p0] <= cap (one element pair has at most
             * (6 * 8^6)^2 tasks; a chunk always takes at least one pair) */
            int64_t hcut[3];
            k_chunk_cut<<<1, 1, 0, s>>>(toff, npairs, p0, cap, (int64_t*)W.cut.p);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(hcut, W.cut.p, sizeof hcut, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            const int64_t p1 = hcut[0], nchunk = hcut[2] - hcut[1];
            if (nchunk > cap) return set_err(PSIVOX_E_ARG, "annihilation: one element pair exceeds the task buffer");
            if (nchunk > 0) {
                k_tet_pairs<<<grid_blocks(p1 - p0, 128), 128, 0, s>>>(pairs, p0, p1, offA, offB, ta, tbp, nullptr,
                                                                      toff, hcut[1], tasks);
                HIPCHK(hipGetLastError());
                HIPCHK(hipEventRecord(W.ev0, s));
                k_annihilate<<<grid_blocks(nchunk, 128), 128, 0, s>>>(ta, tbp, tasks, nchunk, g, mf, status, counters);
                HIPCHK(hipGetLastError());
                HIPCHK(hipEventRecord(W.ev1, s));
                HIPCHK(hipEventSynchronize(W.ev1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
                kms += ms;
            }
            p0 = p1;
        }
        /* elements inside the window */
        std::vector<int8_t> hv((size_t)ne);
        HIPCHK(hipMemcpyAsync(hv.data(), W.an_valid.p, (size_t)ne, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int8_t v : hv) nused += v != 0;
        if (st) st->ntets = nA;
    }
    unsigned long long hcnt[4] = {0, 0, 0, 0};
    unsigned hstatus = 0;
    HIPCHK(hipMemcpyAsync(hcnt, counters, sizeof hcnt, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&hstatus, status, sizeof hstatus, hipMemcpyDeviceToHost, s));
    if (!grid->on_device && (rc = copy_d2h(grid->fields[PSIVOX_M], mf, (size_t)nvox * sizeof(double), s))) return rc;
    HIPCHK(hipStreamSynchronize(s));
    if (hstatus & ST_OVERFLOW)
        return set_err(PSIVOX_E_OVERFLOW, "a polyhedron exceeded its vertex budget (degenerate input?)");
    if (hstatus & ST_BADGEOM) return set_err(PSIVOX_E_OVERFLOW, "inconsistent polyhedron graph (degenerate input?)");
    if (st) {
        st->nelem_used = nused;
        if (ne == 0) st->ntets = 0;
        st->nghosts = ntask_total; /* tet pairs */
        st->nfrag = (int64_t)hcnt[0];
        st->kernel_ms = kms;
        st->kernel = PSIVOX_KERNEL_PLAIN;
        st->reserved = 0;
        st->nwide = 0;
        st->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
    return 0;
}

int run_one(Workspace& W, const psivox_mesh* mesh, psivox_grid* grid, const psivox_opts* o, psivox_stats* st,
            hipStream_t s, int dev) {
    auto t_start = std::chrono::steady_clock::now();
    int rc = check_mesh(mesh);
    if (rc) return rc;
    GridInfo g;
    rc = fill_grid_info(grid, o->xv_true_covariance, g);
    if (rc) return rc;
    if (o->mode == PSIVOX_MODE_ANNIHILATION) return run_annihilation(W, mesh, grid, g, o, st, s, dev);
    if (o->mode != PSIVOX_MODE_DENSITY) return set_err(PSIVOX_E_ARG, "invalid mode %d", o->mode);
    int kernel = o->kernel == PSIVOX_KERNEL_AUTO ? PSIVOX_KERNEL_FAST : o->kernel;
    if (kernel != PSIVOX_KERNEL_PLAIN && kernel != PSIVOX_KERNEL_FAST && kernel != PSIVOX_KERNEL_FAST64)
        return set_err(PSIVOX_E_ARG, "invalid kernel %d", o->kernel);
    /* the automatic choice sends wide tets through the FP64 column kernel;
     * FAST64 sends every tet there */
    const bool route = o->kernel == PSIVOX_KERNEL_AUTO;
    const int first_pass = route ? PASS_NARROW : (kernel == PSIVOX_KERNEL_FAST64 ? PASS_ALL64 : PASS_ALL);

    MeshBox mb;
    mb.periodic = mesh->periodic ? 1 : 0;
    mb.full_window = mb.periodic;
    for (int a = 0; a < 3; ++a) {
        mb.lo[a] = mesh->box[0][a];
        mb.hi[a] = mesh->box[1][a];
        if (mb.lo[a] != g.w0[a] || mb.hi[a] != g.w1[a]) mb.full_window = 0;
    }
    const int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int vpe = mesh->elemtype;

    /* mesh arrays on the device */
    MeshDev md{};
    if ((rc = mesh_to_device(W, mesh, dev, s, md))) return rc;

    /* grid fields on the device.  Large host fields go up beside the
     * kernels (SideUpload). */
    size_t field_bytes = 0;
    for (int f = 0; f < 6; ++f)
        if (grid->fields[f]) field_bytes += (size_t)nvox * FIELD_WIDTH[f] * sizeof(double);
    const bool side_up = !grid->on_device && field_bytes >= STAGE_MIN;
    double* fd[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int f = 0; f < 6; ++f) {
        if (!grid->fields[f]) continue;
        const size_t bytes = (size_t)nvox * FIELD_WIDTH[f] * sizeof(double);
        if (grid->on_device) {
            fd[f] = grid->fields[f];
        } else if (side_up) {
            if ((rc = W.fields[f].get(bytes, dev))) return rc;
            fd[f] = (double*)W.fields[f].p;
        } else {
            const void* p;
            if ((rc = upload(W.fields[f], grid->fields[f], bytes, dev, s, &p))) return rc;
            fd[f] = (double*)p;
        }
    }
    FieldPtrs fptrs;
    for (int f = 0; f < 6; ++f) fptrs.f[f] = fd[f];
    if ((rc = W.acc.get((size_t)nvox * g.nacc * sizeof(double), dev))) return rc;
    double* acc = (double*)W.acc.p;
    if ((rc = W.counters.get(64, dev))) return rc;
    unsigned long long* counters = (unsigned long long*)W.counters.p;
    unsigned* status = (unsigned*)(counters + 4);
    HIPCHK(hipMemsetAsync(counters, 0, 64, s));

    SideUpload side;
    /* The deposit starts from a zeroed accumulator and the fields' present
     * contents are merged in at the end (launch_merge_finalize): cells that
     * receive no mass keep their contents bit for bit, and large host
     * fields can travel while the kernels run. */
    auto start_acc = [&]() -> int {
        HIPCHK(hipMemsetAsync(acc, 0, (size_t)nvox * g.nacc * sizeof(double), s));
        return 0;
    };
    auto finish_acc = [&]() -> int {
        if (side_up) {
            int r = side.finish(W, s);
            if (r) return r;
        }
        launch_merge_finalize(g, acc, fptrs, nvox, s);
        HIPCHK(hipGetLastError());
        return 0;
    };
    if (side_up && (rc = side.start(W, dev, fd, grid, nvox))) return rc;
    if ((rc = start_acc())) return rc;

    int maxlvl = o->refine_max_lvl < 0 ? 0 : (o->refine_max_lvl > MAXLVL ? MAXLVL : o->refine_max_lvl);
    if (vpe == 4) maxlvl = 0;
    const double dxmin = std::min(g.d[0], std::min(g.d[1], g.d[2]));
    const double tol = o->refine_tolerance * dxmin;
    const double tol2 = tol * tol;

    int64_t batch = o->batch_elements > 0 ? o->batch_elements : (int64_t)1 << 22;
    const int64_t ne = mesh->nelem;
    int64_t ntets_total = 0, nwide_total = 0;
    double kms = 0.0;
    int fell_back = 0;

    /* refinement counts and offsets (levels > 0) */
    int64_t* counts = nullptr;
    std::vector<int64_t> hoffs;
    if (maxlvl > 0 && ne > 0) {
        if ((rc = W.counts.get((size_t)(ne + 1) * sizeof(int64_t), dev))) return rc;
        if ((rc = W.offs.get((size_t)(ne + 1) * sizeof(int64_t), dev))) return rc;
        counts = (int64_t*)W.counts.p;
        HIPCHK(hipMemsetAsync(counts + ne, 0, sizeof(int64_t), s));
        k_expand_count<<<grid_blocks(ne, 128), 128, 0, s>>>(md, g, mb, ne, maxlvl, tol2, counts);
        HIPCHK(hipGetLastError());
        size_t tmp_bytes = 0;
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, (int64_t*)W.offs.p, ne + 1, s));
        if ((rc = W.scan_tmp.get(tmp_bytes, dev))) return rc;
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(W.scan_tmp.p, tmp_bytes, counts, (int64_t*)W.offs.p, ne + 1, s));
        hoffs.resize(ne + 1);
        HIPCHK(hipMemcpyAsync(hoffs.data(), W.offs.p, (ne + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    const int64_t max_tets_batch = (int64_t)6 << 22; /* 25M tets = 5.2 GB of records per batch */

    int64_t e0 = 0;
    while (e0 < ne) {
        int64_t e1, nt;
        if (maxlvl > 0) {
            /* largest e1 with offs[e1] - offs[e0] <= max_tets_batch (at least one element) */
            e1 = e0 + 1;
            int64_t lo = e0 + 1, hi = ne;
            while (lo <= hi) {
                int64_t mid = (lo + hi) / 2;
                if (hoffs[mid] - hoffs[e0] <= max_tets_batch) {
                    e1 = mid;
                    lo = mid + 1;
                } else
                    hi = mid - 1;
            }
            nt = hoffs[e1] - hoffs[e0];
        } else {
            e1 = std::min(ne, e0 + batch);
            nt = (e1 - e0) * (vpe == 4 ? 1 : 6);
        }
        if (nt > 0) {
            if ((rc = W.tets.get((size_t)nt * TREC * sizeof(double), dev))) return rc;
            double* tets = (double*)W.tets.p;
            if (maxlvl > 0) {
                k_expand_emit<<<grid_blocks(e1 - e0, 128), 128, 0, s>>>(md, g, mb, e0, e1, maxlvl, tol2,
                                                                        (const int64_t*)W.offs.p, hoffs[e0], tets);
            } else {
                launch_expand_fixed(md, g, mb, e0, e1 - e0, tets, 0.0, s);
            }
            HIPCHK(hipGetLastError());
            if (kernel != PSIVOX_KERNEL_PLAIN && !(mb.periodic && mb.full_window)) {
                k_targets<<<grid_blocks(nt, 64), 64, 0, s>>>(tets, nt, g, mb);
                HIPCHK(hipGetLastError());
            }
            if (kernel == PSIVOX_KERNEL_PLAIN) {
                if (W.timing) HIPCHK(hipEventRecord(W.ev0, s));
                k_vox_plain<<<grid_blocks(nt, 256), 256, 0, s>>>(tets, nt, g, mb, acc, status, counters);
                HIPCHK(hipGetLastError());
            } else {
                /* column tasks, then the per-tet residual pass; under the
                 * automatic choice the wide tets take a second pass in FP64
                 * (a tet that cannot take the column path stops the passes:
                 * the status check below restarts with the FP64 kernel) */
                const PassGeo G{g, mb, acc, nullptr, nullptr};
                int64_t nwide = 0;
                for (int pass = first_pass;; pass = PASS_WIDE) {
                    bool fell = false;
                    const double* ptets;
                    int64_t pnt;
                    if ((rc = pass_tets(W, pass, tets, nt, G, nwide, counters, s, dev, &ptets, &pnt))) return rc;
                    rc = for_task_chunks(
                        W, ptets, pnt, G, pass, status, counters, s, dev, &fell,
                        pass == PASS_NARROW ? &nwide : nullptr,
                        [&](int64_t a0, int64_t a1, const int64_t* colstart, int64_t base, const uint64_t* tasks,
                            int64_t nsub) -> int {
                            int rc2;
                            if ((rc2 = W.colsum.get((size_t)nsub * sizeof(double), dev))) return rc2;
                            if ((rc2 = W.colkey.get((size_t)nsub * sizeof(uint64_t), dev))) return rc2;
                            if (g.order > 0 && (rc2 = W.colmom.get((size_t)nsub * 16 * sizeof(float), dev)))
                                return rc2;
                            if ((rc2 = W.perm.get(order_bytes(nsub), dev))) return rc2;
                            TaskOut out{(double*)W.colsum.p, (uint64_t*)W.colkey.p, (float*)W.colmom.p};
                            rc2 = launch_cols(ptets, tasks, nsub, G, pass_fp64(pass), out, W.perm.p, status, counters, s,
                                              W.timing ? W.ev0 : nullptr);
                            if (rc2)
                                return set_err(PSIVOX_E_HIP, "column kernel launch failed: %s",
                                               hipGetErrorString((hipError_t)rc2));
                            if (W.timing) {
                                HIPCHK(hipEventRecord(W.ev1, s));
                                HIPCHK(hipEventSynchronize(W.ev1));
                                float ms = 0.f;
                                HIPCHK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
                                kms += ms;
                            }
                            if ((rc2 = W.tetb.get((size_t)nt * sizeof(int64_t), dev)) ||
                                (rc2 = W.slist.get((size_t)(WLIST_HEAD + nt) * sizeof(uint32_t), dev)))
                                return rc2;
                            rc2 = launch_residual(ptets, a0, a1, G, colstart, base, tasks, out, (double*)W.colcnt.p,
                                                  (int64_t*)W.tetb.p, (uint32_t*)W.slist.p, s);
                            if (rc2)
                                return set_err(PSIVOX_E_HIP, "residual kernel launch failed: %s",
                                               hipGetErrorString((hipError_t)rc2));
                            return 0;
                        });
                    if (rc) return rc;
                    if (fell || pass != PASS_NARROW || nwide == 0) break;
                }
                nwide_total += nwide;
            }
            if (kernel == PSIVOX_KERNEL_PLAIN && W.timing) {
                HIPCHK(hipEventRecord(W.ev1, s));
                HIPCHK(hipEventSynchronize(W.ev1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
                kms += ms;
            }
            ntets_total += nt;
            if (kernel != PSIVOX_KERNEL_PLAIN) {
                unsigned hs = 0;
                HIPCHK(hipMemcpyAsync(&hs, status, sizeof hs, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                if (hs & ST_FALLBACK) {
                    /* a tet exceeded the FP32 path's vertex budget: redo the
                     * whole call with the FP64 kernel from the untouched fields */
                    kernel = PSIVOX_KERNEL_PLAIN;
                    fell_back = 1;
                    HIPCHK(hipMemsetAsync(counters, 0, 64, s));
                    if ((rc = start_acc())) return rc;
                    e0 = 0;
                    kms = 0.0;
                    ntets_total = 0;
                    nwide_total = 0;
                    continue;
                }
            }
        }
        e0 = e1;
        if (o->progress && o->progress(o->progress_ctx, e0, ne)) {
            /* leave what was deposited so far, as the reference does on Ctrl-C */
            if ((rc = finish_acc())) return rc;
            for (int f = 0; f < 6; ++f)
                if (grid->fields[f] && !grid->on_device)
                    (void)hipMemcpyAsync(grid->fields[f], fd[f], (size_t)nvox * FIELD_WIDTH[f] * sizeof(double),
                                         hipMemcpyDeviceToHost, s);
            (void)hipStreamSynchronize(s);
            return set_err(PSIVOX_E_CANCELLED, "cancelled after %lld of %lld elements", (long long)e0, (long long)ne);
        }
    }

    if ((rc = finish_acc())) return rc;
    for (int f = 0; f < 6; ++f)
        if (grid->fields[f] && !grid->on_device &&
            (rc = copy_d2h(grid->fields[f], fd[f], (size_t)nvox * FIELD_WIDTH[f] * sizeof(double), s)))
            return rc;
    unsigned long long hc[8];
    HIPCHK(hipMemcpyAsync(hc, counters, 64, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    unsigned hstatus = (unsigned)(hc[4] & 0xffffffffu);
    if (hstatus & ST_OVERFLOW)
        return set_err(PSIVOX_E_OVERFLOW, "a polyhedron exceeded its vertex budget (degenerate input?)");
    if (hstatus & ST_BADGEOM) return set_err(PSIVOX_E_OVERFLOW, "inconsistent polyhedron graph (degenerate input?)");
    if (st) {
        st->nfrag = (int64_t)hc[0];
        st->nghosts = (int64_t)hc[1];
        st->ntets = ntets_total;
        st->kernel_ms = kms;
        st->kernel = kernel;
        st->reserved = fell_back;
        st->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        st->nelem_used = -1;
        st->nwide = nwide_total;
    }
    return 0;
}

/*
 * The adjoint of the fields m, v, vv of psivox_voxels (refine level 0):
 * dl_dpos += dL/dpos, dl_dvel += dL/dvel for the upstream gradients
 * dl_df[PSIVOX_M, V, VV] (null: zero; x, xx, xv must be null).  With only
 * dL/dm the weights are dL/dm itself and the velocities get nothing; with
 * dL/dv or dL/dvv grid->fields must hold the forward's m, v (and vv).
 */
int run_adjoint(Workspace& W, const psivox_mesh* mesh, const psivox_grid* grid, const double* const* dl_df,
                double* dl_dpos, double* dl_dvel, const psivox_opts* o, hipStream_t s, int dev) {
    int rc = check_mesh(mesh);
    if (rc) return rc;
    if (dl_df[PSIVOX_X] || dl_df[PSIVOX_XX] || dl_df[PSIVOX_XV])
        return set_err(PSIVOX_E_MODE, "the adjoint takes gradients of m, v and vv only");
    const bool full = dl_df[PSIVOX_V] || dl_df[PSIVOX_VV];
    if (!dl_df[PSIVOX_M] && !full) return set_err(PSIVOX_E_ARG, "no field gradient given");
    if (full && (!grid->fields[PSIVOX_M] || !grid->fields[PSIVOX_V] || (dl_df[PSIVOX_VV] && !grid->fields[PSIVOX_VV])))
        return set_err(PSIVOX_E_ARG, "the adjoint of v and vv needs the forward's fields m, v (and vv) in grid->fields");
    GridInfo g;
    for (int a = 0; a < 3; ++a) {
        if (grid->n[a] <= 0) return set_err(PSIVOX_E_ARG, "grid n[%d] = %d must be positive", a, grid->n[a]);
        /* adjoint tasks carry the plane index in 17 bits, planes 0..n */
        if (grid->n[a] >= ADJ_MAXPLANE)
            return set_err(PSIVOX_E_ARG, "the mass adjoint takes at most %d cells per axis (n[%d] = %d)",
                           ADJ_MAXPLANE - 1, a, grid->n[a]);
        if (!(grid->d[a] > 0.0)) return set_err(PSIVOX_E_ARG, "grid d[%d] must be positive", a);
        g.n[a] = grid->n[a];
        g.w0[a] = grid->window[0][a];
        g.w1[a] = grid->window[1][a];
        g.d[a] = grid->d[a];
    }
    g.flags = 0;
    set_layout(g);
    MeshBox mb;
    mb.periodic = mesh->periodic ? 1 : 0;
    mb.full_window = mb.periodic;
    for (int a = 0; a < 3; ++a) {
        mb.lo[a] = mesh->box[0][a];
        mb.hi[a] = mesh->box[1][a];
        if (mb.lo[a] != g.w0[a] || mb.hi[a] != g.w1[a]) mb.full_window = 0;
    }
    const int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int vpe = mesh->elemtype;
    const int per = vpe == 4 ? 1 : 6;
    MeshDev md{};
    if ((rc = mesh_to_device(W, mesh, dev, s, md))) return rc;
    /* the cells' weights on the device (adj_in: forward m, v, vv, then the
     * gradients of m, v, vv, then dL/dvel) */
    auto dev_field = [&](int slot, const double* src, int width, const double** out) -> int {
        *out = src;
        if (!src || grid->on_device) return 0;
        const void* p;
        int r = upload(W.adj_in[slot], src, (size_t)nvox * width * sizeof(double), dev, s, &p);
        if (r) return r;
        *out = (const double*)p;
        return 0;
    };
    const double* w = nullptr;
    if (!full) {
        if ((rc = dev_field(3, dl_df[PSIVOX_M], 1, &w))) return rc;
    } else {
        const double *fm, *fv, *fvv = nullptr, *gm = nullptr, *gv = nullptr, *gvv = nullptr;
        if ((rc = dev_field(0, grid->fields[PSIVOX_M], 1, &fm))) return rc;
        if ((rc = dev_field(1, grid->fields[PSIVOX_V], 3, &fv))) return rc;
        if (dl_df[PSIVOX_VV] && (rc = dev_field(2, grid->fields[PSIVOX_VV], 9, &fvv))) return rc;
        if ((rc = dev_field(3, dl_df[PSIVOX_M], 1, &gm))) return rc;
        if ((rc = dev_field(4, dl_df[PSIVOX_V], 3, &gv))) return rc;
        if ((rc = dev_field(5, dl_df[PSIVOX_VV], 9, &gvv))) return rc;
        if ((rc = W.adj_v.get((size_t)nvox * ADJ_NW * sizeof(double), dev))) return rc;
        k_adj_weights<<<grid_blocks(nvox, 256), 256, 0, s>>>(nvox, fm, fv, fvv, gm, gv, gvv, (double*)W.adj_v.p);
        HIPCHK(hipGetLastError());
        w = (const double*)W.adj_v.p;
    }
    /* the gradients on the device */
    const size_t gbytes = (size_t)mesh->npart * 3 * sizeof(double);
    double* dpos = dl_dpos;
    double* dvel = full ? dl_dvel : nullptr;
    if (dpos && !mesh->on_device && mesh->npart > 0) {
        const void* p;
        if ((rc = upload(W.adj_g, dl_dpos, gbytes, dev, s, &p))) return rc;
        dpos = (double*)p;
    }
    if (dvel && !mesh->on_device && mesh->npart > 0) {
        const void* p;
        if ((rc = upload(W.adj_in[6], dl_dvel, gbytes, dev, s, &p))) return rc;
        dvel = (double*)p;
    }
    if ((rc = W.counters.get(64, dev))) return rc;
    unsigned long long* counters = (unsigned long long*)W.counters.p;
    unsigned* status = (unsigned*)(counters + 4);
    HIPCHK(hipMemsetAsync(counters, 0, 64, s));
    const int64_t ne = mesh->nelem;
    const int64_t batch = o->batch_elements > 0 ? o->batch_elements : (int64_t)1 << 22;
    AdjArgs A;
    A.w = w;
    A.md = md;
    A.per = per;
    A.dpos = dpos;
    A.dvel = dvel;
    A.prm = nullptr;
    /* refinement: the forward's expansion (counts, offsets), each tet with
     * its parameter record, so that the gradients of its vertices reach the
     * element's nodes; the refinement decisions are piecewise constant in
     * the positions and carry no gradient */
    const int maxlvl = vpe == 4 ? 0 : std::min(o->refine_max_lvl, (int)MAXLVL);
    const double dxmin = std::min(g.d[0], std::min(g.d[1], g.d[2]));
    const double tol2 = (o->refine_tolerance * dxmin) * (o->refine_tolerance * dxmin);
    std::vector<int64_t> hoffs;
    if (maxlvl > 0 && ne > 0) {
        if ((rc = W.counts.get((size_t)(ne + 1) * sizeof(int64_t), dev))) return rc;
        if ((rc = W.offs.get((size_t)(ne + 1) * sizeof(int64_t), dev))) return rc;
        int64_t* counts = (int64_t*)W.counts.p;
        HIPCHK(hipMemsetAsync(counts + ne, 0, sizeof(int64_t), s));
        k_expand_count<<<grid_blocks(ne, 128), 128, 0, s>>>(md, g, mb, ne, maxlvl, tol2, counts);
        HIPCHK(hipGetLastError());
        size_t tmp_bytes = 0;
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, (int64_t*)W.offs.p, ne + 1, s));
        if ((rc = W.scan_tmp.get(tmp_bytes, dev))) return rc;
        HIPCHK(hipcub::DeviceScan::ExclusiveSum(W.scan_tmp.p, tmp_bytes, counts, (int64_t*)W.offs.p, ne + 1, s));
        hoffs.resize(ne + 1);
        HIPCHK(hipMemcpyAsync(hoffs.data(), W.offs.p, (ne + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    const int64_t max_tets_batch = (int64_t)6 << 22;
    for (int64_t e0 = 0; e0 < ne;) {
        int64_t e1, nt;
        if (maxlvl > 0) {
            /* largest e1 with offs[e1] - offs[e0] <= max_tets_batch (at least one element) */
            e1 = e0 + 1;
            int64_t lo = e0 + 1, hi = ne;
            while (lo <= hi) {
                const int64_t mid = (lo + hi) / 2;
                if (hoffs[mid] - hoffs[e0] <= max_tets_batch) {
                    e1 = mid;
                    lo = mid + 1;
                } else
                    hi = mid - 1;
            }
            nt = hoffs[e1] - hoffs[e0];
        } else {
            e1 = std::min(ne, e0 + batch);
            nt = (e1 - e0) * per;
        }
        A.e0 = e0;
        if ((rc = W.tets.get((size_t)std::max<int64_t>(nt, 1) * TREC * sizeof(double), dev))) return rc;
        double* tets = (double*)W.tets.p;
        if (maxlvl > 0) {
            if ((rc = W.adj_prm.get((size_t)std::max<int64_t>(nt, 1) * PREC * sizeof(uint64_t), dev))) return rc;
            A.prm = (const uint64_t*)W.adj_prm.p;
            k_expand_emit<<<grid_blocks(e1 - e0, 128), 128, 0, s>>>(md, g, mb, e0, e1, maxlvl, tol2,
                                                                    (const int64_t*)W.offs.p, hoffs[e0], tets,
                                                                    (uint64_t*)W.adj_prm.p);
        } else {
            launch_expand_fixed(md, g, mb, e0, e1 - e0, tets, 0.0, s);
        }
        HIPCHK(hipGetLastError());
        if (nt == 0) {
            e0 = e1;
            continue;
        }
        if (dpos) {
            /* plane tasks: count, scan, emit (chunks of tets whose tasks fit) */
            if ((rc = W.colcnt.get((size_t)(nt + 1) * sizeof(int64_t), dev))) return rc;
            if ((rc = W.colstart.get((size_t)(nt + 1) * sizeof(int64_t), dev))) return rc;
            int64_t* cnt = (int64_t*)W.colcnt.p;
            int64_t* start = (int64_t*)W.colstart.p;
            HIPCHK(hipMemsetAsync(cnt + nt, 0, sizeof(int64_t), s));
            k_adj_count<<<grid_blocks(nt, 256), 256, 0, s>>>(tets, nt, g, mb, cnt, status);
            HIPCHK(hipGetLastError());
            size_t tb = 0;
            HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, start, nt + 1, s));
            if ((rc = W.scan_tmp.get(tb, dev))) return rc;
            HIPCHK(hipcub::DeviceScan::ExclusiveSum(W.scan_tmp.p, tb, cnt, start, nt + 1, s));
            std::vector<int64_t> hs((size_t)nt + 1);
            HIPCHK(hipMemcpyAsync(hs.data(), start, (nt + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            const int64_t cap = (int64_t)1 << 28;
            if ((rc = W.tasks.get((size_t)std::min(hs[nt], cap) * sizeof(uint64_t) + 8, dev))) return rc;
            if ((rc = W.asorted.get((size_t)std::min(hs[nt], cap) * sizeof(uint64_t) + 8, dev))) return rc;
            uint64_t* tasks = (uint64_t*)W.tasks.p;
            uint64_t* sorted = (uint64_t*)W.asorted.p;
            for (int64_t a0 = 0; a0 < nt;) {
                int64_t a1 = std::upper_bound(hs.begin() + a0 + 1, hs.end(), hs[a0] + cap) - hs.begin() - 1;
                if (a1 <= a0) a1 = a0 + 1;
                const int64_t nsub = hs[a1] - hs[a0];
                if (nsub > 0) {
                    k_adj_emit<<<grid_blocks(a1 - a0, 256), 256, 0, s>>>(tets, a0, a1, g, mb, start, hs[a0], tasks);
                    HIPCHK(hipGetLastError());
                    k_adj_order<<<(unsigned)((nsub + ORDW - 1) / ORDW), 1024, 0, s>>>(tets, tasks, nsub, g, mb, sorted);
                    HIPCHK(hipGetLastError());
                    if (full)
                        k_adj<true><<<grid_blocks(nsub, FBLOCK), FBLOCK, 0, s>>>(tets, sorted, nsub, g, mb, A);
                    else
                        k_adj<false><<<grid_blocks(nsub, FBLOCK), FBLOCK, 0, s>>>(tets, sorted, nsub, g, mb, A);
                    HIPCHK(hipGetLastError());
                }
                a0 = a1;
            }
        }
        if (dvel) {
            /* column tasks of the batch, the velocity part per task */
            bool fell = false;
            const PassGeo G{g, mb, nullptr, nullptr, nullptr};
            rc = for_task_chunks(W, tets, nt, G, PASS_ALL64, status, counters, s, dev, &fell, nullptr,
                                 [&](int64_t, int64_t, const int64_t*, int64_t, const uint64_t* tasks, int64_t nsub) -> int {
                                     /* the forward's cost order and a persistent grid */
                                     int rc2;
                                     if ((rc2 = W.perm.get(order_bytes(nsub), dev))) return rc2;
                                     uint64_t* sorted = order_sorted(W.perm.p, nsub);
                                     unsigned* next = perm_counter(W.perm.p, nsub);
                                     k_col_order<false><<<(unsigned)((nsub + ORDW - 1) / ORDW), ORDNT, 0, s>>>(
                                         tets, tasks, nsub, g, mb, nullptr, order_perm(W.perm.p), sorted, nullptr, next);
                                     const int64_t hold = resident_blocks(k_adj_vel, 0);
                                     const int64_t nwave = (nsub + FBLOCK - 1) / FBLOCK;
                                     k_adj_vel<<<(unsigned)(nwave < hold ? nwave : hold), FBLOCK, 0, s>>>(
                                         tets, sorted, next, nsub, g, mb, A, status);
                                     HIPCHK(hipGetLastError());
                                     return 0;
                                 });
            if (rc) return rc;
            if (fell) return set_err(PSIVOX_E_OVERFLOW, "adjoint: a tet cannot take the column path");
        }
        if (o->progress && o->progress(o->progress_ctx, e1, ne)) {
            (void)hipStreamSynchronize(s);
            return set_err(PSIVOX_E_CANCELLED, "cancelled after %lld of %lld elements", (long long)e1, (long long)ne);
        }
        e0 = e1;
    }
    if (dpos && dpos != dl_dpos) HIPCHK(hipMemcpyAsync(dl_dpos, dpos, gbytes, hipMemcpyDeviceToHost, s));
    if (dvel && dvel != dl_dvel) HIPCHK(hipMemcpyAsync(dl_dvel, dvel, gbytes, hipMemcpyDeviceToHost, s));
    unsigned hst = 0;
    HIPCHK(hipMemcpyAsync(&hst, status, sizeof hst, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hst & ST_BADGEOM) return set_err(PSIVOX_E_OVERFLOW, "adjoint: a tet has non-finite vertices");
    if (hst & ST_OVERFLOW) return set_err(PSIVOX_E_OVERFLOW, "adjoint: a column exceeded its vertex budget");
    return 0;
}

/*
 * Many (mesh, grid) pairs in one device pass.  Meshes are grouped up to the
 * tet-buffer capacity; per group: raw conversion and expansion per mesh
 * (records tagged with the mesh's index), one count / scan / emit over all
 * its tets, one ordering and one column kernel over all its tasks (each
 * wavefront reads its meshes' geometry from a table), one residual pass,
 * and finalize per mesh.  The pass needs one field set for the whole batch,
 * refinement level 0 and the column kernel; other batches, and a group in
 * which a tet needs the FP64 kernel, run mesh by mesh.
 */
int run_batched(Workspace& W, int nmesh, const psivox_mesh* meshes, psivox_grid* grids, const psivox_opts* o,
                psivox_stats* stats, hipStream_t s, int dev) {
    auto t_start = std::chrono::steady_clock::now();
    int rc;
    const int kernel = o->kernel == PSIVOX_KERNEL_AUTO ? PSIVOX_KERNEL_FAST : o->kernel;
    const bool route = o->kernel == PSIVOX_KERNEL_AUTO;
    std::vector<GridInfo> G(nmesh);
    std::vector<MeshBox> MB(nmesh);
    bool same = kernel == PSIVOX_KERNEL_FAST && o->mode == PSIVOX_MODE_DENSITY;
    for (int i = 0; i < nmesh; ++i) {
        if ((rc = check_mesh(&meshes[i]))) return rc;
        if ((rc = fill_grid_info(&grids[i], o->xv_true_covariance, G[i]))) return rc;
        MB[i].periodic = meshes[i].periodic ? 1 : 0;
        MB[i].full_window = MB[i].periodic;
        for (int a = 0; a < 3; ++a) {
            MB[i].lo[a] = meshes[i].box[0][a];
            MB[i].hi[a] = meshes[i].box[1][a];
            if (MB[i].lo[a] != G[i].w0[a] || MB[i].hi[a] != G[i].w1[a]) MB[i].full_window = 0;
        }
        if (G[i].flags != G[0].flags) same = false;
        if (o->refine_max_lvl > 0 && meshes[i].elemtype != 4) same = false;
    }
    auto one = [&](int i) -> int {
        return run_one(W, &meshes[i], &grids[i], o, stats ? &stats[i] : nullptr, s, dev);
    };
    if (!same) {
        for (int i = 0; i < nmesh; ++i)
            if ((rc = one(i))) return rc;
        return 0;
    }
    const int64_t max_tets = (int64_t)6 << 22;
    auto per_of = [&](int i) -> int64_t { return meshes[i].elemtype == 4 ? 1 : 6; };
    int64_t elems_total = 0, elems_done = 0;
    for (int i = 0; i < nmesh; ++i) elems_total += meshes[i].nelem;
    int i0 = 0;
    while (i0 < nmesh) {
        /* the group [i0, i1) */
        int i1 = i0;
        int64_t nt = 0;
        /* at most 65535 meshes a group: the ordering hands the column kernel
         * each task's mesh as 16 bits */
        while (i1 < nmesh && i1 - i0 < 65535 && (i1 == i0 || nt + meshes[i1].nelem * per_of(i1) <= max_tets)) {
            nt += meshes[i1].nelem * per_of(i1);
            i1++;
        }
        if (i1 == i0 + 1 || nt == 0) {
            for (int i = i0; i < i1; ++i)
                if ((rc = one(i))) return rc;
            for (int i = i0; i < i1; ++i) elems_done += meshes[i].nelem;
            i0 = i1;
            continue;
        }
        const int nm = i1 - i0;
        /* device layout of the group: mesh arrays, fields, accumulators */
        std::vector<MeshDev> md(nm);
        std::vector<int64_t> T(nm + 1, 0), acc_off(nm + 1, 0), fld_off(nm + 1, 0);
        size_t mesh_bytes = 0;
        auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
        for (int q = 0; q < nm; ++q) {
            const psivox_mesh& m = meshes[i0 + q];
            const GridInfo& g = G[i0 + q];
            const int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
            T[q + 1] = T[q] + m.nelem * per_of(i0 + q);
            acc_off[q + 1] = acc_off[q] + nvox * g.nacc;
            int64_t fw = 0;
            for (int f = 0; f < 6; ++f)
                if (grids[i0 + q].fields[f] && !grids[i0 + q].on_device) fw += FIELD_WIDTH[f];
            fld_off[q + 1] = fld_off[q] + nvox * fw;
            if (!m.on_device)
                mesh_bytes += al(m.npart * 24) * (m.vel ? 2 : 1) + al(m.nelem * 8) + al(m.nelem * m.elemtype * 4);
        }
        if ((rc = W.bmesh.get(mesh_bytes, dev))) return rc;
        if ((rc = W.bfield.get((size_t)fld_off[nm] * sizeof(double), dev))) return rc;
        if ((rc = W.bacc.get((size_t)acc_off[nm] * sizeof(double), dev))) return rc;
        if ((rc = W.tets.get((size_t)nt * TREC * sizeof(double), dev))) return rc;
        if ((rc = W.bmcnt.get((size_t)(mcnt_words(nm) + 8) * sizeof(unsigned long long), dev))) return rc;
        if ((rc = W.counters.get(64, dev))) return rc;
        unsigned long long* counters = (unsigned long long*)W.counters.p;
        unsigned* status = (unsigned*)(counters + 4);
        unsigned long long* mcnt = (unsigned long long*)W.bmcnt.p;
        HIPCHK(hipMemsetAsync(counters, 0, 64, s));
        HIPCHK(hipMemsetAsync(mcnt, 0, (size_t)mcnt_words(nm) * sizeof(unsigned long long), s));
        double* tets = (double*)W.tets.p;
        double* bacc = (double*)W.bacc.p;
        std::vector<FieldPtrs> fp(nm);
        std::vector<std::array<double*, 6>> fdev(nm);
        char* mp = (char*)W.bmesh.p;
        for (int q = 0; q < nm; ++q) {
            const int i = i0 + q;
            const psivox_mesh& m = meshes[i];
            md[q] = MeshDev{};
            md[q].vpe = m.elemtype;
            if (m.on_device) {
                md[q].pos = m.pos;
                md[q].vel = m.vel;
                md[q].mass = m.mass;
                md[q].conn = m.conn;
            } else if (m.nelem > 0) {
                if ((rc = copy_h2d(mp, m.pos, m.npart * 24, s))) return rc;
                md[q].pos = (const double*)mp;
                mp += al(m.npart * 24);
                if (m.vel) {
                    if ((rc = copy_h2d(mp, m.vel, m.npart * 24, s))) return rc;
                    md[q].vel = (const double*)mp;
                    mp += al(m.npart * 24);
                }
                if ((rc = copy_h2d(mp, m.mass, m.nelem * 8, s))) return rc;
                md[q].mass = (const double*)mp;
                mp += al(m.nelem * 8);
                if ((rc = copy_h2d(mp, m.conn, m.nelem * m.elemtype * 4, s))) return rc;
                md[q].conn = (const int32_t*)mp;
                mp += al(m.nelem * m.elemtype * 4);
            }
            if ((rc = check_conn(W, md[q].conn, m.nelem, m.elemtype, m.npart, dev, s))) return rc;
            const GridInfo& g = G[i];
            const int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
            double* fbase = (double*)W.bfield.p + fld_off[q];
            for (int f = 0; f < 6; ++f) {
                fdev[q][f] = nullptr;
                if (!grids[i].fields[f]) continue;
                if (grids[i].on_device) {
                    fdev[q][f] = grids[i].fields[f];
                } else {
                    fdev[q][f] = fbase;
                    if ((rc = copy_h2d(fbase, grids[i].fields[f], (size_t)nvox * FIELD_WIDTH[f] * sizeof(double), s)))
                        return rc;
                    fbase += nvox * FIELD_WIDTH[f];
                }
            }
            for (int f = 0; f < 6; ++f) fp[q].f[f] = fdev[q][f];
            HIPCHK(hipMemsetAsync(bacc + acc_off[q], 0, (size_t)nvox * g.nacc * sizeof(double), s));
            if (m.nelem > 0) {
                launch_expand_fixed(md[q], g, MB[i], 0, m.nelem, tets + TREC * T[q], (double)q, s);
                HIPCHK(hipGetLastError());
                if (!(MB[i].periodic && MB[i].full_window)) {
                    k_targets<<<grid_blocks(T[q + 1] - T[q], 64), 64, 0, s>>>(tets + TREC * T[q], T[q + 1] - T[q], g,
                                                                             MB[i]);
                    HIPCHK(hipGetLastError());
                }
            }
        }
        std::vector<BatchGeo> hb(nm);
        for (int q = 0; q < nm; ++q) hb[q] = BatchGeo{G[i0 + q], MB[i0 + q], bacc + acc_off[q]};
        if ((rc = W.btable.get(nm * sizeof(BatchGeo), dev))) return rc;
        HIPCHK(hipMemcpyAsync(W.btable.p, hb.data(), nm * sizeof(BatchGeo), hipMemcpyHostToDevice, s));
        const PassGeo PG{G[i0], MB[i0], nullptr, (const BatchGeo*)W.btable.p, mcnt};
        double kms = 0.0;
        bool fell = false;
        int64_t nwide = 0;
        /* one ordering and one column kernel over the tasks of all the
         * group's meshes per chunk (and pass) */
        for (int pass = route ? PASS_NARROW : PASS_ALL;; pass = PASS_WIDE) {
            const double* ptets;
            int64_t pnt;
            if ((rc = pass_tets(W, pass, tets, nt, PG, nwide, counters, s, dev, &ptets, &pnt))) return rc;
            rc = for_task_chunks(
                W, ptets, pnt, PG, pass, status, counters, s, dev, &fell, pass == PASS_NARROW ? &nwide : nullptr,
                [&](int64_t a0, int64_t a1, const int64_t* colstart, int64_t base, const uint64_t* tasks,
                    int64_t nsub) -> int {
                    int rc2;
                    if ((rc2 = W.colsum.get((size_t)nsub * sizeof(double), dev))) return rc2;
                    if ((rc2 = W.colkey.get((size_t)nsub * sizeof(uint64_t), dev))) return rc2;
                    if (G[i0].order > 0 && (rc2 = W.colmom.get((size_t)nsub * 16 * sizeof(float), dev)))
                        return rc2;
                    if ((rc2 = W.perm.get(order_bytes(nsub), dev))) return rc2;
                    TaskOut out{(double*)W.colsum.p, (uint64_t*)W.colkey.p, (float*)W.colmom.p};
                    rc2 = launch_cols(ptets, tasks, nsub, PG, pass == PASS_WIDE, out, W.perm.p, status, counters, s,
                                      W.ev0);
                    if (rc2)
                        return set_err(PSIVOX_E_HIP, "column kernel launch failed: %s",
                                       hipGetErrorString((hipError_t)rc2));
                    HIPCHK(hipEventRecord(W.ev1, s));
                    if ((rc2 = W.tetb.get((size_t)nt * sizeof(int64_t), dev)) ||
                        (rc2 = W.slist.get((size_t)(WLIST_HEAD + nt) * sizeof(uint32_t), dev)))
                        return rc2;
                    rc2 = launch_residual(ptets, a0, a1, PG, colstart, base, tasks, out, (double*)W.colcnt.p,
                                          (int64_t*)W.tetb.p, (uint32_t*)W.slist.p, s);
                    if (rc2)
                        return set_err(PSIVOX_E_HIP, "residual kernel launch failed: %s",
                                       hipGetErrorString((hipError_t)rc2));
                    HIPCHK(hipEventSynchronize(W.ev1));
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
                    kms += ms;
                    return 0;
                });
            if (rc) return rc;
            if (fell || pass != PASS_NARROW || nwide == 0) break;
        }
        unsigned hs = 0;
        HIPCHK(hipMemcpyAsync(&hs, status, sizeof hs, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (fell || (hs & ST_FALLBACK)) {
            /* the fields are untouched until finalize: redo the group mesh by
             * mesh (each falls back to the FP64 kernel where it must) */
            for (int i = i0; i < i1; ++i)
                if ((rc = one(i))) return rc;
        } else {
            std::vector<unsigned long long> hm((size_t)mcnt_words(nm));
            for (int q = 0; q < nm; ++q) {
                const int i = i0 + q;
                const GridInfo& g = G[i];
                const int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
                launch_merge_finalize(g, bacc + acc_off[q], fp[q], nvox, s);
                HIPCHK(hipGetLastError());
                for (int f = 0; f < 6; ++f)
                    if (grids[i].fields[f] && !grids[i].on_device &&
                        (rc = copy_d2h(grids[i].fields[f], fdev[q][f], (size_t)nvox * FIELD_WIDTH[f] * sizeof(double),
                                       s)))
                        return rc;
            }
            HIPCHK(hipMemcpyAsync(hm.data(), mcnt, (size_t)mcnt_words(nm) * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (stats) {
                const double wall =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
                for (int q = 0; q < nm; ++q) {
                    psivox_stats& st = stats[i0 + q];
                    int64_t sum[3] = {0, 0, 0};
                    for (int slot = 0; slot < 3; ++slot)
                        for (int k = 0; k < MSTRIPE; ++k) sum[slot] += (int64_t)hm[mcnt_word(q, slot, k)];
                    st.nfrag = sum[0];
                    st.nghosts = sum[1];
                    st.nwide = sum[2];
                    st.ntets = T[q + 1] - T[q];
                    st.kernel_ms = kms;
                    st.total_ms = wall;
                    st.kernel = PSIVOX_KERNEL_FAST;
                    st.reserved = 0;
                    st.nelem_used = -1;
                }
            }
        }
        for (int i = i0; i < i1; ++i) elems_done += meshes[i].nelem;
        if (o->progress && o->progress(o->progress_ctx, elems_done, elems_total))
            return set_err(PSIVOX_E_CANCELLED, "cancelled after %lld of %lld elements", (long long)elems_done,
                           (long long)elems_total);
        i0 = i1;
    }
    return 0;
}

/*
 * Scratch of the allocation-free entry point (psivox_voxels_noalloc): every
 * device buffer beyond the caller's inputs and outputs is carved from this
 * array.  Each GPU the library runs on holds its own copy, at its own
 * address, until the process ends.  Host code never passes the array's name
 * to a kernel or a copy: it takes the device address once per device from
 * hipGetSymbolAddress (scratch_address) and uses that.
 */
constexpr size_t NOALLOC_SCRATCH = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH];

/* the scratch array's address on device dev (current), looked up once */
int scratch_address(int dev, char** out) {
    static std::mutex mu;
    static void* ptr[256];
    if (dev < 0 || dev >= 256) return set_err(PSIVOX_E_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(PSIVOX_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        ptr[dev] = p;
    }
    *out = (char*)ptr[dev];
    return 0;
}

/* the thread's arena while an allocation-free call runs */
struct ArenaScope {
    explicit ArenaScope(Arena* a) { t_arena = a; }
    ~ArenaScope() { t_arena = nullptr; }
};

Workspace g_ws_noalloc;

/* the x-slab and y-slab step tables in the current device's memory, once
 * per device (every thread's calls see them once the first upload returned) */
int upload_tables(int dev) {
    static std::mutex mu;
    static bool done[256];
    if (dev < 0 || dev >= 256) return set_err(PSIVOX_E_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(mu);
    if (done[dev]) return 0;
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(slab_steps), host_slab_steps(), sizeof(SlabStep) * 81, 0,
                             hipMemcpyHostToDevice));
    if (host_y2_count() != Y2STEPS) return set_err(PSIVOX_E_HIP, "y-slab table size mismatch");
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(y2_steps), host_y2_steps(), sizeof(Y2Step) * Y2STEPS, 0,
                             hipMemcpyHostToDevice));
    done[dev] = true;
    return 0;
}

/*
 * psivox_voxels without allocating: device mesh and grid, density mode,
 * refinement level 0.  The grid accumulator, the tet records of an element
 * batch and a chunk of column tasks are laid out in the device's scratch
 * array: the batch takes up to two fifths of what the accumulator leaves,
 * the task chunk the rest.  Runs on opts->stream (NULL: the null stream),
 * records no events (kernel_ms is 0).
 */
int run_noalloc(const psivox_mesh* mesh, psivox_grid* grid, const psivox_opts* o, psivox_stats* st) {
    int rc = check_mesh(mesh);
    if (rc) return rc;
    if (!mesh->on_device || !grid->on_device)
        return set_err(PSIVOX_E_ARG, "psivox_voxels_noalloc takes a device mesh and a device grid (on_device = 1)");
    if (o->mode != PSIVOX_MODE_DENSITY) return set_err(PSIVOX_E_MODE, "psivox_voxels_noalloc: density mode only");
    if (o->refine_max_lvl > 0 && mesh->elemtype != PSIVOX_SIMPLEX)
        return set_err(PSIVOX_E_MODE, "psivox_voxels_noalloc: refinement level 0 only");
    GridInfo g;
    if ((rc = fill_grid_info(grid, o->xv_true_covariance, g))) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_err(PSIVOX_E_NODEV, "no GPU visible");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if ((rc = upload_tables(dev))) return rc;
    char* base = nullptr;
    if ((rc = scratch_address(dev, &base))) return rc;
    hipStream_t s = (hipStream_t)o->stream;

    Workspace& W = g_ws_noalloc;
    W = Workspace{};
    W.timing = false;
    W.compact_wide = false; /* the FP64 pass runs in place: its scratch is laid out before the call */
    Arena A{base, NOALLOC_SCRATCH, 0};
    ArenaScope scope(&A);
    const int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
    if ((rc = W.acc.get((size_t)nvox * g.nacc * sizeof(double), dev))) return rc;
    if ((rc = W.counters.get(64, dev)) || (rc = W.connchk.get(sizeof(unsigned), dev)) ||
        (rc = W.cut.get(4 * sizeof(int64_t), dev)))
        return rc;
    /* element batch: tet records, per-tet counts, starts, residual slots
     * and the spread list in up to two fifths of the rest */
    const int per = mesh->elemtype == PSIVOX_SIMPLEX ? 1 : 6;
    const size_t per_tet = TREC * sizeof(double) + 3 * sizeof(int64_t) + sizeof(uint32_t);
    const int64_t nt_room = (int64_t)((A.size - A.off) * 2 / 5 / per_tet);
    int64_t ne_batch = std::max<int64_t>(1, nt_room / per);
    if (mesh->nelem > 0) ne_batch = std::min<int64_t>(ne_batch, mesh->nelem);
    const int64_t nt = ne_batch * per;
    size_t tb = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (const int64_t*)nullptr, (int64_t*)nullptr, nt + 1, s));
    if ((rc = W.tets.get((size_t)nt * TREC * sizeof(double), dev)) ||
        (rc = W.colcnt.get((size_t)(nt + 1) * sizeof(int64_t), dev)) ||
        (rc = W.colstart.get((size_t)(nt + 1) * sizeof(int64_t), dev)) ||
        (rc = W.tetb.get((size_t)nt * sizeof(int64_t), dev)) ||
        (rc = W.slist.get((size_t)(WLIST_HEAD + nt) * sizeof(uint32_t), dev)) || (rc = W.scan_tmp.get(tb, dev)))
        return rc;
    /* a chunk of column tasks: task words, mass sums, largest-fragment
     * keys and moments, and the ordering's scratch (order_bytes: the
     * processing order, the task words in that order, the mesh ids) */
    const size_t per_task = 4 * sizeof(uint64_t) + sizeof(uint32_t) + sizeof(uint16_t) +
                            (g.order > 0 ? 16 * sizeof(float) : 0);
    const size_t left = A.size - A.off;
    const int64_t cap = left > 8192 ? (int64_t)((left - 8192) / per_task) : 0;
    if (cap < 4096)
        return set_err(PSIVOX_E_NOMEM, "psivox_voxels_noalloc: the grid's accumulator (%zu bytes) leaves too little of the "
                                       "%zu bytes of scratch", (size_t)nvox * g.nacc * sizeof(double), A.size);
    W.task_cap = std::min<int64_t>(cap, (int64_t)1 << 28);
    if ((rc = W.tasks.get((size_t)W.task_cap * sizeof(uint64_t), dev)) ||
        (rc = W.colsum.get((size_t)W.task_cap * sizeof(double), dev)) ||
        (rc = W.colkey.get((size_t)W.task_cap * sizeof(uint64_t), dev)) ||
        (g.order > 0 && (rc = W.colmom.get((size_t)W.task_cap * 16 * sizeof(float), dev))) ||
        (rc = W.perm.get(order_bytes(W.task_cap), dev)))
        return rc;
    psivox_opts ob = *o;
    ob.batch_elements = ne_batch;
    return run_one(W, mesh, grid, &ob, st, s, dev);
}

int prepare(Workspace& W, const psivox_opts* o, hipStream_t* s, int* dev) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_err(PSIVOX_E_NODEV, "no GPU visible");
    HIPCHK(hipGetDevice(dev));
    int rc = upload_tables(*dev);
    if (rc) return rc;
    if (o && o->stream) {
        *s = (hipStream_t)o->stream;
    } else {
        if (!W.stream || W.stream_dev != *dev) {
            HIPCHK(hipStreamCreateWithFlags(&W.stream, hipStreamNonBlocking));
            W.stream_dev = *dev;
        }
        *s = W.stream;
    }
    if (!W.ev0) {
        HIPCHK(hipEventCreate(&W.ev0));
        HIPCHK(hipEventCreate(&W.ev1));
    }
    return 0;
}

}  // namespace

/* The element-box index behind psivox_elem_index_* and psivox_sample_vdf:
 * the mesh on the device (copies of a host mesh, or the caller's device
 * arrays), every element's wrapped box and whether it passed the window
 * test, and grow-only buffers for queries. */
struct psivox_elem_index {
    int dev = -1;
    int64_t nelem = 0;
    MeshDev md{};
    DevBuf pos, vel, mass, conn, box, valid, flags, cand, nsel, tmp, counts, offs, tets, out;
};

namespace {

int index_stream(const void* stream, hipStream_t* s, int* dev) {
    psivox_opts o{};
    o.stream = (void*)stream;
    return prepare(g_ws, &o, s, dev);
}

/* the elements whose box meets q into idx->cand (ascending); *nc = count */
int index_select(psivox_elem_index* X, const QueryBox& q, hipStream_t s, int64_t* nc) {
    int rc;
    const int64_t ne = X->nelem;
    *nc = 0;
    if (ne == 0) return 0;
    if ((rc = X->flags.get((size_t)ne, X->dev)) || (rc = X->cand.get((size_t)ne * sizeof(int32_t), X->dev)) ||
        (rc = X->nsel.get(sizeof(int64_t), X->dev)))
        return rc;
    k_box_meets<<<grid_blocks(ne, 256), 256, 0, s>>>((const double*)X->box.p, (const int8_t*)X->valid.p, ne, q,
                                                     (uint8_t*)X->flags.p);
    HIPCHK(hipGetLastError());
    hipcub::CountingInputIterator<int32_t> ids(0);
    size_t tb = 0;
    HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, tb, ids, (const uint8_t*)X->flags.p, (int32_t*)X->cand.p,
                                         (int64_t*)X->nsel.p, ne, s));
    if ((rc = X->tmp.get(tb, X->dev))) return rc;
    HIPCHK(hipcub::DeviceSelect::Flagged(X->tmp.p, tb, ids, (const uint8_t*)X->flags.p, (int32_t*)X->cand.p,
                                         (int64_t*)X->nsel.p, ne, s));
    HIPCHK(hipMemcpyAsync(nc, X->nsel.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int index_build(const psivox_mesh* mesh, const double window[2][3], hipStream_t s, int dev,
                psivox_elem_index* X) {
    int rc = check_mesh(mesh);
    if (rc) return rc;
    if (mesh->nelem > INT32_MAX) return set_err(PSIVOX_E_ARG, "the element index takes at most 2^31-1 elements");
    X->dev = dev;
    X->nelem = mesh->nelem;
    X->md = MeshDev{};
    X->md.vpe = mesh->elemtype;
    if (mesh->nelem == 0) return 0;
    if (mesh->on_device) {
        X->md.pos = mesh->pos;
        X->md.vel = mesh->vel;
        X->md.mass = mesh->mass;
        X->md.conn = mesh->conn;
    } else {
        const void* p;
        if ((rc = upload(X->pos, mesh->pos, (size_t)mesh->npart * 3 * sizeof(double), dev, s, &p))) return rc;
        X->md.pos = (const double*)p;
        if (mesh->vel) {
            if ((rc = upload(X->vel, mesh->vel, (size_t)mesh->npart * 3 * sizeof(double), dev, s, &p))) return rc;
            X->md.vel = (const double*)p;
        }
        if ((rc = upload(X->mass, mesh->mass, (size_t)mesh->nelem * sizeof(double), dev, s, &p))) return rc;
        X->md.mass = (const double*)p;
        if ((rc = upload(X->conn, mesh->conn, (size_t)mesh->nelem * mesh->elemtype * sizeof(int32_t), dev, s, &p)))
            return rc;
        X->md.conn = (const int32_t*)p;
    }
    if ((rc = check_conn(g_ws, X->md.conn, mesh->nelem, mesh->elemtype, mesh->npart, dev, s))) return rc;
    GridInfo g;
    memset(&g, 0, sizeof g);
    MeshBox mb;
    mb.periodic = mesh->periodic ? 1 : 0;
    mb.full_window = 0;
    for (int a = 0; a < 3; ++a) {
        g.w0[a] = window[0][a];
        g.w1[a] = window[1][a];
        mb.lo[a] = mesh->box[0][a];
        mb.hi[a] = mesh->box[1][a];
    }
    if ((rc = X->box.get((size_t)mesh->nelem * 6 * sizeof(double), dev)) ||
        (rc = X->valid.get((size_t)mesh->nelem, dev)))
        return rc;
    k_elem_boxes<<<grid_blocks(mesh->nelem, 128), 128, 0, s>>>(X->md, g, mb, mesh->nelem, (double*)X->box.p,
                                                               (int8_t*)X->valid.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

/* the multistream sample at x (psi_sample_vdf, src/cpsi.c:177-275): streams
 * in ascending element order, tets in refinement order */
int sample_vdf(psivox_elem_index* X, const double* x, double reftol, int maxlvl, double* rho, double* vel,
               int64_t cap, int64_t* n, hipStream_t s) {
    int rc;
    *n = 0;
    QueryBox q;
    for (int a = 0; a < 3; ++a) q.lo[a] = q.hi[a] = x[a];
    int64_t nc = 0;
    if ((rc = index_select(X, q, s, &nc))) return rc;
    if (nc == 0) return 0;
    if (maxlvl > MAXLVL) maxlvl = MAXLVL; /* src/refine.c:89-92 */
    if (maxlvl < 0) maxlvl = 0;
    GridInfo g; /* unbounded window, no wrap: every candidate expands as it lies */
    memset(&g, 0, sizeof g);
    MeshBox mb;
    mb.periodic = 0;
    mb.full_window = 0;
    for (int a = 0; a < 3; ++a) {
        g.w0[a] = -HUGE_VAL;
        g.w1[a] = HUGE_VAL;
        mb.lo[a] = mb.hi[a] = 0.0;
    }
    const double tol2 = reftol * reftol; /* position units here (src/cpsi.c:197) */
    if ((rc = X->counts.get((size_t)nc * sizeof(int64_t), X->dev)) ||
        (rc = X->offs.get((size_t)nc * sizeof(int64_t), X->dev)))
        return rc;
    k_vdf_count<<<grid_blocks(nc, 64), 64, 0, s>>>(X->md, g, mb, (const int32_t*)X->cand.p, nc, maxlvl, tol2,
                                                   (int64_t*)X->counts.p);
    HIPCHK(hipGetLastError());
    std::vector<int64_t> cnt((size_t)nc), off((size_t)nc);
    HIPCHK(hipMemcpyAsync(cnt.data(), X->counts.p, (size_t)nc * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int64_t nt = 0;
    for (int64_t c = 0; c < nc; ++c) {
        off[(size_t)c] = nt;
        nt += cnt[(size_t)c];
    }
    if (nt == 0) return 0;
    if ((rc = X->tets.get((size_t)nt * TREC * sizeof(double), X->dev)) ||
        (rc = X->out.get((size_t)nt * 5 * sizeof(double), X->dev)))
        return rc;
    HIPCHK(hipMemcpyAsync(X->offs.p, off.data(), (size_t)nc * sizeof(int64_t), hipMemcpyHostToDevice, s));
    k_vdf_emit<<<grid_blocks(nc, 64), 64, 0, s>>>(X->md, g, mb, (const int32_t*)X->cand.p, nc, maxlvl, tol2,
                                                  (const int64_t*)X->offs.p, (double*)X->tets.p);
    HIPCHK(hipGetLastError());
    k_vdf_sample<<<grid_blocks(nt, 256), 256, 0, s>>>((const double*)X->tets.p, nt, q, (double*)X->out.p);
    HIPCHK(hipGetLastError());
    std::vector<double> res((size_t)nt * 5);
    HIPCHK(hipMemcpyAsync(res.data(), X->out.p, res.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int64_t k = 0;
    for (int64_t t = 0; t < nt; ++t) {
        const double* r = &res[(size_t)t * 5];
        if (r[0] == 0.0) continue;
        if (k < cap) {
            if (rho) rho[k] = r[1];
            if (vel)
                for (int a = 0; a < 3; ++a) vel[3 * k + a] = r[2 + a];
        }
        k++;
    }
    *n = k;
    return 0;
}

}  // namespace

/* ------------------------------------------------------------------ */
/* C ABI                                                               */
/* ------------------------------------------------------------------ */

extern "C" {

int psivox_abi_version(void) { return PSIVOX_ABI_VERSION; }

const char* psivox_last_error(void) { return g_err.c_str(); }

int psivox_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int psivox_voxels(const psivox_mesh* mesh, psivox_grid* grid, const psivox_opts* opts, psivox_stats* stats) {
    if (!mesh || !grid) return set_err(PSIVOX_E_ARG, "mesh and grid are required");
    psivox_opts o{};
    o.refine_tolerance = 1.0;
    if (opts) o = *opts;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    hipStream_t s;
    int dev;
    int rc = prepare(g_ws, &o, &s, &dev);
    if (rc) return rc;
    return run_one(g_ws, mesh, grid, &o, stats, s, dev);
}

int psivox_voxels_noalloc(const psivox_mesh* mesh, psivox_grid* grid, const psivox_opts* opts, psivox_stats* stats) {
    if (!mesh || !grid) return set_err(PSIVOX_E_ARG, "mesh and grid are required");
    psivox_opts o{};
    o.refine_tolerance = 1.0;
    if (opts) o = *opts;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    return run_noalloc(mesh, grid, &o, stats);
}

int psivox_voxels_batched(int32_t nmesh, const psivox_mesh* meshes, psivox_grid* grids, const psivox_opts* opts,
                          psivox_stats* stats) {
    if (nmesh < 0 || (nmesh > 0 && (!meshes || !grids))) return set_err(PSIVOX_E_ARG, "bad batch arguments");
    if (nmesh == 0) return 0;
    psivox_opts o{};
    o.refine_tolerance = 1.0;
    if (opts) o = *opts;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    hipStream_t s;
    int dev;
    int rc = prepare(g_ws, &o, &s, &dev);
    if (rc) return rc;
    return run_batched(g_ws, nmesh, meshes, grids, &o, stats, s, dev);
}

int psivox_mass_adjoint(const psivox_mesh* mesh, const psivox_grid* grid, const double* dl_dm, double* dl_dpos,
                        const psivox_opts* opts) {
    if (!mesh || !grid) return set_err(PSIVOX_E_ARG, "mesh and grid are required");
    if (!dl_dm || !dl_dpos) return set_err(PSIVOX_E_ARG, "dl_dm and dl_dpos are required");
    psivox_opts o{};
    o.refine_tolerance = 1.0;
    if (opts) o = *opts;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    hipStream_t s;
    int dev;
    int rc = prepare(g_ws, &o, &s, &dev);
    if (rc) return rc;
    const double* df[6] = {dl_dm, nullptr, nullptr, nullptr, nullptr, nullptr};
    return run_adjoint(g_ws, mesh, grid, df, dl_dpos, nullptr, &o, s, dev);
}

int psivox_adjoint(const psivox_mesh* mesh, const psivox_grid* grid, const double* const* dl_dfields, double* dl_dpos,
                   double* dl_dvel, const psivox_opts* opts) {
    if (!mesh || !grid || !dl_dfields) return set_err(PSIVOX_E_ARG, "mesh, grid and dl_dfields are required");
    if (!dl_dpos && !dl_dvel) return set_err(PSIVOX_E_ARG, "dl_dpos or dl_dvel is required");
    psivox_opts o{};
    o.refine_tolerance = 1.0;
    if (opts) o = *opts;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    hipStream_t s;
    int dev;
    int rc = prepare(g_ws, &o, &s, &dev);
    if (rc) return rc;
    return run_adjoint(g_ws, mesh, grid, dl_dfields, dl_dpos, dl_dvel, &o, s, dev);
}

int psivox_elem_index_build(const psivox_mesh* mesh, const double window[2][3], void* stream,
                            psivox_elem_index** out) {
    if (!out || !window) return set_err(PSIVOX_E_ARG, "window and out are required");
    *out = nullptr;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    hipStream_t s;
    int dev;
    int rc = index_stream(stream, &s, &dev);
    if (rc) return rc;
    psivox_elem_index* X = new (std::nothrow) psivox_elem_index;
    if (!X) return set_err(PSIVOX_E_NOMEM, "out of host memory");
    if ((rc = index_build(mesh, window, s, dev, X))) {
        psivox_elem_index_free(X);
        return rc;
    }
    *out = X;
    return 0;
}

void psivox_elem_index_free(psivox_elem_index* idx) {
    if (!idx) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (idx->dev >= 0) (void)hipSetDevice(idx->dev);
    DevBuf* bufs[] = {&idx->pos, &idx->vel, &idx->mass, &idx->conn, &idx->box, &idx->valid, &idx->flags,
                      &idx->cand, &idx->nsel, &idx->tmp, &idx->counts, &idx->offs, &idx->tets, &idx->out};
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    (void)hipSetDevice(cur);
    delete idx;
}

int psivox_elem_index_query(psivox_elem_index* idx, const double box[2][3], int32_t* out, int64_t cap, int64_t* n,
                            void* stream) {
    if (!idx || !box || !n || cap < 0 || (cap > 0 && !out)) return set_err(PSIVOX_E_ARG, "bad query arguments");
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    hipStream_t s;
    int dev;
    int rc = index_stream(stream, &s, &dev);
    if (rc) return rc;
    if (dev != idx->dev && idx->nelem > 0)
        return set_err(PSIVOX_E_ARG, "the index lives on device %d, the current device is %d", idx->dev, dev);
    QueryBox q;
    for (int a = 0; a < 3; ++a) {
        q.lo[a] = box[0][a];
        q.hi[a] = box[1][a];
    }
    int64_t nc = 0;
    if ((rc = index_select(idx, q, s, &nc))) return rc;
    const int64_t k = nc < cap ? nc : cap;
    if (k > 0) {
        HIPCHK(hipMemcpyAsync(out, idx->cand.p, (size_t)k * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    *n = nc;
    return 0;
}

int psivox_sample_vdf(psivox_elem_index* idx, const double point[3], double refine_tolerance,
                      int32_t refine_max_lvl, double* rho, double* vel, int64_t cap, int64_t* n, void* stream) {
    if (!idx || !point || !n || cap < 0) return set_err(PSIVOX_E_ARG, "bad sample arguments");
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    hipStream_t s;
    int dev;
    int rc = index_stream(stream, &s, &dev);
    if (rc) return rc;
    if (dev != idx->dev && idx->nelem > 0)
        return set_err(PSIVOX_E_ARG, "the index lives on device %d, the current device is %d", idx->dev, dev);
    return sample_vdf(idx, point, refine_tolerance, refine_max_lvl, rho, vel, cap, n, s);
}

}  // extern "C"
