/*
 * psivox.hip -- MI355X (gfx950) voxelizer behind PSI.voxels() and its C ABI
 * (include/psivox.h).
 *
 * Device pipeline for one call (all on one HIP stream):
 *   1. k_to_raw     grid fields (reference layout) -> additive raw moments
 *   2. k_expand     element -> wrapped, refined tets (count/scan/emit for
 *                   refine_max_lvl > 0, fixed 6-slot output otherwise)
 *   3. k_vox_plain  one thread per tet: FP64 clip + sweep + atomic deposit
 *      k_vox_fast   binned FP32 voxel-local path (psivox_fast.h)
 *   4. k_finalize   raw moments -> mass, means, covariances in place
 * Fields travel host<->device only when the caller passes host pointers.
 */
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/psivox.h"
#include "psivox_adjoint.h"
#include "psivox_fast.h"
#include "psivox_geom.h"
#include "psivox_plain.h"
#include "psivox_refine.h"
#include "psivox_moments.h"

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

/* existing grid contents -> raw moments (merge into what is there, as the
 * reference's in-place online update does) */
__global__ void k_to_raw(GridInfo g, FieldPtrs f, double* __restrict__ acc, int64_t nvox) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nvox) return;
    cell_to_raw(g, f, acc, c);
}

/* raw moments -> reference layout; cells that received nothing keep their
 * previous contents */
__global__ void k_finalize(GridInfo g, const double* __restrict__ acc, FieldPtrs f, int64_t nvox) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nvox) return;
    cell_finalize(g, acc, f, c);
}

struct MeshDev {
    const double* pos;
    const double* vel;
    const double* mass;
    const int32_t* conn;
    int vpe;
};

/* refine_max_lvl == 0: fixed slots (6 per element, 1 for simplices);
 * elements outside the window leave invalid records */
__global__ void k_expand_fixed(MeshDev m, GridInfo g, MeshBox mb, int64_t e0, int64_t ne, double* __restrict__ tets) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne) return;
    const int per = (m.vpe == 4) ? 1 : 6;
    double P[27][3], U[27][3];
    const int64_t e = e0 + t;
    for (int v = 0; v < m.vpe; ++v) {
        int64_t id = m.conn[e * m.vpe + v];
        for (int a = 0; a < 3; ++a) {
            P[v][a] = m.pos[3 * id + a];
            U[v][a] = m.vel ? m.vel[3 * id + a] : 0.0;
        }
    }
    double* out = tets + (int64_t)TREC * per * t;
    if (!wrap_and_test(P, m.vpe, g, mb)) {
        for (int s = 0; s < per; ++s) out[TREC * s + 25] = 0.0;
        return;
    }
    const double mass = m.mass[e];
    if (m.vpe == 4) {
        for (int v = 0; v < 4; ++v)
            for (int a = 0; a < 3; ++a) {
                out[3 * v + a] = P[v][a];
                out[12 + 3 * v + a] = U[v][a];
            }
        out[24] = mass;
        out[25] = 1.0;
        return;
    }
    if (m.vpe == 8) {
        double p27[27][3], u27[27][3];
        lin27((const double (*)[3])P, p27);
        lin27((const double (*)[3])U, u27);
        emit6(out, (const double (*)[3])p27, (const double (*)[3])u27, mass);
    } else {
        emit6(out, (const double (*)[3])P, (const double (*)[3])U, mass);
    }
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
                              const int64_t* __restrict__ offsets, int64_t base, double* __restrict__ tets) {
    int64_t e = e0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    RefineScratch S;
    expand_element(m.pos, m.vel, m.mass, m.conn, m.vpe, e, g, mb, maxlvl, tol2, S,
                   tets + (int64_t)TREC * (offsets[e] - base));
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
        atomicAdd(out, ta);
        atomicAdd(out + 1, tb);
    }
}

/* one thread per tet record, FP64 throughout */
__global__ __launch_bounds__(256) void k_vox_plain(const double* __restrict__ tets, int64_t ntet, GridInfo g,
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

/* column tasks per tet record (0 for invalid records); a tet the column
 * path cannot take flags the FP64 fallback */
__global__ __launch_bounds__(256) void k_col_count(const double* __restrict__ tets, int64_t ntet, GridInfo g,
                                                   MeshBox mb, int64_t* __restrict__ cnt, unsigned* status,
                                                   unsigned long long* counters) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int nc = 0;
    if (t < ntet) {
        const double* r = tets + (int64_t)TREC * t;
        int64_t n = 0;
        if (r[25] != 0.0) {
            n = tet_columns(r, t, g, mb, nullptr, &nc);
            if (n < 0) {
                atomicOr(status, ST_FALLBACK);
                n = 0;
            }
        }
        cnt[t] = n;
    }
    block_count(0, nc, counters);
}

/* write the column tasks of tets [t0, t1) at colstart[t] - base */
__global__ __launch_bounds__(256) void k_col_emit(const double* __restrict__ tets, int64_t t0, int64_t t1, GridInfo g,
                                                  MeshBox mb, const int64_t* __restrict__ colstart, int64_t base,
                                                  uint64_t* __restrict__ tasks) {
    const int64_t t = t0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1) return;
    const double* r = tets + (int64_t)TREC * t;
    if (r[25] == 0.0 || colstart[t + 1] == colstart[t]) return;
    tet_columns(r, t, g, mb, tasks + (colstart[t] - base), nullptr);
}

/* per-task outputs for the residual pass: FP64 fragment-mass sum, the
 * largest fragment's key and (ORDER > 0) its moments and origin */
struct TaskOut {
    double* sum;
    uint64_t* key;
    float* mom; /* 16 floats per task: m0, s1[3], s2[6], O[3], unused */
};

/* one lane per column task: FP32 column clip in this lane's LDS slice,
 * cell moments, FP64 deposit */
template <int ORDER>
__global__ __launch_bounds__(FBLOCK, 2) void k_vox_cols(const double* __restrict__ tets,
                                                        const uint64_t* __restrict__ tasks, int64_t ntask, GridInfo g,
                                                        MeshBox mb, double* __restrict__ acc, TaskOut out,
                                                        unsigned* status, unsigned long long* counters) {
    __shared__ float sq[3][NVF][FBLOCK];
    __shared__ uint32_t snb[NVF][FBLOCK];
    const int lane = threadIdx.x;
    LPoly P{&sq[0][0][lane], &sq[1][0][lane], &sq[2][0][lane], &snb[0][lane], 0};
    const int64_t c = (int64_t)blockIdx.x * FBLOCK + lane;
    int64_t nf = 0;
    if (c < ntask) {
        const uint64_t w = tasks[c];
        const double* r = tets + (int64_t)TREC * (int64_t)(w & 0xffffffffull);
        AtomicSink sink{acc};
        double sum;
        BestFrag B;
        if (!column_task<ORDER>(sink, g, mb, r, w, P, nf, sum, B)) atomicOr(status, ST_FALLBACK);
        out.sum[c] = sum;
        out.key[c] = B.kl >= 0 ? best_key(B.M.m0, B.kl) : 0ull;
        if (ORDER > 0 && B.kl >= 0) {
            float4* m = (float4*)(out.mom + 16 * c);
            m[0] = make_float4(B.M.m0, B.M.s1[0], B.M.s1[1], B.M.s1[2]);
            m[1] = make_float4(B.M.s2[0], B.M.s2[1], B.M.s2[2], B.M.s2[3]);
            m[2] = make_float4(B.M.s2[4], B.M.s2[5], B.O[0], B.O[1]);
            m[3] = make_float4(B.O[2], 0.f, 0.f, 0.f);
        }
    }
    block_count(nf, 0, counters);
}

/* one thread per tet of [t0, t1): reduce its tasks' outputs and deposit the
 * mass residual into its largest fragment */
template <int ORDER>
__global__ __launch_bounds__(256) void k_residual(const double* __restrict__ tets, int64_t t0, int64_t t1,
                                                  GridInfo g, MeshBox mb, const int64_t* __restrict__ colstart,
                                                  int64_t base, const uint64_t* __restrict__ tasks, TaskOut out,
                                                  double* __restrict__ acc) {
    const int64_t t = t0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1) return;
    const int64_t c0 = colstart[t] - base, c1 = colstart[t + 1] - base;
    if (c1 <= c0) return;
    double sum = 0.0;
    uint64_t kbest = 0;
    int64_t cbest = -1;
    for (int64_t c = c0; c < c1; ++c) {
        sum += out.sum[c];
        const uint64_t k = out.key[c];
        if (k > kbest) {
            kbest = k;
            cbest = c;
        }
    }
    if (cbest < 0) return;
    BestFrag B;
    B.kl = (int)(kbest & 0xffffffffull);
    const uint32_t mbits = (uint32_t)(kbest >> 32);
    memcpy(&B.M.m0, &mbits, 4);
    if (ORDER > 0) {
        const float4* m = (const float4*)(out.mom + 16 * cbest);
        const float4 a = m[0], b = m[1], c = m[2], d = m[3];
        B.M.s1[0] = a.y;
        B.M.s1[1] = a.z;
        B.M.s1[2] = a.w;
        B.M.s2[0] = b.x;
        B.M.s2[1] = b.y;
        B.M.s2[2] = b.z;
        B.M.s2[3] = b.w;
        B.M.s2[4] = c.x;
        B.M.s2[5] = c.y;
        B.O[0] = c.z;
        B.O[1] = c.w;
        B.O[2] = d.x;
    }
    AtomicSink sink{acc};
    tet_residual<ORDER>(sink, g, mb, tets + (int64_t)TREC * t, tasks[cbest], B, sum);
}

static int launch_cols(const double* tets, const uint64_t* tasks, int64_t ntask, const GridInfo& g, const MeshBox& mb,
                       double* acc, TaskOut out, unsigned* status, unsigned long long* counters, hipStream_t s) {
    const unsigned nb = (unsigned)((ntask + FBLOCK - 1) / FBLOCK);
    if (g.order == 0)
        k_vox_cols<0><<<nb, FBLOCK, 0, s>>>(tets, tasks, ntask, g, mb, acc, out, status, counters);
    else if (g.order == 1)
        k_vox_cols<1><<<nb, FBLOCK, 0, s>>>(tets, tasks, ntask, g, mb, acc, out, status, counters);
    else
        k_vox_cols<2><<<nb, FBLOCK, 0, s>>>(tets, tasks, ntask, g, mb, acc, out, status, counters);
    return (int)hipGetLastError();
}

static int launch_residual(const double* tets, int64_t t0, int64_t t1, const GridInfo& g, const MeshBox& mb,
                           const int64_t* colstart, int64_t base, const uint64_t* tasks, TaskOut out, double* acc,
                           hipStream_t s) {
    const unsigned nb = (unsigned)((t1 - t0 + 255) / 256);
    if (g.order == 0)
        k_residual<0><<<nb, 256, 0, s>>>(tets, t0, t1, g, mb, colstart, base, tasks, out, acc);
    else if (g.order == 1)
        k_residual<1><<<nb, 256, 0, s>>>(tets, t0, t1, g, mb, colstart, base, tasks, out, acc);
    else
        k_residual<2><<<nb, 256, 0, s>>>(tets, t0, t1, g, mb, colstart, base, tasks, out, acc);
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
    const double* w; /* dL/dm */
    MeshDev md;
    int64_t e0;      /* first element of the batch */
    int per;         /* tets per element */
    double* dpos;    /* dL/dpos, npart x 3 */
};

/* one lane per (tet, copy, axis, plane): FP64 section integrals scattered
 * to the tet's particles */
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
    adj_plane_task<FBLOCK>(g, mb, tets + (int64_t)TREC * t, w, A.w, R, o);
    const int64_t e = A.e0 + t / A.per;
    const int tl = (int)(t % A.per);
    for (int v = 0; v < 4; ++v) {
        if (o[v] == 0.0) continue;
        const int64_t id = A.md.conn[e * A.md.vpe + tet_node(A.md.vpe, tl, v)];
        unsafeAtomicAdd(A.dpos + 3 * id + ax, o[v]);
    }
}

/* ------------------------------------------------------------------ */
/* host side                                                           */
/* ------------------------------------------------------------------ */

namespace {

/* grow-only device buffers reused across calls on this thread's device */
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    int dev = -1;
    int get(size_t bytes, int device) {
        if (p && n >= bytes && dev == device) return 0;
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
    DevBuf pos, vel, mass, conn, fields[6], acc, tets, counts, offs, scan_tmp, counters, colcnt, colstart, tasks,
        colsum, colkey, colmom, ptets, adj_w, adj_g;
    hipStream_t stream = nullptr;
    int stream_dev = -1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
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

int upload(DevBuf& b, const void* src, size_t bytes, int dev, hipStream_t s, const void** out) {
    int rc = b.get(bytes, dev);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
    *out = b.p;
    return 0;
}

inline unsigned grid_blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

/* the mesh arrays as device pointers (uploaded when they are on the host) */
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
        return 0;
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
    return 0;
}

/*
 * Column tasks of tet records [0, nt): count, scan, and for each chunk of
 * tets whose tasks fit the task buffer emit them and call
 * fn(a0, a1, colstart, base, ntask_chunk).  Returns PSIVOX_E_* or 0; *fell
 * is set (and nothing run) when a tet cannot take the column path.
 */
template <class Fn>
int for_task_chunks(Workspace& W, const double* tets, int64_t nt, const GridInfo& g, const MeshBox& mb,
                    unsigned* status, unsigned long long* counters, hipStream_t s, int dev, bool* fell, Fn fn) {
    int rc;
    *fell = false;
    if ((rc = W.colcnt.get((size_t)(nt + 1) * sizeof(int64_t), dev))) return rc;
    if ((rc = W.colstart.get((size_t)(nt + 1) * sizeof(int64_t), dev))) return rc;
    int64_t* colcnt = (int64_t*)W.colcnt.p;
    int64_t* colstart = (int64_t*)W.colstart.p;
    HIPCHK(hipMemsetAsync(colcnt + nt, 0, sizeof(int64_t), s));
    k_col_count<<<grid_blocks(nt, 256), 256, 0, s>>>(tets, nt, g, mb, colcnt, status, counters);
    HIPCHK(hipGetLastError());
    size_t tb = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, colcnt, colstart, nt + 1, s));
    if ((rc = W.scan_tmp.get(tb, dev))) return rc;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(W.scan_tmp.p, tb, colcnt, colstart, nt + 1, s));
    int64_t ntask = 0;
    unsigned hs = 0;
    HIPCHK(hipMemcpyAsync(&ntask, colstart + nt, sizeof ntask, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&hs, status, sizeof hs, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hs & ST_FALLBACK) {
        *fell = true;
        return 0;
    }
    if (ntask == 0) return 0;
    /* tet sub-ranges whose tasks fit the task buffers */
    const int64_t cap = (int64_t)1 << 28;
    std::vector<int64_t> cuts{0, nt};
    std::vector<int64_t> hcs;
    if (ntask > cap) {
        hcs.resize(nt + 1);
        HIPCHK(hipMemcpy(hcs.data(), colstart, (nt + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        cuts.assign(1, 0);
        while (cuts.back() < nt) {
            const int64_t a0 = cuts.back();
            int64_t a1 = std::upper_bound(hcs.begin() + a0 + 1, hcs.end(), hcs[a0] + cap) - hcs.begin() - 1;
            if (a1 <= a0) a1 = a0 + 1; /* a single tet has fewer than TASK_MAXPERTET < cap tasks */
            cuts.push_back(a1);
        }
    }
    const int64_t tcap = std::min(ntask, cap);
    if ((rc = W.tasks.get((size_t)tcap * sizeof(uint64_t), dev))) return rc;
    uint64_t* tasks = (uint64_t*)W.tasks.p;
    for (size_t q = 0; q + 1 < cuts.size(); ++q) {
        const int64_t a0 = cuts[q], a1 = cuts[q + 1];
        const int64_t base = hcs.empty() ? 0 : hcs[a0];
        const int64_t nsub = hcs.empty() ? ntask : hcs[a1] - hcs[a0];
        if (nsub == 0) continue;
        k_col_emit<<<grid_blocks(a1 - a0, 256), 256, 0, s>>>(tets, a0, a1, g, mb, colstart, base, tasks);
        HIPCHK(hipGetLastError());
        if ((rc = fn(a0, a1, (const int64_t*)colstart, base, (const uint64_t*)tasks, nsub))) return rc;
    }
    return 0;
}

/* One (mesh, grid) voxelization on stream s with workspace W. */
int run_one(Workspace& W, const psivox_mesh* mesh, psivox_grid* grid, const psivox_opts* o, psivox_stats* st,
            hipStream_t s, int dev) {
    auto t_start = std::chrono::steady_clock::now();
    int rc = check_mesh(mesh);
    if (rc) return rc;
    GridInfo g;
    rc = fill_grid_info(grid, o->xv_true_covariance, g);
    if (rc) return rc;
    if (o->mode == PSIVOX_MODE_ANNIHILATION)
        return set_err(PSIVOX_E_MODE, "mode 'annihilation' is not supported by the GPU build");
    if (o->mode != PSIVOX_MODE_DENSITY) return set_err(PSIVOX_E_ARG, "invalid mode %d", o->mode);
    int kernel = o->kernel == PSIVOX_KERNEL_AUTO ? PSIVOX_KERNEL_FAST : o->kernel;
    if (kernel != PSIVOX_KERNEL_PLAIN && kernel != PSIVOX_KERNEL_FAST)
        return set_err(PSIVOX_E_ARG, "invalid kernel %d", o->kernel);

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

    /* grid fields on the device */
    double* fd[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int f = 0; f < 6; ++f) {
        if (!grid->fields[f]) continue;
        if (grid->on_device) {
            fd[f] = grid->fields[f];
        } else {
            const void* p;
            if ((rc = upload(W.fields[f], grid->fields[f], (size_t)nvox * FIELD_WIDTH[f] * sizeof(double), dev, s, &p)))
                return rc;
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

    k_to_raw<<<grid_blocks(nvox, 256), 256, 0, s>>>(g, fptrs, acc, nvox);
    HIPCHK(hipGetLastError());

    int maxlvl = o->refine_max_lvl < 0 ? 0 : (o->refine_max_lvl > MAXLVL ? MAXLVL : o->refine_max_lvl);
    if (vpe == 4) maxlvl = 0;
    const double dxmin = std::min(g.d[0], std::min(g.d[1], g.d[2]));
    const double tol = o->refine_tolerance * dxmin;
    const double tol2 = tol * tol;

    int64_t batch = o->batch_elements > 0 ? o->batch_elements : (int64_t)1 << 22;
    const int64_t ne = mesh->nelem;
    int64_t ntets_total = 0;
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
                k_expand_fixed<<<grid_blocks(e1 - e0, 128), 128, 0, s>>>(md, g, mb, e0, e1 - e0, tets);
            }
            HIPCHK(hipGetLastError());
            if (kernel == PSIVOX_KERNEL_FAST && !(mb.periodic && mb.full_window)) {
                k_targets<<<grid_blocks(nt, 64), 64, 0, s>>>(tets, nt, g, mb);
                HIPCHK(hipGetLastError());
            }
            if (kernel == PSIVOX_KERNEL_PLAIN) {
                HIPCHK(hipEventRecord(W.ev0, s));
                k_vox_plain<<<grid_blocks(nt, 256), 256, 0, s>>>(tets, nt, g, mb, acc, status, counters);
                HIPCHK(hipGetLastError());
            } else {
                /* column tasks, then the per-tet residual pass */
                bool fell = false;
                rc = for_task_chunks(
                    W, tets, nt, g, mb, status, counters, s, dev, &fell,
                    [&](int64_t a0, int64_t a1, const int64_t* colstart, int64_t base, const uint64_t* tasks,
                        int64_t nsub) -> int {
                        int rc2;
                        if ((rc2 = W.colsum.get((size_t)nsub * sizeof(double), dev))) return rc2;
                        if ((rc2 = W.colkey.get((size_t)nsub * sizeof(uint64_t), dev))) return rc2;
                        if (g.order > 0 && (rc2 = W.colmom.get((size_t)nsub * 16 * sizeof(float), dev))) return rc2;
                        TaskOut out{(double*)W.colsum.p, (uint64_t*)W.colkey.p, (float*)W.colmom.p};
                        HIPCHK(hipEventRecord(W.ev0, s));
                        rc2 = launch_cols(tets, tasks, nsub, g, mb, acc, out, status, counters, s);
                        if (rc2)
                            return set_err(PSIVOX_E_HIP, "column kernel launch failed: %s",
                                           hipGetErrorString((hipError_t)rc2));
                        HIPCHK(hipEventRecord(W.ev1, s));
                        HIPCHK(hipEventSynchronize(W.ev1));
                        float ms = 0.f;
                        HIPCHK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
                        kms += ms;
                        rc2 = launch_residual(tets, a0, a1, g, mb, colstart, base, tasks, out, acc, s);
                        if (rc2)
                            return set_err(PSIVOX_E_HIP, "residual kernel launch failed: %s",
                                           hipGetErrorString((hipError_t)rc2));
                        return 0;
                    });
                if (rc) return rc;
                (void)fell; /* the status check below restarts with the FP64 kernel */
            }
            if (kernel == PSIVOX_KERNEL_PLAIN) {
                HIPCHK(hipEventRecord(W.ev1, s));
                HIPCHK(hipEventSynchronize(W.ev1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
                kms += ms;
            }
            ntets_total += nt;
            if (kernel == PSIVOX_KERNEL_FAST) {
                unsigned hs = 0;
                HIPCHK(hipMemcpyAsync(&hs, status, sizeof hs, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                if (hs & ST_FALLBACK) {
                    /* a tet exceeded the FP32 path's vertex budget: redo the
                     * whole call with the FP64 kernel from the untouched fields */
                    kernel = PSIVOX_KERNEL_PLAIN;
                    fell_back = 1;
                    HIPCHK(hipMemsetAsync(counters, 0, 64, s));
                    k_to_raw<<<grid_blocks(nvox, 256), 256, 0, s>>>(g, fptrs, acc, nvox);
                    HIPCHK(hipGetLastError());
                    e0 = 0;
                    kms = 0.0;
                    ntets_total = 0;
                    continue;
                }
            }
        }
        e0 = e1;
        if (o->progress && o->progress(o->progress_ctx, e0, ne)) {
            /* leave what was deposited so far, as the reference does on Ctrl-C */
            k_finalize<<<grid_blocks(nvox, 256), 256, 0, s>>>(g, acc, fptrs, nvox);
            for (int f = 0; f < 6; ++f)
                if (grid->fields[f] && !grid->on_device)
                    (void)hipMemcpyAsync(grid->fields[f], fd[f], (size_t)nvox * FIELD_WIDTH[f] * sizeof(double),
                                         hipMemcpyDeviceToHost, s);
            (void)hipStreamSynchronize(s);
            return set_err(PSIVOX_E_CANCELLED, "cancelled after %lld of %lld elements", (long long)e0, (long long)ne);
        }
    }

    k_finalize<<<grid_blocks(nvox, 256), 256, 0, s>>>(g, acc, fptrs, nvox);
    HIPCHK(hipGetLastError());
    for (int f = 0; f < 6; ++f)
        if (grid->fields[f] && !grid->on_device)
            HIPCHK(hipMemcpyAsync(grid->fields[f], fd[f], (size_t)nvox * FIELD_WIDTH[f] * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
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
    }
    return 0;
}

/* dL/dpos += adjoint of the mass deposit (refine level 0) */
int run_adjoint(Workspace& W, const psivox_mesh* mesh, const psivox_grid* grid, const double* dl_dm, double* dl_dpos,
                const psivox_opts* o, hipStream_t s, int dev) {
    int rc = check_mesh(mesh);
    if (rc) return rc;
    if (!dl_dm || !dl_dpos) return set_err(PSIVOX_E_ARG, "dl_dm and dl_dpos are required");
    if (o->refine_max_lvl > 0 && mesh->elemtype != 4)
        return set_err(PSIVOX_E_MODE, "the mass adjoint supports refine_max_lvl = 0 only");
    GridInfo g;
    for (int a = 0; a < 3; ++a) {
        if (grid->n[a] <= 0) return set_err(PSIVOX_E_ARG, "grid n[%d] = %d must be positive", a, grid->n[a]);
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
    const double* w = dl_dm;
    if (!grid->on_device) {
        const void* p;
        if ((rc = upload(W.adj_w, dl_dm, (size_t)nvox * sizeof(double), dev, s, &p))) return rc;
        w = (const double*)p;
    }
    double* dpos = dl_dpos;
    const size_t gbytes = (size_t)mesh->npart * 3 * sizeof(double);
    if (!mesh->on_device && mesh->npart > 0) {
        const void* p;
        if ((rc = upload(W.adj_g, dl_dpos, gbytes, dev, s, &p))) return rc;
        dpos = (double*)p;
    }
    if ((rc = W.counters.get(64, dev))) return rc;
    unsigned long long* counters = (unsigned long long*)W.counters.p;
    unsigned* status = (unsigned*)(counters + 4);
    HIPCHK(hipMemsetAsync(counters, 0, 64, s));
    const int64_t ne = mesh->nelem;
    const int64_t batch = o->batch_elements > 0 ? o->batch_elements : (int64_t)1 << 22;
    for (int64_t e0 = 0; e0 < ne; e0 += batch) {
        const int64_t e1 = std::min(ne, e0 + batch);
        const int64_t nt = (e1 - e0) * per;
        if ((rc = W.tets.get((size_t)nt * TREC * sizeof(double), dev))) return rc;
        double* tets = (double*)W.tets.p;
        k_expand_fixed<<<grid_blocks(e1 - e0, 128), 128, 0, s>>>(md, g, mb, e0, e1 - e0, tets);
        HIPCHK(hipGetLastError());
        /* tasks: count, scan, emit (chunks of tets whose tasks fit) */
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
        uint64_t* tasks = (uint64_t*)W.tasks.p;
        AdjArgs A;
        A.w = w;
        A.md = md;
        A.e0 = e0;
        A.per = per;
        A.dpos = dpos;
        for (int64_t a0 = 0; a0 < nt;) {
            int64_t a1 = std::upper_bound(hs.begin() + a0 + 1, hs.end(), hs[a0] + cap) - hs.begin() - 1;
            if (a1 <= a0) a1 = a0 + 1;
            const int64_t nsub = hs[a1] - hs[a0];
            if (nsub > 0) {
                k_adj_emit<<<grid_blocks(a1 - a0, 256), 256, 0, s>>>(tets, a0, a1, g, mb, start, hs[a0], tasks);
                HIPCHK(hipGetLastError());
                k_adj<<<grid_blocks(nsub, FBLOCK), FBLOCK, 0, s>>>(tets, tasks, nsub, g, mb, A);
                HIPCHK(hipGetLastError());
            }
            a0 = a1;
        }
        if (o->progress && o->progress(o->progress_ctx, e1, ne)) {
            (void)hipStreamSynchronize(s);
            return set_err(PSIVOX_E_CANCELLED, "cancelled after %lld of %lld elements", (long long)e1, (long long)ne);
        }
    }
    if (!mesh->on_device && mesh->npart > 0)
        HIPCHK(hipMemcpyAsync(dl_dpos, dpos, gbytes, hipMemcpyDeviceToHost, s));
    unsigned hst = 0;
    HIPCHK(hipMemcpyAsync(&hst, status, sizeof hst, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hst & ST_BADGEOM) return set_err(PSIVOX_E_OVERFLOW, "mass adjoint: a tet has non-finite vertices");
    return 0;
}

int prepare(Workspace& W, const psivox_opts* o, hipStream_t* s, int* dev) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_err(PSIVOX_E_NODEV, "no GPU visible");
    HIPCHK(hipGetDevice(dev));
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

int psivox_voxels_batched(int32_t nmesh, const psivox_mesh* meshes, psivox_grid* grids, const psivox_opts* opts,
                          psivox_stats* stats) {
    if (nmesh < 0 || (nmesh > 0 && (!meshes || !grids))) return set_err(PSIVOX_E_ARG, "bad batch arguments");
    for (int i = 0; i < nmesh; ++i) {
        int rc = psivox_voxels(&meshes[i], &grids[i], opts, stats ? &stats[i] : nullptr);
        if (rc) return rc;
    }
    return 0;
}

int psivox_mass_adjoint(const psivox_mesh* mesh, const psivox_grid* grid, const double* dl_dm, double* dl_dpos,
                        const psivox_opts* opts) {
    if (!mesh || !grid) return set_err(PSIVOX_E_ARG, "mesh and grid are required");
    psivox_opts o{};
    o.refine_tolerance = 1.0;
    if (opts) o = *opts;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    hipStream_t s;
    int dev;
    int rc = prepare(g_ws, &o, &s, &dev);
    if (rc) return rc;
    return run_adjoint(g_ws, mesh, grid, dl_dm, dl_dpos, &o, s, dev);
}

}  // extern "C"
