/*
 * FP64 atomic-add throughput of the deposit pattern (a measurement program,
 * not part of the library).
 *
 * The column kernel deposits each fragment as a row of nacc FP64 adds into
 * one cell's consecutive slots of the accumulator (nacc = 10 for m, m v,
 * m vv), and a wavefront issues the staged rows slot by slot: one atomic
 * instruction covers 64 / nacc rows of different cells.  This program
 * times that pattern alone, on an accumulator of 256^3 cells x nacc slots,
 * with one-wavefront workgroups as the column kernel:
 *   random  - every row at a random cell;
 *   column  - each row position steps through consecutive cells, as a
 *             lane walks its column;
 *   store   - the random pattern with plain stores instead of atomics.
 * nacc = 1 and 8 bound the rate by requests: a row of 8 aligned doubles
 * is one 64-byte request, a row of 10 two.
 *
 *   ./pypsi_amd/atomic_rate      (one JSON line per case)
 */
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr uint32_t NCELL = 256u * 256u * 256u;
constexpr int NBLOCK = 2048; /* 8 one-wavefront workgroups per CU */
constexpr int ITERS = 4096;

__device__ uint32_t mix(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

enum { RANDOM = 0, COLUMN = 1, STORE = 2 };

template <int MODE>
__global__ __launch_bounds__(64) void k_rows(double* acc, int nacc) {
    const int lane = (int)threadIdx.x;
    const int nrow = 64 / nacc, r = lane / nacc, s = lane - r * nacc;
    if (r >= nrow) return;
    for (int it = 0; it < ITERS; ++it) {
        const uint32_t cell = MODE == COLUMN
                                  ? (mix(blockIdx.x * 64u + (uint32_t)r) + (uint32_t)it) % NCELL
                                  : mix((blockIdx.x * 1315423911u) ^ ((uint32_t)it * 2654435761u) ^ (uint32_t)r) % NCELL;
        double* p = acc + (size_t)cell * (size_t)nacc + (size_t)s;
        if (MODE == STORE)
            *p = 1.0;
        else
            unsafeAtomicAdd(p, 1.0);
    }
}

#define CHK(x)                                                                      \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
            return 1;                                                               \
        }                                                                           \
    } while (0)

}  // namespace

int main() {
    const char* names[3] = {"random", "column", "store"};
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    for (int nacc : {10, 8, 1}) {
        double* acc = nullptr;
        const size_t bytes = (size_t)NCELL * (size_t)nacc * sizeof(double);
        CHK(hipMalloc(&acc, bytes));
        CHK(hipMemset(acc, 0, bytes));
        for (int mode = RANDOM; mode <= STORE; ++mode) {
            float best = 0.f;
            for (int rep = 0; rep < 3; ++rep) {
                CHK(hipEventRecord(e0, nullptr));
                if (mode == RANDOM)
                    k_rows<RANDOM><<<NBLOCK, 64>>>(acc, nacc);
                else if (mode == COLUMN)
                    k_rows<COLUMN><<<NBLOCK, 64>>>(acc, nacc);
                else
                    k_rows<STORE><<<NBLOCK, 64>>>(acc, nacc);
                CHK(hipGetLastError());
                CHK(hipEventRecord(e1, nullptr));
                CHK(hipEventSynchronize(e1));
                float ms = 0.f;
                CHK(hipEventElapsedTime(&ms, e0, e1));
                if (rep == 0 || ms < best) best = ms;
            }
            const double ops = (double)NBLOCK * ITERS * (double)((64 / nacc) * nacc);
            printf("{\"nacc\": %d, \"pattern\": \"%s\", \"ops\": %.0f, \"ms\": %.3f, \"gops_per_s\": %.1f}\n", nacc,
                   names[mode], ops, best, ops / best * 1e-6);
        }
        CHK(hipFree(acc));
    }
    return 0;
}
