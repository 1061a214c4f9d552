/*
 * psivox_fast.h -- binned FP32 voxel-local voxelization (work in progress:
 * until it lands, the automatic kernel choice is the FP64 plain kernel and
 * an explicit request for the fast kernel is refused).
 */
#ifndef PSIVOX_FAST_H
#define PSIVOX_FAST_H

#include "psivox_geom.h"

#define PSIVOX_KERNEL_AUTO_CHOICE PSIVOX_KERNEL_PLAIN

static inline int launch_fast(const double*, int64_t, const pvx::GridInfo&, const pvx::MeshBox&, double*, unsigned*,
                              unsigned long long*, hipStream_t) {
    return (int)hipErrorNotSupported;
}

#endif
