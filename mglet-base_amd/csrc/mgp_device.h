// mgp_device.h -- device helpers shared by the kernel sources.
#pragma once
#include <hip/hip_runtime.h>

#include "mgp_internal.h"

namespace mgp {

__device__ __forceinline__ void unpack(uint32_t c, int &i, int &j, int &k) {
    i = c & 1023;
    j = (c >> 10) & 1023;
    k = c >> 20;
}

// Face f of block nb (the neighbour's own dimensions place its faces;
// only the tangential dimensions match across a CON face).
template <typename T>
__device__ __forceinline__ const T *nbr_face(const T *base,
                                             const Block *blocks, int nb,
                                             int f) {
    if (nb < 0) return nullptr;
    const Block &N = blocks[nb];
    return base + N.fc + face_base(f, N.ni, N.nj, N.nk);
}

// Orders the LDS accesses of one wavefront: lanes of a wave see each
// other's earlier LDS writes without a hardware barrier, the fences only
// keep the compiler from moving memory operations across this point.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace mgp
