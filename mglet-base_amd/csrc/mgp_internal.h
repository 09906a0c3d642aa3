// mgp_internal.h -- data structures shared by the host orchestration
// (mgp_api.cpp) and the gfx950 kernels (mgp_kernels.hip).
//
// Device layout ("wavefront order", WO): every library-owned per-cell
// array stores the interior cells of a block contiguously, ordered by
// hyperplane p = i + j + k (0-based interior indices) and, inside a plane,
// by i then j -- the same order as the reference's SIP access index
// `iacc` (src/flow/sip_hyperplane_mod.F90:269-361).  A sweep step over one
// hyperplane therefore reads a contiguous run of every array.  Ghost
// values live in per-block face buffers:
//   face 0/1 (-x/+x) indexed k + NK*j, face 2/3 (-y/+y) k + NK*i,
//   face 4/5 (-z/+z) j + NJ*i.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mgp {

constexpr int kMaxN = 1000;   // interior cells per direction (10-bit packing)

// One distinct interior block shape (NI, NJ, NK).
struct Shape {
    int ni, nj, nk;
    int np;            // number of hyperplanes NI + NJ + NK - 2
    int ncell;         // NI * NJ * NK
    int64_t cell_off;  // offset into the packed cell table (i | j<<10 | k<<20)
    int64_t pstart_off;// offset into the plane start table (np + 1 entries)
    int64_t slot_off;  // offset into the spatial->slot table (ncell entries)
    int fsize;         // face buffer entries per block: 2(NK NJ + NK NI + NJ NI)
};

// Per-block record (one per grid, in grid-id order).
struct Block {
    int shape;
    int ni, nj, nk;
    int level;
    int flags;         // bit 0: BP == 1 on the block and its ghost layer
    int nbr[6];        // 0-based block index feeding each face ghost, -1 none
    int bc[6];         // first BC code per face
    int64_t wo;        // offset of the block's interior cells in WO arrays
    int64_t sp;        // offset of the block in reference-layout 3-D fields
    int64_t fc;        // offset of the block's face buffers
    int64_t o1x, o1y, o1z; // offsets of the block in 1-D fields
};

// Face buffer sub-offsets inside one block's face region.
__host__ __device__ inline int face_base(int f, int ni, int nj, int nk) {
    const int ax = nk * nj, ay = nk * ni, az = nj * ni;
    switch (f) {
    case 0: return 0;
    case 1: return ax;
    case 2: return 2 * ax;
    case 3: return 2 * ax + ay;
    case 4: return 2 * ax + 2 * ay;
    default: return 2 * ax + 2 * ay + az;
    }
}

// One box copy of the ghost-layer exchange (reference layout).
struct CopyTask {
    int64_t src;       // offset of the first source element
    int64_t dst;       // offset of the first destination element
    int ni, nj, nk;    // box extent
    int s_sj, s_si;    // source strides of j and i (k stride is 1)
    int d_sj, d_si;    // destination strides
};

}  // namespace mgp
