// mgp_internal.h -- data structures shared by the host orchestration
// (mgp_api.cpp) and the gfx950 kernels (mgp_kernels.hip).
//
// Device layout ("wavefront order", WO): every library-owned per-cell
// array stores the interior cells of a block contiguously, ordered by
// hyperplane p = i + j + k (0-based interior indices) and, inside a plane,
// by i then j -- the same order as the reference's SIP access index
// `iacc` (src/flow/sip_hyperplane_mod.F90:269-361).  A sweep step over one
// hyperplane therefore reads a contiguous run of every array.  Ghost
// values live in per-block face buffers: face 0/1 (-x/+x) hold (k, j),
// face 2/3 (-y/+y) (k, i), face 4/5 (-z/+z) (j, i), each anti-diagonal-
// major (see face_index).
#pragma once
#include <rocprofiler-sdk-roctx/roctx.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mgp {

// roctx range around a host call (a C-ABI entry point or a kernel launch):
// shown by rocprofv3 --marker-trace, a few tens of ns without a profiler.
struct RoctxRange {
    explicit RoctxRange(const char *name) { roctxRangePushA(name); }
    ~RoctxRange() { roctxRangePop(); }
    RoctxRange(const RoctxRange &) = delete;
    RoctxRange &operator=(const RoctxRange &) = delete;
};
#define MGP_RANGE_CAT2(a, b) a##b
#define MGP_RANGE_CAT(a, b) MGP_RANGE_CAT2(a, b)
#define MGP_RANGE(name) ::mgp::RoctxRange MGP_RANGE_CAT(mgp_range_, __LINE__)(name)

constexpr int kMaxN = 1000;   // interior cells per direction (10-bit packing)
constexpr int kMaxN1 = 64;    // ... supported by the plane kernels (LDS rows)

// One distinct interior block shape (NI, NJ, NK).
struct Shape {
    int ni, nj, nk;
    int np;            // number of hyperplanes NI + NJ + NK - 2
    int ncell;         // NI * NJ * NK
    int64_t cell_off;  // offset into the packed cell table (i | j<<10 | k<<20)
    int64_t pstart_off;// offset into the plane start table (np + 1 entries)
    int64_t slot_off;  // offset into the spatial->slot table (ncell entries)
    int fsize;         // face buffer entries per block: 2(NK NJ + NK NI + NJ NI)
    // per wave count (kPlanWaves), item width (64, 128 cells) and plan
    // kind: offset of the shape's items, and items per wave (a multiple of
    // kPlanPad)
    int plan_item[4][2][5];
    int plan_nitem[4][2][5];
    int plan_ent[5];      // per plan kind: offset of the shape's entries
};

// Plans of the plane kernels (built per shape on the host).  A plan is a
// sequence of items; item n = {x, y} with x = cpos | ccnt << 16 | ecnt <<
// 24 and y = epos | (p + 8) << 21: the compute part handles WO cells
// [cpos, cpos + ccnt) of one hyperplane, the entry part moves entries
// [epos, epos + ecnt) of the plan's entry table into LDS, and p is the step
// the item belongs to (it selects the LDS ring slots).  With NW waves per
// block the items of each step are dealt to the waves in turn, and steps are
// padded with no-op items of the same p so that every wave sees the same
// sequence of steps; item q of wave w is stored at q NW + w.
enum PlanKind { kPlanFwdLap = 0, kPlanFwdZero, kPlanBwd, kPlanOut, kPlanSor, kPlanKinds };
constexpr int kPlanPad = 8;
constexpr int kPlanWaves[4] = {1, 2, 4, 8};
// Entry: LDS index (12 bits), source kind (3 bits: own face 0-5 of the
// ghost, 6 = the block's WO cells), target (bit 15: 0 = window ring slot,
// 1 = recurrence buffer) and the offset into the source (16 bits).
// Face-store entry of the plane kernels (kFaceSlots per plan item, one per
// lane, built on the host): the LDS position of a value the item computed
// (12 bits: in the step's recurrence or window slot, or the item's staging
// row), the value's offset in the block's face region (14 bits) and bit 26
// for an upper face (the forward sweep's second destination); kFaceNone
// where the lane stores nothing.  The plan builder rejects an item with
// more face cells than lanes (none of 24^3 has more than 45).
constexpr int kFaceSlots = 64;
constexpr uint32_t kFaceNone = 0xffffffffu;
__host__ __device__ inline uint32_t face_entry(int src, int dst, int hi) {
    return (uint32_t)src | (uint32_t)dst << 12 | (uint32_t)hi << 26;
}
__host__ __device__ inline uint32_t plan_entry(int lds, int kind, int tgt, int off) {
    return (uint32_t)lds | ((uint32_t)kind << 12) | ((uint32_t)tgt << 15) |
           ((uint32_t)off << 16);
}

// Per-block record (one per grid, in grid-id order).
struct Block {
    int shape;
    int ni, nj, nk;
    int level;
    int flags;         // bit 0: BP == 1 on the block and its ghost layer
    int nbr[6];        // 0-based block index feeding each face ghost, -1 none
    int par;           // bit f: face f is PAR (ghost from bound_pressure)
    int parent;        // 0-based parent block, -1 none
    int pos[3];        // iposition, jposition, kposition in the parent
    int bc[6];         // first BC code per face
    int64_t wo;        // offset of the block's interior cells in WO arrays
    // offset of the block's SIP factors (and transposed factors) in WO
    // order: blocks whose factors are equal bit for bit share one copy
    int64_t fo;
    int64_t sp;        // offset of the block in reference-layout 3-D fields
    int64_t fc;        // offset of the block's face buffers
    int64_t sh;        // offset of the block's ghost-shell buffers
    int64_t o1x, o1y, o1z; // offsets of the block in 1-D fields
    int wallmask;      // bit f: face f is a wall (GSAP fold-in)
    double wall[6];    // raw boundary coefficient of face f (aw(0), ae(ni-1), ...)
};

// Position of face cell (u, v), u in [0, a), v in [0, b), inside one face
// buffer.  Faces are stored anti-diagonal-major: diagonal d = u + v, then v
// ascending.  A sweep step over hyperplane p touches exactly one diagonal
// of each face, so those accesses are contiguous.
__host__ __device__ inline int diag_start(int d, int a, int b) {
    const int m = a < b ? a : b, M = a < b ? b : a;
    if (d <= m) return d * (d + 1) / 2;
    const int base1 = m * (m + 1) / 2;
    if (d <= M) return base1 + (d - m) * m;
    const int base2 = base1 + (M - m) * m;
    return base2 + (d - M) * (a + b - 1) - (M + d - 1) * (d - M) / 2;
}
__host__ __device__ inline int face_index(int u, int v, int a, int b) {
    const int d = u + v;
    const int vmin = d - (a - 1) > 0 ? d - (a - 1) : 0;
    return diag_start(d, a, b) + (v - vmin);
}
// x faces hold (k, j), y faces (k, i), z faces (j, i)
__host__ __device__ inline int fidx_x(int k, int j, int nj, int nk) { return face_index(k, j, nk, nj); }
__host__ __device__ inline int fidx_y(int k, int i, int ni, int nk) { return face_index(k, i, nk, ni); }
__host__ __device__ inline int fidx_z(int j, int i, int ni, int nj) { return face_index(j, i, nj, ni); }

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

// The first ghost layer of a block ("shell": cells (i, j, k) in [-1, n]^3
// with at least one index outside [0, n)), enumerated slab by slab in i.
__host__ __device__ inline int shell_size(int ni, int nj, int nk) {
    return 2 * (nj + 2) * (nk + 2) + ni * (2 * (nk + 2) + 2 * nj);
}
__host__ __device__ inline int shell_index(int i, int j, int k, int ni, int nj,
                                           int nk) {
    const int a = (nj + 2) * (nk + 2);
    if (i < 0) return (j + 1) * (nk + 2) + (k + 1);
    if (i >= ni) return a + (j + 1) * (nk + 2) + (k + 1);
    const int base = 2 * a + i * (2 * (nk + 2) + 2 * nj);
    if (j < 0) return base + (k + 1);
    if (j >= nj) return base + (nk + 2) + (k + 1);
    return base + 2 * (nk + 2) + 2 * j + (k < 0 ? 0 : 1);
}

// One box copy of the ghost-layer exchange (reference layout).
struct CopyTask {
    int64_t src;       // offset of the first source element
    int64_t dst;       // offset of the first destination element
    int ni, nj, nk;    // box extent
    int s_sj, s_si;    // source strides of j and i (k stride is 1)
    int d_sj, d_si;    // destination strides
};

// The exchange by receiving block: one workgroup per receiving block of a
// stage copies all of the block's ghost entries of that stage.  An entry
// is (destination offset in the receiving block, source offset in the
// block of its slot | slot << 24); slot q is the source block of the
// block's q-th box copy.  In pair tables an entry moves two values
// consecutive along k.  Blocks with the same copy pattern share a table.
struct PullBlock {
    int64_t dst;       // offset of the receiving block (reference layout)
    int tab, n;        // first entry of its table, number of entries
    int64_t src[26];   // offset of the source block of each slot
};
constexpr int kPullThreads = 256;
constexpr int kPullSlotBits = 24;   // source offsets below 2^24 (blocks <= 68^3)

}  // namespace mgp
