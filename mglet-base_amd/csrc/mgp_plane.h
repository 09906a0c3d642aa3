// mgp_plane.h -- device helpers and the plan engine shared by the plane
// kernels (mgp_sweeps.hip, mgp_fused.hip): one wave64 owns one block and
// walks the items of a host-built plan over the block's i+j+k hyperplanes.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "mgp_device.h"
#include "mgp_launch.h"

namespace mgp {

constexpr int kWave = 64;

// Loads at a 32-bit element offset from a uniform base: the byte offset
// stays an unsigned 32-bit value, so the load takes the base from SGPRs.
template <typename T>
__device__ __forceinline__ T ld(const T *base, int s) {
    return *(const T *)((const char *)base + (uint32_t)s * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st(T *base, int s, T v) {
    *(T *)((char *)base + (uint32_t)s * (uint32_t)sizeof(T)) = v;
}

// Cell word of the plane kernels (per shape and WO cell, built on the
// host): i | j << 6 | k << 12 | flags << 18 with flags bit 0..5 = (i == 0,
// i == NI-1, j == 0, j == NJ-1, k == 0, k == NK-1).
__device__ __forceinline__ void cell_ijk(uint32_t w, int &i, int &j, int &k) {
    i = w & 63;
    j = (w >> 6) & 63;
    k = (w >> 12) & 63;
}
__device__ __forceinline__ uint32_t cell_flags(uint32_t w) { return w >> 18; }

// Load through a pointer read from LDS: marks it as global memory (a
// generic pointer would become a flat load, which the compiler can only
// wait for by draining every outstanding load).
template <typename T>
__device__ __forceinline__ T gld(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}

// The wave's index in its workgroup, as a uniform value: the plan item
// table offset by it is a uniform address, so the item descriptors load
// through the scalar cache instead of a vector register ring.
__device__ __forceinline__ int wave_index() {
    return (int)(__builtin_amdgcn_readfirstlane(threadIdx.x) / kWave);
}

// Index of a per-block wave count in the plan tables (kPlanWaves).
template <int NW>
constexpr int plan_nw() {
    static_assert(NW == 1 || NW == 2 || NW == 4 || NW == 8, "no plans for this wave count");
    return NW == 1 ? 0 : (NW == 2 ? 1 : (NW == 4 ? 2 : 3));
}

// HILF window ring of the kernels that read its 7-point stencil: plane q
// lives in slot (q + 1) mod NS.  With one wave per block three slots
// suffice when the plan stores each window entry of plane p+2 (slot of
// plane p-1) only after the item that computes the last cell of plane p
// reading plane p-1 at that position; the FP64 kernels use them (one more
// resident wave per CU: FP64 forward sweep 5.89 -> 5.54 ms on 32768 blocks
// of 24^3).  The FP32 kernels keep four slots and the plain entry order,
// which their window loads coalesce better (the scheduled order cost them
// 3 %, more than the extra waves gained).  Several waves per block keep
// four slots.
template <typename T, int NW>
constexpr int win_slots() { return NW == 1 && sizeof(T) == 8 ? 3 : 4; }
template <int NS>
__device__ __forceinline__ int win_slot(int q) {
    // q >= -4: an unsigned remainder by a constant (a multiply-high), which
    // the kernels' four uses per step share
    return NS == 4 ? ((q + 1) & 3) : (int)((unsigned)(q + 1 + 3 * NS) % (unsigned)NS);
}

// Barrier of the waves of a block over LDS only: orders LDS accesses
// across it without waiting for the global loads in flight (a full
// __syncthreads would drain the prefetch pipeline at every plane).
__device__ __forceinline__ void block_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Per-wave table of entry sources, by entry kind.  Absent sources point at
// the zero buffer, so every entry load is valid and yields 0 there.
template <typename T>
struct SrcTab {
    const T *h[8];        // window (target 0) values
    const T *r[8];        // recurrence buffer (target 1) values
    const uint8_t *b[8];  // window BP
};

// r - a b.  FP64 keeps the reference's separate multiply and subtract (the
// build compiles without contraction, for bitwise parity); the FP32 sweeps
// of the mixed-precision form, which the FP64 defect correction makes
// converge to the FP64 solution and which are not held to bitwise parity,
// use one fused multiply-add.
__device__ __forceinline__ double fms(double r, double a, double b) { return r - a * b; }
__device__ __forceinline__ float fms(float r, float a, float b) { return __builtin_fmaf(-a, b, r); }

// The 7-point stencil of HILF (and BP) at one cell, read from the window
// planes p-1, p, p+1, then res = -(A phi) from it, src/flow/
// laplacephi_mod.F90:149-162 (laplace_cell below in two steps, so a kernel
// can issue the reads of several cells before it waits for any).
template <typename T, typename A, bool MASKED>
struct Stencil {
    A pw, pe, ps, pn, pb, pt, pc;
    A bw, be, bs, bn, bb, bt, bc;
    bool bp;   // with BP (a block with blocked cells)
    __device__ __forceinline__ void read(const T *Hm, const T *H0, const T *Hp,
                                         const uint8_t *Bm, const uint8_t *B0,
                                         const uint8_t *Bp, int c0, int HW, bool withbp) {
        pw = (A)Hm[c0 - HW]; pe = (A)Hp[c0 + HW];
        ps = (A)Hm[c0 - 1]; pn = (A)Hp[c0 + 1];
        pb = (A)Hm[c0]; pt = (A)Hp[c0]; pc = (A)H0[c0];
        bp = MASKED && withbp;
        if (bp) {
            bc = (A)B0[c0];
            bw = (A)Bm[c0 - HW]; be = (A)Bp[c0 + HW];
            bs = (A)Bm[c0 - 1]; bn = (A)Bp[c0 + 1];
            bb = (A)Bm[c0]; bt = (A)Bp[c0];
        }
    }
    __device__ __forceinline__ A eval(A aw, A ae, A as, A an, A ab, A at, A ap) const;
};

// res = -(A phi) at one cell, src/flow/laplacephi_mod.F90:149-162, with
// phi and BP read from the window planes p-1, p, p+1.
template <typename T, typename A, bool MASKED>
__device__ __forceinline__ A laplace_cell(const T *Hm, const T *H0,
                                          const T *Hp, const uint8_t *Bm,
                                          const uint8_t *B0,
                                          const uint8_t *Bp, int c0, int HW,
                                          A aw, A ae, A as, A an, A ab, A at,
                                          A ap) {
    const A pw = (A)Hm[c0 - HW], pe = (A)Hp[c0 + HW];
    const A psv = (A)Hm[c0 - 1], pn = (A)Hp[c0 + 1];
    const A pb = (A)Hm[c0], pt = (A)Hp[c0], pc = (A)H0[c0];
    A r;
    if (MASKED) {
        const A bc = (A)B0[c0];
        const A bw = (A)Bm[c0 - HW], be = (A)Bp[c0 + HW];
        const A bs = (A)Bm[c0 - 1], bn = (A)Bp[c0 + 1];
        const A bb = (A)Bm[c0], bt = (A)Bp[c0];
        r = -(aw * pw * bw * bc);
        r = r - ae * pe * bc * be;
        r = r - as * psv * bs * bc;
        r = r - an * pn * bc * bn;
        r = r - ab * pb * bb * bc;
        r = r - at * pt * bc * bt;
    } else {
        r = -(aw * pw);
        r = fms(r, ae, pe);
        r = fms(r, as, psv);
        r = fms(r, an, pn);
        r = fms(r, ab, pb);
        r = fms(r, at, pt);
    }
    r = fms(r, ap, pc);
    return r;
}

template <typename T, typename A, bool MASKED>
__device__ __forceinline__ A Stencil<T, A, MASKED>::eval(A aw, A ae, A as, A an, A ab, A at,
                                                         A ap) const {
    A r;
    if (MASKED && bp) {
        r = -(aw * pw * bw * bc);
        r = r - ae * pe * bc * be;
        r = r - as * ps * bs * bc;
        r = r - an * pn * bc * bn;
        r = r - ab * pb * bb * bc;
        r = r - at * pt * bc * bt;
    } else {
        r = -(aw * pw);
        r = fms(r, ae, pe);
        r = fms(r, as, ps);
        r = fms(r, an, pn);
        r = fms(r, ab, pb);
        r = fms(r, at, pt);
    }
    r = fms(r, ap, pc);
    return r;
}

// Length of the LDS rows of 1-D data: the largest block dimension, even
// (so that FP32 rows keep the FP64 rows behind them 8-byte aligned).
__host__ __device__ inline int row_len(const Topo &t) {
    const int m = t.max_ni > t.max_nj ? t.max_ni : t.max_nj;
    return ((m > t.max_nk ? m : t.max_nk) + 1) & ~1;
}

// 1-D coefficients of a block in LDS as pairs per direction,
// [x: (w, e)][y: (s, n)][z: (b, t)] with nr pairs each, so that a cell reads
// its six coefficients with three paired LDS loads.
template <typename C>
__device__ __forceinline__ void coef_rows(C *c, const C *aw, const C *ae,
                                          const C *as, const C *an,
                                          const C *ab, const C *at,
                                          const Block &B, int nr) {
    for (int e = threadIdx.x; e < 6 * nr; e += blockDim.x) {
        const int r = e / nr, q = e % nr;
        const int n = r < 2 ? B.ni : (r < 4 ? B.nj : B.nk);
        if (q >= n) continue;
        const C *src = r == 0 ? aw + B.o1x : r == 1 ? ae + B.o1x
                     : r == 2 ? as + B.o1y : r == 3 ? an + B.o1y
                     : r == 4 ? ab + B.o1z : at + B.o1z;
        c[2 * (nr * (r >> 1) + q) + (r & 1)] = src[q + 2];
    }
}

template <typename C>
struct alignas(2 * sizeof(C)) CoefPair {
    C lo, hi;
};
// The six coefficients of cell (i, j, k) from the pairs of coef_rows.
template <typename C>
struct Coef6 {
    C w, e, s, n, b, t;
    __device__ __forceinline__ Coef6(const C *c, int nr, int i, int j, int k) {
        const CoefPair<C> x = ((const CoefPair<C> *)c)[i];
        const CoefPair<C> y = ((const CoefPair<C> *)c)[nr + j];
        const CoefPair<C> z = ((const CoefPair<C> *)c)[2 * nr + k];
        w = x.lo; e = x.hi; s = y.lo; n = y.hi; b = z.lo; t = z.hi;
    }
};

// Separable part of GSAP in LDS: [x y z][nr] of 2 / (d(i-1) d(i)).
__device__ __forceinline__ void ap_rows(double *c, const double *ax,
                                        const double *ay, const double *az,
                                        const Block &B, int nr) {
    for (int e = threadIdx.x; e < 3 * nr; e += blockDim.x) {
        const int r = e / nr, q = e % nr;
        const int n = r == 0 ? B.ni : (r == 1 ? B.nj : B.nk);
        if (q >= n) continue;
        const double *src = r == 0 ? ax + B.o1x : r == 1 ? ay + B.o1y : az + B.o1z;
        c[e] = src[q + 2];
    }
}

// The block's wall fold-in terms, copied to registers before a plane loop:
// read through the Block reference inside the loop they would be reloaded
// at every step (the compiler cannot prove them unaliased by the kernel's
// stores), and each reload would wait for every load in flight.
struct WallFold {
    uint32_t mask;
    double w[6];
    __device__ __forceinline__ void load(const Block &B) {
        mask = (uint32_t)B.wallmask;
#pragma unroll
        for (int f = 0; f < 6; ++f) w[f] = B.wall[f];
    }
};

// GSAP of a cell of a block without blocked cells (BP = 1 on the block
// and its ghost layer), in the order of src/ib/noib_mod.F90:154-244: the
// separable part, the six a_nb (1 - BP BP_nb) terms (all +0 here, exact),
// then the wall fold-ins of the boundary cells (a_nb BP BP_nb = a_nb).
// Equal bit for bit to the stored GSAP.
__device__ __forceinline__ double ap_cell(const double *ar, int nr, int i, int j,
                                          int k, uint32_t fl, const WallFold &B) {
    double a = -ar[i];
    a = a - ar[nr + j];
    a = a - ar[2 * nr + k];
    const uint32_t w = fl & B.mask;
    if (w) {
        if (w & 1) a = a + B.w[0];
        if (w & 2) a = a + B.w[1];
        if (w & 4) a = a + B.w[2];
        if (w & 8) a = a + B.w[3];
        if (w & 16) a = a + B.w[4];
        if (w & 32) a = a + B.w[5];
    }
    return a;
}

// Window sources of HILF for the 6 faces: the neighbour's boundary plane
// (CON), the block's own PAR ghosts from bound_pressure, or none.
template <typename T>
__device__ __forceinline__ const T *window_face(const T *hface, const T *hpar,
                                                const Block *blocks,
                                                const Block &B, int f) {
    const T *g = nbr_face(hface, blocks, B.nbr[f], f ^ 1);
    if (!g && ((B.par >> f) & 1))
        g = hpar + B.fc + face_base(f, B.ni, B.nj, B.nk);
    return g;
}

// Fills the source table: lanes 0-7 window values, 8-15 recurrence
// values, 16-23 window BP.
template <typename T>
__device__ __forceinline__ void fill_tab(SrcTab<T> *tab, const Topo &t,
                                         const Block &B, const T *hface,
                                         const T *hpar, const T *hilf,
                                         const T *rsrc, int rmask, Mask m) {
    const T *zero = (const T *)t.zero;
    const int l = threadIdx.x;
    if (l < 8) {
        const T *g = nullptr;
        if (l < 6 && hface) g = window_face(hface, hpar, t.blocks, B, l);
        if (l == 6) g = hilf + B.wo;
        tab->h[l] = g ? g : zero;
    } else if (l < 16) {
        const int f = l - 8;
        const T *g = nullptr;
        if (f < 6 && rsrc && ((rmask >> f) & 1))
            g = nbr_face(rsrc, t.blocks, B.nbr[f], f ^ 1);
        tab->r[f] = g ? g : zero;
    } else if (l < 24) {
        const int f = l - 16;
        const uint8_t *g = nullptr;
        if (f < 6 && m.bpf) g = m.bpf + B.fc + face_base(f, B.ni, B.nj, B.nk);
        if (f == 6 && m.bp) g = m.bp + B.wo;
        tab->b[f] = g ? g : (const uint8_t *)t.zero;
    }
}

// Row l of an item of n rows, lanes past the end repeating its last row
// (row 0 when n = 0): min(l, max(n, 1) - 1), one VALU instruction with the
// bound in a scalar register instead of a compare and two selects.
__device__ __forceinline__ int clamp_row(int l, int n) {
    const uint32_t lim = (uint32_t)__builtin_amdgcn_readfirstlane(n > 0 ? n - 1 : 0);
    return (int)min((uint32_t)l, lim);
}

// Fields of a face-store entry (face_entry, mgp_internal.h).
__device__ __forceinline__ int face_src(uint32_t f) { return f & 0xfff; }
__device__ __forceinline__ int face_dst(uint32_t f) { return (f >> 12) & 0x3fff; }
__device__ __forceinline__ bool face_hi(uint32_t f) { return (f >> 26) & 1; }

// The plan engine: runs the items of one block through functor K.  An
// item covers up to 64 M cells (M per lane) and 64 M entries.  K provides
// Data, load(Data&, s), kFar with Far, far_load(Far&, s) and merge(Data&,
// const Far&), compute(const Data&, s, p, l) (l = the cell's row
// in the item, lane + 64 m), store(entry, value, bp, p), and, with
// K::kFaces, face_get(entry, p) and face_put(entry, value): the boundary
// values the item computed, read back from LDS by the lanes of its
// face-store entries at the end of the item and stored to the face planes,
// once each, at the end of the next item, so the store does not wait for
// the read (a per-cell store behind six flag tests costs about as much as
// the rest of a backward sweep).
//
// Pipeline (W = 2D items per loop trip, every ring slot named statically,
// so no register holding a load in flight is ever copied -- a copy would
// make the wave wait for that load):
//   * descriptor of item q: vector load at issue(q - 2D), ring iv[q mod W];
//   * entry words of item q: loaded at issue(q - D), ring ew[q mod W];
//   * data and entry values of item q: loaded at issue(q), ring q mod D,
//     consumed by compute(q), which runs D items later;
//   * with K::kFar, the fields of K::Far (the ones HBM serves, when the
//     rest come from L2): loaded at issue(q - D), ring q mod W, merged into
//     item q's Data (K::merge) just before compute(q), so they get twice
//     the prefetch distance without the registers of a deeper ring for
//     every field;
//   * the LDS source pointers of item q's entries are looked up in
//     issue(q), just before its entry loads (held through a compute they
//     would cost 2 M registers).
// All loads go through the in-order vector memory counter, so each item
// waits only for its own loads.  The loop starts W items early with no-op
// items instead of a prologue, so it is entered with no load outstanding
// and the compiler's wait counts at the loop head are those of the steady
// state.  Indices outside the plan are clamped to its closing no-op item.
template <int D, int M, int NW, bool EBP, typename T, class K>
__device__ __forceinline__ void run_plan(K &k, const uint2 *__restrict__ items,
                                         int nitem,
                                         const uint32_t *__restrict__ ents,
                                         const SrcTab<T> *tab,
                                         const uint32_t *__restrict__ fents = nullptr) {
    constexpr int W = 2 * D;
    const int lane = threadIdx.x & (kWave - 1);
    const int last = nitem - 1;
    typename K::Data cd[D][M];
    uint32_t ew[W][M];
    T ev[D][M];
    uint8_t eb[D][M];
    uint2 iv[W];
    uint32_t ix[D], iy[D];
    uint32_t fw[D];   // face-store entries of the items in flight
    uint32_t fwp = kFaceNone;   // ... of the previous item, and its value
    T fvp = T(0);
    const T *pn[M];
    const uint8_t *bn[M];
    typename K::Far fd[K::kFar ? W : 1][M];
    if constexpr (K::kFar) {
#pragma unroll
        for (int r = 0; r < W; ++r)
#pragma unroll
            for (int m = 0; m < M; ++m) fd[r][m] = typename K::Far{};
    }
    // this wave's items: every NW-th of the plan, from its own first one
    auto ldi = [&](int q) -> uint2 { return ld(items, ((q < 0 || q > last) ? last : q) * NW); };
    auto lookup = [&](int m, uint32_t e) {
        const int kind = (e >> 12) & 7;
        pn[m] = (e & 0x8000u) ? tab->r[kind] : tab->h[kind];
        if (EBP) bn[m] = tab->b[kind];
    };
    // issue(q), q = q0 + w + D: r = q mod W, u = q mod D
    auto issue = [&](int u, int r, int q) {
        const uint32_t x = __builtin_amdgcn_readfirstlane(iv[r].x);
        const uint32_t y = __builtin_amdgcn_readfirstlane(iv[r].y);
        ix[u] = x;
        iy[u] = y;
        const int cpos = x & 0xffff, cc = (x >> 16) & 0xff;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int l = lane + 64 * m;
            k.load(cd[u][m], cpos + clamp_row(l, cc));
        }
#pragma unroll
        for (int m = 0; m < M; ++m) {
            lookup(m, ew[r][m]);
            const int off = (int)(ew[r][m] >> 16);
            ev[u][m] = gld(pn[m] + off);
            if (EBP) eb[u][m] = gld(bn[m] + off);
        }
        if constexpr (K::kFaces) {
            const int qc = (q < 0 || q > last) ? last : q;
            fw[u] = ld(fents, (qc * NW) * kFaceSlots + lane);
        }
        const int r2 = (r + D) % W;
        const uint32_t x2 = __builtin_amdgcn_readfirstlane(iv[r2].x);
        const uint32_t y2 = __builtin_amdgcn_readfirstlane(iv[r2].y);
        const int ec2 = x2 >> 24, ep2 = y2 & 0x1fffff;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int l = lane + 64 * m;
            ew[r2][m] = ld(ents, ep2 + clamp_row(l, ec2));
        }
        if constexpr (K::kFar) {   // item q + D
            const int cp2 = x2 & 0xffff, cc2 = (x2 >> 16) & 0xff;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int l = lane + 64 * m;
                k.far_load(fd[r2 % (K::kFar ? W : 1)][m],
                           cp2 + clamp_row(l, cc2));
            }
        }
        iv[r] = ldi(q + W);
    };
#pragma unroll
    for (int r = 0; r < W; ++r) {
        uint2 v = ldi(r < D ? r : -1);
        v.x = __builtin_amdgcn_readfirstlane(v.x);  // completes the load here
        v.y = __builtin_amdgcn_readfirstlane(v.y);
        iv[r] = v;
#pragma unroll
        for (int m = 0; m < M; ++m) ew[r][m] = 0u;  // entry word 0: a valid address
    }
#pragma unroll
    for (int u = 0; u < D; ++u) {
        ix[u] = 0u;
        iy[u] = 8u << 21;
        fw[u] = kFaceNone;
    }
    for (int q0 = -W; q0 < nitem; q0 += W) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const int u = w % D;
            const uint32_t x = ix[u], y = iy[u];
            const int cpos = x & 0xffff, cc = (x >> 16) & 0xff, ec = x >> 24;
            const int p = (int)(y >> 21) - 8;
            if constexpr (K::kFar) {
#pragma unroll
                for (int m = 0; m < M; ++m) k.merge(cd[u][m], fd[w % (K::kFar ? W : 1)][m]);
            }
            bool split = false;
            if constexpr (K::kSplit && M > 1) {
                // every row holds cells: all LDS reads of the item first,
                // then the arithmetic (stores of lanes past the chunk off)
                split = cc > 64 * (M - 1);
                if (split) {
                    typename K::G g[M];
#pragma unroll
                    for (int m = 0; m < M; ++m) g[m] = k.gather(cd[u][m], cpos + lane + 64 * m, p);
#pragma unroll
                    for (int m = 0; m < M; ++m)
                        k.finish(cd[u][m], g[m], cpos + lane + 64 * m, p, lane + 64 * m,
                                 lane + 64 * m < cc);
                }
            }
            if (!split) {
#pragma unroll
                for (int m = 0; m < M; ++m)
                    if (lane + 64 * m < cc) k.compute(cd[u][m], cpos + lane + 64 * m, p, lane + 64 * m);
            }
#pragma unroll
            for (int m = 0; m < M; ++m)
                if (lane + 64 * m < ec)
                    k.store(ew[w][m], ev[u][m], EBP ? eb[u][m] : (uint8_t)0, p);
            if constexpr (K::kFaces) {
                if (fwp != kFaceNone) k.face_put(fwp, fvp);
                fwp = fw[u];
                if (fwp != kFaceNone) fvp = k.face_get(fwp, p);
            }
            static_assert(kFaceSlots == kWave, "one face-store entry per lane");
            wave_sync();
            // several waves per block: the next step reads what all waves
            // wrote in this one (the next item was issued D items ago)
            const bool step_ends = NW > 1 && (iy[(w + 1) % D] >> 21) != (y >> 21);
            issue(u, (w + D) % W, q0 + w + D);
            if (step_ends) block_sync_lds();
        }
    }
}

}  // namespace mgp
