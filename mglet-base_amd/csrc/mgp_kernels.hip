// mgp_kernels.hip -- gfx950 kernels of the pressure projection outside the
// inner iteration: coefficient setup (GSAP, SIP factors in hyperplane
// order), divergence, max-norm, exit test, layout conversion, pressure
// correction and the ghost-layer copies.  The plane kernels of the inner
// and outer iterations are in mgp_sweeps.hip.
//
// Per-cell arrays are stored hyperplane-major (WO layout, see
// mgp_internal.h).  All kernels are compiled with -ffp-contract=off: each
// expression is evaluated in the reference's association order without
// FMA, which makes the FP64 path bitwise equal to the CPU oracle.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mgp_device.h"
#include "mgp_launch.h"

namespace mgp {

#define TINY ((double)1.0e-20f)

constexpr int kThreads = 256;

// ---------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T *p, int64_t n, T v) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        p[e] = v;
}

template <typename T>
void launch_fill(T *p, int64_t n, T v, hipStream_t s) {
    MGP_RANGE("mgp:fill");
    if (n <= 0) return;
    int64_t nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    k_fill<T><<<(unsigned)nb, 256, 0, s>>>(p, n, v);
}
void launch_fill_u8(uint8_t *p, int64_t n, uint8_t v, hipStream_t s) {
    MGP_RANGE("mgp:fill_u8");
    launch_fill<uint8_t>(p, n, v, s);
}
template void launch_fill<double>(double *, int64_t, double, hipStream_t);
template void launch_fill<float>(float *, int64_t, float, hipStream_t);
template void launch_fill<int>(int *, int64_t, int, hipStream_t);

template <typename T>
__global__ void k_convert(const double *src, T *dst, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = (T)src[e];
}
template <typename T>
void launch_convert(const double *src, T *dst, int64_t n, hipStream_t s) {
    MGP_RANGE("mgp:convert");
    if (n <= 0) return;
    int64_t nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    k_convert<T><<<(unsigned)nb, 256, 0, s>>>(src, dst, n);
}
template void launch_convert<float>(const double *, float *, int64_t, hipStream_t);
template void launch_convert<double>(const double *, double *, int64_t, hipStream_t);

// BP of cell (i, j, k) of a block (0-based interior indices, one of them
// possibly -1 or N for the ghost layer).
__device__ __forceinline__ double bp_at(const Block &B, const int *slot_of,
                                        Mask m, int i, int j, int k) {
    if (B.flags & 1) return 1.0;
    const int ni = B.ni, nj = B.nj, nk = B.nk;
    const uint8_t *f = m.bpf + B.fc;
    if (i < 0) return f[face_base(0, ni, nj, nk) + fidx_x(k, j, nj, nk)];
    if (i >= ni) return f[face_base(1, ni, nj, nk) + fidx_x(k, j, nj, nk)];
    if (j < 0) return f[face_base(2, ni, nj, nk) + fidx_y(k, i, ni, nk)];
    if (j >= nj) return f[face_base(3, ni, nj, nk) + fidx_y(k, i, ni, nk)];
    if (k < 0) return f[face_base(4, ni, nj, nk) + fidx_z(j, i, ni, nj)];
    if (k >= nk) return f[face_base(5, ni, nj, nk) + fidx_z(j, i, ni, nj)];
    return m.bp[B.wo + slot_of[k + nk * (j + nj * i)]];
}

// Workgroup barrier that orders LDS only.  __syncthreads() also fences
// global memory, which on CDNA waits for every outstanding global load
// (s_waitcnt vmcnt(0)) and would drain the sweeps' prefetches each step;
// all cross-lane traffic inside a sweep goes through LDS.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Work index of workgroup w (linear id) of `total`: the MI355X dispatcher
// hands workgroup w to XCD w mod 8, each XCD with its own L2, so work
// index W runs on one XCD with W - 1 and W + 1 (XCD x takes the x-th
// contiguous range): workgroups whose data share cache lines or rows meet
// in one L2 instead of fetching them once per XCD.
__device__ __forceinline__ int xcd_work(int w, int total) {
    constexpr int kXcd = 8;
    const int x = w % kXcd, i = w / kXcd;
    const int q = total / kXcd, r = total % kXcd;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// ---------------------------------------------------------------------
// GSAP: src/ib/noib_mod.F90:154-244.  One thread per WO cell.
__global__ void k_ap(Topo t, const double *dx_, const double *dy_,
                     const double *dz_, const double *aw_, const double *ae_,
                     const double *as_, const double *an_, const double *ab_,
                     const double *at_, Mask m, double *ap) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int *slot_of = t.slot_of + S.slot_off;
    const uint32_t *cells = t.cells + S.cell_off;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < S.ncell;
         s += gridDim.x * blockDim.x) {
        int i, j, k;
        unpack(cells[s], i, j, k);
        // 1-based reference index r = i + 3; dx(r) is dx_[o1x + r - 1]
        const double *dx = dx_ + B.o1x + 2, *dy = dy_ + B.o1y + 2,
                     *dz = dz_ + B.o1z + 2;
        const double *aw = aw_ + B.o1x + 2, *ae = ae_ + B.o1x + 2;
        const double *as = as_ + B.o1y + 2, *an = an_ + B.o1y + 2;
        const double *ab = ab_ + B.o1z + 2, *at = at_ + B.o1z + 2;
        double a = -(2.0 / (dx[i - 1] * dx[i])) - 2.0 / (dy[j - 1] * dy[j]) -
                   2.0 / (dz[k - 1] * dz[k]);
        const double c = bp_at(B, slot_of, m, i, j, k);
        const double bw = bp_at(B, slot_of, m, i - 1, j, k);
        const double be = bp_at(B, slot_of, m, i + 1, j, k);
        const double bs = bp_at(B, slot_of, m, i, j - 1, k);
        const double bn = bp_at(B, slot_of, m, i, j + 1, k);
        const double bb = bp_at(B, slot_of, m, i, j, k - 1);
        const double bt = bp_at(B, slot_of, m, i, j, k + 1);
        a = a + aw[i] * (1.0 - bw * c);
        a = a + ae[i] * (1.0 - c * be);
        a = a + as[j] * (1.0 - bs * c);
        a = a + an[j] * (1.0 - c * bn);
        a = a + ab[k] * (1.0 - bb * c);
        a = a + at[k] * (1.0 - c * bt);
        const int *bc = B.bc;
        if (i == 0 && (bc[0] == 2 || bc[0] == 5 || bc[0] == 6 || bc[0] == 19))
            a = a + aw[0] * (bw * c);
        if (i == B.ni - 1 && (bc[1] == 2 || bc[1] == 5 || bc[1] == 6))
            a = a + ae[i] * (c * be);
        if (j == 0 && (bc[2] == 2 || bc[2] == 5 || bc[2] == 6 || bc[2] == 19))
            a = a + as[0] * (bs * c);
        if (j == B.nj - 1 && (bc[3] == 2 || bc[3] == 5 || bc[3] == 6))
            a = a + an[j] * (c * bn);
        if (k == 0 && (bc[4] == 2 || bc[4] == 5 || bc[4] == 6 || bc[4] == 19))
            a = a + ab[0] * (bb * c);
        if (k == B.nk - 1 && (bc[5] == 2 || bc[5] == 5 || bc[5] == 6))
            a = a + at[k] * (c * bt);
        ap[B.wo + s] = a;
    }
}

void launch_ap(const Topo &t, const double *dx, const double *dy,
               const double *dz, const double *aw, const double *ae,
               const double *as, const double *an, const double *ab,
               const double *at, Mask m, double *ap, hipStream_t s) {
    MGP_RANGE("mgp:ap");
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    dim3 grid((maxc + kThreads - 1) / kThreads, t.nlist);
    k_ap<<<grid, kThreads, 0, s>>>(t, dx, dy, dz, aw, ae, as, an, ab, at, m,
                                   ap);
}

// ---------------------------------------------------------------------
// SIP incomplete LU in hyperplane order: src/flow/sip_hyperplane_mod.F90:
// 106-257 (identical values to src/flow/sip_classic_mod.F90:17-127).
// The upper factors of plane p-1 are kept per pencil in LDS.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_lu(
    Topo t, const double *ap, const double *aw_, const double *ae_,
    const double *as_, const double *an_, const double *ab_,
    const double *at_, Mask m, double ALFA, T *lw_o, T *ls_o, T *lb_o,
    T *lpr_o, T *ue_o, T *un_o, T *ut_o) {
    extern __shared__ double lds_lu[];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int NI = B.ni, NJ = B.nj;
    const int RW = NJ + 1, RS = (NJ + 1) * (NI + 1);
    double *U = lds_lu;  // [2][3][RS]
    const uint32_t *cells = t.cells + S.cell_off;
    const int *ps = t.pstart + S.pstart_off;
    const int *slot_of = t.slot_of + S.slot_off;
    const double *aw = aw_ + B.o1x + 2, *ae = ae_ + B.o1x + 2;
    const double *as = as_ + B.o1y + 2, *an = an_ + B.o1y + 2;
    const double *ab = ab_ + B.o1z + 2, *at = at_ + B.o1z + 2;
    for (int e = threadIdx.x; e < 6 * RS; e += blockDim.x) U[e] = 0.0;
    lds_barrier();
    for (int p = 0; p < S.np; ++p) {
        const double *pe = U + ((p - 1) & 1) * 3 * RS;
        const double *pn = pe + RS, *pt = pn + RS;
        double *ce = U + (p & 1) * 3 * RS;
        double *cn = ce + RS, *ct = cn + RS;
        const int cnt = ps[p + 1] - ps[p];
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
            const int s = ps[p] + q;
            int i, j, k;
            unpack(cells[s], i, j, k);
            const int c0 = (j + 1) + RW * (i + 1);
            const int cim = c0 - RW, cjm = c0 - 1;
            const double e_im = pe[cim], n_im = pn[cim], t_im = pt[cim];
            const double e_jm = pe[cjm], n_jm = pn[cjm], t_jm = pt[cjm];
            const double e_km = pe[c0], n_km = pn[c0], t_km = pt[c0];
            double bu_m = 1.0, bv_m = 1.0, bw_m = 1.0, bu_c = 1.0, bv_c = 1.0,
                   bw_c = 1.0;
            if (!(B.flags & 1)) {
                const double c = bp_at(B, slot_of, m, i, j, k);
                bu_m = bp_at(B, slot_of, m, i - 1, j, k) * c;
                bv_m = bp_at(B, slot_of, m, i, j - 1, k) * c;
                bw_m = bp_at(B, slot_of, m, i, j, k - 1) * c;
                bu_c = c * bp_at(B, slot_of, m, i + 1, j, k);
                bv_c = c * bp_at(B, slot_of, m, i, j + 1, k);
                bw_c = c * bp_at(B, slot_of, m, i, j, k + 1);
            }
            double lw = aw[i] * bu_m / (1.0 + ALFA * (n_im + t_im));
            double ls = as[j] * bv_m / (1.0 + ALFA * (e_jm + t_jm));
            double lb = ab[k] * bw_m / (1.0 + ALFA * (n_km + e_km));
            const double p1 = ALFA * (lb * e_km + ls * e_jm);
            const double p2 = ALFA * (lb * n_km + lw * n_im);
            const double p3 = ALFA * (lw * t_im + ls * t_jm);
            const double lpr = 1.0 / (ap[B.wo + s] + p1 + p2 + p3 -
                                      lb * t_km - lw * e_im - ls * n_jm +
                                      TINY);
            const double ue = (ae[i] * bu_c - p1) * lpr;
            const double un = (an[j] * bv_c - p2) * lpr;
            const double ut = (at[k] * bw_c - p3) * lpr;
            lw = lw * lpr;
            ls = ls * lpr;
            lb = lb * lpr;
            ce[c0] = ue;
            cn[c0] = un;
            ct[c0] = ut;
            const int64_t o = B.wo + s;
            lw_o[o] = (T)lw;
            ls_o[o] = (T)ls;
            lb_o[o] = (T)lb;
            lpr_o[o] = (T)lpr;
            ue_o[o] = (T)ue;
            un_o[o] = (T)un;
            ut_o[o] = (T)ut;
        }
        lds_barrier();
    }
}

template <typename T>
void launch_lu(const Topo &t, const double *ap, const double *aw,
               const double *ae, const double *as, const double *an,
               const double *ab, const double *at, Mask m, double alfa,
               T *const f[7], hipStream_t s) {
    MGP_RANGE("mgp:lu");
    const size_t lds = 6 * (size_t)(t.max_nj + 1) * (t.max_ni + 1) *
                       sizeof(double);
    k_lu<T><<<t.nlist, kThreads, lds, s>>>(t, ap, aw, ae, as, an, ab, at, m,
                                           alfa, f[0], f[1], f[2], f[3], f[4], f[5],
                                           f[6]);
}
template void launch_lu<double>(const Topo &, const double *, const double *,
                                const double *, const double *, const double *,
                                const double *, const double *, Mask, double,
                                double *const[7], hipStream_t);
template void launch_lu<float>(const Topo &, const double *, const double *,
                               const double *, const double *, const double *,
                               const double *, const double *, Mask, double,
                               float *const[7], hipStream_t);

// ---------------------------------------------------------------------
// Shared factor storage.  The SIP factors of a block depend only on its
// shape, 1-D coefficients, GSAP and BP; blocks whose factors and GSAP come
// out equal bit for bit (every interior block of a uniform grid, for
// instance) keep one copy of both, which the plane kernels then read from
// L2 instead of HBM.  Classes are found by a hash of each block's factor
// and GSAP bits (and BP bytes), confirmed by an exact comparison with the
// class representative.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t fbits(double v) { return (uint64_t)__double_as_longlong(v); }
__device__ __forceinline__ uint64_t fbits(float v) { return (uint64_t)__float_as_uint(v); }

template <typename T>
struct F7 {
    const T *f[7];
    const double *ap;   // GSAP (FP64)
};

// One workgroup per block: order-independent sum of mixed (position,
// factor, bits) words over the block's factors at Block::wo.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fhash(Topo t, F7<T> F, Mask m,
                                                   unsigned long long *out) {
    __shared__ unsigned long long red[kThreads];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const int n = t.shapes[B.shape].ncell;
    uint64_t acc = 0;
    for (int s = threadIdx.x; s < n; s += kThreads) {
#pragma unroll
        for (int q = 0; q < 7; ++q)
            acc += mix64(fbits(F.f[q][B.wo + s]) ^ mix64((uint64_t)s * 9 + q + 1));
        acc += mix64(fbits(F.ap[B.wo + s]) ^ mix64((uint64_t)s * 9 + 8));
        if (m.bp) acc += mix64((uint64_t)m.bp[B.wo + s] ^ mix64((uint64_t)s * 9 + 9));
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = red[0] ^ mix64((uint64_t)B.shape + 0x9e3779b97f4a7c15ull);
}

// bad[b] = 1 when block b's factors (or BP bytes) differ from those of its
// representative rep[b] (both at Block::wo); blocks that represent
// themselves are skipped.
template <typename T>
__global__ void k_fcmp(Topo t, F7<T> F, Mask m, const int *rep, int *bad) {
    const int b = t.blist[blockIdx.y];
    const int r = rep[b];
    if (r == b) return;
    const Block &B = t.blocks[b];
    const Block &R = t.blocks[r];
    const int n = t.shapes[B.shape].ncell;
    bool diff = B.shape != R.shape;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n && !diff;
         s += gridDim.x * blockDim.x) {
#pragma unroll
        for (int q = 0; q < 7; ++q)
            diff = diff || fbits(F.f[q][B.wo + s]) != fbits(F.f[q][R.wo + s]);
        diff = diff || fbits(F.ap[B.wo + s]) != fbits(F.ap[R.wo + s]);
        if (m.bp) diff = diff || m.bp[B.wo + s] != m.bp[R.wo + s];
    }
    if (diff) bad[b] = 1;
}

// Copies each representative's factors and GSAP from Block::wo of the old
// arrays to Block::fo of the new ones (GSAP in FP64, and in the working
// precision when that is FP32).
template <typename T>
__global__ void k_fgather(Topo t, F7<T> F, F7<T> G, float *ap32) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const int n = t.shapes[B.shape].ncell;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += gridDim.x * blockDim.x) {
#pragma unroll
        for (int q = 0; q < 7; ++q) const_cast<T *>(G.f[q])[B.fo + s] = F.f[q][B.wo + s];
        const double a = F.ap[B.wo + s];
        const_cast<double *>(G.ap)[B.fo + s] = a;
        if (ap32) ap32[B.fo + s] = (float)a;
    }
}

template <typename T>
void launch_fhash(const Topo &t, const T *const f[7], const double *ap, Mask m,
                  unsigned long long *out, hipStream_t s) {
    MGP_RANGE("mgp:fhash");
    F7<T> F;
    for (int q = 0; q < 7; ++q) F.f[q] = f[q];
    F.ap = ap;
    k_fhash<T><<<t.nlist, kThreads, 0, s>>>(t, F, m, out);
}
template <typename T>
void launch_fcmp(const Topo &t, const T *const f[7], const double *ap, Mask m,
                 const int *rep, int *bad, hipStream_t s) {
    MGP_RANGE("mgp:fcmp");
    F7<T> F;
    for (int q = 0; q < 7; ++q) F.f[q] = f[q];
    F.ap = ap;
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    dim3 grid((maxc + kThreads - 1) / kThreads, t.nlist);
    k_fcmp<T><<<grid, kThreads, 0, s>>>(t, F, m, rep, bad);
}
template <typename T>
void launch_fgather(const Topo &t, const T *const f[7], const double *ap,
                    T *const g[7], double *gap, float *gap32, hipStream_t s) {
    MGP_RANGE("mgp:fgather");
    F7<T> F, G;
    for (int q = 0; q < 7; ++q) {
        F.f[q] = f[q];
        G.f[q] = g[q];
    }
    F.ap = ap;
    G.ap = gap;
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    dim3 grid((maxc + kThreads - 1) / kThreads, t.nlist);
    k_fgather<T><<<grid, kThreads, 0, s>>>(t, F, G, gap32);
}
#define MGP_INST_FSHARE(T)                                                          \
    template void launch_fhash<T>(const Topo &, const T *const[7], const double *, Mask, \
                                  unsigned long long *, hipStream_t);                 \
    template void launch_fcmp<T>(const Topo &, const T *const[7], const double *, Mask, \
                                 const int *, int *, hipStream_t);                    \
    template void launch_fgather<T>(const Topo &, const T *const[7], const double *,    \
                                    T *const[7], double *, float *, hipStream_t);
MGP_INST_FSHARE(double)
MGP_INST_FSHARE(float)
#undef MGP_INST_FSHARE

// ---------------------------------------------------------------------
// Layout changes between the reference layout and WO go through LDS, one
// workgroup per block: the reference-layout side is read or written in its
// own order (k fastest, coalesced), the WO side in WO order, and LDS holds
// up to kStage WO slots of the block at a time (blocks with more cells take
// several passes).
constexpr int kStageThreads = 1024;

// RHS = fak * div(u, v, w): src/ib/noib_mod.F90:304-354 (noib) and the
// ghostcell variant src/ib/gc_mod.F90:278-347 (+ fak*SDIV, the caller's
// geometry source or zero, then x BP); u, v, w, SDIV and BP in the reference
// layout, RHS in WO.
template <typename T>
__global__ __launch_bounds__(kStageThreads) void k_divcal(
    Topo t, const double *u, const double *v, const double *w,
    const double *rddx_, const double *rddy_, const double *rddz_,
    const double *bp_sp, const double *sdiv_sp, double fak, double *rhs64,
    T *rhs_t) {
    extern __shared__ double st_dv[];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int *slot_of = t.slot_of + S.slot_off;
    const int NJ = B.nj, NK = B.nk, n = S.ncell;
    const int64_t kk = B.nk + 4, jj = B.nj + 4;
    const double *rddx = rddx_ + B.o1x + 2, *rddy = rddy_ + B.o1y + 2,
                 *rddz = rddz_ + B.o1z + 2;
    for (int base = 0; base < n; base += kStage) {
        const int hi = n < base + kStage ? n : base + kStage;
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int sl = slot_of[e];   // e = k + NK (j + NJ i)
            if (sl < base || sl >= hi) continue;
            const int k = e % NK, j = (e / NK) % NJ, i = e / (NK * NJ);
            const int64_t c = B.sp + (k + 2) + (j + 2) * kk + (i + 2) * kk * jj;
            double d = fak * ((u[c] - u[c - kk * jj]) * rddx[i] +
                              (v[c] - v[c - kk]) * rddy[j] +
                              (w[c] - w[c - 1]) * rddz[k]);
            if (bp_sp) {
                d = d + fak * (sdiv_sp ? sdiv_sp[c] : 0.0);
                d = bp_sp[c] * d;
            }
            st_dv[sl - base] = d;
        }
        __syncthreads();
        for (int q = threadIdx.x; q < hi - base; q += blockDim.x) {
            const double d = st_dv[q];
            rhs64[B.wo + base + q] = d;
            if (rhs_t) rhs_t[B.wo + base + q] = (T)d;
        }
        __syncthreads();
    }
}

static size_t stage_bytes(const Topo &t) {
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    return (size_t)(maxc < kStage ? maxc : kStage) * sizeof(double);
}

template <typename T>
void launch_divcal(const Topo &t, const double *u, const double *v,
                   const double *w, const double *rddx, const double *rddy,
                   const double *rddz, const double *bp_sp, const double *sdiv_sp,
                   double fak, double *rhs64, T *rhs_t, hipStream_t s) {
    MGP_RANGE("mgp:divcal");
    k_divcal<T><<<t.nlist, kStageThreads, stage_bytes(t), s>>>(
        t, u, v, w, rddx, rddy, rddz, bp_sp, sdiv_sp, fak, rhs64, rhs_t);
}
template void launch_divcal<double>(const Topo &, const double *,
                                    const double *, const double *,
                                    const double *, const double *,
                                    const double *, const double *,
                                    const double *, double, double *, double *,
                                    hipStream_t);
template void launch_divcal<float>(const Topo &, const double *,
                                   const double *, const double *,
                                   const double *, const double *,
                                   const double *, const double *,
                                   const double *, double, double *, float *,
                                   hipStream_t);

// ---------------------------------------------------------------------
// Block-wide max of non-negative doubles folded into a global maximum.
// Non-negative IEEE doubles order like their bit patterns.
__device__ void block_max_to_global(double m, unsigned long long *g) {
    __shared__ double red[kThreads / 64];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(m, off, 64);
        m = (o > m || o != o) ? o : m;   // a NaN from either side stays
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q)
            r = (red[q] > r || red[q] != red[q]) ? red[q] : r;
        atomicMax(g, (unsigned long long)__double_as_longlong(r));
    }
}

// max |RHS| over interior cells with the level weight 2^(maxlevel-level):
// src/flow/pressuresolver_mod.F90:623-689.
__global__ void k_maxabs(Topo t, const double *rhs64, Outer o) {
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    double m = 0.0;
    for (int s = threadIdx.x; s < S.ncell; s += blockDim.x) {
        const double a = fabs(rhs64[B.wo + s]);
        m = (a > m || a != a) ? a : m;
    }
    m = m * ldexp(1.0, t.maxlevel - B.level);
    block_max_to_global(m, o.gmax);
}

__global__ void k_maxabs_store(Outer o) {
    const double m = __longlong_as_double((long long)*o.gmax);
    *o.maxrhs0 = m;
    *o.maxrhs = m;
    *o.gmax = 0ull;
    *o.ctr = 0;
    *o.converged = 0;
    *o.nonfinite = (m != m || m == INFINITY) ? 1 : 0;
    *o.done = *o.nonfinite;
}

void launch_maxabs(const Topo &t, const double *rhs64, Outer o,
                   hipStream_t s) {
    MGP_RANGE("mgp:maxabs");
    launch_fill<unsigned long long>(o.gmax, 1, 0ull, s);
    k_maxabs<<<t.nlist, kThreads, 0, s>>>(t, rhs64, o);
    k_maxabs_store<<<1, 1, 0, s>>>(o);
}
template void launch_fill<unsigned long long>(unsigned long long *, int64_t,
                                              unsigned long long, hipStream_t);

// max |RHS| of the outer iteration on a multi-level grid, taken after the
// restriction of the RHS to the parents (src/flow/pressuresolver_mod.F90:
// 288-298)
void launch_maxabs_acc(const Topo &t, const double *rhs64, Outer o,
                       hipStream_t s) {
    MGP_RANGE("mgp:maxabs_acc");
    launch_fill<unsigned long long>(o.gmax, 1, 0ull, s);
    k_maxabs<<<t.nlist, kThreads, 0, s>>>(t, rhs64, o);
}

// ---------------------------------------------------------------------
// SOR_RAP = divide0(1, AP): src/flow/sor_mod.F90:18-49,
// src/core/simdfunctions_mod.F90:70-82.
template <typename T>
__global__ void k_rap(const double *ap, T *rap, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const double a = ap[e];
        rap[e] = (T)(a == 0.0 ? 0.0 : 1.0 / a);
    }
}

template <typename T>
void launch_rap(const double *ap, T *rap, int64_t n, hipStream_t s) {
    MGP_RANGE("mgp:rap");
    if (n <= 0) return;
    int64_t nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    k_rap<T><<<(unsigned)nb, 256, 0, s>>>(ap, rap, n);
}
template void launch_rap<double>(const double *, double *, int64_t, hipStream_t);
template void launch_rap<float>(const double *, float *, int64_t, hipStream_t);

// Exit test of the outer loop (src/flow/pressuresolver_mod.F90:274-301).
__global__ void k_check(Outer o, int nouter, int nouter_min, double epcorr,
                        double eps_rel, double prefak, double *hist,
                        int hist_cap) {
    if (*o.done) return;
    const double m = __longlong_as_double((long long)*o.gmax);
    *o.gmax = 0ull;
    const int it = *o.ctr + 1;
    *o.ctr = it;
    *o.maxrhs = m;
    if (hist && it <= hist_cap) hist[it - 1] = m;
    const bool bad = (m != m) || m == INFINITY;
    bool conv = false;
    if (it >= nouter_min)
        conv = eps_rel > 0.0 ? (m <= eps_rel * *o.maxrhs0)
                             : (m / prefak < epcorr);
    if (conv) *o.converged = 1;
    if (bad) *o.nonfinite = 1;
    if (conv || bad || it >= nouter) *o.done = 1;
}

void launch_check(Outer o, int nouter, int nouter_min, double epcorr,
                  double eps_rel, double prefak, double *hist, int hist_cap,
                  hipStream_t s) {
    MGP_RANGE("mgp:check");
    k_check<<<1, 1, 0, s>>>(o, nouter, nouter_min, epcorr, eps_rel, prefak,
                            hist, hist_cap);
}

// ---------------------------------------------------------------------
// WO -> reference layout, through LDS as k_divcal: the interior, with the
// ghosts left as they are (ZG false) or set to zero (ZG true: the block's
// whole region is written in reference order, in whole lines, where the
// block fits one stage).
template <typename T, bool ZG>
__global__ __launch_bounds__(kStageThreads) void k_wo_to_sp(Topo t, const T *src,
                                                            double *dst) {
    extern __shared__ double st_ws[];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int *slot_of = t.slot_of + S.slot_off;
    const int NI = B.ni, NJ = B.nj, NK = B.nk, n = S.ncell;
    const int64_t kk = B.nk + 4, jj = B.nj + 4;
    if (ZG) {
        const int tot = (int)(kk * jj) * (NI + 4);
        const bool one = n <= kStage;
        if (one) {
            for (int q = threadIdx.x; q < n; q += blockDim.x)
                st_ws[q] = (double)src[B.wo + q];
            __syncthreads();
        }
        for (int p = threadIdx.x; p < tot; p += blockDim.x) {
            const int k = p % (int)kk, r = p / (int)kk, j = r % (int)jj, i = r / (int)jj;
            const bool in = k >= 2 && k < NK + 2 && j >= 2 && j < NJ + 2 && i >= 2 && i < NI + 2;
            if (in && !one) continue;   // the passes below
            dst[B.sp + p] = in ? st_ws[slot_of[(k - 2) + NK * ((j - 2) + NJ * (i - 2))]] : 0.0;
        }
        if (one) return;
    }
    for (int base = 0; base < n; base += kStage) {
        const int hi = n < base + kStage ? n : base + kStage;
        for (int q = threadIdx.x; q < hi - base; q += blockDim.x)
            st_ws[q] = (double)src[B.wo + base + q];
        __syncthreads();
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int sl = slot_of[e];
            if (sl < base || sl >= hi) continue;
            const int k = e % NK, j = (e / NK) % NJ, i = e / (NK * NJ);
            dst[B.sp + (k + 2) + (j + 2) * kk + (i + 2) * kk * jj] = st_ws[sl - base];
        }
        __syncthreads();
    }
}
template <typename T>
void launch_wo_to_sp(const Topo &t, const T *src, double *dst,
                     hipStream_t s, bool zero_ghosts) {
    MGP_RANGE("mgp:wo_to_sp");
    if (zero_ghosts)
        k_wo_to_sp<T, true><<<t.nlist, kStageThreads, stage_bytes(t), s>>>(t, src, dst);
    else
        k_wo_to_sp<T, false><<<t.nlist, kStageThreads, stage_bytes(t), s>>>(t, src, dst);
}
template void launch_wo_to_sp<double>(const Topo &, const double *, double *,
                                      hipStream_t, bool);
template void launch_wo_to_sp<float>(const Topo &, const float *, double *,
                                     hipStream_t, bool);
template void launch_wo_to_sp<uint8_t>(const Topo &, const uint8_t *,
                                       double *, hipStream_t, bool);

// Factors back to the reference's storage: hyperplane access order in the
// first ncell entries of the block (src/flow/sip_hyperplane_mod.F90:
// 224-230) or spatial order for the classic solver.
template <typename T>
__global__ void k_to_ref_factors(Topo t, const T *wo, double *ref,
                                 int spatial) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const uint32_t *cells = t.cells + S.cell_off;
    const int64_t kk = B.nk + 4, jj = B.nj + 4;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < S.ncell;
         s += gridDim.x * blockDim.x) {
        int i, j, k;
        unpack(cells[s], i, j, k);
        const int64_t d = spatial ? (k + 2) + (j + 2) * kk + (i + 2) * kk * jj
                                  : (int64_t)s;
        ref[B.sp + d] = (double)wo[B.fo + s];
    }
}
template <typename T>
void launch_to_ref_factors(const Topo &t, const T *wo, double *ref,
                           int spatial, hipStream_t s) {
    MGP_RANGE("mgp:to_ref_factors");
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    dim3 grid((maxc + kThreads - 1) / kThreads, t.nlist);
    k_to_ref_factors<T><<<grid, kThreads, 0, s>>>(t, wo, ref, spatial);
}
template void launch_to_ref_factors<double>(const Topo &, const double *,
                                            double *, int, hipStream_t);
template void launch_to_ref_factors<float>(const Topo &, const float *,
                                           double *, int, hipStream_t);

// ---------------------------------------------------------------------
// Velocity and pressure correction (src/flow/pressuresolver_mod.F90:
// 787-879) on the reference layout.  The threads of a block walk its rows
// (i, j) in 2..ii-1 x 2..jj-1, all kk positions of a row, so every HBM
// line a row touches is used whole: a field is read and written back over
// the full rows it updates anywhere (positions outside the update range
// keep their value), and rows it never updates are not touched.  Element
// indices are split with a float reciprocal and one correction step
// (exact for the row and block sizes here).
__device__ __forceinline__ int div_small(int e, int d, float rd) {
    int q = (int)((float)e * rd);
    const int r = e - q * d;
    if (r < 0) --q;
    else if (r >= d) ++q;
    return q;
}

// The update of one position (i, j, k 1-based, c its offset in the block)
// of a block (sp, ox, oy, oz: its offsets in the 3-D and 1-D fields):
// r = P, U, V, W, each the stored value where that field is not updated
// along k.  corr_load_t reads every operand whatever the position, so a
// wave waits for memory once and takes one path: all of them lie inside
// the block (rows i + 1 and j + 1 exist for i <= ii - 1, j <= jj - 1; the
// k + 1 neighbour is clamped at the end of the row, where W is not
// updated).
struct CorrArgs {
    double *u, *v, *w, *p;
    const double *dp, *rdx, *rdy, *rdz, *bp;
    double rfak;
};

// the operands of one position, and the update from them
struct CorrIn {
    double pv, uv, vv, wv, d0, di, dj, dk, rdx, rdy, rdz, b0, bi, bj, bk;
};

template <bool BP>
__device__ __forceinline__ CorrIn corr_load_t(const CorrArgs &a, int64_t sp, int ox, int oy,
                                              int oz, int i, int j, int k, int64_t c, int jj,
                                              int kk) {
    const int64_t sj = kk, si = (int64_t)kk * jj;
    const int64_t ck = c + (k < kk ? 1 : 0);
    const double *dp = a.dp + sp;
    CorrIn x;
    x.pv = a.p[sp + c];
    x.uv = a.u[sp + c];
    x.vv = a.v[sp + c];
    x.wv = a.w[sp + c];
    x.d0 = dp[c];
    x.di = dp[c + si];
    x.dj = dp[c + sj];
    x.dk = dp[ck];
    x.rdx = a.rdx[ox + i - 1];
    x.rdy = a.rdy[oy + j - 1];
    x.rdz = a.rdz[oz + k - 1];
    if (BP) {
        const double *bp = a.bp + sp;
        x.b0 = bp[c];
        x.bi = bp[c + si];
        x.bj = bp[c + sj];
        x.bk = bp[ck];
    } else {
        x.b0 = x.bi = x.bj = x.bk = 1.0;
    }
    return x;
}

template <bool BP>
__device__ __forceinline__ void corr_eval_t(const CorrArgs &a, const CorrIn &x, int k, int kk,
                                            double r[4]) {
    const double rfak = a.rfak;
    const bool kin = k >= 3 && k <= kk - 2;
    if (BP) {
        r[0] = k >= 2 && k <= kk - 1 ? x.pv + x.d0 * x.b0 : x.pv;
        r[1] = kin ? x.uv + (x.d0 - x.di) * x.b0 * x.bi * x.rdx * rfak : x.uv;
        r[2] = kin ? x.vv + (x.d0 - x.dj) * x.b0 * x.bj * x.rdy * rfak : x.vv;
        r[3] = k >= 2 && k <= kk - 2 ? x.wv + (x.d0 - x.dk) * x.b0 * x.bk * x.rdz * rfak : x.wv;
    } else {
        r[0] = k >= 2 && k <= kk - 1 ? x.pv + x.d0 : x.pv;
        r[1] = kin ? x.uv + (x.d0 - x.di) * x.rdx * rfak : x.uv;
        r[2] = kin ? x.vv + (x.d0 - x.dj) * x.rdy * rfak : x.vv;
        r[3] = k >= 2 && k <= kk - 2 ? x.wv + (x.d0 - x.dk) * x.rdz * rfak : x.wv;
    }
}

// The plain update of one position: each field read only on the rows it
// updates (corr_store's row ranges), the same expressions as corr_eval_t
__device__ __forceinline__ void corr_cell(const CorrArgs &a, int64_t sp, int ox, int oy,
                                          int oz, int i, int j, int k, int64_t c, int ii,
                                          int jj, int kk, double r[4]) {
    const int64_t sj = kk, si = (int64_t)kk * jj;
    const double *u = a.u + sp, *v = a.v + sp, *w = a.w + sp, *p = a.p + sp;
    const double *dp = a.dp + sp;
    const double *bp = a.bp ? a.bp + sp : nullptr;
    const double rfak = a.rfak;
    const bool kin = k >= 3 && k <= kk - 2;
    {   // P: i, j, k in 2..n-1
        const double pv = p[c];
        r[0] = pv;
        if (k >= 2 && k <= kk - 1) r[0] = bp ? pv + dp[c] * bp[c] : pv + dp[c];
    }
    if (i <= ii - 2 && j >= 3 && j <= jj - 2) {   // U rows
        const double uv = u[c];
        r[1] = uv;
        if (kin) {
            const double rdx = a.rdx[ox + i - 1];
            r[1] = bp ? uv + (dp[c] - dp[c + si]) * bp[c] * bp[c + si] * rdx * rfak
                      : uv + (dp[c] - dp[c + si]) * rdx * rfak;
        }
    }
    if (i >= 3 && i <= ii - 2 && j <= jj - 2) {   // V rows
        const double vv = v[c];
        r[2] = vv;
        if (kin) {
            const double rdy = a.rdy[oy + j - 1];
            r[2] = bp ? vv + (dp[c] - dp[c + sj]) * bp[c] * bp[c + sj] * rdy * rfak
                      : vv + (dp[c] - dp[c + sj]) * rdy * rfak;
        }
    }
    if (i >= 3 && i <= ii - 2 && j >= 3 && j <= jj - 2) {   // W rows
        const double wv = w[c];
        r[3] = wv;
        if (k >= 2 && k <= kk - 2) {
            const double rdz = a.rdz[oz + k - 1];
            r[3] = bp ? wv + (dp[c] - dp[c + 1]) * bp[c] * bp[c + 1] * rdz * rfak
                      : wv + (dp[c] - dp[c + 1]) * rdz * rfak;
        }
    }
}

__device__ __forceinline__ void corr_store(const CorrArgs &a, int64_t sp, int i, int j,
                                           int64_t c, int ii, int jj, const double r[4]) {
    a.p[sp + c] = r[0];
    if (i <= ii - 2 && j >= 3 && j <= jj - 2) a.u[sp + c] = r[1];      // U rows
    if (i >= 3 && i <= ii - 2 && j <= jj - 2) a.v[sp + c] = r[2];      // V rows
    if (i >= 3 && i <= ii - 2 && j >= 3 && j <= jj - 2) a.w[sp + c] = r[3];   // W rows
}

__global__ __launch_bounds__(kThreads) void k_mgpcorr(Topo t, CorrArgs a) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const int kk = B.nk + 4, jj = B.nj + 4, ii = B.ni + 4;
    const int nrj = jj - 2, nel = (ii - 2) * nrj * kk;
    const int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= nel) return;
    const int row = div_small(e, kk, 1.0f / (float)kk);
    const int k = e - row * kk + 1;
    const int ri = div_small(row, nrj, 1.0f / (float)nrj);
    const int i = ri + 2, j = row - ri * nrj + 2;
    const int64_t c = (int64_t)(k - 1) + (int64_t)(j - 1) * kk + (int64_t)(i - 1) * kk * jj;
    double r[4];
    corr_cell(a, B.sp, B.o1x, B.o1y, B.o1z, i, j, k, c, ii, jj, kk, r);
    corr_store(a, B.sp, i, j, c, ii, jj, r);
}

// The update along z-chains of blocks (closing exchange of single-level
// grids whose k-faces join blocks of one size in cycles): workgroup
// (x, chain) takes the same whole rows of every block of its chain in
// order, and writes the k-face ghosts of interior rows (k = 0, 1 and
// nk + 2, nk + 3, 0-based) as part of those rows, with the values the
// exchange would copy there.  Step q stores block q; it has computed block
// q + 1 (from loads issued one step earlier, so they arrive during the
// previous step's stores) and issued the loads of block q + 2.  Block q's
// k-high ghosts are block q + 1's results at k = 2, 3 and its k-low ghosts
// block q - 1's at nk, nk + 1, both passed through LDS; the first block's
// k-low ghosts are evaluated from the last block's cells before anything
// is stored, and the last block's k-high ghosts are the first block's
// results.  Every value is the update the exchange would copy, the same
// expressions on the same operands, and every store is part of a whole
// row; the exchange that follows leaves these cells out.  Lanes load and
// store only their own positions (the k-low ghost lanes of the first
// block aside, whose source block is stored last), so the barrier orders
// LDS only.  Lanes beyond the workgroup's rows repeat the work of its
// first lane (same loads, same LDS slots, the same values stored to the
// same places): every lane takes one path with a fixed number of loads
// per step.  Each field is stored only on the rows it updates (storing
// all four on every row, which keeps the store count fixed too, measured
// 4 % slower on 32768 blocks of 24^3: it moves more bytes).
template <bool BP>
__global__ __launch_bounds__(kThreads) void k_mgpcorr_chain(Topo t, const int *chain,
                                                           const int2 *chains, CorrArgs a) {
    constexpr int S = 2 * kCorrRows;
    __shared__ double carry[4][4][S];   // block b's results at nk, nk + 1: [b mod 4]
    __shared__ double nxt[2][4][S];     // block b's results at 2, 3: [b mod 2]
    __shared__ double first[4][S];      // block 0's results at 2, 3
    __shared__ int64_t rsp[kCorrChainMax];   // the chain's block offsets
    __shared__ int rox[kCorrChainMax], roy[kCorrChainMax], roz[kCorrChainMax];
    // row groups of one chain, and then the next chain, on one XCD: the
    // DP rows at i + 1 and the cache lines two neighbouring row groups share
    // come from its L2 (11.10 -> 10.71 ms on 32768 blocks of 24^3)
    const int W = xcd_work(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int bx = W % gridDim.x;
    const int2 ch = chains[W / gridDim.x];
    const Block &B0 = t.blocks[chain[ch.x]];
    const int nk = B0.nk, kk = nk + 4, jj = B0.nj + 4, ii = B0.ni + 4;
    const int nrj = jj - 2, nrow = (ii - 2) * nrj;
    const int R = min(kThreads / kk, kCorrRows);
    if (bx * R >= nrow) return;   // whole workgroup
    int lr = div_small(threadIdx.x, kk, 1.0f / (float)kk);
    int k = threadIdx.x - lr * kk + 1;
    if (lr >= R || bx * R + lr >= nrow) lr = 0, k = 1;   // repeat the first lane
    const int row = bx * R + lr;
    const int ri = div_small(row, nrj, 1.0f / (float)nrj);
    const int i = ri + 2, j = row - ri * nrj + 2;
    const int64_t c = (int64_t)(k - 1) + (int64_t)(j - 1) * kk + (int64_t)(i - 1) * kk * jj;
    const bool inner = i >= 3 && i <= ii - 2 && j >= 3 && j <= jj - 2;
    const bool glo = inner && k <= 2, ghi = inner && k >= kk - 1;
    const int slot = 2 * lr;
    for (int q = threadIdx.x; q < ch.y; q += kThreads) {
        const Block &Bq = t.blocks[chain[ch.x + q]];
        rsp[q] = Bq.sp;
        rox[q] = Bq.o1x;
        roy[q] = Bq.o1y;
        roz[q] = Bq.o1z;
    }
    __syncthreads();
    auto load = [&](int qb, int ke, int64_t ce) {
        return corr_load_t<BP>(a, rsp[qb], rox[qb], roy[qb], roz[qb], i, j, ke, ce, jj, kk);
    };
    auto save = [&](int qb, const double r[4]) {
        if (inner && (k == kk - 3 || k == kk - 2))
            for (int f = 0; f < 4; ++f) carry[qb & 3][f][slot + k - (kk - 3)] = r[f];
        if (inner && (k == 3 || k == 4))
            for (int f = 0; f < 4; ++f) {
                nxt[qb & 1][f][slot + k - 3] = r[f];
                if (qb == 0) first[f][slot + k - 3] = r[f];
            }
    };
    double rc[4], rn[4] = {0.0, 0.0, 0.0, 0.0};
    CorrIn x;
    {   // block 0; its k-low ghost lanes evaluate the last block at k + nk
        const int ke = glo ? k + nk : k;
        const CorrIn x0 = glo ? load(ch.y - 1, ke, c + nk) : load(0, k, c);
        corr_eval_t<BP>(a, x0, ke, kk, rc);
        save(0, rc);
        if (ch.y > 1) x = load(1, k, c);
    }
    for (int q = 0; q < ch.y; ++q) {
        if (q + 1 < ch.y) {
            corr_eval_t<BP>(a, x, k, kk, rn);
            save(q + 1, rn);
            if (q + 2 < ch.y) x = load(q + 2, k, c);
        }
        lds_barrier();
        {
            if (glo && q > 0)
                for (int f = 0; f < 4; ++f) rc[f] = carry[(q - 1) & 3][f][slot + k - 1];
            if (ghi)
                for (int f = 0; f < 4; ++f)
                    rc[f] = q + 1 < ch.y ? nxt[(q + 1) & 1][f][slot + k - (kk - 1)]
                                         : first[f][slot + k - (kk - 1)];
            corr_store(a, rsp[q], i, j, c, ii, jj, rc);
            for (int f = 0; f < 4; ++f) rc[f] = rn[f];
        }
    }
}

void launch_mgpcorr(const Topo &t, double *u, double *v, double *w,
                    double *p, const double *dp, const double *rdx,
                    const double *rdy, const double *rdz, double rfak,
                    const double *bp, const CorrChains *cc, hipStream_t s) {
    MGP_RANGE("mgp:mgpcorr");
    const CorrArgs a{u, v, w, p, dp, rdx, rdy, rdz, bp, rfak};
    if (cc) {
        if (bp)
            k_mgpcorr_chain<true><<<dim3(cc->nx, cc->n), kThreads, 0, s>>>(t, cc->chain,
                                                                           cc->chains, a);
        else
            k_mgpcorr_chain<false><<<dim3(cc->nx, cc->n), kThreads, 0, s>>>(t, cc->chain,
                                                                            cc->chains, a);
    } else {
        const int maxe = (t.max_ni + 2) * (t.max_nj + 2) * (t.max_nk + 4);
        k_mgpcorr<<<dim3((maxe + kThreads - 1) / kThreads, t.nlist), kThreads, 0, s>>>(t, a);
    }
}

// ---------------------------------------------------------------------
// Ghost-layer exchange (connect_self_single, src/core/connect/conn1_mod.
// F90:674-704, all fields at once) by receiving block: the workgroup of a
// block reads its slots' source offsets into LDS and walks the block's
// entry table, which lists the ghost positions in memory order, so
// consecutive lanes write consecutive addresses wherever the ghost region
// is contiguous (rows along k).  Every load of an entry is issued before
// its stores.
struct FieldList {
    double *f[6];
};

template <int NF, int MODE>
__global__ __launch_bounds__(kPullThreads) void k_pull(const PullBlock *pb, const uint2 *ents,
                                                      FieldList fl) {
    __shared__ int64_t src[26];
    // (receivers in chain order spread over the XCDs: mapping consecutive
    // ones to one XCD, as the correction does, ran 5 % slower)
    const PullBlock &B = pb[blockIdx.x];
    if (threadIdx.x < 26) src[threadIdx.x] = B.src[threadIdx.x];
    const int64_t dst = B.dst;
    const int n = B.n;
    const uint2 *e = ents + B.tab;
    __syncthreads();
    constexpr uint32_t kMask = (1u << kPullSlotBits) - 1;
    for (int q = threadIdx.x; q < n; q += kPullThreads) {
        const uint2 E = e[q];
        const int64_t so = src[E.y >> kPullSlotBits] + (E.y & kMask);
        const int64_t d = dst + E.x;
        if (MODE == 2) {
            double2 v[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) v[f] = *(const double2 *)(fl.f[f] + so);
#pragma unroll
            for (int f = 0; f < NF; ++f) *(double2 *)(fl.f[f] + d) = v[f];
        } else if (MODE == 1) {
            double v[NF][2];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                v[f][0] = fl.f[f][so];
                v[f][1] = fl.f[f][so + 1];
            }
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                fl.f[f][d] = v[f][0];
                fl.f[f][d + 1] = v[f][1];
            }
        } else {
            double v[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) v[f] = fl.f[f][so];
#pragma unroll
            for (int f = 0; f < NF; ++f) fl.f[f][d] = v[f];
        }
    }
}

template <int NF>
void launch_pull_nf(const PullBlock *blocks, int nblock, const uint2 *ents, int mode,
                    const FieldList &fl, hipStream_t s) {
    if (mode == 2) k_pull<NF, 2><<<nblock, kPullThreads, 0, s>>>(blocks, ents, fl);
    else if (mode == 1) k_pull<NF, 1><<<nblock, kPullThreads, 0, s>>>(blocks, ents, fl);
    else k_pull<NF, 0><<<nblock, kPullThreads, 0, s>>>(blocks, ents, fl);
}

void launch_pull(const PullBlock *blocks, int nblock, const uint2 *ents, int mode,
                 double *const *fields, int nfield, hipStream_t s) {
    MGP_RANGE("mgp:pull");
    if (nblock <= 0 || nfield <= 0) return;
    FieldList fl{};
    for (int f = 0; f < nfield && f < 6; ++f) fl.f[f] = fields[f];
    switch (nfield) {
    case 1: launch_pull_nf<1>(blocks, nblock, ents, mode, fl, s); break;
    case 2: launch_pull_nf<2>(blocks, nblock, ents, mode, fl, s); break;
    case 3: launch_pull_nf<3>(blocks, nblock, ents, mode, fl, s); break;
    case 4: launch_pull_nf<4>(blocks, nblock, ents, mode, fl, s); break;
    case 5: launch_pull_nf<5>(blocks, nblock, ents, mode, fl, s); break;
    default: launch_pull_nf<6>(blocks, nblock, ents, mode, fl, s); break;
    }
}

}  // namespace mgp
