// mgp_kernels.hip -- gfx950 kernels of the pressure projection.
//
// Design (per block, one workgroup of 256 lanes = four wave64s):
//   * SIP sweeps walk the i+j+k hyperplanes of the block (the reference's
//     dependency order, src/flow/sip_hyperplane_mod.F90:450-569).  Every
//     per-cell array is stored hyperplane-major (WO layout, see
//     mgp_internal.h), so each step reads one contiguous run per array.
//   * The recurrence state lives in LDS as a 2-D "pencil" array R[j][i]
//     holding the latest value along each k-pencil: cell (k, j, i) of plane
//     p reads (k-1, j, i), (k, j-1, i), (k, j, i-1) from R of plane p-1.
//     R is double-buffered, so one barrier per plane suffices.
//   * The Laplacian fused into the forward sweep reads a 4-slot ring of
//     hyperplanes of HILF, also stored per pencil ([j][i]) in LDS.
//   * Ghost values come from face buffers written by the neighbouring
//     blocks' previous kernels; the reference's halo copies of RES and DP
//     (src/flow/pressuresolver_mod.F90:414, 453-457) become reads of those
//     buffers, with the same lagging.
// All kernels are compiled with -ffp-contract=off: each expression is
// evaluated in the reference's association order without FMA, which makes
// the FP64 path bitwise equal to the CPU oracle.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mgp_launch.h"

namespace mgp {

#define ALFA ((double)0.92f)
#define TINY ((double)1.0e-20f)

constexpr int kThreads = 256;

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

// ---------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T *p, int64_t n, T v) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        p[e] = v;
}

template <typename T>
void launch_fill(T *p, int64_t n, T v, hipStream_t s) {
    if (n <= 0) return;
    int64_t nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    k_fill<T><<<(unsigned)nb, 256, 0, s>>>(p, n, v);
}
void launch_fill_u8(uint8_t *p, int64_t n, uint8_t v, hipStream_t s) {
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
    const double *at_, Mask m, T *lw_o, T *ls_o, T *lb_o, T *lpr_o, T *ue_o,
    T *un_o, T *ut_o) {
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
               const double *ab, const double *at, Mask m, T *const f[7],
               hipStream_t s) {
    const size_t lds = 6 * (size_t)(t.max_nj + 1) * (t.max_ni + 1) *
                       sizeof(double);
    k_lu<T><<<t.nlist, kThreads, lds, s>>>(t, ap, aw, ae, as, an, ab, at, m,
                                           f[0], f[1], f[2], f[3], f[4], f[5],
                                           f[6]);
}
template void launch_lu<double>(const Topo &, const double *, const double *,
                                const double *, const double *, const double *,
                                const double *, const double *, Mask,
                                double *const[7], hipStream_t);
template void launch_lu<float>(const Topo &, const double *, const double *,
                               const double *, const double *, const double *,
                               const double *, const double *, Mask,
                               float *const[7], hipStream_t);

// ---------------------------------------------------------------------
// RHS = fak * div(u, v, w): src/ib/noib_mod.F90:304-354 (noib) and the
// ghostcell variant src/ib/gc_mod.F90:278-347 (+ fak*SDIV with SDIV = 0,
// then x BP).  One thread per WO cell; u, v, w in the reference layout.
template <typename T>
__global__ void k_divcal(Topo t, const double *u, const double *v,
                         const double *w, const double *rddx_,
                         const double *rddy_, const double *rddz_,
                         const double *bp_sp, double fak, double *rhs64,
                         T *rhs_t) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const uint32_t *cells = t.cells + S.cell_off;
    const int64_t kk = B.nk + 4, jj = B.nj + 4;
    const double *rddx = rddx_ + B.o1x + 2, *rddy = rddy_ + B.o1y + 2,
                 *rddz = rddz_ + B.o1z + 2;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < S.ncell;
         s += gridDim.x * blockDim.x) {
        int i, j, k;
        unpack(cells[s], i, j, k);
        const int64_t c = B.sp + (k + 2) + (j + 2) * kk + (i + 2) * kk * jj;
        double d = fak * ((u[c] - u[c - kk * jj]) * rddx[i] +
                          (v[c] - v[c - kk]) * rddy[j] +
                          (w[c] - w[c - 1]) * rddz[k]);
        if (bp_sp) {
            d = d + fak * 0.0;
            d = bp_sp[c] * d;
        }
        rhs64[B.wo + s] = d;
        if (rhs_t) rhs_t[B.wo + s] = (T)d;
    }
}

template <typename T>
void launch_divcal(const Topo &t, const double *u, const double *v,
                   const double *w, const double *rddx, const double *rddy,
                   const double *rddz, const double *bp_sp, double fak,
                   double *rhs64, T *rhs_t, hipStream_t s) {
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    dim3 grid((maxc + kThreads - 1) / kThreads, t.nlist);
    k_divcal<T><<<grid, kThreads, 0, s>>>(t, u, v, w, rddx, rddy, rddz,
                                          bp_sp, fak, rhs64, rhs_t);
}
template void launch_divcal<double>(const Topo &, const double *,
                                    const double *, const double *,
                                    const double *, const double *,
                                    const double *, const double *, double,
                                    double *, double *, hipStream_t);
template void launch_divcal<float>(const Topo &, const double *,
                                   const double *, const double *,
                                   const double *, const double *,
                                   const double *, const double *, double,
                                   double *, float *, hipStream_t);

// ---------------------------------------------------------------------
// Block-wide max of non-negative doubles folded into a global maximum.
// Non-negative IEEE doubles order like their bit patterns.
__device__ void block_max_to_global(double m, unsigned long long *g) {
    __shared__ double red[kThreads / 64];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(m, off, 64);
        m = (o > m || m != m) ? o : m;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q)
            r = (red[q] > r || r != r) ? red[q] : r;
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
    launch_fill<unsigned long long>(o.gmax, 1, 0ull, s);
    k_maxabs<<<t.nlist, kThreads, 0, s>>>(t, rhs64, o);
}

// ---------------------------------------------------------------------
// Plane-window helpers.  H holds 4 hyperplanes of a field, each stored per
// pencil [j][i] with a one-cell border: index (j+1) + (NJ+2)(i+1).  Plane q
// (q = -1 .. NP) goes to slot (q+1)&3.  Ghost entries (one index out of
// range by one) come from face buffers.
template <typename T>
struct Window {
    const T *zlo, *zhi, *ylo, *yhi, *xlo, *xhi;  // neighbour faces or null
};

// Ghost sources of HILF per face: the neighbour's boundary plane (CON),
// the block's own PAR ghosts from bound_pressure, or none (zero).
template <typename T>
__device__ __forceinline__ Window<T> hilf_window(const T *hface,
                                                 const T *hpar,
                                                 const Block *blocks,
                                                 const Block &B) {
    Window<T> w;
    const T *g[6];
    for (int f = 0; f < 6; ++f) {
        g[f] = nbr_face(hface, blocks, B.nbr[f], f ^ 1);
        if (!g[f] && ((B.par >> f) & 1))
            g[f] = hpar + B.fc + face_base(f, B.ni, B.nj, B.nk);
    }
    w.xlo = g[0];
    w.xhi = g[1];
    w.ylo = g[2];
    w.yhi = g[3];
    w.zlo = g[4];
    w.zhi = g[5];
    return w;
}

// res = -(A phi) at one cell, src/flow/laplacephi_mod.F90:149-162, with
// phi and BP read from the plane windows of planes p-1, p, p+1.
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
        r = r - ae * pe;
        r = r - as * psv;
        r = r - an * pn;
        r = r - ab * pb;
        r = r - at * pt;
    }
    r = r - ap * pc;
    return r;
}

// Software pipeline shared by the plane-sweep kernels.  During step p a
// workgroup (a) issues the global loads for step p+1 (per-cell data) and
// for plane p+3 of the HILF window, (b) computes plane p from registers and
// LDS, (c) moves the window plane loaded during step p-1 into LDS, then
// (d) hits the step's single LDS-only barrier.  The step loop is unrolled
// twice with ping-pong register buffers, so no register holding a load in
// flight is ever copied (a copy would force the wave to wait for it), and
// every load is unconditional (clamped index, selected afterwards), which
// keeps the compiler's wait counts exact.  The 1-D coefficients of the
// block sit in LDS.
constexpr int kMaxPlanes = 512;
constexpr int kMaxN1 = 64;      // interior cells per direction (4 ni <= 256)
constexpr uint32_t kNone = 0xffffffffu;

// 1-D coefficient rows of a block in LDS: [w e s n b t][kMaxN1]
template <typename T>
__device__ __forceinline__ void coef_to_lds(T *c, const T *aw, const T *ae,
                                            const T *as, const T *an,
                                            const T *ab, const T *at,
                                            const Block &B) {
    for (int e = threadIdx.x; e < 6 * kMaxN1; e += blockDim.x) {
        const int r = e / kMaxN1, q = e % kMaxN1;
        const int n = r < 2 ? B.ni : (r < 4 ? B.nj : B.nk);
        if (q >= n) continue;
        const T *src = r == 0 ? aw + B.o1x : r == 1 ? ae + B.o1x
                     : r == 2 ? as + B.o1y : r == 3 ? an + B.o1y
                     : r == 4 ? ab + B.o1z : at + B.o1z;
        c[e] = src[q + 2];
    }
}

// One hyperplane of a window: up to MS interior cells and one ghost entry
// per lane (4 NI + 2 NJ <= 256 is enforced at init).
template <typename T, int MS>
struct WinReg {
    uint32_t c[MS];
    T v[MS];
    uint8_t b[MS];
    int gi;
    T g;
    uint8_t gb;
};

template <typename T, int MS, bool MASK>
__device__ __forceinline__ void win_load(WinReg<T, MS> &w, int q, const Block &B,
                                         int np, const int *sps,
                                         const uint32_t *cells, const T *field,
                                         const Window<T> &hw, Mask m) {
    const int NI = B.ni, NJ = B.nj, NK = B.nk, HW = NJ + 2;
    const bool inplane = q >= 0 && q < np;
    const int base = inplane ? sps[q] : 0;
    const int cnt = inplane ? sps[q + 1] - base : 0;
#pragma unroll
    for (int mm = 0; mm < MS; ++mm) {
        const int x = threadIdx.x + mm * kThreads;
        const bool ok = x < cnt;
        const int s = base + (ok ? x : 0);
        const uint32_t cv = cells[s];
        const T fv = field[B.wo + s];
        w.c[mm] = ok ? cv : kNone;
        w.v[mm] = fv;
        if (MASK) w.b[mm] = m.bp[B.wo + s];
    }
    // ghost entry (one index out of range by one)
    const int e = threadIdx.x;
    int idx = -1, fo = 0, ff = 0;
    const T *src = nullptr;
    int i, j, k;
    if (e < NI) {
        i = e; j = q + 1 - i;
        if (j >= 0 && j < NJ) { idx = (j + 1) + HW * (i + 1); src = hw.zlo; fo = fidx_z(j, i, NI, NJ); ff = 4; }
    } else if (e < 2 * NI) {
        i = e - NI; j = q - NK - i;
        if (j >= 0 && j < NJ) { idx = (j + 1) + HW * (i + 1); src = hw.zhi; fo = fidx_z(j, i, NI, NJ); ff = 5; }
    } else if (e < 3 * NI) {
        i = e - 2 * NI; k = q + 1 - i;
        if (k >= 0 && k < NK) { idx = HW * (i + 1); src = hw.ylo; fo = fidx_y(k, i, NI, NK); ff = 2; }
    } else if (e < 4 * NI) {
        i = e - 3 * NI; k = q - NJ - i;
        if (k >= 0 && k < NK) { idx = (NJ + 1) + HW * (i + 1); src = hw.yhi; fo = fidx_y(k, i, NI, NK); ff = 3; }
    } else if (e < 4 * NI + NJ) {
        j = e - 4 * NI; k = q + 1 - j;
        if (k >= 0 && k < NK) { idx = j + 1; src = hw.xlo; fo = fidx_x(k, j, NJ, NK); ff = 0; }
    } else if (e < 4 * NI + 2 * NJ) {
        j = e - 4 * NI - NJ; k = q - NI - j;
        if (k >= 0 && k < NK) { idx = (j + 1) + HW * (NI + 1); src = hw.xhi; fo = fidx_x(k, j, NJ, NK); ff = 1; }
    }
    const bool has = idx >= 0 && src != nullptr;
    const T gv = *(has ? src + fo : field + B.wo);
    w.gi = idx;
    w.g = has ? gv : T(0);
    if (MASK) {
        const uint8_t *bf = m.bpf + B.fc;
        const uint8_t bv = *(idx >= 0 ? bf + face_base(ff, NI, NJ, NK) + fo : bf);
        w.gb = idx >= 0 ? bv : 0;
    }
}

template <typename T, int MS, bool MASK>
__device__ __forceinline__ void win_store(const WinReg<T, MS> &w, T *Hq,
                                          uint8_t *Bq, int HW) {
#pragma unroll
    for (int mm = 0; mm < MS; ++mm)
        if (w.c[mm] != kNone) {
            int i, j, k;
            unpack(w.c[mm], i, j, k);
            Hq[(j + 1) + HW * (i + 1)] = w.v[mm];
            if (MASK) Bq[(j + 1) + HW * (i + 1)] = w.b[mm];
        }
    if (w.gi >= 0) {
        Hq[w.gi] = w.g;
        if (MASK) Bq[w.gi] = w.gb;
    }
}

// ghost value of a sweep's recurrence border, prefetched one step ahead
template <typename T>
struct Border {
    T v;
    int idx;
};

template <typename T>
struct FwdSlot {
    uint32_t c;
    T rhs, lpr, lb, ls, lw, ap;
};

// ---------------------------------------------------------------------
// Fused inner-iteration first half: RES = -(A HILF) (laplacephi_level),
// then sipiter1 (src/flow/sip_hyperplane_mod.F90:450-505, classic order
// src/flow/sip_classic_mod.F90:143-166).  Low-side ghosts of RES are the
// neighbours' upper planes of the previous exchange (rhi[rd]); the upper
// planes of this sweep go to rhi[1-rd] unless this is the last inner
// iteration (the reference's conn(..., forward=-1) then skips faces
// 1/3/5), the lower planes always go to rlo for the backward sweep.
// LAP = false is the first inner iteration of an outer iteration, where
// HILF = 0 and the Laplacian term vanishes.
template <typename T, bool CLASSIC, bool MASKED, bool LAP, int MS>
__global__ __launch_bounds__(kThreads) void k_forward(
    Topo t, Fields<T> F, Mask m, int rd, int write_hi, const int *done) {
    if (*done) return;
    extern __shared__ __align__(16) unsigned char lds_fw[];
    __shared__ int sps[kMaxPlanes + 1];
    __shared__ T cf[6 * kMaxN1];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int NI = B.ni, NJ = B.nj, NK = B.nk, NP = S.np;
    const int RW = NJ + 1, RS = (NJ + 1) * (NI + 1);
    const int HW = NJ + 2;
    const int mRS = (t.max_nj + 1) * (t.max_ni + 1);
    const int mHS = (t.max_nj + 2) * (t.max_ni + 2);
    T *R = (T *)lds_fw;
    T *H = R + 2 * mRS;
    uint8_t *BH = (uint8_t *)(H + 4 * mHS);
    const uint32_t *cells = t.cells + S.cell_off;
    for (int e = threadIdx.x; e <= NP; e += blockDim.x)
        sps[e] = t.pstart[S.pstart_off + e];
    const int64_t wo = B.wo;
    const T *rh = F.rhi[rd];
    T *rw = write_hi ? F.rhi[rd ^ 1] : nullptr;
    const T *rzlo = nbr_face(rh, t.blocks, B.nbr[4], 5);
    const T *rylo = nbr_face(rh, t.blocks, B.nbr[2], 3);
    const T *rxlo = nbr_face(rh, t.blocks, B.nbr[0], 1);
    const Window<T> hw = hilf_window(F.hface, F.hpar, t.blocks, B);
    coef_to_lds(cf, F.aw, F.ae, F.as, F.an, F.ab, F.at, B);
    const T *aw = cf, *ae = cf + kMaxN1, *as = cf + 2 * kMaxN1,
            *an = cf + 3 * kMaxN1, *ab = cf + 4 * kMaxN1, *at = cf + 5 * kMaxN1;
    const T *lw = F.f[0] + wo, *ls = F.f[1] + wo, *lb = F.f[2] + wo,
            *lpr = F.f[3] + wo;
    const T *rhs = F.rhs + wo, *apv = F.ap + wo;
    T *res = F.res + wo;
    const T *dummy = rhs;
    const int64_t fc = B.fc;
    const int fb0 = face_base(0, NI, NJ, NK), fb1 = face_base(1, NI, NJ, NK),
              fb2 = face_base(2, NI, NJ, NK), fb3 = face_base(3, NI, NJ, NK),
              fb4 = face_base(4, NI, NJ, NK), fb5 = face_base(5, NI, NJ, NK);
    const bool masked = MASKED && !(B.flags & 1);

    // pencils start from the (lagged) ghost below k = 0
    for (int e = threadIdx.x; e < NJ * NI; e += blockDim.x) {
        const int j = e % NJ, i = e / NJ;
        const T v = rzlo ? rzlo[fidx_z(j, i, NI, NJ)] : T(0);
        R[(j + 1) + RW * (i + 1)] = v;
        R[RS + (j + 1) + RW * (i + 1)] = v;
    }
    // border values j = -1 / i = -1 read at step pp from R[(pp-1)&1]
    auto border_load = [&](int pp, Border<T> &bo) {
        const int e = threadIdx.x;
        int idx = -1;
        const T *src = dummy;
        if (e < NI) {
            const int i = e, k = pp - i;
            if (k >= 0 && k < NK) { idx = RW * (i + 1); if (rylo) src = rylo + fidx_y(k, i, NI, NK); }
        } else if (e < NI + NJ) {
            const int j = e - NI, k = pp - j;
            if (k >= 0 && k < NK) { idx = j + 1; if (rxlo) src = rxlo + fidx_x(k, j, NJ, NK); }
        }
        const T v = *src;
        bo.idx = idx;
        bo.v = src != dummy ? v : T(0);
    };
    __syncthreads();  // sps, cf
    auto load_slots = [&](int q, FwdSlot<T> (&d)[MS]) {
        const bool ok = q < NP;
        const int base = ok ? sps[q] : 0;
        const int cnt = ok ? sps[q + 1] - base : 0;
#pragma unroll
        for (int mm = 0; mm < MS; ++mm) {
            const int x = threadIdx.x + mm * kThreads;
            const bool in = x < cnt;
            const int s = base + (in ? x : 0);
            const uint32_t cv = cells[s];
            d[mm].rhs = rhs[s];
            d[mm].lpr = lpr[s];
            d[mm].lb = lb[s];
            d[mm].ls = ls[s];
            d[mm].lw = lw[s];
            if (LAP) d[mm].ap = apv[s];
            d[mm].c = in ? cv : kNone;
        }
    };
    auto wload = [&](WinReg<T, MS> &w, int q) {
        if (MASKED && masked) win_load<T, MS, true>(w, q, B, NP, sps, cells, F.hilf, hw, m);
        else win_load<T, MS, false>(w, q, B, NP, sps, cells, F.hilf, hw, m);
    };
    auto wstore = [&](const WinReg<T, MS> &w, int q) {
        const int sl = (q + 1) & 3;
        if (MASKED && masked) win_store<T, MS, true>(w, H + sl * mHS, BH + sl * mHS, HW);
        else win_store<T, MS, false>(w, H + sl * mHS, nullptr, HW);
    };
    {
        Border<T> b0;
        border_load(0, b0);
        if (b0.idx >= 0) R[RS + b0.idx] = b0.v;
    }
    Border<T> ba;
    border_load(1, ba);
    WinReg<T, MS> wa;
    if (LAP) {
        for (int q = -1; q <= 1; ++q) {
            wload(wa, q);
            wstore(wa, q);
        }
        wload(wa, 2);
    }
    FwdSlot<T> sa[MS];
    load_slots(0, sa);
    lds_barrier();

    // one step per iteration; the register rotation at its end waits for
    // the prefetched loads, but keeps the kernel at 6-7 waves per SIMD
    for (int p = 0; p < NP; ++p) {
        FwdSlot<T> sn[MS];
        Border<T> bn;
        WinReg<T, MS> wn;
        load_slots(p + 1, sn);
        border_load(p + 2, bn);
        if (LAP) wload(wn, p + 3);
        const T *Rp = R + ((p - 1) & 1) * RS;
        T *Rc = R + (p & 1) * RS;
        const T *Hm = H + (p & 3) * mHS, *H0 = H + ((p + 1) & 3) * mHS,
                *Hp = H + ((p + 2) & 3) * mHS;
        const uint8_t *Bm = BH + (p & 3) * mHS, *B0 = BH + ((p + 1) & 3) * mHS,
                      *Bp = BH + ((p + 2) & 3) * mHS;
        const int base = sps[p];
#pragma unroll
        for (int mm = 0; mm < MS; ++mm) {
            if (sa[mm].c == kNone) continue;
            const int s = base + threadIdx.x + mm * kThreads;
            int i, j, k;
            unpack(sa[mm].c, i, j, k);
            T r = T(0);
            if (LAP) {
                const int c0 = (j + 1) + HW * (i + 1);
                if (MASKED && masked)
                    r = laplace_cell<T, T, true>(Hm, H0, Hp, Bm, B0, Bp, c0, HW, aw[i], ae[i],
                                                 as[j], an[j], ab[k], at[k], sa[mm].ap);
                else
                    r = laplace_cell<T, T, false>(Hm, H0, Hp, Bm, B0, Bp, c0, HW, aw[i], ae[i],
                                                  as[j], an[j], ab[k], at[k], sa[mm].ap);
            }
            const int c1 = (j + 1) + RW * (i + 1);
            r = (sa[mm].rhs + r) * sa[mm].lpr;
            if (CLASSIC) {
                r = r - sa[mm].lw * Rp[c1 - RW];
                r = r - sa[mm].ls * Rp[c1 - 1];
                r = r - sa[mm].lb * Rp[c1];
            } else {
                r = r - sa[mm].lb * Rp[c1];
                r = r - sa[mm].ls * Rp[c1 - 1];
                r = r - sa[mm].lw * Rp[c1 - RW];
            }
            Rc[c1] = r;
            res[s] = r;
            if (i == 0) F.rlo[fc + fb0 + fidx_x(k, j, NJ, NK)] = r;
            if (j == 0) F.rlo[fc + fb2 + fidx_y(k, i, NI, NK)] = r;
            if (k == 0) F.rlo[fc + fb4 + fidx_z(j, i, NI, NJ)] = r;
            if (rw) {
                if (i == NI - 1) rw[fc + fb1 + fidx_x(k, j, NJ, NK)] = r;
                if (j == NJ - 1) rw[fc + fb3 + fidx_y(k, i, NI, NK)] = r;
                if (k == NK - 1) rw[fc + fb5 + fidx_z(j, i, NI, NJ)] = r;
            }
        }
        if (ba.idx >= 0) Rc[ba.idx] = ba.v;
        if (LAP) wstore(wa, p + 2);
        lds_barrier();
#pragma unroll
        for (int mm = 0; mm < MS; ++mm) sa[mm] = sn[mm];
        wa = wn;
        ba = bn;
    }
}

template <typename T, bool CLASSIC, bool MASKED, bool LAP>
static void launch_fwd_ms(int ms, const Topo &t, const Fields<T> &f, Mask m,
                          int rd, int write_hi, const int *done, size_t lds,
                          hipStream_t s) {
    switch (ms) {
    case 1: k_forward<T, CLASSIC, MASKED, LAP, 1><<<t.nlist, kThreads, lds, s>>>(t, f, m, rd, write_hi, done); break;
    case 2: k_forward<T, CLASSIC, MASKED, LAP, 2><<<t.nlist, kThreads, lds, s>>>(t, f, m, rd, write_hi, done); break;
    case 3: k_forward<T, CLASSIC, MASKED, LAP, 3><<<t.nlist, kThreads, lds, s>>>(t, f, m, rd, write_hi, done); break;
    default: k_forward<T, CLASSIC, MASKED, LAP, 4><<<t.nlist, kThreads, lds, s>>>(t, f, m, rd, write_hi, done); break;
    }
}

// slots per lane: the largest hyperplane of any block shape / 256
inline int slots_per_lane(const Topo &t) {
    const int ms = (t.max_plane + kThreads - 1) / kThreads;
    return ms < 1 ? 1 : ms;
}

template <typename T>
void launch_forward(const Topo &t, const Fields<T> &f, Mask m, int rd,
                    int write_hi, int hilf_zero, int classic, const int *done,
                    hipStream_t s) {
    const size_t mRS = (size_t)(t.max_nj + 1) * (t.max_ni + 1);
    const size_t mHS = (size_t)(t.max_nj + 2) * (t.max_ni + 2);
    const bool masked = m.bp != nullptr;
    const size_t lds = 2 * mRS * sizeof(T) +
                       (hilf_zero ? 0 : 4 * mHS * sizeof(T) + (masked ? 4 * mHS : 0)) + 16;
    const int ms = slots_per_lane(t);
    const bool lap = !hilf_zero;
#define MGP_FWD(C, M, L) launch_fwd_ms<T, C, M, L>(ms, t, f, m, rd, write_hi, done, lds, s)
    if (classic) {
        if (masked) { if (lap) MGP_FWD(true, true, true); else MGP_FWD(true, true, false); }
        else { if (lap) MGP_FWD(true, false, true); else MGP_FWD(true, false, false); }
    } else {
        if (masked) { if (lap) MGP_FWD(false, true, true); else MGP_FWD(false, true, false); }
        else { if (lap) MGP_FWD(false, false, true); else MGP_FWD(false, false, false); }
    }
#undef MGP_FWD
}
template void launch_forward<double>(const Topo &, const Fields<double> &,
                                     Mask, int, int, int, int, const int *,
                                     hipStream_t);
template void launch_forward<float>(const Topo &, const Fields<float> &, Mask,
                                    int, int, int, int, const int *,
                                    hipStream_t);

// ---------------------------------------------------------------------
template <typename T>
struct BwdSlot {
    uint32_t c;
    T res, ue, un, ut, h;
};

// Inner-iteration second half: sipiter2 (src/flow/sip_hyperplane_mod.F90:
// 508-569; classic order src/flow/sip_classic_mod.F90:169-198), then
// HILF += RES.  Upper ghosts of RES are the neighbours' lower planes of
// this iteration's forward sweep.  Also writes HILF's boundary planes,
// which stand in for the reference's conn(ilevel, 1, s1=dp) (:414).
// HZ: HILF is zero on entry (first inner iteration of an outer iteration).
template <typename T, bool CLASSIC, bool HZ, int MS>
__global__ __launch_bounds__(kThreads) void k_backward(Topo t, Fields<T> F,
                                                       const int *done) {
    if (*done) return;
    extern __shared__ __align__(16) unsigned char lds_bw[];
    __shared__ int sps[kMaxPlanes + 1];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int NI = B.ni, NJ = B.nj, NK = B.nk, NP = S.np;
    const int RW = NJ + 1, RS = (NJ + 1) * (NI + 1);
    T *R = (T *)lds_bw;
    const uint32_t *cells = t.cells + S.cell_off;
    for (int e = threadIdx.x; e <= NP; e += blockDim.x)
        sps[e] = t.pstart[S.pstart_off + e];
    const int64_t wo = B.wo;
    const T *rzhi = nbr_face((const T *)F.rlo, t.blocks, B.nbr[5], 4);
    const T *ryhi = nbr_face((const T *)F.rlo, t.blocks, B.nbr[3], 2);
    const T *rxhi = nbr_face((const T *)F.rlo, t.blocks, B.nbr[1], 0);
    const T *ue = F.f[4] + wo, *un = F.f[5] + wo, *ut = F.f[6] + wo;
    const T *res = F.res + wo;
    T *hilf = F.hilf + wo;
    const T *dummy = res;
    const int64_t fc = B.fc;
    const int fb0 = face_base(0, NI, NJ, NK), fb1 = face_base(1, NI, NJ, NK),
              fb2 = face_base(2, NI, NJ, NK), fb3 = face_base(3, NI, NJ, NK),
              fb4 = face_base(4, NI, NJ, NK), fb5 = face_base(5, NI, NJ, NK);

    for (int e = threadIdx.x; e < NJ * NI; e += blockDim.x) {
        const int j = e % NJ, i = e / NJ;
        const T v = rzhi ? rzhi[fidx_z(j, i, NI, NJ)] : T(0);
        R[j + RW * i] = v;
        R[RS + j + RW * i] = v;
    }
    // border values j = NJ / i = NI read at step pp from R[(pp+1)&1]
    auto border_load = [&](int pp, Border<T> &bo) {
        const int e = threadIdx.x;
        int idx = -1;
        const T *src = dummy;
        if (e < NI) {
            const int i = e, k = pp - (NJ - 1) - i;
            if (k >= 0 && k < NK) { idx = NJ + RW * i; if (ryhi) src = ryhi + fidx_y(k, i, NI, NK); }
        } else if (e < NI + NJ) {
            const int j = e - NI, k = pp - j - (NI - 1);
            if (k >= 0 && k < NK) { idx = j + RW * NI; if (rxhi) src = rxhi + fidx_x(k, j, NJ, NK); }
        }
        const T v = *src;
        bo.idx = idx;
        bo.v = src != dummy ? v : T(0);
    };
    __syncthreads();  // sps
    auto load_slots = [&](int q, BwdSlot<T> (&d)[MS]) {
        const bool ok = q >= 0;
        const int base = ok ? sps[q] : 0;
        const int cnt = ok ? sps[q + 1] - base : 0;
#pragma unroll
        for (int mm = 0; mm < MS; ++mm) {
            const int x = threadIdx.x + mm * kThreads;
            const bool in = x < cnt;
            const int s = base + (in ? x : 0);
            const uint32_t cv = cells[s];
            d[mm].res = res[s];
            d[mm].ue = ue[s];
            d[mm].un = un[s];
            d[mm].ut = ut[s];
            if (!HZ) d[mm].h = hilf[s];
            d[mm].c = in ? cv : kNone;
        }
    };
    {
        Border<T> b0;
        border_load(NP - 1, b0);
        if (b0.idx >= 0) R[(NP & 1) * RS + b0.idx] = b0.v;
    }
    Border<T> ba, bb;
    border_load(NP - 2, ba);
    BwdSlot<T> sa[MS], sb[MS];
    load_slots(NP - 1, sa);
    lds_barrier();

    auto step = [&](int p, BwdSlot<T> (&cur)[MS], BwdSlot<T> (&nxt)[MS],
                    Border<T> &bcur, Border<T> &bnxt) {
        load_slots(p - 1, nxt);
        border_load(p - 2, bnxt);
        const T *Rp = R + ((p + 1) & 1) * RS;
        T *Rc = R + (p & 1) * RS;
        const int base = sps[p];
#pragma unroll
        for (int mm = 0; mm < MS; ++mm) {
            if (cur[mm].c == kNone) continue;
            const int s = base + threadIdx.x + mm * kThreads;
            int i, j, k;
            unpack(cur[mm].c, i, j, k);
            const int c1 = j + RW * i;
            T r = cur[mm].res;
            if (CLASSIC) {
                r = r - cur[mm].un * Rp[c1 + 1];
                r = r - cur[mm].ue * Rp[c1 + RW];
                r = r - cur[mm].ut * Rp[c1];
            } else {
                r = r - cur[mm].ut * Rp[c1];
                r = r - cur[mm].un * Rp[c1 + 1];
                r = r - cur[mm].ue * Rp[c1 + RW];
            }
            Rc[c1] = r;
            const T h = (HZ ? T(0) : cur[mm].h) + r;
            hilf[s] = h;
            if (i == 0) F.hface[fc + fb0 + fidx_x(k, j, NJ, NK)] = h;
            if (i == NI - 1) F.hface[fc + fb1 + fidx_x(k, j, NJ, NK)] = h;
            if (j == 0) F.hface[fc + fb2 + fidx_y(k, i, NI, NK)] = h;
            if (j == NJ - 1) F.hface[fc + fb3 + fidx_y(k, i, NI, NK)] = h;
            if (k == 0) F.hface[fc + fb4 + fidx_z(j, i, NI, NJ)] = h;
            if (k == NK - 1) F.hface[fc + fb5 + fidx_z(j, i, NI, NJ)] = h;
        }
        if (bcur.idx >= 0) Rc[bcur.idx] = bcur.v;
        lds_barrier();
    };
    for (int p = NP - 1; p >= 0; p -= 2) {
        step(p, sa, sb, ba, bb);
        if (p - 1 >= 0) step(p - 1, sb, sa, bb, ba);
    }
}

template <typename T, bool CLASSIC, bool HZ>
static void launch_bwd_ms(int ms, const Topo &t, const Fields<T> &f,
                          const int *done, size_t lds, hipStream_t s) {
    switch (ms) {
    case 1: k_backward<T, CLASSIC, HZ, 1><<<t.nlist, kThreads, lds, s>>>(t, f, done); break;
    case 2: k_backward<T, CLASSIC, HZ, 2><<<t.nlist, kThreads, lds, s>>>(t, f, done); break;
    case 3: k_backward<T, CLASSIC, HZ, 3><<<t.nlist, kThreads, lds, s>>>(t, f, done); break;
    default: k_backward<T, CLASSIC, HZ, 4><<<t.nlist, kThreads, lds, s>>>(t, f, done); break;
    }
}

template <typename T>
void launch_backward(const Topo &t, const Fields<T> &f, int hilf_zero,
                     int classic, const int *done, hipStream_t s) {
    const size_t mRS = (size_t)(t.max_nj + 1) * (t.max_ni + 1);
    const size_t lds = 2 * mRS * sizeof(T) + 16;
    const int ms = slots_per_lane(t);
    if (classic) {
        if (hilf_zero) launch_bwd_ms<T, true, true>(ms, t, f, done, lds, s);
        else launch_bwd_ms<T, true, false>(ms, t, f, done, lds, s);
    } else {
        if (hilf_zero) launch_bwd_ms<T, false, true>(ms, t, f, done, lds, s);
        else launch_bwd_ms<T, false, false>(ms, t, f, done, lds, s);
    }
}
template void launch_backward<double>(const Topo &, const Fields<double> &,
                                      int, int, const int *, hipStream_t);
template void launch_backward<float>(const Topo &, const Fields<float> &, int,
                                     int, const int *, hipStream_t);

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
    if (n <= 0) return;
    int64_t nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    k_rap<T><<<(unsigned)nb, 256, 0, s>>>(ap, rap, n);
}
template void launch_rap<double>(const double *, double *, int64_t, hipStream_t);
template void launch_rap<float>(const double *, float *, int64_t, hipStream_t);

template <typename T>
struct SorSlot {
    uint32_t c;
    T rhs, rap;
};

constexpr int kSorRing = 8;  // window planes p-3 .. p+2 live in LDS

// Red-black SOR of HILF (relax, src/flow/sor_mod.F90:102-203), both colours
// in one pass over the block's hyperplanes.  The reference's first colour
// (rb = 0: i+j+k even in its 1-based indices, :124-130) is exactly the set
// of odd planes p = i+j+k of the 0-based interior, the second colour the
// even planes, and every 7-point neighbour of a cell lies on plane p-1 or
// p+1.  Step p relaxes plane p when p is odd (its neighbours are old even
// planes) and plane p-2 when p is even (its neighbours p-3, p-1 were
// relaxed at earlier steps), so each value is read and written once.
// Ghost values stay those of the previous exchange (fin) for both colours,
// as in the reference, where conn(DP) follows relax (:414); the new
// boundary planes go to fout.  With an odd cell count along k the loop
// bounds of :124-130 never reach the cells k = nk-1; they keep their value.
// HZ: HILF is zero on entry (first inner iteration of an outer iteration).
template <typename T, bool MASKED, bool HZ, int MS>
__global__ __launch_bounds__(kThreads) void k_sor(Topo t, Fields<T> F, Mask m,
                                                  const T *fin, T *fout,
                                                  double omega,
                                                  const int *done) {
    if (*done) return;
    extern __shared__ __align__(16) unsigned char lds_so[];
    __shared__ int sps[kMaxPlanes + 1];
    __shared__ T cf[6 * kMaxN1];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int NI = B.ni, NJ = B.nj, NK = B.nk, NP = S.np, HW = NJ + 2;
    const int mHS = (t.max_nj + 2) * (t.max_ni + 2);
    T *H = (T *)lds_so;
    uint8_t *BH = (uint8_t *)(H + kSorRing * mHS);
    const uint32_t *cells = t.cells + S.cell_off;
    for (int e = threadIdx.x; e <= NP; e += blockDim.x)
        sps[e] = t.pstart[S.pstart_off + e];
    const int64_t wo = B.wo;
    const Window<T> hw = hilf_window(fin, F.hpar, t.blocks, B);
    coef_to_lds(cf, F.aw, F.ae, F.as, F.an, F.ab, F.at, B);
    const T *aw = cf, *ae = cf + kMaxN1, *as = cf + 2 * kMaxN1,
            *an = cf + 3 * kMaxN1, *ab = cf + 4 * kMaxN1, *at = cf + 5 * kMaxN1;
    const bool masked = MASKED && !(B.flags & 1);
    const T om = (T)omega, om1 = (T)(1.0 - omega);
    const bool keep_top = (NK & 1) != 0;
    const int64_t fc = B.fc;
    const int fb0 = face_base(0, NI, NJ, NK), fb1 = face_base(1, NI, NJ, NK),
              fb2 = face_base(2, NI, NJ, NK), fb3 = face_base(3, NI, NJ, NK),
              fb4 = face_base(4, NI, NJ, NK), fb5 = face_base(5, NI, NJ, NK);
    __syncthreads();  // sps, cf

    auto plane_of_step = [](int p) { return (p & 1) ? p : p - 2; };
    auto load_slots = [&](int u, SorSlot<T> (&d)[MS]) {
        const bool ok = u >= 0 && u < NP;
        const int base = ok ? sps[u] : 0;
        const int cnt = ok ? sps[u + 1] - base : 0;
#pragma unroll
        for (int mm = 0; mm < MS; ++mm) {
            const int x = threadIdx.x + mm * kThreads;
            const bool in = x < cnt;
            const int s = base + (in ? x : 0);
            const uint32_t cv = cells[s];
            d[mm].rhs = F.rhs[wo + s];
            d[mm].rap = F.rap[wo + s];
            d[mm].c = in ? cv : kNone;
        }
    };
    auto wload = [&](WinReg<T, MS> &w, int q) {
        if (MASKED && masked) win_load<T, MS, true>(w, q, B, NP, sps, cells, (const T *)F.hilf, hw, m);
        else win_load<T, MS, false>(w, q, B, NP, sps, cells, (const T *)F.hilf, hw, m);
        if (HZ) {
#pragma unroll
            for (int mm = 0; mm < MS; ++mm) w.v[mm] = T(0);
            w.g = T(0);
        }
    };
    auto wstore = [&](const WinReg<T, MS> &w, int q) {
        const int sl = (q + 1) & (kSorRing - 1);
        if (MASKED && masked) win_store<T, MS, true>(w, H + sl * mHS, BH + sl * mHS, HW);
        else win_store<T, MS, false>(w, H + sl * mHS, nullptr, HW);
    };
    WinReg<T, MS> wa;
    for (int q = -1; q <= 1; ++q) {
        wload(wa, q);
        wstore(wa, q);
    }
    wload(wa, 2);
    SorSlot<T> sa[MS];
    load_slots(plane_of_step(0), sa);
    lds_barrier();
    for (int p = 0; p <= NP + 1; ++p) {
        SorSlot<T> sn[MS];
        WinReg<T, MS> wn;
        load_slots(plane_of_step(p + 1), sn);
        wload(wn, p + 3);
        const int u = plane_of_step(p);
        if (u >= 0 && u < NP) {
            const int su = (u + 1) & (kSorRing - 1);
            T *H0 = H + su * mHS;
            const T *Hm = H + (u & (kSorRing - 1)) * mHS;
            const T *Hp = H + ((u + 2) & (kSorRing - 1)) * mHS;
            const uint8_t *B0 = BH + su * mHS;
            const uint8_t *Bm = BH + (u & (kSorRing - 1)) * mHS;
            const uint8_t *Bp = BH + ((u + 2) & (kSorRing - 1)) * mHS;
            const int base = sps[u];
#pragma unroll
            for (int mm = 0; mm < MS; ++mm) {
                if (sa[mm].c == kNone) continue;
                int i, j, k;
                unpack(sa[mm].c, i, j, k);
                const int c0 = (j + 1) + HW * (i + 1);
                const T old = H0[c0];
                T nv = old;
                if (!(keep_top && k == NK - 1)) {
                    const T pw = Hm[c0 - HW], pe = Hp[c0 + HW];
                    const T psv = Hm[c0 - 1], pn = Hp[c0 + 1];
                    const T pb = Hm[c0], pt = Hp[c0];
                    T r;
                    if (MASKED && masked) {
                        const T bc = (T)B0[c0];
                        const T cw = (T)Bm[c0 - HW] * bc * aw[i];
                        const T ce = bc * (T)Bp[c0 + HW] * ae[i];
                        const T cs = (T)Bm[c0 - 1] * bc * as[j];
                        const T cn = bc * (T)Bp[c0 + 1] * an[j];
                        const T cb = (T)Bm[c0] * bc * ab[k];
                        const T ct = bc * (T)Bp[c0] * at[k];
                        r = cw * pw + ce * pe + cs * psv + cn * pn + cb * pb + ct * pt;
                    } else {
                        r = aw[i] * pw + ae[i] * pe + as[j] * psv + an[j] * pn +
                            ab[k] * pb + at[k] * pt;
                    }
                    r = (r - sa[mm].rhs) * sa[mm].rap;
                    nv = om1 * old - om * r;
                }
                H0[c0] = nv;
                F.hilf[wo + base + threadIdx.x + mm * kThreads] = nv;
                if (i == 0) fout[fc + fb0 + fidx_x(k, j, NJ, NK)] = nv;
                if (i == NI - 1) fout[fc + fb1 + fidx_x(k, j, NJ, NK)] = nv;
                if (j == 0) fout[fc + fb2 + fidx_y(k, i, NI, NK)] = nv;
                if (j == NJ - 1) fout[fc + fb3 + fidx_y(k, i, NI, NK)] = nv;
                if (k == 0) fout[fc + fb4 + fidx_z(j, i, NI, NJ)] = nv;
                if (k == NK - 1) fout[fc + fb5 + fidx_z(j, i, NI, NJ)] = nv;
            }
        }
        wstore(wa, p + 2);
        lds_barrier();
#pragma unroll
        for (int mm = 0; mm < MS; ++mm) sa[mm] = sn[mm];
        wa = wn;
    }
}

template <typename T, bool MASKED, bool HZ>
static void launch_sor_ms(int ms, const Topo &t, const Fields<T> &f, Mask m,
                          const T *fin, T *fout, double omega,
                          const int *done, size_t lds, hipStream_t s) {
    switch (ms) {
    case 1: k_sor<T, MASKED, HZ, 1><<<t.nlist, kThreads, lds, s>>>(t, f, m, fin, fout, omega, done); break;
    case 2: k_sor<T, MASKED, HZ, 2><<<t.nlist, kThreads, lds, s>>>(t, f, m, fin, fout, omega, done); break;
    case 3: k_sor<T, MASKED, HZ, 3><<<t.nlist, kThreads, lds, s>>>(t, f, m, fin, fout, omega, done); break;
    default: k_sor<T, MASKED, HZ, 4><<<t.nlist, kThreads, lds, s>>>(t, f, m, fin, fout, omega, done); break;
    }
}

template <typename T>
void launch_sor(const Topo &t, const Fields<T> &f, Mask m, const T *fin,
                T *fout, double omega, int hilf_zero, const int *done,
                hipStream_t s) {
    const size_t mHS = (size_t)(t.max_nj + 2) * (t.max_ni + 2);
    const bool masked = m.bp != nullptr;
    const size_t lds = kSorRing * mHS * sizeof(T) + (masked ? kSorRing * mHS : 0) + 16;
    const int ms = slots_per_lane(t);
    if (masked) {
        if (hilf_zero) launch_sor_ms<T, true, true>(ms, t, f, m, fin, fout, omega, done, lds, s);
        else launch_sor_ms<T, true, false>(ms, t, f, m, fin, fout, omega, done, lds, s);
    } else {
        if (hilf_zero) launch_sor_ms<T, false, true>(ms, t, f, m, fin, fout, omega, done, lds, s);
        else launch_sor_ms<T, false, false>(ms, t, f, m, fin, fout, omega, done, lds, s);
    }
}
template void launch_sor<double>(const Topo &, const Fields<double> &, Mask,
                                 const double *, double *, double, int,
                                 const int *, hipStream_t);
template void launch_sor<float>(const Topo &, const Fields<float> &, Mask,
                                const float *, float *, double, int,
                                const int *, hipStream_t);

// ---------------------------------------------------------------------
struct OutSlot {
    uint32_t c;
    double ap, rhs, dp;
};

// FP64 outer update (src/flow/pressuresolver_mod.F90:258-279): RES =
// -(A HILF) with the FP64 operator, RHS += RES, max |RHS|, DP += HILF.
// In the mixed-precision mode the sweeps' RHS copy is refreshed here.
template <typename T, bool MASKED, int MS>
__global__ __launch_bounds__(kThreads) void k_outer(Topo t, Fields<T> F,
                                                    Mask m, Outer o,
                                                    T *rhs_t) {
    if (*o.done) return;
    extern __shared__ __align__(16) unsigned char lds_ou[];
    __shared__ int sps[kMaxPlanes + 1];
    __shared__ double cf[6 * kMaxN1];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int NJ = B.nj, NP = S.np, HW = NJ + 2;
    const int mHS = (t.max_nj + 2) * (t.max_ni + 2);
    T *H = (T *)lds_ou;
    uint8_t *BH = (uint8_t *)(H + 4 * mHS);
    const uint32_t *cells = t.cells + S.cell_off;
    for (int e = threadIdx.x; e <= NP; e += blockDim.x)
        sps[e] = t.pstart[S.pstart_off + e];
    const int64_t wo = B.wo;
    const Window<T> hw = hilf_window(F.hface, F.hpar, t.blocks, B);
    coef_to_lds(cf, o.aw, o.ae, o.as, o.an, o.ab, o.at, B);
    const double *aw = cf, *ae = cf + kMaxN1, *as = cf + 2 * kMaxN1,
                 *an = cf + 3 * kMaxN1, *ab = cf + 4 * kMaxN1, *at = cf + 5 * kMaxN1;
    const bool masked = MASKED && !(B.flags & 1);
    double mx = 0.0;
    __syncthreads();  // sps, cf
    auto load_slots = [&](int q, OutSlot (&d)[MS]) {
        const bool ok = q < NP;
        const int base = ok ? sps[q] : 0;
        const int cnt = ok ? sps[q + 1] - base : 0;
#pragma unroll
        for (int mm = 0; mm < MS; ++mm) {
            const int x = threadIdx.x + mm * kThreads;
            const bool in = x < cnt;
            const int64_t s = wo + base + (in ? x : 0);
            const uint32_t cv = cells[base + (in ? x : 0)];
            d[mm].ap = o.ap64[s];
            d[mm].rhs = o.rhs64[s];
            d[mm].dp = o.dp64[s];
            d[mm].c = in ? cv : kNone;
        }
    };
    auto wload = [&](WinReg<T, MS> &w, int q) {
        if (MASKED && masked) win_load<T, MS, true>(w, q, B, NP, sps, cells, (const T *)F.hilf, hw, m);
        else win_load<T, MS, false>(w, q, B, NP, sps, cells, (const T *)F.hilf, hw, m);
    };
    auto wstore = [&](const WinReg<T, MS> &w, int q) {
        const int sl = (q + 1) & 3;
        if (MASKED && masked) win_store<T, MS, true>(w, H + sl * mHS, BH + sl * mHS, HW);
        else win_store<T, MS, false>(w, H + sl * mHS, nullptr, HW);
    };
    WinReg<T, MS> wa;
    for (int q = -1; q <= 1; ++q) {
        wload(wa, q);
        wstore(wa, q);
    }
    wload(wa, 2);
    OutSlot sa[MS];
    load_slots(0, sa);
    lds_barrier();
    for (int p = 0; p < NP; ++p) {
        OutSlot sn[MS];
        WinReg<T, MS> wn;
        load_slots(p + 1, sn);
        wload(wn, p + 3);
        const T *Hm = H + (p & 3) * mHS, *H0 = H + ((p + 1) & 3) * mHS,
                *Hp = H + ((p + 2) & 3) * mHS;
        const uint8_t *Bm = BH + (p & 3) * mHS, *B0 = BH + ((p + 1) & 3) * mHS,
                      *Bp = BH + ((p + 2) & 3) * mHS;
        const int base = sps[p];
#pragma unroll
        for (int mm = 0; mm < MS; ++mm) {
            if (sa[mm].c == kNone) continue;
            const int64_t s = wo + base + threadIdx.x + mm * kThreads;
            int i, j, k;
            unpack(sa[mm].c, i, j, k);
            const int c0 = (j + 1) + HW * (i + 1);
            double r;
            if (MASKED && masked)
                r = laplace_cell<T, double, true>(Hm, H0, Hp, Bm, B0, Bp, c0, HW, aw[i], ae[i],
                                                  as[j], an[j], ab[k], at[k], sa[mm].ap);
            else
                r = laplace_cell<T, double, false>(Hm, H0, Hp, Bm, B0, Bp, c0, HW, aw[i], ae[i],
                                                   as[j], an[j], ab[k], at[k], sa[mm].ap);
            const double rn = sa[mm].rhs + r;
            o.rhs64[s] = rn;
            if (rhs_t) rhs_t[s] = (T)rn;
            const double a = fabs(rn);
            mx = (a > mx || a != a) ? a : mx;
            o.dp64[s] = sa[mm].dp + (double)H0[c0];
        }
        wstore(wa, p + 2);
        lds_barrier();
#pragma unroll
        for (int mm = 0; mm < MS; ++mm) sa[mm] = sn[mm];
        wa = wn;
    }
    mx = mx * ldexp(1.0, t.maxlevel - B.level);
    block_max_to_global(mx, o.gmax);
}

template <typename T, bool MASKED>
static void launch_out_ms(int ms, const Topo &t, const Fields<T> &f, Mask m,
                          Outer o, T *rhs_t, size_t lds, hipStream_t s) {
    switch (ms) {
    case 1: k_outer<T, MASKED, 1><<<t.nlist, kThreads, lds, s>>>(t, f, m, o, rhs_t); break;
    case 2: k_outer<T, MASKED, 2><<<t.nlist, kThreads, lds, s>>>(t, f, m, o, rhs_t); break;
    case 3: k_outer<T, MASKED, 3><<<t.nlist, kThreads, lds, s>>>(t, f, m, o, rhs_t); break;
    default: k_outer<T, MASKED, 4><<<t.nlist, kThreads, lds, s>>>(t, f, m, o, rhs_t); break;
    }
}

template <typename T>
void launch_outer(const Topo &t, const Fields<T> &f, Mask m, Outer o,
                  T *rhs_t, hipStream_t s) {
    const size_t mHS = (size_t)(t.max_nj + 2) * (t.max_ni + 2);
    const bool masked = m.bp != nullptr;
    const size_t lds = 4 * mHS * sizeof(T) + (masked ? 4 * mHS : 0) + 16;
    const int ms = slots_per_lane(t);
    if (masked) launch_out_ms<T, true>(ms, t, f, m, o, rhs_t, lds, s);
    else launch_out_ms<T, false>(ms, t, f, m, o, rhs_t, lds, s);
}
template void launch_outer<double>(const Topo &, const Fields<double> &, Mask,
                                   Outer, double *, hipStream_t);
template void launch_outer<float>(const Topo &, const Fields<float> &, Mask,
                                  Outer, float *, hipStream_t);

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
    k_check<<<1, 1, 0, s>>>(o, nouter, nouter_min, epcorr, eps_rel, prefak,
                            hist, hist_cap);
}

// ---------------------------------------------------------------------
// WO -> reference layout (interior only; ghosts are left as they are).
template <typename T>
__global__ void k_wo_to_sp(Topo t, const T *src, double *dst) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const uint32_t *cells = t.cells + S.cell_off;
    const int64_t kk = B.nk + 4, jj = B.nj + 4;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < S.ncell;
         s += gridDim.x * blockDim.x) {
        int i, j, k;
        unpack(cells[s], i, j, k);
        dst[B.sp + (k + 2) + (j + 2) * kk + (i + 2) * kk * jj] =
            (double)src[B.wo + s];
    }
}
template <typename T>
void launch_wo_to_sp(const Topo &t, const T *src, double *dst,
                     hipStream_t s) {
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    dim3 grid((maxc + kThreads - 1) / kThreads, t.nlist);
    k_wo_to_sp<T><<<grid, kThreads, 0, s>>>(t, src, dst);
}
template void launch_wo_to_sp<double>(const Topo &, const double *, double *,
                                      hipStream_t);
template void launch_wo_to_sp<float>(const Topo &, const float *, double *,
                                     hipStream_t);
template void launch_wo_to_sp<uint8_t>(const Topo &, const uint8_t *,
                                       double *, hipStream_t);

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
        ref[B.sp + d] = (double)wo[B.wo + s];
    }
}
template <typename T>
void launch_to_ref_factors(const Topo &t, const T *wo, double *ref,
                           int spatial, hipStream_t s) {
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
// 787-879) on the reference layout; one thread per stored cell.
__global__ void k_mgpcorr(Topo t, double *u_, double *v_, double *w_,
                          double *p_, const double *dp_, const double *rdx_,
                          const double *rdy_, const double *rdz_,
                          double rfak, const double *bp_) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const int kk = B.nk + 4, jj = B.nj + 4, ii = B.ni + 4;
    const int64_t n = (int64_t)kk * jj * ii;
    double *u = u_ + B.sp, *v = v_ + B.sp, *w = w_ + B.sp, *p = p_ + B.sp;
    const double *dp = dp_ + B.sp;
    const double *bp = bp_ ? bp_ + B.sp : nullptr;
    const double *rdx = rdx_ + B.o1x - 1, *rdy = rdy_ + B.o1y - 1,
                 *rdz = rdz_ + B.o1z - 1;  // 1-based
    const int64_t sj = kk, si = (int64_t)kk * jj;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % kk) + 1;
        const int j = (int)((e / kk) % jj) + 1;
        const int i = (int)(e / si) + 1;
        if (i >= 2 && i <= ii - 1 && j >= 2 && j <= jj - 1 && k >= 2 &&
            k <= kk - 1)
            p[e] = bp ? p[e] + dp[e] * bp[e] : p[e] + dp[e];
        if (i >= 2 && i <= ii - 2 && j >= 3 && j <= jj - 2 && k >= 3 &&
            k <= kk - 2)
            u[e] = bp ? u[e] + (dp[e] - dp[e + si]) * bp[e] * bp[e + si] *
                                   rdx[i] * rfak
                      : u[e] + (dp[e] - dp[e + si]) * rdx[i] * rfak;
        if (i >= 3 && i <= ii - 2 && j >= 2 && j <= jj - 2 && k >= 3 &&
            k <= kk - 2)
            v[e] = bp ? v[e] + (dp[e] - dp[e + sj]) * bp[e] * bp[e + sj] *
                                   rdy[j] * rfak
                      : v[e] + (dp[e] - dp[e + sj]) * rdy[j] * rfak;
        if (i >= 3 && i <= ii - 2 && j >= 3 && j <= jj - 2 && k >= 2 &&
            k <= kk - 2)
            w[e] = bp ? w[e] + (dp[e] - dp[e + 1]) * bp[e] * bp[e + 1] *
                                   rdz[k] * rfak
                      : w[e] + (dp[e] - dp[e + 1]) * rdz[k] * rfak;
    }
}

void launch_mgpcorr(const Topo &t, double *u, double *v, double *w,
                    double *p, const double *dp, const double *rdx,
                    const double *rdy, const double *rdz, double rfak,
                    const double *bp, hipStream_t s) {
    const int maxc = (t.max_ni + 4) * (t.max_nj + 4) * (t.max_nk + 4);
    dim3 grid((maxc + kThreads - 1) / kThreads, t.nlist);
    k_mgpcorr<<<grid, kThreads, 0, s>>>(t, u, v, w, p, dp, rdx, rdy, rdz,
                                        rfak, bp);
}

// ---------------------------------------------------------------------
// Ghost-layer box copies (connect_self_single, src/core/connect/
// conn1_mod.F90:674-704): one workgroup per task, all fields.
struct FieldList {
    double *f[6];
    int n;
};

__global__ void k_copy(const CopyTask *tasks, FieldList fl) {
    const CopyTask T = tasks[blockIdx.x];
    const int n = T.ni * T.nj * T.nk;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const int k = e % T.nk, j = (e / T.nk) % T.nj, i = e / (T.nk * T.nj);
        const int64_t so = T.src + k + (int64_t)j * T.s_sj + (int64_t)i * T.s_si;
        const int64_t d = T.dst + k + (int64_t)j * T.d_sj + (int64_t)i * T.d_si;
        for (int f = 0; f < fl.n; ++f) fl.f[f][d] = fl.f[f][so];
    }
}

void launch_copy(const CopyTask *tasks, int ntask, double *const *fields,
                 int nfield, hipStream_t s) {
    if (ntask <= 0) return;
    FieldList fl;
    fl.n = nfield;
    for (int f = 0; f < nfield && f < 6; ++f) fl.f[f] = fields[f];
    k_copy<<<ntask, 128, 0, s>>>(tasks, fl);
}

}  // namespace mgp
