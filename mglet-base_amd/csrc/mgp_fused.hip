// mgp_fused.hip -- the FP64 outer update of one outer iteration fused with
// the first SIP forward sweep of the next one (single-level SIP solves).
//
// The outer update (src/flow/pressuresolver_mod.F90:258-279) computes
// RHS += -(A HILF) in FP64, max |RHS| and DP += HILF.  The next outer
// iteration starts with HILF = 0 (:253-257), so its first forward sweep
// (sipiter1 without the Laplacian, src/flow/sip_hyperplane_mod.F90:450-505)
// needs only the new RHS, which this kernel has in a register: it runs that
// sweep right after the update of each cell, in the same hyperplane walk,
// instead of storing RHS and reading it back in a second pass.  The window
// (HILF planes with their face ghosts) and the recurrence ghosts come from
// the same plan as the forward sweep with the Laplacian.  Each value is
// computed by the same expressions as the separate kernels, so the results
// are bitwise those of k_outer followed by k_forward.
//
// The update belongs to outer iteration n, the sweep to n + 1: when the
// exit test after this kernel stops the loop, the sweep's RES and face
// planes are left unused.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "mgp_device.h"
#include "mgp_launch.h"
#include "mgp_plane.h"

namespace mgp {

// LATE (FP64 with factors shared between blocks, served by L2): the
// sweep's factors load at the start of compute, where the FP64 update
// hides their latency, instead of D items ahead (8 registers fewer per
// pipeline slot; on 32768 blocks of 24^3 5.92 -> 5.09 ms with 4 waves of
// 64-cell items; FP32 measured slower that way).
template <typename T, bool CLASSIC, bool MASKED, int NS, bool LATE = false>
struct OutFwdK {
    static constexpr bool kFaces = true;
    static constexpr bool kSplit = false;
    static constexpr bool kFar = false;
    struct Far {};
    struct Early {
        T lpr, lb, ls, lw;
    };
    struct Late {};
    struct Data : std::conditional_t<LATE, Late, Early> {
        uint32_t c;
        double ap, rhs, dp;
    };
    const uint32_t *cells;
    // outer update
    const double *ar;
    WallFold wf;
    const double *ap;
    double *rhs, *dp;
    T *rhs_t;
    const double *cf;
    // forward sweep
    const T *lpr, *lb, *ls, *lw;
    T *res, *rlo, *rw;
    T *R, *H;
    uint8_t *BH;
    int HW, RW, RS, mHS, nr;
    bool masked;
    double mx;

    __device__ __forceinline__ void load(Data &d, int s) const {
        d.c = ld(cells, s);
        if (MASKED) d.ap = ld(ap, s);
        d.rhs = ld((const double *)rhs, s);
        d.dp = ld((const double *)dp, s);
        if constexpr (!LATE) {
            d.lpr = ld(lpr, s);
            d.lb = ld(lb, s);
            d.ls = ld(ls, s);
            d.lw = ld(lw, s);
        }
    }
    __device__ __forceinline__ void compute(const Data &d, int s, int p, int) {
        T flpr, flb, fls, flw;
        if constexpr (LATE) {
            flpr = ld(lpr, s);
            flb = ld(lb, s);
            fls = ld(ls, s);
            flw = ld(lw, s);
        } else {
            flpr = d.lpr;
            flb = d.lb;
            fls = d.ls;
            flw = d.lw;
        }
        int i, j, k;
        cell_ijk(d.c, i, j, k);
        // outer update of iteration n (as k_outer)
        const int c0 = (j + 1) + __mul24(HW, i + 1);
        const int sm = win_slot<NS>(p - 1) * mHS, s0 = win_slot<NS>(p) * mHS,
                  sp = win_slot<NS>(p + 1) * mHS;
        const double apv = MASKED ? d.ap : ap_cell(ar, nr, i, j, k, cell_flags(d.c), wf);
        const Coef6<double> co(cf, nr, i, j, k);
        double q;
        if (MASKED && masked)
            q = laplace_cell<T, double, true>(H + sm, H + s0, H + sp, BH + sm, BH + s0,
                                              BH + sp, c0, HW, co.w, co.e, co.s, co.n, co.b, co.t,
                                              apv);
        else
            q = laplace_cell<T, double, false>(H + sm, H + s0, H + sp, nullptr, nullptr,
                                               nullptr, c0, HW, co.w, co.e, co.s, co.n, co.b, co.t,
                                               apv);
        const double rn = d.rhs + q;
        st(rhs, s, rn);
        if (rhs_t) st(rhs_t, s, (T)rn);
        const double a = fabs(rn);
        mx = (a > mx || a != a) ? a : mx;
        st(dp, s, d.dp + (double)H[s0 + c0]);
        // first forward sweep of iteration n + 1 (as k_forward with HILF = 0)
        const int c1 = (j + 1) + __mul24(RW, i + 1);
        const T *Rp = R + ((p - 1) & 1) * RS;
        T *Rc = R + (p & 1) * RS;
        T r = T(0);
        r = ((T)rn + r) * flpr;
        if (CLASSIC) {
            r = fms(r, flw, Rp[c1 - RW]);
            r = fms(r, fls, Rp[c1 - 1]);
            r = fms(r, flb, Rp[c1]);
        } else {
            r = fms(r, flb, Rp[c1]);
            r = fms(r, fls, Rp[c1 - 1]);
            r = fms(r, flw, Rp[c1 - RW]);
        }
        Rc[c1] = r;
        st(res, s, r);
    }
    // RES of a boundary cell into the lower or upper planes (as k_forward)
    __device__ __forceinline__ T face_get(uint32_t f, int p) const {
        return R[(p & 1) * RS + face_src(f)];
    }
    __device__ __forceinline__ void face_put(uint32_t f, T v) const {
        T *dst = face_hi(f) ? rw : rlo;
        if (dst) st(dst, face_dst(f), v);
    }
    __device__ __forceinline__ void store(uint32_t e, T v, uint8_t b, int p) const {
        const int x = e & 0xfff;
        if (e & 0x8000u) {
            R[(p & 1) * RS + x] = v;
        } else {
            const int sl = win_slot<NS>(p + 2) * mHS;
            H[sl + x] = v;
            if (MASKED && masked) BH[sl + x] = b;
        }
    }
};

template <typename T, bool CLASSIC, bool MASKED, int D, int M, int NW, bool LATE = false>
__global__ __launch_bounds__(kWave * NW) void k_outer_fwd(
    const uint2 *__restrict__ items, Topo t, Fields<T> F, Mask m, Outer o,
    T *rhs_t, int rd, int write_hi) {
    if (*o.done) return;
    extern __shared__ __align__(16) unsigned char lds_of[];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int mRS = (t.max_nj + 1) * (t.max_ni + 1);
    const int mHS = (t.max_nj + 2) * (t.max_ni + 2);
    SrcTab<T> *tab = (SrcTab<T> *)lds_of;
    double *cf = (double *)(lds_of + sizeof(SrcTab<T>));
    const int nr = row_len(t);
    double *ar = cf + 6 * nr;
    T *R = (T *)(ar + (MASKED ? 0 : 3 * nr));
    T *H = R + 2 * mRS;
    constexpr int NS = win_slots<T, NW>();
    uint8_t *BH = (uint8_t *)(H + NS * mHS);
    // window sources (HILF and its face ghosts) and the recurrence sources
    // (the neighbours' upper RES planes feed faces 0, 2, 4)
    fill_tab<T>(tab, t, B, F.hface, F.hpar, F.hilf, F.rhi[rd], 0x15, m);
    coef_rows<double>(cf, o.aw, o.ae, o.as, o.an, o.ab, o.at, B, nr);
    if (!MASKED) ap_rows(ar, F.apx, F.apy, F.apz, B, nr);
    if (NW > 1) __syncthreads(); else wave_sync();
    const int NI = B.ni, NJ = B.nj;
    const int64_t wo = B.wo, fc = B.fc;
    OutFwdK<T, CLASSIC, MASKED, NS, LATE> k;
    k.cells = t.cellx + S.cell_off;
    k.ar = ar;
    k.wf.load(B);
    k.ap = o.ap64 + wo;
    k.rhs = o.rhs64 + wo;
    k.dp = o.dp64 + wo;
    k.rhs_t = rhs_t ? rhs_t + wo : nullptr;
    k.cf = cf;
    k.lw = F.f[0] + B.fo;
    k.ls = F.f[1] + B.fo;
    k.lb = F.f[2] + B.fo;
    k.lpr = F.f[3] + B.fo;
    k.res = F.res + wo;
    k.rlo = F.rlo + fc;
    k.rw = write_hi ? F.rhi[rd ^ 1] + fc : nullptr;
    k.R = R;
    k.H = H;
    k.BH = BH;
    k.HW = NJ + 2;
    k.RW = NJ + 1;
    k.RS = (NJ + 1) * (NI + 1);
    k.mHS = mHS;
    k.nr = nr;
    k.masked = MASKED && !(B.flags & 1);
    k.mx = 0.0;
    constexpr int nw = plan_nw<NW>();
    const int ib = S.plan_item[nw][M - 1][kPlanFwdLap] + wave_index();
    run_plan<D, M, NW, MASKED, T>(k, items + ib, S.plan_nitem[nw][M - 1][kPlanFwdLap],
                                  t.ents + S.plan_ent[kPlanFwdLap], tab,
                                  t.fents + (int64_t)ib * kWave);
    double mx = k.mx * ldexp(1.0, t.maxlevel - B.level);
    for (int off = 32; off > 0; off >>= 1) {
        const double x = __shfl_down(mx, off, kWave);
        mx = (x > mx || x != x) ? x : mx;   // a NaN from either side stays
    }
    if ((threadIdx.x & (kWave - 1)) == 0)   // one per wave
        atomicMax(o.gmax, (unsigned long long)__double_as_longlong(mx));
}

template <typename T, bool CLASSIC, bool MASKED>
static void ofw_launch(const Topo &t, const Fields<T> &f, Mask m, Outer o, T *rhs_t,
                       int rd, int write_hi, size_t lds, size_t wslot, hipStream_t s) {
#define MGP_OFW0(D, M, NW, L) \
    k_outer_fwd<T, CLASSIC, MASKED, D, M, NW, L><<<t.nlist, kWave * NW, \
        lds + win_slots<T, NW>() * wslot, s>>>( \
        t.items, t, f, m, o, rhs_t, rd, write_hi)
    // FP64 with factors shared between blocks: late factor loads
#define MGP_OFW(D, M, NW) \
    if constexpr (sizeof(T) == 8 && !MASKED) { \
        if (f.aps) { \
            MGP_OFW0(D, M, NW, true); \
            break; \
        } \
    } \
    MGP_OFW0(D, M, NW, false)
    switch (plan_shape()) {
    case 1: MGP_OFW(4, 1, 1); break;
    case 3: MGP_OFW(3, 2, 1); break;
    case 4: MGP_OFW(6, 2, 1); break;
    case 5: MGP_OFW(8, 1, 1); break;
    case 6: MGP_OFW(4, 1, 4); break;
    case 7: MGP_OFW(4, 1, 8); break;
    case 8: MGP_OFW(3, 2, 2); break;
    case 9: MGP_OFW(4, 1, 2); break;
    default: MGP_OFW(2, 2, 1); break;
    }
#undef MGP_OFW
#undef MGP_OFW0
}

template <typename T>
void launch_outer_fwd(const Topo &t, const Fields<T> &f, Mask m, Outer o, T *rhs_t,
                      int rd, int write_hi, int classic, hipStream_t s) {
    MGP_RANGE("mgp:outer_fwd");
    const size_t mRS = (size_t)(t.max_nj + 1) * (t.max_ni + 1);
    const size_t mHS = (size_t)(t.max_nj + 2) * (t.max_ni + 2);
    const bool masked = m.bp != nullptr;
    const size_t nr = row_len(t);
    const size_t lds = sizeof(SrcTab<T>) + 6 * nr * sizeof(double) +
                       (masked ? 0 : 3 * nr * sizeof(double)) + 2 * mRS * sizeof(T);
    const size_t wslot = mHS * sizeof(T) + (masked ? mHS : 0);
    if (classic) {
        if (masked) ofw_launch<T, true, true>(t, f, m, o, rhs_t, rd, write_hi, lds, wslot, s);
        else ofw_launch<T, true, false>(t, f, m, o, rhs_t, rd, write_hi, lds, wslot, s);
    } else {
        if (masked) ofw_launch<T, false, true>(t, f, m, o, rhs_t, rd, write_hi, lds, wslot, s);
        else ofw_launch<T, false, false>(t, f, m, o, rhs_t, rd, write_hi, lds, wslot, s);
    }
}
template void launch_outer_fwd<double>(const Topo &, const Fields<double> &, Mask, Outer,
                                       double *, int, int, int, hipStream_t);
template void launch_outer_fwd<float>(const Topo &, const Fields<float> &, Mask, Outer,
                                      float *, int, int, int, hipStream_t);

}  // namespace mgp
