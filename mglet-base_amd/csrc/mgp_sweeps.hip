// mgp_sweeps.hip -- the plane kernels of the pressure iteration on gfx950:
// the SIP forward sweep with the fused Laplacian, the SIP backward sweep,
// red-black SOR and the FP64 defect update.
//
// One wave64 owns one block and walks its i+j+k hyperplanes (the
// reference's dependency order, src/flow/sip_hyperplane_mod.F90:450-569).
// A single wave needs no workgroup barrier between planes: its LDS
// accesses are processed in program order.
//
// The work of a block is a host-built plan (mgp_api.cpp, build_plans): a
// sequence of items, each one 64-lane chunk of a hyperplane to compute and
// one 64-lane chunk of entries to move into LDS (window values of HILF
// with their ghosts, recurrence ghosts from the neighbours' face buffers).
// The engine keeps D items of global loads in flight: while item n is
// computed, the loads of items n+1 .. n+D-1 are outstanding, and the entry
// words of items n+D .. n+2D-1 as well.  Every item issues the same
// number of loads (lanes beyond a chunk read a clamped, valid address), so
// the compiler's wait counts stay exact across the unrolled loop.
//
// LDS per wave: source table, 1-D coefficient rows, the pencil arrays R
// (two planes of the recurrence, indexed [i][j]) and the HILF window ring
// H (4 planes, 8 for SOR, indexed [i+1][j+1], with BP bytes when masked).
// All arithmetic follows the reference's association order (the build
// uses -ffp-contract=off), so the FP64 path is bitwise equal to the oracle.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "mgp_device.h"
#include "mgp_launch.h"
#include "mgp_plane.h"

namespace mgp {

// ---------------------------------------------------------------------
// Fused inner-iteration first half: RES = -(A HILF) (laplacephi_level),
// then sipiter1 (src/flow/sip_hyperplane_mod.F90:450-505, classic order
// src/flow/sip_classic_mod.F90:143-166).  Low-side ghosts of RES are the
// neighbours' upper planes of the previous exchange (rhi[rd]); the upper
// planes of this sweep go to rhi[1-rd] unless this is the last inner
// iteration (the reference's conn(..., forward=-1) then skips faces
// 1/3/5), the lower planes always go to rlo for the backward sweep.
// LAP = false is the first inner iteration of an outer iteration, where
// HILF = 0 and the Laplacian term vanishes.  LATE: the factors and GSAP,
// which L2 serves when every block shares one copy, are loaded when the
// item gathers its LDS values instead of D items ahead, which frees 10
// (FP32) or 20 (FP64) registers per pipeline slot (on 32768 blocks of
// 24^3 the multi-wave FP32 shapes run 2-2.5 % faster, FP64 with 4 waves of
// 64-cell items or 2 waves of 128-cell items 16-19 %; from HBM, without
// sharing, they would stall instead).
template <typename T, bool CLASSIC, bool MASKED, bool LAP, int NS, bool LATE = false>
struct FwdK {
    // launches without blocked cells take GSAP from the copy shared with the
    // factors (an L2 hit), or rebuild it from its separable part when no
    // factors are shared (saves an HBM load)
    static constexpr bool APC = LAP && !MASKED;
    static constexpr bool kFaces = true;
    static constexpr bool kSplit = true;
    static constexpr bool kFar = false;
    struct Far {};
    struct Early {
        T lpr, lb, ls, lw, ap;
    };
    struct Late {};
    struct Data : std::conditional_t<LATE, Late, Early> {
        uint32_t c;
        T rhs;
    };
    // LDS values of one cell (gather), then its arithmetic and stores
    // (finish), so the engine can issue two cells' reads before waiting
    struct G {
        Stencil<T, T, MASKED> st;
        T w, e, s, n, b, t, ap;   // 1-D coefficients, GSAP
        T rk, rj, ri;             // recurrence neighbours k-1, j-1, i-1
        T lpr, lb, ls, lw;        // factors
        int c1;
    };
    const uint32_t *cells;
    const T *rhs, *lpr, *lb, *ls, *lw, *ap;
    T *res, *rlo, *rw;
    const T *cf;
    const double *ar;
    WallFold wf;
    T *R, *H;
    uint8_t *BH;
    int NI, NJ, NK, HW, RW, RS, mHS, nr;
    bool masked, apsh;

    // every item issues the same loads (the GSAP load reads zeros when the
    // value is rebuilt instead), so the wait counts stay exact
    __device__ __forceinline__ void load(Data &d, int s) const {
        d.c = ld(cells, s);
        d.rhs = ld(rhs, s);
        if constexpr (!LATE) {
            d.lpr = ld(lpr, s);
            d.lb = ld(lb, s);
            d.ls = ld(ls, s);
            d.lw = ld(lw, s);
            if (LAP) d.ap = ld(ap, s);
        }
    }
    __device__ __forceinline__ G gather(const Data &d, int s, int p) const {
        G g;
        T apl = T(0);
        if constexpr (LATE) {
            g.lpr = ld(lpr, s);
            g.lb = ld(lb, s);
            g.ls = ld(ls, s);
            g.lw = ld(lw, s);
            if (LAP) apl = ld(ap, s);
        } else {
            g.lpr = d.lpr;
            g.lb = d.lb;
            g.ls = d.ls;
            g.lw = d.lw;
            if (LAP) apl = d.ap;
        }
        int i, j, k;
        cell_ijk(d.c, i, j, k);
        if (LAP) {
            // LDS positions stay below 2^24: full-rate 24-bit multiplies
            const int c0 = (j + 1) + __mul24(HW, i + 1);
            const int sm = win_slot<NS>(p - 1) * mHS, s0 = win_slot<NS>(p) * mHS,
                      sp = win_slot<NS>(p + 1) * mHS;
            const Coef6<T> co(cf, nr, i, j, k);
            g.w = co.w; g.e = co.e; g.s = co.s; g.n = co.n; g.b = co.b; g.t = co.t;
            g.ap = (APC && !apsh) ? (T)ap_cell(ar, nr, i, j, k, cell_flags(d.c), wf) : apl;
            g.st.read(H + sm, H + s0, H + sp, BH + sm, BH + s0, BH + sp, c0, HW,
                      MASKED && masked);
        }
        g.c1 = (j + 1) + __mul24(RW, i + 1);
        const T *Rp = R + ((p - 1) & 1) * RS;
        g.rk = Rp[g.c1];
        g.rj = Rp[g.c1 - 1];
        g.ri = Rp[g.c1 - RW];
        return g;
    }
    __device__ __forceinline__ void finish(const Data &d, const G &g, int s, int p, int,
                                           bool ok) const {
        T r = T(0);
        if (LAP) r = g.st.eval(g.w, g.e, g.s, g.n, g.b, g.t, g.ap);
        r = (d.rhs + r) * g.lpr;
        if (CLASSIC) {
            r = fms(r, g.lw, g.ri);
            r = fms(r, g.ls, g.rj);
            r = fms(r, g.lb, g.rk);
        } else {
            r = fms(r, g.lb, g.rk);
            r = fms(r, g.ls, g.rj);
            r = fms(r, g.lw, g.ri);
        }
        if (ok) {
            R[(p & 1) * RS + g.c1] = r;
            st(res, s, r);
        }
    }
    __device__ __forceinline__ void compute(const Data &d, int s, int p, int l) const {
        finish(d, gather(d, s, p), s, p, l, true);
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
    // RES of a boundary cell (in the step's recurrence plane) into the
    // lower planes, or the upper planes unless this is the last inner
    // iteration
    __device__ __forceinline__ T face_get(uint32_t f, int p) const {
        return R[(p & 1) * RS + face_src(f)];
    }
    __device__ __forceinline__ void face_put(uint32_t f, T v) const {
        T *dst = face_hi(f) ? rw : rlo;
        if (dst) st(dst, face_dst(f), v);
    }
};

template <typename T, bool CLASSIC, bool MASKED, bool LAP, int D, int M, int NW,
          bool LATE = false>
__global__ __launch_bounds__(kWave * NW) void k_forward(
    const uint2 *__restrict__ items, Topo t, Fields<T> F, Mask m, int rd,
    int write_hi, const int *done) {
    if (*done) return;
    extern __shared__ __align__(16) unsigned char lds_fw[];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int mRS = (t.max_nj + 1) * (t.max_ni + 1);
    const int mHS = (t.max_nj + 2) * (t.max_ni + 2);
    constexpr int NS = win_slots<T, NW>();
    constexpr bool APC = FwdK<T, CLASSIC, MASKED, LAP, NS>::APC;
    SrcTab<T> *tab = (SrcTab<T> *)lds_fw;
    T *cf = (T *)(lds_fw + sizeof(SrcTab<T>));
    const int nr = row_len(t);
    double *ar = (double *)(cf + 6 * nr);
    T *R = (T *)(ar + (APC ? 3 * nr : 0));
    T *H = R + 2 * mRS;
    uint8_t *BH = (uint8_t *)(H + NS * mHS);
    // R sources: the neighbours' upper planes feed faces 0, 2, 4
    fill_tab<T>(tab, t, B, LAP ? F.hface : nullptr, F.hpar, F.hilf, F.rhi[rd],
                0x15, m);
    if (LAP) coef_rows<T>(cf, F.aw, F.ae, F.as, F.an, F.ab, F.at, B, nr);
    if (APC) ap_rows(ar, F.apx, F.apy, F.apz, B, nr);
    if (NW > 1) __syncthreads(); else wave_sync();
    const int NI = B.ni, NJ = B.nj, NK = B.nk;
    const int64_t wo = B.wo, fc = B.fc;
    FwdK<T, CLASSIC, MASKED, LAP, NS, LATE> k;
    k.cells = t.cellx + S.cell_off;
    k.rhs = F.rhs + wo;
    k.lw = F.f[0] + B.fo;
    k.ls = F.f[1] + B.fo;
    k.lb = F.f[2] + B.fo;
    k.lpr = F.f[3] + B.fo;
    k.apsh = APC && F.aps != nullptr;
    k.ap = !APC ? F.ap + wo : (F.aps ? F.aps + B.fo : (const T *)t.zero);
    k.res = F.res + wo;
    k.rlo = F.rlo + fc;
    k.rw = write_hi ? F.rhi[rd ^ 1] + fc : nullptr;
    k.cf = cf;
    k.ar = ar;
    k.wf.load(B);
    k.R = R;
    k.H = H;
    k.BH = BH;
    k.NI = NI;
    k.NJ = NJ;
    k.NK = NK;
    k.HW = NJ + 2;
    k.RW = NJ + 1;
    k.RS = (NJ + 1) * (NI + 1);
    k.mHS = mHS;
    k.nr = nr;
    k.masked = MASKED && !(B.flags & 1);
    const int kind = LAP ? kPlanFwdLap : kPlanFwdZero;
    constexpr int nw = plan_nw<NW>();
    const int ib = S.plan_item[nw][M - 1][kind] + wave_index();
    run_plan<D, M, NW, MASKED && LAP, T>(k, items + ib, S.plan_nitem[nw][M - 1][kind],
                                         t.ents + S.plan_ent[kind], tab,
                                         t.fents + (int64_t)ib * kWave);
}

// ---------------------------------------------------------------------
// Inner-iteration second half: sipiter2 (src/flow/sip_hyperplane_mod.F90:
// 508-569; classic order src/flow/sip_classic_mod.F90:169-198), then
// HILF += RES.  Upper ghosts of RES are the neighbours' lower planes of
// this iteration's forward sweep.  Also writes HILF's boundary planes,
// which stand in for the reference's conn(ilevel, 1, s1=dp) (:414).
// HZ: HILF is zero on entry (first inner iteration of an outer iteration).
template <typename T, bool CLASSIC, bool HZ, bool LATE = false>
struct BwdK {
    static constexpr bool kFaces = true;
    static constexpr bool kSplit = true;
    // RES and HILF come from HBM, the cell words and factors (shared) from
    // L2: the former get the engine's far prefetch.  LATE (FP64 with
    // factors shared between blocks): the factors load when the item
    // gathers (best FP64 shape 3.89 -> 3.80 ms on 32768 blocks of 24^3;
    // FP32 measured slower that way)
    static constexpr bool kFar = true;
    struct Early {
        T ue, un, ut;
    };
    struct Late {};
    struct Data : std::conditional_t<LATE, Late, Early> {
        uint32_t c;
        T res, h;
    };
    struct Far {
        T res, h;
    };
    struct G {
        T rt, rn, re;   // recurrence neighbours k+1, j+1, i+1
        T ue, un, ut;   // factors
        int c1;
    };
    const uint32_t *cells;
    const T *res, *ue, *un, *ut;
    T *hilf, *hface;
    T *R, *stg;   // recurrence planes; this wave's staging row of HILF
    int NI, NJ, NK, RW, RS;

    __device__ __forceinline__ void load(Data &d, int s) const {
        d.c = ld(cells, s);
        if constexpr (!LATE) {
            d.ue = ld(ue, s);
            d.un = ld(un, s);
            d.ut = ld(ut, s);
        }
    }
    __device__ __forceinline__ void far_load(Far &f, int s) const {
        f.res = ld(res, s);
        f.h = HZ ? T(0) : ld((const T *)hilf, s);
    }
    __device__ __forceinline__ void merge(Data &d, const Far &f) const {
        d.res = f.res;
        d.h = f.h;
    }
    __device__ __forceinline__ G gather(const Data &d, int s, int p) const {
        G g;
        if constexpr (LATE) {
            g.ue = ld(ue, s);
            g.un = ld(un, s);
            g.ut = ld(ut, s);
        } else {
            g.ue = d.ue;
            g.un = d.un;
            g.ut = d.ut;
        }
        int i, j, k;
        cell_ijk(d.c, i, j, k);
        g.c1 = j + __mul24(RW, i);
        const T *Rp = R + ((p + 1) & 1) * RS;
        g.rt = Rp[g.c1];
        g.rn = Rp[g.c1 + 1];
        g.re = Rp[g.c1 + RW];
        return g;
    }
    __device__ __forceinline__ void finish(const Data &d, const G &g, int s, int p, int l,
                                           bool ok) const {
        T r = d.res;
        if (CLASSIC) {
            r = fms(r, g.un, g.rn);
            r = fms(r, g.ue, g.re);
            r = fms(r, g.ut, g.rt);
        } else {
            r = fms(r, g.ut, g.rt);
            r = fms(r, g.un, g.rn);
            r = fms(r, g.ue, g.re);
        }
        const T h = (HZ ? T(0) : d.h) + r;
        if (ok) {
            R[(p & 1) * RS + g.c1] = r;
            st(hilf, s, h);
            stg[l] = h;
        }
    }
    __device__ __forceinline__ void compute(const Data &d, int s, int p, int l) const {
        finish(d, gather(d, s, p), s, p, l, true);
    }
    __device__ __forceinline__ void store(uint32_t e, T v, uint8_t, int p) const {
        R[(p & 1) * RS + (e & 0xfff)] = v;
    }
    // HILF of a boundary cell into the block's boundary planes
    __device__ __forceinline__ T face_get(uint32_t f, int) const { return stg[face_src(f)]; }
    __device__ __forceinline__ void face_put(uint32_t f, T v) const { st(hface, face_dst(f), v); }
};

template <typename T, bool CLASSIC, bool HZ, int D, int M, int NW, bool LATE = false>
__global__ __launch_bounds__(kWave * NW) void k_backward(
    const uint2 *__restrict__ items, Topo t, Fields<T> F, const int *done) {
    if (*done) return;
    extern __shared__ __align__(16) unsigned char lds_bw[];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    SrcTab<T> *tab = (SrcTab<T> *)lds_bw;
    T *R = (T *)(lds_bw + sizeof(SrcTab<T>));
    const int mRS = (t.max_nj + 1) * (t.max_ni + 1);
    // R sources: the neighbours' lower planes feed faces 1, 3, 5
    fill_tab<T>(tab, t, B, nullptr, nullptr, F.hilf, F.rlo, 0x2a, Mask{});
    if (NW > 1) __syncthreads(); else wave_sync();
    const int NI = B.ni, NJ = B.nj, NK = B.nk;
    const int64_t wo = B.wo;
    BwdK<T, CLASSIC, HZ, LATE> k;
    k.cells = t.cellx + S.cell_off;
    k.res = F.res + wo;
    k.ue = F.f[4] + B.fo;
    k.un = F.f[5] + B.fo;
    k.ut = F.f[6] + B.fo;
    k.hilf = F.hilf + wo;
    k.hface = F.hface + B.fc;
    k.R = R;
    k.stg = R + 2 * mRS + (threadIdx.x / kWave) * (kWave * M);
    k.NI = NI;
    k.NJ = NJ;
    k.NK = NK;
    k.RW = NJ + 1;
    k.RS = (NJ + 1) * (NI + 1);
    constexpr int nw = plan_nw<NW>();
    const int ib = S.plan_item[nw][M - 1][kPlanBwd] + wave_index();
    run_plan<D, M, NW, false, T>(k, items + ib, S.plan_nitem[nw][M - 1][kPlanBwd],
                                 t.ents + S.plan_ent[kPlanBwd], tab,
                                 t.fents + (int64_t)ib * kWave);
}

// ---------------------------------------------------------------------
// Red-black SOR of HILF (relax, src/flow/sor_mod.F90:102-203), both colours
// in one pass over the block's hyperplanes.  The reference's first colour
// (rb = 0: i+j+k even in its 1-based indices, :124-130) is exactly the set
// of odd planes p = i+j+k of the 0-based interior, the second colour the
// even planes, and every 7-point neighbour of a cell lies on plane p-1 or
// p+1.  Step p relaxes plane p when p is odd (its neighbours are old even
// planes) and plane p-2 when p is even (its neighbours p-3, p-1 were
// relaxed at earlier steps), so each value is read and written once; the
// window ring holds 8 planes for that.  Ghost values stay those of the
// previous exchange (fin) for both colours, as in the reference, where
// conn(DP) follows relax (:414); the new boundary planes go to fout.  With
// an odd cell count along k the loop bounds of :124-130 never reach the
// cells k = nk-1; they keep their value.
// HZ: HILF is zero on entry (first inner iteration of an outer iteration).
template <typename T, bool MASKED, bool HZ>
struct SorK {
    static constexpr bool kFaces = true;
    static constexpr bool kSplit = false;
    static constexpr bool kFar = false;
    struct Far {};
    struct Data {
        uint32_t c;
        T rhs, rap;
    };
    const uint32_t *cells;
    const T *rhs, *rap;
    T *hilf, *fout;
    const T *cf;
    const double *ar;
    WallFold wf;
    T *H;
    uint8_t *BH;
    int NI, NJ, NK, HW, mHS, nr;
    T om, om1;
    bool masked, keep_top;

    __device__ __forceinline__ void load(Data &d, int s) const {
        d.c = ld(cells, s);
        d.rhs = ld(rhs, s);
        if (MASKED) d.rap = ld(rap, s);
    }
    __device__ __forceinline__ void compute(const Data &d, int s, int p, int) const {
        const int u = (p & 1) ? p : p - 2;
        int i, j, k;
        cell_ijk(d.c, i, j, k);
        const int c0 = (j + 1) + __mul24(HW, i + 1);
        T *H0 = H + ((u + 1) & 7) * mHS;
        const T *Hm = H + (u & 7) * mHS, *Hp = H + ((u + 2) & 7) * mHS;
        const T old = H0[c0];
        T nv = old;
        if (!(keep_top && k == NK - 1)) {
            const Coef6<T> co(cf, nr, i, j, k);
            const T pw = Hm[c0 - HW], pe = Hp[c0 + HW];
            const T psv = Hm[c0 - 1], pn = Hp[c0 + 1];
            const T pb = Hm[c0], pt = Hp[c0];
            T r;
            if (MASKED && masked) {
                const uint8_t *B0 = BH + ((u + 1) & 7) * mHS;
                const uint8_t *Bm = BH + (u & 7) * mHS, *Bp = BH + ((u + 2) & 7) * mHS;
                const T bc = (T)B0[c0];
                const T cw = (T)Bm[c0 - HW] * bc * co.w;
                const T ce = bc * (T)Bp[c0 + HW] * co.e;
                const T cs = (T)Bm[c0 - 1] * bc * co.s;
                const T cn = bc * (T)Bp[c0 + 1] * co.n;
                const T cb = (T)Bm[c0] * bc * co.b;
                const T ct = bc * (T)Bp[c0] * co.t;
                r = cw * pw + ce * pe + cs * psv + cn * pn + cb * pb + ct * pt;
            } else {
                r = co.w * pw + co.e * pe + co.s * psv + co.n * pn + co.b * pb + co.t * pt;
            }
            T rap;
            if (MASKED) {
                rap = d.rap;
            } else {
                // SOR_RAP = divide0(1, GSAP), src/flow/sor_mod.F90:18-49
                const double a = ap_cell(ar, nr, i, j, k, cell_flags(d.c), wf);
                rap = (T)(a == 0.0 ? 0.0 : 1.0 / a);
            }
            r = (r - d.rhs) * rap;
            nv = om1 * old - om * r;
        }
        H0[c0] = nv;
        st(hilf, s, nv);
    }
    // the relaxed value of a boundary cell (in its window slot) into the
    // new boundary planes
    __device__ __forceinline__ T face_get(uint32_t f, int p) const {
        const int u = (p & 1) ? p : p - 2;
        return H[((u + 1) & 7) * mHS + face_src(f)];
    }
    __device__ __forceinline__ void face_put(uint32_t f, T v) const { st(fout, face_dst(f), v); }
    __device__ __forceinline__ void store(uint32_t e, T v, uint8_t b, int p) const {
        const int x = e & 0xfff;
        const int sl = ((p + 3) & 7) * mHS;
        H[sl + x] = HZ ? T(0) : v;
        if (MASKED && masked) BH[sl + x] = b;
    }
};

template <typename T, bool MASKED, bool HZ, int D, int M>
__global__ __launch_bounds__(kWave) void k_sor(
    const uint2 *__restrict__ items, Topo t, Fields<T> F, Mask m, const T *fin,
    T *fout, double omega, const int *done) {
    if (*done) return;
    extern __shared__ __align__(16) unsigned char lds_so[];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int mHS = (t.max_nj + 2) * (t.max_ni + 2);
    SrcTab<T> *tab = (SrcTab<T> *)lds_so;
    T *cf = (T *)(lds_so + sizeof(SrcTab<T>));
    const int nr = row_len(t);
    double *ar = (double *)(cf + 6 * nr);
    T *H = (T *)(ar + (MASKED ? 0 : 3 * nr));
    uint8_t *BH = (uint8_t *)(H + 8 * mHS);
    fill_tab<T>(tab, t, B, fin, F.hpar, F.hilf, nullptr, 0, m);
    coef_rows<T>(cf, F.aw, F.ae, F.as, F.an, F.ab, F.at, B, nr);
    if (!MASKED) ap_rows(ar, F.apx, F.apy, F.apz, B, nr);
    wave_sync();
    const int NI = B.ni, NJ = B.nj, NK = B.nk;
    const int64_t wo = B.wo;
    SorK<T, MASKED, HZ> k;
    k.cells = t.cellx + S.cell_off;
    k.rhs = F.rhs + wo;
    k.rap = F.rap + wo;
    k.hilf = F.hilf + wo;
    k.fout = fout + B.fc;
    k.cf = cf;
    k.ar = ar;
    k.wf.load(B);
    k.H = H;
    k.BH = BH;
    k.NI = NI;
    k.NJ = NJ;
    k.NK = NK;
    k.HW = NJ + 2;
    k.mHS = mHS;
    k.nr = nr;
    k.om = (T)omega;
    k.om1 = (T)(1.0 - omega);
    k.masked = MASKED && !(B.flags & 1);
    k.keep_top = (NK & 1) != 0;
    const int ib = S.plan_item[0][M - 1][kPlanSor];
    run_plan<D, M, 1, MASKED, T>(k, items + ib, S.plan_nitem[0][M - 1][kPlanSor],
                                 t.ents + S.plan_ent[kPlanSor], tab,
                                 t.fents + (int64_t)ib * kWave);
}

// ---------------------------------------------------------------------
// FP64 outer update (src/flow/pressuresolver_mod.F90:258-279): RES =
// -(A HILF) with the FP64 operator, RHS += RES, max |RHS|, DP += HILF.
// In the mixed-precision mode the sweeps' RHS copy is refreshed here.
template <typename T, bool MASKED, int NS>
struct OutK {
    static constexpr bool kFaces = false;
    static constexpr bool kSplit = false;
    static constexpr bool kFar = false;
    struct Far {};
    struct Data {
        uint32_t c;
        double ap, rhs, dp;
    };
    const uint32_t *cells;
    const double *ar;
    WallFold wf;
    const double *ap;
    double *rhs, *dp;
    T *rhs_t;
    const double *cf;
    T *H;
    uint8_t *BH;
    int HW, mHS, nr;
    bool masked;
    double mx;

    __device__ __forceinline__ void load(Data &d, int s) const {
        d.c = ld(cells, s);
        if (MASKED) d.ap = ld(ap, s);
        d.rhs = ld((const double *)rhs, s);
        d.dp = ld((const double *)dp, s);
    }
    __device__ __forceinline__ void compute(const Data &d, int s, int p, int) {
        int i, j, k;
        cell_ijk(d.c, i, j, k);
        const int c0 = (j + 1) + __mul24(HW, i + 1);
        const int sm = win_slot<NS>(p - 1) * mHS, s0 = win_slot<NS>(p) * mHS,
                  sp = win_slot<NS>(p + 1) * mHS;
        const double apv = MASKED ? d.ap : ap_cell(ar, nr, i, j, k, cell_flags(d.c), wf);
        const Coef6<double> co(cf, nr, i, j, k);
        double r;
        if (MASKED && masked)
            r = laplace_cell<T, double, true>(H + sm, H + s0, H + sp, BH + sm, BH + s0,
                                              BH + sp, c0, HW, co.w, co.e, co.s, co.n, co.b, co.t,
                                              apv);
        else
            r = laplace_cell<T, double, false>(H + sm, H + s0, H + sp, nullptr, nullptr,
                                               nullptr, c0, HW, co.w, co.e, co.s, co.n, co.b, co.t,
                                               apv);
        const double rn = d.rhs + r;
        st(rhs, s, rn);
        if (rhs_t) st(rhs_t, s, (T)rn);
        const double a = fabs(rn);
        mx = (a > mx || a != a) ? a : mx;
        st(dp, s, d.dp + (double)H[s0 + c0]);
    }
    __device__ __forceinline__ void store(uint32_t e, T v, uint8_t b, int p) const {
        const int x = e & 0xfff;
        const int sl = win_slot<NS>(p + 2) * mHS;
        H[sl + x] = v;
        if (MASKED && masked) BH[sl + x] = b;
    }
};

template <typename T, bool MASKED, int D, int M, int NW>
__global__ __launch_bounds__(kWave * NW) void k_outer(
    const uint2 *__restrict__ items, Topo t, Fields<T> F, Mask m, Outer o,
    T *rhs_t) {
    if (*o.done) return;
    extern __shared__ __align__(16) unsigned char lds_ou[];
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const int mHS = (t.max_nj + 2) * (t.max_ni + 2);
    SrcTab<T> *tab = (SrcTab<T> *)lds_ou;
    double *cf = (double *)(lds_ou + sizeof(SrcTab<T>));
    const int nr = row_len(t);
    double *ar = cf + 6 * nr;
    constexpr int NS = win_slots<T, NW>();
    T *H = (T *)(ar + (MASKED ? 0 : 3 * nr));
    uint8_t *BH = (uint8_t *)(H + NS * mHS);
    fill_tab<T>(tab, t, B, F.hface, F.hpar, F.hilf, nullptr, 0, m);
    coef_rows<double>(cf, o.aw, o.ae, o.as, o.an, o.ab, o.at, B, nr);
    if (!MASKED) ap_rows(ar, F.apx, F.apy, F.apz, B, nr);
    if (NW > 1) __syncthreads(); else wave_sync();
    const int64_t wo = B.wo;
    OutK<T, MASKED, NS> k;
    k.cells = t.cellx + S.cell_off;
    k.ar = ar;
    k.wf.load(B);
    k.ap = o.ap64 + wo;
    k.rhs = o.rhs64 + wo;
    k.dp = o.dp64 + wo;
    k.rhs_t = rhs_t ? rhs_t + wo : nullptr;
    k.cf = cf;
    k.H = H;
    k.BH = BH;
    k.HW = B.nj + 2;
    k.mHS = mHS;
    k.nr = nr;
    k.masked = MASKED && !(B.flags & 1);
    k.mx = 0.0;
    constexpr int nw = plan_nw<NW>();
    run_plan<D, M, NW, MASKED, T>(k, items + S.plan_item[nw][M - 1][kPlanOut] + wave_index(),
                                  S.plan_nitem[nw][M - 1][kPlanOut], t.ents + S.plan_ent[kPlanOut],
                                  tab);
    double mx = k.mx * ldexp(1.0, t.maxlevel - B.level);
    for (int off = 32; off > 0; off >>= 1) {
        const double x = __shfl_down(mx, off, kWave);
        mx = (x > mx || x != x) ? x : mx;   // a NaN from either side stays
    }
    if ((threadIdx.x & (kWave - 1)) == 0)   // one per wave
        atomicMax(o.gmax, (unsigned long long)__double_as_longlong(mx));
}

// ---------------------------------------------------------------------
// Launchers.  The pipeline shape is a tuning knob, set_plan_shape(): 1 =
// 4 items of 64 cells in flight, 2 = 2 items of 128 cells (two cells per
// lane, independent, so their latencies overlap), 3 = 3 items of 128,
// 4 = 6 items of 128, 5 = 8 items of 64, all one wave per block; 6 and 7 =
// 4 items of 64 in flight per wave with 4 and 8 waves per block, which
// share each hyperplane's items and meet at an LDS barrier per plane.
// Launches with too few blocks to fill the chip are bound by one wave's
// instruction stream; the deep and the multi-wave shapes are for them.
// SOR relaxes HILF in place in global memory, which a prefetching wave
// must not race, so it runs one wave per block in every shape.
// The backward sweep has a shape of its own (set_plan_shapes), tuned apart:
// it holds fewer registers and no HILF window, so it can prefer deeper
// pipelines than the forward sweep and the outer update.
static int g_shape = 1, g_shape_bwd = 1;

void set_plan_shape(int v) { set_plan_shapes(v, v); }
void set_plan_shapes(int fwd, int bwd) {
    g_shape = (fwd >= 1 && fwd <= kNumShapes) ? fwd : 1;
    g_shape_bwd = (bwd >= 1 && bwd <= kNumShapes) ? bwd : 1;
}
int plan_shape() { return g_shape; }

template <typename T>
static size_t tab_bytes() { return sizeof(SrcTab<T>); }

// lds: the kernel's LDS without the HILF window ring; wslot: bytes of one
// ring slot (win_slots<T, NW>() of them, 0 without a window)
template <typename T, bool CLASSIC, bool MASKED, bool LAP>
static void fwd_launch(const Topo &t, const Fields<T> &f, Mask m, int rd,
                       int write_hi, const int *done, size_t lds, size_t wslot,
                       hipStream_t s) {
#define MGP_FWD1(D, M, NW) \
    k_forward<T, CLASSIC, MASKED, LAP, D, M, NW><<<t.nlist, kWave * NW, \
        lds + win_slots<T, NW>() * wslot, s>>>(t.items, t, f, m, rd, write_hi, done)
    // late factor loads for the multi-wave shapes when every block's
    // factors come from a shared copy (L2); measured slower for 8 waves,
    // and in FP64 for 64-cell items with 2 waves
#define MGP_FWDL(D, M, NW, OK) \
    if constexpr (OK) { \
        if (f.aps) { \
            k_forward<T, CLASSIC, MASKED, LAP, D, M, NW, true><<<t.nlist, kWave * NW, \
                lds + win_slots<T, NW>() * wslot, s>>>(t.items, t, f, m, rd, write_hi, done); \
            break; \
        } \
    } \
    MGP_FWD1(D, M, NW)
    constexpr bool kLate = !MASKED && LAP;
    constexpr bool kLate32 = kLate && sizeof(T) == 4;
    switch (g_shape) {
    case 1: MGP_FWD1(4, 1, 1); break;
    case 3: MGP_FWD1(3, 2, 1); break;
    case 4: MGP_FWD1(6, 2, 1); break;
    case 5: MGP_FWD1(8, 1, 1); break;
    case 6: MGP_FWDL(4, 1, 4, kLate); break;
    case 7: MGP_FWD1(4, 1, 8); break;
    case 8: MGP_FWDL(3, 2, 2, kLate); break;
    case 9: MGP_FWDL(4, 1, 2, kLate32); break;
    default: MGP_FWD1(2, 2, 1); break;
    }
#undef MGP_FWDL
#undef MGP_FWD1
}

template <typename T>
void launch_forward(const Topo &t, const Fields<T> &f, Mask m, int rd,
                    int write_hi, int hilf_zero, int classic, const int *done,
                    hipStream_t s) {
    MGP_RANGE("mgp:forward");
    const size_t mRS = (size_t)(t.max_nj + 1) * (t.max_ni + 1);
    const size_t mHS = (size_t)(t.max_nj + 2) * (t.max_ni + 2);
    const bool masked = m.bp != nullptr;
    const bool lap = !hilf_zero;
    const size_t nr = row_len(t);
    const size_t lds = tab_bytes<T>() + 6 * nr * sizeof(T) + 2 * mRS * sizeof(T) +
                       (lap && !masked ? 3 * nr * sizeof(double) : 0);
    const size_t wslot = lap ? mHS * sizeof(T) + (masked ? mHS : 0) : 0;
#define MGP_FWD(C, M, L) fwd_launch<T, C, M, L>(t, f, m, rd, write_hi, done, lds, wslot, s)
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

template <typename T, bool CLASSIC, bool HZ>
static void bwd_launch(const Topo &t, const Fields<T> &f, const int *done,
                       size_t lds, hipStream_t s) {
    // each wave stages its item's HILF values (64 M) for the face stores
#define MGP_BWD0(D, M, NW, L) \
    k_backward<T, CLASSIC, HZ, D, M, NW, L><<<t.nlist, kWave * NW, \
        lds + (size_t)NW * kWave * M * sizeof(T), s>>>(t.items, t, f, done)
    // FP64 with factors shared between blocks: late factor loads
#define MGP_BWD(D, M, NW) \
    if constexpr (sizeof(T) == 8) { \
        if (f.aps) { \
            MGP_BWD0(D, M, NW, true); \
            break; \
        } \
    } \
    MGP_BWD0(D, M, NW, false)
    switch (g_shape_bwd) {
    case 1: MGP_BWD(4, 1, 1); break;
    case 3: MGP_BWD(3, 2, 1); break;
    case 4: MGP_BWD(6, 2, 1); break;
    case 5: MGP_BWD(8, 1, 1); break;
    case 6: MGP_BWD(4, 1, 4); break;
    case 7: MGP_BWD(4, 1, 8); break;
    case 8: MGP_BWD(3, 2, 2); break;
    case 9: MGP_BWD(4, 1, 2); break;
    default: MGP_BWD(2, 2, 1); break;
    }
#undef MGP_BWD
#undef MGP_BWD0
}

template <typename T>
void launch_backward(const Topo &t, const Fields<T> &f, int hilf_zero,
                     int classic, const int *done, hipStream_t s) {
    MGP_RANGE("mgp:backward");
    const size_t mRS = (size_t)(t.max_nj + 1) * (t.max_ni + 1);
    const size_t lds = tab_bytes<T>() + 2 * mRS * sizeof(T);
    if (classic) {
        if (hilf_zero) bwd_launch<T, true, true>(t, f, done, lds, s);
        else bwd_launch<T, true, false>(t, f, done, lds, s);
    } else {
        if (hilf_zero) bwd_launch<T, false, true>(t, f, done, lds, s);
        else bwd_launch<T, false, false>(t, f, done, lds, s);
    }
}
template void launch_backward<double>(const Topo &, const Fields<double> &,
                                      int, int, const int *, hipStream_t);
template void launch_backward<float>(const Topo &, const Fields<float> &, int,
                                     int, const int *, hipStream_t);

template <typename T, bool MASKED, bool HZ>
static void sor_launch(const Topo &t, const Fields<T> &f, Mask m, const T *fin,
                       T *fout, double omega, const int *done, size_t lds,
                       hipStream_t s) {
    switch (g_shape) {
    case 1: k_sor<T, MASKED, HZ, 4, 1><<<t.nlist, kWave, lds, s>>>(t.items, t, f, m, fin, fout, omega, done); break;
    case 3:
    case 8: k_sor<T, MASKED, HZ, 3, 2><<<t.nlist, kWave, lds, s>>>(t.items, t, f, m, fin, fout, omega, done); break;
    case 4: k_sor<T, MASKED, HZ, 6, 2><<<t.nlist, kWave, lds, s>>>(t.items, t, f, m, fin, fout, omega, done); break;
    case 5: k_sor<T, MASKED, HZ, 8, 1><<<t.nlist, kWave, lds, s>>>(t.items, t, f, m, fin, fout, omega, done); break;
    case 6:   // several waves per block: SOR relaxes in place, one wave
    case 7:
    case 9: k_sor<T, MASKED, HZ, 4, 1><<<t.nlist, kWave, lds, s>>>(t.items, t, f, m, fin, fout, omega, done); break;
    default: k_sor<T, MASKED, HZ, 2, 2><<<t.nlist, kWave, lds, s>>>(t.items, t, f, m, fin, fout, omega, done); break;
    }
}

template <typename T>
void launch_sor(const Topo &t, const Fields<T> &f, Mask m, const T *fin,
                T *fout, double omega, int hilf_zero, const int *done,
                hipStream_t s) {
    MGP_RANGE("mgp:sor");
    const size_t mHS = (size_t)(t.max_nj + 2) * (t.max_ni + 2);
    const bool masked = m.bp != nullptr;
    const size_t nr = row_len(t);
    const size_t lds = tab_bytes<T>() + 6 * nr * sizeof(T) + 8 * mHS * sizeof(T) +
                       (masked ? 8 * mHS : 3 * nr * sizeof(double));
    if (masked) {
        if (hilf_zero) sor_launch<T, true, true>(t, f, m, fin, fout, omega, done, lds, s);
        else sor_launch<T, true, false>(t, f, m, fin, fout, omega, done, lds, s);
    } else {
        if (hilf_zero) sor_launch<T, false, true>(t, f, m, fin, fout, omega, done, lds, s);
        else sor_launch<T, false, false>(t, f, m, fin, fout, omega, done, lds, s);
    }
}
template void launch_sor<double>(const Topo &, const Fields<double> &, Mask,
                                 const double *, double *, double, int,
                                 const int *, hipStream_t);
template void launch_sor<float>(const Topo &, const Fields<float> &, Mask,
                                const float *, float *, double, int,
                                const int *, hipStream_t);

template <typename T, bool MASKED>
static void out_launch(const Topo &t, const Fields<T> &f, Mask m, Outer o,
                       T *rhs_t, size_t lds, size_t wslot, hipStream_t s) {
#define MGP_OUT(D, M, NW) \
    k_outer<T, MASKED, D, M, NW><<<t.nlist, kWave * NW, lds + win_slots<T, NW>() * wslot, s>>>( \
        t.items, t, f, m, o, rhs_t)
    switch (g_shape) {
    case 1: MGP_OUT(4, 1, 1); break;
    case 3: MGP_OUT(3, 2, 1); break;
    case 4: MGP_OUT(6, 2, 1); break;
    case 5: MGP_OUT(8, 1, 1); break;
    case 6: MGP_OUT(4, 1, 4); break;
    case 7: MGP_OUT(4, 1, 8); break;
    case 8: MGP_OUT(3, 2, 2); break;
    case 9: MGP_OUT(4, 1, 2); break;
    default: MGP_OUT(2, 2, 1); break;
    }
#undef MGP_OUT
}

template <typename T>
void launch_outer(const Topo &t, const Fields<T> &f, Mask m, Outer o,
                  T *rhs_t, hipStream_t s) {
    MGP_RANGE("mgp:outer");
    const size_t mHS = (size_t)(t.max_nj + 2) * (t.max_ni + 2);
    const bool masked = m.bp != nullptr;
    const size_t nr = row_len(t);
    const size_t lds = tab_bytes<T>() + 6 * nr * sizeof(double) +
                       (masked ? 0 : 3 * nr * sizeof(double));
    const size_t wslot = mHS * sizeof(T) + (masked ? mHS : 0);
    if (masked) out_launch<T, true>(t, f, m, o, rhs_t, lds, wslot, s);
    else out_launch<T, false>(t, f, m, o, rhs_t, lds, wslot, s);
}
template void launch_outer<double>(const Topo &, const Fields<double> &, Mask,
                                   Outer, double *, hipStream_t);
template void launch_outer<float>(const Topo &, const Fields<float> &, Mask,
                                  Outer, float *, hipStream_t);

}  // namespace mgp
