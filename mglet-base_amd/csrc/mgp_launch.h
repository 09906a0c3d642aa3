// mgp_launch.h -- host-side launchers of the gfx950 kernels in
// mgp_kernels.hip, called by the orchestration in mgp_api.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include "mgp_internal.h"

namespace mgp {

// Read-only per-level device view of the topology and shape tables.
struct Topo {
    const Block *blocks;   // [nblocks]
    const int *blist;      // block indices of the launch
    int nlist;             // number of blocks in the launch
    const Shape *shapes;
    const uint32_t *cells; // packed (i | j<<10 | k<<20) per WO slot
    const int *pstart;     // plane starts per shape
    const int *slot_of;    // spatial (k + NK(j + NJ i)) -> WO slot
    int max_ni, max_nj, max_nk;
    int max_plane;         // largest hyperplane over all shapes
    int maxlevel;
    const uint2 *items;    // plan items of all shapes, widths and kinds
    const uint32_t *ents;  // plan entries
    const uint32_t *cellx; // cell words of the plane kernels (per WO slot)
    const uint32_t *fents; // face-store entries, 64 per plan item (parallel to items)
    const void *zero;      // kZeroBytes of zeros: source of absent ghosts
};
constexpr size_t kZeroBytes = 65536 * sizeof(double);

// Field set of one precision T (double or float) in WO layout.
template <typename T>
struct Fields {
    T *hilf;                    // correction HILF
    T *rhs;                     // right-hand side read by the sweeps
    T *res;                     // forward-sweep result RES
    const T *ap;                // GSAP
    const T *rap;               // SOR_RAP = divide0(1, GSAP) (SOR only)
    const T *f[7];              // LW LS LB LPR UE UN UT, at Block::fo
    const T *aw, *ae, *as, *an, *ab, *at;  // 1-D coefficients (ref. layout)
    // 2 / (d(i-1) d(i)) per direction (ref. 1-D layout, FP64): the separable
    // part of GSAP, from which the plane kernels rebuild GSAP of blocks
    // without blocked cells
    const double *apx, *apy, *apz;
    // GSAP kept with the shared factors (at Block::fo), or null when no
    // factors are shared; the plane kernels then rebuild it on chip
    const T *aps;
    T *hface;                   // HILF boundary planes (own faces)
    T *hpar;                    // HILF ghosts of the block's PAR faces
    T *rhi[2];                  // RES upper planes, double-buffered
    T *rlo;                     // RES lower planes
};

// Shared state of the FP64 outer (defect-correction) loop.
struct Outer {
    double *rhs64;       // RHS in FP64 (WO)
    double *dp64;        // DP in FP64 (WO)
    const double *ap64;  // GSAP in FP64 (WO)
    const double *aw, *ae, *as, *an, *ab, *at;  // FP64 1-D coefficients
    unsigned long long *gmax;  // running max |RHS| (bit pattern)
    double *maxrhs0;     // max |RHS| before the first iteration
    double *maxrhs;      // max |RHS| after the last outer iteration
    int *ctr;            // outer iterations done
    int *done;           // 1: stop (converged, nouter reached or non-finite)
    int *converged;      // 1: exit test fired
    int *nonfinite;      // 1: NaN/Inf residual seen
};

struct Mask {
    const uint8_t *bp;   // BP per WO cell (1 fluid, 0 blocked)
    const uint8_t *bpf;  // BP of the block's own ghost layer, face layout
};

void launch_fill_u8(uint8_t *p, int64_t n, uint8_t v, hipStream_t s);
template <typename T>
void launch_fill(T *p, int64_t n, T v, hipStream_t s);

void launch_ap(const Topo &t, const double *dx, const double *dy,
               const double *dz, const double *aw_raw, const double *ae_raw,
               const double *as_raw, const double *an_raw,
               const double *ab_raw, const double *at_raw, Mask m,
               double *ap, hipStream_t s);

template <typename T>
void launch_lu(const Topo &t, const double *ap, const double *aw,
               const double *ae, const double *as, const double *an,
               const double *ab, const double *at, Mask m, double alfa,
               T *const f[7], hipStream_t s);

// shared factor storage: per-block hash of the factors and GSAP at
// Block::wo, exact comparison with the representative rep[b], copy of the
// representatives' factors and GSAP (FP64, and FP32 when gap32 is given) to
// Block::fo of new arrays
template <typename T>
void launch_fhash(const Topo &t, const T *const f[7], const double *ap, Mask m,
                  unsigned long long *out, hipStream_t s);
template <typename T>
void launch_fcmp(const Topo &t, const T *const f[7], const double *ap, Mask m,
                 const int *rep, int *bad, hipStream_t s);
template <typename T>
void launch_fgather(const Topo &t, const T *const f[7], const double *ap,
                    T *const g[7], double *gap, float *gap32, hipStream_t s);

template <typename T>
void launch_convert(const double *src, T *dst, int64_t n, hipStream_t s);

template <typename T>
void launch_divcal(const Topo &t, const double *u, const double *v,
                   const double *w, const double *rddx, const double *rddy,
                   const double *rddz, const double *bp_sp, const double *sdiv_sp,
                   double fak, double *rhs64, T *rhs_t, hipStream_t s);

void launch_maxabs(const Topo &t, const double *rhs64, Outer o, hipStream_t s);

template <typename T>
void launch_forward(const Topo &t, const Fields<T> &f, Mask m, int rd,
                    int write_hi, int hilf_zero, int classic, const int *done,
                    hipStream_t s);

template <typename T>
void launch_backward(const Topo &t, const Fields<T> &f, int hilf_zero,
                     int classic, const int *done, hipStream_t s);

template <typename T>
void launch_rap(const double *ap, T *rap, int64_t n, hipStream_t s);

template <typename T>
void launch_sor(const Topo &t, const Fields<T> &f, Mask m, const T *fin,
                T *fout, double omega, int hilf_zero, const int *done,
                hipStream_t s);

// pipeline shapes of the plane kernels (items in flight per wave x cells
// per item, waves per block): 1: 4x64, 2: 2x128, 3: 3x128, 4: 6x128,
// 5: 8x64 (one wave); 6: 4x64 (4 waves), 7: 4x64 (8 waves), 8: 3x128
// (2 waves), 9: 4x64 (2 waves)
constexpr int kNumShapes = 9;
void set_plan_shape(int v);                // every plane kernel
void set_plan_shapes(int fwd, int bwd);    // the backward sweep apart
int plan_shape();

template <typename T>
void launch_outer(const Topo &t, const Fields<T> &f, Mask m, Outer o,
                  T *rhs_t, hipStream_t s);

void launch_check(Outer o, int nouter, int nouter_min, double epcorr,
                  double eps_rel, double prefak, double *hist, int hist_cap,
                  hipStream_t s);

// the outer update of one outer iteration fused with the first forward
// sweep (HILF = 0) of the next (mgp_fused.hip); rd and write_hi as for
// launch_forward
template <typename T>
void launch_outer_fwd(const Topo &t, const Fields<T> &f, Mask m, Outer o, T *rhs_t,
                      int rd, int write_hi, int classic, hipStream_t s);

// zero_ghosts: also set every non-interior position of each block to 0
template <typename T>
void launch_wo_to_sp(const Topo &t, const T *src, double *dst,
                     hipStream_t s, bool zero_ghosts = false);

// z-chains of blocks for the velocity and pressure correction that also
// writes the k-face ghosts of interior rows (k_mgpcorr_chain): chain[] holds
// block indices, chains[c] = (first entry, length) of chain c, each a cycle
// along +z of at most kCorrChainMax blocks of one size with nk >= 2; nx
// workgroups per chain.
// WO slots the layout-change kernels (divcal, WO -> reference copy) stage
// in LDS per pass: 16384 doubles, 128 KB of the MI355X's 160 KB per
// workgroup, so a block of up to 16384 cells (24^3 = 13824) takes one
// pass.  Each pass reads (divcal) or writes (the copy) the parts of the
// reference layout's k-lines whose cells fall in its slot range, so two
// passes move every line twice (8192 slots: divcal 5.74 ms, 4.42 ms with
// one pass on 32768 blocks of 24^3).
constexpr int kStage = 16384;
constexpr int kCorrThreads = 256;   // workgroup of the correction kernels
constexpr int kCorrChainMax = 256;  // longest chain
constexpr int kCorrRows = 16;       // most rows per workgroup of the chain update
struct CorrChains {
    const int *chain;
    const int2 *chains;
    int n, nx;
};

// cc = nullptr: plain update of the blocks of t.blist; otherwise the chain
// update, after which the exchange must leave out the k-face copies of
// interior rows.
void launch_mgpcorr(const Topo &t, double *u, double *v, double *w,
                    double *p, const double *dp, const double *rdx,
                    const double *rdy, const double *rdz, double rfak,
                    const double *bp_sp, const CorrChains *cc, hipStream_t s);

// Arguments of the momentum right-hand side (reference layout, FP64).
struct MomArgs {
    double *uo, *vo, *wo;
    const double *u, *v, *w, *p, *g;
    const double *dx, *dy, *dz, *ddx, *ddy, *ddz;
    const double *rdx, *rdy, *rdz, *rddx, *rddy, *rddz;
    const double *rpdx, *rpdy, *rpdz;   // 1 / (rho dx) etc.
    double rrho, gmol, iles;            // 1 / rho
    double gp[3];      // body-force pressure gradient (gradp)
    double rho;
    double ww[5];      // Werner-Wengle constants cwb, cpo4, cpo5, cpo6, cpo8
};
void launch_tstle4(const Topo &t, const MomArgs &a, hipStream_t s);
// tstle4_par at the PAR faces of the npar blocks of plist (after tstle4)
void launch_tstle4_par(const Topo &t, const MomArgs &a, const int *plist, int npar,
                       hipStream_t s);
void launch_rkstep(double *p, double *dp, const double *rhsp, int64_t n,
                   double frhs, double dtfu, hipStream_t s);

// One stage of the exchange: `nblock` receiving blocks.  mode 0: an entry
// moves one value; 1: two values (scalar accesses); 2: two values as one
// 16-byte access (every field pointer 16-byte aligned).
void launch_pull(const PullBlock *blocks, int nblock, const uint2 *ents, int mode,
                 double *const *fields, int nfield, hipStream_t s);

template <typename T>
void launch_to_ref_factors(const Topo &t, const T *wo, double *ref,
                           int spatial, hipStream_t s);

// coupling of grid levels (mgp_levels.hip)
// `done`: the outer loop's stop flag (the kernels return when it is set)
template <typename T>
void launch_ctof(const Topo &t, const Fields<T> &f, T *fout, T *hsh,
                 double *dsh, const int *done, hipStream_t s);
void launch_shell_to_sp(const Topo &t, const double *dsh, double *sp,
                        hipStream_t s);
template <typename T>
void launch_parent(const Topo &t, const Fields<T> &f, T *pbuf, const int *done,
                   hipStream_t s);
template <typename T>
void launch_bound(const Topo &t, const Fields<T> &f, const T *pbuf, Mask m,
                  int bpv, const double *const metric[6], const int *done,
                  hipStream_t s);
// mode 0 plain restriction, 1 / 2 the ghostcell 'P' / 'R' restrictions
template <typename S, typename D>
void launch_restrict(const Topo &t, const S *src, D *dst, D *face,
                     float *dst2, Mask m, int mode, const double *const dd[3],
                     const int *done, hipStream_t s);
void launch_parent_sp(const Topo &t, const double *f, double *pbuf,
                      hipStream_t s);
void launch_bound_sp(const Topo &t, double *p, const double *pbuf,
                     const double *bp, int bpv, const double *const metric[6],
                     hipStream_t s);
// closing sync of mgpoisl (reference layout): U, V, W of the parents into
// the PAR face buffers, then the ghost values of U, V, W, P at the FIX,
// OP1, NOS, SLI and PAR faces (bbo: buffer offset per grid and face, -1
// none; pinf: OP1 pressure per grid and face; bp may be null: all fluid)
void launch_parent_uvw(const Topo &t, const double *u, const double *v, const double *w,
                       double *ubuf, double *vbuf, double *wbuf, const int64_t *bbo,
                       hipStream_t s);
void launch_bound_flow(const Topo &t, double *u, double *v, double *w, double *p,
                       const double *ubuf, const double *vbuf, const double *wbuf,
                       const int64_t *bbo, const double *pinf, const double *bp,
                       const double *ddx, const double *ddy, const double *ddz, double rho,
                       hipStream_t s);
// gc: ghostcell restriction of P (bp may be null: all fluid)
// One transfer of par_ftoc_norm: the face-normal velocity on PAR face
// `face` (0..5) of block `fine` into block `dest` (its parent, or the
// parent's neighbour across that face) at the 1-based normal index dnorm.
struct ParFtocTask {
    int fine, face, dest, dnorm;
};
void launch_par_ftoc(const Block *blocks, const ParFtocTask *tasks, int ntask, double *u,
                     double *v, double *w, const double *const dd[3], int sum,
                     hipStream_t s);

void launch_ftoc_uvwp(const Topo &t, double *u, double *v, double *w,
                      double *p, int gc, const double *bp,
                      const double *const dd[3], hipStream_t s);
void launch_maxabs_acc(const Topo &t, const double *rhs64, Outer o,
                       hipStream_t s);  // harmless after the stop: k_check skips

// adjoint projection (mgp_adjoint.hip)
template <typename T>
void launch_tfactors(const Topo &t, Mask m, const T *const f[7],
                     T *const g[6], hipStream_t s);
void launch_adj_clean(const Topo &t, double *u, double *v, double *w,
                      const double *bp, hipStream_t s);
template <typename T>
void launch_gradT(const Topo &t, const double *u, const double *v,
                  const double *w, const double *rdx, const double *rdy,
                  const double *rdz, const double *bp, double *rhs64,
                  T *rhs_t, hipStream_t s);
void launch_divT(const Topo &t, double *u, double *v, double *w,
                 const double *y, const double *rddx, const double *rddy,
                 const double *rddz, const double *bp, hipStream_t s);

}  // namespace mgp
