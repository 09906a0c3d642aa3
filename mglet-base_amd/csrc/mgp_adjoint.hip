// mgp_adjoint.hip -- kernels of the adjoint projection.
//
// The projection maps the interior face velocities u to
//     u' = u - G A^-1 D u
// (D = divcal without rho/dt, A = the GSA* operator, G = the mgpcorr
// stencil without dt/rho; the two scale factors cancel).  Its adjoint is
//     ubar = w - D^T A^-T G^T w.
// The kernels here compute G^T w (the right-hand side of the adjoint
// Poisson problem), D^T y, and the factors that let the unchanged SIP sweep
// kernels apply (LU)^-T = L^-T U^-T of every block, so A^T y = G^T w is
// solved by the same outer/inner iteration as the forward problem.
//
// Faces: u(i) is the east face of cell i (divcal uses u(i) - u(i-1),
// src/ib/noib_mod.F90:336-343).  Two kinds of interior faces belong to the
// boundary conditions rather than to the unknowns and are neither inputs
// nor outputs of the adjoint: faces on a block boundary without a
// connected neighbour (set by the caller's bound_flow), and, with the
// ghostcell operator, faces with BP = 0 on either side (mgpcorr leaves
// them unchanged).  On the remaining faces the projection's Jacobian does
// not depend on the constant left undetermined by a singular Poisson
// problem, so its transpose is well defined.
#include <hip/hip_runtime.h>

#include "mgp_launch.h"

namespace mgp {

namespace {

constexpr int kThreadsAdj = 256;

__device__ __forceinline__ void unpack_cell(uint32_t c, int &i, int &j,
                                            int &k) {
    i = c & 1023;
    j = (c >> 10) & 1023;
    k = c >> 20;
}

// Factors of the transposed block preconditioner.  M = L U with L lower
// (diagonal 1/LPR, off-diagonals LW/LPR, LS/LPR, LB/LPR) and U unit upper
// (UE, UN, UT).  With the scaled factors
//   LWt_c = LPR_c UE_w / LPR_w   LSt_c = LPR_c UN_s / LPR_s
//   LBt_c = LPR_c UT_b / LPR_b   UEt_c = LPR_c LW_e / LPR_e
//   UNt_c = LPR_c LS_n / LPR_n   UTt_c = LPR_c LB_t / LPR_t
// (neighbours inside the block, zero across its faces) the forward sweep
// z_c = LPR_c r_c - LBt z_b - LSt z_s - LWt z_w yields z = LPR (U^-T r) and
// the backward sweep y_c = z_c - UTt y_t - UNt y_n - UEt y_e yields
// y = L^-T U^-T r.  Blocked cells get zero factors, which keeps their
// values at zero as in the forward solve.
template <typename T>
__global__ void k_tfactors(Topo t, Mask m, const T *lw, const T *ls,
                           const T *lb, const T *lpr, const T *ue,
                           const T *un, const T *ut, T *lwt, T *lst, T *lbt,
                           T *uet, T *unt, T *utt) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const uint32_t *cells = t.cells + S.cell_off;
    const int *slot = t.slot_of + S.slot_off;
    const int NI = B.ni, NJ = B.nj, NK = B.nk;
    // factors and transposed factors at the block's factor offset (blocks
    // with equal factors share them, and their transposed factors are then
    // equal too: they involve the block's own cells only)
    const int64_t fo = B.fo;
    const bool masked = m.bp != nullptr && !(B.flags & 1);
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < S.ncell;
         s += gridDim.x * blockDim.x) {
        int i, j, k;
        unpack_cell(cells[s], i, j, k);
        const int64_t o = fo + s;
        double a = 0.0, bs = 0.0, c = 0.0, d = 0.0, e = 0.0, f = 0.0;
        if (!(masked && m.bp[B.wo + s] == 0)) {
            const double p = (double)lpr[o];
            auto at = [&](int ii, int jj, int kk) {
                return fo + slot[kk + NK * (jj + NJ * ii)];
            };
            if (i > 0) { const int64_t q = at(i - 1, j, k); a = p * (double)ue[q] / (double)lpr[q]; }
            if (j > 0) { const int64_t q = at(i, j - 1, k); bs = p * (double)un[q] / (double)lpr[q]; }
            if (k > 0) { const int64_t q = at(i, j, k - 1); c = p * (double)ut[q] / (double)lpr[q]; }
            if (i < NI - 1) { const int64_t q = at(i + 1, j, k); d = p * (double)lw[q] / (double)lpr[q]; }
            if (j < NJ - 1) { const int64_t q = at(i, j + 1, k); e = p * (double)ls[q] / (double)lpr[q]; }
            if (k < NK - 1) { const int64_t q = at(i, j, k + 1); f = p * (double)lb[q] / (double)lpr[q]; }
        }
        lwt[o] = (T)a;
        lst[o] = (T)bs;
        lbt[o] = (T)c;
        uet[o] = (T)d;
        unt[o] = (T)e;
        utt[o] = (T)f;
    }
}

// Faces of cell (i, j, k) (0-based interior) that are boundary conditions.
__device__ __forceinline__ void bc_faces(const Block &B, const double *bp,
                                         int64_t c, int64_t si, int64_t sj,
                                         int i, int j, int k, bool &xu,
                                         bool &xv, bool &xw) {
    xu = i == B.ni - 1 && B.nbr[1] < 0;
    xv = j == B.nj - 1 && B.nbr[3] < 0;
    xw = k == B.nk - 1 && B.nbr[5] < 0;
    if (bp) {
        const double bc = bp[c];
        xu = xu || bc * bp[c + si] == 0.0;
        xv = xv || bc * bp[c + sj] == 0.0;
        xw = xw || bc * bp[c + 1] == 0.0;
    }
}

// Zero the ghost layers of u, v, w and the boundary-condition faces
// (reference layout).
__global__ void k_adj_clean(Topo t, double *u_, double *v_, double *w_,
                            const double *bp_) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const int kk = B.nk + 4, jj = B.nj + 4, ii = B.ni + 4;
    const int64_t n = (int64_t)kk * jj * ii;
    double *u = u_ + B.sp, *v = v_ + B.sp, *w = w_ + B.sp;
    const double *bp = bp_ ? bp_ + B.sp : nullptr;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % kk) + 1;
        const int j = (int)((e / kk) % jj) + 1;
        const int i = (int)(e / ((int64_t)kk * jj)) + 1;
        const bool in = i >= 3 && i <= ii - 2 && j >= 3 && j <= jj - 2 &&
                        k >= 3 && k <= kk - 2;
        bool xu = true, xv = true, xw = true;
        if (in) bc_faces(B, bp, e, (int64_t)kk * jj, kk, i - 3, j - 3, k - 3, xu, xv, xw);
        if (xu) u[e] = 0.0;
        if (xv) v[e] = 0.0;
        if (xw) w[e] = 0.0;
    }
}

// rhs_c = (G^T w)_c = -rdx(i) b(i,i+1) u(i) + rdx(i-1) b(i-1,i) u(i-1) + ...
// (transpose of src/flow/pressuresolver_mod.F90:787-879; the ghost u(i-1)
// at the block's low face holds the neighbour's face after a 1-layer
// exchange, or zero).
template <typename T>
__global__ void k_gradT(Topo t, const double *u_, const double *v_,
                        const double *w_, const double *rdx_,
                        const double *rdy_, const double *rdz_,
                        const double *bp_, double *rhs64, T *rhs_t) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const uint32_t *cells = t.cells + S.cell_off;
    const int64_t kk = B.nk + 4, jj = B.nj + 4, si = kk * jj, sj = kk;
    const double *rdx = rdx_ + B.o1x + 2, *rdy = rdy_ + B.o1y + 2,
                 *rdz = rdz_ + B.o1z + 2;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < S.ncell;
         s += gridDim.x * blockDim.x) {
        int i, j, k;
        unpack_cell(cells[s], i, j, k);
        const int64_t c = B.sp + (k + 2) + (j + 2) * kk + (i + 2) * si;
        double bxe = 1.0, bxw = 1.0, bye = 1.0, byw = 1.0, bze = 1.0,
               bzw = 1.0;
        if (bp_) {
            const double bc = bp_[c];
            bxe = bc * bp_[c + si];
            bxw = bp_[c - si] * bc;
            bye = bc * bp_[c + sj];
            byw = bp_[c - sj] * bc;
            bze = bc * bp_[c + 1];
            bzw = bp_[c - 1] * bc;
        }
        const double g = -rdx[i] * bxe * u_[c] + rdx[i - 1] * bxw * u_[c - si]
                         - rdy[j] * bye * v_[c] + rdy[j - 1] * byw * v_[c - sj]
                         - rdz[k] * bze * w_[c] + rdz[k - 1] * bzw * w_[c - 1];
        rhs64[B.wo + s] = g;
        if (rhs_t) rhs_t[B.wo + s] = (T)g;
    }
}

// u(i) <- u(i) - (D^T y)(i) = u(i) - (rddx(i) y_i - rddx(i+1) y_{i+1}) on
// every interior face (transpose of src/ib/noib_mod.F90:304-354; with the
// ghostcell operator y is first multiplied by BP).  y carries its face
// ghosts from a 1-layer exchange.
__global__ void k_divT(Topo t, double *u_, double *v_, double *w_,
                       const double *y_, const double *rddx_,
                       const double *rddy_, const double *rddz_,
                       const double *bp_) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const uint32_t *cells = t.cells + S.cell_off;
    const int64_t kk = B.nk + 4, jj = B.nj + 4, si = kk * jj, sj = kk;
    const double *rddx = rddx_ + B.o1x + 2, *rddy = rddy_ + B.o1y + 2,
                 *rddz = rddz_ + B.o1z + 2;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < S.ncell;
         s += gridDim.x * blockDim.x) {
        int i, j, k;
        unpack_cell(cells[s], i, j, k);
        const int64_t c = B.sp + (k + 2) + (j + 2) * kk + (i + 2) * si;
        auto yb = [&](int64_t q) { return bp_ ? bp_[q] * y_[q] : y_[q]; };
        const double yc = yb(c);
        bool xu, xv, xw;
        bc_faces(B, bp_, c, si, sj, i, j, k, xu, xv, xw);
        u_[c] = xu ? 0.0 : u_[c] - (rddx[i] * yc - rddx[i + 1] * yb(c + si));
        v_[c] = xv ? 0.0 : v_[c] - (rddy[j] * yc - rddy[j + 1] * yb(c + sj));
        w_[c] = xw ? 0.0 : w_[c] - (rddz[k] * yc - rddz[k + 1] * yb(c + 1));
    }
}

}  // namespace

template <typename T>
void launch_tfactors(const Topo &t, Mask m, const T *const f[7],
                     T *const g[6], hipStream_t s) {
    MGP_RANGE("mgp:tfactors");
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    dim3 grid((maxc + kThreadsAdj - 1) / kThreadsAdj, t.nlist);
    k_tfactors<T><<<grid, kThreadsAdj, 0, s>>>(t, m, f[0], f[1], f[2], f[3],
                                               f[4], f[5], f[6], g[0], g[1],
                                               g[2], g[3], g[4], g[5]);
}
template void launch_tfactors<double>(const Topo &, Mask, const double *const[7],
                                      double *const[6], hipStream_t);
template void launch_tfactors<float>(const Topo &, Mask, const float *const[7],
                                     float *const[6], hipStream_t);

void launch_adj_clean(const Topo &t, double *u, double *v, double *w,
                      const double *bp, hipStream_t s) {
    MGP_RANGE("mgp:adj_clean");
    const int maxc = (t.max_ni + 4) * (t.max_nj + 4) * (t.max_nk + 4);
    dim3 grid((maxc + kThreadsAdj - 1) / kThreadsAdj, t.nlist);
    k_adj_clean<<<grid, kThreadsAdj, 0, s>>>(t, u, v, w, bp);
}

template <typename T>
void launch_gradT(const Topo &t, const double *u, const double *v,
                  const double *w, const double *rdx, const double *rdy,
                  const double *rdz, const double *bp, double *rhs64,
                  T *rhs_t, hipStream_t s) {
    MGP_RANGE("mgp:gradT");
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    dim3 grid((maxc + kThreadsAdj - 1) / kThreadsAdj, t.nlist);
    k_gradT<T><<<grid, kThreadsAdj, 0, s>>>(t, u, v, w, rdx, rdy, rdz, bp,
                                            rhs64, rhs_t);
}
template void launch_gradT<double>(const Topo &, const double *, const double *,
                                   const double *, const double *, const double *,
                                   const double *, const double *, double *,
                                   double *, hipStream_t);
template void launch_gradT<float>(const Topo &, const double *, const double *,
                                  const double *, const double *, const double *,
                                  const double *, const double *, double *,
                                  float *, hipStream_t);

void launch_divT(const Topo &t, double *u, double *v, double *w,
                 const double *y, const double *rddx, const double *rddy,
                 const double *rddz, const double *bp, hipStream_t s) {
    MGP_RANGE("mgp:divT");
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    dim3 grid((maxc + kThreadsAdj - 1) / kThreadsAdj, t.nlist);
    k_divT<<<grid, kThreadsAdj, 0, s>>>(t, u, v, w, y, rddx, rddy, rddz, bp);
}

}  // namespace mgp
