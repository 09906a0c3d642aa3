// mgp_momentum.hip -- the RK stage around the projection on gfx950: the
// momentum right-hand side of tstle4 (convection, diffusion, pressure
// gradient) and the 2N-storage Runge-Kutta update.
//
// Both work in the reference layout (the caller's fields), one thread per
// stored position with k fastest, so every neighbour access of a wave is a
// contiguous run.  Expressions keep the reference's association order
// (Fortran evaluates operators of equal precedence left to right; a leading
// minus applies to the whole product) and the build uses -ffp-contract=off,
// so the result is bitwise that of the CPU restatement.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mgp_launch.h"

namespace mgp {

constexpr int kMomThreads = 256;

__device__ __forceinline__ double gpair(double a, double b, double gmol) {
    const double s = a + b;
    return a * b / (s > gmol ? s : gmol);
}

// swcle3d_one (src/flow/tstle4_mod.F90:1848-1859) with tauwin
// (src/flow/wernerwengle_mod.F90:64-87): the Werner-Wengle wall shear stress
// of velocity uq at wall distance dd, over rho dd.  pow is the device
// library's, so these terms agree with a CPU evaluation to rounding, not
// bit for bit.
__device__ __forceinline__ double swc(double dd, double uq, const MomArgs &a) {
    const double vz = copysign(1.0, uq), un = fabs(uq);
    double tau;
    if (un >= a.ww[3] / dd) {
        const double ddsb = pow(dd, -a.ww[0]);
        tau = vz * a.rho * pow(ddsb * (a.ww[1] * un + a.ww[2] / dd), a.ww[4]);
    } else {
        tau = vz * 2.0 * a.gmol * un / dd;
    }
    return tau / a.rho / dd;
}

// The swcle3d terms (src/flow/tstle4_mod.F90:1756-1845) of velocity
// component comp (0 u, 1 v, 2 w) at (k, j, i): NOS walls act on the
// tangential components of the first cell, each face loop running over
// 2..n-1 of the tangential indices, in the reference's face order.
__device__ __forceinline__ double wall_terms(double val, int comp, int k, int j, int i,
                                             const int *bc, int kk, int jj, int ii,
                                             double vel, const double *ddx, const double *ddy,
                                             const double *ddz, const MomArgs &a) {
    const bool tk = k >= 2 && k <= kk - 1, tj = j >= 2 && j <= jj - 1, ti = i >= 2 && i <= ii - 1;
    if (comp != 0) {
        if (bc[0] == 5 && i == 3 && tj && tk) val = val - swc(ddx[i], vel, a);
        if (bc[1] == 5 && i == ii - 2 && tj && tk) val = val - swc(ddx[i], vel, a);
    }
    if (comp != 1) {
        if (bc[2] == 5 && j == 3 && ti && tk) val = val - swc(ddy[j], vel, a);
        if (bc[3] == 5 && j == jj - 2 && ti && tk) val = val - swc(ddy[j], vel, a);
    }
    if (comp != 2) {
        if (bc[4] == 5 && k == 3 && ti && tj) val = val - swc(ddz[k], vel, a);
        if (bc[5] == 5 && k == kk - 2 && ti && tj) val = val - swc(ddz[k], vel, a);
    }
    return val;
}

// tstle4 without immersed boundaries or PAR faces (src/flow/tstle4_mod.F90:
// 13-529): uo = convection (tstle4_kon :139-265), minus the wall stress next
// to NOS walls (swcle3d, called first in tstle4_diff :315), + diffusion
// (tstle4_diff :268-458, the transpose part of the stress when the LES
// model is on) - pressure gradient and body-force gradient (tstle4_gradp
// :461-529); zero outside the loop ranges (but for the wall terms, whose
// loops include ghost rows), as the reference zeroes uo before filling it.
// The transporting velocities are u, v, w.
// 1/rho and 1/(rho dx) come from host tables: the same correctly rounded
// quotients, without a division per position.
// One-dimensional grid of nchunk workgroups per block, consecutive
// workgroups walking one block, so the i-1 plane a workgroup reads was
// fetched by one a few workgroups earlier.
__global__ __launch_bounds__(kMomThreads) void k_tstle4(Topo t, MomArgs a, int nchunk,
                                                         int first) {
    const int b = t.blist[first + (int)(blockIdx.x / nchunk)];
    const Block &B = t.blocks[b];
    const int kk = B.nk + 4, jj = B.nj + 4, ii = B.ni + 4;
    const int64_t n = (int64_t)kk * jj * ii;
    const int64_t e = (int64_t)(blockIdx.x % nchunk) * kMomThreads + threadIdx.x;
    if (e >= n) return;
    const int k = (int)(e % kk) + 1, j = (int)((e / kk) % jj) + 1,
              i = (int)(e / ((int64_t)kk * jj)) + 1;
    const int64_t sj = kk, si = (int64_t)kk * jj;
    const int64_t c = B.sp + e;
    const double *u = a.u, *v = a.v, *w = a.w, *p = a.p, *gg = a.g;
    const double *ut = u, *vt = v, *wt = w;
    // 1-based 1-D metrics of the block
    const double *dx = a.dx + B.o1x - 1, *ddx = a.ddx + B.o1x - 1;
    const double *dy = a.dy + B.o1y - 1, *ddy = a.ddy + B.o1y - 1;
    const double *dz = a.dz + B.o1z - 1, *ddz = a.ddz + B.o1z - 1;
    const double *rdx = a.rdx + B.o1x - 1, *rddx = a.rddx + B.o1x - 1;
    const double *rdy = a.rdy + B.o1y - 1, *rddy = a.rddy + B.o1y - 1;
    const double *rdz = a.rdz + B.o1z - 1, *rddz = a.rddz + B.o1z - 1;
    // CON (7) and OP1 (3) faces extend the loops
    const int nfu = B.bc[0] == 3, nbu = B.bc[1] == 7 || B.bc[1] == 3;
    const int nrv = B.bc[2] == 3, nlv = B.bc[3] == 7 || B.bc[3] == 3;
    const int nbw = B.bc[4] == 3, ntw = B.bc[5] == 7 || B.bc[5] == 3;
    const double rrho = a.rrho, gmol = a.gmol, iles = a.iles;
    const bool nos = B.bc[0] == 5 || B.bc[1] == 5 || B.bc[2] == 5 || B.bc[3] == 5 ||
                     B.bc[4] == 5 || B.bc[5] == 5;
    const double *rpx = a.rpdx + B.o1x - 1, *rpy = a.rpdy + B.o1y - 1,
                 *rpz = a.rpdz + B.o1z - 1;
    const bool inj = j >= 3 && j <= jj - 2, ink = k >= 3 && k <= kk - 2,
               ini = i >= 3 && i <= ii - 2;
#define U(o) u[c + (o)]
#define V(o) v[c + (o)]
#define W(o) w[c + (o)]
#define G(o) gg[c + (o)]
    double ro = 0.0;
    if (i >= 3 - nfu && i <= ii - 3 + nbu && inj && ink) {
        const double ax = ddy[j] * ddz[k], ay = dx[i] * ddz[k], az = dx[i] * ddy[j];
        const double fe = ax * (ut[c] + (ut[c + si] - ut[c]) * 0.5 * dx[i] / ddx[i + 1]);
        const double fw = ax * (ut[c - si] + (ut[c] - ut[c - si]) * 0.5 * dx[i - 1] / ddx[i]);
        const double fn = ay * (vt[c] + vt[c + si]) * 0.5;
        const double fs = ay * (vt[c - sj] + vt[c - sj + si]) * 0.5;
        const double ft = az * (wt[c] + wt[c + si]) * 0.5;
        const double fb = az * (wt[c - 1] + wt[c - 1 + si]) * 0.5;
        double qe = 0.5 * fe * (U(0) + U(si));
        double qw = 0.5 * fw * (U(-si) + U(0));
        double qn = 0.5 * fn * (U(0) + U(sj));
        double qs = 0.5 * fs * (U(-sj) + U(0));
        double qt = 0.5 * ft * (U(0) + U(1));
        double qb = 0.5 * fb * (U(-1) + U(0));
        ro = -((qe - qw + qn - qs + qt - qb) * rdx[i] * rddy[j] * rddz[k]);
        if (nos) ro = wall_terms(ro, 0, k, j, i, B.bc, kk, jj, ii, U(0), ddx, ddy, ddz, a);
        const double ge = G(si), gw = G(0);
        const double gn = gpair(G(0), G(sj), gmol) + gpair(G(si), G(si + sj), gmol);
        const double gs = gpair(G(-sj), G(0), gmol) + gpair(G(si - sj), G(si), gmol);
        const double gt = gpair(G(0), G(1), gmol) + gpair(G(si), G(si + 1), gmol);
        const double gb = gpair(G(-1), G(0), gmol) + gpair(G(si - 1), G(si), gmol);
        qe = -(ge * ax * rddx[i + 1] * (U(si) - U(0)));
        qw = -(gw * ax * rddx[i] * (U(0) - U(-si)));
        qn = -(gn * ay * rdy[j] * (U(sj) - U(0)));
        qs = -(gs * ay * rdy[j - 1] * (U(0) - U(-sj)));
        qt = -(gt * az * rdz[k] * (U(1) - U(0)));
        qb = -(gb * az * rdz[k - 1] * (U(0) - U(-1)));
        const double st = ((ge * (U(si) - U(0)) * rddx[i + 1]) - (gw * (U(0) - U(-si)) * rddx[i])) * ax
                          + ((gn * (V(si) - V(0))) - (gs * (V(si - sj) - V(-sj)))) * ddz[k]
                          + ((gt * (W(si) - W(0))) - (gb * (W(si - 1) - W(-1)))) * ddy[j];
        const double qc = st * iles;
        const double fak = rrho * rddy[j] * rdx[i] * rddz[k];
        ro = ro - fak * (qe - qw + qn - qs + qt - qb - qc);
        ro = ro - rpx[i] * (p[c + si] - p[c] + a.gp[0] * dx[i]);
    } else if (nos) {
        ro = wall_terms(ro, 0, k, j, i, B.bc, kk, jj, ii, U(0), ddx, ddy, ddz, a);
    }
    a.uo[c] = ro;
    ro = 0.0;
    if (ini && j >= 3 - nrv && j <= jj - 3 + nlv && ink) {
        const double ax = dy[j] * ddz[k], ay = ddx[i] * ddz[k], az = ddx[i] * dy[j];
        const double fe = ax * (ut[c] + ut[c + sj]) * 0.5;
        const double fw = ax * (ut[c - si] + ut[c - si + sj]) * 0.5;
        const double fn = ay * (vt[c] + (vt[c + sj] - vt[c]) * 0.5 * dy[j] / ddy[j + 1]);
        const double fs = ay * (vt[c - sj] + (vt[c] - vt[c - sj]) * 0.5 * dy[j - 1] / ddy[j]);
        const double ft = az * (wt[c] + wt[c + sj]) * 0.5;
        const double fb = az * (wt[c - 1] + wt[c - 1 + sj]) * 0.5;
        double qe = 0.5 * fe * (V(0) + V(si));
        double qw = 0.5 * fw * (V(-si) + V(0));
        double qn = 0.5 * fn * (V(0) + V(sj));
        double qs = 0.5 * fs * (V(-sj) + V(0));
        double qt = 0.5 * ft * (V(0) + V(1));
        double qb = 0.5 * fb * (V(-1) + V(0));
        ro = -((qe - qw + qn - qs + qt - qb) * rddx[i] * rdy[j] * rddz[k]);
        if (nos) ro = wall_terms(ro, 1, k, j, i, B.bc, kk, jj, ii, V(0), ddx, ddy, ddz, a);
        const double ge = gpair(G(0), G(si), gmol) + gpair(G(sj), G(sj + si), gmol);
        const double gw = gpair(G(-si), G(0), gmol) + gpair(G(sj - si), G(sj), gmol);
        const double gn = G(sj), gs = G(0);
        const double gt = gpair(G(0), G(1), gmol) + gpair(G(sj), G(sj + 1), gmol);
        const double gb = gpair(G(-1), G(0), gmol) + gpair(G(sj - 1), G(sj), gmol);
        qe = -(ge * ax * rdx[i] * (V(si) - V(0)));
        qw = -(gw * ax * rdx[i - 1] * (V(0) - V(-si)));
        qn = -(gn * ay * rddy[j + 1] * (V(sj) - V(0)));
        qs = -(gs * ay * rddy[j] * (V(0) - V(-sj)));
        qt = -(gt * az * rdz[k] * (V(1) - V(0)));
        qb = -(gb * az * rdz[k - 1] * (V(0) - V(-1)));
        const double st = ((ge * (U(sj) - U(0))) - (gw * (U(sj - si) - U(-si)))) * ddz[k]
                          + ((gn * (V(sj) - V(0)) * rddy[j + 1]) - (gs * (V(0) - V(-sj)) * rddy[j])) * ay
                          + ((gt * (W(sj) - W(0))) - (gb * (W(sj - 1) - W(-1)))) * ddx[i];
        const double qc = st * iles;
        const double fak = rrho * rddx[i] * rdy[j] * rddz[k];
        ro = ro - fak * (qe - qw + qn - qs + qt - qb - qc);
        ro = ro - rpy[j] * (p[c + sj] - p[c] + a.gp[1] * dy[j]);
    } else if (nos) {
        ro = wall_terms(ro, 1, k, j, i, B.bc, kk, jj, ii, V(0), ddx, ddy, ddz, a);
    }
    a.vo[c] = ro;
    ro = 0.0;
    if (ini && inj && k >= 3 - nbw && k <= kk - 3 + ntw) {
        const double ax = ddy[j] * dz[k], ay = ddx[i] * dz[k], az = ddx[i] * ddy[j];
        const double fe = ax * (ut[c] + ut[c + 1]) * 0.5;
        const double fw = ax * (ut[c - si] + ut[c - si + 1]) * 0.5;
        const double fn = ay * (vt[c] + vt[c + 1]) * 0.5;
        const double fs = ay * (vt[c - sj] + vt[c - sj + 1]) * 0.5;
        const double ft = az * (wt[c] + (wt[c + 1] - wt[c]) * 0.5 * dz[k] / ddz[k + 1]);
        const double fb = az * (wt[c - 1] + (wt[c] - wt[c - 1]) * 0.5 * dz[k - 1] / ddz[k]);
        double qe = 0.5 * fe * (W(0) + W(si));
        double qw = 0.5 * fw * (W(-si) + W(0));
        double qn = 0.5 * fn * (W(0) + W(sj));
        double qs = 0.5 * fs * (W(-sj) + W(0));
        double qt = 0.5 * ft * (W(0) + W(1));
        double qb = 0.5 * fb * (W(-1) + W(0));
        ro = -((qe - qw + qn - qs + qt - qb) * rddx[i] * rddy[j] * rdz[k]);
        if (nos) ro = wall_terms(ro, 2, k, j, i, B.bc, kk, jj, ii, W(0), ddx, ddy, ddz, a);
        const double ge = gpair(G(0), G(si), gmol) + gpair(G(1), G(1 + si), gmol);
        const double gw = gpair(G(-si), G(0), gmol) + gpair(G(1 - si), G(1), gmol);
        const double gn = gpair(G(0), G(sj), gmol) + gpair(G(1), G(1 + sj), gmol);
        const double gs = gpair(G(-sj), G(0), gmol) + gpair(G(1 - sj), G(1), gmol);
        const double gt = G(1), gb = G(0);
        qe = -(ge * ax * rdx[i] * (W(si) - W(0)));
        qw = -(gw * ax * rdx[i - 1] * (W(0) - W(-si)));
        qn = -(gn * ay * rdy[j] * (W(sj) - W(0)));
        qs = -(gs * ay * rdy[j - 1] * (W(0) - W(-sj)));
        qt = -(gt * az * rddz[k + 1] * (W(1) - W(0)));
        qb = -(gb * az * rddz[k] * (W(0) - W(-1)));
        const double st = ((ge * (U(1) - U(0))) - (gw * (U(1 - si) - U(-si)))) * ddy[j]
                          + ((gn * (V(1) - V(0))) - (gs * (V(1 - sj) - V(-sj)))) * ddx[i]
                          + ((gt * (W(1) - W(0)) * rddz[k + 1]) - (gb * (W(0) - W(-1)) * rddz[k])) * az;
        const double qc = st * iles;
        const double fak = rrho * rddx[i] * rddy[j] * rdz[k];
        ro = ro - fak * (qe - qw + qn - qs + qt - qb - qc);
        ro = ro - rpz[k] * (p[c + 1] - p[c] + a.gp[2] * dz[k]);
    } else if (nos) {
        ro = wall_terms(ro, 2, k, j, i, B.bc, kk, jj, ii, W(0), ddx, ddy, ddz, a);
    }
    a.wo[c] = ro;
#undef U
#undef V
#undef W
#undef G
}

void launch_tstle4(const Topo &t, const MomArgs &a, hipStream_t s) {
    MGP_RANGE("mgp:tstle4");
    const int maxc = (t.max_ni + 4) * (t.max_nj + 4) * (t.max_nk + 4);
    const int nchunk = (maxc + kMomThreads - 1) / kMomThreads;
    // launches of at most 2^31 work-items
    const int per = (int)std::max<int64_t>(1, (INT64_C(1) << 31) / ((int64_t)nchunk * kMomThreads));
    for (int first = 0; first < t.nlist; first += per) {
        const int nb = std::min(per, t.nlist - first);
        k_tstle4<<<(unsigned)((int64_t)nchunk * nb), kMomThreads, 0, s>>>(t, a, nchunk, first);
    }
}

// ---------------------------------------------------------------------
// tstle4_par (src/flow/tstle4_mod.F90:532-1753): the convection terms next
// to the PAR faces of a refined grid, added to uo, vo, wo after tstle4 (the
// reference calls it per grid after kon, diff and gradp, :114-116; it reads
// only u, v, w and writes only the grid's own uo, vo, wo, so it can follow
// the whole tstle4 launch).  The reference writes every face out by hand;
// all instances share one form, evaluated here for a face with normal axis
// n, side lo (faces 1, 3, 5) or hi (2, 4, 6) and tangential momentum
// component c, with o the third axis:
//  1. upwind at the second layer (:571-776), faces 1..6: along n the
//     positions iu = 4 (lo) or N-2 (hi), f = area (t(iu-1; a) +
//     t(iu-1; a+1)) 0.5, q = 0.5 (f -|f|) (0.5 c(iu) - 1.5 c(iu-1) +
//     c(iu-2)) 0.5 0.5 (lo) or 0.5 (f +|f|) (0.5 c(iu-1) - 1.5 c(iu) +
//     c(iu+1)) 0.5 0.5 (hi); co(iu) += F (-q), co(iu-1) += F q;
//  2. the momentum-conserving flux (:778-1750), faces 2, 1, 6, 5, 4, 3:
//     the standard flux qs = 0.5 f (c(L) + c(G)) minus the coarse-consistent
//     flux qn, whose f is the mean of the normal transport velocity over the
//     2 x 4 fine faces of one coarse face (pairs (b, b+1) = (3, 4), (5, 6),
//     ..; even a; the outermost a-1 or a+2 replaced by doubling the inner
//     end at a = 2 and Nc-2), odd a taking the mean of its even neighbours;
//     co(L) += F (qs - qn) (lo) or F (qn - qs) (hi).
// area = d_c(a) dd_o(b), F = -(rX rY) rZ of the updated cell (rX = rdx for
// u else rddx, ...), t the normal component itself (no immersed boundary).
// One workgroup per grid with PAR faces walks the faces in the reference's
// order, a barrier between faces (edge cells take terms of several faces,
// in that order); inside a face every position updates its own cells.
struct ParGrid {
    int n[3];                          // ii, jj, kk
    int64_t st[3];                     // strides of i, j, k
    double *out[3];
    const double *f[3];
    const double *d[3], *dd[3], *rd[3], *rdd[3];   // 1-based
    __device__ __forceinline__ int64_t ix(int a0, int v0, int a1, int v1, int a2, int v2) const {
        return (int64_t)(v0 - 1) * st[a0] + (int64_t)(v1 - 1) * st[a1] + (int64_t)(v2 - 1) * st[a2];
    }
    // F of component comp at the position x[a0] = v0, x[a1] = v1, x[a2] = v2
    __device__ __forceinline__ double fac(int comp, int a0, int v0, int a1, int v1, int a2,
                                          int v2) const {
        int x[3];
        x[a0] = v0;
        x[a1] = v1;
        x[a2] = v2;
        const double rx = comp == 0 ? rd[0][x[0]] : rdd[0][x[0]];
        const double ry = comp == 1 ? rd[1][x[1]] : rdd[1][x[1]];
        const double rz = comp == 2 ? rd[2][x[2]] : rdd[2][x[2]];
        return -1.0 * rx * ry * rz;
    }
};

__device__ void par_upwind(const ParGrid &P, int n, int hi) {
    const int N = P.n[n], iu = hi ? N - 2 : 4;
    const int c0 = n == 0 ? 1 : 0, c1 = n == 2 ? 1 : 2;   // the tangential components
    const int na0 = P.n[c0] - 4, na1 = P.n[c1] - 4;
    const int nb0 = P.n[3 - n - c0] - 4, nb1 = P.n[3 - n - c1] - 4;
    const int tot0 = na0 * nb0, tot = tot0 + na1 * nb1;
    const double *tn = P.f[n];
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const bool second = e >= tot0;
        const int c = second ? c1 : c0, o = 3 - n - c;
        const int r = second ? e - tot0 : e, nb = second ? nb1 : nb0;
        const int a = 3 + r / nb, b = 3 + r % nb;
        const double *cf = P.f[c];
        const double area = P.d[c][a] * P.dd[o][b];
        const double f = area * (tn[P.ix(n, iu - 1, c, a, o, b)] +
                                 tn[P.ix(n, iu - 1, c, a + 1, o, b)]) * 0.5;
        double q;
        if (!hi)
            q = 0.5 * (f - fabs(f)) *
                (0.5 * cf[P.ix(n, iu, c, a, o, b)] - 1.5 * cf[P.ix(n, iu - 1, c, a, o, b)] +
                 cf[P.ix(n, iu - 2, c, a, o, b)]) * 0.5 * 0.5;
        else
            q = 0.5 * (f + fabs(f)) *
                (0.5 * cf[P.ix(n, iu - 1, c, a, o, b)] - 1.5 * cf[P.ix(n, iu, c, a, o, b)] +
                 cf[P.ix(n, iu + 1, c, a, o, b)]) * 0.5 * 0.5;
        const double F = P.fac(c, n, iu, c, a, o, b);
        double *co = P.out[c];
        const int64_t x0 = P.ix(n, iu, c, a, o, b), x1 = P.ix(n, iu - 1, c, a, o, b);
        co[x0] = co[x0] + F * (-q);
        co[x1] = co[x1] + F * q;
    }
}

// qn at an even a (the pair of b containing b; 0 where no pair reaches b)
__device__ __forceinline__ double par_qn_even(const ParGrid &P, int n, int c, int o, int L,
                                              int G, int T, int a, int b) {
    const int Nc = P.n[c], No = P.n[o];
    const int b0 = ((b - 3) & 1) ? b - 1 : b;
    if (b0 > No - 3) return 0.0;
    const double *tn = P.f[n], *cf = P.f[c];
#define TN(x, y) tn[P.ix(n, T, c, x, o, y)]
    double sum;
    if (a == 2)
        sum = 2.0 * TN(a, b0) + 2.0 * TN(a, b0 + 1) + TN(a + 1, b0) + TN(a + 1, b0 + 1) +
              TN(a + 2, b0) + TN(a + 2, b0 + 1);
    else if (a == Nc - 2)
        sum = TN(a - 1, b0) + TN(a - 1, b0 + 1) + TN(a, b0) + TN(a, b0 + 1) +
              2.0 * TN(a + 1, b0) + 2.0 * TN(a + 1, b0 + 1);
    else
        sum = TN(a - 1, b0) + TN(a - 1, b0 + 1) + TN(a, b0) + TN(a, b0 + 1) + TN(a + 1, b0) +
              TN(a + 1, b0 + 1) + TN(a + 2, b0) + TN(a + 2, b0 + 1);
#undef TN
    const double f = P.d[c][a] * P.dd[o][b0] * sum * 0.125;
    return 0.5 * f * (cf[P.ix(n, L, c, a, o, b)] + cf[P.ix(n, G, c, a, o, b)]);
}

__device__ void par_conserve(const ParGrid &P, int n, int hi, int c0, int c1) {
    const int N = P.n[n];
    const int L = hi ? N - 2 : 3, G = hi ? N - 1 : 2, T = hi ? L : G;
    const int na0 = P.n[c0] - 3, na1 = P.n[c1] - 3;
    const int nb0 = P.n[3 - n - c0] - 4, nb1 = P.n[3 - n - c1] - 4;
    const int tot0 = na0 * nb0, tot = tot0 + na1 * nb1;
    const double *tn = P.f[n];
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const bool second = e >= tot0;
        const int c = second ? c1 : c0, o = 3 - n - c;
        const int r = second ? e - tot0 : e, nb = second ? nb1 : nb0;
        const int a = 2 + r / nb, b = 3 + r % nb;
        const double *cf = P.f[c];
        const double f = P.d[c][a] * P.dd[o][b] *
                         (tn[P.ix(n, T, c, a, o, b)] + tn[P.ix(n, T, c, a + 1, o, b)]) * 0.5;
        const double qs = 0.5 * f * (cf[P.ix(n, L, c, a, o, b)] + cf[P.ix(n, G, c, a, o, b)]);
        const double qn = (a & 1) ? 0.5 * (par_qn_even(P, n, c, o, L, G, T, a - 1, b) +
                                           par_qn_even(P, n, c, o, L, G, T, a + 1, b))
                                  : par_qn_even(P, n, c, o, L, G, T, a, b);
        const double F = P.fac(c, n, L, c, a, o, b);
        const double dq = hi ? qn - qs : qs - qn;
        double *co = P.out[c];
        const int64_t x = P.ix(n, L, c, a, o, b);
        co[x] = co[x] + F * dq;
    }
}

__global__ __launch_bounds__(kMomThreads) void k_tstle4_par(Topo t, MomArgs a,
                                                            const int *plist) {
    const Block &B = t.blocks[plist[blockIdx.x]];
    ParGrid P;
    P.n[0] = B.ni + 4;
    P.n[1] = B.nj + 4;
    P.n[2] = B.nk + 4;
    P.st[2] = 1;
    P.st[1] = P.n[2];
    P.st[0] = (int64_t)P.n[2] * P.n[1];
    P.out[0] = a.uo + B.sp;
    P.out[1] = a.vo + B.sp;
    P.out[2] = a.wo + B.sp;
    P.f[0] = a.u + B.sp;
    P.f[1] = a.v + B.sp;
    P.f[2] = a.w + B.sp;
    P.d[0] = a.dx + B.o1x - 1;  P.dd[0] = a.ddx + B.o1x - 1;
    P.d[1] = a.dy + B.o1y - 1;  P.dd[1] = a.ddy + B.o1y - 1;
    P.d[2] = a.dz + B.o1z - 1;  P.dd[2] = a.ddz + B.o1z - 1;
    P.rd[0] = a.rdx + B.o1x - 1;  P.rdd[0] = a.rddx + B.o1x - 1;
    P.rd[1] = a.rdy + B.o1y - 1;  P.rdd[1] = a.rddy + B.o1y - 1;
    P.rd[2] = a.rdz + B.o1z - 1;  P.rdd[2] = a.rddz + B.o1z - 1;
    for (int f = 0; f < 6; ++f)
        if ((B.par >> f) & 1) {
            par_upwind(P, f >> 1, f & 1);
            __syncthreads();
        }
    // faces 2, 1, 6, 5, 4, 3 with the component pairs (w, v), (u, w), (u, v)
    const int order[6] = {1, 0, 5, 4, 3, 2};
    for (int q = 0; q < 6; ++q) {
        const int f = order[q], n = f >> 1;
        if (!((B.par >> f) & 1)) continue;
        par_conserve(P, n, f & 1, n == 0 ? 2 : 0, n == 2 ? 1 : (n == 0 ? 1 : 2));
        __syncthreads();
    }
}

void launch_tstle4_par(const Topo &t, const MomArgs &a, const int *plist, int npar,
                       hipStream_t s) {
    MGP_RANGE("mgp:tstle4_par");
    if (npar > 0) k_tstle4_par<<<npar, kMomThreads, 0, s>>>(t, a, plist);
}

// src/core/rungekutta_mod.F90:315-333 over every stored position
__global__ void k_rkstep(double *p, double *dp, const double *rhsp,
                         int64_t n, double frhs, double dtfu) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const double d = frhs * dp[e] + rhsp[e];
        dp[e] = d;
        p[e] = p[e] + dtfu * d;
    }
}

void launch_rkstep(double *p, double *dp, const double *rhsp, int64_t n,
                   double frhs, double dtfu, hipStream_t s) {
    MGP_RANGE("mgp:rkstep");
    if (n <= 0) return;
    int64_t nb = (n + kMomThreads - 1) / kMomThreads;
    if (nb > 65536) nb = 65536;
    k_rkstep<<<(unsigned)nb, kMomThreads, 0, s>>>(p, dp, rhsp, n, frhs, dtfu);
}

}  // namespace mgp
