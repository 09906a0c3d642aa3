// mgp_bound.hip -- the closing sync of mgpoisl on gfx950: the velocity
// parent (PAR face buffers of U, V, W from the parent grid) and bound_flow
// (ghost values of U, V, W, P at FIX, OP1, NOS, SLI and PAR faces), which
// the reference runs per level before the closing exchange
// (src/flow/pressuresolver_mod.F90:338-348).
//
// Fields are in the reference layout (k fastest, two ghost layers).  One
// workgroup owns one grid.  The faces of a grid are processed in the
// reference's order 1..6 with a barrier between them: neighbouring faces
// write the same edge positions, and a later face reads values an earlier
// one wrote (src/core/bound_mod.F90:36-76).  Inside a face the positions
// are independent, except that the OP1 pressure reads the face's new
// velocities (a barrier before the pressure phase).  Expressions keep the
// reference's evaluation order (the build has no contraction), so the
// results are bitwise the oracle's.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mgp_device.h"
#include "mgp_launch.h"

namespace mgp {

namespace {

constexpr int kBoundThreads = 256;
constexpr int kFIX = 2, kOP1 = 3, kNOS = 5, kSLI = 6, kPAR = 8;

struct Grid3 {
    double *u, *v, *w, *p;
    const double *bp;   // BP of the grid (reference layout) or null (1)
    int kk, jj, ii;
    __device__ __forceinline__ int64_t ix(int k, int j, int i) const {
        return (int64_t)(k - 1) + (int64_t)(j - 1) * kk + (int64_t)(i - 1) * kk * jj;
    }
    __device__ __forceinline__ double b(int k, int j, int i) const {
        return bp ? bp[ix(k, j, i)] : 1.0;
    }
};

// bfront (src/flow/bound_flow_mod.F90:23-212): faces 1 (f = 0) and 2.
__device__ void bfront(int f, int ctyp, const Grid3 &G, const double *ub, const double *vb,
                       const double *wb, double pinf, const double *ddy, const double *ddz,
                       double rho) {
    const int kk = G.kk, jj = G.jj, ii = G.ii;
    const int i2 = f == 0 ? 2 : ii - 1, i3 = f == 0 ? 3 : ii - 2, i4 = f == 0 ? 4 : ii - 3;
    const int istag1 = f == 0 ? 1 : ii - 1, istag2 = f == 0 ? 2 : ii - 2;
    const double dir = f == 0 ? -1.0 : 1.0;
    double *u = G.u, *v = G.v, *w = G.w, *p = G.p;
    auto BF = [&](const double *b, int k, int j) { return b[(k - 1) + (int64_t)(j - 1) * kk]; };
    const int t = threadIdx.x, nt = blockDim.x;
    switch (ctyp) {
    case kFIX:
        for (int e = t; e < jj * kk; e += nt) {
            const int k = 1 + e % kk, j = 1 + e / kk;
            u[G.ix(k, j, istag2)] = BF(ub, k, j);
            v[G.ix(k, j, i2)] = 2.0 * BF(vb, k, j) - v[G.ix(k, j, i3)];
            w[G.ix(k, j, i2)] = 2.0 * BF(wb, k, j) - w[G.ix(k, j, i3)];
        }
        break;
    case kOP1:
        for (int e = t; e < (jj - 1) * (kk - 1); e += nt) {
            const int k = 1 + e % (kk - 1), j = 1 + e / (kk - 1);
            u[G.ix(k, j, istag1)] = u[G.ix(k, j, istag2)];
            double flag = copysign(1.0, dir * (u[G.ix(k, j, istag2)] + u[G.ix(k, j + 1, istag2)]));
            flag = 0.5 * (flag + 1.0);
            v[G.ix(k, j, i2)] = flag * v[G.ix(k, j, i3)] +
                                (1.0 - flag) * (2.0 * BF(vb, k, j) - v[G.ix(k, j, i3)]);
            flag = copysign(1.0, dir * (u[G.ix(k, j, istag2)] + u[G.ix(k + 1, j, istag2)]));
            flag = 0.5 * (flag + 1.0);
            w[G.ix(k, j, i2)] = flag * w[G.ix(k, j, i3)] +
                                (1.0 - flag) * (2.0 * BF(wb, k, j) - w[G.ix(k, j, i3)]);
        }
        break;
    case kNOS:
        for (int e = t; e < jj * kk; e += nt) {
            const int k = 1 + e % kk, j = 1 + e / kk;
            u[G.ix(k, j, istag2)] = 0.0;
            v[G.ix(k, j, i2)] = 0.0;
            w[G.ix(k, j, i2)] = 0.0;
        }
        break;
    case kSLI:
        for (int e = t; e < jj * kk; e += nt) {
            const int k = 1 + e % kk, j = 1 + e / kk;
            u[G.ix(k, j, istag2)] = 0.0;
            v[G.ix(k, j, i2)] = v[G.ix(k, j, i3)];
            w[G.ix(k, j, i2)] = w[G.ix(k, j, i3)];
        }
        break;
    case kPAR: {
        // tangential components (positions 2..n-1) and the normal one (2x2
        // groups from 3): disjoint writes, no barrier between them
        const int nt1 = (jj - 2) * (kk - 2);
        for (int e = t; e < nt1; e += nt) {
            const int k = 2 + e % (kk - 2), j = 2 + e / (kk - 2);
            const double sbv = G.b(k, j, i2) * G.b(k, j + 1, i2);
            const double sbw = G.b(k, j, i2) * G.b(k + 1, j, i2);
            v[G.ix(k, j, i2)] = (BF(vb, k, j) - 0.5 * (v[G.ix(k, j, i3)] - v[G.ix(k, j, i4)])) * sbv +
                                (1.0 - sbv) * v[G.ix(k, j, i2)];
            w[G.ix(k, j, i2)] = (BF(wb, k, j) - 0.5 * (w[G.ix(k, j, i3)] - w[G.ix(k, j, i4)])) * sbw +
                                (1.0 - sbw) * w[G.ix(k, j, i2)];
        }
        const int nk2 = (kk - 3) / 2, nj2 = (jj - 3) / 2;
        for (int e = t; e < nk2 * nj2; e += nt) {
            const int k = 3 + 2 * (e % nk2), j = 3 + 2 * (e / nk2);
            const double sbu = G.b(k, j, i2);
            double sb11, sb12, sb13, sb14;
            if (sbu > 0.5) {
                sb11 = G.b(k, j, i2) * G.b(k, j, i3);
                sb12 = G.b(k, j + 1, i2) * G.b(k, j + 1, i3);
                sb13 = G.b(k + 1, j, i2) * G.b(k + 1, j, i3);
                sb14 = G.b(k + 1, j + 1, i2) * G.b(k + 1, j + 1, i3);
            } else {
                sb11 = fmin(G.b(k, j, i3) + G.b(k, j, i2) + G.b(k, j, i4) + G.b(k, j - 1, i3) +
                            G.b(k, j + 1, i3) + G.b(k - 1, j, i3) + G.b(k + 1, j, i3), 1.0);
                sb12 = fmin(G.b(k, j + 1, i3) + G.b(k, j + 1, i2) + G.b(k, j + 1, i4) +
                            G.b(k, j, i3) + G.b(k, j + 2, i3) + G.b(k - 1, j + 1, i3) +
                            G.b(k + 1, j + 1, i3), 1.0);
                sb13 = fmin(G.b(k + 1, j, i3) + G.b(k + 1, j, i2) + G.b(k + 1, j, i4) +
                            G.b(k + 1, j - 1, i3) + G.b(k + 1, j + 1, i3) + G.b(k, j, i3) +
                            G.b(k + 2, j, i3), 1.0);
                sb14 = fmin(G.b(k + 1, j + 1, i3) + G.b(k + 1, j + 1, i2) + G.b(k + 1, j + 1, i4) +
                            G.b(k + 1, j, i3) + G.b(k + 1, j + 2, i3) + G.b(k, j + 1, i3) +
                            G.b(k + 2, j + 1, i3), 1.0);
            }
            const double dy0 = ddy[j - 1], dy1 = ddy[j], dz0 = ddz[k - 1], dz1 = ddz[k];
            double fak = 1.0 / ((dy0 + dy1) * (dz0 + dz1)) *
                         (sb11 * dy0 * dz0 + sb12 * dy1 * dz0 + sb13 * dy0 * dz1 + sb14 * dy1 * dz1);
            if (fak < 0.1) fak = 1.0;
            fak = 1.0 / fak;
            const double u11 = u[G.ix(k, j, istag2)], u12 = u[G.ix(k, j + 1, istag2)];
            const double u13 = u[G.ix(k + 1, j, istag2)], u14 = u[G.ix(k + 1, j + 1, istag2)];
            const double ubfine = (u11 * (1.0 - sb11) * dy0 * dz0 + u12 * (1.0 - sb12) * dy1 * dz0 +
                                   u13 * (1.0 - sb13) * dy0 * dz1 + u14 * (1.0 - sb14) * dy1 * dz1) /
                                  ((dy0 + dy1) * (dz0 + dz1));
            u[G.ix(k, j, istag2)] = (BF(ub, k, j) - ubfine) * fak * sb11 + (1.0 - sb11) * u11 * sbu;
            u[G.ix(k, j + 1, istag2)] = (BF(ub, k, j + 1) - ubfine) * fak * sb12 + (1.0 - sb12) * u12 * sbu;
            u[G.ix(k + 1, j, istag2)] = (BF(ub, k + 1, j) - ubfine) * fak * sb13 + (1.0 - sb13) * u13 * sbu;
            u[G.ix(k + 1, j + 1, istag2)] = (BF(ub, k + 1, j + 1) - ubfine) * fak * sb14 +
                                            (1.0 - sb14) * u14 * sbu;
        }
        break;
    }
    }
    __syncthreads();
    if (ctyp == kOP1) {
        for (int e = t; e < (jj - 1) * (kk - 1); e += nt) {
            const int k = 2 + e % (kk - 1), j = 2 + e / (kk - 1);
            const double vi = 0.25 * (v[G.ix(k, j - 1, i2)] + v[G.ix(k, j, i2)] +
                                      v[G.ix(k, j - 1, i3)] + v[G.ix(k, j, i3)]);
            const double wi = 0.25 * (w[G.ix(k - 1, j, i2)] + w[G.ix(k, j, i2)] +
                                      w[G.ix(k - 1, j, i3)] + w[G.ix(k, j, i3)]);
            const double us = u[G.ix(k, j, istag2)];
            const double umagsqr = us * us + vi * vi + wi * wi;
            p[G.ix(k, j, i2)] = pinf + fmin(0.0, 0.5 * rho * copysign(umagsqr, dir * us));
        }
    } else if (ctyp == kFIX || ctyp == kNOS || ctyp == kSLI) {
        for (int e = t; e < jj * kk; e += nt) {
            const int k = 1 + e % kk, j = 1 + e / kk;
            p[G.ix(k, j, i2)] = p[G.ix(k, j, i3)];
        }
    }
}

// bright (src/flow/bound_flow_mod.F90:215-494): faces 3 (f = 2) and 4.
// The tangential w of a PAR face is extrapolated from w(j2) (as written in
// the reference; bfront and bbottom use the j3/k3 cell).
__device__ void bright(int f, int ctyp, const Grid3 &G, const double *ub, const double *vb,
                       const double *wb, double pinf, const double *ddx, const double *ddz,
                       double rho) {
    const int kk = G.kk, jj = G.jj, ii = G.ii;
    const int j2 = f == 2 ? 2 : jj - 1, j3 = f == 2 ? 3 : jj - 2, j4 = f == 2 ? 4 : jj - 3;
    const int jstag1 = f == 2 ? 1 : jj - 1, jstag2 = f == 2 ? 2 : jj - 2;
    const double dir = f == 2 ? -1.0 : 1.0;
    double *u = G.u, *v = G.v, *w = G.w, *p = G.p;
    auto BF = [&](const double *b, int k, int i) { return b[(k - 1) + (int64_t)(i - 1) * kk]; };
    const int t = threadIdx.x, nt = blockDim.x;
    switch (ctyp) {
    case kFIX:
        for (int e = t; e < ii * kk; e += nt) {
            const int k = 1 + e % kk, i = 1 + e / kk;
            u[G.ix(k, j2, i)] = 2.0 * BF(ub, k, i) - u[G.ix(k, j3, i)];
            v[G.ix(k, jstag2, i)] = BF(vb, k, i);
            w[G.ix(k, j2, i)] = 2.0 * BF(wb, k, i) - w[G.ix(k, j3, i)];
        }
        break;
    case kOP1:
        for (int e = t; e < (ii - 1) * (kk - 1); e += nt) {
            const int k = 1 + e % (kk - 1), i = 1 + e / (kk - 1);
            double flag = copysign(1.0, dir * (v[G.ix(k, jstag2, i)] + v[G.ix(k, jstag2, i + 1)]));
            flag = 0.5 * (flag + 1.0);
            u[G.ix(k, j2, i)] = flag * u[G.ix(k, j3, i)] +
                                (1.0 - flag) * (2.0 * BF(ub, k, i) - u[G.ix(k, j3, i)]);
            v[G.ix(k, jstag1, i)] = v[G.ix(k, jstag2, i)];
            flag = copysign(1.0, dir * (v[G.ix(k, jstag2, i)] + v[G.ix(k + 1, jstag2, i)]));
            flag = 0.5 * (flag + 1.0);
            w[G.ix(k, j2, i)] = flag * w[G.ix(k, j3, i)] +
                                (1.0 - flag) * (2.0 * BF(wb, k, i) - w[G.ix(k, j3, i)]);
        }
        break;
    case kNOS:
        for (int e = t; e < ii * kk; e += nt) {
            const int k = 1 + e % kk, i = 1 + e / kk;
            u[G.ix(k, j2, i)] = 0.0;
            v[G.ix(k, jstag2, i)] = 0.0;
            w[G.ix(k, j2, i)] = 0.0;
        }
        break;
    case kSLI:
        for (int e = t; e < ii * kk; e += nt) {
            const int k = 1 + e % kk, i = 1 + e / kk;
            u[G.ix(k, j2, i)] = u[G.ix(k, j3, i)];
            v[G.ix(k, jstag2, i)] = 0.0;
            w[G.ix(k, j2, i)] = w[G.ix(k, j3, i)];
        }
        break;
    case kPAR: {
        const int nt1 = (ii - 2) * (kk - 2);
        for (int e = t; e < nt1; e += nt) {
            const int k = 2 + e % (kk - 2), i = 2 + e / (kk - 2);
            const double sbu = G.b(k, j2, i) * G.b(k, j2, i + 1);
            const double sbw = G.b(k, j2, i) * G.b(k + 1, j2, i);
            u[G.ix(k, j2, i)] = (BF(ub, k, i) - 0.5 * (u[G.ix(k, j3, i)] - u[G.ix(k, j4, i)])) * sbu +
                                (1.0 - sbu) * u[G.ix(k, j2, i)];
            w[G.ix(k, j2, i)] = (BF(wb, k, i) - 0.5 * (w[G.ix(k, j2, i)] - w[G.ix(k, j4, i)])) * sbw +
                                (1.0 - sbw) * w[G.ix(k, j2, i)];
        }
        const int nk2 = (kk - 3) / 2, ni2 = (ii - 3) / 2;
        for (int e = t; e < nk2 * ni2; e += nt) {
            const int k = 3 + 2 * (e % nk2), i = 3 + 2 * (e / nk2);
            const double sbv = G.b(k, j2, i);
            double sb11, sb12, sb13, sb14;
            if (sbv > 0.5) {
                sb11 = G.b(k, j2, i) * G.b(k, j3, i);
                sb12 = G.b(k, j2, i + 1) * G.b(k, j3, i + 1);
                sb13 = G.b(k + 1, j2, i) * G.b(k + 1, j3, i);
                sb14 = G.b(k + 1, j2, i + 1) * G.b(k + 1, j3, i + 1);
            } else {
                sb11 = fmin(G.b(k, j3, i) + G.b(k, j2, i) + G.b(k, j4, i) + G.b(k, j3, i - 1) +
                            G.b(k, j3, i + 1) + G.b(k - 1, j3, i) + G.b(k + 1, j3, i), 1.0);
                sb12 = fmin(G.b(k, j3, i + 1) + G.b(k, j2, i + 1) + G.b(k, j4, i + 1) +
                            G.b(k, j3, i) + G.b(k, j3, i + 2) + G.b(k - 1, j3, i + 1) +
                            G.b(k + 1, j3, i + 1), 1.0);
                sb13 = fmin(G.b(k + 1, j3, i) + G.b(k + 1, j2, i) + G.b(k + 1, j4, i) +
                            G.b(k + 1, j3, i - 1) + G.b(k + 1, j3, i + 1) + G.b(k, j3, i) +
                            G.b(k + 2, j3, i), 1.0);
                sb14 = fmin(G.b(k + 1, j3, i + 1) + G.b(k + 1, j2, i + 1) + G.b(k + 1, j4, i + 1) +
                            G.b(k + 1, j3, i) + G.b(k + 1, j3, i + 2) + G.b(k, j3, i + 1) +
                            G.b(k + 2, j3, i + 1), 1.0);
            }
            const double dz0 = ddz[k - 1], dz1 = ddz[k], dx0 = ddx[i - 1], dx1 = ddx[i];
            double fak = 1.0 / ((dz0 + dz1) * (dx0 + dx1)) *
                         (sb11 * dz0 * dx0 + sb12 * dz0 * dx1 + sb13 * dz1 * dx0 + sb14 * dz1 * dx1);
            if (fak < 0.1) fak = 1.0;
            fak = 1.0 / fak;
            const double v11 = v[G.ix(k, jstag2, i)], v12 = v[G.ix(k, jstag2, i + 1)];
            const double v13 = v[G.ix(k + 1, jstag2, i)], v14 = v[G.ix(k + 1, jstag2, i + 1)];
            const double vbfine = (v11 * (1.0 - sb11) * dz0 * dx0 + v12 * (1.0 - sb12) * dz0 * dx1 +
                                   v13 * (1.0 - sb13) * dz1 * dx0 + v14 * (1.0 - sb14) * dz1 * dx1) /
                                  ((dz0 + dz1) * (dx0 + dx1));
            v[G.ix(k, jstag2, i)] = (BF(vb, k, i) - vbfine) * fak * sb11 + v11 * (1.0 - sb11) * sbv;
            v[G.ix(k, jstag2, i + 1)] = (BF(vb, k, i + 1) - vbfine) * fak * sb12 + v12 * (1.0 - sb12) * sbv;
            v[G.ix(k + 1, jstag2, i)] = (BF(vb, k + 1, i) - vbfine) * fak * sb13 + v13 * (1.0 - sb13) * sbv;
            v[G.ix(k + 1, jstag2, i + 1)] = (BF(vb, k + 1, i + 1) - vbfine) * fak * sb14 +
                                            v14 * (1.0 - sb14) * sbv;
        }
        break;
    }
    }
    __syncthreads();
    if (ctyp == kOP1) {
        for (int e = t; e < (ii - 1) * (kk - 1); e += nt) {
            const int k = 2 + e % (kk - 1), i = 2 + e / (kk - 1);
            const double ui = 0.25 * (u[G.ix(k, j2, i - 1)] + u[G.ix(k, j2, i)] +
                                      u[G.ix(k, j3, i - 1)] + u[G.ix(k, j3, i)]);
            const double wi = 0.25 * (w[G.ix(k - 1, j2, i)] + w[G.ix(k, j2, i)] +
                                      w[G.ix(k - 1, j3, i)] + w[G.ix(k, j3, i)]);
            const double vs = v[G.ix(k, jstag2, i)];
            const double umagsqr = ui * ui + vs * vs + wi * wi;
            p[G.ix(k, j2, i)] = pinf + fmin(0.0, 0.5 * rho * copysign(umagsqr, dir * vs));
        }
    } else if (ctyp == kFIX || ctyp == kNOS || ctyp == kSLI) {
        for (int e = t; e < ii * kk; e += nt) {
            const int k = 1 + e % kk, i = 1 + e / kk;
            p[G.ix(k, j2, i)] = p[G.ix(k, j3, i)];
        }
    }
}

// bbottom (src/flow/bound_flow_mod.F90:497-731): faces 5 (f = 4) and 6.
__device__ void bbottom(int f, int ctyp, const Grid3 &G, const double *ub, const double *vb,
                        const double *wb, double pinf, const double *ddx, const double *ddy,
                        double rho) {
    const int kk = G.kk, jj = G.jj, ii = G.ii;
    const int k2 = f == 4 ? 2 : kk - 1, k3 = f == 4 ? 3 : kk - 2, k4 = f == 4 ? 4 : kk - 3;
    const int kstag1 = f == 4 ? 1 : kk - 1, kstag2 = f == 4 ? 2 : kk - 2;
    const double dir = f == 4 ? -1.0 : 1.0;
    double *u = G.u, *v = G.v, *w = G.w, *p = G.p;
    auto BF = [&](const double *b, int j, int i) { return b[(j - 1) + (int64_t)(i - 1) * jj]; };
    const int t = threadIdx.x, nt = blockDim.x;
    switch (ctyp) {
    case kFIX:
        for (int e = t; e < ii * jj; e += nt) {
            const int j = 1 + e % jj, i = 1 + e / jj;
            u[G.ix(k2, j, i)] = 2.0 * BF(ub, j, i) - u[G.ix(k3, j, i)];
            v[G.ix(k2, j, i)] = 2.0 * BF(vb, j, i) - v[G.ix(k3, j, i)];
            w[G.ix(kstag2, j, i)] = BF(wb, j, i);
        }
        break;
    case kOP1:
        for (int e = t; e < (ii - 1) * (jj - 1); e += nt) {
            const int j = 1 + e % (jj - 1), i = 1 + e / (jj - 1);
            double flag = copysign(1.0, dir * (w[G.ix(kstag2, j, i)] + w[G.ix(kstag2, j, i + 1)]));
            flag = 0.5 * (flag + 1.0);
            u[G.ix(k2, j, i)] = flag * u[G.ix(k3, j, i)] +
                                (1.0 - flag) * (2 * BF(ub, j, i) - u[G.ix(k3, j, i)]);
            flag = copysign(1.0, dir * (w[G.ix(kstag2, j, i)] + w[G.ix(kstag2, j + 1, i)]));
            flag = 0.5 * (flag + 1.0);
            v[G.ix(k2, j, i)] = flag * v[G.ix(k3, j, i)] +
                                (1.0 - flag) * (2.0 * BF(vb, j, i) - v[G.ix(k3, j, i)]);
            w[G.ix(kstag1, j, i)] = w[G.ix(kstag2, j, i)];
        }
        break;
    case kNOS:
        for (int e = t; e < ii * jj; e += nt) {
            const int j = 1 + e % jj, i = 1 + e / jj;
            u[G.ix(k2, j, i)] = 0.0;
            v[G.ix(k2, j, i)] = 0.0;
            w[G.ix(kstag2, j, i)] = 0.0;
        }
        break;
    case kSLI:
        for (int e = t; e < ii * jj; e += nt) {
            const int j = 1 + e % jj, i = 1 + e / jj;
            u[G.ix(k2, j, i)] = u[G.ix(k3, j, i)];
            v[G.ix(k2, j, i)] = v[G.ix(k3, j, i)];
            w[G.ix(kstag2, j, i)] = 0.0;
        }
        break;
    case kPAR: {
        const int nt1 = (ii - 2) * (jj - 2);
        for (int e = t; e < nt1; e += nt) {
            const int j = 2 + e % (jj - 2), i = 2 + e / (jj - 2);
            const double sbu = G.b(k2, j, i) * G.b(k2, j, i + 1);
            const double sbv = G.b(k2, j, i) * G.b(k2, j + 1, i);
            u[G.ix(k2, j, i)] = (BF(ub, j, i) - 0.5 * (u[G.ix(k3, j, i)] - u[G.ix(k4, j, i)])) * sbu +
                                (1.0 - sbu) * u[G.ix(k2, j, i)];
            v[G.ix(k2, j, i)] = (BF(vb, j, i) - 0.5 * (v[G.ix(k3, j, i)] - v[G.ix(k4, j, i)])) * sbv +
                                (1.0 - sbv) * v[G.ix(k2, j, i)];
        }
        const int nj2 = (jj - 3) / 2, ni2 = (ii - 3) / 2;
        for (int e = t; e < nj2 * ni2; e += nt) {
            const int j = 3 + 2 * (e % nj2), i = 3 + 2 * (e / nj2);
            const double sbw = G.b(k2, j, i);
            double sb11, sb12, sb13, sb14;
            if (sbw > 0.5) {
                sb11 = G.b(k2, j, i) * G.b(k3, j, i);
                sb12 = G.b(k2, j, i + 1) * G.b(k3, j, i + 1);
                sb13 = G.b(k2, j + 1, i) * G.b(k3, j + 1, i);
                sb14 = G.b(k2, j + 1, i + 1) * G.b(k3, j + 1, i + 1);
            } else {
                sb11 = fmin(G.b(k3, j, i) + G.b(k2, j, i) + G.b(k4, j, i) + G.b(k3, j, i - 1) +
                            G.b(k3, j, i + 1) + G.b(k3, j - 1, i) + G.b(k3, j + 1, i), 1.0);
                sb12 = fmin(G.b(k3, j, i + 1) + G.b(k2, j, i + 1) + G.b(k4, j, i + 1) +
                            G.b(k3, j, i) + G.b(k3, j, i + 2) + G.b(k3, j - 1, i + 1) +
                            G.b(k3, j + 1, i + 1), 1.0);
                sb13 = fmin(G.b(k3, j + 1, i) + G.b(k2, j + 1, i) + G.b(k4, j + 1, i) +
                            G.b(k3, j + 1, i - 1) + G.b(k3, j + 1, i + 1) + G.b(k3, j, i) +
                            G.b(k3, j + 2, i), 1.0);
                sb14 = fmin(G.b(k3, j + 1, i + 1) + G.b(k2, j + 1, i + 1) + G.b(k4, j + 1, i + 1) +
                            G.b(k3, j + 1, i) + G.b(k3, j + 1, i + 2) + G.b(k3, j, i + 1) +
                            G.b(k3, j + 2, i + 1), 1.0);
            }
            const double dy0 = ddy[j - 1], dy1 = ddy[j], dx0 = ddx[i - 1], dx1 = ddx[i];
            double fak = 1. / ((dy0 + dy1) * (dx0 + dx1)) *
                         (sb11 * dy0 * dx0 + sb12 * dy0 * dx1 + sb13 * dy1 * dx0 + sb14 * dy1 * dx1);
            if (fak < 0.1) fak = 1.0;
            fak = 1.0 / fak;
            const double w11 = w[G.ix(kstag2, j, i)], w12 = w[G.ix(kstag2, j, i + 1)];
            const double w13 = w[G.ix(kstag2, j + 1, i)], w14 = w[G.ix(kstag2, j + 1, i + 1)];
            const double wbfine = (w11 * (1.0 - sb11) * dy0 * dx0 + w12 * (1.0 - sb12) * dy0 * dx1 +
                                   w13 * (1.0 - sb13) * dy1 * dx0 + w14 * (1.0 - sb14) * dy1 * dx1) /
                                  ((dy0 + dy1) * (dx0 + dx1));
            w[G.ix(kstag2, j, i)] = (BF(wb, j, i) - wbfine) * fak * sb11 + w11 * (1.0 - sb11) * sbw;
            w[G.ix(kstag2, j, i + 1)] = (BF(wb, j, i + 1) - wbfine) * fak * sb12 + w12 * (1.0 - sb12) * sbw;
            w[G.ix(kstag2, j + 1, i)] = (BF(wb, j + 1, i) - wbfine) * fak * sb13 + w13 * (1.0 - sb13) * sbw;
            w[G.ix(kstag2, j + 1, i + 1)] = (BF(wb, j + 1, i + 1) - wbfine) * fak * sb14 +
                                            w14 * (1.0 - sb14) * sbw;
        }
        break;
    }
    }
    __syncthreads();
    if (ctyp == kOP1) {
        for (int e = t; e < (ii - 1) * (jj - 1); e += nt) {
            const int j = 2 + e % (jj - 1), i = 2 + e / (jj - 1);
            const double ui = 0.25 * (u[G.ix(k2, j, i - 1)] + u[G.ix(k2, j, i)] +
                                      u[G.ix(k3, j, i - 1)] + u[G.ix(k3, j, i)]);
            const double vi = 0.25 * (v[G.ix(k2, j - 1, i)] + v[G.ix(k2, j, i)] +
                                      v[G.ix(k3, j - 1, i)] + v[G.ix(k3, j, i)]);
            const double ws = w[G.ix(kstag2, j, i)];
            const double umagsqr = ui * ui + vi * vi + ws * ws;
            p[G.ix(k2, j, i)] = pinf + fmin(0.0, 0.5 * rho * copysign(umagsqr, dir * ws));
        }
    } else if (ctyp == kFIX || ctyp == kNOS || ctyp == kSLI) {
        for (int e = t; e < ii * jj; e += nt) {
            const int j = 1 + e % jj, i = 1 + e / jj;
            p[G.ix(k2, j, i)] = p[G.ix(k3, j, i)];
        }
    }
}

__device__ bool bound_type(int c) { return c == kFIX || c == kOP1 || c == kNOS || c == kSLI || c == kPAR; }

// bound_flow%bound(ilevel, u, v, w, p) over the launch's grids: faces in
// the order 1..6, the first boundary condition of each.
__global__ __launch_bounds__(kBoundThreads) void k_bound_flow(
    Topo t, double *u, double *v, double *w, double *p, const double *ubuf, const double *vbuf,
    const double *wbuf, const int64_t *bbo, const double *pinf, const double *bp,
    const double *ddx, const double *ddy, const double *ddz, double rho) {
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    Grid3 G;
    G.u = u + B.sp;
    G.v = v + B.sp;
    G.w = w + B.sp;
    G.p = p + B.sp;
    G.bp = bp ? bp + B.sp : nullptr;
    G.kk = B.nk + 4;
    G.jj = B.nj + 4;
    G.ii = B.ni + 4;
    for (int f = 0; f < 6; ++f) {
        const int ctyp = B.bc[f];
        if (!bound_type(ctyp)) continue;
        const int64_t o = bbo[6 * b + f];
        const double *ub = o >= 0 ? ubuf + o : nullptr, *vb = o >= 0 ? vbuf + o : nullptr,
                     *wb = o >= 0 ? wbuf + o : nullptr;
        const double pf = pinf[6 * b + f];
        if (f < 2)
            bfront(f, ctyp, G, ub, vb, wb, pf, ddy + B.o1y, ddz + B.o1z, rho);
        else if (f < 4)
            bright(f, ctyp, G, ub, vb, wb, pf, ddx + B.o1x, ddz + B.o1z, rho);
        else
            bbottom(f, ctyp, G, ub, vb, wb, pf, ddx + B.o1x, ddy + B.o1y, rho);
        __syncthreads();
    }
}

// parent(ilevel, u, v, w): the PAR face buffers of a fine grid from the
// coarse slab of start_and_stop with icomp (the normal component at a
// back face starts one coarse index lower, src/ib/parent/parent_core_mod.
// F90:213-267) through prolong1 / prolong2 (src/ib/parent/parent1_mod.F90:
// 515-583): fine face position (a, b) takes slab value (2 + (a-1)/2,
// 2 + (b-1)/2), and along a direction the component is staggered in
// (stag, :185-210) the odd positions take the mean with the slab value
// before.  At most one direction of a face is staggered for a component,
// so each output is one copy or one mean, as in the two passes of the
// reference.
__global__ __launch_bounds__(kBoundThreads) void k_parent_uvw(
    Topo t, const double *u, const double *v, const double *w, double *ubuf, double *vbuf,
    double *wbuf, const int64_t *bbo) {
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    if (B.parent < 0) return;
    const Block &P = t.blocks[B.parent];
    const int kk = B.nk + 4, jj = B.nj + 4, ii = B.ni + 4;
    const int kkc = P.nk + 4, jjc = P.nj + 4;
    const int ipos = B.pos[0], jpos = B.pos[1], kpos = B.pos[2];
    for (int f = 0; f < 6; ++f) {
        if (B.bc[f] != kPAR) continue;
        const int64_t o = bbo[6 * b + f];
        for (int comp = 1; comp <= 3; ++comp) {
            const double *fc = (comp == 1 ? u : comp == 2 ? v : w) + P.sp;
            double *buf = (comp == 1 ? ubuf : comp == 2 ? vbuf : wbuf) + o;
            const int icomp = (f < 2 && comp == 1) || (f >= 2 && f < 4 && comp == 2) ||
                              (f >= 4 && comp == 3) ? comp : 0;
            int ks = kpos - 2, js = jpos - 2, is = ipos - 2;
            switch (f) {
            case 0: is = ipos - 1; break;
            case 1: is = ipos + (ii - 4) / 2 - (icomp == 1); break;
            case 2: js = jpos - 1; break;
            case 3: js = jpos + (jj - 4) / 2 - (icomp == 2); break;
            case 4: ks = kpos - 1; break;
            default: ks = kpos + (kk - 4) / 2 - (icomp == 3); break;
            }
            int n1, n2, s1, s2;
            if (f < 2) {
                n1 = kk; n2 = jj; s1 = comp == 3; s2 = comp == 2;
            } else if (f < 4) {
                n1 = kk; n2 = ii; s1 = comp == 3; s2 = comp == 1;
            } else {
                n1 = jj; n2 = ii; s1 = comp == 2; s2 = comp == 1;
            }
            // slab value (x, y), 1-based in the packed face
            auto in = [&](int x, int y) -> double {
                int k, j, i;
                if (f < 2) { k = ks + x - 1; j = js + y - 1; i = is; }
                else if (f < 4) { k = ks + x - 1; j = js; i = is + y - 1; }
                else { k = ks; j = js + x - 1; i = is + y - 1; }
                return fc[(int64_t)(k - 1) + (int64_t)(j - 1) * kkc + (int64_t)(i - 1) * kkc * jjc];
            };
            auto tmp = [&](int a, int y) -> double {
                const int ac = 2 + (a - 1) / 2;
                return (s1 && (a & 1)) ? 0.5 * (in(ac, y) + in(ac - 1, y)) : in(ac, y);
            };
            for (int e = threadIdx.x; e < n1 * n2; e += blockDim.x) {
                const int a = 1 + e % n1, bb = 1 + e / n1;
                const int bc = 2 + (bb - 1) / 2;
                buf[e] = (s2 && (bb & 1)) ? 0.5 * (tmp(a, bc) + tmp(a, bc - 1)) : tmp(a, bc);
            }
        }
    }
}

}  // namespace

void launch_parent_uvw(const Topo &t, const double *u, const double *v, const double *w,
                       double *ubuf, double *vbuf, double *wbuf, const int64_t *bbo,
                       hipStream_t s) {
    MGP_RANGE("mgp:parent_uvw");
    if (t.nlist <= 0) return;
    k_parent_uvw<<<t.nlist, kBoundThreads, 0, s>>>(t, u, v, w, ubuf, vbuf, wbuf, bbo);
}

void launch_bound_flow(const Topo &t, double *u, double *v, double *w, double *p,
                       const double *ubuf, const double *vbuf, const double *wbuf,
                       const int64_t *bbo, const double *pinf, const double *bp,
                       const double *ddx, const double *ddy, const double *ddz, double rho,
                       hipStream_t s) {
    MGP_RANGE("mgp:bound_flow");
    if (t.nlist <= 0) return;
    k_bound_flow<<<t.nlist, kBoundThreads, 0, s>>>(t, u, v, w, p, ubuf, vbuf, wbuf, bbo, pinf,
                                                   bp, ddx, ddy, ddz, rho);
}

}  // namespace mgp
