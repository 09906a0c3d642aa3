// mgp_levels.hip -- coupling of grid levels (nested refinement).
//
// A grid above the coarsest level has a parent one level coarser; it covers
// n/2 parent cells per direction starting at the parent's 1-based index
// pos (iposition, jposition, kposition).  Faces of a refined grid without a
// same-level neighbour are PAR faces: their ghost values come from
// bound_pressure, which blends the grid's own interior with the parent's
// values across the face (face buffers filled by `parent`).
//
// Working fields live in WO layout (see mgp_internal.h); a block's PAR-face
// ghosts are kept in its own face region of `hpar` (same face layout as the
// boundary planes in `hface`), the parent's values per 2x2 face group at
// the group's first cell in `pbuf`.  The post-solve steps on DP, U, V, W, P
// run on the reference layout, where ghosts are stored in place.
//
// Every expression keeps the reference's evaluation order (the file is
// built with -ffp-contract=off), so FP64 results equal the oracle's.
#include <hip/hip_runtime.h>

#include "mgp_launch.h"

namespace mgp {

namespace {

constexpr int kThreadsLv = 256;

__device__ __forceinline__ void unpack_c(uint32_t c, int &i, int &j, int &k) {
    i = c & 1023;
    j = (c >> 10) & 1023;
    k = c >> 20;
}

// position of face cell (tangential coordinates of cell (i, j, k)) in the
// face region of a block
__device__ __forceinline__ int face_slot(int f, int i, int j, int k, int ni,
                                         int nj, int nk) {
    const int b = face_base(f, ni, nj, nk);
    return f < 2 ? b + fidx_x(k, j, nj, nk)
         : f < 4 ? b + fidx_y(k, i, ni, nk)
                 : b + fidx_z(j, i, ni, nj);
}

__device__ __forceinline__ int slot_at(const Topo &t, const Block &B, int i,
                                       int j, int k) {
    const Shape &S = t.shapes[B.shape];
    return t.slot_of[S.slot_off + k + B.nk * (j + B.nj * i)];
}

// Value of a WO field at 0-based (i, j, k) of block B, one index possibly
// in the ghost layer (-1 or n): interior from the field, ghosts from the
// neighbour's boundary plane (CON), the block's own PAR ghost, or zero.
template <typename T>
__device__ T wo_at(const Topo &t, const Block &B, const T *field,
                   const T *hface, const T *hpar, int i, int j, int k) {
    const int ni = B.ni, nj = B.nj, nk = B.nk;
    int f = -1;
    if (i < 0) f = 0;
    else if (i >= ni) f = 1;
    else if (j < 0) f = 2;
    else if (j >= nj) f = 3;
    else if (k < 0) f = 4;
    else if (k >= nk) f = 5;
    if (f < 0) return field[B.wo + slot_at(t, B, i, j, k)];
    if (B.nbr[f] >= 0) {
        const Block &N = t.blocks[B.nbr[f]];
        // the neighbour's opposite boundary plane, same tangential indices
        return hface[N.fc + face_slot(f ^ 1, i, j, k, N.ni, N.nj, N.nk)];
    }
    if ((B.par >> f) & 1) return hpar[B.fc + face_slot(f, i, j, k, ni, nj, nk)];
    return T(0);
}

// BP at 0-based (i, j, k) (ghosts from the block's own ghost layer)
__device__ __forceinline__ double bp_of(const Topo &t, const Block &B, Mask m,
                                        int i, int j, int k) {
    if (!m.bp || (B.flags & 1)) return 1.0;
    const int ni = B.ni, nj = B.nj, nk = B.nk;
    int f = -1;
    if (i < 0) f = 0;
    else if (i >= ni) f = 1;
    else if (j < 0) f = 2;
    else if (j >= nj) f = 3;
    else if (k < 0) f = 4;
    else if (k >= nk) f = 5;
    if (f < 0) return m.bp[B.wo + slot_at(t, B, i, j, k)];
    return m.bpf[B.fc + face_slot(f, i, j, k, ni, nj, nk)];
}

// pressureftocone_A / _B (src/flow/bound_pressure_mod.F90:536-608) and one
// 2x2 group of bfront / bright / bbottom and their _bp forms (:256-534).
// P(k, j, i) and BP(k, j, i) take 1-based reference indices; d*, dd* are
// 1-based metric rows; pbv is the parent's value for the group; put(k, j,
// i, v) stores one ghost value.  f is the 0-based face, (q1, q2) the
// group's 1-based first cell along the face's two tangential directions
// ((j, k) for x faces, (i, k) for y faces, (i, j) for z faces).
template <bool BPV, typename PF, typename BF, typename SF>
__device__ void bound_group(int f, int q1, int q2, int kk, int jj, int ii,
                            const double *dx, const double *dy,
                            const double *dz, const double *ddx,
                            const double *ddy, const double *ddz, double pbv,
                            PF P, BF BP, SF put) {
    int i, j, k;          // first cell of the 2x2x2 group (1-based)
    int g2, g3, gs2;      // ghost, first interior plane, staggered index
    if (f < 2) {
        g2 = f == 0 ? 2 : ii - 1;
        g3 = f == 0 ? 3 : ii - 2;
        const int g4 = f == 0 ? 4 : ii - 3;
        gs2 = f == 0 ? 2 : ii - 2;
        i = g3 < g4 ? g3 : g4;
        j = q1;
        k = q2;
    } else if (f < 4) {
        g2 = f == 2 ? 2 : jj - 1;
        g3 = f == 2 ? 3 : jj - 2;
        const int g4 = f == 2 ? 4 : jj - 3;
        gs2 = f == 2 ? 2 : jj - 2;
        j = g3 < g4 ? g3 : g4;
        i = q1;
        k = q2;
    } else {
        g2 = f == 4 ? 2 : kk - 1;
        g3 = f == 4 ? 3 : kk - 2;
        const int g4 = f == 4 ? 4 : kk - 3;
        gs2 = f == 4 ? 2 : kk - 2;
        k = g3 < g4 ? g3 : g4;
        i = q1;
        j = q2;
    }
    double pcnew, bpc = 1.0;
    if (!BPV) {
        double sump = 0.0, sumvol = 0.0;
        for (int l = 0; l <= 1; ++l)
            for (int m = 0; m <= 1; ++m)
                for (int n = 0; n <= 1; ++n) {
                    const double vol = ddz[k + n] * ddy[j + m] * ddx[i + l];
                    sump = sump + P(k + n, j + m, i + l) * vol;
                    sumvol = sumvol + vol;
                }
        pcnew = sump / sumvol;
    } else {
        double sumbp = 0.0;
        for (int l = 0; l <= 1; ++l)
            for (int m = 0; m <= 1; ++m)
                for (int n = 0; n <= 1; ++n)
                    sumbp = sumbp + BP(k + n, j + m, i + l);
        bpc = sumbp < 1.0 ? sumbp : 1.0;
        if (bpc < 0.5) {
            pcnew = 0.0;
        } else {
            double sump = 0.0, sumvol = 0.0;
            for (int l = 0; l <= 1; ++l)
                for (int m = 0; m <= 1; ++m)
                    for (int n = 0; n <= 1; ++n) {
                        const double vol = BP(k + n, j + m, i + l) * ddz[k + n] *
                                           ddy[j + m] * ddx[i + l];
                        sump = sump + P(k + n, j + m, i + l) * vol;
                        sumvol = sumvol + vol;
                    }
            pcnew = sump / sumvol;
        }
    }
    double fak = 1.0;
    if (BPV) {
        if (bpc < 0.5) pcnew = pbv;
        double sb11, sb12, sb13, sb14;
        if (f < 2) {
            sb11 = BP(k, j, g2) * BP(k, j, g3);
            sb12 = BP(k, j + 1, g2) * BP(k, j + 1, g3);
            sb13 = BP(k + 1, j, g2) * BP(k + 1, j, g3);
            sb14 = BP(k + 1, j + 1, g2) * BP(k + 1, j + 1, g3);
            fak = (sb11 * ddy[j] * ddz[k] + sb12 * ddy[j + 1] * ddz[k] +
                   sb13 * ddy[j] * ddz[k + 1] + sb14 * ddy[j + 1] * ddz[k + 1]) /
                  ((ddy[j] + ddy[j + 1]) * (ddz[k] + ddz[k + 1]));
        } else if (f < 4) {
            sb11 = BP(k, g2, i) * BP(k, g3, i);
            sb12 = BP(k, g2, i + 1) * BP(k, g3, i + 1);
            sb13 = BP(k + 1, g2, i) * BP(k + 1, g3, i);
            sb14 = BP(k + 1, g2, i + 1) * BP(k + 1, g3, i + 1);
            fak = (sb11 * ddx[i] * ddz[k] + sb12 * ddx[i + 1] * ddz[k] +
                   sb13 * ddx[i] * ddz[k + 1] + sb14 * ddx[i + 1] * ddz[k + 1]) /
                  ((ddx[i] + ddx[i + 1]) * (ddz[k] + ddz[k + 1]));
        } else {
            sb11 = BP(g2, j, i) * BP(g3, j, i);
            sb12 = BP(g2, j, i + 1) * BP(g3, j, i + 1);
            sb13 = BP(g2, j + 1, i) * BP(g3, j + 1, i);
            sb14 = BP(g2, j + 1, i + 1) * BP(g3, j + 1, i + 1);
            fak = (sb11 * ddx[i] * ddy[j] + sb12 * ddx[i + 1] * ddy[j] +
                   sb13 * ddx[i] * ddy[j + 1] + sb14 * ddx[i + 1] * ddy[j + 1]) /
                  ((ddx[i] + ddx[i + 1]) * (ddy[j] + ddy[j + 1]));
        }
        if (fak < 0.1) fak = 1.0;
        fak = 1.0 / fak;
    }
    if (f < 2) {
        for (int m = 0; m <= 1; ++m)
            for (int n = 0; n <= 1; ++n)
                put(k + n, j + m, g2, BPV
                    ? P(k + n, j + m, g3) + fak * dx[gs2] / (ddx[g3] + ddx[g2]) * (pbv - pcnew)
                    : P(k + n, j + m, g3) + dx[gs2] / (ddx[g3] + ddx[g2]) * (pbv - pcnew));
    } else if (f < 4) {
        for (int l = 0; l <= 1; ++l)
            for (int n = 0; n <= 1; ++n)
                put(k + n, g2, i + l, BPV
                    ? P(k + n, g3, i + l) + fak * dy[gs2] / (ddy[g3] + ddy[g2]) * (pbv - pcnew)
                    : P(k + n, g3, i + l) + dy[gs2] / (ddy[g3] + ddy[g2]) * (pbv - pcnew));
    } else {
        for (int l = 0; l <= 1; ++l)
            for (int m = 0; m <= 1; ++m)
                put(g2, j + m, i + l, BPV
                    ? P(g3, j + m, i + l) + fak * dz[gs2] / (ddz[g3] + ddz[g2]) * (pbv - pcnew)
                    : P(g3, j + m, i + l) + dz[gs2] / (ddz[g3] + ddz[g2]) * (pbv - pcnew));
    }
}

// groups of a block's PAR faces: face f has n1 x n2 groups of 2x2 cells
__device__ __forceinline__ void face_groups(const Block &B, int f, int &n1,
                                            int &n2) {
    if (f < 2) { n1 = B.nj / 2; n2 = B.nk / 2; }
    else if (f < 4) { n1 = B.ni / 2; n2 = B.nk / 2; }
    else { n1 = B.ni / 2; n2 = B.nj / 2; }
}

// 0-based coordinates of the first fine cell of group (a, b) on face f
__device__ __forceinline__ void group_cell(const Block &B, int f, int a, int b,
                                           int &i, int &j, int &k) {
    if (f < 2) { i = f == 0 ? 0 : B.ni - 1; j = 2 * a; k = 2 * b; }
    else if (f < 4) { j = f == 2 ? 0 : B.nj - 1; i = 2 * a; k = 2 * b; }
    else { k = f == 4 ? 0 : B.nk - 1; i = 2 * a; j = 2 * b; }
}

// 0-based parent coordinates of the parent cell across face f next to
// group (a, b) (start_and_stop, src/ib/parent/parent_core_mod.F90:214-262,
// with the prolongation's jc = 2 + (j-1)/2, parent1_mod.F90:523-583)
__device__ __forceinline__ void parent_cell(const Block &B, int f, int a,
                                            int b, int &ci, int &cj, int &ck) {
    const int pi = B.pos[0] - 3, pj = B.pos[1] - 3, pk = B.pos[2] - 3;
    if (f < 2) {
        ci = f == 0 ? pi - 1 : pi + B.ni / 2; cj = pj + a; ck = pk + b;
    } else if (f < 4) {
        cj = f == 2 ? pj - 1 : pj + B.nj / 2; ci = pi + a; ck = pk + b;
    } else {
        ck = f == 4 ? pk - 1 : pk + B.nk / 2; ci = pi + a; cj = pj + b;
    }
}

// Parent value at 0-based (ci, cj, ck), each index in [-1, n]: interior,
// face ghosts as in wo_at, other ghosts (edges, corners, faces without a
// neighbour or PAR source) from the parent's shell, which holds what ctof
// injected there (zero on the coarsest level).
template <typename T>
__device__ T parent_at(const Topo &t, const Block &P, const Fields<T> &F,
                       const T *hsh, int ci, int cj, int ck) {
    const int out = (ci < 0 || ci >= P.ni) + (cj < 0 || cj >= P.nj) +
                    (ck < 0 || ck >= P.nk);
    if (out == 0) return F.hilf[P.wo + slot_at(t, P, ci, cj, ck)];
    if (out == 1) {
        int f;
        if (ci < 0) f = 0;
        else if (ci >= P.ni) f = 1;
        else if (cj < 0) f = 2;
        else if (cj >= P.nj) f = 3;
        else if (ck < 0) f = 4;
        else f = 5;
        if (P.nbr[f] >= 0 || ((P.par >> f) & 1))
            return wo_at<T>(t, P, F.hilf, F.hface, F.hpar, ci, cj, ck);
    }
    return hsh[P.sh + shell_index(ci, cj, ck, P.ni, P.nj, P.nk)];
}

// ---------------------------------------------------------------------
// ctof (src/ib/ctof/ctof1_mod.F90:232-331): every fine cell and the first
// ghost layer <- parent cell (pos - 3 + floor((i + 2) / 2) - 1, ...).  The
// boundary planes go to `fout`, where the next sweep of the neighbours
// reads its ghosts; the first ghost layer goes to the shell `hsh` and is
// summed into `dsh` (DP accumulates HILF over the whole array,
// src/flow/pressuresolver_mod.F90:882-908).
template <typename T>
__global__ void k_ctof(Topo t, Fields<T> F, T *fout, T *hsh, double *dsh,
                       const int *done) {
    if (*done) return;
    T *hilf = F.hilf;
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const Shape &S = t.shapes[B.shape];
    const Block &P = t.blocks[B.parent];
    const uint32_t *cells = t.cells + S.cell_off;
    const int ni = B.ni, nj = B.nj, nk = B.nk;
    const int nsh = shell_size(ni, nj, nk);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nsh;
         e += gridDim.x * blockDim.x) {
        // decode the shell position e (inverse of shell_index)
        const int a = (nj + 2) * (nk + 2);
        int i, j, k;
        if (e < 2 * a) {
            i = e < a ? -1 : ni;
            const int r = e % a;
            j = r / (nk + 2) - 1;
            k = r % (nk + 2) - 1;
        } else {
            const int slab = 2 * (nk + 2) + 2 * nj;
            const int r0 = e - 2 * a;
            i = r0 / slab;
            const int r = r0 % slab;
            if (r < nk + 2) { j = -1; k = r - 1; }
            else if (r < 2 * (nk + 2)) { j = nj; k = r - (nk + 2) - 1; }
            else { j = (r - 2 * (nk + 2)) / 2; k = ((r - 2 * (nk + 2)) & 1) ? nk : -1; }
        }
        const T v = parent_at<T>(t, P, F, hsh, B.pos[0] - 4 + (i + 2) / 2,
                                 B.pos[1] - 4 + (j + 2) / 2, B.pos[2] - 4 + (k + 2) / 2);
        hsh[B.sh + e] = v;
        dsh[B.sh + e] = dsh[B.sh + e] + (double)v;
    }
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < S.ncell;
         s += gridDim.x * blockDim.x) {
        int i, j, k;
        unpack_c(cells[s], i, j, k);
        const T v = hilf[P.wo + slot_at(t, P, B.pos[0] - 3 + i / 2,
                                        B.pos[1] - 3 + j / 2, B.pos[2] - 3 + k / 2)];
        hilf[B.wo + s] = v;
        if (i == 0) fout[B.fc + face_slot(0, i, j, k, ni, nj, nk)] = v;
        if (i == ni - 1) fout[B.fc + face_slot(1, i, j, k, ni, nj, nk)] = v;
        if (j == 0) fout[B.fc + face_slot(2, i, j, k, ni, nj, nk)] = v;
        if (j == nj - 1) fout[B.fc + face_slot(3, i, j, k, ni, nj, nk)] = v;
        if (k == 0) fout[B.fc + face_slot(4, i, j, k, ni, nj, nk)] = v;
        if (k == nk - 1) fout[B.fc + face_slot(5, i, j, k, ni, nj, nk)] = v;
    }
}

// parent for HILF: per PAR face group, the parent's value across the face
template <typename T>
__global__ void k_parent(Topo t, const T *hilf, const T *hface,
                         const T *hpar, T *pbuf, const int *done) {
    if (*done) return;
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    if (!B.par) return;
    const Block &P = t.blocks[B.parent];
    for (int f = 0; f < 6; ++f) {
        if (!((B.par >> f) & 1)) continue;
        int n1, n2;
        face_groups(B, f, n1, n2);
        for (int e = threadIdx.x; e < n1 * n2; e += blockDim.x) {
            const int a = e / n2, c = e % n2;
            int ci, cj, ck, i, j, k;
            parent_cell(B, f, a, c, ci, cj, ck);
            group_cell(B, f, a, c, i, j, k);
            pbuf[B.fc + face_slot(f, i, j, k, B.ni, B.nj, B.nk)] =
                wo_at<T>(t, P, hilf, hface, hpar, ci, cj, ck);
        }
    }
}

// bound_pressure for HILF: the PAR ghosts of every block of the launch
template <typename T, bool BPV>
__global__ void k_bound(Topo t, const T *hilf, T *hpar, const T *pbuf,
                        Mask m, const double *dx_, const double *dy_,
                        const double *dz_, const double *ddx_,
                        const double *ddy_, const double *ddz_, const int *done) {
    if (*done) return;
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    if (!B.par) return;
    const int ni = B.ni, nj = B.nj, nk = B.nk;
    const int ii = ni + 4, jj = nj + 4, kk = nk + 4;
    const double *dx = dx_ + B.o1x - 1, *dy = dy_ + B.o1y - 1, *dz = dz_ + B.o1z - 1;
    const double *ddx = ddx_ + B.o1x - 1, *ddy = ddy_ + B.o1y - 1,
                 *ddz = ddz_ + B.o1z - 1;
    auto P = [&](int k, int j, int i) {
        return (double)hilf[B.wo + slot_at(t, B, i - 3, j - 3, k - 3)];
    };
    auto BP = [&](int k, int j, int i) { return bp_of(t, B, m, i - 3, j - 3, k - 3); };
    for (int f = 0; f < 6; ++f) {
        if (!((B.par >> f) & 1)) continue;
        int n1, n2;
        face_groups(B, f, n1, n2);
        auto put = [&](int k, int j, int i, double v) {
            int fi = i - 3, fj = j - 3, fk = k - 3;
            if (f < 2) fi = 0;
            else if (f < 4) fj = 0;
            else fk = 0;
            hpar[B.fc + face_slot(f, fi, fj, fk, ni, nj, nk)] = (T)v;
        };
        for (int e = threadIdx.x; e < n1 * n2; e += blockDim.x) {
            const int a = e / n2, c = e % n2;
            int i, j, k;
            group_cell(B, f, a, c, i, j, k);
            const double pbv = (double)pbuf[B.fc + face_slot(f, i, j, k, ni, nj, nk)];
            bound_group<BPV>(f, 3 + 2 * a, 3 + 2 * c, kk, jj, ii, dx, dy, dz, ddx,
                             ddy, ddz, pbv, P, BP, put);
        }
    }
}

// Restriction of a WO field into the parent cells (src/ib/ftoc_mod.F90:
// 581-626).  MODE 0: restrict_s (src/ib/noib_restrict_mod.F90:255-298);
// with the ghostcell operator MODE 1: restrict_p, the BP-weighted mean, 0
// for a fully blocked group (src/ib/gc_restrict_mod.F90:59-164), MODE 2:
// restrict_r, the BP-weighted sum over the full group volume (:167-259).
// Parent cells on the parent's boundary also refresh its boundary planes
// (`face`, may be null) and the optional second copy `dst2` (the sweeps'
// reduced-precision RHS).
template <typename S, typename D, int MODE>
__global__ void k_restrict(Topo t, const S *src, D *dst, D *face, float *dst2,
                           Mask m, const double *ddx_, const double *ddy_,
                           const double *ddz_, const int *done) {
    if (*done) return;
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const Block &P = t.blocks[B.parent];
    const int gi = B.ni / 2, gj = B.nj / 2, gk = B.nk / 2;
    const double *ddx = ddx_ + B.o1x - 1, *ddy = ddy_ + B.o1y - 1,
                 *ddz = ddz_ + B.o1z - 1;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < gi * gj * gk;
         e += gridDim.x * blockDim.x) {
        const int a = e / (gj * gk), c = (e / gk) % gj, d = e % gk;
        const int i0 = 2 * a, j0 = 2 * c, k0 = 2 * d;           // 0-based
        const int i = i0 + 3, j = j0 + 3, k = k0 + 3;           // 1-based
        auto ff = [&](int kq, int jq, int iq) {
            return (double)src[B.wo + slot_at(t, B, iq - 3, jq - 3, kq - 3)];
        };
        auto bq = [&](int kq, int jq, int iq) { return bp_of(t, B, m, iq - 3, jq - 3, kq - 3); };
        double v;
        if (MODE == 0) {
            const double spv =
                ff(k, j, i) * ddz[k] * ddy[j] * ddx[i] +
                ff(k, j, i + 1) * ddz[k] * ddy[j] * ddx[i + 1] +
                ff(k, j + 1, i) * ddz[k] * ddy[j + 1] * ddx[i] +
                ff(k, j + 1, i + 1) * ddz[k] * ddy[j + 1] * ddx[i + 1] +
                ff(k + 1, j, i) * ddz[k + 1] * ddy[j] * ddx[i] +
                ff(k + 1, j, i + 1) * ddz[k + 1] * ddy[j] * ddx[i + 1] +
                ff(k + 1, j + 1, i) * ddz[k + 1] * ddy[j + 1] * ddx[i] +
                ff(k + 1, j + 1, i + 1) * ddz[k + 1] * ddy[j + 1] * ddx[i + 1];
            const double sv = (ddx[i] + ddx[i + 1]) * (ddy[j] + ddy[j + 1]) *
                              (ddz[k] + ddz[k + 1]);
            v = spv / sv;
        } else if (MODE == 1) {
            double vol[8], pv[8];
            for (int n = 0; n <= 1; ++n) {
                const int kq = k + n, q = 4 * n;
                vol[q] = bq(kq, j, i) * ddz[kq] * ddy[j] * ddx[i];
                vol[q + 1] = bq(kq, j, i + 1) * ddz[kq] * ddy[j] * ddx[i + 1];
                vol[q + 2] = bq(kq, j + 1, i) * ddz[kq] * ddy[j + 1] * ddx[i];
                vol[q + 3] = bq(kq, j + 1, i + 1) * ddz[kq] * ddy[j + 1] * ddx[i + 1];
                pv[q] = ff(kq, j, i) * vol[q];
                pv[q + 1] = ff(kq, j, i + 1) * vol[q + 1];
                pv[q + 2] = ff(kq, j + 1, i) * vol[q + 2];
                pv[q + 3] = ff(kq, j + 1, i + 1) * vol[q + 3];
            }
            const double spv = pv[0] + pv[1] + pv[2] + pv[3] + pv[4] + pv[5] + pv[6] + pv[7];
            const double sv = vol[0] + vol[1] + vol[2] + vol[3] + vol[4] + vol[5] + vol[6] +
                              vol[7];
            v = sv > 0.0 ? spv / sv : 0.0;
        } else {
            double pv[8];
            for (int n = 0; n <= 1; ++n) {
                const int kq = k + n, q = 4 * n;
                pv[q] = ff(kq, j, i) * bq(kq, j, i) * ddz[kq] * ddy[j] * ddx[i];
                pv[q + 1] = ff(kq, j, i + 1) * bq(kq, j, i + 1) * ddz[kq] * ddy[j] * ddx[i + 1];
                pv[q + 2] = ff(kq, j + 1, i) * bq(kq, j + 1, i) * ddz[kq] * ddy[j + 1] * ddx[i];
                pv[q + 3] = ff(kq, j + 1, i + 1) * bq(kq, j + 1, i + 1) * ddz[kq] * ddy[j + 1] *
                            ddx[i + 1];
            }
            const double spv = pv[0] + pv[1] + pv[2] + pv[3] + pv[4] + pv[5] + pv[6] + pv[7];
            const double sv = (ddx[i] + ddx[i + 1]) * (ddy[j] + ddy[j + 1]) *
                              (ddz[k] + ddz[k + 1]);
            v = spv / sv;
        }
        const int ci = B.pos[0] - 3 + a, cj = B.pos[1] - 3 + c, ck = B.pos[2] - 3 + d;
        const int64_t o = P.wo + slot_at(t, P, ci, cj, ck);
        dst[o] = (D)v;
        if (dst2) dst2[o] = (float)v;
        if (face) {
            const int pn[3] = {P.ni, P.nj, P.nk};
            const int cc[3] = {ci, cj, ck};
            for (int f = 0; f < 6; ++f) {
                const int ax = f / 2;
                const bool on = (f & 1) ? cc[ax] == pn[ax] - 1 : cc[ax] == 0;
                if (on) face[P.fc + face_slot(f, ci, cj, ck, P.ni, P.nj, P.nk)] = (D)v;
            }
        }
    }
}

// ---------------------------------------------------------------------
// Reference layout (FP64): DP after the outer loop, U, V, W, P.
__device__ __forceinline__ int64_t ref_ix(const Block &B, int k, int j, int i) {
    const int64_t kk = B.nk + 4, jj = B.nj + 4;
    return B.sp + (k - 1) + (j - 1) * kk + (i - 1) * kk * jj;
}

// parent (src/ib/parent/parent1_mod.F90:296-583) of a reference-layout
// scalar: the group values of every PAR face
__global__ void k_parent_sp(Topo t, const double *f_, double *pbuf) {
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    if (!B.par) return;
    const Block &P = t.blocks[B.parent];
    for (int f = 0; f < 6; ++f) {
        if (!((B.par >> f) & 1)) continue;
        int n1, n2;
        face_groups(B, f, n1, n2);
        for (int e = threadIdx.x; e < n1 * n2; e += blockDim.x) {
            const int a = e / n2, c = e % n2;
            int ci, cj, ck, i, j, k;
            parent_cell(B, f, a, c, ci, cj, ck);
            group_cell(B, f, a, c, i, j, k);
            pbuf[B.fc + face_slot(f, i, j, k, B.ni, B.nj, B.nk)] =
                f_[ref_ix(P, ck + 3, cj + 3, ci + 3)];
        }
    }
}

template <bool BPV>
__global__ void k_bound_sp(Topo t, double *p_, const double *pbuf,
                           const double *bp_, const double *dx_,
                           const double *dy_, const double *dz_,
                           const double *ddx_, const double *ddy_,
                           const double *ddz_) {
    const int b = t.blist[blockIdx.x];
    const Block &B = t.blocks[b];
    if (!B.par) return;
    const int ni = B.ni, nj = B.nj, nk = B.nk;
    const int ii = ni + 4, jj = nj + 4, kk = nk + 4;
    const double *dx = dx_ + B.o1x - 1, *dy = dy_ + B.o1y - 1, *dz = dz_ + B.o1z - 1;
    const double *ddx = ddx_ + B.o1x - 1, *ddy = ddy_ + B.o1y - 1,
                 *ddz = ddz_ + B.o1z - 1;
    auto P = [&](int k, int j, int i) { return p_[ref_ix(B, k, j, i)]; };
    auto BP = [&](int k, int j, int i) { return bp_ ? bp_[ref_ix(B, k, j, i)] : 1.0; };
    auto put = [&](int k, int j, int i, double v) { p_[ref_ix(B, k, j, i)] = v; };
    for (int f = 0; f < 6; ++f) {
        if (!((B.par >> f) & 1)) continue;
        int n1, n2;
        face_groups(B, f, n1, n2);
        for (int e = threadIdx.x; e < n1 * n2; e += blockDim.x) {
            const int a = e / n2, c = e % n2;
            int i, j, k;
            group_cell(B, f, a, c, i, j, k);
            const double pbv = pbuf[B.fc + face_slot(f, i, j, k, ni, nj, nk)];
            bound_group<BPV>(f, 3 + 2 * a, 3 + 2 * c, kk, jj, ii, dx, dy, dz, ddx,
                             ddy, ddz, pbv, P, BP, put);
        }
        // groups of one face write only their own ghost cells; the next face
        // reads interior cells only, so no barrier is needed between faces
    }
}

// ftoc(ilevel, u, v, w, p): restrict_u/v/w/s (src/ib/noib_restrict_mod.F90:
// 135-298) with the staggered parent offsets of restrict_recieve_open
// (src/ib/ftoc_mod.F90:581-626); one thread per parent target.  Two
// same-level neighbours with the same parent both restrict the normal
// velocity of their shared face (one from its ghost face, one from its last
// interior face); the reference unpacks grid after grid, so the grid with
// the higher id wins, and only that one writes here.
__device__ __forceinline__ bool shared_face_loses(const Topo &t, int b,
                                                  const Block &B, int f) {
    const int nb = B.nbr[f];
    return nb > b && t.blocks[nb].parent == B.parent;
}
__global__ void k_ftoc_uvwp(Topo t, double *u, double *v, double *w,
                            double *p, const double *ddx_, const double *ddy_,
                            const double *ddz_, int gc, const double *bp) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const Block &P = t.blocks[B.parent];
    const int ipos = B.pos[0], jpos = B.pos[1], kpos = B.pos[2];
    const double *ddx = ddx_ + B.o1x - 1, *ddy = ddy_ + B.o1y - 1,
                 *ddz = ddz_ + B.o1z - 1;
    // U: i = 2 .. ii-2 step 2 (ni/2 + 1 values), j, k = 3 .. n-3 step 2
    const int nu = (B.ni / 2 + 1) * (B.nj / 2) * (B.nk / 2);
    const int nv = (B.ni / 2) * (B.nj / 2 + 1) * (B.nk / 2);
    const int nw = (B.ni / 2) * (B.nj / 2) * (B.nk / 2 + 1);
    const int np = (B.ni / 2) * (B.nj / 2) * (B.nk / 2);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nu + nv + nw + np;
         e += gridDim.x * blockDim.x) {
        if (e < nu) {
            const int gk = B.nk / 2, gj = B.nj / 2;
            const int i = 2 + 2 * (e / (gj * gk)), j = 3 + 2 * ((e / gk) % gj),
                      k = 3 + 2 * (e % gk);
            const double *ff = u;
            const double sa = ff[ref_ix(B, k, j, i)] * ddy[j] * ddz[k] +
                              ff[ref_ix(B, k, j + 1, i)] * ddy[j + 1] * ddz[k] +
                              ff[ref_ix(B, k + 1, j, i)] * ddy[j] * ddz[k + 1] +
                              ff[ref_ix(B, k + 1, j + 1, i)] * ddy[j + 1] * ddz[k + 1];
            const double a = (ddy[j] + ddy[j + 1]) * (ddz[k] + ddz[k + 1]);
            if (!((i == 2 && shared_face_loses(t, b, B, 0)) ||
                  (i == B.ni + 2 && shared_face_loses(t, b, B, 1))))
                u[ref_ix(P, kpos + (k - 3) / 2, jpos + (j - 3) / 2, ipos + (i - 2) / 2 - 1)] = sa / a;
        } else if (e < nu + nv) {
            const int q = e - nu;
            const int gk = B.nk / 2, gj = B.nj / 2 + 1;
            const int i = 3 + 2 * (q / (gj * gk)), j = 2 + 2 * ((q / gk) % gj),
                      k = 3 + 2 * (q % gk);
            const double *ff = v;
            const double sa = ff[ref_ix(B, k, j, i)] * ddx[i] * ddz[k] +
                              ff[ref_ix(B, k, j, i + 1)] * ddx[i + 1] * ddz[k] +
                              ff[ref_ix(B, k + 1, j, i)] * ddx[i] * ddz[k + 1] +
                              ff[ref_ix(B, k + 1, j, i + 1)] * ddx[i + 1] * ddz[k + 1];
            const double a = (ddx[i] + ddx[i + 1]) * (ddz[k] + ddz[k + 1]);
            if (!((j == 2 && shared_face_loses(t, b, B, 2)) ||
                  (j == B.nj + 2 && shared_face_loses(t, b, B, 3))))
                v[ref_ix(P, kpos + (k - 3) / 2, jpos + (j - 2) / 2 - 1, ipos + (i - 3) / 2)] = sa / a;
        } else if (e < nu + nv + nw) {
            const int q = e - nu - nv;
            const int gk = B.nk / 2 + 1, gj = B.nj / 2;
            const int i = 3 + 2 * (q / (gj * gk)), j = 3 + 2 * ((q / gk) % gj),
                      k = 2 + 2 * (q % gk);
            const double *ff = w;
            const double sa = ff[ref_ix(B, k, j, i)] * ddx[i] * ddy[j] +
                              ff[ref_ix(B, k, j, i + 1)] * ddx[i + 1] * ddy[j] +
                              ff[ref_ix(B, k, j + 1, i)] * ddx[i] * ddy[j + 1] +
                              ff[ref_ix(B, k, j + 1, i + 1)] * ddx[i + 1] * ddy[j + 1];
            const double a = (ddx[i] + ddx[i + 1]) * (ddy[j] + ddy[j + 1]);
            if (!((k == 2 && shared_face_loses(t, b, B, 4)) ||
                  (k == B.nk + 2 && shared_face_loses(t, b, B, 5))))
                w[ref_ix(P, kpos + (k - 2) / 2 - 1, jpos + (j - 3) / 2, ipos + (i - 3) / 2)] = sa / a;
        } else {
            const int q = e - nu - nv - nw;
            const int gk = B.nk / 2, gj = B.nj / 2;
            const int i = 3 + 2 * (q / (gj * gk)), j = 3 + 2 * ((q / gk) % gj),
                      k = 3 + 2 * (q % gk);
            auto ff = [&](int kq, int jq, int iq) { return p[ref_ix(B, kq, jq, iq)]; };
            double r;
            if (!gc) {
                const double spv =
                    ff(k, j, i) * ddz[k] * ddy[j] * ddx[i] +
                    ff(k, j, i + 1) * ddz[k] * ddy[j] * ddx[i + 1] +
                    ff(k, j + 1, i) * ddz[k] * ddy[j + 1] * ddx[i] +
                    ff(k, j + 1, i + 1) * ddz[k] * ddy[j + 1] * ddx[i + 1] +
                    ff(k + 1, j, i) * ddz[k + 1] * ddy[j] * ddx[i] +
                    ff(k + 1, j, i + 1) * ddz[k + 1] * ddy[j] * ddx[i + 1] +
                    ff(k + 1, j + 1, i) * ddz[k + 1] * ddy[j + 1] * ddx[i] +
                    ff(k + 1, j + 1, i + 1) * ddz[k + 1] * ddy[j + 1] * ddx[i + 1];
                const double sv = (ddx[i] + ddx[i + 1]) * (ddy[j] + ddy[j + 1]) *
                                  (ddz[k] + ddz[k + 1]);
                r = spv / sv;
            } else {
                // ghostcell restrict_p (src/ib/gc_restrict_mod.F90:59-164)
                auto bq = [&](int kq, int jq, int iq) {
                    return bp ? bp[ref_ix(B, kq, jq, iq)] : 1.0;
                };
                double vol[8], pv[8];
                for (int n = 0; n <= 1; ++n) {
                    const int kq = k + n, q = 4 * n;
                    vol[q] = bq(kq, j, i) * ddz[kq] * ddy[j] * ddx[i];
                    vol[q + 1] = bq(kq, j, i + 1) * ddz[kq] * ddy[j] * ddx[i + 1];
                    vol[q + 2] = bq(kq, j + 1, i) * ddz[kq] * ddy[j + 1] * ddx[i];
                    vol[q + 3] = bq(kq, j + 1, i + 1) * ddz[kq] * ddy[j + 1] * ddx[i + 1];
                    pv[q] = ff(kq, j, i) * vol[q];
                    pv[q + 1] = ff(kq, j, i + 1) * vol[q + 1];
                    pv[q + 2] = ff(kq, j + 1, i) * vol[q + 2];
                    pv[q + 3] = ff(kq, j + 1, i + 1) * vol[q + 3];
                }
                const double spv = pv[0] + pv[1] + pv[2] + pv[3] + pv[4] + pv[5] + pv[6] +
                                   pv[7];
                const double sv = vol[0] + vol[1] + vol[2] + vol[3] + vol[4] + vol[5] +
                                  vol[6] + vol[7];
                r = sv > 0.0 ? spv / sv : 0.0;
            }
            p[ref_ix(P, kpos + (k - 3) / 2, jpos + (j - 3) / 2, ipos + (i - 3) / 2)] = r;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------
template <typename T>
void launch_ctof(const Topo &t, const Fields<T> &f, T *fout, T *hsh,
                 double *dsh, const int *done, hipStream_t s) {
    MGP_RANGE("mgp:ctof");
    const int maxc = t.max_ni * t.max_nj * t.max_nk;
    dim3 grid((maxc + kThreadsLv - 1) / kThreadsLv, t.nlist);
    k_ctof<T><<<grid, kThreadsLv, 0, s>>>(t, f, fout, hsh, dsh, done);
}

// DP's first ghost layer from the shell sums (reference layout)
__global__ void k_shell_to_sp(Topo t, const double *dsh, double *sp) {
    const int b = t.blist[blockIdx.y];
    const Block &B = t.blocks[b];
    const int ni = B.ni, nj = B.nj, nk = B.nk;
    const int nsh = shell_size(ni, nj, nk);
    const int a = (nj + 2) * (nk + 2);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nsh;
         e += gridDim.x * blockDim.x) {
        int i, j, k;
        if (e < 2 * a) {
            i = e < a ? -1 : ni;
            const int r = e % a;
            j = r / (nk + 2) - 1;
            k = r % (nk + 2) - 1;
        } else {
            const int slab = 2 * (nk + 2) + 2 * nj;
            const int r0 = e - 2 * a;
            i = r0 / slab;
            const int r = r0 % slab;
            if (r < nk + 2) { j = -1; k = r - 1; }
            else if (r < 2 * (nk + 2)) { j = nj; k = r - (nk + 2) - 1; }
            else { j = (r - 2 * (nk + 2)) / 2; k = ((r - 2 * (nk + 2)) & 1) ? nk : -1; }
        }
        sp[ref_ix(B, k + 3, j + 3, i + 3)] = dsh[B.sh + e];
    }
}

void launch_shell_to_sp(const Topo &t, const double *dsh, double *sp,
                        hipStream_t s) {
    MGP_RANGE("mgp:shell_to_sp");
    const int maxs = shell_size(t.max_ni, t.max_nj, t.max_nk);
    dim3 grid((maxs + kThreadsLv - 1) / kThreadsLv, t.nlist);
    k_shell_to_sp<<<grid, kThreadsLv, 0, s>>>(t, dsh, sp);
}

template <typename T>
void launch_parent(const Topo &t, const Fields<T> &f, T *pbuf, const int *done,
                   hipStream_t s) {
    MGP_RANGE("mgp:parent");
    k_parent<T><<<t.nlist, kThreadsLv, 0, s>>>(t, f.hilf, f.hface, f.hpar, pbuf, done);
}

template <typename T>
void launch_bound(const Topo &t, const Fields<T> &f, const T *pbuf, Mask m,
                  int bpv, const double *const metric[6], const int *done,
                  hipStream_t s) {
    MGP_RANGE("mgp:bound");
    if (bpv)
        k_bound<T, true><<<t.nlist, kThreadsLv, 0, s>>>(
            t, f.hilf, f.hpar, pbuf, m, metric[0], metric[1], metric[2],
            metric[3], metric[4], metric[5], done);
    else
        k_bound<T, false><<<t.nlist, kThreadsLv, 0, s>>>(
            t, f.hilf, f.hpar, pbuf, m, metric[0], metric[1], metric[2],
            metric[3], metric[4], metric[5], done);
}

template <typename S, typename D>
void launch_restrict(const Topo &t, const S *src, D *dst, D *face,
                     float *dst2, Mask m, int mode, const double *const dd[3],
                     const int *done, hipStream_t s) {
    MGP_RANGE("mgp:restrict");
    const int maxg = (t.max_ni / 2) * (t.max_nj / 2) * (t.max_nk / 2);
    dim3 grid((maxg + kThreadsLv - 1) / kThreadsLv, t.nlist);
    if (mode == 1)
        k_restrict<S, D, 1><<<grid, kThreadsLv, 0, s>>>(t, src, dst, face, dst2, m,
                                                        dd[0], dd[1], dd[2], done);
    else if (mode == 2)
        k_restrict<S, D, 2><<<grid, kThreadsLv, 0, s>>>(t, src, dst, face, dst2, m,
                                                        dd[0], dd[1], dd[2], done);
    else
        k_restrict<S, D, 0><<<grid, kThreadsLv, 0, s>>>(t, src, dst, face, dst2, m,
                                                        dd[0], dd[1], dd[2], done);
}

void launch_parent_sp(const Topo &t, const double *f, double *pbuf,
                      hipStream_t s) {
    MGP_RANGE("mgp:parent_sp");
    k_parent_sp<<<t.nlist, kThreadsLv, 0, s>>>(t, f, pbuf);
}

void launch_bound_sp(const Topo &t, double *p, const double *pbuf,
                     const double *bp, int bpv, const double *const metric[6],
                     hipStream_t s) {
    MGP_RANGE("mgp:bound_sp");
    if (bpv)
        k_bound_sp<true><<<t.nlist, kThreadsLv, 0, s>>>(
            t, p, pbuf, bp, metric[0], metric[1], metric[2], metric[3],
            metric[4], metric[5]);
    else
        k_bound_sp<false><<<t.nlist, kThreadsLv, 0, s>>>(
            t, p, pbuf, nullptr, metric[0], metric[1], metric[2], metric[3],
            metric[4], metric[5]);
}

void launch_ftoc_uvwp(const Topo &t, double *u, double *v, double *w,
                      double *p, int gc, const double *bp,
                      const double *const dd[3], hipStream_t s) {
    MGP_RANGE("mgp:ftoc_uvwp");
    const int maxn = 4 * (t.max_ni / 2 + 1) * (t.max_nj / 2 + 1) * (t.max_nk / 2 + 1);
    dim3 grid((maxn + kThreadsLv - 1) / kThreadsLv, t.nlist);
    k_ftoc_uvwp<<<grid, kThreadsLv, 0, s>>>(t, u, v, w, p, dd[0], dd[1], dd[2], gc, bp);
}

// par_ftoc_norm (src/ib/par_ftoc_mod.F90:35-475): one workgroup per task
// (a PAR face of a fine grid and one coarse grid receiving it).  Each
// coarse face position (a, b) takes its 2 x 2 group of the fine face plane:
// the plain sum, or the area-weighted mean (pack_single :234-337, terms in
// its order: along t1 then t2 on x and y faces, t2 then t1 on z faces).
__global__ void k_par_ftoc(const Block *blocks, const ParFtocTask *tasks, double *u,
                           double *v, double *w, const double *ddx_, const double *ddy_,
                           const double *ddz_, int sum) {
    const ParFtocTask T = tasks[blockIdx.x];
    const Block &F = blocks[T.fine], &C = blocks[T.dest];
    const int n = T.face >> 1;
    const int t1 = n == 0 ? 1 : 0, t2 = n == 2 ? 1 : 2;
    const int nf[3] = {F.ni + 4, F.nj + 4, F.nk + 4};
    const int nc[3] = {C.ni + 4, C.nj + 4, C.nk + 4};
    const int64_t sf[3] = {(int64_t)nf[2] * nf[1], nf[2], 1};
    const int64_t sc[3] = {(int64_t)nc[2] * nc[1], nc[2], 1};
    const double *dd[3] = {ddx_ + F.o1x - 1, ddy_ + F.o1y - 1, ddz_ + F.o1z - 1};
    const double *w1 = dd[t1], *w2 = dd[t2];
    const double *ff = (n == 0 ? u : n == 1 ? v : w) + F.sp;
    double *fc = (n == 0 ? u : n == 1 ? v : w) + C.sp;
    const int fn = (T.face & 1) ? nf[n] - 2 : 2;   // fine face plane
    const int na = (nf[t1] - 4) / 2, nb = (nf[t2] - 4) / 2;
    for (int e = threadIdx.x; e < na * nb; e += blockDim.x) {
        const int a = e / nb, b = e % nb;
        const int p1 = 3 + 2 * a, p2 = 3 + 2 * b;
        const int64_t x0 = (int64_t)(fn - 1) * sf[n] + (int64_t)(p1 - 1) * sf[t1] +
                           (int64_t)(p2 - 1) * sf[t2];
        const double f00 = ff[x0], f10 = ff[x0 + sf[t1]], f01 = ff[x0 + sf[t2]],
                     f11 = ff[x0 + sf[t1] + sf[t2]];
        double val;
        if (sum) {
            val = n == 2 ? f00 + f01 + f10 + f11 : f00 + f10 + f01 + f11;
        } else {
            const double a00 = f00 * w1[p1] * w2[p2], a10 = f10 * w1[p1 + 1] * w2[p2],
                         a01 = f01 * w1[p1] * w2[p2 + 1], a11 = f11 * w1[p1 + 1] * w2[p2 + 1];
            const double ua = n == 2 ? a00 + a01 + a10 + a11 : a00 + a10 + a01 + a11;
            val = ua / ((w1[p1] + w1[p1 + 1]) * (w2[p2] + w2[p2 + 1]));
        }
        const int64_t xc = (int64_t)(T.dnorm - 1) * sc[n] +
                           (int64_t)(F.pos[t1] + a - 1) * sc[t1] +
                           (int64_t)(F.pos[t2] + b - 1) * sc[t2];
        fc[xc] = val;
    }
}

void launch_par_ftoc(const Block *blocks, const ParFtocTask *tasks, int ntask, double *u,
                     double *v, double *w, const double *const dd[3], int sum,
                     hipStream_t s) {
    MGP_RANGE("mgp:par_ftoc");
    if (ntask > 0)
        k_par_ftoc<<<ntask, kThreadsLv, 0, s>>>(blocks, tasks, u, v, w, dd[0], dd[1], dd[2],
                                                sum);
}

template void launch_ctof<double>(const Topo &, const Fields<double> &, double *, double *,
                                  double *, const int *, hipStream_t);
template void launch_ctof<float>(const Topo &, const Fields<float> &, float *, float *,
                                 double *, const int *, hipStream_t);
template void launch_parent<double>(const Topo &, const Fields<double> &, double *,
                                    const int *, hipStream_t);
template void launch_parent<float>(const Topo &, const Fields<float> &, float *, const int *,
                                   hipStream_t);
template void launch_bound<double>(const Topo &, const Fields<double> &, const double *, Mask,
                                   int, const double *const[6], const int *, hipStream_t);
template void launch_bound<float>(const Topo &, const Fields<float> &, const float *, Mask,
                                  int, const double *const[6], const int *, hipStream_t);
template void launch_restrict<double, double>(const Topo &, const double *, double *,
                                              double *, float *, Mask, int,
                                              const double *const[3], const int *,
                                              hipStream_t);
template void launch_restrict<float, float>(const Topo &, const float *, float *, float *,
                                            float *, Mask, int, const double *const[3],
                                            const int *, hipStream_t);

}  // namespace mgp
