/*
 * mgp_oracle.c -- CPU restatement (FP64) of MGLET-base's pressure-projection
 * hot path.  TEST INFRASTRUCTURE ONLY: the checker for the HIP library and
 * the timed CPU baseline.  The product path never links or calls this file.
 *
 * Each function cites the reference routine it restates.  The restatement
 * keeps the reference's loop bounds, evaluation order and literals:
 *   - `alfa = 0.92` and `1.0e-20` are default-real (real32) literals in the
 *     reference, so in an FP64 build they hold (double)0.92f and
 *     (double)1.0e-20f (src/flow/sip_hyperplane_mod.F90:123, 202;
 *     src/flow/sip_classic_mod.F90:26, 82);
 *   - Fortran evaluates `a - b*c - d*e` left to right; so does C here, and
 *     the file is built with -ffp-contract=off (no FMA contraction).
 * Parity status of each function is recorded in the tests that pin it.
 */
#include "mgp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* alfa is the reference constant; tests may change it to pin the
 * factorisation against closed forms (alfa = 0 gives ILU(0)). */
static double ALFA = (double)0.92f;
void og_set_alfa(double a) { ALFA = a; }
#define TINY ((double)1.0e-20f)

/* 1-based (k, j, i) -> 0-based offset inside one block */
#define IX(k, j, i) ((int64_t)((k)-1) + (int64_t)((j)-1) * kk + \
                     (int64_t)((i)-1) * kk * jj)

#define GDIMS(t, g)                 \
    const int kk = (t)->kk[(g)-1];  \
    const int jj = (t)->jj[(g)-1];  \
    const int ii = (t)->ii[(g)-1];  \
    (void)kk; (void)jj; (void)ii

/* ------------------------------------------------------------------ */
/* Face / edge / corner tables: src/core/connect/connect_core_mod.F90:37-165.
 * These tables are part of the data format (face numbering) and must be
 * the same as the reference's. */
static const int FACELIST[26][4] = {
    {1, 1, 0, 0}, {1, 2, 0, 0}, {1, 3, 0, 0}, {1, 4, 0, 0}, {1, 5, 0, 0},
    {1, 6, 0, 0}, {2, 1, 3, 0}, {2, 1, 4, 0}, {2, 1, 5, 0}, {2, 1, 6, 0},
    {2, 2, 3, 0}, {2, 2, 4, 0}, {2, 2, 5, 0}, {2, 2, 6, 0}, {2, 3, 5, 0},
    {2, 3, 6, 0}, {2, 4, 5, 0}, {2, 4, 6, 0}, {3, 1, 3, 5}, {3, 1, 3, 6},
    {3, 1, 4, 5}, {3, 1, 4, 6}, {3, 2, 3, 5}, {3, 2, 3, 6}, {3, 2, 4, 5},
    {3, 2, 4, 6}};
static const int FACENBR[26] = {2,  1,  4,  3,  6,  5,  12, 11, 14,
                                13, 8,  7,  10, 9,  18, 17, 16, 15,
                                26, 25, 24, 23, 22, 21, 20, 19};
static const int RESCUE_DIR[26][7] = {
    {1, 0, 0, 0, 0, 0, 0},      {2, 0, 0, 0, 0, 0, 0},
    {3, 0, 0, 0, 0, 0, 0},      {4, 0, 0, 0, 0, 0, 0},
    {5, 0, 0, 0, 0, 0, 0},      {6, 0, 0, 0, 0, 0, 0},
    {7, 1, 3, 0, 0, 0, 0},      {8, 4, 1, 0, 0, 0, 0},
    {9, 1, 5, 0, 0, 0, 0},      {10, 6, 1, 0, 0, 0, 0},
    {11, 2, 3, 0, 0, 0, 0},     {12, 2, 4, 0, 0, 0, 0},
    {13, 2, 5, 0, 0, 0, 0},     {14, 2, 6, 0, 0, 0, 0},
    {15, 3, 5, 0, 0, 0, 0},     {16, 6, 3, 0, 0, 0, 0},
    {17, 4, 5, 0, 0, 0, 0},     {18, 4, 6, 0, 0, 0, 0},
    {19, 7, 9, 15, 1, 3, 5},    {20, 7, 10, 16, 1, 3, 6},
    {21, 8, 9, 17, 1, 4, 5},    {22, 8, 10, 18, 1, 4, 6},
    {23, 11, 13, 15, 2, 3, 5},  {24, 11, 14, 16, 2, 3, 6},
    {25, 12, 13, 17, 2, 4, 5},  {26, 12, 14, 18, 2, 4, 6}};
static const int RESCUE_NBR[26][7] = {
    {2, 0, 0, 0, 0, 0, 0},        {1, 0, 0, 0, 0, 0, 0},
    {4, 0, 0, 0, 0, 0, 0},        {3, 0, 0, 0, 0, 0, 0},
    {6, 0, 0, 0, 0, 0, 0},        {5, 0, 0, 0, 0, 0, 0},
    {12, 11, 8, 0, 0, 0, 0},      {11, 7, 12, 0, 0, 0, 0},
    {14, 13, 10, 0, 0, 0, 0},     {13, 9, 14, 0, 0, 0, 0},
    {8, 7, 12, 0, 0, 0, 0},       {7, 8, 11, 0, 0, 0, 0},
    {10, 9, 14, 0, 0, 0, 0},      {9, 10, 13, 0, 0, 0, 0},
    {18, 17, 16, 0, 0, 0, 0},     {17, 15, 18, 0, 0, 0, 0},
    {16, 15, 18, 0, 0, 0, 0},     {15, 16, 17, 0, 0, 0, 0},
    {26, 25, 24, 22, 23, 21, 20}, {25, 26, 23, 21, 24, 22, 19},
    {24, 23, 26, 20, 25, 19, 22}, {23, 24, 25, 19, 26, 20, 21},
    {22, 21, 20, 26, 19, 25, 24}, {21, 22, 19, 25, 20, 26, 23},
    {20, 19, 22, 24, 21, 23, 26}, {19, 20, 21, 23, 22, 24, 25}};

static int is_con(int code) { return code == 7 || code == 19; }

/* ------------------------------------------------------------------ */
/* src/ib/noib_mod.F90:81-250 (giteig) */
void og_giteig(const og_topo *t, const double *dx_, const double *dy_,
               const double *dz_, const double *bp_, double *aw_,
               double *ae_, double *as_, double *an_, double *ab_,
               double *at_, double *ap_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        const double *dx = dx_ + t->ip1x[g - 1] - 1; /* 1-based */
        const double *dy = dy_ + t->ip1y[g - 1] - 1;
        const double *dz = dz_ + t->ip1z[g - 1] - 1;
        double *aw = aw_ + t->ip1x[g - 1] - 1, *ae = ae_ + t->ip1x[g - 1] - 1;
        double *as = as_ + t->ip1y[g - 1] - 1, *an = an_ + t->ip1y[g - 1] - 1;
        double *ab = ab_ + t->ip1z[g - 1] - 1, *at = at_ + t->ip1z[g - 1] - 1;
        const double *bp = bp_ + t->ip3[g - 1];
        double *ap = ap_ + t->ip3[g - 1];
        const int32_t *bc = t->bc + 6 * (g - 1);

        for (int i = 3; i <= ii - 2; ++i) {
            ae[i] = 2.0 / ((dx[i - 1] + dx[i]) * dx[i]);
            aw[i] = 2.0 / ((dx[i - 1] + dx[i]) * dx[i - 1]);
        }
        for (int j = 3; j <= jj - 2; ++j) {
            an[j] = 2.0 / ((dy[j - 1] + dy[j]) * dy[j]);
            as[j] = 2.0 / ((dy[j - 1] + dy[j]) * dy[j - 1]);
        }
        for (int k = 3; k <= kk - 2; ++k) {
            at[k] = 2.0 / ((dz[k - 1] + dz[k]) * dz[k]);
            ab[k] = 2.0 / ((dz[k - 1] + dz[k]) * dz[k - 1]);
        }
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k)
                    ap[IX(k, j, i)] = -(2.0 / (dx[i - 1] * dx[i]))
                                      - 2.0 / (dy[j - 1] * dy[j])
                                      - 2.0 / (dz[k - 1] * dz[k]);
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k) {
                    const double c = bp[IX(k, j, i)];
                    double a = ap[IX(k, j, i)];
                    a = a + aw[i] * (1.0 - bp[IX(k, j, i - 1)] * c);
                    a = a + ae[i] * (1.0 - c * bp[IX(k, j, i + 1)]);
                    a = a + as[j] * (1.0 - bp[IX(k, j - 1, i)] * c);
                    a = a + an[j] * (1.0 - c * bp[IX(k, j + 1, i)]);
                    a = a + ab[k] * (1.0 - bp[IX(k - 1, j, i)] * c);
                    a = a + at[k] * (1.0 - c * bp[IX(k + 1, j, i)]);
                    ap[IX(k, j, i)] = a;
                }
        const int nfro = bc[0], nbac = bc[1], nrgt = bc[2], nlft = bc[3],
                  nbot = bc[4], ntop = bc[5];
        if (nfro == 2 || nfro == 5 || nfro == 6 || nfro == 19) {
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k)
                    ap[IX(k, j, 3)] = ap[IX(k, j, 3)]
                        + aw[3] * (bp[IX(k, j, 2)] * bp[IX(k, j, 3)]);
            aw[3] = 0.0;
        }
        if (nbac == 2 || nbac == 5 || nbac == 6) {
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k)
                    ap[IX(k, j, ii - 2)] = ap[IX(k, j, ii - 2)]
                        + ae[ii - 2] * (bp[IX(k, j, ii - 2)] * bp[IX(k, j, ii - 1)]);
            ae[ii - 2] = 0.0;
        }
        if (nrgt == 2 || nrgt == 5 || nrgt == 6 || nrgt == 19) {
            for (int i = 3; i <= ii - 2; ++i)
                for (int k = 3; k <= kk - 2; ++k)
                    ap[IX(k, 3, i)] = ap[IX(k, 3, i)]
                        + as[3] * (bp[IX(k, 2, i)] * bp[IX(k, 3, i)]);
            as[3] = 0.0;
        }
        if (nlft == 2 || nlft == 5 || nlft == 6) {
            for (int i = 3; i <= ii - 2; ++i)
                for (int k = 3; k <= kk - 2; ++k)
                    ap[IX(k, jj - 2, i)] = ap[IX(k, jj - 2, i)]
                        + an[jj - 2] * (bp[IX(k, jj - 2, i)] * bp[IX(k, jj - 1, i)]);
            an[jj - 2] = 0.0;
        }
        if (nbot == 2 || nbot == 5 || nbot == 6 || nbot == 19) {
            for (int i = 3; i <= ii - 2; ++i)
                for (int j = 3; j <= jj - 2; ++j)
                    ap[IX(3, j, i)] = ap[IX(3, j, i)]
                        + ab[3] * (bp[IX(2, j, i)] * bp[IX(3, j, i)]);
            ab[3] = 0.0;
        }
        if (ntop == 2 || ntop == 5 || ntop == 6) {
            for (int i = 3; i <= ii - 2; ++i)
                for (int j = 3; j <= jj - 2; ++j)
                    ap[IX(kk - 2, j, i)] = ap[IX(kk - 2, j, i)]
                        + at[kk - 2] * (bp[IX(kk - 2, j, i)] * bp[IX(kk - 1, j, i)]);
            at[kk - 2] = 0.0;
        }
    }
}

/* ------------------------------------------------------------------ */
/* src/flow/sip_hyperplane_mod.F90:269-447: MIP(m) plane offsets and IDX
 * (1-based linear index, ascending inside each plane).  mip/idx are full
 * 3-D intfields; this block's MIP uses entries 1..ii+jj+kk of its region. */
static int cmp_i64(const void *a, const void *b)
{
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

int og_hp_tables(const og_topo *t, int64_t *mip_, int64_t *idx_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        int64_t *mip = mip_ + t->ip3[g - 1] - 1; /* 1-based */
        int64_t *idx = idx_ + t->ip3[g - 1] - 1;
        const int n3dmin = 9, n3dmax = (ii - 2) + (jj - 2) + (kk - 2);
        for (int m = 1; m <= ii + jj + kk; ++m) mip[m] = 0;
        for (int64_t s = 1; s <= (int64_t)ii * jj * kk; ++s) idx[s] = -1;
        for (int m = n3dmin; m <= n3dmax; ++m) {
            int ic1 = 0;
            const int64_t lm = mip[m];
            const int nkmin = (m - (ii - 2) - (jj - 2) > 3) ? m - (ii - 2) - (jj - 2) : 3;
            const int nkmax = (m - 6 < kk - 2) ? m - 6 : kk - 2;
            for (int k = nkmin; k <= nkmax; ++k) {
                const int nimin = (m - (jj - 2) - k > 3) ? m - (jj - 2) - k : 3;
                const int nimax = (m - k - 3 < ii - 2) ? m - k - 3 : ii - 2;
                for (int i = nimin; i <= nimax; ++i) {
                    const int j = m - k - i;
                    if (i < 3 || i > ii - 2 || j < 3 || j > jj - 2 ||
                        k < 3 || k > kk - 2)
                        return -1;
                    ++ic1;
                    idx[lm + ic1] = (int64_t)k + (int64_t)(j - 1) * kk
                                    + (int64_t)(i - 1) * kk * jj;
                }
            }
            mip[m + 1] = mip[m] + ic1;
        }
        /* sip_hyperplane_sort_grid: ascending inside each plane */
        for (int m = n3dmin; m <= n3dmax; ++m) {
            const int64_t len = mip[m + 1] - mip[m];
            qsort(&idx[mip[m] + 1], (size_t)len, sizeof(int64_t), cmp_i64);
            for (int64_t s = mip[m] + 1; s < mip[m + 1]; ++s)
                if (idx[s + 1] <= idx[s]) return -2;
        }
        /* sip_hyperplane_check_grid */
        int64_t sum = 0;
        char *seen = calloc((size_t)ii * jj * kk, 1);
        for (int m = n3dmin; m <= n3dmax; ++m)
            for (int64_t s = mip[m] + 1; s <= mip[m + 1]; ++s) {
                const int64_t l = idx[s] - 1;
                const int k = (int)(l % kk) + 1;
                const int j = (int)((l / kk) % jj) + 1;
                const int i = (int)(l / ((int64_t)kk * jj)) + 1;
                if (k + j + i != m) { free(seen); return -3; }
                if (seen[l]++) { free(seen); return -4; }
                ++sum;
            }
        free(seen);
        if (sum != (int64_t)(kk - 4) * (jj - 4) * (ii - 4)) return -5;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* src/flow/sip_hyperplane_mod.F90:106-257 */
void og_lu_hp(const og_topo *t, const double *aw_, const double *ae_,
              const double *as_, const double *an_, const double *ab_,
              const double *at_, const double *ap_, const double *bp_,
              const int64_t *mip_, const int64_t *idx_, double *lw_,
              double *ls_, double *lb_, double *lpr_, double *ue_,
              double *un_, double *ut_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        const int64_t n = (int64_t)kk * jj * ii;
        const int64_t o3 = t->ip3[g - 1];
        const double *aw = aw_ + t->ip1x[g - 1] - 1, *ae = ae_ + t->ip1x[g - 1] - 1;
        const double *as = as_ + t->ip1y[g - 1] - 1, *an = an_ + t->ip1y[g - 1] - 1;
        const double *ab = ab_ + t->ip1z[g - 1] - 1, *at = at_ + t->ip1z[g - 1] - 1;
        const double *ap = ap_ + o3, *bp = bp_ + o3;
        const int64_t *mip = mip_ + o3 - 1, *idxs = idx_ + o3 - 1;
        double *lw = lw_ + o3 - 1, *ls = ls_ + o3 - 1, *lb = lb_ + o3 - 1,
               *lpr = lpr_ + o3 - 1, *ue = ue_ + o3 - 1, *un = un_ + o3 - 1,
               *ut = ut_ + o3 - 1;
        /* tmp arrays, 1-based, zero-initialised (:137-143) */
        double *tw = calloc((size_t)n + 1, sizeof(double));
        double *ts = calloc((size_t)n + 1, sizeof(double));
        double *tb = calloc((size_t)n + 1, sizeof(double));
        double *tp = calloc((size_t)n + 1, sizeof(double));
        double *te = calloc((size_t)n + 1, sizeof(double));
        double *tn = calloc((size_t)n + 1, sizeof(double));
        double *tt = calloc((size_t)n + 1, sizeof(double));
        const int n3dmin = 9, n3dmax = (ii - 2) + (jj - 2) + (kk - 2);
        for (int m = n3dmin; m <= n3dmax; ++m) {
            const int64_t lm = mip[m], lp = mip[m + 1] - lm;
            for (int64_t ip = 1; ip <= lp; ++ip) {
                const int64_t iacc = lm + ip;
                const int64_t id = idxs[iacc];
                const int64_t km = id - 1, jm = id - kk,
                              im = id - (int64_t)kk * jj;
                const int64_t l = id - 1;
                const int k = (int)(l % kk) + 1;
                const int j = (int)((l / kk) % jj) + 1;
                const int i = (int)(l / ((int64_t)kk * jj)) + 1;
                const double bu_m = bp[IX(k, j, i - 1)] * bp[IX(k, j, i)];
                const double bv_m = bp[IX(k, j - 1, i)] * bp[IX(k, j, i)];
                const double bw_m = bp[IX(k - 1, j, i)] * bp[IX(k, j, i)];
                const double bu_c = bp[IX(k, j, i)] * bp[IX(k, j, i + 1)];
                const double bv_c = bp[IX(k, j, i)] * bp[IX(k, j + 1, i)];
                const double bw_c = bp[IX(k, j, i)] * bp[IX(k + 1, j, i)];
                tw[id] = aw[i] * bu_m / (1.0 + ALFA * (tn[im] + tt[im]));
                ts[id] = as[j] * bv_m / (1.0 + ALFA * (te[jm] + tt[jm]));
                tb[id] = ab[k] * bw_m / (1.0 + ALFA * (tn[km] + te[km]));
                const double p1 = ALFA * (tb[id] * te[km] + ts[id] * te[jm]);
                const double p2 = ALFA * (tb[id] * tn[km] + tw[id] * tn[im]);
                const double p3 = ALFA * (tw[id] * tt[im] + ts[id] * tt[jm]);
                tp[id] = 1.0 / (ap[IX(k, j, i)] + p1 + p2 + p3
                                - tb[id] * tt[km] - tw[id] * te[im]
                                - ts[id] * tn[jm] + TINY);
                te[id] = (ae[i] * bu_c - p1) * tp[id];
                tn[id] = (an[j] * bv_c - p2) * tp[id];
                tt[id] = (at[k] * bw_c - p3) * tp[id];
                tw[id] = tw[id] * tp[id];
                ts[id] = ts[id] * tp[id];
                tb[id] = tb[id] * tp[id];
                lw[iacc] = tw[id]; ls[iacc] = ts[id]; lb[iacc] = tb[id];
                lpr[iacc] = tp[id]; ue[iacc] = te[id]; un[iacc] = tn[id];
                ut[iacc] = tt[id];
            }
        }
        free(tw); free(ts); free(tb); free(tp); free(te); free(tn); free(tt);
    }
}

/* src/flow/sip_classic_mod.F90:17-127 (factor fields start zeroed) */
void og_lu_classic(const og_topo *t, const double *aw_, const double *ae_,
                   const double *as_, const double *an_, const double *ab_,
                   const double *at_, const double *ap_, const double *bp_,
                   double *lw_, double *ls_, double *lb_, double *lpr_,
                   double *ue_, double *un_, double *ut_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        const int64_t o3 = t->ip3[g - 1];
        const double *aw = aw_ + t->ip1x[g - 1] - 1, *ae = ae_ + t->ip1x[g - 1] - 1;
        const double *as = as_ + t->ip1y[g - 1] - 1, *an = an_ + t->ip1y[g - 1] - 1;
        const double *ab = ab_ + t->ip1z[g - 1] - 1, *at = at_ + t->ip1z[g - 1] - 1;
        const double *ap = ap_ + o3, *bp = bp_ + o3;
        double *lw = lw_ + o3, *ls = ls_ + o3, *lb = lb_ + o3,
               *lpr = lpr_ + o3, *ue = ue_ + o3, *un = un_ + o3,
               *ut = ut_ + o3;
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k) {
                    const int64_t c = IX(k, j, i), im = IX(k, j, i - 1),
                                  jm = IX(k, j - 1, i), km = IX(k - 1, j, i);
                    const double bu_m = bp[im] * bp[c];
                    const double bv_m = bp[jm] * bp[c];
                    const double bw_m = bp[km] * bp[c];
                    lw[c] = aw[i] * bu_m / (1.0 + ALFA * (un[im] + ut[im]));
                    ls[c] = as[j] * bv_m / (1.0 + ALFA * (ue[jm] + ut[jm]));
                    lb[c] = ab[k] * bw_m / (1.0 + ALFA * (un[km] + ue[km]));
                    const double p1 = ALFA * (lb[c] * ue[km] + ls[c] * ue[jm]);
                    const double p2 = ALFA * (lb[c] * un[km] + lw[c] * un[im]);
                    const double p3 = ALFA * (lw[c] * ut[im] + ls[c] * ut[jm]);
                    lpr[c] = 1.0 / (ap[c] + p1 + p2 + p3 - lb[c] * ut[km]
                                    - lw[c] * ue[im] - ls[c] * un[jm] + TINY);
                    ue[c] = (ae[i] * (bp[c] * bp[IX(k, j, i + 1)]) - p1) * lpr[c];
                    un[c] = (an[j] * (bp[c] * bp[IX(k, j + 1, i)]) - p2) * lpr[c];
                    ut[c] = (at[k] * (bp[c] * bp[IX(k + 1, j, i)]) - p3) * lpr[c];
                }
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j) {
                for (int k = 3; k <= kk - 2; ++k) lw[IX(k, j, i)] *= lpr[IX(k, j, i)];
                for (int k = 3; k <= kk - 2; ++k) ls[IX(k, j, i)] *= lpr[IX(k, j, i)];
                for (int k = 3; k <= kk - 2; ++k) lb[IX(k, j, i)] *= lpr[IX(k, j, i)];
            }
    }
}

/* ------------------------------------------------------------------ */
/* src/flow/laplacephi_mod.F90:134-165 */
static void laplacephi_grid(int kk, int jj, int ii, double *res,
                            const double *phi, const double *aw,
                            const double *ae, const double *an,
                            const double *as, const double *at,
                            const double *ab, const double *ap,
                            const double *bp)
{
    for (int i = 3; i <= ii - 2; ++i)
        for (int j = 3; j <= jj - 2; ++j)
            for (int k = 3; k <= kk - 2; ++k) {
                const int64_t c = IX(k, j, i);
                double r = -(aw[i] * phi[IX(k, j, i - 1)] * bp[IX(k, j, i - 1)] * bp[c]);
                r = r - ae[i] * phi[IX(k, j, i + 1)] * bp[c] * bp[IX(k, j, i + 1)];
                r = r - as[j] * phi[IX(k, j - 1, i)] * bp[IX(k, j - 1, i)] * bp[c];
                r = r - an[j] * phi[IX(k, j + 1, i)] * bp[c] * bp[IX(k, j + 1, i)];
                r = r - ab[k] * phi[IX(k - 1, j, i)] * bp[IX(k - 1, j, i)] * bp[c];
                r = r - at[k] * phi[IX(k + 1, j, i)] * bp[c] * bp[IX(k + 1, j, i)];
                r = r - ap[c] * phi[c];
                res[c] = r;
            }
}

void og_laplacephi(const og_topo *t, int level, double *res, const double *phi,
                   const double *aw, const double *ae, const double *as,
                   const double *an, const double *ab, const double *at,
                   const double *ap, const double *bp)
{
#pragma omp parallel for schedule(dynamic)
    for (int g = 1; g <= t->ngrid; ++g) {
        if (level >= 0 && t->level[g - 1] != level) continue;
        GDIMS(t, g);
        const int64_t o3 = t->ip3[g - 1];
        laplacephi_grid(kk, jj, ii, res + o3, phi + o3,
                        aw + t->ip1x[g - 1] - 1, ae + t->ip1x[g - 1] - 1,
                        an + t->ip1y[g - 1] - 1, as + t->ip1y[g - 1] - 1,
                        at + t->ip1z[g - 1] - 1, ab + t->ip1z[g - 1] - 1,
                        ap + o3, bp + o3);
    }
}

/* ------------------------------------------------------------------ */
/* src/flow/sip_hyperplane_mod.F90:450-505 */
void og_sipiter1_hp(const og_topo *t, int level, const double *rhs_,
                    double *res_, const double *lw_, const double *ls_,
                    const double *lb_, const double *lpr_, const int64_t *mip_,
                    const int64_t *idx_)
{
#pragma omp parallel for schedule(dynamic)
    for (int g = 1; g <= t->ngrid; ++g) {
        if (level >= 0 && t->level[g - 1] != level) continue;
        GDIMS(t, g);
        const int64_t o = t->ip3[g - 1] - 1; /* 1-based views */
        const double *rhs = rhs_ + o, *lw = lw_ + o, *ls = ls_ + o,
                     *lb = lb_ + o, *lpr = lpr_ + o;
        double *res = res_ + o;
        const int64_t *mip = mip_ + o, *idxs = idx_ + o;
        const int n3dmin = 9, n3dmax = (ii - 2) + (jj - 2) + (kk - 2);
        for (int m = n3dmin; m <= n3dmax; ++m) {
            const int64_t lm = mip[m], lp = mip[m + 1] - lm;
            for (int64_t ip = 1; ip <= lp; ++ip) {
                const int64_t iacc = lm + ip;
                const int64_t id = idxs[iacc];
                double r = (rhs[id] + res[id]) * lpr[iacc];
                r = r - lb[iacc] * res[id - 1] - ls[iacc] * res[id - kk]
                      - lw[iacc] * res[id - (int64_t)kk * jj];
                res[id] = r;
            }
        }
    }
}

/* src/flow/sip_hyperplane_mod.F90:508-569 */
void og_sipiter2_hp(const og_topo *t, int level, double *phi_, double *res_,
                    const double *ue_, const double *un_, const double *ut_,
                    const int64_t *mip_, const int64_t *idx_)
{
#pragma omp parallel for schedule(dynamic)
    for (int g = 1; g <= t->ngrid; ++g) {
        if (level >= 0 && t->level[g - 1] != level) continue;
        GDIMS(t, g);
        const int64_t o = t->ip3[g - 1] - 1;
        const double *ue = ue_ + o, *un = un_ + o, *ut = ut_ + o;
        double *res = res_ + o, *phi = phi_ + o;
        const int64_t *mip = mip_ + o, *idxs = idx_ + o;
        const int n3dmin = 9, n3dmax = (ii - 2) + (jj - 2) + (kk - 2);
        for (int m = n3dmax; m >= n3dmin; --m) {
            const int64_t lm = mip[m], lp = mip[m + 1] - lm;
            for (int64_t ip = 1; ip <= lp; ++ip) {
                const int64_t iacc = lm + ip;
                const int64_t id = idxs[iacc];
                double r = res[id];
                r = r - ut[iacc] * res[id + 1] - un[iacc] * res[id + kk]
                      - ue[iacc] * res[id + (int64_t)kk * jj];
                res[id] = r;
            }
        }
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k) {
                    const int64_t id = (int64_t)k + (int64_t)(j - 1) * kk
                                       + (int64_t)(i - 1) * kk * jj;
                    phi[id] = phi[id] + res[id];
                }
    }
}

/* src/flow/sip_classic_mod.F90:143-166 */
void og_sipiter1_cl(const og_topo *t, int level, const double *rhs_,
                    double *res_, const double *lw_, const double *ls_,
                    const double *lb_, const double *lpr_)
{
#pragma omp parallel for schedule(dynamic)
    for (int g = 1; g <= t->ngrid; ++g) {
        if (level >= 0 && t->level[g - 1] != level) continue;
        GDIMS(t, g);
        const int64_t o = t->ip3[g - 1];
        const double *rhs = rhs_ + o, *lw = lw_ + o, *ls = ls_ + o,
                     *lb = lb_ + o, *lpr = lpr_ + o;
        double *res = res_ + o;
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j) {
                for (int k = 3; k <= kk - 2; ++k) {
                    const int64_t c = IX(k, j, i);
                    double r = (rhs[c] + res[c]) * lpr[c];
                    r = r - lw[c] * res[IX(k, j, i - 1)]
                          - ls[c] * res[IX(k, j - 1, i)];
                    res[c] = r;
                }
                for (int k = 3; k <= kk - 2; ++k) {
                    const int64_t c = IX(k, j, i);
                    res[c] = res[c] - lb[c] * res[IX(k - 1, j, i)];
                }
            }
    }
}

/* src/flow/sip_classic_mod.F90:169-198 */
void og_sipiter2_cl(const og_topo *t, int level, double *phi_, double *res_,
                    const double *ue_, const double *un_, const double *ut_)
{
#pragma omp parallel for schedule(dynamic)
    for (int g = 1; g <= t->ngrid; ++g) {
        if (level >= 0 && t->level[g - 1] != level) continue;
        GDIMS(t, g);
        const int64_t o = t->ip3[g - 1];
        const double *ue = ue_ + o, *un = un_ + o, *ut = ut_ + o;
        double *res = res_ + o, *phi = phi_ + o;
        for (int i = ii - 2; i >= 3; --i)
            for (int j = jj - 2; j >= 3; --j) {
                for (int k = 3; k <= kk - 2; ++k) {
                    const int64_t c = IX(k, j, i);
                    res[c] = res[c] - un[c] * res[IX(k, j + 1, i)]
                                    - ue[c] * res[IX(k, j, i + 1)];
                }
                for (int k = kk - 2; k >= 3; --k) {
                    const int64_t c = IX(k, j, i);
                    res[c] = res[c] - ut[c] * res[IX(k + 1, j, i)];
                }
            }
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k) {
                    const int64_t c = IX(k, j, i);
                    phi[c] = phi[c] + res[c];
                }
    }
}

/* ------------------------------------------------------------------ */
/* src/flow/sor_mod.F90:18-49 (divide0: src/core/simdfunctions_mod.F90:70-82) */
void og_sor_init(const og_topo *t, const double *ap_, double *rap_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        const double *ap = ap_ + t->ip3[g - 1];
        double *rap = rap_ + t->ip3[g - 1];
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k) {
                    const double b = ap[IX(k, j, i)];
                    rap[IX(k, j, i)] = (b == 0.0) ? 0.0 : 1.0 / b;
                }
    }
}

/* src/flow/sor_mod.F90:102-203 */
void og_relax(const og_topo *t, int level, double *dp_, const double *rhs_,
              const double *aw_, const double *ae_, const double *as_,
              const double *an_, const double *ab_, const double *at_,
              const double *rap_, const double *bp_, double omg)
{
#pragma omp parallel for schedule(dynamic)
    for (int g = 1; g <= t->ngrid; ++g) {
        if (level >= 0 && t->level[g - 1] != level) continue;
        GDIMS(t, g);
        const int64_t o3 = t->ip3[g - 1];
        double *dp = dp_ + o3;
        const double *rhs = rhs_ + o3, *rap = rap_ + o3;
        const double *bp = bp_ ? bp_ + o3 : NULL;
        const double *gw = aw_ + t->ip1x[g - 1] - 1, *ge = ae_ + t->ip1x[g - 1] - 1;
        const double *gs = as_ + t->ip1y[g - 1] - 1, *gn = an_ + t->ip1y[g - 1] - 1;
        const double *gb = ab_ + t->ip1z[g - 1] - 1, *gt = at_ + t->ip1z[g - 1] - 1;
        for (int rb = 0; rb <= 1; ++rb)
            for (int i = 3; i <= ii - 2; ++i)
                for (int j = 3; j <= jj - 2; ++j) {
                    int kstart, kstop;
                    if ((i + j) % 2 == rb) { kstart = 4; kstop = kk - 2; }
                    else { kstart = 3; kstop = kk - 3; }
                    for (int k = kstart; k <= kstop; k += 2) {
                        const int64_t c = IX(k, j, i);
                        double res;
                        if (bp) {
                            const double aw = bp[IX(k, j, i - 1)] * bp[c] * gw[i];
                            const double ae = bp[c] * bp[IX(k, j, i + 1)] * ge[i];
                            const double as = bp[IX(k, j - 1, i)] * bp[c] * gs[j];
                            const double an = bp[c] * bp[IX(k, j + 1, i)] * gn[j];
                            const double ab = bp[IX(k - 1, j, i)] * bp[c] * gb[k];
                            const double at = bp[c] * bp[IX(k + 1, j, i)] * gt[k];
                            res = (aw * dp[IX(k, j, i - 1)] + ae * dp[IX(k, j, i + 1)]
                                   + as * dp[IX(k, j - 1, i)] + an * dp[IX(k, j + 1, i)]
                                   + ab * dp[IX(k - 1, j, i)] + at * dp[IX(k + 1, j, i)]
                                   - rhs[c]) * rap[c];
                        } else {
                            res = (gw[i] * dp[IX(k, j, i - 1)] + ge[i] * dp[IX(k, j, i + 1)]
                                   + gs[j] * dp[IX(k, j - 1, i)] + gn[j] * dp[IX(k, j + 1, i)]
                                   + gb[k] * dp[IX(k - 1, j, i)] + gt[k] * dp[IX(k + 1, j, i)]
                                   - rhs[c]) * rap[c];
                        }
                        dp[c] = (1.0 - omg) * dp[c] - omg * res;
                    }
                }
    }
}

/* ------------------------------------------------------------------ */
/* src/ib/noib_mod.F90:304-354 */
void og_divcal(const og_topo *t, double *div_, const double *u_,
               const double *v_, const double *w_, const double *rddx_,
               const double *rddy_, const double *rddz_, double fak,
               const double *bp_, const double *sdiv_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        const int64_t o3 = t->ip3[g - 1];
        double *div = div_ + o3;
        const double *u = u_ + o3, *v = v_ + o3, *w = w_ + o3;
        const double *rddx = rddx_ + t->ip1x[g - 1] - 1;
        const double *rddy = rddy_ + t->ip1y[g - 1] - 1;
        const double *rddz = rddz_ + t->ip1z[g - 1] - 1;
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k) {
                    const int64_t c = IX(k, j, i);
                    div[c] = fak * ((u[c] - u[IX(k, j, i - 1)]) * rddx[i]
                                    + (v[c] - v[IX(k, j - 1, i)]) * rddy[j]
                                    + (w[c] - w[IX(k - 1, j, i)]) * rddz[k]);
                }
        if (sdiv_) {
            const double *sdiv = sdiv_ + o3;
            for (int i = 3; i <= ii - 2; ++i)
                for (int j = 3; j <= jj - 2; ++j)
                    for (int k = 3; k <= kk - 2; ++k)
                        div[IX(k, j, i)] = div[IX(k, j, i)] + fak * sdiv[IX(k, j, i)];
        }
        if (bp_) {
            const double *bp = bp_ + o3;
            for (int64_t c = 0; c < (int64_t)ii * jj * kk; ++c)
                div[c] = bp[c] * div[c];
        }
    }
}

/* src/flow/pressuresolver_mod.F90:787-879 */
void og_mgpcorr(const og_topo *t, double *u_, double *v_, double *w_,
                double *p_, const double *dp_, const double *rdx_,
                const double *rdy_, const double *rdz_, double rfak,
                const double *bp_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        const int64_t o3 = t->ip3[g - 1];
        double *u = u_ + o3, *v = v_ + o3, *w = w_ + o3, *p = p_ + o3;
        const double *dp = dp_ + o3;
        const double *rdx = rdx_ + t->ip1x[g - 1] - 1;
        const double *rdy = rdy_ + t->ip1y[g - 1] - 1;
        const double *rdz = rdz_ + t->ip1z[g - 1] - 1;
        if (bp_) {
            const double *bp = bp_ + o3;
            for (int i = 2; i <= ii - 1; ++i)
                for (int j = 2; j <= jj - 1; ++j)
                    for (int k = 2; k <= kk - 1; ++k)
                        p[IX(k, j, i)] = p[IX(k, j, i)] + dp[IX(k, j, i)] * bp[IX(k, j, i)];
            for (int i = 2; i <= ii - 2; ++i)
                for (int j = 3; j <= jj - 2; ++j)
                    for (int k = 3; k <= kk - 2; ++k)
                        u[IX(k, j, i)] = u[IX(k, j, i)]
                            + (dp[IX(k, j, i)] - dp[IX(k, j, i + 1)])
                              * bp[IX(k, j, i)] * bp[IX(k, j, i + 1)] * rdx[i] * rfak;
            for (int i = 3; i <= ii - 2; ++i)
                for (int j = 2; j <= jj - 2; ++j)
                    for (int k = 3; k <= kk - 2; ++k)
                        v[IX(k, j, i)] = v[IX(k, j, i)]
                            + (dp[IX(k, j, i)] - dp[IX(k, j + 1, i)])
                              * bp[IX(k, j, i)] * bp[IX(k, j + 1, i)] * rdy[j] * rfak;
            for (int i = 3; i <= ii - 2; ++i)
                for (int j = 3; j <= jj - 2; ++j)
                    for (int k = 2; k <= kk - 2; ++k)
                        w[IX(k, j, i)] = w[IX(k, j, i)]
                            + (dp[IX(k, j, i)] - dp[IX(k + 1, j, i)])
                              * bp[IX(k, j, i)] * bp[IX(k + 1, j, i)] * rdz[k] * rfak;
        } else {
            for (int i = 2; i <= ii - 1; ++i)
                for (int j = 2; j <= jj - 1; ++j)
                    for (int k = 2; k <= kk - 1; ++k)
                        p[IX(k, j, i)] = p[IX(k, j, i)] + dp[IX(k, j, i)];
            for (int i = 2; i <= ii - 2; ++i)
                for (int j = 3; j <= jj - 2; ++j)
                    for (int k = 3; k <= kk - 2; ++k)
                        u[IX(k, j, i)] = u[IX(k, j, i)]
                            + (dp[IX(k, j, i)] - dp[IX(k, j, i + 1)]) * rdx[i] * rfak;
            for (int i = 3; i <= ii - 2; ++i)
                for (int j = 2; j <= jj - 2; ++j)
                    for (int k = 3; k <= kk - 2; ++k)
                        v[IX(k, j, i)] = v[IX(k, j, i)]
                            + (dp[IX(k, j, i)] - dp[IX(k, j + 1, i)]) * rdy[j] * rfak;
            for (int i = 3; i <= ii - 2; ++i)
                for (int j = 3; j <= jj - 2; ++j)
                    for (int k = 2; k <= kk - 2; ++k)
                        w[IX(k, j, i)] = w[IX(k, j, i)]
                            + (dp[IX(k, j, i)] - dp[IX(k + 1, j, i)]) * rdz[k] * rfak;
        }
    }
}

/* src/flow/pressuresolver_mod.F90:623-689 */
double og_maxabscal(const og_topo *t, const double *phi_, double *maxlvl)
{
    double maxabs = 0.0;
    if (maxlvl)
        for (int l = t->minlevel; l <= t->maxlevel; ++l)
            maxlvl[l - t->minlevel] = 0.0;
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        const double *phi = phi_ + t->ip3[g - 1];
        double m = 0.0;
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k) {
                    const double a = fabs(phi[IX(k, j, i)]);
                    /* NaN propagates so that a diverging solve is seen */
                    m = (a > m || a != a) ? a : m;
                }
        const double s = ldexp(1.0, t->maxlevel - t->level[g - 1]);
        const double ms = fabs(m * s);
        maxabs = (ms > maxabs || ms != ms) ? ms : maxabs;
        if (maxlvl) {
            double *ml = &maxlvl[t->level[g - 1] - t->minlevel];
            *ml = (m > *ml) ? m : *ml;
        }
    }
    return maxabs;
}

/* src/flow/pressuresolver_mod.F90:722-742 */
void og_rescal(const og_topo *t, double *rhs_, const double *res_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        double *rhs = rhs_ + t->ip3[g - 1];
        const double *res = res_ + t->ip3[g - 1];
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k)
                    rhs[IX(k, j, i)] = rhs[IX(k, j, i)] + res[IX(k, j, i)];
    }
}

/* src/flow/pressuresolver_mod.F90:882-908 */
void og_accumulate(int64_t n, double *dp, const double *hilf)
{
    for (int64_t c = 0; c < n; ++c) dp[c] = dp[c] + hilf[c];
}

/* ------------------------------------------------------------------ */
/* src/core/connect/connect_core_mod.F90:373-433 */
static void get_nbrs(const og_topo *t, int iface, const int32_t *nbrs,
                     int *nbrgrid, int *nbrface)
{
    *nbrgrid = 0;
    *nbrface = 0;
    for (int r = 0; r < 7; ++r) {
        const int dir = RESCUE_DIR[iface - 1][r];
        if (dir == 0) break;
        if (nbrs[dir - 1] > 0) {
            *nbrgrid = nbrs[dir - 1];
            *nbrface = RESCUE_NBR[iface - 1][r];
            if (*nbrface > 18) {
                const int32_t *bcn = t->bc + 6 * (*nbrgrid - 1);
                const int f1 = FACELIST[*nbrface - 1][1];
                const int f2 = FACELIST[*nbrface - 1][2];
                const int f3 = FACELIST[*nbrface - 1][3];
                if (!is_con(bcn[f1 - 1]) && !is_con(bcn[f2 - 1]) &&
                    !is_con(bcn[f3 - 1])) {
                    *nbrgrid = 0;
                    *nbrface = 0;
                    continue;
                }
            }
            break;
        }
    }
}

/* src/core/connect/connect_core_mod.F90:167-370 (one rank: the sort by
 * sending rank is a stable no-op, sendconns == recvconns). */
int og_connect_build(const og_topo *t, og_conn *out, int cap)
{
    int n = 0;
    for (int g = 1; g <= t->ngrid; ++g) {
        const int32_t *nb = t->nbr + 26 * (g - 1);
        const int32_t *bc = t->bc + 6 * (g - 1);
        for (int iface = 1; iface <= 26; ++iface) {
            int want;
            if (iface <= 6) {
                want = is_con(bc[iface - 1]);
            } else if (iface <= 18) {
                want = is_con(bc[FACELIST[iface - 1][1] - 1]) ||
                       is_con(bc[FACELIST[iface - 1][2] - 1]);
            } else {
                want = is_con(bc[FACELIST[iface - 1][1] - 1]) ||
                       is_con(bc[FACELIST[iface - 1][2] - 1]) ||
                       is_con(bc[FACELIST[iface - 1][3] - 1]);
            }
            if (!want) continue;
            int ng, nf;
            get_nbrs(t, iface, nb, &ng, &nf);
            if (ng == 0) continue;
            if (n >= cap) return -1;
            out[n].recv_grid = g;
            out[n].send_grid = ng;
            out[n].recv_face = iface;
            out[n].send_face = nf;
            ++n;
        }
    }
    return n;
}

/* src/core/connect/connect_core_mod.F90:475-548 */
static void start_and_stop_face(const og_topo *t, int g, int iface, int *is,
                                int *ie, int *js, int *je, int *ks, int *ke,
                                int nplane, int nghost)
{
    GDIMS(t, g);
    switch (iface) {
    case 1: *is = 3 - nplane; *ie = 2; break;
    case 2: *is = ii - 1; *ie = ii - (2 - nplane); break;
    case 3: *js = 3 - nplane; *je = 2; break;
    case 4: *js = jj - 1; *je = jj - (2 - nplane); break;
    case 5: *ks = 3 - nplane; *ke = 2; break;
    case 6: *ks = kk - 1; *ke = kk - (2 - nplane); break;
    }
    if (nghost) {
        switch (iface) {
        case 1: *is += nplane; *ie += nplane; break;
        case 2: *is -= nplane; *ie -= nplane; break;
        case 3: *js += nplane; *je += nplane; break;
        case 4: *js -= nplane; *je -= nplane; break;
        case 5: *ks += nplane; *ke += nplane; break;
        case 6: *ks -= nplane; *ke -= nplane; break;
        }
    }
}

/* src/core/connect/connect_core_mod.F90:454-472 */
static void start_and_stop(const og_topo *t, int g, int iface, int *is,
                           int *ie, int *js, int *je, int *ks, int *ke,
                           int nplane, int nghost)
{
    GDIMS(t, g);
    *is = 3; *ie = ii - 2; *js = 3; *je = jj - 2; *ks = 3; *ke = kk - 2;
    for (int f = 1; f <= FACELIST[iface - 1][0]; ++f)
        start_and_stop_face(t, g, FACELIST[iface - 1][f], is, ie, js, je, ks,
                            ke, nplane, nghost);
}

/* src/core/connect/connect_core_mod.F90:575-602 */
static void corr_start_stop(const og_topo *t, int g, int fsend, int frecv,
                            int *is, int *ie, int *js, int *je, int *ks,
                            int *ke, int nplane)
{
    if (frecv <= 6) return;
    for (int a = 1; a <= FACELIST[fsend - 1][0]; ++a)
        for (int b = 1; b <= FACELIST[frecv - 1][0]; ++b)
            if (FACELIST[fsend - 1][a] == FACELIST[frecv - 1][b])
                start_and_stop_face(t, g, FACELIST[fsend - 1][a], is, ie, js,
                                    je, ks, ke, nplane, 0);
}

/* src/core/connect/connect_core_mod.F90:640-682 */
static int decide(const og_topo *t, const og_conn *c, int vertices, int fwd,
                  int minlvl, int maxlvl)
{
    int ex = 1;
    const int fr = c->recv_face;
    const int lvl = t->level[c->recv_grid - 1];
    if (lvl > maxlvl || lvl < minlvl) ex = 0;
    if (!vertices && fr > 6) ex = 0;
    if (fwd == 1 && (fr == 2 || fr == 4 || fr == 6)) ex = 0;
    else if (fwd == -1 && (fr == 1 || fr == 3 || fr == 5)) ex = 0;
    return ex;
}

/* src/core/connect/conn1_mod.F90:31-129 (one rank), connect_self :561-704 */
int og_exchange(const og_topo *t, const og_conn *conns, int nconn, int level,
            int layers, int corners, int forward, double **fields, int nfield)
{
    if (layers < 1 || layers > 2) return -1;
    if (forward != 0 && corners) return -2;
    const int minl = level >= 0 ? level : t->minlevel;
    const int maxl = level >= 0 ? level : t->maxlevel;
    for (int s = 0; s < nconn; ++s) {
        const og_conn *c = &conns[s];
        if (!decide(t, c, corners, forward, minl, maxl)) continue;
        const int gd = c->recv_grid, gs = c->send_grid;
        int is, ie, js, je, ks, ke, isd, ied, jsd, jed, ksd, ked;
        start_and_stop(t, gs, FACENBR[c->recv_face - 1], &is, &ie, &js, &je,
                       &ks, &ke, layers, 1);
        corr_start_stop(t, gs, c->send_face, c->recv_face, &is, &ie, &js, &je,
                        &ks, &ke, layers);
        start_and_stop(t, gd, c->recv_face, &isd, &ied, &jsd, &jed, &ksd,
                       &ked, layers, 0);
        if ((ie - is + 1) * (je - js + 1) * (ke - ks + 1) !=
            (ied - isd + 1) * (jed - jsd + 1) * (ked - ksd + 1))
            return -3;
        const int koff = ks - ksd, joff = js - jsd, ioff = is - isd;
        const int kks = t->kk[gs - 1], jjs = t->jj[gs - 1];
        const int kkd = t->kk[gd - 1], jjd = t->jj[gd - 1];
        for (int f = 0; f < nfield; ++f) {
            const double *src = fields[f] + t->ip3[gs - 1];
            double *dst = fields[f] + t->ip3[gd - 1];
            for (int i = isd; i <= ied; ++i)
                for (int j = jsd; j <= jed; ++j)
                    for (int k = ksd; k <= ked; ++k)
                        dst[(k - 1) + (int64_t)(j - 1) * kkd +
                            (int64_t)(i - 1) * kkd * jjd] =
                            src[(k + koff - 1) + (int64_t)(j + joff - 1) * kks +
                                (int64_t)(i + ioff - 1) * kks * jjs];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Multi-level coupling (levels are nested: a grid's parent is one level
 * coarser; iposition.. is the parent's 1-based index of the first parent
 * cell the grid covers, src/core/gridio_mod.F90:15-37). */
#define PAR_BC 8

/* PAR face buffers: offsets (-1 = none) and total length; one buffer per
 * PAR face of kk*jj / kk*ii / jj*ii values (src/core/pointers_mod.F90:
 * 64-92). */
int64_t og_par_buffers(const og_topo *t, int64_t *ipbb)
{
    int64_t n = 0;
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        for (int f = 1; f <= 6; ++f) {
            int64_t *o = &ipbb[6 * (g - 1) + f - 1];
            if (t->bc[6 * (g - 1) + f - 1] != PAR_BC) { *o = -1; continue; }
            *o = n;
            n += f <= 2 ? (int64_t)kk * jj : f <= 4 ? (int64_t)kk * ii : (int64_t)jj * ii;
        }
    }
    return n;
}

#define POS(g, a) (t->pos[3 * ((g)-1) + (a)])

/* ctof1: write_buffer box (src/ib/ctof/ctof1_mod.F90:232-262) and
 * unpack_single (:308-331): piecewise-constant injection of the parent's
 * values into the whole fine array, ghost layers included. */
void og_ctof(const og_topo *t, int level, double *ff_, const double *fc_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        if (t->level[g - 1] != level) continue;
        const int c = t->iparent[g - 1];
        if (c == 0) continue;
        GDIMS(t, g);
        const int ipos = POS(g, 0), jpos = POS(g, 1), kpos = POS(g, 2);
        const int64_t kkc = t->kk[c - 1], jjc = t->jj[c - 1];
        double *ff = ff_ + t->ip3[g - 1];
        const double *fc = fc_ + t->ip3[c - 1];
        for (int i = 1; i <= ii; ++i)
            for (int j = 1; j <= jj; ++j)
                for (int k = 1; k <= kk; ++k) {
                    const int ic = ipos - 1 + (i - 1) / 2;
                    const int jc = jpos - 1 + (j - 1) / 2;
                    const int kc = kpos - 1 + (k - 1) / 2;
                    ff[IX(k, j, i)] = fc[(kc - 1) + (jc - 1) * kkc + (ic - 1) * kkc * jjc];
                }
    }
}

/* parent for one scalar field: the coarse slab of start_and_stop
 * (src/ib/parent/parent_core_mod.F90:214-262, icomp = 0), packed as in
 * parent1 pack_single (src/ib/parent/parent1_mod.F90:296-338) and prolonged
 * to the fine face with prolong1/prolong2, istag = 0 (:523-583): fine face
 * cell (a, b) takes slab cell (2 + (a-1)/2, 2 + (b-1)/2). */
void og_parent(const og_topo *t, int level, const double *f_, double *pbuf,
               const int64_t *ipbb)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        if (t->level[g - 1] != level) continue;
        const int c = t->iparent[g - 1];
        if (c == 0) continue;
        GDIMS(t, g);
        const int ipos = POS(g, 0), jpos = POS(g, 1), kpos = POS(g, 2);
        const int64_t kkc = t->kk[c - 1], jjc = t->jj[c - 1];
        const double *fc = f_ + t->ip3[c - 1];
#define FC(k, j, i) fc[((k)-1) + ((j)-1) * kkc + ((i)-1) * kkc * jjc]
        for (int f = 1; f <= 6; ++f) {
            const int64_t o = ipbb[6 * (g - 1) + f - 1];
            if (o < 0) continue;
            double *pb = pbuf + o;
            int ks = kpos - 2, js = jpos - 2, is = ipos - 2;
            switch (f) {
            case 1: is = ipos - 1; break;
            case 2: is = ipos + (ii - 4) / 2; break;
            case 3: js = jpos - 1; break;
            case 4: js = jpos + (jj - 4) / 2; break;
            case 5: ks = kpos - 1; break;
            case 6: ks = kpos + (kk - 4) / 2; break;
            }
            if (f <= 2) {
                for (int j = 1; j <= jj; ++j)
                    for (int k = 1; k <= kk; ++k)
                        pb[(k - 1) + (int64_t)(j - 1) * kk] =
                            FC(ks + 1 + (k - 1) / 2, js + 1 + (j - 1) / 2, is);
            } else if (f <= 4) {
                for (int i = 1; i <= ii; ++i)
                    for (int k = 1; k <= kk; ++k)
                        pb[(k - 1) + (int64_t)(i - 1) * kk] =
                            FC(ks + 1 + (k - 1) / 2, js, is + 1 + (i - 1) / 2);
            } else {
                for (int i = 1; i <= ii; ++i)
                    for (int j = 1; j <= jj; ++j)
                        pb[(j - 1) + (int64_t)(i - 1) * jj] =
                            FC(ks, js + 1 + (j - 1) / 2, is + 1 + (i - 1) / 2);
            }
        }
#undef FC
    }
}

/* pressureftocone_A / _B: src/flow/bound_pressure_mod.F90:536-608 */
static void pftocone(int k, int j, int i, int kk, int jj, const double *p,
                     const double *bp, const double *ddx, const double *ddy,
                     const double *ddz, double *pc, double *bpc)
{
    if (!bp) {
        *bpc = 1.0;
        double sump = 0.0, sumvol = 0.0;
        for (int l = 0; l <= 1; ++l)
            for (int m = 0; m <= 1; ++m)
                for (int n = 0; n <= 1; ++n) {
                    const double vol = ddz[k + n] * ddy[j + m] * ddx[i + l];
                    sump = sump + p[IX(k + n, j + m, i + l)] * vol;
                    sumvol = sumvol + vol;
                }
        *pc = sump / sumvol;
        return;
    }
    double sumbp = 0.0;
    for (int l = 0; l <= 1; ++l)
        for (int m = 0; m <= 1; ++m)
            for (int n = 0; n <= 1; ++n)
                sumbp = sumbp + bp[IX(k + n, j + m, i + l)];
    *bpc = sumbp < 1.0 ? sumbp : 1.0;
    if (*bpc < 0.5) {
        *pc = 0.0;
        return;
    }
    double sump = 0.0, sumvol = 0.0;
    for (int l = 0; l <= 1; ++l)
        for (int m = 0; m <= 1; ++m)
            for (int n = 0; n <= 1; ++n) {
                const double vol = bp[IX(k + n, j + m, i + l)] * ddz[k + n] *
                                   ddy[j + m] * ddx[i + l];
                sump = sump + p[IX(k + n, j + m, i + l)] * vol;
                sumvol = sumvol + vol;
            }
    *pc = sump / sumvol;
}

/* bound_pressure: src/flow/bound_pressure_mod.F90:106-534 (bfront,
 * bright, bbottom and their _bp forms); bp NULL selects the A forms.
 * Metric pointers are 1-based. */
void og_bound_pressure(const og_topo *t, int level, double *p_,
                       const double *pbuf, const int64_t *ipbb,
                       const double *bp_, const double *dx_,
                       const double *dy_, const double *dz_,
                       const double *ddx_, const double *ddy_,
                       const double *ddz_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        if (t->level[g - 1] != level) continue;
        GDIMS(t, g);
        double *p = p_ + t->ip3[g - 1];
        const double *bp = bp_ ? bp_ + t->ip3[g - 1] : NULL;
        const double *dx = dx_ + t->ip1x[g - 1] - 1, *ddx = ddx_ + t->ip1x[g - 1] - 1;
        const double *dy = dy_ + t->ip1y[g - 1] - 1, *ddy = ddy_ + t->ip1y[g - 1] - 1;
        const double *dz = dz_ + t->ip1z[g - 1] - 1, *ddz = ddz_ + t->ip1z[g - 1] - 1;
        for (int f = 1; f <= 6; ++f) {
            const int64_t o = ipbb[6 * (g - 1) + f - 1];
            if (o < 0) continue;
            const double *pb = pbuf + o;
            if (f <= 2) {
                const int i2 = f == 1 ? 2 : ii - 1, i3 = f == 1 ? 3 : ii - 2;
                const int i4 = f == 1 ? 4 : ii - 3, is2 = f == 1 ? 2 : ii - 2;
                const int i = i3 < i4 ? i3 : i4;
                for (int j = 3; j <= jj - 2; j += 2)
                    for (int k = 3; k <= kk - 2; k += 2) {
                        double pcnew, bpc, fak = 1.0;
                        pftocone(k, j, i, kk, jj, p, bp, ddx, ddy, ddz, &pcnew, &bpc);
                        const double pbv = pb[(k - 1) + (int64_t)(j - 1) * kk];
                        if (bp) {
                            if (bpc < 0.5) pcnew = pbv;
                            const double sb11 = bp[IX(k, j, i2)] * bp[IX(k, j, i3)];
                            const double sb12 = bp[IX(k, j + 1, i2)] * bp[IX(k, j + 1, i3)];
                            const double sb13 = bp[IX(k + 1, j, i2)] * bp[IX(k + 1, j, i3)];
                            const double sb14 = bp[IX(k + 1, j + 1, i2)] * bp[IX(k + 1, j + 1, i3)];
                            fak = (sb11 * ddy[j] * ddz[k] + sb12 * ddy[j + 1] * ddz[k] +
                                   sb13 * ddy[j] * ddz[k + 1] + sb14 * ddy[j + 1] * ddz[k + 1]) /
                                  ((ddy[j] + ddy[j + 1]) * (ddz[k] + ddz[k + 1]));
                            if (fak < 0.1) fak = 1.0;
                            fak = 1.0 / fak;
                        }
                        for (int m = 0; m <= 1; ++m)
                            for (int n = 0; n <= 1; ++n)
                                p[IX(k + n, j + m, i2)] = bp
                                    ? p[IX(k + n, j + m, i3)] + fak * dx[is2] / (ddx[i3] + ddx[i2]) * (pbv - pcnew)
                                    : p[IX(k + n, j + m, i3)] + dx[is2] / (ddx[i3] + ddx[i2]) * (pbv - pcnew);
                    }
            } else if (f <= 4) {
                const int j2 = f == 3 ? 2 : jj - 1, j3 = f == 3 ? 3 : jj - 2;
                const int j4 = f == 3 ? 4 : jj - 3, js2 = f == 3 ? 2 : jj - 2;
                const int j = j3 < j4 ? j3 : j4;
                for (int i = 3; i <= ii - 2; i += 2)
                    for (int k = 3; k <= kk - 2; k += 2) {
                        double pcnew, bpc, fak = 1.0;
                        pftocone(k, j, i, kk, jj, p, bp, ddx, ddy, ddz, &pcnew, &bpc);
                        const double pbv = pb[(k - 1) + (int64_t)(i - 1) * kk];
                        if (bp) {
                            if (bpc < 0.5) pcnew = pbv;
                            const double sb11 = bp[IX(k, j2, i)] * bp[IX(k, j3, i)];
                            const double sb12 = bp[IX(k, j2, i + 1)] * bp[IX(k, j3, i + 1)];
                            const double sb13 = bp[IX(k + 1, j2, i)] * bp[IX(k + 1, j3, i)];
                            const double sb14 = bp[IX(k + 1, j2, i + 1)] * bp[IX(k + 1, j3, i + 1)];
                            fak = (sb11 * ddx[i] * ddz[k] + sb12 * ddx[i + 1] * ddz[k] +
                                   sb13 * ddx[i] * ddz[k + 1] + sb14 * ddx[i + 1] * ddz[k + 1]) /
                                  ((ddx[i] + ddx[i + 1]) * (ddz[k] + ddz[k + 1]));
                            if (fak < 0.1) fak = 1.0;
                            fak = 1.0 / fak;
                        }
                        for (int l = 0; l <= 1; ++l)
                            for (int n = 0; n <= 1; ++n)
                                p[IX(k + n, j2, i + l)] = bp
                                    ? p[IX(k + n, j3, i + l)] + fak * dy[js2] / (ddy[j3] + ddy[j2]) * (pbv - pcnew)
                                    : p[IX(k + n, j3, i + l)] + dy[js2] / (ddy[j3] + ddy[j2]) * (pbv - pcnew);
                    }
            } else {
                const int k2 = f == 5 ? 2 : kk - 1, k3 = f == 5 ? 3 : kk - 2;
                const int k4 = f == 5 ? 4 : kk - 3, ks2 = f == 5 ? 2 : kk - 2;
                const int k = k3 < k4 ? k3 : k4;
                for (int i = 3; i <= ii - 2; i += 2)
                    for (int j = 3; j <= jj - 2; j += 2) {
                        double pcnew, bpc, fak = 1.0;
                        pftocone(k, j, i, kk, jj, p, bp, ddx, ddy, ddz, &pcnew, &bpc);
                        const double pbv = pb[(j - 1) + (int64_t)(i - 1) * jj];
                        if (bp) {
                            if (bpc < 0.5) pcnew = pbv;
                            const double sb11 = bp[IX(k2, j, i)] * bp[IX(k3, j, i)];
                            const double sb12 = bp[IX(k2, j, i + 1)] * bp[IX(k3, j, i + 1)];
                            const double sb13 = bp[IX(k2, j + 1, i)] * bp[IX(k3, j + 1, i)];
                            const double sb14 = bp[IX(k2, j + 1, i + 1)] * bp[IX(k3, j + 1, i + 1)];
                            fak = (sb11 * ddx[i] * ddy[j] + sb12 * ddx[i + 1] * ddy[j] +
                                   sb13 * ddx[i] * ddy[j + 1] + sb14 * ddx[i + 1] * ddy[j + 1]) /
                                  ((ddx[i] + ddx[i + 1]) * (ddy[j] + ddy[j + 1]));
                            if (fak < 0.1) fak = 1.0;
                            fak = 1.0 / fak;
                        }
                        for (int l = 0; l <= 1; ++l)
                            for (int m = 0; m <= 1; ++m)
                                p[IX(k2, j + m, i + l)] = bp
                                    ? p[IX(k3, j + m, i + l)] + fak * dz[ks2] / (ddz[k3] + ddz[k2]) * (pbv - pcnew)
                                    : p[IX(k3, j + m, i + l)] + dz[ks2] / (ddz[k3] + ddz[k2]) * (pbv - pcnew);
                    }
            }
        }
    }
}

/* ftoc of a cell-centred scalar into the parent cells under the grid
 * (restrict_recieve_open, src/ib/ftoc_mod.F90:581-626).  bp NULL: noib
 * restrict_s, the volume-weighted mean of each 2x2x2 fine group
 * (src/ib/noib_restrict_mod.F90:255-298).  With bp (ghostcell,
 * src/ib/gc_restrict_mod.F90): flag 'P' restrict_p, the BP-weighted mean
 * (0 for a fully blocked group, :59-164); flag 'R' restrict_r, the sum of
 * BP-weighted values over the full group volume (:167-259). */
void og_ftoc_s(const og_topo *t, int level, const double *ff_, double *fc_,
               const double *ddx_, const double *ddy_, const double *ddz_,
               const double *bp_, char flag)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        if (t->level[g - 1] != level) continue;
        const int c = t->iparent[g - 1];
        if (c == 0) continue;
        GDIMS(t, g);
        const int ipos = POS(g, 0), jpos = POS(g, 1), kpos = POS(g, 2);
        const int64_t kkc = t->kk[c - 1], jjc = t->jj[c - 1];
        const double *ff = ff_ + t->ip3[g - 1];
        const double *bp = bp_ ? bp_ + t->ip3[g - 1] : NULL;
        double *fc = fc_ + t->ip3[c - 1];
        const double *ddx = ddx_ + t->ip1x[g - 1] - 1, *ddy = ddy_ + t->ip1y[g - 1] - 1,
                     *ddz = ddz_ + t->ip1z[g - 1] - 1;
        for (int i = 3; i <= ii - 3; i += 2)
            for (int j = 3; j <= jj - 3; j += 2)
                for (int k = 3; k <= kk - 3; k += 2) {
                    double v;
                    if (!bp) {
                        const double spv =
                            ff[IX(k, j, i)] * ddz[k] * ddy[j] * ddx[i] +
                            ff[IX(k, j, i + 1)] * ddz[k] * ddy[j] * ddx[i + 1] +
                            ff[IX(k, j + 1, i)] * ddz[k] * ddy[j + 1] * ddx[i] +
                            ff[IX(k, j + 1, i + 1)] * ddz[k] * ddy[j + 1] * ddx[i + 1] +
                            ff[IX(k + 1, j, i)] * ddz[k + 1] * ddy[j] * ddx[i] +
                            ff[IX(k + 1, j, i + 1)] * ddz[k + 1] * ddy[j] * ddx[i + 1] +
                            ff[IX(k + 1, j + 1, i)] * ddz[k + 1] * ddy[j + 1] * ddx[i] +
                            ff[IX(k + 1, j + 1, i + 1)] * ddz[k + 1] * ddy[j + 1] * ddx[i + 1];
                        const double sv = (ddx[i] + ddx[i + 1]) * (ddy[j] + ddy[j + 1]) *
                                          (ddz[k] + ddz[k + 1]);
                        v = spv / sv;
                    } else if (flag == 'P') {
                        double vol[8], pv[8];
                        int q = 0;
                        for (int n = 0; n <= 1; ++n) {      /* order of :96-125 */
                            const int kq = k + n;
                            vol[q] = bp[IX(kq, j, i)] * ddz[kq] * ddy[j] * ddx[i];
                            vol[q + 1] = bp[IX(kq, j, i + 1)] * ddz[kq] * ddy[j] * ddx[i + 1];
                            vol[q + 2] = bp[IX(kq, j + 1, i)] * ddz[kq] * ddy[j + 1] * ddx[i];
                            vol[q + 3] = bp[IX(kq, j + 1, i + 1)] * ddz[kq] * ddy[j + 1] * ddx[i + 1];
                            pv[q] = ff[IX(kq, j, i)] * vol[q];
                            pv[q + 1] = ff[IX(kq, j, i + 1)] * vol[q + 1];
                            pv[q + 2] = ff[IX(kq, j + 1, i)] * vol[q + 2];
                            pv[q + 3] = ff[IX(kq, j + 1, i + 1)] * vol[q + 3];
                            q += 4;
                        }
                        const double spv = pv[0] + pv[1] + pv[2] + pv[3] + pv[4] + pv[5] +
                                           pv[6] + pv[7];
                        const double sv = vol[0] + vol[1] + vol[2] + vol[3] + vol[4] +
                                          vol[5] + vol[6] + vol[7];
                        v = sv > 0.0 ? spv / sv : 0.0;
                    } else {
                        double pv[8];
                        int q = 0;
                        for (int n = 0; n <= 1; ++n) {
                            const int kq = k + n;
                            pv[q] = ff[IX(kq, j, i)] * bp[IX(kq, j, i)] * ddz[kq] * ddy[j] * ddx[i];
                            pv[q + 1] = ff[IX(kq, j, i + 1)] * bp[IX(kq, j, i + 1)] * ddz[kq] *
                                        ddy[j] * ddx[i + 1];
                            pv[q + 2] = ff[IX(kq, j + 1, i)] * bp[IX(kq, j + 1, i)] * ddz[kq] *
                                        ddy[j + 1] * ddx[i];
                            pv[q + 3] = ff[IX(kq, j + 1, i + 1)] * bp[IX(kq, j + 1, i + 1)] *
                                        ddz[kq] * ddy[j + 1] * ddx[i + 1];
                            q += 4;
                        }
                        const double spv = pv[0] + pv[1] + pv[2] + pv[3] + pv[4] + pv[5] +
                                           pv[6] + pv[7];
                        const double sv = (ddx[i] + ddx[i + 1]) * (ddy[j] + ddy[j + 1]) *
                                          (ddz[k] + ddz[k + 1]);
                        v = spv / sv;
                    }
                    const int ic = ipos + (i - 3) / 2, jc = jpos + (j - 3) / 2,
                              kc = kpos + (k - 3) / 2;
                    fc[(kc - 1) + (jc - 1) * kkc + (ic - 1) * kkc * jjc] = v;
                }
    }
}

/* ftoc(ilevel, u, v, w, p): restrict_u/v/w/s (src/ib/noib_restrict_mod.F90:
 * 135-298) with the staggered offsets of restrict_recieve_open
 * (src/ib/ftoc_mod.F90:581-626). */
void og_ftoc_uvwp(const og_topo *t, int level, double *u_, double *v_,
                  double *w_, double *p_, const double *ddx_,
                  const double *ddy_, const double *ddz_, const double *bp_)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        if (t->level[g - 1] != level) continue;
        const int c = t->iparent[g - 1];
        if (c == 0) continue;
        GDIMS(t, g);
        const int ipos = POS(g, 0), jpos = POS(g, 1), kpos = POS(g, 2);
        const int64_t kkc = t->kk[c - 1], jjc = t->jj[c - 1];
        const int64_t of = t->ip3[g - 1], oc = t->ip3[c - 1];
        const double *ddx = ddx_ + t->ip1x[g - 1] - 1, *ddy = ddy_ + t->ip1y[g - 1] - 1,
                     *ddz = ddz_ + t->ip1z[g - 1] - 1;
#define CI(kc, jc, ic) (oc + ((kc)-1) + ((jc)-1) * kkc + ((ic)-1) * kkc * jjc)
        {   /* U */
            const double *ff = u_ + of;
            for (int i = 2; i <= ii - 2; i += 2)
                for (int j = 3; j <= jj - 3; j += 2)
                    for (int k = 3; k <= kk - 3; k += 2) {
                        const double sa = ff[IX(k, j, i)] * ddy[j] * ddz[k] +
                                          ff[IX(k, j + 1, i)] * ddy[j + 1] * ddz[k] +
                                          ff[IX(k + 1, j, i)] * ddy[j] * ddz[k + 1] +
                                          ff[IX(k + 1, j + 1, i)] * ddy[j + 1] * ddz[k + 1];
                        const double a = (ddy[j] + ddy[j + 1]) * (ddz[k] + ddz[k + 1]);
                        u_[CI(kpos + (k - 3) / 2, jpos + (j - 3) / 2,
                              ipos + (i - 2) / 2 - 1)] = sa / a;
                    }
        }
        {   /* V */
            const double *ff = v_ + of;
            for (int i = 3; i <= ii - 3; i += 2)
                for (int j = 2; j <= jj - 2; j += 2)
                    for (int k = 3; k <= kk - 3; k += 2) {
                        const double sa = ff[IX(k, j, i)] * ddx[i] * ddz[k] +
                                          ff[IX(k, j, i + 1)] * ddx[i + 1] * ddz[k] +
                                          ff[IX(k + 1, j, i)] * ddx[i] * ddz[k + 1] +
                                          ff[IX(k + 1, j, i + 1)] * ddx[i + 1] * ddz[k + 1];
                        const double a = (ddx[i] + ddx[i + 1]) * (ddz[k] + ddz[k + 1]);
                        v_[CI(kpos + (k - 3) / 2, jpos + (j - 2) / 2 - 1,
                              ipos + (i - 3) / 2)] = sa / a;
                    }
        }
        {   /* W */
            const double *ff = w_ + of;
            for (int i = 3; i <= ii - 3; i += 2)
                for (int j = 3; j <= jj - 3; j += 2)
                    for (int k = 2; k <= kk - 2; k += 2) {
                        const double sa = ff[IX(k, j, i)] * ddx[i] * ddy[j] +
                                          ff[IX(k, j, i + 1)] * ddx[i + 1] * ddy[j] +
                                          ff[IX(k, j + 1, i)] * ddx[i] * ddy[j + 1] +
                                          ff[IX(k, j + 1, i + 1)] * ddx[i + 1] * ddy[j + 1];
                        const double a = (ddx[i] + ddx[i + 1]) * (ddy[j] + ddy[j + 1]);
                        w_[CI(kpos + (k - 2) / 2 - 1, jpos + (j - 3) / 2,
                              ipos + (i - 3) / 2)] = sa / a;
                    }
        }
#undef CI
    }
    og_ftoc_s(t, level, p_, p_, ddx_, ddy_, ddz_, bp_, 'P');
}

/* ------------------------------------------------------------------ */
/* ------------------------------------------------------------------ */
/* Closing sync of mgpoisl: parent(u, v, w), bound_flow, connect per level
 * (src/flow/pressuresolver_mod.F90:338-348). */

#define FIX_BC 2
#define OP1_BC 3
#define NOS_BC 5
#define SLI_BC 6

/* Velocity boundary buffers (the field buffers at ipbb, src/core/
 * pointers_mod.F90:64-92) of every face whose flow boundary reads them:
 * FIX, OP1 (the caller's prescribed values) and PAR (filled by parent),
 * kk*jj / kk*ii / jj*ii values per face, first index fastest.  bbo[6*ngrid]
 * gets the offsets (-1 = none); returns the total length. */
int64_t og_bound_buffers(const og_topo *t, int64_t *bbo)
{
    int64_t n = 0;
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        for (int f = 1; f <= 6; ++f) {
            const int b = t->bc[6 * (g - 1) + f - 1];
            int64_t *o = &bbo[6 * (g - 1) + f - 1];
            if (b != FIX_BC && b != OP1_BC && b != PAR_BC) { *o = -1; continue; }
            *o = n;
            n += f <= 2 ? (int64_t)kk * jj : f <= 4 ? (int64_t)kk * ii : (int64_t)jj * ii;
        }
    }
    return n;
}

/* prolong1 / prolong2 (src/ib/parent/parent1_mod.F90:515-583) on a packed
 * coarse face `in` (jjc x iic) into the fine face `out` (jj x ii) through
 * `tmp` (jj x iic). */
static void prolong12(int jjc, int iic, int jj, int ii, const double *in,
                      double *tmp, double *out, int stag1, int stag2)
{
#define IN(j, i) in[((j)-1) + (int64_t)((i)-1) * jjc]
#define TM(j, i) tmp[((j)-1) + (int64_t)((i)-1) * jj]
#define OU(j, i) out[((j)-1) + (int64_t)((i)-1) * jj]
    for (int ic = 1; ic <= iic; ++ic)
        for (int j = 1; j <= jj; j += 2) {
            const int jc = 2 + (j - 1) / 2;
            TM(j, ic) = stag1 ? 0.5 * (IN(jc, ic) + IN(jc - 1, ic)) : IN(jc, ic);
            TM(j + 1, ic) = IN(jc, ic);
        }
    for (int i = 1; i <= ii; i += 2) {
        const int ic = 2 + (i - 1) / 2;
        for (int j = 1; j <= jj; ++j) {
            OU(j, i) = stag2 ? 0.5 * (TM(j, ic) + TM(j, ic - 1)) : TM(j, ic);
            OU(j, i + 1) = TM(j, ic);
        }
    }
#undef IN
#undef TM
#undef OU
}

/* parent(ilevel, u, v, w) into the PAR face buffers: the coarse slab of
 * start_and_stop with icomp (the normal component at a back face starts one
 * index lower, src/ib/parent/parent_core_mod.F90:213-267), packed k, j, i
 * fastest first (pack_single, src/ib/parent/parent1_mod.F90:296-338) and
 * prolonged with the staggering of stag (:185-210): at x faces (k, j) v is
 * staggered in the second, w in the first direction; at y faces (k, i) u in
 * the second, w in the first; at z faces (j, i) u in the second, v in the
 * first. */
void og_parent_uvw(const og_topo *t, int level, const double *u_, const double *v_,
                   const double *w_, double *ubuf, double *vbuf, double *wbuf,
                   const int64_t *bbo)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        if (t->level[g - 1] != level) continue;
        const int c = t->iparent[g - 1];
        if (c == 0) continue;
        GDIMS(t, g);
        const int ipos = POS(g, 0), jpos = POS(g, 1), kpos = POS(g, 2);
        const int64_t kkc = t->kk[c - 1], jjc = t->jj[c - 1];
        const int kkc2 = kk / 2 + 2, jjc2 = jj / 2 + 2, iic2 = ii / 2 + 2;
        const int mx = (kk > jj ? kk : jj) > ii ? (kk > jj ? kk : jj) : ii;
        const int mxc = mx / 2 + 2;
        double *in = malloc(sizeof(double) * (size_t)mxc * mxc);
        double *tmp = malloc(sizeof(double) * (size_t)mx * mxc);
        for (int f = 1; f <= 6; ++f) {
            if (t->bc[6 * (g - 1) + f - 1] != PAR_BC) continue;
            const int64_t o = bbo[6 * (g - 1) + f - 1];
            for (int comp = 1; comp <= 3; ++comp) {
                const double *fc = (comp == 1 ? u_ : comp == 2 ? v_ : w_) + t->ip3[c - 1];
                double *buf = (comp == 1 ? ubuf : comp == 2 ? vbuf : wbuf) + o;
                const int icomp = (f <= 2 && comp == 1) || (f > 2 && f <= 4 && comp == 2) ||
                                  (f > 4 && comp == 3) ? comp : 0;
                int ks = kpos - 2, js = jpos - 2, is = ipos - 2;
                switch (f) {
                case 1: is = ipos - 1; break;
                case 2: is = ipos + (ii - 4) / 2 - (icomp == 1); break;
                case 3: js = jpos - 1; break;
                case 4: js = jpos + (jj - 4) / 2 - (icomp == 2); break;
                case 5: ks = kpos - 1; break;
                default: ks = kpos + (kk - 4) / 2 - (icomp == 3); break;
                }
                int n1, n2, f1, f2, s1 = 0, s2 = 0;
                if (f <= 2) {
                    n1 = kkc2; n2 = jjc2; f1 = kk; f2 = jj;
                    s1 = comp == 3; s2 = comp == 2;
                } else if (f <= 4) {
                    n1 = kkc2; n2 = iic2; f1 = kk; f2 = ii;
                    s1 = comp == 3; s2 = comp == 1;
                } else {
                    n1 = jjc2; n2 = iic2; f1 = jj; f2 = ii;
                    s1 = comp == 2; s2 = comp == 1;
                }
#define FC(k, j, i) fc[((k)-1) + ((j)-1) * kkc + ((i)-1) * kkc * jjc]
                for (int b = 1; b <= n2; ++b)
                    for (int a = 1; a <= n1; ++a)
                        in[(a - 1) + (int64_t)(b - 1) * n1] =
                            f <= 2 ? FC(ks + a - 1, js + b - 1, is)
                            : f <= 4 ? FC(ks + a - 1, js, is + b - 1)
                                     : FC(ks, js + a - 1, is + b - 1);
#undef FC
                prolong12(n1, n2, f1, f2, in, tmp, buf, s1, s2);
            }
        }
        free(in);
        free(tmp);
    }
}

/* par_ftoc_norm(ilevel, u, v, w, sum) (src/ib/par_ftoc_mod.F90:35-60): for
 * every PAR face of the grids of `level`, the face-normal velocity on the
 * fine face plane (start_and_stop :764-810: position 2 or n-2 along the
 * normal, 3..n-2 across) packed per 2x2 group (pack_single :234-337: the
 * plain sum, or the ddx/ddy/ddz-area-weighted mean) and unpacked into the
 * coarse face positions (unpack_single :383-475) of the parent, at the
 * parent's own boundary face when the fine face lies on it, else inside
 * the parent (ipos-1 for a low face, the last covered index for a high
 * face); a fine face on the parent's boundary is also sent to the parent's
 * neighbour across it, at that grid's opposite boundary face (the
 * connection table of init_par_ftoc :513-745: pos == 3 and one_connect,
 * pos > 3).  Each coarse position is written by one message, so the
 * reference's arrival order does not matter; here grids and faces go in
 * order. */
static void par_ftoc_pack(const og_topo *t, int g, int f, const double *u_, const double *v_,
                          const double *w_, const double *ddx_, const double *ddy_,
                          const double *ddz_, int sum, double *buf)
{
    GDIMS(t, g);
    const double *ff = (f <= 2 ? u_ : f <= 4 ? v_ : w_) + t->ip3[g - 1];
    const double *ddx = ddx_ + t->ip1x[g - 1] - 1, *ddy = ddy_ + t->ip1y[g - 1] - 1,
                 *ddz = ddz_ + t->ip1z[g - 1] - 1;
    int ks = 3, ke = kk - 2, js = 3, je = jj - 2, is = 3, ie = ii - 2;
    switch (f) {
    case 1: is = ie = 2; break;
    case 2: is = ie = ii - 2; break;
    case 3: js = je = 2; break;
    case 4: js = je = jj - 2; break;
    case 5: ks = ke = 2; break;
    default: ks = ke = kk - 2; break;
    }
    int64_t n = 0;
    if (f <= 2) {
        const int i = is;
        for (int j = js; j <= je; j += 2)
            for (int k = ks; k <= ke; k += 2) {
                if (sum)
                    buf[n++] = ff[IX(k, j, i)] + ff[IX(k, j + 1, i)] + ff[IX(k + 1, j, i)] +
                               ff[IX(k + 1, j + 1, i)];
                else {
                    const double ua = ff[IX(k, j, i)] * ddy[j] * ddz[k] +
                                      ff[IX(k, j + 1, i)] * ddy[j + 1] * ddz[k] +
                                      ff[IX(k + 1, j, i)] * ddy[j] * ddz[k + 1] +
                                      ff[IX(k + 1, j + 1, i)] * ddy[j + 1] * ddz[k + 1];
                    const double a = (ddy[j] + ddy[j + 1]) * (ddz[k] + ddz[k + 1]);
                    buf[n++] = ua / a;
                }
            }
    } else if (f <= 4) {
        const int j = js;
        for (int i = is; i <= ie; i += 2)
            for (int k = ks; k <= ke; k += 2) {
                if (sum)
                    buf[n++] = ff[IX(k, j, i)] + ff[IX(k, j, i + 1)] + ff[IX(k + 1, j, i)] +
                               ff[IX(k + 1, j, i + 1)];
                else {
                    const double ua = ff[IX(k, j, i)] * ddx[i] * ddz[k] +
                                      ff[IX(k, j, i + 1)] * ddx[i + 1] * ddz[k] +
                                      ff[IX(k + 1, j, i)] * ddx[i] * ddz[k + 1] +
                                      ff[IX(k + 1, j, i + 1)] * ddx[i + 1] * ddz[k + 1];
                    const double a = (ddx[i] + ddx[i + 1]) * (ddz[k] + ddz[k + 1]);
                    buf[n++] = ua / a;
                }
            }
    } else {
        const int k = ks;
        for (int i = is; i <= ie; i += 2)
            for (int j = js; j <= je; j += 2) {
                if (sum)
                    buf[n++] = ff[IX(k, j, i)] + ff[IX(k, j + 1, i)] + ff[IX(k, j, i + 1)] +
                               ff[IX(k, j + 1, i + 1)];
                else {
                    const double ua = ff[IX(k, j, i)] * ddx[i] * ddy[j] +
                                      ff[IX(k, j + 1, i)] * ddx[i] * ddy[j + 1] +
                                      ff[IX(k, j, i + 1)] * ddx[i + 1] * ddy[j] +
                                      ff[IX(k, j + 1, i + 1)] * ddx[i + 1] * ddy[j + 1];
                    const double a = (ddx[i] + ddx[i + 1]) * (ddy[j] + ddy[j + 1]);
                    buf[n++] = ua / a;
                }
            }
    }
}

static void par_ftoc_unpack(const og_topo *t, int gf, int gc, int f, int frecv, double *u_,
                            double *v_, double *w_, const double *buf)
{
    const int iif = t->ii[gf - 1], jjf = t->jj[gf - 1], kkf = t->kk[gf - 1];
    const int ipos = POS(gf, 0), jpos = POS(gf, 1), kpos = POS(gf, 2);
    int is = ipos, ie = ipos + (iif - 4) / 2 - 1;
    int js = jpos, je = jpos + (jjf - 4) / 2 - 1;
    int ks = kpos, ke = kpos + (kkf - 4) / 2 - 1;
    if (frecv < 0) {
        switch (f) {
        case 1: is = ie = ipos - 1; break;
        case 2: is = ie; break;
        case 3: js = je = jpos - 1; break;
        case 4: js = je; break;
        case 5: ks = ke = kpos - 1; break;
        default: ks = ke; break;
        }
    } else {
        const int iic = t->ii[gc - 1], jjc = t->jj[gc - 1], kkc = t->kk[gc - 1];
        switch (frecv) {
        case 1: is = ie = 2; break;
        case 2: is = ie = iic - 2; break;
        case 3: js = je = 2; break;
        case 4: js = je = jjc - 2; break;
        case 5: ks = ke = 2; break;
        default: ks = ke = kkc - 2; break;
        }
    }
    double *fc = (f <= 2 ? u_ : f <= 4 ? v_ : w_) + t->ip3[gc - 1];
    const int64_t kk = t->kk[gc - 1], jj = t->jj[gc - 1];
    int64_t n = 0;
    for (int i = is; i <= ie; ++i)
        for (int j = js; j <= je; ++j)
            for (int k = ks; k <= ke; ++k) fc[IX(k, j, i)] = buf[n++];
}

void og_par_ftoc_norm(const og_topo *t, int level, double *u, double *v, double *w,
                      const double *ddx, const double *ddy, const double *ddz, int sum)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        if (t->level[g - 1] != level) continue;
        for (int f = 1; f <= 6; ++f) {
            if (t->bc[6 * (g - 1) + f - 1] != PAR_BC) continue;
            const int c = t->iparent[g - 1];
            const int ax = (f - 1) / 2, odd = (f & 1) != 0;
            const int pos = POS(g, ax);
            const int nf = ax == 0 ? t->ii[g - 1] : ax == 1 ? t->jj[g - 1] : t->kk[g - 1];
            const int nc = ax == 0 ? t->ii[c - 1] : ax == 1 ? t->jj[c - 1] : t->kk[c - 1];
            const int one_connect = !(nf > nc);
            int frecv = f;
            if (pos == 3 && one_connect) {
                if (!odd) frecv = -1;
            } else if (pos > 3) {
                if (odd) frecv = -1;
            }
            const int64_t area = (int64_t)((t->ii[g - 1] - 4) / 2 + 2) *
                                 ((t->jj[g - 1] - 4) / 2 + 2) * ((t->kk[g - 1] - 4) / 2 + 2);
            double *buf = malloc(sizeof(double) * (size_t)area);
            par_ftoc_pack(t, g, f, u, v, w, ddx, ddy, ddz, sum, buf);
            par_ftoc_unpack(t, g, c, f, frecv, u, v, w, buf);
            if (frecv > 0) {
                /* the parent's neighbour across face f, receiving at its
                 * opposite face */
                const int nb = t->nbr[26 * (c - 1) + f - 1];
                if (nb > 0) par_ftoc_unpack(t, g, nb, f, odd ? f + 1 : f - 1, u, v, w, buf);
            }
            free(buf);
        }
    }
}


#define BPV(k, j, i) (bp ? bp[IX(k, j, i)] : 1.0)

/* bfront (src/flow/bound_flow_mod.F90:23-212): faces 1 and 2 */
static void bfront(int f, int ctyp, int kk, int jj, int ii, double *u, double *v,
                   double *w, double *p, const double *ub, const double *vb,
                   const double *wb, double pinf, const double *bp, const double *ddy,
                   const double *ddz, double rho)
{
    const int i2 = f == 1 ? 2 : ii - 1, i3 = f == 1 ? 3 : ii - 2, i4 = f == 1 ? 4 : ii - 3;
    const int istag1 = f == 1 ? 1 : ii - 1, istag2 = f == 1 ? 2 : ii - 2;
    const double dir = f == 1 ? -1.0 : 1.0;
#define BF(b, k, j) b[((k)-1) + (int64_t)((j)-1) * kk]
#define DY(j) ddy[(j)-1]
#define DZ(k) ddz[(k)-1]
    switch (ctyp) {
    case FIX_BC:
        for (int j = 1; j <= jj; ++j)
            for (int k = 1; k <= kk; ++k) {
                u[IX(k, j, istag2)] = BF(ub, k, j);
                v[IX(k, j, i2)] = 2.0 * BF(vb, k, j) - v[IX(k, j, i3)];
                w[IX(k, j, i2)] = 2.0 * BF(wb, k, j) - w[IX(k, j, i3)];
            }
        break;
    case OP1_BC:
        for (int j = 1; j <= jj - 1; ++j)
            for (int k = 1; k <= kk - 1; ++k) {
                u[IX(k, j, istag1)] = u[IX(k, j, istag2)];
                double flag = copysign(1.0, dir * (u[IX(k, j, istag2)] + u[IX(k, j + 1, istag2)]));
                flag = 0.5 * (flag + 1.0);
                v[IX(k, j, i2)] = flag * v[IX(k, j, i3)] +
                                  (1.0 - flag) * (2.0 * BF(vb, k, j) - v[IX(k, j, i3)]);
                flag = copysign(1.0, dir * (u[IX(k, j, istag2)] + u[IX(k + 1, j, istag2)]));
                flag = 0.5 * (flag + 1.0);
                w[IX(k, j, i2)] = flag * w[IX(k, j, i3)] +
                                  (1.0 - flag) * (2.0 * BF(wb, k, j) - w[IX(k, j, i3)]);
            }
        break;
    case NOS_BC:
        for (int j = 1; j <= jj; ++j)
            for (int k = 1; k <= kk; ++k) {
                u[IX(k, j, istag2)] = 0.0;
                v[IX(k, j, i2)] = 0.0;
                w[IX(k, j, i2)] = 0.0;
            }
        break;
    case SLI_BC:
        for (int j = 1; j <= jj; ++j)
            for (int k = 1; k <= kk; ++k) {
                u[IX(k, j, istag2)] = 0.0;
                v[IX(k, j, i2)] = v[IX(k, j, i3)];
                w[IX(k, j, i2)] = w[IX(k, j, i3)];
            }
        break;
    case PAR_BC:
        for (int j = 2; j <= jj - 1; ++j)
            for (int k = 2; k <= kk - 1; ++k) {
                const double sbv = BPV(k, j, i2) * BPV(k, j + 1, i2);
                const double sbw = BPV(k, j, i2) * BPV(k + 1, j, i2);
                v[IX(k, j, i2)] = (BF(vb, k, j) - 0.5 * (v[IX(k, j, i3)] - v[IX(k, j, i4)])) * sbv +
                                  (1.0 - sbv) * v[IX(k, j, i2)];
                w[IX(k, j, i2)] = (BF(wb, k, j) - 0.5 * (w[IX(k, j, i3)] - w[IX(k, j, i4)])) * sbw +
                                  (1.0 - sbw) * w[IX(k, j, i2)];
            }
        for (int j = 3; j <= jj - 2; j += 2)
            for (int k = 3; k <= kk - 2; k += 2) {
                const double sbu = BPV(k, j, i2);
                double sb11, sb12, sb13, sb14;
                if (sbu > 0.5) {
                    sb11 = BPV(k, j, i2) * BPV(k, j, i3);
                    sb12 = BPV(k, j + 1, i2) * BPV(k, j + 1, i3);
                    sb13 = BPV(k + 1, j, i2) * BPV(k + 1, j, i3);
                    sb14 = BPV(k + 1, j + 1, i2) * BPV(k + 1, j + 1, i3);
                } else {
                    sb11 = fmin(BPV(k, j, i3) + BPV(k, j, i2) + BPV(k, j, i4) + BPV(k, j - 1, i3) +
                                BPV(k, j + 1, i3) + BPV(k - 1, j, i3) + BPV(k + 1, j, i3), 1.0);
                    sb12 = fmin(BPV(k, j + 1, i3) + BPV(k, j + 1, i2) + BPV(k, j + 1, i4) +
                                BPV(k, j, i3) + BPV(k, j + 2, i3) + BPV(k - 1, j + 1, i3) +
                                BPV(k + 1, j + 1, i3), 1.0);
                    sb13 = fmin(BPV(k + 1, j, i3) + BPV(k + 1, j, i2) + BPV(k + 1, j, i4) +
                                BPV(k + 1, j - 1, i3) + BPV(k + 1, j + 1, i3) + BPV(k, j, i3) +
                                BPV(k + 2, j, i3), 1.0);
                    sb14 = fmin(BPV(k + 1, j + 1, i3) + BPV(k + 1, j + 1, i2) + BPV(k + 1, j + 1, i4) +
                                BPV(k + 1, j, i3) + BPV(k + 1, j + 2, i3) + BPV(k, j + 1, i3) +
                                BPV(k + 2, j + 1, i3), 1.0);
                }
                double fak = 1.0 / ((DY(j) + DY(j + 1)) * (DZ(k) + DZ(k + 1))) *
                             (sb11 * DY(j) * DZ(k) + sb12 * DY(j + 1) * DZ(k) +
                              sb13 * DY(j) * DZ(k + 1) + sb14 * DY(j + 1) * DZ(k + 1));
                if (fak < 0.1) fak = 1.0;
                fak = 1.0 / fak;
                const double u11 = u[IX(k, j, istag2)], u12 = u[IX(k, j + 1, istag2)];
                const double u13 = u[IX(k + 1, j, istag2)], u14 = u[IX(k + 1, j + 1, istag2)];
                const double ubfine = (u11 * (1.0 - sb11) * DY(j) * DZ(k) +
                                       u12 * (1.0 - sb12) * DY(j + 1) * DZ(k) +
                                       u13 * (1.0 - sb13) * DY(j) * DZ(k + 1) +
                                       u14 * (1.0 - sb14) * DY(j + 1) * DZ(k + 1)) /
                                      ((DY(j) + DY(j + 1)) * (DZ(k) + DZ(k + 1)));
                u[IX(k, j, istag2)] = (BF(ub, k, j) - ubfine) * fak * sb11 + (1.0 - sb11) * u11 * sbu;
                u[IX(k, j + 1, istag2)] = (BF(ub, k, j + 1) - ubfine) * fak * sb12 + (1.0 - sb12) * u12 * sbu;
                u[IX(k + 1, j, istag2)] = (BF(ub, k + 1, j) - ubfine) * fak * sb13 + (1.0 - sb13) * u13 * sbu;
                u[IX(k + 1, j + 1, istag2)] = (BF(ub, k + 1, j + 1) - ubfine) * fak * sb14 +
                                              (1.0 - sb14) * u14 * sbu;
            }
        break;
    }
    if (ctyp == OP1_BC) {
        for (int j = 2; j <= jj; ++j)
            for (int k = 2; k <= kk; ++k) {
                const double vi = 0.25 * (v[IX(k, j - 1, i2)] + v[IX(k, j, i2)] +
                                          v[IX(k, j - 1, i3)] + v[IX(k, j, i3)]);
                const double wi = 0.25 * (w[IX(k - 1, j, i2)] + w[IX(k, j, i2)] +
                                          w[IX(k - 1, j, i3)] + w[IX(k, j, i3)]);
                const double us = u[IX(k, j, istag2)];
                const double umagsqr = us * us + vi * vi + wi * wi;
                p[IX(k, j, i2)] = pinf + fmin(0.0, 0.5 * rho * copysign(umagsqr, dir * us));
            }
    } else if (ctyp == FIX_BC || ctyp == NOS_BC || ctyp == SLI_BC) {
        for (int j = 1; j <= jj; ++j)
            for (int k = 1; k <= kk; ++k) p[IX(k, j, i2)] = p[IX(k, j, i3)];
    }
#undef BF
#undef DY
#undef DZ
}

/* bright (src/flow/bound_flow_mod.F90:215-494): faces 3 and 4.  The
 * tangential w of a PAR face is extrapolated from w(j2), not w(j3) as in
 * bfront and bbottom: restated as written. */
static void bright(int f, int ctyp, int kk, int jj, int ii, double *u, double *v,
                   double *w, double *p, const double *ub, const double *vb,
                   const double *wb, double pinf, const double *bp, const double *ddx,
                   const double *ddz, double rho)
{
    const int j2 = f == 3 ? 2 : jj - 1, j3 = f == 3 ? 3 : jj - 2, j4 = f == 3 ? 4 : jj - 3;
    const int jstag1 = f == 3 ? 1 : jj - 1, jstag2 = f == 3 ? 2 : jj - 2;
    const double dir = f == 3 ? -1.0 : 1.0;
#define BF(b, k, i) b[((k)-1) + (int64_t)((i)-1) * kk]
#define DX(i) ddx[(i)-1]
#define DZ(k) ddz[(k)-1]
    switch (ctyp) {
    case FIX_BC:
        for (int i = 1; i <= ii; ++i)
            for (int k = 1; k <= kk; ++k) {
                u[IX(k, j2, i)] = 2.0 * BF(ub, k, i) - u[IX(k, j3, i)];
                v[IX(k, jstag2, i)] = BF(vb, k, i);
                w[IX(k, j2, i)] = 2.0 * BF(wb, k, i) - w[IX(k, j3, i)];
            }
        break;
    case OP1_BC:
        for (int i = 1; i <= ii - 1; ++i)
            for (int k = 1; k <= kk - 1; ++k) {
                double flag = copysign(1.0, dir * (v[IX(k, jstag2, i)] + v[IX(k, jstag2, i + 1)]));
                flag = 0.5 * (flag + 1.0);
                u[IX(k, j2, i)] = flag * u[IX(k, j3, i)] +
                                  (1.0 - flag) * (2.0 * BF(ub, k, i) - u[IX(k, j3, i)]);
                v[IX(k, jstag1, i)] = v[IX(k, jstag2, i)];
                flag = copysign(1.0, dir * (v[IX(k, jstag2, i)] + v[IX(k + 1, jstag2, i)]));
                flag = 0.5 * (flag + 1.0);
                w[IX(k, j2, i)] = flag * w[IX(k, j3, i)] +
                                  (1.0 - flag) * (2.0 * BF(wb, k, i) - w[IX(k, j3, i)]);
            }
        break;
    case NOS_BC:
        for (int i = 1; i <= ii; ++i)
            for (int k = 1; k <= kk; ++k) {
                u[IX(k, j2, i)] = 0.0;
                v[IX(k, jstag2, i)] = 0.0;
                w[IX(k, j2, i)] = 0.0;
            }
        break;
    case SLI_BC:
        for (int i = 1; i <= ii; ++i)
            for (int k = 1; k <= kk; ++k) {
                u[IX(k, j2, i)] = u[IX(k, j3, i)];
                v[IX(k, jstag2, i)] = 0.0;
                w[IX(k, j2, i)] = w[IX(k, j3, i)];
            }
        break;
    case PAR_BC:
        for (int i = 2; i <= ii - 1; ++i)
            for (int k = 2; k <= kk - 1; ++k) {
                const double sbu = BPV(k, j2, i) * BPV(k, j2, i + 1);
                const double sbw = BPV(k, j2, i) * BPV(k + 1, j2, i);
                u[IX(k, j2, i)] = (BF(ub, k, i) - 0.5 * (u[IX(k, j3, i)] - u[IX(k, j4, i)])) * sbu +
                                  (1.0 - sbu) * u[IX(k, j2, i)];
                w[IX(k, j2, i)] = (BF(wb, k, i) - 0.5 * (w[IX(k, j2, i)] - w[IX(k, j4, i)])) * sbw +
                                  (1.0 - sbw) * w[IX(k, j2, i)];
            }
        for (int i = 3; i <= ii - 2; i += 2)
            for (int k = 3; k <= kk - 2; k += 2) {
                const double sbv = BPV(k, j2, i);
                double sb11, sb12, sb13, sb14;
                if (sbv > 0.5) {
                    sb11 = BPV(k, j2, i) * BPV(k, j3, i);
                    sb12 = BPV(k, j2, i + 1) * BPV(k, j3, i + 1);
                    sb13 = BPV(k + 1, j2, i) * BPV(k + 1, j3, i);
                    sb14 = BPV(k + 1, j2, i + 1) * BPV(k + 1, j3, i + 1);
                } else {
                    sb11 = fmin(BPV(k, j3, i) + BPV(k, j2, i) + BPV(k, j4, i) + BPV(k, j3, i - 1) +
                                BPV(k, j3, i + 1) + BPV(k - 1, j3, i) + BPV(k + 1, j3, i), 1.0);
                    sb12 = fmin(BPV(k, j3, i + 1) + BPV(k, j2, i + 1) + BPV(k, j4, i + 1) +
                                BPV(k, j3, i) + BPV(k, j3, i + 2) + BPV(k - 1, j3, i + 1) +
                                BPV(k + 1, j3, i + 1), 1.0);
                    sb13 = fmin(BPV(k + 1, j3, i) + BPV(k + 1, j2, i) + BPV(k + 1, j4, i) +
                                BPV(k + 1, j3, i - 1) + BPV(k + 1, j3, i + 1) + BPV(k, j3, i) +
                                BPV(k + 2, j3, i), 1.0);
                    sb14 = fmin(BPV(k + 1, j3, i + 1) + BPV(k + 1, j2, i + 1) + BPV(k + 1, j4, i + 1) +
                                BPV(k + 1, j3, i) + BPV(k + 1, j3, i + 2) + BPV(k, j3, i + 1) +
                                BPV(k + 2, j3, i + 1), 1.0);
                }
                double fak = 1.0 / ((DZ(k) + DZ(k + 1)) * (DX(i) + DX(i + 1))) *
                             (sb11 * DZ(k) * DX(i) + sb12 * DZ(k) * DX(i + 1) +
                              sb13 * DZ(k + 1) * DX(i) + sb14 * DZ(k + 1) * DX(i + 1));
                if (fak < 0.1) fak = 1.0;
                fak = 1.0 / fak;
                const double v11 = v[IX(k, jstag2, i)], v12 = v[IX(k, jstag2, i + 1)];
                const double v13 = v[IX(k + 1, jstag2, i)], v14 = v[IX(k + 1, jstag2, i + 1)];
                const double vbfine = (v11 * (1.0 - sb11) * DZ(k) * DX(i) +
                                       v12 * (1.0 - sb12) * DZ(k) * DX(i + 1) +
                                       v13 * (1.0 - sb13) * DZ(k + 1) * DX(i) +
                                       v14 * (1.0 - sb14) * DZ(k + 1) * DX(i + 1)) /
                                      ((DZ(k) + DZ(k + 1)) * (DX(i) + DX(i + 1)));
                v[IX(k, jstag2, i)] = (BF(vb, k, i) - vbfine) * fak * sb11 + v11 * (1.0 - sb11) * sbv;
                v[IX(k, jstag2, i + 1)] = (BF(vb, k, i + 1) - vbfine) * fak * sb12 + v12 * (1.0 - sb12) * sbv;
                v[IX(k + 1, jstag2, i)] = (BF(vb, k + 1, i) - vbfine) * fak * sb13 + v13 * (1.0 - sb13) * sbv;
                v[IX(k + 1, jstag2, i + 1)] = (BF(vb, k + 1, i + 1) - vbfine) * fak * sb14 +
                                              v14 * (1.0 - sb14) * sbv;
            }
        break;
    }
    if (ctyp == OP1_BC) {
        for (int i = 2; i <= ii; ++i)
            for (int k = 2; k <= kk; ++k) {
                const double ui = 0.25 * (u[IX(k, j2, i - 1)] + u[IX(k, j2, i)] +
                                          u[IX(k, j3, i - 1)] + u[IX(k, j3, i)]);
                const double wi = 0.25 * (w[IX(k - 1, j2, i)] + w[IX(k, j2, i)] +
                                          w[IX(k - 1, j3, i)] + w[IX(k, j3, i)]);
                const double vs = v[IX(k, jstag2, i)];
                const double umagsqr = ui * ui + vs * vs + wi * wi;
                p[IX(k, j2, i)] = pinf + fmin(0.0, 0.5 * rho * copysign(umagsqr, dir * vs));
            }
    } else if (ctyp == FIX_BC || ctyp == NOS_BC || ctyp == SLI_BC) {
        for (int i = 1; i <= ii; ++i)
            for (int k = 1; k <= kk; ++k) p[IX(k, j2, i)] = p[IX(k, j3, i)];
    }
#undef BF
#undef DX
#undef DZ
}

/* bbottom (src/flow/bound_flow_mod.F90:497-731): faces 5 and 6 */
static void bbottom(int f, int ctyp, int kk, int jj, int ii, double *u, double *v,
                    double *w, double *p, const double *ub, const double *vb,
                    const double *wb, double pinf, const double *bp, const double *ddx,
                    const double *ddy, double rho)
{
    const int k2 = f == 5 ? 2 : kk - 1, k3 = f == 5 ? 3 : kk - 2, k4 = f == 5 ? 4 : kk - 3;
    const int kstag1 = f == 5 ? 1 : kk - 1, kstag2 = f == 5 ? 2 : kk - 2;
    const double dir = f == 5 ? -1.0 : 1.0;
#define BF(b, j, i) b[((j)-1) + (int64_t)((i)-1) * jj]
#define DX(i) ddx[(i)-1]
#define DY(j) ddy[(j)-1]
    switch (ctyp) {
    case FIX_BC:
        for (int i = 1; i <= ii; ++i)
            for (int j = 1; j <= jj; ++j) {
                u[IX(k2, j, i)] = 2.0 * BF(ub, j, i) - u[IX(k3, j, i)];
                v[IX(k2, j, i)] = 2.0 * BF(vb, j, i) - v[IX(k3, j, i)];
                w[IX(kstag2, j, i)] = BF(wb, j, i);
            }
        break;
    case OP1_BC:
        for (int i = 1; i <= ii - 1; ++i)
            for (int j = 1; j <= jj - 1; ++j) {
                double flag = copysign(1.0, dir * (w[IX(kstag2, j, i)] + w[IX(kstag2, j, i + 1)]));
                flag = 0.5 * (flag + 1.0);
                u[IX(k2, j, i)] = flag * u[IX(k3, j, i)] +
                                  (1.0 - flag) * (2 * BF(ub, j, i) - u[IX(k3, j, i)]);
                flag = copysign(1.0, dir * (w[IX(kstag2, j, i)] + w[IX(kstag2, j + 1, i)]));
                flag = 0.5 * (flag + 1.0);
                v[IX(k2, j, i)] = flag * v[IX(k3, j, i)] +
                                  (1.0 - flag) * (2.0 * BF(vb, j, i) - v[IX(k3, j, i)]);
                w[IX(kstag1, j, i)] = w[IX(kstag2, j, i)];
            }
        break;
    case NOS_BC:
        for (int i = 1; i <= ii; ++i)
            for (int j = 1; j <= jj; ++j) {
                u[IX(k2, j, i)] = 0.0;
                v[IX(k2, j, i)] = 0.0;
                w[IX(kstag2, j, i)] = 0.0;
            }
        break;
    case SLI_BC:
        for (int i = 1; i <= ii; ++i)
            for (int j = 1; j <= jj; ++j) {
                u[IX(k2, j, i)] = u[IX(k3, j, i)];
                v[IX(k2, j, i)] = v[IX(k3, j, i)];
                w[IX(kstag2, j, i)] = 0.0;
            }
        break;
    case PAR_BC:
        for (int i = 2; i <= ii - 1; ++i)
            for (int j = 2; j <= jj - 1; ++j) {
                const double sbu = BPV(k2, j, i) * BPV(k2, j, i + 1);
                const double sbv = BPV(k2, j, i) * BPV(k2, j + 1, i);
                u[IX(k2, j, i)] = (BF(ub, j, i) - 0.5 * (u[IX(k3, j, i)] - u[IX(k4, j, i)])) * sbu +
                                  (1.0 - sbu) * u[IX(k2, j, i)];
                v[IX(k2, j, i)] = (BF(vb, j, i) - 0.5 * (v[IX(k3, j, i)] - v[IX(k4, j, i)])) * sbv +
                                  (1.0 - sbv) * v[IX(k2, j, i)];
            }
        for (int i = 3; i <= ii - 2; i += 2)
            for (int j = 3; j <= jj - 2; j += 2) {
                const double sbw = BPV(k2, j, i);
                double sb11, sb12, sb13, sb14;
                if (sbw > 0.5) {
                    sb11 = BPV(k2, j, i) * BPV(k3, j, i);
                    sb12 = BPV(k2, j, i + 1) * BPV(k3, j, i + 1);
                    sb13 = BPV(k2, j + 1, i) * BPV(k3, j + 1, i);
                    sb14 = BPV(k2, j + 1, i + 1) * BPV(k3, j + 1, i + 1);
                } else {
                    sb11 = fmin(BPV(k3, j, i) + BPV(k2, j, i) + BPV(k4, j, i) + BPV(k3, j, i - 1) +
                                BPV(k3, j, i + 1) + BPV(k3, j - 1, i) + BPV(k3, j + 1, i), 1.0);
                    sb12 = fmin(BPV(k3, j, i + 1) + BPV(k2, j, i + 1) + BPV(k4, j, i + 1) +
                                BPV(k3, j, i) + BPV(k3, j, i + 2) + BPV(k3, j - 1, i + 1) +
                                BPV(k3, j + 1, i + 1), 1.0);
                    sb13 = fmin(BPV(k3, j + 1, i) + BPV(k2, j + 1, i) + BPV(k4, j + 1, i) +
                                BPV(k3, j + 1, i - 1) + BPV(k3, j + 1, i + 1) + BPV(k3, j, i) +
                                BPV(k3, j + 2, i), 1.0);
                    sb14 = fmin(BPV(k3, j + 1, i + 1) + BPV(k2, j + 1, i + 1) + BPV(k4, j + 1, i + 1) +
                                BPV(k3, j + 1, i) + BPV(k3, j + 1, i + 2) + BPV(k3, j, i + 1) +
                                BPV(k3, j + 2, i + 1), 1.0);
                }
                double fak = 1. / ((DY(j) + DY(j + 1)) * (DX(i) + DX(i + 1))) *
                             (sb11 * DY(j) * DX(i) + sb12 * DY(j) * DX(i + 1) +
                              sb13 * DY(j + 1) * DX(i) + sb14 * DY(j + 1) * DX(i + 1));
                if (fak < 0.1) fak = 1.0;
                fak = 1.0 / fak;
                const double w11 = w[IX(kstag2, j, i)], w12 = w[IX(kstag2, j, i + 1)];
                const double w13 = w[IX(kstag2, j + 1, i)], w14 = w[IX(kstag2, j + 1, i + 1)];
                const double wbfine = (w11 * (1.0 - sb11) * DY(j) * DX(i) +
                                       w12 * (1.0 - sb12) * DY(j) * DX(i + 1) +
                                       w13 * (1.0 - sb13) * DY(j + 1) * DX(i) +
                                       w14 * (1.0 - sb14) * DY(j + 1) * DX(i + 1)) /
                                      ((DY(j) + DY(j + 1)) * (DX(i) + DX(i + 1)));
                w[IX(kstag2, j, i)] = (BF(wb, j, i) - wbfine) * fak * sb11 + w11 * (1.0 - sb11) * sbw;
                w[IX(kstag2, j, i + 1)] = (BF(wb, j, i + 1) - wbfine) * fak * sb12 + w12 * (1.0 - sb12) * sbw;
                w[IX(kstag2, j + 1, i)] = (BF(wb, j + 1, i) - wbfine) * fak * sb13 + w13 * (1.0 - sb13) * sbw;
                w[IX(kstag2, j + 1, i + 1)] = (BF(wb, j + 1, i + 1) - wbfine) * fak * sb14 +
                                              w14 * (1.0 - sb14) * sbw;
            }
        break;
    }
    if (ctyp == OP1_BC) {
        for (int i = 2; i <= ii; ++i)
            for (int j = 2; j <= jj; ++j) {
                const double ui = 0.25 * (u[IX(k2, j, i - 1)] + u[IX(k2, j, i)] +
                                          u[IX(k3, j, i - 1)] + u[IX(k3, j, i)]);
                const double vi = 0.25 * (v[IX(k2, j - 1, i)] + v[IX(k2, j, i)] +
                                          v[IX(k3, j - 1, i)] + v[IX(k3, j, i)]);
                const double ws = w[IX(kstag2, j, i)];
                const double umagsqr = ui * ui + vi * vi + ws * ws;
                p[IX(k2, j, i)] = pinf + fmin(0.0, 0.5 * rho * copysign(umagsqr, dir * ws));
            }
    } else if (ctyp == FIX_BC || ctyp == NOS_BC || ctyp == SLI_BC) {
        for (int i = 1; i <= ii; ++i)
            for (int j = 1; j <= jj; ++j) p[IX(k2, j, i)] = p[IX(k3, j, i)];
    }
#undef BF
#undef DX
#undef DY
}

/* bound_flow%bound(ilevel, u, v, w, p): src/core/bound_mod.F90:36-76 (the
 * faces of each grid in order 1..6; the first boundary condition of each
 * face).  pinf[6*ngrid]: the OP1 pressure parameter (NULL = 0). */
void og_bound_flow(const og_topo *t, int level, double *u_, double *v_, double *w_,
                   double *p_, const double *ubuf, const double *vbuf, const double *wbuf,
                   const int64_t *bbo, const double *pinf, const double *bp_,
                   const double *ddx_, const double *ddy_, const double *ddz_, double rho)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        if (level >= 0 && t->level[g - 1] != level) continue;
        GDIMS(t, g);
        const int64_t o3 = t->ip3[g - 1];
        double *u = u_ + o3, *v = v_ + o3, *w = w_ + o3, *p = p_ + o3;
        const double *bp = bp_ ? bp_ + o3 : NULL;
        const double *ddx = ddx_ + t->ip1x[g - 1], *ddy = ddy_ + t->ip1y[g - 1],
                     *ddz = ddz_ + t->ip1z[g - 1];
        for (int f = 1; f <= 6; ++f) {
            const int ctyp = t->bc[6 * (g - 1) + f - 1];
            if (ctyp != FIX_BC && ctyp != OP1_BC && ctyp != NOS_BC && ctyp != SLI_BC &&
                ctyp != PAR_BC)
                continue;
            const int64_t o = bbo[6 * (g - 1) + f - 1];
            const double *ub = o >= 0 ? ubuf + o : NULL, *vb = o >= 0 ? vbuf + o : NULL,
                         *wb = o >= 0 ? wbuf + o : NULL;
            const double pf = pinf ? pinf[6 * (g - 1) + f - 1] : 0.0;
            if (f <= 2)
                bfront(f, ctyp, kk, jj, ii, u, v, w, p, ub, vb, wb, pf, bp, ddy, ddz, rho);
            else if (f <= 4)
                bright(f, ctyp, kk, jj, ii, u, v, w, p, ub, vb, wb, pf, bp, ddx, ddz, rho);
            else
                bbottom(f, ctyp, kk, jj, ii, u, v, w, p, ub, vb, wb, pf, bp, ddx, ddy, rho);
        }
    }
}
#undef BPV

static int64_t field_len(const og_topo *t)
{
    const int g = t->ngrid;
    return t->ip3[g - 1] + (int64_t)t->kk[g - 1] * t->jj[g - 1] * t->ii[g - 1];
}

/* Everything the multi-level steps need besides the SIP operands. */
typedef struct ml_ctx {
    const int64_t *ipbb;
    double *pbuf;
    const double *bp;   /* NULL: bound_pressure A forms (no IB)          */
    const double *dx, *dy, *dz, *ddx, *ddy, *ddz;
} ml_ctx;

/* src/flow/pressuresolver_mod.F90:423-465 (sip) and :361-420 (mgpoisit) */
static void mgpoisit(const og_topo *t, const og_conn *c, int nconn,
                     const og_params *prm, int ilevel, double *dp,
                     const double *rhs, double *res, const double *bp,
                     const double *aw, const double *ae, const double *as,
                     const double *an, const double *ab, const double *at,
                     const double *ap, const double *lw, const double *ls,
                     const double *lb, const double *lpr, const double *ue,
                     const double *un, const double *ut, const int64_t *mip,
                     const int64_t *idx, const double *rap, const ml_ctx *ml)
{
    for (int iloop = 1; iloop <= prm->ninner; ++iloop) {
        og_bound_pressure(t, ilevel, dp, ml->pbuf, ml->ipbb, ml->bp, ml->dx,
                          ml->dy, ml->dz, ml->ddx, ml->ddy, ml->ddz);
        const int use_sor = prm->ityp == 1 &&
                            (ilevel > t->minlevel || prm->sor_all_levels);
        if (use_sor) {
            og_relax(t, ilevel, dp, rhs, aw, ae, as, an, ab, at, rap,
                     prm->ibtype ? bp : NULL, prm->omega);
        } else {
            og_laplacephi(t, ilevel, res, dp, aw, ae, as, an, ab, at, ap, bp);
            if (prm->ityp == 2)
                og_sipiter1_cl(t, ilevel, rhs, res, lw, ls, lb, lpr);
            else
                og_sipiter1_hp(t, ilevel, rhs, res, lw, ls, lb, lpr, mip, idx);
            double *fr[1] = {res};
            og_exchange(t, c, nconn, ilevel, 1, 0,
                    iloop < prm->ninner ? 0 : -1, fr, 1);
            if (prm->ityp == 2)
                og_sipiter2_cl(t, ilevel, dp, res, ue, un, ut);
            else
                og_sipiter2_hp(t, ilevel, dp, res, ue, un, ut, mip, idx);
        }
        double *fd[1] = {dp};
        og_exchange(t, c, nconn, ilevel, 1, 0, 0, fd, 1);
    }
    og_bound_pressure(t, ilevel, dp, ml->pbuf, ml->ipbb, ml->bp, ml->dx,
                      ml->dy, ml->dz, ml->ddx, ml->ddy, ml->ddz);
}

/* src/flow/pressuresolver_mod.F90:160-357 */
int og_mgpoisl(const og_topo *t, const og_conn *c, int nconn,
               const og_params *prm, double *u, double *v, double *w,
               double *p, double dt, const double *bp,
               const double *aw, const double *ae, const double *as,
               const double *an, const double *ab, const double *at,
               const double *ap, const double *lw, const double *ls,
               const double *lb, const double *lpr, const double *ue,
               const double *un, const double *ut, const int64_t *mip,
               const int64_t *idx, const double *rap,
               const double *rdx, const double *rdy, const double *rdz,
               const double *rddx, const double *rddy, const double *rddz,
               const double *dx, const double *dy, const double *dz,
               const double *ddx, const double *ddy, const double *ddz,
               double *dp, double *hilf, double *rhs, double *res,
               og_stats *st)
{
    const int64_t n = field_len(t);
    memset(dp, 0, (size_t)n * sizeof(double));
    memset(hilf, 0, (size_t)n * sizeof(double));
    memset(rhs, 0, (size_t)n * sizeof(double));
    memset(res, 0, (size_t)n * sizeof(double));

    int64_t *ipbb = malloc(sizeof(int64_t) * 6 * (size_t)t->ngrid);
    const int64_t nbuf = og_par_buffers(t, ipbb);
    /* face buffers of HILF and of DP (fields carry their own buffers) */
    double *pbh = calloc((size_t)(nbuf > 0 ? nbuf : 1), sizeof(double));
    double *pbd = calloc((size_t)(nbuf > 0 ? nbuf : 1), sizeof(double));
    ml_ctx mh = {ipbb, pbh, prm->ibtype ? bp : NULL, dx, dy, dz, ddx, ddy, ddz};
    ml_ctx md = mh;
    md.pbuf = pbd;

    const double prefak = prm->rho / dt;
    if (prm->ibtype == 0) {
        og_divcal(t, rhs, u, v, w, rddx, rddy, rddz, prefak, NULL, NULL);
    } else {
        /* ghostcell divcal (src/ib/gc_mod.F90:278-347) adds fak*SDIV (the
         * caller's geometry source, zero without STL cut cells), then
         * multiplies by BP */
        double *zero = prm->sdiv ? NULL : calloc((size_t)n, sizeof(double));
        og_divcal(t, rhs, u, v, w, rddx, rddy, rddz, prefak, bp,
                  prm->sdiv ? prm->sdiv : zero);
        free(zero);
    }
    const double *bpr = prm->ibtype ? bp : NULL;   /* gc_restrict with the IB */
    for (int l = t->maxlevel; l >= t->minlevel; --l)
        og_ftoc_s(t, l, rhs, rhs, ddx, ddy, ddz, bpr, 'R');
    st->maxrhs0 = og_maxabscal(t, rhs, NULL);

    int ipc = 0, ipcount;
    double maxrhs = st->maxrhs0;
    for (ipcount = 1; ipcount <= prm->nouter; ++ipcount) {
        for (int l = t->minlevel; l <= t->maxlevel; ++l) {
            og_ctof(t, l, hilf, hilf);
            og_parent(t, l, hilf, pbh, ipbb);
            mgpoisit(t, c, nconn, prm, l, hilf, rhs, res, bp, aw, ae, as, an,
                     ab, at, ap, lw, ls, lb, lpr, ue, un, ut, mip, idx, rap, &mh);
        }
        for (int l = t->maxlevel; l >= t->minlevel; --l)
            og_ftoc_s(t, l, hilf, hilf, ddx, ddy, ddz, bpr, 'P');
        double *fh[1] = {hilf};
        og_exchange(t, c, nconn, -1, 1, 0, 0, fh, 1);
        og_laplacephi(t, -1, res, hilf, aw, ae, as, an, ab, at, ap, bp);
        og_rescal(t, rhs, res);
        for (int l = t->maxlevel; l >= t->minlevel + 1; --l)
            og_ftoc_s(t, l, rhs, rhs, ddx, ddy, ddz, bpr, 'R');
        maxrhs = og_maxabscal(t, rhs, NULL);
        og_accumulate(n, dp, hilf);
        memset(hilf, 0, (size_t)n * sizeof(double));
        ipc += prm->ninner;
        if (st->history) st->history[ipcount - 1] = maxrhs;
        if (ipcount >= prm->nouter_min) {
            int done;
            if (prm->eps_rel > 0.0)
                done = maxrhs <= prm->eps_rel * st->maxrhs0;
            else
                done = maxrhs / prefak < prm->epcorr;
            if (done) { ++ipcount; break; }
        }
    }
    st->ipc = ipc;
    st->nouter_done = ipcount - 1;
    st->maxrhs = maxrhs;

    for (int l = t->minlevel; l <= t->maxlevel; ++l) {
        og_parent(t, l, dp, pbd, ipbb);
        og_bound_pressure(t, l, dp, pbd, ipbb, md.bp, dx, dy, dz, ddx, ddy, ddz);
    }
    og_mgpcorr(t, u, v, w, p, dp, rdx, rdy, rdz, dt / prm->rho,
               prm->ibtype ? bp : NULL);
    for (int l = t->maxlevel; l >= t->minlevel; --l)
        og_ftoc_uvwp(t, l, u, v, w, p, ddx, ddy, ddz, bpr);
    /* per level, coarse to fine: parent, then bound_flow to fill the
     * ghost layers at PAR and wall faces, then the exchange inside the
     * level, which uses them (src/flow/pressuresolver_mod.F90:338-348); a
     * caller that runs these itself stops after ftoc */
    if (!prm->sync_by_caller) {
        int64_t *bbo = malloc(sizeof(int64_t) * 6 * (size_t)t->ngrid);
        const int64_t nb = og_bound_buffers(t, bbo);
        const size_t nbz = (size_t)(nb > 0 ? nb : 1);
        double *ub = calloc(nbz, sizeof(double)), *vb = calloc(nbz, sizeof(double)),
               *wb = calloc(nbz, sizeof(double));
        if (prm->ubuf) memcpy(ub, prm->ubuf, (size_t)nb * sizeof(double));
        if (prm->vbuf) memcpy(vb, prm->vbuf, (size_t)nb * sizeof(double));
        if (prm->wbuf) memcpy(wb, prm->wbuf, (size_t)nb * sizeof(double));
        double *fs[4] = {u, v, w, p};
        for (int l = t->minlevel; l <= t->maxlevel; ++l) {
            og_parent_uvw(t, l, u, v, w, ub, vb, wb, bbo);
            og_bound_flow(t, l, u, v, w, p, ub, vb, wb, bbo, prm->pinf, bp, ddx, ddy, ddz,
                          prm->rho);
            og_exchange(t, c, nconn, l, 2, 1, 0, fs, 4);
        }
        free(ub);
        free(vb);
        free(wb);
        free(bbo);
    }
    free(pbh);
    free(pbd);
    free(ipbb);
    return 0;
}

/* ------------------------------------------------------------------ */
double og_time_sip(const og_topo *t, const og_conn *c, int nconn, int niter,
                   int nthreads, const double *rhs, double *hilf, double *res,
                   const double *aw, const double *ae, const double *as,
                   const double *an, const double *ab, const double *at,
                   const double *ap, const double *bp, const double *lw,
                   const double *ls, const double *lb, const double *lpr,
                   const double *ue, const double *un, const double *ut,
                   const int64_t *mip, const int64_t *idx)
{
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    for (int it = 1; it <= niter; ++it) {
        og_laplacephi(t, -1, res, hilf, aw, ae, as, an, ab, at, ap, bp);
        og_sipiter1_hp(t, -1, rhs, res, lw, ls, lb, lpr, mip, idx);
        double *fr[1] = {res};
        og_exchange(t, c, nconn, -1, 1, 0, 0, fr, 1);
        og_sipiter2_hp(t, -1, hilf, res, ue, un, ut, mip, idx);
        double *fh[1] = {hilf};
        og_exchange(t, c, nconn, -1, 1, 0, 0, fh, 1);
    }
    clock_gettime(CLOCK_MONOTONIC, &b);
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

/* ------------------------------------------------------------------ */
/* Momentum right-hand side of one RK stage, tstle4 without immersed
 * boundaries and without PAR faces (src/flow/tstle4_mod.F90:13-120):
 * uo = convection (tstle4_kon :139-265), minus the Werner-Wengle wall
 * stress next to NOS walls (swcle3d, :1756-1845, called at the start of
 * tstle4_diff :315), + diffusion (tstle4_diff :268-458) - pressure gradient
 * (tstle4_gradp :461-529), then tstle4_par at PAR faces (:532-1753, see
 * og_tstle4_par).  ut = u etc. (noib).  Each expression keeps the
 * reference's association order (Fortran evaluates operators of equal
 * precedence left to right; a leading minus applies to the whole product). */
static double gpair(double a, double b, double gmol)
{
    const double s = a + b;
    return a * b / (s > gmol ? s : gmol);
}

/* Werner-Wengle constants (init_wernerwengle, src/flow/wernerwengle_mod.F90:
 * 13-29): ww[0..4] = cwb, cpo4, cpo5, cpo6, cpo8 */
void og_ww_constants(double rho, double gmol, double ww[5])
{
    const double cwa = 8.3, cwb = 1.0 / 7.0;
    const double cpo1 = 1.0 - cwb, cpo2 = 1.0 + cwb;
    ww[0] = cwb;
    ww[1] = cpo2 / cwa * pow(gmol / rho, cwb);
    ww[2] = 0.5 * cpo1 * pow(cwa, cpo2 / cpo1) * pow(gmol / rho, cpo2);
    ww[3] = 0.5 * gmol / rho * pow(cwa, 2.0 / cpo1);
    ww[4] = 2.0 / cpo2;
}

/* swcle3d_one (src/flow/tstle4_mod.F90:1848-1859) with tauwin
 * (src/flow/wernerwengle_mod.F90:64-87): the wall shear stress of the
 * velocity uq at wall distance dd, over rho dd */
double og_swc(double dd, double uq, double rho, double gmol, const double ww[5])
{
    const double vz = copysign(1.0, uq), un = fabs(uq);
    double tau;
    if (un >= ww[3] / dd) {            /* fully turbulent */
        const double ddsb = pow(dd, -ww[0]);
        tau = vz * rho * pow(ddsb * (ww[1] * un + ww[2] / dd), ww[4]);
    } else {                           /* laminar sublayer */
        tau = vz * 2.0 * gmol * un / dd;
    }
    return tau / rho / dd;
}

/* The swcle3d terms that velocity component comp (0 u, 1 v, 2 w) at
 * (k, j, i) receives from the NOS walls of its grid, in the reference's
 * face order (front, back, right, left, bottom, top; each face loop runs
 * over 2..n-1 of the two tangential indices). */
static double swc_terms(double val, int comp, int k, int j, int i, const int32_t *bc,
                        int kk, int jj, int ii, double vel, const double *ddx,
                        const double *ddy, const double *ddz, double rho, double gmol,
                        const double ww[5])
{
    const int tk = k >= 2 && k <= kk - 1, tj = j >= 2 && j <= jj - 1, ti = i >= 2 && i <= ii - 1;
    if (comp != 0) {                   /* v and w on x walls */
        if (bc[0] == 5 && i == 3 && tj && tk) val = val - og_swc(ddx[i], vel, rho, gmol, ww);
        if (bc[1] == 5 && i == ii - 2 && tj && tk) val = val - og_swc(ddx[i], vel, rho, gmol, ww);
    }
    if (comp != 1) {                   /* u and w on y walls */
        if (bc[2] == 5 && j == 3 && ti && tk) val = val - og_swc(ddy[j], vel, rho, gmol, ww);
        if (bc[3] == 5 && j == jj - 2 && ti && tk) val = val - og_swc(ddy[j], vel, rho, gmol, ww);
    }
    if (comp != 2) {                   /* u and v on z walls */
        if (bc[4] == 5 && k == 3 && ti && tj) val = val - og_swc(ddz[k], vel, rho, gmol, ww);
        if (bc[5] == 5 && k == kk - 2 && ti && tj) val = val - og_swc(ddz[k], vel, rho, gmol, ww);
    }
    return val;
}

/* ------------------------------------------------------------------ */
/* tstle4_par (src/flow/tstle4_mod.F90:532-1753): the convection terms next
 * to PAR faces, added to uo, vo, wo of one grid after kon, diff and gradp.
 * The reference writes each face out by hand; every instance has the same
 * form, restated here once for a face with normal axis n (0 x, 1 y, 2 z),
 * side lo (faces 1, 3, 5) or hi (2, 4, 6), and a tangential momentum
 * component c, with o the third axis.  Positions along n: L the first
 * interior layer (3, or N-2), G the ghost layer beside it (2, or N-1), T
 * the face between them (G at lo faces, L at hi faces), where the normal
 * transport velocity lives.
 *
 * 1. Upwind at the second layer (:571-776), faces in the order 1..6, both
 *    tangential components over a, b in 3..N-2: with i = 4 (lo) or N-2
 *    (hi) along n, f = area*(t(i-1; a) + t(i-1; a+1))*0.5 and
 *    q = 0.5*(f -|f|)*(0.5*c(i) - 1.5*c(i-1) + c(i-2))*0.5*0.5 (lo), or
 *    q = 0.5*(f +|f|)*(0.5*c(i-1) - 1.5*c(i) + c(i+1))*0.5*0.5 (hi);
 *    co(i) += F*(-q), co(i-1) += F*q, with F of position i for both.
 * 2. The momentum-conserving flux (:778-1750), faces in the order 2, 1, 6,
 *    5, 4, 3; x faces w then v, y faces u then w, z faces u then v.  Over
 *    a in 2..Nc-2 (the component's own axis) and b in 3..No-2: the standard
 *    flux qs = 0.5*f*(c(L) + c(G)), f = area*(t(T; a) + t(T; a+1))*0.5;
 *    the coarse-consistent flux qn at even a, shared by the pair (b, b+1)
 *    (b = 3, 5, ..): f = area(a, b)*S*0.125 with S the sum of t(T) over
 *    a-1..a+2 and b, b+1, the outermost a-1 or a+2 replaced by doubling the
 *    inner end at a = 2 and a = Nc-2; odd a take the mean of the even
 *    neighbours; then co(L) += F*(qs - qn) (lo) or F*(qn - qs) (hi).
 *    (The reference's LEFT/RIGHT w cells also compute qn at odd a, which
 *    the mean then overwrites.)
 * area = d_c(a)*dd_o(b) (the momentum cell's face); F = -(rX*rY)*rZ of the
 * updated cell, rX = rdx for u else rddx, rY = rdy for v else rddy, rZ =
 * rdz for w else rddz, grouped (x*y)*z as the reference's fkdt*rz.  t is
 * the transport velocity, u, v, w themselves (no immersed boundary). */
typedef struct par_grid {
    int kk, jj, ii, n[3];
    double *out[3];
    const double *f[3];
    const double *d[3], *dd[3], *rd[3], *rdd[3];   /* 1-based */
} par_grid;

static int64_t par_ix(const par_grid *P, int ax0, int v0, int ax1, int v1, int ax2, int v2)
{
    int x[3];
    x[ax0] = v0;
    x[ax1] = v1;
    x[ax2] = v2;
    return (int64_t)(x[2] - 1) + (int64_t)(x[1] - 1) * P->kk + (int64_t)(x[0] - 1) * P->kk * P->jj;
}

static double par_cellfac(const par_grid *P, int comp, int ax0, int v0, int ax1, int v1,
                          int ax2, int v2)
{
    int x[3];
    x[ax0] = v0;
    x[ax1] = v1;
    x[ax2] = v2;
    const double rx = comp == 0 ? P->rd[0][x[0]] : P->rdd[0][x[0]];
    const double ry = comp == 1 ? P->rd[1][x[1]] : P->rdd[1][x[1]];
    const double rz = comp == 2 ? P->rd[2][x[2]] : P->rdd[2][x[2]];
    double fk = -1.0 * rx * ry;
    fk = fk * 1.0;                 /* wkon */
    return fk * rz;
}

static void par_upwind(par_grid *P, int n, int hi)
{
    const int N = P->n[n];
    const int iu = hi ? N - 2 : 4;
    const double *tn = P->f[n];
    for (int c = 0; c < 3; ++c) {
        if (c == n) continue;
        const int o = 3 - n - c;
        const double *cf = P->f[c];
        double *co = P->out[c];
        for (int a = 3; a <= P->n[c] - 2; ++a)
            for (int b = 3; b <= P->n[o] - 2; ++b) {
                const double area = P->d[c][a] * P->dd[o][b];
                const double f = area * (tn[par_ix(P, n, iu - 1, c, a, o, b)] +
                                         tn[par_ix(P, n, iu - 1, c, a + 1, o, b)]) * 0.5;
                double q;
                if (!hi)
                    q = 0.5 * (f - fabs(f)) *
                        (0.5 * cf[par_ix(P, n, iu, c, a, o, b)] - 1.5 * cf[par_ix(P, n, iu - 1, c, a, o, b)] +
                         cf[par_ix(P, n, iu - 2, c, a, o, b)]) * 0.5 * 0.5;
                else
                    q = 0.5 * (f + fabs(f)) *
                        (0.5 * cf[par_ix(P, n, iu - 1, c, a, o, b)] - 1.5 * cf[par_ix(P, n, iu, c, a, o, b)] +
                         cf[par_ix(P, n, iu + 1, c, a, o, b)]) * 0.5 * 0.5;
                const double F = par_cellfac(P, c, n, iu, c, a, o, b);
                const int64_t x0 = par_ix(P, n, iu, c, a, o, b), x1 = par_ix(P, n, iu - 1, c, a, o, b);
                co[x0] = co[x0] + F * (-q);
                co[x1] = co[x1] + F * q;
            }
    }
}

static void par_conserve(par_grid *P, int n, int hi, int c)
{
    const int N = P->n[n], o = 3 - n - c;
    const int L = hi ? N - 2 : 3, G = hi ? N - 1 : 2, T = hi ? L : G;
    const int Nc = P->n[c], No = P->n[o];
    const double *tn = P->f[n], *cf = P->f[c];
    double *co = P->out[c];
    double *qs = calloc((size_t)(Nc + 1) * (No + 1), sizeof(double));
    double *qn = calloc((size_t)(Nc + 1) * (No + 1), sizeof(double));
#define Q(arr, a, b) arr[(size_t)(a) * (No + 1) + (b)]
#define TN(a, b) tn[par_ix(P, n, T, c, a, o, b)]
#define CL(a, b) cf[par_ix(P, n, L, c, a, o, b)]
#define CG(a, b) cf[par_ix(P, n, G, c, a, o, b)]
    for (int a = 2; a <= Nc - 2; ++a)
        for (int b = 3; b <= No - 2; ++b) {
            const double f = P->d[c][a] * P->dd[o][b] * (TN(a, b) + TN(a + 1, b)) * 0.5;
            Q(qs, a, b) = 0.5 * f * (CL(a, b) + CG(a, b));
        }
    for (int a = 2; a <= Nc - 2; a += 2)
        for (int b = 3; b <= No - 3; b += 2) {
            const double area = P->d[c][a] * P->dd[o][b];
            double sum;
            if (a == 2)
                sum = 2.0 * TN(a, b) + 2.0 * TN(a, b + 1) + TN(a + 1, b) + TN(a + 1, b + 1) +
                      TN(a + 2, b) + TN(a + 2, b + 1);
            else if (a == Nc - 2)
                sum = TN(a - 1, b) + TN(a - 1, b + 1) + TN(a, b) + TN(a, b + 1) +
                      2.0 * TN(a + 1, b) + 2.0 * TN(a + 1, b + 1);
            else
                sum = TN(a - 1, b) + TN(a - 1, b + 1) + TN(a, b) + TN(a, b + 1) + TN(a + 1, b) +
                      TN(a + 1, b + 1) + TN(a + 2, b) + TN(a + 2, b + 1);
            const double f = area * sum * 0.125;
            Q(qn, a, b) = 0.5 * f * (CL(a, b) + CG(a, b));
            Q(qn, a, b + 1) = 0.5 * f * (CL(a, b + 1) + CG(a, b + 1));
        }
    for (int a = 3; a <= Nc - 3; a += 2)
        for (int b = 3; b <= No - 2; ++b)
            Q(qn, a, b) = 0.5 * (Q(qn, a - 1, b) + Q(qn, a + 1, b));
    for (int a = 2; a <= Nc - 2; ++a)
        for (int b = 3; b <= No - 2; ++b) {
            const int64_t x = par_ix(P, n, L, c, a, o, b);
            const double F = par_cellfac(P, c, n, L, c, a, o, b);
            const double dq = hi ? Q(qn, a, b) - Q(qs, a, b) : Q(qs, a, b) - Q(qn, a, b);
            co[x] = co[x] + F * dq;
        }
#undef Q
#undef TN
#undef CL
#undef CG
    free(qs);
    free(qn);
}

static void par_grid_run(par_grid *P, const int32_t *bc)
{
    /* faces 1..6 = (n, hi): (0,0) (0,1) (1,0) (1,1) (2,0) (2,1) */
    for (int f = 0; f < 6; ++f)
        if (bc[f] == 8) par_upwind(P, f / 2, f & 1);
    static const int order[6] = {1, 0, 5, 4, 3, 2};
    static const int comps[3][2] = {{2, 1}, {0, 2}, {0, 1}};
    for (int q = 0; q < 6; ++q) {
        const int f = order[q];
        if (bc[f] != 8) continue;
        par_conserve(P, f / 2, f & 1, comps[f / 2][0]);
        par_conserve(P, f / 2, f & 1, comps[f / 2][1]);
    }
}

static void par_setup(par_grid *P, const og_topo *t, int g, double *uo_, double *vo_,
                      double *wo_, const double *u_, const double *v_, const double *w_,
                      const double *dx_, const double *dy_, const double *dz_,
                      const double *ddx_, const double *ddy_, const double *ddz_,
                      const double *rdx_, const double *rdy_, const double *rdz_,
                      const double *rddx_, const double *rddy_, const double *rddz_)
{
    GDIMS(t, g);
    const int64_t o3 = t->ip3[g - 1];
    const int64_t ox = t->ip1x[g - 1] - 1, oy = t->ip1y[g - 1] - 1, oz = t->ip1z[g - 1] - 1;
    P->kk = kk;
    P->jj = jj;
    P->ii = ii;
    P->n[0] = ii;
    P->n[1] = jj;
    P->n[2] = kk;
    P->out[0] = uo_ + o3;
    P->out[1] = vo_ + o3;
    P->out[2] = wo_ + o3;
    P->f[0] = u_ + o3;
    P->f[1] = v_ + o3;
    P->f[2] = w_ + o3;
    P->d[0] = dx_ + ox; P->d[1] = dy_ + oy; P->d[2] = dz_ + oz;
    P->dd[0] = ddx_ + ox; P->dd[1] = ddy_ + oy; P->dd[2] = ddz_ + oz;
    P->rd[0] = rdx_ + ox; P->rd[1] = rdy_ + oy; P->rd[2] = rdz_ + oz;
    P->rdd[0] = rddx_ + ox; P->rdd[1] = rddy_ + oy; P->rdd[2] = rddz_ + oz;
}

void og_tstle4_par(const og_topo *t, double *uo, double *vo, double *wo,
                   const double *u, const double *v, const double *w,
                   const double *dx, const double *dy, const double *dz,
                   const double *ddx, const double *ddy, const double *ddz,
                   const double *rdx, const double *rdy, const double *rdz,
                   const double *rddx, const double *rddy, const double *rddz)
{
    for (int g = 1; g <= t->ngrid; ++g) {
        par_grid P;
        par_setup(&P, t, g, uo, vo, wo, u, v, w, dx, dy, dz, ddx, ddy, ddz, rdx, rdy, rdz,
                  rddx, rddy, rddz);
        par_grid_run(&P, t->bc + 6 * (g - 1));
    }
}

int og_tstle4(const og_topo *t, double *uo_, double *vo_, double *wo_,
              const double *u_, const double *v_, const double *w_,
              const double *p_, const double *g_, double rho, double gmol,
              int ilesmodel, const double *gradp,
              const double *dx_, const double *dy_, const double *dz_,
              const double *ddx_, const double *ddy_, const double *ddz_,
              const double *rdx_, const double *rdy_, const double *rdz_,
              const double *rddx_, const double *rddy_, const double *rddz_)
{
    const double iles = ilesmodel == 0 ? 0.0 : 1.0;
    double ww[5];
    og_ww_constants(rho, gmol, ww);
    for (int g = 1; g <= t->ngrid; ++g) {
        GDIMS(t, g);
        const int64_t o3 = t->ip3[g - 1];
        double *uo = uo_ + o3, *vo = vo_ + o3, *wo = wo_ + o3;
        const double *u = u_ + o3, *v = v_ + o3, *w = w_ + o3;
        const double *ut = u, *vt = v, *wt = w;
        const double *p = p_ + o3, *gg = g_ + o3;
        const double *dx = dx_ + t->ip1x[g - 1] - 1, *ddx = ddx_ + t->ip1x[g - 1] - 1;
        const double *dy = dy_ + t->ip1y[g - 1] - 1, *ddy = ddy_ + t->ip1y[g - 1] - 1;
        const double *dz = dz_ + t->ip1z[g - 1] - 1, *ddz = ddz_ + t->ip1z[g - 1] - 1;
        const double *rdx = rdx_ + t->ip1x[g - 1] - 1, *rddx = rddx_ + t->ip1x[g - 1] - 1;
        const double *rdy = rdy_ + t->ip1y[g - 1] - 1, *rddy = rddy_ + t->ip1y[g - 1] - 1;
        const double *rdz = rdz_ + t->ip1z[g - 1] - 1, *rddz = rddz_ + t->ip1z[g - 1] - 1;
        const int32_t *bc = t->bc + 6 * (g - 1);
        /* nfro..ntop = faces 1..6; CON = 7 and OP1 = 3 extend the loops */
        const int nfu = bc[0] == 3, nbu = bc[1] == 7 || bc[1] == 3;
        const int nrv = bc[2] == 3, nlv = bc[3] == 7 || bc[3] == 3;
        const int nbw = bc[4] == 3, ntw = bc[5] == 7 || bc[5] == 3;
        const double gpx = gradp ? gradp[0] : 0.0, gpy = gradp ? gradp[1] : 0.0,
                     gpz = gradp ? gradp[2] : 0.0;
        for (int64_t c = 0; c < (int64_t)ii * jj * kk; ++c) uo[c] = vo[c] = wo[c] = 0.0;
#define G(k, j, i) gg[IX(k, j, i)]
        /* u momentum */
        for (int i = 3 - nfu; i <= ii - 3 + nbu; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3; k <= kk - 2; ++k) {
                    const int64_t c = IX(k, j, i);
                    double ax = ddy[j] * ddz[k], ay = dx[i] * ddz[k], az = dx[i] * ddy[j];
                    double fe = ax * (ut[c] + (ut[IX(k, j, i + 1)] - ut[c]) * 0.5 * dx[i] / ddx[i + 1]);
                    double fw = ax * (ut[IX(k, j, i - 1)] + (ut[c] - ut[IX(k, j, i - 1)]) * 0.5 * dx[i - 1] / ddx[i]);
                    double fn = ay * (vt[c] + vt[IX(k, j, i + 1)]) * 0.5;
                    double fs = ay * (vt[IX(k, j - 1, i)] + vt[IX(k, j - 1, i + 1)]) * 0.5;
                    double ft = az * (wt[c] + wt[IX(k, j, i + 1)]) * 0.5;
                    double fb = az * (wt[IX(k - 1, j, i)] + wt[IX(k - 1, j, i + 1)]) * 0.5;
                    double qe = 0.5 * fe * (u[c] + u[IX(k, j, i + 1)]);
                    double qw = 0.5 * fw * (u[IX(k, j, i - 1)] + u[c]);
                    double qn = 0.5 * fn * (u[c] + u[IX(k, j + 1, i)]);
                    double qs = 0.5 * fs * (u[IX(k, j - 1, i)] + u[c]);
                    double qt = 0.5 * ft * (u[c] + u[IX(k + 1, j, i)]);
                    double qb = 0.5 * fb * (u[IX(k - 1, j, i)] + u[c]);
                    uo[c] = -((qe - qw + qn - qs + qt - qb) * rdx[i] * rddy[j] * rddz[k]);
                    uo[c] = swc_terms(uo[c], 0, k, j, i, bc, kk, jj, ii, u[c], ddx, ddy, ddz, rho, gmol, ww);
                    /* diffusion */
                    const double ge = G(k, j, i + 1), gw = G(k, j, i);
                    const double gn = gpair(G(k, j, i), G(k, j + 1, i), gmol) + gpair(G(k, j, i + 1), G(k, j + 1, i + 1), gmol);
                    const double gs = gpair(G(k, j - 1, i), G(k, j, i), gmol) + gpair(G(k, j - 1, i + 1), G(k, j, i + 1), gmol);
                    const double gt = gpair(G(k, j, i), G(k + 1, j, i), gmol) + gpair(G(k, j, i + 1), G(k + 1, j, i + 1), gmol);
                    const double gb = gpair(G(k - 1, j, i), G(k, j, i), gmol) + gpair(G(k - 1, j, i + 1), G(k, j, i + 1), gmol);
                    qe = -(ge * ax * rddx[i + 1] * (u[IX(k, j, i + 1)] - u[c]));
                    qw = -(gw * ax * rddx[i] * (u[c] - u[IX(k, j, i - 1)]));
                    qn = -(gn * ay * rdy[j] * (u[IX(k, j + 1, i)] - u[c]));
                    qs = -(gs * ay * rdy[j - 1] * (u[c] - u[IX(k, j - 1, i)]));
                    qt = -(gt * az * rdz[k] * (u[IX(k + 1, j, i)] - u[c]));
                    qb = -(gb * az * rdz[k - 1] * (u[c] - u[IX(k - 1, j, i)]));
                    const double st = ((ge * (u[IX(k, j, i + 1)] - u[c]) * rddx[i + 1])
                                       - (gw * (u[c] - u[IX(k, j, i - 1)]) * rddx[i])) * ax
                                      + ((gn * (v[IX(k, j, i + 1)] - v[c]))
                                         - (gs * (v[IX(k, j - 1, i + 1)] - v[IX(k, j - 1, i)]))) * ddz[k]
                                      + ((gt * (w[IX(k, j, i + 1)] - w[c]))
                                         - (gb * (w[IX(k - 1, j, i + 1)] - w[IX(k - 1, j, i)]))) * ddy[j];
                    const double qc = st * iles;
                    const double fak = 1.0 / rho * rddy[j] * rdx[i] * rddz[k];
                    uo[c] = uo[c] - fak * (qe - qw + qn - qs + qt - qb - qc);
                    uo[c] = uo[c] - 1.0 / (rho * dx[i]) * (p[IX(k, j, i + 1)] - p[c] + gpx * dx[i]);
                }
        /* v momentum */
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3 - nrv; j <= jj - 3 + nlv; ++j)
                for (int k = 3; k <= kk - 2; ++k) {
                    const int64_t c = IX(k, j, i);
                    double ax = dy[j] * ddz[k], ay = ddx[i] * ddz[k], az = ddx[i] * dy[j];
                    double fe = ax * (ut[c] + ut[IX(k, j + 1, i)]) * 0.5;
                    double fw = ax * (ut[IX(k, j, i - 1)] + ut[IX(k, j + 1, i - 1)]) * 0.5;
                    double fn = ay * (vt[c] + (vt[IX(k, j + 1, i)] - vt[c]) * 0.5 * dy[j] / ddy[j + 1]);
                    double fs = ay * (vt[IX(k, j - 1, i)] + (vt[c] - vt[IX(k, j - 1, i)]) * 0.5 * dy[j - 1] / ddy[j]);
                    double ft = az * (wt[c] + wt[IX(k, j + 1, i)]) * 0.5;
                    double fb = az * (wt[IX(k - 1, j, i)] + wt[IX(k - 1, j + 1, i)]) * 0.5;
                    double qe = 0.5 * fe * (v[c] + v[IX(k, j, i + 1)]);
                    double qw = 0.5 * fw * (v[IX(k, j, i - 1)] + v[c]);
                    double qn = 0.5 * fn * (v[c] + v[IX(k, j + 1, i)]);
                    double qs = 0.5 * fs * (v[IX(k, j - 1, i)] + v[c]);
                    double qt = 0.5 * ft * (v[c] + v[IX(k + 1, j, i)]);
                    double qb = 0.5 * fb * (v[IX(k - 1, j, i)] + v[c]);
                    vo[c] = -((qe - qw + qn - qs + qt - qb) * rddx[i] * rdy[j] * rddz[k]);
                    vo[c] = swc_terms(vo[c], 1, k, j, i, bc, kk, jj, ii, v[c], ddx, ddy, ddz, rho, gmol, ww);
                    const double ge = gpair(G(k, j, i), G(k, j, i + 1), gmol) + gpair(G(k, j + 1, i), G(k, j + 1, i + 1), gmol);
                    const double gw = gpair(G(k, j, i - 1), G(k, j, i), gmol) + gpair(G(k, j + 1, i - 1), G(k, j + 1, i), gmol);
                    const double gn = G(k, j + 1, i), gs = G(k, j, i);
                    const double gt = gpair(G(k, j, i), G(k + 1, j, i), gmol) + gpair(G(k, j + 1, i), G(k + 1, j + 1, i), gmol);
                    const double gb = gpair(G(k - 1, j, i), G(k, j, i), gmol) + gpair(G(k - 1, j + 1, i), G(k, j + 1, i), gmol);
                    qe = -(ge * ax * rdx[i] * (v[IX(k, j, i + 1)] - v[c]));
                    qw = -(gw * ax * rdx[i - 1] * (v[c] - v[IX(k, j, i - 1)]));
                    qn = -(gn * ay * rddy[j + 1] * (v[IX(k, j + 1, i)] - v[c]));
                    qs = -(gs * ay * rddy[j] * (v[c] - v[IX(k, j - 1, i)]));
                    qt = -(gt * az * rdz[k] * (v[IX(k + 1, j, i)] - v[c]));
                    qb = -(gb * az * rdz[k - 1] * (v[c] - v[IX(k - 1, j, i)]));
                    const double st = ((ge * (u[IX(k, j + 1, i)] - u[c]))
                                       - (gw * (u[IX(k, j + 1, i - 1)] - u[IX(k, j, i - 1)]))) * ddz[k]
                                      + ((gn * (v[IX(k, j + 1, i)] - v[c]) * rddy[j + 1])
                                         - (gs * (v[c] - v[IX(k, j - 1, i)]) * rddy[j])) * ay
                                      + ((gt * (w[IX(k, j + 1, i)] - w[c]))
                                         - (gb * (w[IX(k - 1, j + 1, i)] - w[IX(k - 1, j, i)]))) * ddx[i];
                    const double qc = st * iles;
                    const double fak = 1.0 / rho * rddx[i] * rdy[j] * rddz[k];
                    vo[c] = vo[c] - fak * (qe - qw + qn - qs + qt - qb - qc);
                    vo[c] = vo[c] - 1.0 / (rho * dy[j]) * (p[IX(k, j + 1, i)] - p[c] + gpy * dy[j]);
                }
        /* w momentum */
        for (int i = 3; i <= ii - 2; ++i)
            for (int j = 3; j <= jj - 2; ++j)
                for (int k = 3 - nbw; k <= kk - 3 + ntw; ++k) {
                    const int64_t c = IX(k, j, i);
                    double ax = ddy[j] * dz[k], ay = ddx[i] * dz[k], az = ddx[i] * ddy[j];
                    double fe = ax * (ut[c] + ut[IX(k + 1, j, i)]) * 0.5;
                    double fw = ax * (ut[IX(k, j, i - 1)] + ut[IX(k + 1, j, i - 1)]) * 0.5;
                    double fn = ay * (vt[c] + vt[IX(k + 1, j, i)]) * 0.5;
                    double fs = ay * (vt[IX(k, j - 1, i)] + vt[IX(k + 1, j - 1, i)]) * 0.5;
                    double ft = az * (wt[c] + (wt[IX(k + 1, j, i)] - wt[c]) * 0.5 * dz[k] / ddz[k + 1]);
                    double fb = az * (wt[IX(k - 1, j, i)] + (wt[c] - wt[IX(k - 1, j, i)]) * 0.5 * dz[k - 1] / ddz[k]);
                    double qe = 0.5 * fe * (w[c] + w[IX(k, j, i + 1)]);
                    double qw = 0.5 * fw * (w[IX(k, j, i - 1)] + w[c]);
                    double qn = 0.5 * fn * (w[c] + w[IX(k, j + 1, i)]);
                    double qs = 0.5 * fs * (w[IX(k, j - 1, i)] + w[c]);
                    double qt = 0.5 * ft * (w[c] + w[IX(k + 1, j, i)]);
                    double qb = 0.5 * fb * (w[IX(k - 1, j, i)] + w[c]);
                    wo[c] = -((qe - qw + qn - qs + qt - qb) * rddx[i] * rddy[j] * rdz[k]);
                    wo[c] = swc_terms(wo[c], 2, k, j, i, bc, kk, jj, ii, w[c], ddx, ddy, ddz, rho, gmol, ww);
                    const double ge = gpair(G(k, j, i), G(k, j, i + 1), gmol) + gpair(G(k + 1, j, i), G(k + 1, j, i + 1), gmol);
                    const double gw = gpair(G(k, j, i - 1), G(k, j, i), gmol) + gpair(G(k + 1, j, i - 1), G(k + 1, j, i), gmol);
                    const double gn = gpair(G(k, j, i), G(k, j + 1, i), gmol) + gpair(G(k + 1, j, i), G(k + 1, j + 1, i), gmol);
                    const double gs = gpair(G(k, j - 1, i), G(k, j, i), gmol) + gpair(G(k + 1, j - 1, i), G(k + 1, j, i), gmol);
                    const double gt = G(k + 1, j, i), gb = G(k, j, i);
                    qe = -(ge * ax * rdx[i] * (w[IX(k, j, i + 1)] - w[c]));
                    qw = -(gw * ax * rdx[i - 1] * (w[c] - w[IX(k, j, i - 1)]));
                    qn = -(gn * ay * rdy[j] * (w[IX(k, j + 1, i)] - w[c]));
                    qs = -(gs * ay * rdy[j - 1] * (w[c] - w[IX(k, j - 1, i)]));
                    qt = -(gt * az * rddz[k + 1] * (w[IX(k + 1, j, i)] - w[c]));
                    qb = -(gb * az * rddz[k] * (w[c] - w[IX(k - 1, j, i)]));
                    const double st = ((ge * (u[IX(k + 1, j, i)] - u[c]))
                                       - (gw * (u[IX(k + 1, j, i - 1)] - u[IX(k, j, i - 1)]))) * ddy[j]
                                      + ((gn * (v[IX(k + 1, j, i)] - v[c]))
                                         - (gs * (v[IX(k + 1, j - 1, i)] - v[IX(k, j - 1, i)]))) * ddx[i]
                                      + ((gt * (w[IX(k + 1, j, i)] - w[c]) * rddz[k + 1])
                                         - (gb * (w[c] - w[IX(k - 1, j, i)]) * rddz[k])) * az;
                    const double qc = st * iles;
                    const double fak = 1.0 / rho * rddx[i] * rddy[j] * rdz[k];
                    wo[c] = wo[c] - fak * (qe - qw + qn - qs + qt - qb - qc);
                    wo[c] = wo[c] - 1.0 / (rho * dz[k]) * (p[IX(k + 1, j, i)] - p[c] + gpz * dz[k]);
                }
#undef G
        /* swcle3d at the positions outside the loop ranges above (its face
         * loops include the ghost rows 2 and n-1), where uo is still zero */
        for (int i = 1; i <= ii; ++i)
            for (int j = 1; j <= jj; ++j)
                for (int k = 1; k <= kk; ++k) {
                    const int64_t c = IX(k, j, i);
                    const int ik = k >= 3 && k <= kk - 2, ij = j >= 3 && j <= jj - 2,
                              iI = i >= 3 && i <= ii - 2;
                    if (!(i >= 3 - nfu && i <= ii - 3 + nbu && ij && ik))
                        uo[c] = swc_terms(uo[c], 0, k, j, i, bc, kk, jj, ii, u[c], ddx, ddy, ddz, rho, gmol, ww);
                    if (!(iI && j >= 3 - nrv && j <= jj - 3 + nlv && ik))
                        vo[c] = swc_terms(vo[c], 1, k, j, i, bc, kk, jj, ii, v[c], ddx, ddy, ddz, rho, gmol, ww);
                    if (!(iI && ij && k >= 3 - nbw && k <= kk - 3 + ntw))
                        wo[c] = swc_terms(wo[c], 2, k, j, i, bc, kk, jj, ii, w[c], ddx, ddy, ddz, rho, gmol, ww);
                }
        /* tstle4_par after kon, diff and gradp of the grid (:114-116) */
        par_grid P;
        par_setup(&P, t, g, uo_, vo_, wo_, u_, v_, w_, dx_, dy_, dz_, ddx_, ddy_, ddz_, rdx_,
                  rdy_, rdz_, rddx_, rddy_, rddz_);
        par_grid_run(&P, bc);
    }
    return 0;
}

/* src/core/rungekutta_mod.F90:315-333 */
void og_rkstep(int64_t n, double *p, double *dp, const double *rhsp,
               double frhs, double dtfu)
{
    for (int64_t i = 0; i < n; ++i) {
        dp[i] = frhs * dp[i] + rhsp[i];
        p[i] = p[i] + dtfu * dp[i];
    }
}
