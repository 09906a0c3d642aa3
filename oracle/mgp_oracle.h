/*
 * mgp_oracle.h -- CPU restatement of MGLET-base's pressure-projection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP library is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or as the timed CPU baseline.
 * Nothing in the product path links or calls it.
 *
 * Everything here works on the reference's own memory layout:
 *   - every 3-D field is one flat array, block `g` (1-based grid id) starts
 *     at ip3[g-1] (0-based), dims kk x jj x ii with k fastest and two ghost
 *     layers on each side (src/core/pointers_mod.F90:48-92);
 *   - 1-D metric/coefficient fields start at ip1x/ip1y/ip1z[g-1];
 *   - indices inside a block are the reference's 1-based (k, j, i).
 * All arithmetic is FP64 and is compiled with -ffp-contract=off so every
 * expression is evaluated in the reference's association order without
 * fused multiply-adds.
 */
#ifndef MGP_ORACLE_H
#define MGP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Block topology: the subset of GRIDINFO / itypboconds / pointers the path
 * reads (src/core/gridio_mod.F90:15-37, src/core/grids_mod.F90:253-290). */
typedef struct og_topo {
    int32_t ngrid;
    int32_t minlevel, maxlevel;
    const int32_t *kk, *jj, *ii;       /* [ngrid] dims incl. 2+2 ghosts   */
    const int32_t *level;              /* [ngrid]                          */
    const int32_t *nbr;                /* [ngrid*26] grid ids, 0 = none    */
    const int32_t *bc;                 /* [ngrid*6]  itypboconds(1,f,g)    */
    const int64_t *ip3;                /* [ngrid] 0-based 3-D offsets      */
    const int64_t *ip1x, *ip1y, *ip1z; /* [ngrid] 0-based 1-D offsets      */
    const int32_t *iparent;            /* [ngrid] parent grid id, 0 = none */
    const int32_t *pos;                /* [ngrid*3] iposition, jposition,  */
                                       /*   kposition (1-based, in parent) */
} og_topo;

/* One connection record (src/core/connect/connect_core_mod.F90:19-31).   */
typedef struct og_conn {
    int32_t recv_grid, send_grid, recv_face, send_face;
} og_conn;

typedef struct og_params {
    int32_t ityp;         /* 0 hyperplane SIP, 1 SOR (levels>min), 2 classic */
    int32_t ninner, nouter, nouter_min;
    double epcorr;        /* absolute criterion: maxrhs/prefak < epcorr    */
    double eps_rel;       /* >0: stop when maxrhs <= eps_rel*maxrhs0        */
    double omega;         /* SOR relaxation factor                          */
    double rho;
    int32_t ibtype;       /* 0 noib, 1 ghostcell (divcal*BP, masked pcorr)  */
    int32_t sor_all_levels; /* 1: run SOR also on minlevel                 */
    const double *sdiv;   /* ghostcell divergence source SDIV (3-D field),  */
                          /*   NULL = zero (no STL cut cells)              */
    int32_t sync_by_caller; /* 1: stop after ftoc(u, v, w, p); 0: run the  */
                          /*   closing parent, bound_flow, connect         */
    const double *ubuf, *vbuf, *wbuf; /* FIX/OP1 boundary values in the   */
                          /*   og_bound_buffers layout (NULL = zero)       */
    const double *pinf;   /* [ngrid*6] OP1 pressure parameter (NULL = 0)   */
} og_params;

typedef struct og_stats {
    int32_t ipc;          /* inner iterations run (reference `ipc`)        */
    int32_t nouter_done;  /* outer iterations run                          */
    double maxrhs0;       /* level-weighted max|RHS| before iterating      */
    double maxrhs;        /* after the last outer iteration                */
    double *history;      /* optional [nouter] maxrhs after each outer it. */
} og_stats;

/* src/ib/noib_mod.F90:81-250 */
void og_giteig(const og_topo *t, const double *dx, const double *dy,
               const double *dz, const double *bp, double *aw, double *ae,
               double *as, double *an, double *ab, double *at, double *ap);

/* src/flow/sip_hyperplane_mod.F90:269-447 (MIP / IDX, 1-based values) */
int og_hp_tables(const og_topo *t, int64_t *mip, int64_t *idx);

/* Test hook: change SIP alfa (reference value (double)0.92f). */
void og_set_alfa(double a);

/* src/flow/sip_hyperplane_mod.F90:106-257 (factors in access order) */
void og_lu_hp(const og_topo *t, const double *aw, const double *ae,
              const double *as, const double *an, const double *ab,
              const double *at, const double *ap, const double *bp,
              const int64_t *mip, const int64_t *idx, double *lw, double *ls,
              double *lb, double *lpr, double *ue, double *un, double *ut);

/* src/flow/sip_classic_mod.F90:17-127 (factors in spatial order) */
void og_lu_classic(const og_topo *t, const double *aw, const double *ae,
                   const double *as, const double *an, const double *ab,
                   const double *at, const double *ap, const double *bp,
                   double *lw, double *ls, double *lb, double *lpr,
                   double *ue, double *un, double *ut);

/* src/flow/laplacephi_mod.F90:14-165; level < 0 means all grids */
void og_laplacephi(const og_topo *t, int level, double *res, const double *phi,
                   const double *aw, const double *ae, const double *as,
                   const double *an, const double *ab, const double *at,
                   const double *ap, const double *bp);

/* src/flow/sip_hyperplane_mod.F90:450-569 */
void og_sipiter1_hp(const og_topo *t, int level, const double *rhs,
                    double *res, const double *lw, const double *ls,
                    const double *lb, const double *lpr, const int64_t *mip,
                    const int64_t *idx);
void og_sipiter2_hp(const og_topo *t, int level, double *phi, double *res,
                    const double *ue, const double *un, const double *ut,
                    const int64_t *mip, const int64_t *idx);

/* src/flow/sip_classic_mod.F90:143-198 */
void og_sipiter1_cl(const og_topo *t, int level, const double *rhs,
                    double *res, const double *lw, const double *ls,
                    const double *lb, const double *lpr);
void og_sipiter2_cl(const og_topo *t, int level, double *phi, double *res,
                    const double *ue, const double *un, const double *ut);

/* src/flow/sor_mod.F90:18-49 and :102-203; bp may be NULL (unmasked) */
void og_sor_init(const og_topo *t, const double *ap, double *rap);
void og_relax(const og_topo *t, int level, double *dp, const double *rhs,
              const double *aw, const double *ae, const double *as,
              const double *an, const double *ab, const double *at,
              const double *rap, const double *bp, double omg);

/* src/ib/noib_mod.F90:304-354 (bp/sdiv NULL for the noib call) */
void og_divcal(const og_topo *t, double *div, const double *u,
               const double *v, const double *w, const double *rddx,
               const double *rddy, const double *rddz, double fak,
               const double *bp, const double *sdiv);

/* src/flow/pressuresolver_mod.F90:745-879 (bp NULL = unmasked branch) */
void og_mgpcorr(const og_topo *t, double *u, double *v, double *w, double *p,
                const double *dp, const double *rdx, const double *rdy,
                const double *rdz, double rfak, const double *bp);

/* src/flow/pressuresolver_mod.F90:623-742, 882-908 */
double og_maxabscal(const og_topo *t, const double *phi, double *maxlvl);
void og_rescal(const og_topo *t, double *rhs, const double *res);
void og_accumulate(int64_t n, double *dp, const double *hilf);

/* src/core/connect/connect_core_mod.F90:167-433: returns the number of
 * connections written to `out` (capacity `cap`), -1 on overflow. */
int og_connect_build(const og_topo *t, og_conn *out, int cap);

/* src/core/connect/conn1_mod.F90:31-129, 561-704 and connect_core_mod
 * start_and_stop / corr_start_stop / decide, for one rank.
 * level < 0: all levels.  fields: nfield pointers to 3-D fields. */
int og_exchange(const og_topo *t, const og_conn *c, int nconn, int level,
            int layers, int corners, int forward, double **fields,
            int nfield);

/* Multi-level coupling (grids need iparent/pos).  PAR face buffers:
 * og_par_buffers fills ipbb[6*ngrid] (-1 = no buffer) and returns the
 * total length.  Metric pointers are the flat 1-D fields. */
int64_t og_par_buffers(const og_topo *t, int64_t *ipbb);
/* src/ib/ctof/ctof1_mod.F90:232-331 */
void og_ctof(const og_topo *t, int level, double *ff, const double *fc);
/* src/ib/parent/parent1_mod.F90:296-583 (one scalar field) */
void og_parent(const og_topo *t, int level, const double *f, double *pbuf,
               const int64_t *ipbb);
/* src/flow/bound_pressure_mod.F90:106-608 (bp NULL: A forms) */
void og_bound_pressure(const og_topo *t, int level, double *p,
                       const double *pbuf, const int64_t *ipbb,
                       const double *bp, const double *dx, const double *dy,
                       const double *dz, const double *ddx, const double *ddy,
                       const double *ddz);
/* src/ib/noib_restrict_mod.F90:255-298 (bp NULL) or the ghostcell
 * src/ib/gc_restrict_mod.F90 restrict_p / restrict_r (flag 'P' / 'R'),
 * into src/ib/ftoc_mod.F90:581-626 */
void og_ftoc_s(const og_topo *t, int level, const double *ff, double *fc,
               const double *ddx, const double *ddy, const double *ddz,
               const double *bp, char flag);
/* ftoc(ilevel, u, v, w, p): src/ib/noib_restrict_mod.F90:135-298; with bp
 * the pressure uses the ghostcell restrict_p */
void og_ftoc_uvwp(const og_topo *t, int level, double *u, double *v,
                  double *w, double *p, const double *ddx, const double *ddy,
                  const double *ddz, const double *bp);

/* Velocity boundary buffers of the FIX, OP1 and PAR faces (the field
 * buffers at ipbb, src/core/pointers_mod.F90:64-92): kk*jj / kk*ii / jj*ii
 * values per face, first index fastest; bbo[6*ngrid] offsets (-1 = none);
 * returns the total length. */
int64_t og_bound_buffers(const og_topo *t, int64_t *bbo);
/* parent(ilevel, u, v, w): src/ib/parent/parent1_mod.F90:296-583 with the
 * staggered prolongation (prolong1/2, istag 1) and start_and_stop's icomp
 * (src/ib/parent/parent_core_mod.F90:185-267), into the PAR face buffers */
void og_parent_uvw(const og_topo *t, int level, const double *u, const double *v,
                   const double *w, double *ubuf, double *vbuf, double *wbuf,
                   const int64_t *bbo);
/* par_ftoc_norm(ilevel, u, v, w, sum): src/ib/par_ftoc_mod.F90:35-475, the
 * face-normal velocity at the PAR faces of the grids of `level` into the
 * coarse face positions of the parent (and of the parent's neighbour where
 * the face lies on the parent's boundary): area-weighted mean (sum = 0) or
 * plain sum of each 2x2 group */
void og_par_ftoc_norm(const og_topo *t, int level, double *u, double *v, double *w,
                      const double *ddx, const double *ddy, const double *ddz, int sum);
/* bound_flow%bound(ilevel, u, v, w, p): src/flow/bound_flow_mod.F90 (FIX,
 * OP1, NOS, SLI, PAR) over src/core/bound_mod.F90:36-76; level < 0: all;
 * bp NULL = 1 */
void og_bound_flow(const og_topo *t, int level, double *u, double *v, double *w,
                   double *p, const double *ubuf, const double *vbuf, const double *wbuf,
                   const int64_t *bbo, const double *pinf, const double *bp,
                   const double *ddx, const double *ddy, const double *ddz, double rho);

/* src/flow/pressuresolver_mod.F90:160-357, single rank, all levels,
 * ending per level with parent(u, v, w), bound_flow and the 2-layer
 * connect with corners (unless prm->sync_by_caller).  Work arrays dp,
 * hilf, rhs, res are full 3-D fields. */
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
               og_stats *st);

/* Momentum right-hand side of one RK stage without immersed boundaries
 * (src/flow/tstle4_mod.F90:13-529: convection, Werner-Wengle wall stress
 * next to NOS walls, diffusion with viscosity field g, pressure gradient
 * with the optional body-force gradient gradp[3] on every grid).  uo, vo,
 * wo are zeroed and then filled, tstle4_par included at PAR faces. */
/* Werner-Wengle constants cwb, cpo4, cpo5, cpo6, cpo8
 * (src/flow/wernerwengle_mod.F90:13-29) and the wall term swcle3d_one
 * (src/flow/tstle4_mod.F90:1848-1859, tauwin :64-87 of wernerwengle_mod). */
void og_ww_constants(double rho, double gmol, double ww[5]);
double og_swc(double dd, double uq, double rho, double gmol, const double ww[5]);

int og_tstle4(const og_topo *t, double *uo, double *vo, double *wo,
              const double *u, const double *v, const double *w,
              const double *p, const double *g, double rho, double gmol,
              int ilesmodel, const double *gradp,
              const double *dx, const double *dy, const double *dz,
              const double *ddx, const double *ddy, const double *ddz,
              const double *rdx, const double *rdy, const double *rdz,
              const double *rddx, const double *rddy, const double *rddz);

/* src/core/rungekutta_mod.F90:315-333: dp = frhs dp + rhsp; p += dtfu dp */
/* tstle4_par alone (src/flow/tstle4_mod.F90:532-1753): adds the
 * upwind and momentum-conserving convection terms at the PAR faces of every
 * grid to uo, vo, wo (no immersed boundary: the transport velocity is u, v,
 * w).  Metrics 1-D, reference layout. */
void og_tstle4_par(const og_topo *t, double *uo, double *vo, double *wo,
                   const double *u, const double *v, const double *w,
                   const double *dx, const double *dy, const double *dz,
                   const double *ddx, const double *ddy, const double *ddz,
                   const double *rdx, const double *rdy, const double *rdz,
                   const double *rddx, const double *rddy, const double *rddz);

void og_rkstep(int64_t n, double *p, double *dp, const double *rhsp,
               double frhs, double dtfu);

/* Timed inner-iteration loop used as the CPU baseline: `niter` fused
 * (laplacephi + sipiter1 + conn + sipiter2 + conn) SIP iterations over all
 * grids with `nthreads` OpenMP threads.  Returns wall seconds. */
double og_time_sip(const og_topo *t, const og_conn *c, int nconn, int niter,
                   int nthreads, const double *rhs, double *hilf, double *res,
                   const double *aw, const double *ae, const double *as,
                   const double *an, const double *ab, const double *at,
                   const double *ap, const double *bp, const double *lw,
                   const double *ls, const double *lb, const double *lpr,
                   const double *ue, const double *un, const double *ut,
                   const int64_t *mip, const int64_t *idx);

#ifdef __cplusplus
}
#endif
#endif
