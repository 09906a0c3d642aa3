/*
 * mgp.h -- C ABI of the MI355X pressure-projection library (libmgp.so).
 *
 * The library replaces the kernels behind MGLET-base's Fortran module
 * `pressuresolver_mod` (src/flow/pressuresolver_mod.F90:45, PUBLIC
 * init_pressuresolver / mgpoisl / finish_pressuresolver) together with the
 * coefficient setup of `ib%giteig` (src/ib/noib_mod.F90:81-250) and the
 * intra-rank ghost-layer copies `conn` / `connect`
 * (src/core/connect/conn1_mod.F90:31-129, connect2_mod.F90:75-138).
 * The Fortran side binds these symbols with ISO_C_BINDING BIND(C)
 * (mglet-base_amd/fortran/mgp_pressuresolver_mod.F90).
 *
 * Conventions (mirroring src/core/mglet_precision.h:12-39 with
 * _MGLET_DOUBLE_PRECISION_ and 32-bit integers):
 *   - every entry point returns an int status, 0 = success (the reference's
 *     C helpers report errors through an integer, e.g.
 *     src/core/jsoncpp_wrapper.h:10-33); mgp_last_error() explains it;
 *   - 3-D fields use the reference layout: one flat array per field, block
 *     igrid starting at ip3[igrid-1] (1-based, as ip3d in
 *     src/core/pointers_mod.F90:48-92), dims kk x jj x ii, k fastest,
 *     two ghost layers per side; 1-D metric fields start at ip1x/y/z;
 *   - the caller owns every array it passes; the library owns its device
 *     state and never frees caller memory.
 */
#ifndef MGP_H
#define MGP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double mgletreal;
typedef int32_t mgletint;

enum {
    MGP_OK = 0,
    MGP_ERR_ARG = 1,       /* invalid argument or configuration            */
    MGP_ERR_HIP = 2,       /* HIP runtime error (message in last_error)    */
    MGP_ERR_TOPOLOGY = 3,  /* unsupported or inconsistent block topology   */
    MGP_ERR_NONFINITE = 4, /* residual norm became NaN/Inf                 */
    MGP_ERR_STATE = 5      /* call out of order (e.g. solve before init)   */
};

/* Block topology: GRIDINFO subset (src/core/gridio_mod.F90:15-37),
 * itypboconds(1, iface, igrid) (src/core/grids_mod.F90:253-290) and the
 * pointer offsets (src/core/pointers_mod.F90:23-103), all per grid. */
typedef struct mgp_grids {
    mgletint ngrid;
    mgletint minlevel, maxlevel;
    const mgletint *kk, *jj, *ii;   /* [ngrid] dims incl. ghost layers     */
    const mgletint *level;          /* [ngrid]                              */
    const mgletint *nbrgrid;        /* [ngrid*26] neighbour ids, 0 = none   */
    const mgletint *itypbc;         /* [ngrid*6] first BC code per face     */
    const int64_t *ip3;             /* [ngrid] 1-based 3-D offsets          */
    const int64_t *ip1x, *ip1y, *ip1z; /* [ngrid] 1-based 1-D offsets       */
    const mgletint *iparent;        /* [ngrid] parent grid id, 0 = none, or */
                                    /*   NULL for a single level            */
    const mgletint *position;       /* [ngrid*3] iposition, jposition,      */
                                    /*   kposition in the parent (GRIDINFO) */
} mgp_grids;

/* /flow/pressuresolver keys (src/flow/pressuresolver_mod.F90:57-142) plus
 * the build's own switches. */
typedef struct mgp_config {
    mgletint ityp;        /* 0 "sip" (hyperplane order), 1 "sor", 2 "classic" */
    mgletint ninner;      /* inner iterations per outer iteration (>= 0)    */
    mgletint nouter;      /* maximum outer iterations (>= 0)                */
    mgletint nouter_min;  /* outer iterations before the exit test (>= 0)   */
    mgletreal epcorr;     /* exit when maxrhs / (rho/dt) < epcorr           */
    mgletreal omega;      /* SOR relaxation factor, 0 < omega < 2           */
    mgletreal rho;        /* density (src/flow/flowcore_mod.F90:43-98)      */
    mgletint ibtype;      /* 0 "noib", 1 "ghostcell" (divcal x BP, masked   */
                          /*   velocity correction)                          */
    mgletint precision;   /* 0 all FP64; 1 FP32 inner sweeps inside the     */
                          /*   FP64 defect-correction (outer) loop          */
    mgletreal eps_rel;    /* > 0: exit when maxrhs <= eps_rel * maxrhs at   */
                          /*   iteration 0 instead of the epcorr test       */
    mgletint sor_all_levels; /* 1: "sor" also relaxes the coarsest level    */
    mgletint check_every; /* outer iterations per captured graph (>= 1)     */
    mgletreal alfa;       /* SIP alfa, 0 <= alfa <= 1; 0 selects the        */
                          /*   reference's hard-coded default-real 0.92,    */
                          /*   (double)0.92f (src/flow/sip_hyperplane_mod.  */
                          /*   F90:123, 202); a negative value gives 0      */
    mgletint sync_by_caller; /* 0: mgpoisl ends with the reference's       */
                          /*   closing sync per level (parent of U, V, W,   */
                          /*   bound_flow, connect; src/flow/              */
                          /*   pressuresolver_mod.F90:338-348); 1: it stops */
                          /*   after ftoc and the caller runs them          */
} mgp_config;

typedef struct mgp_stats {
    mgletint ipc;         /* inner iterations run (reference `ipc`)         */
    mgletint nouter_done; /* outer iterations run                           */
    mgletint converged;   /* 1 if the exit test fired before nouter          */
    mgletreal maxrhs0;    /* level-weighted max |RHS| before iterating      */
    mgletreal maxrhs;     /* the same after the last outer iteration        */
    double solve_ms;      /* device time of the outer-iteration loop        */
} mgp_stats;

typedef struct mgp_handle mgp_handle;

/* init_pressuresolver (src/flow/pressuresolver_mod.F90:47-145) together
 * with ib%giteig (src/ib/noib_mod.F90:81-250): builds the connection
 * tables, the Poisson coefficients and the SIP factors on the GPU.
 * dx..ddz are the 1-D metric fields DX, DY, DZ, DDX, DDY, DDZ; bp is the
 * 3-D BP field (1 fluid, 0 blocked) or NULL for all-fluid.  Host memory. */
int mgp_init(const mgp_grids *grids, const mgp_config *cfg,
             const mgletreal *dx, const mgletreal *dy, const mgletreal *dz,
             const mgletreal *ddx, const mgletreal *ddy, const mgletreal *ddz,
             const mgletreal *bp, mgp_handle **out);

/* Same as mgp_init with the BP field given as an int8 mask per cell of the
 * 3-D layout: 1 = blocked (BP = 0), 0 = fluid (BP = 1). */
int mgp_init_mask(const mgp_grids *grids, const mgp_config *cfg,
                  const mgletreal *dx, const mgletreal *dy,
                  const mgletreal *dz, const mgletreal *ddx,
                  const mgletreal *ddy, const mgletreal *ddz,
                  const int8_t *blocked, mgp_handle **out);

/* The ghostcell divergence source SDIV (3-D field, reference layout, host
 * memory), which the ghostcell divcal adds as fak*SDIV before the BP factor
 * (src/ib/gc_mod.F90:278-347; the caller computes it from its STL cut cells,
 * src/flow/gc_flowstencils_mod.F90:1486-1530).  NULL resets it to zero, the
 * value without cut cells.  Only for handles with ibtype 1. */
int mgp_set_sdiv(mgp_handle *h, const mgletreal *sdiv);

/* Velocity boundary buffers that bound_flow reads (the field buffers at
 * ipbb, src/core/pointers_mod.F90:64-92, read by f1%get_buffer in
 * src/flow/bound_flow_mod.F90:62, 298, 536): one kk*jj (faces 1, 2), kk*ii
 * (3, 4) or jj*ii (5, 6) array per FIX, OP1 or PAR face, first index
 * fastest, faces in grid and face order.  offsets[6*(igrid-1) + iface-1]
 * (may be NULL) gets each face's start, -1 for faces without a buffer;
 * *n the total length. */
int mgp_bound_layout(mgp_handle *h, int64_t *offsets, int64_t *n);

/* The caller's FIX / OP1 boundary values in that layout (host memory, the
 * values U%buffers, V%buffers, W%buffers hold) and the OP1 pressure per
 * grid and face (pinf of get_bcprms, src/flow/bound_flow_mod.F90:229-231;
 * [ngrid*6]).  NULL sets an array to zero (the start value).  PAR entries
 * are rewritten by the closing parent of every mgpoisl. */
int mgp_set_bound(mgp_handle *h, const mgletreal *ubuf, const mgletreal *vbuf,
                  const mgletreal *wbuf, const mgletreal *pinf);

/* mgpoisl(u, v, w, p, dt, ittot, irk) (src/flow/pressuresolver_mod.F90:
 * 160-357): projection of host arrays u, v, w, p in place. */
int mgp_mgpoisl(mgp_handle *h, mgletreal *u, mgletreal *v, mgletreal *w,
                mgletreal *p, mgletreal dt, mgletint ittot, mgletint irk,
                mgp_stats *stats);

/* The same with u, v, w, p already resident in device memory. */
int mgp_mgpoisl_device(mgp_handle *h, mgletreal *u, mgletreal *v,
                       mgletreal *w, mgletreal *p, mgletreal dt,
                       mgletint ittot, mgletint irk, mgp_stats *stats);

/* Adjoint of the projection.  No counterpart in the reference: for
 * differentiable callers (gradient-based optimisation and data assimilation
 * through the time loop).  mgpoisl maps the interior face velocities u to
 * u' = u - G A^-1 D u (D: ib%divcal, src/ib/noib_mod.F90:304-354; A: the
 * GSA* operator; G: mgpcorr, src/flow/pressuresolver_mod.F90:787-879).  On
 * entry ubar, vbar, wbar hold dJ/du' on interior faces, on exit dJ/du =
 * ubar - D^T A^-T G^T ubar.  The adjoint Poisson problem A^T y = G^T ubar is
 * solved with the handle's solver settings and the transposed SIP factors
 * (U^T, then L^T).  Ghost layers and faces on block boundaries without a
 * connected neighbour (the caller's boundary conditions) are zero on exit.
 * One grid level only; overwrites the state read by mgp_get_field("DP"). */
int mgp_adjoint(mgp_handle *h, mgletreal *ubar, mgletreal *vbar,
                mgletreal *wbar, mgletreal dt, mgp_stats *stats);

/* The same with ubar, vbar, wbar resident in device memory. */
int mgp_adjoint_device(mgp_handle *h, mgletreal *ubar, mgletreal *vbar,
                       mgletreal *wbar, mgletreal dt, mgp_stats *stats);

/* finish_pressuresolver (src/flow/pressuresolver_mod.F90:148-157). */
int mgp_finish(mgp_handle *h);

/* Text of the last error on this thread ("" if none). */
const char *mgp_last_error(void);

/* Change ninner / nouter / nouter_min / epcorr / eps_rel / check_every of
 * an initialised handle (the factors are kept). */
int mgp_set_iterations(mgp_handle *h, mgletint ninner, mgletint nouter,
                       mgletint nouter_min, mgletreal epcorr,
                       mgletreal eps_rel, mgletint check_every);

/* Use `stream` (a hipStream_t, default priority) for all work of h;
 * NULL restores the handle's own stream. */
int mgp_set_stream(mgp_handle *h, void *stream);

/* Copy a library field to host memory in the reference layout.  Names:
 * "GSAP", "BP", "SIPLW", "SIPLS", "SIPLB", "SIPLPR", "SIPUE", "SIPUN",
 * "SIPUT" (the factors in hyperplane access order, as
 * src/flow/sip_hyperplane_mod.F90:224-230 stores them, or spatial order
 * for ityp 2), "DP", "RHS" (state after the last mgpoisl), "DPG" (DP with
 * the ghost layer mgpcorr reads: exchanged, bound_pressure at PAR faces,
 * accumulated injected values elsewhere).  1-D fields
 * "GSAW", "GSAE", "GSAS", "GSAN", "GSAB", "GSAT" use the 1-D layout. */
int mgp_get_field(mgp_handle *h, const char *name, mgletreal *out);

/* The same for one grid (igrid 1-based; 0 = every grid, as mgp_get_field):
 * out gets the kk x jj x ii values of the grid, the block the reference's
 * field_t%get_ptr(arr, igrid) points at (src/core/field_t/realfield_mod.F90:
 * 113-135).  3-D fields only. */
int mgp_get_field_grid(mgp_handle *h, const char *name, mgletint igrid,
                       mgletreal *out);

/* Ghost-layer exchange of device fields in the reference layout
 * (conn / connect: src/core/connect/conn1_mod.F90:31-129,
 * connect2_mod.F90:75-138).  level < 0: all levels; layers 1 or 2;
 * corners 0/1; forward 0, 1 or -1. */
int mgp_connect_device(mgp_handle *h, mgletint level, mgletint layers,
                       mgletint corners, mgletint forward,
                       mgletreal **fields, mgletint nfield);

/* ib%divcal (src/ib/noib_mod.F90:253-354): div = fak * div(u, v, w) over
 * the interior of device fields (+ fak*SDIV, then x BP, for ibtype 1). */
int mgp_divcal_device(mgp_handle *h, mgletreal *div, const mgletreal *u,
                      const mgletreal *v, const mgletreal *w,
                      mgletreal fak);

/* mgp_divcal_device without allocating: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * intermediate divergence lives in a static device array of 8,388,608 FP64
 * values (64 MiB, one per GPU, held until the process ends); a handle with
 * more interior cells gets MGP_ERR_ARG.  Every call on a GPU shares that
 * array: calls that overlap on one GPU (other streams, other threads)
 * race. */
int mgp_divcal_device_noalloc(mgp_handle *h, mgletreal *div, const mgletreal *u,
                              const mgletreal *v, const mgletreal *w, mgletreal fak);

/* The velocity boundary conditions of one level (device fields, reference
 * layout): parent (src/ib/parent/parent_mod.F90:16-52; prolong1/prolong2
 * with the staggered positions, src/ib/parent/parent1_mod.F90:515-583)
 * fills the library's PAR face buffers of the level's refined grids from
 * their parents' U, V, W, then bound_flow%bound (src/flow/bound_flow_mod.F90)
 * sets the ghost values at FIX, OP1, NOS, SLI and PAR faces (FIX / OP1
 * values and OP1 pressures from mgp_set_bound).  The pair the reference
 * runs per level before the divergence of an RK stage, after the normal
 * exchange (src/flow/timeintegration_mod.F90:119-123), and inside
 * mgpoisl's closing sync, before the two-layer exchange
 * (src/flow/pressuresolver_mod.F90:338-348). */
int mgp_flow_bc_device(mgp_handle *h, mgletint level, mgletreal *u, mgletreal *v,
                       mgletreal *w, mgletreal *p);

/* par_ftoc_norm(ilevel, u, v, w, sum) (src/ib/par_ftoc_mod.F90:35-475) on
 * device fields: the face-normal velocity at the PAR faces of the level's
 * grids, per 2x2 group of fine faces the area-weighted mean (sum = 0) or
 * the plain sum (sum != 0), into the coincident coarse face positions of
 * the parent, and of the parent's neighbour where the face lies on the
 * parent's boundary.  Called from the ghostcell flux setup
 * (src/flow/gc_flowstencils_mod.F90:1420) after ftoc, fine to coarse. */
int mgp_par_ftoc_device(mgp_handle *h, mgletint level, mgletreal *u, mgletreal *v,
                        mgletreal *w, mgletint sum);

/* The RK stage around the projection (device fields, reference layout).
 *
 * tstle4 (src/flow/tstle4_mod.F90:13-120, no immersed boundary): uo, vo, wo
 * = convection + diffusion with viscosity field g (gmol the molecular
 * viscosity in its harmonic means; ilesmodel != 0 adds the transpose part
 * of the stress) - grad p / rho - gradp[3] / rho (gradp may be NULL), zero
 * outside the loop ranges; next to NOS walls the Werner-Wengle wall stress
 * of swcle3d (:1756-1845) is subtracted from the tangential components.
 * rho is the handle's.  At the PAR faces of refined grids tstle4_par
 * (:532-1753) adds its upwind and momentum-conserving convection terms.
 * Replaces the call at src/flow/timeintegration_mod.F90:98. */
int mgp_tstle4_device(mgp_handle *h, mgletreal *uo, mgletreal *vo, mgletreal *wo,
                      const mgletreal *u, const mgletreal *v, const mgletreal *w,
                      const mgletreal *p, const mgletreal *g, mgletreal gmol,
                      mgletint ilesmodel, const mgletreal *gradp);

/* rkstep (src/core/rungekutta_mod.F90:315-333) over every stored position
 * of the handle's 3-D fields: dp = frhs dp + rhsp, p = p + dtfu dp.
 * Replaces the calls at src/flow/timeintegration_mod.F90:105-107. */
int mgp_rkstep_device(mgp_handle *h, mgletreal *p, mgletreal *dp,
                      const mgletreal *rhsp, mgletreal frhs, mgletreal dtfu);

/* Times `niter` inner SIP iterations on the current RHS with HIP events
 * around each kernel; writes the mean kernel durations in ms:
 * ms[0] forward sweep, ms[1] backward sweep, ms[2] one outer update. */
int mgp_time_kernels(mgp_handle *h, mgletint niter, double *ms);

/* Pipeline shape of the plane kernels (the sweeps, SOR and the outer
 * update): 1 = four 64-cell items of loads in flight per wave, 2 = two
 * 128-cell items, 3 = three 128-cell items, 4 = six 128-cell items, 5 =
 * eight 64-cell items, all one wave per block; 6 and 7 = four 64-cell
 * items per wave with 4 and 8 waves sharing each block; 8 and 9 = three
 * 128-cell and four 64-cell items per wave with 2 waves sharing each block
 * (SOR keeps one wave).  request -1 only reports; 0 runs the autotuner:
 * it times one inner iteration per shape (ninner x forward + backward, or
 * SOR, plus one outer update) on h's own blocks and keeps the fastest;
 * 1..MGP_NUM_PLAN_SHAPES sets the shape.  Results do not depend on the
 * shape.  chosen gets the shape in use, ms[MGP_NUM_PLAN_SHAPES] (may be
 * NULL) the timed ms per shape with every kernel in that shape (0 when not
 * timed).  The autotuner gives the backward sweep, timed on its own, the
 * shape fastest for it (mgp_plan_shape_backward); a request of 1..N sets
 * both.  mgp_init autotunes unless MGP_PLAN_SHAPE=1..MGP_NUM_PLAN_SHAPES
 * is set. */
#define MGP_NUM_PLAN_SHAPES 9
int mgp_plan_shape(mgp_handle *h, mgletint request, mgletint *chosen,
                   double *ms);

/* The backward sweep's pipeline shape in use (a tuning query with no
 * counterpart in the reference). */
int mgp_plan_shape_backward(mgp_handle *h, mgletint *chosen);

/* Level-weighted max |RHS| after each outer iteration of the last mgpoisl
 * call (the quantity of the exit test, src/flow/pressuresolver_mod.F90:
 * 274-301); writes min(count, cap) values and the count. */
int mgp_get_history(mgp_handle *h, mgletreal *out, mgletint cap,
                    mgletint *count);

/* Number of interior cells and bytes of device memory held by h. */
int mgp_info(mgp_handle *h, int64_t *ncells, int64_t *device_bytes);

/* Shared SIP factor storage (no reference counterpart; the reference keeps
 * LW..UT per grid, src/flow/sip_hyperplane_mod.F90:106-257): mgp_init
 * groups the blocks whose factors (and BP) are equal bit for bit and keeps
 * one copy per group.  nclasses gets the number of groups (= the number of
 * grids when nothing is shared or MGP_SHARED_FACTORS=0 was set at init),
 * ncells_stored the factor cells kept per factor array. */
int mgp_factor_classes(mgp_handle *h, mgletint *nclasses, int64_t *ncells_stored);

#ifdef __cplusplus
}
#endif
#endif /* MGP_H */
