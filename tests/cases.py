"""Seeded inputs shared by the tests: topologies and staggered velocity
fields in the reference layout (see mgp.grids)."""
import numpy as np

from mgp.grids import NOS, stretched_box, uniform_box


def topology(name):
    if name == "self12":        # one self-periodic block (TaylorGreen topology)
        return uniform_box((1, 1, 1), (12, 12, 12), length=(2 * np.pi,) * 3)
    if name == "box2":          # 2x2x2 periodic blocks of 8^3
        return uniform_box((2, 2, 2), (8, 8, 8))
    if name == "aniso":         # uneven block dims, walls in y, periodic x/z
        return uniform_box((2, 2, 3), (8, 6, 10), length=(1.0, 0.5, 1.5),
                           periodic=(True, False, True), wall=NOS)
    if name == "walls":         # closed box
        return uniform_box((2, 2, 2), (6, 8, 6), periodic=(False, False, False))
    if name == "odd":           # odd interior counts, periodic
        return uniform_box((2, 1, 2), (7, 6, 5))
    if name == "stretched":     # geometric spacing along y (walls), periodic x/z
        return stretched_box((2, 2, 2), (6, 8, 6), length=(1.0, 1.0, 1.0), ratio=1.15)
    if name == "c2":            # 8x8x8 periodic blocks of 24^3 (192^3)
        return uniform_box((8, 8, 8), (24, 24, 24))
    raise KeyError(name)


def apply_walls(g, u, v, w):
    """Zero normal velocity on wall faces (what the caller's bound_flow does
    for NOS walls, src/flow/bound_flow_mod.F90): u at the -x wall lives at
    ghost index i=2, at the +x wall at the last interior index ii-2."""
    for b in range(1, g.ngrid + 1):
        bc = g.bc[b - 1]
        for f, (arr, ax) in enumerate(((u, 0), (u, 0), (v, 1), (v, 1), (w, 2), (w, 2))):
            if bc[f] in (7, 19):
                continue
            view = g.block_view(arr, b)
            n = view.shape[ax]
            idx = 1 if f % 2 == 0 else n - 3
            sl = [slice(None)] * 3
            sl[ax] = idx
            view[tuple(sl)] = 0.0


def velocity(g, o, seed, amp=1.0):
    """U(-amp, amp) staggered velocity on interior cells, zero normal
    velocity on walls, ghosts filled by the oracle's 2-layer exchange with
    corners (as the caller's connect does before mgpoisl)."""
    rng = np.random.default_rng(seed)
    inner = g.interior_mask()
    u, v, w = (np.where(inner, rng.uniform(-amp, amp, g.n3d), 0.0) for _ in range(3))
    apply_walls(g, u, v, w)
    o.exchange([u, v, w], layers=2, corners=True)
    return u, v, w


def blocked_mask(g, seed, frac=0.1):
    """int8 mask (1 = blocked) on interior cells, ghosts from the exchange."""
    rng = np.random.default_rng(seed)
    inner = g.interior_mask()
    m = np.where(inner, (rng.random(g.n3d) < frac).astype(np.float64), 0.0)
    return m


def adjoint_case(oracle_mod, name, ibtype, seed):
    """Topology, oracle, oracle state and the sparse operators of the
    projection (oracle.adjoint) for the adjoint checks."""
    from oracle.adjoint import Operators
    g = topology(name)
    o = oracle_mod.Oracle(g)
    bp = None
    if ibtype:
        bp = 1.0 - blocked_mask(g, seed + 100, 0.12)
        o.exchange([bp], layers=2, corners=True)
    st = o.setup(bp=bp)
    return g, o, st, Operators(o, st, ibtype), bp


def two_level(periodic=False, n=8):
    """2x2x2 blocks of n^3 on level 1; the first block refined into 8
    children of n^3 (level 2), whose outer faces inside the domain are PAR."""
    from mgp.grids import refined_box
    return refined_box((2, 2, 2), n, [(n, lambda B: tuple(B["origin"]) == (0, 0, 0))],
                       periodic=(periodic,) * 3)


def centres(g):
    """Cell-centre coordinates (x, y, z) of every stored position."""
    X = np.zeros(g.n3d)
    Y = np.zeros(g.n3d)
    Z = np.zeros(g.n3d)
    for b in range(1, g.ngrid + 1):
        ii, jj, kk = g.ii[b - 1], g.jj[b - 1], g.kk[b - 1]
        x = g.x[g.ip1x[b - 1]:g.ip1x[b - 1] + ii]
        y = g.y[g.ip1y[b - 1]:g.ip1y[b - 1] + jj]
        z = g.z[g.ip1z[b - 1]:g.ip1z[b - 1] + kk]
        g.block_view(X, b)[...] = x[:, None, None]
        g.block_view(Y, b)[...] = y[None, :, None]
        g.block_view(Z, b)[...] = z[None, None, :]
    return X, Y, Z


def three_level():
    """Three levels with mixed block sizes: 2x2x2 blocks of 8^3; the first
    refined into 4^3 children (64, each covering 2^3 parent cells); two
    neighbouring ones refined into one 8^3 child each (covering the whole
    4^3 parent), so the finest level has CON and PAR faces."""
    from mgp.grids import refined_box
    pick = {(4, 4, 4), (8, 4, 4)}
    return refined_box((2, 2, 2), 8,
                       [(4, lambda B: tuple(B["origin"]) == (0, 0, 0)),
                        (8, lambda B: tuple(B["origin"]) in pick)],
                       periodic=(True, True, True))


def sphere_mask(g, radius=0.2, centre=(0.5, 0.5, 0.5)):
    """int8 mask (1 = blocked) from the signed distance to a sphere at the
    cell centres, ghosts included (every block owns its ghost cells' centres)."""
    X, Y, Z = centres(g)
    d = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2) - radius
    return (d < 0).astype(np.int8)


def velocity_levels(g, o, seed, amp=1.0, blocked=None):
    """U(-amp, amp) face velocity on a nested hierarchy whose pressure
    problem is consistent on every level: faces next to blocked cells, faces
    on the planes of PAR faces and parent faces between covered and
    uncovered parent cells carry no flux, the parent faces under a refined
    grid hold the restriction of the fine faces (ftoc), and the ghosts of
    every level are exchanged."""
    rng = np.random.default_rng(seed)
    inner = g.interior_mask()
    u, v, w = (np.where(inner, rng.uniform(-amp, amp, g.n3d), 0.0) for _ in range(3))
    apply_walls(g, u, v, w)
    if blocked is not None:
        close_blocked_faces(g, blocked, u, v, w)
    covered = g.zeros()
    for b in range(1, g.ngrid + 1):
        if g.iparent[b - 1]:
            n = (g.ii[b - 1] - 4, g.jj[b - 1] - 4, g.kk[b - 1] - 4)
            p0 = g.position[b - 1] - 1
            g.block_view(covered, g.iparent[b - 1])[
                p0[0]:p0[0] + n[0] // 2, p0[1]:p0[1] + n[1] // 2, p0[2]:p0[2] + n[2] // 2] = 1.0
    for lvl in range(g.minlevel, g.maxlevel + 1):
        o.exchange([covered], level=lvl, layers=2, corners=True)
    for b in range(1, g.ngrid + 1):
        c = g.block_view(covered, b)
        for f, (arr, ax) in enumerate(((u, 0), (u, 0), (v, 1), (v, 1), (w, 2), (w, 2))):
            a = g.block_view(arr, b)
            if f % 2 == 0:
                sl0, sl1 = [slice(None)] * 3, [slice(None)] * 3
                sl0[ax], sl1[ax] = slice(0, -1), slice(1, None)
                a[tuple(sl0)] *= (c[tuple(sl0)] == c[tuple(sl1)])
            if g.bc[b - 1, f] == 8:
                sl = [slice(None)] * 3
                sl[ax] = 1 if f % 2 == 0 else a.shape[ax] - 3
                a[tuple(sl)] = 0.0
    p = g.zeros()
    for lvl in range(g.maxlevel, g.minlevel, -1):
        o.ftoc_uvwp(lvl, u, v, w, p)
    for lvl in range(g.minlevel, g.maxlevel + 1):
        o.exchange([u, v, w], level=lvl, layers=2, corners=True)
    return u, v, w


def config3(width=0.25):
    """Multi-level batch with an immersed sphere: a level-1 box of 4x4x4
    blocks of 24^3 on the unit cube (periodic), refined everywhere into
    16^3 blocks (level 2), and around the sphere of radius 0.2 at the centre
    into 32^3 blocks (level 3, one per level-2 block whose centre lies
    within `width` of the sphere surface).  The default width gives 2464
    blocks (64 of 24^3, 1728 of 16^3, 672 of 32^3; 29,982,720 cells)."""
    from mgp.grids import refined_box
    shell = lambda B: abs(np.linalg.norm(B["centre"] - 0.5) - 0.2) < width
    return refined_box((4, 4, 4), 24, [(16, lambda B: True), (32, shell)])


def close_blocked_faces(g, blocked, u, v, w):
    """Zero the face velocities next to blocked cells (the immersed
    boundary's no-through-flow condition)."""
    fluid = 1.0 - blocked.astype(np.float64)
    for b in range(1, g.ngrid + 1):
        f = g.block_view(fluid, b)
        for arr, ax in ((u, 0), (v, 1), (w, 2)):
            a = g.block_view(arr, b)
            sl0 = [slice(None)] * 3
            sl1 = [slice(None)] * 3
            sl0[ax] = slice(0, -1)
            sl1[ax] = slice(1, None)
            open_ = f[tuple(sl0)] * f[tuple(sl1)]
            a[tuple(sl0)] *= open_


def mask_levels(g, o, blocked):
    """BP of a nested hierarchy as the immersed-boundary setup leaves it
    (src/ib/gc_blockbp_mod.F90:299): the finest cells from the geometry,
    parent cells under a refined grid fluid if any of their 8 children is
    (ftoc 'E', the max over each 2x2x2 group, src/ib/gc_restrict_mod.F90:
    289-320), ghosts exchanged per level.  Returns (bp, blocked int8)."""
    bp = 1.0 - blocked.astype(np.float64)
    for lvl in range(g.maxlevel, g.minlevel, -1):
        for b in range(1, g.ngrid + 1):
            if g.level[b - 1] != lvl:
                continue
            f = g.block_view(bp, b)[2:-2, 2:-2, 2:-2]
            n = f.shape
            m = f.reshape(n[0] // 2, 2, n[1] // 2, 2, n[2] // 2, 2).max(axis=(1, 3, 5))
            pv = g.block_view(bp, g.iparent[b - 1])
            p0 = g.position[b - 1] - 1
            pv[p0[0]:p0[0] + n[0] // 2, p0[1]:p0[1] + n[1] // 2, p0[2]:p0[2] + n[2] // 2] = m
    for lvl in range(g.minlevel, g.maxlevel + 1):
        o.exchange([bp], level=lvl, layers=2, corners=True)
        # ghost layers of PAR faces one level finer: the parent's cells
        # across the face (piecewise constant, as parent prolongs them)
        for b in range(1, g.ngrid + 1):
            if g.level[b - 1] != lvl + 1:
                continue
            fv, pv = g.block_view(bp, b), g.block_view(bp, g.iparent[b - 1])
            p0 = g.position[b - 1] - 1
            n = fv.shape
            for f in range(6):
                if g.bc[b - 1, f] != 8:
                    continue
                ax = f // 2
                idx = [np.arange(2, n[a] - 2) for a in range(3)]
                cidx = [p0[a] + (idx[a] - 2) // 2 for a in range(3)]
                for layer in (0, 1):
                    idx[ax] = np.array([1 - layer if f % 2 == 0 else n[ax] - 2 + layer])
                    cidx[ax] = np.array([p0[ax] - 1 if f % 2 == 0
                                         else p0[ax] + (n[ax] - 4) // 2])
                    fv[np.ix_(*idx)] = pv[np.ix_(*cidx)]
    return bp, (bp == 0.0).astype(np.int8)


# ---------------------------------------------------------------- reference topologies
BC_CODES = {"FIX": 2, "OP1": 3, "NOS": 5, "SLI": 6, "CON": 7, "PAR": 8, "GRA": 15,
            "SCA": 16, "REF": 17, "NRE": 18, "CO1": 19, "NIX": 99}
# cases of the reference's tests/*/grids.h5 whose blocks the plane kernels
# take (every interior count 1..64, at most 65536 cells, (ni+2)(nj+2) <= 4096,
# face region <= 16384; sphere's 40^3 blocks need 87 KB of LDS per workgroup
# in FP64, which the MI355X's 160 KB allows; BlasiusBL, Inflow-ramp and
# NaturalConvection are one cell thick along y, or along y and z)
SUPPORTED_TOPOLOGIES = ["BlasiusBL", "Checkblock", "GeostrophicWind", "HotAndColdSpheres",
                        "Inflow-ramp", "NaturalConvection", "Parker_micro", "Santarelli",
                        "ScalarDecay", "Statistics", "TaylorGreen", "sphere"]


def golden_topologies():
    import json
    from pathlib import Path
    with open(Path(__file__).resolve().parent / "golden" / "topologies.json") as fh:
        return json.load(fh)


def reference_topology(name):
    """Grids of the reference's test case `name` (tests/<name>/grids.h5 as
    kept in tests/golden/topologies.json): dims, levels, parents, positions,
    NBRGRID and the first boundary condition of each face (the one the
    pressure path reads, itypboconds(1, iface, igrid)); uniform spacing per
    block from its bounding box."""
    from mgp.grids import Grids
    recs = golden_topologies()[name]["grids"]
    assert [r["IGRID"] for r in recs] == list(range(1, len(recs) + 1))
    dims = [(r["II"] - 4, r["JJ"] - 4, r["KK"] - 4) for r in recs]
    box = np.array([r["BBOX"] for r in recs]).reshape(-1, 3, 2)
    spacing = (box[:, :, 1] - box[:, :, 0]) / np.array(dims)
    levels = [r["LEVEL"] for r in recs]
    iparent = [max(r["IPARENT"], 0) for r in recs]
    position = [(max(r["IPOSITION"], 0), max(r["JPOSITION"], 0), max(r["KPOSITION"], 0))
                for r in recs]
    bc = [[BC_CODES[t] for t in r["BC"]] for r in recs]
    return Grids(dims, levels, [r["NBRGRID"] for r in recs], bc, spacing, box[:, :, 0],
                 iparent=iparent, position=position)


def lattice_ids(g):
    """Per stored position: 1 + X + NX*(Y + NY*Z), (X, Y, Z) the cell's index
    on its level's global lattice, wrapped periodically over the domain the
    coarsest level spans; interior cells hold it, ghosts hold 0.  Returns
    (field, expected) where expected holds the id at every position."""
    lo = np.array([g.origin[b] for b in range(g.ngrid) if g.level[b] == g.minlevel]).min(axis=0)
    hi = np.array([g.origin[b] + g.spacing[b] * (g.ii[b] - 4, g.jj[b] - 4, g.kk[b] - 4)
                   for b in range(g.ngrid) if g.level[b] == g.minlevel]).max(axis=0)
    field, expect = g.zeros(), g.zeros()
    for b in range(1, g.ngrid + 1):
        h = g.spacing[b - 1]
        N = np.rint((hi - lo) / h).astype(np.int64)
        o = np.rint((g.origin[b - 1] - lo) / h).astype(np.int64)
        n = (g.ii[b - 1], g.jj[b - 1], g.kk[b - 1])
        X, Y, Z = ((o[a] + np.arange(n[a]) - 2) % N[a] for a in range(3))
        ids = 1 + X[:, None, None] + N[0] * (Y[None, :, None] + N[1] * Z[None, None, :])
        g.block_view(expect, b)[...] = ids
        g.block_view(field, b)[2:-2, 2:-2, 2:-2] = ids[2:-2, 2:-2, 2:-2]
    return field, expect
