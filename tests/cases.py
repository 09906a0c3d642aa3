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
