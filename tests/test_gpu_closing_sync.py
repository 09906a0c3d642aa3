"""GPU parity of mgpoisl's closing sync (src/flow/pressuresolver_mod.F90:
338-348): per level the parent of U, V, W into the PAR face buffers,
bound_flow at FIX, OP1, NOS, SLI and PAR faces (src/flow/bound_flow_mod.F90)
and the 2-layer exchange.  U, V, W, P after the call, ghosts included, are
bitwise the oracle's, which runs the three steps in the reference order."""
import numpy as np
import pytest

from cases import BC_CODES, mask_levels, sphere_mask, three_level, topology, two_level, \
    velocity, velocity_levels

pytestmark = pytest.mark.gpu

# wall type per face of the domain boundary: -x, +x, -y, +y, -z, +z
MIXED = ("FIX", "OP1", "NOS", "SLI", "OP1", "FIX")


def _set_walls(g, kinds):
    """Boundary faces that are neither CON nor PAR get the given codes."""
    keep = (BC_CODES["CON"], BC_CODES["PAR"], 19)
    for b in range(g.ngrid):
        for f in range(6):
            if g.bc[b, f] not in keep:
                g.bc[b, f] = BC_CODES[kinds[f]]


def _bound_values(o, g, seed):
    rng = np.random.default_rng(seed)
    return {"u": rng.uniform(-1, 1, max(o.nbb, 1)), "v": rng.uniform(-1, 1, max(o.nbb, 1)),
            "w": rng.uniform(-1, 1, max(o.nbb, 1)), "pinf": rng.uniform(-1, 1, 6 * g.ngrid)}


def _compare(oracle_mod, g, seed, levels, bound_seed=None, sync_by_caller=False,
             blocked=None, ibtype=0, ninner=3, nouter=3):
    import mgp
    o = oracle_mod.Oracle(g)
    bp = None
    if blocked is not None:
        bp, blocked = mask_levels(g, o, blocked)
    st = o.setup(bp=bp)
    if levels:
        u, v, w = velocity_levels(g, o, seed, blocked=blocked)
    else:
        u, v, w = velocity(g, o, seed)
    p = np.zeros(g.n3d)
    ins = [a.copy() for a in (u, v, w, p)]
    bound = _bound_values(o, g, bound_seed) if bound_seed is not None else None
    so = o.mgpoisl(st, u, v, w, p, 1e-3, ninner=ninner, nouter=nouter, ibtype=ibtype,
                   sync_by_caller=int(sync_by_caller), bound=bound)
    s = mgp.PressureSolver(g, ninner=ninner, nouter=nouter, bp=bp,
                           ib={0: "noib", 1: "ghostcell"}[ibtype],
                           sync_by_caller=sync_by_caller)
    off, n = s.bound_layout()
    assert n == o.nbb and np.array_equal(off, o.bbo)
    if bound is not None:
        s.set_bound(bound["u"][:n], bound["v"][:n], bound["w"][:n], bound["pinf"])
    sg = s.mgpoisl(*ins, 1e-3)
    assert sg["maxrhs"] == so["maxrhs"]
    for a, b, nm in zip((u, v, w, p), ins, "uvwp"):
        assert np.array_equal(a, b), (nm, np.abs(a - b).max())
    s.close()
    return (u, v, w, p)


def test_mixed_wall_types_bitwise(oracle_mod):
    """FIX, OP1, NOS and SLI faces on one closed box with prescribed
    boundary velocities and OP1 pressures."""
    g = topology("walls")
    _set_walls(g, MIXED)
    _compare(oracle_mod, g, 301, levels=False, bound_seed=302)


@pytest.mark.parametrize("name", ["walls", "aniso"])
def test_nos_walls_bitwise(oracle_mod, name):
    _compare(oracle_mod, topology(name), 303, levels=False)


@pytest.mark.parametrize("kinds", [None, MIXED])
def test_two_level_par_and_walls_bitwise(oracle_mod, kinds):
    """PAR faces take U, V, W from the parent (staggered prolongation, flux
    correction of the normal component), meeting FIX / OP1 / SLI / NOS
    faces at their edges."""
    g = two_level(periodic=False)
    if kinds is not None:
        _set_walls(g, kinds)
    _compare(oracle_mod, g, 304, levels=True, bound_seed=305 if kinds else None)


def test_three_level_par_bitwise(oracle_mod):
    _compare(oracle_mod, three_level(), 306, levels=True)


def test_two_level_masked_par_bitwise(oracle_mod):
    """Blocked cells next to the PAR faces: the BP weights of the normal
    flux correction and of the tangential extrapolation."""
    g = two_level(periodic=False)
    blocked = sphere_mask(g, radius=0.22, centre=(0.5, 0.5, 0.5))
    _compare(oracle_mod, g, 307, levels=True, blocked=blocked, ibtype=1)


def test_sync_by_caller_stops_after_ftoc(oracle_mod):
    """sync_by_caller: both stop after ftoc (bitwise), and the closing sync
    does change the ghost values (the comparison above is not vacuous)."""
    g = two_level(periodic=False)
    _set_walls(g, MIXED)
    a = _compare(oracle_mod, g, 308, levels=True, sync_by_caller=True)
    g = two_level(periodic=False)
    _set_walls(g, MIXED)
    b = _compare(oracle_mod, g, 308, levels=True, bound_seed=309)
    assert any(not np.array_equal(x, y) for x, y in zip(a, b))
