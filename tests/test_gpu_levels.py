"""GPU parity on nested grid levels (ctof, parent, bound_pressure, ftoc;
src/flow/pressuresolver_mod.F90:160-420): bitwise FP64 equality with the
oracle at a fixed iteration count, and the FP32-refined form converged."""
import numpy as np
import pytest

from cases import blocked_mask, mask_levels, sphere_mask, three_level, two_level, velocity_levels

pytestmark = pytest.mark.gpu


def _run(oracle_mod, g, seed, ityp=0, ninner=3, nouter=3, ibtype=0, blocked=None,
         eps_rel=0.0, precision="fp64", omega=1.1, sdiv_seed=None):
    import mgp
    o = oracle_mod.Oracle(g)
    bp = None
    if blocked is not None:
        bp, blocked = mask_levels(g, o, blocked)
    st = o.setup(bp=bp, ityp=ityp)
    u, v, w = velocity_levels(g, o, seed, blocked=blocked)
    p = np.zeros(g.n3d)
    ins = [a.copy() for a in (u, v, w, p)]
    sdiv = None
    if sdiv_seed is not None:    # geometry source on interior cells of every level
        rng = np.random.default_rng(sdiv_seed)
        sdiv = np.where(g.interior_mask(), rng.standard_normal(g.n3d), 0.0)
    so = o.mgpoisl(st, u, v, w, p, 1e-3, ityp=ityp, ninner=ninner, nouter=nouter,
                   eps_rel=eps_rel, ibtype=ibtype, omega=omega, sdiv=sdiv)
    s = mgp.PressureSolver(g, type={0: "sip", 1: "sor", 2: "classic"}[ityp], ninner=ninner,
                           nouter=nouter, eps_rel=eps_rel, omega=omega,
                           ib={0: "noib", 1: "ghostcell"}[ibtype], precision=precision,
                           bp=bp)
    if sdiv is not None:
        s.set_sdiv(sdiv)
    sg = s.mgpoisl(*ins, 1e-3)
    return so, sg, s, (u, v, w, p), ins


def _equal(so, sg, s, ref, got):
    assert sg["maxrhs0"] == so["maxrhs0"]
    assert sg["maxrhs"] == so["maxrhs"]
    assert np.array_equal(s.history(), so["history"])
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("ityp", [0, 2])
def test_two_level_bitwise(oracle_mod, periodic, ityp):
    g = two_level(periodic=periodic)
    so, sg, s, ref, got = _run(oracle_mod, g, 81, ityp=ityp)
    _equal(so, sg, s, ref, got)


def test_two_level_sor_on_the_fine_level(oracle_mod):
    """type "sor": SOR on the refined level, SIP on the coarsest
    (src/flow/pressuresolver_mod.F90:403-412)."""
    g = two_level()
    so, sg, s, ref, got = _run(oracle_mod, g, 82, ityp=1, ninner=4, omega=1.2)
    _equal(so, sg, s, ref, got)


def test_two_level_masked_ghostcell_bitwise(oracle_mod):
    """Random blocked cells with BP consistent across levels (a parent cell
    is fluid when any child is); ghostcell restriction gc_restrict."""
    g = two_level(periodic=True)
    so, sg, s, ref, got = _run(oracle_mod, g, 84, ibtype=1,
                               blocked=blocked_mask(g, 83, 0.12).astype(np.int8))
    _equal(so, sg, s, ref, got)


@pytest.mark.parametrize("ninner", [1, 2, 5])
def test_three_level_mixed_sizes_bitwise(oracle_mod, ninner):
    g = three_level()
    so, sg, s, ref, got = _run(oracle_mod, g, 85, ninner=ninner, nouter=3)
    _equal(so, sg, s, ref, got)


@pytest.mark.parametrize("sdiv_seed", [None, 88])
def test_sphere_three_levels_masked_bitwise(oracle_mod, sdiv_seed):
    """Blocked cells inside a sphere (int8 mask from the signed distance)
    across three levels, ghostcell operator, with and without a geometry
    source SDIV."""
    g = three_level()
    blocked = sphere_mask(g, radius=0.3, centre=(0.25, 0.25, 0.25))
    so, sg, s, ref, got = _run(oracle_mod, g, 86, ibtype=1, blocked=blocked, nouter=2,
                               sdiv_seed=sdiv_seed)
    _equal(so, sg, s, ref, got)


@pytest.mark.parametrize("masked", [False, True])
def test_multilevel_converged_fp32_refined(oracle_mod, masked):
    g = three_level()
    blocked = sphere_mask(g, radius=0.3, centre=(0.25, 0.25, 0.25)) if masked else None
    so, sg, s, ref, got = _run(oracle_mod, g, 87, nouter=2000, eps_rel=1e-10,
                               precision="fp32", ibtype=int(masked), blocked=blocked)
    assert so["maxrhs"] <= 1e-10 * so["maxrhs0"]
    assert sg["converged"] and sg["maxrhs"] <= 1e-10 * sg["maxrhs0"]
    inner = g.interior_mask()
    a, b = got[3][inner], ref[3][inner]
    a, b = a - a.mean(), b - b.mean()
    assert np.abs(a - b).max() <= 1e-7 * np.abs(b).max()
    for x, y in zip(got[:3], ref[:3]):
        assert np.abs(x - y).max() <= 1e-7 * np.abs(y).max()


def test_config3_sphere_full_size(oracle_mod):
    """Multi-level batch with the immersed sphere at full size (level-1 box of
    4x4x4 blocks of 24^3, 1728 blocks of 16^3, 672 of 32^3 around the sphere
    of radius 0.2: 2464 blocks): FP64 bitwise equal to the oracle over two
    outer iterations; the FP32-refined solve converged to 1e-10 against the
    FP64 one (error reported in the assertion message)."""
    import mgp
    from cases import config3
    g = config3()
    assert g.ngrid == 2464
    o = oracle_mod.Oracle(g)
    bp, blocked = mask_levels(g, o, sphere_mask(g))
    st = o.setup(bp=bp)
    u, v, w = velocity_levels(g, o, 3, blocked=blocked)
    ins = [a.copy() for a in (u, v, w, np.zeros(g.n3d))]
    ref = [a.copy() for a in ins]
    so = o.mgpoisl(st, *ref, 1e-3, ninner=5, nouter=2, ibtype=1)
    s64 = mgp.PressureSolver(g, ninner=5, nouter=2, ib="ghostcell", bp=bp)
    got = [a.copy() for a in ins]
    sg = s64.mgpoisl(*got, 1e-3)
    _equal(so, sg, s64, ref, got)
    s64.set_iterations(nouter=3000, eps_rel=1e-10)
    r64 = [a.copy() for a in ins]
    st64 = s64.mgpoisl(*r64, 1e-3)
    s32 = mgp.PressureSolver(g, ninner=5, nouter=3000, eps_rel=1e-10, ib="ghostcell", bp=bp,
                             precision="fp32")
    r32 = [a.copy() for a in ins]
    st32 = s32.mgpoisl(*r32, 1e-3)
    assert st64["converged"] and st32["converged"]
    inner = g.interior_mask()
    a, b = r32[3][inner], r64[3][inner]
    err_p = np.abs((a - a.mean()) - (b - b.mean())).max() / np.abs(b - b.mean()).max()
    err_u = max(np.abs(x - y).max() for x, y in zip(r32[:3], r64[:3])) / \
        max(np.abs(y).max() for y in r64[:3])
    print(f"[config3] {g.ngrid} grids: FP64 {st64['nouter_done']} outer iterations "
          f"{st64['solve_ms']:.1f} ms, FP32-refined {st32['nouter_done']} outer iterations "
          f"{st32['solve_ms']:.1f} ms; FP32-refined vs FP64 (bitwise the oracle's "
          f"iteration): pressure {err_p:.3e}, velocity {err_u:.3e} (relative max)", flush=True)
    assert err_p < 1e-7 and err_u < 1e-7, (err_p, err_u, st64["nouter_done"],
                                           st32["nouter_done"])


def test_config3_int8_mask_entry(oracle_mod):
    """The int8 blocked-cell entry (mgp_init_mask) on the full config-3
    sphere batch: the caller passes 1 = blocked at every stored position,
    with the parent cells under refined grids and the ghost layers already
    set the way the immersed-boundary setup leaves BP (a parent cell is
    fluid when any child is, ghosts exchanged per level, PAR ghosts from
    the parent: src/ib/gc_blockbp_mod.F90:299).  The library's BP, its
    factors and two outer iterations equal the float-BP entry's and the
    oracle's bit for bit."""
    import mgp
    from cases import config3
    g = config3()
    o = oracle_mod.Oracle(g)
    bp, blocked = mask_levels(g, o, sphere_mask(g))
    assert blocked.dtype == np.int8 and np.array_equal(1.0 - blocked, bp)
    st = o.setup(bp=bp)
    u, v, w = velocity_levels(g, o, 3, blocked=blocked)
    ins = [a.copy() for a in (u, v, w, np.zeros(g.n3d))]
    ref = [a.copy() for a in ins]
    so = o.mgpoisl(st, *ref, 1e-3, ninner=5, nouter=2, ibtype=1)
    s8 = mgp.PressureSolver(g, ninner=5, nouter=2, ib="ghostcell", blocked=blocked)
    sf = mgp.PressureSolver(g, ninner=5, nouter=2, ib="ghostcell", bp=bp)
    inner = g.interior_mask()
    assert np.array_equal(s8.get_field("BP")[inner], bp[inner])
    for k in ("BP", "GSAP", "SIPLPR", "SIPUT"):
        assert np.array_equal(s8.get_field(k), sf.get_field(k)), k
    got = [a.copy() for a in ins]
    sg = s8.mgpoisl(*got, 1e-3)
    _equal(so, sg, s8, ref, got)
    gf = [a.copy() for a in ins]
    sf.mgpoisl(*gf, 1e-3)
    for a, b, nm in zip(gf, got, "uvwp"):
        assert np.array_equal(a, b), nm
