"""GPU parity on the reference's own test grids (tests/*/grids.h5 topologies
kept in tests/golden/topologies.json): connect per level and the whole
mgpoisl (src/flow/pressuresolver_mod.F90:160-420), bitwise against the
oracle in FP64, and the FP32-refined form converged."""
import numpy as np
import pytest

from cases import SUPPORTED_TOPOLOGIES, reference_topology, velocity_levels

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", SUPPORTED_TOPOLOGIES)
def test_connect_reference_topology(oracle_mod, name):
    import torch
    import mgp
    g = reference_topology(name)
    print(f"[topology] {name}: {g.ngrid} grids, levels {g.minlevel}..{g.maxlevel}", flush=True)
    o = oracle_mod.Oracle(g)
    s = mgp.PressureSolver(g)
    rng = np.random.default_rng(91)
    for lvl in range(g.minlevel, g.maxlevel + 1):
        for layers, corners, fwd in ((1, False, 0), (1, False, 1), (1, False, -1),
                                     (2, True, 0), (2, False, 0)):
            a = np.where(g.interior_mask(), rng.standard_normal(g.n3d), 0.0)
            b = a.copy()
            o.exchange([a], level=lvl, layers=layers, corners=corners, forward=fwd)
            t = torch.from_numpy(b).cuda()
            s.connect([t], level=lvl, layers=layers, corners=corners, forward=fwd)
            assert np.array_equal(t.cpu().numpy(), a), (name, lvl, layers, corners, fwd)
    s.close()


def _solve(oracle_mod, g, seed, ityp=0, ninner=3, nouter=2, omega=1.1, shape=None,
           sor_all_levels=0):
    import mgp
    o = oracle_mod.Oracle(g)
    st = o.setup(ityp=ityp)
    u, v, w = velocity_levels(g, o, seed)
    p = np.zeros(g.n3d)
    got = [a.copy() for a in (u, v, w, p)]
    so = o.mgpoisl(st, u, v, w, p, 1e-3, ityp=ityp, ninner=ninner, nouter=nouter, omega=omega,
                   sor_all_levels=sor_all_levels)
    s = mgp.PressureSolver(g, type={0: "sip", 1: "sor", 2: "classic"}[ityp], ninner=ninner,
                           nouter=nouter, omega=omega, sor_all_levels=sor_all_levels)
    if shape is not None:
        assert s.plan_shape(shape)[0] == shape
    sg = s.mgpoisl(*got, 1e-3)
    assert sg["maxrhs0"] == so["maxrhs0"] and sg["maxrhs"] == so["maxrhs"]
    assert np.array_equal(s.history(), so["history"])
    for a, b, nm in zip((u, v, w, p), got, "uvwp"):
        assert np.array_equal(a, b), (nm, float(np.abs(a - b).max()))
    s.close()


@pytest.mark.parametrize("ityp", [0, 2])
@pytest.mark.parametrize("name", SUPPORTED_TOPOLOGIES)
def test_mgpoisl_reference_topology_bitwise(oracle_mod, name, ityp):
    g = reference_topology(name)
    print(f"[topology] {name} ityp={ityp}", flush=True)
    _solve(oracle_mod, g, 92 + ityp, ityp=ityp)


@pytest.mark.parametrize("name", ["Parker_micro", "Checkblock", "Statistics", "BlasiusBL"])
def test_mgpoisl_reference_topology_sor_bitwise(oracle_mod, name):
    """type "sor": SOR on the refined levels of the reference's nested grids."""
    g = reference_topology(name)
    print(f"[topology] {name} sor", flush=True)
    _solve(oracle_mod, g, 95, ityp=1, ninner=4, omega=1.2)


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5, 6, 7, 8, 9])
@pytest.mark.parametrize("ityp", [0, 1, 2])
def test_thin_block_plan_shapes_bitwise(oracle_mod, shape, ityp):
    """Every pipeline shape on NaturalConvection's 32 x 1 x 32 block, whose
    plan items end early (every cell lies on both y faces, so 64 cells would
    need more than an item's 64 face-store entries): SIP, SOR (on its only
    level) and classic SIP, bitwise against the oracle."""
    g = reference_topology("NaturalConvection")
    print(f"[topology] NaturalConvection shape {shape} ityp={ityp}", flush=True)
    _solve(oracle_mod, g, 97, ityp=ityp, ninner=3, nouter=2, omega=1.2, shape=shape,
           sor_all_levels=int(ityp == 1))


@pytest.mark.parametrize("name", ["Parker_micro", "Santarelli", "Checkblock",
                                  "HotAndColdSpheres", "TaylorGreen", "GeostrophicWind",
                                  "sphere", "NaturalConvection"])
def test_fp32_refined_converged_reference_topology(oracle_mod, name):
    """FP32 sweeps inside the FP64 correction reach 1e-10 on the reference's
    grids and agree with the FP64 solve (the pressure modulo its mean where
    no outlet fixes the constant)."""
    import mgp
    g = reference_topology(name)
    print(f"[topology] {name} converged", flush=True)
    o = oracle_mod.Oracle(g)
    u, v, w = velocity_levels(g, o, 96)
    ins = [a.copy() for a in (u, v, w, np.zeros(g.n3d))]
    out = {}
    for prec in ("fp64", "fp32"):
        s = mgp.PressureSolver(g, ninner=5, nouter=4000, eps_rel=1e-10, precision=prec)
        f = [a.copy() for a in ins]
        r = s.mgpoisl(*f, 1e-3)
        s.close()
        assert r["converged"] and r["maxrhs"] <= 1e-10 * r["maxrhs0"], (prec, r)
        out[prec] = (f, r)
    (f32, r32), (f64, r64) = out["fp32"], out["fp64"]
    inner = g.interior_mask()
    a, b = f32[3][inner], f64[3][inner]
    if not (g.bc == 3).any():
        a, b = a - a.mean(), b - b.mean()
    err_p = np.abs(a - b).max() / np.abs(b).max()
    err_u = max(np.abs(a - b).max() for a, b in zip(f32[:3], f64[:3])) / \
        max(np.abs(b).max() for b in f64[:3])
    assert err_p < 1e-7 and err_u < 1e-7, (err_p, err_u, r32["nouter_done"], r64["nouter_done"])
