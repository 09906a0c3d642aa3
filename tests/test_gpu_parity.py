"""GPU parity: libmgp (through its C ABI) against the CPU oracle.

FP64 results must be bitwise equal: both sides evaluate every expression in
the reference's order without FMA, and the GPU sweeps follow the same
dependency order (hyperplanes), so only the storage layout differs.
"""
import numpy as np
import pytest

from cases import blocked_mask, topology, velocity

pytestmark = pytest.mark.gpu

TOPOS = ["self12", "box2", "aniso", "walls", "odd"]


def _solver(g, **kw):
    import mgp
    return mgp.PressureSolver(g, **kw)


def _run_both(oracle_mod, name, seed, ityp=0, ninner=5, nouter=3, ibtype=0,
              frac=0.0, eps_rel=0.0, precision="fp64", nouter_min=0, omega=1.1,
              sor_all_levels=0):
    g = topology(name)
    o = oracle_mod.Oracle(g)
    bp = None
    if frac > 0:
        m = blocked_mask(g, seed + 100, frac)
        bp = 1.0 - m
        o.exchange([bp], layers=2, corners=True)
    st = o.setup(bp=bp, ityp=ityp)
    u, v, w = velocity(g, o, seed)
    p = np.zeros(g.n3d)
    ins = [a.copy() for a in (u, v, w, p)]
    so = o.mgpoisl(st, u, v, w, p, 1e-3, ityp=ityp, ninner=ninner, nouter=nouter,
                   nouter_min=nouter_min, eps_rel=eps_rel, ibtype=ibtype, omega=omega,
                   sor_all_levels=sor_all_levels)
    s = _solver(g, type={0: "sip", 1: "sor", 2: "classic"}[ityp], ninner=ninner,
                nouter=nouter, nouter_min=nouter_min, eps_rel=eps_rel,
                ib={0: "noib", 1: "ghostcell"}[ibtype], precision=precision, bp=bp,
                omega=omega, sor_all_levels=sor_all_levels)
    gu, gv, gw, gp = ins
    sg = s.mgpoisl(gu, gv, gw, gp, 1e-3)
    return g, o, st, so, s, sg, (u, v, w, p), (gu, gv, gw, gp)


@pytest.mark.parametrize("name", TOPOS)
def test_coefficients_bitwise(oracle_mod, name):
    g = topology(name)
    o = oracle_mod.Oracle(g)
    st = o.setup()
    s = _solver(g)
    assert np.array_equal(s.get_field("GSAP"), st["coef"]["ap"])
    for k in ("GSAW", "GSAE", "GSAS", "GSAN", "GSAB", "GSAT"):
        assert np.array_equal(s.get_field(k), st["coef"][k[2:].lower()]), k
    for k, f in (("SIPLW", "lw"), ("SIPLS", "ls"), ("SIPLB", "lb"), ("SIPLPR", "lpr"),
                 ("SIPUE", "ue"), ("SIPUN", "un"), ("SIPUT", "ut")):
        assert np.array_equal(s.get_field(k), st["fac"][f]), k


@pytest.mark.parametrize("name", TOPOS)
@pytest.mark.parametrize("ityp", [0, 2])
def test_mgpoisl_fixed_iterations_bitwise(oracle_mod, name, ityp):
    g, o, st, so, s, sg, ref, got = _run_both(oracle_mod, name, seed=2, ityp=ityp)
    assert sg["ipc"] == so["ipc"] and sg["nouter_done"] == so["nouter_done"]
    assert sg["maxrhs0"] == so["maxrhs0"]
    assert sg["maxrhs"] == so["maxrhs"]
    assert np.array_equal(s.history(), so["history"])
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm
    assert np.array_equal(s.get_field("DP"), so["dp"] * g.interior_mask())


@pytest.mark.parametrize("ninner", [1, 2, 4])
def test_ninner_variants_bitwise(oracle_mod, ninner):
    g, o, st, so, s, sg, ref, got = _run_both(oracle_mod, "box2", seed=3, ninner=ninner,
                                              nouter=4)
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm


def test_masked_ghostcell_bitwise(oracle_mod):
    g, o, st, so, s, sg, ref, got = _run_both(oracle_mod, "box2", seed=4, ibtype=1,
                                              frac=0.15)
    assert np.array_equal(s.get_field("GSAP"), st["coef"]["ap"])
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm


def test_converged_fp64_and_fp32(oracle_mod):
    g, o, st, so, s, sg, ref, got = _run_both(oracle_mod, "box2", seed=5, nouter=400,
                                              eps_rel=1e-10)
    assert so["maxrhs"] <= 1e-10 * so["maxrhs0"]
    assert sg["nouter_done"] == so["nouter_done"]
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    # FP32 sweeps inside the FP64 outer loop reach the same residual
    s32 = _solver(g, nouter=400, eps_rel=1e-10, precision="fp32")
    u, v, w = velocity(g, o, 5)
    p = np.zeros(g.n3d)
    r32 = s32.mgpoisl(u, v, w, p, 1e-3)
    assert r32["converged"] and r32["maxrhs"] <= 1e-10 * r32["maxrhs0"]
    inner = g.interior_mask()
    pr = ref[3][inner] - ref[3][inner].mean()
    pq = p[inner] - p[inner].mean()
    assert np.abs(pq - pr).max() <= 1e-8 * np.abs(pr).max()


@pytest.mark.parametrize("name", TOPOS)
@pytest.mark.parametrize("ninner", [1, 4, 5])
def test_sor_bitwise(oracle_mod, name, ninner):
    """SOR (relax, src/flow/sor_mod.F90:102-203) on every level, bitwise."""
    g, o, st, so, s, sg, ref, got = _run_both(oracle_mod, name, seed=8, ityp=1,
                                              ninner=ninner, nouter=3, omega=1.2,
                                              sor_all_levels=1)
    assert sg["ipc"] == so["ipc"] and sg["nouter_done"] == so["nouter_done"]
    assert sg["maxrhs"] == so["maxrhs"]
    assert np.array_equal(s.history(), so["history"])
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm


def test_sor_masked_bitwise(oracle_mod):
    g, o, st, so, s, sg, ref, got = _run_both(oracle_mod, "box2", seed=9, ityp=1,
                                              ninner=3, nouter=4, omega=1.3,
                                              sor_all_levels=1, ibtype=1, frac=0.15)
    assert sg["maxrhs"] == so["maxrhs"]
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm


def test_sor_single_level_runs_sip(oracle_mod):
    """type "sor" relaxes only levels above the coarsest (mgpoisit,
    src/flow/pressuresolver_mod.F90:403-412): on one level it is SIP."""
    g, o, st, so, s, sg, ref, got = _run_both(oracle_mod, "box2", seed=10, ityp=1,
                                              ninner=2, nouter=2)
    _, _, _, so0, _, _, ref0, _ = _run_both(oracle_mod, "box2", seed=10, ityp=0,
                                            ninner=2, nouter=2)
    assert so["maxrhs"] == so0["maxrhs"] and sg["maxrhs"] == so["maxrhs"]
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm


def test_sor_converges_fp64_and_fp32(oracle_mod):
    g = topology("box2")
    o = oracle_mod.Oracle(g)
    st = o.setup()
    u, v, w = velocity(g, o, 12)
    p = np.zeros(g.n3d)
    so = o.mgpoisl(st, u.copy(), v.copy(), w.copy(), p, 1e-3, ityp=1, omega=1.2,
                   sor_all_levels=1, ninner=4, nouter=3000, eps_rel=1e-10)
    assert so["maxrhs"] <= 1e-10 * so["maxrhs0"]
    for prec in ("fp64", "fp32"):
        s = _solver(g, type="sor", omega=1.2, sor_all_levels=True, ninner=4,
                    nouter=3000, eps_rel=1e-10, precision=prec)
        gp = np.zeros(g.n3d)
        r = s.mgpoisl(u.copy(), v.copy(), w.copy(), gp, 1e-3)
        assert r["converged"] and r["maxrhs"] <= 1e-10 * r["maxrhs0"], prec
        if prec == "fp64":
            assert r["nouter_done"] == so["nouter_done"]
            assert np.array_equal(gp, p)
        else:
            # periodic: the pressure is fixed up to a constant
            inner = g.interior_mask()
            a, b = gp[inner] - gp[inner].mean(), p[inner] - p[inner].mean()
            assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max()


def test_connect_identity(oracle_mod):
    import torch
    g = topology("aniso")
    o = oracle_mod.Oracle(g)
    s = _solver(g)
    for layers, corners, fwd in ((1, False, 0), (1, False, 1), (1, False, -1),
                                 (2, True, 0), (2, False, 0)):
        a = np.arange(g.n3d, dtype=np.float64) * g.interior_mask()
        b = a.copy()
        o.exchange([a], layers=layers, corners=corners, forward=fwd)
        t = torch.from_numpy(b).cuda()
        s.connect([t], layers=layers, corners=corners, forward=fwd)
        assert np.array_equal(t.cpu().numpy(), a), (layers, corners, fwd)


def test_single_hip_runtime():
    """torch and libmgp share one HIP runtime in the process."""
    import torch
    import mgp
    g = topology("box2")
    s = _solver(g)
    x = torch.ones(4, device="cuda")
    assert float(x.sum()) == 4.0
    maps = open("/proc/self/maps").read()
    libs = {ln.split()[-1] for ln in maps.splitlines() if "libamdhip64" in ln}
    assert len(libs) == 1, libs
    s.close()


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5, 6, 7, 8, 9])
@pytest.mark.parametrize("case", ["sip-walls", "classic-masked", "sor-odd", "sor-masked"])
def test_plan_shapes_bitwise(oracle_mod, shape, case):
    """Every pipeline shape of the plane kernels (items of 64 or 128 cells,
    2 to 8 in flight, 1, 2, 4 or 8 waves per block) gives the oracle's FP64 result bit for bit."""
    kind, name = case.split("-")
    ityp = {"sip": 0, "classic": 2, "sor": 1}[kind]
    topo = {"walls": "walls", "masked": "box2", "odd": "odd"}[name]
    frac = 0.15 if name == "masked" else 0.0
    g = topology(topo)
    o = oracle_mod.Oracle(g)
    bp = None
    if frac > 0:
        bp = 1.0 - blocked_mask(g, 107, frac)
        o.exchange([bp], layers=2, corners=True)
    ibtype = 1 if frac > 0 else 0
    st = o.setup(bp=bp, ityp=ityp)
    u, v, w = velocity(g, o, 7)
    p = np.zeros(g.n3d)
    ins = [a.copy() for a in (u, v, w, p)]
    kw = dict(sor_all_levels=1) if ityp == 1 else {}
    so = o.mgpoisl(st, u, v, w, p, 1e-3, ityp=ityp, ninner=3, nouter=3, ibtype=ibtype, **kw)
    s = _solver(g, type={0: "sip", 1: "sor", 2: "classic"}[ityp], ninner=3, nouter=3,
                ib={0: "noib", 1: "ghostcell"}[ibtype], bp=bp, **kw)
    assert s.plan_shape(shape)[0] == shape
    sg = s.mgpoisl(*ins, 1e-3)
    assert sg["maxrhs"] == so["maxrhs"]
    for a, b, nm in zip((u, v, w, p), ins, "uvwp"):
        assert np.array_equal(a, b), nm


def test_plan_shape_autotuner():
    """The autotuner times every shape on the handle's blocks and keeps the
    fastest."""
    import mgp
    g = topology("box2")
    s = _solver(g)
    shape, ms = s.plan_shape(0)
    assert 1 <= shape <= mgp.NUM_PLAN_SHAPES
    assert np.all(ms > 0) and ms[shape - 1] == ms.min()
    assert s.plan_shape()[0] == shape


@pytest.mark.parametrize("ibtype", [0, 1])
def test_divcal_device_bitwise(oracle_mod, ibtype):
    """mgp_divcal_device: noib, and ghostcell with BP and the source SDIV."""
    import torch
    g = topology("aniso")
    o = oracle_mod.Oracle(g)
    bp = sdiv = None
    if ibtype:
        bp = 1.0 - blocked_mask(g, 150, 0.1)
        o.exchange([bp], layers=2, corners=True)
        sdiv = np.where(g.interior_mask(), np.random.default_rng(151).standard_normal(g.n3d),
                        0.0)
    u, v, w = velocity(g, o, 152)
    ref = g.zeros()
    o.divcal(ref, u, v, w, 3.5, bp=bp, sdiv=sdiv)
    s = _solver(g, ib={0: "noib", 1: "ghostcell"}[ibtype], bp=bp)
    if ibtype:
        s.set_sdiv(sdiv)
    got = s.divcal(*[torch.from_numpy(a).cuda() for a in (u, v, w)], 3.5).cpu().numpy()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("name", ["box2", "walls"])
@pytest.mark.parametrize("ityp", [0, 2])
def test_sdiv_source_bitwise(oracle_mod, name, ityp):
    """Ghostcell projection with a nonzero geometry source SDIV (seeded
    normal values on interior cells): FP64 bitwise equal to the oracle over
    three outer iterations; resetting SDIV gives the solve without it."""
    import mgp
    g = topology(name)
    o = oracle_mod.Oracle(g)
    bp = 1.0 - blocked_mask(g, 140, 0.1)
    o.exchange([bp], layers=2, corners=True)
    st = o.setup(bp=bp, ityp=ityp)
    u, v, w = velocity(g, o, 141)
    sdiv = np.where(g.interior_mask(), np.random.default_rng(142).standard_normal(g.n3d), 0.0)
    ins = [a.copy() for a in (u, v, w, np.zeros(g.n3d))]
    ref = [a.copy() for a in ins]
    so = o.mgpoisl(st, *ref, 1e-3, ityp=ityp, ninner=5, nouter=3, ibtype=1, sdiv=sdiv)
    s = mgp.PressureSolver(g, type={0: "sip", 2: "classic"}[ityp], ninner=5, nouter=3,
                           ib="ghostcell", bp=bp)
    s.set_sdiv(sdiv)
    got = [a.copy() for a in ins]
    sg = s.mgpoisl(*got, 1e-3)
    assert sg["maxrhs0"] == so["maxrhs0"] and sg["maxrhs"] == so["maxrhs"]
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm
    s.set_sdiv(None)
    ref0 = [a.copy() for a in ins]
    so0 = o.mgpoisl(st, *ref0, 1e-3, ityp=ityp, ninner=5, nouter=3, ibtype=1)
    got0 = [a.copy() for a in ins]
    sg0 = s.mgpoisl(*got0, 1e-3)
    assert sg0["maxrhs0"] == so0["maxrhs0"] != so["maxrhs0"]
    assert np.array_equal(ref0[3], got0[3])


def test_sdiv_rejected_without_ghostcell():
    import mgp
    g = topology("box2")
    s = _solver(g)
    with pytest.raises(mgp.MgpError):
        s.set_sdiv(np.zeros(g.n3d))


@pytest.mark.parametrize("name,classes", [("box2", 1), ("aniso", 2), ("walls", 8), ("c2", 1)])
def test_shared_factor_classes(name, classes):
    """Blocks with equal SIP factors share one copy: every block of a
    periodic uniform box is alike; walls make classes of their own."""
    g = topology(name)
    s = _solver(g)
    n, cells = s.factor_classes()
    assert n == classes, (name, n)
    per_block = (g.ii - 4) * (g.jj - 4) * (g.kk - 4)
    assert cells == per_block[:classes].sum()
    s.close()


def test_shared_factor_classes_masked():
    """Random blocked cells make every block's factors differ."""
    g = topology("box2")
    bp = 1.0 - blocked_mask(g, 7, 0.1)
    s = _solver(g, bp=bp)
    assert s.factor_classes() == (g.ngrid, g.ncells)
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_shared_factors_change_nothing(oracle_mod, monkeypatch, precision):
    """One copy per block (MGP_SHARED_FACTORS=0) and shared copies give the
    same bits."""
    g = topology("box2")
    o = oracle_mod.Oracle(g)
    u, v, w = velocity(g, o, 61)
    outs = []
    for env in ("0", "1"):
        monkeypatch.setenv("MGP_SHARED_FACTORS", env)
        s = _solver(g, ninner=5, nouter=6, precision=precision)
        assert s.factor_classes()[0] == (g.ngrid if env == "0" else 1)
        f = [a.copy() for a in (u, v, w, np.zeros(g.n3d))]
        s.mgpoisl(*f, 1e-3)
        outs.append((f, s.history()))
        s.close()
    (a, ha), (b, hb) = outs
    assert np.array_equal(ha, hb)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_config2_sor_omega_17_bitwise(oracle_mod):
    """The 8x8x8-block case (192^3) with SOR at omega = 1.7 on every level:
    the reference's block-lagged red-black SOR diverges there (the residual
    grows about 5x per outer iteration); the GPU follows the oracle bit for
    bit while it does."""
    g = topology("c2")
    o = oracle_mod.Oracle(g)
    st = o.setup(ityp=1)
    u, v, w = velocity(g, o, 7)
    ins = [a.copy() for a in (u, v, w)] + [np.zeros(g.n3d)]
    ref = [a.copy() for a in ins]
    so = o.mgpoisl(st, *ref, 1e-3, ityp=1, ninner=5, nouter=4, omega=1.7, sor_all_levels=1)
    s = _solver(g, type="sor", ninner=5, nouter=4, omega=1.7, sor_all_levels=1)
    got = [a.copy() for a in ins]
    sg = s.mgpoisl(*got, 1e-3)
    h = s.history()
    print(f"[config2 sor 1.7] max|RHS| {so['maxrhs0']:.4e} -> {h}", flush=True)
    assert np.array_equal(h, so["history"]) and sg["maxrhs"] == so["maxrhs"]
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm
    assert h[-1] > 10 * h[0]


def test_default_omega_sor_bitwise(oracle_mod):
    """SOR with the default omega: the library's default is the reference's
    (double)1.1f (src/flow/pressuresolver_mod.F90:135), and a run that names
    no omega on either side is bitwise the oracle's."""
    g = topology("box2")
    o = oracle_mod.Oracle(g)
    st = o.setup(ityp=1)
    u, v, w = velocity(g, o, 71)
    ins = [a.copy() for a in (u, v, w)] + [np.zeros(g.n3d)]
    ref = [a.copy() for a in ins]
    so = o.mgpoisl(st, *ref, 1e-3, ityp=1, ninner=4, nouter=3, sor_all_levels=1)
    s = _solver(g, type="sor", ninner=4, nouter=3, sor_all_levels=True)
    assert s.cfg.omega == float(np.float32(1.1))
    got = [a.copy() for a in ins]
    s.mgpoisl(*got, 1e-3)
    assert np.array_equal(s.history(), so["history"])
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm


@pytest.mark.parametrize("alfa", [0.9, 0.5])
def test_alfa_config_bitwise(oracle_mod, alfa):
    """The SIP alfa of the configuration reaches the factorisation: factors
    and a fixed-count solve equal the oracle's run at the same alfa.  (The
    reference hard-codes (double)0.92f; other values are the build's own
    setting, pinned here only against the oracle's restatement.)"""
    import ctypes as C
    g = topology("walls")
    o = oracle_mod.Oracle(g)
    lib = oracle_mod.lib()
    lib.og_set_alfa(C.c_double(alfa))
    try:
        st = o.setup()
        u, v, w = velocity(g, o, 72)
        ins = [a.copy() for a in (u, v, w)] + [np.zeros(g.n3d)]
        ref = [a.copy() for a in ins]
        so = o.mgpoisl(st, *ref, 1e-3, ninner=5, nouter=3)
    finally:
        lib.og_set_alfa(C.c_double(float(np.float32(0.92))))
    s = _solver(g, ninner=5, nouter=3, alfa=alfa)
    for k, f in (("SIPLW", "lw"), ("SIPLPR", "lpr"), ("SIPUT", "ut")):
        assert np.array_equal(s.get_field(k), st["fac"][f]), k
    got = [a.copy() for a in ins]
    s.mgpoisl(*got, 1e-3)
    assert np.array_equal(s.history(), so["history"])
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm
    s0 = _solver(g, ninner=5, nouter=3)
    assert not np.array_equal(s0.get_field("SIPLW"), st["fac"]["lw"])


@pytest.mark.parametrize("where", ["nan", "inf"])
def test_nonfinite_velocity_reports_error(where):
    """A non-finite velocity makes the residual norm non-finite: mgpoisl
    stops with MGP_ERR_NONFINITE (status 4) instead of iterating on it,
    and the handle stays usable for the next call."""
    import mgp
    g = topology("box2")
    import oracle
    o = oracle.Oracle(g)
    u, v, w = velocity(g, o, 73)
    bad = u.copy()
    bad[np.flatnonzero(g.interior_mask())[100]] = np.nan if where == "nan" else np.inf
    s = _solver(g, ninner=3, nouter=5)
    with pytest.raises(mgp.MgpError, match="status 4"):
        s.mgpoisl(bad, v.copy(), w.copy(), np.zeros(g.n3d), 1e-3)
    r = s.mgpoisl(u.copy(), v.copy(), w.copy(), np.zeros(g.n3d), 1e-3)
    assert np.isfinite(r["maxrhs"]) and r["nouter_done"] == 5


@pytest.mark.parametrize("nb,n", [((1, 1, 1), (12, 12, 12)), ((2, 2, 2), (8, 8, 8)),
                                  ((2, 1, 2), (7, 6, 5)), ((2, 1, 3), (6, 5, 2))])
def test_closing_exchange_push_bitwise(oracle_mod, monkeypatch, nb, n):
    """On single-level periodic grids the velocity and pressure correction
    stores the k-face ghosts of interior rows for the closing exchange
    itself (src/flow/pressuresolver_mod.F90:323-348: mgpcorr, then
    connect); MGP_KPUSH=0 keeps the plain correction and full exchange.
    Every ghost starts with a value no exchange made, so a ghost the push
    misses keeps it.  Cases: one self-periodic block (a block pushes into
    itself), 2x2x2 blocks, odd sizes, nk = 2 (one cell feeds both
    neighbours)."""
    import mgp
    from mgp.grids import uniform_box
    g = uniform_box(nb, n)
    o = oracle_mod.Oracle(g)
    st = o.setup()
    rng = np.random.default_rng(21)
    u, v, w, p = (rng.uniform(-1, 1, g.n3d) for _ in range(4))
    ins = [a.copy() for a in (u, v, w, p)]
    o.mgpoisl(st, u, v, w, p, 1e-3, ninner=3, nouter=2)
    outs = []
    for push in ("1", "0"):
        monkeypatch.setenv("MGP_KPUSH", push)
        s = _solver(g, ninner=3, nouter=2)
        f = [a.copy() for a in ins]
        s.mgpoisl(*f, 1e-3)
        s.close()
        outs.append(f)
    for a, b, c, nm in zip((u, v, w, p), outs[0], outs[1], "uvwp"):
        assert np.array_equal(a, b), ("push", nm, float(np.abs(a - b).max()))
        assert np.array_equal(a, c), ("plain", nm)
