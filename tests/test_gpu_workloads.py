"""GPU checks on the benchmark workloads themselves, at their full sizes:

* 8x8x8 periodic blocks of 24^3 (192^3, seed 2): hyperplane SIP bitwise
  against the oracle at a fixed iteration count, and the FP32-refined form
  converged to 1e-10 against the FP64 form;
* 32x32x32 periodic blocks of 24^3 (452,984,832 cells, k^-5/3 velocity,
  seed 4): the one shared copy of the SIP factors against the oracle's
  factors of one periodic 24^3 block, shared against per-block factors bit
  for bit, and the FP32-refined residual history against the FP64 one;
* the adjoint projection on 8x8x4 periodic blocks of 24^3 (3,538,944
  cells, seeds 5, 6, 7): the dot-product identity, central differences at a
  fixed iteration count, FP32-refined against FP64 gradients and a
  checkpointed 20-step rollout against the plain one.

Reference: src/flow/pressuresolver_mod.F90:160-357 (mgpoisl),
src/flow/sip_hyperplane_mod.F90:106-257 (factors)."""
import numpy as np
import pytest

from cases import topology, velocity

pytestmark = pytest.mark.gpu

C4_BLOCKS = (32, 32, 32)
N24 = (24, 24, 24)


def _solver(g, **kw):
    import mgp
    return mgp.PressureSolver(g, **kw)


# ---------------------------------------------------------------- 192^3, SIP
def test_c2_sip_fixed_count_bitwise(oracle_mod):
    """Hyperplane SIP on the 8x8x8-block case with the seed-2 U(-1,1)
    velocity: u, v, w, p, the residual history and ipc equal the oracle's
    bit for bit after 3 outer iterations of 5 inner."""
    g = topology("c2")
    o = oracle_mod.Oracle(g)
    st = o.setup()
    u, v, w = velocity(g, o, 2)
    ins = [a.copy() for a in (u, v, w)] + [np.zeros(g.n3d)]
    ref = [a.copy() for a in ins]
    so = o.mgpoisl(st, *ref, 1e-3, ninner=5, nouter=3)
    s = _solver(g, ninner=5, nouter=3)
    got = [a.copy() for a in ins]
    sg = s.mgpoisl(*got, 1e-3)
    assert sg["ipc"] == so["ipc"] == 15
    assert np.array_equal(s.history(), so["history"])
    for a, b, nm in zip(ref, got, "uvwp"):
        assert np.array_equal(a, b), nm
    s.close()


def test_c2_sip_fp32_refined_converges_with_fp64():
    """Both forms reach a 1e-10 relative residual after the same number of
    inner iterations, and the FP32-refined pressure and velocity agree with
    the FP64 ones to 1e-12 (relative max; the pressure modulo its mean)."""
    import torch
    g = topology("c2")
    import oracle
    o = oracle.Oracle(g)
    u, v, w = velocity(g, o, 2)
    out = {}
    for prec in ("fp64", "fp32"):
        s = _solver(g, ninner=5, nouter=4000, eps_rel=1e-10, precision=prec, check_every=8)
        f = [torch.from_numpy(a.copy()).cuda() for a in (u, v, w)] + \
            [torch.zeros(g.n3d, dtype=torch.float64, device="cuda")]
        r = s.mgpoisl(*f, 1e-3)
        assert r["converged"] and r["maxrhs"] <= 1e-10 * r["maxrhs0"], (prec, r)
        out[prec] = (r, [t.cpu().numpy() for t in f])
        print(f"[c2 sip {prec}] ipc {r['ipc']} outer {r['nouter_done']} "
              f"solve {r['solve_ms']:.1f} ms", flush=True)
        s.close()
    (r64, f64), (r32, f32) = out["fp64"], out["fp32"]
    assert r32["ipc"] == r64["ipc"]
    inner = g.interior_mask()
    for a, b, nm in zip(f64, f32, "uvwp"):
        if nm == "p":
            a, b = a[inner] - a[inner].mean(), b[inner] - b[inner].mean()
        err = np.abs(a - b).max() / np.abs(a).max()
        print(f"[c2 sip] fp32-refined vs fp64 {nm}: {err:.2e}", flush=True)
        assert err <= 1e-12, nm


# ---------------------------------------------------------------- 32768 blocks
@pytest.fixture(scope="module")
def c4():
    """The 32768-block workload: topology and the seed-4 k^-5/3 velocity
    (CUDA tensors, ghosts filled by the library's 2-layer exchange)."""
    import torch
    import bench
    from mgp.grids import uniform_box
    g = uniform_box(C4_BLOCKS, N24)
    f = bench.make_velocity(g, C4_BLOCKS, N24, seed=4)
    yield g, f
    del f
    torch.cuda.empty_cache()


def _c4_solver(g, monkeypatch, shared=True, **kw):
    # one fixed pipeline shape: results do not depend on it, and the
    # autotuner's timing runs would only lengthen the test
    monkeypatch.setenv("MGP_PLAN_SHAPE", "3")
    monkeypatch.setenv("MGP_SHARED_FACTORS", "1" if shared else "0")
    return _solver(g, **kw)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_c4_shared_factors_equal_oracle_block(oracle_mod, c4, monkeypatch, prec):
    """Every block of the periodic 32x32x32 layout has the factors of one
    self-periodic 24^3 block with the same spacing; the library keeps one
    copy (one class of 13,824 cells), and that copy, read back through three
    different grids, equals the oracle's factors of the single block (cast
    to FP32 for the FP32-refined form) bit for bit."""
    from mgp.grids import uniform_box
    g, _ = c4
    one = uniform_box((1, 1, 1), N24, length=(1.0 / 32,) * 3)
    assert one.spacing[0].tolist() == g.spacing[0].tolist()
    st = oracle_mod.Oracle(one).setup()
    s = _c4_solver(g, monkeypatch, ninner=5, nouter=1, precision=prec)
    assert s.factor_classes() == (1, 13824)
    for igrid in (1, 4711, g.ngrid):
        for name, k in (("SIPLW", "lw"), ("SIPLS", "ls"), ("SIPLB", "lb"), ("SIPLPR", "lpr"),
                        ("SIPUE", "ue"), ("SIPUN", "un"), ("SIPUT", "ut")):
            want = st["fac"][k]
            if prec == "fp32":
                want = want.astype(np.float32).astype(np.float64)
            assert np.array_equal(s.get_field(name, igrid), want), (igrid, name)
        assert np.array_equal(s.get_field("GSAP", igrid), st["coef"]["ap"]), igrid
    s.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_c4_shared_and_unshared_factors_bitwise(c4, monkeypatch, prec):
    """Two outer iterations on the full 32768-block batch with one factor
    copy per block and with the shared copy give the same bits in u, v, w,
    p and in the residual history."""
    import torch
    g, f0 = c4
    outs = []
    for shared in (False, True):
        s = _c4_solver(g, monkeypatch, shared=shared, ninner=5, nouter=2, precision=prec)
        assert s.factor_classes()[0] == (1 if shared else g.ngrid)
        if not outs:
            s.connect(f0, layers=2, corners=True)
        f = [t.clone() for t in f0] + [torch.zeros_like(f0[0])]
        r = s.mgpoisl(*f, 1e-3)
        assert r["nouter_done"] == 2
        outs.append((f, s.history()))
        s.close()
    (a, ha), (b, hb) = outs
    assert np.array_equal(ha, hb)
    for x, y, nm in zip(a, b, "uvwp"):
        assert torch.equal(x, y), nm


def test_c4_fp32_refined_history_follows_fp64(c4, monkeypatch):
    """Over 50 outer iterations the level-weighted max |RHS| of the
    FP32-refined form stays within 1e-5 (relative) of the FP64 form's after
    every outer iteration: the FP32 inner sweeps do not change the
    convergence of the defect correction."""
    import torch
    g, f0 = c4
    hist = {}
    for prec in ("fp64", "fp32"):
        s = _c4_solver(g, monkeypatch, ninner=5, nouter=50, precision=prec, check_every=50)
        s.connect(f0, layers=2, corners=True)
        f = [t.clone() for t in f0] + [torch.zeros_like(f0[0])]
        r = s.mgpoisl(*f, 1e-3)
        assert r["nouter_done"] == 50
        hist[prec] = s.history() / r["maxrhs0"]
        s.close()
        del f
    ratio = hist["fp32"] / hist["fp64"]
    print(f"[c4] relative residual after 50: fp64 {hist['fp64'][-1]:.6e}, "
          f"fp32-refined {hist['fp32'][-1]:.6e}, ratio range "
          f"[{ratio.min():.7f}, {ratio.max():.7f}]", flush=True)
    assert np.all(np.abs(ratio - 1.0) <= 1e-5)
    assert hist["fp64"][-1] < 0.5 * hist["fp64"][0]


# ---------------------------------------------------------------- adjoint, 256 blocks
C5_FORMS = ("fp64", "fp32")


@pytest.fixture(scope="module")
def c5():
    """8x8x4 periodic blocks of 24^3 on [0,1]x[0,1]x[0,0.5]: U(-1,1) input
    velocity (numpy default_rng(5)), N(0,1) weights (rng 6) and directions
    (rng 7) on interior cells, as CUDA tensors."""
    import torch
    from mgp.grids import uniform_box
    g = uniform_box((8, 8, 4), N24, length=(1.0, 1.0, 0.5))
    inner = torch.from_numpy(g.interior_mask()).cuda()

    def field(rng, kind):
        a = rng.uniform(-1, 1, g.n3d) if kind == "u" else rng.standard_normal(g.n3d)
        return torch.where(inner, torch.from_numpy(a).cuda(), 0.0)

    r5, r6, r7 = (np.random.default_rng(k) for k in (5, 6, 7))
    u0 = [field(r5, "u") for _ in range(3)]
    wts = [field(r6, "w") for _ in range(3)]
    dirs = [[field(r7, "w") for _ in range(3)] for _ in range(8)]
    return g, u0, wts, dirs, {}


def _c5_gradient(c5, prec):
    """J = sum w * P(u) and dJ/du through torch.autograd (the library's
    adjoint), solves converged to 1e-12; cached per form."""
    import torch
    import mgp.autograd as ag
    g, u0, wts, _, cache = c5
    if prec not in cache:
        import time
        s = _solver(g, ninner=5, nouter=4000, eps_rel=1e-12, precision=prec, check_every=8)
        x = [t.clone().requires_grad_(True) for t in u0]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = ag.project(s, *x, 1e-3)
        J = sum((a * b).sum() for a, b in zip(wts, out[:3]))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        assert s.last_stats["converged"] and s.last_adjoint_stats["converged"]
        print(f"[c5 {prec}] forward {s.last_stats['nouter_done']} outer {t1 - t0:.2f} s, "
              f"adjoint {s.last_adjoint_stats['nouter_done']} outer {t2 - t1:.2f} s",
              flush=True)
        cache[prec] = (s, float(J.detach()), [t.grad.detach().clone() for t in x])
    return cache[prec]


@pytest.mark.parametrize("prec", C5_FORMS)
def test_c5_dot_product_identity(c5, prec):
    """The projection is linear in u, so J(u) = <w, P u> = <P^T w, u> =
    <dJ/du, u>: forward solve and adjoint solve on the GPU agree."""
    _, u0, _, _, _ = c5
    s, J, grad = _c5_gradient(c5, prec)
    lhs = J
    rhs = float(sum((gr * u).sum() for gr, u in zip(grad, u0)))
    scale = float(sum((gr.abs() * u.abs()).sum() for gr, u in zip(grad, u0)))
    print(f"[c5 {prec}] <w,Pu> = {lhs:.15e}, <P^T w,u> = {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-10 * scale


@pytest.mark.parametrize("prec", C5_FORMS)
def test_c5_central_differences(c5, prec):
    """dJ/du against central differences (h = 1e-6) in 8 random directions,
    with the outer-iteration count fixed at what the gradient's forward
    solve needed (so that J is one fixed linear map)."""
    import torch
    import mgp.autograd as ag
    g, u0, wts, dirs, _ = c5
    s, _, grad = _c5_gradient(c5, prec)
    nfix = s.last_stats["nouter_done"]
    s.set_iterations(nouter=nfix, eps_rel=0.0)

    def J(f):
        with torch.no_grad():
            o = ag.project(s, *f, 1e-3)
        return float(sum((a * b).sum() for a, b in zip(wts, o[:3])))

    h = 1e-6
    errs = []
    for d in dirs:
        fd = (J([a + h * b for a, b in zip(u0, d)]) - J([a - h * b for a, b in zip(u0, d)])) / (2 * h)
        ad = float(sum((t * b).sum() for t, b in zip(grad, d)))
        errs.append(abs(fd - ad) / max(abs(ad), 1e-300))
    s.set_iterations(nouter=4000, eps_rel=1e-12)
    print(f"[c5 {prec}] {nfix} outer fixed, FD rel err max {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-6


def test_c5_fp32_refined_gradient_matches_fp64(c5):
    """Both forms converge to 1e-12, so their gradients agree far below the
    FP32 rounding of the inner sweeps."""
    _, _, g64 = _c5_gradient(c5, "fp64")
    _, _, g32 = _c5_gradient(c5, "fp32")
    err = max(float((a - b).abs().max()) for a, b in zip(g32, g64)) / \
        max(float(b.abs().max()) for b in g64)
    print(f"[c5] fp32-refined gradient vs fp64: {err:.2e}", flush=True)
    assert err <= 1e-11


def test_c5_checkpointed_rollout_bitwise(c5):
    """20 steps of u <- P(u + 0.1 tanh u) at a fixed 60 outer iterations per
    projection: the gradient with a checkpoint every 5 steps (the segments'
    forward passes recomputed during the backward pass) equals the one
    that keeps every step, bit for bit."""
    import torch
    import mgp.autograd as ag
    g, u0, wts, _, _ = c5
    s = _solver(g, ninner=5, nouter=60, check_every=60)

    def step(a, b, c):
        p = ag.project(s, a + 0.1 * torch.tanh(a), b + 0.1 * torch.tanh(b),
                       c + 0.1 * torch.tanh(c), 1e-3)
        return p[0], p[1], p[2]

    grads = []
    for every in (0, 5):
        x = [t.clone().requires_grad_(True) for t in u0]
        out = ag.rollout(step, x, 20, checkpoint_every=every)
        sum((a * b).sum() for a, b in zip(wts, out)).backward()
        grads.append([t.grad.clone() for t in x])
    for a, b in zip(*grads):
        assert torch.equal(a, b)
    s.close()
