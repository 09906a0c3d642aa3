"""GPU adjoint projection (mgp_adjoint / mgp_adjoint_device) and its torch
autograd front end, against the sparse oracle (oracle.adjoint), the
dot-product identity with the GPU projection, and central differences.

The adjoint has no counterpart in the reference; tolerances are those of a
converged solve (eps_rel 1e-13): 1e-9 relative to the largest entry."""
import numpy as np
import pytest

from cases import adjoint_case, topology, velocity

pytestmark = pytest.mark.gpu

CASES = [("box2", 0), ("walls", 0), ("stretched", 0), ("aniso", 0), ("box2", 1)]


def _solver(g, ibtype=0, bp=None, **kw):
    import mgp
    kw.setdefault("nouter", 4000)
    kw.setdefault("eps_rel", 1e-13)
    return mgp.PressureSolver(g, ib={0: "noib", 1: "ghostcell"}[ibtype], bp=bp, **kw)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name,ibtype", CASES)
def test_adjoint_matches_sparse_oracle(oracle_mod, name, ibtype, prec):
    g, o, st, ops, bp = adjoint_case(oracle_mod, name, ibtype, 40)
    s = _solver(g, ibtype, bp, precision=prec)
    rng = np.random.default_rng(41)
    wx = rng.standard_normal(ops.nface)
    wu, wv, ww = ops.scatter(wx)
    r = s.adjoint(wu, wv, ww, 1e-3)
    assert r["converged"] and r["maxrhs"] <= 1e-13 * r["maxrhs0"]
    ref = ops.adjoint(wx)
    got = ops.gather(wu, wv, ww)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()
    # nothing outside the unknown faces
    rest = [a.copy() for a in (wu, wv, ww)]
    for a, pos in zip(rest, ops.faces):
        a[pos] = 0.0
    assert all(not a.any() for a in rest)


@pytest.mark.parametrize("name,ibtype", [("walls", 0), ("stretched", 0), ("box2", 1)])
def test_adjoint_dot_product_with_gpu_projection(oracle_mod, name, ibtype):
    """<w, P u> = <P^T w, u>, both sides on the GPU."""
    g, o, st, ops, bp = adjoint_case(oracle_mod, name, ibtype, 42)
    s = _solver(g, ibtype, bp)
    rng = np.random.default_rng(43)
    x = rng.uniform(-1, 1, ops.nface)
    wx = rng.standard_normal(ops.nface)
    u, v, w = ops.scatter(x)
    o.exchange([u, v, w], layers=2, corners=True)
    r = s.mgpoisl(u, v, w, g.zeros(), 1e-3)
    assert r["converged"]
    wu, wv, ww = ops.scatter(wx)
    s.adjoint(wu, wv, ww, 1e-3)
    lhs = wx @ ops.gather(u, v, w)
    rhs = ops.gather(wu, wv, ww) @ x
    assert abs(lhs - rhs) <= 1e-10 * np.abs(wx).sum() * np.abs(x).max()


def _ghost_transpose(o, g, wfull):
    """Oracle-side transpose of the 2-layer exchange with corners."""
    inner = g.interior_mask()
    ids = np.where(inner, np.arange(g.n3d, dtype=np.float64), -1.0)
    o.exchange([ids], layers=2, corners=True)
    dst = np.flatnonzero((ids >= 0) & ~inner)
    src = ids[dst].astype(np.int64)
    acc = wfull.copy()
    np.add.at(acc, src, wfull[dst])
    acc[~inner] = 0.0
    return acc


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_autograd_projection_gradient(oracle_mod, prec):
    """dJ/du of J = sum(w * u') over every stored entry (ghosts included)
    through torch.autograd equals the oracle's transposed projection."""
    import torch
    import mgp.autograd as ag
    g, o, st, ops, bp = adjoint_case(oracle_mod, "stretched", 0, 44)
    s = _solver(g, precision=prec)
    rng = np.random.default_rng(45)
    x = rng.uniform(-1, 1, ops.nface)
    wf = [rng.standard_normal(g.n3d) for _ in range(3)]
    uvw = [torch.from_numpy(a).cuda().requires_grad_(True) for a in ops.scatter(x)]
    out = ag.project(s, *uvw, 1e-3)
    J = sum((torch.from_numpy(wa).cuda() * t).sum() for wa, t in zip(wf, out[:3]))
    J.backward()
    ref = ops.adjoint(ops.gather(*[_ghost_transpose(o, g, a) for a in wf]))
    grads = [t.grad.cpu().numpy() for t in uvw]
    assert np.abs(ops.gather(*grads) - ref).max() <= 1e-9 * np.abs(ref).max()
    assert s.last_adjoint_stats["converged"]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_central_differences_fixed_iterations(prec):
    """J = <w, P(u)> against central differences (h = 1e-6) in random
    directions, with the iteration count fixed during the difference."""
    import torch
    import mgp.autograd as ag
    from mgp.grids import uniform_box
    g = uniform_box((4, 4, 2), (12, 12, 12), length=(2.0, 2.0, 1.0))
    s = _solver(g, precision=prec, nouter=3000, eps_rel=1e-13)
    rng = np.random.default_rng(46)
    inner = torch.from_numpy(g.interior_mask()).cuda()
    u0 = [torch.where(inner, torch.from_numpy(rng.uniform(-1, 1, g.n3d)).cuda(), 0.0)
          for _ in range(3)]
    wts = [torch.where(inner, torch.from_numpy(rng.standard_normal(g.n3d)).cuda(), 0.0)
           for _ in range(3)]
    x = [t.clone().requires_grad_(True) for t in u0]
    out = ag.project(s, *x, 1e-3)
    sum((a * b).sum() for a, b in zip(wts, out[:3])).backward()
    n_fix = s.last_stats["nouter_done"]
    s.set_iterations(nouter=n_fix, eps_rel=0.0)
    h = 1e-6

    def J(fields):
        with torch.no_grad():
            o3 = ag.project(s, *fields, 1e-3)
        return float(sum((a * b).sum() for a, b in zip(wts, o3[:3])))

    for d in range(8):
        dirs = [torch.where(inner, torch.from_numpy(rng.standard_normal(g.n3d)).cuda(), 0.0)
                for _ in range(3)]
        fd = (J([a + h * b for a, b in zip(u0, dirs)]) -
              J([a - h * b for a, b in zip(u0, dirs)])) / (2 * h)
        ad = float(sum((t.grad * b).sum() for t, b in zip(x, dirs)))
        assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1.0), (d, fd, ad)


def test_checkpointed_rollout_matches_plain():
    """20 nonlinear steps u <- P(u + dt tanh(u)): the gradient with
    checkpoints every 5 steps equals the one that keeps every step."""
    import torch
    import mgp.autograd as ag
    from mgp.grids import uniform_box
    g = uniform_box((2, 2, 2), (8, 8, 8))
    s = _solver(g, nouter=3000, eps_rel=1e-13)
    rng = np.random.default_rng(47)
    inner = torch.from_numpy(g.interior_mask()).cuda()
    u0 = [torch.where(inner, torch.from_numpy(rng.uniform(-1, 1, g.n3d)).cuda(), 0.0)
          for _ in range(3)]
    wts = [torch.from_numpy(rng.standard_normal(g.n3d)).cuda() for _ in range(3)]

    def step(u, v, w):
        a, b, c, _ = ag.project(s, u + 0.1 * torch.tanh(u), v + 0.1 * torch.tanh(v),
                                w + 0.1 * torch.tanh(w), 1e-3)
        return a, b, c

    grads = []
    for every in (0, 5):
        x = [t.clone().requires_grad_(True) for t in u0]
        out = ag.rollout(step, x, 20, checkpoint_every=every)
        sum((a * b).sum() for a, b in zip(wts, out)).backward()
        grads.append([t.grad.clone() for t in x])
    for a, b in zip(*grads):
        assert torch.allclose(a, b, rtol=0, atol=1e-10 * float(a.abs().max()))
