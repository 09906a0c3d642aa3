"""GPU par_ftoc_norm (src/ib/par_ftoc_mod.F90:35-475) through the C ABI
(mgp_par_ftoc_device): the face-normal velocity at PAR faces into the
coincident coarse face positions, bitwise equal to the oracle on random
fields, area-weighted and summed, level by level from fine to coarse."""
import numpy as np
import pytest

from cases import config3, reference_topology, three_level, two_level

pytestmark = pytest.mark.gpu


def _grids(name):
    if name == "two_level":
        return two_level(periodic=False)
    if name == "two_level_periodic":
        return two_level(periodic=True)
    if name == "three_level":
        return three_level()
    if name == "config3":
        return config3()
    return reference_topology(name)


@pytest.mark.parametrize("name", ["two_level", "two_level_periodic", "three_level",
                                  "Parker_micro", "Checkblock", "config3"])
@pytest.mark.parametrize("summed", [False, True])
def test_par_ftoc_bitwise(oracle_mod, name, summed):
    import torch
    import mgp
    g = _grids(name)
    print(f"[par_ftoc] {name} sum={summed}: {g.ngrid} grids", flush=True)
    o = oracle_mod.Oracle(g)
    rng = np.random.default_rng(43)
    u, v, w = (rng.uniform(-1, 1, g.n3d) for _ in range(3))
    t = [torch.from_numpy(a.copy()).cuda() for a in (u, v, w)]
    ref = [a.copy() for a in (u, v, w)]
    s = mgp.PressureSolver(g)
    for level in range(g.maxlevel, g.minlevel, -1):
        o.par_ftoc_norm(level, u, v, w, sum=summed)
        s.par_ftoc_norm(*t, level, sum=summed)
    s.close()
    changed = 0
    for a, b, r, nm in zip((u, v, w), t, ref, "uvw"):
        got = b.cpu().numpy()
        assert np.array_equal(a, got), (nm, float(np.abs(a - got).max()))
        changed += int((a != r).sum())
    assert changed > 0
