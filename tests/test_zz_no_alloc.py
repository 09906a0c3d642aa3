"""mgp_divcal_device_noalloc (the divergence through the library's static
scratch array, no device allocation) against mgp_divcal_device (which
allocates its intermediate): bitwise equal on two calls in a row, the second
on a smaller grid, on the same GPU; handles beyond the scratch size are
rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from cases import blocked_mask, topology

pytestmark = pytest.mark.gpu


def _fields(g, seed):
    rng = np.random.default_rng(seed)
    inner = g.interior_mask()
    return [np.where(inner, rng.uniform(-1, 1, g.n3d), 0.0) for _ in range(3)]


def _noalloc(s, tu, tv, tw, fak):
    import torch
    out = torch.zeros_like(tu)
    s._bind_stream()
    rc = s.lib.mgp_divcal_device_noalloc(s.h, out.data_ptr(), tu.data_ptr(), tv.data_ptr(),
                                         tw.data_ptr(), C.c_double(fak))
    assert rc == 0, s.lib.mgp_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _case(g, seed, fak, **kw):
    import mgp
    import torch
    s = mgp.PressureSolver(g, **kw)
    tu, tv, tw = (torch.from_numpy(a).cuda() for a in _fields(g, seed))
    ref = s.divcal(tu, tv, tw, fak).cpu().numpy()
    return s, (tu, tv, tw), ref


def test_noalloc_divcal_matches_allocating_entry():
    import mgp
    big = mgp.uniform_box((4, 4, 4), (16, 16, 16))
    small = topology("walls")
    s1, f1, ref1 = _case(big, 401, 2.5)
    got1 = _noalloc(s1, *f1, 2.5)
    assert np.array_equal(got1, ref1)
    # second call, smaller grid, ghostcell form with BP and SDIV: the
    # scratch keeps the first call's values beyond the smaller grid's cells
    bp = 1.0 - blocked_mask(small, 402, 0.15)
    s2, f2, _ = _case(small, 403, 3.5, ib="ghostcell", bp=bp)
    rng = np.random.default_rng(404)
    s2.set_sdiv(np.where(small.interior_mask(), rng.standard_normal(small.n3d), 0.0))
    ref2 = s2.divcal(*f2, 3.5).cpu().numpy()
    got2 = _noalloc(s2, *f2, 3.5)
    assert np.array_equal(got2, ref2)
    assert np.count_nonzero(got2) > 0
    # and the first handle again, after the smaller one overwrote the start
    assert np.array_equal(_noalloc(s1, *f1, 2.5), ref1)
    s1.close()
    s2.close()


def test_noalloc_divcal_rejects_grids_beyond_the_scratch():
    import mgp
    import torch
    g = mgp.uniform_box((8, 8, 8), (26, 26, 26))     # 8,998,912 cells > 2**23
    s = mgp.PressureSolver(g)
    t = torch.zeros(g.n3d, dtype=torch.float64, device="cuda")
    rc = s.lib.mgp_divcal_device_noalloc(s.h, t.data_ptr(), t.data_ptr(), t.data_ptr(),
                                         t.data_ptr(), C.c_double(1.0))
    assert rc == 1
    assert b"scratch" in s.lib.mgp_last_error()
    s.close()
