"""The oracle's tstle4_par (src/flow/tstle4_mod.F90:532-1753), which restates
the reference's hand-written per-face code in one form for all faces, is
pinned here by

* independent straight-line restatements of two faces in the reference's
  statement order (FRONT, a low x face, with its w and v cells; TOP, a high
  z face, with its u and v cells), compared bitwise on random fields and
  random metrics;
* mirror symmetry between the low and high face of each axis (bitwise, on a
  uniform grid, where the reference's per-face expressions are mirror
  images of each other);
* rotation of the axes (x, y, z) -> (y, z, x), which maps FRONT to BOTTOM and
  BOTTOM to RIGHT (to rounding: the cell factor's grouping is not rotation
  invariant);
* a uniform velocity, for which both the upwind term and the difference of
  the coarse-consistent and the standard flux vanish exactly.
"""
import numpy as np
import pytest

from mgp.grids import uniform_box

FRONT, BACK, RIGHT, LEFT, BOTTOM, TOP = range(6)


def _block(n=(8, 10, 12), par=(), seed=0, random_metrics=True):
    g = uniform_box((1, 1, 1), n)
    for f in par:
        g.bc[0, f] = 8
    if random_metrics:
        rng = np.random.default_rng(seed + 1000)
        for a in ("dx", "dy", "dz", "ddx", "ddy", "ddz"):
            getattr(g, a)[:] = rng.uniform(0.5, 1.5, getattr(g, a).size)
        g.rdx, g.rdy, g.rdz = 1.0 / g.dx, 1.0 / g.dy, 1.0 / g.dz
        g.rddx, g.rddy, g.rddz = 1.0 / g.ddx, 1.0 / g.ddy, 1.0 / g.ddz
    return g


def _fields(g, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1, 1, g.n3d) for _ in range(3)]


def _par(oracle_mod, g, u, v, w):
    o = oracle_mod.Oracle(g)
    uo, vo, wo = g.zeros(), g.zeros(), g.zeros()
    o.tstle4_par(uo, vo, wo, u, v, w)
    return uo, vo, wo


class _A:
    """1-based (k, j, i) access to block 1 of a flat field."""

    def __init__(self, g, f):
        self.v = g.block_view(f, 1)

    def __call__(self, k, j, i):
        return self.v[i - 1, j - 1, k - 1]

    def add(self, k, j, i, x):
        self.v[i - 1, j - 1, k - 1] = self.v[i - 1, j - 1, k - 1] + x


def _front_literal(g, u, v, w):
    """FRONT (face 1) in the reference's statement order: the upwind terms
    at i = 4 (:571-606), then the momentum-conserving w cell and v cell
    (:935-1089)."""
    uo, vo, wo = g.zeros(), g.zeros(), g.zeros()
    U, V, W = _A(g, u), _A(g, v), _A(g, w)
    VO, WO = _A(g, vo), _A(g, wo)
    kk, jj = int(g.kk[0]), int(g.jj[0])
    dy, dz, ddy, ddz = (np.r_[0.0, a] for a in (g.dy, g.dz, g.ddy, g.ddz))   # 1-based
    rdy, rdz, rddx, rddy, rddz = (np.r_[0.0, a] for a in (g.rdy, g.rdz, g.rddx, g.rddy, g.rddz))
    i = 4
    for j in range(3, jj - 1):
        fkdtv = -1.0 * rddx[i] * rdy[j] * 1.0
        fkdtw = -1.0 * rddx[i] * rddy[j] * 1.0
        for k in range(3, kk - 1):
            avx = dy[j] * ddz[k]
            awx = ddy[j] * dz[k]
            fvw = avx * (U(k, j, i - 1) + U(k, j + 1, i - 1)) * 0.5
            q = 0.5 * (fvw - abs(fvw)) * (0.5 * V(k, j, i) - 1.5 * V(k, j, i - 1) + V(k, j, i - 2)) * 0.5 * 0.5
            VO.add(k, j, i, fkdtv * rddz[k] * (-q))
            VO.add(k, j, i - 1, fkdtv * rddz[k] * (+q))
            fww = awx * (U(k, j, i - 1) + U(k + 1, j, i - 1)) * 0.5
            q = 0.5 * (fww - abs(fww)) * (0.5 * W(k, j, i) - 1.5 * W(k, j, i - 1) + W(k, j, i - 2)) * 0.5 * 0.5
            WO.add(k, j, i, fkdtw * rdz[k] * (-q))
            WO.add(k, j, i - 1, fkdtw * rdz[k] * (+q))
    i = 3
    # w cell: standard flux at i-1, new flux at i, distribution, update
    wcw = {}
    for j in range(3, jj - 1):
        for k in range(2, kk - 1):
            fww = ddy[j] * dz[k] * (U(k, j, i - 1) + U(k + 1, j, i - 1)) * 0.5
            wcw[(k, j, i - 1)] = 0.5 * fww * (W(k, j, i) + W(k, j, i - 1))
    for j in range(3, jj - 1, 2):
        ks = [2] + list(range(4, kk - 3, 2)) + [kk - 2]
        for k in ks:
            awx = ddy[j] * dz[k]
            t = lambda kq, jq: U(kq, jq, i - 1)
            if k == 2:
                s = 2 * t(k, j) + 2 * t(k, j + 1) + t(k + 1, j) + t(k + 1, j + 1) + t(k + 2, j) + t(k + 2, j + 1)
            elif k == kk - 2:
                s = t(k - 1, j) + t(k - 1, j + 1) + t(k, j) + t(k, j + 1) + 2 * t(k + 1, j) + 2 * t(k + 1, j + 1)
            else:
                s = (t(k - 1, j) + t(k - 1, j + 1) + t(k, j) + t(k, j + 1) + t(k + 1, j) + t(k + 1, j + 1)
                     + t(k + 2, j) + t(k + 2, j + 1))
            fww = awx * s * 0.125
            wcw[(k, j, i)] = 0.5 * fww * (W(k, j, i) + W(k, j, i - 1))
            wcw[(k, j + 1, i)] = 0.5 * fww * (W(k, j + 1, i) + W(k, j + 1, i - 1))
    for j in range(3, jj - 1):
        for k in range(2, kk - 3, 2):
            wcw[(k + 1, j, i)] = 0.5 * (wcw[(k, j, i)] + wcw[(k + 2, j, i)])
    for j in range(3, jj - 1):
        fkdtw = -1.0 * rddx[i] * rddy[j] * 1.0
        for k in range(2, kk - 1):
            WO.add(k, j, i, fkdtw * rdz[k] * (wcw[(k, j, i - 1)] - wcw[(k, j, i)]))
    # v cell
    wcv = {}
    for j in range(2, jj - 1):
        for k in range(3, kk - 1):
            fvw = ddz[k] * dy[j] * (U(k, j, i - 1) + U(k, j + 1, i - 1)) * 0.5
            wcv[(k, j, i - 1)] = 0.5 * fvw * (V(k, j, i) + V(k, j, i - 1))
    js = [2] + list(range(4, jj - 3, 2)) + [jj - 2]
    for j in js:
        for k in range(3, kk - 2, 2):
            avx = ddz[k] * dy[j]
            t = lambda kq, jq: U(kq, jq, i - 1)
            if j == 2:
                s = 2 * t(k, j) + 2 * t(k + 1, j) + t(k, j + 1) + t(k + 1, j + 1) + t(k, j + 2) + t(k + 1, j + 2)
            elif j == jj - 2:
                s = t(k, j - 1) + t(k + 1, j - 1) + t(k, j) + t(k + 1, j) + 2 * t(k, j + 1) + 2 * t(k + 1, j + 1)
            else:
                s = (t(k, j - 1) + t(k + 1, j - 1) + t(k, j) + t(k + 1, j) + t(k, j + 1) + t(k + 1, j + 1)
                     + t(k, j + 2) + t(k + 1, j + 2))
            fvw = avx * s * 0.125
            wcv[(k, j, i)] = 0.5 * fvw * (V(k, j, i) + V(k, j, i - 1))
            wcv[(k + 1, j, i)] = 0.5 * fvw * (V(k + 1, j, i) + V(k + 1, j, i - 1))
    for j in range(2, jj - 3, 2):
        for k in range(3, kk - 1):
            wcv[(k, j + 1, i)] = 0.5 * (wcv[(k, j, i)] + wcv[(k, j + 2, i)])
    for j in range(2, jj - 1):
        fkdtv = -1.0 * rddx[i] * rdy[j] * 1.0
        for k in range(3, kk - 1):
            VO.add(k, j, i, fkdtv * rddz[k] * (-wcv[(k, j, i)] + wcv[(k, j, i - 1)]))
    return uo, vo, wo


def _top_literal(g, u, v, w):
    """TOP (face 6) in the reference's statement order: the upwind terms at
    k = kk-2 (:738-776), then the momentum-conserving u cell and v cell
    (:1092-1253)."""
    uo, vo, wo = g.zeros(), g.zeros(), g.zeros()
    U, V, W = _A(g, u), _A(g, v), _A(g, w)
    UO, VO = _A(g, uo), _A(g, vo)
    kk, jj, ii = int(g.kk[0]), int(g.jj[0]), int(g.ii[0])
    dx, dy, ddx, ddy = (np.r_[0.0, a] for a in (g.dx, g.dy, g.ddx, g.ddy))
    rdx, rdy, rddx, rddy, rddz = (np.r_[0.0, a] for a in (g.rdx, g.rdy, g.rddx, g.rddy, g.rddz))
    k = kk - 2
    for i in range(3, ii - 1):
        for j in range(3, jj - 1):
            fkdtu = -1.0 * rddy[j] * rdx[i] * 1.0
            fkdtv = -1.0 * rddx[i] * rdy[j] * 1.0
            auz = dx[i] * ddy[j]
            avz = ddx[i] * dy[j]
            fub = auz * (W(k - 1, j, i) + W(k - 1, j, i + 1)) * 0.5
            q = 0.5 * (fub + abs(fub)) * (0.5 * U(k - 1, j, i) - 1.5 * U(k, j, i) + U(k + 1, j, i)) * 0.5 * 0.5
            UO.add(k, j, i, fkdtu * rddz[k] * (-q))
            UO.add(k - 1, j, i, fkdtu * rddz[k] * (+q))
            fvb = avz * (W(k - 1, j, i) + W(k - 1, j + 1, i)) * 0.5
            q = 0.5 * (fvb + abs(fvb)) * (0.5 * V(k - 1, j, i) - 1.5 * V(k, j, i) + V(k + 1, j, i)) * 0.5 * 0.5
            VO.add(k, j, i, fkdtv * rddz[k] * (-q))
            VO.add(k - 1, j, i, fkdtv * rddz[k] * (+q))
    # u cell
    wcu = {}
    for i in range(2, ii - 1):
        for j in range(3, jj - 1):
            fut = dx[i] * ddy[j] * (W(k, j, i) + W(k, j, i + 1)) * 0.5
            wcu[(k + 1, j, i)] = 0.5 * fut * (U(k, j, i) + U(k + 1, j, i))
    t = lambda jq, iq: W(k, jq, iq)
    for i in [2] + list(range(4, ii - 3, 2)) + [ii - 2]:
        for j in range(3, jj - 2, 2):
            auz = dx[i] * ddy[j]
            if i == 2:
                s = 2 * t(j, i) + 2 * t(j + 1, i) + t(j, i + 1) + t(j + 1, i + 1) + t(j, i + 2) + t(j + 1, i + 2)
            elif i == ii - 2:
                s = t(j, i - 1) + t(j + 1, i - 1) + t(j, i) + t(j + 1, i) + 2 * t(j, i + 1) + 2 * t(j + 1, i + 1)
            else:
                s = (t(j, i - 1) + t(j + 1, i - 1) + t(j, i) + t(j + 1, i) + t(j, i + 1) + t(j + 1, i + 1)
                     + t(j, i + 2) + t(j + 1, i + 2))
            fut = auz * s * 0.125
            wcu[(k, j, i)] = 0.5 * fut * (U(k, j, i) + U(k + 1, j, i))
            wcu[(k, j + 1, i)] = 0.5 * fut * (U(k, j + 1, i) + U(k + 1, j + 1, i))
    for i in range(2, ii - 3, 2):
        for j in range(3, jj - 1):
            wcu[(k, j, i + 1)] = 0.5 * (wcu[(k, j, i)] + wcu[(k, j, i + 2)])
    for i in range(2, ii - 1):
        for j in range(3, jj - 1):
            fkdtu = -1.0 * rddy[j] * rdx[i] * 1.0
            UO.add(k, j, i, fkdtu * rddz[k] * (wcu[(k, j, i)] - wcu[(k + 1, j, i)]))
    # v cell
    wcv = {}
    for i in range(3, ii - 1):
        for j in range(2, jj - 1):
            fvt = ddx[i] * dy[j] * (W(k, j, i) + W(k, j + 1, i)) * 0.5
            wcv[(k + 1, j, i)] = 0.5 * fvt * (V(k, j, i) + V(k + 1, j, i))
    for i in range(3, ii - 2, 2):
        for j in [2] + list(range(4, jj - 3, 2)) + [jj - 2]:
            avz = ddx[i] * dy[j]
            if j == 2:
                s = 2 * t(j, i) + 2 * t(j, i + 1) + t(j + 1, i) + t(j + 1, i + 1) + t(j + 2, i) + t(j + 2, i + 1)
            elif j == jj - 2:
                s = t(j - 1, i) + t(j - 1, i + 1) + t(j, i) + t(j, i + 1) + 2 * t(j + 1, i) + 2 * t(j + 1, i + 1)
            else:
                s = (t(j - 1, i) + t(j - 1, i + 1) + t(j, i) + t(j, i + 1) + t(j + 1, i) + t(j + 1, i + 1)
                     + t(j + 2, i) + t(j + 2, i + 1))
            fvt = avz * s * 0.125
            wcv[(k, j, i)] = 0.5 * fvt * (V(k, j, i) + V(k + 1, j, i))
            wcv[(k, j, i + 1)] = 0.5 * fvt * (V(k, j, i + 1) + V(k + 1, j, i + 1))
    for i in range(3, ii - 1):
        for j in range(2, jj - 3, 2):
            wcv[(k, j + 1, i)] = 0.5 * (wcv[(k, j, i)] + wcv[(k, j + 2, i)])
    for i in range(3, ii - 1):
        for j in range(2, jj - 1):
            fkdtv = -1.0 * rddx[i] * rdy[j] * 1.0
            VO.add(k, j, i, fkdtv * rddz[k] * (wcv[(k, j, i)] - wcv[(k + 1, j, i)]))
    return uo, vo, wo


@pytest.mark.parametrize("seed", [1, 2])
def test_front_face_matches_statement_order_restatement(oracle_mod, seed):
    g = _block(par=(FRONT,), seed=seed)
    u, v, w = _fields(g, seed)
    got = _par(oracle_mod, g, u, v, w)
    ref = _front_literal(g, u, v, w)
    for a, b, nm in zip(got, ref, "uvw"):
        assert np.array_equal(a, b), nm
    assert np.count_nonzero(got[1]) > 0 and np.count_nonzero(got[2]) > 0
    assert not np.any(got[0])


@pytest.mark.parametrize("seed", [3, 4])
def test_top_face_matches_statement_order_restatement(oracle_mod, seed):
    g = _block(par=(TOP,), seed=seed)
    u, v, w = _fields(g, seed)
    got = _par(oracle_mod, g, u, v, w)
    ref = _top_literal(g, u, v, w)
    for a, b, nm in zip(got, ref, "uvw"):
        assert np.array_equal(a, b), nm
    assert not np.any(got[2])


def _mirror(g, f, axis, odd):
    """Reflect a field of block 1 across the block's centre plane normal to
    `axis`; `odd`: the velocity component along that axis (staggered:
    position p goes to n - p, and the sign flips)."""
    out = f.copy()
    src = g.block_view(f, 1)
    dst = g.block_view(out, 1)
    n = src.shape[axis]
    if odd:
        idx = np.r_[np.arange(n - 2, -1, -1), n - 1]   # p -> n-2-p (0-based), last kept
        moved = np.take(src, idx, axis=axis)
        dst[...] = -moved
    else:
        dst[...] = np.flip(src, axis=axis)
    return out


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_low_and_high_faces_mirror(oracle_mod, axis):
    """On a uniform grid the high face's terms are the mirror image of the
    low face's: uo etc. of the mirrored fields equal the mirrored uo etc."""
    g = _block(n=(8, 8, 8), par=(2 * axis,), random_metrics=False)
    gm = _block(n=(8, 8, 8), par=(2 * axis + 1,), random_metrics=False)
    u, v, w = _fields(g, 10 + axis)
    comps = [u, v, w]
    mir = [_mirror(g, c, axis, odd=(q == axis)) for q, c in enumerate(comps)]
    lo = _par(oracle_mod, g, *comps)
    hi = _par(oracle_mod, gm, *mir)
    for q in range(3):
        expect = _mirror(g, lo[q], axis, odd=(q == axis))
        inner = g.block_view(np.ones(g.n3d, bool), 1).copy()
        # the staggered component's last position has no mirror partner
        if q == axis:
            sl = [slice(None)] * 3
            sl[axis] = -1
            inner[tuple(sl)] = False
        a = g.block_view(hi[q], 1)[inner]
        b = g.block_view(expect, 1)[inner]
        assert np.array_equal(a, b), "uvw"[q]
        assert np.count_nonzero(a) > 0 or q == axis


def _rotate(g, f):
    """(x, y, z) -> (y, z, x) on a cube: new[i', j', k'] = old[k', i', j']."""
    out = f.copy()
    g.block_view(out, 1)[...] = g.block_view(f, 1).transpose(1, 2, 0)
    return out


@pytest.mark.parametrize("face,rotated", [(FRONT, BOTTOM), (BOTTOM, RIGHT), (BACK, TOP),
                                          (TOP, LEFT)])
def test_axis_rotation(oracle_mod, face, rotated):
    """Rotating the fields and the PAR face (x -> z', y -> x', z -> y')
    rotates the terms; the cell factor's (x*y)*z grouping makes this exact
    only to rounding."""
    g = _block(n=(8, 8, 8), par=(face,), random_metrics=False)
    gr = _block(n=(8, 8, 8), par=(rotated,), random_metrics=False)
    u, v, w = _fields(g, 20 + face)
    base = _par(oracle_mod, g, u, v, w)
    # new u' = old v, v' = old w, w' = old u, positions rotated
    ur, vr, wr = _rotate(g, v), _rotate(g, w), _rotate(g, u)
    rot = _par(oracle_mod, gr, ur, vr, wr)
    expect = [_rotate(g, base[1]), _rotate(g, base[2]), _rotate(g, base[0])]
    for a, b, nm in zip(rot, expect, "uvw"):
        scale = np.abs(b).max() + 1e-300
        assert np.abs(a - b).max() <= 1e-13 * scale, nm
    assert any(np.count_nonzero(a) for a in rot)


def test_uniform_velocity_adds_nothing(oracle_mod):
    """Constant u, v, w on a uniform grid: the upwind term's 0.5c - 1.5c + c
    and the difference of the coarse-consistent and the standard flux are
    0 (on a stretched grid the pair's shared flux uses the metrics of its
    first cell, so they differ there)."""
    g = _block(n=(8, 10, 12), par=tuple(range(6)), random_metrics=False)
    u, v, w = (np.full(g.n3d, c) for c in (0.75, -0.5, 0.25))
    uo, vo, wo = _par(oracle_mod, g, u, v, w)
    assert not np.any(uo) and not np.any(vo) and not np.any(wo)
