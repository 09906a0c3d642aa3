"""The oracle's par_ftoc_norm (src/ib/par_ftoc_mod.F90:35-475) is pinned by
what the transfer must do geometrically, independent of the reference's
connection table:

* placement: with sentinels on the coarse level, the positions written are
  exactly the coarse face positions that coincide with a fine PAR face
  (same staggered coordinate along the normal, periodic images included,
  cell centres inside the fine face), found by brute force over every
  coarse grid;
* the area weights: a linear field is restricted exactly (its area-weighted
  mean over a 2 x 2 group is its value at the coarse face centre), and the
  coarse flux value x coarse face area equals the sum of the four fine
  fluxes for a random field;
* the sum variant: four times a constant.
"""
import numpy as np
import pytest

from cases import reference_topology, three_level, two_level

SENTINEL = 1.0e30


def _grids(name):
    if name == "two_level":
        return two_level(periodic=False)
    if name == "two_level_periodic":
        return two_level(periodic=True)
    if name == "three_level":
        return three_level()
    return reference_topology(name)


def _coords(g, b, comp):
    """Coordinates of the stored positions of component comp (0 u, 1 v, 2 w)
    of block b: three 1-D arrays indexed by the 0-based i, j, k."""
    cs = []
    for a, (c, s, o) in enumerate(((g.x, g.xstag, g.ip1x), (g.y, g.ystag, g.ip1y),
                                   (g.z, g.zstag, g.ip1z))):
        n = (g.ii, g.jj, g.kk)[a][b - 1]
        arr = s if a == comp else c
        cs.append(arr[o[b - 1]:o[b - 1] + n])
    return cs


def _domain(g):
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for b in range(1, g.ngrid + 1):
        if g.level[b - 1] != g.minlevel:
            continue
        for a, (s, o, n) in enumerate(((g.xstag, g.ip1x, g.ii), (g.ystag, g.ip1y, g.jj),
                                       (g.zstag, g.ip1z, g.kk))):
            st = s[o[b - 1]:o[b - 1] + n[b - 1]]
            lo[a] = min(lo[a], st[1])
            hi[a] = max(hi[a], st[-3])
    return hi - lo


def _expected(g, level):
    """(grid, comp, i, j, k) of every coarse face position on level-1 that
    coincides with a PAR face of a grid on `level`, brute force."""
    L = _domain(g)
    out = set()
    for gf in range(1, g.ngrid + 1):
        if g.level[gf - 1] != level:
            continue
        for f in range(6):
            if g.bc[gf - 1, f] != 8:
                continue
            n = f // 2
            cf = _coords(g, gf, n)
            nf = len(cf[n])
            X = cf[n][1 if f % 2 == 0 else nf - 3]       # the fine face plane
            t = [a for a in range(3) if a != n]
            ext = [(cf[a][1], cf[a][len(cf[a]) - 3]) for a in t]   # interior extent
            for gc in range(1, g.ngrid + 1):
                if g.level[gc - 1] != level - 1:
                    continue
                cc = _coords(g, gc, n)
                nc = [len(c) for c in cc]
                d = (cc[n] - X) / L[n]
                hit = np.abs(d - np.round(d)) < 1e-9
                for ic in np.nonzero(hit)[0]:
                    if ic < 1 or ic > nc[n] - 3:
                        continue
                    r1 = [q for q in range(2, nc[t[0]] - 2)
                          if ext[0][0] < cc[t[0]][q] < ext[0][1]]
                    r2 = [q for q in range(2, nc[t[1]] - 2)
                          if ext[1][0] < cc[t[1]][q] < ext[1][1]]
                    for q1 in r1:
                        for q2 in r2:
                            idx = [0, 0, 0]
                            idx[n], idx[t[0]], idx[t[1]] = ic, q1, q2
                            out.add((gc, n, *idx))
    return out


def _linear(g, level_fill, coef):
    """u, v, w = coef . (1, x, y, z) at each component's positions on the
    grids of level_fill, SENTINEL elsewhere."""
    fields = [np.full(g.n3d, SENTINEL) for _ in range(3)]
    for b in range(1, g.ngrid + 1):
        if g.level[b - 1] != level_fill:
            continue
        for comp in range(3):
            x, y, z = _coords(g, b, comp)
            g.block_view(fields[comp], b)[...] = (coef[0] + coef[1] * x[:, None, None] +
                                                  coef[2] * y[None, :, None] +
                                                  coef[3] * z[None, None, :])
    return fields


@pytest.mark.parametrize("name", ["two_level", "two_level_periodic", "three_level",
                                  "Parker_micro", "Checkblock"])
def test_par_ftoc_writes_the_coincident_coarse_faces(oracle_mod, name):
    g = _grids(name)
    o = oracle_mod.Oracle(g)
    coef = (0.3, 1.7, -0.8, 0.45)
    seen = 0
    for level in range(g.minlevel + 1, g.maxlevel + 1):
        u, v, w = _linear(g, level, coef)
        o.par_ftoc_norm(level, u, v, w)
        got = set()
        for comp, fld in enumerate((u, v, w)):
            for b in range(1, g.ngrid + 1):
                if g.level[b - 1] != level - 1:
                    continue
                view = g.block_view(fld, b)
                for i, j, k in zip(*np.nonzero(view != SENTINEL)):
                    got.add((b, comp, i, j, k))
        exp = _expected(g, level)
        assert got == exp, (name, level, len(got), len(exp))
        seen += len(got)
    assert seen > 0


@pytest.mark.parametrize("name", ["two_level", "three_level"])
def test_par_ftoc_linear_values(oracle_mod, name):
    """Away from periodic images (non-periodic grid, or fields without an
    x-dependence across the normal), the restricted value of a linear field
    is the field at the coarse face centre."""
    g = _grids(name)
    o = oracle_mod.Oracle(g)
    coef = (0.3, 0.0, -0.8, 0.45) if name == "three_level" else (0.3, 1.7, -0.8, 0.45)
    for level in range(g.minlevel + 1, g.maxlevel + 1):
        for comp in range(3):
            c = list(coef)
            if name == "three_level":
                c = [0.3, 0.6, -0.8, 0.45]
                c[1 + comp] = 0.0      # no dependence along the normal
            u, v, w = _linear(g, level, c)
            o.par_ftoc_norm(level, u, v, w)
            fld = (u, v, w)[comp]
            for b in range(1, g.ngrid + 1):
                if g.level[b - 1] != level - 1:
                    continue
                view = g.block_view(fld, b)
                x, y, z = _coords(g, b, comp)
                for i, j, k in zip(*np.nonzero(view != SENTINEL)):
                    want = c[0] + c[1] * x[i] + c[2] * y[j] + c[3] * z[k]
                    assert abs(view[i, j, k] - want) <= 1e-13 * (1 + abs(want)), \
                        (b, comp, i, j, k, view[i, j, k], want)


@pytest.mark.parametrize("name", ["two_level", "three_level", "Parker_micro"])
def test_par_ftoc_conserves_flux(oracle_mod, name):
    """coarse value x coarse face area (the coarse ddy ddz, ddx ddz, ddx ddy)
    = sum of the four fine values x fine face areas, random fields."""
    g = _grids(name)
    o = oracle_mod.Oracle(g)
    rng = np.random.default_rng(41)
    dd = (g.ddx, g.ddy, g.ddz)
    offs = (g.ip1x, g.ip1y, g.ip1z)
    for level in range(g.minlevel + 1, g.maxlevel + 1):
        u, v, w = (rng.uniform(-1, 1, g.n3d) for _ in range(3))
        before = [a.copy() for a in (u, v, w)]
        o.par_ftoc_norm(level, u, v, w)
        for (gc, n, ic, jc, kc) in _expected(g, level):
            t = [a for a in range(3) if a != n]
            area_c = 1.0
            idx_c = (ic, jc, kc)
            for a in t:
                area_c *= dd[a][offs[a][gc - 1] + idx_c[a]]
            # the fine grid whose PAR face this is: search the level's grids
            flux = None
            for gf in range(1, g.ngrid + 1):
                if g.level[gf - 1] != level or g.iparent[gf - 1] == 0:
                    continue
                cf = _coords(g, gf, n)
                ccoord = _coords(g, gc, n)
                centre = [ccoord[a][idx_c[a]] for a in range(3)]
                ok = True
                sel = []
                for a in t:
                    q = [p for p in range(2, len(cf[a]) - 2)
                         if abs(cf[a][p] - centre[a]) < 0.75 * dd[a][offs[a][gf - 1] + p]]
                    if len(q) != 2:
                        ok = False
                        break
                    sel.append(q)
                if not ok:
                    continue
                L = _domain(g)
                nf = len(cf[n])
                planes = [p for p in (1, nf - 3) if g.bc[gf - 1, 2 * n + (p != 1)] == 8 and
                          abs(((cf[n][p] - centre[n]) / L[n]) -
                              np.round((cf[n][p] - centre[n]) / L[n])) < 1e-9]
                if not planes:
                    continue
                view = g.block_view(before[n], gf)
                flux = 0.0
                for q1 in sel[0]:
                    for q2 in sel[1]:
                        idx = [0, 0, 0]
                        idx[n], idx[t[0]], idx[t[1]] = planes[0], q1, q2
                        fa = dd[t[0]][offs[t[0]][gf - 1] + q1] * dd[t[1]][offs[t[1]][gf - 1] + q2]
                        flux += view[tuple(idx)] * fa
                break
            assert flux is not None
            got = g.block_view((u, v, w)[n], gc)[ic, jc, kc] * area_c
            assert abs(got - flux) <= 1e-13 * max(1.0, abs(flux)), (gc, n, ic, jc, kc)


def test_par_ftoc_sum_variant(oracle_mod):
    g = _grids("three_level")
    o = oracle_mod.Oracle(g)
    for level in range(g.minlevel + 1, g.maxlevel + 1):
        u, v, w = _linear(g, level, (0.375, 0.0, 0.0, 0.0))
        o.par_ftoc_norm(level, u, v, w, sum=True)
        written = 0
        for fld in (u, v, w):
            for b in range(1, g.ngrid + 1):
                if g.level[b - 1] == level - 1:
                    view = g.block_view(fld, b)
                    vals = view[view != SENTINEL]
                    assert np.all(vals == 1.5)
                    written += vals.size
        assert written > 0
