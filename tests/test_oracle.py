"""Pin the CPU oracle before trusting it.

The reference's own tests hold no values for this path (its cases only check
exit codes), so each oracle function is pinned here by what the mathematics
fixes: closed forms, library routines on assembled matrices, brute force on
tiny inputs and exact invariants.  A plausible slip (wrong neighbour, wrong
sign, transposed coefficient, off-by-one plane) fails at least one test.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from cases import apply_walls, topology, velocity
from mgp.grids import FIX, NOS, uniform_box


def _ids(g):
    """interior index map: (g, i, j, k) 0-based interior -> flat 3-D index"""
    out = []
    for b in range(1, g.ngrid + 1):
        ii, jj, kk = g.ii[b - 1], g.jj[b - 1], g.kk[b - 1]
        base = g.ip3[b - 1]
        for i in range(ii - 4):
            for j in range(jj - 4):
                for k in range(kk - 4):
                    out.append(base + (k + 2) + (j + 2) * kk + (i + 2) * kk * jj)
    return np.array(out)


# ---------------------------------------------------------------- giteig
def test_giteig_uniform_closed_form(oracle_mod):
    g = uniform_box((2, 1, 1), (4, 4, 4), length=(2.0, 1.0, 1.0))   # h = 1/4
    o = oracle_mod.Oracle(g)
    c = o.giteig(g.ones())
    inner = g.interior_mask()
    assert np.all(c["ap"][inner] == -6.0 * 16.0)
    for k in ("aw", "ae", "as", "an", "ab", "at"):
        nz = c[k][c[k] != 0]
        assert np.all(nz == 16.0), k


def test_giteig_wall_fold_in_and_blocked_cells(oracle_mod):
    g = uniform_box((1, 1, 1), (4, 4, 4), periodic=(False, True, True), wall=NOS)
    o = oracle_mod.Oracle(g)
    bp = g.ones()
    v = g.block_view(bp, 1)
    v[3, 3, 3] = 0.0                         # one blocked cell (interior i=j=k=2)
    c = o.giteig(bp)
    a = g.block_view(c["ap"], 1)
    h2 = 16.0
    # low/high x walls: coefficient folded into AP and zeroed (noib_mod :178-198)
    assert c["aw"][2] == 0.0 and c["ae"][g.ii[0] - 3] == 0.0
    assert a[2, 4, 4] == -6 * h2 + h2        # wall neighbour folded back
    # blocked cell: all six links cut -> AP = -6/h^2 + 6/h^2 = 0 exactly
    assert a[3, 3, 3] == 0.0
    # fluid neighbour of the blocked cell loses one link
    assert a[4, 3, 3] == -6 * h2 + h2


# ---------------------------------------------------------------- tables
def test_hp_tables_brute_force(oracle_mod):
    g = uniform_box((1, 1, 1), (5, 3, 4))
    mip, idx = oracle_mod.Oracle(g).hp_tables()
    ii, jj, kk = g.ii[0], g.jj[0], g.kk[0]
    pos = 0
    for m in range(9, (ii - 2) + (jj - 2) + (kk - 2) + 1):
        cells = sorted(k + (j - 1) * kk + (i - 1) * kk * jj
                       for i in range(3, ii - 1) for j in range(3, jj - 1)
                       for k in range(3, kk - 1) if i + j + k == m)
        assert mip[m - 1] == pos                       # MIP(m), 1-based m
        assert list(idx[pos:pos + len(cells)]) == cells
        pos += len(cells)
    assert pos == 5 * 3 * 4


# ---------------------------------------------------------------- LU
def _assemble(g, o, c, lo, f):
    """A, and the unit triangular factors of the sweeps, for a 1-block grid
    whose faces are all walls (no ghost coupling)."""
    ids = _ids(g)
    n = len(ids)
    ii, jj, kk = g.ii[0], g.jj[0], g.kk[0]
    ni, nj, nk = ii - 4, jj - 4, kk - 4
    num = lambda i, j, k: k + nk * (j + nj * i)
    A = sp.lil_matrix((n, n))
    Lu = sp.lil_matrix((n, n))
    Uu = sp.lil_matrix((n, n))
    lpr = np.zeros(n)
    mip, idx = o.hp_tables()
    acc = {int(idx[s]) - 1: s for s in range(n)}        # linear -> access index
    ap = g.block_view(c["ap"], 1)
    for i in range(ni):
        for j in range(nj):
            for k in range(nk):
                r = num(i, j, k)
                lin = (k + 2) + (j + 2) * kk + (i + 2) * kk * jj
                s = acc[lin]
                A[r, r] = ap[i + 2, j + 2, k + 2]
                lpr[r] = f["lpr"][s]
                nb = ((i - 1, j, k, c["aw"][i + 2], "lw", Lu), (i + 1, j, k, c["ae"][i + 2], "ue", Uu),
                      (i, j - 1, k, c["as"][j + 2], "ls", Lu), (i, j + 1, k, c["an"][j + 2], "un", Uu),
                      (i, j, k - 1, c["ab"][k + 2], "lb", Lu), (i, j, k + 1, c["at"][k + 2], "ut", Uu))
                for a, b, d, coef, fk, M in nb:
                    if 0 <= a < ni and 0 <= b < nj and 0 <= d < nk:
                        A[r, num(a, b, d)] = coef
                        M[r, num(a, b, d)] = f[fk][s]
    I = sp.identity(n, format="csr")
    return A.tocsr(), (I + Lu).tocsr(), (I + Uu).tocsr(), lpr, ids


@pytest.fixture
def closed_box():
    return uniform_box((1, 1, 1), (4, 3, 5), length=(1.0, 0.75, 1.25),
                       periodic=(False, False, False))


def test_lu_alfa0_is_ilu0(oracle_mod, closed_box):
    """alfa = 0: L U equals A on the 7-point pattern (textbook ILU(0))."""
    g = closed_box
    o = oracle_mod.Oracle(g)
    lib = oracle_mod.lib()
    lib.og_set_alfa(C.c_double(0.0))
    try:
        c = o.giteig(g.ones())
        mip, idx = o.hp_tables()
        f = o.lu_hp(c, g.ones(), mip, idx)
    finally:
        lib.og_set_alfa(C.c_double(float(np.float32(0.92))))
    A, L1, U1, lpr, _ = _assemble(g, o, c, None, f)
    M = (sp.diags(1.0 / lpr) @ L1 @ U1).toarray()
    Ad = A.toarray()
    pat = Ad != 0
    assert np.allclose(M[pat], Ad[pat], rtol=1e-12, atol=0)
    assert np.abs(M[~pat]).max() > 0            # fill-in exists off the pattern


def test_lu_alfa1_row_sums(oracle_mod, closed_box):
    """alfa = 1: Stone's compensation makes M and A agree on constants."""
    g = closed_box
    o = oracle_mod.Oracle(g)
    lib = oracle_mod.lib()
    lib.og_set_alfa(C.c_double(1.0))
    try:
        c = o.giteig(g.ones())
        mip, idx = o.hp_tables()
        f = o.lu_hp(c, g.ones(), mip, idx)
    finally:
        lib.og_set_alfa(C.c_double(float(np.float32(0.92))))
    A, L1, U1, lpr, _ = _assemble(g, o, c, None, f)
    M = sp.diags(1.0 / lpr) @ L1 @ U1
    one = np.ones(A.shape[0])
    assert np.allclose(M @ one, A @ one, rtol=0, atol=1e-9 * np.abs(A).max())


def test_lu_classic_equals_hyperplane(oracle_mod):
    """Both restatements produce identical factors; only storage differs."""
    g = topology("aniso")
    o = oracle_mod.Oracle(g)
    c = o.giteig(g.ones())
    mip, idx = o.hp_tables()
    fh = o.lu_hp(c, g.ones(), mip, idx)
    fc = o.lu_classic(c, g.ones())
    for b in range(1, g.ngrid + 1):
        base = g.ip3[b - 1]
        n = (g.ii[b - 1] - 4) * (g.jj[b - 1] - 4) * (g.kk[b - 1] - 4)
        lin = idx[base:base + n] - 1 + base
        for k in fh:
            assert np.array_equal(fh[k][base:base + n], fc[k][lin]), k


def test_sweeps_equal_triangular_solves(oracle_mod, closed_box):
    g = closed_box
    o = oracle_mod.Oracle(g)
    st = o.setup()
    A, L1, U1, lpr, ids = _assemble(g, o, st["coef"], None, st["fac"])
    rng = np.random.default_rng(11)
    rhs = g.zeros()
    rhs[ids] = rng.standard_normal(len(ids))
    for ityp in (0, 2):
        f = st["fac"] if ityp == 0 else o.lu_classic(st["coef"], st["bp"])
        hilf, res = g.zeros(), g.zeros()
        o.laplacephi(res, hilf, st["coef"], st["bp"])
        if ityp == 0:
            o.sipiter1_hp(rhs, res, f, st["mip"], st["idx"])
            o.sipiter2_hp(hilf, res, f, st["mip"], st["idx"])
        else:
            o.sipiter1_cl(rhs, res, f)
            o.sipiter2_cl(hilf, res, f)
        y = spla.spsolve_triangular(L1, lpr * rhs[ids], lower=True)
        x = spla.spsolve_triangular(U1, y, lower=False)
        assert np.allclose(hilf[ids], x, rtol=1e-13, atol=1e-13 * np.abs(x).max()), ityp


# ---------------------------------------------------------------- stencils
def test_laplacian_eigenfunction(oracle_mod):
    g = uniform_box((1, 1, 1), (8, 6, 10), length=(2.0, 1.5, 2.5))
    o = oracle_mod.Oracle(g)
    st = o.setup()
    m = (1, 2, 3)
    L = (2.0, 1.5, 2.5)
    h = [L[a] / n for a, n in enumerate((8, 6, 10))]
    phi = g.zeros()
    v = g.block_view(phi, 1)
    X = g.x[:g.ii[0]][:, None, None]
    Y = g.y[:g.jj[0]][None, :, None]
    Z = g.z[:g.kk[0]][None, None, :]
    v[...] = (np.sin(2 * np.pi * m[0] * X / L[0]) * np.cos(2 * np.pi * m[1] * Y / L[1])
              * np.sin(2 * np.pi * m[2] * Z / L[2]))
    lam = sum(-4.0 / h[a] ** 2 * np.sin(np.pi * m[a] * h[a] / L[a]) ** 2 for a in range(3))
    res = g.zeros()
    o.laplacephi(res, phi, st["coef"], st["bp"])
    inner = g.interior_mask()
    assert np.allclose(res[inner], -lam * phi[inner], atol=1e-11 * abs(lam))


def test_divcal_discrete_taylor_green(oracle_mod):
    """The staggered Taylor-Green field is exactly divergence free."""
    g = uniform_box((2, 2, 2), (8, 8, 8), length=(2 * np.pi,) * 3)
    o = oracle_mod.Oracle(g)
    u, v, w = g.zeros(), g.zeros(), g.zeros()
    for b in range(1, g.ngrid + 1):
        sx = slice(g.ip1x[b - 1], g.ip1x[b - 1] + g.ii[b - 1])
        sy = slice(g.ip1y[b - 1], g.ip1y[b - 1] + g.jj[b - 1])
        sz = slice(g.ip1z[b - 1], g.ip1z[b - 1] + g.kk[b - 1])
        xs, ys, zs = g.xstag[sx][:, None, None], g.ystag[sy][None, :, None], g.zstag[sz][None, None, :]
        xc, yc, zc = g.x[sx][:, None, None], g.y[sy][None, :, None], g.z[sz][None, None, :]
        g.block_view(u, b)[...] = np.sin(xs) * np.cos(yc) * np.cos(zc)
        g.block_view(v, b)[...] = -np.cos(xc) * np.sin(ys) * np.cos(zc)
    div = g.zeros()
    o.divcal(div, u, v, w, 1.0)
    assert np.abs(div).max() < 1e-13


def test_divergence_gradient_adjoint(oracle_mod):
    """<D u, p> = -<u, G p> on a periodic uniform grid (volume weights equal),
    which pins the index shifts of divcal against those of mgpcorr."""
    g = topology("box2")
    o = oracle_mod.Oracle(g)
    u, v, w = velocity(g, o, 21)
    rng = np.random.default_rng(22)
    inner = g.interior_mask()
    p = np.where(inner, rng.standard_normal(g.n3d), 0.0)
    o.exchange([p], layers=1)
    div = g.zeros()
    o.divcal(div, u, v, w, 1.0)
    gu, gv, gw, gp = g.zeros(), g.zeros(), g.zeros(), g.zeros()
    o.mgpcorr(gu, gv, gw, gp, p, -1.0)      # gu = (p_{i+1} - p_i) / dx
    lhs = np.dot(div[inner], p[inner])
    rhs = -(np.dot(u[inner], gu[inner]) + np.dot(v[inner], gv[inner]) +
            np.dot(w[inner], gw[inner]))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


# ---------------------------------------------------------------- halo
def _global_ids(g, nb, n):
    a = g.zeros()
    gid = np.zeros(g.n3d, dtype=np.int64)
    for b in range(1, g.ngrid + 1):
        bx, by, bz = (b - 1) % nb[0], ((b - 1) // nb[0]) % nb[1], (b - 1) // (nb[0] * nb[1])
        v = g.block_view(a, b)
        for i in range(n[0]):
            for j in range(n[1]):
                for k in range(n[2]):
                    X, Y, Z = bx * n[0] + i, by * n[1] + j, bz * n[2] + k
                    v[i + 2, j + 2, k + 2] = 1 + X + nb[0] * n[0] * (Y + nb[1] * n[1] * Z)
    return a


@pytest.mark.parametrize("layers,corners", [(1, False), (2, False), (2, True)])
def test_exchange_identity_periodic(oracle_mod, layers, corners):
    nb, n = (3, 2, 2), (4, 3, 5)
    g = uniform_box(nb, n)
    o = oracle_mod.Oracle(g)
    a = _global_ids(g, nb, n)
    o.exchange([a], layers=layers, corners=corners)
    N = [nb[d] * n[d] for d in range(3)]
    for b in range(1, g.ngrid + 1):
        bx, by, bz = (b - 1) % nb[0], ((b - 1) // nb[0]) % nb[1], (b - 1) // (nb[0] * nb[1])
        v = g.block_view(a, b)
        for i in range(-2, n[0] + 2):
            for j in range(-2, n[1] + 2):
                for k in range(-2, n[2] + 2):
                    out = [(c < 0 or c >= n[d]) for d, c in enumerate((i, j, k))]
                    dist = [max(-c, c - n[d] + 1, 0) for d, c in enumerate((i, j, k))]
                    if not any(out):
                        continue
                    got = v[i + 2, j + 2, k + 2]
                    reach = max(dist) <= layers and (corners or sum(out) == 1)
                    if not reach:
                        assert got == 0.0
                        continue
                    X = (bx * n[0] + i) % N[0]
                    Y = (by * n[1] + j) % N[1]
                    Z = (bz * n[2] + k) % N[2]
                    assert got == 1 + X + N[0] * (Y + N[1] * Z), (b, i, j, k)


def test_exchange_forward_flags(oracle_mod):
    """forward=1 fills only the low-side ghosts (faces 1/3/5), forward=-1
    only the high-side ghosts (connect_core_mod.F90:672-681)."""
    nb, n = (2, 2, 2), (3, 3, 3)
    g = uniform_box(nb, n)
    o = oracle_mod.Oracle(g)
    for fwd, lo_filled in ((1, True), (-1, False)):
        a = _global_ids(g, nb, n)
        o.exchange([a], layers=1, forward=fwd)
        v = g.block_view(a, 1)
        assert (v[1, 3, 3] != 0) == lo_filled and (v[-2, 3, 3] != 0) == (not lo_filled)
        assert (v[3, 1, 3] != 0) == lo_filled and (v[3, 3, -2] != 0) == (not lo_filled)


# ---------------------------------------------------------------- projection
def test_projection_annihilates_gradients(oracle_mod):
    """u = -G phi is removed entirely and dt*DP recovers -phi (mod const)."""
    g = topology("box2")
    o = oracle_mod.Oracle(g)
    st = o.setup()
    rng = np.random.default_rng(31)
    inner = g.interior_mask()
    phi = np.where(inner, rng.standard_normal(g.n3d), 0.0)
    o.exchange([phi], layers=1)
    u, v, w, q = g.zeros(), g.zeros(), g.zeros(), g.zeros()
    o.mgpcorr(u, v, w, q, phi, 1.0)          # u = (phi_i - phi_{i+1}) / dx
    o.exchange([u, v, w], layers=2, corners=True)
    p = g.zeros()
    dt = 1e-2
    r = o.mgpoisl(st, u, v, w, p, dt, nouter=400, eps_rel=1e-12)
    assert r["maxrhs"] <= 1e-12 * r["maxrhs0"]
    for a in (u, v, w):
        assert np.abs(a[inner]).max() < 1e-9
    dp = r["dp"][inner] * dt            # u_new = u - dt G dp = -G (phi + dt dp)
    assert np.allclose(dp - dp.mean(), -(phi[inner] - phi[inner].mean()), atol=1e-9)


def test_projection_idempotent(oracle_mod):
    g = topology("aniso")
    o = oracle_mod.Oracle(g)
    st = o.setup()
    u, v, w = velocity(g, o, 41)
    p = g.zeros()
    o.mgpoisl(st, u, v, w, p, 1e-3, nouter=400, eps_rel=1e-12)
    apply_walls(g, u, v, w)                  # the caller's bound_flow
    u1, v1, w1 = u.copy(), v.copy(), w.copy()
    p2 = g.zeros()
    o.mgpoisl(st, u, v, w, p2, 1e-3, nouter=400, eps_rel=1e-12)
    inner = g.interior_mask()
    scale = np.abs(u1[inner]).max()
    for a, b in ((u, u1), (v, v1), (w, w1)):
        assert np.abs(a[inner] - b[inner]).max() < 1e-9 * scale


@pytest.mark.parametrize("periodic", [False, True])
def test_sor_red_black_dense(oracle_mod, periodic):
    """relax against red-black SOR written with the assembled operator A
    (diagonal AP, off-diagonals GSA*): x_c <- (1-w) x_c + w (b_c - sum_nb
    A_c,nb x_nb) / A_cc, visiting cells in the loop order of
    src/flow/sor_mod.F90:122-131 (Fortran indices, rb = 0 then 1).  Ghost
    values are frozen at their values before the call, which matters for a
    self-periodic block; with an odd interior count along k the loop bounds
    never visit k = kk-2, which must keep its value."""
    g = uniform_box((1, 1, 1), (4, 3, 5), length=(1.0, 0.75, 1.25),
                    periodic=(periodic,) * 3)
    o = oracle_mod.Oracle(g)
    st = o.setup()
    c = st["coef"]
    ii, jj, kk = g.ii[0], g.jj[0], g.kk[0]
    ni, nj, nk = ii - 4, jj - 4, kk - 4
    rng = np.random.default_rng(17)
    inner = g.interior_mask()
    dp = np.where(inner, rng.standard_normal(g.n3d), 0.0)
    rhs = np.where(inner, rng.standard_normal(g.n3d), 0.0)
    o.exchange([dp])                       # ghosts of the previous conn
    omega = 1.3
    x = g.block_view(dp.copy(), 1)
    x0 = x.copy()
    b = g.block_view(rhs, 1)
    ap = g.block_view(c["ap"], 1)
    visited = np.zeros((ii, jj, kk), bool)
    for rb in (0, 1):
        for i in range(3, ii - 1):
            for j in range(3, jj - 1):
                ks = (4, kk - 2) if (i + j) % 2 == rb else (3, kk - 3)
                for k in range(ks[0], ks[1] + 1, 2):
                    I, J, K = i - 1, j - 1, k - 1           # 0-based array index
                    s = 0.0
                    for (a, bb, d, coef) in ((I - 1, J, K, c["aw"][i - 1]),
                                             (I + 1, J, K, c["ae"][i - 1]),
                                             (I, J - 1, K, c["as"][j - 1]),
                                             (I, J + 1, K, c["an"][j - 1]),
                                             (I, J, K - 1, c["ab"][k - 1]),
                                             (I, J, K + 1, c["at"][k - 1])):
                        ghost = not (2 <= a < ii - 2 and 2 <= bb < jj - 2 and 2 <= d < kk - 2)
                        s += coef * (x0[a, bb, d] if ghost else x[a, bb, d])
                    x[I, J, K] = (1 - omega) * x[I, J, K] + omega * (b[I, J, K] - s) / ap[I, J, K]
                    visited[I, J, K] = True
    got = dp.copy()
    o.relax(got, rhs, c, st["rap"], omega)
    gv = g.block_view(got, 1)
    sl = (slice(2, ii - 2), slice(2, jj - 2), slice(2, kk - 2))
    assert np.allclose(gv[sl], x[sl], rtol=1e-13, atol=1e-13 * np.abs(x[sl]).max())
    assert nk % 2 == 1
    assert not visited[sl][:, :, nk - 1].any() and visited[sl][:, :, :nk - 1].all()
    assert np.array_equal(gv[sl][:, :, nk - 1], x0[sl][:, :, nk - 1])
    assert np.array_equal(got[~inner], dp[~inner])      # ghosts untouched


def test_sor_fixed_point_and_solution(oracle_mod):
    """SOR leaves the exact solution unchanged, and SOR and SIP converge to
    the same pressure (sor_mod.F90:102-203 against sip)."""
    g = topology("walls")
    o = oracle_mod.Oracle(g)
    st = o.setup()
    u, v, w = velocity(g, o, 51)
    ps, pr = g.zeros(), g.zeros()
    us = [a.copy() for a in (u, v, w)]
    r1 = o.mgpoisl(st, *us, ps, 1e-3, nouter=2000, eps_rel=1e-11)
    ur = [a.copy() for a in (u, v, w)]
    r2 = o.mgpoisl(st, *ur, pr, 1e-3, ityp=1, omega=1.2, sor_all_levels=1,
                   nouter=5000, eps_rel=1e-11)
    assert r1["maxrhs"] <= 1e-11 * r1["maxrhs0"] and r2["maxrhs"] <= 1e-11 * r2["maxrhs0"]
    inner = g.interior_mask()
    a, b = ps[inner] - ps[inner].mean(), pr[inner] - pr[inner].mean()
    assert np.abs(a - b).max() < 1e-8 * np.abs(a).max()
    # fixed point: relax on the converged DP with its own RHS
    rhs = g.zeros()
    o.divcal(rhs, u, v, w, 1e3)
    dp = r1["dp"].copy()
    before = dp.copy()
    o.relax(dp, rhs, st["coef"], st["rap"], 1.3)
    assert np.abs(dp[inner] - before[inner]).max() < 1e-7 * np.abs(before[inner]).max()


# ---------------------------------------------------------------- adjoint
def _adjoint_case(oracle_mod, name, ibtype, seed):
    from cases import adjoint_case
    return adjoint_case(oracle_mod, name, ibtype, seed)[:4]


@pytest.mark.parametrize("name,ibtype", [("box2", 0), ("walls", 0), ("stretched", 0),
                                         ("aniso", 0), ("box2", 1)])
def test_adjoint_operators_equal_oracle_kernels(oracle_mod, name, ibtype):
    """The assembled D, A, G reproduce divcal, laplacephi and mgpcorr."""
    g, o, st, ops = _adjoint_case(oracle_mod, name, ibtype, 31)
    rng = np.random.default_rng(32)
    inner = g.interior_mask()
    cells = ops.cells
    # A phi = -laplacephi(phi)
    phi = np.where(inner, rng.standard_normal(g.n3d), 0.0)
    o.exchange([phi])
    res = g.zeros()
    o.laplacephi(res, phi, st["coef"], st["bp"])
    a = ops.A @ phi[cells]
    assert np.allclose(a, -res[cells], rtol=1e-13, atol=1e-13 * np.abs(res).max())
    # D u = divcal(u) on unknown faces (boundary-condition faces zero)
    x = rng.standard_normal(ops.nface)
    u, v, w = ops.scatter(x)
    o.exchange([u, v, w])
    div = g.zeros()
    o.divcal(div, u, v, w, 1.0, bp=st["bp"] if ibtype else None,
             sdiv=g.zeros() if ibtype else None)
    d = ops.D @ x
    assert np.allclose(d, div[cells], rtol=1e-13, atol=1e-13 * np.abs(div).max())
    # u' - u = -G p (mgpcorr with rfak = 1)
    p = np.where(inner, rng.standard_normal(g.n3d), 0.0)
    o.exchange([p])
    uu, vv, ww, pp = g.zeros(), g.zeros(), g.zeros(), g.zeros()
    o.mgpcorr(uu, vv, ww, pp, p, 1.0, bp=st["bp"] if ibtype else None)
    gp = ops.G @ p[cells]
    assert np.allclose(-gp, ops.gather(uu, vv, ww), rtol=1e-13, atol=1e-13 * np.abs(gp).max())


def test_stretched_operator_is_not_symmetric(oracle_mod):
    g, o, st, ops = _adjoint_case(oracle_mod, "stretched", 0, 33)
    A = ops.A
    assert abs(A - A.T).max() > 1e-3 * abs(A).max()


@pytest.mark.parametrize("name,ibtype", [("walls", 0), ("stretched", 0), ("box2", 1)])
def test_adjoint_dot_product_with_oracle_projection(oracle_mod, name, ibtype):
    """<w, P u> = <P^T w, u> with P the oracle's converged mgpoisl and P^T
    the sparse adjoint."""
    g, o, st, ops = _adjoint_case(oracle_mod, name, ibtype, 34)
    rng = np.random.default_rng(35)
    x = rng.uniform(-1, 1, ops.nface)
    wx = rng.standard_normal(ops.nface)
    u, v, w = ops.scatter(x)
    o.exchange([u, v, w], layers=2, corners=True)
    r = o.mgpoisl(st, u, v, w, g.zeros(), 1e-3, nouter=4000, eps_rel=1e-13, ibtype=ibtype)
    assert r["maxrhs"] <= 1e-13 * r["maxrhs0"]
    lhs = wx @ ops.gather(u, v, w)
    rhs = ops.adjoint(wx) @ x
    assert abs(lhs - rhs) <= 1e-10 * np.abs(wx).sum() * np.abs(x).max()
    # the adjoint is a projection too
    y = ops.adjoint(wx)
    assert np.allclose(ops.adjoint(y), y, rtol=0, atol=1e-10 * np.abs(y).max())


# ---------------------------------------------------------------- levels
def test_ctof_injects_the_parent_cell(oracle_mod):
    from cases import centres, two_level
    g = two_level()
    o = oracle_mod.Oracle(g)
    X, Y, Z = centres(g)
    coarse = 0.3 + 1.7 * X - 0.9 * Y + 2.3 * Z
    coarse[g.ip3[8]:] = 0.0                      # fine grids: unknown
    o.exchange([coarse], layers=2, corners=True)
    o.ctof(2, coarse)
    # every fine cell (ghosts included) holds the value of the coarse cell
    # whose volume contains the fine cell centre
    h1 = g.spacing[0, 0]
    for b in range(9, 17):
        fv = g.block_view(coarse, b)
        xf, yf, zf = (a[g.ip1x[b - 1]:g.ip1x[b - 1] + g.ii[b - 1]] if k == 0 else
                      a[g.ip1y[b - 1]:g.ip1y[b - 1] + g.jj[b - 1]] if k == 1 else
                      a[g.ip1z[b - 1]:g.ip1z[b - 1] + g.kk[b - 1]]
                      for k, a in enumerate((g.x, g.y, g.z)))
        cx = (np.floor(xf / h1) + 0.5) * h1
        cy = (np.floor(yf / h1) + 0.5) * h1
        cz = (np.floor(zf / h1) + 0.5) * h1
        want = 0.3 + 1.7 * cx[:, None, None] - 0.9 * cy[None, :, None] + 2.3 * cz[None, None, :]
        assert np.allclose(fv, want, rtol=0, atol=1e-12), b


def test_ftoc_restriction_exact_for_linear_fields(oracle_mod):
    """Volume-weighted means of linear cell and face fields over 2x2x2 /
    2x2 fine groups equal the field at the coarse cell / face centre."""
    from cases import centres, two_level
    g = two_level()
    o = oracle_mod.Oracle(g)
    X, Y, Z = centres(g)
    lin = lambda x, y, z: 0.3 + 1.7 * x - 0.9 * y + 2.3 * z
    f = lin(X, Y, Z)
    o.ftoc(2, f)
    ref = lin(X, Y, Z)
    assert np.allclose(f, ref, rtol=0, atol=1e-12)
    # staggered: u at x-faces (X + dx/2), v at y-faces, w at z-faces
    hx = np.repeat([g.spacing[b, 0] for b in range(g.ngrid)],
                   [int(g.ii[b] * g.jj[b] * g.kk[b]) for b in range(g.ngrid)])
    u = lin(X + 0.5 * hx, Y, Z)
    v = lin(X, Y + 0.5 * hx, Z)
    w = lin(X, Y, Z + 0.5 * hx)
    u0, v0, w0 = u.copy(), v.copy(), w.copy()
    p = f.copy()
    o.ftoc_uvwp(2, u, v, w, p)
    for a, b in ((u, u0), (v, v0), (w, w0)):
        assert np.allclose(a, b, rtol=0, atol=1e-12)
    # the parent cells under the children are overwritten: a field that is
    # 1 on the fine level and 0 elsewhere fills the refined parent with 1
    z = g.zeros()
    z[g.ip3[8]:] = 1.0
    o.ftoc(2, z)
    assert np.all(g.block_view(z, 1)[2:-2, 2:-2, 2:-2] == 1.0)
    assert not g.block_view(z, 2).any()


def test_parent_and_bound_pressure_exact_for_linear_pressure(oracle_mod):
    """With a linear pressure on both levels, parent's face buffers and
    bound_pressure reproduce the linear field at the PAR-face ghost centres
    (A form, and the B form with BP = 1)."""
    from cases import centres, two_level
    g = two_level()
    o = oracle_mod.Oracle(g)
    X, Y, Z = centres(g)
    lin = 0.3 + 1.7 * X - 0.9 * Y + 2.3 * Z
    for bp in (None, g.ones()):
        p = np.where(g.interior_mask(), lin, 0.0)
        o.exchange([p], layers=2, corners=True)
        pb = o.par_buffer()
        o.parent(2, p, pb)
        o.bound_pressure(2, p, pb, bp=bp)
        nchecked = 0
        for b in range(9, 17):
            for f in range(6):
                if g.bc[b - 1, f] != 8:
                    continue
                pv, lv = g.block_view(p, b), g.block_view(lin, b)
                sl = [slice(2, -2)] * 3
                sl[f // 2] = 1 if f % 2 == 0 else -2
                assert np.allclose(pv[tuple(sl)], lv[tuple(sl)], rtol=0, atol=1e-12), (b, f)
                nchecked += 1
        assert nchecked == 12               # 4 children per refined-region face


def test_bound_pressure_blocked_group_keeps_interior(oracle_mod):
    """B form: a 2x2x2 group with BP = 0 takes pc from the buffer, so the
    ghost equals the interior value (bound_pressure_mod.F90:320-330)."""
    from cases import two_level
    g = two_level()
    o = oracle_mod.Oracle(g)
    rng = np.random.default_rng(61)
    p = np.where(g.interior_mask(), rng.standard_normal(g.n3d), 0.0)
    o.exchange([p], layers=2, corners=True)
    bp = g.ones()
    b = 9 + 1                                 # child at (1, 0, 0): PAR on its +x face
    assert g.bc[b - 1, 1] == 8
    v = g.block_view(bp, b)
    v[-4:-2, 2:4, 2:4] = 0.0                  # group at the +x face, j, k = 3..4
    pb = o.par_buffer()
    o.parent(2, p, pb)
    o.bound_pressure(2, p, pb, bp=bp)
    pv = g.block_view(p, b)
    assert np.array_equal(pv[-2, 2:4, 2:4], pv[-3, 2:4, 2:4])


def test_multilevel_full_refinement_matches_single_level(oracle_mod):
    """A fine level that covers the whole periodic domain has no PAR faces;
    converged, its pressure and velocity equal those of a single-level
    solve on the fine blocks alone (the coarse level only accelerates)."""
    from mgp.grids import refined_box, uniform_box
    from cases import velocity
    gm = refined_box((2, 2, 2), 8, [(8, lambda B: True)])
    gs = uniform_box((4, 4, 4), (8, 8, 8))
    om, os_ = oracle_mod.Oracle(gm), oracle_mod.Oracle(gs)
    u, v, w = velocity(gs, os_, 71)
    # copy the single-level fields into the fine level of the hierarchy
    off = gm.ip3[8]
    fine = [gm.zeros() for _ in range(3)]
    order = []
    for b in range(9, gm.ngrid + 1):
        o_ = gm.blocks[b - 1]["origin"] // 8
        order.append(1 + o_[0] + 4 * (o_[1] + 4 * o_[2]))
    for arr_m, arr_s in zip(fine, (u, v, w)):
        for bm, bs in zip(range(9, gm.ngrid + 1), order):
            gm.block_view(arr_m, bm)[...] = gs.block_view(arr_s, bs)
    stm, sts = om.setup(), os_.setup()
    pm, ps = gm.zeros(), gs.zeros()
    rm = om.mgpoisl(stm, *fine, pm, 1e-3, nouter=3000, eps_rel=1e-12)
    rs = os_.mgpoisl(sts, u, v, w, ps, 1e-3, nouter=3000, eps_rel=1e-12)
    assert rm["maxrhs"] <= 1e-12 * rm["maxrhs0"] and rs["maxrhs"] <= 1e-12 * rs["maxrhs0"]
    assert rm["nouter_done"] < rs["nouter_done"]          # the coarse level helps
    inner = gs.interior_mask()
    for arr_m, arr_s in zip(fine + [pm], (u, v, w, ps)):
        got = np.concatenate([gm.block_view(arr_m, bm)[2:-2, 2:-2, 2:-2].ravel()
                              for bm in range(9, gm.ngrid + 1)])
        ref = np.concatenate([gs.block_view(arr_s, bs)[2:-2, 2:-2, 2:-2].ravel()
                              for bs in order])
        if arr_s is ps:
            got, ref = got - got.mean(), ref - ref.mean()
        assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_ghostcell_restriction_forms(oracle_mod):
    """gc_restrict restrict_p / restrict_r (src/ib/gc_restrict_mod.F90):
    with BP = 1 both equal the plain volume-weighted mean (to rounding); a
    fully blocked group restricts to 0 under 'P'; 'R' sums only fluid
    cells over the full group volume."""
    from cases import two_level
    g = two_level()
    o = oracle_mod.Oracle(g)
    rng = np.random.default_rng(91)
    f = rng.standard_normal(g.n3d)
    ref = f.copy()
    o.ftoc(2, ref)
    for flag in ("P", "R"):
        a = f.copy()
        o.ftoc(2, a, bp=g.ones(), flag=flag)
        assert np.allclose(a, ref, rtol=1e-14, atol=1e-14)
    bp = g.ones()
    child = g.block_view(bp, 9)
    child[2:4, 2:4, 2:4] = 0.0                 # first 2x2x2 group blocked
    child[4:6, 2:4, 2:4] = 0.0
    child[4, 2, 2] = 1.0                       # second group: one fluid cell
    fv = g.block_view(f, 9)
    a = f.copy()
    o.ftoc(2, a, bp=bp, flag="P")
    pv = g.block_view(a, 1)
    assert pv[2, 2, 2] == 0.0
    assert np.isclose(pv[3, 2, 2], fv[4, 2, 2], rtol=1e-14)
    a = f.copy()
    o.ftoc(2, a, bp=bp, flag="R")
    pv = g.block_view(a, 1)
    assert pv[2, 2, 2] == 0.0
    assert np.isclose(pv[3, 2, 2], fv[4, 2, 2] / 8.0, rtol=1e-14)


# --- momentum right-hand side (tstle4) and the RK step -----------------
def _self_block(n=12, length=2 * np.pi):
    return uniform_box((1, 1, 1), (n, n, n), length=(length,) * 3)


def _fill(g, o, fields_interior):
    """Interior arrays (n, n, n) indexed [i, j, k] into block 1, ghosts by
    the oracle's exchange."""
    out = []
    for a in fields_interior:
        f = g.zeros()
        g.block_view(f, 1)[2:-2, 2:-2, 2:-2] = a
        out.append(f)
    o.exchange(out, layers=2, corners=True)
    return out


def _curl_field(n, h, seed):
    """Discrete curl of a random vector potential on a periodic n^3 grid:
    exactly divergence free up to rounding."""
    rng = np.random.default_rng(seed)
    ax_, ay_, az_ = (rng.uniform(-1, 1, (n, n, n)) for _ in range(3))
    d = lambda a, axis: (a - np.roll(a, 1, axis=axis)) / h    # backward difference
    u = d(az_, 1) - d(ay_, 2)
    v = d(ax_, 2) - d(az_, 0)
    w = d(ay_, 0) - d(ax_, 1)
    return u, v, w


def test_tstle4_uniform_flow_is_steady(oracle_mod):
    g = _self_block(8)
    o = oracle_mod.Oracle(g)
    n = 8
    u, v, w = _fill(g, o, [np.full((n, n, n), 0.7), np.full((n, n, n), -0.3),
                           np.full((n, n, n), 1.1)])
    p = g.zeros()
    visc = np.full(g.n3d, 1e-2)
    for iles in (0, 1):
        uo, vo, wo = o.tstle4(u, v, w, p, visc, 1.0, 1e-2, ilesmodel=iles)
        assert np.abs(uo).max() == 0 and np.abs(vo).max() == 0 and np.abs(wo).max() == 0


def test_tstle4_convection_conserves_energy(oracle_mod):
    """The skew-symmetric convective form (tstle4_kon) does no work on a
    divergence-free periodic field: sum of u . conv(u) over the momentum
    cells vanishes (viscosity 0, pressure 0)."""
    n = 12
    g = _self_block(n)
    o = oracle_mod.Oracle(g)
    h = 2 * np.pi / n
    ui, vi, wi = _curl_field(n, h, 11)
    u, v, w = _fill(g, o, [ui, vi, wi])
    div = g.zeros()
    o.divcal(div, u, v, w, 1.0)
    assert np.abs(div).max() < 1e-12 * np.abs(ui).max() / h
    uo, vo, wo = o.tstle4(u, v, w, g.zeros(), g.zeros(), 1.0, 1e-30)
    inner = g.interior_mask()
    terms = np.concatenate([(u * uo)[inner], (v * vo)[inner], (w * wo)[inner]])
    assert abs(terms.sum()) < 1e-13 * np.abs(terms).sum()
    assert np.abs(uo[inner]).max() > 0.1      # the term itself is not trivial


def test_tstle4_diffusion_eigenmode(oracle_mod):
    """u = sin(m x) along x: the viscous part of the right-hand side (the
    result minus the inviscid one) is the 3-point second difference
    (ilesmodel 0) and twice that when the transpose part of the stress is on
    (ilesmodel 1, u varies along x only)."""
    n, m = 16, 3
    g = _self_block(n)
    o = oracle_mod.Oracle(g)
    h = 2 * np.pi / n
    xs = (np.arange(n) + 1.0) * h                # u faces of the interior cells
    ui = np.broadcast_to(np.sin(m * xs)[:, None, None], (n, n, n)).copy()
    zero = np.zeros((n, n, n))
    u, v, w = _fill(g, o, [ui, zero, zero])
    nu, rho = 2e-2, 1.3
    visc = np.full(g.n3d, nu)
    lam = -(nu / rho) * 4.0 / h ** 2 * np.sin(m * h / 2) ** 2
    inner = g.interior_mask()
    conv = o.tstle4(u, v, w, g.zeros(), g.zeros(), rho, 1e-30)   # viscosity 0
    for iles, f in ((0, 1.0), (1, 2.0)):
        uo, vo, wo = o.tstle4(u, v, w, g.zeros(), visc, rho, nu, ilesmodel=iles)
        d = uo - conv[0]
        assert np.allclose(d[inner], f * lam * u[inner], rtol=0, atol=1e-11 * abs(lam))
        assert np.abs(vo - conv[1]).max() < 1e-12 and np.abs(wo - conv[2]).max() < 1e-12


def test_tstle4_pressure_gradient(oracle_mod):
    """A linear pressure p = a x on interior cells: -1/rho dp/dx on the u
    faces between them, plus the body-force gradient gradp."""
    n = 8
    g = uniform_box((1, 1, 1), (n, n, n), length=(1.0, 1.0, 1.0), periodic=(False, True, True),
                    wall=FIX)
    o = oracle_mod.Oracle(g)
    a, rho = 0.75, 2.0
    xc = g.x[2:2 + n]
    p = g.zeros()
    g.block_view(p, 1)[2:-2, 2:-2, 2:-2] = (a * xc)[:, None, None]
    o.exchange([p], layers=2, corners=True)     # periodic in y and z
    z = g.zeros()
    uo, vo, wo = o.tstle4(z, z, z, p, z, rho, 1e-30, gradp=[0.0, 0.25, 0.0])
    # faces between two interior cells: i = 3 .. ii-3 (walls at x)
    du = g.block_view(uo, 1)[2:n + 1, 2:-2, 2:-2]
    assert np.allclose(du, -a / rho, rtol=1e-12)
    dv = g.block_view(vo, 1)[2:-2, 2:-2, 2:-2]
    assert np.allclose(dv, -0.25 / rho, rtol=1e-12)


def test_rkstep_first_stage(oracle_mod):
    g = _self_block(4)
    o = oracle_mod.Oracle(g)
    rng = np.random.default_rng(3)
    p0, rhs = rng.standard_normal(g.n3d), rng.standard_normal(g.n3d)
    p, dp = p0.copy(), np.full(g.n3d, 5.0)
    o.rkstep(p, dp, rhs, 0.0, 0.25)           # Williamson stage 1: a = 0
    assert np.array_equal(dp, rhs) and np.array_equal(p, p0 + 0.25 * rhs)
    p2, dp2 = p.copy(), dp.copy()
    o.rkstep(p2, dp2, rhs, -5.0 / 9.0, 0.5)
    assert np.array_equal(dp2, -5.0 / 9.0 * dp + rhs)


def test_sdiv_source_cancelling_the_divergence_is_a_no_op(oracle_mod):
    """Ghostcell mgpoisl with the geometry source SDIV = -div(u) on every
    interior cell: the right-hand side fak*(div u + SDIV)*BP is zero, so the
    projection leaves U, V, W and P as they were; with SDIV = 0 it does not."""
    from cases import blocked_mask, topology, velocity
    g = topology("box2")
    o = oracle_mod.Oracle(g)
    bp = 1.0 - blocked_mask(g, 7, 0.1)
    o.exchange([bp], layers=2, corners=True)
    st = o.setup(bp=bp)
    u, v, w = velocity(g, o, 8)
    div = g.zeros()
    o.divcal(div, u, v, w, 1.0)
    sdiv = -div
    p = np.full(g.n3d, 0.25)
    ins = [a.copy() for a in (u, v, w, p)]
    so = o.mgpoisl(st, u, v, w, p, 1e-3, ninner=5, nouter=2, ibtype=1, sdiv=sdiv)
    assert so["maxrhs0"] == 0.0
    inner = g.interior_mask()
    for a, b in zip((u, v, w, p), ins):
        assert np.array_equal(a[inner], b[inner])
    ref = [a.copy() for a in ins]
    o.mgpoisl(st, *ref, 1e-3, ninner=5, nouter=2, ibtype=1)
    assert np.abs(ref[3] - ins[3])[inner].max() > 0


# ---------------------------------------------------------------- wall stress
@pytest.mark.parametrize("rho,gmol", [(1.0, 1e-2), (1.2, 1.8e-5), (1000.0, 1e-3)])
@pytest.mark.parametrize("dd", [1e-3, 0.01, 0.37])
def test_wall_law_continuous_at_regime_switch(oracle_mod, rho, gmol, dd):
    """The Werner-Wengle wall stress (tauwin, src/flow/wernerwengle_mod.F90:
    64-87) switches from the laminar to the power-law form at
    |u| = cpo6/dd; the two forms meet there only with the constants of
    init_wernerwengle (:13-29), so a wrong constant shows as a jump."""
    ww = oracle_mod.ww_constants(rho, gmol)
    us = ww[3] / dd
    lo = oracle_mod.wall_term(dd, us * (1 - 1e-12), rho, gmol)
    hi = oracle_mod.wall_term(dd, us * (1 + 1e-12), rho, gmol)
    assert abs(hi - lo) <= 1e-10 * abs(lo)
    # laminar sublayer: 2 nu u / dd^2, odd in u
    u = 0.25 * us
    assert np.isclose(oracle_mod.wall_term(dd, u, rho, gmol), 2 * gmol / rho * u / dd**2,
                      rtol=1e-14, atol=0)
    assert oracle_mod.wall_term(dd, -u, rho, gmol) == -oracle_mod.wall_term(dd, u, rho, gmol)
    # far in the turbulent regime tau grows as |u|^(2/(1+1/7)) = |u|^1.75
    big = oracle_mod.wall_term(dd, 1e8 * us, rho, gmol)
    assert abs(oracle_mod.wall_term(dd, 2e8 * us, rho, gmol) / big - 2**1.75) < 1e-7


def test_tstle4_nos_wall_terms(oracle_mod):
    """NOS walls (swcle3d, src/flow/tstle4_mod.F90:1756-1845) change tstle4
    only by the wall stress of the tangential components in the first cell
    next to each wall (including its ghost rows 2 and n-1), each wall
    subtracting wall_term(wall distance, velocity) once: NOS minus FIX walls
    (same loops, no wall stress)."""
    from mgp.grids import FIX, NOS, uniform_box
    gn = uniform_box((2, 2, 2), (6, 8, 6), periodic=(False, False, False), wall=NOS)
    gf = uniform_box((2, 2, 2), (6, 8, 6), periodic=(False, False, False), wall=FIX)
    on, of = oracle_mod.Oracle(gn), oracle_mod.Oracle(gf)
    rng = np.random.default_rng(41)
    u, v, w = (rng.uniform(-3, 3, gn.n3d) for _ in range(3))
    p = rng.standard_normal(gn.n3d)
    visc = rng.uniform(1e-3, 2e-2, gn.n3d)
    rho, gmol = 1.2, 1e-3
    rn = on.tstle4(u, v, w, p, visc, rho, gmol)
    rf = of.tstle4(u, v, w, p, visc, rho, gmol)
    hits = 0
    for comp, (a, b, vel) in enumerate(zip(rn, rf, (u, v, w))):
        for blk in range(1, gn.ngrid + 1):
            an, af, vv = gn.block_view(a, blk), gn.block_view(b, blk), gn.block_view(vel, blk)
            ii, jj, kk = an.shape
            bc = gn.bc[blk - 1]
            expect = np.zeros_like(an)
            dx = gn.ddx[gn.ip1x[blk - 1]:gn.ip1x[blk - 1] + ii]
            dy = gn.ddy[gn.ip1y[blk - 1]:gn.ip1y[blk - 1] + jj]
            dz = gn.ddz[gn.ip1z[blk - 1]:gn.ip1z[blk - 1] + kk]
            for ax, (lo_f, hi_f) in enumerate(((0, 1), (2, 3), (4, 5))):
                if ax == comp:
                    continue        # the wall-normal component has no wall stress
                for f, plane in ((lo_f, 2), (hi_f, an.shape[ax] - 3)):
                    if bc[f] != NOS:
                        continue
                    sl = [slice(1, n - 1) for n in an.shape]
                    sl[ax] = plane
                    dd = (dx, dy, dz)[ax][plane]
                    for idx in zip(*np.nonzero(np.ones_like(an[tuple(sl)], dtype=bool))):
                        full = list(idx)
                        full.insert(ax, None)
                        pos = tuple(plane if q is None else s.start + q
                                    for q, s in zip(full, sl))
                        expect[pos] -= oracle_mod.wall_term(dd, vv[pos], rho, gmol)
                        hits += 1
            diff = an - af
            assert np.allclose(diff, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max()), \
                (comp, blk)
    assert hits > 0


def test_sor_lagged_ghosts_diverge_at_high_omega(oracle_mod):
    """sor relaxes every grid completely (both colours) before conn exchanges
    the ghosts (src/flow/pressuresolver_mod.F90:403-414, sor_mod.F90:52-101),
    so the blocks see their neighbours' values of the previous relaxation.
    On a periodic 2x2x2-block box this block-lagged red-black SOR converges
    at omega = 1.2 and diverges at omega = 1.7."""
    g = topology("box2")
    o = oracle_mod.Oracle(g)
    st = o.setup(ityp=1)
    u, v, w = velocity(g, o, 7)
    hist = {}
    for om in (1.2, 1.7):
        f = [a.copy() for a in (u, v, w)] + [np.zeros(g.n3d)]
        hist[om] = o.mgpoisl(st, *f, 1e-3, ityp=1, ninner=5, nouter=20, omega=om,
                             sor_all_levels=1)["history"]
    assert hist[1.2][-1] < 1e-3 * hist[1.2][0]
    assert hist[1.7][-1] > 1e6 * hist[1.7][0]


# ---------------------------------------------------------------- closing sync
def _linear_uvw(g, cu=(3.0, 5.0), cv=(-2.0, 7.0), cw=(1.0, 3.0)):
    """u, v, w linear along their own staggered axis (u = a + b x at the
    u-face positions, constant in y and z, ...), on every grid and every
    stored position; dyadic grids keep the arithmetic exact."""
    u, v, w = g.zeros(), g.zeros(), g.zeros()
    for b in range(1, g.ngrid + 1):
        sx = slice(g.ip1x[b - 1], g.ip1x[b - 1] + g.ii[b - 1])
        sy = slice(g.ip1y[b - 1], g.ip1y[b - 1] + g.jj[b - 1])
        sz = slice(g.ip1z[b - 1], g.ip1z[b - 1] + g.kk[b - 1])
        g.block_view(u, b)[...] = (cu[0] + cu[1] * g.xstag[sx])[:, None, None]
        g.block_view(v, b)[...] = (cv[0] + cv[1] * g.ystag[sy])[None, :, None]
        g.block_view(w, b)[...] = (cw[0] + cw[1] * g.zstag[sz])[None, None, :]
    return u, v, w


def _face_buf(o, g, buf, b, f):
    """The 2-D view of grid b's buffer of face f (1-based): (k, j), (k, i) or
    (j, i), first index fastest."""
    off = int(o.bbo[6 * (b - 1) + f - 1])
    assert off >= 0
    kk, jj, ii = int(g.kk[b - 1]), int(g.jj[b - 1]), int(g.ii[b - 1])
    n1, n2 = ((kk, jj), (kk, jj), (kk, ii), (kk, ii), (jj, ii), (jj, ii))[f - 1]
    return buf[off:off + n1 * n2].reshape(n2, n1).T


@pytest.mark.parametrize("topo", ["two_level", "three_level"])
def test_parent_uvw_staggered_prolongation_exact(oracle_mod, topo):
    """parent(u, v, w) (src/ib/parent/parent1_mod.F90:296-583): for fields
    linear along their own staggered axis, the prolonged PAR face buffers
    hold the field's value at the fine face positions exactly -- the normal
    component the value on the block's boundary plane (start_and_stop's
    icomp shift at back faces), a tangential component staggered along the
    face the linear interpolation at the fine faces (prolong istag 1)."""
    from cases import three_level, two_level
    g = two_level(periodic=False) if topo == "two_level" else three_level()
    o = oracle_mod.Oracle(g)
    u, v, w = _linear_uvw(g)
    nchecked = 0
    for lvl in range(g.minlevel + 1, g.maxlevel + 1):
        ub, vb, wb = o.bound_buffer(), o.bound_buffer(), o.bound_buffer()
        o.parent_uvw(lvl, u, v, w, ub, vb, wb)
        for b in range(1, g.ngrid + 1):
            if g.level[b - 1] != lvl:
                continue
            xs = g.xstag[g.ip1x[b - 1]:g.ip1x[b - 1] + g.ii[b - 1]]
            ys = g.ystag[g.ip1y[b - 1]:g.ip1y[b - 1] + g.jj[b - 1]]
            zs = g.zstag[g.ip1z[b - 1]:g.ip1z[b - 1] + g.kk[b - 1]]
            ii, jj, kk = len(xs), len(ys), len(zs)
            for f in range(1, 7):
                if g.bc[b - 1, f - 1] != 8:
                    continue
                U, V, W = (_face_buf(o, g, x, b, f) for x in (ub, vb, wb))
                if f <= 2:
                    xb = xs[1] if f == 1 else xs[ii - 3]
                    assert np.all(U == 3.0 + 5.0 * xb)
                    assert np.array_equal(V, np.broadcast_to((-2.0 + 7.0 * ys)[None, :], V.shape))
                    assert np.array_equal(W, np.broadcast_to((1.0 + 3.0 * zs)[:, None], W.shape))
                elif f <= 4:
                    yb = ys[1] if f == 3 else ys[jj - 3]
                    assert np.array_equal(U, np.broadcast_to((3.0 + 5.0 * xs)[None, :], U.shape))
                    assert np.all(V == -2.0 + 7.0 * yb)
                    assert np.array_equal(W, np.broadcast_to((1.0 + 3.0 * zs)[:, None], W.shape))
                else:
                    zb = zs[1] if f == 5 else zs[kk - 3]
                    assert np.array_equal(U, np.broadcast_to((3.0 + 5.0 * xs)[None, :], U.shape))
                    assert np.array_equal(V, np.broadcast_to((-2.0 + 7.0 * ys)[:, None], V.shape))
                    assert np.all(W == 1.0 + 3.0 * zb)
                nchecked += 1
    assert nchecked >= 8


def test_par_tangential_ghosts_linear_exact(oracle_mod):
    """parent then bound_flow at PAR faces (src/flow/bound_flow_mod.F90,
    PAR branches): with fluid everywhere, a tangential component linear
    along the face normal gets its exact value at the ghost cell centre
    (the buffer holds the coarse value one fine cell further out, the
    extrapolation goes half a cell back) -- except the y faces' w, which the
    reference extrapolates from w(j2) instead of w(j3) (bright): that ghost
    gets the value on the face plane instead.  One run per face
    orientation, the two tangential components linear along its normal."""
    from cases import two_level
    g = two_level(periodic=False)
    o = oracle_mod.Oracle(g)
    nchecked = 0
    for ax in range(3):
        u, v, w, p = g.zeros(), g.zeros(), g.zeros(), g.zeros()
        for b in range(1, g.ngrid + 1):
            c = (g.x, g.y, g.z)[ax]
            off = (g.ip1x, g.ip1y, g.ip1z)[ax][b - 1]
            n = (g.ii, g.jj, g.kk)[ax][b - 1]
            shape = [1, 1, 1]
            shape[ax] = n
            t = c[off:off + n].reshape(shape)
            comps = [(u, 1.0, 2.0), (v, -3.0, 6.0), (w, 5.0, -4.0)]
            for q, (arr, a0, a1) in enumerate(comps):
                if q != ax:
                    g.block_view(arr, b)[...] = a0 + a1 * t
        before = [a.copy() for a in (u, v, w)]
        ub, vb, wb = o.bound_buffer(), o.bound_buffer(), o.bound_buffer()
        o.parent_uvw(2, u, v, w, ub, vb, wb)
        o.bound_flow(2, u, v, w, p, ub, vb, wb)
        for b in range(1, g.ngrid + 1):
            if g.level[b - 1] != 2:
                continue
            F = [g.block_view(a, b) for a in (u, v, w)]
            F0 = [g.block_view(a, b) for a in before]
            n = F[0].shape[ax]
            for hi in (0, 1):
                if g.bc[b - 1, 2 * ax + hi] != 8:
                    continue
                gi = n - 2 if hi else 1          # ghost layer i2 (0-based)
                # interior tangential positions, without the normal-velocity
                # planes of the other faces (their PAR correction runs first
                # for faces with a smaller number and moves those values)
                sl = [slice(2, -3)] * 3
                sl[ax] = gi
                sl = tuple(sl)
                for q in range(3):
                    if q == ax:
                        continue
                    if ax == 1 and q == 2:       # bright's w: the face-plane value
                        nb = list(sl)
                        nb[ax] = gi + (-1 if hi else 1)
                        want = 0.5 * (F0[q][sl] + F0[q][tuple(nb)])
                    else:
                        want = F0[q][sl]
                    assert np.array_equal(F[q][sl], want), (ax, hi, q)
                nchecked += 1
    assert nchecked >= 6


def test_par_normal_correction_conserves_flux(oracle_mod):
    """The PAR normal-velocity correction (bound_flow_mod, PAR branches):
    over each 2x2 group of fine boundary faces whose ghost cell is fluid,
    the area-weighted sum of the corrected normal velocities equals the
    parent's (prolonged, piecewise constant) value times the group's area,
    whichever of the four faces are blocked -- the blocked ones keep their
    values, the open ones absorb the difference."""
    from cases import two_level
    g = two_level(periodic=False)
    o = oracle_mod.Oracle(g)
    rng = np.random.default_rng(301)
    u, v, w, p = (rng.uniform(-1, 1, g.n3d) for _ in range(4))
    bp = (rng.random(g.n3d) > 0.3).astype(np.float64)
    ub, vb, wb = o.bound_buffer(), o.bound_buffer(), o.bound_buffer()
    o.parent_uvw(2, u, v, w, ub, vb, wb)
    o.bound_flow(2, u, v, w, p, ub, vb, wb, bp=bp)
    ngroups = 0
    for b in range(1, g.ngrid + 1):
        if g.level[b - 1] != 2:
            continue
        ii, jj, kk = int(g.ii[b - 1]), int(g.jj[b - 1]), int(g.kk[b - 1])
        d = [g.ddx[g.ip1x[b - 1]:][:ii], g.ddy[g.ip1y[b - 1]:][:jj], g.ddz[g.ip1z[b - 1]:][:kk]]
        B = g.block_view(bp, b)
        for f in range(1, 7):
            if g.bc[b - 1, f - 1] != 8:
                continue
            ax, hi = (f - 1) // 2, (f - 1) % 2
            F = g.block_view((u, v, w)[ax], b)
            buf = _face_buf(o, g, (ub, vb, wb)[ax], b, f)
            n = (ii, jj, kk)[ax]
            c2, c3 = (n - 2, n - 3) if hi else (1, 2)     # ghost i2, interior i3
            cs = n - 3 if hi else 1                       # normal-velocity plane
            t1, t2 = [a for a in range(3) if a != ax]     # face axes, buffer (t2-fast? no)
            # buffer index order: x faces (k, j), y (k, i), z (j, i)
            first, second = (2, 1) if ax == 0 else ((2, 0) if ax == 1 else (1, 0))
            for s in range(2, (ii, jj, kk)[second] - 3, 2):
                for r in range(2, (ii, jj, kk)[first] - 3, 2):
                    def at(arr, nrm, rr, ss):
                        idx = [0, 0, 0]
                        idx[ax], idx[first], idx[second] = nrm, rr, ss
                        return arr[tuple(idx)]
                    if at(B, c2, r, s) <= 0.5:
                        continue
                    cells = [(r, s), (r, s + 1), (r + 1, s), (r + 1, s + 1)]
                    if not any(at(B, c2, rr, ss) * at(B, c3, rr, ss) for rr, ss in cells):
                        continue
                    area = [d[first][rr] * d[second][ss] for rr, ss in cells]
                    flux = sum(a * at(F, cs, rr, ss) for a, (rr, ss) in zip(area, cells))
                    assert buf[r, s] == buf[r + 1, s + 1]
                    want = buf[r, s] * sum(area)
                    assert abs(flux - want) <= 1e-13 * max(1.0, abs(want) / sum(area)) * sum(area)
                    ngroups += 1
    assert ngroups > 50


def test_bound_flow_wall_types(oracle_mod):
    """bound_flow at the wall faces of one closed block, faces FIX, OP1,
    NOS, SLI, FIX, OP1 (src/flow/bound_flow_mod.F90): FIX puts the
    prescribed normal velocity on the wall and makes the tangential ghost
    and interior average to the prescribed value; OP1 copies the normal
    velocity outward, keeps tangential ghosts zero-gradient at outflow and
    averaging to the prescribed value at inflow, and sets the ghost
    pressure to pinf at outflow and below it at inflow; NOS zeroes the
    wall velocities, SLI the normal one and copies the tangential ones; the
    ghost pressure of FIX, NOS and SLI copies the interior.  Dyadic values
    keep every relation exact; positions next to edges are left out
    (faces later in the order 1..6 overwrite them)."""
    from mgp.grids import uniform_box
    g = uniform_box((1, 1, 1), (8, 8, 8), periodic=(False, False, False))
    g.bc[0, :] = [2, 3, 5, 6, 2, 3]
    o = oracle_mod.Oracle(g)
    rng = np.random.default_rng(302)
    dy = lambda n: rng.integers(-32, 33, n) / 8.0
    u, v, w, p = (dy(g.n3d) for _ in range(4))
    ub, vb, wb = (dy(o.nbb) for _ in range(3))
    pinf = dy(6)
    rho = 1.0
    o.bound_flow(-1, u, v, w, p, ub, vb, wb, pinf=pinf, rho=rho)
    U, V, W, P = (g.block_view(a, 1) for a in (u, v, w, p))
    n = U.shape[0]
    s = slice(2, n - 3)     # tangential range without the faces' normal planes
    # face 1 (FIX, -x): i2 = 1, i3 = 2, istag2 = 1 (0-based)
    B = [_face_buf(o, g, x, 1, 1) for x in (ub, vb, wb)]   # (k, j)
    assert np.array_equal(U[1, s, s], B[0][s, s].T)
    assert np.array_equal(0.5 * (V[1, s, s] + V[2, s, s]), B[1][s, s].T)
    assert np.array_equal(0.5 * (W[1, s, s] + W[2, s, s]), B[2][s, s].T)
    assert np.array_equal(P[1, s, s], P[2, s, s])
    # face 2 (OP1, +x): istag1 = n-2, istag2 = n-3, i2 = n-2, i3 = n-3
    assert np.array_equal(U[n - 2, s, s], U[n - 3, s, s])
    B = [_face_buf(o, g, x, 1, 2) for x in (ub, vb, wb)]
    for j in range(2, n - 3):
        for k in range(2, n - 3):
            # SIGN(1, +0.0) is +1: a zero flux counts as outflow here
            out_v = U[n - 3, j, k] + U[n - 3, j + 1, k] >= 0
            out_w = U[n - 3, j, k] + U[n - 3, j, k + 1] >= 0
            if out_v:
                assert V[n - 2, j, k] == V[n - 3, j, k]
            else:
                assert 0.5 * (V[n - 2, j, k] + V[n - 3, j, k]) == B[1][k, j]
            if out_w:
                assert W[n - 2, j, k] == W[n - 3, j, k]
            else:
                assert 0.5 * (W[n - 2, j, k] + W[n - 3, j, k]) == B[2][k, j]
            if U[n - 3, j, k] >= 0:
                assert P[n - 2, j, k] == pinf[1]
            elif U[n - 3, j, k] < 0:
                assert P[n - 2, j, k] < pinf[1]
    # face 3 (NOS, -y): jstag2 = j2 = 1
    assert not V[s, 1, s].any() and not U[s, 1, s].any() and not W[s, 1, s].any()
    assert np.array_equal(P[s, 1, s], P[s, 2, s])
    # face 4 (SLI, +y): jstag2 = n-3, j2 = n-2
    assert not V[s, n - 3, s].any()
    assert np.array_equal(U[s, n - 2, s], U[s, n - 3, s])
    assert np.array_equal(W[s, n - 2, s], W[s, n - 3, s])
    assert np.array_equal(P[s, n - 2, s], P[s, n - 3, s])
    # face 5 (FIX, -z): kstag2 = k2 = 1; buffer (j, i)
    B = [_face_buf(o, g, x, 1, 5) for x in (ub, vb, wb)]
    assert np.array_equal(W[s, s, 1], B[2][s, s].T)
    assert np.array_equal(0.5 * (U[s, s, 1] + U[s, s, 2]), B[0][s, s].T)
    assert np.array_equal(0.5 * (V[s, s, 1] + V[s, s, 2]), B[1][s, s].T)
    # face 6 (OP1, +z): normal copied outward
    assert np.array_equal(W[s, s, n - 2], W[s, s, n - 3])
