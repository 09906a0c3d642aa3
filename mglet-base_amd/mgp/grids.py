"""Block topology and metrics in MGLET-base's own terms.

A `Grids` object holds what the reference's GRIDINFO, boundary-condition
table and pointer module hold for the pressure path:

* per block (grid id 1..ngrid): dims kk, jj, ii including two ghost layers
  on each side, level, the 26 neighbour grid ids (`nbrgrid`,
  src/core/gridio_mod.F90:15-37) and the first boundary-condition code of
  each of the six faces (`itypboconds(1, iface, igrid)`, codes from
  src/core/grids_mod.F90:253-290: FIX=2, NOS=5, SLI=6, CON=7, PAR=8, CO1=19);
* the flat-array offsets ip3 / ip1x / ip1y / ip1z
  (src/core/pointers_mod.F90:23-103), stored 0-based;
* the 1-D metric fields X, DX, DDX (and Y, Z analogues) with the reference's
  meaning: X(i) cell centre, DX(i) = X(i+1) - X(i), DDX(i) cell width,
  XSTAG(i) = X(i) + DX(i)/2 (src/core/corefields_mod.F90:154-191).

Faces 1..6 are -x, +x, -y, +y, -z, +z; directions 7..26 are the edges and
corners in the order of `FACELIST` (src/core/connect/connect_core_mod.F90:
37-63).  Blocks are numbered x fastest, then y, then z.
"""
from __future__ import annotations

import numpy as np

CON, FIX, NOS, SLI, PAR, CO1 = 7, 2, 5, 6, 8, 19

# elementary faces of each of the 26 directions (connect_core_mod.F90:37-63)
FACELIST = [
    (1,), (2,), (3,), (4,), (5,), (6,),
    (1, 3), (1, 4), (1, 5), (1, 6), (2, 3), (2, 4), (2, 5), (2, 6),
    (3, 5), (3, 6), (4, 5), (4, 6),
    (1, 3, 5), (1, 3, 6), (1, 4, 5), (1, 4, 6),
    (2, 3, 5), (2, 3, 6), (2, 4, 5), (2, 4, 6),
]
_FACE_OFFSET = {1: (-1, 0, 0), 2: (1, 0, 0), 3: (0, -1, 0), 4: (0, 1, 0),
                5: (0, 0, -1), 6: (0, 0, 1)}
DIR_OFFSET = [tuple(sum(_FACE_OFFSET[f][a] for f in fl) for a in range(3))
              for fl in FACELIST]


class Grids:
    """Topology + metrics of a set of blocks (one rank owns them all)."""

    def __init__(self, dims, level, nbr, bc, spacing, origin, minlevel=None,
                 maxlevel=None, iparent=None, position=None):
        dims = np.asarray(dims, dtype=np.int32).reshape(-1, 3)  # (nx, ny, nz) interior
        self.ngrid = dims.shape[0]
        self.ii = (dims[:, 0] + 4).astype(np.int32)
        self.jj = (dims[:, 1] + 4).astype(np.int32)
        self.kk = (dims[:, 2] + 4).astype(np.int32)
        self.level = np.asarray(level, dtype=np.int32)
        self.nbr = np.ascontiguousarray(np.asarray(nbr, dtype=np.int32).reshape(self.ngrid, 26))
        self.bc = np.ascontiguousarray(np.asarray(bc, dtype=np.int32).reshape(self.ngrid, 6))
        # parent grid id (0: none) and iposition/jposition/kposition: the
        # parent's 1-based index (ghosts included) of the first parent cell
        # the grid covers (GRIDINFO, src/core/gridio_mod.F90:15-37)
        self.iparent = (np.zeros(self.ngrid, np.int32) if iparent is None
                        else np.asarray(iparent, dtype=np.int32))
        self.position = (np.zeros((self.ngrid, 3), np.int32) if position is None
                         else np.ascontiguousarray(np.asarray(position, dtype=np.int32)
                                                   .reshape(self.ngrid, 3)))
        self.minlevel = int(self.level.min()) if minlevel is None else minlevel
        self.maxlevel = int(self.level.max()) if maxlevel is None else maxlevel
        if np.any(np.diff(self.level) < 0):
            raise ValueError("grids must be numbered level by level")
        n3 = self.ii.astype(np.int64) * self.jj * self.kk
        self.ip3 = np.concatenate([[0], np.cumsum(n3)[:-1]]).astype(np.int64)
        self.ip1x = np.concatenate([[0], np.cumsum(self.ii)[:-1]]).astype(np.int64)
        self.ip1y = np.concatenate([[0], np.cumsum(self.jj)[:-1]]).astype(np.int64)
        self.ip1z = np.concatenate([[0], np.cumsum(self.kk)[:-1]]).astype(np.int64)
        self.n3d = int(n3.sum())
        self.n1x, self.n1y, self.n1z = int(self.ii.sum()), int(self.jj.sum()), int(self.kk.sum())
        self.ncells = int((dims[:, 0].astype(np.int64) * dims[:, 1] * dims[:, 2]).sum())
        # uniform metrics per block: spacing[g] = (hx, hy, hz), origin[g] = low corner
        spacing = np.asarray(spacing, dtype=np.float64).reshape(self.ngrid, 3)
        origin = np.asarray(origin, dtype=np.float64).reshape(self.ngrid, 3)
        self.spacing, self.origin = spacing, origin
        self.x, self.dx, self.ddx = self._metric(0, self.ii, self.ip1x, self.n1x)
        self.y, self.dy, self.ddy = self._metric(1, self.jj, self.ip1y, self.n1y)
        self.z, self.dz, self.ddz = self._metric(2, self.kk, self.ip1z, self.n1z)
        self.rdx, self.rdy, self.rdz = (1.0 / self.dx, 1.0 / self.dy, 1.0 / self.dz)
        self.rddx, self.rddy, self.rddz = (1.0 / self.ddx, 1.0 / self.ddy, 1.0 / self.ddz)
        self.xstag = self.x + 0.5 * self.dx
        self.ystag = self.y + 0.5 * self.dy
        self.zstag = self.z + 0.5 * self.dz

    def _metric(self, axis, n, ip1, ntot):
        x = np.zeros(ntot)
        d = np.zeros(ntot)
        for g in range(self.ngrid):
            h = self.spacing[g, axis]
            o = self.origin[g, axis]
            idx = np.arange(1, n[g] + 1)
            x[ip1[g]:ip1[g] + n[g]] = o + (idx - 2.5) * h
            d[ip1[g]:ip1[g] + n[g]] = h
        return x, d, d.copy()

    # ------------------------------------------------------------------
    def block_view(self, field, g):
        """(ii, jj, kk)-indexed view of block g (1-based id) of a flat field;
        view[i-1, j-1, k-1] is the reference's field(k, j, i)."""
        g0 = g - 1
        n = int(self.ii[g0]) * int(self.jj[g0]) * int(self.kk[g0])
        blk = field[self.ip3[g0]:self.ip3[g0] + n]
        return blk.reshape(int(self.ii[g0]), int(self.jj[g0]), int(self.kk[g0]))

    def interior_mask(self):
        """bool flat mask of interior cells (k, j, i in [3, n-2])."""
        m = np.zeros(self.n3d, dtype=bool)
        for g in range(1, self.ngrid + 1):
            self.block_view(m, g)[2:-2, 2:-2, 2:-2] = True
        return m

    def zeros(self, dtype=np.float64):
        return np.zeros(self.n3d, dtype=dtype)

    def ones(self, dtype=np.float64):
        return np.ones(self.n3d, dtype=dtype)


def uniform_box(nb, n, length=(1.0, 1.0, 1.0), periodic=(True, True, True),
                wall=NOS, origin=(0.0, 0.0, 0.0), level=0):
    """nb = (nbx, nby, nbz) blocks of n = (nx, ny, nz) interior cells each,
    covering a box of the given length; outer faces are periodic (CON to
    the wrapped neighbour, a self-connection when nb is 1 along that axis)
    or walls with code `wall`."""
    nbx, nby, nbz = nb
    nx, ny, nz = n
    ngrid = nbx * nby * nbz
    h = (length[0] / (nbx * nx), length[1] / (nby * ny), length[2] / (nbz * nz))
    dims = np.tile([nx, ny, nz], (ngrid, 1))
    nbr = np.zeros((ngrid, 26), dtype=np.int32)
    bc = np.zeros((ngrid, 6), dtype=np.int32)
    org = np.zeros((ngrid, 3))

    def gid(bx, by, bz):
        return 1 + bx + nbx * (by + nby * bz)

    for bz in range(nbz):
        for by in range(nby):
            for bx in range(nbx):
                g = gid(bx, by, bz) - 1
                org[g] = (origin[0] + bx * nx * h[0], origin[1] + by * ny * h[1],
                          origin[2] + bz * nz * h[2])
                pos = (bx, by, bz)
                nbs = (nbx, nby, nbz)
                for d, off in enumerate(DIR_OFFSET):
                    q = []
                    ok = True
                    for a in range(3):
                        c = pos[a] + off[a]
                        if c < 0 or c >= nbs[a]:
                            if periodic[a]:
                                c %= nbs[a]
                            else:
                                ok = False
                        q.append(c)
                    nbr[g, d] = gid(*q) if ok else 0
                for f in range(6):
                    a = f // 2
                    hi = f % 2 == 1
                    at_edge = (pos[a] == nbs[a] - 1) if hi else (pos[a] == 0)
                    if at_edge and not periodic[a]:
                        bc[g, f] = wall
                    else:
                        bc[g, f] = CON
    return Grids(dims, np.full(ngrid, level), nbr, bc, np.tile(h, (ngrid, 1)), org)


def stretched_box(nb, n, length=(1.0, 1.0, 1.0), ratio=1.15, axis=1, wall=NOS):
    """uniform_box with the cell-centre spacing along `axis` growing by
    `ratio` from cell to cell (walls on that axis, periodic elsewhere).
    The metrics keep the reference's relations: DX(i) = X(i+1) - X(i),
    faces midway between centres, DDX(i) = (DX(i-1) + DX(i)) / 2; ghost
    metrics continue the global progression, so neighbouring blocks agree
    on their shared faces.  The Poisson operator is then not symmetric."""
    periodic = [True, True, True]
    periodic[axis] = False
    g = uniform_box(nb, n, length, tuple(periodic), wall)
    nn = nb[axis] * n[axis]
    # global centre spacings DX for global cells c = -3 .. nn + 2
    cg = np.arange(-3, nn + 3)
    d = ratio ** cg.astype(np.float64)
    d *= length[axis] / d[3:3 + nn].sum()
    xc = np.concatenate([[0.0], np.cumsum(d)])       # centres of cells -3 .. nn + 3
    ip1 = (g.ip1x, g.ip1y, g.ip1z)[axis]
    nloc = (g.ii, g.jj, g.kk)[axis]
    x = np.zeros_like((g.x, g.y, g.z)[axis])
    dx = np.zeros_like(x)
    ddx = np.zeros_like(x)
    nbx, nby = nb[0], nb[1]
    for b in range(g.ngrid):
        pos = (b % nbx, (b // nbx) % nby, b // (nbx * nby))[axis]
        m = int(nloc[b])
        q = np.arange(1, m + 1)                        # local 1-based index
        c = pos * n[axis] + (q - 3)                    # global cell, interior from 0
        sl = slice(ip1[b], ip1[b] + m)
        x[sl] = xc[c + 3]
        dx[sl] = xc[c + 4] - xc[c + 3]
        ddx[sl] = 0.5 * ((xc[c + 3] - xc[c + 2]) + (xc[c + 4] - xc[c + 3]))
    names = ("x", "y", "z")[axis]
    setattr(g, names, x)
    setattr(g, "d" + names, dx)
    setattr(g, "dd" + names, ddx)
    setattr(g, "rd" + names, 1.0 / dx)
    setattr(g, "rdd" + names, 1.0 / ddx)
    setattr(g, names + "stag", x + 0.5 * dx)
    return g


def refined_box(base_nb, base_n, levels, length=(1.0, 1.0, 1.0),
                periodic=(True, True, True), wall=NOS, base_level=1):
    """Nested grid hierarchy.  The base level is a box of `base_nb` blocks of
    `base_n` cells (an int or a 3-tuple).  Each entry of `levels` is
    (n, refine): every block of the previous level for which
    refine(block) is true is covered by children of n cells per direction
    (cell size halved, so a child covers n/2 parent cells per direction and
    the children tile the parent).  A block is a dict with keys level,
    origin (first interior cell in cells of its own level), n, centre.
    Faces with a same-level neighbour are CON, faces on a non-periodic
    domain boundary are walls, the remaining faces of refined blocks are
    PAR (parent-coupled)."""
    n0 = np.array(base_n if np.ndim(base_n) else (base_n,) * 3, dtype=np.int64)
    nb0 = np.array(base_nb, dtype=np.int64)
    L = np.asarray(length, dtype=np.float64)
    cells0 = nb0 * n0
    h0 = L / cells0
    blocks = []
    for bz in range(nb0[2]):
        for by in range(nb0[1]):
            for bx in range(nb0[0]):
                o = np.array((bx, by, bz)) * n0
                blocks.append(dict(level=base_level, origin=o, n=n0.copy(), parent=-1))
    prev = list(range(len(blocks)))
    for li, (n, refine) in enumerate(levels, start=1):
        nn = np.array(n if np.ndim(n) else (n,) * 3, dtype=np.int64)
        lvl = base_level + li
        h = h0 / 2 ** li
        cur = []
        for pidx in prev:
            P = blocks[pidx]
            hp = h0 / 2 ** (li - 1)
            P["centre"] = (P["origin"] + P["n"] / 2) * hp
            if not refine(P):
                continue
            if np.any((2 * P["n"]) % nn):
                raise ValueError("children must tile their parent")
            m = 2 * P["n"] // nn
            for cz in range(m[2]):
                for cy in range(m[1]):
                    for cx in range(m[0]):
                        o = 2 * P["origin"] + np.array((cx, cy, cz)) * nn
                        blocks.append(dict(level=lvl, origin=o, n=nn.copy(), parent=pidx))
                        cur.append(len(blocks) - 1)
        prev = cur
    for B in blocks:
        hl = h0 / 2 ** (B["level"] - base_level)
        B["centre"] = (B["origin"] + B["n"] / 2) * hl
    ngrid = len(blocks)
    # lattice lookup per level
    index = {}
    for g, B in enumerate(blocks):
        index[(B["level"], tuple(B["origin"] // B["n"]))] = g
    nbr = np.zeros((ngrid, 26), dtype=np.int32)
    bc = np.zeros((ngrid, 6), dtype=np.int32)
    for g, B in enumerate(blocks):
        li = B["level"] - base_level
        span = cells0 * 2 ** li // B["n"]           # lattice size of the level
        pos = B["origin"] // B["n"]
        for d, off in enumerate(DIR_OFFSET):
            q = pos + np.array(off)
            ok = True
            for a in range(3):
                if q[a] < 0 or q[a] >= span[a]:
                    if periodic[a]:
                        q[a] %= span[a]
                    else:
                        ok = False
            j = index.get((B["level"], tuple(q))) if ok else None
            nbr[g, d] = 0 if j is None else j + 1
        for f in range(6):
            a = f // 2
            edge = (pos[a] == span[a] - 1) if f % 2 else (pos[a] == 0)
            if nbr[g, f]:
                bc[g, f] = CON
            elif edge and not periodic[a]:
                bc[g, f] = wall
            else:
                bc[g, f] = PAR
    dims = np.array([B["n"] for B in blocks])
    level = np.array([B["level"] for B in blocks])
    spacing = np.array([h0 / 2 ** (B["level"] - base_level) for B in blocks])
    origin = np.array([B["origin"] * s for B, s in zip(blocks, spacing)])
    iparent = np.array([B["parent"] + 1 for B in blocks])
    position = np.zeros((ngrid, 3), dtype=np.int32)
    for g, B in enumerate(blocks):
        if B["parent"] >= 0:
            P = blocks[B["parent"]]
            position[g] = B["origin"] // 2 - P["origin"] + 3
    g = Grids(dims, level, nbr, bc, spacing, origin, iparent=iparent, position=position)
    g.blocks = blocks
    return g
