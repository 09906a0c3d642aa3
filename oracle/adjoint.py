"""Sparse restatement of the projection's operators, for checking the
adjoint projection.

TEST INFRASTRUCTURE ONLY (like the rest of oracle/): used by tests/ as the
checker, never by the product path.

The projection restricted to the interior face velocities that are unknowns
(not boundary conditions) is u' = u - G A^+ D u with
  D  divcal without rho/dt (src/ib/noib_mod.F90:304-354), times BP for the
     ghostcell operator (src/ib/gc_mod.F90:278-347);
  A  the GSA* operator of laplacephi (src/flow/laplacephi_mod.F90:134-165):
     (A phi)_c = sum_nb a_nb BP_c BP_nb phi_nb + AP_c phi_c;
  G  mgpcorr without dt/rho (src/flow/pressuresolver_mod.F90:787-879):
     (G p) at face u(i) = (p(i+1) - p(i)) rdx(i) [BP(i) BP(i+1)].
Boundary-condition faces are faces on a block boundary without a connected
neighbour and, with the ghostcell operator, faces with BP = 0 on either
side.  Ghost values are resolved to the interior cells they copy by running
the oracle's exchange on a field of cell ids, so every matrix acts on
interior unknowns only.  The adjoint w - D^T A^-T G^T w is then computed
with a sparse least-squares solve (scipy lsqr), which shares nothing with
the SIP iteration it checks.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _positions(g):
    """Fortran indices (i, j, k), block id and strides of every stored
    position of the flat 3-D layout."""
    n3 = g.n3d
    I = np.empty(n3, np.int64)
    J = np.empty(n3, np.int64)
    K = np.empty(n3, np.int64)
    blk = np.empty(n3, np.int64)
    si = np.empty(n3, np.int64)
    sj = np.empty(n3, np.int64)
    for b in range(g.ngrid):
        ii, jj, kk = int(g.ii[b]), int(g.jj[b]), int(g.kk[b])
        e = np.arange(ii * jj * kk)
        sl = slice(g.ip3[b], g.ip3[b] + ii * jj * kk)
        K[sl] = e % kk + 1
        J[sl] = (e // kk) % jj + 1
        I[sl] = e // (kk * jj) + 1
        blk[sl] = b
        si[sl] = kk * jj
        sj[sl] = kk
    return I, J, K, blk, si, sj


class Operators:
    """D, A, G of one topology as scipy sparse matrices on interior
    unknowns.  Face unknowns are numbered axis by axis; `faces[a]` holds
    the flat positions of the unknown faces of axis a (the face of a cell is
    its high-side face)."""

    def __init__(self, o, st, ibtype=0):
        g = o.g
        self.g = g
        n3 = g.n3d
        inner = g.interior_mask()
        I, J, K, blk, si, sj = _positions(g)
        ids = np.where(inner, np.arange(n3, dtype=np.float64), -1.0)
        o.exchange([ids], layers=1)
        src = ids.astype(np.int64)                 # interior position or -1
        cells = np.flatnonzero(inner)
        self.cells = cells
        ncell = len(cells)
        cidx = -np.ones(n3, np.int64)
        cidx[cells] = np.arange(ncell)

        def cell_of(pos):
            r = src[pos]
            return np.where(r >= 0, cidx[np.maximum(r, 0)], -1)

        bp = st["bp"]
        use_bp = ibtype == 1
        strides = (si[cells], sj[cells], np.ones(ncell, np.int64))
        oned = [(g.ip1x, I), (g.ip1y, J), (g.ip1z, K)]
        rd = (g.rdx, g.rdy, g.rdz)
        rdd = (g.rddx, g.rddy, g.rddz)

        def metric(arr, a, off):
            ip1, idx = oned[a]
            return arr[ip1[blk[cells]] + idx[cells] - 1 + off]

        # unknown faces
        self.faces = []
        fcol = []
        nf = 0
        for a in range(3):
            hi = cells + strides[a]
            ok = src[hi] >= 0
            if use_bp:
                ok &= bp[cells] * bp[hi] != 0.0
            col = -np.ones(n3, np.int64)
            pos = cells[ok]
            col[pos] = nf + np.arange(len(pos))
            nf += len(pos)
            self.faces.append(pos)
            fcol.append(col)
        self.nface = nf

        # D: div_c = sum_a (u_a(c) - u_a(c - s_a)) rdd_a(c) [x BP_c]
        rows, cols, vals = [], [], []
        w = bp[cells] if use_bp else np.ones(ncell)
        for a in range(3):
            r = metric(rdd[a], a, 0) * w
            c_own = fcol[a][cells]
            m = c_own >= 0
            rows.append(np.arange(ncell)[m]); cols.append(c_own[m]); vals.append(r[m])
            lo = src[cells - strides[a]]             # position whose face is u(c - s_a)
            c_lo = np.where(lo >= 0, fcol[a][np.maximum(lo, 0)], -1)
            m = c_lo >= 0
            rows.append(np.arange(ncell)[m]); cols.append(c_lo[m]); vals.append(-r[m])
        self.D = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                               shape=(ncell, nf))

        # G: (G p)_f = (p(c + s_a) - p(c)) rd_a(c) [x BP_c BP_{c+s_a}]
        rows, cols, vals = [], [], []
        for a in range(3):
            pos = self.faces[a]
            cc = cidx[pos]
            r = rd[a][oned[a][0][blk[pos]] + oned[a][1][pos] - 1]
            if use_bp:
                r = r * bp[pos] * bp[pos + strides[a][cc]]
            hi = cell_of(pos + strides[a][cc])
            f = fcol[a][pos]
            rows += [f, f]; cols += [cc, hi]; vals += [-r, r]
        self.G = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                               shape=(nf, ncell))

        # A: (A phi)_c = sum_nb a_nb BP_c BP_nb phi_nb + AP_c phi_c
        coef = st["coef"]
        rows, cols, vals = [np.arange(ncell)], [np.arange(ncell)], [coef["ap"][cells]]
        names = (("aw", "ae"), ("as", "an"), ("ab", "at"))
        for a in range(3):
            for side, name in ((-1, names[a][0]), (1, names[a][1])):
                nbpos = cells + side * strides[a]
                nb = cell_of(nbpos)
                cf = metric(coef[name], a, 0) * bp[cells] * bp[nbpos]
                m = (nb >= 0) & (cf != 0.0)
                rows.append(np.arange(ncell)[m]); cols.append(nb[m]); vals.append(cf[m])
        self.A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                               shape=(ncell, ncell))

    # ------------------------------------------------------------------
    def gather(self, u, v, w):
        """Unknown face values of the flat fields u, v, w."""
        return np.concatenate([f[p] for f, p in zip((u, v, w), self.faces)])

    def scatter(self, x):
        """Flat fields with x on the unknown faces and zero elsewhere."""
        out = [self.g.zeros() for _ in range(3)]
        o = 0
        for a in range(3):
            n = len(self.faces[a])
            out[a][self.faces[a]] = x[o:o + n]
            o += n
        return out

    def project(self, x, tol=1e-15):
        """u - G A^+ D u on the unknown faces."""
        dp = spla.lsqr(self.A, self.D @ x, atol=tol, btol=tol, iter_lim=200000)[0]
        return x - self.G @ dp

    def adjoint(self, wx, tol=1e-15):
        """w - D^T A^-T G^T w on the unknown faces."""
        y = spla.lsqr(self.A.T.tocsr(), self.G.T @ wx, atol=tol, btol=tol,
                      iter_lim=200000)[0]
        return wx - self.D.T @ y
