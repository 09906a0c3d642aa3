"""ctypes front end of the CPU oracle (oracle/mgp_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker or as the timed CPU baseline.
The product path (mglet-base_amd/) never imports this package.

All arrays use the reference's flat per-field layout (see mgp.grids.Grids).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "libmgp_oracle.so"
_lib = None

# The reference's default omega: default_value=1.1 is a real32 literal
# widened to real64 (src/flow/pressuresolver_mod.F90:135).
OMEGA_DEFAULT = float(np.float32(1.1))

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class OgTopo(C.Structure):
    _fields_ = [("ngrid", C.c_int32), ("minlevel", C.c_int32), ("maxlevel", C.c_int32),
                ("kk", _i32p), ("jj", _i32p), ("ii", _i32p), ("level", _i32p),
                ("nbr", _i32p), ("bc", _i32p), ("ip3", _i64p), ("ip1x", _i64p),
                ("ip1y", _i64p), ("ip1z", _i64p), ("iparent", _i32p), ("pos", _i32p)]


class OgConn(C.Structure):
    _fields_ = [("recv_grid", C.c_int32), ("send_grid", C.c_int32),
                ("recv_face", C.c_int32), ("send_face", C.c_int32)]


class OgParams(C.Structure):
    _fields_ = [("ityp", C.c_int32), ("ninner", C.c_int32), ("nouter", C.c_int32),
                ("nouter_min", C.c_int32), ("epcorr", C.c_double), ("eps_rel", C.c_double),
                ("omega", C.c_double), ("rho", C.c_double), ("ibtype", C.c_int32),
                ("sor_all_levels", C.c_int32), ("sdiv", _dp), ("sync_by_caller", C.c_int32),
                ("ubuf", _dp), ("vbuf", _dp), ("wbuf", _dp), ("pinf", _dp)]


class OgStats(C.Structure):
    _fields_ = [("ipc", C.c_int32), ("nouter_done", C.c_int32), ("maxrhs0", C.c_double),
                ("maxrhs", C.c_double), ("history", _dp)]


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(_LIB_PATH))
        _lib.og_maxabscal.restype = C.c_double
        _lib.og_time_sip.restype = C.c_double
        _lib.og_par_buffers.restype = C.c_int64
        _lib.og_bound_buffers.restype = C.c_int64
        _lib.og_swc.restype = C.c_double
        _lib.og_swc.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        _lib.og_ww_constants.argtypes = [C.c_double, C.c_double, _dp]
    return _lib


def ww_constants(rho, gmol):
    """Werner-Wengle constants (cwb, cpo4, cpo5, cpo6, cpo8),
    src/flow/wernerwengle_mod.F90:13-29."""
    ww = np.zeros(5)
    lib().og_ww_constants(rho, gmol, _p(ww))
    return ww


def wall_term(dd, u, rho, gmol):
    """swcle3d_one (src/flow/tstle4_mod.F90:1848-1859): Werner-Wengle wall
    shear stress (tauwin) of velocity u at wall distance dd, over rho dd."""
    ww = ww_constants(rho, gmol)
    return lib().og_swc(dd, u, rho, gmol, _p(ww))


def _p(a, ct=_dp):
    return a.ctypes.data_as(ct)


class Oracle:
    """FP64 restatement of the reference pressure path over one `Grids`."""

    def __init__(self, grids):
        self.g = grids
        self.lib = lib()
        g = grids
        self._keep = [g.kk, g.jj, g.ii, g.level, g.nbr, g.bc, g.ip3, g.ip1x, g.ip1y, g.ip1z,
                      g.iparent, g.position]
        self.topo = OgTopo(g.ngrid, g.minlevel, g.maxlevel, _p(g.kk, _i32p), _p(g.jj, _i32p),
                           _p(g.ii, _i32p), _p(g.level, _i32p), _p(g.nbr, _i32p),
                           _p(g.bc, _i32p), _p(g.ip3, _i64p), _p(g.ip1x, _i64p),
                           _p(g.ip1y, _i64p), _p(g.ip1z, _i64p), _p(g.iparent, _i32p),
                           _p(g.position, _i32p))
        self.ipbb = np.zeros(6 * g.ngrid, dtype=np.int64)
        self.nbuf = int(self.lib.og_par_buffers(C.byref(self.topo), _p(self.ipbb, _i64p)))
        # velocity boundary buffers of the FIX, OP1 and PAR faces
        self.bbo = np.zeros(6 * g.ngrid, dtype=np.int64)
        self.nbb = int(self.lib.og_bound_buffers(C.byref(self.topo), _p(self.bbo, _i64p)))
        cap = 26 * g.ngrid + 26
        self.conns = (OgConn * cap)()
        n = self.lib.og_connect_build(C.byref(self.topo), self.conns, cap)
        if n < 0:
            raise RuntimeError("connection table overflow")
        self.nconn = n

    # -- setup ---------------------------------------------------------
    def giteig(self, bp):
        g = self.g
        c = {k: np.zeros(n) for k, n in (("aw", g.n1x), ("ae", g.n1x), ("as", g.n1y),
                                          ("an", g.n1y), ("ab", g.n1z), ("at", g.n1z))}
        c["ap"] = np.zeros(g.n3d)
        self.lib.og_giteig(C.byref(self.topo), _p(g.dx), _p(g.dy), _p(g.dz), _p(bp),
                           _p(c["aw"]), _p(c["ae"]), _p(c["as"]), _p(c["an"]),
                           _p(c["ab"]), _p(c["at"]), _p(c["ap"]))
        return c

    def hp_tables(self):
        mip = np.zeros(self.g.n3d, dtype=np.int64)
        idx = np.zeros(self.g.n3d, dtype=np.int64)
        rc = self.lib.og_hp_tables(C.byref(self.topo), _p(mip, _i64p), _p(idx, _i64p))
        if rc != 0:
            raise RuntimeError(f"hyperplane table check failed ({rc})")
        return mip, idx

    def lu_hp(self, c, bp, mip, idx):
        f = {k: np.zeros(self.g.n3d) for k in ("lw", "ls", "lb", "lpr", "ue", "un", "ut")}
        self.lib.og_lu_hp(C.byref(self.topo), _p(c["aw"]), _p(c["ae"]), _p(c["as"]),
                          _p(c["an"]), _p(c["ab"]), _p(c["at"]), _p(c["ap"]), _p(bp),
                          _p(mip, _i64p), _p(idx, _i64p), *[_p(f[k]) for k in
                          ("lw", "ls", "lb", "lpr", "ue", "un", "ut")])
        return f

    def lu_classic(self, c, bp):
        f = {k: np.zeros(self.g.n3d) for k in ("lw", "ls", "lb", "lpr", "ue", "un", "ut")}
        self.lib.og_lu_classic(C.byref(self.topo), _p(c["aw"]), _p(c["ae"]), _p(c["as"]),
                               _p(c["an"]), _p(c["ab"]), _p(c["at"]), _p(c["ap"]), _p(bp),
                               *[_p(f[k]) for k in ("lw", "ls", "lb", "lpr", "ue", "un", "ut")])
        return f

    def sor_init(self, c):
        rap = np.zeros(self.g.n3d)
        self.lib.og_sor_init(C.byref(self.topo), _p(c["ap"]), _p(rap))
        return rap

    # -- kernels -------------------------------------------------------
    def laplacephi(self, res, phi, c, bp, level=-1):
        self.lib.og_laplacephi(C.byref(self.topo), level, _p(res), _p(phi), _p(c["aw"]),
                               _p(c["ae"]), _p(c["as"]), _p(c["an"]), _p(c["ab"]),
                               _p(c["at"]), _p(c["ap"]), _p(bp))

    def sipiter1_hp(self, rhs, res, f, mip, idx, level=-1):
        self.lib.og_sipiter1_hp(C.byref(self.topo), level, _p(rhs), _p(res), _p(f["lw"]),
                                _p(f["ls"]), _p(f["lb"]), _p(f["lpr"]), _p(mip, _i64p),
                                _p(idx, _i64p))

    def sipiter2_hp(self, phi, res, f, mip, idx, level=-1):
        self.lib.og_sipiter2_hp(C.byref(self.topo), level, _p(phi), _p(res), _p(f["ue"]),
                                _p(f["un"]), _p(f["ut"]), _p(mip, _i64p), _p(idx, _i64p))

    def sipiter1_cl(self, rhs, res, f, level=-1):
        self.lib.og_sipiter1_cl(C.byref(self.topo), level, _p(rhs), _p(res), _p(f["lw"]),
                                _p(f["ls"]), _p(f["lb"]), _p(f["lpr"]))

    def sipiter2_cl(self, phi, res, f, level=-1):
        self.lib.og_sipiter2_cl(C.byref(self.topo), level, _p(phi), _p(res), _p(f["ue"]),
                                _p(f["un"]), _p(f["ut"]))

    def relax(self, dp, rhs, c, rap, omega, bp=None, level=-1):
        self.lib.og_relax(C.byref(self.topo), level, _p(dp), _p(rhs), _p(c["aw"]),
                          _p(c["ae"]), _p(c["as"]), _p(c["an"]), _p(c["ab"]), _p(c["at"]),
                          _p(rap), _p(bp) if bp is not None else None, C.c_double(omega))

    def divcal(self, div, u, v, w, fak, bp=None, sdiv=None):
        g = self.g
        self.lib.og_divcal(C.byref(self.topo), _p(div), _p(u), _p(v), _p(w), _p(g.rddx),
                           _p(g.rddy), _p(g.rddz), C.c_double(fak),
                           _p(bp) if bp is not None else None,
                           _p(sdiv) if sdiv is not None else None)

    def tstle4(self, u, v, w, p, g_visc, rho, gmol, ilesmodel=0, gradp=None):
        """Momentum right-hand side (src/flow/tstle4_mod.F90:13-1753, noib,
        with the Werner-Wengle wall stress next to NOS walls and tstle4_par
        at PAR faces): returns (uo, vo, wo)."""
        g = self.g
        uo, vo, wo = g.zeros(), g.zeros(), g.zeros()
        gp = np.asarray(gradp, dtype=np.float64) if gradp is not None else None
        rc = self.lib.og_tstle4(C.byref(self.topo), _p(uo), _p(vo), _p(wo), _p(u), _p(v),
                                _p(w), _p(p), _p(g_visc), C.c_double(rho), C.c_double(gmol),
                                ilesmodel, _p(gp) if gp is not None else None,
                                _p(g.dx), _p(g.dy), _p(g.dz), _p(g.ddx), _p(g.ddy), _p(g.ddz),
                                _p(g.rdx), _p(g.rdy), _p(g.rdz), _p(g.rddx), _p(g.rddy),
                                _p(g.rddz))
        if rc != 0:
            raise ValueError(f"og_tstle4 failed ({rc})")
        return uo, vo, wo

    def tstle4_par(self, uo, vo, wo, u, v, w):
        """tstle4_par alone (src/flow/tstle4_mod.F90:532-1753): adds the PAR
        face convection terms to uo, vo, wo in place."""
        g = self.g
        self.lib.og_tstle4_par(C.byref(self.topo), _p(uo), _p(vo), _p(wo), _p(u), _p(v),
                               _p(w), _p(g.dx), _p(g.dy), _p(g.dz), _p(g.ddx), _p(g.ddy),
                               _p(g.ddz), _p(g.rdx), _p(g.rdy), _p(g.rdz), _p(g.rddx),
                               _p(g.rddy), _p(g.rddz))

    def rkstep(self, p, dp, rhsp, frhs, dtfu):
        """src/core/rungekutta_mod.F90:315-333, in place on p and dp."""
        self.lib.og_rkstep(C.c_int64(p.size), _p(p), _p(dp), _p(rhsp), C.c_double(frhs),
                           C.c_double(dtfu))

    def mgpcorr(self, u, v, w, p, dp, rfak, bp=None):
        g = self.g
        self.lib.og_mgpcorr(C.byref(self.topo), _p(u), _p(v), _p(w), _p(p), _p(dp),
                            _p(g.rdx), _p(g.rdy), _p(g.rdz), C.c_double(rfak),
                            _p(bp) if bp is not None else None)

    def maxabscal(self, phi):
        ml = np.zeros(self.g.maxlevel - self.g.minlevel + 1)
        m = self.lib.og_maxabscal(C.byref(self.topo), _p(phi), _p(ml))
        return m, ml

    def exchange(self, fields, level=-1, layers=1, corners=False, forward=0):
        arr = (_dp * len(fields))(*[_p(f) for f in fields])
        rc = self.lib.og_exchange(C.byref(self.topo), self.conns, self.nconn, level, layers,
                                  int(corners), forward, arr, len(fields))
        if rc != 0:
            raise RuntimeError(f"og_exchange failed ({rc})")

    def connections(self):
        return np.array([(c.recv_grid, c.send_grid, c.recv_face, c.send_face)
                         for c in self.conns[:self.nconn]], dtype=np.int32).reshape(-1, 4)

    # -- whole projection ----------------------------------------------
    def setup(self, bp=None, ityp=0):
        """giteig + tables + LU (+ SOR_RAP); returns a state dict."""
        if bp is None:
            bp = self.g.ones()
        c = self.giteig(bp)
        mip, idx = self.hp_tables()
        f = self.lu_classic(c, bp) if ityp == 2 else self.lu_hp(c, bp, mip, idx)
        rap = self.sor_init(c)
        return dict(bp=bp, coef=c, fac=f, mip=mip, idx=idx, rap=rap)

    def mgpoisl(self, st, u, v, w, p, dt, ityp=0, ninner=5, nouter=9, nouter_min=0,
                epcorr=0.0, eps_rel=0.0, omega=OMEGA_DEFAULT, rho=1.0, ibtype=0,
                sor_all_levels=0, sdiv=None, sync_by_caller=0, bound=None):
        """sdiv: the ghostcell divergence source (ibtype 1), None = zero.
        sync_by_caller: stop after ftoc(u, v, w, p) instead of running the
        closing parent, bound_flow and connect.  bound: dict with the FIX /
        OP1 boundary values "u", "v", "w" (bound_buffer() layout) and
        "pinf" ([ngrid*6] OP1 pressure), any of them optional."""
        g = self.g
        bd = bound or {}
        bu, bv, bw = (np.ascontiguousarray(bd[k], dtype=np.float64) if k in bd else None
                      for k in ("u", "v", "w"))
        pf = np.ascontiguousarray(bd["pinf"], dtype=np.float64) if "pinf" in bd else None
        prm = OgParams(ityp, ninner, nouter, nouter_min, epcorr, eps_rel, omega, rho, ibtype,
                       sor_all_levels, _p(sdiv) if sdiv is not None else None,
                       int(sync_by_caller), *[_p(a) if a is not None else None
                                              for a in (bu, bv, bw, pf)])
        hist = np.zeros(max(nouter, 1))
        stats = OgStats(0, 0, 0.0, 0.0, _p(hist))
        work = [np.zeros(g.n3d) for _ in range(4)]
        c, f = st["coef"], st["fac"]
        rc = self.lib.og_mgpoisl(
            C.byref(self.topo), self.conns, self.nconn, C.byref(prm), _p(u), _p(v), _p(w),
            _p(p), C.c_double(dt), _p(st["bp"]), _p(c["aw"]), _p(c["ae"]), _p(c["as"]),
            _p(c["an"]), _p(c["ab"]), _p(c["at"]), _p(c["ap"]),
            *[_p(f[k]) for k in ("lw", "ls", "lb", "lpr", "ue", "un", "ut")],
            _p(st["mip"], _i64p), _p(st["idx"], _i64p), _p(st["rap"]), _p(g.rdx), _p(g.rdy),
            _p(g.rdz), _p(g.rddx), _p(g.rddy), _p(g.rddz), _p(g.dx), _p(g.dy), _p(g.dz),
            _p(g.ddx), _p(g.ddy), _p(g.ddz), *[_p(a) for a in work],
            C.byref(stats))
        if rc != 0:
            raise RuntimeError(f"og_mgpoisl failed ({rc})")
        return dict(ipc=stats.ipc, nouter_done=stats.nouter_done, maxrhs0=stats.maxrhs0,
                    maxrhs=stats.maxrhs, history=hist[:stats.nouter_done].copy(),
                    dp=work[0], rhs=work[2])

    # -- multi-level coupling ------------------------------------------
    def par_buffer(self):
        return np.zeros(max(self.nbuf, 1))

    def bound_buffer(self):
        """A zeroed velocity boundary buffer (FIX, OP1 and PAR faces)."""
        return np.zeros(max(self.nbb, 1))

    def parent_uvw(self, level, u, v, w, ub, vb, wb):
        self.lib.og_parent_uvw(C.byref(self.topo), level, _p(u), _p(v), _p(w), _p(ub),
                               _p(vb), _p(wb), _p(self.bbo, _i64p))

    def bound_flow(self, level, u, v, w, p, ub, vb, wb, pinf=None, bp=None, rho=1.0):
        g = self.g
        pf = None if pinf is None else np.ascontiguousarray(pinf, dtype=np.float64)
        self.lib.og_bound_flow(C.byref(self.topo), level, _p(u), _p(v), _p(w), _p(p), _p(ub),
                               _p(vb), _p(wb), _p(self.bbo, _i64p),
                               _p(pf) if pf is not None else None,
                               _p(bp) if bp is not None else None, _p(g.ddx), _p(g.ddy),
                               _p(g.ddz), C.c_double(rho))

    def ctof(self, level, f):
        self.lib.og_ctof(C.byref(self.topo), level, _p(f), _p(f))

    def parent(self, level, f, pbuf):
        self.lib.og_parent(C.byref(self.topo), level, _p(f), _p(pbuf), _p(self.ipbb, _i64p))

    def bound_pressure(self, level, p, pbuf, bp=None):
        g = self.g
        self.lib.og_bound_pressure(C.byref(self.topo), level, _p(p), _p(pbuf),
                                   _p(self.ipbb, _i64p), _p(bp) if bp is not None else None,
                                   _p(g.dx), _p(g.dy), _p(g.dz), _p(g.ddx), _p(g.ddy),
                                   _p(g.ddz))

    def ftoc(self, level, f, bp=None, flag="P"):
        g = self.g
        self.lib.og_ftoc_s(C.byref(self.topo), level, _p(f), _p(f), _p(g.ddx), _p(g.ddy),
                           _p(g.ddz), _p(bp) if bp is not None else None,
                           C.c_char(flag.encode()))

    def ftoc_uvwp(self, level, u, v, w, p, bp=None):
        g = self.g
        self.lib.og_ftoc_uvwp(C.byref(self.topo), level, _p(u), _p(v), _p(w), _p(p),
                              _p(g.ddx), _p(g.ddy), _p(g.ddz),
                              _p(bp) if bp is not None else None)

    def par_ftoc_norm(self, level, u, v, w, sum=False):
        g = self.g
        self.lib.og_par_ftoc_norm(C.byref(self.topo), level, _p(u), _p(v), _p(w),
                                  _p(g.ddx), _p(g.ddy), _p(g.ddz), int(bool(sum)))

    def time_sip(self, st, rhs, niter, nthreads=0):
        """Seconds for `niter` hyperplane-SIP inner iterations (CPU baseline)."""
        g = self.g
        hilf = np.zeros(g.n3d)
        res = np.zeros(g.n3d)
        c, f = st["coef"], st["fac"]
        return self.lib.og_time_sip(
            C.byref(self.topo), self.conns, self.nconn, niter, nthreads, _p(rhs), _p(hilf),
            _p(res), _p(c["aw"]), _p(c["ae"]), _p(c["as"]), _p(c["an"]), _p(c["ab"]),
            _p(c["at"]), _p(c["ap"]), _p(st["bp"]),
            *[_p(f[k]) for k in ("lw", "ls", "lb", "lpr", "ue", "un", "ut")],
            _p(st["mip"], _i64p), _p(st["idx"], _i64p))


def cpu_count():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1
