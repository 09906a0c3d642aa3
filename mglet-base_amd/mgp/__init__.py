"""mgp -- Python front end of the MI355X pressure-projection library.

`PressureSolver` mirrors the Fortran module `pressuresolver_mod`
(src/flow/pressuresolver_mod.F90:45): construction = init_pressuresolver
(+ ib%giteig), `mgpoisl` = mgpoisl, `close` = finish_pressuresolver.  The
keyword arguments are the `/flow/pressuresolver` keys with the reference's
defaults and validation (:57-142).  Fields are flat float64 arrays in the
reference layout (see mgp.grids): numpy arrays are copied through host
memory, torch CUDA tensors are used in place.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import MgpConfig, MgpError, MgpGrids, MgpStats
from .grids import Grids, uniform_box  # noqa: F401

__all__ = ["PressureSolver", "MgpError", "Grids", "uniform_box"]

_TYPES = {"sip": 0, "sor": 1, "classic": 2}
_IB = {"noib": 0, "ghostcell": 1}
_PREC = {"fp64": 0, "fp32": 1}
NUM_PLAN_SHAPES = 9       # MGP_NUM_PLAN_SHAPES


def _ptr(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


# The reference reads omega with default_value=1.1, a default-real (real32)
# literal that config_mod widens with REAL(default_value, real64)
# (src/flow/pressuresolver_mod.F90:135): the FP64 default is (double)1.1f.
OMEGA_DEFAULT = float(np.float32(1.1))


def _check_host(arrays, n):
    for a in arrays:
        if (not isinstance(a, np.ndarray) or a.dtype != np.float64
                or not a.flags.c_contiguous or a.size != n):
            raise MgpError("fields must be contiguous float64 arrays of the 3-D length")


def _check_device(tensors, n):
    """Every tensor handed to a device entry point must be a contiguous
    CUDA float64 tensor of the 3-D length: the kernels index n values from
    the raw pointer, so anything else would be an out-of-bounds or host
    access on the GPU.  Raises before any library call."""
    try:
        import torch
    except ImportError as e:  # pragma: no cover - torch is part of the image
        raise MgpError("device fields need torch") from e
    for a in tensors:
        if (not isinstance(a, torch.Tensor) or a.dtype != torch.float64 or not a.is_cuda
                or not a.is_contiguous() or a.numel() != n):
            raise MgpError("fields must be contiguous CUDA float64 tensors of the 3-D length")


class PressureSolver:
    def __init__(self, grids: Grids, type="sip", ninner=5, nouter=9, nouter_min=0,
                 epcorr=0.0, omega=OMEGA_DEFAULT, rho=1.0, ib="noib", precision="fp64",
                 eps_rel=0.0, check_every=4, bp=None, blocked=None, sor_all_levels=False,
                 alfa=0.0, sync_by_caller=False):
        self.lib = _lib.load()
        self.grids = g = grids
        t = type.lower()
        if t not in _TYPES:
            raise MgpError(f"Invalid pressure solver type: {type}")
        self._keep = [g.kk, g.jj, g.ii, g.level, g.nbr, g.bc,
                      g.ip3 + 1, g.ip1x + 1, g.ip1y + 1, g.ip1z + 1,
                      np.ascontiguousarray(g.iparent, dtype=np.int32),
                      np.ascontiguousarray(g.position, dtype=np.int32)]
        k = self._keep
        nested = bool(np.any(g.iparent))
        self._grids = MgpGrids(g.ngrid, g.minlevel, g.maxlevel,
                               _ptr(k[0], C.c_int32), _ptr(k[1], C.c_int32),
                               _ptr(k[2], C.c_int32), _ptr(k[3], C.c_int32),
                               _ptr(k[4], C.c_int32), _ptr(k[5], C.c_int32),
                               _ptr(k[6], C.c_int64), _ptr(k[7], C.c_int64),
                               _ptr(k[8], C.c_int64), _ptr(k[9], C.c_int64),
                               _ptr(k[10], C.c_int32) if nested else None,
                               _ptr(k[11], C.c_int32) if nested else None)
        self.cfg = MgpConfig(_TYPES[t], ninner, nouter, nouter_min, epcorr, omega, rho,
                             _IB[ib], _PREC[precision], eps_rel, int(bool(sor_all_levels)),
                             check_every, alfa, int(bool(sync_by_caller)))
        h = C.c_void_p()
        if blocked is not None:
            blocked = np.ascontiguousarray(blocked, dtype=np.int8)
            rc = self.lib.mgp_init_mask(C.byref(self._grids), C.byref(self.cfg), _ptr(g.dx),
                                        _ptr(g.dy), _ptr(g.dz), _ptr(g.ddx), _ptr(g.ddy),
                                        _ptr(g.ddz), _ptr(blocked, C.c_int8), C.byref(h))
        else:
            bpa = None if bp is None else np.ascontiguousarray(bp, dtype=np.float64)
            rc = self.lib.mgp_init(C.byref(self._grids), C.byref(self.cfg), _ptr(g.dx),
                                   _ptr(g.dy), _ptr(g.dz), _ptr(g.ddx), _ptr(g.ddy),
                                   _ptr(g.ddz), _ptr(bpa) if bpa is not None else None,
                                   C.byref(h))
        _lib.check(rc, "mgp_init")
        self.h = h

    # ------------------------------------------------------------------
    def set_iterations(self, ninner=None, nouter=None, nouter_min=None, epcorr=None,
                       eps_rel=None, check_every=None):
        c = self.cfg
        c.ninner = c.ninner if ninner is None else ninner
        c.nouter = c.nouter if nouter is None else nouter
        c.nouter_min = c.nouter_min if nouter_min is None else nouter_min
        c.epcorr = c.epcorr if epcorr is None else epcorr
        c.eps_rel = c.eps_rel if eps_rel is None else eps_rel
        c.check_every = c.check_every if check_every is None else check_every
        _lib.check(self.lib.mgp_set_iterations(self.h, c.ninner, c.nouter, c.nouter_min,
                                               c.epcorr, c.eps_rel, c.check_every),
                   "mgp_set_iterations")

    def mgpoisl(self, u, v, w, p, dt, ittot=0, irk=0):
        """Project (u, v, w) and update p in place; returns the stats."""
        st = MgpStats()
        if isinstance(u, np.ndarray):
            _check_host((u, v, w, p), self.grids.n3d)
            rc = self.lib.mgp_mgpoisl(self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(p), dt, ittot,
                                      irk, C.byref(st))
        else:
            _check_device((u, v, w, p), self.grids.n3d)
            self._bind_stream()
            rc = self.lib.mgp_mgpoisl_device(self.h, u.data_ptr(), v.data_ptr(), w.data_ptr(),
                                             p.data_ptr(), dt, ittot, irk, C.byref(st))
        _lib.check(rc, "mgp_mgpoisl")
        return dict(ipc=st.ipc, nouter_done=st.nouter_done, converged=bool(st.converged),
                    maxrhs0=st.maxrhs0, maxrhs=st.maxrhs, solve_ms=st.solve_ms)

    def adjoint(self, ubar, vbar, wbar, dt):
        """Adjoint of `mgpoisl` with respect to the interior face velocities:
        (ubar, vbar, wbar) = dJ/du' on entry, dJ/du on exit (in place).  The
        adjoint Poisson problem uses this solver's iteration settings."""
        st = MgpStats()
        if isinstance(ubar, np.ndarray):
            _check_host((ubar, vbar, wbar), self.grids.n3d)
            rc = self.lib.mgp_adjoint(self.h, _ptr(ubar), _ptr(vbar), _ptr(wbar), dt,
                                      C.byref(st))
        else:
            _check_device((ubar, vbar, wbar), self.grids.n3d)
            self._bind_stream()
            rc = self.lib.mgp_adjoint_device(self.h, ubar.data_ptr(), vbar.data_ptr(),
                                             wbar.data_ptr(), dt, C.byref(st))
        _lib.check(rc, "mgp_adjoint")
        return dict(ipc=st.ipc, nouter_done=st.nouter_done, converged=bool(st.converged),
                    maxrhs0=st.maxrhs0, maxrhs=st.maxrhs, solve_ms=st.solve_ms)

    def _bind_stream(self):
        # the library works on its own stream and synchronises it before
        # returning; torch work queued before the call must be complete
        import torch
        torch.cuda.current_stream().synchronize()

    def get_field(self, name, igrid=None):
        """A library field in the reference layout; with igrid (1-based)
        only that grid's kk x jj x ii block of a 3-D field."""
        g = self.grids
        if igrid is not None:
            if not 1 <= igrid <= g.ngrid:
                raise MgpError(f"invalid grid {igrid}")
            b = igrid - 1
            out = np.zeros(int(g.kk[b]) * int(g.jj[b]) * int(g.ii[b]))
            _lib.check(self.lib.mgp_get_field_grid(self.h, name.encode(), igrid, _ptr(out)),
                       "mgp_get_field_grid")
            return out
        n = {"GSAW": g.n1x, "GSAE": g.n1x, "GSAS": g.n1y, "GSAN": g.n1y,
             "GSAB": g.n1z, "GSAT": g.n1z}.get(name, g.n3d)
        out = np.zeros(n)
        _lib.check(self.lib.mgp_get_field(self.h, name.encode(), _ptr(out)), "mgp_get_field")
        return out

    def connect(self, fields, level=-1, layers=1, corners=False, forward=0):
        """Ghost-layer exchange of CUDA float64 tensors (reference layout)."""
        _check_device(fields, self.grids.n3d)
        self._bind_stream()
        arr = (C.c_void_p * len(fields))(*[f.data_ptr() for f in fields])
        _lib.check(self.lib.mgp_connect_device(self.h, level, layers, int(corners), forward,
                                               arr, len(fields)), "mgp_connect_device")

    def flow_bc(self, u, v, w, p, level):
        """parent, then bound_flow%bound, of one level on CUDA float64
        tensors (reference layout): U, V, W at the PAR faces of the level's
        refined grids from their parents, ghost values at FIX, OP1, NOS,
        SLI and PAR faces (src/ib/parent/parent_mod.F90:16-52,
        src/flow/bound_flow_mod.F90)."""
        _check_device((u, v, w, p), self.grids.n3d)
        self._bind_stream()
        _lib.check(self.lib.mgp_flow_bc_device(self.h, level, u.data_ptr(), v.data_ptr(),
                                               w.data_ptr(), p.data_ptr()),
                   "mgp_flow_bc_device")

    def par_ftoc_norm(self, u, v, w, level, sum=False):
        """par_ftoc_norm of one level on CUDA float64 tensors (reference
        layout): the face-normal velocity at the PAR faces of the level's
        grids into the coincident coarse faces (src/ib/par_ftoc_mod.F90:
        35-475); area-weighted mean of each 2x2 group, or the sum."""
        _check_device((u, v, w), self.grids.n3d)
        self._bind_stream()
        _lib.check(self.lib.mgp_par_ftoc_device(self.h, level, u.data_ptr(), v.data_ptr(),
                                                w.data_ptr(), int(bool(sum))),
                   "mgp_par_ftoc_device")

    def divcal(self, u, v, w, fak=1.0):
        import torch
        _check_device((u, v, w), self.grids.n3d)
        self._bind_stream()
        div = torch.zeros_like(u)
        _lib.check(self.lib.mgp_divcal_device(self.h, div.data_ptr(), u.data_ptr(),
                                              v.data_ptr(), w.data_ptr(), fak),
                   "mgp_divcal_device")
        return div

    def set_sdiv(self, sdiv):
        """The ghostcell divergence source SDIV (flat float64 array of the
        3-D length, copied to the device); None resets it to zero."""
        if sdiv is None:
            _lib.check(self.lib.mgp_set_sdiv(self.h, None), "mgp_set_sdiv")
            return
        a = np.ascontiguousarray(sdiv, dtype=np.float64)
        if a.size != self.grids.n3d:
            raise MgpError("SDIV must have the 3-D length")
        _lib.check(self.lib.mgp_set_sdiv(self.h, _ptr(a)), "mgp_set_sdiv")

    def bound_layout(self):
        """(offsets [ngrid*6] int64, n): where the FIX, OP1 and PAR faces'
        velocity boundary values start in the buffers of set_bound (-1:
        none) and their total length."""
        n = C.c_int64()
        _lib.check(self.lib.mgp_bound_layout(self.h, None, C.byref(n)), "mgp_bound_layout")
        off = np.zeros(6 * self.grids.ngrid, dtype=np.int64)
        _lib.check(self.lib.mgp_bound_layout(self.h, _ptr(off, C.c_int64), C.byref(n)),
                   "mgp_bound_layout")
        return off, n.value

    def set_bound(self, u=None, v=None, w=None, pinf=None):
        """FIX / OP1 boundary values of U, V, W (bound_layout() order, float64
        of its length) and the OP1 pressure per grid and face ([ngrid*6]);
        None sets an array to zero."""
        _, n = self.bound_layout()
        arrs = []
        for a, m, nm in ((u, n, "u"), (v, n, "v"), (w, n, "w"),
                         (pinf, 6 * self.grids.ngrid, "pinf")):
            if a is None:
                arrs.append(None)
                continue
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.size != m:
                raise MgpError(f"set_bound: {nm} must have {m} values")
            arrs.append(a)
        _lib.check(self.lib.mgp_set_bound(self.h, *[_ptr(a) if a is not None else None
                                                    for a in arrs]), "mgp_set_bound")

    def tstle4(self, u, v, w, p, g, gmol, ilesmodel=0, gradp=None):
        """Momentum right-hand side of one RK stage (tstle4 with tstle4_par
        at PAR faces, no immersed boundary) on CUDA float64 tensors; returns
        (uo, vo, wo)."""
        import torch
        _check_device((u, v, w, p, g), self.grids.n3d)
        gp = None
        if gradp is not None:
            gradp = np.ascontiguousarray(gradp, dtype=np.float64)
            if gradp.size != 3:
                raise MgpError("gradp must hold three values")
            gp = _ptr(gradp)
        self._bind_stream()
        uo, vo, wo = (torch.empty_like(u) for _ in range(3))
        _lib.check(self.lib.mgp_tstle4_device(self.h, uo.data_ptr(), vo.data_ptr(),
                                              wo.data_ptr(), u.data_ptr(), v.data_ptr(),
                                              w.data_ptr(), p.data_ptr(), g.data_ptr(),
                                              gmol, ilesmodel, gp),
                   "mgp_tstle4_device")
        return uo, vo, wo

    def rkstep(self, p, dp, rhsp, frhs, dtfu):
        """2N-storage Runge-Kutta update in place: dp = frhs dp + rhsp,
        p += dtfu dp (CUDA float64 tensors of the 3-D length)."""
        _check_device((p, dp, rhsp), self.grids.n3d)
        self._bind_stream()
        _lib.check(self.lib.mgp_rkstep_device(self.h, p.data_ptr(), dp.data_ptr(),
                                              rhsp.data_ptr(), frhs, dtfu),
                   "mgp_rkstep_device")

    def time_kernels(self, niter=10):
        ms = np.zeros(3)
        _lib.check(self.lib.mgp_time_kernels(self.h, niter, _ptr(ms)), "mgp_time_kernels")
        return dict(forward_ms=ms[0], backward_ms=ms[1], outer_ms=ms[2])

    def plan_shape(self, request=-1):
        """Pipeline shape of the plane kernels.  request -1: report; 0: run
        the autotuner again; 1..NUM_PLAN_SHAPES: set.  Returns (shape, ms per
        shape as last timed by the autotuner, 0 where not timed)."""
        chosen = C.c_int32()
        ms = np.zeros(NUM_PLAN_SHAPES)
        _lib.check(self.lib.mgp_plan_shape(self.h, request, C.byref(chosen), _ptr(ms)),
                   "mgp_plan_shape")
        return chosen.value, ms

    def plan_shape_backward(self):
        """The backward sweep's pipeline shape (the autotuner picks it apart
        from the others; a set request sets both)."""
        chosen = C.c_int32()
        _lib.check(self.lib.mgp_plan_shape_backward(self.h, C.byref(chosen)),
                   "mgp_plan_shape_backward")
        return chosen.value

    def history(self):
        """max |RHS| after each outer iteration of the last mgpoisl call."""
        cap = max(1, self.cfg.nouter)
        out = np.zeros(cap)
        n = C.c_int32()
        _lib.check(self.lib.mgp_get_history(self.h, _ptr(out), cap, C.byref(n)),
                   "mgp_get_history")
        return out[:n.value].copy()

    def info(self):
        a, b = C.c_int64(), C.c_int64()
        _lib.check(self.lib.mgp_info(self.h, C.byref(a), C.byref(b)), "mgp_info")
        return dict(ncells=a.value, device_bytes=b.value)

    def factor_classes(self):
        """(groups of blocks sharing one copy of the SIP factors, factor
        cells stored per factor array)."""
        a, b = C.c_int32(), C.c_int64()
        _lib.check(self.lib.mgp_factor_classes(self.h, C.byref(a), C.byref(b)),
                   "mgp_factor_classes")
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            _lib.check(self.lib.mgp_finish(self.h), "mgp_finish")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
