"""One Runge-Kutta stage of the flow solver around the projection
(src/flow/timeintegration_mod.F90:22-142), on the library's device kernels:
tstle4 (with tstle4_par at the PAR faces of refined grids), the 2N-storage
update of U, V, W, then per level the velocity exchange the divergence needs,
parent and bound_flow (src/flow/timeintegration_mod.F90:119-123), then
mgpoisl.

Covered: nested grids, walls and periodic faces, no immersed boundary, no
LES model (the viscosity field G is the caller's), no buoyancy or Coriolis
term, no time-dependent inflow.  The reference exchanges only the
face-normal component of each velocity before the divergence
(`connect(..., normal=.true., forward=1)`); here all three go through the
forward exchange, which fills the same ghosts the divergence reads and ghosts
that mgpoisl's closing two-layer exchange overwrites.
"""
from __future__ import annotations

# Williamson's low-storage RK3 (src/core/rungekutta_mod.F90:77-92)
WILLIAMSON = {
    "a": (0.0, -5.0 / 9.0, -153.0 / 128.0),
    "b": (1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0),
    "c": (0.0, 1.0 / 3.0, 3.0 / 4.0),
}


def coeffs(irk, scheme=WILLIAMSON):
    """(frhs, fu, dtrk) of stage irk (1-based): src/core/rungekutta_mod.F90:
    240-268 (get_coeffs_a/_b)."""
    n = len(scheme["a"])
    if irk < 1 or irk > n:
        raise ValueError("irk out of range")
    frhs = scheme["a"][irk - 1]
    fu = scheme["b"][irk - 1]
    dtrk = scheme["c"][irk] if irk < n else 1.0
    return frhs, fu, dtrk


def rk_stage(solver, u, v, w, p, du, dv, dw, g, dt, irk, gmol, ilesmodel=0,
             gradp=None, ittot=0, scheme=WILLIAMSON):
    """Advance U, V, W, P (CUDA float64 tensors, reference layout) by stage
    irk; DU, DV, DW are the 2N-storage increments (zero before stage 1).
    Returns mgpoisl's statistics."""
    frhs, fu, dtrk = coeffs(irk, scheme)
    uo, vo, wo = solver.tstle4(u, v, w, p, g, gmol, ilesmodel, gradp)
    solver.rkstep(u, du, uo, frhs, dt * fu)
    solver.rkstep(v, dv, vo, frhs, dt * fu)
    solver.rkstep(w, dw, wo, frhs, dt * fu)
    for level in range(solver.grids.minlevel, solver.grids.maxlevel + 1):
        solver.connect([u, v, w], level=level, layers=1, forward=1)
        solver.flow_bc(u, v, w, p, level)
    return solver.mgpoisl(u, v, w, p, dtrk * dt, ittot=ittot, irk=irk)
