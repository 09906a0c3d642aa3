"""Host-side argument checks of the Python binding and the reference's
default-real literals (no GPU: nothing here reaches a kernel)."""
import inspect
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _bare_solver(g):
    """A PressureSolver without a device handle: every device entry point
    must reject bad tensors before it touches the library."""
    import mgp
    s = object.__new__(mgp.PressureSolver)
    s.grids = g
    s.h = None
    s.lib = None     # any library call would fail with AttributeError
    return s


def test_device_wrappers_reject_bad_tensors():
    import torch
    import mgp
    g = mgp.uniform_box((1, 1, 2), (4, 4, 4))
    s = _bare_solver(g)
    n = g.n3d
    good_shape_cpu = torch.zeros(n, dtype=torch.float64)
    bad = [good_shape_cpu,                                # host tensor
           torch.zeros(n, dtype=torch.float32),           # wrong dtype
           torch.zeros(n - 1, dtype=torch.float64),       # too short
           torch.zeros(2 * n, dtype=torch.float64)[::2],  # not contiguous
           np.zeros(n)]                                   # not a tensor
    calls = [lambda t: s.connect([t]),
             lambda t: s.divcal(t, t, t),
             lambda t: s.tstle4(t, t, t, t, t, 1e-2),
             lambda t: s.rkstep(t, t, t, 0.0, 1.0),
             lambda t: s.mgpoisl(t, t, t, t, 1e-3) if not isinstance(t, np.ndarray)
             else s.mgpoisl(t[:-1], t, t, t, 1e-3),
             lambda t: s.adjoint(t, t, t, 1e-3) if not isinstance(t, np.ndarray)
             else s.adjoint(t.astype(np.float32), t, t, 1e-3)]
    for call in calls:
        for t in bad:
            with pytest.raises(mgp.MgpError):
                call(t)


def test_omega_default_is_the_real32_literal():
    """src/flow/pressuresolver_mod.F90:135 reads omega with default_value=1.1,
    a default-real literal widened to real64: (double)1.1f, not 1.1."""
    import mgp
    import oracle
    want = float(np.float32(1.1))
    assert want != 1.1
    sig = inspect.signature(mgp.PressureSolver.__init__)
    assert sig.parameters["omega"].default == want
    assert inspect.signature(oracle.Oracle.mgpoisl).parameters["omega"].default == want
    src = (ROOT / "mglet-base_amd" / "fortran" / "mgp_bindc_mod.F90").read_text()
    assert "omega = REAL(1.1, C_DOUBLE)" in src


def test_config_tail_field_layout():
    """mgp_config ends with alfa (0 = the reference's (double)0.92f) and
    sync_by_caller (0 = the library runs the closing sync); the ctypes and
    Fortran mirrors carry both in that order."""
    import ctypes
    from mgp import _lib
    assert _lib.MgpConfig._fields_[-2:] == [("alfa", ctypes.c_double),
                                            ("sync_by_caller", ctypes.c_int32)]
    src = (ROOT / "mglet-base_amd" / "fortran" / "mgp_bindc_mod.F90").read_text()
    assert src.index("REAL(C_DOUBLE) :: alfa = 0.0_C_DOUBLE") < \
        src.index("INTEGER(C_INT) :: sync_by_caller = 0") < src.index("END TYPE mgp_config_t")
    hdr = (ROOT / "include" / "mgp.h").read_text()
    assert hdr.index("mgletreal alfa;") < hdr.index("mgletint sync_by_caller;") < \
        hdr.index("} mgp_config;")


def test_roctx_ranges_on_entry_points_and_launchers():
    """Every C-ABI entry point that does work and every kernel launcher opens
    a roctx range (src/core/timer_mod.F90:114 times the same calls), and the
    library links the roctx runtime."""
    import re
    import subprocess
    csrc = ROOT / "mglet-base_amd" / "csrc"
    api = (csrc / "mgp_api.cpp").read_text()
    body = api[api.index('extern "C" {'):]
    entries = re.findall(r'^int (mgp_\w+)\([^{]*\{\s*\n\s*MGP_RANGE\("(\w+)"\)', body, re.M)
    declared = re.findall(r'^int (mgp_\w+)\(', body, re.M)
    assert declared and [e[0] for e in entries] == declared
    assert all(a == b for a, b in entries)
    for f in csrc.glob("*.hip"):
        s = f.read_text()
        for m in re.finditer(r'^void (launch_\w+)\(', s, re.M):
            if m.group(1) == "launch_pull_nf":      # called by launch_pull only
                continue
            b = s.index("{", m.end())
            assert "MGP_RANGE(" in s[b:b + 80], (f.name, m.group(1))
    from mgp import _lib
    out = subprocess.run(["readelf", "-d", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    assert "roctx" in out
