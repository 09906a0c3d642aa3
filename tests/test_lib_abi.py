"""The C-ABI library loads and exports every symbol include/mgp.h declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import subprocess

import pytest


def test_header_symbols_exported():
    from mgp import _lib
    lib = _lib.load()
    names = _lib.exported_symbols()
    assert len(names) >= 13
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    dyn = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= dyn


def test_errors_without_handle():
    from mgp import _lib
    lib = _lib.load()
    assert lib.mgp_finish(None) == 5
    assert b"not initialised" in lib.mgp_last_error()
    assert lib.mgp_mgpoisl(None, None, None, None, None, 1.0, 0, 0, None) == 5
    n = C.c_int64(-1)
    assert lib.mgp_bound_layout(None, None, C.byref(n)) == 5
    assert lib.mgp_set_bound(None, None, None, None, None) == 5


def test_invalid_config_rejected():
    """Parameter validation mirrors src/flow/pressuresolver_mod.F90:59-107."""
    import numpy as np
    import mgp
    from mgp import _lib
    g = mgp.uniform_box((1, 1, 1), (4, 4, 4))
    for kw, msg in ((dict(ninner=-1), "ninner"), (dict(nouter=-1), "nouter"),
                    (dict(type="sor", omega=2.0), "omega"), (dict(epcorr=-1.0), "epcorr"),
                    (dict(alfa=1.5), "alfa")):
        with pytest.raises(_lib.MgpError, match=msg):
            mgp.PressureSolver(g, **kw)
    with pytest.raises(_lib.MgpError, match="type"):
        mgp.PressureSolver(g, type="multigrid")
