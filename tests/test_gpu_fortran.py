"""The Fortran side of the drop-in: a Fortran program calls init_pressuresolver
/ mgpoisl / finish_pressuresolver of mgp_pressuresolver_mod, which binds
libmgp through ISO_C_BINDING; its output must equal the oracle bitwise."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cases import topology, velocity

pytestmark = pytest.mark.gpu
DRIVER = Path(__file__).resolve().parent.parent / "mglet-base_amd" / "_build" / "mgp_driver"


def _write_case(path, g, u, v, w, p, ityp, ninner, nouter, dt):
    with open(path, "wb") as f:
        np.array([g.ngrid, g.minlevel, g.maxlevel, g.n3d, g.n1x, g.n1y, g.n1z],
                 dtype=np.int64).tofile(f)
        for a in (g.kk, g.jj, g.ii, g.level, g.nbr.reshape(-1), g.bc.reshape(-1)):
            np.asarray(a, dtype=np.int32).tofile(f)
        for a in (g.ip3, g.ip1x, g.ip1y, g.ip1z):
            (np.asarray(a, dtype=np.int64) + 1).tofile(f)
        for a in (g.dx, g.dy, g.dz, g.ddx, g.ddy, g.ddz, g.ones(), u, v, w, p):
            np.asarray(a, dtype=np.float64).tofile(f)
        np.array([ityp, ninner, nouter, 0, 0, 0], dtype=np.int32).tofile(f)
        np.array([0.0, 1.1, 1.0, 0.0, dt], dtype=np.float64).tofile(f)


@pytest.mark.parametrize("name,ityp", [("box2", 0), ("aniso", 2)])
def test_fortran_driver_bitwise(oracle_mod, tmp_path, name, ityp):
    if not DRIVER.exists():
        pytest.skip("mgp_driver not built (flang)")
    g = topology(name)
    o = oracle_mod.Oracle(g)
    st = o.setup(ityp=ityp)
    u, v, w = velocity(g, o, 7)
    p = g.zeros()
    _write_case(tmp_path / "case.bin", g, u, v, w, p, ityp, 5, 3, 1e-3)
    so = o.mgpoisl(st, u, v, w, p, 1e-3, ityp=ityp, ninner=5, nouter=3)
    r = subprocess.run([str(DRIVER), str(tmp_path / "case.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float64, count=4 * g.n3d)
    for k, a in enumerate((u, v, w, p)):
        assert np.array_equal(out[k * g.n3d:(k + 1) * g.n3d], a), "uvwp"[k]
    ints = np.fromfile(tmp_path / "out.bin", dtype=np.int32, count=3, offset=4 * g.n3d * 8)
    assert ints[0] == so["ipc"]
