"""Write tests/golden/topologies.json from the reference's own test grids.

The reference's test cases ship their block topologies as HDF5 files
(tests/<case>/grids.h5: the GRIDINFO compound of src/core/gridio_mod.F90:
15-52 and one boundary-condition compound per face, FRONT/BACK/RIGHT/LEFT/
BOTTOM/TOP, read by src/core/grids_mod.F90:203-250).  This script reads
them with the HDF5 command-line dumper (there is no h5py in this image) and
keeps, per case and grid, what the pressure path needs: IGRID, II, JJ, KK,
LEVEL, IPARENT, I/J/KPOSITION, BBOX, NBRGRID and the boundary-condition type
names of the six faces.  Only data is written; it runs where the reference
is mounted, and the tests read the JSON it leaves.

    python tests/golden/make_topologies.py [/path/to/reference]
"""
from __future__ import annotations

import json
import re
import subprocess
import sys
from pathlib import Path

H5DUMP = "/opt/conda/bin/h5dump"
# face datasets in face order 1..6 (-x, +x, -y, +y, -z, +z); the names are
# the reference's own (src/core/grids_mod.F90:203-250)
FACE_SETS = ["FRONT", "BACK", "RIGHT", "LEFT", "BOTTOM", "TOP"]
KEEP = ["IGRID", "II", "JJ", "KK", "LEVEL", "IPARENT", "IPOSITION",
        "JPOSITION", "KPOSITION", "BBOX", "NBRGRID"]
_NUM = re.compile(r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?")


def dump(path: Path, dset: str) -> tuple[str, str]:
    out = subprocess.run([H5DUMP, "-d", "/" + dset, "-y", "-w", "0", str(path)],
                         check=True, capture_output=True, text=True).stdout
    head, _, data = out.partition("DATA {")
    # attributes of the dataset follow its data
    return head, data.split("ATTRIBUTE")[0]


def members(head: str) -> list[tuple[str, int]]:
    """(name, element count) of each compound member, in order."""
    out = []
    for dims, name in re.findall(
            r'(?:H5T_ARRAY \{ ((?:\[\d+\])+) (?:[^{}]|\{[^{}]*\})*?\}|H5T_\w+) "(\w+)";', head):
        n = 1
        for d in re.findall(r"\[(\d+)\]", dims):
            n *= int(d)
        out.append((name, n))
    return out


def read_gridinfo(path: Path) -> list[dict]:
    head, data = dump(path, "GRIDINFO")
    mem = members(head)
    per = sum(n for _, n in mem)
    vals = [float(x) for x in _NUM.findall(data)]
    if len(vals) % per:
        raise ValueError(f"{path}: GRIDINFO has {len(vals)} numbers, not a multiple of {per}")
    grids = []
    for r in range(len(vals) // per):
        rec, pos = {}, r * per
        for name, n in mem:
            v = vals[pos:pos + n]
            pos += n
            if name in KEEP:
                if name == "BBOX":
                    rec[name] = v
                else:
                    iv = [int(x) for x in v]
                    rec[name] = iv if n > 1 else iv[0]
        grids.append(rec)
    return grids


def read_faces(path: Path, ngrid: int) -> list[list[str]]:
    """First boundary-condition type name of each face, per grid."""
    types = [[""] * 6 for _ in range(ngrid)]
    for f, ds in enumerate(FACE_SETS):
        _, data = dump(path, ds)
        recs = re.findall(r'\[ "([^"]*)", "([^"]*)" \]', data)
        if len(recs) != ngrid:
            raise ValueError(f"{path}:{ds}: {len(recs)} records for {ngrid} grids")
        for g, (t1, _t2) in enumerate(recs):
            types[g][f] = t1.split("\\000")[0].strip()
    return types


def main(ref: Path) -> None:
    cases = {}
    for gpath in sorted((ref / "tests").glob("*/grids.h5")):
        grids = read_gridinfo(gpath)
        faces = read_faces(gpath, len(grids))
        for g, t in zip(grids, faces):
            g["BC"] = t
        cases[gpath.parent.name] = {"source": f"tests/{gpath.parent.name}/grids.h5",
                                    "grids": grids}
    out = Path(__file__).resolve().parent / "topologies.json"
    with open(out, "w") as fh:
        json.dump(cases, fh, separators=(",", ":"))
        fh.write("\n")
    for name, c in cases.items():
        print(f"{name}: {len(c['grids'])} grids")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
