"""The reference's own test grids (tests/*/grids.h5, kept as data in
tests/golden/topologies.json) against the direction tables and the oracle's
connection indexing (src/core/connect/connect_core_mod.F90:37-433,
conn1_mod.F90:31-129, connect2_mod.F90:75-138)."""
import numpy as np
import pytest

from cases import SUPPORTED_TOPOLOGIES, golden_topologies, lattice_ids, reference_topology
from mgp.grids import DIR_OFFSET, FACELIST


def test_fixture_holds_every_reference_case():
    topo = golden_topologies()
    assert len(topo) == 13
    assert {k: len(v["grids"]) for k, v in topo.items()}["Checkblock"] == 344
    for name, case in topo.items():
        lv = [r["LEVEL"] for r in case["grids"]]
        assert lv == sorted(lv), name          # numbered level by level


@pytest.mark.parametrize("name", sorted(golden_topologies()))
def test_nbrgrid_matches_bounding_boxes(name):
    """Every NBRGRID entry of the reference's files names the block that
    touches the grid in that direction: offset by the grid's own extent along
    each axis of the direction (FACELIST), modulo the periodic domain, same
    level, same size.  This pins the 26-direction table against the
    reference's data."""
    recs = golden_topologies()[name]["grids"]
    box = {r["IGRID"]: np.array(r["BBOX"]).reshape(3, 2) for r in recs}
    lvl = {r["IGRID"]: r["LEVEL"] for r in recs}
    lo = np.min([box[g][:, 0] for g in box if lvl[g] == min(lvl.values())], axis=0)
    hi = np.max([box[g][:, 1] for g in box if lvl[g] == min(lvl.values())], axis=0)
    span = hi - lo
    nchecked = 0
    for r in recs:
        b = box[r["IGRID"]]
        ext = b[:, 1] - b[:, 0]
        for d, nb in enumerate(r["NBRGRID"]):
            if nb <= 0:
                continue
            want = b[:, 0] + np.array(DIR_OFFSET[d]) * ext
            got = box[nb][:, 0]
            diff = (want - got) / span
            assert np.allclose(diff - np.rint(diff), 0.0, atol=1e-9), (name, r["IGRID"], d + 1, nb)
            assert np.allclose(box[nb][:, 1] - box[nb][:, 0], ext, rtol=1e-12)
            assert lvl[nb] == r["LEVEL"]
            nchecked += 1
    assert nchecked > 0 or len(recs) == 1


@pytest.mark.parametrize("name", SUPPORTED_TOPOLOGIES)
@pytest.mark.parametrize("layers,corners", [(1, False), (2, True)])
def test_exchange_on_reference_topology(oracle_mod, name, layers, corners):
    """Exchange of lattice ids per level on the reference's grids: every
    ghost the connect fills holds the id of the same-level cell at that place
    (wrapped periodically), and every ghost of a CON face with a neighbour is
    filled (first `layers` layers, face-tangential interior range)."""
    g = reference_topology(name)
    o = oracle_mod.Oracle(g)
    a, expect = lattice_ids(g)
    for lvl in range(g.minlevel, g.maxlevel + 1):
        o.exchange([a], level=lvl, layers=layers, corners=corners)
    filled = a != 0
    bad = filled & (a != expect)
    assert not bad.any(), f"{name}: {int(bad.sum())} ghosts hold another cell's id"
    for b in range(1, g.ngrid + 1):
        v = g.block_view(a, b)
        for f in range(6):
            if g.bc[b - 1, f] not in (7, 19) or g.nbr[b - 1, f] <= 0:
                continue
            ax = f // 2
            # a sender one cell thick along the axis sends its own ghost as
            # the second layer, which the lattice starts empty
            nsend = g.block_view(a, int(g.nbr[b - 1, f])).shape[ax] - 4
            for layer in range(min(layers, nsend)):
                sl = [slice(2, -2)] * 3
                sl[ax] = 1 - layer if f % 2 == 0 else v.shape[ax] - 2 + layer
                assert (v[tuple(sl)] != 0).all(), (name, b, f + 1, layer)


def _block_ok(ni, nj, nk):
    """The block sizes mgp_init takes (before its LDS check against the
    device): interior counts 1..64, at most 65536 cells (16-bit cell
    positions), (ni+2)(nj+2) <= 4096 (12-bit LDS positions) and a face
    region of at most 16384 values (14-bit offsets)."""
    return (min(ni, nj, nk) >= 1 and max(ni, nj, nk) <= 64 and ni * nj * nk <= 65536 and
            (ni + 2) * (nj + 2) <= 4096 and 2 * (nk * nj + nk * ni + nj * ni) <= 16384)


def test_unsupported_reference_topologies_are_named():
    """The cases outside SUPPORTED_TOPOLOGIES have a block the plane kernels
    do not take: one interior cell along an axis (the 2-D cases)."""
    for name, case in golden_topologies().items():
        dims = [(r["II"] - 4, r["JJ"] - 4, r["KK"] - 4) for r in case["grids"]]
        ok = all(_block_ok(*d) for d in dims)
        assert ok == (name in SUPPORTED_TOPOLOGIES), name


def test_facelist_offsets():
    assert len(FACELIST) == 26 and len(set(DIR_OFFSET)) == 26
