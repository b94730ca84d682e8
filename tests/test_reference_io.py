"""The VTK and globals I/O (sphdem_vtk) against files the reference itself
wrote, kept under tests/golden/reference_io/ as data fixtures:
  singleDEM3D_initDataRestart0000000_0.vtu   (2,501 particles, a restart piece)
  singleDEM3D_initDataRestart0000000.pvtu    (its parallel index)
  singleDEMdrop3d_initDataRestart0000000_0.vtu (7,594 particles)
  singleDEM3D_initDataGlobals0.dat, singleDEMdrop3d_initDataGlobals0.dat,
  singleDEM3D_resultsGlobals0.dat            (globals files, ioGlobals.cpp)
(copies of the reference's singleDEM3D/ and singleDEMdrop3d/ set-up
outputs; singleDEM/ holds byte-identical copies of the singleDEM3D ones).

VTK wrote each DataArray's RangeMin/RangeMax attribute from the values it
encoded (for a vector array the range of the tuple norms), so decoding is
checked against numbers the reference's own writer produced.  CPU only."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle
import sphdem as sd
import sphdem_vtk as vtk

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_io")
VTUS = ["singleDEM3D_initDataRestart0000000_0.vtu", "singleDEMdrop3d_initDataRestart0000000_0.vtu"]


def ranges(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {el.get("Name"): (float(el.get("RangeMin")), float(el.get("RangeMax")))
           for el in piece.findall("PointData/DataArray")}
    p = piece.find("Points/DataArray")
    out["<points>"] = (float(p.get("RangeMin")), float(p.get("RangeMax")))
    return int(piece.get("NumberOfPoints")), out


def close(a, b):
    # VTK prints the ranges with 11 significant digits
    return abs(a - b) <= 1e-10 * max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("name", VTUS)
def test_reference_restart_pieces_decode_to_their_recorded_ranges(name):
    path = os.path.join(HERE, name)
    n, rng = ranges(path)
    pts, data = vtk.read_vtu(path)
    assert pts.shape == (n, 3) and pts.dtype == np.float32
    norms = np.sqrt((pts.astype(np.float64) ** 2).sum(1))
    lo, hi = rng["<points>"]
    assert abs(norms.min() - lo) <= 1e-6 * max(hi, 1e-30)
    assert abs(norms.max() - hi) <= 1e-6 * max(hi, 1e-30)
    assert set(data) == set(rng) - {"<points>"}
    for k, a in data.items():
        assert len(a) == n, k
        v = np.sqrt((a.astype(np.float64) ** 2).sum(1)) if a.ndim == 2 else a.astype(np.float64)
        assert close(v.min(), rng[k][0]) and close(v.max(), rng[k][1]), (k, v.min(), v.max(),
                                                                          rng[k])
    # tags are the particles' numbers 1..n; types are the reference's iamTypes
    assert np.array_equal(np.sort(data["tag"]), np.arange(1, n + 1))
    assert set(np.unique(data["iam"])) <= {sd.SPH, sd.SPH_BOUNDARY, sd.BOUNDARY, sd.DEM}
    assert (data["iam"] == sd.DEM).any()


def test_reference_index_file_lists_its_piece_and_arrays():
    pieces, layout = vtk.read_pvtu(os.path.join(HERE, "singleDEM3D_initDataRestart0000000.pvtu"))
    assert pieces == ["initDataRestart0000000_0.vtu"]
    assert layout == vtk.point_data_layout(os.path.join(HERE, VTUS[0]))


def test_restart_written_here_has_the_reference_layout(tmp_path):
    """writeRestart's arrays (io_data_vtk.cpp:283-300) in its order, types and
    component counts, as in the reference's file (which predates norm3 and
    still carries the since-removed alpha array), with iam, v and norm1
    active."""
    ref = vtk.point_data_layout(os.path.join(HERE, VTUS[1]))
    c = sd.sediment_case(nf=6, n_dem=4)
    o = oracle.Oracle(c)
    out = str(tmp_path / "caseRestart0000000_0.vtu")
    vtk.write_restart(o, out)
    mine = vtk.point_data_layout(out)
    assert [x[0] for x in mine] == [x[0] for x in ref if x[0] != "alpha"] + ["norm3"]
    common = {x[0]: x for x in ref}
    for x in mine:
        if x[0] in common:
            assert x == common[x[0]], x
    pd = ET.parse(out).getroot().find("UnstructuredGrid/Piece/PointData")
    assert (pd.get("Scalars"), pd.get("Vectors"), pd.get("Normals")) == ("iam", "v", "norm1")
    # and it reads back: positions exact (Float64 points), the carried fields exact
    pts, d = vtk.read_vtu(out)
    assert np.array_equal(pts, o.get("r"), equal_nan=True)
    for k in ("dens", "h", "mass", "v", "vhat", "iam", "tag"):
        assert np.array_equal(d[k], o.get(k), equal_nan=d[k].dtype.kind == "f"), k
    # the index file of a one-rank set names the piece and the same arrays
    idx = vtk.write_restart_set(o, str(tmp_path / "case"), 0)
    pieces, layout = vtk.read_pvtu(idx)
    assert pieces == ["caseRestart0000000_0.vtu"]
    assert layout == mine


def test_reference_restart_becomes_a_case():
    """readRestart (io_data_vtk.cpp:20-115) restores r, v, vhat, dens, h, mass,
    iam and tag; the case holds the particles in tag order."""
    path = os.path.join(HERE, VTUS[1])
    pts, d = vtk.read_vtu(path)
    P = sd.sediment_case(nf=6, n_dem=1).params
    c = vtk.case_from_restart(path, P, sd.HOOKS_SEDIMENT)
    order = np.argsort(d["tag"])
    assert c.n == len(pts)
    assert np.array_equal(c.r, pts[order].astype(np.float64))
    for k in ("v", "vhat", "dens", "h", "mass"):
        assert np.array_equal(getattr(c, k), d[k][order]), k
    assert np.array_equal(c.iam, d["iam"][order])


def test_reference_globals_files_parse_with_our_header():
    for name in ("singleDEM3D_initDataGlobals0.dat", "singleDEMdrop3d_initDataGlobals0.dat"):
        path = os.path.join(HERE, name)
        with open(path) as f:
            header = f.readline().rstrip("\n")
        # ioGlobals.cpp:17: our writer's header is the reference's, column for column
        assert header == vtk.GLOBALS_HEADER
        tab = vtk.read_globals(path)
        assert tab.shape[0] == 1
        # the named columns, then custom[] values (the set-up wrote 0 steps)
        ncol = len(vtk.GLOBALS_HEADER.split()) - 1
        assert tab.shape[1] >= ncol
        assert tab[0, 0] == 0 and tab[0, 1] == 0 and tab[0, 2] == 0
    empty = os.path.join(HERE, "singleDEM3D_resultsGlobals0.dat")
    with open(empty) as f:
        assert f.readline().rstrip("\n") == vtk.GLOBALS_HEADER
    assert vtk.read_globals(empty).size == 0
