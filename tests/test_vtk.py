"""VTK XML (.vtu) particle files: the reference's output/restart format
(io_data_vtk.cpp), written and read without VTK."""
import os

import numpy as np
import pytest

import sphdem as sd
import sphdem_vtk as vtk


def test_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(3)
    n = 20000  # several 32 KiB compression blocks per array
    pts = rng.normal(size=(n, 3))
    data = {"dens": rng.uniform(900, 1100, n), "v": rng.normal(size=(n, 3)),
            "iam": rng.integers(0, 12, n).astype(np.int32), "tag": np.arange(1, n + 1)}
    for compress in (True, False):
        p = tmp_path / f"x{int(compress)}.vtu"
        vtk.write_vtu(str(p), pts, data, compress=compress)
        q, d = vtk.read_vtu(str(p))
        assert np.array_equal(q, pts)
        for k, a in data.items():
            assert np.array_equal(d[k], a), k
            assert d[k].dtype == (np.int32 if k in ("iam", "tag") else np.float64)


def test_reads_ascii_files(tmp_path):
    """An ascii-mode file as ParaView writes it."""
    text = """<?xml version="1.0"?>
<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">
  <UnstructuredGrid>
    <Piece NumberOfPoints="2" NumberOfCells="0">
      <PointData>
        <DataArray type="Float64" Name="dens" format="ascii">1000.5 999.25</DataArray>
        <DataArray type="Int32" Name="iam" format="ascii">0 11</DataArray>
        <DataArray type="Float64" Name="v" NumberOfComponents="3" format="ascii">1 2 3 4 5 6</DataArray>
      </PointData>
      <Points>
        <DataArray type="Float64" NumberOfComponents="3" format="ascii">0 0 0 1 1 1</DataArray>
      </Points>
      <Cells>
        <DataArray type="Int32" Name="connectivity" format="ascii"></DataArray>
        <DataArray type="Int32" Name="offsets" format="ascii"></DataArray>
        <DataArray type="UInt8" Name="types" format="ascii"></DataArray>
      </Cells>
    </Piece>
  </UnstructuredGrid>
</VTKFile>
"""
    p = tmp_path / "a.vtu"
    p.write_text(text)
    q, d = vtk.read_vtu(str(p))
    assert q.tolist() == [[0, 0, 0], [1, 1, 1]]
    assert d["dens"].tolist() == [1000.5, 999.25]
    assert d["iam"].tolist() == [0, 11]
    assert d["v"].shape == (2, 3)


def test_files_are_well_formed_xml(tmp_path):
    import xml.etree.ElementTree as ET

    p = tmp_path / "w.vtu"
    vtk.write_vtu(str(p), np.zeros((5, 3)), {"h": np.ones(5)})
    root = ET.parse(str(p)).getroot()
    assert root.get("compressor") == "vtkZLibDataCompressor"
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.get("NumberOfPoints") == "5"
    assert [e.get("Name") for e in piece.findall("Cells/DataArray")] == [
        "connectivity", "offsets", "types"]


@pytest.mark.gpu
def test_output_and_restart_on_the_gpu(tmp_path):
    """Output of a live run read back as a restart: the restored particles are
    the written ones (what the reference's readRestart restores)."""
    c = sd.sediment_case(nf=12, n_dem=60)
    sim = sd.Sim(c)
    sim.run(2)
    p = tmp_path / "Restart2_0.vtu"
    vtk.write_particles(sim, str(p))
    rc = vtk.case_from_restart(str(p), c.params, c.hookset)
    keep = sim.get("iam") != sd.DELETED
    assert np.array_equal(rc.r, sim.get("r")[keep])
    assert np.array_equal(rc.dens, sim.get("dens")[keep])
    s2 = sd.Sim(rc)
    assert np.array_equal(s2.get("r"), rc.r)
    s2.run(1)
    assert np.isfinite(s2.get("dens")).all()


def test_globals_line_columns():
    class G:
        sphStep, time = 7, 0.25
        custom = [float(i) for i in range(9)]

    o = {k: 0.0 for k in ("angMom", "eKE", "eBForce", "eElast", "eElastExact", "eViscF",
                          "eViscB", "eFF", "eTotal", "maxV", "aveDens", "varDens",
                          "aveDensFromMass", "maxF", "maxFF", "rmsFF", "edElastdt",
                          "edViscFdt", "edFFdt")}
    o.update(linMom=[1.5, -2.0, 0.0], n=10, nSph=8, eKE=3.0)
    cols = vtk.globals_line(2, G(), o).split()
    assert len(cols) == len(vtk.GLOBALS_HEADER.split()) - 1 + 9
    assert cols[:4] == ["2", "7", "0.25", "1.5"] and cols[6] == "3.0"
    assert cols[24:26] == ["10", "8"]


@pytest.mark.gpu
def test_globals_file_from_a_run(tmp_path):
    sim = sd.Sim(sd.sediment_case(nf=12, n_dem=60))
    p = tmp_path / "Globals0.dat"
    for k in range(2):
        sim.run(1)
        vtk.append_globals(str(p), k, sim)
    t = vtk.read_globals(str(p))
    assert t.shape[0] == 2 and t[1, 1] == 2
