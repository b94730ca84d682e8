"""VTK XML UnstructuredGrid (.vtu) particle files: the output and restart
format of the reference (io_data_vtk.cpp:20-115 reads a restart, :231-411
writes output; both through vtkXMLUnstructuredGrid{Reader,Writer} in binary
data mode).  VTK itself is not available here, so this is a self-contained
writer and reader of the file format: points, one vertex cell per particle and
named point-data arrays, each array inline, base64-encoded and zlib-compressed
in 32 KiB blocks (the XML format's "binary" mode with vtkZLibDataCompressor and
UInt32 block headers).  The reader also accepts uncompressed binary and ascii
arrays, as ParaView may write them.

Two layouts follow the reference's writers: the restart file of
writeRestart (io_data_vtk.cpp:231-411: FP64 dens h eBForce eViscF eViscB eFF u
mass colour, Int32 iam tag, FP64 3-vectors v vhat vort norm1 norm2 norm3, with
iam, v and norm1 the active scalars, vectors and normals, one piece per rank
under a .pvtu index) and the output file of writeOutput (dens, h, u, mass, iam,
tag, v, vhat, press, porosity, shepSum, fdrag, fb, fp, fv).  Positions are
written as Float64 (the reference's vtkPoints default to Float32, which rounds
a restart's positions).
"""
from __future__ import annotations

import base64
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np

BLOCK = 32768
_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
              np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64",
              np.dtype(np.uint8): "UInt8"}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


def _encode(a: np.ndarray, compress: bool) -> str:
    raw = np.ascontiguousarray(a).tobytes()
    if not compress:
        return base64.b64encode(struct.pack("<I", len(raw)) + raw).decode("ascii")
    blocks = [raw[i:i + BLOCK] for i in range(0, len(raw), BLOCK)] or [b""]
    comp = [zlib.compress(b) for b in blocks]
    last = len(blocks[-1]) if raw else 0
    header = struct.pack(f"<{3 + len(comp)}I", len(comp), BLOCK, last, *[len(c) for c in comp])
    return (base64.b64encode(header) + base64.b64encode(b"".join(comp))).decode("ascii")


def _array_xml(name, a, ncomp, compress, indent):
    t = _VTK_TYPES[np.dtype(a.dtype)]
    attrs = f'type="{t}" Name="{name}" format="binary"'
    if ncomp > 1:
        attrs += f' NumberOfComponents="{ncomp}"'
    return f'{indent}<DataArray {attrs}>\n{indent}  {_encode(a, compress)}\n{indent}</DataArray>\n'


def write_vtu(path: str, points: np.ndarray, point_data: dict, compress: bool = True,
              active: dict | None = None):
    """points: (n, 3) float64; point_data: name -> (n,) or (n, k) array;
    active: PointData attributes naming the active arrays (Scalars, Vectors,
    Normals)."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    n = len(pts)
    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    attrs = "".join(f' {k}="{v}"' for k, v in (active or {}).items())
    out = ['<?xml version="1.0"?>\n',
           f'<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian"'
           f' header_type="UInt32"{comp}>\n',
           "  <UnstructuredGrid>\n",
           f'    <Piece NumberOfPoints="{n}" NumberOfCells="{n}">\n',
           f"      <PointData{attrs}>\n"]
    for name, a in point_data.items():
        a = np.asarray(a)
        if a.dtype == np.float32 or a.dtype == np.float64:
            pass
        elif np.issubdtype(a.dtype, np.integer):
            a = a.astype(np.int32)
        else:
            raise TypeError(f"{name}: unsupported dtype {a.dtype}")
        ncomp = a.shape[1] if a.ndim == 2 else 1
        out.append(_array_xml(name, a, ncomp, compress, "        "))
    out.append("      </PointData>\n      <Points>\n")
    out.append(_array_xml("Points", pts, 3, compress, "        "))
    out.append("      </Points>\n      <Cells>\n")
    idx = np.arange(n, dtype=np.int32)
    out.append(_array_xml("connectivity", idx, 1, compress, "        "))
    out.append(_array_xml("offsets", idx + 1, 1, compress, "        "))
    out.append(_array_xml("types", np.ones(n, dtype=np.uint8), 1, compress, "        "))
    out.append("      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n")
    with open(path, "w") as f:
        f.write("".join(out))


def _b64len(nbytes: int) -> int:
    return 4 * ((nbytes + 2) // 3)


def _decode(text: str, dtype, compressed: bool, header_fmt: str) -> np.ndarray:
    text = "".join(text.split())
    hsize = struct.calcsize("<" + header_fmt)
    if not compressed:
        raw = base64.b64decode(text)
        nbytes = struct.unpack("<" + header_fmt, raw[:hsize])[0]
        return np.frombuffer(raw[hsize:hsize + nbytes], dtype=dtype).copy()
    first = base64.b64decode(text[:_b64len(hsize * 3)])
    nblocks = struct.unpack("<" + header_fmt, first[:hsize])[0]
    hlen = _b64len(hsize * (3 + nblocks))
    header = struct.unpack(f"<{3 + nblocks}{header_fmt}", base64.b64decode(text[:hlen]))
    sizes = header[3:]
    data = base64.b64decode(text[hlen:])
    parts, pos = [], 0
    for s in sizes:
        parts.append(zlib.decompress(data[pos:pos + s]))
        pos += s
    return np.frombuffer(b"".join(parts), dtype=dtype).copy()


def point_data_layout(path: str) -> list:
    """(name, VTK type, components) of every PointData array of a .vtu file,
    in file order, without decoding them."""
    root = ET.parse(path).getroot()
    return [(el.get("Name"), el.get("type"), int(el.get("NumberOfComponents", "1")))
            for el in root.findall("UnstructuredGrid/Piece/PointData/DataArray")]


def read_vtu(path: str):
    """Returns (points (n, 3), {name: array}) of the first piece."""
    root = ET.parse(path).getroot()
    if root.get("type") != "UnstructuredGrid":
        raise ValueError(f"{path}: not an UnstructuredGrid file")
    if root.get("byte_order", "LittleEndian") != "LittleEndian":
        raise ValueError(f"{path}: big-endian files are not supported")
    compressed = root.get("compressor") is not None
    header_fmt = "Q" if root.get("header_type") == "UInt64" else "I"
    piece = root.find("UnstructuredGrid/Piece")
    n = int(piece.get("NumberOfPoints"))

    def arr(el):
        dtype = _NP_TYPES[el.get("type")]
        ncomp = int(el.get("NumberOfComponents", "1"))
        fmt = el.get("format", "ascii")
        if fmt == "ascii":
            a = np.array(el.text.split(), dtype=dtype)
        elif fmt == "binary":
            a = _decode(el.text or "", dtype, compressed, header_fmt)
        else:
            raise ValueError(f"{path}: appended data is not supported")
        return a.reshape(n, ncomp) if ncomp > 1 else a

    points = arr(piece.find("Points/DataArray"))
    data = {el.get("Name"): arr(el) for el in piece.findall("PointData/DataArray")}
    return points, data


OUTPUT_FIELDS = ("dens", "h", "u", "mass", "v", "vhat", "press", "porosity", "shepSum", "fdrag",
                 "fb", "fp", "fv")


def write_particles(sim, path: str, compress: bool = True):
    """The reference's output/restart file of a live simulation (particles in
    upload order; tag = index + 1)."""
    data = {k: sim.get(k) for k in OUTPUT_FIELDS}
    data["iam"] = sim.get("iam").astype(np.int32)
    data["tag"] = sim.get("tag").astype(np.int32)
    write_vtu(path, sim.get("r"), data, compress)


# writeRestart's arrays in its order (io_data_vtk.cpp:283-300, 3D), with the
# Cparticle field each holds; fields the library does not carry (colour,
# vorticity, the boundary normals, the per-particle energies other than the
# viscous one) are written as zeros, as for particles that never set them
RESTART_ARRAYS = (("dens", 1, np.float64), ("h", 1, np.float64), ("eBForce", 1, np.float64),
                  ("eViscF", 1, np.float64), ("eViscB", 1, np.float64), ("eFF", 1, np.float64),
                  ("u", 1, np.float64), ("mass", 1, np.float64), ("colour", 1, np.float64),
                  ("iam", 1, np.int32), ("tag", 1, np.int32), ("v", 3, np.float64),
                  ("vhat", 3, np.float64), ("vort", 3, np.float64), ("norm1", 3, np.float64),
                  ("norm2", 3, np.float64), ("norm3", 3, np.float64))
RESTART_ACTIVE = {"Scalars": "iam", "Vectors": "v", "Normals": "norm1"}
_CARRIED = {"dens", "h", "u", "mass", "iam", "tag", "v", "vhat"}


def restart_arrays(get, n: int) -> dict:
    """The restart file's point data from `get(name)` (a live Sim's or any
    object with the library's field names)."""
    out = {}
    for name, k, dt in RESTART_ARRAYS:
        if name in _CARRIED:
            out[name] = np.asarray(get(name)).astype(dt)
        else:
            out[name] = np.zeros((n, k) if k > 1 else n, dtype=dt)
    return out


def write_restart(src, path: str, compress: bool = True):
    """The reference's restart file (writeRestart) of a live simulation or the
    oracle: the .vtu of one piece, particles in upload order."""
    n = len(src.get("iam"))
    write_vtu(path, src.get("r"), restart_arrays(src.get, n), compress, RESTART_ACTIVE)


def write_pvtu(path: str, pieces: list, layout: list, active: dict | None = None):
    """The parallel index file vtkXMLPUnstructuredGridWriter writes beside the
    pieces: the arrays' names, types and components, and one Piece per rank."""
    attrs = "".join(f' {k}="{v}"' for k, v in (active or {}).items())
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="PUnstructuredGrid" version="0.1" byte_order="LittleEndian"'
           ' compressor="vtkZLibDataCompressor">\n',
           '  <PUnstructuredGrid GhostLevel="0">\n', f"    <PPointData{attrs}>\n"]
    for name, vtype, k in layout:
        nc = f' NumberOfComponents="{k}"' if k > 1 else ""
        out.append(f'      <PDataArray type="{vtype}" Name="{name}"{nc}/>\n')
    out.append("    </PPointData>\n    <PPoints>\n"
               '      <PDataArray type="Float64" NumberOfComponents="3"/>\n    </PPoints>\n')
    for p in pieces:
        out.append(f'    <Piece Source="{p}"/>\n')
    out.append("  </PUnstructuredGrid>\n</VTKFile>\n")
    with open(path, "w") as f:
        f.write("".join(out))


def read_pvtu(path: str):
    """(piece file names, [(name, VTK type, components)]) of a .pvtu index."""
    root = ET.parse(path).getroot()
    if root.get("type") != "PUnstructuredGrid":
        raise ValueError(f"{path}: not a PUnstructuredGrid file")
    g = root.find("PUnstructuredGrid")
    layout = [(el.get("Name"), el.get("type"), int(el.get("NumberOfComponents", "1")))
              for el in g.findall("PPointData/PDataArray")]
    return [p.get("Source") for p in g.findall("Piece")], layout


def write_restart_set(src, base: str, timestep: int, compress: bool = True):
    """A one-rank restart as the reference names it: base + "Restart" +
    %07d + ".pvtu" indexing base + "Restart" + %07d + "_0.vtu"
    (io_data_vtk.cpp:26-28, 381)."""
    import os

    stem = f"{base}Restart{timestep:07d}"
    write_restart(src, stem + "_0.vtu", compress)
    layout = [(nm, "Int32" if dt == np.int32 else "Float64", k) for nm, k, dt in RESTART_ARRAYS]
    write_pvtu(stem + ".pvtu", [os.path.basename(stem) + "_0.vtu"], layout, RESTART_ACTIVE)
    return stem + ".pvtu"


def case_from_restart(path: str, params, hookset: int, name: str = "restart"):
    """A Case from a restart file: what the reference's readRestart
    (io_data_vtk.cpp:20-115) restores -- position, v, vhat, dens, h, mass, iam,
    tag -- with the particles ordered by tag; particles the library deleted
    (iam 255) are dropped, as the reference container no longer holds them."""
    import sphdem as sd

    pts, d = read_vtu(path)
    order = np.argsort(d["tag"], kind="stable") if "tag" in d else np.arange(len(pts))
    keep = order[d["iam"][order] != sd.DELETED]
    return sd.Case(name, params, pts[keep].copy(), d["v"][keep].copy(), d["vhat"][keep].copy(),
                   d["mass"][keep].copy(), d["h"][keep].copy(), d["dens"][keep].copy(),
                   d["iam"][keep].astype(np.int32), hookset, {"restart": path})


# ---- the globals file: header of ioGlobals.cpp:17, lines of myWriteLine (ioGlobals.cpp:42-50)
GLOBALS_HEADER = ("# OutStep SPHStep Time linMomX linMomY angMom KE eBForce eElast eElastExact "
                  "eViscF eViscB eFF eTotal wtTotalStep wtTotalForceCalc wtTotalDddtCalc "
                  "wtTotalReset wtTotalUpdateDomain wtTotalMPI maxV aveDens varDens "
                  "aveDensFromMass n nSph maxF maxFF rmsFF edElastdt edViscFdt edFFdt")


def globals_line(outstep: int, g, o: dict, wall_times=(0.0,) * 6) -> str:
    """The columns CioGlobals::myWriteLine writes (the wall-clock timers are
    the caller's), followed by custom[]."""
    cols = [outstep, g.sphStep, g.time, o["linMom"][0], o["linMom"][1], o["angMom"], o["eKE"],
            o["eBForce"], o["eElast"], o["eElastExact"], o["eViscF"], o["eViscB"], o["eFF"],
            o["eTotal"], *wall_times, o["maxV"], o["aveDens"], o["varDens"],
            o["aveDensFromMass"], o["n"], o["nSph"], o["maxF"], o["maxFF"], o["rmsFF"],
            o["edElastdt"], o["edViscFdt"], o["edFFdt"], *list(g.custom)]
    return " ".join(repr(x) if isinstance(x, float) else str(x) for x in cols)


def append_globals(path: str, outstep: int, sim, wall_times=(0.0,) * 6):
    """Appends the output step's line to a Globals file (created with its
    header line if missing), from the live simulation's diagnostics."""
    import os

    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write(GLOBALS_HEADER + "\n")
        f.write(globals_line(outstep, sim.globals(), sim.output_vars(), wall_times) + "\n")


def read_globals(path: str) -> np.ndarray:
    """The numeric table of a Globals file (header skipped; a file with only
    the header gives an empty table)."""
    with open(path) as f:
        rows = [ln for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
    if not rows:
        return np.zeros((0, len(GLOBALS_HEADER.split()) - 1))
    return np.loadtxt(rows, ndmin=2)
