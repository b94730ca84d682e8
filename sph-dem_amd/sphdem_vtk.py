"""VTK XML UnstructuredGrid (.vtu) particle files: the output and restart
format of the reference (io_data_vtk.cpp:20-115 reads a restart, :231-411
writes output; both through vtkXMLUnstructuredGrid{Reader,Writer} in binary
data mode).  VTK itself is not available here, so this is a self-contained
writer and reader of the file format: points, one vertex cell per particle and
named point-data arrays, each array inline, base64-encoded and zlib-compressed
in 32 KiB blocks (the XML format's "binary" mode with vtkZLibDataCompressor and
UInt32 block headers).  The reader also accepts uncompressed binary and ascii
arrays, as ParaView may write them.

Array names follow the reference's writer: dens, h, u, mass, iam, tag, v, vhat,
press, porosity, shepSum, fdrag, fb, fp, fv.
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


def write_vtu(path: str, points: np.ndarray, point_data: dict, compress: bool = True):
    """points: (n, 3) float64; point_data: name -> (n,) or (n, k) array."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    n = len(pts)
    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    out = ['<?xml version="1.0"?>\n',
           f'<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian"'
           f' header_type="UInt32"{comp}>\n',
           "  <UnstructuredGrid>\n",
           f'    <Piece NumberOfPoints="{n}" NumberOfCells="{n}">\n',
           "      <PointData>\n"]
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
    """The numeric table of a Globals file (header skipped)."""
    return np.loadtxt(path, comments="#", ndmin=2)
