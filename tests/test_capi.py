"""The C-ABI library loads and exports every entry point the headers under
include/ declare (no GPU calls).  CPU only."""
import ctypes
import os
import re

import pytest

import sphdem as sd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    """Entry points of every header under include/."""
    inc = os.path.join(ROOT, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc))
                   if f.endswith(".h"))
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sphdem_\w+)\s*\(", text, re.M)))


def decl_names(body, types):
    """Member names of a C struct body, in order (handles `double a[3], b[3];`)."""
    out = []
    body = body[body.index("{") + 1:]
    for m in re.finditer(r"^[ \t]*(\w+)[ \t]+([^;]+);", body, re.M):
        if m.group(1) not in types:
            continue
        for d in m.group(2).split(","):
            out.append(re.match(r"\s*(\w+)", d).group(1))
    return out


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "sphdem_create" in syms and "sphdem_run_steps" in syms
    assert len(syms) >= 30


def test_library_exports_every_declared_symbol():
    if not os.path.exists(sd.LIB_PATH):
        pytest.fail("libsphdem.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(sd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_params_struct_layout_matches_header():
    # every field of sphdem_params in the header is mirrored, in order
    text = open(os.path.join(ROOT, "include", "sphdem_c.h")).read()
    body = text[text.index("typedef struct sphdem_params"):text.index("} sphdem_params;")]
    names = decl_names(body, ("uint32_t", "int32_t", "double"))
    mirrored = [f[0] for f in sd.Params._fields_]
    assert mirrored == names
    g = text[text.index("typedef struct sphdem_globals"):text.index("} sphdem_globals;")]
    gnames = decl_names(g, ("int64_t", "double"))
    assert [f[0] for f in sd.Globals._fields_] == gnames


def test_library_refuses_to_run_without_a_gpu():
    """No CPU fallback: creating a context without a gfx950 device fails."""
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    lib = sd.load_library()
    P = sd.sediment_case(nf=12, n_dem=0).params
    h = ctypes.c_void_p()
    rc = lib.sphdem_create(ctypes.byref(P), 0, 100, ctypes.byref(h))
    assert rc != 0


def test_more_than_one_viscosity_variant_is_refused():
    """calcViscForce takes one variant; two flags are an argument error,
    reported before any device work."""
    lib = sd.load_library()
    P = sd.dambreak_case(nf=6).params
    P.flags |= sd.F_VISC_CLEARY  # on top of VISC_MORRIS
    h = ctypes.c_void_p()
    assert lib.sphdem_create(ctypes.byref(P), 0, 100, ctypes.byref(h)) == -1
