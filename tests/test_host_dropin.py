"""The C++ host drop-in (sph-dem_amd/host): the reference's CdataLL /
CsphIncompress / CcustomSim surface over the C ABI, with the sediment3D case
written against it (host functors through traverse<F, If>)."""
import os
import subprocess

import numpy as np
import pytest

import sphdem as sd

CASE_DIR = os.path.join(os.path.dirname(sd.LIB_PATH), "host", "cases", "sediment3D")
RUN = os.path.join(CASE_DIR, "run")
DUMP = os.path.join(CASE_DIR, "dump_params")
CHECK = os.path.join(CASE_DIR, "check_traversals")


def host_case():
    """The case the committed customConstants.h was generated from."""
    return sd.sediment_case(nf=12, n_dem=60)


def built(path):
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built; run __graft_entry__.build()")
    return path


def test_committed_constants_match_the_generator():
    c = host_case()
    want = sd.custom_constants_header(
        c.params, "SPH-DEM sedimentation, 12^3 fluid + 60 DEM (the host test case)",
        max_particles=2 * c.n)
    with open(os.path.join(CASE_DIR, "customConstants.h")) as f:
        assert f.read() == want


def test_constants_bridge_reproduces_the_parameters():
    """paramsFromConstants() over the case's customConstants.h gives the same
    sphdem_params, byte for byte, as the Python binding builds."""
    out = subprocess.run([built(DUMP)], capture_output=True, text=True, timeout=60, check=True)
    got = bytes.fromhex(out.stdout.strip())
    assert got == bytes(host_case().params)


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="checks the behaviour without a GPU")
def test_run_without_gpu_exits_like_the_reference(tmp_path):
    """No gfx950 device: a message on stderr and exit(-1), as CdataLL's fatal
    errors do in the reference (dataLL.cpp:200-211); no CPU fallback."""
    ic = tmp_path / "ic.bin"
    sd.write_ic(host_case(), str(ic))
    r = subprocess.run([built(RUN), str(ic), "1", str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 255
    assert "gfx950" in r.stderr


@pytest.mark.gpu
def test_host_run_matches_the_library_run(tmp_path):
    """Three steps of the sediment3D case through the C++ host (host functors
    for the CcustomSim hooks, pulled from and pushed to the GPU around each
    traverse) against the same steps through sphdem_run_steps with the device
    hook set: the hooks do the same arithmetic, so the states agree to the
    last bits (tolerance 1e-12 normwise)."""
    c = host_case()
    ic = tmp_path / "ic.bin"
    out = tmp_path / "state.bin"
    sd.write_ic(c, str(ic))
    r = subprocess.run([built(RUN), str(ic), "3", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    st = sd.read_state(str(out))
    sim = sd.Sim(c)
    sim.run(3)
    g = sim.globals()
    assert st["time"] == g.time and st["dt"] == g.dt
    for name, _ in sd.STATE_FIELDS:
        a, b = st[name], sim.get(name)
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / scale if scale else np.abs(a).max()
        assert err <= 1e-12, f"{name}: {err}"


@pytest.mark.gpu
def test_host_pair_traversals_walk_the_library_sets(tmp_path):
    """neighbours<F>(), neighboursGroup<F, If>(), neighboursGroupCoupling<F, If>()
    and traverse<T, F, If>(T&) of the C++ drop-in with host functors: every pair
    of the library's sets is visited once, periodic partners arrive as shifted
    ghost copies (all within the neighbour criterion), and a write made inside
    a group functor reaches the GPU.  Then reverseSyncParticlesBetweenProcs of
    a scatter, the field-by-field lazy upload, calcOutputVars,
    getParticleFromTag, deletion and insertion (dataLL.h:185-257)."""
    c = host_case()
    ic = tmp_path / "ic.bin"
    sd.write_ic(c, str(ic))
    r = subprocess.run([built(CHECK), str(ic)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = dict(kv.split("=") for kv in r.stdout.split())
    sim = sd.Sim(c)
    sim.reset()
    n_sph, n_dem = sim.pair_count(sd.PAIRS_SPH), sim.pair_count(sd.PAIRS_DEM)
    images = int((sim.pairs(sd.PAIRS_SPH)[:, 2] != 13).sum())
    assert images > 0  # the case is periodic in x and y: some partners are ghosts
    assert int(got["pairs"]) == n_sph + n_dem
    assert float(got["worst"]) <= 1.0 + 1e-12
    grp = int(((c.iam == sd.SPH) | (c.iam == sd.SPH_BOUNDARY)).sum())
    assert int(got["groups"]) == grp
    assert int(got["groupnb"]) == n_sph == int(got["written"])
    assert int(got["coupling"]) == sim.pair_count(sd.PAIRS_DEM2SPH) > 0
    assert float(got["cplworst"]) <= 1.0 + 1e-12
    assert float(got["mass"]) == pytest.approx(float(c.mass.sum()), rel=1e-12)
    # reverseSyncParticlesBetweenProcs carried the ghosts' scatter back: each
    # particle counts its appearances in the lists (not its own list length:
    # lattice pairs at exactly 2h through a periodic image are accepted from
    # one side only when the spacing is not dyadic, as in the reference)
    app = np.bincount(sim.pairs(sd.PAIRS_SPH)[:, 1], minlength=c.n).astype(np.int64)
    assert int(got["scattersum"]) == int(app.sum()) == n_sph
    assert int(got["scatterw"]) == int((app * np.arange(1, c.n + 1)).sum())
    # a traversal that wrote f uploaded f alone
    assert int(got["fielduploads"]) == 1
    # calcOutputVars through CsphIncompress counts the container
    assert int(got["nsph"]) == int((c.iam == sd.SPH).sum())
    assert int(got["outn"]) == c.n
    # getParticleFromTag, markForDeletion + deleteParticles, insertNewParticle
    assert int(got["tagok"]) == 1 and int(got["ndel"]) == 1
    assert int(got["ndev"]) == c.n and int(got["insertok"]) == 1
    sim.close()


REF_CASE = "/root/reference/sediment3D"


@pytest.mark.skipif(not os.path.isdir(REF_CASE),
                    reason="the reference's case directory is present only where it was surveyed")
def test_reference_case_code_builds_unchanged(tmp_path):
    """The reference's sediment3D plug-in (customSim.cpp, customOutput.cpp with
    their headers, customConstants.h and parameters.h, compiled from where they
    lie) and a driver in the shape of run.cpp:15-160 build and link against the
    drop-in (sph-dem_amd/host/sphdem) and libsphdem.so with include paths only:
    CdataLL(particleContainer&, CglobalVars&, bool), CsphIncompress with
    calcOutputVars, CcustomSimBase / CcustomOutputBase, Cio_data_vtk, Nsph and
    the traversal templates the plug-in instantiates."""
    host = os.path.join(os.path.dirname(sd.LIB_PATH), "host", "sphdem")
    driver = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "run_shaped.cpp")
    objs = []
    for src in (os.path.join(REF_CASE, "customSim.cpp"), os.path.join(REF_CASE, "customOutput.cpp"),
                driver):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        r = subprocess.run(["g++", "-std=c++17", "-O1", "-c", f"-I{REF_CASE}", f"-I{host}", src,
                            "-o", obj], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-4000:]
        objs.append(obj)
    exe = str(tmp_path / "run")
    libdir = os.path.dirname(sd.LIB_PATH)
    r = subprocess.run(["g++", *objs, "-o", exe, f"-L{libdir}", "-lsphdem", "-lz",
                        f"-Wl,-rpath,{libdir}"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert os.path.exists(exe)

