"""The device arithmetic helpers (sph-dem_amd/csrc/sphdem_math.h), built for the
host, against glibc and the oracle.  CPU only."""
import ctypes
import ctypes.util
import os

import numpy as np
import pytest

import oracle
import sphdem as sd

HOSTMATH = os.path.join(os.path.dirname(sd.LIB_PATH), "libsphdem_hostmath.so")


@pytest.fixture(scope="module")
def hm():
    if not os.path.exists(HOSTMATH):
        pytest.fail("libsphdem_hostmath.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(HOSTMATH)
    lib.sphdem_host_pow_third.restype = ctypes.c_double
    lib.sphdem_host_pow_third.argtypes = [ctypes.c_double]
    for nm in ("wendW", "wendF", "wendPdWh"):
        f = getattr(lib, "sphdem_host_" + nm)
        f.restype = ctypes.c_double
        f.argtypes = [ctypes.c_double] * 3
    return lib


@pytest.fixture(scope="module")
def libm():
    m = ctypes.CDLL(ctypes.util.find_library("m"))
    m.pow.restype = ctypes.c_double
    m.pow.argtypes = [ctypes.c_double, ctypes.c_double]
    return m


def test_pow_third_is_correctly_rounded_and_matches_glibc(hm, libm):
    """h = HFAC*pow(m/rho, 1.0/NDIM) (sphIncompress.h:699).  The device's
    pow_third is correctly rounded; glibc's pow has errors up to ~0.504 ulp, so
    the two agree bitwise except where glibc itself rounds the wrong way, and
    there they differ by one ulp.  Exact decimal arithmetic decides."""
    import decimal
    import math

    decimal.getcontext().prec = 60
    rng = np.random.default_rng(7)
    xs = np.concatenate([
        10.0 ** rng.uniform(-12, 2, 100000),
        (1.0 / 64) ** 3 * (1 + rng.uniform(-0.05, 0.05, 50000)),  # m/rho near dx^3
        np.array([2.0 ** -18, 1e-6, 8e-6, 1.0, 8.0, 27.0, 1e-9, 0.001]),
    ])
    third = 1.0 / 3
    Y = decimal.Decimal(third)
    mism = 0
    for x in xs:
        a, b = hm.sphdem_host_pow_third(x), libm.pow(x, third)
        if a == b:
            continue
        mism += 1
        ex = decimal.Decimal(x) ** Y
        ulp = decimal.Decimal(math.ulp(a))
        assert abs(decimal.Decimal(a) - ex) <= ulp / 2          # ours is correctly rounded
        assert abs(decimal.Decimal(b) - ex) > ulp / 2           # glibc's is not, here
        assert abs(a - b) == math.ulp(a)
    # exact-lattice case m/rho = dx^3 = 2^-18: both give 2^-6 + 1 ulp
    assert hm.sphdem_host_pow_third(2.0 ** -18) == libm.pow(2.0 ** -18, third) == 2.0 ** -6 + 2.0 ** -58
    assert mism < 2e-3 * len(xs)
    # and a spot check of correct rounding on agreeing arguments
    for x in xs[:2000]:
        a = hm.sphdem_host_pow_third(x)
        ex = decimal.Decimal(x) ** Y
        assert abs(decimal.Decimal(a) - ex) <= decimal.Decimal(math.ulp(a)) / 2


def test_neighbour_threshold_square(libm):
    """is_neighbr compares with pow(2*max(h_a,h_b), 2) (particle.h:237); the
    device uses the exact product x*x (correctly rounded).  glibc's pow(x, 2)
    rounds the wrong way on near-midpoint squares (< 0.2% of arguments, by one
    ulp), so the neighbour predicates can differ only for a squared distance
    equal to the larger of the two thresholds."""
    import decimal
    import math

    decimal.getcontext().prec = 50
    rng = np.random.default_rng(3)
    xs = 10.0 ** rng.uniform(-6, 1, 100000)
    mism = 0
    for x in xs:
        a, b = x * x, libm.pow(x, 2.0)
        if a != b:
            mism += 1
            ex = decimal.Decimal(x) ** 2
            assert abs(decimal.Decimal(a) - ex) <= decimal.Decimal(math.ulp(a)) / 2
            assert abs(a - b) == math.ulp(a)
    assert mism < 2e-3 * len(xs)


def test_device_wendland_matches_oracle(hm):
    """Device W, F, pdWh (sph.h:74-81, 199-205, 285-312) against the oracle's
    glibc-pow versions: bitwise equal for nearly all arguments; the rest differ
    by rounding amplified only by the formula's own cancellation near q = 0."""
    c = sd.dambreak_case(nf=12)
    o = oracle.Oracle(c)
    wcon = c.params.wcon
    rng = np.random.default_rng(5)
    same = total = 0
    for q, h in zip(rng.uniform(0, 2.1, 3000), rng.uniform(0.001, 2.0, 3000)):
        for dev, ref in ((hm.sphdem_host_wendW, o.W), (hm.sphdem_host_wendF, o.F),
                         (hm.sphdem_host_wendPdWh, o.pdWh)):
            a, b = dev(wcon, q, h), ref(q, h)
            total += 1
            same += a == b
            tol = 1e-14 if q > 0.05 else 1e-12
            assert a == pytest.approx(b, rel=tol, abs=1e-300)
    assert same > 0.97 * total
