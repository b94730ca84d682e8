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


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_reciprocal_division_is_correctly_rounded(hm):
    """divr(x, y, 1/y) (Markstein: one product, an exact FMA remainder, one FMA
    correction) equals the IEEE quotient x/y bit for bit, including divisors
    with all-ones significands and near powers of two, where a plain x*(1/y)
    is off by an ulp in a sizeable fraction of cases."""
    rng = np.random.default_rng(17)
    n = 2_000_000
    x = rng.uniform(1, 2, n) * np.exp2(rng.integers(-60, 60, n)) * rng.choice([-1, 1], n)
    y = rng.uniform(1, 2, n) * np.exp2(rng.integers(-60, 60, n))
    yb = y.view(np.uint64)
    yb[::7] |= np.uint64((1 << 52) - 1)                      # all-ones significand
    yb[1::11] &= ~np.uint64((1 << 52) - 1)                   # powers of two ...
    yb[1::11] |= rng.integers(0, 256, yb[1::11].size).astype(np.uint64)  # ... and just above
    x[::13] = 0.0
    out = np.empty(n)
    hm.sphdem_host_divr_arr(_ptr(x), _ptr(y), _ptr(out), ctypes.c_long(n))
    exact = x / y
    assert np.array_equal(out.view(np.uint64), exact.view(np.uint64))
    xz = np.array([-0.0, 0.0])
    yz = np.array([3.0, 3.0])
    oz = np.empty(2)
    hm.sphdem_host_divr_arr(_ptr(xz), _ptr(yz), _ptr(oz), ctypes.c_long(2))
    assert np.array_equal(oz, [0.0, 0.0])  # -0/y comes out as +0
    naive = x * (1.0 / y)
    assert np.count_nonzero(naive != exact) > n // 100  # the correction step matters


@pytest.mark.parametrize("name", ["wendF", "wendW", "wendPdWh"])
def test_hoisted_wendland_forms_are_bitwise_equal(hm, name):
    """The pair kernels use the Wendland functions with 1/h, 1/h^3, 3/h^4 and
    1/h^5 computed once per particle (and divisions by h as divr): bit for bit
    the functions they replace."""
    rng = np.random.default_rng(23)
    n = 400_000
    q = rng.uniform(0, 2.05, n)
    q[::50] = 2.0
    q[1::50] = 0.0
    h = rng.uniform(1e-3, 2.0, n)
    wcon = 21.0 / (16.0 * 3.14159265)
    a = np.empty(n)
    b = np.empty(n)
    getattr(hm, f"sphdem_host_{name}i_arr")(ctypes.c_double(wcon), _ptr(q), _ptr(h), _ptr(a),
                                            ctypes.c_long(n))
    getattr(hm, f"sphdem_host_{name}_arr")(ctypes.c_double(wcon), _ptr(q), _ptr(h), _ptr(b),
                                           ctypes.c_long(n))
    # bit for bit, except that a zero may come out as +0 where the plain form
    # gives -0 (at q = 2 exactly); signed zeros vanish in the pair sums
    same = (a.view(np.uint64) == b.view(np.uint64)) | ((a == 0) & (b == 0))
    assert same.all(), f"{np.count_nonzero(~same)} differ, e.g. q={q[~same][:3]} h={h[~same][:3]}"
