"""The sweeps' division z / d = quot_fast(z, d, 1/d) (hs_sweep.hpp), compiled for the
host from the same source (libhs_check.so), against IEEE division.

The FP64 sweeps claim the reference's arithmetic bit for bit, and the
reference divides (`y[i] = (y[i] - acc) / D[i]`, l1_gauss_seidel.jl:1068-1095).
quot_fast forms q0 = z*r, e = fma(-q0, d, z), q = fma(e, r, q0) from the
correctly rounded reciprocal r; by Markstein's theorem (P. Markstein, "Computation
of elementary functions on the IBM RISC System/6000 processor", IBM J. Res. Dev.
34(1), 1990; J.-M. Muller et al., Handbook of Floating-Point Arithmetic, 2nd ed.,
Theorem 4.10) q is the correctly rounded z/d when r is within half an ulp of 1/d,
q0 within one ulp of z/d and no step overflows or underflows.  The sweeps take
this path when every divisor of a partition lies in [2^-400, 2^400]; the claimed
domain is |z/d| in [2^-500, 2^500].  Checked here, bit for bit:
* random operands over that whole exponent range;
* divisors whose significand is all ones, powers of two, 1 + ulp, the guard's
  edges 2^-400 and 2^400;
* quotients at and next to binade boundaries (z = d * 2^k and its neighbours),
  at halfway cases of the rounding, and quotients of exact integers;
and outside the domain the documented differences: an overflowing quotient
gives NaN instead of +-inf, the sign of a zero remainder z = -0 is lost.
"""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "hybridsmoothers.jl_amd", "hybridsmoothers", "libhs_check.so")


@pytest.fixture(scope="module")
def chk():
    if not os.path.exists(LIB):
        import sys
        sys.path.insert(0, os.path.join(ROOT, "hybridsmoothers.jl_amd"))
        import build
        build.build_check()
    L = C.CDLL(LIB)
    f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.hs_check_quot_f64.argtypes = [C.c_int64, f64, f64, f64, i32]
    L.hs_check_quot_f32.argtypes = [C.c_int64, f32, f32, f32, i32]
    return L


def _quot(chk, z, d):
    z = np.ascontiguousarray(z, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    q = np.empty_like(z)
    safe = np.empty(z.size, np.int32)
    chk.hs_check_quot_f64(z.size, z, d, q, safe)
    return q, safe.astype(bool)


def _same_bits(a, b):
    return np.asarray(a).view(np.uint64) == np.asarray(b).view(np.uint64)


def _in_domain(z, d):
    with np.errstate(over="ignore", under="ignore", divide="ignore", invalid="ignore"):
        q = np.abs(z / d)
    return (q >= 2.0 ** -500) & (q <= 2.0 ** 500)


def test_random_operands_over_the_domain(chk):
    rng = np.random.default_rng(0)
    n = 2_000_000
    d = rng.uniform(1, 2, n) * np.exp2(rng.integers(-400, 400, n)) * rng.choice([-1, 1], n)
    z = rng.uniform(1, 2, n) * np.exp2(rng.integers(-880, 880, n)) * rng.choice([-1, 1], n)
    keep = _in_domain(z, d)
    z, d = z[keep], d[keep]
    assert z.size > 500_000
    q, safe = _quot(chk, z, d)
    assert safe.all()
    bad = ~_same_bits(q, z / d)
    assert not bad.any(), (z[bad][:4], d[bad][:4], q[bad][:4], (z / d)[bad][:4])


def _special_divisors():
    ones = np.frombuffer(np.uint64(0x3FFFFFFFFFFFFFFF).tobytes(), np.float64)[0]  # 2 - 2^-52
    out = [1.0, 1.0 + 2.0 ** -52, 1.5, 3.0, 0.1, 2.0 ** -400, 2.0 ** 400, ones]
    for e in (-399, -200, -53, -1, 0, 1, 52, 200, 398):
        out += [ones * 2.0 ** e, 2.0 ** e, (1 + 2.0 ** -52) * 2.0 ** e,
                np.nextafter(2.0 ** e, 0.0)]
    return np.array(out + [-v for v in out])


def test_adversarial_divisors_and_binade_edges(chk):
    rng = np.random.default_rng(1)
    ds = _special_divisors()
    zs = []
    for d in ds:
        for k in (-450, -100, -1, 0, 1, 60, 450):
            edge = d * 2.0 ** k  # z/d exactly a power of two
            zs += [edge, np.nextafter(edge, np.inf), np.nextafter(edge, -np.inf)]
        # quotients 1 - 2^-53 and 1 + 2^-52 (the rounding ties around 1)
        zs += [d * (1 - 2.0 ** -53), d * (1 + 2.0 ** -52), d * (1 + 2.0 ** -53)]
        # exact integer quotients and random significands
        zs += list(d * rng.integers(1, 1 << 20, 8).astype(np.float64))
        zs += list(d * rng.uniform(-8, 8, 16))
    z = np.array(zs)
    d = np.repeat(ds, z.size // ds.size)
    assert z.size == d.size
    keep = _in_domain(z, d)
    q, safe = _quot(chk, z[keep], d[keep])
    assert safe.all()
    bad = ~_same_bits(q, z[keep] / d[keep])
    assert not bad.any(), (z[keep][bad][:4], d[keep][bad][:4])


def test_halfway_and_all_ones_significands(chk):
    """Dividends whose significand is all ones or 1 + ulp over all-ones divisors:
    the products z*r sit next to rounding boundaries, where a missing
    correction step shows."""
    ones = np.frombuffer(np.uint64(0x3FFFFFFFFFFFFFFF).tobytes(), np.float64)[0]
    rng = np.random.default_rng(2)
    base = np.array([ones, 1.0, 1 + 2.0 ** -52, 1.5 - 2.0 ** -52, 1.5, 1.5 + 2.0 ** -52])
    z = np.concatenate([base * 2.0 ** e for e in rng.integers(-300, 300, 400)])
    d = np.concatenate([np.roll(base, k) * 2.0 ** e
                        for k, e in zip(range(400), rng.integers(-300, 300, 400))])
    keep = _in_domain(z, d)
    q, _ = _quot(chk, z[keep], d[keep])
    assert _same_bits(q, z[keep] / d[keep]).all()


def test_outside_the_domain(chk):
    """The documented differences: the sweeps do not claim these operands."""
    q, safe = _quot(chk, np.array([1e300, -1e300]), np.array([1e-100, 1e-100]))
    assert safe.all() and np.isnan(q).all()  # IEEE gives +-inf
    q, _ = _quot(chk, np.array([-0.0, 0.0]), np.array([3.0, 3.0]))
    assert q[0] == 0.0 and not np.signbit(q[0])  # IEEE gives -0
    assert q[1] == 0.0 and not np.signbit(q[1])
    # divisors outside the guard are refused (the wave divides instead)
    _, safe = _quot(chk, np.ones(4), np.array([2.0 ** -401, 2.0 ** 401, 0.0, np.inf]))
    assert not safe.any()


def test_fp32(chk):
    rng = np.random.default_rng(3)
    n = 500_000
    d = (rng.uniform(1, 2, n) * np.exp2(rng.integers(-60, 60, n))).astype(np.float32)
    z = (rng.uniform(-2, 2, n) * np.exp2(rng.integers(-40, 40, n))).astype(np.float32)
    q = np.empty_like(z)
    safe = np.empty(n, np.int32)
    chk.hs_check_quot_f32(n, z, d, q, safe)
    assert safe.all()
    assert (q.view(np.uint32) == (z / d).view(np.uint32)).all()
