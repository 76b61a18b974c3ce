"""The FP64 logarithm of the pointwise kernels (classicaldft_amd/csrc/fast_math.hpp,
flog): its argument reduction, truncated atanh series and split ln 2, restated
in numpy with the same constants and operation order, are within 2 ulp of
numpy's log over the whole binary64 range, at the octave boundaries, near 1
and on subnormals.  The device code is the same algorithm; the GPU parity tests
compare every evaluation that uses it with the oracle (std::log) to 1e-12."""
import numpy as np

LN2_HI = float.fromhex("0x1.62e42fefa3800p-1")
LN2_LO = float.fromhex("0x1.ef35793c76730p-45")


def flog(x):
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)
    lo = m < 0.70710678118654752440
    m = np.where(lo, 2.0 * m, m)
    e = np.where(lo, e - 1, e)
    f = m - 1.0
    d = f + 2.0
    r = 1.0 / d
    q0 = f * r
    s = q0 + (f - q0 * d) * r
    z = s * s
    q = np.full_like(z, 1.0 / 19.0)
    for c in (17, 15, 13, 11, 9, 7, 5, 3):
        q = q * z + 1.0 / c
    s2 = s + s
    lm = s2 * z * q + s2
    de = e.astype(np.float64)
    return de * LN2_HI + (de * LN2_LO + lm)


def test_split_ln2_is_exact_for_every_exponent():
    # the head has at most 42 significant bits, so e * head is exact for |e| <= 1075
    mant = int(LN2_HI * 2**53)
    assert mant % (1 << 11) == 0
    from fractions import Fraction

    ln2 = Fraction("0.693147180559945309417232121458176568075500134360255254120680009")
    assert abs(Fraction(LN2_HI) + Fraction(LN2_LO) - ln2) < Fraction(1, 2**100)


def test_flog_within_two_ulp():
    rng = np.random.default_rng(7)
    x = np.concatenate([
        10.0 ** rng.uniform(-307, 307, 400_000),
        1.0 + rng.uniform(-0.3, 0.5, 400_000),
        1.0 + rng.uniform(-1e-9, 1e-9, 50_000),
        rng.uniform(0.70, 0.72, 50_000),
        rng.uniform(1.40, 1.42, 50_000),
        np.array([np.sqrt(0.5), np.nextafter(np.sqrt(0.5), 0.0), np.sqrt(2.0), 0.5, 2.0, 1e-300,
                  5e-324, 2.2250738585072014e-308, 1.7976931348623157e308]),
    ])
    x = x[x != 1.0]
    ref = np.log(x)
    ulp = np.abs(flog(x) - ref) / np.spacing(np.abs(ref))
    assert ulp.max() <= 2.0, float(ulp.max())
    assert flog(np.array([1.0]))[0] == 0.0
