"""The output's number formatter (FormatG6, csrc/results.cpp) against printf's
"%g" and std::to_chars(general, 6), which is how the reference's
std::ostream writes the identity, E-value and bit score
(src/aligner_common.h:372-389).  The formatter takes a fast path that leaves
values near a rounding tie to the library, so the cases below cover exact
ties (dyadic values and half integers), powers of ten and their neighbours,
every binary exponent (random bit patterns), and the output's own kinds of
values.  CPU only (the product's host code in oracle/libhostpipe.so)."""
import ctypes

import numpy as np
import pytest

import oracle_py

N = 400_000


def mismatches(values):
    v = np.ascontiguousarray(values, dtype=np.float64)
    out = np.zeros(v.size * 32, dtype=np.uint8)
    first = ctypes.c_uint64()
    bad = oracle_py.hostpipe().gzh_format_g6_check(v.ctypes.data, v.size, out.ctypes.data,
                                                   ctypes.byref(first))
    texts = [bytes(out[i * 32:(i + 1) * 32]).split(b"\0", 1)[0].decode() for i in range(min(v.size, 2000))]
    return bad, (v[first.value] if bad else None), texts


def cases(rng):
    bits = rng.integers(0, 2 ** 63, N, dtype=np.uint64)
    bits |= rng.integers(0, 2, N, dtype=np.uint64) << np.uint64(63)
    rand_bits = bits.view(np.float64)
    p10 = 10.0 ** np.arange(-307, 309)
    with np.errstate(over="ignore"):  # 9.99... * 1e308 is inf, which is a case too
        near = np.concatenate([np.nextafter(p10, 0), p10, np.nextafter(p10, np.inf),
                               p10 * 0.9999995, p10 * 9.999995, p10 * 9.9999950001])
    length = rng.integers(1, 3000, N)
    diff = rng.integers(0, 3000, N) % length
    score = rng.integers(-50, 20000, N)
    lam, K = np.float32(0.267), np.float32(0.041)
    return {
        "random bit patterns": rand_bits[np.isfinite(rand_bits)],
        "dyadic (ties at six digits)": rng.integers(1, 2 ** 24, N) / 2.0 ** rng.integers(0, 40, N),
        "half integers": np.concatenate([rng.integers(1, 10 ** 9, N) + 0.5,
                                         rng.integers(1, 10 ** 8, N) * 10.0 + 5]),
        "powers of ten": np.concatenate([near, -near]),
        "identity": (1.0 - (diff.astype(np.float32) / length.astype(np.float32)).astype(np.float64)) * 100,
        "bit score": ((score.astype(np.float32) * lam - np.float32(np.log(K))) /
                      np.float32(np.log(2.0))).astype(np.float64),
        "E-value": rng.integers(1, 2 ** 40, N) * float(K) * np.exp(-score * float(lam)),
        "special": np.array([0.0, -0.0, 1e-300, 9.99999e-301, 5e-324, 2.2250738585072014e-308,
                             1.7976931348623157e308, 1e300, 9.999995e299, np.inf, -np.inf, np.nan]),
    }


@pytest.mark.parametrize("seed", [1, 2])
def test_format_matches_printf_g(seed):
    for name, values in cases(np.random.default_rng(seed)).items():
        bad, example, _ = mismatches(values)
        assert bad == 0, (name, example)


def test_format_known_texts():
    values = [100.0, 93.75, 79 / 64, 123456.5, 1234565.0, 0.0001, 1e-5, 4.2e-27, 3.5e-300,
              1e300 * 1.5, -0.0, 45.8]
    _, _, texts = mismatches(values)
    assert texts == ["100", "93.75", "1.23438", "123456", "1.23456e+06", "0.0001", "1e-05",
                     "4.2e-27", "3.5e-300", "1.5e+300", "-0", "45.8"]
