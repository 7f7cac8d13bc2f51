"""Pins of the oracle's random starters (cppcore/src/kpm/Starter.cpp:48-83).

- the raw std::mt19937 stream: the C++ standard's known answer ([rand.predef]: the 10000th
  invocation of a default-constructed mt19937 yields 4123659995) and numpy's independent MT19937
  implementation seeded the same way (legacy RandomState(5489) uses init_genrand);
- the Rademacher sign rule derived from libstdc++'s generate_canonical: f64 element i is -1 iff
  draw 2i+1 (the high word) is below 2^31; f32 element i is -1 iff float(draw i) < 2^31;
- complex phases lie on the unit circle with angle 2*pi_f*u.
"""
import numpy as np


def test_mt19937_known_answer(oracle):
    s = oracle.mt19937_stream(10000)
    assert int(s[-1]) == 4123659995


def test_mt19937_matches_numpy_stream(oracle):
    s = oracle.mt19937_stream(5000)
    ref = np.random.RandomState(5489).randint(0, 2**32, size=5000, dtype=np.uint64).astype(np.uint32)
    np.testing.assert_array_equal(s, ref)


def test_rademacher_f64_sign_rule(oracle):
    n = 300000
    r = oracle.random_starter(n, 1, np.float64)[0]
    draws = np.random.RandomState(5489).randint(0, 2**32, size=2 * n, dtype=np.uint64)
    expected = np.where(draws[1::2] < 2**31, -1.0, 1.0)
    np.testing.assert_array_equal(r, expected)


def test_rademacher_f32_sign_rule(oracle):
    n = 300000
    r = oracle.random_starter(n, 1, np.float32)[0]
    draws = np.random.RandomState(5489).randint(0, 2**32, size=n, dtype=np.uint64)
    expected = np.where(draws.astype(np.float32) < np.float32(2**31), -1.0, 1.0)
    np.testing.assert_array_equal(r, expected)


def test_columns_continue_one_stream(oracle):
    # block starters draw whole columns one after another from one generator (Starter.hpp:42-50)
    both = oracle.random_starter(1000, 2, np.float64)
    one = oracle.random_starter(2000, 1, np.float64)[0]
    np.testing.assert_array_equal(both.ravel(), one)


def test_complex_phases(oracle):
    n = 20000
    z = oracle.random_starter(n, 1, np.complex128)[0]
    u = oracle.uniform(n, np.float64)
    np.testing.assert_allclose(np.abs(z), 1.0, rtol=0, atol=1e-15)
    theta = float(np.float32(2) * np.float32(np.pi)) * u
    np.testing.assert_allclose(z, np.exp(1j * theta), rtol=0, atol=2e-15)
    assert abs(np.mean(z)) < 0.02  # phases spread uniformly


def test_uniform_is_canonical(oracle):
    # generate_canonical<double,53>: (lo + hi * 2^32) / 2^64 with lo drawn first
    n = 1000
    u = oracle.uniform(n, np.float64)
    d = np.random.RandomState(5489).randint(0, 2**32, size=2 * n, dtype=np.uint64).astype(np.float64)
    np.testing.assert_array_equal(u, (d[0::2] + d[1::2] * 2.0**32) / 2.0**64)


def test_uniform_f32_is_canonical(oracle):
    # generate_canonical<float,24> takes one draw: float(d) / 2^32, clamped below 1
    n = 5000
    u = oracle.uniform(n, np.float32)
    d = np.random.RandomState(5489).randint(0, 2**32, size=n, dtype=np.uint64)
    ref = d.astype(np.float32) / np.float32(2.0**32)
    ref = np.where(ref >= 1, np.nextafter(np.float32(1), np.float32(0)), ref)
    np.testing.assert_array_equal(u, ref)


def test_complex64_phases(oracle):
    # complex<float> starters: exp(i * 2 pi_f * u) with u the float draws (Starter.cpp:72-82)
    n = 20000
    z = oracle.random_starter(n, 2, np.complex64)
    assert z.dtype == np.complex64
    u = oracle.uniform(2 * n, np.float32).reshape(2, n)
    theta = (np.float32(2) * np.float32(np.pi)) * u
    np.testing.assert_allclose(np.abs(z), 1.0, rtol=0, atol=3e-7)
    np.testing.assert_allclose(z, np.exp(1j * theta.astype(np.float64)), rtol=0, atol=4e-7)
    # the two columns continue one stream
    np.testing.assert_array_equal(z.ravel(), oracle.random_starter(2 * n, 1, np.complex64)[0])
