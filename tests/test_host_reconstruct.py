"""The host KuboBastin reconstruction (factorised form, cppcore/include/kpm/reconstruct.hpp:90-143)
against the oracle's direct restatement, which the reference's conductivity fixtures pin."""
import numpy as np
import pytest


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32, np.complex64])
def test_kubo_bastin_matches_oracle(oracle, dtype):
    from pybinding_amd import _native
    native = _native.load_module()
    rng = np.random.default_rng(11)
    M = 34
    mu = rng.standard_normal((M, M))
    if np.iscomplexobj(np.zeros(1, dtype)):
        mu = mu + 1j * rng.standard_normal((M, M))
    mu = (mu / np.arange(1, M + 1)[:, None] / np.arange(1, M + 1)[None, :]).astype(dtype)
    a, b = oracle.scale_factors(-9, 7)
    samples = np.linspace(-9, 7, 97)
    chem = np.linspace(-2, 2, 9)
    for temperature in (0.0, 30.0):
        ref = oracle.kubo_bastin(mu.astype(np.complex128), chem, samples, temperature, a, b)
        got = native.kubo_bastin(mu, list(chem), list(samples), temperature, (a, b))
        tol = 1e-10 if dtype in (np.float64, np.complex128) else 1e-4
        np.testing.assert_allclose(got, ref, rtol=tol, atol=tol * np.abs(ref).max())
