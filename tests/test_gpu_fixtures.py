"""The product path (`kpm()` facade -> C++ host core -> HIP kernels) against the reference's own
regression fixtures, with the reference test's configurations and tolerances
(tests/test_kpm.py:25-47,146-197): LDOS and DOS rtol 1e-3 / atol 1e-6, conductivity rtol 1e-2 /
atol 1e-5.  The models come from the model builder; scaling from the engine's GPU Lanczos."""
import numpy as np
import pytest

import pybinding_amd as pb
from pybinding_amd.model import Model
from fixture_models import golden, ldos_models, dos_models, cond_models

pytestmark = pytest.mark.gpu

LDOS_CONFIGS = [
    {"matrix_format": "CSR", "optimal_size": False, "interleaved": False},
    {"matrix_format": "CSR", "optimal_size": True, "interleaved": False},
    {"matrix_format": "CSR", "optimal_size": False, "interleaved": True},
    {"matrix_format": "ELL", "optimal_size": True, "interleaved": True},
]
DOS_CONFIGS = [
    {"matrix_format": "ELL", "optimal_size": False, "interleaved": False},
    {"matrix_format": "ELL", "optimal_size": True, "interleaved": True},
]


@pytest.mark.parametrize("name", list(ldos_models()))
def test_ldos(name):
    model = Model(*ldos_models()[name])
    expected = golden("ldos[%s]" % name)
    energy = np.linspace(0, 2, 25)
    for c in LDOS_CONFIGS:
        kpm = pb.kpm(model, kernel=pb.lorentz_kernel(), silent=True, **c)
        result = kpm.calc_ldos(energy, broadening=0.15, position=[0, 0.07], reduce=False)
        np.testing.assert_allclose(result.data, expected, rtol=1e-3, atol=1e-6, err_msg=str(c))


@pytest.mark.parametrize("name", list(dos_models()))
def test_dos(name):
    model = Model(*dos_models()[name])
    expected = golden("dos[%s]" % name)
    energy = np.linspace(0, 2, 25)
    for c in DOS_CONFIGS:
        kpm = pb.kpm(model, kernel=pb.lorentz_kernel(), silent=True, **c)
        result = kpm.calc_dos(energy, broadening=0.15)
        np.testing.assert_allclose(result.data, expected, rtol=1e-3, atol=1e-6, err_msg=str(c))


@pytest.mark.parametrize("name", list(cond_models()))
def test_conductivity(name):
    model = Model(*cond_models()[name])
    expected = golden("conductivity[%s]" % name)
    energy = np.linspace(-2, 2, 25)
    for c in DOS_CONFIGS:
        kpm = pb.kpm(model, energy_range=[-9, 9], kernel=pb.lorentz_kernel(), silent=True, **c)
        result = kpm.calc_conductivity(energy, broadening=0.5, temperature=0, num_points=200)
        np.testing.assert_allclose(result.data, expected, rtol=1e-2, atol=1e-5, err_msg=str(c))
