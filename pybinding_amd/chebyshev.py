"""Kernel polynomial method on the MI355X engine -- the `pb.kpm(model, ...)` interface.

Mirrors the reference's Python API (pybinding/chebyshev.py:65-445): `kpm()` returns a `KPM` whose
`moments`, `calc_greens`, `calc_ldos`, `calc_spatial_ldos`, `calc_dos`, `deferred_ldos` and
`calc_conductivity` keep the reference's signatures, return types and error messages.  Every
moment is computed by the HIP engine (``_kpmx`` -> ``libkpmx.so``); there is no CPU path.
"""
import warnings

import numpy as np
import scipy.sparse

from . import results
from ._native import load_module

__all__ = ['KPM', 'kpm', 'kpm_cuda', 'kpm_hip', 'SpatialLDOS',
           'jackson_kernel', 'lorentz_kernel', 'dirichlet_kernel']


class SpatialLDOS:
    """Holds the results of :meth:`KPM.calc_spatial_ldos` (pybinding/chebyshev.py:21-62)"""

    def __init__(self, data, energy, structure):
        self.data = data
        self.energy = energy
        self.structure = structure

    def structure_map(self, energy):
        idx = np.argmin(abs(self.energy - energy))
        return self.structure.with_data(self.data[idx])

    def ldos(self, position, sublattice=""):
        idx = self.structure.find_nearest(position, sublattice)
        return results.Series(self.energy, self.data[:, idx],
                              labels=dict(variable="E (eV)", data="LDOS", columns="orbitals"))


def _as_hamiltonian(model):
    """The Hamiltonian CSR of a Model (or of a bare sparse matrix)."""
    if scipy.sparse.issparse(model):
        return scipy.sparse.csr_matrix(model)
    return scipy.sparse.csr_matrix(model.hamiltonian)


class KPM:
    """Front object of a KPM calculation on the GPU engine (interface of
    pybinding/chebyshev.py:65-317).  Made by :func:`kpm`, not constructed by hand."""

    def __init__(self, impl, model=None):
        from .model import Model
        if isinstance(impl, Model):
            raise TypeError("You're probably looking for `pb.kpm()` (lowercase).")
        self.impl = impl
        self._model = model

    @property
    def model(self):
        """Model this calculation was made from (None for a bare sparse matrix)"""
        return self._model

    @property
    def system(self):
        """Sites of the model (positions, sublattices, orbitals): `model.system`"""
        return self._model.system

    @property
    def scaling_factors(self) -> tuple:
        """(a, b) such that the scaled Hamiltonian is (H - b) / a, with its spectrum inside [-1, 1]"""
        return self.impl.scaling_factors

    @property
    def kernel(self):
        """Kernel whose coefficients damp the moments before reconstruction"""
        return self.impl.kernel

    @property
    def stats(self):
        """Stats of the last calculation (num_moments, nnz, eps, moments_time, ...)"""
        return self.impl.stats

    def report(self, shortform=False):
        """Summary of the most recent moment computation: spectrum bounds, number of moments,
        processed non-zeros per second and time; `shortform` gives it on one line"""
        s = self.impl.stats
        lo, hi = self.impl.energy_range
        if shortform:
            return "{:.2f}, {:.2f} | {} @ {:.3g}eps [{:.3g}s]".format(lo, hi, s.num_moments, s.eps, s.moments_time)
        return ("- Spectrum bounds ({:.2f}, {:.2f} eV)\n- KPM calculated {} moments at {:.4g} non-zero "
                "elements per second | {:.4g}s".format(lo, hi, s.num_moments, s.eps, s.moments_time))

    def __call__(self, *args, **kwargs):
        warnings.warn("Use .calc_greens() instead", DeprecationWarning)
        return self.calc_greens(*args, **kwargs)

    # ------------------------------------------------------------------------------ argument checks
    def _check_vector(self, name, v):
        """KPM::moments argument checks (cppcore/src/KPM.cpp:19-47)"""
        size = self.impl.size
        if v.size != size:
            raise RuntimeError("Size mismatch between the model Hamiltonian and the given "
                               "argument '{}'".format(name))
        if not self.impl.is_complex and np.any(np.imag(v) != 0):
            raise RuntimeError("The model Hamiltonian is real, but the given argument "
                               "'{}' is complex".format(name))

    def moments(self, num_moments, alpha, beta=None, op=None):
        r"""Chebyshev moments :math:`\mu_n = <\beta|op \cdot T_n(H)|\alpha>` of the scaled Hamiltonian,
        damped by the kernel (pybinding/chebyshev.py:117-149).  `beta` defaults to `alpha`, `op`
        to the identity."""
        alpha = np.asarray(alpha).ravel()
        self._check_vector("alpha", alpha)
        if beta is None or len(beta) == 0:
            beta = np.zeros(0, dtype=complex)
        else:
            beta = np.asarray(beta).ravel()
            self._check_vector("beta", beta)
        if op is not None:
            op = scipy.sparse.csr_matrix(op, dtype=complex)
            size = self.impl.size
            if op.shape != (size, size):
                raise RuntimeError("Size mismatch between the model Hamiltonian and the given "
                                   "argument 'operator'")
            if not self.impl.is_complex and np.any(op.data.imag != 0):
                raise RuntimeError("The model Hamiltonian is real, but the given argument "
                                   "'operator' is complex")
            if op.nnz == 0:
                op = None
        return self.impl.moments(num_moments, alpha.astype(complex), beta.astype(complex), op)

    def calc_greens(self, i, j, energy, broadening):
        """Green's function of a single Hamiltonian element, or of a list of columns `j`
        (cppcore/src/KPM.cpp:103-129)."""
        size = self.impl.size
        energy = np.asarray(energy, dtype=float)
        cols = list(np.atleast_1d(j))
        # valid indices are 0 .. size - 1 (the reference's check also lets `size` through, which
        # would index one element past the end of the reorder map)
        if i < 0 or i >= size or any(c < 0 or c >= size for c in cols):
            raise RuntimeError("KPM::calc_greens(i,j): invalid value for i or j.")
        gs = self.impl.calc_greens_vector(int(i), [int(c) for c in cols], energy, float(broadening))
        return gs if np.ndim(j) > 0 else gs[0]

    def calc_ldos(self, energy, broadening, position, sublattice="", reduce=True):
        """Local density of states at the site nearest to `position` (cppcore/src/KPM.cpp:55-65)"""
        energy = np.asarray(energy, dtype=float)
        system = self.system
        site = system.find_nearest(position, sublattice)
        ham_idx = system.to_hamiltonian_indices(site)
        ldos = self.impl.calc_ldos_indices(energy, float(broadening), [int(i) for i in ham_idx])
        if reduce and ldos.shape[1] > 1:
            ldos = ldos.sum(axis=1, keepdims=True)
        return results.Series(energy, ldos.squeeze(), labels=dict(variable="E (eV)", data="LDOS",
                                                                  columns="orbitals"))

    def calc_spatial_ldos(self, energy, broadening, shape, sublattice=""):
        """LDOS as a function of energy and space (cppcore/src/KPM.cpp:67-94)"""
        system = self.system
        if system.is_multiorbital:
            raise RuntimeError("This function doesn't currently support multi-orbital models")
        energy = np.asarray(energy, dtype=float)
        contains = shape.contains(*system.positions)
        idx = np.nonzero(contains)[0]
        if sublattice:
            idx = idx[system.sublattices_str[idx] == sublattice]
        ldos = self.impl.calc_ldos_indices(energy, float(broadening), [int(i) for i in idx])
        smap = system[contains]
        if sublattice:
            smap = smap[smap.sub == sublattice]
        return SpatialLDOS(ldos, energy, smap)

    def calc_dos(self, energy, broadening, num_random=1):
        """Density of states (stochastic trace over `num_random` random vectors)"""
        energy = np.asarray(energy, dtype=float)
        dos = self.impl.calc_dos(energy, float(broadening), int(num_random))
        return results.Series(energy, dos, labels=dict(variable="E (eV)", data="DOS"))

    def deferred_ldos(self, energy, broadening, position, sublattice=""):
        """Same as :meth:`calc_ldos` but evaluated later (reference: a Deferred for parallel sweeps)"""
        return results.Deferred(lambda: self.calc_ldos(energy, broadening, position, sublattice).data)

    def calc_conductivity(self, chemical_potential, broadening, temperature,
                          direction="xx", volume=1.0, num_random=1, num_points=1000):
        """Kubo-Bastin electrical conductivity (cppcore/src/KPM.cpp:131-148)"""
        if len(direction) != 2 or any(c not in "xyz" for c in direction):
            raise RuntimeError("Invalid direction: must be 'xx', 'xy', 'zz', or similar.")
        system = self.system
        pos = system.expanded_positions if system.is_multiorbital else system.positions
        coord = dict(zip("xyz", pos))
        chemical_potential = np.asarray(chemical_potential, dtype=float)
        data = self.impl.calc_conductivity(chemical_potential, float(broadening), float(temperature),
                                           np.asarray(coord[direction[0]], dtype=np.float32),
                                           np.asarray(coord[direction[1]], dtype=np.float32),
                                           int(num_random), int(num_points)).real
        if volume != 1.0:
            data /= volume
        return results.Series(chemical_potential, data,
                              labels=dict(variable=r"$\mu$ (eV)", data=r"$\sigma (e^2/h)$"))

    # ------------------------------------------------------------------------------ engine extras
    def dos_moments(self, num_moments, num_random=1):
        """Undamped mean diagonal moments of `num_random` reference random starters"""
        return self.impl.dos_moments(int(num_moments), int(num_random))


class _ComputeProgressReporter:
    """Prints progress of long moment computations (pybinding/chebyshev.py:320-336)"""

    def __init__(self):
        self.total = 0
        self.done = 0

    def __call__(self, delta, total):
        if total == 1:
            return
        if delta < 0:
            print("Computing KPM moments...")
            self.total, self.done = total, 0
        elif delta == total:
            pass
        else:
            self.done += delta


def _site_order(model, site_order):
    """The GPU site order for whole-system calculations (see pybinding_amd.ordering)."""
    if site_order is None or isinstance(site_order, str) and site_order == "bfs":
        return None
    if not isinstance(site_order, str):
        return np.asarray(site_order, dtype=np.int32)
    if site_order != "auto":
        raise ValueError("site_order must be 'auto', 'bfs' or a permutation")
    system = getattr(model, "system", None)
    if system is None:
        return None
    from .ordering import locality_order
    x, y, _ = system.expanded_positions if system.is_multiorbital else system.positions
    return locality_order(x, y)


def kpm(model, energy_range=None, kernel="default", num_threads="auto", silent=False, **kwargs):
    """The Kernel Polynomial Method on the GPU (pybinding/chebyshev.py:339-378)

    Parameters
    ----------
    model : Model or scipy.sparse matrix
        Model which will provide the Hamiltonian matrix.
    energy_range : Optional[Tuple[float, float]]
        Spectrum bounds; found by the Lanczos procedure (on the GPU) when not given.
    kernel : Kernel
        :func:`jackson_kernel` (default), :func:`lorentz_kernel` or :func:`dirichlet_kernel`.
    num_threads : int
        Accepted for compatibility; the work runs on the GPU.
    silent : bool
        Don't show any progress messages.
    **kwargs
        matrix_format ("ELL" or "CSR"), optimal_size, interleaved, lanczos_precision,
        device (GPU index), kubo_precision ("fp64", "fp32_split" or "bf16x3"), autotune (time the
        step kernel's variants on the first block of random vectors of each width and keep the
        fastest; the choice stays with the device for later calculations),
        site_order ("auto": a Hilbert-curve order of the model's site positions for whole-system
        calculations, "bfs": the reference's BFS reorder only, or an explicit permutation).
    """
    native = load_module()
    order = _site_order(model, kwargs.pop("site_order", "auto"))
    h = _as_hamiltonian(model)
    h.sort_indices()
    if kernel == "default":
        kernel = None
    if "progress_callback" not in kwargs and not silent:
        kwargs["progress_callback"] = _ComputeProgressReporter()
    if silent:
        kwargs.pop("progress_callback", None)
    impl = native.kpm(h.shape, h.indptr.astype(np.int32), h.indices.astype(np.int32), h.data,
                      tuple(energy_range) if energy_range is not None else (0.0, 0.0), kernel, **kwargs)
    if order is not None:
        impl.set_locality_order(order)
    return KPM(impl, model if not scipy.sparse.issparse(model) else None)


kpm_hip = kpm


def kpm_cuda(model, energy_range=None, kernel="default", **kwargs):
    """The reference's GPU entry point name (pybinding/chebyshev.py:381-404); runs on the HIP engine."""
    return kpm(model, energy_range, kernel, **kwargs)


def jackson_kernel():
    """The Jackson kernel -- a good general-purpose kernel (pybinding/chebyshev.py:407-415)"""
    return load_module().jackson_kernel()


def lorentz_kernel(lambda_value=4.0):
    """The Lorentz kernel -- best for Green's function (pybinding/chebyshev.py:418-431)"""
    return load_module().lorentz_kernel(lambda_value)


def dirichlet_kernel():
    """The Dirichlet kernel -- returns raw moments (pybinding/chebyshev.py:434-445)"""
    return load_module().dirichlet_kernel()
