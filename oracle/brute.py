"""ORACLE -- test infrastructure only.

Closed forms that pin the recurrences independently of how they are coded: with the dense
eigendecomposition H~ = V diag(l) V^H (H~ = H2/2 with eigenvalues in [-1, 1]),

    T_n(H~) = V diag(cos(n arccos l)) V^H,

so every moment the reference defines has a direct formula:
    diagonal      mu_n = sum_k |<v_k|r>|^2 T_n(l_k)                     (mu_0 halved)
    generic       mu_n = <beta| op T_n(H~) |alpha>                       (mu_0 halved)
    vectors       T_n(H~) r                                               (row 0 halved)
"""
import numpy as np


def _eig(h2):
    h = np.asarray(h2.toarray() if hasattr(h2, "toarray") else h2, dtype=np.complex128) / 2
    lam, vec = np.linalg.eigh(h)
    lam = np.clip(lam, -1.0, 1.0)
    return lam, vec


def cheb(n, lam):
    return np.cos(np.arange(n)[:, None] * np.arccos(lam)[None, :])


def moments_eig(h2, alpha, num_moments, beta=None, op=None):
    """mu_n = <beta| op T_n(H2/2) |alpha>, mu_0 halved (reference collector convention)."""
    lam, vec = _eig(h2)
    alpha = np.asarray(alpha, dtype=np.complex128)
    beta = alpha if beta is None else np.asarray(beta, dtype=np.complex128)
    a_k = vec.conj().T @ alpha                       # <v_k|alpha>
    if op is not None:
        opd = np.asarray(op.toarray() if hasattr(op, "toarray") else op, dtype=np.complex128)
        b_k = (vec.conj().T @ (opd.conj().T @ beta))  # <v_k| op^H beta>
    else:
        b_k = vec.conj().T @ beta
    t = cheb(num_moments, lam)                       # (M, K)
    mu = t @ (b_k.conj() * a_k)
    mu[0] *= 0.5
    return mu


def vectors_eig(h2, alpha, num_moments):
    """Rows T_n(H2/2) alpha, row 0 halved (DenseMatrixCollector convention)."""
    lam, vec = _eig(h2)
    a_k = vec.conj().T @ np.asarray(alpha, dtype=np.complex128)
    t = cheb(num_moments, lam)
    out = (t * a_k[None, :]) @ vec.T
    out[0] *= 0.5
    return out
