"""Shared builders for the parity tests: the workloads of the benchmark configurations at test sizes."""
import numpy as np

from xerus_amd.workloads import laplace_cores  # noqa: F401

SEED = 0xBAADF00D  # the reference's unit-test seed (src/xerus/test/test.cpp:105)


def apply_dense_laplace(v, d, n, potentials=None):
    L = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    out = np.zeros_like(v)
    for k in range(d):
        H = L + (np.diag(potentials[k]) if potentials is not None else 0)
        out += np.moveaxis(np.tensordot(H, v, axes=([1], [k])), 0, k)
    return out


def exact_laplace_solve(b, d, n):
    """x* = (⊗S)·diag(1/Σλ)·(⊗S)ᵀ·b for the Dirichlet Laplacian (sine eigenbasis)."""
    L = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    lam, S = np.linalg.eigh(L)
    y = b.copy()
    for k in range(d):
        y = np.moveaxis(np.tensordot(S.T, y, axes=([1], [k])), 0, k)
    y /= sum(np.meshgrid(*([lam] * d), indexing="ij"))
    for k in range(d):
        y = np.moveaxis(np.tensordot(S, y, axes=([1], [k])), 0, k)
    return y
