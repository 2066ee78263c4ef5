"""TEST INFRASTRUCTURE ONLY — numpy restatement of xerus's RankOneMeasurementSet and its ADF path.

Follows xerus 3.0.1:
- ``src/xerus/measurments.cpp:253-265`` (conversion from single points: unit vectors, same order),
  ``:329-344`` + ``:549-558`` (``sort``: the first differing entry decides, the larger entry goes first),
  ``:346-355`` (``normalize``), ``:392-424`` (``measure``: x(u) = Σ x[i_0..i_{d-1}]·Π u_k[i_k]),
  ``:469-505`` (``test``: ‖b − x(u)‖/‖b‖), ``:513-536`` (``create_random_positions``: each vector
  ``Tensor::random({n_k})``, measurement-major, then ``sort(true)``);
- ``src/xerus/algorithms/adf.cpp:233-288`` (stacks through the mixed matrices M_k = Σ_j u_k[j]·G_k[:, j, :]),
  ``:333-357`` (rank-one dyadic product: Δ[j] += r·u_k[j]·F ⊗ B), ``:413-465`` (step denominator, all at slot 0),
  ``:481-487`` (``update_x``: one step ‖Δ‖²/‖AΔ‖² for the whole core) and ``:489-542`` (the sweep, with the
  core moved by QR; the orthogonal gauge does not change the iterates as full tensors).

The random draws come from the FP64 oracle's engine (``oracle.normal_draws``), which mirrors xerus's
``misc::randomEngine`` + ``defaultNormalDistribution``. Only ``tests/`` import this module.
"""
from __future__ import annotations

import numpy as np


def offsets(dims):
    return np.concatenate([[0], np.cumsum(dims)]).astype(int)


def sort_rows(U, values=None, positions_only=True):
    """RankOneMeasurementSet::sort: rows ordered so that at the first differing entry the larger comes first."""
    order = sorted(range(U.shape[0]), key=lambda i: tuple(-U[i]))
    U = U[order]
    if values is not None and not positions_only:
        values = values[order]
    return U, values


def random_set(P, dims, draws):
    """create_random_positions from P·Σn normal draws (measurement-major, mode-minor)."""
    U = np.asarray(draws, dtype=np.float64).reshape(P, int(sum(dims)))
    U, _ = sort_rows(U)
    return U, np.zeros(P)


def from_single_point(positions, values, dims):
    o = offsets(dims)
    P = positions.shape[0]
    U = np.zeros((P, o[-1]))
    for k in range(len(dims)):
        U[np.arange(P), o[k] + positions[:, k]] = 1.0
    return U, np.asarray(values, dtype=np.float64).copy()


def normalize(U, values, dims):
    o = offsets(dims)
    U = U.copy()
    values = np.asarray(values, dtype=np.float64).copy()
    for k in range(len(dims)):
        nrm = np.linalg.norm(U[:, o[k]:o[k + 1]], axis=1)
        U[:, o[k]:o[k + 1]] /= nrm[:, None]
        values /= nrm
    return U, values


def _mixed(core, u):  # M_k for every measurement: (P, r_l, r_r)
    return np.einsum("pj,ajc->pac", u, core)


def left_stacks(cores, U, dims):
    """F[k] = M_0·…·M_{k-1} for every measurement, k = 0..d (F[0] = 1)."""
    o = offsets(dims)
    P = U.shape[0]
    F = [np.ones((P, 1))]
    for k, c in enumerate(cores):
        F.append(np.einsum("pa,pac->pc", F[-1], _mixed(c, U[:, o[k]:o[k + 1]])))
    return F


def right_stacks(cores, U, dims):
    """B[k] = M_k·…·M_{d-1}·1, k = 0..d (B[d] = 1)."""
    o = offsets(dims)
    P = U.shape[0]
    d = len(cores)
    B = [None] * (d + 1)
    B[d] = np.ones((P, 1))
    for k in range(d - 1, -1, -1):
        B[k] = np.einsum("pac,pc->pa", _mixed(cores[k], U[:, o[k]:o[k + 1]]), B[k + 1])
    return B


def value(cores, U, dims):
    return left_stacks(cores, U, dims)[-1][:, 0]


def test(cores, U, values, dims):
    return np.linalg.norm(values - value(cores, U, dims)) / np.linalg.norm(values)


def projected_gradient(cores, U, values, dims, k):
    """xerus's Δ_k (adf.cpp:359-396, rank-one dyadic product :333-357) in the core layout (r_l, n, r_r)."""
    o = offsets(dims)
    F = left_stacks(cores, U, dims)
    B = right_stacks(cores, U, dims)
    res = values - F[-1][:, 0]
    return np.einsum("p,pa,pj,pc->ajc", res, F[k], U[:, o[k]:o[k + 1]], B[k + 1])


def step_denominator(cores, U, dims, k, D):
    """Σ_i (x(u_i) with core k replaced by D)² — calculate_slicewise_norm_A_projGrad summed over its slots."""
    c2 = list(cores)
    c2[k] = D
    return float(np.sum(value(c2, U, dims) ** 2))


def _qr_right(cores, k):  # move the core from k to k+1 (keepRank QR)
    c = cores[k]
    rl, n, rr = c.shape
    Q, R = np.linalg.qr(c.reshape(rl * n, rr))
    cores[k] = Q.reshape(rl, n, Q.shape[1])
    cores[k + 1] = np.einsum("ab,bjc->ajc", R, cores[k + 1])


def _to_core_zero(cores):
    for k in range(len(cores) - 1, 0, -1):
        c = cores[k]
        rl, n, rr = c.shape
        Q, R = np.linalg.qr(c.reshape(rl, n * rr).T)  # c = Rᵀ·Qᵀ
        cores[k] = Q.T.reshape(Q.shape[1], n, rr)
        cores[k - 1] = np.einsum("ajb,cb->ajc", cores[k - 1], R)


def adf_solve(cores, U, values, dims, max_iterations, target=1e-8, min_decrease=0.999):
    """ADFVariant::solve_with_current_ranks for rank-one measurements (no rank increase: maxRanks = ranks)."""
    cores = [np.array(c, dtype=np.float64) for c in cores]
    d = len(cores)
    norm_b = np.linalg.norm(values)
    res_norm = last = np.finfo(float).max
    dec = [0.0, 0.0, 0.0]
    it = 0
    while max_iterations == 0 or it < max_iterations:
        _to_core_zero(cores)
        last, res_norm = res_norm, np.linalg.norm(values - value(cores, U, dims)) / norm_b
        dec4 = dec[2]
        dec = [res_norm / last, dec[0], dec[1]]
        if res_norm < target or dec[0] * dec[1] * dec[2] * dec4 > min_decrease ** 4:
            break
        for k in range(d):
            D = projected_gradient(cores, U, values, dims, k)
            step = float(np.sum(D * D)) / step_denominator(cores, U, dims, k, D)
            cores[k] = cores[k] + step * D
            if k + 1 < d:
                _qr_right(cores, k)
        it += 1
    return cores, res_norm, it


def full(cores):
    t = cores[0]
    for c in cores[1:]:
        t = np.tensordot(t, c, axes=([-1], [0]))
    return t.reshape(t.shape[1:-1])
