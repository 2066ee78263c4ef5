"""ALS building blocks on device buffers (xt_als_env_left/right, xt_als_local_apply, xt_als_local_cg) against dense
evaluations of their definitions (als.cpp:184-200 and :239-247 for the stack pushes, :383-402 for the local
operator) and, for the local CG, against a dense solve of the densified local operator (what xerus's
lapack_solver, als.cpp:43-75, solves)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    from xerus_amd._lib import DeviceArray
    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=np.float64))


def _shapes(rng, a=5, n=4, c=6, p=3, q=2):
    L = rng.standard_normal((a, p, a))
    R = rng.standard_normal((c, q, c))
    A = rng.standard_normal((p, n, n, q))
    X = rng.standard_normal((a, n, c))
    return L, R, A, X


def test_env_updates(xe):
    from xerus_amd._lib import DeviceArray, call
    print("env updates", flush=True)
    rng = np.random.default_rng(3)
    L, R, A, X = _shapes(rng)
    a, n, c, p, q = 5, 4, 6, 3, 2
    dL, dR, dA, dX = _dev(L), _dev(R), _dev(A), _dev(X)  # named: a temporary's buffer would be freed before the call
    out = DeviceArray((c, q, c), np.float64)
    call("xt_als_env_left", dL.ptr, dX.ptr, dA.ptr, a, n, c, p, q, out.ptr)
    ref = np.einsum("apb,anc,pnmq,bmd->cqd", L, X, A, X)
    assert np.abs(out.numpy() - ref).max() < 1e-12 * np.abs(ref).max()
    out2 = DeviceArray((a, p, a), np.float64)
    call("xt_als_env_right", dR.ptr, dX.ptr, dA.ptr, a, n, c, p, q, out2.ptr)
    ref2 = np.einsum("cqd,anc,pnmq,bmd->apb", R, X, A, X)
    assert np.abs(out2.numpy() - ref2).max() < 1e-12 * np.abs(ref2).max()


def test_local_apply(xe):
    from xerus_amd._lib import DeviceArray, call
    print("local apply", flush=True)
    rng = np.random.default_rng(4)
    L, R, A, x = _shapes(rng)
    a, n, c, p, q = 5, 4, 6, 3, 2
    L2 = rng.standard_normal((a, p, a))
    R2 = rng.standard_normal((c, q, c))
    dL, dA, dR, dx = _dev(L2), _dev(A), _dev(R2), _dev(x)
    y = DeviceArray((a, n, c), np.float64)
    call("xt_als_local_apply", dL.ptr, dA.ptr, dR.ptr, a, n, c, p, q, dx.ptr, y.ptr)
    ref = np.einsum("apb,pnmq,cqd,bmd->anc", L2, A, R2, x)
    assert np.abs(y.numpy() - ref).max() < 1e-12 * np.abs(ref).max()


def _spd(rng, k):
    M = rng.standard_normal((k, k))
    return M @ M.T + k * np.eye(k)


@pytest.mark.parametrize("a,n,c", [(4, 3, 5), (16, 8, 12)])
def test_local_cg_solves_the_dense_system(xe, a, n, c):
    """A_loc = L_0⊗A_00⊗R_0 + L_1⊗A_11⊗R_1 with SPD factors (SPD); x from a zero start against the dense solve."""
    from xerus_amd._lib import DeviceArray, call
    print(f"local cg a={a} n={n} c={c}", flush=True)
    rng = np.random.default_rng(a * n * c)
    p = q = 2
    L = np.stack([_spd(rng, a), _spd(rng, a)], axis=1)                # (a, p, a)
    R = np.stack([_spd(rng, c), _spd(rng, c)], axis=1)                # (c, q, c)
    A = np.zeros((p, n, n, q))
    A[0, :, :, 0], A[1, :, :, 1] = _spd(rng, n), _spd(rng, n)
    b = rng.standard_normal((a, n, c))
    dL, dA, dR, db = _dev(L), _dev(A), _dev(R), _dev(b)
    dx = _dev(np.zeros((a, n, c)))
    it = ctypes.c_size_t()
    call("xt_als_local_cg", dL.ptr, dA.ptr, dR.ptr, a, n, c, p, q, db.ptr, dx.ptr, 1e-13, 5000, ctypes.byref(it))
    M = np.einsum("apb,pnmq,cqd->ancbmd", L, A, R).reshape(a * n * c, a * n * c)
    ref = np.linalg.solve(M, b.ravel())
    x = dx.numpy().ravel()
    assert it.value > 0
    assert np.linalg.norm(M @ x - b.ravel()) <= 1e-12 * np.linalg.norm(b)
    assert np.abs(x - ref).max() < 1e-9 * np.abs(ref).max()
    # a converged start returns at once
    it2 = ctypes.c_size_t()
    call("xt_als_local_cg", dL.ptr, dA.ptr, dR.ptr, a, n, c, p, q, db.ptr, dx.ptr, 1e-9, 5000, ctypes.byref(it2))
    assert it2.value == 0


def test_building_blocks_reject_empty_dimensions(xe):
    from xerus_amd._lib import DeviceArray, call
    d = DeviceArray((4,), np.float64)
    with pytest.raises(xe.XttError):
        call("xt_als_local_apply", d.ptr, d.ptr, d.ptr, 1, 0, 1, 1, 1, d.ptr, d.ptr)
