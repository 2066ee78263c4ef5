"""GPU check of the allocation-free entry point xt_tt_sample_residual_adjoint_f32_noalloc against the allocating
xt_tt_sample_residual_adjoint_f32 whose kernels it shares: two calls in a row, the second on a smaller problem (fewer
samples, lower order), each bit-identical to the allocating call (residuals, gradients, FP64 sums), with the library's
cached device bytes unchanged by the allocation-free calls. Shapes beyond the scratch array are refused."""
import ctypes

import numpy as np
import pytest

from helpers import SEED

pytestmark = pytest.mark.gpu


def _problem(oracle, d, n, r, P, seed):
    oracle.seed(seed)
    x = oracle.TT.random([n] * d, [r] * (d - 1))
    truth = oracle.TT.random([n] * d, [r] * (d - 1))
    m = oracle.Measurements.random(P, [n] * d)
    m.measure(truth)
    cores = [c.astype(np.float32) for c in x.cores()]
    return cores, [1] + [c.shape[2] for c in cores], m.positions(d), m.values()


def _cache_bytes(call):
    use, cached = ctypes.c_size_t(), ctypes.c_size_t()
    call("xt_cache_bytes", ctypes.byref(use), ctypes.byref(cached))
    return use.value + cached.value


def test_noalloc_matches_allocating_call(xe, oracle):
    from xerus_amd._lib import DeviceArray, call
    for step, (d, n, r, P) in enumerate([(10, 10, 8, 20000), (6, 7, 5, 3001)]):
        print(f"no-alloc call {step}: d={d} n={n} r={r} P={P}", flush=True)
        cores, rk, pos, vals = _problem(oracle, d, n, r, P, SEED + 11 + step)
        flat = DeviceArray.from_numpy(np.concatenate([c.ravel() for c in cores]).astype(np.float32))
        dpos = DeviceArray.from_numpy(pos.astype(np.uint32))
        db = DeviceArray.from_numpy(vals.astype(np.float32))
        ns, rs = (ctypes.c_int * d)(*[n] * d), (ctypes.c_int * (d + 1))(*rk)
        # the allocating call
        resid_a, grad_a = DeviceArray((P,), np.float32), DeviceArray(flat.shape, np.float32)
        sums_a = (ctypes.c_double * 2)()
        call("xt_tt_sample_residual_adjoint_f32", d, ns, rs, flat.ptr, dpos.ptr, P, db.ptr, resid_a.ptr, sums_a, grad_a.ptr, None)
        # the allocation-free call: outputs pre-filled with garbage, sums on the device
        resid_b = DeviceArray.from_numpy(np.full(P, np.nan, np.float32))
        grad_b = DeviceArray.from_numpy(np.full(flat.shape, 7.0, np.float32))
        sums_b = DeviceArray.from_numpy(np.full(2, -1.0))
        before = _cache_bytes(call)
        call("xt_tt_sample_residual_adjoint_f32_noalloc", d, ns, rs, flat.ptr, dpos.ptr, P, db.ptr, resid_b.ptr, sums_b.ptr,
             grad_b.ptr, None)
        call("xt_synchronize")
        assert _cache_bytes(call) == before
        np.testing.assert_array_equal(resid_b.numpy(), resid_a.numpy())
        np.testing.assert_array_equal(grad_b.numpy(), grad_a.numpy())
        np.testing.assert_array_equal(sums_b.numpy(), np.array(sums_a[:]))


def test_noalloc_refuses_shapes_outside_its_kernel(xe):
    """Shapes the one-pass kernel does not take (rank above 8, a mode above 16, order above 16) are refused before
    anything is launched. Every buffer passed is a real one of the size the shape asks for (the scratch array itself
    fits every shape the kernel takes: at most 14592 core floats against room for 16314)."""
    from xerus_amd._lib import XT_ERR_INVALID, XT_ERR_UNSUPPORTED, DeviceArray, lib
    buf = DeviceArray((1 << 16,), np.float32)
    pos = DeviceArray((1 << 16,), np.uint32)
    for d, nn, r, want in ((4, 10, 12, XT_ERR_UNSUPPORTED), (4, 40, 8, XT_ERR_UNSUPPORTED), (17, 4, 4, XT_ERR_INVALID)):
        ns, rs = (ctypes.c_int * d)(*[nn] * d), (ctypes.c_int * (d + 1))(*([1] + [r] * (d - 1) + [1]))
        rc = lib().xt_tt_sample_residual_adjoint_f32_noalloc(d, ns, rs, buf.ptr, pos.ptr, 10, buf.ptr, buf.ptr, None, buf.ptr, None)
        assert rc == want, (d, nn, r, rc, lib().xt_last_error())
