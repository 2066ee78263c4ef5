"""ctypes binding of libxtt.so (the C ABI declared in include/xtt.h).

The library is built in-tree (``make -C xerus_amd/csrc``); loading fails loudly when it is missing — there is
no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxtt.so")

XT_OK, XT_ERR_INVALID, XT_ERR_HIP, XT_ERR_NUMERIC, XT_ERR_NOMEM, XT_ERR_UNSUPPORTED = range(6)
XT_F64, XT_F32, XT_BF16 = 0, 1, 2

_sz = C.c_size_t
_ps = C.POINTER(C.c_size_t)
_pd = C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)
_pi = C.POINTER(C.c_int)
_pu32 = C.POINTER(C.c_uint32)
_vp = C.c_void_p
# void (*)(void* user, size_t iteration, double elapsed_us, double residual, const size_t* ranks, size_t nranks, size_t flags)
PERF_CALLBACK = C.CFUNCTYPE(None, _vp, C.c_size_t, C.c_double, C.c_double, C.POINTER(C.c_size_t), C.c_size_t, C.c_size_t)


class XtALSOptions(C.Structure):
    _fields_ = [
        ("sites", C.c_uint),
        ("numHalfSweeps", C.c_size_t),
        ("convergenceEpsilon", C.c_double),
        ("useResidualForEndCriterion", C.c_int),
        ("preserveCorePosition", C.c_int),
        ("assumeSPD", C.c_int),
        ("precision", C.c_int),
        ("localTolerance", C.c_double),
        ("maxLocalIterations", C.c_int),
        ("innerTolerance", C.c_double),
        ("maxRefinements", C.c_int),
    ]


class XtALSStats(C.Structure):
    _fields_ = [
        ("halfSweeps", C.c_size_t),
        ("localSolves", C.c_size_t),
        ("cgIterations", C.c_size_t),
        ("refinements", C.c_size_t),
        ("matvecs64", C.c_size_t),
        ("matvecsLow", C.c_size_t),
        ("flopsMatvec", C.c_double),
        ("flopsEnv", C.c_double),
    ]


# name → (restype, argtypes); every entry point returns an int status
SIGNATURES = {
    "xt_last_error": (C.c_char_p, []),
    "xt_version": (C.c_char_p, []),
    "xt_synchronize": (C.c_int, []),
    "xt_release_cache": (C.c_int, []),
    "xt_cache_bytes": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "xt_graph_replays": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "xt_gemm_tune": (C.c_int, [C.c_int, _sz, _sz, _sz, _sz, C.c_int, C.c_int, C.c_int, _pi, _pi, _pd, _pi, _pi, _pd]),
    "xt_gemm_tune_set": (C.c_int, [C.c_int, _sz, _sz, _sz, _sz, C.c_int, C.c_int, C.c_int, C.c_int]),
    "xt_gemm_tune_clear": (C.c_int, []),
    "xt_gemm_tune_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_size_t)]),
    "xt_malloc": (C.c_int, [_sz, C.POINTER(_vp)]),
    "xt_free": (C.c_int, [_vp]),
    "xt_memcpy_htod": (C.c_int, [_vp, _vp, _sz]),
    "xt_memcpy_dtoh": (C.c_int, [_vp, _vp, _sz]),
    "xt_profile_enable": (C.c_int, [C.c_int]),
    "xt_profile_read": (C.c_int, [_sz, C.c_char_p, _sz, _ps, _pd, _pd, _pd, _ps]),
    "xt_seed": (C.c_int, [C.c_uint64]),
    "xt_normal_draws": (C.c_int, [_sz, _pd]),
    "xt_uniform_draws": (C.c_int, [_sz, _sz, _ps]),
    "xt_gemm": (C.c_int, [C.c_int, C.c_int, _sz, _sz, _sz, C.c_double, _vp, _vp, _vp, _vp]),
    "xt_gemm_f32": (C.c_int, [C.c_int, C.c_int, _sz, _sz, _sz, C.c_float, _vp, _vp, _vp, _vp]),
    "xt_gemm_bf16": (C.c_int, [C.c_int, C.c_int, _sz, _sz, _sz, C.c_float, _vp, _vp, _vp, _vp]),
    "xt_qr": (C.c_int, [_vp, _sz, _sz, _vp, _vp]),
    "xt_rq": (C.c_int, [_vp, _sz, _sz, _vp, _vp]),
    "xt_qc": (C.c_int, [_vp, _sz, _sz, _vp, _vp, _ps]),
    "xt_cq": (C.c_int, [_vp, _sz, _sz, _vp, _vp, _ps]),
    "xt_solve": (C.c_int, [_vp, _sz, _sz, _vp, _sz, _vp, C.POINTER(C.c_int)]),
    "xt_als_env_left": (C.c_int, [_vp, _vp, _vp, _sz, _sz, _sz, _sz, _sz, _vp]),
    "xt_als_env_right": (C.c_int, [_vp, _vp, _vp, _sz, _sz, _sz, _sz, _sz, _vp]),
    "xt_als_local_apply": (C.c_int, [_vp, _vp, _vp, _sz, _sz, _sz, _sz, _sz, _vp, _vp]),
    "xt_als_local_cg": (C.c_int, [_vp, _vp, _vp, _sz, _sz, _sz, _sz, _sz, _vp, _vp, C.c_double, C.c_int, C.POINTER(C.c_size_t)]),
    "xt_svd": (C.c_int, [_vp, _sz, _sz, _vp, _vp, _vp]),
    "xt_reshuffle": (C.c_int, [_sz, _ps, _ps, _vp, _vp]),
    "xt_one_norm": (C.c_int, [_vp, _sz, _pd]),
    "xt_two_norm": (C.c_int, [_vp, _sz, _pd]),
    "xt_matrix_vector_product": (C.c_int, [_vp, _sz, C.c_double, _vp, _sz, C.c_int, _vp, _vp]),
    "xt_dyadic_vector_product": (C.c_int, [_vp, _sz, _sz, C.c_double, _vp, _vp, _vp]),
    "xt_solve_least_squares": (C.c_int, [_vp, _vp, _sz, _sz, _vp, _sz]),
    "xt_dot": (C.c_int, [_vp, _vp, _sz, _pd]),
    "xt_tt_new": (C.c_int, [C.c_int, _sz, _ps, C.POINTER(_vp)]),
    "xt_tt_random": (C.c_int, [C.c_int, _sz, _ps, _ps, C.POINTER(_vp)]),
    "xt_tt_from_tensor": (C.c_int, [C.c_int, _sz, _ps, _pd, C.c_double, _ps, C.POINTER(_vp)]),
    "xt_tt_ones": (C.c_int, [C.c_int, _sz, _ps, C.POINTER(_vp)]),
    "xt_tt_identity": (C.c_int, [_sz, _ps, C.POINTER(_vp)]),
    "xt_tt_kronecker": (C.c_int, [C.c_int, _sz, _ps, C.POINTER(_vp)]),
    "xt_tt_dirac": (C.c_int, [C.c_int, _sz, _ps, _ps, C.POINTER(_vp)]),
    "xt_tt_transpose": (C.c_int, [_vp]),
    "xt_tt_copy": (C.c_int, [_vp, C.POINTER(_vp)]),
    "xt_tt_free": (C.c_int, [_vp]),
    "xt_tt_order": (C.c_int, [_vp, _ps]),
    "xt_tt_is_operator": (C.c_int, [_vp, _pi]),
    "xt_tt_dims": (C.c_int, [_vp, _ps]),
    "xt_tt_ranks": (C.c_int, [_vp, _ps]),
    "xt_tt_component_dims": (C.c_int, [_vp, _sz, _ps]),
    "xt_tt_get_component": (C.c_int, [_vp, _sz, _pd]),
    "xt_tt_set_component": (C.c_int, [_vp, _sz, _sz, _ps, _pd]),
    "xt_tt_component_ptr": (C.c_int, [_vp, _sz, C.POINTER(_vp)]),
    "xt_tt_canonicalization": (C.c_int, [_vp, _pi, _ps]),
    "xt_tt_assume_core_position": (C.c_int, [_vp, _sz]),
    "xt_tt_move_core": (C.c_int, [_vp, _sz, C.c_int]),
    "xt_tt_round": (C.c_int, [_vp, _ps, C.c_double]),
    "xt_tt_frob_norm": (C.c_int, [_vp, _pd]),
    "xt_tt_dot": (C.c_int, [_vp, _vp, _pd]),
    "xt_tt_add": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "xt_tt_scale": (C.c_int, [_vp, C.c_double]),
    "xt_tt_apply": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "xt_tt_full": (C.c_int, [_vp, _pd]),
    "xt_tt_residual": (C.c_int, [_vp, _vp, _vp, _pd]),
    "xt_tt_save": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "xt_tt_load": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "xt_als_default_options": (C.c_int, [C.POINTER(XtALSOptions), C.c_uint, C.c_int]),
    "xt_als_solve": (C.c_int, [C.POINTER(XtALSOptions), _vp, _vp, _vp, _pd, C.POINTER(XtALSStats)]),
    "xt_als_round_batch": (C.c_int, [C.POINTER(XtALSOptions), _sz, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), _sz,
                                     C.c_double, C.c_int, _pd, C.POINTER(XtALSStats)]),
    "xt_r1meas_random": (C.c_int, [_sz, _sz, _ps, C.POINTER(_vp)]),
    "xt_r1meas_from_single_point": (C.c_int, [_vp, _sz, _ps, C.POINTER(_vp)]),
    "xt_r1meas_new": (C.c_int, [_sz, _sz, _ps, _pd, _pd, C.POINTER(_vp)]),
    "xt_r1meas_free": (C.c_int, [_vp]),
    "xt_r1meas_size": (C.c_int, [_vp, _ps]),
    "xt_r1meas_vectors": (C.c_int, [_vp, _pd]),
    "xt_r1meas_values": (C.c_int, [_vp, _pd]),
    "xt_r1meas_set_values": (C.c_int, [_vp, _pd]),
    "xt_r1meas_sort": (C.c_int, [_vp, C.c_int]),
    "xt_r1meas_normalize": (C.c_int, [_vp]),
    "xt_r1meas_measure": (C.c_int, [_vp, _vp]),
    "xt_r1meas_test": (C.c_int, [_vp, _vp, _pd]),
    "xt_adf_gradient_r1": (C.c_int, [_vp, _vp, C.c_int, _pd, _pd]),
    "xt_adf_solve_r1": (C.c_int, [_vp, _vp, _ps, _sz, C.c_double, C.c_double, C.c_int, _pd, _ps]),
    "xt_meas_random": (C.c_int, [_sz, _sz, _ps, C.POINTER(_vp)]),
    "xt_meas_new": (C.c_int, [_sz, _sz, _pu32, _pd, C.POINTER(_vp)]),
    "xt_meas_free": (C.c_int, [_vp]),
    "xt_meas_size": (C.c_int, [_vp, _ps]),
    "xt_meas_positions": (C.c_int, [_vp, _pu32]),
    "xt_meas_positions_ptr": (C.c_int, [_vp, C.POINTER(_vp)]),
    "xt_meas_values": (C.c_int, [_vp, _pd]),
    "xt_meas_set_values": (C.c_int, [_vp, _pd]),
    "xt_meas_measure": (C.c_int, [_vp, _vp]),
    "xt_meas_stack_plan": (C.c_int, [_vp, _ps, _ps, _pu32, _pu32, _pu32, _pu32, _ps]),
    "xt_adf_gradient": (C.c_int, [_vp, _vp, C.c_int, _pd, _pd]),
    "xt_adf_solve": (C.c_int, [_vp, _vp, _ps, _sz, C.c_double, C.c_double, C.c_int, _pd, _ps]),
    "xt_perf_sink": (C.c_int, [PERF_CALLBACK, _vp]),
    "xt_tt_sample_f32": (C.c_int, [C.c_int, _pi, _pi, _vp, _vp, C.c_int, _vp, _vp]),
    "xt_tt_sample_residual_f32": (C.c_int, [C.c_int, _pi, _pi, _vp, _vp, C.c_int, _vp, _vp, _pd, _vp]),
    "xt_tt_sample_adjoint_f32": (C.c_int, [C.c_int, _pi, _pi, _vp, _vp, C.c_int, _vp, _vp, _vp]),
    "xt_tt_sample_residual_adjoint_f32": (C.c_int, [C.c_int, _pi, _pi, _vp, _vp, C.c_int, _vp, _vp, _pd, _vp, _vp]),
    "xt_tt_sample_residual_adjoint_f32_noalloc": (C.c_int, [C.c_int, _pi, _pi, _vp, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None


class XttError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[xtt error {code}] {msg}")
        self.code = code


def _foreign_hip_runtime():
    """Path of a HIP runtime mapped into this process from outside the ROCm installation (torch bundles one),
    or None. libxtt resolves libamdhip64.so.7 to an already mapped runtime of that name."""
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                path = line.split()[-1] if line.strip() else ""
                if "libamdhip64" in path and not path.startswith("/opt/rocm"):
                    return path
    except OSError:
        pass
    return None


_on_torch_runtime = False


def lib() -> C.CDLL:
    """Load libxtt.so once; raises if it has not been built.

    libxtt binds to the HIP runtime already mapped into the process when it is loaded. Loaded first (the usual
    case for xerus_amd), that is the ROCm runtime it was built with; loaded after torch, it is torch's bundled
    runtime, and then libxtt's kernels work on torch's device tensors and streams (xerus_amd.autograd relies on
    it). A process keeps one runtime: torch imported after libxtt runs on the ROCm runtime without a usable
    device, so tensors there must stay on the CPU."""
    global _lib, _on_torch_runtime
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libxtt.so not built: {LIB_PATH} missing (run `make -C xerus_amd/csrc`)")
        _on_torch_runtime = _foreign_hip_runtime() is not None
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        if not os.environ.get("XTT_NO_TUNING"):  # the measured GEMM tile table, when one is committed
            from . import tuning
            if os.path.exists(tuning.DEFAULT_TABLE):
                tuning.load()
    return _lib


def on_torch_runtime() -> bool:
    """Whether libxtt was loaded on torch's HIP runtime (device tensors of torch are then valid libxtt pointers)."""
    return _lib is not None and _on_torch_runtime


def check(rc: int) -> None:
    if rc != XT_OK:
        raise XttError(rc, lib().xt_last_error().decode())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def profile_enable(on: bool) -> None:
    call("xt_profile_enable", int(on))


def profile_read(max_records: int = 4096):
    """Per-kernel {tag: (launches, total_ms, algorithmic_flops, algorithmic_bytes)} since the last read (synchronises)."""
    tag_len = 96
    tags = C.create_string_buffer(max_records * tag_len)
    launches = (C.c_size_t * max_records)()
    ms = (C.c_double * max_records)()
    flops = (C.c_double * max_records)()
    nbytes = (C.c_double * max_records)()
    count = C.c_size_t()
    check(lib().xt_profile_read(max_records, tags, tag_len, launches, ms, flops, nbytes, C.byref(count)))
    out = {}
    for i in range(min(count.value, max_records)):
        tag = tags.raw[i * tag_len:(i + 1) * tag_len].split(b"\0", 1)[0].decode()
        out[tag] = (int(launches[i]), float(ms[i]), float(flops[i]), float(nbytes[i]))
    return out


class DeviceArray:
    """A device buffer owned through libxtt (xt_malloc), for callers that do not bring their own allocator."""

    def __init__(self, shape, dtype):
        import numpy as np
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        call("xt_malloc", self.nbytes, C.byref(p))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        call("xt_memcpy_htod", d.ptr, a.ctypes.data, d.nbytes)
        return d

    def numpy(self):
        import numpy as np
        out = np.empty(self.shape, dtype=self.dtype)
        call("xt_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p and _lib is not None:
            _lib.xt_free(p)
            self.ptr = None


def graph_replays():
    """(ALS call, ADF sweep) replays of recorded calls accepted since the process started."""
    out = [C.c_size_t() for _ in range(2)]
    call("xt_graph_replays", *[C.byref(o) for o in out])
    return tuple(o.value for o in out)
