"""TEST INFRASTRUCTURE ONLY — ctypes view of the oracle (``oracle/libxtt_oracle.so``).

The oracle is an FP64 CPU restatement of xerus 3.0.1's Tensor-Train path (see the C++ sources beside this
file for the reference file:line each routine follows). Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this module; the product (``xerus_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libxtt_oracle.so")
_lib = None

_sz = C.c_size_t
_pd = C.POINTER(C.c_double)
_ps = C.POINTER(C.c_size_t)
_pu32 = C.POINTER(C.c_uint32)
_vp = C.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        sig = {
            "or_last_error": (C.c_char_p, []),
            "or_seed": (None, [C.c_uint64]),
            "or_normal_draws": (None, [_sz, _pd]),
            "or_uniform_draws": (None, [_sz, _sz, _ps]),
            "or_raw_draws": (None, [_sz, C.POINTER(C.c_uint64)]),
            "or_gemm": (C.c_int, [C.c_int, C.c_int, _sz, _sz, _sz, C.c_double, _pd, _pd, _pd]),
            "or_qr": (C.c_int, [_pd, _sz, _sz, _pd, _pd]),
            "or_rq": (C.c_int, [_pd, _sz, _sz, _pd, _pd]),
            "or_qc": (C.c_int, [_pd, _sz, _sz, _pd, _pd, _ps]),
            "or_cq": (C.c_int, [_pd, _sz, _sz, _pd, _pd, _ps]),
            "or_svd": (C.c_int, [_pd, _sz, _sz, _pd, _pd, _pd]),
            "or_solve": (C.c_int, [_pd, _sz, _pd, _pd, C.POINTER(C.c_int)]),
            "or_reshuffle": (C.c_int, [_sz, _ps, _ps, _pd, _pd]),
            "or_contract": (C.c_int, [_sz, _ps, _pd, C.c_int, _sz, _ps, _pd, C.c_int, _sz, _pd]),
            "or_svd_truncated": (C.c_int, [_sz, _ps, _pd, _sz, _sz, C.c_double, _ps, _pd, _pd, _pd]),
            "or_tt_new": (_vp, [C.c_int, _sz, _ps]),
            "or_tt_random": (_vp, [C.c_int, _sz, _ps, _ps]),
            "or_tt_copy": (_vp, [_vp]),
            "or_tt_free": (None, [_vp]),
            "or_tt_order": (_sz, [_vp]),
            "or_tt_is_operator": (C.c_int, [_vp]),
            "or_tt_dims": (None, [_vp, _ps]),
            "or_tt_ranks": (None, [_vp, _ps]),
            "or_tt_core_dims": (None, [_vp, _sz, _ps]),
            "or_tt_get_core": (None, [_vp, _sz, _pd]),
            "or_tt_set_core": (C.c_int, [_vp, _sz, _sz, _ps, _pd]),
            "or_tt_state": (None, [_vp, C.POINTER(C.c_int), _ps]),
            "or_tt_set_state": (None, [_vp, C.c_int, _sz]),
            "or_tt_move_core": (C.c_int, [_vp, _sz, C.c_int]),
            "or_tt_round": (C.c_int, [_vp, _ps, C.c_double]),
            "or_tt_frob_norm": (C.c_double, [_vp]),
            "or_tt_dot": (C.c_double, [_vp, _vp]),
            "or_tt_add": (_vp, [_vp, _vp]),
            "or_tt_scale": (None, [_vp, C.c_double]),
            "or_tt_apply": (_vp, [_vp, _vp]),
            "or_tt_full": (C.c_int, [_vp, _pd]),
            "or_tt_residual": (C.c_double, [_vp, _vp, _vp]),
            "or_als": (C.c_int, [_vp, _vp, _vp, C.c_uint, _sz, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_double, _sz, _pd, _ps]),
            "or_meas_random": (_vp, [_sz, _sz, _ps]),
            "or_meas_new": (_vp, [_sz, _sz, _pu32, _pd]),
            "or_meas_free": (None, [_vp]),
            "or_meas_size": (_sz, [_vp]),
            "or_meas_positions": (None, [_vp, _pu32]),
            "or_meas_values": (None, [_vp, _pd]),
            "or_meas_set_values": (None, [_vp, _pd]),
            "or_meas_measure": (C.c_int, [_vp, _vp]),
            "or_adf_plan": (C.c_int, [_vp, _ps, _ps, _pu32, _pu32, _ps]),
            "or_adf_gradient": (C.c_int, [_vp, _vp, _pd, _pd]),
            "or_adf_solve": (C.c_int, [_vp, _vp, _ps, _sz, C.c_double, C.c_double, _pd, _ps]),
            "or_adf_core_step": (C.c_int, [_vp, _vp, _sz]),
            "or_tt_from_tensor": (_vp, [C.c_int, _sz, _ps, _pd, C.c_double, _ps]),
            "or_tt_special": (_vp, [C.c_int, C.c_int, _sz, _ps, _ps]),
            "or_tt_transpose": (C.c_int, [_vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().or_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_pd)


def _s(a):
    a = np.ascontiguousarray(a, dtype=np.uintp)
    return a, a.ctypes.data_as(_ps)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------------ RNG

def seed(s: int) -> None:
    lib().or_seed(s)


def normal_draws(n: int) -> np.ndarray:
    out = np.empty(n)
    lib().or_normal_draws(n, _d(out))
    return out


def uniform_draws(n: int, hi: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uintp)
    lib().or_uniform_draws(n, hi, out.ctypes.data_as(_ps))
    return out


def raw_draws(n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().or_raw_draws(n, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


# ------------------------------------------------------------------ dense kernels

def gemm(A, B, ta=False, tb=False, alpha=1.0):
    A, B = f64(A), f64(B)
    m = A.shape[1] if ta else A.shape[0]
    k = A.shape[0] if ta else A.shape[1]
    n = B.shape[0] if tb else B.shape[1]
    C_ = np.empty((m, n))
    _check(lib().or_gemm(int(ta), int(tb), m, n, k, alpha, _d(A), _d(B), _d(C_)))
    return C_


def qr(A):
    A = f64(A); m, n = A.shape; k = min(m, n)
    Q, R = np.empty((m, k)), np.empty((k, n))
    _check(lib().or_qr(_d(A), m, n, _d(Q), _d(R)))
    return Q, R


def rq(A):
    A = f64(A); m, n = A.shape; k = min(m, n)
    R, Q = np.empty((m, k)), np.empty((k, n))
    _check(lib().or_rq(_d(A), m, n, _d(R), _d(Q)))
    return R, Q


def qc(A):
    A = f64(A); m, n = A.shape; k = min(m, n)
    Q, Cm, r = np.zeros(m * k), np.zeros(k * n), C.c_size_t()
    _check(lib().or_qc(_d(A), m, n, _d(Q), _d(Cm), C.byref(r)))
    r = r.value
    return Q[: m * r].reshape(m, r), Cm[: r * n].reshape(r, n), r


def cq(A):
    A = f64(A); m, n = A.shape; k = min(m, n)
    Cm, Q, r = np.zeros(m * k), np.zeros(k * n), C.c_size_t()
    _check(lib().or_cq(_d(A), m, n, _d(Cm), _d(Q), C.byref(r)))
    r = r.value
    return Cm[: m * r].reshape(m, r), Q[: r * n].reshape(r, n), r


def svd(A):
    A = f64(A); m, n = A.shape; k = min(m, n)
    U, S, Vt = np.empty((m, k)), np.empty(k), np.empty((k, n))
    _check(lib().or_svd(_d(A), m, n, _d(U), _d(S), _d(Vt)))
    return U, S, Vt


def svd_truncated(T, split, max_rank=0, eps=0.0):
    T = f64(T); dims, pd_ = _s(T.shape)
    m = int(np.prod(T.shape[:split])); n = int(np.prod(T.shape[split:])); k = min(m, n)
    U, S, Vt, r = np.empty(m * k), np.empty(k), np.empty(k * n), C.c_size_t()
    _check(lib().or_svd_truncated(T.ndim, pd_, _d(T), split, max_rank, eps, C.byref(r), _d(U), _d(S), _d(Vt)))
    r = r.value
    return U[: m * r].reshape(T.shape[:split] + (r,)), S[:r].copy(), Vt[: r * n].reshape((r,) + T.shape[split:])


def solve(A, b):
    A, b = f64(A), f64(b)
    x, meth = np.empty(b.shape[0]), C.c_int()
    _check(lib().or_solve(_d(A), A.shape[0], _d(b), _d(x), C.byref(meth)))
    return x, meth.value


def reshuffle(T, shuffle):
    T = f64(T); dims, pd_ = _s(T.shape); sh, psh = _s(shuffle)
    nd = [0] * T.ndim
    for i, s in enumerate(shuffle):
        nd[s] = T.shape[i]
    out = np.empty(nd)
    _check(lib().or_reshuffle(T.ndim, pd_, psh, _d(T), _d(out)))
    return out


def contract(L, ltrans, R, rtrans, modes):
    L, R = f64(L), f64(R)
    ld, pl = _s(L.shape); rd, pr = _s(R.shape)
    lrem = L.shape[modes:] if ltrans else L.shape[: L.ndim - modes]
    rrem = R.shape[: R.ndim - modes] if rtrans else R.shape[modes:]
    out = np.empty(tuple(lrem) + tuple(rrem))
    _check(lib().or_contract(L.ndim, pl, _d(L), int(ltrans), R.ndim, pr, _d(R), int(rtrans), modes, _d(out)))
    return out


# ------------------------------------------------------------------ TT


class TT:
    """Owning handle of an oracle TTTensor / TTOperator."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError(lib().or_last_error().decode())
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_tt_free(self.h)
            self.h = None

    @staticmethod
    def zero(dims, op=False):
        d, p = _s(dims)
        return TT(lib().or_tt_new(int(op), len(dims), p))

    @staticmethod
    def random(dims, ranks, op=False):
        d, p = _s(dims); r, pr = _s(ranks if len(ranks) else [0])
        return TT(lib().or_tt_random(int(op), len(dims), p, pr))

    @staticmethod
    def from_cores(cores, op=False, canonicalized=False, core_position=0):
        d = len(cores)
        dims = [c.shape[1] for c in cores] + ([c.shape[2] for c in cores] if op else [])
        t = TT.zero(dims, op)
        for k, c in enumerate(cores):
            t.set_core(k, c)
        t.set_state(canonicalized, core_position)
        return t

    @staticmethod
    def from_tensor(full, op=False, eps=8 * np.finfo(float).eps, max_ranks=None):
        """TTNetwork(const Tensor&, eps, maxRanks) (ttNetwork.cpp:112-160); an operator's tensor is (n…, m…)."""
        full = f64(full); dims, p = _s(full.shape)
        mr = None
        if max_ranks is not None:
            d = full.ndim // (2 if op else 1)
            arr, mr = _s([max_ranks] * (d - 1) if np.isscalar(max_ranks) else max_ranks)
        return TT(lib().or_tt_from_tensor(int(op), full.ndim, p, _d(full), eps, mr))

    @staticmethod
    def ones(dims, op=False):
        d, p = _s(dims)
        return TT(lib().or_tt_special(0, int(op), len(dims), p, None))

    @staticmethod
    def identity(dims):
        d, p = _s(dims)
        return TT(lib().or_tt_special(1, 1, len(dims), p, None))

    @staticmethod
    def kronecker(dims, op=False):
        d, p = _s(dims)
        return TT(lib().or_tt_special(2, int(op), len(dims), p, None))

    @staticmethod
    def dirac(dims, position, op=False):
        d, p = _s(dims); q, pq = _s(position)
        return TT(lib().or_tt_special(3, int(op), len(dims), p, pq))

    def transpose(self):
        _check(lib().or_tt_transpose(self.h))

    def copy(self):
        return TT(lib().or_tt_copy(self.h))

    @property
    def order(self):
        return lib().or_tt_order(self.h)

    @property
    def is_operator(self):
        return bool(lib().or_tt_is_operator(self.h))

    @property
    def dims(self):
        n = self.order * (2 if self.is_operator else 1)
        out = np.empty(n, dtype=np.uintp)
        lib().or_tt_dims(self.h, out.ctypes.data_as(_ps))
        return [int(v) for v in out]

    def ranks(self):
        d = self.order
        out = np.zeros(max(d - 1, 1), dtype=np.uintp)
        lib().or_tt_ranks(self.h, out.ctypes.data_as(_ps))
        return [int(v) for v in out[: d - 1]]

    def core(self, k):
        nd = 4 if self.is_operator else 3
        dims = np.empty(nd, dtype=np.uintp)
        lib().or_tt_core_dims(self.h, k, dims.ctypes.data_as(_ps))
        out = np.empty(tuple(int(v) for v in dims))
        lib().or_tt_get_core(self.h, k, _d(out))
        return out

    def cores(self):
        return [self.core(k) for k in range(self.order)]

    def set_core(self, k, c):
        c = f64(c); dims, p = _s(c.shape)
        _check(lib().or_tt_set_core(self.h, k, c.ndim, p, _d(c)))

    def state(self):
        c, p = C.c_int(), C.c_size_t()
        lib().or_tt_state(self.h, C.byref(c), C.byref(p))
        return bool(c.value), p.value

    def set_state(self, canonicalized, pos):
        lib().or_tt_set_state(self.h, int(canonicalized), pos)

    def move_core(self, pos, keep_rank=False):
        _check(lib().or_tt_move_core(self.h, pos, int(keep_rank)))

    def round(self, max_ranks=None, eps=8 * np.finfo(float).eps):
        if max_ranks is None:
            max_ranks = [np.iinfo(np.uintp).max] * (self.order - 1)
        elif np.isscalar(max_ranks):
            max_ranks = [int(max_ranks)] * (self.order - 1)
        r, p = _s(max_ranks if len(max_ranks) else [0])
        _check(lib().or_tt_round(self.h, p, eps))

    def frob_norm(self):
        return lib().or_tt_frob_norm(self.h)

    def dot(self, other):
        return lib().or_tt_dot(self.h, other.h)

    def __add__(self, other):
        return TT(lib().or_tt_add(self.h, other.h))

    def scale(self, s):
        lib().or_tt_scale(self.h, s)

    def apply(self, x):
        return TT(lib().or_tt_apply(self.h, x.h))

    def full(self):
        out = np.empty(int(np.prod(self.dims)))
        _check(lib().or_tt_full(self.h, _d(out)))
        return out.reshape(self.dims)


def residual(A: TT, x: TT, b: TT) -> float:
    return lib().or_tt_residual(A.h, x.h, b.h)


def als(A, x, b, sites=1, num_half_sweeps=0, eps=1e-6, use_residual=False, preserve_core=True, spd=True,
        local_solver="dense", cg_tol=1e-13, cg_max_iter=4000):
    energy = C.c_double()
    stats = np.zeros(3, dtype=np.uintp)
    _check(lib().or_als(A.h if A is not None else None, x.h, b.h, sites, num_half_sweeps, eps, int(use_residual),
                        int(preserve_core), int(spd), 0 if local_solver == "dense" else 1, cg_tol, cg_max_iter,
                        C.byref(energy), stats.ctypes.data_as(_ps)))
    return energy.value, {"half_sweeps": int(stats[0]), "local_solves": int(stats[1]), "cg_iterations": int(stats[2])}


# ------------------------------------------------------------------ measurements / ADF


class Measurements:
    def __init__(self, handle):
        if not handle:
            raise RuntimeError(lib().or_last_error().decode())
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_meas_free(self.h)
            self.h = None

    @staticmethod
    def random(P, dims):
        d, p = _s(dims)
        return Measurements(lib().or_meas_random(P, len(dims), p))

    @staticmethod
    def from_arrays(positions, values):
        pos = np.ascontiguousarray(positions, dtype=np.uint32)
        val = f64(values)
        return Measurements(lib().or_meas_new(pos.shape[0], pos.shape[1], pos.ctypes.data_as(_pu32), _d(val)))

    def __len__(self):
        return lib().or_meas_size(self.h)

    def positions(self, d):
        out = np.empty((len(self), d), dtype=np.uint32)
        lib().or_meas_positions(self.h, out.ctypes.data_as(_pu32))
        return out

    def values(self):
        out = np.empty(len(self))
        lib().or_meas_values(self.h, _d(out))
        return out

    def set_values(self, v):
        v = f64(v)
        lib().or_meas_set_values(self.h, _d(v))

    def measure(self, x: TT):
        _check(lib().or_meas_measure(self.h, x.h))


def adf_plan(m: Measurements, d: int):
    fc, bc = np.zeros(d, dtype=np.uintp), np.zeros(d, dtype=np.uintp)
    nu = C.c_size_t()
    _check(lib().or_adf_plan(m.h, fc.ctypes.data_as(_ps), bc.ctypes.data_as(_ps), None, None, C.byref(nu)))
    fi, bi = np.empty(int(fc.sum()), dtype=np.uint32), np.empty(int(bc.sum()), dtype=np.uint32)
    _check(lib().or_adf_plan(m.h, fc.ctypes.data_as(_ps), bc.ctypes.data_as(_ps), fi.ctypes.data_as(_pu32),
                             bi.ctypes.data_as(_pu32), C.byref(nu)))
    return fc.astype(int), bc.astype(int), fi, bi, nu.value


def adf_gradient(x: TT, m: Measurements):
    cores = x.cores()
    out = np.empty(sum(c.size for c in cores))
    rn = C.c_double()
    _check(lib().or_adf_gradient(x.h, m.h, _d(out), C.byref(rn)))
    grads, o = [], 0
    for c in cores:
        grads.append(out[o: o + c.size].reshape(c.shape))
        o += c.size
    return grads, rn.value


def adf_core_step(x: TT, m: Measurements, k: int) -> None:
    """One ADF sweep step at component k on x as it is (gradient, slice norms, update_x; adf.cpp:360-478)."""
    _check(lib().or_adf_core_step(x.h, m.h, k))


def adf_solve(x: TT, m: Measurements, max_ranks=None, max_iterations=0, target=1e-8, min_decrease=0.999):
    res, it = C.c_double(), C.c_size_t()
    mr = None
    if max_ranks is not None:
        arr, mr = _s(max_ranks)
    _check(lib().or_adf_solve(x.h, m.h, mr, max_iterations, target, min_decrease, C.byref(res), C.byref(it)))
    return res.value, it.value
