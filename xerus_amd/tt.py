"""xerus-compatible Python surface of the TT path, backed by libxtt (device-resident cores on the MI355X).

Names and argument meaning follow xerus's Python module (src/xerus/python/ttnetwork.cpp:32-165,
leastSquares.cpp:97-138, recovery.cpp:31-160): ``TTTensor``/``TTOperator`` with ``random``,
``get_component``/``set_component``, ``ranks``, ``move_core``, ``round``, ``frob_norm``; the ``ALSVariant``
presets ``ALS``, ``ALS_SPD``, ``DMRG``, ``DMRG_SPD``; ``SinglePointMeasurementSet``; ``ADFVariant`` / ``ADF``.
Errors raise :class:`XttError` (xerus throws ``generic_error``).
"""
from __future__ import annotations

import copy as _copy
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import XT_BF16, XT_F32, XT_F64, XttError, call

EPSILON = 8 * np.finfo(float).eps  # include/xerus/basic.h:50
_SIZE_MAX = np.iinfo(np.uintp).max


def _sizes(seq):
    a = np.ascontiguousarray(list(seq), dtype=np.uintp)
    return a, a.ctypes.data_as(C.POINTER(C.c_size_t))


def seed(s: int) -> None:
    """misc::randomEngine.seed(s) (the seed the reference's tests use is 0xBAADF00D)."""
    call("xt_seed", int(s))


def normal_draws(n: int) -> np.ndarray:
    out = np.empty(n)
    call("xt_normal_draws", n, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def uniform_draws(n: int, hi: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uintp)
    call("xt_uniform_draws", n, hi, out.ctypes.data_as(C.POINTER(C.c_size_t)))
    return out


class _TTNetwork:
    _is_operator = False

    def __init__(self, arg=None, *, _handle=None):
        if _handle is not None:
            self._h = _handle
            return
        if arg is None:
            arg = [1]
        if isinstance(arg, (int, np.integer)):
            arg = [1] * int(arg)
        dims = list(arg)
        _, p = _sizes(dims)
        h = C.c_void_p()
        call("xt_tt_new", int(self._is_operator), len(dims), p, C.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib._lib is not None:
            _lib._lib.xt_tt_free(h)
            self._h = None

    # ------------------------------------------------------------------ construction
    @classmethod
    def random(cls, dims, ranks):
        """TTNetwork::random(dims, ranks) (ttNetwork.h:128-155): N(0,1) cores from the shared engine."""
        dims = list(dims)
        if isinstance(ranks, (int, np.integer)):
            ranks = [int(ranks)] * (len(dims) // (2 if cls._is_operator else 1) - 1)
        _, pd = _sizes(dims)
        _, pr = _sizes(list(ranks) or [0])
        h = C.c_void_p()
        call("xt_tt_random", int(cls._is_operator), len(dims), pd, pr, C.byref(h))
        return cls(_handle=h)

    @classmethod
    def from_components(cls, cores, core_position=None):
        """Build from a list of component arrays; canonical at `core_position` if given (else not canonical)."""
        cores = [np.asarray(c, dtype=np.float64) for c in cores]
        if cls._is_operator:
            dims = [c.shape[1] for c in cores] + [c.shape[2] for c in cores]
        else:
            dims = [c.shape[1] for c in cores]
        t = cls(dims)
        for k, c in enumerate(cores):
            t.set_component(k, c)
        if core_position is not None:
            t.assume_core_position(core_position)
        return t

    @classmethod
    def from_tensor(cls, tensor, eps: float = EPSILON, maxRanks=None):
        """TTNetwork(const Tensor&, eps, maxRanks) (ttNetwork.cpp:112-160): TT-SVD of a dense array (an operator's
        array is indexed (n_0..n_{d-1}, m_0..m_{d-1})); maxRanks a list, a number or None (no cap)."""
        full = np.ascontiguousarray(tensor, dtype=np.float64)
        _, pd = _sizes(full.shape)
        d = full.ndim // (2 if cls._is_operator else 1)
        pr = None
        if maxRanks is not None:
            ranks = [int(maxRanks)] * (d - 1) if np.isscalar(maxRanks) else list(maxRanks)
            _keep, pr = _sizes(ranks or [0])
        h = C.c_void_p()
        call("xt_tt_from_tensor", int(cls._is_operator), full.ndim, pd, full.ctypes.data_as(C.POINTER(C.c_double)), float(eps),
             pr, C.byref(h))
        return cls(_handle=h)

    @classmethod
    def ones(cls, dims):
        """TTNetwork::ones (ttNetwork.cpp:170-191): the all-ones tensor, rank 1, canonical at 0."""
        _, pd = _sizes(list(dims))
        h = C.c_void_p()
        call("xt_tt_ones", int(cls._is_operator), len(list(dims)), pd, C.byref(h))
        return cls(_handle=h)

    @classmethod
    def kronecker(cls, dims):
        """TTNetwork::kronecker (ttNetwork.cpp:224-254): ones where every index is equal."""
        _, pd = _sizes(list(dims))
        h = C.c_void_p()
        call("xt_tt_kronecker", int(cls._is_operator), len(list(dims)), pd, C.byref(h))
        return cls(_handle=h)

    @classmethod
    def dirac(cls, dims, position):
        """TTNetwork::dirac (ttNetwork.cpp:257-283): a single 1 at `position` (a multi-index or a linear index)."""
        dims = list(dims)
        if isinstance(position, (int, np.integer)):
            position = list(np.unravel_index(int(position), dims))
        _, pd = _sizes(dims)
        _, pp = _sizes([int(v) for v in position])
        h = C.c_void_p()
        call("xt_tt_dirac", int(cls._is_operator), len(dims), pd, pp, C.byref(h))
        return cls(_handle=h)

    def copy(self):
        h = C.c_void_p()
        call("xt_tt_copy", self._h, C.byref(h))
        return type(self)(_handle=h)

    __copy__ = copy

    def __deepcopy__(self, memo):
        return self.copy()

    # ------------------------------------------------------------------ structure
    def order(self) -> int:
        o = C.c_size_t()
        call("xt_tt_order", self._h, C.byref(o))
        return o.value

    def degree(self) -> int:
        return self.order() * (2 if self._is_operator else 1)

    @property
    def dimensions(self):
        out = np.empty(self.degree(), dtype=np.uintp)
        call("xt_tt_dims", self._h, out.ctypes.data_as(C.POINTER(C.c_size_t)))
        return [int(v) for v in out]

    def ranks(self):
        d = self.order()
        out = np.zeros(max(d - 1, 1), dtype=np.uintp)
        call("xt_tt_ranks", self._h, out.ctypes.data_as(C.POINTER(C.c_size_t)))
        return [int(v) for v in out[: d - 1]]

    def rank(self, i: int) -> int:
        return self.ranks()[i]

    def component_dims(self, k):
        out = np.empty(4 if self._is_operator else 3, dtype=np.uintp)
        call("xt_tt_component_dims", self._h, k, out.ctypes.data_as(C.POINTER(C.c_size_t)))
        return tuple(int(v) for v in out)

    def get_component(self, k: int) -> np.ndarray:
        out = np.empty(self.component_dims(k))
        call("xt_tt_get_component", self._h, k, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def set_component(self, k: int, T) -> None:
        T = np.ascontiguousarray(T, dtype=np.float64)
        _, p = _sizes(T.shape)
        call("xt_tt_set_component", self._h, k, T.ndim, p, T.ctypes.data_as(C.POINTER(C.c_double)))

    def component_ptr(self, k: int) -> int:
        p = C.c_void_p()
        call("xt_tt_component_ptr", self._h, k, C.byref(p))
        return p.value

    @property
    def canonicalized(self) -> bool:
        return self._state()[0]

    @property
    def corePosition(self) -> int:
        return self._state()[1]

    def _state(self):
        c, p = C.c_int(), C.c_size_t()
        call("xt_tt_canonicalization", self._h, C.byref(c), C.byref(p))
        return bool(c.value), p.value

    def assume_core_position(self, pos: int) -> None:
        call("xt_tt_assume_core_position", self._h, pos)

    # ------------------------------------------------------------------ TT algorithms
    def move_core(self, position: int, keepRank: bool = False) -> None:
        """TTNetwork::move_core (ttNetwork.cpp:581-628)."""
        call("xt_tt_move_core", self._h, position, int(keepRank))

    def canonicalize_left(self) -> None:
        self.move_core(0)

    def canonicalize_right(self) -> None:
        self.move_core(self.order() - 1)

    def round(self, arg=None, eps: float | None = None) -> None:
        """TTNetwork::round (ttNetwork.cpp:643-684): round(eps: float), round(maxRank: int) or
        round(maxRanks: list[, eps])."""
        d = self.order()
        if isinstance(arg, float) and eps is None:
            ranks, e = [_SIZE_MAX] * (d - 1), arg
        elif isinstance(arg, (int, np.integer)):
            if int(arg) <= 0:
                raise XttError(_lib.XT_ERR_INVALID, "MaxRank must be positive")
            ranks, e = [int(arg)] * (d - 1), EPSILON if eps is None else eps
        else:
            ranks = list(arg) if arg is not None else [_SIZE_MAX] * (d - 1)
            e = EPSILON if eps is None else eps
        _, p = _sizes(ranks or [0])
        call("xt_tt_round", self._h, p, float(e))

    def frob_norm(self) -> float:
        out = C.c_double()
        call("xt_tt_frob_norm", self._h, C.byref(out))
        return out.value

    def dot(self, other) -> float:
        out = C.c_double()
        call("xt_tt_dot", self._h, other._h, C.byref(out))
        return out.value

    def full(self) -> np.ndarray:
        dims = self.dimensions
        out = np.empty(int(np.prod(dims)))
        call("xt_tt_full", self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(dims)

    # ------------------------------------------------------------------ arithmetic
    def __add__(self, other):
        h = C.c_void_p()
        call("xt_tt_add", self._h, other._h, C.byref(h))
        return type(self)(_handle=h)

    def __sub__(self, other):
        return self + other * -1.0

    def __mul__(self, s):
        if isinstance(s, (int, float, np.floating, np.integer)):
            t = self.copy()
            call("xt_tt_scale", t._h, float(s))
            return t
        return NotImplemented

    __rmul__ = __mul__

    def __truediv__(self, s):
        return self * (1.0 / float(s))

    @staticmethod
    def reduce_to_maximal_ranks(ranks, dims, is_operator=False):
        """ttNetwork.cpp:369-402."""
        ranks = list(ranks)
        d = len(dims) // (2 if is_operator else 1)
        cur = 1
        for i in range(d - 1):
            cur *= dims[i] * (dims[d + i] if is_operator else 1)
            if cur < ranks[i]:
                ranks[i] = cur
            else:
                cur = ranks[i]
        cur = 1
        for i in range(1, d):
            cur *= dims[d - i] * (dims[2 * d - i] if is_operator else 1)
            if cur < ranks[d - i - 1]:
                ranks[d - i - 1] = cur
            else:
                cur = ranks[d - i - 1]
        return ranks


class TTTensor(_TTNetwork):
    _is_operator = False


class TTOperator(_TTNetwork):
    _is_operator = True

    @classmethod
    def identity(cls, dims):
        """TTOperator::identity (ttNetwork.cpp:194-221): δ(i_k, j_k) in every component; dims = [n…, m…]."""
        _, pd = _sizes(list(dims))
        h = C.c_void_p()
        call("xt_tt_identity", len(list(dims)), pd, C.byref(h))
        return cls(_handle=h)

    def transpose(self) -> None:
        """TTOperator::transpose (ttNetwork.h:443-448): swaps the row and column modes of every component."""
        call("xt_tt_transpose", self._h)

    def __call__(self, x: TTTensor) -> TTTensor:
        """A·x as a TT of rank r_A·r_x (ttStack.cpp:197-310), not canonicalized."""
        h = C.c_void_p()
        call("xt_tt_apply", self._h, x._h, C.byref(h))
        return TTTensor(_handle=h)


class FileFormat:
    """misc::FileFormat (misc/fileIO.h:41): BINARY = exact storage, TSV = tab-separated text."""
    BINARY = 0
    TSV = 1


def save_to_file(obj, filename: str, format: int = FileFormat.BINARY) -> None:
    """misc::save_to_file for a TTTensor / TTOperator (python/misc.cpp:84-90): a xerus datafile, byte compatible
    with xerus's own writer (misc/fileIO.h:105-121)."""
    if not isinstance(obj, _TTNetwork):
        raise TypeError("save_to_file: only TTTensor and TTOperator are stored by this build")
    call("xt_tt_save", obj._h, os.fsencode(filename), int(format))


def load_from_file(filename: str):
    """misc::load_from_file (python/misc.cpp:92-116): the TTTensor or TTOperator a xerus datafile holds."""
    h = C.c_void_p()
    call("xt_tt_load", os.fsencode(filename), C.byref(h))
    is_op = C.c_int()
    call("xt_tt_is_operator", h, C.byref(is_op))
    return TTOperator(_handle=h) if is_op.value else TTTensor(_handle=h)


def frob_norm(t) -> float:
    return t.frob_norm()


def solve(A, b, return_method: bool = False):
    """xerus::solve on dense matrices (tensor.cpp:1654-1704 → blasWrapper::solve, blasLapackWrapper.cpp:542-651), on
    the device: X with A·X = b for A (m × n) and b (m,) or (m, nrhs). xerus's dispatch picks minimum-norm least
    squares for m ≠ n, LU for a non-symmetric A, Cholesky for a symmetric A with a definite diagonal (falling
    back on failure) and a symmetric indefinite solve otherwise. return_method also returns the method taken:
    0 LU, 1 Cholesky, 2 symmetric indefinite, 3 least squares."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if A.ndim != 2 or b.shape[0] != A.shape[0]:
        raise XttError("solve: A must be m × n and b must have m rows")
    m, n = A.shape
    B = np.ascontiguousarray(b.reshape(m, -1))
    nrhs = B.shape[1]
    dA, dB, dX = _lib.DeviceArray.from_numpy(A), _lib.DeviceArray.from_numpy(B), _lib.DeviceArray((n, nrhs), np.float64)
    method = C.c_int()
    call("xt_solve", dA.ptr, m, n, dB.ptr, nrhs, dX.ptr, C.byref(method))
    X = dX.numpy()
    X = X.reshape(n) if b.ndim == 1 else X
    return (X, method.value) if return_method else X


def residual(A: TTOperator, x: TTTensor, b: TTTensor) -> float:
    """‖A·x − b‖/‖b‖ computed stably on the device (TT of A·x − b, Householder R sweep)."""
    out = C.c_double()
    call("xt_tt_residual", A._h, x._h, b._h, C.byref(out))
    return out.value


# ---------------------------------------------------------------------- ALS


# ---------------------------------------------------------------------- PerformanceData


class PerformanceData:
    """PerformanceData (performanceData.h:43-157): the optional trace argument of ALSVariant and ADFVariant.

    ``data`` collects DataPoint(iterationCount, elapsedTime [µs], residual, error, ranks, flags) as xerus's
    PerformanceData::add does (performanceData.cpp:34-80); ``errorFunction(x)``, when given, is evaluated at every
    point with the clock paused. Points come from the library through a callback on the calling thread.
    """

    class DataPoint(tuple):
        __slots__ = ()
        _names = ("iterationCount", "elapsedTime", "residual", "error", "ranks", "flags")

        def __new__(cls, *vals):
            return tuple.__new__(cls, vals)

        def __getattr__(self, name):
            try:
                return self[self._names.index(name)]
            except ValueError:
                raise AttributeError(name) from None

    def __init__(self, printProgress=False, active=True, errorFunction=None):
        self.printProgress = printProgress
        self.active = active
        self.errorFunction = errorFunction
        self.data = []
        self.additionalInformation = ""

    def __bool__(self):
        return self.active

    def reset(self):
        self.data = []
        self.additionalInformation = ""

    def _run(self, fn, x):
        """Run fn() with this object as the thread's sink; x is the tensor the error function is evaluated on."""
        if not self.active:
            return fn()

        def cb(_user, it, us, res, ranks, nranks, flags):
            err = 0.0
            if self.errorFunction is not None:
                err = float(self.errorFunction(x))
            rk = [ranks[i] for i in range(nranks)]
            self.data.append(PerformanceData.DataPoint(int(it), float(us), float(res), err, rk, int(flags)))
            if self.printProgress:
                print(f"Iteration {it:4d} Time: {us * 1e-6:6.2f}s Residual: {res:11.6e} Flags: {flags} Ranks: {rk}")

        hook = _lib.PERF_CALLBACK(cb)
        call("xt_perf_sink", hook, None)
        try:
            return fn()
        finally:
            call("xt_perf_sink", _lib.PERF_CALLBACK(), None)

    def dump_to_file(self, path):
        """One line per point: iteration, time (µs), residual, error, flags, ranks (whitespace separated)."""
        with open(path, "w") as f:
            f.write(self.additionalInformation)
            for p in self.data:
                f.write(f"{p.iterationCount} {p.elapsedTime:.0f} {p.residual!r} {p.error!r} {p.flags} {' '.join(map(str, p.ranks))}\n")


class ALSVariant:
    """ALSVariant (als.h:37-223). Extra fields choose the local solve on the device: ``precision``
    (XT_F64 / XT_F32 / XT_BF16 local matvecs, the latter two with FP64 refinement), ``localTolerance``,
    ``maxLocalIterations``, ``innerTolerance``, ``maxRefinements``. A PerformanceData passed as the fourth
    argument (or ``perfData=``) records the run."""

    FLAG_FINISHED_HALFSWEEP = 1  # als.h:109-110
    FLAG_FINISHED_FULLSWEEP = 3

    def __init__(self, sites: int, numHalfSweeps: int, assumeSPD: bool, useResidual: bool = False):
        if sites <= 0:
            raise XttError(_lib.XT_ERR_INVALID, "sites must be > 0")
        o = _lib.XtALSOptions()
        call("xt_als_default_options", C.byref(o), sites, int(assumeSPD))
        self.sites = sites
        self.numHalfSweeps = numHalfSweeps
        self.convergenceEpsilon = o.convergenceEpsilon
        self.useResidualForEndCriterion = useResidual
        self.preserveCorePosition = True
        self.assumeSPD = assumeSPD
        self.precision = XT_F64
        self.localTolerance = o.localTolerance
        self.maxLocalIterations = o.maxLocalIterations
        self.innerTolerance = 0.0
        self.maxRefinements = o.maxRefinements
        self.last_stats = None

    def copy(self):
        return _copy.copy(self)

    def _options(self):
        o = _lib.XtALSOptions()
        o.sites = self.sites
        o.numHalfSweeps = self.numHalfSweeps
        o.convergenceEpsilon = self.convergenceEpsilon
        o.useResidualForEndCriterion = int(self.useResidualForEndCriterion)
        o.preserveCorePosition = int(self.preserveCorePosition)
        o.assumeSPD = int(self.assumeSPD)
        o.precision = self.precision
        o.localTolerance = self.localTolerance
        o.maxLocalIterations = self.maxLocalIterations
        o.innerTolerance = self.innerTolerance
        o.maxRefinements = self.maxRefinements
        return o

    def __call__(self, *args, **kw):
        """(A, x, b[, epsilon | numHalfSweeps]) or (x, b[, epsilon | numHalfSweeps]); x is updated in place,
        the energy is returned (als.h:152-207)."""
        args = list(args)
        A = args.pop(0) if isinstance(args[0], TTOperator) else None
        x, b = args[0], args[1]
        o = self._options()
        extra = args[2] if len(args) > 2 else kw.get("epsilon", kw.get("numHalfSweeps"))
        if "numHalfSweeps" in kw:
            o.numHalfSweeps = int(kw["numHalfSweeps"])
        elif "epsilon" in kw:
            o.convergenceEpsilon = float(kw["epsilon"])
        elif extra is not None:
            if isinstance(extra, (int, np.integer)) and not isinstance(extra, bool):
                o.numHalfSweeps = int(extra)
            else:
                o.convergenceEpsilon = float(extra)
        energy = C.c_double()
        st = _lib.XtALSStats()
        perf = kw.get("perfData", args[3] if len(args) > 3 else None)
        run = lambda: call("xt_als_solve", C.byref(o), A._h if A is not None else None, x._h, b._h, C.byref(energy),
                           C.byref(st))
        if perf is not None:
            perf._run(run, x)
        else:
            run()
        self.last_stats = {f: getattr(st, f) for f, _ in _lib.XtALSStats._fields_}
        return energy.value


def als_round_batch(variant: "ALSVariant", As, xs, bs, sweeps: int, eps: float, threads: int = 8, residuals: bool = True):
    """Many independent problems: problem p runs ``sweeps`` × (variant(As[p], xs[p], bs[p]), xs[p].round(eps)) on
    one of ``threads`` worker threads, each with its own stream (xs updated in place). ``threads=0`` runs all problems
    in lockstep instead: every step of every problem is one grouped launch (one-site ALS_SPD with an operator and
    XT_F64 or XT_F32 local solves; other options fall back to worker threads). Returns the relative residuals ‖A x − b‖/‖b‖ (or
    None) and the summed solver counters."""
    n = len(xs)
    if not (len(As) == n and len(bs) == n):
        raise ValueError("As, xs and bs must have the same length")
    o = variant._options()
    A_arr = (C.c_void_p * n)(*[a._h if a is not None else None for a in As])
    x_arr = (C.c_void_p * n)(*[x._h for x in xs])
    b_arr = (C.c_void_p * n)(*[b._h for b in bs])
    res = np.zeros(n) if residuals else None
    st = _lib.XtALSStats()
    call("xt_als_round_batch", C.byref(o), n, A_arr, x_arr, b_arr, sweeps, eps, threads,
         res.ctypes.data_as(C.POINTER(C.c_double)) if residuals else None, C.byref(st))
    return res, {f: getattr(st, f) for f, _ in _lib.XtALSStats._fields_}


ALS = ALSVariant(1, 0, False)
ALS_SPD = ALSVariant(1, 0, True)
DMRG = ALSVariant(2, 0, False)
DMRG_SPD = ALSVariant(2, 0, True)


# ---------------------------------------------------------------------- measurements and ADF


class SinglePointMeasurementSet:
    """SinglePointMeasurementSet (measurments.h:43-91): sorted unique positions and measured values."""

    def __init__(self, _handle=None):
        self._h = _handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib._lib is not None:
            _lib._lib.xt_meas_free(h)
            self._h = None

    @classmethod
    def random(cls, numMeasurements: int, arg):
        """random(P, dims) draws positions with the shared engine (measurments.cpp:222-248);
        random(P, solution) additionally measures the TT `solution`."""
        dims = arg.dimensions if isinstance(arg, _TTNetwork) else list(arg)
        _, p = _sizes(dims)
        h = C.c_void_p()
        call("xt_meas_random", numMeasurements, len(dims), p, C.byref(h))
        m = cls(h)
        if isinstance(arg, _TTNetwork):
            m.measure(arg)
        return m

    @classmethod
    def from_arrays(cls, positions, values):
        pos = np.ascontiguousarray(positions, dtype=np.uint32)
        val = np.ascontiguousarray(values, dtype=np.float64)
        h = C.c_void_p()
        call("xt_meas_new", pos.shape[0], pos.shape[1], pos.ctypes.data_as(C.POINTER(C.c_uint32)),
             val.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h))
        return cls(h)

    def size(self) -> int:
        n = C.c_size_t()
        call("xt_meas_size", self._h, C.byref(n))
        return n.value

    __len__ = size

    def positions(self, order: int) -> np.ndarray:
        out = np.empty((self.size(), order), dtype=np.uint32)
        call("xt_meas_positions", self._h, out.ctypes.data_as(C.POINTER(C.c_uint32)))
        return out

    def positions_device_ptr(self) -> int:
        p = C.c_void_p()
        call("xt_meas_positions_ptr", self._h, C.byref(p))
        return p.value

    def stack_plan(self, order: int) -> dict:
        """ADF's stack plan of these positions (adf.cpp:102-191), built on the device: per level k the representative
        sample ids of the forward entries (shared positions 0..k) and backward entries (shared k..d-1), ascending,
        and every sample's entry per level (slots[k, i])."""
        ps = C.POINTER(C.c_size_t)
        pu = C.POINTER(C.c_uint32)
        fc, bc = np.zeros(order, dtype=np.uintp), np.zeros(order, dtype=np.uintp)
        nu = C.c_size_t()
        call("xt_meas_stack_plan", self._h, fc.ctypes.data_as(ps), bc.ctypes.data_as(ps), None, None, None, None, C.byref(nu))
        P = self.size()
        fi, bi = np.empty(int(fc.sum()), dtype=np.uint32), np.empty(int(bc.sum()), dtype=np.uint32)
        fs, bs = np.empty((order, P), dtype=np.uint32), np.empty((order, P), dtype=np.uint32)
        call("xt_meas_stack_plan", self._h, fc.ctypes.data_as(ps), bc.ctypes.data_as(ps), fi.ctypes.data_as(pu), bi.ctypes.data_as(pu),
             fs.ctypes.data_as(pu), bs.ctypes.data_as(pu), C.byref(nu))
        return {"fwd_counts": fc.astype(int), "bwd_counts": bc.astype(int), "fwd_ids": fi, "bwd_ids": bi, "fwd_slots": fs,
                "bwd_slots": bs, "num_unique": nu.value}

    @property
    def measuredValues(self) -> np.ndarray:
        out = np.empty(self.size())
        call("xt_meas_values", self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    @measuredValues.setter
    def measuredValues(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        call("xt_meas_set_values", self._h, v.ctypes.data_as(C.POINTER(C.c_double)))

    def measure(self, solution: TTTensor) -> None:
        call("xt_meas_measure", self._h, solution._h)

    def frob_norm(self) -> float:
        return float(np.linalg.norm(self.measuredValues))


class RankOneMeasurementSet:
    """RankOneMeasurementSet (measurments.h:94-150, python/recovery.cpp:84-124): measurement i is the rank-one
    tensor u_{i,0} ⊗ … ⊗ u_{i,d-1} with measured value b_i. ``RankOneMeasurementSet(sp, dims)`` converts a
    SinglePointMeasurementSet (unit vectors, measurments.cpp:253-265)."""

    def __init__(self, other=None, dims=None, *, _handle=None, _dims=None):
        self._h = None
        if _handle is not None:
            self._h, self._dims = _handle, list(_dims)
            return
        if isinstance(other, RankOneMeasurementSet):  # copy
            self._dims = list(other._dims)
            self._build(other.vectors(), other.measuredValues)
            return
        if isinstance(other, SinglePointMeasurementSet):
            if dims is None:
                raise ValueError("RankOneMeasurementSet(SinglePointMeasurementSet, dims): dims are required")
            _, p = _sizes(dims)
            h = C.c_void_p()
            call("xt_r1meas_from_single_point", other._h, len(dims), p, C.byref(h))
            self._h, self._dims = h, list(dims)
            return
        self._dims = list(dims) if dims is not None else []
        self._build(np.zeros((0, sum(self._dims))), np.zeros(0))

    def _build(self, vectors, values):
        vec = np.ascontiguousarray(vectors, dtype=np.float64)
        val = np.ascontiguousarray(values, dtype=np.float64)
        _, p = _sizes(self._dims)
        h = C.c_void_p()
        call("xt_r1meas_new", val.size, len(self._dims), p, vec.ctypes.data_as(C.POINTER(C.c_double)),
             val.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h))
        old, self._h = self._h, h
        if old and _lib._lib is not None:
            _lib._lib.xt_r1meas_free(old)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib._lib is not None:
            _lib._lib.xt_r1meas_free(h)
            self._h = None

    @classmethod
    def random(cls, numMeasurements: int, arg):
        """random(P, dims): every vector drawn N(0,1) from the shared engine, then sorted (measurments.cpp:513-536);
        random(P, solution) additionally measures the TT `solution`."""
        dims = arg.dimensions if isinstance(arg, _TTNetwork) else list(arg)
        _, p = _sizes(dims)
        h = C.c_void_p()
        call("xt_r1meas_random", numMeasurements, len(dims), p, C.byref(h))
        m = cls(_handle=h, _dims=dims)
        if isinstance(arg, _TTNetwork):
            m.measure(arg)
        return m

    def size(self) -> int:
        n = C.c_size_t()
        call("xt_r1meas_size", self._h, C.byref(n))
        return n.value

    __len__ = size

    def degree(self) -> int:
        return len(self._dims) if self.size() else 0

    def vectors(self) -> np.ndarray:
        """All positions as a P × Σn_k array (mode k's vector at column offset Σ_{j<k} n_j)."""
        out = np.empty((self.size(), sum(self._dims)))
        call("xt_r1meas_vectors", self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def get_position(self, i: int):
        row = self.vectors()[i]
        offs = np.cumsum([0] + self._dims)
        return [row[offs[k]:offs[k + 1]].copy() for k in range(len(self._dims))]

    def set_position(self, i: int, pos) -> None:
        vec = self.vectors()
        vec[i] = np.concatenate([np.asarray(v, dtype=np.float64).ravel() for v in pos])
        self._build(vec, self.measuredValues)

    def get_measuredValue(self, i: int) -> float:
        return float(self.measuredValues[i])

    def set_measuredValue(self, i: int, val: float) -> None:
        v = self.measuredValues
        v[i] = val
        self.measuredValues = v

    def add(self, position, measuredValue: float) -> None:
        """RankOneMeasurementSet::add (measurments.cpp:309-326)."""
        pos = [np.asarray(v, dtype=np.float64).ravel() for v in position]
        if self.size() == 0 and not self._dims:
            self._dims = [len(v) for v in pos]
        if [len(v) for v in pos] != self._dims:
            raise XttError(1, "Inconsitend dimensions obtained.")
        vec = np.vstack([self.vectors(), np.concatenate(pos)[None, :]])
        self._build(vec, np.append(self.measuredValues, measuredValue))

    @property
    def measuredValues(self) -> np.ndarray:
        out = np.empty(self.size())
        call("xt_r1meas_values", self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    @measuredValues.setter
    def measuredValues(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        call("xt_r1meas_set_values", self._h, v.ctypes.data_as(C.POINTER(C.c_double)))

    def sort(self, positionsOnly: bool = False) -> None:
        call("xt_r1meas_sort", self._h, int(positionsOnly))

    def normalize(self) -> None:
        call("xt_r1meas_normalize", self._h)

    def measure(self, solution: TTTensor) -> None:
        call("xt_r1meas_measure", self._h, solution._h)

    def test(self, solution: TTTensor) -> float:
        out = C.c_double()
        call("xt_r1meas_test", self._h, solution._h, C.byref(out))
        return out.value

    def frob_norm(self) -> float:
        return float(np.linalg.norm(self.measuredValues))


def adf_gradient(x: TTTensor, measurements, precision: int = XT_F64):
    """xerus's projected gradients Δ_k = Σ_i (b_i − x_i)·∂x_i/∂G_k for every core at x (adf.cpp:314-396),
    computed by the device adjoint kernel, for a SinglePointMeasurementSet or a RankOneMeasurementSet.
    Returns (list of arrays shaped like the cores, ‖P_Ω(x) − b‖/‖b‖)."""
    shapes = [x.component_dims(k) for k in range(x.order())]
    out = np.empty(sum(int(np.prod(s)) for s in shapes))
    rn = C.c_double()
    fn = "xt_adf_gradient_r1" if isinstance(measurements, RankOneMeasurementSet) else "xt_adf_gradient"
    call(fn, x._h, measurements._h, precision, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(rn))
    grads, o = [], 0
    for s in shapes:
        n = int(np.prod(s))
        grads.append(out[o: o + n].reshape(s))
        o += n
    return grads, rn.value


class ADFVariant:
    """ADFVariant (adf.h:230-266): ``maxIterations``, ``targetResidualNorm``, ``minimalResidualNormDecrease``."""

    def __init__(self, maxIteration: int, targetResidual: float, minimalResidualDecrease: float):
        self.maxIterations = maxIteration
        self.targetResidualNorm = targetResidual
        self.minimalResidualNormDecrease = minimalResidualDecrease
        self.precision = XT_F64
        self.last_iterations = None

    def __call__(self, x: TTTensor, measurements, maxRanks=None, perfData=None) -> float:
        """Returns ‖P_Ω(x) − b‖/‖b‖; x is updated in place (adf.h:248-266). `measurements` is a
        SinglePointMeasurementSet or a RankOneMeasurementSet."""
        if isinstance(maxRanks, PerformanceData):
            maxRanks, perfData = None, maxRanks
        p = None
        if maxRanks is not None:
            _arr, p = _sizes(maxRanks)
        res, it = C.c_double(), C.c_size_t()
        fn = "xt_adf_solve_r1" if isinstance(measurements, RankOneMeasurementSet) else "xt_adf_solve"
        run = lambda: call(fn, x._h, measurements._h, p, self.maxIterations, self.targetResidualNorm,
                           self.minimalResidualNormDecrease, self.precision, C.byref(res), C.byref(it))
        if perfData is not None:
            perfData._run(run, x)
        else:
            run()
        self.last_iterations = it.value
        return res.value


ADF = ADFVariant(0, 1e-8, 0.999)  # adf.cpp:610
