"""GEMM tile autotuning for libxtt's contractions.

Every contraction of the TT path (stack pushes, local operators, R·core products, TT products) is one libxtt GEMM
launch whose tile edge (32 or 64) and split-K slice count come from a heuristic. This module measures instead:

- ``collect_shapes(run)`` runs a workload once with the launch profiler and shape tags on and returns the solo GEMM
  shapes it launched (paired launches have a fixed layout and are not tuned), with launch counts;
- ``tune(shapes)`` times the heuristic's choice and every candidate per shape on the device (``xt_gemm_tune``),
  installs the fastest and returns one row per shape;
- ``save(rows, path)`` / ``load(path)`` write and read the measured table (tab-separated, one shape per line).

``xerus_amd.lib()`` loads ``tables/gemm_gfx950.tsv`` beside this file when it exists (unless XTT_NO_TUNING is set),
so the committed table applies to every process. Shape tags need ``XTT_PROFILE_SHAPES=1`` in the environment before libxtt is loaded.
"""
from __future__ import annotations

import ctypes as C
import os
import re

from . import _lib

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables", "gemm_gfx950.tsv")
COLUMNS = ["precision", "m", "n", "k", "batch", "ta", "tb", "launches", "tile", "slices", "us", "heuristic_tile",
           "heuristic_slices", "heuristic_us"]
_TAG = re.compile(r"^gemm_x4_kernel<(double|float),\d+,\d+> m=(\d+) n=(\d+) k=(\d+) batch=(\d+) t=(\d)(\d)( splitk)?$")


def collect_shapes(run) -> list:
    """Solo GEMM shapes launched by ``run()``: [(precision, m, n, k, batch, ta, tb, launches)], most launched first."""
    if not os.environ.get("XTT_PROFILE_SHAPES"):
        raise RuntimeError("set XTT_PROFILE_SHAPES=1 before libxtt is loaded to collect GEMM shapes")
    _lib.profile_enable(True)
    _lib.profile_read()
    try:
        run()
        rec = _lib.profile_read(1 << 15)
    finally:
        _lib.profile_enable(False)
    shapes = {}
    for tag, (launches, _ms, _fl, _by) in rec.items():
        mt = _TAG.match(tag)
        if not mt:
            continue
        key = (_lib.XT_F64 if mt.group(1) == "double" else _lib.XT_F32,) + tuple(int(mt.group(i)) for i in range(2, 8))
        shapes[key] = shapes.get(key, 0) + launches
    return sorted(((*k, c) for k, c in shapes.items()), key=lambda r: -r[-1])


def tune(shapes, iters: int = 50) -> list:
    """Times and installs the fastest choice per shape; returns rows in COLUMNS order."""
    rows = []
    for prec, m, n, k, batch, ta, tb, launches in shapes:
        t, s, hs, ht = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        us, hus = C.c_double(), C.c_double()
        _lib.call("xt_gemm_tune", prec, m, n, k, batch, ta, tb, iters, C.byref(t), C.byref(s), C.byref(us), C.byref(ht), C.byref(hs),
                  C.byref(hus))
        rows.append([prec, m, n, k, batch, ta, tb, launches, t.value, s.value, round(us.value, 3), ht.value, hs.value,
                     round(hus.value, 3)])
    return rows


def save(rows, path: str = DEFAULT_TABLE, header: str = "") -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        for line in header.splitlines():
            f.write(f"# {line}\n")
        f.write("\t".join(COLUMNS) + "\n")
        for r in rows:
            f.write("\t".join(str(v) for v in r) + "\n")


def read(path: str = DEFAULT_TABLE) -> list:
    rows = []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or line.startswith(COLUMNS[0]) or not line.strip():
                continue
            v = line.split("\t")
            rows.append([int(x) for x in v[:10]] + [float(v[10]), int(v[11]), int(v[12]), float(v[13])])
    return rows


def load(path: str = DEFAULT_TABLE) -> int:
    """Installs the table's choices (rows whose choice differs from the heuristic's) through xt_gemm_tune_load;
    returns how many."""
    n = C.c_size_t()
    _lib.call("xt_gemm_tune_load", path.encode(), C.byref(n))
    return n.value
