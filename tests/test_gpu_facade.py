"""GPU check of the C++ face with xerus's names (include/xtt_xerus.hpp): tests/cpp/facade_check, compiled against
libxtt by tests/cpp/Makefile, runs config 1 as xerus user code would write it (TTOperator from set_component, random
TTTensors from the seeded engine, ALS_SPD with a PerformanceData, round, sum, the Tensor constructor, ones, ADF and
a refused malformed TT). Its energy, residual and ranks are compared with the oracle's config-1 run."""
import os
import subprocess

import numpy as np
import pytest

from helpers import SEED, laplace_cores

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_check")


def test_cpp_facade_config1(xe, oracle):
    print("C++ facade config 1", flush=True)
    assert os.path.exists(BIN), "tests/cpp/facade_check is built by __graft_entry__.build()"
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stderr
    vals = {}
    for line in out.stdout.splitlines():
        key, *rest = line.split()
        vals[key] = rest
    d, n = 8, 4
    Ao = oracle.TT.from_cores(laplace_cores(d, n), op=True)
    oracle.seed(SEED)
    bo = oracle.TT.random([n] * d, [2] * (d - 1))
    bo.scale(1.0 / bo.frob_norm())
    xo = oracle.TT.random([n] * d, [8] * (d - 1))
    eo, _ = oracle.als(Ao, xo, bo, num_half_sweeps=10, eps=0.0)
    assert abs(float(vals["energy"][0]) - eo) < 1e-12 * abs(eo)
    assert abs(float(vals["residual"][0]) - oracle.residual(Ao, xo, bo)) < 1e-6 * oracle.residual(Ao, xo, bo)
    assert [int(v) for v in vals["ranks"]] == xo.ranks()
    assert vals["core"] == [str(xo.state()[1]), "1"]
    assert int(vals["perf_points"][0]) > 10
    assert [int(v) for v in vals["sum_ranks"]] == xo.ranks()  # x + x rounded back to x's ranks
    assert abs(float(vals["sum_norm_ratio"][0]) - 2.0) < 1e-12
    assert vals["from_tensor_ranks"] == ["1", "1"] and float(vals["ones_diff"][0]) < 1e-12
    assert float(vals["adf_residual"][0]) < 1e-3
    assert vals["format_check"] == ["ok"]
