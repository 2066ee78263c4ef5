"""CPU checks of the C ABI: libxtt.so loads, exports every function declared in include/xtt.h, and its
host-side pieces (argument validation, option defaults, the shared random engine) behave like xerus's —
none of these touch the GPU."""
import ctypes as C
import os
import re

import numpy as np

from helpers import SEED

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "xtt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xt_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported(xe):
    from xerus_amd import _lib
    names = declared_functions()
    assert len(names) >= 50
    L = xe.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with include/xtt.h"


def test_null_arguments_rejected(xe):
    L = xe.lib()
    assert L.xt_tt_new(0, 3, None, None) == xe._lib.XT_ERR_INVALID
    assert b"must not be NULL" in L.xt_last_error()
    assert L.xt_tt_order(None, None) == xe._lib.XT_ERR_INVALID
    assert L.xt_als_default_options(None, 1, 1) == xe._lib.XT_ERR_INVALID


def test_als_presets_match_reference(xe):
    """ALS/ALS_SPD sites 1, DMRG/DMRG_SPD sites 2, convergenceEpsilon 1e-6, preserveCorePosition true
    (als.h:134-141, als.cpp:781-788); ADF = ADFVariant(0, 1e-8, 0.999) (adf.cpp:610)."""
    assert (xe.ALS.sites, xe.ALS.assumeSPD) == (1, False)
    assert (xe.ALS_SPD.sites, xe.ALS_SPD.assumeSPD) == (1, True)
    assert (xe.DMRG.sites, xe.DMRG.assumeSPD) == (2, False)
    assert (xe.DMRG_SPD.sites, xe.DMRG_SPD.assumeSPD) == (2, True)
    for v in (xe.ALS, xe.ALS_SPD, xe.DMRG, xe.DMRG_SPD):
        assert v.convergenceEpsilon == 1e-6 and v.preserveCorePosition and v.numHalfSweeps == 0
        assert not v.useResidualForEndCriterion
    assert (xe.ADF.maxIterations, xe.ADF.targetResidualNorm, xe.ADF.minimalResidualNormDecrease) == (0, 1e-8, 0.999)


def test_engine_matches_oracle(xe, oracle):
    """libxtt's host engine and the oracle draw identical streams (same mt19937_64 and libstdc++ distributions)."""
    xe.seed(SEED)
    oracle.seed(SEED)
    assert np.array_equal(xe.normal_draws(500), oracle.normal_draws(500))
    assert np.array_equal(xe.uniform_draws(500, 15), oracle.uniform_draws(500, 15))


def test_reduce_to_maximal_ranks(xe):
    f = xe.TTTensor.reduce_to_maximal_ranks
    assert f([8] * 7, [4] * 8) == [4, 8, 8, 8, 8, 8, 4]
    assert f([64] * 15, [16] * 16) == [16] + [64] * 13 + [16]
    assert f([100, 100], [2, 3, 4]) == [2, 4]


def test_gpus_flag_refuses_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_hip_runtime_selection():
    """libxtt binds to the HIP runtime mapped when it loads: torch's bundled copy when torch came first (then torch's
    device tensors are libxtt pointers), the ROCm installation's when libxtt came first."""
    import subprocess
    import sys
    torch_first = ("import sys; sys.path.insert(0, %r); import torch\n"
                   "import xerus_amd; from xerus_amd import _lib; xerus_amd.lib()\n"
                   "maps = open('/proc/self/maps').read()\n"
                   "print('torch-runtime', _lib.on_torch_runtime(), 'torch/lib/libamdhip64' in maps)\n") % ROOT
    out = subprocess.run([sys.executable, "-c", torch_first], capture_output=True, text=True, timeout=300)
    assert "torch-runtime True True" in out.stdout, out.stdout + out.stderr
    xtt_first = ("import sys; sys.path.insert(0, %r); import xerus_amd; xerus_amd.lib(); import torch\n"
                 "from xerus_amd import _lib\n"
                 "maps = open('/proc/self/maps').read()\n"
                 "print('ok', '/opt/rocm' in maps and 'libxtt' in maps, _lib.on_torch_runtime())\n") % ROOT
    out = subprocess.run([sys.executable, "-c", xtt_first], capture_output=True, text=True, timeout=300)
    assert "ok True False" in out.stdout, out.stdout + out.stderr


def test_torch_binding_rejects_inconsistent_ranks():
    """The sampled-entry kernels take every core offset from the rank chain, so the torch binding refuses cores
    whose neighbouring ranks disagree (cores (1,4,3),(2,4,1) would size the gradient 24 floats for 20) before any
    device buffer is made."""
    import subprocess
    import sys
    prog = ("import sys; sys.path.insert(0, %r)\n"
            "from xerus_amd.autograd import tt_sample\nimport torch\n"
            "pos = torch.zeros((3, 2), dtype=torch.int32)\n"
            "for shapes in ([(1, 4, 3), (2, 4, 1)], [(2, 4, 3), (3, 4, 1)], [(1, 4, 3), (3, 4, 2)]):\n"
            "    try:\n        tt_sample(pos, [torch.zeros(s) for s in shapes]); print('accepted', shapes)\n"
            "    except ValueError as e:\n        print('refused', 'rank' in str(e))\n") % ROOT
    out = subprocess.run([sys.executable, "-c", prog], capture_output=True, text=True, timeout=300)
    assert out.stdout.split("\n")[:3] == ["refused True"] * 3, out.stdout + out.stderr


def test_cpp_faces_compile(tmp_path):
    """The C++ face with xerus's names (include/xtt_xerus.hpp) and the blasWrapper shim shown in INTEGRATION.md compile
    against include/xtt.h (syntax and types; linking and running are the GPU test's)."""
    import subprocess
    inc = os.path.join(ROOT, "include")
    src = tmp_path / "face.cpp"
    src.write_text('#include "xtt_xerus.hpp"\nint main() { xerus::ALSVariant v = xerus::DMRG_SPD; return (int)v.sites - 2; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-I", inc, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    assert blocks
    for i, b in enumerate(blocks):
        f = tmp_path / f"shim{i}.cpp"
        f.write_text(b)
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", inc, str(f)], capture_output=True, text=True)
        assert r.returncode == 0, f"INTEGRATION.md C++ block {i}:\n{r.stderr}"


def test_gemm_tile_table_loads(xe):
    """The committed measured GEMM tile table parses, every choice is one the kernels support, a choice that differs
    from the heuristic's was measured faster, and xt_gemm_tune_load installs exactly those rows (host only)."""
    import ctypes
    from xerus_amd import tuning
    rows = tuning.read()
    assert len(rows) > 100
    changed = 0
    for r in rows:
        prec, m, n, k, batch, ta, tb, launches, tile, slices, us, htile, hslices, hus = r
        assert prec in (0, 1) and min(m, n, k, batch, launches) > 0 and ta in (0, 1) and tb in (0, 1)
        assert tile in (32, 64) and (slices == 0 or (tile == 32 and 2 <= slices <= 256))
        if (tile, slices) != (htile, hslices):
            changed += 1
            assert us < hus
    n_inst = ctypes.c_size_t()
    xe._lib.call("xt_gemm_tune_load", tuning.DEFAULT_TABLE.encode(), ctypes.byref(n_inst))
    assert n_inst.value == changed > 0
