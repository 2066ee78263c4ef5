"""xerus datafiles for TTTensor / TTOperator (misc::save_to_file / load_from_file).

The oracle (oracle/fileio.py) restates the stream writers; it is pinned here against hand-derived files
(the literal TSV text below follows ttNetwork.cpp:57-108 and :1454-1470, tensorNetwork.cpp:1429-1469 and
tensor.cpp:1781-1806 field by field). libxtt must then write the oracle's bytes exactly (binary and TSV) and
read every file the oracle writes, including sparse-component files, which libxtt itself never writes.
"""
import os
import struct

import numpy as np
import pytest

from oracle import fileio as F

SMAX = "18446744073709551615"

# TTTensor dims [2, 3], ranks [1]: component 0 = [1.5, -2], component 1 = [0.25, 3, 1e-300], core at 0.
LITERAL_TT_TSV = (
    "Xerus xerus::TTNetwork<false> datafile.\nFormat: TSV\n"
    "1\t1\t0\t"                                   # TTNetwork version, canonicalized, corePosition
    "1\t2\t2\t3\t\n"                              # TensorNetwork version, dimensions {2, 3}
    "1\t1\t2\t2\t1\t3\t\n\n"                      # external links (node, index position, dimension)
    "4\t\n"                                       # node count
    "1\t0\t1\t0\t1\t"                             # left virtual node: one link to node 1, index 0
    "3\t0\t0\t0\t1\t1\t" + SMAX + "\t0\t2\t0\t2\t0\t1\t"   # component 0: [left, mode 0, right]
    "3\t0\t1\t2\t1\t1\t" + SMAX + "\t1\t3\t0\t3\t0\t1\t"   # component 1
    "1\t0\t2\t2\t1\t\n"                           # right virtual node: link to node 2, index N+1 = 2
    "1\t1\t1\t1\t1\t\n"                           # ones({1})
    "1\t3\t1\t2\t1\t1\t1.5\t-2\t\n"
    "1\t3\t1\t3\t1\t1\t0.25\t3\t1e-300\t\n"
    "1\t1\t1\t1\t1\t\n"
)

# TTOperator dims [2, 1] x [1, 2] (d = 2), ranks [1], not canonicalized: components (1,2,1,1) and (1,1,2,1).
LITERAL_OP_TSV = (
    "Xerus xerus::TTNetwork<true> datafile.\nFormat: TSV\n"
    "1\t0\t0\t"
    "1\t4\t2\t1\t1\t2\t\n"
    "1\t1\t2\t2\t1\t1\t1\t2\t1\t2\t2\t2\t\n\n"
    "4\t\n"
    "1\t0\t1\t0\t1\t"
    "4\t0\t0\t0\t1\t1\t" + SMAX + "\t0\t2\t1\t" + SMAX + "\t2\t1\t0\t2\t0\t1\t"
    "4\t0\t1\t3\t1\t1\t" + SMAX + "\t1\t1\t1\t" + SMAX + "\t3\t2\t0\t3\t0\t1\t"
    "1\t0\t2\t3\t1\t\n"
    "1\t1\t1\t1\t1\t\n"
    "1\t4\t1\t2\t1\t1\t1\t0.5\t-7\t\n"
    "1\t4\t1\t1\t2\t1\t1\t2\t4\t\n"
    "1\t1\t1\t1\t1\t\n"
)


def _lit_tt():
    return [np.array([1.5, -2.0]).reshape(1, 2, 1), np.array([0.25, 3.0, 1e-300]).reshape(1, 3, 1)], [2, 3]


def _lit_op():
    return [np.array([0.5, -7.0]).reshape(1, 2, 1, 1), np.array([2.0, 4.0]).reshape(1, 1, 2, 1)], [2, 1, 1, 2]


def test_oracle_matches_literal_files():
    cores, dims = _lit_tt()
    assert F.write_tt(cores, dims, False, True, 0, tsv=True).decode() == LITERAL_TT_TSV
    cores, dims = _lit_op()
    assert F.write_tt(cores, dims, True, False, 0, tsv=True).decode() == LITERAL_OP_TSV


def test_oracle_binary_layout():
    cores, dims = _lit_tt()
    b = F.write_tt(cores, dims, False, True, 1, tsv=False)
    head = b"Xerus xerus::TTNetwork<false> datafile.\nFormat: Binary\n"
    assert b.startswith(head)
    p = len(head)
    # version 1 (size_t), canonicalized (1-byte bool), corePosition 1, TN version 1, dims {2, 3}
    assert struct.unpack_from("<Q?QQQQQ", b, p) == (1, True, 1, 1, 2, 2, 3)
    # the last component's values end the file right before the right virtual node (version, {1}, dense, 1.0)
    tail = struct.pack("<QQQQd", 1, 1, 1, 1, 1.0)
    assert b.endswith(struct.pack("<ddd", 0.25, 3.0, 1e-300) + tail)
    # every byte is a field. size_t: TT 2, TN version + dims 4, external links 6, node count 1, node links
    # 4 + 10 + 10 + 4, tensors 4 + 6 + 6 + 4; bool: canonicalized + 8 link flags; double: 1 + 2 + 3 + 1
    n_sz = 2 + 4 + 6 + 1 + (4 + 10 + 10 + 4) + (4 + 6 + 6 + 4)
    assert len(b) == len(head) + 8 * n_sz + 9 + 8 * 7


@pytest.mark.parametrize("tsv", [False, True])
@pytest.mark.parametrize("op", [False, True])
def test_oracle_round_trip(op, tsv):
    rng = np.random.default_rng(7 + 2 * op + tsv)
    d = 4
    n = [3, 2, 4, 2]
    m = [2, 2, 1, 3]
    r = [1, 2, 3, 2, 1]
    cores = [rng.standard_normal((r[k], n[k], m[k], r[k + 1]) if op else (r[k], n[k], r[k + 1])) for k in range(d)]
    dims = n + m if op else n
    out = F.read_tt(F.write_tt(cores, dims, op, False, 2, tsv=tsv))
    assert out["op"] == op and out["dims"] == dims and not out["canonicalized"] and out["core_position"] == 2
    for a, b in zip(cores, out["cores"]):
        assert a.shape == b.shape
        if tsv:
            assert np.allclose(a, b, rtol=1e-15, atol=0)
        else:
            assert np.array_equal(a, b)
    # the sparse-component variant stores the same values
    sp = F.read_tt(F.sparse_component_file(cores, dims, op))
    for a, b in zip(cores, sp["cores"]):
        assert np.array_equal(a, b)


def _random_tt(xe, op, seed):
    xe.seed(seed)
    if op:
        return xe.TTOperator.random([3, 2, 4, 2, 2, 2, 1, 3], [2, 3, 2])
    return xe.TTTensor.random([3, 5, 4, 2, 6], [3, 4, 4, 2])


def _host_cores(t):
    return [t.get_component(k) for k in range(t.order())]


@pytest.mark.gpu
@pytest.mark.parametrize("tsv", [False, True])
@pytest.mark.parametrize("op", [False, True])
def test_save_matches_oracle_bytes(xe, tmp_path, op, tsv):
    print(f"save op={op} tsv={tsv}", flush=True)
    t = _random_tt(xe, op, 11 + op)
    path = str(tmp_path / "t.dat")
    xe.save_to_file(t, path, xe.FileFormat.TSV if tsv else xe.FileFormat.BINARY)
    canon, pos = t.canonicalized, t.corePosition
    want = F.write_tt(_host_cores(t), list(t.dimensions), op, canon, pos, tsv=tsv)
    assert open(path, "rb").read() == want


@pytest.mark.gpu
def test_save_literal_file(xe, tmp_path):
    cores, dims = _lit_tt()
    t = xe.TTTensor(dims)
    for k, c in enumerate(cores):
        t.set_component(k, c)
    t.assume_core_position(0)
    path = str(tmp_path / "lit.tsv")
    xe.save_to_file(t, path, xe.FileFormat.TSV)
    assert open(path).read() == LITERAL_TT_TSV


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["binary", "tsv", "sparse"])
@pytest.mark.parametrize("op", [False, True])
def test_load_oracle_files(xe, tmp_path, op, kind):
    print(f"load op={op} kind={kind}", flush=True)
    rng = np.random.default_rng(3 + op)
    n = [3, 2, 4]
    m = [2, 1, 2]
    r = [1, 3, 2, 1]
    cores = [rng.standard_normal((r[k], n[k], m[k], r[k + 1]) if op else (r[k], n[k], r[k + 1])) for k in range(3)]
    cores[1].reshape(-1)[::3] = 0.0  # zeros, so the sparse variant really is sparse
    dims = n + m if op else n
    if kind == "sparse":
        data = F.sparse_component_file(cores, dims, op, True, 1)
    else:
        data = F.write_tt(cores, dims, op, True, 1, tsv=kind == "tsv")
    path = tmp_path / "in.dat"
    path.write_bytes(data)
    t = xe.load_from_file(str(path))
    assert isinstance(t, xe.TTOperator if op else xe.TTTensor)
    assert list(t.dimensions) == dims
    assert (t.canonicalized, t.corePosition) == (True, 1)
    for a, b in zip(cores, _host_cores(t)):
        if kind == "tsv":
            assert np.allclose(a, b, rtol=1e-15, atol=0)
        else:
            assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("op", [False, True])
def test_round_trip_and_full_tensor(xe, tmp_path, op):
    """TT:read_write_file (fileIO.cxx:76-89): same dimensions, canonical flags and values after a round trip."""
    t = _random_tt(xe, op, 21 + op)
    path = str(tmp_path / "rt.dat")
    xe.save_to_file(t, path)
    u = xe.load_from_file(path)
    assert (u.canonicalized, u.corePosition) == (t.canonicalized, t.corePosition)
    assert list(u.dimensions) == list(t.dimensions) and u.ranks() == t.ranks()
    for a, b in zip(_host_cores(t), _host_cores(u)):
        assert np.array_equal(a, b)
    assert np.array_equal(t.full(), u.full())


@pytest.mark.gpu
def test_load_rejects_bad_files(xe, tmp_path):
    cores, dims = _lit_tt()
    good = F.write_tt(cores, dims, False, True, 0, tsv=False)
    bad = {
        "header": good.replace(b"TTNetwork<false>", b"TensorNetwork", 1),
        "format": good.replace(b"Format: Binary", b"Format: Ascii!", 1),
        "truncated": good[:-9],
        "version": good.replace(struct.pack("<Q?Q", 1, True, 0), struct.pack("<Q?Q", 2, True, 0), 1),
    }
    for name, data in bad.items():
        p = tmp_path / f"{name}.dat"
        p.write_bytes(data)
        with pytest.raises(xe.XttError):
            xe.load_from_file(str(p))
    with pytest.raises(xe.XttError):
        xe.load_from_file(str(tmp_path / "missing.dat"))
