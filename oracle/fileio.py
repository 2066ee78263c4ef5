"""TEST INFRASTRUCTURE ONLY — pure-Python restatement of xerus's datafile format for TT networks.

Follows xerus 3.0.1:
- ``misc/fileIO.h:52-99`` (scalar and vector stream writers/readers: raw bytes for BINARY, ``value`` + tab for
  TSV) and ``:105-160`` (``save_to_file`` / ``load_from_file``: the two header lines, the binary reopen);
- ``src/xerus/ttNetwork.cpp:1454-1488`` (TTNetwork: version 1, canonicalized, corePosition, then the network);
- ``src/xerus/tensorNetwork.cpp:1429-1504`` (version 1, dimensions, external links, node links, node tensors,
  with the TSV line breaks);
- ``src/xerus/tensor.cpp:1781-1843`` (version 1, dimensions, representation 1 = dense values /
  2 = (position, value) pairs);
- ``src/xerus/ttNetwork.cpp:57-108`` (the network of a TT: virtual boundary nodes ``ones({1})`` around the
  components, the link layout) and ``:475-492`` (link dimensions follow the component dimensions).

TSV numbers are written as C++ iostreams do with ``setprecision(16)`` in the default float field, i.e. ``%.16g``.
Only ``tests/`` import this module.
"""
from __future__ import annotations

import struct

import numpy as np

SIZE_MAX = (1 << 64) - 1
CLASS = {False: "xerus::TTNetwork<false>", True: "xerus::TTNetwork<true>"}


class _Out:
    def __init__(self, tsv: bool):
        self.tsv = tsv
        self.parts: list[bytes] = []

    def sz(self, v: int):
        self.parts.append(f"{v}\t".encode() if self.tsv else struct.pack("<Q", v))

    def flag(self, v: bool):
        self.parts.append(f"{int(bool(v))}\t".encode() if self.tsv else struct.pack("<?", bool(v)))

    def value(self, v: float):
        self.parts.append(f"{v:.16g}\t".encode() if self.tsv else struct.pack("<d", v))

    def text(self, s: str):
        if self.tsv:
            self.parts.append(s.encode())

    def dims(self, d):
        self.sz(len(d))
        for x in d:
            self.sz(int(x))

    def link(self, external, other, pos, dim):
        self.flag(external)
        self.sz(other)
        self.sz(pos)
        self.sz(dim)

    def tensor(self, arr: np.ndarray):
        self.sz(1)
        self.dims(arr.shape)
        self.sz(1)
        for v in np.asarray(arr, dtype=np.float64).ravel():
            self.value(float(v))


def write_tt(cores, dims, op: bool, canonicalized: bool = True, core_position: int = 0, tsv: bool = False) -> bytes:
    """The bytes ``misc::save_to_file(tt, name, format)`` writes for a TT with these components."""
    d = len(cores)
    N = 2 if op else 1
    assert len(dims) == N * d
    ranks = [1] + [c.shape[-1] for c in cores[:-1]] + [1]
    o = _Out(tsv)
    o.parts.append(f"Xerus {CLASS[op]} datafile.\nFormat: {'TSV' if tsv else 'Binary'}\n".encode())
    o.sz(1)
    o.flag(canonicalized)
    o.sz(core_position)
    o.sz(1)
    o.dims(dims)
    o.text("\n")
    for i in range(d):
        o.sz(i + 1)
        o.sz(1)
        o.sz(dims[i])
    if op:
        for i in range(d):
            o.sz(i + 1)
            o.sz(2)
            o.sz(dims[d + i])
    o.text("\n\n")
    o.sz(d + 2)
    o.text("\n")
    o.sz(1)
    o.link(False, 1, 0, 1)
    for i in range(d):
        o.sz(N + 2)
        o.link(False, i, 0 if i == 0 else N + 1, ranks[i])
        o.link(True, SIZE_MAX, i, dims[i])
        if op:
            o.link(True, SIZE_MAX, i + d, dims[d + i])
        o.link(False, i + 2, 0, ranks[i + 1])
    o.sz(1)
    o.link(False, d, N + 1, 1)
    o.text("\n")
    one = np.ones(1)
    o.tensor(one)
    o.text("\n")
    for c in cores:
        o.tensor(c)
        o.text("\n")
    o.tensor(one)
    o.text("\n")
    return b"".join(o.parts)


class _In:
    def __init__(self, data: bytes, pos: int, tsv: bool):
        self.data, self.pos, self.tsv = data, pos, tsv
        if tsv:
            self.tokens = data[pos:].split()
            self.t = 0

    def _tok(self):
        tok = self.tokens[self.t]
        self.t += 1
        return tok

    def sz(self) -> int:
        if self.tsv:
            return int(self._tok())
        v = struct.unpack_from("<Q", self.data, self.pos)[0]
        self.pos += 8
        return v

    def flag(self) -> bool:
        if self.tsv:
            return bool(int(self._tok()))
        v = struct.unpack_from("<?", self.data, self.pos)[0]
        self.pos += 1
        return v

    def value(self) -> float:
        if self.tsv:
            return float(self._tok())
        v = struct.unpack_from("<d", self.data, self.pos)[0]
        self.pos += 8
        return v

    def dims(self):
        return [self.sz() for _ in range(self.sz())]

    def tensor(self) -> np.ndarray:
        assert self.sz() == 1
        dims = self.dims()
        out = np.zeros(int(np.prod(dims)) if dims else 1)
        rep = self.sz()
        if rep == 1:
            for i in range(out.size):
                out[i] = self.value()
        else:
            assert rep == 2
            for _ in range(self.sz()):
                p = self.sz()
                out[p] = self.value()
        return out.reshape(dims)


def read_tt(data: bytes) -> dict:
    """Parse a TT datafile: components, dims, operator flag, canonicalized, corePosition."""
    nl = data.index(b"\n")
    first = data[:nl].decode()
    op = {f"Xerus {CLASS[False]} datafile.": False, f"Xerus {CLASS[True]} datafile.": True}[first]
    rest = data[nl + 1:]
    nl2 = rest.index(b"\n")
    qual, fmt = rest[:nl2].decode().split()
    assert qual == "Format:"
    inp = _In(data, nl + 1 + nl2 + 1, fmt == "TSV")
    assert inp.sz() == 1
    canonicalized = inp.flag()
    core_position = inp.sz()
    assert inp.sz() == 1
    dims = inp.dims()
    links = [(inp.sz(), inp.sz(), inp.sz()) for _ in dims]
    nodes = inp.sz()
    node_links = []
    for _ in range(nodes):
        node_links.append([(inp.flag(), inp.sz(), inp.sz(), inp.sz()) for _ in range(inp.sz())])
    tensors = [inp.tensor() for _ in range(nodes)]
    return {"op": op, "dims": dims, "canonicalized": canonicalized, "core_position": core_position,
            "external_links": links, "node_links": node_links, "cores": tensors[1:-1], "boundary": (tensors[0], tensors[-1])}


def sparse_component_file(cores, dims, op: bool, canonicalized: bool = True, core_position: int = 0) -> bytes:
    """Binary file whose components are stored as sparse tensors (representation 2, nonzeros only), as xerus
    writes dirac-initialised components (tensor.cpp:1796-1804)."""
    dense = write_tt(cores, dims, op, canonicalized, core_position, tsv=False)
    # rebuild the tensor section: everything after the node links is replaced
    d = len(cores)
    N = 2 if op else 1
    head = _header_and_links_length(dense, d, N, len(dims))
    o = _Out(False)
    one = np.ones(1)
    o.tensor(one)
    for c in cores:
        flat = np.asarray(c, dtype=np.float64).ravel()
        nz = np.flatnonzero(flat)
        o.sz(1)
        o.dims(c.shape)
        o.sz(2)
        o.sz(len(nz))
        for p in nz:
            o.sz(int(p))
            o.value(float(flat[p]))
    o.tensor(one)
    return dense[:head] + b"".join(o.parts)


def _header_and_links_length(data: bytes, d: int, N: int, ndims: int) -> int:
    nl = data.index(b"\n")
    nl2 = data.index(b"\n", nl + 1)
    pos = nl2 + 1
    pos += 8 + 1 + 8          # TT version, canonicalized, corePosition
    pos += 8 + 8 + 8 * ndims  # TN version, dims vector
    pos += 24 * ndims         # external links
    pos += 8                  # node count
    pos += 2 * (8 + 25)       # boundary nodes: degree + one link
    pos += d * (8 + 25 * (N + 2))
    return pos
