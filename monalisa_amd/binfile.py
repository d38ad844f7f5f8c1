"""The reference's binary array and sparse-matrix file format
(src/readWrite/bmArray2Binary.m:6-74, bmBinary2Array.m:6-51,
bmArray2Binary_sparseMat.m:6-70, bmBinary2Array_sparseMat.m:6-58), so that
matrices saved by reference users load into the GPU plans and ours load back
into MATLAB.

An array NAME in directory D is two files:

* ``D/NAME.hdat``: text lines (CRLF) ``ndims``, the sizes separated by
  spaces, the element count, the type (``int``, ``longlong``, ``float``,
  ``double``) and the bytes per element;
* ``D/NAME.dat``: the elements, little-endian, column-major (MATLAB order).

A sparse matrix is a directory holding ``bmArray2Binary_sparseMat_header.txt``
(type, block type, kernel type, r_size, l_size, l_nJump and, for the
cpp_prepared types, nBlock, one per line) and its arrays.
"""
from __future__ import annotations

import os

import numpy as np

_TYPES = {  # file type name -> (numpy dtype, bytes)
    "int": (np.dtype("<i4"), 4),
    "longlong": (np.dtype("<i8"), 8),
    "float": (np.dtype("<f4"), 4),
    "double": (np.dtype("<f8"), 8),
}
_ARG = {"int": "int", "int64": "longlong", "single": "float", "double": "double"}
HEADER = "bmArray2Binary_sparseMat_header.txt"


def _write_lines(path, lines):
    with open(path, "w", newline="") as f:
        for ln in lines:
            f.write(f"{ln}\r\n")


def bmArray2Binary(arr, directory, name, type_):
    """Writes ``arr`` as NAME.hdat / NAME.dat; type_ is the reference's
    argument ('int', 'int64', 'single' or 'double')."""
    if type_ not in _ARG:
        raise ValueError("Type is unknown")
    tname = _ARG[type_]
    dt, nbytes = _TYPES[tname]
    a = np.asarray(arr)
    if a.ndim < 2:  # MATLAB arrays have at least two dimensions (row vectors)
        a = a.reshape(1, -1) if a.size else a.reshape(0, 0)
    shape = a.shape
    os.makedirs(directory, exist_ok=True)
    _write_lines(os.path.join(directory, name + ".hdat"),
                 [len(shape), " ".join(str(int(v)) for v in shape), int(a.size), tname, nbytes])
    with open(os.path.join(directory, name + ".dat"), "wb") as f:
        f.write(np.asarray(a, dtype=dt).ravel(order="F").tobytes())


def bmBinary2Array(directory, name):
    """Reads NAME.hdat / NAME.dat; returns the array with its MATLAB shape
    (column-major) and type (int32, int64, float32, float64)."""
    with open(os.path.join(directory, name + ".hdat")) as f:
        tok = f.read().split()
    nd = int(tok[0])
    shape = tuple(int(v) for v in tok[1:1 + nd])
    n = int(tok[1 + nd])
    tname = tok[2 + nd]
    if tname not in _TYPES:
        raise ValueError(f"{name}.hdat: unknown type {tname!r}")
    dt, nbytes = _TYPES[tname]
    if int(tok[3 + nd]) != nbytes:
        raise ValueError(f"{name}.hdat: {tname} with {tok[3 + nd]} bytes per element")
    data = np.fromfile(os.path.join(directory, name + ".dat"), dtype=dt, count=n)
    if data.size != n:
        raise ValueError(f"{name}.dat holds {data.size} of {n} elements")
    return data.reshape(shape, order="F")


def _jumps(ind):
    """1-based ascending indices -> jump encoding (first entry absolute, 0-based)."""
    ind = np.asarray(ind, dtype=np.int64).ravel()
    j = np.empty(ind.shape, dtype=np.int64)
    if ind.size:
        j[0] = ind[0] - 1
        j[1:] = np.diff(ind)
    return j.astype(np.int32)


def bmBinary2Array_sparseMat(directory):
    """Loads a saved bmSparseMat as a cpp_prepared one-block
    :class:`~monalisa_amd.sparse.BmSparseMat`.  matlab_ind files are
    converted as bmSparseMat.cpp_prepare does (bmSparseMat.m:100-188);
    multi-block files are re-encoded as one block (the first jump of each
    block is absolute in the file)."""
    from .sparse import BmSparseMat
    with open(os.path.join(directory, HEADER)) as f:
        tok = f.read().split()
    stype, kernel_type = tok[0], tok[2]
    if stype not in ("matlab_ind", "l_squeezed_matlab_ind", "cpp_prepared", "l_squeezed_cpp_prepared"):
        raise ValueError(f"sparse matrix type {stype!r} cannot be loaded")
    r_size, l_size, l_nJump = int(tok[3]), int(tok[4]), int(tok[5])
    rd = lambda nm: bmBinary2Array(directory, nm).ravel(order="F")
    r_nJump = rd("r_nJump").astype(np.int32)
    m_val = rd("m_val").astype(np.float32)
    if stype in ("matlab_ind", "l_squeezed_matlab_ind"):
        r_jump = _jumps(rd("r_ind"))
        l_jump = _jumps(rd("l_ind")) if stype == "l_squeezed_matlab_ind" else None
    elif stype in ("cpp_prepared", "l_squeezed_cpp_prepared"):
        nBlock = int(tok[6])
        r_jump = rd("r_jump").astype(np.int64)
        l_jump = rd("l_jump").astype(np.int32) if stype == "l_squeezed_cpp_prepared" else None
        if nBlock > 1:
            if l_jump is not None:
                raise ValueError("multi-block squeezed matrices are not supported")
            mstart = rd("m_block_start").astype(np.int64)
            col = np.empty(r_jump.shape, dtype=np.int64)
            bounds = list(mstart) + [r_jump.size]
            for b in range(nBlock):
                col[bounds[b]:bounds[b + 1]] = np.cumsum(r_jump[bounds[b]:bounds[b + 1]])
            r_jump = np.diff(np.concatenate([[0], col]))
        r_jump = r_jump.astype(np.int32)
    if r_nJump.size != l_nJump:
        raise ValueError("r_nJump does not have l_nJump entries")
    opt = {}
    for nm in ("N_u", "d_u", "nWin", "kernelParam"):
        if os.path.exists(os.path.join(directory, nm + ".hdat")):
            v = rd(nm)
            opt[nm] = v if v.size else None
    return BmSparseMat(r_size, l_size, r_jump, r_nJump, m_val, l_jump, N_u=opt.get("N_u"), d_u=opt.get("d_u"),
                       kernel_type=kernel_type, nWin=opt.get("nWin"), kernelParam=opt.get("kernelParam"))


def bmArray2Binary_sparseMat(s, directory):
    """Saves a cpp_prepared one-block BmSparseMat in the reference layout."""
    os.makedirs(directory, exist_ok=True)
    empty = np.zeros((0, 0))
    bmArray2Binary(s.r_nJump, directory, "r_nJump", "int")
    bmArray2Binary(s.r_jump, directory, "r_jump", "int")
    bmArray2Binary(s.m_val, directory, "m_val", "single")
    bmArray2Binary(s.l_jump if s.l_jump is not None else empty, directory, "l_jump", "int")
    bmArray2Binary(s.N_u if s.N_u is not None else empty, directory, "N_u", "int")
    bmArray2Binary(s.d_u if s.d_u is not None else empty, directory, "d_u", "single")
    bmArray2Binary(s.nWin if s.nWin is not None else empty, directory, "nWin", "int")
    bmArray2Binary(s.kernelParam if s.kernelParam is not None else empty, directory, "kernelParam", "single")
    bmArray2Binary(np.array([s.l_nJump]), directory, "block_length", "int")
    bmArray2Binary(np.array([0]), directory, "l_block_start", "int")
    bmArray2Binary(np.array([0]), directory, "m_block_start", "int64")
    _write_lines(os.path.join(directory, HEADER),
                 [s.type, s.block_type, s.kernel_type, s.r_size, s.l_size, s.l_nJump, 1])
