"""ctypes access to the oracle's C restatement (oracle/_build) and to the
reference build (oracle/_ref).  Test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PORT = os.path.join(ROOT, "oracle", "_build", "libmona_port.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libref_spmv.so")

IP = C.POINTER(C.c_int)
FP = C.POINTER(C.c_float)


def _build():
    if not os.path.exists(PORT):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "port"], check=True)


def port():
    _build()
    h = C.CDLL(PORT)
    sig12 = [C.c_int, IP, IP, FP, C.c_int, IP, C.c_int, FP, FP, C.c_int, FP, FP]
    sig10 = [C.c_int, IP, IP, FP, C.c_int, IP, C.c_int, FP, C.c_int, FP]
    for n in ("port_cC_oBlock_omp", "port_cC_oBlock", "port_cR_oBlock"):
        getattr(h, n).argtypes = sig12
        getattr(h, n).restype = None
    for n in ("port_rC_oBlock", "port_rR_oBlock"):
        getattr(h, n).argtypes = sig10
        getattr(h, n).restype = None
    h.port_r_nJump2index.argtypes = [C.c_int, IP, IP]
    h.port_gridder_n2u_leight.argtypes = [C.c_int, FP, FP, FP, FP, FP, FP, FP, FP, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]
    h.port_gridder_n2u.argtypes = [FP, FP, FP, FP, FP, IP, C.c_int, C.c_int, C.c_int, C.c_int]
    h.port_secret_length.argtypes = [C.c_int, FP, FP, FP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    h.port_secret_length.restype = C.c_longlong
    return h


def ref():
    """The reference's own compiled kernel, or None when it was not built
    (reference sources absent)."""
    if not os.path.exists(REF):
        return None
    h = C.CDLL(REF)
    h.ref_cC_oBlock_omp.argtypes = [C.c_int, IP, IP, FP, C.c_int, IP, C.c_int, FP, FP, C.c_int, FP, FP]
    h.ref_cC_oBlock_omp.restype = None
    return h


def ip(a):
    return None if a is None else a.ctypes.data_as(IP)


def fp(a):
    return a.ctypes.data_as(FP)


def call_cC(fn, M, vr, vi):
    """W = S V (column layout, split planes) with one of the kernels."""
    n_vec = vr.shape[0]
    wr = np.empty((n_vec, M["l_size"]), np.float32)
    wi = np.empty((n_vec, M["l_size"]), np.float32)
    fn(M["r_size"], ip(M["r_jump"]), ip(M["r_nJump"]), fp(M["m_val"]), M["l_size"],
       ip(M.get("l_jump")), M["l_nJump"], fp(vr), fp(vi), n_vec, fp(wr), fp(wi))
    return wr, wi


def random_matrix(rng, l_size, r_size, max_per_row, squeeze=False, empty_frac=0.2):
    """Random cpp_prepared one-block matrix (ascending columns per row,
    running-delta encoding with absolute first entry)."""
    rows = np.arange(l_size)
    if squeeze:
        keep = rng.random(l_size) > empty_frac
        keep[0] = keep[-1] = True
        rows = rows[keep]
    r_nJump, cols, vals = [], [], []
    for _ in rows:
        n = int(rng.integers(0 if not squeeze else 1, max_per_row + 1))
        c = np.sort(rng.choice(r_size, size=min(n, r_size), replace=False))
        r_nJump.append(len(c))
        cols.extend(c.tolist())
        vals.extend(rng.normal(size=len(c)).tolist())
    cols = np.array(cols, dtype=np.int64)
    r_jump = np.diff(np.concatenate([[0], cols])).astype(np.int32)
    M = dict(r_size=int(r_size), l_size=int(l_size), r_nJump=np.array(r_nJump, np.int32),
             r_jump=r_jump, m_val=np.array(vals, np.float32), l_nJump=len(r_nJump), l_jump=None)
    if squeeze:
        M["l_jump"] = np.diff(np.concatenate([[0], rows])).astype(np.int32)
    return M
