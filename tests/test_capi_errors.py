"""Error paths of the host-side C-ABI entry points (no device work): the
sparse-matrix handle rejects exactly the matrices bmSparseMat.check rejects
(src/sparseMat/m/bmSparseMat.m:192-393, restated by the oracle's check()),
with MONA_ERR_ARG and the rule in mona_last_error(); the structure builders
reject bad arguments; a valid handle exports its arrays unchanged."""
import ctypes as C

import numpy as np
import pytest

import oracle_libs as OL
from oracle import mona_oracle as O

MONA_ERR_ARG = 1


def _lib():
    from monalisa_amd import lib
    return lib


def _create(r_size, r_jump, r_nJump, m_val, l_size, l_jump=None, l_nJump=None, nnz=None):
    L = _lib()
    h = C.c_void_p()
    rj = np.ascontiguousarray(r_jump, np.int32)
    rn = np.ascontiguousarray(r_nJump, np.int32)
    mv = np.ascontiguousarray(m_val, np.float32)
    lj = None if l_jump is None else np.ascontiguousarray(l_jump, np.int32)
    st = L.mona_sparse_create(r_size, OL.ip(rj), OL.ip(rn), OL.fp(mv), l_size, OL.ip(lj),
                              rn.size if l_nJump is None else l_nJump, mv.size if nnz is None else nnz,
                              C.byref(h))
    return st, h


def _oracle_fails(r_size, r_jump, r_nJump, m_val, l_size, l_jump=None):
    G = dict(r_size=r_size, l_size=l_size, r_jump=np.asarray(r_jump, np.int32), r_nJump=np.asarray(r_nJump, np.int32),
             m_val=np.asarray(m_val, np.float32), l_nJump=len(r_nJump),
             l_jump=None if l_jump is None else np.asarray(l_jump, np.int32),
             type="cpp_prepared" if l_jump is None else "l_squeezed_cpp_prepared")
    return O.check(G)


def test_valid_handle_round_trips():
    L = _lib()
    M = OL.random_matrix(np.random.default_rng(0), 40, 30, 6, squeeze=True)
    st, h = _create(M["r_size"], M["r_jump"], M["r_nJump"], M["m_val"], M["l_size"], M.get("l_jump"))
    assert st == 0, L.mona_last_error()
    nnz, lnj, rs, ls, sq = C.c_longlong(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
    assert L.mona_sparse_info(h, C.byref(nnz), C.byref(lnj), C.byref(rs), C.byref(ls), C.byref(sq)) == 0
    assert (nnz.value, lnj.value, rs.value, ls.value) == (M["m_val"].size, M["l_nJump"], M["r_size"], M["l_size"])
    rj = np.empty(nnz.value, np.int32)
    rn = np.empty(lnj.value, np.int32)
    mv = np.empty(nnz.value, np.float32)
    lj = np.empty(lnj.value, np.int32)
    assert L.mona_sparse_export(h, OL.ip(rj), OL.ip(rn), OL.fp(mv), OL.ip(lj)) == 0
    assert np.array_equal(rj, M["r_jump"]) and np.array_equal(rn, M["r_nJump"])
    assert np.array_equal(mv.view(np.int32), np.asarray(M["m_val"], np.float32).view(np.int32))
    if sq.value:
        assert np.array_equal(lj, M["l_jump"])
    assert L.mona_sparse_destroy(h) == 0


# (case, rule): each breaks one rule of bmSparseMat.check
CASES = [
    ("one output row", "2.5"),
    ("one column", "2.5"),
    ("negative row count", "2.6"),
    ("no entries", "3.1"),
]


def _case(name):
    r_jump, r_nJump, m_val = [0, 1, 1, 2], [2, 0, 2], [0.5, 0.5, 1.0, 1.0]
    r_size, l_size = 8, 5
    if name == "one output row":
        return dict(r_size=r_size, r_jump=[0, 1], r_nJump=[2], m_val=[1.0, 1.0], l_size=1)
    if name == "one column":
        return dict(r_size=1, r_jump=[0, 0, 0, 0], r_nJump=r_nJump, m_val=m_val, l_size=l_size)
    if name == "negative row count":
        return dict(r_size=r_size, r_jump=[0, 1, 1, 2], r_nJump=[3, -1, 2], m_val=m_val, l_size=l_size)
    if name == "no entries":
        return dict(r_size=r_size, r_jump=[], r_nJump=[0, 0, 0], m_val=[], l_size=l_size)
    raise ValueError(name)


@pytest.mark.parametrize("name,rule", CASES)
def test_invalid_matrix_is_rejected_like_check(name, rule):
    k = _case(name)
    assert rule in _oracle_fails(k["r_size"], k["r_jump"], k["r_nJump"], k["m_val"], k["l_size"])
    st, h = _create(k["r_size"], k["r_jump"], k["r_nJump"], k["m_val"], k["l_size"])
    assert st == MONA_ERR_ARG and not h.value
    assert f"check {rule}" in _lib().mona_last_error().decode()


def test_entry_count_mismatch_is_rejected():
    """sum(r_nJump) != numel(m_val) (rule 3.2)."""
    r_jump, r_nJump, m_val = [0, 1, 1, 2, 0], [2, 0, 2], [0.5, 0.5, 1.0, 1.0, 1.0]
    assert "3.2" in _oracle_fails(8, r_jump, r_nJump, m_val, 5)
    st, _ = _create(8, r_jump, r_nJump, m_val, 5)
    assert st == MONA_ERR_ARG and "check 3.2" in _lib().mona_last_error().decode()


def test_valid_small_matrix_passes_both():
    r_jump, r_nJump, m_val = [0, 1, 1, 2], [2, 0, 2], [0.5, 0.5, 1.0, 1.0]
    assert _oracle_fails(8, r_jump, r_nJump, m_val, 5) == []
    st, h = _create(8, r_jump, r_nJump, m_val, 5)
    assert st == 0
    assert _lib().mona_sparse_destroy(h) == 0


def test_structure_builders_reject_bad_arguments():
    L = _lib()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    t = np.zeros((4, 2), np.float64)
    ve = np.ones(4, np.float64)
    N_u = np.array([8, 8], np.int32)
    dK = np.array([0.5, 0.5], np.float64)
    kp = np.array([1.95, 10.0, 10.0], np.float64)
    hn, hu, hut = C.c_void_p(), C.c_void_p(), C.c_void_p()
    # nothing requested
    assert L.mona_traj2sparsemat(2, 4, dp(t), dp(ve), OL.ip(N_u), dp(dK), 1, 3, dp(kp), 3, 0,
                                 C.byref(hn), C.byref(hu), C.byref(hut)) == MONA_ERR_ARG
    # NULL trajectory
    assert L.mona_traj2sparsemat(2, 4, None, dp(ve), OL.ip(N_u), dp(dK), 1, 3, dp(kp), 3, 2,
                                 C.byref(hn), C.byref(hu), C.byref(hut)) == MONA_ERR_ARG
    # a valid call builds a matrix pair on the host
    st = L.mona_traj2sparsemat(2, 4, dp(t), dp(ve), OL.ip(N_u), dp(dK), 1, 3, dp(kp), 3, 2,
                               C.byref(hn), C.byref(hu), C.byref(hut))
    assert st == 0, L.mona_last_error()
    assert L.mona_sparse_destroy(hu) == 0 and L.mona_sparse_destroy(hut) == 0
    # deformation fields: imDim outside 1..3
    v = np.zeros((4, 16), np.float64)
    n4 = np.array([2, 2, 2, 2], np.int32)
    assert L.mona_deform2sparsemat(4, OL.ip(n4), dp(v), None, 0, 3, C.byref(hn), C.byref(hu),
                                   C.byref(hut)) == MONA_ERR_ARG
