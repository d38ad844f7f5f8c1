"""The reference's binary array / sparse-matrix file format
(src/readWrite/bmArray2Binary*.m, bmBinary2Array*.m): exact bytes of the
header and data files, round trips, and the conversions on loading
(matlab_ind -> cpp_prepared as bmSparseMat.cpp_prepare does, multi-block ->
one block)."""
import os

import numpy as np
import pytest

from monalisa_amd import binfile
from oracle import mona_oracle as O


def test_array_file_bytes(tmp_path):
    d = str(tmp_path)
    binfile.bmArray2Binary(np.array([1, 2, 3]), d, "a", "int")
    assert open(os.path.join(d, "a.hdat"), "rb").read() == b"2\r\n1 3\r\n3\r\nint\r\n4\r\n"
    assert open(os.path.join(d, "a.dat"), "rb").read() == np.array([1, 2, 3], "<i4").tobytes()
    binfile.bmArray2Binary(np.array([[1, 2], [3, 4]]), d, "m", "int64")
    assert open(os.path.join(d, "m.hdat"), "rb").read() == b"2\r\n2 2\r\n4\r\nlonglong\r\n8\r\n"
    # column-major (MATLAB) element order
    assert open(os.path.join(d, "m.dat"), "rb").read() == np.array([1, 3, 2, 4], "<i8").tobytes()
    binfile.bmArray2Binary(np.float64(0.5), d, "s", "single")
    assert open(os.path.join(d, "s.hdat"), "rb").read() == b"2\r\n1 1\r\n1\r\nfloat\r\n4\r\n"
    binfile.bmArray2Binary(np.zeros((0, 0)), d, "e", "double")
    assert open(os.path.join(d, "e.hdat"), "rb").read() == b"2\r\n0 0\r\n0\r\ndouble\r\n8\r\n"
    assert os.path.getsize(os.path.join(d, "e.dat")) == 0


def test_array_round_trip(tmp_path):
    d = str(tmp_path)
    rng = np.random.default_rng(0)
    cases = [(rng.integers(-5, 5, size=(3, 4, 2)), "int", np.int32),
             (rng.normal(size=7), "single", np.float32),
             (rng.normal(size=(5, 2)), "double", np.float64),
             (np.array([2 ** 40, -3]), "int64", np.int64)]
    for i, (a, t, dt) in enumerate(cases):
        binfile.bmArray2Binary(a, d, f"x{i}", t)
        b = binfile.bmBinary2Array(d, f"x{i}")
        assert b.dtype == dt
        np.testing.assert_array_equal(b.reshape(-1, order="F"), np.asarray(a, dt).reshape(-1, order="F"))


def _traj_mats():
    from monalisa_amd.sparse import bmTraj2SparseMat
    from monalisa_amd import synth
    t = synth.radial2d_golden(16, 24, k_max=6.0)
    ve = synth.ve_radial2(t, 24, 16)
    return bmTraj2SparseMat(t, ve, [24, 24], [0.5, 0.5], "kaiser")


@pytest.mark.parametrize("which", ["Gu", "Gut"])
def test_sparse_round_trip(tmp_path, which):
    G = _traj_mats()[which]
    d = str(tmp_path / which)
    binfile.bmArray2Binary_sparseMat(G, d)
    header = open(os.path.join(d, binfile.HEADER), "rb").read().split(b"\r\n")
    assert header[0].decode() == G.type and header[1] == b"one_block"
    H = binfile.bmBinary2Array_sparseMat(d)
    assert (H.r_size, H.l_size, H.l_nJump, H.type) == (G.r_size, G.l_size, G.l_nJump, G.type)
    np.testing.assert_array_equal(H.r_jump, G.r_jump)
    np.testing.assert_array_equal(H.r_nJump, G.r_nJump)
    np.testing.assert_array_equal(H.m_val, G.m_val)
    if G.l_jump is None:
        assert H.l_jump is None
    else:
        np.testing.assert_array_equal(H.l_jump, G.l_jump)


def _write_matlab_ind(d, Gm):
    binfile.bmArray2Binary(Gm["r_nJump"], d, "r_nJump", "int")
    binfile.bmArray2Binary(Gm["r_ind"], d, "r_ind", "int")
    binfile.bmArray2Binary(Gm["m_val"], d, "m_val", "single")
    binfile.bmArray2Binary(Gm["l_ind"] if Gm["l_ind"] is not None else np.zeros((0, 0)), d, "l_ind", "int")
    binfile._write_lines(os.path.join(d, binfile.HEADER),
                         [Gm["type"], "void", "kaiser", Gm["r_size"], Gm["l_size"], Gm["l_nJump"]])


@pytest.mark.parametrize("squeeze", [False, True])
def test_matlab_ind_file_is_cpp_prepared_on_load(tmp_path, squeeze):
    """A matlab_ind file loads as the encoding bmSparseMat.cpp_prepare makes."""
    rng = np.random.default_rng(3)
    r = rng.integers(0, 9, 60)
    c = rng.integers(0, 11, 60)
    keep = np.unique(r * 100 + c, return_index=True)[1]
    r, c = r[keep], c[keep]
    v = rng.normal(size=r.size)
    Gm = O.to_bmsparse(r, c, v, (12, 11))
    if squeeze:
        O.l_squeeze(Gm)
    d = str(tmp_path)
    _write_matlab_ind(d, Gm)
    H = binfile.bmBinary2Array_sparseMat(d)
    O.cpp_prepare(Gm)
    np.testing.assert_array_equal(H.r_jump, Gm["r_jump"])
    np.testing.assert_array_equal(H.r_nJump, Gm["r_nJump"])
    np.testing.assert_array_equal(H.m_val, Gm["m_val"])
    if squeeze:
        np.testing.assert_array_equal(H.l_jump, Gm["l_jump"])
    else:
        assert H.l_jump is None
    assert H.type == Gm["type"]


def test_multi_block_file_loads_as_one_block(tmp_path):
    """Two blocks (the first jump of each absolute) re-encode to the
    one-block jumps of the same matrix."""
    G = _traj_mats()["Gu"]
    col = np.cumsum(G.r_jump.astype(np.int64))
    nnz_first = int(G.r_nJump[:G.l_nJump // 2].sum())
    rj2 = G.r_jump.astype(np.int64).copy()
    rj2[nnz_first] = col[nnz_first]           # absolute first entry of block 2
    d = str(tmp_path)
    binfile.bmArray2Binary(G.r_nJump, d, "r_nJump", "int")
    binfile.bmArray2Binary(rj2, d, "r_jump", "int")
    binfile.bmArray2Binary(G.m_val, d, "m_val", "single")
    binfile.bmArray2Binary(np.zeros((0, 0)), d, "l_jump", "int")
    binfile.bmArray2Binary(np.array([G.l_nJump // 2, G.l_nJump - G.l_nJump // 2]), d, "block_length", "int")
    binfile.bmArray2Binary(np.array([0, G.l_nJump // 2]), d, "l_block_start", "int")
    binfile.bmArray2Binary(np.array([0, nnz_first]), d, "m_block_start", "int64")
    binfile._write_lines(os.path.join(d, binfile.HEADER),
                         ["cpp_prepared", "multi_block", "kaiser", G.r_size, G.l_size, G.l_nJump, 2])
    H = binfile.bmBinary2Array_sparseMat(d)
    np.testing.assert_array_equal(H.r_jump, G.r_jump)


def test_unknown_type_is_refused(tmp_path):
    d = str(tmp_path)
    binfile._write_lines(os.path.join(d, binfile.HEADER), ["void", "void", "void", 1, 1, 1])
    with pytest.raises(ValueError):
        binfile.bmBinary2Array_sparseMat(d)
