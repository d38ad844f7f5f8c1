"""GPU drop-in products vs the reference kernel's golden vectors and the
oracle's C restatement: bit-exact."""
import ctypes as C

import numpy as np
import pytest

import oracle_libs as OL
from test_oracle_port import load_golden

pytestmark = pytest.mark.gpu


def lib():
    from monalisa_amd import lib as L
    return L


def gpu_cC(fn, M, vr, vi):
    n_vec = vr.shape[0]
    wr = np.full((n_vec, M["l_size"]), np.nan, np.float32)
    wi = np.full((n_vec, M["l_size"]), np.nan, np.float32)
    fn(M["r_size"], OL.ip(M["r_jump"]), OL.ip(M["r_nJump"]), OL.fp(M["m_val"]), M["l_size"],
       OL.ip(M.get("l_jump")), M["l_nJump"], OL.fp(np.ascontiguousarray(vr)), OL.fp(np.ascontiguousarray(vi)),
       n_vec, OL.fp(wr), OL.fp(wi))
    return wr, wi


@pytest.mark.parametrize("case", load_golden(), ids=lambda c: c[0])
def test_cC_oBlock_omp_matches_reference_golden(case):
    name, M, vr, vi, wr, wi = case
    print(f"[gpu] cC_oBlock_omp golden case {name}", flush=True)
    L = lib()
    from monalisa_amd._lib import CXX_CC_OBLOCK_OMP
    for fn in (L.bmSparseMat_cC_oBlock_omp, getattr(L, CXX_CC_OBLOCK_OMP), L.bmSparseMat_cC_oBlock,
               L.bmSparseMat_cC_oBlock_2_omp):
        gr, gi = gpu_cC(fn, M, vr, vi)
        assert np.array_equal(gr.view(np.int32), wr.view(np.int32))
        assert np.array_equal(gi.view(np.int32), wi.view(np.int32))


@pytest.mark.parametrize("seed", range(8))
def test_all_layouts_bit_exact_vs_port(seed):
    print(f"[gpu] layouts seed {seed}", flush=True)
    rng = np.random.default_rng(100 + seed)
    M = OL.random_matrix(rng, int(rng.integers(2, 500)), int(rng.integers(2, 700)),
                         int(rng.integers(1, 60)), squeeze=bool(seed % 2))
    n = int(rng.integers(1, 7))
    vr = rng.normal(size=(n, M["r_size"])).astype(np.float32)
    vi = rng.normal(size=(n, M["r_size"])).astype(np.float32)
    P = OL.port()
    L = lib()
    er, ei = OL.call_cC(P.port_cC_oBlock_omp, M, vr, vi)
    gr, gi = gpu_cC(L.bmSparseMat_cC_oBlock_omp, M, vr, vi)
    assert np.array_equal(gr, er) and np.array_equal(gi, ei)
    # row layout
    vrr, vir = np.ascontiguousarray(vr.T), np.ascontiguousarray(vi.T)
    rr = np.full((M["l_size"], n), np.nan, np.float32)
    ri = np.full((M["l_size"], n), np.nan, np.float32)
    L.bmSparseMat_cR_oBlock(M["r_size"], OL.ip(M["r_jump"]), OL.ip(M["r_nJump"]), OL.fp(M["m_val"]),
                            M["l_size"], OL.ip(M["l_jump"]), M["l_nJump"], OL.fp(vrr), OL.fp(vir), n,
                            OL.fp(rr), OL.fp(ri))
    assert np.array_equal(rr.T, er) and np.array_equal(ri.T, ei)
    # real column / row layouts
    rc = np.full((n, M["l_size"]), np.nan, np.float32)
    L.bmSparseMat_rC_oBlock(M["r_size"], OL.ip(M["r_jump"]), OL.ip(M["r_nJump"]), OL.fp(M["m_val"]),
                            M["l_size"], OL.ip(M["l_jump"]), M["l_nJump"], OL.fp(vr), n, OL.fp(rc))
    assert np.array_equal(rc, er)
    rrow = np.full((M["l_size"], n), np.nan, np.float32)
    L.bmSparseMat_rR_oBlock(M["r_size"], OL.ip(M["r_jump"]), OL.ip(M["r_nJump"]), OL.fp(M["m_val"]),
                            M["l_size"], OL.ip(M["l_jump"]), M["l_nJump"], OL.fp(vrr), n, OL.fp(rrow))
    assert np.array_equal(rrow.T, er)


@pytest.mark.parametrize("n", [1, 4, 20])
def test_long_rows_bit_exact_vs_port(n):
    """Rows far longer than 64 entries: the drop-in products still sum each
    row sequentially in stored order (the reference's kernel does), so they
    equal the C restatement bit for bit."""
    print(f"[gpu] long rows n_vec={n}", flush=True)
    rng = np.random.default_rng(900 + n)
    M = OL.random_matrix(rng, 60, 3000, 900, squeeze=False)
    assert max(M["r_nJump"]) > 64
    vr = rng.normal(size=(n, M["r_size"])).astype(np.float32)
    vi = rng.normal(size=(n, M["r_size"])).astype(np.float32)
    er, ei = OL.call_cC(OL.port().port_cC_oBlock_omp, M, vr, vi)
    gr, gi = gpu_cC(lib().bmSparseMat_cC_oBlock_omp, M, vr, vi)
    assert np.array_equal(gr.view(np.int32), er.view(np.int32))
    assert np.array_equal(gi.view(np.int32), ei.view(np.int32))


@pytest.mark.parametrize("squeeze", [False, True])
def test_multiblock_row_layout(squeeze):
    """cR/rR_nBlock_omp (bmSparseMat_cR_nBlock_omp_mex.cpp:97-483,
    bmSparseMat_rR_nBlock_omp_mex.cpp:97-436) on a matrix encoded as
    cpp_prepare('multi_block') writes it (bmSparseMat.m:131-165): the column
    cursor restarts at every block (first r_jump of a block absolute) and,
    for a squeezed matrix, so does the output-row cursor (first l_jump of a
    block absolute; the reference's l_sparsity branch).  Every row is still
    summed in stored order, so the products equal the one-block product bit
    for bit; they also agree with the dense float64 product."""
    print(f"[gpu] multi-block products squeeze={squeeze}", flush=True)
    rng = np.random.default_rng(5 + squeeze)
    M = OL.random_matrix(rng, 120, 90, 10, squeeze=squeeze)
    n = 2
    nl = M["l_nJump"]
    bl = [nl // 3, nl // 2 - nl // 3, nl - nl // 2]
    lbs = np.array([0, bl[0], bl[0] + bl[1]], np.int32)
    rowptr = np.concatenate([[0], np.cumsum(M["r_nJump"])])
    mbs = rowptr[lbs].astype(np.int64)
    col = np.cumsum(M["r_jump"].astype(np.int64))
    rj = M["r_jump"].copy()
    for m in mbs:
        if m < col.size:
            rj[m] = col[m]
    lj = None
    rows = np.arange(nl)
    if squeeze:
        rows = np.cumsum(M["l_jump"].astype(np.int64))
        lj = M["l_jump"].copy()
        lj[lbs] = rows[lbs]
        assert not np.array_equal(lj, M["l_jump"])
    vr = rng.normal(size=(M["r_size"], n)).astype(np.float32)
    vi = rng.normal(size=(M["r_size"], n)).astype(np.float32)
    er, ei = OL.call_cC(OL.port().port_cC_oBlock_omp, M, np.ascontiguousarray(vr.T), np.ascontiguousarray(vi.T))
    L = lib()
    blk = (3, OL.ip(np.array(bl, np.int32)), OL.ip(lbs), mbs.ctypes.data_as(C.POINTER(C.c_longlong)))
    wr = np.full((M["l_size"], n), np.nan, np.float32)
    wi = np.full((M["l_size"], n), np.nan, np.float32)
    L.bmSparseMat_cR_nBlock_omp(M["r_size"], OL.ip(rj), OL.ip(M["r_nJump"]), OL.fp(M["m_val"]), M["l_size"],
                                OL.ip(lj), nl, *blk, OL.fp(vr), OL.fp(vi), n, OL.fp(wr), OL.fp(wi))
    assert np.array_equal(wr.T.view(np.int32), er.view(np.int32))
    assert np.array_equal(wi.T.view(np.int32), ei.view(np.int32))
    rr = np.full((M["l_size"], n), np.nan, np.float32)
    L.bmSparseMat_rR_nBlock_omp(M["r_size"], OL.ip(rj), OL.ip(M["r_nJump"]), OL.fp(M["m_val"]), M["l_size"],
                                OL.ip(lj), nl, *blk, OL.fp(vr), n, OL.fp(rr))
    assert np.array_equal(rr.T.view(np.int32), er.view(np.int32))
    # dense float64 product of the same matrix
    A = np.zeros((M["l_size"], M["r_size"]))
    np.add.at(A, (np.repeat(rows, M["r_nJump"]), col), M["m_val"].astype(np.float64))
    ref = A @ (vr.astype(np.float64) + 1j * vi.astype(np.float64))
    got = wr.astype(np.float64) + 1j * wi.astype(np.float64)
    assert np.linalg.norm(got - ref) <= 1e-6 * np.linalg.norm(ref)


def test_r_nJump2index_gpu():
    from monalisa_amd.sparse import bmSparseMat_r_nJump2index
    r = np.random.default_rng(2).integers(0, 30, size=10000).astype(np.int32)
    assert np.array_equal(bmSparseMat_r_nJump2index(r), np.repeat(np.arange(10000), r))


def test_sparse_apply_device_and_determinism():
    import torch
    from monalisa_amd.sparse import BmSparseMat
    from monalisa_amd._lib import check, ptr, stream_of
    rng = np.random.default_rng(9)
    M = OL.random_matrix(rng, 3000, 4000, 30, squeeze=True)
    S = BmSparseMat(M["r_size"], M["l_size"], M["r_jump"], M["r_nJump"], M["m_val"], M["l_jump"])
    n = 4
    v = (rng.normal(size=(M["r_size"], n)) + 1j * rng.normal(size=(M["r_size"], n))).astype(np.complex64)
    vd = torch.from_numpy(v).cuda()
    w1 = torch.empty((M["l_size"], n), dtype=torch.complex64, device="cuda")
    w2 = torch.empty_like(w1)
    L = lib()
    check(L.mona_sparse_apply(S.handle(), ptr(vd), ptr(w1), n, stream_of()))
    check(L.mona_sparse_apply(S.handle(), ptr(vd), ptr(w2), n, stream_of()))
    torch.cuda.synchronize()
    a = w1.cpu().numpy()
    assert np.array_equal(a.view(np.int32), w2.cpu().numpy().view(np.int32))
    er, ei = OL.call_cC(OL.port().port_cC_oBlock_omp, M, np.ascontiguousarray(v.real.T),
                        np.ascontiguousarray(v.imag.T))
    assert np.array_equal(a.real.T, er) and np.array_equal(a.imag.T, ei)


def test_dropin_matrix_cache_follows_in_place_edits():
    """The drop-in products keep the decoded device copy of a matrix between
    calls (keyed by its host arrays and a hash of their contents): repeated
    calls give identical results, and a value or structure edited in place
    in the same host arrays is decoded again (each result bit-exact against
    the C restatement)."""
    print("[gpu] drop-in matrix cache", flush=True)
    rng = np.random.default_rng(17)
    M = OL.random_matrix(rng, 400, 300, 20, squeeze=True)
    vr = rng.normal(size=(3, M["r_size"])).astype(np.float32)
    vi = rng.normal(size=(3, M["r_size"])).astype(np.float32)
    P = OL.port()
    L = lib()
    for step in range(4):
        er, ei = OL.call_cC(P.port_cC_oBlock_omp, M, vr, vi)
        for _ in range(2):
            gr, gi = gpu_cC(L.bmSparseMat_cC_oBlock_omp, M, vr, vi)
            assert np.array_equal(gr.view(np.int32), er.view(np.int32)), step
            assert np.array_equal(gi.view(np.int32), ei.view(np.int32)), step
        if step == 0:
            M["m_val"][17] *= -3.0                      # a value, same arrays
        elif step == 1:
            c = np.cumsum(M["r_jump"].astype(np.int64))
            k = int(np.argmax(c < M["r_size"] - 1))     # an entry whose column can move right by one
            M["r_jump"][k] += 1
            M["r_jump"][k + 1] -= 1
        elif step == 2:
            k = int(np.argmax(M["l_jump"][1:] >= 2))   # an output row that can move down by one
            M["l_jump"][k] += 1
            M["l_jump"][k + 1] -= 1
