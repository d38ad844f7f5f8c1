"""mona_sparse_apply_noalloc: the product w = S v with the long-row partial
sums in the library's file-scope device array instead of a per-matrix
buffer.  Through the C ABI its output equals mona_sparse_apply's bit for
bit on two calls in a row (the second with a smaller matrix and fewer
vectors, so the array still holds the first call's larger contents); a
matrix not yet on the device and a call whose partial sums exceed the
array are refused by host checks before any launch."""

import numpy as np
import pytest

from monalisa_amd import synth

pytestmark = pytest.mark.gpu

MONA_ERR_ARG, MONA_ERR_STATE = 1, 4
SCRATCH_BYTES = 64 << 20


def _n_chunk(M):
    r = M.r_nJump.astype(np.int64)
    long_rows = r[r > 64]
    return int(((long_rows + 63) // 64).sum())


def _gut(nSeg, nShot, nSample, N):
    from monalisa_amd.sparse import bmTraj2SparseMat
    t = synth.koosh_ball(nSeg, nShot, nSample, k_max=N / 4)
    ve = synth.ve_radial3(t, nSample, nSeg * nShot)
    return bmTraj2SparseMat(t, ve, [N] * 3, [0.5] * 3, "kaiser")["Gut"]


def _rand(shape, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randn(shape, dtype=torch.complex64, device="cuda", generator=g)


def _noalloc(M, v, n_vec):
    import torch
    from monalisa_amd._lib import lib, ptr, stream_of
    w = torch.full((M.l_size, n_vec), float("nan"), dtype=torch.complex64, device="cuda")
    st = lib.mona_sparse_apply_noalloc(M.handle(), ptr(v), ptr(w), n_vec, stream_of())
    torch.cuda.synchronize()
    return st, w


def test_noalloc_equals_apply_twice_then_smaller():
    import torch
    from monalisa_amd._lib import lib, last_error
    from monalisa_amd.sparse import bmSparseMat_apply
    A = _gut(10, 40, 64, 32)        # 25,600 samples on 32^3: the centre cells' rows are long
    B = _gut(6, 20, 48, 24)         # smaller matrix for the second call
    assert _n_chunk(A) > _n_chunk(B) > 0
    print(f"[gpu] noalloc: A {A.l_nJump} rows, {_n_chunk(A)} chunks; B {_n_chunk(B)} chunks", flush=True)

    # a matrix that has not been uploaded is refused (host check, nothing launched)
    vA = _rand((A.r_size, 4), 1)
    st, _ = _noalloc(A, vA, 4)
    assert st == MONA_ERR_STATE and "not on the device" in last_error()

    ref_A = bmSparseMat_apply(A, vA)             # uploads A
    torch.cuda.synchronize()
    st, w_A = _noalloc(A, vA, 4)
    assert st == 0, last_error()
    assert torch.equal(w_A.view(torch.float32), ref_A.view(torch.float32))

    vB = _rand((B.r_size, 2), 2)
    ref_B = bmSparseMat_apply(B, vB)
    torch.cuda.synchronize()
    st, w_B = _noalloc(B, vB, 2)
    assert st == 0, last_error()
    assert torch.equal(w_B.view(torch.float32), ref_B.view(torch.float32))
    # the first matrix again after the smaller call
    st, w_A2 = _noalloc(A, vA, 4)
    assert st == 0 and torch.equal(w_A2.view(torch.float32), ref_A.view(torch.float32))
    assert lib.mona_device_count() >= 1


def test_noalloc_refuses_scratch_beyond_maximum():
    import torch
    from monalisa_amd._lib import last_error
    from monalisa_amd.sparse import bmSparseMat_apply, bmTraj2SparseMat
    t = synth.radial2d_golden(2000, 64, k_max=8.0)
    M = bmTraj2SparseMat(t, synth.ve_radial2(t, 64, 2000), [32, 32], [0.5, 0.5], "kaiser")["Gut"]
    nc = _n_chunk(M)
    n_vec = SCRATCH_BYTES // (8 * nc) + 1
    assert nc * n_vec * 8 > SCRATCH_BYTES
    print(f"[gpu] noalloc limit: {nc} chunks x {n_vec} vectors", flush=True)
    bmSparseMat_apply(M, _rand((M.r_size, 1), 3))   # uploads M
    torch.cuda.synchronize()
    v = torch.zeros((M.r_size, n_vec), dtype=torch.complex64, device="cuda")
    st, _ = _noalloc(M, v, n_vec)
    assert st == MONA_ERR_ARG
    assert "64" in last_error() or "67108864" in last_error()
    # one vector fewer than the limit allows still runs and equals the allocating product
    n_ok = SCRATCH_BYTES // (8 * nc)
    v = _rand((M.r_size, n_ok), 4)
    st, w = _noalloc(M, v, n_ok)
    assert st == 0, last_error()
    assert torch.equal(w.view(torch.float32), bmSparseMat_apply(M, v).view(torch.float32))
