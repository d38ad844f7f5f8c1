"""The C restatement of the reference's native kernels (oracle/port.c) is
pinned bit-exactly against the reference's own compiled kernel and against
the committed golden vectors written by that kernel."""
import os

import numpy as np
import pytest

import oracle_libs as OL
from oracle import mona_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "spmv_golden.npz")


def load_golden():
    z = np.load(GOLD, allow_pickle=False)
    out = []
    for name in z["case_names"]:
        p = str(name) + "__"
        r_size, l_size, l_nJump, n_vec = [int(v) for v in z[p + "sizes"]]
        M = dict(r_size=r_size, l_size=l_size, l_nJump=l_nJump, r_jump=z[p + "r_jump"],
                 r_nJump=z[p + "r_nJump"], m_val=z[p + "m_val"],
                 l_jump=z[p + "l_jump"] if (p + "l_jump") in z.files else None)
        out.append((str(name), M, z[p + "v_re"], z[p + "v_im"], z[p + "w_re"], z[p + "w_im"]))
    return out


@pytest.mark.parametrize("case", load_golden(), ids=lambda c: c[0])
def test_port_matches_golden(case):
    name, M, vr, vi, wr, wi = case
    P = OL.port()
    for fn in (P.port_cC_oBlock_omp, P.port_cC_oBlock):
        gr, gi = OL.call_cC(fn, M, vr, vi)
        assert np.array_equal(gr.view(np.int32), wr.view(np.int32))
        assert np.array_equal(gi.view(np.int32), wi.view(np.int32))


@pytest.mark.parametrize("seed", range(6))
def test_port_matches_reference_build(seed):
    ref = OL.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(seed)
    M = OL.random_matrix(rng, int(rng.integers(2, 200)), int(rng.integers(2, 300)),
                         int(rng.integers(1, 40)), squeeze=bool(seed % 2))
    n_vec = int(rng.integers(1, 6))
    vr = rng.normal(size=(n_vec, M["r_size"])).astype(np.float32)
    vi = rng.normal(size=(n_vec, M["r_size"])).astype(np.float32)
    a = OL.call_cC(OL.port().port_cC_oBlock_omp, M, vr, vi)
    b = OL.call_cC(ref.ref_cC_oBlock_omp, M, vr, vi)
    assert np.array_equal(a[0].view(np.int32), b[0].view(np.int32))
    assert np.array_equal(a[1].view(np.int32), b[1].view(np.int32))


def test_port_vs_dense_float64():
    rng = np.random.default_rng(7)
    M = OL.random_matrix(rng, 50, 70, 15, squeeze=True)
    vr = rng.normal(size=(2, 70)).astype(np.float32)
    vi = rng.normal(size=(2, 70)).astype(np.float32)
    wr, wi = OL.call_cC(OL.port().port_cC_oBlock_omp, M, vr, vi)
    G = dict(M, type="l_squeezed_cpp_prepared")
    A = O.to_dense(G)
    ref = A @ (vr.astype(np.float64) + 1j * vi.astype(np.float64)).T
    got = (wr + 1j * wi).T
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-6


def test_row_layout_variants_match_column_layout():
    """cR/rR/rC restatements equal the cC sums (same per-element sequence)."""
    rng = np.random.default_rng(3)
    P = OL.port()
    for sq in (False, True):
        M = OL.random_matrix(rng, 30, 40, 8, squeeze=sq)
        n = 3
        vr = rng.normal(size=(n, 40)).astype(np.float32)
        vi = rng.normal(size=(n, 40)).astype(np.float32)
        wr, wi = OL.call_cC(P.port_cC_oBlock_omp, M, vr, vi)
        # row layout: element (col, k) at col*n + k
        rr = np.empty((M["l_size"], n), np.float32)
        ri = np.empty((M["l_size"], n), np.float32)
        vrr = np.ascontiguousarray(vr.T)
        vir = np.ascontiguousarray(vi.T)
        P.port_cR_oBlock(M["r_size"], OL.ip(M["r_jump"]), OL.ip(M["r_nJump"]), OL.fp(M["m_val"]),
                         M["l_size"], OL.ip(M["l_jump"]), M["l_nJump"], OL.fp(vrr), OL.fp(vir), n,
                         OL.fp(rr), OL.fp(ri))
        assert np.array_equal(rr.T, wr) and np.array_equal(ri.T, wi)
        rc = np.empty((n, M["l_size"]), np.float32)
        P.port_rC_oBlock(M["r_size"], OL.ip(M["r_jump"]), OL.ip(M["r_nJump"]), OL.fp(M["m_val"]),
                         M["l_size"], OL.ip(M["l_jump"]), M["l_nJump"], OL.fp(vr), n, OL.fp(rc))
        assert np.array_equal(rc, wr)


def test_r_nJump2index_closed_form():
    rng = np.random.default_rng(1)
    r = rng.integers(0, 7, size=50).astype(np.int32)
    out = np.empty(int(r.sum()), np.int32)
    OL.port().port_r_nJump2index(50, OL.ip(r), OL.ip(out))
    assert np.array_equal(out, np.repeat(np.arange(50), r))
    assert np.array_equal(O.r_nJump2index(r), out)


def leight_bruteforce(imDim, a, t, Dn, N, nWin, sigma):
    """Per-cell restatement of the normalised Gaussian n2u (cell loop instead
    of the reference's point loop), float64."""
    nPt = t.shape[1]
    Nu = int(np.prod(N))
    b = np.zeros((Nu, a.shape[1]), np.complex128)
    nrm = np.zeros(Nu)
    h = nWin // 2
    g1 = 1 / (np.sqrt(2 * 3.141592) * sigma)
    for p in range(nPt):
        base = [int(np.floor(t[d, p] + 0.5)) - 1 for d in range(imDim)]
        rr = [t[d, p] - (base[d] + 1) for d in range(imDim)]
        offs = np.stack(np.meshgrid(*([np.arange(-h, h + 1)] * imDim), indexing="ij"), -1).reshape(-1, imDim)
        for o in offs:
            cell = [base[d] + o[d] for d in range(imDim)]
            if any(c < 0 or c >= N[d] for d, c in enumerate(cell)):
                continue
            d2 = sum((rr[d] - o[d]) ** 2 for d in range(imDim))
            w = g1 * np.exp(-d2 / (2 * sigma ** 2)) * Dn[p]
            lin = cell[0] + (cell[1] * N[0] if imDim > 1 else 0) + (cell[2] * N[0] * N[1] if imDim > 2 else 0)
            b[lin] += w * a[p]
            nrm[lin] += w
    nrm[nrm < 1e-5] = 1.0
    return b / nrm[:, None]


@pytest.mark.parametrize("imDim", [1, 2, 3])
def test_port_leight_vs_bruteforce(imDim):
    rng = np.random.default_rng(imDim)
    N = [12, 10, 8][:imDim]
    nPt, nCh = 60, 2
    t = np.stack([rng.uniform(1, N[d], nPt) for d in range(imDim)], 0).astype(np.float32)
    a = (rng.normal(size=(nPt, nCh)) + 1j * rng.normal(size=(nPt, nCh))).astype(np.complex64)
    Dn = rng.uniform(0.5, 1.5, nPt).astype(np.float32)
    ar = np.ascontiguousarray(a.real)
    ai = np.ascontiguousarray(a.imag)
    Nu = int(np.prod(N))
    br = np.empty((Nu, nCh), np.float32)
    bi = np.empty((Nu, nCh), np.float32)
    tz = [t[d] if d < imDim else np.ones(nPt, np.float32) for d in range(3)]
    tz = [np.ascontiguousarray(v) for v in tz]
    Nz = N + [1] * (3 - imDim)
    OL.port().port_gridder_n2u_leight(imDim, OL.fp(ar), OL.fp(ai), OL.fp(br), OL.fp(bi), OL.fp(tz[0]),
                                      OL.fp(tz[1]), OL.fp(tz[2]), OL.fp(Dn), nCh, nPt, Nz[0], Nz[1], Nz[2],
                                      3, 0.61, 10.0)
    ref = leight_bruteforce(imDim, a.astype(np.complex128), t.astype(np.float64), Dn, N, 3, 0.61)
    got = br + 1j * bi
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
