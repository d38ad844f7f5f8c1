"""GPU direct gridders (leight, precomputed weights) and the in-box pair
count against the oracle's C restatement of the reference MEX kernels."""
import numpy as np
import pytest

import oracle_libs as OL

pytestmark = pytest.mark.gpu


def _points(rng, imDim, N, nPt):
    t = [rng.uniform(0.6, N[d] + 0.4, nPt).astype(np.float32) for d in range(imDim)]
    while len(t) < 3:
        t.append(np.ones(nPt, np.float32))
    return [np.ascontiguousarray(v) for v in t]


@pytest.mark.parametrize("imDim,N,nCh,nWin", [(1, [40], 2, 3), (2, [24, 20], 3, 3), (3, [12, 10, 14], 2, 3),
                                              (2, [16, 16], 1, 4), (3, [8, 8, 8], 4, 5)])
def test_leight_vs_port(imDim, N, nCh, nWin):
    from monalisa_amd import lib
    print(f"[gpu] leight imDim={imDim} N={N} nWin={nWin}", flush=True)
    rng = np.random.default_rng(imDim * 10 + nWin)
    nPt = 300
    tx, ty, tz = _points(rng, imDim, N, nPt)
    ar = rng.normal(size=(nPt, nCh)).astype(np.float32)
    ai = rng.normal(size=(nPt, nCh)).astype(np.float32)
    Dn = rng.uniform(0.2, 1.5, nPt).astype(np.float32)
    Nz = N + [1] * (3 - imDim)
    Nu = int(np.prod(N))
    er = np.empty((Nu, nCh), np.float32)
    ei = np.empty((Nu, nCh), np.float32)
    OL.port().port_gridder_n2u_leight(imDim, OL.fp(ar), OL.fp(ai), OL.fp(er), OL.fp(ei), OL.fp(tx), OL.fp(ty),
                                      OL.fp(tz), OL.fp(Dn), nCh, nPt, Nz[0], Nz[1], Nz[2], nWin, 0.61, 10.0)
    gr = np.full((Nu, nCh), np.nan, np.float32)
    gi = np.full((Nu, nCh), np.nan, np.float32)
    if imDim == 1:
        lib.bmGridder_n2u_leight1_mex(OL.fp(ar), OL.fp(ai), OL.fp(gr), OL.fp(gi), OL.fp(tx), OL.fp(Dn), nCh, nPt,
                                      N[0], nWin, 0.61, 10.0)
    elif imDim == 2:
        lib.bmGridder_n2u_leight2_mex(OL.fp(ar), OL.fp(ai), OL.fp(gr), OL.fp(gi), OL.fp(tx), OL.fp(ty), OL.fp(Dn),
                                      nCh, nPt, N[0], N[1], nWin, 0.61, 10.0)
    else:
        lib.bmGridder_n2u_leight3_mex(OL.fp(ar), OL.fp(ai), OL.fp(gr), OL.fp(gi), OL.fp(tx), OL.fp(ty), OL.fp(tz),
                                      OL.fp(Dn), nCh, nPt, N[0], N[1], N[2], nWin, 0.61, 10.0)
    e = er + 1j * ei
    g = gr + 1j * gi
    assert np.isfinite(g).all()
    assert np.linalg.norm(g - e) / np.linalg.norm(e) < 1e-5
    # the untouched cells are exactly zero, as in the reference
    assert np.array_equal(g == 0, e == 0)


@pytest.mark.parametrize("nNb,nCh", [(9, 2), (27, 3)])
def test_precomputed_gridder_bit_exact_vs_port(nNb, nCh):
    from monalisa_amd import lib
    rng = np.random.default_rng(nNb)
    nPt, Nu = 500, 2000
    cells = rng.integers(0, Nu, size=(nPt, nNb))
    flat = cells.reshape(-1)
    nJump = np.diff(np.concatenate([[0], flat])).astype(np.int32)
    w = rng.uniform(0.01, 1.0, size=(nPt, nNb)).astype(np.float32)
    ar = rng.normal(size=(nPt, nCh)).astype(np.float32)
    ai = rng.normal(size=(nPt, nCh)).astype(np.float32)
    er = np.empty((Nu, nCh), np.float32)
    ei = np.empty((Nu, nCh), np.float32)
    OL.port().port_gridder_n2u(OL.fp(ar), OL.fp(ai), OL.fp(er), OL.fp(ei), OL.fp(w), OL.ip(nJump), nNb, nCh, nPt, Nu)
    gr = np.full((Nu, nCh), np.nan, np.float32)
    gi = np.full((Nu, nCh), np.nan, np.float32)
    lib.bmGridder_n2u_mex(OL.fp(ar), OL.fp(ai), OL.fp(gr), OL.fp(gi), OL.fp(w), OL.ip(nJump), nNb, nCh, nPt, Nu)
    assert np.array_equal(gr.view(np.int32), er.view(np.int32))
    assert np.array_equal(gi.view(np.int32), ei.view(np.int32))


@pytest.mark.parametrize("imDim,N,nWin", [(1, [30], 3), (2, [20, 18], 3), (3, [10, 12, 9], 3), (3, [10, 10, 10], 4)])
def test_secret_length_vs_port(imDim, N, nWin):
    import ctypes as C
    from monalisa_amd import lib
    rng = np.random.default_rng(sum(N))
    nPt = 400
    tx, ty, tz = _points(rng, imDim, N, nPt)
    Nz = N + [1] * (3 - imDim)
    ref = OL.port().port_secret_length(imDim, OL.fp(tx), OL.fp(ty), OL.fp(tz), nPt, Nz[0], Nz[1], Nz[2], nWin)
    out = C.c_longlong(-1)
    if imDim == 1:
        lib.bmGriddingMatrix_secret_length1_mex(OL.fp(tx), 1, nPt, N[0], nWin, C.byref(out))
    elif imDim == 2:
        lib.bmGriddingMatrix_secret_length2_mex(OL.fp(tx), OL.fp(ty), 1, nPt, N[0], N[1], nWin, C.byref(out))
    else:
        lib.bmGriddingMatrix_secret_length3_mex(OL.fp(tx), OL.fp(ty), OL.fp(tz), 1, nPt, N[0], N[1], N[2], nWin,
                                                C.byref(out))
    assert out.value == ref and ref > 0
