"""Mixed precision on the GPU: the operators with bf16 storage of the
grid-side arrays and k-space against the float64 oracle, and CG-SENSE by
iterative refinement whose result is checked with the oracle's float64
normal-equation residual."""
import numpy as np
import pytest

from problems import make_problem, oracle_mats, nd, flat, rel
from oracle import mona_oracle as O

pytestmark = pytest.mark.gpu


def _setup(kind="3d32", nCh=4, seed=3):
    import torch
    from test_gpu_ops import _plan
    p = make_problem(kind, seed=seed, nCh=nCh)
    P = _plan(p)
    R = oracle_mats(p)
    Cn = nd(p["C"].astype(np.complex128), p["n_u"], p["nCh"])
    KFC = O.kf(Cn, p["N_u"], p["n_u"], p["dK_u"], p["kernel"])
    KFCc = O.kf_conj(np.conj(Cn), p["N_u"], p["n_u"], p["dK_u"], p["kernel"])
    return p, P, R, KFC, KFCc


@pytest.mark.parametrize("kind", ["3d32", "2d_rect"])
def test_bf16_operators_vs_oracle(kind):
    """bf16 storage (8-bit significand): the operators stay within a few
    2^-9 roundings of the float64 result, and the error is well above the
    float operators' (so the reduced storage is really in use)."""
    import torch
    print(f"[gpu] bf16 operators {kind}", flush=True)
    p, P, R, KFC, KFCc = _setup(kind)
    x = torch.from_numpy(p["x"]).cuda()
    yref = O.shanna(nd(p["x"].astype(np.complex128), p["n_u"]), R["Gu"], KFC, p["N_u"], p["n_u"])
    y16 = P.bmShanna_bf16(x).cpu().numpy()
    y32 = P.bmShanna(x).cpu().numpy()
    e16, e32 = rel(y16, yref), rel(y32, yref)
    assert e32 < 1e-5
    assert 1e-4 < e16 < 1e-2
    rng = np.random.default_rng(1)
    yy = (rng.normal(size=y32.shape) + 1j * rng.normal(size=y32.shape)).astype(np.complex64)
    xref = flat(O.nakatsha(yy.astype(np.complex128), R["Gut"], KFCc, True, p["N_u"], p["n_u"]), p["n_u"])
    x16 = P.bmNakatsha_bf16(torch.from_numpy(yy).cuda()).cpu().numpy()
    assert 1e-4 < rel(x16, xref) < 1e-2


def _normal_residual(x, y, ve, p, R, KFC, KFCc):
    """||M^H HY (y - M x)|| / ||M^H HY y|| in float64 with the oracle."""
    hy = np.minimum(ve, ve.max()).astype(np.float64)[:, None]
    xn = nd(x.astype(np.complex128), p["n_u"])
    r = y.astype(np.complex128) - O.shanna(xn, R["Gu"], KFC, p["N_u"], p["n_u"])
    g = O.nakatsha(hy * r, R["Gut"], KFCc, True, p["N_u"], p["n_u"])
    g0 = O.nakatsha(hy * y.astype(np.complex128), R["Gut"], KFCc, True, p["N_u"], p["n_u"])
    return float(np.linalg.norm(g) / np.linalg.norm(g0))


@pytest.mark.parametrize("bf16", [True, False])
def test_sensa_refined_restores_fp32_level(bf16):
    """The refinement with bf16 products reaches the relative normal-equation
    residual of 20 float CG-SENSE steps (or 1e-5 if that is lower), as the
    float64 oracle measures it on the returned image."""
    import torch
    print(f"[gpu] sensa refined bf16={bf16}", flush=True)
    p, P, R, KFC, KFCc = _setup()
    xt = torch.from_numpy(p["x"]).cuda()
    y = P.bmShanna(xt).cpu().numpy()
    rng = np.random.default_rng(11)
    y = (y + 0.01 * np.abs(y).mean() * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape))).astype(
        np.complex64)
    ve = p["ve"].astype(np.float32)
    yd, ved = torch.from_numpy(y).cuda(), torch.from_numpy(ve).cuda()
    x32 = torch.zeros_like(xt)
    P.bmSensa(x32, yd, ved, 20, nIter=1)
    rho32 = P.sensa_refined(x32.clone(), yd, ved, nInner=1, nOuter=0)[0]
    # the solver's residual of a given image is the oracle's float64 one
    assert abs(_normal_residual(x32.cpu().numpy(), y, ve, p, R, KFC, KFCc) - rho32) < 0.02 * rho32
    tol = max(1e-5, rho32)
    x = torch.zeros_like(xt)
    hist = P.sensa_refined(x, yd, ved, nInner=10, nOuter=40, tol=tol, bf16=bf16)
    assert hist[0] == 1.0
    assert hist[-1] < tol, (tol, hist)
    assert hist.steps <= 40 * 10
    assert _normal_residual(x.cpu().numpy(), y, ve, p, R, KFC, KFCc) < 1.05 * tol


def test_sensa_refined_deterministic():
    import torch
    print("[gpu] sensa refined determinism", flush=True)
    p, P, R, KFC, KFCc = _setup(seed=4)
    xt = torch.from_numpy(p["x"]).cuda()
    y = P.bmShanna(xt)
    ve = torch.from_numpy(p["ve"].astype(np.float32)).cuda()
    outs = []
    for _ in range(2):
        x = torch.zeros_like(xt)
        h = P.sensa_refined(x, y, ve, nInner=5, nOuter=3, tol=0.0, bf16=True)
        torch.cuda.synchronize()
        outs.append((x.cpu().numpy(), h))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]
    assert len(outs[0][1]) == 4 and outs[0][1].steps == 15
