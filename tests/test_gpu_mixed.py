"""Mixed precision on the GPU: the operators with bf16 storage of the
grid-side arrays and k-space against the float64 oracle, and CG-SENSE by
iterative refinement whose result is checked with the oracle's float64
normal-equation residual and its float64 converged solution."""
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


def test_sensa_refined_reaches_1e5():
    """Iterative refinement to a relative normal-equation residual below
    1e-5 (bmSensa.m:39-85's system; the mixed-precision target) on
    consistent data y = M x, with bf16 and with float storage: the solver's
    residual and the float64 oracle's residual of the returned image are
    both below 1e-5.  With bf16 storage and no residual replacement the same
    CG stalls above 1e-5: the float replacement is what restores the
    float-level residual.  The system is ill-conditioned (a residual of
    1e-5 pins the image only to about cond * 1e-5), so the image gaps --
    bf16 against float refinement, and against a float64 oracle CG of the
    same step count -- are printed and bounded loosely."""
    import torch
    print("[gpu] sensa refined to 1e-5: bf16 and float", flush=True)
    p, P, R, KFC, KFCc = _setup()
    xt = torch.from_numpy(p["x"]).cuda()
    y = P.bmShanna(xt).cpu().numpy()
    ve = p["ve"].astype(np.float32)
    yd, ved = torch.from_numpy(y).cuda(), torch.from_numpy(ve).cuda()
    tol = 1e-5
    out = {}
    for bf16 in (True, False):
        x = torch.zeros_like(xt)
        hist = P.sensa_refined(x, yd, ved, nInner=10, nOuter=60, tol=tol, bf16=bf16)
        assert hist[0] == 1.0
        assert hist[-1] < tol, hist
        xg = x.cpu().numpy()
        res64 = _normal_residual(xg, y, ve, p, R, KFC, KFCc)
        assert res64 < 1.05 * tol, res64
        out[bf16] = (xg, hist, res64)
    # float64 oracle CG-SENSE with as many steps as the float refinement took
    # (data scaled by 1e6 so bmSensa's absolute early exit, bmSensa.m:54,
    # does not stop it first; CG is otherwise scale invariant)
    Cn = nd(p["C"].astype(np.complex128), p["n_u"], p["nCh"])
    k = out[False][1].steps
    x64, _ = O.sensa(np.zeros(p["n_u"], complex), 1e6 * y.astype(np.complex128), ve.astype(np.float64), Cn,
                     R["Gu"], R["Gut"], p["N_u"], p["n_u"], p["dK_u"], k, None, 1, kernel=(p["kernel"], None, None))
    x64 = flat(x64, p["n_u"]) / 1e6
    gap16_32 = rel(out[True][0], out[False][0])
    gap32_64 = rel(out[False][0], x64)
    gap16_64 = rel(out[True][0], x64)
    print(f"steps bf16 {out[True][1].steps} float {k}; float64 residuals {out[True][2]:.2e} {out[False][2]:.2e}; "
          f"image gaps bf16-float {gap16_32:.2e}, float-float64 {gap32_64:.2e}, bf16-float64 {gap16_64:.2e}",
          flush=True)
    assert gap16_32 < 5e-3 and gap32_64 < 5e-3 and gap16_64 < 5e-3
    x = torch.zeros_like(xt)
    plain = P.sensa_refined(x, yd, ved, nInner=out[True][1].steps + 100, nOuter=1, tol=0.0, bf16=True)
    assert plain[-1] > tol, plain


def test_sensa_refined_bf16_tracks_fp32_on_noisy_data():
    """On noisy data (the normal equations converge slowly) the bf16 and the
    float refinement reach the same residual after the same replacements,
    to within 10%: bf16 storage perturbs each operator by about 2^-9, which
    moves CG's Krylov sequence, so the two residual histories differ by a
    few percent on a slowly converging system, without a drift between
    them."""
    import torch
    print("[gpu] sensa refined, noisy data: bf16 vs float", flush=True)
    p, P, R, KFC, KFCc = _setup()
    xt = torch.from_numpy(p["x"]).cuda()
    y = P.bmShanna(xt).cpu().numpy()
    rng = np.random.default_rng(11)
    y = (y + 0.01 * np.abs(y).mean() * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape))).astype(
        np.complex64)
    yd, ved = torch.from_numpy(y).cuda(), torch.from_numpy(p["ve"].astype(np.float32)).cuda()
    hs = []
    for bf16 in (False, True):
        x = torch.zeros_like(xt)
        hs.append(P.sensa_refined(x, yd, ved, nInner=10, nOuter=10, tol=0.0, bf16=bf16))
    h32, h16 = np.array(hs[0]), np.array(hs[1])
    assert len(h32) == len(h16) == 11
    assert np.all(np.abs(h16[1:] - h32[1:]) <= 0.1 * h32[1:]), (h32, h16)
    assert h16[-1] <= 1.1 * h32[-1]


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
