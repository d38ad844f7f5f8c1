"""The adjoint gridding from LDS-staged groups of neighbouring grid cells
(the plan's default) against the per-entry gather of the same tile-order
Gut: equal bit for bit (the same per-row operation sequence), with and
without the point weights, on 2-D and 3-D frames, odd and even coil
counts, a sparse frame, rows longer than 64 entries, bf16 storage, the
batched form, and against the float64 oracle."""
import numpy as np
import pytest

from problems import make_problem, oracle_mats, nd, rel
from oracle import mona_oracle as O

pytestmark = pytest.mark.gpu

CASES = [("2d", 3), ("3d", 4), ("3d32", 20), ("3d32s", 2), ("2d_rect", 1), ("3d32", 5), ("3d32", 8)]


def _plan(p):
    import torch
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    G = bmTraj2SparseMat(p["t"], p["ve"], p["N_u"], p["dK_u"], p["kernel"])
    P = Plan(p["N_u"], p["n_u"], p["dK_u"], p["nCh"], p["kernel"])
    P.set_coils(torch.from_numpy(p["C"]).cuda())
    P.set_frame(0, G["Gu"], G["Gut"])
    return P, G


def _both(P, fn):
    P.set_options(staged_adjoint=True)
    a = fn()
    P.set_options(staged_adjoint=False)
    b = fn()
    P.set_options(staged_adjoint=True)
    return a, b


@pytest.mark.parametrize("kind,nCh", CASES)
def test_staged_adjoint_bit_exact(kind, nCh):
    import torch
    print(f"[gpu] staged adjoint {kind} nCh={nCh}", flush=True)
    p = make_problem(kind, seed=7, nCh=nCh)
    P, _ = _plan(p)
    rng = np.random.default_rng(3)
    yy = (rng.normal(size=(P.nPt(), nCh)) + 1j * rng.normal(size=(P.nPt(), nCh))).astype(np.complex64)
    yd = torch.from_numpy(yy).cuda()
    a, b = _both(P, lambda: P.bmNakatsha(yd).cpu().numpy())
    np.testing.assert_array_equal(a, b)
    a, b = _both(P, lambda: P.bmNakatsha(yd, C_flag=False).cpu().numpy())
    np.testing.assert_array_equal(a, b)
    # weighted form (point weights from ve) through one CG-SENSE step
    x0 = torch.from_numpy(p["x"]).cuda()
    ve = torch.from_numpy(p["ve"].astype(np.float32)).cuda()

    def cg():
        x = x0.clone()
        P.bmSensa(x, yd, ve, 2)
        return x.cpu().numpy()
    a, b = _both(P, cg)
    np.testing.assert_array_equal(a, b)
    # and the float64 oracle
    R = oracle_mats(p)
    Cn = nd(p["C"].astype(np.complex128), p["n_u"], nCh)
    KFCc = O.kf_conj(np.conj(Cn), p["N_u"], p["n_u"], p["dK_u"], p["kernel"])
    xr = O.nakatsha(yy.astype(np.complex128), R["Gut"], KFCc, True, p["N_u"], p["n_u"])
    assert rel(P.bmNakatsha(yd).cpu().numpy(), xr.reshape(-1, order="F")) < 1e-5


def test_staged_adjoint_long_rows_and_bf16():
    """A trajectory dense at the k-space centre: rows over 64 entries are
    summed by ranges of consecutive samples beside the staged groups, the
    same in both forms."""
    import torch
    from monalisa_amd import synth
    print("[gpu] staged adjoint, long rows, bf16", flush=True)
    p = make_problem("3d32", seed=6, nCh=20)
    t = synth.koosh_ball(20, 30, 32, k_max=8.0)
    p["t"], p["ve"] = t, synth.ve_radial3(t, 32, 600)
    P, G = _plan(p)
    assert int(G["Gut"].r_nJump.max()) > 64
    rng = np.random.default_rng(5)
    yy = torch.from_numpy((rng.normal(size=(P.nPt(), 20)) + 1j * rng.normal(size=(P.nPt(), 20)))
                          .astype(np.complex64)).cuda()
    a, b = _both(P, lambda: P.bmNakatsha(yy).cpu().numpy())
    np.testing.assert_array_equal(a, b)
    a, b = _both(P, lambda: P.bmNakatsha_bf16(yy).cpu().numpy())
    np.testing.assert_array_equal(a, b)


def test_staged_adjoint_batched():
    """The block-diagonal batched operators stack the frames' groups: the
    batched CG-SENSE equals the per-frame one bit for bit, staged or not,
    and the staged and per-entry forms agree."""
    import torch
    from monalisa_amd import synth
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    print("[gpu] staged adjoint, batched", flush=True)
    rng = np.random.default_rng(8)
    n_u, N_u, dK_u, nCh = [16, 16, 16], [32, 32, 32], [0.5] * 3, 8
    P = Plan(N_u, n_u, dK_u, nCh, "kaiser")
    P.set_coils(torch.from_numpy(synth.coil_maps(n_u, nCh, rng)).cuda())
    t_all = synth.koosh_ball(8, 10, 32, k_max=8.0)
    ve_all = synth.ve_radial3(t_all, 32, 80)
    for f, (a, b) in enumerate(((0, 40), (40, 70), (70, 80))):
        sl = slice(a * 32, b * 32)
        G = bmTraj2SparseMat(t_all[:, sl], ve_all[sl], N_u, dK_u, "kaiser")
        P.set_frame(f, G["Gu"], G["Gut"])
    ys = [torch.from_numpy((rng.normal(size=(P.nPt(f), nCh)) + 1j * rng.normal(size=(P.nPt(f), nCh)))
                           .astype(np.complex64)).cuda() for f in range(3)]
    ves = [torch.from_numpy(np.abs(rng.normal(size=P.nPt(f))).astype(np.float32) + 0.1).cuda() for f in range(3)]
    x0 = torch.zeros((3, int(np.prod(n_u))), dtype=torch.complex64, device="cuda")

    def run():
        xb = x0.clone()
        P.bmSensa_batch(xb, ys, ves, 3)
        xf = []
        for f in range(3):
            x = x0[f].clone()
            P.bmSensa(x, ys[f], ves[f], 3, f=f)
            xf.append(x.cpu().numpy())
        return xb.cpu().numpy(), np.stack(xf)
    (sb, sf), (pb, pf) = _both(P, run)
    np.testing.assert_array_equal(sb, sf)
    np.testing.assert_array_equal(sb, pb)
    np.testing.assert_array_equal(pb, pf)


def test_group_autotune_keeps_results():
    """The group-size autotuner regroups every frame's adjoint gridding; the
    adjoint of each frame and the batched CG-SENSE keep their bits, and the
    batched form restacks the new groups."""
    import torch
    from monalisa_amd import synth
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    print("[gpu] group autotune", flush=True)
    rng = np.random.default_rng(12)
    n_u, N_u, dK_u, nCh = [16, 16, 16], [32, 32, 32], [0.5] * 3, 20
    P = Plan(N_u, n_u, dK_u, nCh, "kaiser")
    P.set_coils(torch.from_numpy(synth.coil_maps(n_u, nCh, rng)).cuda())
    t_all = synth.koosh_ball(10, 12, 32, k_max=8.0)
    ve_all = synth.ve_radial3(t_all, 32, 120)
    for f, (a, b) in enumerate(((0, 60), (60, 120))):
        sl = slice(a * 32, b * 32)
        G = bmTraj2SparseMat(t_all[:, sl], ve_all[sl], N_u, dK_u, "kaiser")
        P.set_frame(f, G["Gu"], G["Gut"])
    ys = [torch.from_numpy((rng.normal(size=(P.nPt(f), nCh)) + 1j * rng.normal(size=(P.nPt(f), nCh)))
                           .astype(np.complex64)).cuda() for f in range(2)]
    ves = [torch.from_numpy(np.abs(rng.normal(size=P.nPt(f))).astype(np.float32) + 0.1).cuda() for f in range(2)]

    def run():
        adj = [P.bmNakatsha(ys[f], f=f).cpu().numpy() for f in range(2)]
        xb = torch.zeros((2, P.nu), dtype=torch.complex64, device="cuda")
        P.bmSensa_batch(xb, ys, ves, 2)
        return adj, xb.cpu().numpy()
    adj0, xb0 = run()
    kb, rows = P.autotune_groups(0)
    assert kb in (12, 16, 24) and rows in (128, 256, 512)
    adj1, xb1 = run()
    for f in range(2):
        np.testing.assert_array_equal(adj1[f], adj0[f])
    np.testing.assert_array_equal(xb1, xb0)


def test_staged_long_rows_batched():
    """Frames dense at the k-space centre (rows over 64 entries, summed by
    ranges of consecutive samples beside the staged groups): the batched
    CG-SENSE, whose stacked matrices stack the frames' ranges, equals the
    per-frame one bit for bit, staged or not."""
    import torch
    from monalisa_amd import synth
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    print("[gpu] staged adjoint, long rows, batched", flush=True)
    rng = np.random.default_rng(21)
    n_u, N_u, dK_u, nCh = [16, 16, 16], [32, 32, 32], [0.5] * 3, 4
    P = Plan(N_u, n_u, dK_u, nCh, "kaiser")
    P.set_coils(torch.from_numpy(synth.coil_maps(n_u, nCh, rng)).cuda())
    t_all = synth.koosh_ball(20, 30, 32, k_max=8.0)
    ve_all = synth.ve_radial3(t_all, 32, 600)
    for f, (a, b) in enumerate(((0, 350), (350, 600))):
        sl = slice(a * 32, b * 32)
        G = bmTraj2SparseMat(t_all[:, sl], ve_all[sl], N_u, dK_u, "kaiser")
        assert int(G["Gut"].r_nJump.max()) > 64
        P.set_frame(f, G["Gu"], G["Gut"])
    ys = [torch.from_numpy((rng.normal(size=(P.nPt(f), nCh)) + 1j * rng.normal(size=(P.nPt(f), nCh)))
                           .astype(np.complex64)).cuda() for f in range(2)]
    ves = [torch.from_numpy(np.abs(rng.normal(size=P.nPt(f))).astype(np.float32) + 0.1).cuda() for f in range(2)]
    x0 = torch.zeros((2, P.nu), dtype=torch.complex64, device="cuda")

    def run():
        xb = x0.clone()
        P.bmSensa_batch(xb, ys, ves, 3)
        xf = []
        for f in range(2):
            x = x0[f].clone()
            P.bmSensa(x, ys[f], ves[f], 3, f=f)
            xf.append(x.cpu().numpy())
        return xb.cpu().numpy(), np.stack(xf)
    (sb, sf), (pb, pf) = _both(P, run)
    np.testing.assert_array_equal(sb, sf)
    np.testing.assert_array_equal(sb, pb)
    np.testing.assert_array_equal(pb, pf)
