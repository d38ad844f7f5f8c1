"""ADMM-TV solvers and the gridded reconstruction on the GPU against the
float64 oracle restatements."""
import numpy as np
import pytest

from problems import make_problem, oracle_mats, nd, flat, rel
from oracle import mona_oracle as O

pytestmark = pytest.mark.gpu


def _data(P, p, seed=1):
    import torch
    rng = np.random.default_rng(seed)
    y = P.bmShanna(torch.from_numpy(p["x"]).cuda()).cpu().numpy()
    y = y + 0.02 * np.abs(y).mean() * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape))
    return y.astype(np.complex64)


@pytest.mark.parametrize("kind", ["2d", "3d"])
def test_steva_vs_oracle(kind):
    import torch
    from test_gpu_ops import _plan
    print(f"[gpu] steva {kind}", flush=True)
    p = make_problem(kind, seed=7, nCh=2)
    P = _plan(p)
    R = oracle_mats(p)
    y = _data(P, p)
    ve = p["ve"].astype(np.float32)
    x0 = np.zeros(p["x"].shape, np.complex64)
    delta, rho, nCGD, nIter = (0.3, 3.0, 4, 3) if kind == "2d" else (0.003, 0.03, 4, 3)
    xg = torch.from_numpy(x0).cuda()
    z, u, hist = P.bmSteva(xg, None, None, torch.from_numpy(y).cuda(), torch.from_numpy(ve).cuda(), delta, rho,
                           nCGD, nIter=nIter)
    Cn = nd(p["C"].astype(np.complex128), p["n_u"], p["nCh"])
    xr, zr, ur, hr = O.steva(np.zeros(p["n_u"], complex), None, None, y.astype(np.complex128),
                             ve.astype(np.float64), Cn, R["Gu"], R["Gut"], p["N_u"], p["n_u"], p["dK_u"], delta, rho,
                             nCGD, None, nIter, kernel=(p["kernel"], None, None))
    assert rel(xg.cpu().numpy(), flat(xr, p["n_u"])) < 1e-4
    zr_flat = np.stack([flat(zr[a], p["n_u"]) for a in range(len(p["n_u"]))])
    assert np.linalg.norm(zr_flat) > 0, "threshold too large: z is identically zero"
    assert rel(z.cpu().numpy(), zr_flat) < 1e-3
    assert np.allclose(hist, np.array(hr), rtol=2e-3)


def _teva_problem(nFr=3):
    import torch
    from monalisa_amd import synth
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    rng = np.random.default_rng(3)
    n_u, N_u, dK_u, nCh = [16, 16], [32, 32], [0.5, 0.5], 2
    C = synth.coil_maps(n_u, nCh, rng)
    P = Plan(N_u, n_u, dK_u, nCh, "kaiser")
    P.set_coils(torch.from_numpy(C).cuda())
    ves, Rs = [], []
    for f in range(nFr):
        t = synth.radial2d_golden(12 + 4 * f, 32, k_max=8.0)
        t = np.roll(t, f, axis=1)
        ve = synth.ve_radial2(t, 32, 12 + 4 * f)
        G = bmTraj2SparseMat(t, ve, N_u, dK_u, "kaiser")
        P.set_frame(f, G["Gu"], G["Gut"])
        ves.append(ve.astype(np.float32))
        Rs.append((t, ve))
    xs = [synth.ellipsoid_phantom(n_u, np.random.default_rng(10 + f)) for f in range(nFr)]
    ys = []
    for f in range(nFr):
        y = P.bmShanna(torch.from_numpy(xs[f]).cuda(), f=f).cpu().numpy()
        ys.append((y + 0.01 * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape))).astype(np.complex64))
    return dict(P=P, C=C, ys=ys, ves=ves, Rs=Rs, n_u=n_u, N_u=N_u, dK_u=dK_u, nCh=nCh, nFr=nFr)


def _teva_run(q, nIter=3, delta=0.05, rho=0.5, nCGD=3):
    import torch
    P = q["P"]
    ve_max = float(max(v.max() for v in q["ves"]))
    x = torch.zeros((q["nFr"], P.nu), dtype=torch.complex64, device="cuda")
    z, u, hist = P.bmTevaMorphosia_chain(x, None, None, [torch.from_numpy(v).cuda() for v in q["ys"]],
                                         [torch.from_numpy(v).cuda() for v in q["ves"]], delta, rho, nCGD,
                                         ve_max, nIter=nIter)
    torch.cuda.synchronize()
    return x.cpu().numpy(), z.cpu().numpy(), u.cpu().numpy(), hist


@pytest.mark.parametrize("sched", ["constant", "ramped"])
def test_teva_chain_vs_oracle(sched):
    """Constant delta/rho, and [start, end] ranges spread linearly over the
    iterations (private_init_regul_param, bmTevaMorphosia_chain.m:226-235)."""
    import torch
    from monalisa_amd import synth
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    print(f"[gpu] teva chain {sched}", flush=True)
    rng = np.random.default_rng(3)
    n_u, N_u, dK_u, nCh, nFr = [16, 16], [32, 32], [0.5, 0.5], 2, 3
    C = synth.coil_maps(n_u, nCh, rng)
    P = Plan(N_u, n_u, dK_u, nCh, "kaiser")
    P.set_coils(torch.from_numpy(C).cuda())
    ts, ves, Gs, Rs = [], [], [], []
    for f in range(nFr):
        t = synth.radial2d_golden(12 + 4 * f, 32, k_max=8.0)
        t = np.roll(t, f, axis=1)
        ve = synth.ve_radial2(t, 32, 12 + 4 * f)
        G = bmTraj2SparseMat(t, ve, N_u, dK_u, "kaiser")
        P.set_frame(f, G["Gu"], G["Gut"])
        ts.append(t); ves.append(ve.astype(np.float32)); Gs.append(G)
        Rs.append(O.traj2sparsemat(t, ve, N_u, dK_u, "kaiser"))
    xs = [synth.ellipsoid_phantom(n_u, np.random.default_rng(10 + f)) for f in range(nFr)]
    ys = []
    for f in range(nFr):
        y = P.bmShanna(torch.from_numpy(xs[f]).cuda(), f=f).cpu().numpy()
        ys.append((y + 0.01 * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape))).astype(np.complex64))
    ve_max = float(max(v.max() for v in ves))
    delta, rho, nCGD, nIter = 0.05, 0.5, 3, 3
    if sched == "ramped":
        delta, rho, nIter = (0.02, 0.1), (0.3, 0.8), 4
    x = torch.zeros((nFr, 256), dtype=torch.complex64, device="cuda")
    z, u, hist = P.bmTevaMorphosia_chain(x, None, None, [torch.from_numpy(v).cuda() for v in ys],
                                         [torch.from_numpy(v).cuda() for v in ves], delta, rho, nCGD, ve_max,
                                         nIter=nIter)
    Cn = nd(C.astype(np.complex128), n_u, nCh)
    xr, zr, ur, hr = O.teva_chain([np.zeros(n_u, complex)] * nFr, None, None, [v.astype(np.complex128) for v in ys],
                                  [v.astype(np.float64) for v in ves], Cn, [r["Gu"] for r in Rs],
                                  [r["Gut"] for r in Rs], N_u, n_u, dK_u, delta, rho, nCGD, ve_max, nIter,
                                  kernel=("kaiser", None, None))
    got = x.cpu().numpy()
    ref = np.stack([flat(v, n_u) for v in xr])
    assert rel(got, ref) < 1e-4
    assert np.allclose(hist, np.array(hr), rtol=2e-3)


def test_teva_batched_equals_per_frame():
    """All frames in one launch per gridding op (block-diagonal matrices,
    frame-batched FFT passes) give the per-frame solver's results bit for bit:
    every row keeps its entries and order, every FFT line its arithmetic."""
    print("[gpu] teva batched vs per-frame", flush=True)
    q = _teva_problem()
    ref = _teva_run(q)
    q["P"].set_options(batch_frames=True)
    got = _teva_run(q)
    for a, b in zip(got[:3], ref[:3]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got[3], ref[3])


@pytest.mark.parametrize("batch", [False, True])
def test_teva_graph_replays_equal_eager(batch):
    """ADMM iterations replayed from a HIP graph equal the eager loop; the
    graph run is repeated with eager operator calls and synchronisations in
    between, and every replay (one history row each) is checked."""
    import torch
    print(f"[gpu] teva graph batch={batch}", flush=True)
    q = _teva_problem()
    P = q["P"]
    P.set_options(batch_frames=batch)
    ref = _teva_run(q, nIter=4)
    P.set_options(graph=True)
    for rep in range(3):
        got = _teva_run(q, nIter=4)
        for a, b in zip(got[:3], ref[:3]):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(got[3], ref[3])
        # eager work on the same plan between graph runs
        xe = torch.ones(P.nu, dtype=torch.complex64, device="cuda")
        y0 = P.bmShanna(xe, f=rep % q["nFr"])
        torch.cuda.synchronize()
        assert torch.isfinite(torch.view_as_real(y0)).all()
    P.set_options(graph=False, batch_frames=False)
    again = _teva_run(q, nIter=4)
    np.testing.assert_array_equal(again[0], ref[0])


def test_autotune_keeps_results():
    import torch
    print("[gpu] autotune", flush=True)
    q = _teva_problem(nFr=1)
    P = q["P"]
    x = torch.from_numpy(synth_x(P.nu)).cuda()
    y_ref = P.bmShanna(x).cpu().numpy()
    tile = P.autotune()
    assert tile in (16, 32, 64, -2048, -4096, -8192)
    y = P.bmShanna(x).cpu().numpy()
    np.testing.assert_allclose(y, y_ref, rtol=0, atol=1e-5 * np.abs(y_ref).max())


def synth_x(nu, seed=5):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=nu) + 1j * rng.normal(size=nu)).astype(np.complex64)


@pytest.mark.parametrize("with_coils", [True, False])
def test_nasha_vs_oracle(with_coils):
    import torch
    from monalisa_amd.sparse import bmTraj2SparseMat
    from test_gpu_ops import _plan
    p = make_problem("2d", seed=4, nCh=3)
    P = _plan(p)
    G = bmTraj2SparseMat(p["t"], p["ve"], p["N_u"], p["dK_u"], p["kernel"], which=("Gn",))
    R = O.traj2sparsemat(p["t"], p["ve"], p["N_u"], p["dK_u"], p["kernel"], which=("Gn",))
    y = _data(P, p)
    Cd = torch.from_numpy(p["C"]).cuda() if with_coils else None
    xg = P.bmNasha(torch.from_numpy(y).cuda(), G["Gn"], Cd).cpu().numpy()
    Cn = nd(p["C"].astype(np.complex128), p["n_u"], p["nCh"]) if with_coils else None
    xr = O.nasha(y.astype(np.complex128), R["Gn"], p["N_u"], p["n_u"], p["dK_u"], Cn, kernel=(p["kernel"], None, None))
    assert rel(xg, flat(xr, p["n_u"])) < 1e-5
