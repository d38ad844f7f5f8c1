"""GPU FFT, deapodisation and operators against the float64 oracle."""
import numpy as np
import pytest

from problems import make_problem, oracle_mats, nd, flat, rel
from oracle import mona_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,nb", [((8,), 3), ((16, 12), 2), ((8, 6, 10), 2), ((64, 64), 1),
                                  ((32, 32, 32), 2), ((192,), 5), ((24, 40), 20), ((30, 20), 1),
                                  ((64, 48, 40), 3), ((16, 256), 4), ((18, 160), 3), ((96, 96, 96), 1)])
@pytest.mark.parametrize("inverse", [False, True])
def test_cfft_vs_numpy(N, nb, inverse):
    import torch
    from monalisa_amd.nufft import cfft
    print(f"[gpu] cfft N={N} nb={nb} inv={inverse}", flush=True)
    rng = np.random.default_rng(sum(N) + nb)
    a = (rng.normal(size=(int(np.prod(N)), nb)) + 1j * rng.normal(size=(int(np.prod(N)), nb))).astype(np.complex64)
    d = torch.from_numpy(a).cuda()
    cfft(d, N, inverse)
    got = d.cpu().numpy()
    ref = nd(a.astype(np.complex128), N, nb)
    ref = (O.cifft if inverse else O.cfft)(ref, axes=range(len(N)))
    assert rel(got, flat(ref, N)) < 2e-6


def _plan(p):
    import torch
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    G = bmTraj2SparseMat(p["t"], p["ve"], p["N_u"], p["dK_u"], p["kernel"])
    P = Plan(p["N_u"], p["n_u"], p["dK_u"], p["nCh"], p["kernel"])
    P.set_coils(torch.from_numpy(p["C"]).cuda())
    P.set_frame(0, G["Gu"], G["Gut"])
    return P


@pytest.mark.parametrize("kind", ["2d", "3d", "2d_rect"])
@pytest.mark.parametrize("kernel", ["kaiser", "gauss"])
def test_deapodisation_vs_oracle(kind, kernel):
    p = make_problem(kind, kernel=kernel)
    print(f"[gpu] deapod {kind} {kernel}", flush=True)
    P = _plan(p)
    K = P.K().cpu().numpy()
    Kref = O.crop(O.bmK(p["N_u"], p["dK_u"], kernel), p["N_u"], p["n_u"])
    assert rel(K, flat(Kref, p["n_u"])) < 1e-6
    KFC, KFCc = [a.cpu().numpy() for a in P.KFC()]
    Cn = nd(p["C"], p["n_u"], p["nCh"])
    assert rel(KFC, flat(O.kf(Cn, p["N_u"], p["n_u"], p["dK_u"], kernel), p["n_u"])) < 1e-6
    assert rel(KFCc, flat(O.kf_conj(np.conj(Cn), p["N_u"], p["n_u"], p["dK_u"], kernel), p["n_u"])) < 1e-6


@pytest.mark.parametrize("kind,nCh", [("2d", 3), ("3d", 3), ("2d_rect", 3), ("3d32", 3), ("3d32", 20),
                                      ("2d", 1), ("3d32s", 4)])
def test_forward_adjoint_vs_oracle(kind, nCh):
    import torch
    p = make_problem(kind, nCh=nCh)
    print(f"[gpu] forward/adjoint {kind} nCh={nCh}", flush=True)
    P = _plan(p)
    R = oracle_mats(p)
    Cn = nd(p["C"], p["n_u"], p["nCh"])
    KFC = O.kf(Cn, p["N_u"], p["n_u"], p["dK_u"], p["kernel"])
    KFCc = O.kf_conj(np.conj(Cn), p["N_u"], p["n_u"], p["dK_u"], p["kernel"])
    x = torch.from_numpy(p["x"]).cuda()
    y = P.bmShanna(x).cpu().numpy()
    yref = O.shanna(nd(p["x"].astype(np.complex128), p["n_u"]), R["Gu"], KFC, p["N_u"], p["n_u"])
    assert rel(y, yref) < 1e-5
    rng = np.random.default_rng(1)
    yy = (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape)).astype(np.complex64)
    xa = P.bmNakatsha(torch.from_numpy(yy).cuda()).cpu().numpy()
    xref = O.nakatsha(yy.astype(np.complex128), R["Gut"], KFCc, True, p["N_u"], p["n_u"])
    assert rel(xa, flat(xref, p["n_u"])) < 1e-5
    xc = P.bmNakatsha(torch.from_numpy(yy).cuda(), C_flag=False).cpu().numpy()
    xcref = O.nakatsha(yy.astype(np.complex128), R["Gut"], KFCc, False, p["N_u"], p["n_u"])
    assert rel(xc, flat(xcref, p["n_u"])) < 1e-5


def test_adjointness_gpu():
    """<A x, y> = <x, A^H y> * (prod dK_u) relation of the operator pair:
    KFC_conj = conj(KFC) * prod(N_u) and the inverse FFT carries 1/prod(N_u),
    so <M x, y> = <x, M^H y> exactly (bmKF.m:27 vs bmKF_conj.m:30)."""
    import torch
    p = make_problem("3d", nCh=2)
    P = _plan(p)
    rng = np.random.default_rng(4)
    x = (rng.normal(size=p["x"].shape) + 1j * rng.normal(size=p["x"].shape)).astype(np.complex64)
    y = P.bmShanna(torch.from_numpy(x).cuda()).cpu().numpy()
    yy = (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape)).astype(np.complex64)
    xa = P.bmNakatsha(torch.from_numpy(yy).cuda()).cpu().numpy()
    lhs = np.vdot(yy.astype(np.complex128), y.astype(np.complex128))
    rhs = np.vdot(xa.astype(np.complex128), x.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_operators_deterministic():
    import torch
    p = make_problem("2d")
    P = _plan(p)
    x = torch.from_numpy(p["x"]).cuda()
    a = P.bmShanna(x).cpu().numpy()
    b = P.bmShanna(x).cpu().numpy()
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    ya = P.bmNakatsha(torch.from_numpy(a).cuda()).cpu().numpy()
    yb = P.bmNakatsha(torch.from_numpy(a).cuda()).cpu().numpy()
    assert np.array_equal(ya.view(np.int32), yb.view(np.int32))


@pytest.mark.parametrize("kind", ["3d", "3d32", "3d32s"])
def test_fft_pass_variants(kind):
    """FFT pass kernels: the LDS Stockham kernel with tiles 2048 / 4096 /
    8192 runs the same butterflies in the same order for every tile (bit-
    identical), so does the register kernel for 16 / 32 / 64 lines per tile;
    the two kernels agree to float rounding."""
    import torch
    print(f"[gpu] fft pass variants {kind}", flush=True)
    p = make_problem(kind, seed=7, nCh=4)
    P = _plan(p)
    x = torch.from_numpy(p["x"]).cuda()

    def run(tile, lines):
        P.set_options(fft_tile=tile, fft_lines=lines)
        y = P.bmShanna(x)
        xa = P.bmNakatsha(y)
        torch.cuda.synchronize()
        return y.cpu().numpy(), xa.cpu().numpy()

    lds = [run(t, -1) for t in (4096, 2048, 8192)]
    reg = [run(4096, l) for l in (16, 32, 64)]
    for y, xa in lds[1:]:
        np.testing.assert_array_equal(y, lds[0][0])
        np.testing.assert_array_equal(xa, lds[0][1])
    for y, xa in reg[1:]:
        np.testing.assert_array_equal(y, reg[0][0])
        np.testing.assert_array_equal(xa, reg[0][1])
    assert rel(reg[0][0], lds[0][0]) < 1e-6
    assert rel(reg[0][1], lds[0][1]) < 1e-6


def test_loaded_matrices_drive_the_operators(tmp_path):
    """Gu / Gut saved in the reference's file format and loaded back give
    the same operator results on the GPU."""
    import torch
    from monalisa_amd import binfile
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    print("[gpu] loaded matrices", flush=True)
    p = make_problem("3d32", seed=2, nCh=3)
    G = bmTraj2SparseMat(p["t"], p["ve"], p["N_u"], p["dK_u"], p["kernel"])
    for k in ("Gu", "Gut"):
        binfile.bmArray2Binary_sparseMat(G[k], str(tmp_path / k))
    L = {k: binfile.bmBinary2Array_sparseMat(str(tmp_path / k)) for k in ("Gu", "Gut")}
    x = torch.from_numpy(p["x"]).cuda()
    outs = []
    for M in (G, L):
        P = Plan(p["N_u"], p["n_u"], p["dK_u"], p["nCh"], p["kernel"])
        P.set_coils(torch.from_numpy(p["C"]).cuda())
        P.set_frame(0, M["Gu"], M["Gut"])
        # loaded matrices carry no trajectory: the built ones are compared
        # on their stored entries too, then matrix-free
        P.set_options(matrix_free=False)
        y = P.bmShanna(x)
        outs.append((y.cpu().numpy(), P.bmNakatsha(y).cpu().numpy()))
        if M is G:
            P.set_options(matrix_free=True)
            y_mf = P.bmShanna(x).cpu().numpy()
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    assert rel(y_mf, outs[1][0]) < 2e-6


def test_set_frame_rejects_a_matrix_of_another_grid():
    """Gu must address the plan's grid: a matrix built for another N_u is
    refused with MONA_ERR_ARG before any device work."""
    import torch
    from monalisa_amd._lib import MonaError
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    p = make_problem("3d32", nCh=2)
    G = bmTraj2SparseMat(p["t"], p["ve"], [24, 24, 24], p["dK_u"], p["kernel"])
    P = Plan(p["N_u"], p["n_u"], p["dK_u"], 2, p["kernel"])
    P.set_coils(torch.from_numpy(p["C"]).cuda())
    with pytest.raises(MonaError) as e:
        P.set_frame(0, G["Gu"], G["Gut"])
    assert e.value.status == 1
