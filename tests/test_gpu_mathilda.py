"""Direct gridders and bmMathilda on the GPU against the float64 oracle,
and the gridder-based reconstruction against the matrix-based bmNasha."""
import numpy as np
import pytest

from problems import flat, nd, rel
from oracle import mona_oracle as O

pytestmark = pytest.mark.gpu


def _problem(imDim, seed=0, nCh=3):
    from monalisa_amd import synth
    rng = np.random.default_rng(seed)
    if imDim == 2:
        t = synth.radial2d_golden(24, 64, k_max=16.0)
        N_u, n_u, dK_u = [64, 64], [32, 32], [0.5, 0.5]
        ve = synth.ve_radial2(t, 64, 24)
    else:
        t = synth.koosh_ball(4, 20, 32, k_max=8.0)
        N_u, n_u, dK_u = [32, 32, 32], [16, 16, 16], [0.5, 0.5, 0.5]
        ve = synth.ve_radial3(t, 32, 80)
    y = (rng.normal(size=(t.shape[1], nCh)) + 1j * rng.normal(size=(t.shape[1], nCh))).astype(np.complex64)
    C = synth.coil_maps(n_u, nCh, rng)
    return t, ve, y, N_u, n_u, dK_u, C


@pytest.mark.parametrize("imDim", [2, 3])
@pytest.mark.parametrize("leight", [False, True])
def test_gridders_vs_oracle(imDim, leight):
    from monalisa_amd.gridders import bmGridder_n2u, bmGridder_n2u_leight
    print(f"[gpu] gridder imDim={imDim} leight={leight}", flush=True)
    t, ve, y, N_u, n_u, dK_u, C = _problem(imDim)
    got = (bmGridder_n2u_leight if leight else bmGridder_n2u)(y, t, ve, N_u, dK_u, "kaiser")
    ref = (O.gridder_n2u_leight if leight else O.gridder_n2u)(y.astype(np.complex128), t, ve, N_u, dK_u, "kaiser")
    assert rel(got, ref) < 1e-5


@pytest.mark.parametrize("imDim", [2, 3])
@pytest.mark.parametrize("leight,coils", [(True, True), (False, True), (True, False)])
def test_mathilda_vs_oracle(imDim, leight, coils):
    from monalisa_amd.gridders import bmMathilda
    print(f"[gpu] mathilda imDim={imDim} leight={leight} coils={coils}", flush=True)
    t, ve, y, N_u, n_u, dK_u, C = _problem(imDim, seed=1)
    x = bmMathilda(y, t, ve, C if coils else None, N_u, n_u, dK_u, "kaiser", leight_flag=leight).cpu().numpy()
    Cn = nd(C.astype(np.complex128), n_u, C.shape[1]) if coils else None
    ref = O.mathilda(y.astype(np.complex128), t, ve, N_u, n_u, dK_u, Cn, kernel=("kaiser", None, None),
                     leight_flag=leight)
    assert rel(x, flat(ref, n_u)) < 1e-5


def test_mathilda_without_leight_equals_nasha():
    """Normalised gridding with bmGridder_n2u's weights is Gn * y: the two
    reconstructions agree to float rounding."""
    import torch
    from monalisa_amd.gridders import bmMathilda
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    print("[gpu] mathilda vs nasha", flush=True)
    t, ve, y, N_u, n_u, dK_u, C = _problem(3, seed=2)
    P = Plan(N_u, n_u, dK_u, C.shape[1], "kaiser")
    Cd = torch.from_numpy(C).cuda()
    P.set_coils(Cd)
    G = bmTraj2SparseMat(t, ve, N_u, dK_u, "kaiser", which=("Gn",))
    xn = P.bmNasha(torch.from_numpy(y).cuda(), G["Gn"], Cd).cpu().numpy()
    xm = bmMathilda(y, t, ve, Cd, N_u, n_u, dK_u, "kaiser", leight_flag=False, plan=P).cpu().numpy()
    assert rel(xm, xn) < 1e-5
