"""The oracle's direct gridders and bmMathilda (src/gridding123/m/
bmGridder_n2u.m, bmGridder_n2u_leight.m, src/fourier123/map_function/
nonCartesian/bmMathilda.m) pinned by: the normalised gridding of constant
data, the gridder-based reconstruction against the matrix-based bmNasha
(the same normalised n2u built two ways), and the leight restatement
against the C port."""
import numpy as np
import pytest

from monalisa_amd import synth
from oracle import mona_oracle as O
import oracle_libs as OL


def _problem(imDim, seed=0, nCh=2):
    rng = np.random.default_rng(seed)
    if imDim == 2:
        t = synth.radial2d_golden(10, 32, k_max=8.0)
        N_u, n_u, dK_u = [32, 32], [16, 16], [0.5, 0.5]
        ve = synth.ve_radial2(t, 32, 10)
    else:
        t = synth.koosh_ball(2, 10, 16, k_max=4.0)
        N_u, n_u, dK_u = [16, 16, 16], [8, 8, 8], [0.5, 0.5, 0.5]
        ve = synth.ve_radial3(t, 16, 20)
    y = rng.normal(size=(t.shape[1], nCh)) + 1j * rng.normal(size=(t.shape[1], nCh))
    return t, ve, y, N_u, n_u, dK_u


@pytest.mark.parametrize("imDim", [2, 3])
@pytest.mark.parametrize("kern", ["kaiser", "gauss"])
def test_gridder_reconstruction_equals_matrix_reconstruction(imDim, kern):
    """Normalised n2u with bmGridder_n2u's weights equals Gn (built by
    bmTraj2SparseMat) applied to the same data: bmMathilda without leight is
    bmNasha."""
    t, ve, y, N_u, n_u, dK_u = _problem(imDim)
    G = O.traj2sparsemat(t, ve, N_u, dK_u, kern, which=("Gn",))
    a = O.mathilda(y, t, ve, N_u, n_u, dK_u, kernel=(kern, None, None), leight_flag=False)
    b = O.nasha(y, G["Gn"], N_u, n_u, dK_u, kernel=(kern, None, None))
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-6 * np.abs(b).max())


@pytest.mark.parametrize("leight", [False, True])
def test_gridders_normalise_constant_data(leight):
    t, ve, y, N_u, n_u, dK_u = _problem(2)
    one = np.ones((t.shape[1], 1), complex)
    g = (O.gridder_n2u_leight if leight else O.gridder_n2u)(one, t, ve, N_u, dK_u, "kaiser")
    touched = np.abs(g[:, 0]) > 0
    np.testing.assert_allclose(g[touched, 0], 1.0, rtol=1e-12)
    assert touched.sum() > 50


def test_leight_restatement_equals_port():
    t, ve, y, N_u, n_u, dK_u = _problem(3, seed=2)
    tg, Dn, keep = O.gridder_prep(t, ve, N_u, dK_u, True)
    yk = y[keep]
    nPt, nCh = yk.shape
    ar, ai = [np.ascontiguousarray(v, dtype=np.float32) for v in (yk.real, yk.imag)]
    Nu = int(np.prod(N_u))
    br, bi = np.zeros((Nu, nCh), np.float32), np.zeros((Nu, nCh), np.float32)
    tx, ty, tz = [np.ascontiguousarray(tg[d], dtype=np.float32) for d in range(3)]
    kp = O.kernel_defaults("kaiser")[2]
    OL.port().port_gridder_n2u_leight(3, OL.fp(ar), OL.fp(ai), OL.fp(br), OL.fp(bi), OL.fp(tx), OL.fp(ty),
                                      OL.fp(tz), OL.fp(np.ascontiguousarray(Dn, dtype=np.float32)), nCh, nPt,
                                      N_u[0], N_u[1], N_u[2], 3, float(kp[0]), float(kp[1]))
    ref = O.gridder_n2u_leight(y, t, ve, N_u, dK_u, "kaiser")
    got = br + 1j * bi
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-6 * np.abs(ref).max())
