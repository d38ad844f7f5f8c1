"""The Morphosia solvers on the GPU (bmTevaDuoMorphosia_chain,
bmTevaMorphosia_sheet, bmSensitivaMorphosia_chain / _sheet,
bmSensitivaDuoMorphosia_chain, bmSensaMorphosia_chain) with deformation
matrices, against the float64 oracle."""
import numpy as np
import pytest

from problems import flat, nd, rel
from oracle import mona_oracle as O

pytestmark = pytest.mark.gpu


def _field(n_u, seed, amp=1.1):
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(n) / n for n in n_u], indexing="ij")
    v = []
    for _ in range(len(n_u)):
        ph = rng.uniform(0, 2 * np.pi, size=2)
        v.append((amp * np.sin(2 * np.pi * grids[0] + ph[0]) * np.cos(2 * np.pi * grids[-1] + ph[1]))
                 .reshape(-1, order="F"))
    return np.stack(v)


def _setup(nFr, seed):
    from test_gpu_admm import _teva_problem
    from monalisa_amd.sparse import bmImDeformField2SparseMat
    q = _teva_problem(nFr)
    T = []
    for s in (0, 1):
        fields = [_field(q["n_u"], seed + 10 * s + f) for f in range(nFr)]
        T.append(([bmImDeformField2SparseMat(v, q["n_u"], torus_flag=True) for v in fields],
                  [O.deform2sparsemat(v, q["n_u"], torus_flag=True) for v in fields]))
    T[1][0][1] = T[1][1][1] = None                 # one frame of the second set keeps the identity
    return q, T


def _oracle_common(q):
    n_u, N_u, dK_u = q["n_u"], q["N_u"], q["dK_u"]
    Rs = [O.traj2sparsemat(t, ve, N_u, dK_u, "kaiser") for t, ve in q["Rs"]]
    return ([v.astype(np.complex128) for v in q["ys"]], [v.astype(np.float64) for v in q["ves"]],
            nd(q["C"].astype(np.complex128), n_u, q["nCh"]), [r["Gu"] for r in Rs], [r["Gut"] for r in Rs],
            N_u, n_u, dK_u)


def _gpu_lists(q):
    import torch
    return [torch.from_numpy(v).cuda() for v in q["ys"]], [torch.from_numpy(v).cuda() for v in q["ves"]]


def _mats(Tset, k):
    """(device list, oracle list) of Gu or Gut for one set."""
    dev, ref = Tset
    return ([None if G is None else G[k] for G in dev], [None if G is None else G[k] for G in ref])


@pytest.mark.parametrize("solver", ["duo", "sheet"])
def test_teva_two_term_solvers_vs_oracle(solver):
    import torch
    print(f"[gpu] morphosia {solver}", flush=True)
    nFr, n1 = (3, None) if solver == "duo" else (6, 2)
    q, T = _setup(nFr, 30)
    P = q["P"]
    ys, ves = _gpu_lists(q)
    ve_max = float(max(v.max() for v in q["ves"]))
    delta, rho, nCGD, nIter = 0.05, 0.5, 3, 3
    (Tu1, rTu1), (Tu1t, rTu1t) = _mats(T[0], "Gu"), _mats(T[0], "Gut")
    (Tu2, rTu2), (Tu2t, rTu2t) = _mats(T[1], "Gu"), _mats(T[1], "Gut")
    x = torch.zeros((nFr, P.nu), dtype=torch.complex64, device="cuda")
    if solver == "duo":
        z1, z2, u1, u2, hist = P.bmTevaDuoMorphosia_chain(x, None, None, None, None, ys, ves, delta, rho, nCGD,
                                                          ve_max, nIter, Tu1, Tu1t, Tu2, Tu2t)
        xr, zr, ur, hr = O.teva_duo_chain([np.zeros(q["n_u"], complex)] * nFr, None, None, None, None,
                                          *_oracle_common(q), delta, rho, nCGD, ve_max, nIter,
                                          kernel=("kaiser", None, None), Tu1=rTu1, Tu1t=rTu1t, Tu2=rTu2, Tu2t=rTu2t)
    else:
        z1, z2, u1, u2, hist = P.bmTevaMorphosia_sheet(x, None, None, None, None, ys, ves, n1, delta, rho, nCGD,
                                                       ve_max, nIter, Tu1, Tu1t, Tu2, Tu2t)
        xr, zr, ur, hr = O.teva_sheet([np.zeros(q["n_u"], complex)] * nFr, None, None, None, None,
                                      *_oracle_common(q), delta, rho, nCGD, ve_max, n1, nIter,
                                      kernel=("kaiser", None, None), Tu1=rTu1, Tu1t=rTu1t, Tu2=rTu2, Tu2t=rTu2t)
    torch.cuda.synchronize()
    n_u = q["n_u"]
    assert rel(x.cpu().numpy(), np.stack([flat(v, n_u) for v in xr])) < 1e-4
    for zg, zo in ((z1, zr[0]), (z2, zr[1])):
        ref = np.stack([flat(v, n_u) for v in zo])
        assert np.linalg.norm(ref) > 0
        assert rel(zg.cpu().numpy(), ref) < 1e-3
    assert np.allclose(hist, np.array(hr), rtol=2e-3)


@pytest.mark.parametrize("solver", ["chain", "sheet", "duo"])
def test_sensitiva_solvers_vs_oracle(solver):
    import torch
    print(f"[gpu] sensitiva {solver}", flush=True)
    nFr, n1 = (6, 3) if solver == "sheet" else (3, None)
    q, T = _setup(nFr, 50)
    P = q["P"]
    ys, ves = _gpu_lists(q)
    ve_max = float(max(v.max() for v in q["ves"]))
    delta, nCGD, nIter = 0.3, 4, 3
    (Tu1, rTu1), (Tu1t, rTu1t) = _mats(T[0], "Gu"), _mats(T[0], "Gut")
    (Tu2, rTu2), (Tu2t, rTu2t) = _mats(T[1], "Gu"), _mats(T[1], "Gut")
    x = torch.zeros((nFr, P.nu), dtype=torch.complex64, device="cuda")
    if solver == "chain":
        hist = P.bmSensitivaMorphosia_chain(x, ys, ves, delta, nCGD, ve_max, nIter, Tu1, Tu1t)
        xr, hr = O.sensitiva_chain([np.zeros(q["n_u"], complex)] * nFr, *_oracle_common(q), delta, nCGD, ve_max,
                                   nIter, kernel=("kaiser", None, None), Tu=rTu1, Tut=rTu1t)
    elif solver == "duo":
        hist = P.bmSensitivaDuoMorphosia_chain(x, ys, ves, delta, nCGD, ve_max, nIter, Tu1, Tu1t, Tu2, Tu2t)
        xr, hr = O.sensitiva_duo_chain([np.zeros(q["n_u"], complex)] * nFr, *_oracle_common(q), delta, nCGD,
                                       ve_max, nIter, kernel=("kaiser", None, None), Tu1=rTu1, Tu1t=rTu1t, Tu2=rTu2,
                                       Tu2t=rTu2t)
    else:
        hist = P.bmSensitivaMorphosia_sheet(x, ys, ves, n1, delta, nCGD, ve_max, nIter, Tu1, Tu1t, Tu2, Tu2t)
        xr, hr = O.sensitiva_sheet([np.zeros(q["n_u"], complex)] * nFr, *_oracle_common(q), delta, nCGD, ve_max,
                                   n1, nIter, kernel=("kaiser", None, None), Tu1=rTu1, Tu1t=rTu1t, Tu2=rTu2,
                                   Tu2t=rTu2t)
    torch.cuda.synchronize()
    assert rel(x.cpu().numpy(), np.stack([flat(v, q["n_u"]) for v in xr])) < 1e-4
    assert np.allclose(hist, np.array(hr), rtol=2e-3)


def test_two_term_solver_batched_and_graph_equal_eager():
    """Frame-batched operators and graph replays give the eager results bit
    for bit for a two-term solver with two deformation sets."""
    import torch
    print("[gpu] morphosia sheet batched + graph", flush=True)
    q, T = _setup(6, 70)
    P = q["P"]
    ys, ves = _gpu_lists(q)
    ve_max = float(max(v.max() for v in q["ves"]))
    Tu1, Tu1t = _mats(T[0], "Gu")[0], _mats(T[0], "Gut")[0]
    Tu2, Tu2t = _mats(T[1], "Gu")[0], _mats(T[1], "Gut")[0]

    def run():
        x = torch.zeros((6, P.nu), dtype=torch.complex64, device="cuda")
        z1, z2, u1, u2, h = P.bmTevaMorphosia_sheet(x, None, None, None, None, ys, ves, 2, 0.05, 0.5, 3, ve_max, 4,
                                                    Tu1, Tu1t, Tu2, Tu2t)
        torch.cuda.synchronize()
        return [v.cpu().numpy() for v in (x, z1, z2, u1, u2)] + [h]

    ref = run()
    P.set_options(batch_frames=True, graph=True)
    got, again = run(), run()
    for a, b, c in zip(got, ref, again):
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(c, b)


def test_chain_through_general_entry_equals_teva_chain():
    """mona_morphosia(MONA_MORPH_TEVA_CHAIN) and mona_teva_chain are one solver."""
    import torch
    print("[gpu] morphosia chain entry", flush=True)
    q, T = _setup(3, 90)
    P = q["P"]
    ys, ves = _gpu_lists(q)
    ve_max = float(max(v.max() for v in q["ves"]))
    Tu, Tut = _mats(T[0], "Gu")[0], _mats(T[0], "Gut")[0]
    x1 = torch.zeros((3, P.nu), dtype=torch.complex64, device="cuda")
    z, u, h1 = P.bmTevaMorphosia_chain(x1, None, None, ys, ves, 0.05, 0.5, 3, ve_max, nIter=2, Tu=Tu, Tut=Tut)
    x2 = torch.zeros_like(x1)
    zs, us, h2 = P._morph(P.MORPH_TEVA_CHAIN, None, x2, None, None, ys, ves, 0.05, 0.5, 3, ve_max, 2, None, 1)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(x1.cpu().numpy(), x2.cpu().numpy())
    np.testing.assert_array_equal(z.cpu().numpy(), zs[0].cpu().numpy())
    np.testing.assert_array_equal(h1, h2)


@pytest.mark.parametrize("batch", [False, True])
def test_sensa_morphosia_vs_oracle(batch):
    """One image through every frame's deformation (bmSensaMorphosia_chain)."""
    import torch
    print(f"[gpu] sensa morphosia batch={batch}", flush=True)
    q, T = _setup(3, 110)
    P = q["P"]
    if batch:
        P.set_options(batch_frames=True)
    ys, ves = _gpu_lists(q)
    ve_max = float(max(v.max() for v in q["ves"]))
    Tu, rTu = _mats(T[0], "Gu")
    Tut, rTut = _mats(T[0], "Gut")
    Tu[2] = Tut[2] = rTu[2] = rTut[2] = None
    nCGD, nIter = 4, 3
    x = torch.zeros(P.nu, dtype=torch.complex64, device="cuda")
    hist = P.bmSensaMorphosia_chain(x, ys, ves, nCGD, ve_max, nIter, Tu, Tut)
    torch.cuda.synchronize()
    common = _oracle_common(q)
    xr, hr = O.sensa_morphosia(np.zeros(q["n_u"], complex), *common, nCGD, ve_max, nIter,
                               kernel=("kaiser", None, None), Tu=rTu, Tut=rTut)
    assert rel(x.cpu().numpy(), flat(xr, q["n_u"])) < 1e-4
    assert np.allclose(hist, np.array(hr), rtol=2e-3)
