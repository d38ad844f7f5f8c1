"""Deformation matrices on the GPU: the products of bmImDeform / bmImDeformT
and the motion-compensated temporal-TV chain (bmTevaMorphosia_chain.m with
Tu, Tut) against the float64 oracle."""
import numpy as np
import pytest

from problems import flat, nd, rel
from oracle import mona_oracle as O

pytestmark = pytest.mark.gpu


def _field(n_u, seed, amp=1.3):
    """Smooth displacement field [imDim, prod(n_u)] (x fastest)."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(n) / n for n in n_u], indexing="ij")
    v = []
    for d in range(len(n_u)):
        ph = rng.uniform(0, 2 * np.pi, size=2)
        f = amp * np.sin(2 * np.pi * grids[0] + ph[0]) * np.cos(2 * np.pi * grids[-1] + ph[1])
        v.append(f.reshape(-1, order="F"))
    return np.stack(v)


@pytest.mark.parametrize("torus", [False, True])
def test_drop_in_product_applies_deformation(torus):
    """The host drop-in product (bmSparseMat_vec) applies Tu as bmImDeform.m does."""
    from monalisa_amd.sparse import bmImDeformField2SparseMat, bmSparseMat_vec
    print(f"[gpu] deform drop-in product torus={torus}", flush=True)
    N_u = [12, 10]
    v = _field(N_u, 3, amp=2.5)
    G = bmImDeformField2SparseMat(v, N_u, torus_flag=torus)
    R = O.deform2sparsemat(v, N_u, torus_flag=torus)
    rng = np.random.default_rng(2)
    x = (rng.normal(size=(120, 1)) + 1j * rng.normal(size=(120, 1))).astype(np.complex64)
    for k in ("Gu", "Gut"):
        np.testing.assert_allclose(bmSparseMat_vec(G[k], x), O.spmv(R[k], x.astype(np.complex128)),
                                   rtol=2e-5, atol=2e-6)


def test_bmImDeform_on_device():
    import torch
    from monalisa_amd.sparse import bmImDeform, bmImDeformT, bmImDeformField2SparseMat
    print("[gpu] bmImDeform", flush=True)
    N_u = [8, 6, 5]
    v = _field(N_u, 4)
    G = bmImDeformField2SparseMat(v, N_u, torus_flag=True)
    R = O.deform2sparsemat(v, N_u, torus_flag=True)
    rng = np.random.default_rng(5)
    x = (rng.normal(size=N_u) + 1j * rng.normal(size=N_u)).astype(np.complex64)
    xf = torch.from_numpy(flat(x, N_u).copy()).cuda()
    got = bmImDeform(G["Gu"], xf).cpu().numpy()
    np.testing.assert_allclose(got, O.spmv(R["Gu"], flat(x, N_u).astype(np.complex128)), rtol=2e-5, atol=2e-6)
    gotT = bmImDeformT(G["Gut"], xf).cpu().numpy()
    np.testing.assert_allclose(gotT, O.spmv(R["Gut"], flat(x, N_u).astype(np.complex128)), rtol=2e-5, atol=2e-6)
    assert bmImDeform(None, xf) is xf


def _problem():
    from test_gpu_admm import _teva_problem
    q = _teva_problem()
    q["fields"] = [_field(q["n_u"], 20 + f) for f in range(q["nFr"])]
    return q


def _deforms(q, torus=True):
    from monalisa_amd.sparse import bmImDeformField2SparseMat
    return [bmImDeformField2SparseMat(v, q["n_u"], torus_flag=torus) for v in q["fields"]]


def _run(q, Ts, nIter=3, delta=0.05, rho=0.5, nCGD=3):
    import torch
    P = q["P"]
    ve_max = float(max(v.max() for v in q["ves"]))
    x = torch.zeros((q["nFr"], P.nu), dtype=torch.complex64, device="cuda")
    Tu = None if Ts is None else [T["Gu"] if T is not None else None for T in Ts]
    Tut = None if Ts is None else [T["Gut"] if T is not None else None for T in Ts]
    z, u, hist = P.bmTevaMorphosia_chain(x, None, None, [torch.from_numpy(v).cuda() for v in q["ys"]],
                                         [torch.from_numpy(v).cuda() for v in q["ves"]], delta, rho, nCGD,
                                         ve_max, nIter=nIter, Tu=Tu, Tut=Tut)
    torch.cuda.synchronize()
    return x.cpu().numpy(), z.cpu().numpy(), u.cpu().numpy(), hist


@pytest.mark.parametrize("torus", [True, False])
def test_teva_chain_with_deformation_vs_oracle(torus):
    print(f"[gpu] teva chain with deformations torus={torus}", flush=True)
    q = _problem()
    Ts = _deforms(q, torus)
    Ts[1] = None                                   # one frame keeps the identity (Tu{f} = [])
    delta, rho, nCGD, nIter = 0.05, 0.5, 3, 3
    x, z, u, hist = _run(q, Ts, nIter, delta, rho, nCGD)
    n_u, N_u, dK_u = q["n_u"], q["N_u"], q["dK_u"]
    Rs = [O.traj2sparsemat(t, ve, N_u, dK_u, "kaiser") for t, ve in q["Rs"]]
    RT = [O.deform2sparsemat(v, n_u, torus_flag=torus) if T is not None else None
          for v, T in zip(q["fields"], Ts)]
    ve_max = float(max(v.max() for v in q["ves"]))
    Cn = nd(q["C"].astype(np.complex128), n_u, q["nCh"])
    xr, zr, ur, hr = O.teva_chain([np.zeros(n_u, complex)] * q["nFr"], None, None,
                                  [v.astype(np.complex128) for v in q["ys"]],
                                  [v.astype(np.float64) for v in q["ves"]], Cn, [r["Gu"] for r in Rs],
                                  [r["Gut"] for r in Rs], N_u, n_u, dK_u, delta, rho, nCGD, ve_max, nIter,
                                  kernel=("kaiser", None, None),
                                  Tu=[r["Gu"] if r else None for r in RT], Tut=[r["Gut"] if r else None for r in RT])
    assert rel(x, np.stack([flat(v, n_u) for v in xr])) < 1e-4
    zr_flat = np.stack([flat(v, n_u) for v in zr])
    assert np.linalg.norm(zr_flat) > 0
    assert rel(z, zr_flat) < 1e-3
    assert np.allclose(hist, np.array(hr), rtol=2e-3)
    # the deformations change the answer
    q["P"].set_deform(0)
    q["P"].set_deform(2)
    x0 = _run(q, None, nIter, delta, rho, nCGD)[0]
    assert rel(x0, x) > 1e-3


def test_identity_deformation_equals_plain_chain():
    """Zero displacement fields give identity matrices: the chain's results
    equal the undeformed chain's bit for bit."""
    print("[gpu] teva chain identity deformations", flush=True)
    from monalisa_amd.sparse import bmImDeformField2SparseMat
    q = _problem()
    ref = _run(q, None)
    Ts = [bmImDeformField2SparseMat(np.zeros_like(v), q["n_u"]) for v in q["fields"]]
    got = _run(q, Ts)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


def test_deformed_chain_batched_and_graph_equal_eager():
    """Frame-batched operators and the graph-replayed iteration body give the
    eager per-frame results bit for bit with deformations too."""
    print("[gpu] teva chain with deformations batched + graph", flush=True)
    q = _problem()
    Ts = _deforms(q)
    ref = _run(q, Ts, nIter=4)
    q["P"].set_options(batch_frames=True, graph=True)
    got = _run(q, Ts, nIter=4)
    again = _run(q, Ts, nIter=4)
    for a, b, c in zip(got, ref, again):
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(c, b)
