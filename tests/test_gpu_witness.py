"""Witness snapshots of the solvers (bmWitnessInfo.m:92-190, the watch calls
of bmSensa.m:85, bmSteva and the Morphosia solvers): a run with a witness
equals the run without it bit for bit; the image kept after outer
iteration c equals frame 0's witness image (3-D: the middle z slice) of a
separate run stopped after c iterations (the outer iterations restart from
x, and ADMM with constant delta / rho repeats the same first c
iterations); the full image is saved at the witness indices; the histories
land in witnessInfo.param."""
import numpy as np
import pytest

from problems import make_problem

pytestmark = pytest.mark.gpu


def test_sensa_witness_3d(tmp_path):
    import torch
    from test_gpu_ops import _plan
    from monalisa_amd.optim import bmWitnessInfo
    print("[gpu] witness: CG-SENSE 3-D", flush=True)
    p = make_problem("3d32", seed=5, nCh=4)
    P = _plan(p)
    y = P.bmShanna(torch.from_numpy(p["x"]).cuda())
    ve = torch.from_numpy(p["ve"].astype(np.float32)).cuda()
    n_u, nIter = p["n_u"], 3

    def run(k, w=None):
        x = torch.zeros(P.nu, dtype=torch.complex64, device="cuda")
        r = P.bmSensa(x, y, ve, 2, nIter=k, witnessInfo=w)
        return x.cpu().numpy(), r

    x_plain, r_plain = run(nIter)
    w = bmWitnessInfo("sensa3", [2], save_dir=str(tmp_path))
    x_w, r_w = run(nIter, w)
    np.testing.assert_array_equal(x_w, x_plain)
    assert r_w == r_plain
    np.testing.assert_array_equal(w.param["residu"], r_plain)
    assert w.param["recon_name"] == "sensa" and w.finalInd == nIter
    assert w.witness_im.shape == (16, 16, nIter + 1) and len(w.witness_time) == nIter + 1
    for c in range(1, nIter + 1):
        xc, _ = run(c)
        np.testing.assert_array_equal(w.witness_im[:, :, c - 1], xc.reshape(n_u, order="F")[:, :, n_u[2] // 2])
    np.testing.assert_array_equal(w.witness_im[:, :, nIter], w.witness_im[:, :, nIter - 1])   # 'final'
    x2, _ = run(2)
    np.testing.assert_array_equal(np.load(tmp_path / "x_sensa3.npy"), x2.reshape(n_u, order="F"))
    assert (tmp_path / "wit_sensa3.npz").exists()


def test_teva_chain_witness_2d_with_graph(tmp_path):
    """The temporal-TV chain (frames [nFr, nu]) with the HIP-graph option on:
    a watched run stays eager and equals the graph run bit for bit; the
    witness image is frame 0's image."""
    import torch
    from test_gpu_admm import _teva_problem
    from monalisa_amd.optim import bmWitnessInfo
    print("[gpu] witness: temporal-TV chain 2-D", flush=True)
    q = _teva_problem(3)
    P = q["P"]
    P.set_options(graph=True)
    ve_max = float(max(v.max() for v in q["ves"]))
    ys = [torch.from_numpy(v).cuda() for v in q["ys"]]
    vs = [torch.from_numpy(v).cuda() for v in q["ves"]]

    def run(k, w=None):
        x = torch.zeros((q["nFr"], P.nu), dtype=torch.complex64, device="cuda")
        _, _, h = P.bmTevaMorphosia_chain(x, None, None, ys, vs, 0.05, 0.5, 3, ve_max, nIter=k, witnessInfo=w)
        torch.cuda.synchronize()
        return x.cpu().numpy(), h

    x_plain, h_plain = run(4)
    w = bmWitnessInfo("teva2", [1, 4], save_dir=str(tmp_path))
    x_w, h_w = run(4, w)
    P.set_options(graph=False)
    np.testing.assert_array_equal(x_w, x_plain)
    np.testing.assert_array_equal(h_w, h_plain)
    np.testing.assert_array_equal(w.param["dafi"], h_plain[:, 1])
    for c in (1, 2, 3, 4):
        xc, _ = run(c)
        np.testing.assert_array_equal(w.witness_im[:, :, c - 1], xc[0].reshape(q["n_u"], order="F"))
    saved = np.load(tmp_path / "x_teva2.npy")          # last witness index reached: 4
    assert saved.shape == (3, 16, 16)
    np.testing.assert_array_equal(saved[2], x_plain[2].reshape(q["n_u"], order="F"))


def test_steva_witness_no_indices():
    """Without witness indices nothing is captured and the histories are
    still recorded."""
    import torch
    from test_gpu_ops import _plan
    from monalisa_amd.optim import bmWitnessInfo
    print("[gpu] witness: spatial TV, no indices", flush=True)
    p = make_problem("2d", seed=7, nCh=2)
    P = _plan(p)
    y = P.bmShanna(torch.from_numpy(p["x"]).cuda())
    ve = torch.from_numpy(p["ve"].astype(np.float32)).cuda()
    w = bmWitnessInfo("steva", [], save_dir=None)
    x = torch.zeros(P.nu, dtype=torch.complex64, device="cuda")
    _, _, h = P.bmSteva(x, None, None, y, ve, 0.3, 3.0, 3, nIter=2, witnessInfo=w)
    assert w.witness_im is None
    np.testing.assert_array_equal(w.param["total_variation"], h[:, 1])
