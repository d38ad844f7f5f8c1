"""Loop control and monitoring objects (src/optim/bmConvergeCondition.m,
bmWitnessInfo.m): the iteration count of `while convCond.check()`, the
thresholds, and the witness record."""
import numpy as np

from monalisa_amd.optim import bmConvergeCondition, bmWitnessInfo, n_iter


def test_while_check_runs_nIter_max_iterations():
    c = bmConvergeCondition(7)
    k = 0
    while c.check():
        k += 1
    assert k == 7 and c.nIter_curr == 7 and not c.nIter_statu
    assert n_iter(bmConvergeCondition(5)) == 5 and n_iter(3) == 3


def test_thresholds_stop_the_loop():
    c = bmConvergeCondition(100, targDist_th=1e-3)
    dists = [1.0, 0.1, 0.01, 1e-4, 1e-5]
    k = 0
    while c.check(dists[k] if k < len(dists) else 0.0):
        k += 1
    assert k == 3                                   # stops once the distance falls below 1e-3
    assert c.targDist_history == [1.0, 0.1, 0.01, 1e-4]
    np.testing.assert_allclose(c.targProgression_history, [0.9, 0.09, 0.0099])
    c2 = bmConvergeCondition(c)
    assert c2.nIter_curr == c.nIter_curr and c2.targDist_history == c.targDist_history


def test_witness_records_histories():
    w = bmWitnessInfo("teva", [3], save_dir=None)
    h = np.arange(9.0).reshape(3, 3)
    w.record(h, ["objective", "dafi", "regul"])
    np.testing.assert_array_equal(w.param["dafi"], [1.0, 4.0, 7.0])


def test_witness_watch_follows_the_reference(tmp_path):
    """bmWitnessInfo.watch (bmWitnessInfo.m:92-190): witness images of frame
    0 appended along the third axis (2-D: the image; 3-D: the z slice
    fix(n_z/2)+1; 1-D: columns), time and index of every call, the full
    image saved at the witness indices only, nothing at all without them."""
    rng = np.random.default_rng(0)
    n3 = [4, 3, 5]
    w = bmWitnessInfo("t3", [2], save_dir=str(tmp_path))
    xs = [rng.normal(size=(2, 60)) for _ in range(3)]          # two frames of 4 x 3 x 5
    for c in (1, 2, 3):
        w.watch(c, xs[c - 1], n3, "loop")
    w.watch(3, xs[2], n3, "final")
    assert w.witness_im.shape == (4, 3, 4) and w.finalInd == 3 and len(w.witness_time) == 4
    for k, c in enumerate((1, 2, 3, 3)):
        np.testing.assert_array_equal(w.witness_im[:, :, k], xs[c - 1][0].reshape(n3, order="F")[:, :, 2])
    saved = np.load(tmp_path / "x_t3.npy")                    # iteration 2 only
    assert saved.shape == (2, 4, 3, 5)
    np.testing.assert_array_equal(saved[1], xs[1][1].reshape(n3, order="F"))
    rec = np.load(tmp_path / "wit_t3.npz")
    assert int(rec["finalInd"]) == 3 and rec["witness_im"].shape == (4, 3, 4)
    w2 = bmWitnessInfo("t2", [1], save_dir=None)
    im = rng.normal(size=12)
    w2.watch(1, [im, im + 1], [4, 3], "loop")
    w2.watch(1, None, [4, 3], "final", image=im + 2)
    np.testing.assert_array_equal(w2.witness_im[:, :, 0], im.reshape(4, 3, order="F"))
    np.testing.assert_array_equal(w2.witness_im[:, :, 1], (im + 2).reshape(4, 3, order="F"))
    w1 = bmWitnessInfo("t1", [1], save_dir=None)
    w1.watch(1, np.arange(5.0), [5], "loop")
    w1.watch(2, np.arange(5.0) + 1, [5], "loop")
    assert w1.witness_im.shape == (5, 2)
    w0 = bmWitnessInfo("none", [], save_dir=str(tmp_path))
    w0.watch(1, np.ones(5), [5], "loop")
    assert w0.witness_im is None and not (tmp_path / "wit_none.npz").exists()
    assert bmWitnessInfo(w).witness_ind == [2]
