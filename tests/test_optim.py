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
    w = bmWitnessInfo("teva", [3])
    h = np.arange(9.0).reshape(3, 3)
    w.record(h, ["objective", "dafi", "regul"], x=np.ones(4), final_ind=3)
    np.testing.assert_array_equal(w.param["dafi"], [1.0, 4.0, 7.0])
    assert 3 in w.witness_im and w.finalInd == 3
