"""Loop control and monitoring objects of the reference's solvers
(src/optim/bmConvergeCondition.m:6-255, src/optim/bmWitnessInfo.m:15-204).

The solvers run their iterations on the device without a host round trip
per iteration, so they use only ``nIter_max`` of a convergence condition --
which is also all the reference's solvers use: they call ``check()`` without
arguments, which leaves every other threshold inactive.  The witness object
receives the per-iteration histories the solvers return (objective, dafi,
regul, or the residual for CG-SENSE) and the final image.
"""
from __future__ import annotations

import time

import numpy as np


class bmConvergeCondition:
    """Iteration limit plus the optional distance / progression / travel
    thresholds of bmConvergeCondition.m:6-150."""

    def __init__(self, nIter_max=500, targDist_th=1e-4, targProgression_th=1e-4, y_travel_th=1e-4,
                 x_travel_th=1e-4, min_denom=1e-5, targProgression_flag=True):
        if isinstance(nIter_max, bmConvergeCondition):
            self.__dict__.update({k: (list(v) if isinstance(v, list) else v) for k, v in nIter_max.__dict__.items()})
            return
        self.nIter_max = int(nIter_max)
        self.targDist_th = targDist_th
        self.targProgression_th = targProgression_th
        self.y_travel_th = y_travel_th
        self.x_travel_th = x_travel_th
        self.min_denom = min_denom
        self.targProgression_flag = targProgression_flag
        self.nIter_curr = 0
        self.targDist_curr = self.targProgression_curr = self.y_travel_curr = self.x_travel_curr = None
        self.targDist_prev = None
        self.nIter_statu = self.targDist_statu = self.targProgression_statu = True
        self.y_travel_statu = self.x_travel_statu = True
        self.targDist_history, self.targProgression_history = [], []
        self.y_travel_history, self.x_travel_history = [], []
        self.startTime = self.endTime = None

    def check(self, targDist_curr=None, y_travel_curr=None, x_travel_curr=None):
        """Advances the iteration count and returns whether to continue
        (bmConvergeCondition.m:100-150)."""
        if self.nIter_curr == 0:
            self.startTime = time.time()
        self.nIter_curr += 1
        if targDist_curr is not None:
            self.targDist_prev = self.targDist_curr
            self.targDist_curr = abs(targDist_curr)
            self.targDist_history.append(self.targDist_curr)
            if self.targDist_prev is not None and self.targProgression_flag:
                self.targProgression_curr = abs(abs(targDist_curr) - abs(self.targDist_prev))
                self.targProgression_history.append(self.targProgression_curr)
        if y_travel_curr is not None:
            self.y_travel_curr = abs(y_travel_curr)
            self.y_travel_history.append(self.y_travel_curr)
        if x_travel_curr is not None:
            self.x_travel_curr = abs(x_travel_curr)
            self.x_travel_history.append(self.x_travel_curr)
        self.nIter_statu = self.nIter_curr <= self.nIter_max
        if self.targDist_curr is not None:
            self.targDist_statu = self.targDist_curr > self.targDist_th
        if self.targProgression_curr is not None:
            self.targProgression_statu = self.targProgression_curr > self.targProgression_th
        if self.y_travel_curr is not None:
            self.y_travel_statu = self.y_travel_curr > self.y_travel_th
        if self.x_travel_curr is not None:
            self.x_travel_statu = self.x_travel_curr > self.x_travel_th
        go = (self.nIter_statu and self.targDist_statu and self.targProgression_statu and self.y_travel_statu
              and self.x_travel_statu)
        if not go:
            self.endTime = time.time()
            if self.nIter_curr == self.nIter_max + 1:
                self.nIter_curr -= 1
        return go


class bmWitnessInfo:
    """Monitoring record of a solver run (bmWitnessInfo.m:15-120): its name,
    the iterations whose images are kept (witness_ind), the kept images and
    the named per-iteration histories (``param``)."""

    def __init__(self, witness_name=None, witness_ind=()):
        if isinstance(witness_name, bmWitnessInfo):
            self.__dict__.update(witness_name.__dict__)
            return
        self.witness_name = witness_name
        self.witness_ind = [int(i) for i in np.ravel(witness_ind)] if witness_ind is not None else []
        self.witness_im = {}
        self.param = {}
        self.creationTime = time.time()
        self.finalCallTime = None
        self.finalInd = None

    def record(self, history, names, x=None, final_ind=None):
        """Stores the columns of history [nIter, len(names)] under names and
        the final image when its iteration is a witness index."""
        h = np.asarray(history)
        if h.ndim == 1:
            h = h[:, None]
        for k, name in enumerate(names):
            self.param[name] = h[:, k].copy()
        self.finalInd = final_ind if final_ind is not None else h.shape[0]
        self.finalCallTime = time.time()
        if x is not None and self.finalInd in self.witness_ind:
            self.witness_im[self.finalInd] = x.detach().cpu().numpy() if hasattr(x, "detach") else np.array(x)


def n_iter(nIter) -> int:
    """The iteration count of an int or a bmConvergeCondition (nIter_max)."""
    return int(nIter.nIter_max) if isinstance(nIter, bmConvergeCondition) else int(nIter)
