"""Loop control and monitoring objects of the reference's solvers
(src/optim/bmConvergeCondition.m:6-255, src/optim/bmWitnessInfo.m:15-204).

The solvers run their iterations on the device without a host round trip
per iteration, so they use only ``nIter_max`` of a convergence condition --
which is also all the reference's solvers use: they call ``check()`` without
arguments, which leaves every other threshold inactive.  The witness object
receives the per-iteration histories the solvers return (objective, dafi,
regul, or the residual for CG-SENSE) and the snapshots the solvers take on
the device after every outer iteration.
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
    """Monitoring record of a solver run (bmWitnessInfo.m:15-190): its name,
    the iterations whose full images are saved (witness_ind, 1-based), the
    witness images of every iteration (witness_im), their times, and the
    named per-iteration histories (``param``).

    The solvers take their snapshots on the device after every outer
    iteration and hand them to ``watch`` when the run returns, in the
    reference's order (one 'loop' call per iteration, then 'final'); the
    times are therefore those of these calls.  Files go to ``save_dir``
    (default: the working directory, as the reference's ``save`` calls) as
    NumPy files in place of MATLAB v7.3 files: ``wit_<name>.npz`` after every
    call, ``x_<name>.npy`` (and ``z_``, ``u_``) at the witness iterations;
    ``save_dir = None`` writes nothing.  The reference's per-iteration
    progress print is left out."""

    def __init__(self, witness_name=None, witness_ind=(), save_dir="."):
        if isinstance(witness_name, bmWitnessInfo):
            self.__dict__.update({k: (list(v) if isinstance(v, list) else v)
                                  for k, v in witness_name.__dict__.items()})
            return
        self.witness_name = witness_name
        self.witness_ind = [int(i) for i in np.ravel(witness_ind)] if witness_ind is not None else []
        self.witness_im = None
        self.witness_time = []
        self.creationTime = time.time()
        self.finalCallTime = self.creationTime
        self.finalInd = None
        self.param = {}
        self.save_dir = save_dir

    def record(self, history, names):
        """Stores the columns of history [nIter, len(names)] under names."""
        h = np.asarray(history)
        if h.ndim == 1:
            h = h[:, None]
        for k, name in enumerate(names):
            self.param[name] = h[:, k].copy()

    @staticmethod
    def _frame0(x):
        if isinstance(x, (list, tuple)):
            x = x[0]
        return np.asarray(x.detach().cpu().numpy() if hasattr(x, "detach") else x)

    def watch(self, curr_ind, x, n_u, loop_or_final, z=None, u=None, image=None):
        """bmWitnessInfo.watch (bmWitnessInfo.m:92-190): with a non-empty
        witness_ind, appends the witness image of x -- frame 0 (the first
        image of a list), whole in 1-D (as a column) and 2-D, its z slice
        fix(n_u(end)/2) + 1 in 3-D, along the third axis -- records the time
        and curr_ind, saves the record, and saves x (and z, u) when curr_ind
        is a witness index.  image: the witness image itself (the solvers'
        device snapshots); x may then be None for iterations whose full
        image was not kept."""
        if not self.witness_ind:
            return
        n_u = [int(v) for v in np.ravel(n_u)]
        d = len(n_u)
        if image is None:
            # frame 0: the first image of a list, the first row of [nFr, nu]
            im = self._frame0(x).reshape(-1, int(np.prod(n_u)))[0].reshape(n_u, order="F")
            if d == 3:
                im = im[:, :, n_u[2] // 2]
        else:
            im = np.asarray(image).reshape(n_u if d < 3 else n_u[:2], order="F")
        if d == 1:
            im = im.reshape(-1, 1)
            self.witness_im = im if self.witness_im is None else np.concatenate([self.witness_im, im], axis=1)
        else:
            im = im[:, :, None]
            self.witness_im = im if self.witness_im is None else np.concatenate([self.witness_im, im], axis=2)
        now = time.time()
        self.witness_time.append(now)
        self.finalCallTime = now
        self.finalInd = int(curr_ind)
        if self.save_dir is None:
            return
        import os
        name = str(self.witness_name)
        rec = {"witness_im": self.witness_im, "witness_time": np.array(self.witness_time),
               "witness_ind": np.array(self.witness_ind), "finalInd": self.finalInd,
               "creationTime": self.creationTime, "finalCallTime": self.finalCallTime}
        for k, v in self.param.items():
            a = np.asarray(v)
            if a.dtype != object:
                rec["param_" + str(k)] = a
        np.savez(os.path.join(self.save_dir, f"wit_{name}.npz"), **rec)
        if int(curr_ind) in self.witness_ind:
            nu = int(np.prod(n_u))
            for tag, v in (("x", x), ("z", z), ("u", u)):
                if v is None:
                    continue
                a = np.stack([self._frame0(f) for f in v]) if isinstance(v, (list, tuple)) else self._frame0(v)
                # images in the block shape n_u (x fastest), frames first
                a = a.reshape(n_u, order="F") if a.size == nu else \
                    np.stack([r.reshape(n_u, order="F") for r in a.reshape(-1, nu)])
                np.save(os.path.join(self.save_dir, f"{tag}_{name}.npy"), a)


def n_iter(nIter) -> int:
    """The iteration count of an int or a bmConvergeCondition (nIter_max)."""
    return int(nIter.nIter_max) if isinstance(nIter, bmConvergeCondition) else int(nIter)
