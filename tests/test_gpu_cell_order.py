"""The order of the grid cells inside a tile in the adjoint-gridding rows
(cubes of 2 cells a side by default, cubes of 4, the tile's raster order,
the Morton order: MONA_CELL_CUBE) changes which rows share a staged group,
never a row's entries or their order, so the adjoint and the CG-SENSE
solve that uses it are equal bit for bit under every order.  The order is
read once per process, so each runs in a child process of its own (one at
a time)."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import numpy as np
import torch
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
from problems import make_problem
from monalisa_amd import synth
from monalisa_amd.nufft import Plan
from monalisa_amd.sparse import bmTraj2SparseMat
out = {{}}
for kind, nCh in (("3d32", 20), ("3d32", 3), ("3d", 4)):
    p = make_problem(kind, seed=11, nCh=nCh)
    G = bmTraj2SparseMat(p["t"], p["ve"], p["N_u"], p["dK_u"], p["kernel"])
    P = Plan(p["N_u"], p["n_u"], p["dK_u"], nCh, p["kernel"])
    P.set_coils(torch.from_numpy(p["C"]).cuda())
    P.set_frame(0, G["Gu"], G["Gut"])
    rng = np.random.default_rng(3)
    yy = torch.from_numpy((rng.normal(size=(P.nPt(), nCh)) + 1j * rng.normal(size=(P.nPt(), nCh)))
                          .astype(np.complex64)).cuda()
    ved = torch.from_numpy(np.asarray(p["ve"], dtype=np.float32).reshape(-1)).cuda()
    xa = P.bmNakatsha(yy)
    x1 = torch.zeros(P.nu, dtype=torch.complex64, device="cuda")
    P.bmSensa(x1, yy, ved, 4)
    torch.cuda.synchronize()
    out[f"xa_{{kind}}_{{nCh}}"] = xa.cpu().numpy()
    out[f"x1_{{kind}}_{{nCh}}"] = x1.cpu().numpy()
np.savez({path!r}, **out)
print("DONE", flush=True)
"""


def _run(cube, path):
    env = dict(os.environ)
    if cube is None:
        env.pop("MONA_CELL_CUBE", None)
    else:
        env["MONA_CELL_CUBE"] = str(cube)
    code = CHILD.format(root=ROOT, tests=os.path.join(ROOT, "tests"), path=str(path))
    print(f"[gpu] cell-order child MONA_CELL_CUBE={cube}", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return dict(np.load(path))


def test_cell_orders_give_identical_adjoints_and_solves(tmp_path):
    ref = _run(None, tmp_path / "default.npz")
    for cube in (8, 4, 0):
        d = _run(cube, tmp_path / f"cube{cube}.npz")
        assert sorted(d) == sorted(ref)
        for k in ref:
            assert np.abs(ref[k]).sum() > 0, k
            np.testing.assert_array_equal(d[k], ref[k], err_msg=f"{k} MONA_CELL_CUBE={cube}")
