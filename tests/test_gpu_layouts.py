"""The y-major 3-D grid and intermediate layouts change where values are
stored, not how any value is computed: the forward and adjoint operators, a
CG-SENSE solve and a batched two-frame solve give the same bits with the
plain [z][y][x] order (MONA_FFT_YZ=0).  The layout is fixed per process, so
each order runs in a child process of its own (one at a time)."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
import torch
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
from problems import make_problem
from monalisa_amd.nufft import Plan
from monalisa_amd.sparse import bmTraj2SparseMat
out = {{}}
for kind, nCh in (("3d32", 4), ("3d32s", 4), ("3d", 3)):
    p = make_problem(kind, nCh=nCh, seed=9)
    G = bmTraj2SparseMat(p["t"], p["ve"], p["N_u"], p["dK_u"], p["kernel"])
    P = Plan(p["N_u"], p["n_u"], p["dK_u"], nCh, p["kernel"])
    P.set_coils(torch.from_numpy(p["C"]).cuda())
    P.set_frame(0, G["Gu"], G["Gut"])
    P.set_frame(1, G["Gu"], G["Gut"])
    x = torch.from_numpy(p["x"]).cuda()
    y = P.bmShanna(x)
    xa = P.bmNakatsha(y)
    ve = torch.from_numpy(np.asarray(p["ve"], np.float32)).cuda()
    xs = torch.zeros_like(x)
    P.bmSensa(xs, y, ve, 3)
    xb = torch.zeros((2, x.numel()), dtype=torch.complex64, device="cuda")
    P.bmSensa_batch(xb, [y, 0.5 * y], [ve, ve], 3)
    torch.cuda.synchronize()
    for name, t in (("y", y), ("xa", xa), ("xs", xs), ("xb", xb)):
        out[kind + "_" + name] = t.cpu().numpy()
np.savez({path!r}, **out)
print("DONE", flush=True)
"""


def _run(yz, path):
    env = dict(os.environ)
    env["MONA_FFT_YZ"] = "1" if yz else "0"
    code = CHILD.format(root=ROOT, tests=os.path.join(ROOT, "tests"), path=str(path))
    print(f"[gpu] layouts child MONA_FFT_YZ={env['MONA_FFT_YZ']}", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return dict(np.load(path))


def test_y_major_layout_is_bit_identical_to_plain_order(tmp_path):
    a = _run(True, tmp_path / "yz.npz")
    b = _run(False, tmp_path / "plain.npz")
    assert a.keys() == b.keys() and len(a) == 12
    for k in a:
        assert np.abs(a[k]).sum() > 0, k
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
