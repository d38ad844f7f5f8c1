"""The long rows of the adjoint gridding (rows over 64 entries, the k-space
centre) summed by ranges of consecutive samples, against the same rows
summed as 64-entry chunks (MONA_RANGE_KB=0) and against the float64 oracle.
Both are sequential sums regrouped into partial sums added in a fixed
order, so they agree to float rounding; the long rows are a large share of
the entries in this trajectory (a koosh ball dense at the centre).  The
range width is fixed per process, so each form runs in a child process of
its own (one at a time)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from problems import make_problem, oracle_mats, nd, rel
from oracle import mona_oracle as O

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
for nCh in (20, 3):
    p = make_problem("3d32", seed=6, nCh=nCh)
    t = synth.koosh_ball(20, 30, 32, k_max=8.0)
    ve = synth.ve_radial3(t, 32, 600)
    G = bmTraj2SparseMat(t, ve, p["N_u"], p["dK_u"], p["kernel"])
    P = Plan(p["N_u"], p["n_u"], p["dK_u"], nCh, p["kernel"])
    P.set_coils(torch.from_numpy(p["C"]).cuda())
    P.set_frame(0, G["Gu"], G["Gut"])
    P.set_frame(1, G["Gu"], G["Gut"])
    rng = np.random.default_rng(5)
    yy = torch.from_numpy((rng.normal(size=(P.nPt(), nCh)) + 1j * rng.normal(size=(P.nPt(), nCh)))
                          .astype(np.complex64)).cuda()
    ved = torch.from_numpy(ve.astype(np.float32)).cuda()
    xa = P.bmNakatsha(yy)
    x1 = torch.zeros(P.nu, dtype=torch.complex64, device="cuda")
    P.bmSensa(x1, yy, ved, 3)
    xb = torch.zeros((2, P.nu), dtype=torch.complex64, device="cuda")
    P.bmSensa_batch(xb, [yy, 0.5 * yy], [ved, ved], 3)
    torch.cuda.synchronize()
    out[f"y{{nCh}}"] = yy.cpu().numpy()
    out[f"xa{{nCh}}"] = xa.cpu().numpy()
    out[f"x1{{nCh}}"] = x1.cpu().numpy()
    out[f"xb{{nCh}}"] = xb.cpu().numpy()
    out[f"nmax{{nCh}}"] = np.array(int(G["Gut"].r_nJump.max()))
np.savez({path!r}, **out)
print("DONE", flush=True)
"""


def _run(range_kb, path):
    env = dict(os.environ)
    if range_kb is None:
        env.pop("MONA_RANGE_KB", None)
    else:
        env["MONA_RANGE_KB"] = str(range_kb)
    code = CHILD.format(root=ROOT, tests=os.path.join(ROOT, "tests"), path=str(path))
    print(f"[gpu] ranges child MONA_RANGE_KB={range_kb}", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return dict(np.load(path))


def test_long_rows_by_ranges_agree_with_chunks_and_the_oracle(tmp_path):
    a = _run(None, tmp_path / "ranges.npz")
    b = _run(0, tmp_path / "chunks.npz")
    c = _run(4, tmp_path / "narrow.npz")  # narrower ranges: more segments per row
    for nCh in (20, 3):
        assert int(a[f"nmax{nCh}"]) > 64
        for k in ("xa", "x1", "xb"):
            key = f"{k}{nCh}"
            assert np.abs(a[key]).sum() > 0, key
            assert rel(a[key], b[key]) < 1e-6, key
            assert rel(c[key], b[key]) < 1e-6, key
        # the batched solve equals the per-frame one bit for bit in each form
        for d in (a, b, c):
            np.testing.assert_array_equal(d[f"xb{nCh}"][0], d[f"x1{nCh}"])
        # the adjoint against the float64 oracle
        p = make_problem("3d32", seed=6, nCh=nCh)
        from monalisa_amd import synth
        p["t"] = synth.koosh_ball(20, 30, 32, k_max=8.0)
        p["ve"] = synth.ve_radial3(p["t"], 32, 600)
        R = oracle_mats(p)
        Cn = nd(p["C"].astype(np.complex128), p["n_u"], nCh)
        KFCc = O.kf_conj(np.conj(Cn), p["N_u"], p["n_u"], p["dK_u"], p["kernel"])
        xr = O.nakatsha(a[f"y{nCh}"].astype(np.complex128), R["Gut"], KFCc, True, p["N_u"], p["n_u"])
        assert rel(a[f"xa{nCh}"], xr.reshape(-1, order="F")) < 1e-5
