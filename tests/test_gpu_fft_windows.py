"""The register FFT passes whose window is exactly the central half (the
zero-filled forward passes read rows [N/4, 3N/4), the cropping adjoint
passes keep them) take that window at compile time: the rows outside are
zero or dropped in the code itself, and the first radix-4 butterflies of the
forward passes skip their zero inputs.  They compute the same transform as
the run-time windows (MONA_FFT_WIN=0) to float rounding: with fewer uses of
some products the compiler fuses them into multiply-adds differently, so the
low bits differ, while a dropped row or a wrong window would differ at O(1).
This compares the two on the lengths that have such passes (N = 16 N2 with N2
a multiple of 4: 64, 128, 192), dense and sparse (masked) frames, float and
bf16 storage (where a value near a rounding boundary of the 8-bit significand
may round the other way), the operators and a short CG-SENSE solve.  The
switch is read once per process, so each setting runs in a child process of
its own (one at a time)."""
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
sys.path.insert(0, {root!r})
from monalisa_amd import synth
from monalisa_amd.nufft import Plan
from monalisa_amd.sparse import bmTraj2SparseMat
# (name, n_u, N_u, trajectory, coils); k_max = N_u / 4 with dK_u = 1/2
cases = [
    ("3d64", [32] * 3, [64] * 3, ("koosh", 10, 20, 64), 4),
    ("3d64s", [32] * 3, [64] * 3, ("koosh", 2, 10, 64), 3),
    ("3d128s", [64] * 3, [128] * 3, ("koosh", 4, 10, 128), 2),
    ("3d192s", [96] * 3, [192] * 3, ("koosh", 4, 10, 192), 2),
    ("2d128", [64, 64], [128, 128], ("radial", 60, 128), 3),
]
out = {{}}
for name, n_u, N_u, tr, nCh in cases:
    print("case", name, flush=True)
    rng = np.random.default_rng(5)
    k_max = N_u[0] / 4.0
    if tr[0] == "koosh":
        t = synth.koosh_ball(tr[1], tr[2], tr[3], k_max=k_max)
        ve = synth.ve_radial3(t, tr[3], tr[1] * tr[2])
    else:
        t = synth.radial2d_golden(tr[1], tr[2], k_max=k_max)
        ve = synth.ve_radial2(t, tr[2], tr[1])
    dK_u = [0.5] * len(N_u)
    C = synth.coil_maps(n_u, nCh, rng)
    x0 = synth.ellipsoid_phantom(n_u, rng)
    G = bmTraj2SparseMat(t, ve, N_u, dK_u, "kaiser")
    P = Plan(N_u, n_u, dK_u, nCh, "kaiser")
    P.set_coils(torch.from_numpy(C).cuda())
    P.set_frame(0, G["Gu"], G["Gut"])
    x = torch.from_numpy(x0).cuda()
    y = P.bmShanna(x)
    yy = torch.from_numpy((rng.normal(size=tuple(y.shape)) + 1j * rng.normal(size=tuple(y.shape)))
                          .astype(np.complex64)).cuda()
    xa = P.bmNakatsha(yy)
    y16 = P.bmShanna_bf16(x)
    x16 = P.bmNakatsha_bf16(yy)
    ved = torch.from_numpy(np.asarray(ve, dtype=np.float32).reshape(-1)).cuda()
    x1 = torch.zeros(P.nu, dtype=torch.complex64, device="cuda")
    P.bmSensa(x1, yy, ved, 3)
    torch.cuda.synchronize()
    for k, v in (("y", y), ("xa", xa), ("y16", y16), ("x16", x16), ("x1", x1)):
        out[name + "_" + k] = v.cpu().numpy()
np.savez({path!r}, **out)
print("DONE", flush=True)
"""


def _run(win, path):
    env = dict(os.environ)
    if win is None:
        env.pop("MONA_FFT_WIN", None)
    else:
        env["MONA_FFT_WIN"] = str(win)
    code = CHILD.format(root=ROOT, path=str(path))
    print(f"[gpu] fft windows child MONA_FFT_WIN={win}", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-1000:] + r.stderr[-2000:]
    return dict(np.load(path))


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def test_compile_time_windows_match_run_time_windows(tmp_path):
    ct = _run(None, tmp_path / "ct.npz")
    rt = _run(0, tmp_path / "rt.npz")
    assert sorted(ct) == sorted(rt)
    errs = {}
    for k in ct:
        assert np.abs(rt[k]).sum() > 0, k
        assert np.isfinite(ct[k]).all(), k
        errs[k] = _rel(ct[k], rt[k])
    print("relative differences", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for k, e in errs.items():
        # float: a few roundings of the transform; bf16 storage: values that
        # round to the other neighbour; 3 CG steps: the float differences
        # carried through the solve
        tol = 2e-3 if k.endswith(("y16", "x16")) else 1e-5 if k.endswith("x1") else 1e-6
        assert e < tol, (k, e)
