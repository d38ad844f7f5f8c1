"""Gridding-matrix construction (bmTraj2SparseMat.m:169-291 with the
conversion, l_squeeze and cpp_prepare of bmSparseMat.m:62-188) by the
library's host builder and its GPU builder, against the oracle's
construction (itself pinned by the brute-force cell loop on the CPU):
integer arrays bit for bit, m_val bit for bit from the host builder and
within one float ulp from the GPU builder (device exp/sqrt/Bessel)."""
import numpy as np
import pytest

from monalisa_amd import synth
from oracle import mona_oracle as O

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(4)
    t2 = synth.radial2d_golden(24, 48, k_max=12.0)
    t3 = synth.koosh_ball(4, 20, 32, k_max=8.0)
    t_out = t2.copy()
    t_out[:, ::7] *= 1.6          # some samples outside the grid: deleted
    t_small = rng.uniform(-1.0, 1.0, size=(2, 50))
    t_small3 = rng.uniform(-1.0, 1.0, size=(3, 40))
    return [
        ("2d kaiser", t2, synth.ve_radial2(t2, 48, 24), [48, 48], [0.5, 0.5], "kaiser", None),
        ("3d kaiser", t3, synth.ve_radial3(t3, 32, 80), [32, 32, 32], [0.5] * 3, "kaiser", None),
        ("2d gauss nWin 4", t2, synth.ve_radial2(t2, 48, 24), [48, 48], [0.5, 0.5], "gauss", 4),
        ("3d gauss nWin 3", t3, synth.ve_radial3(t3, 32, 80), [32, 32, 32], [0.5] * 3, "gauss", 3),
        ("2d deleted samples", t_out, np.ones(t_out.shape[1]), [48, 48], [0.5, 0.5], "kaiser", None),
        ("2d wrapped duplicates", t_small, np.ones(50), [4, 4], [0.5, 0.5], "kaiser", 5),
        ("3d wrapped duplicates", t_small3, rng.uniform(0.5, 2.0, 40), [4, 4, 4], [0.5] * 3, "kaiser", 5),
        ("2d kaiser nWin 5 rect", t2, synth.ve_radial2(t2, 48, 24), [48, 40], [0.5, 0.6], "kaiser", 5),
    ]


CASES = _cases()


def _c0():
    t = synth.radial2d_golden(256, 512, k_max=128.0)
    return ("C0 2D N=256 512^2 256x512", t, synth.ve_radial2(t, 512, 256), [512, 512], [0.5, 0.5],
            "kaiser", None)


def _compare(R, G, which, exact_val):
    for k in which:
        r, g = R[k], G[k]
        assert (g.r_size, g.l_size, g.l_nJump, g.type) == (r["r_size"], r["l_size"], r["l_nJump"], r["type"]), k
        np.testing.assert_array_equal(g.r_jump, r["r_jump"], err_msg=k)
        np.testing.assert_array_equal(g.r_nJump, r["r_nJump"], err_msg=k)
        if r["l_jump"] is None:
            assert g.l_jump is None, k
        else:
            np.testing.assert_array_equal(g.l_jump, r["l_jump"], err_msg=k)
        if exact_val:
            np.testing.assert_array_equal(g.m_val.view(np.int32), r["m_val"].view(np.int32), err_msg=k)
        else:
            ulp = np.abs(g.m_val.view(np.int32).astype(np.int64) - r["m_val"].view(np.int32).astype(np.int64))
            assert ulp.max() <= 1, (k, int(ulp.max()))


@pytest.mark.parametrize("case", range(len(CASES) + 1))
def test_builders_equal_oracle(case):
    from monalisa_amd.sparse import bmTraj2SparseMat
    name, t, ve, N_u, dK_u, kern, nWin = CASES[case] if case < len(CASES) else _c0()
    print(f"[gpu] structure vs oracle: {name}", flush=True)
    which = ("Gn", "Gu", "Gut")
    R = O.traj2sparsemat(t, ve, N_u, dK_u, kern, nWin, which=which)
    H = bmTraj2SparseMat(t, ve, N_u, dK_u, kern, nWin, which=which)
    _compare(R, H, which, exact_val=True)
    D = bmTraj2SparseMat(t, ve, N_u, dK_u, kern, nWin, which=which, device=True)
    _compare(R, D, which, exact_val=False)
