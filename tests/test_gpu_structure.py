"""Gridding-matrix construction on the GPU against the host builder (itself
bit-exact against the oracle's brute-force construction): identical integer
arrays, m_val within float rounding."""
import numpy as np
import pytest

from monalisa_amd import synth

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(4)
    t2 = synth.radial2d_golden(24, 48, k_max=12.0)
    t3 = synth.koosh_ball(4, 20, 32, k_max=8.0)
    t_out = t2.copy()
    t_out[:, ::7] *= 1.6          # some samples outside the grid: deleted
    t_small = rng.uniform(-1.0, 1.0, size=(2, 50))
    return [
        ("2d kaiser", t2, synth.ve_radial2(t2, 48, 24), [48, 48], [0.5, 0.5], "kaiser", None),
        ("3d kaiser", t3, synth.ve_radial3(t3, 32, 80), [32, 32, 32], [0.5] * 3, "kaiser", None),
        ("2d gauss nWin 4", t2, synth.ve_radial2(t2, 48, 24), [48, 48], [0.5, 0.5], "gauss", 4),
        ("2d deleted samples", t_out, np.ones(t_out.shape[1]), [48, 48], [0.5, 0.5], "kaiser", None),
        ("2d wrapped duplicates", t_small, np.ones(50), [4, 4], [0.5, 0.5], "kaiser", 5),
    ]


@pytest.mark.parametrize("case", range(5))
def test_device_build_equals_host(case):
    from monalisa_amd.sparse import bmTraj2SparseMat
    name, t, ve, N_u, dK_u, kern, nWin = _cases()[case]
    print(f"[gpu] structure {name}", flush=True)
    which = ("Gn", "Gu", "Gut")
    H = bmTraj2SparseMat(t, ve, N_u, dK_u, kern, nWin, which=which)
    D = bmTraj2SparseMat(t, ve, N_u, dK_u, kern, nWin, which=which, device=True)
    for k in which:
        h, d = H[k], D[k]
        assert (h.r_size, h.l_size, h.l_nJump, h.type) == (d.r_size, d.l_size, d.l_nJump, d.type), k
        np.testing.assert_array_equal(d.r_jump, h.r_jump)
        np.testing.assert_array_equal(d.r_nJump, h.r_nJump)
        if h.l_jump is None:
            assert d.l_jump is None
        else:
            np.testing.assert_array_equal(d.l_jump, h.l_jump)
        np.testing.assert_allclose(d.m_val, h.m_val, rtol=3e-7, atol=0)
