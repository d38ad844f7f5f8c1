"""Pins of the 3-D input generators the C1-C3 workloads use
(monalisa_amd/synth.py): the spiral-phyllotaxis spoke angles
(src/traj3/bmPhyllotaxisAngle3.m:21-61), the full-radial koosh-ball
trajectory (src/traj3/bmTraj_fullRadial3_phyllotaxis_lineAssym2.m:43-70) and
its volume elements (src/geom3/bmVolumeElement_full_radial3.m:42-49), by a
hand-worked table and by what the geometry fixes: the polar angles are
(pi/2) sqrt(c / n) for the counters c = 1..n, consecutive counters turn by
the golden angle, every spoke is a line through the origin, and the volume
elements of a full-radial trajectory sum to the volume of its ball."""
import math

import numpy as np
import pytest

from monalisa_amd import synth

GOLD = 2 * math.pi - 2 * math.pi / ((1 + math.sqrt(5)) / 2)   # 137.508 degrees


def test_phyllotaxis_hand_table():
    """nSeg = 2, nShot = 3.  The counter runs over segments (outer) and shots
    (inner) while spoke (i, j) sits at index (i-1) + (j-1) nSeg, so the six
    indices carry counters 1, 4, 2, 5, 3, 6."""
    theta, phi = synth.phyllotaxis_angles(2, 3)
    c = np.array([1, 4, 2, 5, 3, 6], float)
    q = math.pi / (2 * math.sqrt(6))
    np.testing.assert_allclose(theta, q * np.sqrt(c), rtol=0, atol=1e-15)
    np.testing.assert_allclose(phi, np.mod(c * GOLD, 2 * math.pi), rtol=0, atol=1e-12)


def test_phyllotaxis_self_navigation():
    """With self-navigation the first segment of every shot points along z
    (theta = phi = 0) and the other spokes spread over n - nShot counters."""
    nSeg, nShot = 4, 5
    theta, phi = synth.phyllotaxis_angles(nSeg, nShot, selfNav=True)
    first = np.arange(nShot) * nSeg
    assert np.all(theta[first] == 0) and np.all(phi[first] == 0)
    rest = np.setdiff1d(np.arange(nSeg * nShot), first)
    n_pure = nSeg * nShot - nShot
    np.testing.assert_allclose(np.sort(theta[rest]), math.pi / 2 * np.sqrt(np.arange(1, n_pure + 1) / n_pure),
                               rtol=1e-14)


@pytest.mark.parametrize("nSeg,nShot", [(25, 400), (7, 13)])
def test_phyllotaxis_spiral(nSeg, nShot):
    """The polar angles are (pi/2) sqrt(c/n), c = 1..n (the north hemisphere,
    the last spoke on the equator), and along increasing polar angle the
    azimuth advances by the golden angle."""
    theta, phi = synth.phyllotaxis_angles(nSeg, nShot)
    n = nSeg * nShot
    order = np.argsort(theta, kind="stable")
    np.testing.assert_allclose(theta[order], math.pi / 2 * np.sqrt(np.arange(1, n + 1) / n), rtol=1e-14)
    assert theta.max() == pytest.approx(math.pi / 2, rel=1e-15)
    step = np.mod(np.diff(phi[order]), 2 * math.pi)
    np.testing.assert_allclose(step, GOLD, atol=1e-9)


def test_koosh_ball_spokes():
    """Every spoke is nSample points on a line through the origin along
    (sin t cos p, sin t sin p, cos t), at radii (s - nSample/2) / nSample
    times 2 k_max dK_n: the first at -k_max, the centre sample exactly 0."""
    nSeg, nShot, nS, k_max, dK = 3, 5, 16, 6.0, 1.5
    t = synth.koosh_ball(nSeg, nShot, nS, k_max, dK_n=dK).reshape(3, nS, nSeg * nShot, order="F")
    theta, phi = synth.phyllotaxis_angles(nSeg, nShot)
    e = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    r = (np.arange(nS) - nS / 2) / nS * 2 * k_max * dK
    np.testing.assert_allclose(t, r[None, :, None] * e[:, None, :], atol=1e-12)
    assert np.all(t[:, nS // 2, :] == 0)
    np.testing.assert_allclose(np.linalg.norm(t[:, 0, :], axis=0), k_max * dK, rtol=1e-14)


def test_volume_elements_fill_the_ball():
    """A full-radial trajectory's volume elements (4 pi |k|^2 dr / (2 nLine),
    the centre sample (4/3) pi (dr/2)^3 / (2 nLine)) sum to the volume of the
    sampled ball (4/3) pi k_max^3 up to the O(1/nSample) quadrature error;
    along a spoke they grow as |k|^2."""
    nSeg, nShot, nS, k_max = 10, 40, 128, 32.0
    t = synth.koosh_ball(nSeg, nShot, nS, k_max)
    nL = nSeg * nShot
    v = synth.ve_radial3(t, nS, nL).reshape(nS, nL, order="F")
    ball = 4 / 3 * math.pi * k_max ** 3
    assert abs(v.sum() / ball - 1) < 3.0 / nS
    dr = 2 * k_max / nS
    np.testing.assert_allclose(v[nS // 2, :], 4 / 3 * math.pi * (dr / 2) ** 3 / (2 * nL), rtol=1e-12)
    k = np.abs(np.arange(nS) - nS / 2) * dr
    m = np.arange(nS) != nS // 2
    np.testing.assert_allclose(v[m, 0] / v[nS // 2 + 1, 0], (k[m] / k[nS // 2 + 1]) ** 2, rtol=1e-12)
