"""Small seeded reconstruction problems shared by the GPU parity tests."""
import numpy as np

from monalisa_amd import synth
from oracle import mona_oracle as O


def make_problem(kind, seed=0, nCh=3, kernel="kaiser"):
    rng = np.random.default_rng(seed)
    if kind == "2d":
        n_u, N_u = [24, 24], [48, 48]
        nSpoke, nSample = 40, 48
        t = synth.radial2d_golden(nSpoke, nSample, k_max=12.0)
        ve = synth.ve_radial2(t, nSample, nSpoke)
    elif kind == "3d":
        n_u, N_u = [12, 12, 12], [24, 24, 24]
        nSeg, nShot, nSample = 10, 20, 24
        t = synth.koosh_ball(nSeg, nShot, nSample, k_max=6.0)
        ve = synth.ve_radial3(t, nSample, nSeg * nShot)
    elif kind == "3d32":
        # N_u = 32 = 16 * 2: the register-resident FFT passes
        n_u, N_u = [16, 16, 16], [32, 32, 32]
        nSeg, nShot, nSample = 10, 20, 32
        t = synth.koosh_ball(nSeg, nShot, nSample, k_max=8.0)
        ve = synth.ve_radial3(t, nSample, nSeg * nShot)
    elif kind == "3d32s":
        # 20 spokes: under 30% of the grid touched, the masked (sparse-frame) path
        n_u, N_u = [16, 16, 16], [32, 32, 32]
        nSeg, nShot, nSample = 2, 10, 32
        t = synth.koosh_ball(nSeg, nShot, nSample, k_max=8.0)
        ve = synth.ve_radial3(t, nSample, nSeg * nShot)
    elif kind == "2d_rect":
        n_u, N_u = [16, 24], [32, 48]
        nSpoke, nSample = 30, 32
        t = synth.radial2d_golden(nSpoke, nSample, k_max=8.0)
        t[1] *= 1.4
        ve = synth.ve_radial2(t, nSample, nSpoke)
    else:
        raise ValueError(kind)
    dK_u = [0.5] * len(N_u)          # FoV = 1, dK_n = 1
    C = synth.coil_maps(n_u, nCh, rng)
    x = synth.ellipsoid_phantom(n_u, rng)
    return dict(kind=kind, n_u=n_u, N_u=N_u, dK_u=dK_u, t=t, ve=ve, C=C, x=x, nCh=nCh,
                kernel=kernel, rng=rng)


def oracle_mats(p):
    return O.traj2sparsemat(p["t"], p["ve"], p["N_u"], p["dK_u"], p["kernel"])


def nd(flat, n_u, nCh=None):
    """[prod(n_u)] or [prod(n_u), nCh] (x fastest) -> n_u (+ nCh) ndarray."""
    shape = tuple(n_u) + ((nCh,) if nCh else ())
    return np.asarray(flat).reshape(shape, order="F")


def flat(a, n_u):
    nd_ = len(n_u)
    return a.reshape((-1,) + a.shape[nd_:], order="F")


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
