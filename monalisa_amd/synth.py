"""Seeded synthetic inputs for the benchmark and test workloads: radial
trajectories, volume elements (density compensation), coil maps and
phantoms.  Every generator takes an explicit ``numpy.random.Generator`` or
seed so that the same call yields the same arrays on any machine.
"""
from __future__ import annotations

import math

import numpy as np

GOLDEN_ANGLE_2D = math.pi * (3.0 - math.sqrt(5.0))  # 111.246... degrees


def phyllotaxis_angles(nSeg, nShot, selfNav=False):
    """Spiral-phyllotaxis spoke angles (src/traj3/bmPhyllotaxisAngle3.m:21-61):
    spoke (i, j) at index (i-1) + (j-1)*nSeg gets phi = mod(c*goldAngle, 2 pi),
    theta = pi/(2 sqrt(nseg_pure)) * sqrt(c) with the counter c running over
    i outer, j inner."""
    gold = (1 + math.sqrt(5)) / 2
    gold_angle = 2 * math.pi - 2 * math.pi / gold
    ntot = nSeg * nShot
    pure = ntot - nShot if selfNav else ntot
    q = math.pi / (2 * math.sqrt(pure))
    phi = np.zeros(ntot)
    theta = np.zeros(ntot)
    c = 1
    for i in range(nSeg):
        for j in range(nShot):
            idx = i + j * nSeg
            if selfNav and i == 0:
                continue
            phi[idx] = math.fmod(c * gold_angle, 2 * math.pi)
            theta[idx] = q * math.sqrt(c)
            c += 1
    return theta, phi


def koosh_ball(nSeg, nShot, nSample, k_max, dK_n=1.0):
    """3-D full-radial phyllotaxis trajectory, [3, nSample * nSeg * nShot]
    (src/traj3/bmTraj_fullRadial3_phyllotaxis_lineAssym2.m:9-71) with
    nSample points per spoke at r in [-1/2, 1/2) (step 1/nSample) scaled to
    k in [-k_max, k_max) * dK_n."""
    theta, phi = phyllotaxis_angles(nSeg, nShot)
    r = (np.arange(nSample) - nSample / 2) / nSample
    R = r[:, None]
    x = R * np.cos(phi)[None, :] * np.sin(theta)[None, :]
    y = R * np.sin(phi)[None, :] * np.sin(theta)[None, :]
    z = R * np.cos(theta)[None, :] * np.ones_like(R)
    t = np.stack([x, y, z], 0) * (2 * k_max) * dK_n       # [3, nSample, nSpoke]
    return t.reshape(3, -1, order="F")


def radial2d_golden(nSpoke, nSample, k_max, dK_n=1.0):
    """2-D full-radial spokes at golden-angle increments, [2, nSample*nSpoke],
    nSample points per spoke at k in [-k_max, k_max) * dK_n."""
    ang = np.arange(nSpoke) * GOLDEN_ANGLE_2D
    r = (np.arange(nSample) - nSample / 2) / nSample * (2 * k_max) * dK_n
    x = r[:, None] * np.cos(ang)[None, :]
    y = r[:, None] * np.sin(ang)[None, :]
    return np.stack([x, y], 0).reshape(2, -1, order="F")


def bmTraj_fullRadial2_lineAssym2(N, nSeg, dK_n):
    """2-D full-radial trajectory with nSeg lines at equally spread angles
    phi = 2 pi s / nSeg (src/traj2/bmTraj_fullRadial2_lineAssym2.m:7-35):
    N points per line at r = (n - fix(N/2)) kMax / fix(N/2), n = 0 .. N-1,
    kMax = N dK_n / 2.  Returns [2, N * nSeg], the points of a line
    consecutive."""
    N, nSeg = int(N), int(nSeg)
    kMax = N * dK_n / 2
    phi = np.linspace(0, 2 * math.pi, nSeg + 1)[:-1]
    r = (np.arange(N) - N // 2) * kMax / (N // 2)
    x = r[:, None] * np.cos(phi)[None, :]
    y = r[:, None] * np.sin(phi)[None, :]
    return np.stack([x, y], 0).reshape(2, -1, order="F")


def ve_radial3(t, nSample, nSpoke):
    """Volume elements of a 3-D full-radial trajectory
    (src/geom3/bmVolumeElement_full_radial3.m:42-49): dr * 4 pi |k|^2 /
    (2 nLine), the centre point of each line gets (4/3) pi dr0^3 / (2 nLine)."""
    k = np.sqrt((t ** 2).sum(0)).reshape(nSample, nSpoke, order="F")
    dr = np.abs(np.diff(k[:, 0]))
    dr = float(np.median(dr)) if dr.size else 1.0
    v = dr * 4 * math.pi * k ** 2 / (2 * nSpoke)
    v0 = (4.0 / 3.0) * math.pi * (dr / 2) ** 3 / (2 * nSpoke)
    v[nSample // 2, :] = v0
    return v.reshape(-1, order="F")


def ve_radial2(t, nSample, nSpoke):
    """Ramp density compensation of 2-D full-radial spokes
    (src/geom2/bmVolumeElement_voronoi_full_radial2.m:11-77 for equally
    spread angles): dr * |k| * pi / nSpoke, centre points pi (dr/2)^2 / nSpoke."""
    k = np.sqrt((t ** 2).sum(0)).reshape(nSample, nSpoke, order="F")
    dr = float(np.median(np.abs(np.diff(k[:, 0]))))
    v = dr * k * math.pi / nSpoke
    v[nSample // 2, :] = math.pi * (dr / 2) ** 2 / nSpoke
    return v.reshape(-1, order="F")


def _volume_element_1d(x):
    """Widths of the 1-D Voronoi cells of the values x (midpoints between
    sorted neighbours; the two end cells mirror their inner half), in the
    order of x (src/geom1/bmVolumeElement1.m)."""
    x = np.asarray(x, dtype=np.float64)
    order = np.argsort(x, kind="stable")
    xs = x[order]
    mid = (xs[1:] + xs[:-1]) / 2
    edges = np.concatenate([[xs[0] - (mid[0] - xs[0])], mid, [xs[-1] + (xs[-1] - mid[-1])]])
    w = np.empty_like(x)
    w[order] = np.diff(edges)
    return w


def bmVolumeElement_voronoi_full_radial2(t, N):
    """Volume elements of a 2-D full-radial trajectory of lines of N points
    (N even, the zero at index N/2 of each line)
    (src/geom2/bmVolumeElement_voronoi_full_radial2.m:11-77): per point the
    radial Voronoi width along its line times |k| times the angular Voronoi
    width of the line's direction among the 2 nLine directions +-e; the
    centre point of every line gets pi (mean dr_0 / 2)^2 / (2 nLine).
    t: [2, N * nLine], points of a line consecutive.  (The reference infers
    N from t; here it is given.)"""
    t = np.asarray(t, dtype=np.float64)
    N = int(N)
    if t.shape[0] != 2 or N % 2 or t.shape[1] % N:
        raise ValueError("2-D trajectory of whole lines of an even number of points required")
    nLine = t.shape[1] // N
    tl = t.reshape(2, N, nLine, order="F")
    # line directions (bmTraj_lineDirection.m): mean unit vector from the
    # first point to the points of the second half
    d = tl[:, (N + 1) // 2 - 1:, :] - tl[:, :1, :]
    d = d / np.sqrt((d ** 2).sum(0, keepdims=True))
    e = d.mean(1)                                               # [2, nLine]
    dr = np.stack([_volume_element_1d(tl[:, :, i].T @ e[:, i]) for i in range(nLine)], 1)   # [N, nLine]
    ee = np.concatenate([e, -e], 1)
    ang = np.arctan2(ee[1], ee[0])
    order = np.argsort(ang, kind="stable")
    a = ang[order]
    cut = (math.pi - a[-1]) + (a[0] + math.pi)
    w = _volume_element_1d(a)
    w[0] = (a[1] - a[0]) / 2 + cut / 2
    w[-1] = (a[-1] - a[-2]) / 2 + cut / 2
    wa = np.empty_like(w)
    wa[order] = w
    ds = np.sqrt((tl ** 2).sum(0)) * wa[None, :nLine]
    v = (dr * ds).reshape(-1, order="F")
    dr0 = dr[N // 2, :].mean() / 2
    v[N // 2::N] = math.pi * dr0 ** 2 / (2 * nLine)
    return v


def coil_maps(n_u, nCh, rng):
    """Smooth Gaussian-magnitude blobs (centres on a ring/sphere around the
    FoV) times a low-order polynomial phase, normalised so that
    sum_c |C_c|^2 = 1.  Returns [prod(n_u), nCh] complex64, x fastest."""
    n_u = [int(v) for v in n_u]
    d = len(n_u)
    grids = np.meshgrid(*[(np.arange(n) - n / 2) / n for n in n_u], indexing="ij")
    C = np.zeros(tuple(n_u) + (nCh,), dtype=np.complex128)
    for c in range(nCh):
        direction = rng.normal(size=d)
        direction /= np.linalg.norm(direction) + 1e-12
        centre = 0.6 * direction
        width = 0.35 + 0.15 * rng.random()
        r2 = sum((g - cc) ** 2 for g, cc in zip(grids, centre))
        mag = np.exp(-r2 / (2 * width ** 2))
        coef = rng.normal(scale=1.5, size=(d + 1,))
        phase = coef[0] + sum(coef[i + 1] * grids[i] for i in range(d))
        phase = phase + 0.5 * rng.normal() * grids[0] * grids[-1]
        C[..., c] = mag * np.exp(1j * phase)
    ss = np.sqrt((np.abs(C) ** 2).sum(-1, keepdims=True))
    C = C / np.maximum(ss, 1e-12)
    return np.ascontiguousarray(C.reshape(-1, nCh, order="F"), dtype=np.complex64)


def ellipsoid_phantom(n_u, rng, n_ellipsoids=8):
    """Piecewise-constant phantom of random ellipsoids inside the FoV,
    [prod(n_u)] complex64 (x fastest)."""
    n_u = [int(v) for v in n_u]
    d = len(n_u)
    grids = np.meshgrid(*[(np.arange(n) - n / 2) / (n / 2) for n in n_u], indexing="ij")
    img = np.zeros(tuple(n_u), dtype=np.complex128)
    inside = sum((g / 0.85) ** 2 for g in grids) <= 1.0
    img[inside] = 1.0
    for _ in range(n_ellipsoids):
        c = rng.uniform(-0.45, 0.45, size=d)
        ax = rng.uniform(0.08, 0.3, size=d)
        val = rng.uniform(-0.6, 0.9) + 1j * rng.uniform(-0.2, 0.2)
        m = sum(((g - cc) / a) ** 2 for g, cc, a in zip(grids, c, ax)) <= 1.0
        img[m] += val
    return img.reshape(-1, order="F").astype(np.complex64)


def add_noise(y, snr_db, rng):
    """Complex Gaussian noise at the given SNR (signal RMS / noise RMS)."""
    y = np.asarray(y)
    rms = float(np.sqrt(np.mean(np.abs(y) ** 2)))
    sigma = rms * 10 ** (-snr_db / 20.0) / math.sqrt(2)
    n = rng.normal(scale=sigma, size=y.shape) + 1j * rng.normal(scale=sigma, size=y.shape)
    return (y + n).astype(np.complex64)


def dirichlet_bins(nSpoke, nBin, alpha, rng):
    """Uneven split of nSpoke spokes into nBin contiguous-by-index bins with
    counts from Dirichlet(alpha) (every bin gets at least one spoke).
    Returns a list of spoke-index arrays."""
    p = rng.dirichlet([alpha] * nBin)
    counts = np.maximum(1, np.floor(p * nSpoke).astype(int))
    while counts.sum() > nSpoke:
        counts[np.argmax(counts)] -= 1
    while counts.sum() < nSpoke:
        counts[np.argmin(counts)] += 1
    perm = rng.permutation(nSpoke)
    out, o = [], 0
    for c in counts:
        out.append(np.sort(perm[o:o + c]))
        o += c
    return out


def _keys_weight(s):
    """Cubic convolution kernel of Keys (1981) with a = -1/2."""
    s = np.abs(s)
    return np.where(s <= 1, (1.5 * s - 2.5) * s * s + 1,
                    np.where(s < 2, ((-0.5 * s + 2.5) * s - 4) * s + 2, 0.0))


def cubic_interp(values, origin, step, points):
    """Cubic-convolution interpolation on a uniform grid (the method MATLAB's
    interpn 'cubic' uses on uniformly spaced grids; Keys 1981, a = -1/2,
    with Keys' boundary rows f(-1) = 3 f(0) - 3 f(1) + f(2) at each end).
    values: d-D array; origin, step: per-axis first coordinate and spacing;
    points: [d, P].  Points outside the grid's coordinate range give NaN,
    as interpn does.  Reproduces polynomials of degree <= 2 exactly."""
    v = np.asarray(values)
    d = v.ndim
    pts = np.asarray(points, dtype=np.float64).reshape(d, -1)
    for a in range(d):  # boundary rows on both ends of every axis
        n = v.shape[a]
        if n < 3:
            raise ValueError("cubic interpolation needs at least 3 points per axis")
        take = lambda i: np.take(v, [i], axis=a)
        lo = 3 * take(0) - 3 * take(1) + take(2)
        hi = 3 * take(n - 1) - 3 * take(n - 2) + take(n - 3)
        v = np.concatenate([lo, v, hi], axis=a)
    P = pts.shape[1]
    out = np.zeros(P, dtype=np.result_type(v.dtype, np.float64))
    inside = np.ones(P, bool)
    base, frac = [], []
    for a in range(d):
        n = np.asarray(values).shape[a]
        u = (pts[a] - origin[a]) / step[a]
        inside &= (u >= 0) & (u <= n - 1)
        i = np.clip(np.floor(u), 0, n - 2).astype(np.int64)
        base.append(i)
        frac.append(u - i)
    for off in np.ndindex(*([4] * d)):
        w = np.ones(P)
        idx = []
        for a in range(d):
            o = off[a] - 1                       # neighbour offsets -1 .. 2
            w = w * _keys_weight(frac[a] - o)
            idx.append(np.clip(base[a] + o + 1, 0, v.shape[a] - 1))   # +1: the boundary row
        out = out + w * v[tuple(idx)]
    out[~inside] = np.nan
    return out


def bmSimulateMriData(h, C, t, N_u, n_u, dK_u):
    """Simulated k-space data (src/mriRecon/function/bmSimulateMriData.m:8-86):
    for every coil, C .* h zero-filled to N_u, the centred DFT along every
    axis divided by N_u and dK_u, then interpolated at the trajectory points
    with cubic convolution on the grid k = (-N/2 .. N/2-1) dK; points outside
    the grid give 0.  h: image [prod(n_u)] (x fastest) or n_u-shaped; C:
    [prod(n_u), nCh] or None (one coil of ones); t: [imDim, nPt].  Returns
    [nPt, nCh] complex64."""
    n_u = [int(a) for a in np.ravel(n_u)]
    N_u = [int(a) for a in np.ravel(N_u)]
    dK_u = [float(a) for a in np.ravel(dK_u)]
    d = len(n_u)
    h = np.asarray(h).reshape(n_u, order="F").astype(np.complex128)
    C = np.ones((int(np.prod(n_u)), 1)) if C is None else np.asarray(C)
    C = C.reshape(int(np.prod(n_u)), -1)
    t = np.asarray(t, dtype=np.float64).reshape(d, -1)
    y = np.zeros((t.shape[1], C.shape[1]), np.complex64)
    origin = [-N / 2 * dk for N, dk in zip(N_u, dK_u)]
    for c in range(C.shape[1]):
        im = C[:, c].reshape(n_u, order="F") * h
        if N_u != n_u:  # centred zero fill (bmImZeroFill.m: rows N/2+1-n/2 .. N/2+n/2)
            z = np.zeros(N_u, np.complex128)
            z[tuple(slice(N // 2 - n // 2, N // 2 - n // 2 + n) for N, n in zip(N_u, n_u))] = im
            im = z
        for a in range(d):
            im = np.fft.fftshift(np.fft.fft(np.fft.ifftshift(im, axes=a), axis=a), axes=a) / N_u[a] / dK_u[a]
        re = cubic_interp(im.real, origin, dK_u, t)
        imag = cubic_interp(im.imag, origin, dK_u, t)
        v = re + 1j * imag
        v[np.isnan(v)] = 0
        y[:, c] = v.astype(np.complex64)
    return y
