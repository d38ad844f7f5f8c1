"""Direct n2u gridders and the gridded reconstruction built on them
(src/gridding123/m/bmGridder_n2u_leight.m:15-141, bmGridder_n2u.m:15-216,
src/fourier123/map_function/nonCartesian/bmMathilda.m:14-125).

The wrappers' preprocessing (grid coordinates, deletion of points outside
the grid box, neighbour weights and index jumps) runs on the host, as the
MATLAB wrappers do; the gridding itself runs in libmona's drop-in gridders
(bmGridder_n2u_leight{1,2,3}_mex, bmGridder_n2u_mex) and the image in
mona_mathilda.

Grids are returned as [prod(N_u), nCh] complex64: cells with x fastest (the
reference's column order), coils fastest within a cell, which is the
reference's [nCh, N_u] array read in memory order.
"""
from __future__ import annotations

import numpy as np

from ._lib import check, lib, ptr, stream_of
from .sparse import _fp, _ip, kernel_defaults


def _points(data_n, t, Dn):
    t = np.asarray(t)
    if t.ndim == 1:
        t = t[None, :]
    if t.shape[0] > 3 or (t.shape[0] > t.shape[1]):
        t = t.T                                  # [imDim, nPt]
    nPt = t.shape[1]
    y = np.asarray(data_n)
    if y.ndim == 1:
        y = y[:, None]
    if y.shape[0] != nPt:
        y = y.T                                  # [nPt, nCh]
    if y.shape[0] != nPt:
        raise ValueError("data_n must have one row (or column) per trajectory point")
    Dn = np.ravel(np.asarray(Dn))
    if Dn.size != nPt:
        raise ValueError("Dn must have one entry per trajectory point")
    return y, t, Dn


def _shift_and_box(t, Dn, N_u, dK_u, dt):
    """t / dK_u + fix(N_u / 2 + 1) per axis, Dn / prod(dK_u), and the mask of
    the points inside [1, N_u] on every axis (bmGridder_n2u.m:52-93)."""
    imDim = t.shape[0]
    t = t.astype(dt, copy=True)
    Dn = Dn.astype(dt, copy=True)
    keep = np.ones(t.shape[1], bool)
    for d in range(imDim):
        t[d] = t[d] / dt.type(dK_u[d])
        Dn = Dn / dt.type(dK_u[d])
        t[d] = t[d] + dt.type(np.fix(N_u[d] / 2 + 1))
        keep &= (t[d] >= 1) & (t[d] <= N_u[d])
    return t[:, keep], Dn[keep], keep


def bmGridder_n2u_leight(data_n, t, Dn, N_u, dK_u, kernelType=None, nWin=None, kernelParam=None):
    """Normalised n2u with Gaussian weights computed on the fly and a clipped
    box (bmGridder_n2u_leight.m:15-141): sigma = kernelParam[0] whatever the
    kernel type, as the reference's leight MEX does."""
    kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
    y, t, Dn = _points(data_n, t, Dn)
    N_u = [int(v) for v in np.ravel(N_u)]
    dK_u = [float(v) for v in np.ravel(dK_u)]
    imDim = t.shape[0]
    if len(N_u) != imDim or len(dK_u) != imDim:
        raise ValueError("N_u and dK_u need one entry per trajectory dimension")
    tg, Dn, keep = _shift_and_box(t, Dn, N_u, dK_u, np.dtype(np.float32))
    y = y[keep]
    nPt, nCh = y.shape
    ar = np.ascontiguousarray(y.real, dtype=np.float32)
    ai = np.ascontiguousarray(y.imag, dtype=np.float32)
    Nu = int(np.prod(N_u))
    br = np.empty((Nu, nCh), np.float32)
    bi = np.empty((Nu, nCh), np.float32)
    tx = [np.ascontiguousarray(tg[d]) for d in range(imDim)]
    Dn = np.ascontiguousarray(Dn, dtype=np.float32)
    kp1, kp2 = float(kernelParam[0]), float(kernelParam[1])
    if imDim == 1:
        lib.bmGridder_n2u_leight1_mex(_fp(ar), _fp(ai), _fp(br), _fp(bi), _fp(tx[0]), _fp(Dn), nCh, nPt, N_u[0],
                                      int(nWin), kp1, kp2)
    elif imDim == 2:
        lib.bmGridder_n2u_leight2_mex(_fp(ar), _fp(ai), _fp(br), _fp(bi), _fp(tx[0]), _fp(tx[1]), _fp(Dn), nCh, nPt,
                                      N_u[0], N_u[1], int(nWin), kp1, kp2)
    else:
        lib.bmGridder_n2u_leight3_mex(_fp(ar), _fp(ai), _fp(br), _fp(bi), _fp(tx[0]), _fp(tx[1]), _fp(tx[2]),
                                      _fp(Dn), nCh, nPt, N_u[0], N_u[1], N_u[2], int(nWin), kp1, kp2)
    return br + 1j * bi


def gridder_weights(t, Dn, N_u, dK_u, kernelType=None, nWin=None, kernelParam=None):
    """Neighbour weights w [nPt, nNb] (float32) and running index jumps
    [nPt * nNb] (int32, first absolute and 0-based) of bmGridder_n2u.m:95-196
    for the in-box points (returned mask); the window of an even nWin spans
    -nWin/2-1 .. nWin/2 as in the reference."""
    from scipy import special
    kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
    nWin = int(nWin)
    t = np.asarray(t, dtype=np.float64)
    Dn = np.ravel(np.asarray(Dn, dtype=np.float64))
    N_u = [int(v) for v in np.ravel(N_u)]
    dK_u = [float(v) for v in np.ravel(dK_u)]
    imDim = t.shape[0]
    tg, Dn, keep = _shift_and_box(t, Dn, N_u, dK_u, np.dtype(np.float64))
    nPt = tg.shape[1]
    if nWin % 2 == 0:
        r = np.arange(-nWin // 2 - 1, nWin // 2 + 1)
        shift = 0.0
    else:
        r = np.arange(-(nWin // 2), nWin // 2 + 1)
        shift = 0.5
    grids = np.meshgrid(*([r] * imDim), indexing="ij")
    c = np.stack([g.reshape(-1, order="F") for g in grids]).astype(np.float64)   # [imDim, nNb], x fastest
    nNb = c.shape[1]
    tf = np.floor(tg + shift)
    tr = tg - tf
    d = np.sqrt(((tr[:, None, :] - c[:, :, None]) ** 2).sum(axis=0))             # [nNb, nPt]
    kp = [float(v) for v in kernelParam]
    if kernelType == "gauss":
        sigma = kp[0]
        w = np.exp(-0.5 * (d / sigma) ** 2) / (sigma * np.sqrt(2 * np.pi))
    else:
        tau, alpha = kp[0], kp[1]
        w = special.i0(alpha * np.sqrt(np.maximum(1 - (d / tau) ** 2, 0))) / special.i0(alpha)
    w = w * Dn[None, :]
    n = tf[:, None, :] + c[:, :, None]
    lin = np.zeros((nNb, nPt), np.int64)
    stride = 1
    for a in range(imDim):
        lin += (np.mod(n[a] - 1, N_u[a]).astype(np.int64)) * stride
        stride *= N_u[a]
    lin = lin.T.reshape(-1)                                                        # point-major
    jumps = np.empty(lin.shape, np.int64)
    if lin.size:
        jumps[0] = lin[0]
        jumps[1:] = np.diff(lin)
    return (np.ascontiguousarray(w.T, dtype=np.float32), jumps.astype(np.int32), keep)


def bmGridder_n2u(data_n, t, Dn, N_u, dK_u, kernelType=None, nWin=None, kernelParam=None):
    """Normalised n2u with precomputed Kaiser-Bessel or Gaussian weights and
    wrapped neighbour indices (bmGridder_n2u.m:15-216)."""
    y, t, Dn = _points(data_n, t, Dn)
    w, jumps, keep = gridder_weights(t, Dn, N_u, dK_u, kernelType, nWin, kernelParam)
    y = y[keep]
    nPt, nCh = y.shape
    nNb = w.shape[1] if w.ndim == 2 else 0
    Nu = int(np.prod(np.ravel(N_u)))
    ar = np.ascontiguousarray(y.real, dtype=np.float32)
    ai = np.ascontiguousarray(y.imag, dtype=np.float32)
    br = np.empty((Nu, nCh), np.float32)
    bi = np.empty((Nu, nCh), np.float32)
    lib.bmGridder_n2u_mex(_fp(ar), _fp(ai), _fp(br), _fp(bi), _fp(w), _ip(jumps), nNb, nCh, nPt, Nu)
    return br + 1j * bi


def bmMathilda(y, t, v, Cmaps=None, N_u=None, n_u=None, dK_u=None, kernelType=None, nWin=None, kernelParam=None,
               leight_flag=True, plan=None, stream=None):
    """Gridded reconstruction with a direct gridder (bmMathilda.m:14-125):
    bmGridder_n2u_leight (leight_flag, the default) or bmGridder_n2u, then
    the centred inverse FFT times prod(N_u dK_u), the deapodisation K of the
    requested kernel, the crop to n_u and, with coil maps, the coil
    combination of bmCoilSense_pinv.  N_u and dK_u are required (the
    reference derives them from t when empty).  Cmaps: [prod(n_u), nCh]
    complex (host or cuda); plan: a Plan of the same geometry with its coil
    maps set, reused when given.  Returns a cuda tensor [prod(n_u)] (or
    [prod(n_u), nCh] without coil maps)."""
    import torch
    from .nufft import Plan
    if N_u is None or dK_u is None:
        raise ValueError("N_u and dK_u are required")
    N_u = [int(a) for a in np.ravel(N_u)]
    if any(a % 2 for a in N_u):
        raise ValueError("N_u must have all components even for the Fourier transform.")
    n_u = N_u if n_u is None else [int(a) for a in np.ravel(n_u)]
    kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
    gridder = bmGridder_n2u_leight if leight_flag else bmGridder_n2u
    grid = gridder(y, t, v, N_u, dK_u, kernelType, nWin, kernelParam)
    nCh = grid.shape[1]
    if plan is None:
        plan = Plan(N_u, n_u, dK_u, nCh, kernelType, nWin, kernelParam)
        cm = Cmaps if Cmaps is not None else np.ones((int(np.prod(n_u)), nCh), np.complex64)
        cm = torch.as_tensor(np.asarray(cm, dtype=np.complex64) if not torch.is_tensor(cm) else cm)
        plan.set_coils(cm.to("cuda", torch.complex64).contiguous())
    nu = int(np.prod(n_u))
    if plan.nCh != nCh or plan.nu != nu or plan.Nu != int(np.prod(N_u)):
        raise ValueError(f"plan geometry (nCh {plan.nCh}, prod(n_u) {plan.nu}, prod(N_u) {plan.Nu}) does not match "
                         f"the data (nCh {nCh}, prod(n_u) {nu}, prod(N_u) {int(np.prod(N_u))})")
    cm_dev = None
    if Cmaps is not None:
        cm_dev = torch.as_tensor(Cmaps).to("cuda", torch.complex64).contiguous()
        if cm_dev.numel() != nu * nCh:
            raise ValueError(f"Cmaps must have prod(n_u) * nCh = {nu * nCh} elements, got {cm_dev.numel()}")
    g = torch.from_numpy(np.ascontiguousarray(grid, dtype=np.complex64)).cuda()
    out = torch.empty((nu,) if cm_dev is not None else (nu, nCh), dtype=torch.complex64, device="cuda")
    check(lib.mona_mathilda(plan._h, ptr(g), ptr(cm_dev), ptr(out), stream_of(stream)))
    return out
