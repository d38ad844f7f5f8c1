"""Device-resident reconstruction plans: the reference's operator and solver
interface (bmShanna / bmNakatsha / bmSensa / bmSteva /
bmTevaMorphosia_chain / bmNasha) on HIP.

torch tensors are used only as device-memory plumbing (allocation and
streams); every operation runs in libmona.so.  Layouts (complex64):

* image ``x``:   [prod(n_u)]  x fastest (Fortran order of the n_u block);
* k-space ``y``: [nPt, nCh];
* coil maps ``C``: [prod(n_u), nCh].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib, check, ptr, stream_of
from .optim import n_iter
from .sparse import BmSparseMat, KERNELS, kernel_defaults


def _require_cuda(t, name):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise TypeError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def _arg(t, name, numel, real=False):
    """Refuses, before any library call, a tensor the kernels would misread
    through its raw pointer: a dtype other than complex64 (float32 for
    real=True, e.g. volume elements), an element count other than numel, a
    non-contiguous layout, or host memory.  numel None skips the count."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    want = torch.float32 if real else torch.complex64
    if t.dtype != want:
        raise ValueError(f"{name} must be {want}, got {t.dtype}")
    if numel is not None and t.numel() != int(numel):
        raise ValueError(f"{name} must have {int(numel)} elements, got {t.numel()}")
    _require_cuda(t, name)


class Plan:
    """Geometry + gridding kernel + coil maps + frames (Gu, Gut pairs)."""

    def __init__(self, N_u, n_u, dK_u, nCh, kernelType=None, nWin=None, kernelParam=None):
        kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
        self.N_u = [int(v) for v in np.ravel(N_u)]
        self.n_u = [int(v) for v in np.ravel(n_u)] if n_u is not None else list(self.N_u)
        self.dK_u = [float(v) for v in np.ravel(dK_u)]
        self.imDim = len(self.N_u)
        self.nCh = int(nCh)
        self.kernelType, self.nWin, self.kernelParam = kernelType, nWin, kernelParam
        Na = (C.c_int * self.imDim)(*self.N_u)
        na = (C.c_int * self.imDim)(*self.n_u)
        da = (C.c_double * self.imDim)(*self.dK_u)
        kp = (C.c_double * len(kernelParam))(*kernelParam)
        h = C.c_void_p()
        check(lib.mona_plan_create(self.imDim, Na, na, da, self.nCh, KERNELS[kernelType], nWin, kp,
                                   len(kernelParam), C.byref(h)))
        self._h = h
        self._frames = {}
        self._deform = {}
        self.nu = int(np.prod(self.n_u))
        self.Nu = int(np.prod(self.N_u))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None:
            lib.mona_plan_destroy(h)
            self._h = None

    def set_coils(self, Cmaps, stream=None):
        _arg(Cmaps, "C", self.nu * self.nCh)
        check(lib.mona_plan_set_coils(self._h, ptr(Cmaps), stream_of(stream)))

    def set_frame(self, f, Gu: BmSparseMat, Gut: BmSparseMat):
        check(lib.mona_plan_set_frame(self._h, int(f), Gu.handle(), Gut.handle()))
        self._frames[f] = (Gu, Gut)   # keep the handles alive

    def set_deform(self, f, Tu: BmSparseMat = None, Tut: BmSparseMat = None, set=0):
        """Deformation matrices of frame f for the Morphosia solvers: set 0
        holds Tu{f}, Tut{f} (Tu1, Tu1t), set 1 holds Tu2, Tu2t
        (bmTevaMorphosia_chain.m:6-13, bmTevaDuoMorphosia_chain.m:6-16);
        None = identity."""
        if (Tu is None) != (Tut is None):
            raise ValueError("Tu and Tut are given together")
        check(lib.mona_plan_set_deform(self._h, int(set), int(f), Tu.handle() if Tu is not None else None,
                                       Tut.handle() if Tut is not None else None))
        self._deform[(set, f)] = (Tu, Tut)   # keep the handles alive

    def _set_deforms(self, nFr, set, Tu, Tut):
        if Tu is None and Tut is None:
            return
        Tu = list(Tu) if Tu is not None else [None] * nFr
        Tut = list(Tut) if Tut is not None else [None] * nFr
        for f in range(nFr):
            self.set_deform(f, Tu[f], Tut[f], set=set)

    def nPt(self, f=0):
        if f not in self._frames:
            raise ValueError(f"frame {f} has no gridding matrices (set_frame)")
        return self._frames[f][0].l_size

    def _check_frames(self, nFr, y_list, ve_list):
        if len(y_list) != nFr or len(ve_list) != nFr:
            raise ValueError(f"{nFr} frames need {nFr} data and weight vectors, got {len(y_list)} and {len(ve_list)}")
        for f in range(nFr):
            _arg(y_list[f], f"y[{f}]", self.nPt(f) * self.nCh)
            _arg(ve_list[f], f"ve[{f}]", self.nPt(f), real=True)

    # -- execution options (no reference counterpart) ----------------------
    OPT_BATCH_FRAMES, OPT_GRAPH, OPT_FFT_TILE, OPT_FFT_LINES, OPT_MATRIX_FREE, OPT_STAGED_ADJOINT = 1, 2, 3, 4, 5, 6

    def set_options(self, batch_frames=None, graph=None, fft_tile=None, fft_lines=None, matrix_free=None,
                    staged_adjoint=None):
        """batch_frames: the temporal-TV solver applies its operators to all
        frames in one launch per gridding op; graph: its ADMM iterations are
        replayed from a HIP graph (constant delta/rho); fft_tile: plain FFT
        pass tile of the LDS kernel (2048, 4096, 8192; 0 = default);
        fft_lines: register-resident passes (N = 16 * N2) with 16 / 32 / 64
        lines per tile on strided axes, 0 = default, -1 = LDS kernel only;
        matrix_free: the forward gridding from the trajectory kept with Gu
        (default on) instead of the stored entries; staged_adjoint: the
        adjoint gridding sums groups of neighbouring grid cells from k-space
        rows staged in LDS (default on; the same results bit for bit)."""
        for key, v in ((self.OPT_BATCH_FRAMES, batch_frames), (self.OPT_GRAPH, graph),
                       (self.OPT_FFT_TILE, fft_tile), (self.OPT_FFT_LINES, fft_lines),
                       (self.OPT_MATRIX_FREE, matrix_free), (self.OPT_STAGED_ADJOINT, staged_adjoint)):
            if v is not None:
                check(lib.mona_plan_set_option(self._h, key, int(v)))

    def _watched(self, witnessInfo, nIter, x, run, names, params):
        """run() with the device witness snapshots on when witnessInfo has
        witness indices; then the reference's watch calls (bmSensa.m:85,
        :92): one 'loop' call per outer iteration with its snapshot, the
        full x kept at the witness indices, and a 'final' call.  names:
        the history columns run() returns (stored in witnessInfo.param with
        the run's parameters)."""
        if witnessInfo is None:
            return run()
        import torch
        watch = bool(witnessInfo.witness_ind) and nIter >= 1
        if watch:
            e = C.c_longlong()
            check(lib.mona_plan_witness_elems(self._h, C.byref(e)))
            sl = torch.empty((nIter, e.value), dtype=torch.complex64, device=x.device)
            ind = sorted(i for i in set(witnessInfo.witness_ind) if 1 <= i <= nIter)
            full = torch.empty((len(ind), x.numel()), dtype=torch.complex64, device=x.device) if ind else None
            ia = (C.c_int * max(len(ind), 1))(*ind)
            check(lib.mona_plan_set_witness(self._h, ptr(sl), int(nIter), ptr(full) if ind else None,
                                            ia if ind else None, len(ind)))
        try:
            out = run()
        finally:
            if watch:
                check(lib.mona_plan_set_witness(self._h, None, 0, None, None, 0))
        hist = out[-1] if isinstance(out, tuple) else out
        witnessInfo.param.update(params)
        witnessInfo.record(np.asarray(hist, dtype=np.float64).reshape(nIter, -1), names)
        if watch:
            torch.cuda.synchronize()
            slh = sl.cpu().numpy()
            fh = full.cpu().numpy() if ind else None
            for c in range(1, nIter + 1):
                xc = fh[ind.index(c)] if c in ind else None
                witnessInfo.watch(c, xc, self.n_u, "loop", image=slh[c - 1])
            witnessInfo.watch(nIter, x, self.n_u, "final")
        return out

    def autotune(self, f=0, stream=None, groups=False):
        """Times forward + adjoint of frame f for each FFT pass variant, keeps
        the fastest and returns the register kernel's lines per tile (or minus
        the LDS kernel's tile).  groups: then also the group sizes of the
        staged adjoint gridding; returns (lines, (stage KiB, rows))."""
        chosen = C.c_longlong()
        check(lib.mona_plan_autotune(self._h, int(f), stream_of(stream), C.byref(chosen)))
        if not groups:
            return chosen.value
        return chosen.value, self.autotune_groups(f, stream)

    def autotune_groups(self, f=0, stream=None):
        """Times the weighted adjoint gridding of frame f for several group
        sizes of its staged form, regroups every frame with the fastest and
        returns it as (stage KiB, rows per group); results are unchanged."""
        kb, rows = C.c_int(), C.c_int()
        check(lib.mona_plan_autotune_groups(self._h, int(f), stream_of(stream), C.byref(kb), C.byref(rows)))
        return kb.value, rows.value

    def K(self):
        import torch
        out = torch.empty(self.nu, dtype=torch.float32, device="cuda")
        check(lib.mona_plan_get_K(self._h, ptr(out), stream_of()))
        return out

    def KFC(self):
        import torch
        a = torch.empty((self.nu, self.nCh), dtype=torch.complex64, device="cuda")
        b = torch.empty_like(a)
        check(lib.mona_plan_get_KFC(self._h, ptr(a), ptr(b), stream_of()))
        return a, b

    # -- operators ---------------------------------------------------------
    def bmShanna(self, x, f=0, out=None, stream=None):
        """y = Gu * cfft(zerofill(x .* KFC))  (bmShanna_MATLAB.m:6-56)."""
        import torch
        _arg(x, "x", self.nu)
        if out is None:
            out = torch.empty((self.nPt(f), self.nCh), dtype=torch.complex64, device=x.device)
        _arg(out, "out", self.nPt(f) * self.nCh)
        check(lib.mona_shanna(self._h, int(f), ptr(x), ptr(out), stream_of(stream)))
        return out

    def bmNakatsha(self, y, f=0, C_flag=True, out=None, stream=None):
        """x = sum_c KFC_conj .* crop(cifft(Gut * y))  (bmNakatsha_MATLAB.m:6-60)."""
        import torch
        _arg(y, "y", self.nPt(f) * self.nCh)
        shape = (self.nu,) if C_flag else (self.nu, self.nCh)
        if out is None:
            out = torch.empty(shape, dtype=torch.complex64, device=y.device)
        _arg(out, "out", int(np.prod(shape)))
        check(lib.mona_nakatsha(self._h, int(f), ptr(y), ptr(out), 1 if C_flag else 0, stream_of(stream)))
        return out

    # -- solvers -----------------------------------------------------------
    def bmSensa(self, x, y, ve, nCGD, ve_max=None, nIter=1, f=0, stream=None, witnessInfo=None):
        """CG-SENSE in place on x (bmSensa.m:6-96); returns the per-outer-
        iteration residuals <y - Mx, HY (y - Mx)>.  witnessInfo: a
        bmWitnessInfo that receives the residuals and the witness images."""
        nIter = n_iter(nIter)
        _arg(x, "x", self.nu)
        _arg(y, "y", self.nPt(f) * self.nCh)
        _arg(ve, "ve", self.nPt(f), real=True)

        def run():
            res = (C.c_double * max(nIter, 1))()
            check(lib.mona_sensa(self._h, int(f), ptr(x), ptr(y), C.cast(ptr(ve), C.POINTER(C.c_float)),
                                 float(ve_max) if ve_max else 0.0, int(nIter), int(nCGD), res,
                                 stream_of(stream)))
            return list(res)[:nIter]
        return self._watched(witnessInfo, nIter, x, run, ["residu"],
                             {"recon_name": "sensa", "dK_u": self.dK_u, "N_u": self.N_u, "n_u": self.n_u,
                              "nCGD": int(nCGD), "ve_max": ve_max})

    def bmSensa_batch(self, x, y_list, ve_list, nCGD, ve_max=None, nIter=1, stream=None):
        """bmSensa on frames 0 .. nFr-1 at once (batched operators, per-frame
        CG), nFr = x.shape[0] <= the plan's frames: x [nFr, nu] in place;
        y_list / ve_list per frame; returns the residuals [nFr][nIter].  Each
        frame's result equals bmSensa on that frame alone."""
        nIter = n_iter(nIter)
        import torch
        nFr = x.shape[0] if isinstance(x, torch.Tensor) and x.dim() > 0 else 0
        if nFr < 1 or any(f not in self._frames for f in range(nFr)):
            raise ValueError(f"x must be [nFr, nu] with nFr frames set (got {tuple(getattr(x, 'shape', ()))})")
        _arg(x, "x", nFr * self.nu)
        if not isinstance(y_list, torch.Tensor):
            self._check_frames(nFr, y_list, ve_list)
        ycat = y_list if isinstance(y_list, torch.Tensor) else torch.cat([v.reshape(-1) for v in y_list])
        vcat = ve_list if isinstance(ve_list, torch.Tensor) else torch.cat([v.reshape(-1) for v in ve_list])
        _arg(ycat, "y", sum(self.nPt(f) for f in range(nFr)) * self.nCh)
        _arg(vcat, "ve", sum(self.nPt(f) for f in range(nFr)), real=True)
        yo = np.cumsum([0] + [self.nPt(f) * self.nCh for f in range(nFr)])[:-1]
        vo = np.cumsum([0] + [self.nPt(f) for f in range(nFr)])[:-1]
        yoff = (C.c_longlong * nFr)(*[int(v) for v in yo])
        voff = (C.c_longlong * nFr)(*[int(v) for v in vo])
        vm = None
        if ve_max is not None:
            vm = (C.c_float * nFr)(*[float(v) for v in np.broadcast_to(np.asarray(ve_max, dtype=np.float64), (nFr,))])
        res = (C.c_double * (max(nIter, 1) * nFr))()
        ycat, vcat = ycat.contiguous(), vcat.contiguous()   # held until the call returns
        check(lib.mona_sensa_batch(self._h, int(nFr), ptr(x), ptr(ycat), yoff, C.cast(ptr(vcat), C.POINTER(C.c_float)),
                                   voff, vm, int(nIter), int(nCGD), res, stream_of(stream)))
        return np.array(list(res)[:nIter * nFr]).reshape(nFr, nIter)

    def device_bytes(self):
        """Device memory held by the plan (buffers and frame matrices)."""
        b = C.c_longlong()
        check(lib.mona_plan_memory(self._h, C.byref(b)))
        return b.value

    def sensa_refined(self, x, y, ve, nInner, nOuter, tol=1e-5, bf16=True, ve_max=None, f=0, stream=None):
        """Mixed-precision CG-SENSE (no reference counterpart): CG on the
        normal equations with bf16-storage (or float) operator products;
        every nInner steps, and as soon as the recursive estimate says
        converged, the residual is replaced by the true float one; stops
        when that, relative to the residual of x = 0, is < tol, or after
        nOuter periodic replacements.  x in place (any start; nOuter = 0 only
        evaluates the relative residual of x); returns the relative residual
        at the start and after every replacement."""
        _arg(x, "x", self.nu)
        _arg(y, "y", self.nPt(f) * self.nCh)
        _arg(ve, "ve", self.nPt(f), real=True)
        rel = (C.c_double * (2 * nOuter + 1))()
        n_outer = C.c_int()
        n_steps = C.c_int()
        check(lib.mona_sensa_ir(self._h, int(f), ptr(x), ptr(y), C.cast(ptr(ve), C.POINTER(C.c_float)),
                                float(ve_max) if ve_max else 0.0, int(nInner), int(nOuter), float(tol),
                                1 if bf16 else 0, rel, C.byref(n_outer), C.byref(n_steps), stream_of(stream)))
        return RefineHistory(list(rel)[:n_outer.value + 1], n_steps.value)

    def bmShanna_bf16(self, x, f=0, stream=None):
        """bmShanna with bf16 storage of the grid-side arrays and k-space."""
        import torch
        _arg(x, "x", self.nu)
        out = torch.empty((self.nPt(f), self.nCh), dtype=torch.complex64, device=x.device)
        check(lib.mona_shanna_bf16(self._h, int(f), ptr(x), ptr(out), stream_of(stream)))
        return out

    def bmNakatsha_bf16(self, y, f=0, stream=None):
        """bmNakatsha (coil-combined) with bf16 storage of the grid-side
        arrays and k-space."""
        import torch
        _arg(y, "y", self.nPt(f) * self.nCh)
        out = torch.empty((self.nu,), dtype=torch.complex64, device=y.device)
        check(lib.mona_nakatsha_bf16(self._h, int(f), ptr(y), ptr(out), stream_of(stream)))
        return out

    def bmSteva(self, x, z, u, y, ve, delta, rho, nCGD, ve_max=None, nIter=1, f=0, stream=None, witnessInfo=None):
        """Spatial-TV ADMM in place on x, z, u (bmSteva.m:6-138).  z, u are
        [imDim, nu] complex64 (given) or None (z = grad x, u = 0)."""
        nIter = n_iter(nIter)
        import torch
        _arg(x, "x", self.nu)
        _arg(y, "y", self.nPt(f) * self.nCh)
        _arg(ve, "ve", self.nPt(f), real=True)
        given = z is not None
        if z is None:
            z = torch.zeros((self.imDim, self.nu), dtype=torch.complex64, device=x.device)
            u = torch.zeros_like(z)
        _arg(z, "z", self.imDim * self.nu)
        _arg(u, "u", self.imDim * self.nu)
        d = (C.c_float * 2)(*_sched(delta))
        r = (C.c_float * 2)(*_sched(rho))

        def run():
            hist = (C.c_double * (2 * max(nIter, 1)))()
            check(lib.mona_steva(self._h, int(f), ptr(x), ptr(z), ptr(u), 1 if given else 0, ptr(y),
                                 C.cast(ptr(ve), C.POINTER(C.c_float)), float(ve_max) if ve_max else 0.0,
                                 d, r, int(nIter), int(nCGD), hist, stream_of(stream)))
            return z, u, np.array(list(hist)[:2 * nIter]).reshape(nIter, 2)
        return self._watched(witnessInfo, nIter, x, run, ["residu", "total_variation"],
                             {"recon_name": "steva", "dK_u": self.dK_u, "N_u": self.N_u, "n_u": self.n_u,
                              "delta": delta, "rho": rho, "nCGD": int(nCGD), "ve_max": ve_max})

    def bmTevaMorphosia_chain(self, x, z, u, y_list, ve_list, delta, rho, nCGD, ve_max, nIter=1,
                              stream=None, Tu=None, Tut=None, witnessInfo=None):
        """Temporal-TV ADMM over all frames (bmTevaMorphosia_chain.m:6-294),
        regul_mode 'normal'.  x, z, u are [nFr, nu] complex64 (z, u None:
        z = F x, u = 0).  Tu, Tut: lists of deformation matrices per frame
        (entries may be None), F x_f = Tu_f x_f - x_{f-1}; when not given,
        the ones set by set_deform (none: identity, Tu = [])."""
        nIter = n_iter(nIter)
        import torch
        nFr = x.shape[0]
        _arg(x, "x", nFr * self.nu)
        self._check_frames(nFr, y_list, ve_list)
        self._set_deforms(nFr, 0, Tu, Tut)
        given = z is not None
        if z is None:
            z = torch.zeros_like(x)
            u = torch.zeros_like(x)
        _arg(z, "z", nFr * self.nu)
        _arg(u, "u", nFr * self.nu)
        ycat = torch.cat([yy.reshape(-1) for yy in y_list])
        vcat = torch.cat([vv.reshape(-1) for vv in ve_list])
        yo = np.cumsum([0] + [yy.numel() for yy in y_list])[:-1]
        vo = np.cumsum([0] + [vv.numel() for vv in ve_list])[:-1]
        yoff = (C.c_longlong * nFr)(*[int(v) for v in yo])
        voff = (C.c_longlong * nFr)(*[int(v) for v in vo])
        d = (C.c_float * 2)(*_sched(delta))
        r = (C.c_float * 2)(*_sched(rho))

        def run():
            hist = (C.c_double * (3 * max(nIter, 1)))()
            check(lib.mona_teva_chain(self._h, ptr(x), ptr(z), ptr(u), 1 if given else 0, ptr(ycat), yoff,
                                      C.cast(ptr(vcat), C.POINTER(C.c_float)), voff, float(ve_max), d, r,
                                      int(nIter), int(nCGD), hist, stream_of(stream)))
            return z, u, np.array(list(hist)[:3 * nIter]).reshape(nIter, 3)
        return self._watched(witnessInfo, nIter, x, run, ["objective", "dafi", "regul"],
                             {"recon_name": "tevaMorphosia_chain", "dK_u": self.dK_u, "N_u": self.N_u,
                              "n_u": self.n_u, "delta": delta, "rho": rho, "nCGD": int(nCGD), "ve_max": ve_max})

    # -- the other Morphosia solvers (mona_morphosia) ------------------------
    MORPH_TEVA_CHAIN, MORPH_TEVA_DUO_CHAIN, MORPH_TEVA_SHEET = 0, 1, 2
    MORPH_SENSITIVA_CHAIN, MORPH_SENSITIVA_SHEET, MORPH_SENSITIVA_DUO_CHAIN = 3, 4, 5
    MORPH_NAMES = ("tevaMorphosia_chain", "tevaDuoMorphosia_chain", "tevaMorphosia_sheet",
                   "sensitivaMorphosia_chain", "sensitivaMorphosia_sheet", "sensitivaDuoMorphosia_chain")

    def _morph(self, solver, nFr1, x, zs, us, y_list, ve_list, delta, rho, nCGD, ve_max, nIter, stream, nT,
               witnessInfo=None):
        nIter = n_iter(nIter)
        import torch
        nFr = x.shape[0]
        _arg(x, "x", nFr * self.nu)
        self._check_frames(nFr, y_list, ve_list)
        if nFr1 is None:
            nFr1 = nFr
        l1 = rho is not None
        z = u = None
        given = False
        if l1:
            given = zs is not None and all(v is not None for v in zs)
            z = torch.stack(list(zs)).contiguous() if given else torch.zeros((nT,) + tuple(x.shape), dtype=x.dtype,
                                                                            device=x.device)
            u = torch.stack(list(us)).contiguous() if given else torch.zeros_like(z)
            _arg(z, "z", nT * nFr * self.nu)
            _arg(u, "u", nT * nFr * self.nu)
        ycat = torch.cat([yy.reshape(-1) for yy in y_list])
        vcat = torch.cat([vv.reshape(-1) for vv in ve_list])
        yo = np.cumsum([0] + [yy.numel() for yy in y_list])[:-1]
        vo = np.cumsum([0] + [vv.numel() for vv in ve_list])[:-1]
        yoff = (C.c_longlong * nFr)(*[int(v) for v in yo])
        voff = (C.c_longlong * nFr)(*[int(v) for v in vo])
        d = (C.c_float * 2)(*_sched(delta))
        r = (C.c_float * 2)(*_sched(rho)) if l1 else None

        def run():
            hist = (C.c_double * (3 * max(nIter, 1)))()
            check(lib.mona_morphosia(self._h, int(solver), int(nFr1), ptr(x), ptr(z), ptr(u), 1 if given else 0,
                                     ptr(ycat), yoff, C.cast(ptr(vcat), C.POINTER(C.c_float)), voff, float(ve_max),
                                     d, r, int(nIter), int(nCGD), hist, stream_of(stream)))
            return np.array(list(hist)[:3 * nIter]).reshape(nIter, 3)
        h = self._watched(witnessInfo, nIter, x, run, ["objective", "dafi", "regul"],
                          {"recon_name": self.MORPH_NAMES[solver], "dK_u": self.dK_u, "N_u": self.N_u,
                           "n_u": self.n_u, "delta": delta, "rho": rho, "nCGD": int(nCGD), "ve_max": ve_max})
        if not l1:
            return h
        return [z[k] for k in range(nT)], [u[k] for k in range(nT)], h

    def bmTevaDuoMorphosia_chain(self, x, z1, z2, u1, u2, y_list, ve_list, delta, rho, nCGD, ve_max, nIter=1,
                                 Tu1=None, Tu1t=None, Tu2=None, Tu2t=None, stream=None, witnessInfo=None):
        """bmTevaDuoMorphosia_chain.m:6-242: two half-weighted temporal
        differences, 0.5 (Tu1_i x_i - x_{i-1}) and 0.5 (Tu2_i x_i - x_{i+1}).
        x [nFr, nu] complex64 is updated in place; z1..u2 None start from
        z = F x, u = 0.  Returns (z1, z2, u1, u2, hist)."""
        nFr = x.shape[0]
        self._set_deforms(nFr, 0, Tu1, Tu1t)
        self._set_deforms(nFr, 1, Tu2, Tu2t)
        zs, us, h = self._morph(self.MORPH_TEVA_DUO_CHAIN, None, x, None if z1 is None else (z1, z2),
                                None if u1 is None else (u1, u2), y_list, ve_list, delta, rho, nCGD, ve_max, nIter,
                                stream, 2, witnessInfo)
        return zs[0], zs[1], us[0], us[1], h

    def bmTevaMorphosia_sheet(self, x, z1, z2, u1, u2, y_list, ve_list, nFr_1, delta, rho, nCGD, ve_max, nIter=1,
                              Tu1=None, Tu1t=None, Tu2=None, Tu2t=None, stream=None, witnessInfo=None):
        """bmTevaMorphosia_sheet.m:16-244: frames on an nFr_1 x nFr_2 cell
        array, frame k = i + nFr_1*j (the y(:) order of x, y_list, ve_list
        and the deformation lists); differences along i (Tu1) and j (Tu2).
        Returns (z1, z2, u1, u2, hist)."""
        nFr = x.shape[0]
        self._set_deforms(nFr, 0, Tu1, Tu1t)
        self._set_deforms(nFr, 1, Tu2, Tu2t)
        zs, us, h = self._morph(self.MORPH_TEVA_SHEET, nFr_1, x, None if z1 is None else (z1, z2),
                                None if u1 is None else (u1, u2), y_list, ve_list, delta, rho, nCGD, ve_max, nIter,
                                stream, 2, witnessInfo)
        return zs[0], zs[1], us[0], us[1], h

    def bmSensitivaMorphosia_chain(self, x, y_list, ve_list, delta, nCGD, ve_max, nIter=1, Tu=None, Tut=None,
                                   stream=None, witnessInfo=None):
        """bmSensitivaMorphosia_chain.m:6-184: quadratic temporal penalty
        delta*HZ*||Tu_i x_i - x_{i-1}||^2; x updated in place; returns hist."""
        self._set_deforms(x.shape[0], 0, Tu, Tut)
        return self._morph(self.MORPH_SENSITIVA_CHAIN, None, x, None, None, y_list, ve_list, delta, None, nCGD,
                           ve_max, nIter, stream, 1, witnessInfo)

    def bmSensitivaMorphosia_sheet(self, x, y_list, ve_list, nFr_1, delta, nCGD, ve_max, nIter=1, Tu1=None,
                                   Tu1t=None, Tu2=None, Tu2t=None, stream=None, witnessInfo=None):
        """bmSensitivaMorphosia_sheet.m:6-187: the quadratic penalty along
        both cell-array axes (frame k = i + nFr_1*j); returns hist."""
        nFr = x.shape[0]
        self._set_deforms(nFr, 0, Tu1, Tu1t)
        self._set_deforms(nFr, 1, Tu2, Tu2t)
        return self._morph(self.MORPH_SENSITIVA_SHEET, nFr_1, x, None, None, y_list, ve_list, delta, None, nCGD,
                           ve_max, nIter, stream, 2, witnessInfo)

    def bmSensitivaDuoMorphosia_chain(self, x, y_list, ve_list, delta, nCGD, ve_max, nIter=1, Tu1=None, Tu1t=None,
                                      Tu2=None, Tu2t=None, stream=None, witnessInfo=None):
        """bmSensitivaDuoMorphosia_chain.m:6-108: the quadratic penalty on
        0.5 (Tu1_i x_i - x_{i-1}) and 0.5 (Tu2_i x_i - x_{i+1}); returns hist."""
        nFr = x.shape[0]
        self._set_deforms(nFr, 0, Tu1, Tu1t)
        self._set_deforms(nFr, 1, Tu2, Tu2t)
        return self._morph(self.MORPH_SENSITIVA_DUO_CHAIN, None, x, None, None, y_list, ve_list, delta, None, nCGD,
                           ve_max, nIter, stream, 2, witnessInfo)

    def bmSensaMorphosia_chain(self, x, y_list, ve_list, nCGD, ve_max, nIter=1, Tu=None, Tut=None, stream=None,
                               witnessInfo=None):
        """bmSensaMorphosia_chain.m:6-104: one image x [nu] (complex64,
        updated in place) seen by every frame through Tu_f; returns the
        residual R after each iteration."""
        nIter = n_iter(nIter)
        import torch
        nFr = len(y_list)
        _arg(x, "x", self.nu)
        self._check_frames(nFr, y_list, ve_list)
        self._set_deforms(nFr, 0, Tu, Tut)
        ycat = torch.cat([yy.reshape(-1) for yy in y_list])
        vcat = torch.cat([vv.reshape(-1) for vv in ve_list])
        yoff = (C.c_longlong * nFr)(*[int(v) for v in np.cumsum([0] + [yy.numel() for yy in y_list])[:-1]])
        voff = (C.c_longlong * nFr)(*[int(v) for v in np.cumsum([0] + [vv.numel() for vv in ve_list])[:-1]])
        def run():
            hist = (C.c_double * max(nIter, 1))()
            check(lib.mona_sensa_morphosia(self._h, ptr(x), ptr(ycat), yoff,
                                           C.cast(ptr(vcat), C.POINTER(C.c_float)), voff, float(ve_max), int(nIter),
                                           int(nCGD), hist, stream_of(stream)))
            return np.array(list(hist)[:nIter])
        return self._watched(witnessInfo, nIter, x, run, ["residu"],
                             {"recon_name": "sensaMorphosia_chain", "dK_u": self.dK_u, "N_u": self.N_u,
                              "n_u": self.n_u, "nCGD": int(nCGD), "ve_max": ve_max})

    def bmNasha(self, y, Gn: BmSparseMat, Cmaps=None, stream=None):
        """Gridded recon (bmNasha.m:6-91)."""
        import torch
        _arg(y, "y", Gn.r_size * self.nCh)
        if Cmaps is not None:
            _arg(Cmaps, "C", self.nu * self.nCh)
        shape = (self.nu,) if Cmaps is not None else (self.nu, self.nCh)
        out = torch.empty(shape, dtype=torch.complex64, device=y.device)
        check(lib.mona_nasha(self._h, Gn.handle(), ptr(y), ptr(Cmaps), ptr(out), stream_of(stream)))
        return out


class RefineHistory(list):
    """Relative residuals of a refined solve (start, then every residual
    replacement); .steps = CG steps taken."""

    def __init__(self, values, steps):
        super().__init__(values)
        self.steps = steps


def _sched(v):
    v = np.abs(np.ravel(np.asarray(v, dtype=np.float64)))
    if v.size == 1:
        return [float(v[0]), float(v[0])]
    return [float(v[0]), float(v[1])]


def cfft(data, N, inverse=False, stream=None):
    """Centred n-D DFT in place on a [prod(N), nBatch] complex64 device
    tensor (fftshift(fft(ifftshift(.))) per axis; inverse scaled 1/N)."""
    N = [int(v) for v in np.ravel(N)]
    if not hasattr(data, "numel") or data.numel() == 0 or data.numel() % int(np.prod(N)):
        raise ValueError(f"data must hold a whole number of {int(np.prod(N))}-element grids")
    _arg(data, "data", None)
    nb = data.numel() // int(np.prod(N))
    Na = (C.c_int * len(N))(*N)
    check(lib.mona_cfft(len(N), Na, nb, ptr(data), 1 if inverse else 0, stream_of(stream)))
    return data


def timing(enable=None, kind=None):
    """Kernel timing: timing(True) starts, timing(kind=k) -> (launches, ms)."""
    if enable is not None:
        check(lib.mona_timing_enable(1 if enable else 0))
    if kind is not None:
        n = C.c_longlong()
        ms = C.c_double()
        check(lib.mona_timing_read(int(kind), C.byref(n), C.byref(ms)))
        return n.value, ms.value
    return None
