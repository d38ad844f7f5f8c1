"""Host-side mirror of the reference's ``bmSparseMat`` class and of the
functions that build and apply it.

* :class:`BmSparseMat` holds the cpp_prepared one-block encoding
  (src/sparseMat/m/bmSparseMat.m:15-47) and runs the reference's structural
  ``check()`` rules (bmSparseMat.m:192-393) before it is handed to the GPU.
* :func:`bmTraj2SparseMat` builds Gn / Gu / Gut from a trajectory
  (src/gridding123/m/bmTraj2SparseMat.m:12-314).
* :func:`bmSparseMat_vec` is the product dispatcher
  (src/sparseMat/m/bmSparseMat_vec.m:11-242) over the drop-in entry points.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib, check

KERNELS = {"gauss": 0, "kaiser": 1}


def kernel_defaults(kernelType=None, nWin=None, kernelParam=None):
    """bmVarargin_kernelType_nWin_kernelParam (src/varargin/
    bmVarargin_kernelType_nWin_kernelParam.m:10-45)."""
    if not kernelType:
        kernelType = "gauss"
    if nWin is None:
        nWin = 3
    if kernelParam is None:
        kernelParam = [0.61, 10.0] if kernelType == "gauss" else [1.95, 10.0, 10.0]
    kernelParam = [float(np.float32(v)) for v in np.ravel(kernelParam)]
    if kernelType == "gauss" and len(kernelParam) == 3:
        raise ValueError("Wrong list of gridding kernel parameters. ")
    if kernelType == "kaiser" and len(kernelParam) == 2:
        raise ValueError("Wrong list of gridding kernel parameters. ")
    return kernelType, int(np.float32(nWin)), kernelParam


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class BmSparseMat:
    """cpp_prepared, one-block bmSparseMat (fields of bmSparseMat.m:15-47)."""

    def __init__(self, r_size, l_size, r_jump, r_nJump, m_val, l_jump=None, N_u=None, d_u=None,
                 kernel_type="void", nWin=None, kernelParam=None):
        self.r_size = int(r_size)
        self.l_size = int(l_size)
        self.r_jump = _i32(r_jump)
        self.r_nJump = _i32(r_nJump)
        self.m_val = _f32(m_val)
        self.l_jump = None if l_jump is None else _i32(l_jump)
        self.l_nJump = int(self.r_nJump.size)
        self.N_u = N_u
        self.d_u = d_u
        self.kernel_type = kernel_type
        self.nWin = nWin
        self.kernelParam = kernelParam
        self.block_type = "one_block"
        self.type = "l_squeezed_cpp_prepared" if self.l_jump is not None else "cpp_prepared"
        self._handle = None

    # bmSparseMat.check (src/sparseMat/m/bmSparseMat.m:192-393), the rules
    # that apply to a cpp_prepared matrix
    def check(self):
        fails = []
        if self.l_size < 2 or self.r_size < 2 or self.l_nJump < 2:
            fails.append("2.5")
        if np.any(self.r_nJump < 0):
            fails.append("2.6")
        s = int(self.r_nJump.astype(np.int64).sum())
        if s == 0:
            fails.append("3.1")
        elif s != self.m_val.size:
            fails.append("3.2")
        if self.r_nJump.size != self.l_nJump:
            fails.append("3.3")
        if self.r_jump.size == 0:
            fails.append("8.2")
        if self.l_jump is None and s != self.r_jump.size:
            fails.append("9.2")
        if self.l_jump is not None and self.l_jump.size != self.l_nJump:
            fails.append("10.2")
        if not (self.r_size < 2**31 - 1 and self.l_size < 2**31 - 1):
            fails.append("11.1")
        if fails:
            raise ValueError("The sparse matrix did NOT pass the check ! (rules %s)" % ",".join(fails))
        return True

    @property
    def nnz(self):
        return int(self.m_val.size)

    def handle(self):
        """Device-resident copy (decoded once, reused by every product)."""
        if self._handle is None:
            self.check()
            h = C.c_void_p()
            check(lib.mona_sparse_create(self.r_size, _ip(self.r_jump), _ip(self.r_nJump),
                                         _fp(self.m_val), self.l_size, _ip(self.l_jump),
                                         self.l_nJump, self.nnz, C.byref(h)))
            self._handle = h
        return self._handle

    @classmethod
    def _from_handle(cls, h, prov):
        nnz = C.c_longlong()
        lnj = C.c_int()
        rs = C.c_int()
        ls = C.c_int()
        sq = C.c_int()
        check(lib.mona_sparse_info(h, C.byref(nnz), C.byref(lnj), C.byref(rs), C.byref(ls), C.byref(sq)))
        r_jump = np.empty(nnz.value, np.int32)
        r_nJump = np.empty(lnj.value, np.int32)
        m_val = np.empty(nnz.value, np.float32)
        l_jump = np.empty(lnj.value, np.int32) if sq.value else None
        check(lib.mona_sparse_export(h, _ip(r_jump), _ip(r_nJump), _fp(m_val), _ip(l_jump)))
        M = cls(rs.value, ls.value, r_jump, r_nJump, m_val, l_jump, **prov)
        M._handle = h
        return M

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None:
            try:
                lib.mona_sparse_destroy(h)
            except Exception:
                pass
            self._handle = None

    def decode(self):
        """(col, rowptr, rows) with absolute 0-based indices."""
        col = np.cumsum(self.r_jump.astype(np.int64))
        rowptr = np.zeros(self.l_nJump + 1, np.int64)
        rowptr[1:] = np.cumsum(self.r_nJump.astype(np.int64))
        rows = (np.cumsum(self.l_jump.astype(np.int64)) if self.l_jump is not None
                else np.arange(self.l_nJump, dtype=np.int64))
        return col, rowptr, rows


def bmTraj2SparseMat(t, ve, N_u, dK_u, kernelType=None, nWin=None, kernelParam=None,
                     which=("Gu", "Gut"), device=False):
    """Gridding matrices for trajectory t ([imDim, nPt] or [nPt, imDim] with
    imDim <= 3 rows), volume elements ve [nPt].  Returns a dict with the
    requested matrices among 'Gn', 'Gu', 'Gut'.  device=True builds them on
    the GPU (same integer arrays; m_val may differ in the last bit)."""
    kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
    t = np.asarray(t, dtype=np.float64)
    if t.ndim == 1:
        t = t[None, :]
    if t.shape[0] <= 3 and t.shape[0] < t.shape[1]:
        t = t.T
    t = np.ascontiguousarray(t)          # [nPt, imDim]
    nPt, imDim = t.shape
    N_u = np.ascontiguousarray(np.ravel(N_u), dtype=np.int32)
    dK_u = np.ascontiguousarray(np.ravel(dK_u), dtype=np.float64)
    kp = np.ascontiguousarray(kernelParam, dtype=np.float64)
    ve_a = None if ve is None else np.ascontiguousarray(np.ravel(ve), dtype=np.float64)
    w = (1 if "Gn" in which else 0) | (2 if ("Gu" in which or "Gut" in which) else 0)
    hn, hu, hut = C.c_void_p(), C.c_void_p(), C.c_void_p()
    dp = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))
    if device:
        from ._lib import stream_of
        check(lib.mona_traj2sparsemat_device(imDim, nPt, dp(t), dp(ve_a), _ip(N_u), dp(dK_u),
                                             KERNELS[kernelType], nWin, dp(kp), len(kp), w,
                                             C.byref(hn), C.byref(hu), C.byref(hut), stream_of()))
    else:
        check(lib.mona_traj2sparsemat(imDim, nPt, dp(t), dp(ve_a), _ip(N_u), dp(dK_u),
                                      KERNELS[kernelType], nWin, dp(kp), len(kp), w,
                                      C.byref(hn), C.byref(hu), C.byref(hut)))
    prov = dict(N_u=[int(v) for v in N_u], d_u=[float(np.float32(v)) for v in dK_u],
                kernel_type=kernelType, nWin=nWin, kernelParam=kernelParam)
    out = {}
    if w & 1:
        out["Gn"] = BmSparseMat._from_handle(hn, prov)
    if w & 2:
        out["Gu"] = BmSparseMat._from_handle(hu, prov)
        out["Gut"] = BmSparseMat._from_handle(hut, prov)
    return out


def bmImDeformField2SparseMat(v, N_u, Dn=None, torus_flag=False, which=("Gu", "Gut")):
    """Interpolation matrices of deformation field v (bmImDeformField2SparseMat.m:14-147):
    v reshapes to [imDim, prod(N_u)] (or is [prod(N_u), imDim]); pixel p
    moves to its 1-based grid position plus v[:, p].  A list of fields gives
    a list of dicts, as the reference's cell input does.  'Gn' needs Dn."""
    if isinstance(v, (list, tuple)):
        Dns = Dn if isinstance(Dn, (list, tuple)) else [Dn] * len(v)
        return [bmImDeformField2SparseMat(vi, N_u, di, torus_flag, which) for vi, di in zip(v, Dns)]
    N_u = np.ascontiguousarray(np.ravel(N_u), dtype=np.int32)
    imDim = N_u.size
    nPt = int(np.prod(N_u.astype(np.int64)))
    v = np.asarray(v, dtype=np.float64)
    if v.size != imDim * nPt:
        raise ValueError("the deformation field must have as many vectors as the number of pixels in the image")
    if v.ndim == 2 and v.shape == (nPt, imDim) and imDim != nPt:
        va = np.ascontiguousarray(v)
    else:
        va = np.ascontiguousarray(v.reshape(-1, order="F").reshape(nPt, imDim))  # [imDim, nPt] column-major
    w = (1 if "Gn" in which else 0) | (2 if ("Gu" in which or "Gut" in which) else 0)
    if (w & 1) and Dn is None:
        raise ValueError("Gn needs Dn")
    dn = None if Dn is None else np.ascontiguousarray(np.ravel(Dn), dtype=np.float64)
    if dn is not None and dn.size != nPt:
        raise ValueError("Dn must have prod(N_u) entries")
    dp = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))
    hn, hu, hut = C.c_void_p(), C.c_void_p(), C.c_void_p()
    check(lib.mona_deform2sparsemat(imDim, _ip(N_u), dp(va), dp(dn), int(bool(torus_flag)), w,
                                    C.byref(hn), C.byref(hu), C.byref(hut)))
    prov = dict(N_u=[int(x) for x in N_u], d_u=[1.0, 1.0, 1.0], kernel_type="bump", nWin=2, kernelParam=None)
    out = {}
    if w & 1:
        out["Gn"] = BmSparseMat._from_handle(hn, prov)
    if w & 2:
        out["Gu"] = BmSparseMat._from_handle(hu, prov)
        out["Gut"] = BmSparseMat._from_handle(hut, prov)
    return out


def bmSparseMat_apply(s: BmSparseMat, v, stream=None):
    """w = S v on the GPU: v a complex64 cuda tensor [r_size] or
    [r_size, n_vec] (vector index fastest); returns [l_size(, n_vec)]."""
    import torch
    from ._lib import ptr, stream_of
    if not (v.is_cuda and v.dtype == torch.complex64):
        raise TypeError("v must be a complex64 cuda tensor")
    if v.shape[0] != s.r_size:
        raise ValueError(f"v has {v.shape[0]} rows, S has r_size {s.r_size}")
    v = v.contiguous()
    n_vec = 1 if v.dim() == 1 else int(np.prod(v.shape[1:]))
    w = torch.empty((s.l_size,) + tuple(v.shape[1:]), dtype=torch.complex64, device=v.device)
    check(lib.mona_sparse_apply(s.handle(), ptr(v), ptr(w), n_vec, stream_of(stream)))
    return w


def bmImDeform(Tu, x, n_u=None):
    """x_out = Tu x on the image grid (src/image123/bmImDeform.m:6-43 with
    K = []); Tu None returns x.  x: complex64 cuda tensor of prod(n_u)
    elements (any shape), the result has x's shape."""
    if Tu is None:
        return x
    return bmSparseMat_apply(Tu, x.reshape(-1)).reshape(x.shape)


def bmImDeformT(Tut, x, n_u=None):
    """x_out = Tut x (src/image123/bmImDeformT.m:6-45 with K = [])."""
    return bmImDeform(Tut, x, n_u)


def bmSparseMat_r_nJump2index(r_nJump):
    """0-based row of each entry (bmSparseMat_r_nJump2index_mex.cpp:70-89)."""
    r = _i32(r_nJump)
    out = np.empty(int(r.astype(np.int64).sum()), np.int32)
    lib.bmSparseMat_r_nJump2index_mex(r.size, _ip(r), _ip(out))
    return out


def bmSparseMat_vec(s: BmSparseMat, v, omp=True, complex_flag=None, T_flag=None):
    """W = S*V through the drop-in entry points (host arrays, synchronous),
    with the layout rules of bmSparseMat_vec.m:60-98: column vectors
    [r_size, n_vec] (T_flag false) or row vectors [n_vec, r_size]."""
    s.check()
    v = np.asarray(v)
    if v.ndim == 1:
        v = v[:, None]
    is_c = np.iscomplexobj(v) if complex_flag is None else bool(complex_flag)
    if T_flag is None:
        T_flag = not (v.shape[0] >= v.shape[1])
    lj = _ip(s.l_jump)
    if not T_flag:
        n_vec = v.shape[1]
        if is_c:
            vr = _f32(np.real(v).T)
            vi = _f32(np.imag(v).T)
            wr = np.empty((n_vec, s.l_size), np.float32)
            wi = np.empty((n_vec, s.l_size), np.float32)
            fn = lib.bmSparseMat_cC_oBlock_omp if omp else lib.bmSparseMat_cC_oBlock
            fn(s.r_size, _ip(s.r_jump), _ip(s.r_nJump), _fp(s.m_val), s.l_size, lj, s.l_nJump,
               _fp(vr), _fp(vi), n_vec, _fp(wr), _fp(wi))
            return (wr + 1j * wi).T
        vr = _f32(np.real(v).T)
        wr = np.empty((n_vec, s.l_size), np.float32)
        lib.bmSparseMat_rC_oBlock(s.r_size, _ip(s.r_jump), _ip(s.r_nJump), _fp(s.m_val), s.l_size, lj,
                                  s.l_nJump, _fp(vr), n_vec, _fp(wr))
        return wr.T
    n_vec = v.shape[0]
    # row layout: element (col, k) at col*n_vec + k  ==  v.T in C order
    if is_c:
        vr = _f32(np.real(v).T)
        vi = _f32(np.imag(v).T)
        wr = np.empty((s.l_size, n_vec), np.float32)
        wi = np.empty((s.l_size, n_vec), np.float32)
        lib.bmSparseMat_cR_oBlock(s.r_size, _ip(s.r_jump), _ip(s.r_nJump), _fp(s.m_val), s.l_size, lj,
                                  s.l_nJump, _fp(vr), _fp(vi), n_vec, _fp(wr), _fp(wi))
        return (wr + 1j * wi).T
    vr = _f32(np.real(v).T)
    wr = np.empty((s.l_size, n_vec), np.float32)
    lib.bmSparseMat_rR_oBlock(s.r_size, _ip(s.r_jump), _ip(s.r_nJump), _fp(s.m_val), s.l_size, lj,
                              s.l_nJump, _fp(vr), n_vec, _fp(wr))
    return wr.T
