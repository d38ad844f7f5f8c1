"""ORACLE -- test infrastructure only.

Float64 NumPy restatement of the reference's reconstruction hot path.  Only
tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py
may import this module, and only as the checker (or, for the baseline, as the
timed CPU stand-in); the product library never calls it.

Every function cites the reference file:line it follows.  What pins each one:

* ``traj2sparsemat`` / ``to_bmsparse`` / ``l_squeeze`` / ``cpp_prepare``
  (tests/test_oracle_structure.py): a brute-force dense construction that
  loops over grid cells and their periodic images instead of neighbour
  offsets (``dense_gridding_matrix_bruteforce``; 1-D/2-D/3-D, Gauss and
  Kaiser, odd and even nWin, deleted points, windows wider than the grid),
  rows of Gu summing to 1, Gut == Gu^T exactly, ascending columns, the
  absolute first jump, Gut's kept rows = the touched cells, every
  ``check()`` rule; the bmSparseMat_r_nJump2index closed form ``np.repeat``
  (tests/test_oracle_port.py).
* ``spmv``: dense float64 product of the decoded matrix; the float32
  sequential form is the C port (port.c), pinned bit-exact against the
  reference's own compiled kernel.
* ``bmK``: the closed form Sum_x w(x) cos(2 pi k.x / N) evaluated directly
  (``bmK_direct``), which must equal the reference's oversampled-grid DFT,
  and the continuous Gaussian transform for a wide window.
* ``cfft``/``cifft``/``zero_fill``/``crop``: an explicit centred DFT matrix
  on tiny sizes; the 1-based window N/2+1-n/2 : N/2+n/2.
* ``shanna``/``nakatsha``: the dense composition Gu F Z diag(KFC) on tiny
  problems, and adjointness <A x, y>_HY = <x, A^H y>_HX in float64.
* ``sensa``: on a tiny problem, CG converges to the dense least-squares
  solution of the normal equations.
* ``back_gradient``/``back_gradient_T``: the transpose identity on inputs whose
  first slice is zero; ``prox_one_norm``: closed-form soft threshold.
* ``deform2sparsemat``: closed forms (zero field -> identity, integer torus
  shift -> permutation, fractional 1-D shift -> the two normalised bump
  weights), a brute-force loop construction (``deform_dense_bruteforce``),
  deletion of points that leave the grid, Gut == Gu^T exactly.
* ``gridder_n2u`` / ``gridder_n2u_leight`` / ``mathilda``
  (tests/test_oracle_mathilda.py): constant data grids to exactly 1 on the
  touched cells; the reconstruction through bmGridder_n2u's weights equals
  the one through bmTraj2SparseMat's Gn (bmNasha); the leight restatement
  equals the C port of the leight MEX.
* ``steva``/``teva_chain``/``teva_duo_chain``/``teva_sheet`` (ADMM-L1,
  tests/test_oracle_admm_limit.py): on tiny problems the ADMM iterates reach
  the minimiser of 0.5 ||M x - y||^2_HY + delta HZ ||F x||_1 computed
  independently (the bounded least-squares dual solved by BVLS, with F
  built from the reference's private_F definitions); with deformations,
  identity matrices reproduce the undeformed chain.
* ``morphosia`` and its solvers (tests/test_oracle_morphosia.py): every
  difference term is the adjoint of its dagF; the duo chain with identities
  is the chain at rho/2; a one-column sheet is the chain; the sheet is
  symmetric in its axes; the quadratic (Sensitiva) solvers and
  ``sensa_morphosia`` converge to the dense least-squares solution, and
  ``sensa_morphosia`` with one identity frame is ``sensa``.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import special

# ---------------------------------------------------------------------------
# kernel defaults  (src/varargin/bmVarargin_kernelType_nWin_kernelParam.m:10-28)
# ---------------------------------------------------------------------------


def kernel_defaults(kernelType=None, nWin=None, kernelParam=None):
    if not kernelType:
        kernelType = 'gauss'
    if nWin is None:
        nWin = 3
    if kernelParam is None:
        kernelParam = [0.61, 10.0] if kernelType == 'gauss' else [1.95, 10.0, 10.0]
    # double(single(.)) round trip of bmVarargin_kernelType_nWin_kernelParam.m:30-31
    nWin = float(np.float32(nWin))
    kernelParam = [float(np.float32(v)) for v in kernelParam]
    if kernelType == 'gauss' and len(kernelParam) == 3:
        raise ValueError('Wrong list of gridding kernel parameters.')
    if kernelType == 'kaiser' and len(kernelParam) == 2:
        raise ValueError('Wrong list of gridding kernel parameters.')
    return kernelType, nWin, kernelParam


def kernel_weight(d, kernelType, kernelParam):
    """Radial gridding kernel (bmTraj2SparseMat.m:191-204, bmK.m:70-83)."""
    d = np.asarray(d, dtype=np.float64)
    if kernelType == 'gauss':
        sigma = kernelParam[0]
        # MATLAB normpdf(x,0,sigma) = exp(-0.5*(x/sigma)^2) / (sqrt(2*pi)*sigma)
        return np.exp(-0.5 * (d / sigma) ** 2) / (math.sqrt(2.0 * math.pi) * sigma)
    tau, alpha = kernelParam[0], kernelParam[1]
    w = np.maximum(1.0 - (d / tau) ** 2, 0.0)
    w = alpha * np.sqrt(w)
    return special.i0(w) / special.i0(alpha)


def neighbour_offsets(imDim, nWin):
    """Neighbour offsets, x fastest (bmTraj2SparseMat.m:134-164)."""
    nWin = int(nWin)
    if nWin % 2 == 0:
        r = np.arange(-nWin // 2 - 1, nWin // 2 + 1)
        shift = 0.0
    else:
        h = nWin // 2
        r = np.arange(-h, h + 1)
        shift = 0.5
    grids = np.meshgrid(*([r] * imDim), indexing='ij')
    c = np.stack([g.reshape(-1, order='F') for g in grids], axis=0).astype(np.float64)
    return c, shift


# ---------------------------------------------------------------------------
# structure construction (src/gridding123/m/bmTraj2SparseMat.m:58-297)
# ---------------------------------------------------------------------------


def traj_entries(t, ve, N_u, dK_u, kernelType=None, nWin=None, kernelParam=None):
    """Per-point neighbour cells and weights, before sparse assembly.

    t: [imDim, nPt] trajectory (physical k units), ve: [nPt] volume elements.
    Returns dict with ind_1 (1-based linear cell, [nNb, nKeep]),
    pt (0-based point index of each kept point), weight [nNb, nKeep],
    Dn [nKeep], Du, keep mask.
    """
    kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
    t = np.array(t, dtype=np.float64, copy=True)
    if t.ndim == 1:
        t = t[None, :]
    imDim, nPt = t.shape
    N_u = [float(np.float32(v)) for v in np.ravel(N_u)]
    dK_u = [float(np.float32(v)) for v in np.ravel(dK_u)]
    Dn = np.array(ve, dtype=np.float64).reshape(-1).copy() if ve is not None else None
    Du = 1.0
    shift = []
    for d in range(imDim):
        t[d] = t[d] / dK_u[d]
        if Dn is not None:
            Dn = Dn / dK_u[d]
        shift.append(math.trunc(N_u[d] / 2 + 1))
        FoV = 1.0 / dK_u[d]
        Du = Du * (FoV / N_u[d])
    t = t + np.array(shift, dtype=np.float64)[:, None]
    delete = np.zeros(nPt, dtype=bool)
    for d in range(imDim):
        delete |= (t[d] < 1) | (t[d] > N_u[d])
    c, fshift = neighbour_offsets(imDim, nWin)
    nNb = c.shape[1]
    t_floor = np.floor(t + fshift)
    t_rest = t - t_floor
    dd = t_rest[:, None, :] - c[:, :, None]          # [imDim, nNb, nPt]
    sq = np.zeros(dd.shape[1:])
    for d in range(imDim):
        sq = sq + dd[d] ** 2
    dist = np.sqrt(sq)
    w = kernel_weight(dist, kernelType, kernelParam)  # [nNb, nPt]
    n = t_floor[:, None, :] + c[:, :, None]
    lin = np.ones((nNb, nPt))
    stride = 1.0
    for d in range(imDim):
        nd = np.mod(n[d] - 1, N_u[d]) + 1
        lin = lin + (nd - 1) * stride
        stride *= N_u[d]
    keep = ~delete
    return dict(ind_1=lin[:, keep], weight=w[:, keep], pt=np.nonzero(keep)[0],
                Dn=None if Dn is None else Dn[keep], Du=Du, keep=keep, nNb=nNb,
                nPt=nPt, Nu_tot=int(np.prod(N_u)), N_u=N_u, dK_u=dK_u,
                kernelType=kernelType, nWin=nWin, kernelParam=kernelParam,
                t_grid=t)


def _assemble(rows, cols, vals, shape):
    """MATLAB sparse(rows, cols, vals): duplicates summed, zeros dropped.
    Returns row-major COO sorted by (row, col) with float64 values."""
    key = rows.astype(np.int64) * shape[1] + cols.astype(np.int64)
    order = np.argsort(key, kind='stable')
    key = key[order]
    vals = vals[order]
    uniq, start = np.unique(key, return_index=True)
    summed = np.add.reduceat(vals, start) if len(vals) else vals
    nz = summed != 0
    uniq, summed = uniq[nz], summed[nz]
    return (uniq // shape[1]).astype(np.int64), (uniq % shape[1]).astype(np.int64), summed


def _row_normalise(r, c, v, nrow):
    """diag(1./sum(A,2)) * A with the row sum taken in ascending column order
    (bmTraj2SparseMat.m:249-254, 275-280)."""
    s = np.zeros(nrow)
    np.add.at(s, r, v)   # rows visited in ascending column order (sorted COO)
    inv = np.zeros(nrow)
    nzr = s != 0
    inv[nzr] = 1.0 / s[nzr]
    return v * inv[r]


def traj2sparsemat(t, ve, N_u, dK_u, kernelType=None, nWin=None, kernelParam=None,
                   which=('Gn', 'Gu', 'Gut')):
    """bmTraj2SparseMat (src/gridding123/m/bmTraj2SparseMat.m:12-314).
    Returns dict of bmSparseMat field dicts (cpp_prepared, one_block)."""
    e = traj_entries(t, ve, N_u, dK_u, kernelType, nWin, kernelParam)
    nNb, nPt, Nu = e['nNb'], e['nPt'], e['Nu_tot']
    rows = np.repeat(e['pt'], nNb)                       # ind_2 - 1
    cols = (e['ind_1'].T.reshape(-1) - 1).astype(np.int64)  # ind_1 - 1
    w = e['weight'].T.reshape(-1)
    prov = dict(N_u=e['N_u'], d_u=e['dK_u'], kernel_type=e['kernelType'],
                nWin=e['nWin'], kernelParam=e['kernelParam'])
    out = {}
    if 'Gn' in which:
        Dn = np.repeat(e['Dn'], nNb)
        # sparse(ind_2, ind_1, w.*Dn)' : [Nu, nPt]
        r, c, v = _assemble(cols, rows, w * Dn, (Nu, nPt))
        v = _row_normalise(r, c, v, Nu)
        G = to_bmsparse(r, c, v, (Nu, nPt), prov)
        l_squeeze(G)
        cpp_prepare(G)
        out['Gn'] = G
    if 'Gu' in which or 'Gut' in which:
        r, c, v = _assemble(rows, cols, w * e['Du'], (nPt, Nu))
        v = _row_normalise(r, c, v, nPt)
        Gu = to_bmsparse(r, c, v, (nPt, Nu), prov)
        Gut = to_bmsparse(c, r, v, (Nu, nPt), prov)
        cpp_prepare(Gu)
        l_squeeze(Gut)
        cpp_prepare(Gut)
        out['Gu'] = Gu
        out['Gut'] = Gut
    return out


def deform_entries(v, N_u, Dn=None, torus_flag=False):
    """bmImDeformField2sparseMat_ind_weight
    (src/gridding123/m/bmImDeformField2sparseMat_ind_weight.m:6-150):
    pixel p (1-based grid position, x fastest) moves to t = x_p + v[:, p];
    its 2^imDim neighbours floor(t) + c, c in {0,1}^imDim (x fastest),
    wrapped mod N_u, weigh exp(-1/(1-d^2)) (inf -> 0) times (d < 1) with d
    the Euclidean distance of t - floor(t) to c.  Points that leave
    [1, N_u] are deleted unless torus_flag.  v: [imDim, nPt] (float64)."""
    N_u = np.asarray(N_u, dtype=np.int64).ravel()
    imDim = N_u.size
    nPt = int(np.prod(N_u))
    v = np.asarray(v, dtype=np.float64).reshape(-1, order='F').reshape(imDim, nPt, order='F')
    grids = np.meshgrid(*[np.arange(1, n + 1, dtype=np.float64) for n in N_u], indexing='ij')
    x = np.stack([g.reshape(-1, order='F') for g in grids])          # [imDim, nPt]
    t = x + v
    delete = np.zeros(nPt, bool)
    for d in range(imDim):
        delete |= (t[d] < 1) | (t[d] > N_u[d])
    c = np.stack([g.reshape(-1, order='F') for g in
                  np.meshgrid(*[np.arange(2)] * imDim, indexing='ij')])  # [imDim, nNb]
    nNb = c.shape[1]
    tf = np.floor(t)
    tr = t - tf
    dd = tr[:, None, :] - c[:, :, None]                               # [imDim, nNb, nPt]
    dist = np.sqrt((dd ** 2).sum(axis=0))
    with np.errstate(divide='ignore', over='ignore'):
        w = np.exp(-1.0 / (1.0 - dist ** 2))
    w[np.isinf(w)] = 0.0
    w = w * (np.abs(dist) < 1)
    n = tf[:, None, :] + c[:, :, None]
    lin = np.zeros((nNb, nPt), np.int64)
    stride = 1
    for d in range(imDim):
        lin += (np.mod(n[d] - 1, N_u[d]).astype(np.int64)) * stride
        stride *= int(N_u[d])
    keep = np.ones(nPt, bool) if torus_flag else ~delete
    Dn = None if Dn is None else np.asarray(Dn, dtype=np.float64).ravel()
    return dict(ind_1=lin[:, keep], weight=w[:, keep], pt=np.nonzero(keep)[0],
                Dn=None if Dn is None else Dn[keep], nNb=nNb, nPt=nPt, Nu_tot=nPt, N_u=N_u)


def deform2sparsemat(v, N_u, Dn=None, torus_flag=False, which=('Gu', 'Gut')):
    """bmImDeformField2SparseMat (src/gridding123/m/bmImDeformField2SparseMat.m:14-147):
    Gu = rownorm(sparse(pt, cell, w)), Gut = Gu', Gn = rownorm(sparse(pt,
    cell, w.*Dn)'), all cpp_prepared without l_squeeze ('bump' kernel,
    d_u = 1, nWin = 2)."""
    e = deform_entries(v, N_u, Dn, torus_flag)
    nNb, nPt, Nu = e['nNb'], e['nPt'], e['Nu_tot']
    rows = np.repeat(e['pt'], nNb)
    cols = e['ind_1'].T.reshape(-1)
    w = e['weight'].T.reshape(-1)
    prov = dict(N_u=e['N_u'], d_u=np.ones(3), kernel_type='bump', nWin=2, kernelParam=None)
    out = {}
    if 'Gn' in which:
        if e['Dn'] is None:
            raise ValueError('Gn needs Dn')
        r, c, val = _assemble(cols, rows, w * np.repeat(e['Dn'], nNb), (Nu, nPt))
        val = _row_normalise(r, c, val, Nu)
        G = to_bmsparse(r, c, val, (Nu, nPt), prov)
        cpp_prepare(G)
        out['Gn'] = G
    if 'Gu' in which or 'Gut' in which:
        r, c, val = _assemble(rows, cols, w, (nPt, Nu))
        val = _row_normalise(r, c, val, nPt)
        Gu = to_bmsparse(r, c, val, (nPt, Nu), prov)
        Gut = to_bmsparse(c, r, val, (Nu, nPt), prov)
        cpp_prepare(Gu)
        cpp_prepare(Gut)
        out['Gu'] = Gu
        out['Gut'] = Gut
    return out


def deform_dense_bruteforce(v, N_u, torus_flag=False, normalise=True):
    """Independent restatement used to pin deform2sparsemat: plain loops over
    pixels and neighbour corners, math.exp, dense accumulation and (unless
    normalise is False) row normalisation (float64)."""
    N_u = [int(n) for n in np.ravel(N_u)]
    imDim = len(N_u)
    nPt = int(np.prod(N_u))
    v = np.asarray(v, dtype=np.float64).reshape(-1, order='F')
    A = np.zeros((nPt, nPt))
    for p in range(nPt):
        q, pos = p, []
        for d in range(imDim):
            pos.append(q % N_u[d] + 1 + v[p * imDim + d])
            q //= N_u[d]
        if not torus_flag and any(pos[d] < 1 or pos[d] > N_u[d] for d in range(imDim)):
            continue
        for corner in range(2 ** imDim):
            sq, cell, stride = 0.0, 0, 1
            for d in range(imDim):
                cd = (corner >> d) & 1
                base = math.floor(pos[d])
                sq += (pos[d] - base - cd) ** 2
                cell += ((base + cd - 1) % N_u[d]) * stride
                stride *= N_u[d]
            if sq < 1.0:
                A[p, cell] += math.exp(-1.0 / (1.0 - sq))
        s = A[p].sum()
        if s and normalise:
            A[p] /= s
    return A


def to_bmsparse(r, c, v, shape, prov=None):
    """bmSparseMat_matlabSparse2bmSparseMat
    (src/sparseMat/m/bmSparseMat_matlabSparse2bmSparseMat.m:6-40):
    entries row by row, ascending column index, m_val rounded to single."""
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    r_nJump = np.bincount(r, minlength=shape[0]).astype(np.int32)
    G = dict(r_size=int(shape[1]), l_size=int(shape[0]), l_nJump=int(shape[0]),
             r_nJump=r_nJump, r_ind=(c + 1).astype(np.int64),
             m_val=v.astype(np.float32), m_val64=v, l_ind=None, l_jump=None,
             r_jump=None, type='matlab_ind', l_squeeze_flag=False,
             block_type='void')
    if prov:
        G.update(prov)
    return G


def l_squeeze(G):
    """bmSparseMat.l_squeeze (src/sparseMat/m/bmSparseMat.m:62-94)."""
    assert G['type'] == 'matlab_ind'
    mask = G['r_nJump'] != 0
    G['r_nJump'] = G['r_nJump'][mask].astype(np.int32)
    G['l_ind'] = (np.arange(1, G['l_size'] + 1)[mask]).astype(np.int64)
    G['l_nJump'] = int(mask.sum())
    G['type'] = 'l_squeezed_matlab_ind'
    G['l_squeeze_flag'] = True


def cpp_prepare(G):
    """bmSparseMat.cpp_prepare('one_block') (src/sparseMat/m/bmSparseMat.m:100-188)
    with bmSparseMat_blockPartition(...,'one') (bmSparseMat_blockPartition.m:26-33)."""
    r_ind = G['r_ind']
    if r_ind.size == 0:
        raise ValueError('The sparse matrix got a zero block. Preparation aborted.')
    rj = np.empty(r_ind.shape, dtype=np.int64)
    rj[0] = r_ind[0] - 1
    rj[1:] = np.diff(r_ind)
    G['r_jump'] = rj.astype(np.int32)
    if G['type'] == 'l_squeezed_matlab_ind':
        li = G['l_ind']
        lj = np.empty(li.shape, dtype=np.int64)
        lj[0] = li[0] - 1
        lj[1:] = np.diff(li)
        G['l_jump'] = lj.astype(np.int32)
        G['type'] = 'l_squeezed_cpp_prepared'
    else:
        G['l_jump'] = None
        G['type'] = 'cpp_prepared'
    G['nBlock'] = 1
    G['l_block_start'] = 0
    G['block_length'] = int(len(G['r_nJump']))
    G['m_block_start'] = 0
    G['block_type'] = 'one_block'
    G['r_ind'] = None
    G['l_ind'] = None


def decode(G):
    """Absolute column per entry, row pointer and output row per stored row
    of a cpp_prepared one-block bmSparseMat."""
    col = np.cumsum(G['r_jump'].astype(np.int64))
    rowptr = np.zeros(len(G['r_nJump']) + 1, dtype=np.int64)
    rowptr[1:] = np.cumsum(G['r_nJump'].astype(np.int64))
    if G.get('l_jump') is not None:
        rows = np.cumsum(G['l_jump'].astype(np.int64))
    else:
        rows = np.arange(len(G['r_nJump']), dtype=np.int64)
    return col, rowptr, rows


def to_dense(G, dtype=np.float64, exact=False):
    col, rowptr, rows = decode(G)
    A = np.zeros((G['l_size'], G['r_size']), dtype=dtype)
    vals = G['m_val'].astype(np.float64)
    rid = np.repeat(rows, np.diff(rowptr))
    np.add.at(A, (rid, col), vals)
    return A


def to_scipy(G):
    from scipy import sparse
    col, rowptr, rows = decode(G)
    rid = np.repeat(rows, np.diff(rowptr))
    return sparse.csr_matrix((G['m_val'].astype(np.float64), (rid, col)),
                             shape=(G['l_size'], G['r_size']))


def spmv(G, V):
    """W = S V for V [r_size, nVec] (float64)."""
    return to_scipy(G) @ V


def r_nJump2index(r_nJump):
    """bmSparseMat_r_nJump2index_mex.cpp:70-89 (0-based)."""
    r = np.asarray(r_nJump, dtype=np.int64)
    return np.repeat(np.arange(len(r), dtype=np.int32), r)


def check(G):
    """The rules of bmSparseMat.check (src/sparseMat/m/bmSparseMat.m:192-393)
    for a cpp_prepared matrix; returns list of failed rule names."""
    fails = []
    rn = G['r_nJump']
    if G['l_size'] < 2 or G['r_size'] < 2 or G['l_nJump'] < 2:
        fails.append('2.5')
    if np.any(rn < 0):
        fails.append('2.6')
    s = int(rn.astype(np.int64).sum())
    if s == 0:
        fails.append('3.1')
    elif s != G['m_val'].size:
        fails.append('3.2')
    if rn.size != G['l_nJump']:
        fails.append('3.3')
    if G['r_jump'] is None:
        fails.append('8.2')
    if G['type'] == 'cpp_prepared':
        if G.get('l_jump') is not None:
            fails.append('9.1')
        if s != G['r_jump'].size:
            fails.append('9.2')
    if G['type'] == 'l_squeezed_cpp_prepared':
        if G.get('l_jump') is None:
            fails.append('10.1')
        elif G['l_jump'].size != G['l_nJump']:
            fails.append('10.2')
    if not (G['r_size'] < 2**31 - 1 and G['l_size'] < 2**31 - 1):
        fails.append('11.1')
    return fails


def dense_gridding_matrix_bruteforce(t, N_u, dK_u, kernelType=None, nWin=None,
                                     kernelParam=None):
    """Independent restatement used to pin traj2sparsemat: loops over every
    cell of the periodic grid and, per axis, over every periodic image of the
    cell (offset g + 1 - base + k N for integer k), keeping the images whose
    offset lies in the window: [-h, h] around round(t) for odd nWin,
    [-nWin/2 - 1, nWin/2] around floor(t) for even nWin
    (bmTraj2SparseMat.m:134-164).  A window wider than the grid visits a
    cell through several images, whose weights add (MATLAB's sparse() sums
    duplicates).  Points outside [1, N] on any axis are dropped
    (bmTraj2SparseMat.m:116-126).  Returns the un-normalised dense [nPt, Nu]
    weight matrix (w * Du)."""
    kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
    t = np.atleast_2d(np.asarray(t, dtype=np.float64))
    imDim, nPt = t.shape
    N = [int(v) for v in np.ravel(N_u)]
    dK = [float(np.float32(v)) for v in np.ravel(dK_u)]
    Nu = int(np.prod(N))
    A = np.zeros((nPt, Nu))
    Du = 1.0
    for d in range(imDim):
        Du *= (1.0 / dK[d]) / N[d]
    nW = int(nWin)
    if nW % 2 == 1:
        lo, hi, fshift = -(nW // 2), nW // 2, 0.5
    else:
        lo, hi, fshift = -nW // 2 - 1, nW // 2, 0.0
    for p in range(nPt):
        tg = [t[d, p] / dK[d] + math.trunc(N[d] / 2 + 1) for d in range(imDim)]
        if any(tg[d] < 1 or tg[d] > N[d] for d in range(imDim)):
            continue
        base = [math.floor(tg[d] + fshift) for d in range(imDim)]   # 1-based
        # per axis: the (cell, offset) pairs of every periodic image in the window
        per_axis = []
        for d in range(imDim):
            cells, offs = [], []
            kmax = (hi - lo) // N[d] + 2
            for g in range(N[d]):
                for k in range(-kmax, kmax + 1):
                    o = g + 1 - base[d] + k * N[d]
                    if lo <= o <= hi:
                        cells.append(g)
                        offs.append(o)
            per_axis.append((np.array(cells, np.int64), np.array(offs, np.float64)))
        mesh_c = np.meshgrid(*[pa[0] for pa in per_axis], indexing='ij')
        mesh_o = np.meshgrid(*[pa[1] for pa in per_axis], indexing='ij')
        lin = np.zeros(mesh_c[0].shape, dtype=np.int64)
        sq = np.zeros(mesh_c[0].shape)
        stride = 1
        for d in range(imDim):
            lin += mesh_c[d] * stride
            stride *= N[d]
            sq = sq + ((tg[d] - base[d]) - mesh_o[d]) ** 2
        w = kernel_weight(np.sqrt(sq), kernelType, kernelParam)
        np.add.at(A[p], lin.ravel(), w.ravel() * Du)
    return A


# ---------------------------------------------------------------------------
# centred DFTs, zero-fill, crop
# ---------------------------------------------------------------------------


def cfft(x, axes):
    """fftshift(fft(ifftshift(x,n),[],n),n) per axis (bmShanna_MATLAB.m:46-50)."""
    for a in axes:
        x = np.fft.fftshift(np.fft.fft(np.fft.ifftshift(x, axes=a), axis=a), axes=a)
    return x


def cifft(x, axes):
    """fftshift(ifft(ifftshift(x,n),[],n),n) per axis (bmNakatsha_MATLAB.m:36-40)."""
    for a in axes:
        x = np.fft.fftshift(np.fft.ifft(np.fft.ifftshift(x, axes=a), axis=a), axes=a)
    return x


def center_slices(N, n):
    """Centre window N/2+1-n/2 : N/2+n/2 (1-based) of bmImZeroFill.m:41-51 /
    bmImCrope.m, as 0-based slices."""
    return tuple(slice(N[d] // 2 - n[d] // 2, N[d] // 2 - n[d] // 2 + n[d])
                 for d in range(len(N)))


def zero_fill(x, N, n):
    out = np.zeros(tuple(N) + x.shape[len(N):], dtype=np.complex128)
    out[center_slices(N, n)] = x
    return out


def crop(x, N, n):
    return x[center_slices(N, n)]


# ---------------------------------------------------------------------------
# deapodisation (src/fourier123/prep_function/bmK.m:6-156, bmKF.m, bmKF_conj.m)
# ---------------------------------------------------------------------------


def bmK(N_u, dK_u, kernelType=None, nWin=None, kernelParam=None):
    """K on the N_u grid (x fastest index = axis 0), float64, before the
    single() cast.  Follows bmK.m:29-151 literally: kernel on the 2x finer
    grid, per-axis box mask, centred DFT, centre crop, 1/|Re| normalised,
    clamp K_max."""
    kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
    N = [int(v) for v in np.ravel(N_u)]
    imDim = len(N)
    osf = 2
    Nos = [round(n * osf) for n in N]
    ax = [np.arange(-n * osf // 2, n * osf // 2) / osf for n in N]
    g = np.meshgrid(*ax, indexing='ij')
    d = np.sqrt(sum(gi ** 2 for gi in g))
    w = kernel_weight(d, kernelType, kernelParam)
    K_max = kernelParam[-1]
    h = int(nWin) // 2
    if imDim == 1:
        # bmK.m:88 masks x < nWin_half (reference quirk reproduced)
        w[(g[0] < h) | (g[0] > h)] = 0
    else:
        for gi in g:
            w[(gi < -h) | (gi > h)] = 0
    F = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(w)))
    F = F / (np.prod(Nos) * np.prod([1.0 / n for n in N]))
    sl = tuple(slice(Nos[i] // 2 - N[i] // 2, Nos[i] // 2 + N[i] // 2) for i in range(imDim))
    K = np.abs(np.real(F[sl]))
    K = K / np.max(np.abs(K))
    K = 1.0 / K
    return np.minimum(K, K_max)


def bmK_direct(N_u, kernelType=None, nWin=None, kernelParam=None):
    """Closed form of the same K: Sum over the window points x (half-integer
    spacing) of w(|x|) cos(2 pi sum_d x_d k_d / N_d), k in [-N/2, N/2)."""
    kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
    N = [int(v) for v in np.ravel(N_u)]
    h = int(nWin) // 2
    xs = np.arange(-2 * h, 2 * h + 1) / 2.0
    pts = np.meshgrid(*([xs] * len(N)), indexing='ij')
    pts = np.stack([p.ravel() for p in pts], axis=1)
    wts = kernel_weight(np.sqrt((pts ** 2).sum(1)), kernelType, kernelParam)
    ks = np.meshgrid(*[np.arange(-n // 2, n // 2) for n in N], indexing='ij')
    phase = sum(pts[:, d][:, None] * ks[d].ravel()[None, :] / N[d] for d in range(len(N)))
    K = np.abs((wts[:, None] * np.cos(2 * np.pi * phase)).sum(0)).reshape([n for n in N])
    K = K / wts.sum()
    return np.minimum(1.0 / K, kernelParam[-1])


def kf(C, N_u, n_u, dK_u, kernelType=None, nWin=None, kernelParam=None):
    """KFC = K.*F.*C, F = 1/prod(N_u)/prod(dK_u)  (bmKF.m:6-35).
    C: [n_u..., nCh] complex.  Returns [n_u..., nCh] complex128."""
    K = crop(bmK(N_u, dK_u, kernelType, nWin, kernelParam), N_u, n_u)
    F = 1.0 / np.prod(N_u) / np.prod(dK_u)
    return K[..., None] * F * C


def kf_conj(C_conj, N_u, n_u, dK_u, kernelType=None, nWin=None, kernelParam=None):
    """KFC_conj = K.*F_conj.*conj(C), F_conj = 1/prod(dK_u) (bmKF_conj.m:6-38)."""
    K = crop(bmK(N_u, dK_u, kernelType, nWin, kernelParam), N_u, n_u)
    return K[..., None] * (1.0 / np.prod(dK_u)) * C_conj


# ---------------------------------------------------------------------------
# operators: images are [n_u..., ] arrays indexed (x, y, z); grids flatten
# x fastest (Fortran order) to match the reference's column layout.
# ---------------------------------------------------------------------------


def _flatF(a, ndim):
    return a.reshape((-1,) + a.shape[ndim:], order='F')


def _unflatF(a, shape):
    return a.reshape(tuple(shape) + a.shape[1:], order='F')


def shanna(x, Gu, KFC, N_u, n_u):
    """Forward model y = Gu * cfft(zero_fill(x .* KFC))
    (src/fourier123/map_function/nonCartesian/bmShanna_MATLAB.m:6-56).
    x: [n_u...] image, KFC [n_u..., nCh].  Returns y [nPt, nCh]."""
    d = len(N_u)
    xc = x[..., None] * KFC
    g = cfft(zero_fill(xc, N_u, n_u), axes=range(d))
    return spmv(Gu, _flatF(g, d))


def nakatsha(y, Gut, KFC_conj, C_flag, N_u, n_u):
    """Adjoint x = sum_c KFC_conj .* crop(cifft(Gut * y))
    (src/fourier123/map_function/nonCartesian/bmNakatsha_MATLAB.m:6-60)."""
    d = len(N_u)
    g = _unflatF(spmv(Gut, y), N_u)
    g = crop(cifft(g, axes=range(d)), N_u, n_u) * KFC_conj
    return g.sum(-1) if C_flag else g


def hx_of(N_u, dK_u):
    """HX = prod(dX_u), dX_u = (1/dK_u)/N_u  (bmSensa.m:26-27)."""
    return float(np.prod([(1.0 / dk) / n for dk, n in zip(dK_u, N_u)]))


def sensa(x, y, ve, C, Gu, Gut, N_u, n_u, dK_u, nCGD, ve_max=None, nIter=1,
          kernel=(None, None, None)):
    """CG-SENSE, float64 scalars and vectors
    (src/fourier123/solver_function/imDim/nonCartesian/bmSensa.m:6-96).
    Returns (x, residuals per outer iteration)."""
    eps = 10 * np.finfo(np.float32).eps
    KFC = kf(C, N_u, n_u, dK_u, *kernel)
    KFCc = kf_conj(np.conj(C), N_u, n_u, dK_u, *kernel)
    HX = hx_of(N_u, dK_u)
    ve = np.broadcast_to(np.asarray(ve, dtype=np.float64).reshape(-1, 1), y.shape)
    if ve_max is None:
        ve_max = ve.max()
    HY = np.minimum(ve, ve_max)
    x = np.array(x, dtype=np.complex128)
    M = lambda v: shanna(v, Gu, KFC, N_u, n_u)
    Mh = lambda r: (1.0 / HX) * nakatsha(HY * r, Gut, KFCc, True, N_u, n_u)
    Rs = []
    for _ in range(nIter):
        res = y - M(x)
        dag = Mh(res)
        sqn = np.real(np.vdot(dag, HX * dag))
        p = dag
        for i in range(1, nCGD + 1):
            if sqn < eps:
                break
            Mp = M(p)
            den = np.real(np.vdot(Mp, HY * Mp))
            a = sqn / den
            x = x + a * p
            if i == nCGD:
                break
            res = res - a * Mp
            dag = Mh(res)
            sqn_new = np.real(np.vdot(dag, HX * dag))
            b = sqn_new / sqn
            sqn = sqn_new
            p = dag + b * p
        r = y - M(x)
        Rs.append(float(np.real(np.vdot(r, HY * r))))
    return x, Rs


def back_gradient(x, dX, axis):
    """bmBackGradient_n (src/optim/bmBackGradient_n.m:36-44): backward
    difference with circular shift, first slice set to 0."""
    g = (x - np.roll(x, 1, axis=axis)) / dX[axis]
    idx = [slice(None)] * x.ndim
    idx[axis] = 0
    g[tuple(idx)] = 0
    return g


def back_gradient_T(x, dX, axis):
    """bmBackGradient_nT (src/optim/bmBackGradient_nT.m:34-44)."""
    g = x.copy()
    idx = [slice(None)] * x.ndim
    idx[axis] = 0
    g[tuple(idx)] = 0
    return (g - np.roll(x, -1, axis=axis)) / dX[axis]


def prox_one_norm(w, s):
    """Soft threshold applied to real and imaginary parts separately
    (src/linSpace/bmProx_oneNorm.m:30-41)."""
    def st(v):
        return (v > s) * (v - s) + (v < -s) * (v + s)
    return st(np.real(w)) + 1j * st(np.imag(w))


def steva(x, z, u, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, rho, nCGD,
          ve_max=None, nIter=1, kernel=(None, None, None)):
    """Spatial-TV ADMM (src/fourier123/solver_function/imDim/nonCartesian/
    bmSteva.m:6-138), float64.  z, u: lists of per-axis images or None."""
    eps = 10 * np.finfo(np.float32).eps
    d = len(N_u)
    KFC = kf(C, N_u, n_u, dK_u, *kernel)
    KFCc = kf_conj(np.conj(C), N_u, n_u, dK_u, *kernel)
    dX = [(1.0 / dk) / n for dk, n in zip(dK_u, N_u)]
    HX = hx_of(N_u, dK_u)
    HZ = HX
    ve = np.broadcast_to(np.asarray(ve, dtype=np.float64).reshape(-1, 1), y.shape)
    if ve_max is None:
        ve_max = ve.max()
    HY = np.minimum(ve, ve_max)
    delta = np.broadcast_to(np.atleast_1d(np.abs(delta)), (nIter,)) if np.size(delta) == 1 \
        else np.linspace(abs(delta[0]), abs(delta[1]), nIter)
    rho = np.broadcast_to(np.atleast_1d(np.abs(rho)), (nIter,)) if np.size(rho) == 1 \
        else np.linspace(abs(rho[0]), abs(rho[1]), nIter)
    x = np.array(x, dtype=np.complex128)
    M = lambda v: shanna(v, Gu, KFC, N_u, n_u)
    Mh = lambda r: (1.0 / HX) * nakatsha(HY * r, Gut, KFCc, True, N_u, n_u)
    Fop = lambda v: np.stack([back_gradient(v, dX, a) for a in range(d)])
    FT = lambda w: sum(back_gradient_T(w[a], dX, a) for a in range(d))
    z = Fop(x) if z is None else np.array(z, dtype=np.complex128)
    u = np.zeros_like(z) if u is None else np.array(u, dtype=np.complex128)
    hist = []
    for c in range(nIter):
        rc = rho[c]
        ry = y - M(x)
        rz = (z - u) - Fop(x)
        dag = Mh(ry) + rc * (1.0 / HX) * FT(HZ * rz)
        p = dag
        sqn = np.real(np.vdot(dag, HX * dag))
        for i in range(1, nCGD + 1):
            if sqn < eps:
                break
            Mp = M(p)
            Fp = Fop(p)
            den = np.real(np.vdot(Mp, HY * Mp)) + np.real(np.vdot(Fp, rc * HZ * Fp))
            a = sqn / den
            x = x + a * p
            if i == nCGD:
                break
            ry = ry - a * Mp
            rz = rz - a * Fp
            dag = Mh(ry) + rc * (1.0 / HX) * FT(HZ * rz)
            sqn_new = np.real(np.vdot(dag, HX * dag))
            b = sqn_new / sqn
            sqn = sqn_new
            p = dag + b * p
        Fx_u = Fop(x) + u
        z = prox_one_norm(Fx_u, delta[c] / rc)
        u = Fx_u - z
        r = y - M(x)
        R = float(np.real(np.vdot(r, HY * r)))
        g = Fop(x)
        TV = HZ * (np.abs(g.real).sum() + np.abs(g.imag).sum())
        hist.append((R, TV))
    return x, z, u, hist


def _frame_shift(f, n1, n2, axis, d):
    """Frame f = i + n1*j of an n1 x n2 cell array (MATLAB order) moved by d
    along axis 0 (i) or 1 (j), cyclically."""
    i, j = f % n1, f // n1
    if axis == 0:
        i = (i + d) % n1
    else:
        j = (j + d) % n2
    return i + n1 * j


def _schedule(v, nIter):
    """private_init_regul_param (bmTevaMorphosia_chain.m:226-235)."""
    return np.broadcast_to(np.atleast_1d(np.abs(v)), (nIter,)) if np.size(v) == 1 \
        else np.linspace(abs(v[0]), abs(v[1]), nIter)


def morph_terms(terms, nFr, n1, HX, HZ):
    """Per term: the unscaled difference Tu_f x_f - x_shift(f, dir), F (its
    scaled form) and dagF (private_F / private_dagF of the Morphosia
    solvers).  Images are n_u arrays, flattened column-major for Tu."""
    n2 = nFr // n1
    deform = lambda G, a: a if G is None else (G @ a.reshape(-1, order='F')).reshape(a.shape, order='F')
    out = []
    for t in terms:
        Tu = [None] * nFr if t.get('Tu') is None else [None if G is None else to_scipy(G) for G in t['Tu']]
        Tut = [None] * nFr if t.get('Tut') is None else [None if G is None else to_scipy(G) for G in t['Tut']]
        fw = [_frame_shift(f, n1, n2, t['axis'], t['dir']) for f in range(nFr)]
        bw = [_frame_shift(f, n1, n2, t['axis'], -t['dir']) for f in range(nFr)]
        sc = float(t.get('scale', 1.0))
        d = (lambda Tu, fw: lambda v: [deform(Tu[i], v[i]) - v[fw[i]] for i in range(nFr)])(Tu, fw)
        F = (lambda d, sc: lambda v: [sc * e for e in d(v)])(d, sc)
        FT = (lambda Tut, bw, sc: lambda w: [(sc / HX) * (deform(Tut[i], HZ * w[i]) - HZ * w[bw[i]])
                                             for i in range(nFr)])(Tut, bw, sc)
        out.append(dict(diff=d, F=F, FT=FT))
    return out


def morphosia(x, z, u, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, rho, nCGD, ve_max, nIter, kernel,
              terms, l1=True, nFr1=None, first_dagF_scaled=True):
    """Shared body of the Morphosia solvers (src/fourier123/solver_function/
    imDim_morphosia/nonCartesian/bmTevaMorphosia_chain.m, bmTevaDuoMorphosia_chain.m,
    bmTevaMorphosia_sheet.m, bmSensitivaMorphosia_chain.m, bmSensitivaMorphosia_sheet.m),
    regul_mode 'normal', float64.

    Frames f = i + nFr1*j of an nFr1 x nFr2 cell array.  Term t has
    F_t x_f = s_t (Tu_t,f x_f - x_shift(f, axis_t, dir_t)) and
    dagF_t z_f = (s_t/HX)(Tut_t,f (HZ z_f) - HZ z_shift(f, axis_t, -dir_t)).
    l1: ADMM with z_t, u_t and the soft threshold (Teva); else the quadratic
    penalty delta*||F x||^2 solved by restarted CG (Sensitiva: z = u = 0,
    delta in place of rho).  regul sums |.| (l1) or |.|^2 of the unscaled
    differences Tu x_f - x_shift.  first_dagF_scaled False leaves the
    penalty weight off the first gradient of each outer iteration
    (bmSensitivaMorphosia_sheet.m:118-119 against :164-165)."""
    eps = 10 * np.finfo(np.float32).eps
    nFr = len(y)
    n1 = nFr if nFr1 is None else nFr1
    KFC = kf(C, N_u, n_u, dK_u, *kernel)
    KFCc = kf_conj(np.conj(C), N_u, n_u, dK_u, *kernel)
    HX = hx_of(N_u, dK_u)
    HZ = HX
    HY = [np.minimum(np.broadcast_to(np.asarray(v, dtype=np.float64).reshape(-1, 1), yy.shape),
                     ve_max) for v, yy in zip(ve, y)]
    delta = _schedule(delta, nIter)
    rho = _schedule(rho, nIter) if l1 else None
    x = [np.array(v, dtype=np.complex128) for v in x]
    M = lambda v: [shanna(v[i], Gu[i], KFC, N_u, n_u) for i in range(nFr)]
    Mh = lambda r: [(1.0 / HX) * nakatsha(HY[i] * r[i], Gut[i], KFCc, True, N_u, n_u)
                    for i in range(nFr)]
    T = morph_terms(terms, nFr, n1, HX, HZ)
    diff = lambda t, v: t['diff'](v)
    Fop = lambda t, v: t['F'](v)
    FT = lambda t, w: t['FT'](w)
    sq = lambda a, H: sum(float(np.real(np.vdot(ai, (H[i] if isinstance(H, list) else H) * ai)))
                          for i, ai in enumerate(a))
    add = lambda a, b: [p + q for p, q in zip(a, b)]
    if l1:
        z = [Fop(t, x) for t in T] if z is None else [[np.array(v, dtype=np.complex128) for v in zt] for zt in z]
        u = [[np.zeros_like(v) for v in x] for _ in T] if u is None else \
            [[np.array(v, dtype=np.complex128) for v in ut] for ut in u]
    hist = []
    for c in range(nIter):
        dl = delta[c]
        w = rho[c] if l1 else dl                   # weight of the penalty terms
        ry = [yy - m for yy, m in zip(y, M(x))]
        if l1:
            rz = [[(z[k][i] - u[k][i]) - f for i, f in enumerate(Fop(t, x))] for k, t in enumerate(T)]
        else:
            rz = [[0 - f for f in Fop(t, x)] for t in T]
        w0 = w if first_dagF_scaled else 1.0
        dag = Mh(ry)
        pen = None
        for k, t in enumerate(T):
            g = [w0 * v for v in FT(t, rz[k])]
            pen = g if pen is None else add(pen, g)
        dag = add(dag, pen)
        p = dag
        sqn = sq(dag, HX)
        for j in range(1, nCGD + 1):
            if sqn < eps:
                break
            Mp = M(p)
            Fp = [Fop(t, p) for t in T]
            den = sq(Mp, HY) + sum(sq(Fp[k], w * HZ) for k in range(len(T)))
            a = sqn / den
            x = [xi + a * pi for xi, pi in zip(x, p)]
            if j == nCGD:
                break
            ry = [r - a * m for r, m in zip(ry, Mp)]
            rz = [[r - a * f for r, f in zip(rz[k], Fp[k])] for k in range(len(T))]
            dag = Mh(ry)
            pen = None
            for k, t in enumerate(T):
                g = [w * v for v in FT(t, rz[k])]
                pen = g if pen is None else add(pen, g)
            dag = add(dag, pen)
            sqn_new = sq(dag, HX)
            b = sqn_new / sqn
            sqn = sqn_new
            p = [dg + b * pi for dg, pi in zip(dag, p)]
        if l1:
            for k, t in enumerate(T):
                Fx_u = [ui + fi for ui, fi in zip(u[k], Fop(t, x))]
                z[k] = [prox_one_norm(v, dl / w) for v in Fx_u]
                u[k] = [a_ - b_ for a_, b_ in zip(Fx_u, z[k])]
        dafi = sq([yy - m for yy, m in zip(y, M(x))], HY)
        if l1:
            regul = sum(HZ * (np.abs(d.real).sum() + np.abs(d.imag).sum()) for t in T for d in diff(t, x))
        else:
            regul = sum(sq(diff(t, x), HZ) for t in T)
        hist.append((0.5 * dafi + 0.5 * dl * regul, dafi, regul))
    return x, (z if l1 else None), (u if l1 else None), hist


def teva_chain(x, z, u, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, rho, nCGD,
               ve_max, nIter=1, kernel=(None, None, None), Tu=None, Tut=None):
    """bmTevaMorphosia_chain (src/fourier123/solver_function/imDim_morphosia/
    nonCartesian/bmTevaMorphosia_chain.m:6-294): one term,
    F x_i = Tu_i x_i - x_{i-1} (private_F, :266-272),
    dagF z_i = (1/HX)(Tut_i (HZ z_i) - HZ z_{i+1}) (private_dagF, :285-291);
    Tu entries None = identity (bmImDeform.m:6-10).  Lists over frames."""
    xo, zo, uo, h = morphosia(x, None if z is None else [z], None if u is None else [u], y, ve, C, Gu, Gut,
                              N_u, n_u, dK_u, delta, rho, nCGD, ve_max, nIter, kernel,
                              [dict(axis=0, dir=-1, scale=1.0, Tu=Tu, Tut=Tut)])
    return xo, zo[0], uo[0], h


def teva_duo_chain(x, z1, z2, u1, u2, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, rho, nCGD, ve_max, nIter=1,
                   kernel=(None, None, None), Tu1=None, Tu1t=None, Tu2=None, Tu2t=None):
    """bmTevaDuoMorphosia_chain (.../bmTevaDuoMorphosia_chain.m:6-242): two
    terms, F1 x_i = 0.5 (Tu1_i x_i - x_{i-1}) (:313-320),
    F2 x_i = 0.5 (Tu2_i x_i - x_{i+1}) (:322-329), adjoints with 0.5/HX
    (:340-359); regul of the unscaled differences (:254-271)."""
    z = None if z1 is None else [z1, z2]
    u = None if u1 is None else [u1, u2]
    xo, zo, uo, h = morphosia(x, z, u, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, rho, nCGD, ve_max, nIter, kernel,
                              [dict(axis=0, dir=-1, scale=0.5, Tu=Tu1, Tut=Tu1t),
                               dict(axis=0, dir=+1, scale=0.5, Tu=Tu2, Tut=Tu2t)])
    return xo, zo, uo, h


def teva_sheet(x, z1, z2, u1, u2, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, rho, nCGD, ve_max, nFr1, nIter=1,
               kernel=(None, None, None), Tu1=None, Tu1t=None, Tu2=None, Tu2t=None):
    """bmTevaMorphosia_sheet (.../bmTevaMorphosia_sheet.m:16-244): frames on
    an nFr1 x nFr2 cell array (lists in y(:) order), F1 along i (:322-332),
    F2 along j (:334-344), both x_{.} - previous, scale 1."""
    z = None if z1 is None else [z1, z2]
    u = None if u1 is None else [u1, u2]
    xo, zo, uo, h = morphosia(x, z, u, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, rho, nCGD, ve_max, nIter, kernel,
                              [dict(axis=0, dir=-1, scale=1.0, Tu=Tu1, Tut=Tu1t),
                               dict(axis=1, dir=-1, scale=1.0, Tu=Tu2, Tut=Tu2t)], nFr1=nFr1)
    return xo, zo, uo, h


def sensitiva_chain(x, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, nCGD, ve_max, nIter=1,
                    kernel=(None, None, None), Tu=None, Tut=None):
    """bmSensitivaMorphosia_chain (.../bmSensitivaMorphosia_chain.m:6-184):
    quadratic temporal penalty delta*HZ*||Tu_i x_i - x_{i-1}||^2, nCGD CG
    steps per outer iteration restarted from the current x (:91-171)."""
    xo, _, _, h = morphosia(x, None, None, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, None, nCGD, ve_max, nIter,
                            kernel, [dict(axis=0, dir=-1, scale=1.0, Tu=Tu, Tut=Tut)], l1=False)
    return xo, h


def sensitiva_sheet(x, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, nCGD, ve_max, nFr1, nIter=1,
                    kernel=(None, None, None), Tu1=None, Tu1t=None, Tu2=None, Tu2t=None):
    """bmSensitivaMorphosia_sheet (.../bmSensitivaMorphosia_sheet.m:6-187):
    the quadratic penalty along both cell-array axes; the first gradient of
    each outer iteration carries the penalty without delta (:118-119), the
    later ones with it (:164-165)."""
    xo, _, _, h = morphosia(x, None, None, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, None, nCGD, ve_max, nIter,
                            kernel, [dict(axis=0, dir=-1, scale=1.0, Tu=Tu1, Tut=Tu1t),
                                     dict(axis=1, dir=-1, scale=1.0, Tu=Tu2, Tut=Tu2t)],
                            l1=False, nFr1=nFr1, first_dagF_scaled=False)
    return xo, h


def sensitiva_duo_chain(x, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, nCGD, ve_max, nIter=1,
                        kernel=(None, None, None), Tu1=None, Tu1t=None, Tu2=None, Tu2t=None):
    """bmSensitivaDuoMorphosia_chain (.../bmSensitivaDuoMorphosia_chain.m:6-108):
    the duo chain's half-weighted differences (:150-162) and adjoints
    (:170-183) under the quadratic penalty; regul of the unscaled
    differences (:117-130)."""
    xo, _, _, h = morphosia(x, None, None, y, ve, C, Gu, Gut, N_u, n_u, dK_u, delta, None, nCGD, ve_max, nIter,
                            kernel, [dict(axis=0, dir=-1, scale=0.5, Tu=Tu1, Tut=Tu1t),
                                     dict(axis=0, dir=+1, scale=0.5, Tu=Tu2, Tut=Tu2t)], l1=False)
    return xo, h


def sensa_morphosia(x, y, ve, C, Gu, Gut, N_u, n_u, dK_u, nCGD, ve_max, nIter=1, kernel=(None, None, None),
                    Tu=None, Tut=None):
    """bmSensaMorphosia_chain (.../bmSensaMorphosia_chain.m:6-104): one image
    x, frame f observes M_f Tu_f x; CG on sum_f ||M_f Tu_f x - y_f||^2_HY,
    nCGD steps per iteration restarted nIter times.  Returns x and R per
    iteration (:93-99)."""
    eps = 10 * np.finfo(np.float32).eps
    nFr = len(y)
    KFC = kf(C, N_u, n_u, dK_u, *kernel)
    KFCc = kf_conj(np.conj(C), N_u, n_u, dK_u, *kernel)
    HX = hx_of(N_u, dK_u)
    HY = [np.minimum(np.broadcast_to(np.asarray(v, dtype=np.float64).reshape(-1, 1), yy.shape), ve_max)
          for v, yy in zip(ve, y)]
    Tu = [None] * nFr if Tu is None else [None if G is None else to_scipy(G) for G in Tu]
    Tut = [None] * nFr if Tut is None else [None if G is None else to_scipy(G) for G in Tut]
    deform = lambda G, a: a if G is None else (G @ a.reshape(-1, order='F')).reshape(a.shape, order='F')
    A = lambda v: [shanna(deform(Tu[i], v), Gu[i], KFC, N_u, n_u) for i in range(nFr)]
    Ah = lambda r: sum((1.0 / HX) * deform(Tut[i], nakatsha(HY[i] * r[i], Gut[i], KFCc, True, N_u, n_u))
                       for i in range(nFr))
    sq = lambda a: sum(float(np.real(np.vdot(ai, HY[i] * ai))) for i, ai in enumerate(a))
    x = np.array(x, dtype=np.complex128)
    hist = []
    for c in range(nIter):
        r = [yy - m for yy, m in zip(y, A(x))]
        dag = Ah(r)
        sqn = HX * float(np.real(np.vdot(dag, dag)))
        p = dag
        for j in range(1, nCGD + 1):
            if sqn < eps:
                break
            Ap = A(p)
            a = sqn / sq(Ap)
            x = x + a * p
            if j == nCGD:
                break
            r = [ri - a * api for ri, api in zip(r, Ap)]
            dag = Ah(r)
            sqn_new = HX * float(np.real(np.vdot(dag, dag)))
            b = sqn_new / sqn
            sqn = sqn_new
            p = dag + b * p
        hist.append(sq([yy - m for yy, m in zip(y, A(x))]))
    return x, hist


def gridder_prep(t, Dn, N_u, dK_u, single):
    """Grid coordinates of bmGridder_n2u(_leight).m:52-93: t / dK_u +
    fix(N_u / 2 + 1) per axis (1-based), Dn / prod(dK_u), points outside
    [1, N_u] removed.  single: the leight wrapper's float32 arithmetic."""
    dt = np.float32 if single else np.float64
    t = np.asarray(t, dtype=dt).copy()
    Dn = np.ravel(np.asarray(Dn, dtype=dt)).copy()
    keep = np.ones(t.shape[1], bool)
    for d in range(t.shape[0]):
        t[d] = t[d] / dt(dK_u[d])
        Dn = Dn / dt(dK_u[d])
        t[d] = t[d] + dt(np.fix(N_u[d] / 2 + 1))
        keep &= (t[d] >= 1) & (t[d] <= N_u[d])
    return t[:, keep], Dn[keep], keep


def _normalised_accumulate(y, cells, w, Nu):
    """b[cell] = sum w a / sum w with sum w < 1e-5 read as 1
    (bmGridder_n2u_mex.cpp:72-201, leight3_mex.cpp:116-358)."""
    nCh = y.shape[1]
    b = np.zeros((Nu, nCh), complex)
    nrm = np.zeros(Nu)
    for k in range(nCh):
        np.add.at(b[:, k], cells.ravel(), (w * y[:, k][:, None]).ravel())
    np.add.at(nrm, cells.ravel(), w.ravel())
    nrm[nrm < 1e-5] = 1.0
    return b / nrm[:, None]


def gridder_n2u(y, t, Dn, N_u, dK_u, kernelType=None, nWin=None, kernelParam=None):
    """bmGridder_n2u (src/gridding123/m/bmGridder_n2u.m:15-216 with
    bmGridder_n2u_mex.cpp:72-201), float64: neighbour offsets
    -fix(nWin/2)..fix(nWin/2) (odd nWin, floor shift 0.5) or -nWin/2-1..nWin/2
    (even), radial KB / Gauss weights times Dn, indices wrapped mod N_u,
    normalised accumulation.  y [nPt, nCh]; returns [prod(N_u), nCh]."""
    kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
    nWin = int(nWin)
    N_u = [int(a) for a in N_u]
    tg, Dn, keep = gridder_prep(t, Dn, N_u, dK_u, False)
    y = np.asarray(y)[keep]
    imDim, nPt = tg.shape
    if nWin % 2 == 0:
        r, shift = np.arange(-nWin // 2 - 1, nWin // 2 + 1), 0.0
    else:
        r, shift = np.arange(-(nWin // 2), nWin // 2 + 1), 0.5
    c = np.stack([g.reshape(-1, order='F') for g in np.meshgrid(*([r] * imDim), indexing='ij')])
    tf = np.floor(tg + shift)
    d = np.sqrt((((tg - tf)[:, None, :] - c[:, :, None]) ** 2).sum(axis=0))
    w = kernel_weight(d, kernelType, kernelParam) * Dn[None, :]
    n = tf[:, None, :] + c[:, :, None]
    lin = np.zeros(d.shape, np.int64)
    stride = 1
    for a in range(imDim):
        lin += np.mod(n[a] - 1, N_u[a]).astype(np.int64) * stride
        stride *= N_u[a]
    return _normalised_accumulate(y, lin.T, w.T, int(np.prod(N_u)))


def gridder_n2u_leight(y, t, Dn, N_u, dK_u, kernelType=None, nWin=None, kernelParam=None):
    """bmGridder_n2u_leight (src/gridding123/m/bmGridder_n2u_leight.m:15-141
    with leight3_mex.cpp:116-358), float64 after the wrapper's float32
    coordinates: Gaussian weights (sigma = kernelParam[0], normalised with
    the reference's 3.141592) on the box of nWin cells from trunc(t + 0.5)
    (odd nWin) or trunc(t) (even), clipped at the grid edge."""
    kernelType, nWin, kernelParam = kernel_defaults(kernelType, nWin, kernelParam)
    nWin = int(nWin)
    N_u = [int(a) for a in N_u]
    tg, Dn, keep = gridder_prep(t, Dn, N_u, dK_u, True)
    tg = tg.astype(np.float64)
    Dn = Dn.astype(np.float64)
    y = np.asarray(y)[keep]
    imDim, nPt = tg.shape
    shift, nh = (0.5, (nWin - 1) // 2) if nWin % 2 else (0.0, nWin // 2 - 1)
    sigma = float(np.float32(kernelParam[0]))
    g1 = 1.0 / (np.sqrt(2 * 3.141592) * sigma)
    g2 = -1.0 / (2 * sigma * sigma)
    u = np.arange(-nh, -nh + nWin)
    offs = np.stack([g.reshape(-1, order='F') for g in np.meshgrid(*([u] * imDim), indexing='ij')])  # [d, nNb]
    cf = np.trunc(tg + shift)
    rest = tg - cf
    cell0 = cf.astype(np.int64) - 1
    nb = cell0[:, None, :] + offs[:, :, None]                       # [d, nNb, nPt] 0-based
    inside = np.ones(nb.shape[1:], bool)
    for a in range(imDim):
        inside &= (nb[a] >= 0) & (nb[a] <= N_u[a] - 1)
    dsq = ((rest[:, None, :] - offs[:, :, None]) ** 2).sum(axis=0)
    w = g1 * np.exp(g2 * dsq) * Dn[None, :] * inside
    lin = np.zeros(dsq.shape, np.int64)
    stride = 1
    for a in range(imDim):
        lin += np.clip(nb[a], 0, N_u[a] - 1) * stride
        stride *= N_u[a]
    return _normalised_accumulate(y, lin.T, w.T, int(np.prod(N_u)))


def mathilda(y, t, v, N_u, n_u, dK_u, C=None, kernel=(None, None, None), leight_flag=True):
    """bmMathilda (src/fourier123/map_function/nonCartesian/bmMathilda.m:14-125):
    gridder, bmIDF (centred ifft times prod(N_u dK_u)), .* K of the requested
    kernel, crop, optional coil combination (bmCoilSense_pinv.m:6-12)."""
    d = len(N_u)
    grid = (gridder_n2u_leight if leight_flag else gridder_n2u)(y, t, v, N_u, dK_u, *kernel)
    K = bmK(N_u, dK_u, *kernel)
    g = _unflatF(grid, N_u)
    g = cifft(g, axes=range(d)) * (np.prod(N_u) * np.prod(dK_u))
    g = g * K[..., None]
    g = crop(g, N_u, n_u)
    if C is not None:
        g = (np.conj(C) * g).sum(-1) / (np.abs(C) ** 2).sum(-1)
    return g


def nasha(y, Gn, N_u, n_u, dK_u, C=None, kernel=(None, None, None)):
    """Gridded reconstruction (src/fourier123/map_function/nonCartesian/
    bmNasha.m:6-91): n2u with Gn, bmIDF (centred ifft times prod(N_u dK_u)),
    times K, crop, optional coil combine (bmCoilSense_pinv.m:6-12)."""
    d = len(N_u)
    K = bmK(N_u, dK_u, *kernel)
    g = _unflatF(spmv(Gn, y), N_u)
    g = cifft(g, axes=range(d)) * (np.prod(N_u) * np.prod(dK_u))
    g = g * K[..., None]
    g = crop(g, N_u, n_u)
    if C is not None:
        g = (np.conj(C) * g).sum(-1) / (np.abs(C) ** 2).sum(-1)
    return g
