"""The oracle's Morphosia solvers (bmTevaMorphosia_chain, bmTevaDuoMorphosia_chain,
bmTevaMorphosia_sheet, bmSensitivaMorphosia_chain / _sheet) pinned by what
the mathematics fixes: the adjoint pairs of every difference term, the
reductions of one solver to another, and the quadratic solvers' convergence
to the dense least-squares solution."""
import numpy as np
import pytest

from monalisa_amd import synth
from oracle import mona_oracle as O


def _problem(nFr, n_u=(8, 8), nCh=2, seed=0):
    rng = np.random.default_rng(seed)
    n_u = list(n_u)
    N_u, dK_u = [2 * n for n in n_u], [0.5, 0.5]
    C = synth.coil_maps(n_u, nCh, rng)
    Cn = C.reshape(tuple(n_u) + (nCh,), order="F").astype(np.complex128)
    Gu, Gut, ys, ves = [], [], [], []
    KFC = O.kf(Cn, N_u, n_u, dK_u, "kaiser")
    for f in range(nFr):
        t = np.roll(synth.radial2d_golden(6 + 2 * f, 2 * n_u[0], k_max=n_u[0] / 2), f, axis=1)
        ve = synth.ve_radial2(t, 2 * n_u[0], 6 + 2 * f)
        G = O.traj2sparsemat(t, ve, N_u, dK_u, "kaiser")
        x = synth.ellipsoid_phantom(n_u, np.random.default_rng(10 + f)).reshape(n_u, order="F")
        y = O.shanna(x.astype(np.complex128), G["Gu"], KFC, N_u, n_u)
        y = y + 0.01 * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape))
        Gu.append(G["Gu"]); Gut.append(G["Gut"]); ys.append(y); ves.append(ve)
    return dict(n_u=n_u, N_u=N_u, dK_u=dK_u, C=Cn, Gu=Gu, Gut=Gut, y=ys, ve=ves, nFr=nFr,
                ve_max=float(max(v.max() for v in ves)), kernel=("kaiser", None, None))


def _fields(n_u, nFr, seed):
    from oracle.mona_oracle import deform2sparsemat
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nFr):
        v = rng.uniform(-0.8, 0.8, size=(2, int(np.prod(n_u))))
        out.append(deform2sparsemat(v, n_u, torus_flag=True))
    return out


def _args(q):
    return (q["y"], q["ve"], q["C"], q["Gu"], q["Gut"], q["N_u"], q["n_u"], q["dK_u"])


def _x0(q):
    return [np.zeros(q["n_u"], complex)] * q["nFr"]


@pytest.mark.parametrize("axis,d,n1", [(0, -1, 6), (0, +1, 6), (0, -1, 2), (1, -1, 2), (1, +1, 3)])
def test_difference_terms_are_adjoint(axis, d, n1):
    """<F x, z>_HZ = HX <x, dagF z> for every term (shift direction, cell-array
    axis, deformation and its transpose, the 0.5 scale)."""
    nFr, n_u = 6, [5, 4]
    T = _fields(n_u, nFr, 3)
    HX, HZ = 0.37, 0.37
    t = O.morph_terms([dict(axis=axis, dir=d, scale=0.5, Tu=[G["Gu"] for G in T], Tut=[G["Gut"] for G in T])],
                      nFr, n1, HX, HZ)[0]
    rng = np.random.default_rng(1)
    x = [rng.normal(size=n_u) + 1j * rng.normal(size=n_u) for _ in range(nFr)]
    z = [rng.normal(size=n_u) + 1j * rng.normal(size=n_u) for _ in range(nFr)]
    lhs = sum(np.vdot(a, HZ * b) for a, b in zip(t["F"](x), z))
    rhs = HX * sum(np.vdot(a, b) for a, b in zip(x, t["FT"](z)))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_duo_chain_with_identities_is_chain_at_half_rho():
    """Two half-scaled terms over the same differences: the duo chain's ADMM
    is the single chain's with rho/2 (z = 2 z1, u = 2 u1), so x agrees."""
    q = _problem(3)
    delta, rho, nCGD, nIter = 0.02, 0.4, 3, 3
    xd, zd, ud, hd = O.teva_duo_chain(_x0(q), None, None, None, None, *_args(q), delta, rho, nCGD, q["ve_max"],
                                      nIter, kernel=q["kernel"])
    xc, zc, uc, hc = O.teva_chain(_x0(q), None, None, *_args(q), delta, rho / 2, nCGD, q["ve_max"], nIter,
                                  kernel=q["kernel"])
    for a, b in zip(xd, xc):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * np.abs(b).max())
    for a, b in zip(zd[0], zc):
        np.testing.assert_allclose(2 * a, b, rtol=0, atol=1e-9 * np.abs(b).max())
    np.testing.assert_allclose([h[2] for h in hd], [2 * h[2] for h in hc], rtol=1e-9)


def test_sheet_with_one_column_is_chain():
    """nFr2 = 1 and no second deformation: the second term is identically 0
    and the sheet is the chain."""
    q = _problem(3)
    T = _fields(q["n_u"], 3, 5)
    Tu, Tut = [G["Gu"] for G in T], [G["Gut"] for G in T]
    delta, rho, nCGD, nIter = 0.02, 0.4, 3, 2
    xs, zs, us, hs = O.teva_sheet(_x0(q), None, None, None, None, *_args(q), delta, rho, nCGD, q["ve_max"], 3,
                                  nIter, kernel=q["kernel"], Tu1=Tu, Tu1t=Tut)
    xc, zc, uc, hc = O.teva_chain(_x0(q), None, None, *_args(q), delta, rho, nCGD, q["ve_max"], nIter,
                                  kernel=q["kernel"], Tu=Tu, Tut=Tut)
    for a, b in zip(xs, xc):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())
    assert all(np.all(v == 0) for v in zs[1])
    np.testing.assert_allclose(np.array(hs), np.array(hc), rtol=1e-12)


def test_sheet_is_symmetric_in_its_axes():
    """Transposing the cell array and swapping the two terms gives the same
    images (frame (i, j) of one is frame (j, i) of the other)."""
    q = _problem(6)
    T1, T2 = _fields(q["n_u"], 6, 7), _fields(q["n_u"], 6, 8)
    n1, n2 = 2, 3
    perm = [(f % n1) * n2 + f // n1 for f in range(6)]      # (i, j) -> index in the n2 x n1 sheet
    inv = np.argsort(perm)
    pick = lambda L: [L[k] for k in inv]
    delta, rho, nCGD, nIter = 0.02, 0.4, 3, 2
    xa, _, _, ha = O.teva_sheet(_x0(q), None, None, None, None, *_args(q), delta, rho, nCGD, q["ve_max"], n1,
                                nIter, kernel=q["kernel"], Tu1=[G["Gu"] for G in T1], Tu1t=[G["Gut"] for G in T1],
                                Tu2=[G["Gu"] for G in T2], Tu2t=[G["Gut"] for G in T2])
    args = (pick(q["y"]), pick(q["ve"]), q["C"], pick(q["Gu"]), pick(q["Gut"]), q["N_u"], q["n_u"], q["dK_u"])
    xb, _, _, hb = O.teva_sheet(_x0(q), None, None, None, None, *args, delta, rho, nCGD, q["ve_max"], n2,
                                nIter, kernel=q["kernel"], Tu1=pick([G["Gu"] for G in T2]),
                                Tu1t=pick([G["Gut"] for G in T2]), Tu2=pick([G["Gu"] for G in T1]),
                                Tu2t=pick([G["Gut"] for G in T1]))
    for f in range(6):
        np.testing.assert_allclose(xa[f], xb[perm[f]], rtol=0, atol=1e-10 * np.abs(xa[f]).max())
    np.testing.assert_allclose(np.array(ha), np.array(hb), rtol=1e-10)


def _dense_ops(q, terms, n1):
    """Dense M (stacked over frames, HY-weighted) and the term matrices."""
    nu = int(np.prod(q["n_u"]))
    nFr = q["nFr"]
    KFC = O.kf(q["C"], q["N_u"], q["n_u"], q["dK_u"], "kaiser")
    HX = O.hx_of(q["N_u"], q["dK_u"])
    blocks, hy, ys = [], [], []
    for f in range(nFr):
        cols = []
        for k in range(nu):
            e = np.zeros(nu, complex)
            e[k] = 1
            cols.append(O.shanna(e.reshape(q["n_u"], order="F"), q["Gu"][f], KFC, q["N_u"], q["n_u"]).reshape(-1))
        blocks.append(np.stack(cols, axis=1))
        h = np.minimum(np.broadcast_to(q["ve"][f].reshape(-1, 1), q["y"][f].shape), q["ve_max"])
        hy.append(h.reshape(-1))
        ys.append(q["y"][f].reshape(-1))
    from scipy.linalg import block_diag
    M = block_diag(*blocks)
    Fs = []
    for t in O.morph_terms(terms, nFr, n1, HX, HX):
        cols = []
        for k in range(nFr * nu):
            e = np.zeros(nFr * nu, complex)
            e[k] = 1
            xs = [e[f * nu:(f + 1) * nu].reshape(q["n_u"], order="F") for f in range(nFr)]
            cols.append(np.concatenate([v.reshape(-1, order="F") for v in t["diff"](xs)]))
        Fs.append(np.stack(cols, axis=1))
    return M, np.concatenate(hy), np.concatenate(ys), Fs, HX


@pytest.mark.parametrize("sheet", [False, True])
def test_sensitiva_converges_to_least_squares(sheet):
    """Restarted CG on the quadratic penalty reaches the minimiser of
    ||M x - y||^2_HY + delta HZ sum_t ||Tu x_f - x_shift||^2 (delta = 1 for
    the sheet, whose first gradients carry the penalty without delta)."""
    nFr = 4 if sheet else 3
    q = _problem(nFr, n_u=(6, 6), seed=2)
    # data scaled up so that the solvers' absolute stop (||dag||^2_HX <
    # 10 eps_single) does not end the iterations before convergence
    q["y"] = [1e4 * v for v in q["y"]]
    T1 = _fields(q["n_u"], nFr, 9)
    Tu1, Tu1t = [G["Gu"] for G in T1], [G["Gut"] for G in T1]
    if sheet:
        T2 = _fields(q["n_u"], nFr, 10)
        Tu2, Tu2t = [G["Gu"] for G in T2], [G["Gut"] for G in T2]
        delta, n1 = 1.0, 2
        terms = [dict(axis=0, dir=-1, Tu=Tu1, Tut=Tu1t), dict(axis=1, dir=-1, Tu=Tu2, Tut=Tu2t)]
        x, hist = O.sensitiva_sheet(_x0(q), *_args(q), delta, 12, q["ve_max"], n1, 25, kernel=q["kernel"],
                                    Tu1=Tu1, Tu1t=Tu1t, Tu2=Tu2, Tu2t=Tu2t)
    else:
        delta, n1 = 0.3, nFr
        terms = [dict(axis=0, dir=-1, Tu=Tu1, Tut=Tu1t)]
        x, hist = O.sensitiva_chain(_x0(q), *_args(q), delta, 12, q["ve_max"], 25, kernel=q["kernel"],
                                    Tu=Tu1, Tut=Tu1t)
    M, hy, y, Fs, HX = _dense_ops(q, terms, n1)
    A = M.conj().T @ (hy[:, None] * M) + delta * HX * sum(F.conj().T @ F for F in Fs)
    xs = np.linalg.solve(A, M.conj().T @ (hy * y))
    got = np.concatenate([v.reshape(-1, order="F") for v in x])
    assert np.linalg.norm(got - xs) < 1e-6 * np.linalg.norm(xs)
    # the last objective is the dense one's minimum
    r = M @ xs - y
    obj = 0.5 * np.real(np.vdot(r, hy * r)) + 0.5 * delta * HX * sum(np.linalg.norm(F @ xs) ** 2 for F in Fs)
    assert abs(hist[-1][0] - obj) < 1e-8 * obj


def test_sensitiva_duo_converges_to_least_squares():
    q = _problem(3, n_u=(6, 6), seed=4)
    q["y"] = [1e4 * v for v in q["y"]]
    T1, T2 = _fields(q["n_u"], 3, 11), _fields(q["n_u"], 3, 12)
    Tu1, Tu1t = [G["Gu"] for G in T1], [G["Gut"] for G in T1]
    Tu2, Tu2t = [G["Gu"] for G in T2], [G["Gut"] for G in T2]
    delta = 0.4
    x, hist = O.sensitiva_duo_chain(_x0(q), *_args(q), delta, 12, q["ve_max"], 25, kernel=q["kernel"],
                                    Tu1=Tu1, Tu1t=Tu1t, Tu2=Tu2, Tu2t=Tu2t)
    terms = [dict(axis=0, dir=-1, Tu=Tu1, Tut=Tu1t), dict(axis=0, dir=+1, Tu=Tu2, Tut=Tu2t)]
    M, hy, y, Fs, HX = _dense_ops(q, terms, 3)
    # F_t = 0.5 * difference: the penalty is delta HZ sum_t ||0.5 D_t x||^2
    A = M.conj().T @ (hy[:, None] * M) + delta * HX * 0.25 * sum(F.conj().T @ F for F in Fs)
    xs = np.linalg.solve(A, M.conj().T @ (hy * y))
    got = np.concatenate([v.reshape(-1, order="F") for v in x])
    assert np.linalg.norm(got - xs) < 1e-6 * np.linalg.norm(xs)
    # regul: squared norms of the unscaled differences
    r = M @ xs - y
    regul = HX * sum(np.linalg.norm(F @ xs) ** 2 for F in Fs)
    np.testing.assert_allclose(hist[-1][2], regul, rtol=1e-6)
    np.testing.assert_allclose(hist[-1][1], np.real(np.vdot(r, hy * r)), rtol=1e-6)


def test_sensa_morphosia_converges_to_least_squares():
    """One image seen through every frame's deformation: the minimiser of
    sum_f ||M_f Tu_f x - y_f||^2_HY."""
    q = _problem(3, n_u=(6, 6), seed=6)
    q["y"] = [1e4 * v for v in q["y"]]
    T = _fields(q["n_u"], 3, 13)
    Tu, Tut = [G["Gu"] for G in T], [G["Gut"] for G in T]
    Tu[1] = Tut[1] = None
    x, hist = O.sensa_morphosia(np.zeros(q["n_u"], complex), *_args(q), 12, q["ve_max"], 20, kernel=q["kernel"],
                                Tu=Tu, Tut=Tut)
    M, hy, y, _, HX = _dense_ops(q, [], 3)
    nu = 36
    D = [np.eye(nu) if G is None else O.to_dense(G) for G in Tu]
    A = M @ np.vstack(D)                                   # block rows M_f Tu_f
    xs = np.linalg.solve(A.conj().T @ (hy[:, None] * A), A.conj().T @ (hy * y))
    assert np.linalg.norm(x.reshape(-1, order="F") - xs) < 1e-6 * np.linalg.norm(xs)
    r = A @ xs - y
    np.testing.assert_allclose(hist[-1], np.real(np.vdot(r, hy * r)), rtol=1e-6)


def test_sensa_morphosia_of_one_identity_frame_is_sensa():
    q = _problem(1, seed=8)
    xm, hm = O.sensa_morphosia(np.zeros(q["n_u"], complex), *_args(q), 4, q["ve_max"], 2, kernel=q["kernel"])
    xs, hs = O.sensa(np.zeros(q["n_u"], complex), q["y"][0], q["ve"][0], q["C"], q["Gu"][0], q["Gut"][0], q["N_u"],
                     q["n_u"], q["dK_u"], 4, q["ve_max"], 2, kernel=q["kernel"])
    np.testing.assert_allclose(xm, xs, rtol=0, atol=1e-12 * np.abs(xs).max())
