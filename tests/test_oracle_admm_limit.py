"""The oracle's ADMM-L1 solvers (bmSteva, bmTevaMorphosia_chain,
bmTevaDuoMorphosia_chain, bmTevaMorphosia_sheet) pinned by their limit.

The x step of each solver minimises 0.5 ||M x - y||^2_HY + (rho/2) HZ
||F x - (z - u)||^2 and the z step soft-thresholds F x + u at delta/rho, so
a fixed point minimises

    J(x) = 0.5 ||M x - y||^2_HY + delta HZ sum_t (|Re F_t x|_1 + |Im F_t x|_1).

The test computes that minimiser independently: on real and imaginary parts
the problem is min 0.5 x'Ax - b'x + lam ||K x||_1, whose dual is the
bounded least-squares problem min_{|v| <= lam} ||R^-T (b - K'v)||^2
(A = R'R), solved exactly by BVLS; then x = A^-1 (b - K'v).  A wrong
threshold, a dropped HZ, a wrong sign or shift in F or its adjoint moves the
ADMM limit away from it."""
import numpy as np
import pytest
from scipy.optimize import lsq_linear

from oracle import mona_oracle as O
from test_oracle_morphosia import _args, _dense_ops, _fields, _problem, _x0


def _real(A):
    return np.block([[A.real, -A.imag], [A.imag, A.real]])


def _l1_minimiser(M, hy, y, K, lam):
    """argmin 0.5 ||M x - y||^2_hy + lam (|Re K x|_1 + |Im K x|_1), complex x."""
    Mr, Kr = _real(M), _real(K)
    hyr = np.concatenate([hy, hy])
    yr = np.concatenate([y.real, y.imag])
    A = Mr.T @ (hyr[:, None] * Mr)
    b = Mr.T @ (hyr * yr)
    Rt_inv = np.linalg.inv(np.linalg.cholesky(A))          # R^-T, A = R'R
    res = lsq_linear(Rt_inv @ Kr.T, Rt_inv @ b, bounds=(-lam, lam), method="bvls", tol=1e-15, max_iter=10000)
    x = np.linalg.solve(A, b - Kr.T @ res.x)
    n = x.size // 2
    J = lambda c: 0.5 * np.sum(hyr * (Mr @ c - yr) ** 2) + lam * np.abs(Kr @ c).sum()
    return x[:n] + 1j * x[n:], J


def _stack(x):
    return np.concatenate([v.reshape(-1, order="F") for v in x])


def _difference(Tu, n1, n2, axis, nu):
    """Dense K with (K x)_f = Tu_f x_f - x_prev(f), prev = the cyclic
    predecessor of frame f = i + n1 j along i (axis 0) or j (axis 1), as
    private_F of bmTevaMorphosia_chain.m:266-272 (and its sheet / duo forms;
    the duo's second term takes the successor, axis 0 with n1 = nFr, given
    here as n2 = -1)."""
    nFr = n1 * abs(n2)
    K = np.zeros((nFr * nu, nFr * nu))
    for f in range(nFr):
        i, j = f % n1, f // n1
        if n2 < 0:
            g = (i + 1) % n1
        elif axis == 0:
            g = (i - 1) % n1 + n1 * j
        else:
            g = i + n1 * ((j - 1) % n2)
        K[f * nu:(f + 1) * nu, f * nu:(f + 1) * nu] = O.to_dense(Tu[f])
        K[f * nu:(f + 1) * nu, g * nu:(g + 1) * nu] -= np.eye(nu)
    return K


def test_steva_converges_to_l1_minimiser():
    q = _problem(1, n_u=(6, 6), seed=7)
    q["y"] = [1e4 * v for v in q["y"]]
    M, hy, y, _, HX = _dense_ops(q, [], 1)
    n_u, N_u, dK_u = q["n_u"], q["N_u"], q["dK_u"]
    dX = [(1.0 / dk) / n for dk, n in zip(dK_u, N_u)]
    nu = int(np.prod(n_u))
    cols = []
    for k in range(nu):
        e = np.zeros(nu)
        e[k] = 1
        e = e.reshape(n_u, order="F")
        cols.append(np.concatenate([O.back_gradient(e, dX, a).reshape(-1, order="F") for a in range(2)]))
    K = np.stack(cols, axis=1).astype(complex)
    delta, rho = 30.0, 0.02
    xs, J = _l1_minimiser(M, hy, y, K, delta * HX)
    # the L1 term is active: some gradients are exactly zero at the minimiser
    g = K @ xs
    inner = np.abs(g) > 0
    assert 0 < np.mean(np.abs(g[inner]) < 1e-9 * np.abs(g).max())
    x, z, u, hist = O.steva(np.zeros(n_u, complex), None, None, q["y"][0], q["ve"][0], q["C"], q["Gu"][0],
                            q["Gut"][0], N_u, n_u, dK_u, delta, rho, 20, q["ve_max"], 600, kernel=q["kernel"])
    got = x.reshape(-1, order="F")
    assert np.linalg.norm(got - xs) < 1e-6 * np.linalg.norm(xs)
    # z converges to F x and the reported TV to HZ ||F x||_1
    np.testing.assert_allclose(z.reshape(2, -1, order="F").reshape(-1), g, rtol=0, atol=1e-5 * np.abs(g).max())
    np.testing.assert_allclose(hist[-1][1], HX * (np.abs(g.real).sum() + np.abs(g.imag).sum()), rtol=1e-6)


@pytest.mark.parametrize("solver", ["chain", "duo", "sheet"])
def test_teva_morphosia_converges_to_l1_minimiser(solver):
    nFr = 4 if solver == "sheet" else 3
    q = _problem(nFr, n_u=(6, 6), seed=5)
    q["y"] = [1e4 * v for v in q["y"]]
    T1, T2 = _fields(q["n_u"], nFr, 14), _fields(q["n_u"], nFr, 15)
    Tu1, Tu1t = [G["Gu"] for G in T1], [G["Gut"] for G in T1]
    Tu2, Tu2t = [G["Gu"] for G in T2], [G["Gut"] for G in T2]
    delta, rho, nIter = 30.0, 0.3, 300
    if solver == "chain":
        n1 = nFr
        x, *_ = O.teva_chain(_x0(q), None, None, *_args(q), delta, rho, 20, q["ve_max"], nIter,
                             kernel=q["kernel"], Tu=Tu1, Tut=Tu1t)
    elif solver == "duo":
        n1 = nFr
        x, *_ = O.teva_duo_chain(_x0(q), None, None, None, None, *_args(q), delta, rho / 2, 20, q["ve_max"],
                                 nIter, kernel=q["kernel"], Tu1=Tu1, Tu1t=Tu1t, Tu2=Tu2, Tu2t=Tu2t)
    else:
        n1 = 2
        x, *_ = O.teva_sheet(_x0(q), None, None, None, None, *_args(q), delta, rho, 20, q["ve_max"], 2, nIter,
                             kernel=q["kernel"], Tu1=Tu1, Tu1t=Tu1t, Tu2=Tu2, Tu2t=Tu2t)
    M, hy, y, _, HX = _dense_ops(q, [], n1)
    nu = int(np.prod(q["n_u"]))
    if solver == "chain":
        K = _difference(Tu1, nFr, 1, 0, nu)
    elif solver == "duo":
        K = 0.5 * np.vstack([_difference(Tu1, nFr, 1, 0, nu), _difference(Tu2, nFr, -1, 0, nu)])
    else:
        K = np.vstack([_difference(Tu1, 2, nFr // 2, 0, nu), _difference(Tu2, 2, nFr // 2, 1, nu)])
    xs, _ = _l1_minimiser(M, hy, y, K.astype(complex), delta * HX)
    got = _stack(x)
    assert np.linalg.norm(got - xs) < 1e-5 * np.linalg.norm(xs)
