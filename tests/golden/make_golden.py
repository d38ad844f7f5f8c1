"""Writes tests/golden/spmv_golden.npz: seeded jump-encoded matrices and
vectors with the products computed by the reference's own kernel
src/sparseMat/cpp/bmSparseMat_cC_oBlock_omp.cpp, built from the reference
sources into oracle/_ref by oracle/Makefile.  Run here, where the reference
is mounted:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_libs as OL  # noqa: E402
from oracle import mona_oracle as O  # noqa: E402


def cases():
    rng = np.random.default_rng(20260101)
    out = []
    # random structures: plain, squeezed, long rows, many vectors
    out.append(("rand_plain", OL.random_matrix(rng, 40, 33, 9), 3))
    out.append(("rand_squeezed", OL.random_matrix(rng, 64, 50, 12, squeeze=True), 2))
    out.append(("rand_long", OL.random_matrix(rng, 8, 300, 200), 1))
    out.append(("rand_nvec5", OL.random_matrix(rng, 30, 30, 5), 5))
    # gridding matrices of a small 2-D radial trajectory (Gu and Gut)
    ang =np.repeat(np.arange(6) * 2.39996, 16)
    r = np.tile(np.linspace(-7.5, 7.0, 16), 6)
    t = np.stack([r * np.cos(ang), r * np.sin(ang)], 0) * 0.5
    G = O.traj2sparsemat(t, np.ones(96), [16, 16], [0.5, 0.5], "kaiser")
    for k in ("Gu", "Gut", "Gn"):
        g = G[k]
        M = dict(r_size=g["r_size"], l_size=g["l_size"], r_nJump=g["r_nJump"], r_jump=g["r_jump"],
                 m_val=g["m_val"], l_nJump=g["l_nJump"], l_jump=g["l_jump"])
        out.append(("radial2d_" + k, M, 2))
    return out, rng


def main():
    ref = OL.ref()
    if ref is None:
        raise SystemExit("oracle/_ref not built: run make -C oracle ref (needs the reference sources)")
    cs, rng = cases()
    data = {}
    for name, M, n_vec in cs:
        vr = rng.normal(size=(n_vec, M["r_size"])).astype(np.float32)
        vi = rng.normal(size=(n_vec, M["r_size"])).astype(np.float32)
        wr, wi = OL.call_cC(ref.ref_cC_oBlock_omp, M, vr, vi)
        p = name + "__"
        for k in ("r_jump", "r_nJump", "m_val"):
            data[p + k] = M[k]
        if M["l_jump"] is not None:
            data[p + "l_jump"] = M["l_jump"]
        data[p + "sizes"] = np.array([M["r_size"], M["l_size"], M["l_nJump"], n_vec], np.int64)
        data[p + "v_re"], data[p + "v_im"] = vr, vi
        data[p + "w_re"], data[p + "w_im"] = wr, wi
    data["case_names"] = np.array([c[0] for c in cs])
    np.savez_compressed(os.path.join(HERE, "spmv_golden.npz"), **data)
    print("wrote", len(cs), "cases")


if __name__ == "__main__":
    main()
