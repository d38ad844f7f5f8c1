"""The 3-D benchmark workload at its full size (koosh ball 10,000 spokes x
256 samples, 2.56M points, N=128 on the 256^3 grid, 20 coils), checked
through properties that hold at any size: the GPU-built gridding matrices
equal the host-built ones (integer arrays exactly, weights to float
rounding), and the operator pair is linear and adjoint,
<M x, y> = <x, M^H y>."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    from monalisa_amd import synth
    from monalisa_amd.sparse import bmTraj2SparseMat
    n = 128
    t = synth.koosh_ball(25, 400, 256, k_max=n / 2)
    ve = synth.ve_radial3(t, 256, 10000)
    H = bmTraj2SparseMat(t, ve, [2 * n] * 3, [0.5] * 3, "kaiser")
    return dict(n=n, t=t, ve=ve, H=H)


def test_full_size_device_build_equals_host(c1):
    from monalisa_amd.sparse import bmTraj2SparseMat
    print("[gpu] full-size structure: device vs host build", flush=True)
    n = c1["n"]
    D = bmTraj2SparseMat(c1["t"], c1["ve"], [2 * n] * 3, [0.5] * 3, "kaiser", device=True)
    for k in ("Gu", "Gut"):
        h, d = c1["H"][k], D[k]
        assert (h.r_size, h.l_size, h.l_nJump) == (d.r_size, d.l_size, d.l_nJump), k
        if k == "Gu":
            assert h.r_jump.size == 69117381
        assert np.array_equal(d.r_jump, h.r_jump), k
        assert np.array_equal(d.r_nJump, h.r_nJump), k
        if h.l_jump is not None:
            assert np.array_equal(d.l_jump, h.l_jump), k
        np.testing.assert_allclose(d.m_val, h.m_val, rtol=3e-7, atol=0)


def test_full_size_operators_are_linear_and_adjoint(c1):
    import torch
    from monalisa_amd import synth
    from monalisa_amd.nufft import Plan
    print("[gpu] full-size operators: linearity and adjointness", flush=True)
    n, nCh = c1["n"], 20
    n_u, N_u = [n] * 3, [2 * n] * 3
    rng = np.random.default_rng(7)
    P = Plan(N_u, n_u, [0.5] * 3, nCh, "kaiser")
    P.set_coils(torch.from_numpy(synth.coil_maps(n_u, nCh, rng)).cuda())
    P.set_frame(0, c1["H"]["Gu"], c1["H"]["Gut"])
    nu, nPt = n ** 3, P.nPt()
    gen = torch.Generator(device="cuda").manual_seed(3)
    cx = lambda *s: torch.complex(torch.randn(*s, generator=gen, device="cuda"),
                                  torch.randn(*s, generator=gen, device="cuda"))
    x1, x2, y = cx(nu), cx(nu), cx(nPt, nCh)
    a, b = 0.75 - 0.5j, -1.25 + 2.0j
    Mx1, Mx2 = P.bmShanna(x1), P.bmShanna(x2)
    Mc = P.bmShanna((a * x1 + b * x2).contiguous())
    lin = torch.linalg.vector_norm(Mc - (a * Mx1 + b * Mx2)) / torch.linalg.vector_norm(Mc)
    assert lin.item() < 1e-5
    Mhy = P.bmNakatsha(y)
    lhs = torch.vdot(y.reshape(-1).to(torch.complex128), Mx1.reshape(-1).to(torch.complex128))
    rhs = torch.vdot(Mhy.to(torch.complex128), x1.to(torch.complex128))
    assert (abs(lhs - rhs) / abs(lhs)).item() < 1e-5


def test_full_size_noise_free_cg_residuals(c1):
    """bmSensa at the full workload on noise-free data y = M x_phantom, from
    x0 = 0 with nCGD = 1 .. 20 (bmSensa.m:39-85): CG on the normal
    equations minimises <r, HY r> over a growing Krylov space, so the final
    residual of k steps never increases with k.  The image error against
    the phantom is printed beside the one of 30 dB noisy data."""
    import math
    import torch
    from monalisa_amd import synth
    from monalisa_amd.nufft import Plan
    print("[gpu] full-size CG-SENSE on noise-free data, nCGD = 1..20", flush=True)
    n, nCh = c1["n"], 20
    n_u, N_u = [n] * 3, [2 * n] * 3
    rng = np.random.default_rng(1234)
    C = synth.coil_maps(n_u, nCh, rng)
    ph = synth.ellipsoid_phantom(n_u, rng)
    P = Plan(N_u, n_u, [0.5] * 3, nCh, "kaiser")
    P.set_coils(torch.from_numpy(C).cuda())
    P.set_frame(0, c1["H"]["Gu"], c1["H"]["Gut"])
    y = P.bmShanna(torch.from_numpy(ph).cuda())
    ve = torch.from_numpy(c1["ve"].astype(np.float32)).cuda()
    yy = float(torch.sum(ve[:, None] * torch.abs(y) ** 2).item())   # <y, HY y> (ve_max = max ve)

    def err(x):
        v = x.cpu().numpy()
        fit = np.vdot(v, ph) / np.vdot(v, v)
        return float(np.linalg.norm(fit * v - ph) / np.linalg.norm(ph))

    res, errs = [], []
    for k in range(1, 21):
        x = torch.zeros(n ** 3, dtype=torch.complex64, device="cuda")
        res.append(P.bmSensa(x, y, ve, k)[0])
        errs.append(err(x))
    rel = [r / yy for r in res]
    print("noise-free relative residual <r,HY r>/<y,HY y> for k = 1..20:", " ".join(f"{v:.3e}" for v in rel))
    print("noise-free image error vs phantom for k = 1..20:", " ".join(f"{v:.3f}" for v in errs))
    for k in range(1, 20):
        assert res[k] <= res[k - 1] * (1 + 1e-5), (k, res[k - 1], res[k])
    assert rel[-1] < 0.1 * rel[0]
    # the same 20 steps on 30 dB noisy data (the benchmark's input)
    gen = torch.Generator(device="cuda").manual_seed(1234)
    rms = torch.sqrt(torch.mean(torch.abs(y) ** 2)).item()
    sigma = rms * 10 ** (-30 / 20) / math.sqrt(2)
    yn = (y + torch.complex(torch.randn(y.shape, generator=gen, device="cuda"),
                            torch.randn(y.shape, generator=gen, device="cuda")) * sigma).contiguous()
    x = torch.zeros(n ** 3, dtype=torch.complex64, device="cuda")
    P.bmSensa(x, yn, ve, 20)
    print(f"image error vs phantom after 20 steps: noise-free {errs[-1]:.3f}, 30 dB noise {err(x):.3f}")
    assert errs[-1] < errs[0]
