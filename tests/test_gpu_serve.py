"""Serving loop on the GPU: every result is the same alone and in the mix
(batched with other requests of its protocol, overlapped with batches of
other protocols on other streams)."""
import numpy as np
import pytest

from problems import make_problem

pytestmark = pytest.mark.gpu


def _protocols(kind3d="3d32"):
    from monalisa_amd.serve import make_protocol
    out = {}
    for key, kind, nCh in (("2d", "2d_rect", 3), ("3d", kind3d, 4)):
        p = make_problem(kind, seed=21, nCh=nCh)
        out[key] = (p, make_protocol(p["t"], p["ve"], p["C"], p["N_u"], p["n_u"], p["dK_u"], p["kernel"]))
    return out


def _requests(protos, n, seed=5):
    import torch
    from monalisa_amd.nufft import Plan
    rng = np.random.default_rng(seed)
    ys = {}
    for key, (p, pr) in protos.items():
        P = Plan(p["N_u"], p["n_u"], p["dK_u"], p["nCh"], p["kernel"])
        P.set_coils(pr.C)
        P.set_frame(0, pr.G["Gu"], pr.G["Gut"])
        ys[key] = P.bmShanna(torch.from_numpy(p["x"]).cuda())
    reqs = []
    for i in range(n):
        key = "2d" if rng.random() < 0.5 else "3d"
        y0 = ys[key]
        noise = rng.normal(size=(2,) + tuple(y0.shape)).astype(np.float32)
        y = (y0 + 0.05 * y0.abs().mean() * torch.complex(torch.from_numpy(noise[0]), torch.from_numpy(noise[1])).cuda())
        reqs.append((key, y.contiguous()))
    return reqs


def _alone(protos, key, y, nCGD):
    import torch
    from monalisa_amd.nufft import Plan
    p, pr = protos[key]
    P = Plan(pr.N_u, pr.n_u, pr.dK_u, pr.nCh, pr.kernel)
    P.set_coils(pr.C)
    P.set_frame(0, pr.G["Gu"], pr.G["Gut"])
    x = torch.zeros(pr.nu, dtype=torch.complex64, device="cuda")
    P.bmSensa(x, y, pr.ve, nCGD, nIter=1)
    torch.cuda.synchronize()
    return x.cpu().numpy()


@pytest.mark.parametrize("kind", ["3d32", "3d32s"])
def test_sensa_batch_equals_sensa_alone(kind):
    """Dense frames (one adjoint row per cell) and sparse frames (cell-masked
    FFT passes) alike."""
    import torch
    from monalisa_amd.nufft import Plan
    print(f"[gpu] sensa_batch vs sensa {kind}", flush=True)
    protos = _protocols(kind3d=kind)
    reqs = [r for r in _requests(protos, 12, seed=6) if r[0] == "3d"][:3]
    p, pr = protos["3d"]
    P = Plan(pr.N_u, pr.n_u, pr.dK_u, pr.nCh, pr.kernel)
    P.set_coils(pr.C)
    for f in range(len(reqs)):
        P.set_frame(f, pr.G["Gu"], pr.G["Gut"])
    x = torch.zeros((len(reqs), pr.nu), dtype=torch.complex64, device="cuda")
    P.bmSensa_batch(x, [y for _, y in reqs], [pr.ve] * len(reqs), 6, nIter=2)
    got = x.cpu().numpy()
    for f, (key, y) in enumerate(reqs):
        Pa = Plan(pr.N_u, pr.n_u, pr.dK_u, pr.nCh, pr.kernel)
        Pa.set_coils(pr.C)
        Pa.set_frame(0, pr.G["Gu"], pr.G["Gut"])
        xa = torch.zeros(pr.nu, dtype=torch.complex64, device="cuda")
        Pa.bmSensa(xa, y, pr.ve, 6, nIter=2)
        np.testing.assert_array_equal(got[f], xa.cpu().numpy())


def test_served_results_equal_alone():
    import time
    from monalisa_amd.serve import GpuServer
    print("[gpu] serving mix", flush=True)
    protos = _protocols()
    reqs = _requests(protos, 16)
    # mixed CG step counts within each protocol: a batch never mixes them
    nCGDs = [5 if i % 4 else 3 for i in range(len(reqs))]
    alone = [_alone(protos, k, y, n) for (k, y), n in zip(reqs, nCGDs)]
    srv = GpuServer(hbm_budget=8 << 30, max_batch=3, workers=3)
    for key, (_, pr) in protos.items():
        srv.register(key, pr)
    srv.start()
    handles = []
    for i, (k, y) in enumerate(reqs):
        handles.append(srv.submit(k, y, nCGDs[i]))
        if i % 4 == 3:
            time.sleep(0.002)
    assert srv.drain(timeout=120)
    srv.stop()
    assert any(b > 1 for _, b in srv.batches), srv.batches
    for h, a in zip(handles, alone):
        assert h.error is None, h.error
        np.testing.assert_array_equal(h.x.cpu().numpy(), a)
