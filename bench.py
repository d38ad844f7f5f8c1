"""Benchmark: 3-D golden-angle koosh-ball CG-SENSE, N=128 (256^3 grid),
10,000 spokes x 256 samples (2.56M points), 20 coils, KB width 3, 20 CG
iterations (bmSensa with nIter=1, nCGD=20), FP32 vectors with FP64 CG
scalars, on one MI355X.

Prints ONE JSON line (rank 0).  value = CG-SENSE reconstructions per second
with all inputs resident in HBM.  Extra fields: NUFFT pairs/s, the roofline
of the dominant gridding kernel (HIP events around every launch in the timed
region) with the FFT op's beside it (roofline.fft_op: one transform, all
passes, the largest share of a recon), and the CPU baseline (the reference's own sparse kernel built from
its sources, plus scipy FFT standing in for MATLAB's fft) timed on this
host on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # the workload the metric is quoted on
    "c1": dict(n=128, nCh=20, nSeg=25, nShot=400, nSample=256, nCGD=20),
    # small shape for quick checks
    "small": dict(n=64, nCh=8, nSeg=10, nShot=200, nSample=128, nCGD=10),
    # C1 with bf16 storage: CG-SENSE by iterative refinement to a relative
    # normal-equation residual < 1e-5 (consistent data; see bench_mixed)
    "c3": dict(n=128, nCh=20, nSeg=25, nShot=400, nSample=256, nCGD=20, mixed=True, nInner=10, nOuter=40,
               tol=1e-5),
    "c3small": dict(n=32, nCh=8, nSeg=10, nShot=100, nSample=64, nCGD=10, mixed=True, nInner=10, nOuter=40,
                    tol=1e-5),
    # recon serving: Poisson arrivals over a 2D / 3D protocol mix, 10-step CG-SENSE each
    "c4": dict(serve=True, nCGD=10, n_sat=96, n_poisson=240, load=0.8, max_batch=4, workers=3,
               protocols=[("2d128", 2, 128, 4, 256, None), ("2d256", 2, 256, 8, 512, None),
                          ("2d384", 2, 384, 2, 384, None), ("3d64", 3, 64, 8, 2000, (20, 100)),
                          ("3d96", 3, 96, 16, 5000, (25, 200)), ("3d128", 3, 128, 20, 10000, (25, 400))]),
    "c4small": dict(serve=True, nCGD=10, n_sat=24, n_poisson=48, load=0.8, max_batch=3, workers=3,
                    protocols=[("2d128", 2, 128, 4, 256, None), ("3d48", 3, 48, 8, 1000, (10, 100))]),
    # batched cardiac bins, temporal-TV ADMM (30 outer x 5 CG), 20 Dirichlet(2) bins
    "c2": dict(n=96, nCh=16, nSeg=20, nShot=400, nSample=192, nBin=20, alpha=2.0, nIter=30, nCGD=5),
    "c2small": dict(n=48, nCh=4, nSeg=10, nShot=100, nSample=96, nBin=6, alpha=2.0, nIter=4, nCGD=3),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(cfg, seed=1234):
    import torch
    from monalisa_amd import synth
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    rng = np.random.default_rng(seed)
    n, nCh = cfg["n"], cfg["nCh"]
    n_u, N_u, dK_u = [n] * 3, [2 * n] * 3, [0.5] * 3
    nSpoke = cfg["nSeg"] * cfg["nShot"]
    t0 = time.time()
    t = synth.koosh_ball(cfg["nSeg"], cfg["nShot"], cfg["nSample"], k_max=n / 2)
    ve = synth.ve_radial3(t, cfg["nSample"], nSpoke)
    t1 = time.time()
    G = bmTraj2SparseMat(t, ve, N_u, dK_u, "kaiser")
    t_host = time.time() - t1
    t1 = time.time()
    bmTraj2SparseMat(t, ve, N_u, dK_u, "kaiser", device=True)
    torch.cuda.synchronize()
    t_dev = time.time() - t1
    log(f"[bench] trajectory + Gu/Gut ({t.shape[1]} points, nnz {G['Gu'].nnz}) in {time.time() - t0:.1f}s "
        f"(matrices: host {t_host:.2f}s, device {t_dev:.2f}s)")
    C = synth.coil_maps(n_u, nCh, rng)
    phantom = synth.ellipsoid_phantom(n_u, rng)
    P = Plan(N_u, n_u, dK_u, nCh, "kaiser")
    Cd = torch.from_numpy(C).cuda()
    P.set_coils(Cd)
    P.set_frame(0, G["Gu"], G["Gut"])
    xd = torch.from_numpy(phantom).cuda()
    y = P.bmShanna(xd)
    # complex Gaussian noise at SNR 30 dB, seeded on the device
    gen = torch.Generator(device="cuda").manual_seed(seed)
    rms = torch.sqrt(torch.mean(torch.abs(y) ** 2)).item()
    sigma = rms * 10 ** (-30 / 20) / math.sqrt(2)
    noise = torch.complex(torch.randn(y.shape, generator=gen, device="cuda"),
                          torch.randn(y.shape, generator=gen, device="cuda")) * sigma
    y = (y + noise).contiguous()
    ve_d = torch.from_numpy(ve.astype(np.float32)).cuda()
    torch.cuda.synchronize()
    return dict(P=P, G=G, y=y, ve=ve_d, ve_np=ve, C=C, n_u=n_u, N_u=N_u, dK_u=dK_u, nCh=nCh,
                phantom=phantom, xd=xd, P_pts=t.shape[1], build_s={"host": t_host, "device": t_dev})


def build_binned_problem(cfg, seed=4321):
    """Koosh-ball spokes split into bins with Dirichlet(alpha) counts; bin f
    images a phantom whose inner structure varies with f (a cardiac-like
    cycle); y_f = forward(x_f) + 30 dB noise.  x0 is 3 CG-SENSE iterations
    per bin; delta = 0.01 max|x0|, rho = 10 delta."""
    import torch
    from monalisa_amd import synth
    from monalisa_amd.nufft import Plan
    from monalisa_amd.sparse import bmTraj2SparseMat
    rng = np.random.default_rng(seed)
    n, nCh, nSample, nBin = cfg["n"], cfg["nCh"], cfg["nSample"], cfg["nBin"]
    n_u, N_u, dK_u = [n] * 3, [2 * n] * 3, [0.5] * 3
    nSpoke = cfg["nSeg"] * cfg["nShot"]
    t0 = time.time()
    t = synth.koosh_ball(cfg["nSeg"], cfg["nShot"], nSample, k_max=n / 2).reshape(3, nSample, nSpoke, order="F")
    bins = synth.dirichlet_bins(nSpoke, nBin, cfg["alpha"], rng)
    C = synth.coil_maps(n_u, nCh, rng)
    P = Plan(N_u, n_u, dK_u, nCh, "kaiser")
    P.set_coils(torch.from_numpy(C).cuda())
    base = synth.ellipsoid_phantom(n_u, rng)
    moving = synth.ellipsoid_phantom(n_u, rng, n_ellipsoids=3) - synth.ellipsoid_phantom(n_u, rng, n_ellipsoids=0)
    Gs, ves, ys, xs = [], [], [], []
    gen = torch.Generator(device="cuda").manual_seed(seed)
    t_mat = 0.0
    for f, sp in enumerate(bins):
        tb = np.ascontiguousarray(t[:, :, sp].reshape(3, -1, order="F"))
        ve = synth.ve_radial3(tb, nSample, len(sp))
        tm = time.time()
        G = bmTraj2SparseMat(tb, ve, N_u, dK_u, "kaiser", device=True)
        t_mat += time.time() - tm
        P.set_frame(f, G["Gu"], G["Gut"])
        xf = (base + 0.3 * np.sin(2 * np.pi * f / nBin) * moving).astype(np.complex64)
        y = P.bmShanna(torch.from_numpy(xf).cuda(), f=f)
        rms = torch.sqrt(torch.mean(torch.abs(y) ** 2)).item()
        sigma = rms * 10 ** (-30 / 20) / math.sqrt(2)
        y = (y + torch.complex(torch.randn(y.shape, generator=gen, device="cuda"),
                               torch.randn(y.shape, generator=gen, device="cuda")) * sigma).contiguous()
        Gs.append(G); ves.append(torch.from_numpy(ve.astype(np.float32)).cuda()); ys.append(y); xs.append(xf)
    ve_max = float(max(v.max().item() for v in ves))
    x0 = torch.zeros((nBin, P.nu), dtype=torch.complex64, device="cuda")
    for f in range(nBin):
        xf = torch.zeros(P.nu, dtype=torch.complex64, device="cuda")
        P.bmSensa(xf, ys[f], ves[f], 3, ve_max=ve_max, f=f)
        x0[f] = xf
    delta = 0.01 * float(torch.abs(x0).max().item())
    torch.cuda.synchronize()
    log(f"[bench] {nBin} bins, points per bin {min(len(b) for b in bins) * nSample}.."
        f"{max(len(b) for b in bins) * nSample}, built in {time.time() - t0:.1f}s")
    return dict(P=P, G=Gs, ys=ys, ves=ves, ve_max=ve_max, x0=x0, delta=delta, rho=10 * delta, xs=xs,
                n_u=n_u, N_u=N_u, nCh=nCh, nBin=nBin, bins=bins, build_s={"device_all_bins": t_mat})


def bench_binned(args, cfg):
    """Temporal-TV ADMM over all bins: the batched form (one launch per
    gridding op over every bin, HIP-graph-replayed iterations, autotuned FFT
    tile) against the per-bin eager form, same inputs."""
    import torch
    prob = build_binned_problem(cfg)
    P = prob["P"]

    def recon():
        x = prob["x0"].clone()
        P.bmTevaMorphosia_chain(x, None, None, prob["ys"], prob["ves"], prob["delta"], prob["rho"], cfg["nCGD"],
                                prob["ve_max"], nIter=cfg["nIter"])
        return x

    def timed(steps, warmup):
        for _ in range(warmup):
            recon()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            x = recon()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, x

    # the tuned FFT variant and adjoint-gridding group sizes serve both
    # forms, so they compute the same sums
    tile, groups = P.autotune(0, groups=True)
    P.set_options(batch_frames=False, graph=False)
    s_plain, x_plain = timed(max(1, args.steps // 2), args.warmup)
    P.set_options(batch_frames=True, graph=True)
    s_batch, x_batch = timed(args.steps, args.warmup)
    a, b = x_batch.cpu().numpy(), x_plain.cpu().numpy()
    rel = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    xs = np.stack(prob["xs"])
    fit = np.vdot(a.ravel(), xs.ravel()) / np.vdot(a.ravel(), a.ravel())
    img_err = float(np.linalg.norm(fit * a - xs) / np.linalg.norm(xs))
    nb = prob["nBin"]
    out = {
        "metric": "temporal-TV ADMM recons/s (all bins) at 3D N=%d %dch, %d bins" % (cfg["n"], cfg["nCh"], nb),
        "value": 1.0 / s_batch,
        "unit": "recons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_batch * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 vectors, fp64 CG/ADMM scalars)",
        "data": "synthetic (seeded koosh-ball spokes in Dirichlet bins, moving ellipsoid phantom, 30 dB noise)",
        "config": {"workload": f"binned 3D temporal-TV ADMM n_u={cfg['n']}^3 N_u={2 * cfg['n']}^3, "
                               f"{cfg['nSeg'] * cfg['nShot']} spokes x {cfg['nSample']} samples in {nb} "
                               f"Dirichlet({cfg['alpha']}) bins, {cfg['nCh']} coils, "
                               f"{cfg['nIter']} outer x {cfg['nCGD']} CG",
                   "parallelism": "single GPU"},
        "per_bin_eager_recons_per_s": 1.0 / s_plain,
        "batched_over_per_bin": s_plain / s_batch,
        "batched_vs_per_bin_rel_l2": rel,
        "image_rel_err_vs_phantom": img_err,
        "gridding_matrix_build_s": prob["build_s"],
        "fft_tile_autotuned": tile,
        "adjoint_groups_autotuned": {"stage_kib": groups[0], "rows": groups[1]},
        "points_per_bin": [len(b) * cfg["nSample"] for b in prob["bins"]],
    }
    print(json.dumps(out), flush=True)


def _fp64_operators(P, G, x, yk, n_u, N_u, nCh):
    """Float64 host products of the forward operator at x and the adjoint at
    yk with the plan's own gridding matrices and coil weights (KFC and
    KFC_conj read back from the plan): Gu F Z (x .* KFC) and
    sum_c KFC_conj .* crop(F^-1 Gut yk) (bmShanna_MATLAB.m:6-56,
    bmNakatsha_MATLAB.m:6-60), one coil at a time."""
    import scipy.fft as sfft
    from scipy import sparse

    def csr(M):
        col, rowptr, rows = M.decode()
        rid = np.repeat(rows, np.diff(rowptr))
        return sparse.csr_matrix((M.m_val.astype(np.float64), (rid, col)), shape=(M.l_size, M.r_size))

    Gu, Gut = csr(G["Gu"]), csr(G["Gut"])
    kfc, kfcc = P.KFC()
    kfc, kfcc = kfc.cpu().numpy().astype(np.complex128), kfcc.cpu().numpy().astype(np.complex128)
    x = x.cpu().numpy().astype(np.complex128)
    yk = yk.cpu().numpy().astype(np.complex128)
    shp = tuple(N_u[::-1])               # C order: x fastest
    sl = tuple(slice(N // 2 - n // 2, N // 2 - n // 2 + n) for N, n in zip(N_u[::-1], n_u[::-1]))
    w = os.cpu_count() or 8
    yf = np.zeros((G["Gu"].l_size, nCh), np.complex128)
    xa = np.zeros(int(np.prod(n_u)), np.complex128)
    for c in range(nCh):
        z = np.zeros(shp, np.complex128)
        z[sl] = (x * kfc[:, c]).reshape(tuple(n_u[::-1]))
        g = sfft.fftshift(sfft.fftn(sfft.ifftshift(z), workers=w))
        yf[:, c] = Gu @ g.reshape(-1)
        g = (Gut @ yk[:, c]).reshape(shp)
        g = sfft.fftshift(sfft.ifftn(sfft.ifftshift(g), workers=w))
        xa += kfcc[:, c] * g[sl].reshape(-1)
    return yf, xa


def bench_mixed(args, cfg):
    """Mixed precision at the full workload (3D N=128, 20 coils).

    * Target (the configuration's): iterative refinement until the relative
      normal-equation residual is below 1e-5, CG with bf16-storage products
      and float residual replacement every nInner steps; timed beside the
      same refinement with float products.  The noisy benchmark data make
      these normal equations converge far too slowly for 1e-5 (the float
      refinement is still above 5e-4 after 6,000 steps), so the target is
      run on consistent data, y = M x_phantom.
    * Noisy data: both refinements after the same fixed number of steps.
    * bf16 storage without residual replacement: where it stalls.
    * Forward and adjoint operators (FFT + gridding, bf16 and float storage)
      against float64 host products with the same Gu, Gut and coil weights."""
    import torch
    prob = build_problem(cfg)
    P, yn, ve = prob["P"], prob["y"], prob["ve"]
    nu = int(np.prod(prob["n_u"]))
    y0 = P.bmShanna(prob["xd"])          # consistent data
    tol = cfg["tol"]

    def refined(y, bf16, tol, nOuter=None, nInner=None):
        x = torch.zeros(nu, dtype=torch.complex64, device="cuda")
        h = P.sensa_refined(x, y, ve, nInner or cfg["nInner"], nOuter or cfg["nOuter"], tol, bf16=bf16)
        return x, h

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            x, h = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, x, h

    s16, x16, h16 = timed(lambda: refined(y0, True, tol), args.steps)
    s32, x32, h32 = timed(lambda: refined(y0, False, tol), max(1, args.steps // 2))
    # bf16 products with no residual replacement: the level it stalls at
    _, hplain = refined(y0, True, 0.0, nOuter=1, nInner=h16.steps + 100)
    # noisy benchmark data: the same fixed budget of steps for both
    budget = cfg["nInner"] * 10
    _, hn16 = refined(yn, True, 0.0, nOuter=10)
    _, hn32 = refined(yn, False, 0.0, nOuter=10)
    a, b = x16.cpu().numpy(), x32.cpu().numpy()
    # operators against float64 host products (same matrices and coil weights)
    xo = prob["xd"]
    gen = torch.Generator(device="cuda").manual_seed(5)
    yk = torch.complex(torch.randn(yn.shape, generator=gen, device="cuda"),
                       torch.randn(yn.shape, generator=gen, device="cuda")).contiguous()
    t0 = time.time()
    yf64, xa64 = _fp64_operators(P, prob["G"], xo, yk, prob["n_u"], prob["N_u"], prob["nCh"])
    log(f"[bench] float64 host operators in {time.time() - t0:.1f}s")

    def err(g, r):
        g = g.cpu().numpy().astype(np.complex128)
        return float(np.linalg.norm(g.reshape(r.shape) - r) / np.linalg.norm(r))

    ops = {"forward_bf16": err(P.bmShanna_bf16(xo), yf64), "forward_fp32": err(P.bmShanna(xo), yf64),
           "adjoint_bf16": err(P.bmNakatsha_bf16(yk), xa64), "adjoint_fp32": err(P.bmNakatsha(yk), xa64)}
    ph = prob["phantom"]

    def img_err(v):
        fit = np.vdot(v, ph) / np.vdot(v, v)
        return float(np.linalg.norm(fit * v - ph) / np.linalg.norm(ph))

    out = {
        "metric": "mixed-precision CG-SENSE recons/s at 3D N=%d %dch (bf16 storage, float residual "
                  "replacement until relative residual < %g)" % (cfg["n"], cfg["nCh"], tol),
        "value": 1.0 / s16,
        "unit": "recons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s16 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 storage / fp32 arithmetic (fp64 CG scalars, fp32 residual replacement)",
        "data": "synthetic (seeded koosh-ball trajectory, Gaussian-blob coil maps, ellipsoid phantom; "
                "consistent data y = M x for the 1e-5 target, 30 dB noise for the fixed-budget comparison)",
        "config": {"workload": f"3D koosh-ball CG-SENSE n_u={cfg['n']}^3 N_u={2 * cfg['n']}^3 "
                               f"{cfg['nSeg'] * cfg['nShot']} spokes x {cfg['nSample']} samples, {cfg['nCh']} coils; "
                               f"refinement to relative normal-equation residual < {tol}, float residual "
                               f"replacement every {cfg['nInner']} steps",
                   "parallelism": "single GPU"},
        "target_rel_residual": tol,
        "cg_steps_bf16": h16.steps,
        "cg_steps_fp32_refined": h32.steps,
        "final_rel_residual_bf16": h16[-1],
        "final_rel_residual_fp32_refined": h32[-1],
        "fp32_refined_recons_per_s": 1.0 / s32,
        "bf16_without_replacement_rel_residual": {"steps": hplain.steps, "value": hplain[-1]},
        "noisy_data_fixed_budget": {"steps": budget, "rel_residual_bf16": hn16[-1], "rel_residual_fp32": hn32[-1]},
        "image_rel_l2_bf16_vs_fp32_refined": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
        "operator_rel_err_vs_fp64_host": ops,
        "image_rel_err_vs_phantom": {"bf16_refined": img_err(a), "fp32_refined": img_err(b)},
    }
    print(json.dumps(out), flush=True)


def bench_serving(args, cfg, quiet=False):
    """Recon serving: each protocol (geometry, trajectory, coil maps) is
    registered once; requests (y = forward(phantom) + seeded noise) run a
    10-step CG-SENSE.  Served form: continuous batching (same-protocol
    batches, concurrent batches on separate streams, HBM-budget admission).
    Plain per-request form: the same loop with one worker and batches of
    one (FIFO, no overlap).  Saturated throughput: all requests queued at
    once.  Latency: Poisson arrivals at `load` x the plain form's capacity."""
    import torch
    from monalisa_amd import synth
    from monalisa_amd.nufft import Plan
    from monalisa_amd.serve import GpuServer, make_protocol, percentiles
    rng = np.random.default_rng(77)
    protos, y0 = {}, {}
    t0 = time.time()
    for key, dim, n, nCh, nSpoke, segshot in cfg["protocols"]:
        n_u, N_u, dK_u = [n] * dim, [2 * n] * dim, [0.5] * dim
        nSample = 2 * n
        if dim == 2:
            t = synth.radial2d_golden(nSpoke, nSample, k_max=n / 2)
            ve = synth.ve_radial2(t, nSample, nSpoke)
        else:
            t = synth.koosh_ball(segshot[0], segshot[1], nSample, k_max=n / 2)
            ve = synth.ve_radial3(t, nSample, nSpoke)
        C = synth.coil_maps(n_u, nCh, rng)
        pr = make_protocol(t, ve, C, N_u, n_u, dK_u, "kaiser")
        P = Plan(N_u, n_u, dK_u, nCh, "kaiser")
        P.set_coils(pr.C)
        P.set_frame(0, pr.G["Gu"], pr.G["Gut"])
        y0[key] = P.bmShanna(torch.from_numpy(synth.ellipsoid_phantom(n_u, rng)).cuda())
        del P
        protos[key] = pr
    torch.cuda.synchronize()
    log(f"[bench] {len(protos)} protocols registered in {time.time() - t0:.1f}s")
    keys = list(protos)
    gen = torch.Generator(device="cuda").manual_seed(5)

    def make_requests(n, seed):
        r = np.random.default_rng(seed)
        out = []
        for _ in range(n):
            k = keys[int(r.integers(len(keys)))]
            y = y0[k]
            sig = float(y.abs().mean()) * 0.03
            noise = torch.complex(torch.randn(y.shape, generator=gen, device="cuda"),
                                  torch.randn(y.shape, generator=gen, device="cuda")) * sig
            out.append((k, (y + noise).contiguous()))
        torch.cuda.synchronize()
        return out

    free, total = torch.cuda.mem_get_info()
    budget = int(0.6 * free)

    def run(workers, max_batch, reqs, arrivals=None):
        srv = GpuServer(budget, max_batch=max_batch, workers=workers)
        for k, pr in protos.items():
            srv.register(k, pr)
        srv.start()
        # warm: one saturated pass caches the plans this form uses
        for k, y in reqs[:max(len(keys) * max_batch, 8)]:
            srv.submit(k, y, cfg["nCGD"])
        srv.drain()
        srv.completed.clear()
        srv.batches.clear()
        t_start = time.perf_counter()
        handles = []
        for i, (k, y) in enumerate(reqs):
            if arrivals is not None:
                wait = t_start + arrivals[i] - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
            handles.append(srv.submit(k, y, cfg["nCGD"]))
        srv.drain()
        elapsed = time.perf_counter() - t_start
        peak = srv.peak_bytes
        batches = list(srv.batches)
        srv.stop()
        bad = [h for h in handles if h.error is not None]
        if bad:
            raise RuntimeError(f"{len(bad)} requests failed: {bad[0].error!r}")
        lat = [h.latency for h in handles]
        return dict(recons_per_s=len(handles) / elapsed, latency_ms={k: v * 1e3 for k, v in percentiles(lat).items()},
                    mean_batch=float(np.mean([b for _, b in batches])) if batches else 0.0, peak_bytes=peak)

    sat = make_requests(cfg["n_sat"], 1)
    plain_sat = run(1, 1, sat)
    served_sat = run(cfg["workers"], cfg["max_batch"], sat)
    lam = cfg["load"] * plain_sat["recons_per_s"]
    pois = make_requests(cfg["n_poisson"], 2)
    arr = np.cumsum(np.random.default_rng(3).exponential(1.0 / lam, size=len(pois)))
    plain_lat = run(1, 1, pois, arr)
    served_lat = run(cfg["workers"], cfg["max_batch"], pois, arr)
    out = {
        "metric": "served CG-SENSE recons/s (%d-step CG, 2D/3D protocol mix) with p50/p99 latency" % cfg["nCGD"],
        "value": served_sat["recons_per_s"],
        "unit": "recons/s",
        "n_gpus": 1,
        "steps": cfg["n_sat"],
        "warmup": 1,
        "ms_per_step": 1e3 / served_sat["recons_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 vectors, fp64 CG scalars)",
        "data": "synthetic (seeded radial / koosh-ball protocols, Gaussian-blob coil maps, phantoms + noise)",
        "config": {"workload": "recon serving, protocols " + ", ".join(
            f"{k}: {d}D N={n} {c} coils {sp} spokes" for k, d, n, c, sp, _ in cfg["protocols"]),
                   "parallelism": "single GPU", "max_batch": cfg["max_batch"], "workers": cfg["workers"],
                   "hbm_budget_bytes": budget},
        "plain_per_request_recons_per_s": plain_sat["recons_per_s"],
        "served_over_plain": served_sat["recons_per_s"] / plain_sat["recons_per_s"],
        "offered_load_req_per_s": lam,
        "latency_ms": {"served": served_lat["latency_ms"], "plain": plain_lat["latency_ms"]},
        "throughput_at_load": {"served": served_lat["recons_per_s"], "plain": plain_lat["recons_per_s"]},
        "mean_batch": {"saturated": served_sat["mean_batch"], "at_load": served_lat["mean_batch"]},
        "peak_hbm_bytes": max(served_sat["peak_bytes"], served_lat["peak_bytes"]),
    }
    if not quiet:
        print(json.dumps(out), flush=True)
    return out


def measured_traffic(config, op="n2u"):
    """HBM bytes per launch of an op (n2u: one adjoint gridding; fft: one
    transform, all passes) from the committed rocprofv3 counter summary of
    this workload: FETCH_SIZE (scaled by the factor a plain copy of known size
    gives) + WRITE_SIZE."""
    path = os.path.join(ROOT, "profiles", f"{config}_{op}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return {"bytes": d["fetch_bytes"] + d["write_bytes"],
            "source": f"profiles/{config}_{op}_traffic.json (rocprofv3 FETCH_SIZE x {d['fetch_factor']} + WRITE_SIZE)"}


def fft_bytes(prob):
    """Algorithmic HBM bytes of one centred 3-D transform with its fused
    prologue / epilogue: one read and one write of the nCh-coil grid plus the
    image and coil-map side (16 nCh prod(N_u) + 8 prod(n_u) (1 + nCh))."""
    nCh = prob["nCh"]
    G = int(np.prod(prob["N_u"]))
    nu = int(np.prod(prob["n_u"]))
    return 16 * nCh * G + 8 * nu * (1 + nCh)


def gridding_bytes(prob):
    """Algorithmic HBM bytes per launch of the two gridding ops (complex64 =
    8 B; P samples, U touched cells = Gut.l_nJump, G = prod(N_u), d = imDim):
    the forward gridding runs matrix-free, 8 nCh (U + P) + 4 d P (each
    touched cell and each sample once, the sample coordinates); the adjoint
    runs on the stored Gut, 8 nCh (P + G) + 4 (d + 1) P (matrix-free form,
    the dense grid the FFT reads) plus the jump-CSR form's 8 nnz + 8 U."""
    Gu, Gut = prob["G"]["Gu"], prob["G"]["Gut"]
    nCh = prob["nCh"]
    P = Gu.l_size
    U = Gut.l_nJump
    G = int(np.prod(prob["N_u"]))
    d = len(prob["N_u"])
    nnz = Gu.nnz
    u2n = 8 * nCh * (U + P) + 4 * d * P
    n2u = 8 * nCh * (P + G) + 4 * (d + 1) * P + 8 * nnz + 8 * U
    return u2n, n2u, U


def cpu_baseline(prob, threads):
    """One forward and one adjoint at the full workload on the host cores:
    sparse products by the reference's own bmSparseMat_cC_oBlock_omp (built
    from its sources into oracle/_ref; the oracle's C restatement when that
    build is absent), centred FFTs by scipy.fft (workers = threads),
    elementwise by numpy.  A CG-SENSE recon costs 22 forwards + 20 adjoints."""
    import ctypes as C
    import scipy.fft as sfft
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_libs as OL
    from oracle import mona_oracle as O
    ref = OL.ref()
    kind = "reference" if ref is not None else "port"
    fn = ref.ref_cC_oBlock_omp if ref is not None else OL.port().port_cC_oBlock_omp
    P = prob["P"]
    Gu, Gut = prob["G"]["Gu"], prob["G"]["Gut"]
    n_u, N_u, nCh = prob["n_u"], prob["N_u"], prob["nCh"]
    KFC, KFCc = [a.cpu().numpy() for a in P.KFC()]
    x = prob["phantom"]
    Mu = dict(r_size=Gu.r_size, l_size=Gu.l_size, r_jump=Gu.r_jump, r_nJump=Gu.r_nJump, m_val=Gu.m_val,
              l_nJump=Gu.l_nJump, l_jump=None)
    Mt = dict(r_size=Gut.r_size, l_size=Gut.l_size, r_jump=Gut.r_jump, r_nJump=Gut.r_nJump,
              m_val=Gut.m_val, l_nJump=Gut.l_nJump, l_jump=Gut.l_jump)
    omp_env = os.environ.get("OMP_NUM_THREADS")

    def cfft_cpu(g, inverse):
        axes = (1, 2, 3)
        g = sfft.ifftshift(g, axes=axes)
        g = (sfft.ifftn if inverse else sfft.fftn)(g, axes=axes, workers=threads, overwrite_x=True)
        return sfft.fftshift(g, axes=axes)

    sl = tuple(slice(N // 2 - n // 2, N // 2 - n // 2 + n) for N, n in zip(N_u[::-1], n_u[::-1]))

    def forward():
        # x.*KFC -> zero-fill -> cfft -> Gu
        xc = (x[:, None] * KFC).T.reshape((nCh,) + tuple(n_u[::-1]))   # [c, z, y, x]
        g = np.zeros((nCh,) + tuple(N_u[::-1]), np.complex64)
        g[(slice(None),) + sl] = xc
        g = cfft_cpu(g, False).astype(np.complex64)
        gr = np.ascontiguousarray(g.real.reshape(nCh, -1))
        gi = np.ascontiguousarray(g.imag.reshape(nCh, -1))
        del g
        return OL.call_cC(fn, Mu, gr, gi)

    def adjoint(yr, yi):
        # Gut -> cifft -> crop -> .*KFC_conj -> coil sum
        wr, wi = OL.call_cC(fn, Mt, yr, yi)
        g = (wr + 1j * wi).astype(np.complex64).reshape((nCh,) + tuple(N_u[::-1]))
        del wr, wi
        g = cfft_cpu(g, True)
        g = g[(slice(None),) + sl].reshape(nCh, -1).T
        return (g * KFCc).sum(1)

    def timed(fn_, *a):
        t0 = time.perf_counter()
        r = fn_(*a)
        return time.perf_counter() - t0, r

    # one warm-up of each, then the median of 5 timed runs (monotonic clock)
    _, (yr, yi) = timed(forward)
    timed(adjoint, yr, yi)
    tf = [timed(forward)[0] for _ in range(5)]
    ta = [timed(adjoint, yr, yi)[0] for _ in range(5)]
    t_fwd, t_adj = float(np.median(tf)), float(np.median(ta))
    recon_s = 22 * t_fwd + 20 * t_adj
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(value=1.0 / recon_s, unit="recons/s", cores=threads, kind=kind,
                sample=(f"forward and adjoint at the full workload, one warm-up each then the median of 5 "
                        f"(forward {t_fwd:.2f} s, adjoint {t_adj:.2f} s), scaled to a 20-iteration CG-SENSE "
                        f"(22 forward + 20 adjoint, vector updates not counted); sparse products: "
                        f"{'the reference kernel built from its sources' if kind == 'reference' else 'the oracle C restatement'} "
                        f"(OpenMP, min(2*nCh, threads) threads), FFT: scipy.fft with {threads} workers"),
                nufft_pairs_per_s=1.0 / (t_fwd + t_adj),
                forward_s=tf, adjoint_s=ta, nproc=os.cpu_count(), cpu_model=model,
                omp_num_threads=omp_env, fft_workers=threads)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c1", choices=sorted(CONFIGS))
    ap.add_argument("--pairs", type=int, default=10, help="NUFFT pairs timed after the recons")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-latency leg of the default line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="start with a 1 GiB device copy (a known-size kernel for counter calibration)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from monalisa_amd import nufft
    cfg = CONFIGS[args.config]
    if args.calibrate_copy:
        a = torch.ones(1 << 27, dtype=torch.complex64, device="cuda")
        b = a.clone()
        torch.cuda.synchronize()
        del a, b
    if "nBin" in cfg:
        bench_binned(args, cfg)
        return
    if cfg.get("mixed"):
        bench_mixed(args, cfg)
        return
    if cfg.get("serve"):
        bench_serving(args, cfg)
        return
    prob = build_problem(cfg)
    P, y, ve = prob["P"], prob["y"], prob["ve"]
    nu = int(np.prod(prob["n_u"]))
    stream = torch.cuda.current_stream()

    def recon():
        x = torch.zeros(nu, dtype=torch.complex64, device="cuda")
        P.bmSensa(x, y, ve, cfg["nCGD"], nIter=1, stream=stream)
        return x

    for _ in range(args.warmup):
        recon()
    torch.cuda.synchronize()
    nufft.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x = recon()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n_u2n, ms_u2n = nufft.timing(kind=0)
    n_n2u, ms_n2u = nufft.timing(kind=1)
    n_fft, ms_fft = nufft.timing(kind=2)
    nufft.timing(False)

    # NUFFT pairs (forward + adjoint) after the timed recons
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.pairs):
        yy = P.bmShanna(x)
        P.bmNakatsha(yy)
    torch.cuda.synchronize()
    pair_s = (time.perf_counter() - t1) / max(args.pairs, 1)

    # image quality vs the phantom (sanity, not a parity claim)
    xr = x.cpu().numpy()
    ph = prob["phantom"]
    a = np.vdot(xr, ph) / np.vdot(xr, xr)
    img_err = float(np.linalg.norm(a * xr - ph) / np.linalg.norm(ph))

    u2n_b, n2u_b, U = gridding_bytes(prob)
    avg_u2n = ms_u2n / max(n_u2n, 1)
    avg_n2u = ms_n2u / max(n_n2u, 1)
    if ms_n2u >= ms_u2n:
        dom, bytes_, avg = ("n2u (adjoint gridding: LDS-staged groups of grid cells + long rows by ranges of "
                            "consecutive samples)", n2u_b, avg_n2u)
    else:
        dom, bytes_, avg = "u2n (matrix-free forward gridding, x-runs)", u2n_b, avg_u2n
    achieved = bytes_ / (avg * 1e-3) / 1e9
    ms_per = elapsed / args.steps * 1e3
    # the FFT op (the largest share of a recon): each timed region is one
    # transform (all its passes)
    n_tr = n_fft
    fft_b = fft_bytes(prob)
    avg_tr = ms_fft / max(n_tr, 1)
    fft_ach = fft_b / (avg_tr * 1e-3) / 1e9 if n_tr else None
    out = {
        "metric": "NUFFT pairs/s + CG-SENSE recons/s at 3D N=128 20ch; p50/p99 served latency",
        "value": args.steps / elapsed,
        "unit": "recons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 vectors, fp64 CG scalars)",
        "data": "synthetic (seeded koosh-ball trajectory, Gaussian-blob coil maps, ellipsoid phantom, 30 dB noise)",
        "config": {"workload": f"3D koosh-ball CG-SENSE n_u={cfg['n']}^3 N_u={2 * cfg['n']}^3 "
                               f"{cfg['nSeg'] * cfg['nShot']} spokes x {cfg['nSample']} samples "
                               f"({prob['P_pts']} points), {cfg['nCh']} coils, KB nWin=3, "
                               f"{cfg['nCGD']} CG iterations",
                   "parallelism": "single GPU"},
        "nufft_pairs_per_s": 1.0 / pair_s,
        # "other": the rest of a recon's time (CG vector updates, FP64
        # reductions, launch gaps): step time minus the timed operator kernels
        "kernel_ms_per_recon": {"u2n": ms_u2n / args.steps, "n2u": ms_n2u / args.steps,
                                "fft": ms_fft / args.steps,
                                "other": ms_per - (ms_u2n + ms_n2u + ms_fft) / args.steps},
        "launches": {"u2n": n_u2n, "n2u": n_n2u},
        "image_rel_err_vs_phantom": img_err,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": None,
                     "bytes_per_launch": bytes_, "avg_launch_ms": avg,
                     "u2n_GBps": u2n_b / (avg_u2n * 1e-3) / 1e9 if n_u2n else None,
                     "n2u_GBps": n2u_b / (avg_n2u * 1e-3) / 1e9 if n_n2u else None,
                     "u2n": {"kernel": "u2n (matrix-free forward gridding, x-runs)", "bytes_per_launch": u2n_b,
                             "avg_launch_ms": avg_u2n,
                             "achieved": u2n_b / (avg_u2n * 1e-3) / 1e9 if n_u2n else None,
                             "frac": u2n_b / (avg_u2n * 1e-3) / 1e9 / 8000.0 if n_u2n else None, "traffic": None},
                     "touched_cells_U": U,
                     "fft_op": {"achieved": fft_ach, "peak": 8000.0, "unit": "GB/s",
                                "frac": fft_ach / 8000.0 if fft_ach else None, "traffic": None,
                                "bytes_per_transform": fft_b, "avg_transform_ms": avg_tr, "transforms": n_tr}},
        "cpu_baseline": None,
    }
    # measured HBM traffic (committed counter summaries) and the rate on it
    tr = measured_traffic(args.config, "n2u" if ms_n2u >= ms_u2n else "u2n")
    if tr is not None:
        out["roofline"]["traffic"] = tr["bytes"]
        out["roofline"]["traffic_source"] = tr["source"]
        out["roofline"]["achieved_on_traffic"] = tr["bytes"] / (avg * 1e-3) / 1e9
    tr = measured_traffic(args.config, "u2n")
    if tr is not None and n_u2n:
        out["roofline"]["u2n"]["traffic"] = tr["bytes"]
        out["roofline"]["u2n"]["traffic_source"] = tr["source"]
        out["roofline"]["u2n"]["achieved_on_traffic"] = tr["bytes"] / (avg_u2n * 1e-3) / 1e9
    tr = measured_traffic(args.config, "fft")
    if tr is not None:
        out["roofline"]["fft_op"]["traffic"] = tr["bytes"]
        out["roofline"]["fft_op"]["traffic_source"] = tr["source"]
    if not args.no_cpu:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        try:
            cb = cpu_baseline(prob, threads)
            out["cpu_baseline"] = cb
            out["speedup_vs_cpu"] = out["value"] / cb["value"]
        except Exception as e:  # reported, never silently replaced
            out["cpu_baseline"] = {"error": repr(e)}
    if not args.no_serve and args.config == "c1":
        # served latency over the 2D / 3D protocol mix (a shorter run of
        # the recon-serving benchmark)
        del prob, P, y, ve, x, yy
        torch.cuda.synchronize()
        sv = bench_serving(args, dict(CONFIGS["c4"], n_sat=48, n_poisson=120), quiet=True)
        out["served"] = {"recons_per_s": sv["value"], "plain_per_request_recons_per_s":
                         sv["plain_per_request_recons_per_s"], "latency_ms": sv["latency_ms"],
                         "offered_load_req_per_s": sv["offered_load_req_per_s"],
                         "workload": sv["config"]["workload"] + ", 10-step CG-SENSE each"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
