"""Recon-serving loop: continuous batching of CG-SENSE requests on one GPU.

Requests name a registered *protocol*: a geometry, gridding kernel,
trajectory (its Gu / Gut matrices), density weights and coil maps, all kept
device-resident.  The dispatcher admits waiting requests as workers and HBM
budget free up: requests of one protocol are batched (up to ``max_batch``)
into one :meth:`Plan.bmSensa_batch` call (one block-diagonal gridding launch
and frame-batched FFTs per operator application), and several batches run
at once, each on its own worker thread and HIP stream, so their work
overlaps on the device.  Admission is by HBM budget: a batch starts only if
the plans it needs (resident plans count at their measured size, new ones
at an estimate) and its request buffers fit; idle plans are evicted
oldest-first to make room.

A request's result does not depend on what it was batched or overlapped
with: every kernel is deterministic and bmSensa_batch gives each frame the
result bmSensa gives it alone.

The scheduling logic is independent of the device (``backend`` argument);
:class:`GpuBackend` is the libmona one.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
import time
from typing import Any, Dict, List, Optional

import numpy as np


@dataclasses.dataclass
class Request:
    """One reconstruction: ``y`` [nPt, nCh] (device) for a protocol."""
    rid: int
    key: str
    y: Any
    nCGD: int = 10
    t_submit: float = 0.0
    t_start: float = 0.0
    t_done: float = 0.0
    batch: int = 0
    ready: Any = None     # backend object the batch waits on before reading y
    x: Any = None
    residual: Any = None
    error: Optional[BaseException] = None
    done: threading.Event = dataclasses.field(default_factory=threading.Event)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class ReconServer:
    """Continuous-batching CG-SENSE server.

    backend: object with ``batch_capacity(protocol)`` (largest batch worth
    forming for it), ``plan_bytes_estimate(protocol, capacity)``,
    ``request_bytes(protocol)``, ``make_plan(protocol, capacity, stream)``,
    ``plan_bytes(plan)``, ``run(plan, protocol, requests, stream)`` (any
    batch up to the plan's capacity), ``new_stream()``, ``release(plan)``.
    Plans are cached per protocol, built for its batch capacity.
    """

    def __init__(self, backend, hbm_budget: int, max_batch: int = 4, workers: int = 3):
        self.backend = backend
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.n_workers = int(workers)
        self.protocols: Dict[str, Any] = {}
        self._queues: Dict[str, collections.deque] = collections.OrderedDict()
        self._cv = threading.Condition()
        self._idle_plans: Dict[str, List[Any]] = collections.defaultdict(list)
        self._plan_bytes: Dict[int, int] = {}
        self._plan_last_use: Dict[int, float] = {}
        self._busy_bytes = 0          # plans in use + request buffers in flight
        self._free_workers = self.n_workers
        self._jobs: collections.deque = collections.deque()
        self._stop = False
        self._next_id = 0
        self._threads: List[threading.Thread] = []
        self.completed: List[Request] = []
        self.peak_bytes = 0
        self.batches: List[tuple] = []   # (key, B) of every launched batch

    # -- protocols and requests -----------------------------------------
    def register(self, key: str, protocol) -> None:
        with self._cv:
            self.protocols[key] = protocol
            self._queues.setdefault(key, collections.deque())

    def submit(self, key: str, y, nCGD: int = 10, ready=None) -> Request:
        if key not in self.protocols:
            raise KeyError(f"unknown protocol {key!r}")
        with self._cv:
            r = Request(self._next_id, key, y, nCGD, t_submit=time.perf_counter(), ready=ready)
            self._next_id += 1
            self._queues[key].append(r)
            self._cv.notify_all()
        return r

    # -- lifecycle --------------------------------------------------------
    def start(self) -> "ReconServer":
        self._stop = False
        d = threading.Thread(target=self._dispatch, name="mona-dispatch", daemon=True)
        self._threads = [d]
        for i in range(self.n_workers):
            self._threads.append(threading.Thread(target=self._work, name=f"mona-worker{i}", daemon=True))
        for t in self._threads:
            t.start()
        return self

    def stop(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        self._threads = []
        for plans in self._idle_plans.values():
            for p in plans:
                self.backend.release(p)
        self._idle_plans.clear()

    def drain(self, timeout: Optional[float] = None) -> bool:
        """Waits until every submitted request is done."""
        end = None if timeout is None else time.perf_counter() + timeout
        with self._cv:
            while any(self._queues.values()) or self._jobs or self._free_workers < self.n_workers:
                rem = None if end is None else end - time.perf_counter()
                if rem is not None and rem <= 0:
                    return False
                self._cv.wait(rem)
        return True

    # -- accounting -------------------------------------------------------
    def resident_bytes(self) -> int:
        idle = sum(self._plan_bytes[id(p)] for ps in self._idle_plans.values() for p in ps)
        return idle + self._busy_bytes

    def _evict_for(self, need: int, keep=None) -> bool:
        """Frees idle plans (not `keep`), least recently used first, until
        `need` more bytes fit the budget."""
        while self.resident_bytes() + need > self.budget:
            cands = [(self._plan_last_use[id(p)], k, p) for k, ps in self._idle_plans.items() for p in ps
                     if p is not keep]
            if not cands:
                return False
            _, k, p = min(cands, key=lambda c: c[0])
            self._idle_plans[k].remove(p)
            self._plan_bytes.pop(id(p))
            self._plan_last_use.pop(id(p))
            self.backend.release(p)
        return True

    # -- dispatcher -------------------------------------------------------
    def capacity(self, key: str) -> int:
        return max(1, min(self.max_batch, int(self.backend.batch_capacity(self.protocols[key]))))

    def _next_batch(self):
        """Oldest waiting request first; its protocol's waiting requests
        (up to the protocol's batch capacity) form the batch.  It starts
        when a plan of the protocol is idle or a new one fits the budget
        (idle plans of other protocols evicted oldest-first)."""
        heads = [(q[0].t_submit, k) for k, q in self._queues.items() if q]
        if not heads:
            return None
        _, key = min(heads)
        q = self._queues[key]
        proto = self.protocols[key]
        cap = self.capacity(key)
        alone = self.backend.plan_bytes_estimate(proto, cap) + self.backend.request_bytes(proto)
        if alone > self.budget and not self._idle_plans[key]:
            # can never be admitted: fail it instead of blocking the queue
            r = q.popleft()
            r.error = MemoryError(f"protocol {key!r} needs {alone} bytes, budget {self.budget}")
            r.t_start = r.t_done = time.perf_counter()
            self.completed.append(r)
            r.done.set()
            return None
        # a batch runs one CG step count: the head's nCGD, same-nCGD
        # requests of the protocol in queue order (the others stay queued)
        B = min(cap, sum(1 for r in q if r.nCGD == q[0].nCGD))
        plan = self._idle_plans[key][-1] if self._idle_plans[key] else None
        need = (0 if plan is not None else self.backend.plan_bytes_estimate(proto, cap)) + \
            B * self.backend.request_bytes(proto)
        if self._evict_for(need, keep=plan):
            return key, B
        return None

    @staticmethod
    def _take(q, B):
        """The first B requests of q with the head's nCGD, in queue order;
        the others keep their places."""
        n = q[0].nCGD
        out, keep = [], []
        while q and len(out) < B:
            r = q.popleft()
            (out if r.nCGD == n else keep).append(r)
        q.extendleft(reversed(keep))
        return out

    def _dispatch(self):
        with self._cv:
            while not self._stop:
                launched = False
                if self._free_workers > 0:
                    nb = self._next_batch()
                    if nb is not None:
                        key, B = nb
                        proto = self.protocols[key]
                        reqs = self._take(self._queues[key], B)
                        plan = self._idle_plans[key].pop() if self._idle_plans[key] else None
                        if plan is not None:
                            pb = self._plan_bytes[id(plan)]
                        else:
                            pb = self.backend.plan_bytes_estimate(proto, self.capacity(key))
                        rb = B * self.backend.request_bytes(proto)
                        self._busy_bytes += pb + rb
                        self.peak_bytes = max(self.peak_bytes, self.resident_bytes())
                        self._free_workers -= 1
                        self._jobs.append((key, B, plan, pb, rb, reqs))
                        self.batches.append((key, B))
                        launched = True
                        self._cv.notify_all()
                if not launched:
                    self._cv.wait(0.05)

    # -- workers ----------------------------------------------------------
    def _work(self):
        stream = self.backend.new_stream()
        while True:
            with self._cv:
                while not self._jobs and not self._stop:
                    self._cv.wait(0.05)
                if self._stop and not self._jobs:
                    return
                key, B, plan, pb, rb, reqs = self._jobs.popleft()
            proto = self.protocols[key]
            t0 = time.perf_counter()
            for r in reqs:
                r.t_start = t0
                r.batch = B
            err = None
            try:
                if plan is None:
                    plan = self.backend.make_plan(proto, self.capacity(key), stream)
                self.backend.run(plan, proto, reqs, stream)
            except BaseException as e:  # reported on the requests
                err = e
            t1 = time.perf_counter()
            with self._cv:
                self._busy_bytes -= pb + rb
                if plan is not None and err is None:
                    real = self.backend.plan_bytes(plan)
                    self._plan_bytes[id(plan)] = real
                    self._plan_last_use[id(plan)] = t1
                    self._idle_plans[key].append(plan)
                    if self.resident_bytes() > self.budget:
                        self._evict_for(0)
                elif plan is not None:
                    self.backend.release(plan)
                self._free_workers += 1
                for r in reqs:
                    r.t_done = t1
                    r.error = err
                    self.completed.append(r)
                    r.done.set()
                self._cv.notify_all()


# ---------------------------------------------------------------- GPU backend
@dataclasses.dataclass
class Protocol:
    """Device-resident acquisition protocol."""
    N_u: list
    n_u: list
    dK_u: list
    nCh: int
    kernel: str
    G: dict               # bmTraj2SparseMat output (Gu, Gut handles)
    C: Any                # coil maps [nu, nCh] complex64 (device)
    ve: Any               # volume elements [nPt] float32 (device)
    nPt: int = 0

    @property
    def nu(self) -> int:
        return int(np.prod(self.n_u))


def make_protocol(t, ve, C, N_u, n_u, dK_u, kernel="kaiser") -> Protocol:
    """Builds the gridding matrices (on the GPU) and uploads coil maps and
    density weights; the plans of the protocol share them."""
    import torch
    from .sparse import bmTraj2SparseMat
    G = bmTraj2SparseMat(t, ve, N_u, dK_u, kernel, device=True)
    return Protocol(list(map(int, np.ravel(N_u))), list(map(int, np.ravel(n_u))), list(map(float, np.ravel(dK_u))),
                    int(C.shape[1]), kernel, G, torch.from_numpy(np.ascontiguousarray(C, np.complex64)).cuda(),
                    torch.from_numpy(np.asarray(ve, np.float32)).cuda(), nPt=int(G["Gu"].l_size))


class GpuBackend:
    """libmona execution of the batches (Plan.bmSensa_batch).

    work_per_batch: oversampled-grid elements x coils a batch should reach;
    a protocol whose single request already exceeds it is not batched (its
    gridding and FFT launches fill the GPU alone, and batching would only
    lengthen the jobs others wait behind)."""

    def __init__(self, work_per_batch: float = 1.5e8, max_capacity: int = 8):
        self.work_per_batch = work_per_batch
        self.max_capacity = max_capacity

    def batch_capacity(self, proto: "Protocol") -> int:
        work = float(np.prod(proto.N_u)) * proto.nCh
        return int(max(1, min(self.max_capacity, self.work_per_batch // work)))

    def new_stream(self):
        import torch
        return torch.cuda.Stream()

    @staticmethod
    def plan_bytes_estimate(proto: Protocol, B: int) -> int:
        G = int(np.prod(proto.N_u))
        nnz = proto.G["Gu"].nnz
        grid = 8 * G * proto.nCh
        mats = 2 * nnz * 16 + 8 * proto.nPt * 2     # Gu_s, Gut_s (col, val, packed) + row pointers
        per_frame = grid * 2 + 8 * proto.nPt * proto.nCh * 4 + 8 * proto.nu * 4
        return int(1.1 * (mats * (1 + B) + grid + B * per_frame + 16 * proto.nu * proto.nCh)) + (64 << 20)

    @staticmethod
    def request_bytes(proto: Protocol) -> int:
        return 8 * proto.nu

    def make_plan(self, proto: Protocol, B: int, stream):
        import torch
        from .nufft import Plan
        with torch.cuda.stream(stream):
            P = Plan(proto.N_u, proto.n_u, proto.dK_u, proto.nCh, proto.kernel)
            P.set_coils(proto.C, stream=stream)
            for f in range(B):
                P.set_frame(f, proto.G["Gu"], proto.G["Gut"])
        return P

    @staticmethod
    def plan_bytes(plan) -> int:
        return plan.device_bytes()

    def run(self, plan, proto: Protocol, reqs, stream):
        import torch
        B = len(reqs)
        with torch.cuda.stream(stream):
            for r in reqs:
                if r.ready is not None:
                    stream.wait_event(r.ready)
            x = torch.zeros((B, proto.nu), dtype=torch.complex64, device="cuda")
            ycat = torch.cat([r.y.reshape(-1) for r in reqs])
            vcat = proto.ve.repeat(B)
            res = plan.bmSensa_batch(x, ycat, vcat, reqs[0].nCGD, nIter=1, stream=stream)
            stream.synchronize()
        for i, r in enumerate(reqs):
            r.x = x[i]
            r.residual = float(res[i, 0])

    @staticmethod
    def release(plan) -> None:
        plan.__del__()


class GpuServer(ReconServer):
    """ReconServer on the GPU backend: submit() takes a device tensor and
    orders the batch after the caller's stream."""

    def __init__(self, hbm_budget: int, max_batch: int = 4, workers: int = 3):
        super().__init__(GpuBackend(), hbm_budget, max_batch, workers)

    def submit(self, key: str, y, nCGD: int = 10, ready=None) -> Request:
        import torch
        from .nufft import _arg
        if key not in self.protocols:
            raise KeyError(f"unknown protocol {key!r}")
        pr = self.protocols[key]
        # a mis-sized y would shift every later request of a batch
        _arg(y, "y", pr.ve.numel() * pr.nCh)
        if int(nCGD) < 1:
            raise ValueError("nCGD must be >= 1")
        if ready is None:
            ready = torch.cuda.Event()
            ready.record(torch.cuda.current_stream())
        return super().submit(key, y, nCGD, ready)


def percentiles(values, ps=(50, 99)):
    v = np.sort(np.asarray(values, dtype=np.float64))
    return {f"p{p}": float(np.percentile(v, p)) if v.size else None for p in ps}
