"""The benchmark entry point on its reduced configurations: every form
(CG-SENSE, batched temporal TV, mixed precision, serving) runs and prints
one well-formed result line."""
import contextlib
import io
import json
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("config", ["small", "c2small", "c3small", "c4small"])
def test_bench_config_runs(config):
    print(f"[gpu] bench {config}", flush=True)
    sys.path.insert(0, ROOT)
    try:
        import bench
    finally:
        sys.path.pop(0)
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        bench.main(["--config", config, "--no-cpu", "--no-serve", "--steps", "1", "--warmup", "0", "--pairs", "1"])
    lines = [ln for ln in out.getvalue().splitlines() if ln.startswith("{")]
    assert len(lines) == 1, out.getvalue()[-2000:]
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "dtype",
              "config"):
        assert k in d, k
    assert d["value"] > 0 and d["n_gpus"] == 1
    if config == "small":
        assert d["roofline"]["achieved"] > 0 and 0 < d["roofline"]["frac"] < 1
    if config == "c2small":
        # one batched, graph-replayed launch per gridding op over every bin
        # gives the per-bin eager results bit for bit
        assert d["batched_vs_per_bin_rel_l2"] == 0.0
