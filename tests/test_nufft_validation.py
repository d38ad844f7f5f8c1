"""The public operators and solvers refuse, before any library call, the
tensors their kernels would misread through raw pointers: a complex128 (or
other) dtype, a wrong element count, host memory.  Runs on the CPU: the
plan is a stand-in whose library handle must never be touched."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from monalisa_amd import nufft  # noqa: E402


class _Forbidden:
    def __getattr__(self, name):
        raise AssertionError(f"library called ({name}) before the arguments were checked")


class _Mat:
    def __init__(self, l_size, r_size):
        self.l_size, self.r_size = l_size, r_size


@pytest.fixture
def plan(monkeypatch):
    monkeypatch.setattr(nufft, "lib", _Forbidden())
    P = nufft.Plan.__new__(nufft.Plan)
    P._h = None
    P.imDim, P.nCh, P.nu, P.Nu = 2, 3, 16, 64
    P._frames = {0: (_Mat(10, 64), _Mat(64, 10)), 1: (_Mat(7, 64), _Mat(64, 7))}
    P._deform = {}
    return P


def c64(n):
    return torch.zeros(n, dtype=torch.complex64)


def f32(n):
    return torch.zeros(n, dtype=torch.float32)


CASES = [
    ("bmShanna complex128 x", lambda P: P.bmShanna(torch.zeros(16, dtype=torch.complex128))),
    ("bmShanna short x", lambda P: P.bmShanna(c64(15))),
    ("bmShanna short out", lambda P: P.bmShanna(c64(16), out=c64(29))),
    ("bmNakatsha long y", lambda P: P.bmNakatsha(c64(31))),
    ("bmNakatsha frame 1 y of frame 0", lambda P: P.bmNakatsha(c64(30), f=1)),
    ("bmNakatsha C_flag=0 out", lambda P: P.bmNakatsha(c64(30), C_flag=False, out=c64(16))),
    ("bmNakatsha unknown frame", lambda P: P.bmNakatsha(c64(30), f=5)),
    ("bmSensa complex ve", lambda P: P.bmSensa(c64(16), c64(30), c64(10), 3)),
    ("bmSensa short ve", lambda P: P.bmSensa(c64(16), c64(30), f32(9), 3)),
    ("bmSensa float y", lambda P: P.bmSensa(c64(16), f32(60), f32(10), 3)),
    ("bmSensa_batch short y", lambda P: P.bmSensa_batch(c64(32).reshape(2, 16), [c64(30), c64(20)],
                                                        [f32(10), f32(7)], 3)),
    ("bmSensa_batch too many frames", lambda P: P.bmSensa_batch(c64(48).reshape(3, 16), [c64(30)] * 3,
                                                                [f32(10)] * 3, 3)),
    ("sensa_refined short x", lambda P: P.sensa_refined(c64(8), c64(30), f32(10), 5, 2)),
    ("bmShanna_bf16 short x", lambda P: P.bmShanna_bf16(c64(8))),
    ("bmNakatsha_bf16 short y", lambda P: P.bmNakatsha_bf16(c64(8))),
    ("bmSteva short z", lambda P: P.bmSteva(c64(16), c64(16), c64(32), c64(30), f32(10), 0.1, 1.0, 2)),
    ("teva_chain y of another frame", lambda P: P.bmTevaMorphosia_chain(c64(32).reshape(2, 16), None, None,
                                                                         [c64(30), c64(30)], [f32(10), f32(7)],
                                                                         0.1, 1.0, 2, 1.0)),
    ("teva_chain missing frame data", lambda P: P.bmTevaMorphosia_chain(c64(32).reshape(2, 16), None, None,
                                                                         [c64(30)], [f32(10)], 0.1, 1.0, 2, 1.0)),
    ("sensitiva_chain complex128 y", lambda P: P.bmSensitivaMorphosia_chain(
        c64(32).reshape(2, 16), [torch.zeros(30, dtype=torch.complex128), c64(21)], [f32(10), f32(7)],
        0.1, 2, 1.0)),
    ("sensa_morphosia short x", lambda P: P.bmSensaMorphosia_chain(c64(15), [c64(30)], [f32(10)], 2, 1.0)),
    ("nasha short y", lambda P: P.bmNasha(c64(29), _Mat(64, 10))),
    ("set_coils short maps", lambda P: P.set_coils(c64(47))),
]


@pytest.mark.parametrize("name,call", CASES, ids=[c[0] for c in CASES])
def test_refused_before_library_call(plan, name, call):
    with pytest.raises((ValueError, TypeError)):
        call(plan)


def test_host_tensor_of_right_shape_is_refused(plan):
    """Right dtype and size, but host memory: TypeError, still no call."""
    with pytest.raises(TypeError, match="device"):
        plan.bmShanna(c64(16))


def test_cfft_partial_grid_refused(monkeypatch):
    monkeypatch.setattr(nufft, "lib", _Forbidden())
    with pytest.raises(ValueError):
        nufft.cfft(c64(70), [8, 8])
