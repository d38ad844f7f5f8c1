"""ctypes binding of libmona.so (declarations in include/mona.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmona.so")


class MonaError(RuntimeError):
    """A libmona entry point returned a non-zero status."""

    def __init__(self, status, msg):
        super().__init__(f"libmona status {status}: {msg}")
        self.status = status


P = C.c_void_p
I = C.c_int
F = C.c_float
D = C.c_double
LL = C.c_longlong
IP = C.POINTER(C.c_int)
FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
LLP = C.POINTER(C.c_longlong)
PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "mona_last_error": (C.c_char_p, []),
    "mona_device_count": (I, []),
    "mona_version": (C.c_char_p, []),
    # drop-in entry points
    "bmSparseMat_cC_oBlock_omp": (None, [I, IP, IP, FP, I, IP, I, FP, FP, I, FP, FP]),
    "bmSparseMat_cC_oBlock": (None, [I, IP, IP, FP, I, IP, I, FP, FP, I, FP, FP]),
    "bmSparseMat_cC_oBlock_2_omp": (None, [I, IP, IP, FP, I, IP, I, FP, FP, I, FP, FP]),
    "bmSparseMat_cR_oBlock": (None, [I, IP, IP, FP, I, IP, I, FP, FP, I, FP, FP]),
    "bmSparseMat_rC_oBlock": (None, [I, IP, IP, FP, I, IP, I, FP, I, FP]),
    "bmSparseMat_rR_oBlock": (None, [I, IP, IP, FP, I, IP, I, FP, I, FP]),
    "bmSparseMat_cR_nBlock_omp": (None, [I, IP, IP, FP, I, IP, I, I, IP, IP, LLP, FP, FP, I, FP, FP]),
    "bmSparseMat_rR_nBlock_omp": (None, [I, IP, IP, FP, I, IP, I, I, IP, IP, LLP, FP, I, FP]),
    "bmSparseMat_r_nJump2index_mex": (None, [I, IP, IP]),
    "bmGridder_n2u_leight1_mex": (None, [FP, FP, FP, FP, FP, FP, I, I, I, I, F, F]),
    "bmGridder_n2u_leight2_mex": (None, [FP, FP, FP, FP, FP, FP, FP, I, I, I, I, I, F, F]),
    "bmGridder_n2u_leight3_mex": (None, [FP, FP, FP, FP, FP, FP, FP, FP, I, I, I, I, I, I, F, F]),
    "bmGridder_n2u_mex": (None, [FP, FP, FP, FP, FP, IP, I, I, I, I]),
    "bmGriddingMatrix_secret_length1_mex": (None, [FP, I, I, I, I, LLP]),
    "bmGriddingMatrix_secret_length2_mex": (None, [FP, FP, I, I, I, I, I, LLP]),
    "bmGriddingMatrix_secret_length3_mex": (None, [FP, FP, FP, I, I, I, I, I, I, LLP]),
    # handle API
    "mona_sparse_create": (I, [I, IP, IP, FP, I, IP, I, LL, PP]),
    "mona_sparse_destroy": (I, [P]),
    "mona_sparse_info": (I, [P, LLP, IP, IP, IP, IP]),
    "mona_sparse_export": (I, [P, IP, IP, FP, IP]),
    "mona_sparse_apply": (I, [P, P, P, I, P]),
    "mona_sparse_apply_noalloc": (I, [P, P, P, I, P]),
    "mona_traj2sparsemat": (I, [I, I, DP, DP, IP, DP, I, I, DP, I, I, PP, PP, PP]),
    "mona_traj2sparsemat_device": (I, [I, I, DP, DP, IP, DP, I, I, DP, I, I, PP, PP, PP, P]),
    "mona_deform2sparsemat": (I, [I, IP, DP, DP, I, I, PP, PP, PP]),
    "mona_plan_create": (I, [I, IP, IP, DP, I, I, I, DP, I, PP]),
    "mona_plan_destroy": (I, [P]),
    "mona_plan_set_coils": (I, [P, P, P]),
    "mona_plan_set_frame": (I, [P, I, P, P]),
    "mona_plan_get_K": (I, [P, P, P]),
    "mona_plan_get_KFC": (I, [P, P, P, P]),
    "mona_shanna": (I, [P, I, P, P, P]),
    "mona_nakatsha": (I, [P, I, P, P, I, P]),
    "mona_cfft": (I, [I, IP, I, P, I, P]),
    "mona_sensa": (I, [P, I, P, P, P, F, I, I, DP, P]),
    "mona_steva": (I, [P, I, P, P, P, I, P, P, F, FP, FP, I, I, DP, P]),
    "mona_plan_set_deform": (I, [P, I, I, P, P]),
    "mona_morphosia": (I, [P, I, I, P, P, P, I, P, LLP, P, LLP, F, FP, FP, I, I, DP, P]),
    "mona_sensa_morphosia": (I, [P, P, P, LLP, P, LLP, F, I, I, DP, P]),
    "mona_teva_chain": (I, [P, P, P, P, I, P, LLP, P, LLP, F, FP, FP, I, I, DP, P]),
    "mona_mathilda": (I, [P, P, P, P, P]),
    "mona_nasha": (I, [P, P, P, P, P, P]),
    "mona_sensa_batch": (I, [P, I, P, P, LLP, FP, LLP, FP, I, I, DP, P]),
    "mona_plan_memory": (I, [P, LLP]),
    "mona_sensa_ir": (I, [P, I, P, P, FP, F, I, I, D, I, DP, IP, IP, P]),
    "mona_shanna_bf16": (I, [P, I, P, P, P]),
    "mona_nakatsha_bf16": (I, [P, I, P, P, P]),
    "mona_plan_set_option": (I, [P, I, LL]),
    "mona_plan_set_witness": (I, [P, P, I, P, C.POINTER(C.c_int), I]),
    "mona_plan_autotune_groups": (I, [P, I, P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mona_plan_witness_elems": (I, [P, C.POINTER(C.c_longlong)]),
    "mona_plan_autotune": (I, [P, I, P, LLP]),
    "mona_timing_enable": (I, [I]),
    "mona_timing_read": (I, [I, LLP, DP]),
}

# C++-linkage twin of bmSparseMat_cC_oBlock_omp (the reference's header
# declares that one library symbol without extern "C"), by its mangled name.
CXX_CC_OBLOCK_OMP = "_Z25bmSparseMat_cC_oBlock_ompiPiS_PfiS_iS0_S0_iS0_S0_"
CXX_SIGNATURES = {CXX_CC_OBLOCK_OMP: SIGNATURES["bmSparseMat_cC_oBlock_omp"]}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C monalisa_amd/csrc)")
    # torch bundles its own HIP runtime (soname libamdhip64.so.7).  It must be
    # in the process before libmona so that libmona's libamdhip64.so.7
    # dependency resolves to that same runtime: two HIP runtimes in one
    # process do not share streams or device pointers.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    h = C.CDLL(LIB_PATH)
    for name, (res, args) in list(SIGNATURES.items()) + list(CXX_SIGNATURES.items()):
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    return h


lib = _load()


def last_error() -> str:
    m = lib.mona_last_error()
    return m.decode() if m else ""


def check(status: int):
    if status != 0:
        raise MonaError(status, last_error())
    return status


def ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_of(torch_stream=None):
    import torch
    s = torch_stream if torch_stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
