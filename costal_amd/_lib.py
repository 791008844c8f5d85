"""Loader for ``libcostal_hip.so``, the gfx950 library behind ``include/costal_hip.h``.

There is deliberately no CPU fallback: if the library is not built, or no GPU is present, calls
fail loudly with the library's own error.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcostal_hip.so")

# error codes (include/costal_hip.h)
OK = 0
E_SHAPE, E_BAD_K, E_TOO_MANY_ROWS, E_NON_FINITE, E_UNAVAILABLE, E_DEVICE, E_UNSUPPORTED, E_INVALID = range(1, 9)

SELECT = {"cosine": 0, "tanimoto": 1}
PRECISION = {"fp32": 0, "fp16": 1, "fp16x3": 2}
WEIGHT = {"tanimoto": 0, "cosine": 1, "jaccard": 2}

# every symbol the header declares
EXPORTS = (
    "costal_knn_exact",
    "costal_knn_exact_device",
    "costal_knn_bruteforce",
    "costal_reweight",
    "costal_reweight_device",
    "costal_fuzzy_union",
    "costal_last_error",
    "costal_version",
    "costal_device_info",
    "costal_mlp_param_count",
    "costal_mlp_act_floats",
    "costal_mlp_predict_device",
    "costal_mlp_grad_device",
    "costal_mlp_adam_device",
    "costal_mlp_last_error",
    "costal_knn_batched_device",
    "costal_knn_plan",
    "costal_csr_from_edges",
    "costal_mfma_probe",
    "costal_knn_exact_device_noalloc",
)


class KnnParams(C.Structure):
    _fields_ = [
        ("k", C.c_uint64),
        ("margin", C.c_int64),
        ("select", C.c_uint32),
        ("precision", C.c_uint32),
        ("row_block", C.c_uint64),
        ("col_block", C.c_uint64),
        ("col_splits", C.c_uint32),
        ("no_prune", C.c_uint32),
        ("row_threshold", C.c_void_p),
        ("mlp_params", C.c_void_p),
        ("mlp_offset", C.c_float),
        ("mlp_bf16", C.c_uint32),
        ("prune_clusters", C.c_uint32),
        ("mlp_offset2", C.c_float),
        ("prune_tile", C.c_uint32),
    ]


class KnnStats(C.Structure):
    _fields_ = [
        ("fallback_rows", C.c_uint64),
        ("row_block", C.c_uint64),
        ("col_block", C.c_uint64),
        ("blocks_skipped", C.c_uint64),
        ("blocks_total", C.c_uint64),
        ("candidates", C.c_uint64),
        ("col_splits", C.c_uint64),
        ("short_rows", C.c_uint64),
        ("path", C.c_uint64),
        ("inserts", C.c_uint64),
        ("rowscan_rows", C.c_uint64),
        ("device_ms", C.c_double),
        ("prep_ms", C.c_double),
        ("coarse_ms", C.c_double),
        ("rescore_ms", C.c_double),
        ("fallback_ms", C.c_double),
        ("nonfinite_row", C.c_uint64),
        ("nonfinite_sample", C.c_uint64),
        ("coarse_err_max", C.c_double),
        ("coarse_eps", C.c_double),
        ("order_ms", C.c_double),
    ]


class Sample(C.Structure):
    _fields_ = [
        ("x", C.c_void_p),
        ("n", C.c_uint64),
        ("d", C.c_uint64),
        ("idx", C.c_void_p),
        ("cos", C.c_void_p),
        ("norms", C.c_void_p),
        ("weights", C.c_void_p),
        ("margin", C.c_int64),
        ("col_splits", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


NO_ROW = (1 << 64) - 1


class CostalError(ValueError):
    """An error returned by the library; ``code`` is one of the E_* values.  For E_NON_FINITE,
    ``row`` (and for a batch ``sample``) name the first row holding a NaN or inf."""

    def __init__(self, code: int, msg: str, row: int | None = None, sample: int | None = None):
        super().__init__(msg)
        self.code = code
        self.row = row
        self.sample = sample


_lib = None


def _one_hip_runtime() -> None:
    """Load the HIP runtime that PyTorch ships, when PyTorch is installed, before the library.

    The library needs ``libamdhip64.so.7`` and binds to an already loaded runtime of that soname;
    PyTorch's extensions need ``libamdhip64.so`` from their own directory, a copy with the same
    soname.  Loaded in the other order (the library first), the process would hold two HIP (and
    HSA) runtimes and PyTorch would find no GPU.  With PyTorch's copy loaded first by path, both
    bind to it, whichever of the two touches the GPU first.  Nothing is initialised here."""
    import importlib.util

    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    for d in (spec.submodule_search_locations or []) if spec else []:
        path = os.path.join(d, "lib", "libamdhip64.so")
        if os.path.exists(path):
            C.CDLL(path, mode=C.RTLD_GLOBAL)
            return


def load():
    """Load the library (once). Raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc --offload-arch=gfx950)"
        )
    _one_hip_runtime()
    L = C.CDLL(LIB_PATH)
    P, U64, U32 = C.c_void_p, C.c_uint64, C.c_uint32
    L.costal_last_error.restype = C.c_char_p
    L.costal_version.restype = C.c_char_p
    L.costal_device_info.argtypes = [C.c_char_p, C.c_size_t]
    L.costal_knn_exact.argtypes = [P, U64, U64, C.POINTER(KnnParams), P, P, P, P, C.POINTER(KnnStats)]
    L.costal_knn_exact_device.argtypes = [P, U64, U64, C.POINTER(KnnParams), P, P, P, P,
                                          C.POINTER(KnnStats), P]
    L.costal_knn_bruteforce.argtypes = [P, U64, U64, U64, U32, P, P, P]
    L.costal_knn_exact_device_noalloc.argtypes = [P, U64, U64, C.POINTER(KnnParams), P, P, P, P,
                                                  C.POINTER(KnnStats), P]
    L.costal_mfma_probe.argtypes = [P]
    L.costal_csr_from_edges.argtypes = [U64, U64, P, P, P, P, P, P, C.POINTER(C.c_uint64)]
    L.costal_knn_plan.argtypes = [U64, U64, C.POINTER(KnnParams), C.POINTER(KnnStats)]
    L.costal_knn_batched_device.argtypes = [C.POINTER(Sample), U32, C.POINTER(KnnParams), C.POINTER(KnnStats), P]
    L.costal_reweight.argtypes = [U64, U64, P, P, P, U32, P]
    L.costal_reweight_device.argtypes = [U64, U64, P, P, P, U32, P, P]
    L.costal_fuzzy_union.argtypes = [U64, U64, P, P, P, P, P, C.POINTER(C.c_uint64)]
    F = C.c_float
    L.costal_mlp_param_count.restype = U64
    L.costal_mlp_param_count.argtypes = [U64]
    L.costal_mlp_act_floats.restype = U64
    L.costal_mlp_act_floats.argtypes = [U64]
    L.costal_mlp_predict_device.argtypes = [P, U64, U64, P, U32, F, F, U32, P, P]
    L.costal_mlp_grad_device.argtypes = [P, P, U64, U64, P, P, P, P, P]
    L.costal_mlp_adam_device.argtypes = [U64, P, P, P, P, F, F, F, F, U64, P]
    L.costal_mlp_last_error.restype = C.c_char_p
    for name in EXPORTS:
        getattr(L, name).restype = getattr(L, name).restype or C.c_int
    L.costal_last_error.restype = C.c_char_p
    L.costal_version.restype = C.c_char_p
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != OK:
        raise CostalError(rc, load().costal_last_error().decode())


def version() -> str:
    return load().costal_version().decode()


def device_info() -> str:
    buf = C.create_string_buffer(256)
    check(load().costal_device_info(buf, 256))
    return buf.value.decode()
