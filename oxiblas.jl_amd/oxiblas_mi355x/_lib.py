"""ctypes binding of liboxiblas_mi355x.so (the C ABI in include/oxiblas_mi355x.h).

The library is required: importing a compute entry point without it raises
at once, there is no CPU fallback.
"""
import ctypes as C
import os
import threading

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("OXIBLAS_MI355X_LIB", os.path.join(_PKG_DIR, "liboxiblas_mi355x.so"))

OXI_OK = 0
OXI_ERR_HIP = 1
OXI_ERR_ALLOC = 2
OXI_ERR_UNSUPPORTED = 3
OXI_ERR_SINGULAR = 4
OXI_ERR_TIMEOUT = 5
OXI_ERR_NOT_FINITE = 6
OXI_ERR_OVERFLOW = 7
OXI_ERR_NOT_POSDEF = 8

_i64 = C.c_int64
_vp = C.c_void_p
_ch = C.c_char
_dbl = C.c_double
_flt = C.c_float

OXI_PREC_BF16 = 0
OXI_PREC_FP16 = 1
OXI_PREC_FP8 = 2
OXI_PREC_FP32 = 3


class IRInfo(C.Structure):
    """oxi_ir_info: what mixed_solve reports (src/lapack/extended_lapack.jl:219-225)."""
    _fields_ = [
        ("iterations", C.c_int64),
        ("inner_iterations", C.c_int64),
        ("converged", C.c_int32),
        ("fallback", C.c_int32),
        ("reason", C.c_int32),
        ("solver", C.c_int32),
        ("backward_error", C.c_double),
        ("normwise_backward_error", C.c_double),
        ("residual_norm", C.c_double),
        ("factor_ms", C.c_double),
        ("refine_ms", C.c_double),
    ]

class RefineInfo(C.Structure):
    """oxi_refine_info: what refine_solve / expert_solve report
    (src/lapack/expert_refinement.jl:103, advanced_factorizations.jl:644)."""
    _fields_ = [
        ("iterations", C.c_int64),
        ("converged", C.c_int32),
        ("reserved", C.c_int32),
        ("backward_error", C.c_double),
        ("residual_norm", C.c_double),
        ("rcond_estimate", C.c_double),
        ("forward_error_estimate", C.c_double),
    ]


# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "oxi_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "oxi_destroy": (C.c_int, [_vp]),
    "oxi_set_stream": (C.c_int, [_vp, _vp]),
    "oxi_sync": (C.c_int, [_vp]),
    "oxi_last_error": (C.c_char_p, [_vp]),
    "oxi_version": (C.c_char_p, []),
    "oxi_malloc": (C.c_int, [_vp, C.POINTER(_vp), C.c_uint64]),
    "oxi_free": (C.c_int, [_vp, _vp]),
    "oxi_memcpy": (C.c_int, [_vp, _vp, _vp, C.c_uint64, C.c_int]),
    "oxi_dgemm": (C.c_int, [_vp, _ch, _ch, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl, _vp, _i64]),
    "oxi_dgemm_strided_batched": (C.c_int, [_vp, _ch, _ch, _i64, _i64, _i64, _dbl, _vp, _i64, _i64, _vp, _i64, _i64,
                                            _dbl, _vp, _i64, _i64, _i64]),
    "oxi_dgemm_batched": (C.c_int, [_vp, _ch, _ch, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl, _vp, _i64,
                                    _i64]),
    "oxi_dsyrk": (C.c_int, [_vp, _ch, _ch, _i64, _i64, _dbl, _vp, _i64, _dbl, _vp, _i64]),
    "oxi_dlogabsdet": (C.c_int, [_vp, _i64, _vp, _i64, _vp, C.POINTER(_dbl), C.POINTER(_dbl)]),
    "oxi_dgetri": (C.c_int, [_vp, _i64, _vp, _i64, _vp, _vp, _i64]),
    "oxi_dbackward_error": (C.c_int, [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                                      C.POINTER(_dbl)]),
    "oxi_dinverse_norm_estimate": (C.c_int, [_vp, _i64, _vp, _i64, _vp, C.c_int32, C.POINTER(_dbl)]),
    "oxi_drefine_solve": (C.c_int, [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _i64, C.c_int32,
                                    _dbl, C.POINTER(RefineInfo)]),
    "oxi_dexpert_solve": (C.c_int, [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, C.c_int32, _dbl, C.c_int32,
                                    _vp, _i64, _vp, _vp, _vp, C.POINTER(RefineInfo)]),
    "oxi_dpotrf": (C.c_int, [_vp, _ch, _i64, _vp, _i64, C.POINTER(_i64)]),
    "oxi_dpotrs": (C.c_int, [_vp, _ch, _i64, _i64, _vp, _i64, _vp, _i64]),
    "oxi_dtrsm": (C.c_int, [_vp, _ch, _ch, _ch, _ch, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, C.POINTER(_i64)]),
    "oxi_dlaswp": (C.c_int, [_vp, _i64, _vp, _i64, _i64, _i64, _vp, C.c_int]),
    "oxi_dlaswp_noalloc": (C.c_int, [_vp, _i64, _vp, _i64, _i64, _i64, _vp, C.c_int]),
    "oxi_dgetrf": (C.c_int, [_vp, _i64, _i64, _vp, _i64, _vp, _i64, C.POINTER(_i64)]),
    "oxi_dgetrf_copy": (C.c_int, [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, C.POINTER(_i64)]),
    "oxi_dgetrs": (C.c_int, [_vp, _ch, _i64, _i64, _vp, _i64, _vp, _vp, _i64]),
    "oxi_dgetrs_plan_create": (C.c_int, [_vp, _i64, _vp, _i64, _vp, C.POINTER(_vp)]),
    "oxi_dgetrs_plan": (C.c_int, [_vp, _vp, _ch, _i64, _vp, _i64]),
    "oxi_dgetrs_plan_destroy": (C.c_int, [_vp, _vp]),
    "oxi_dgesv_adjoint": (C.c_int, [_vp, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64]),
    "oxi_sgemm": (C.c_int, [_vp, _ch, _ch, _i64, _i64, _i64, _flt, _vp, _i64, _vp, _i64, _flt, _vp, _i64]),
    "oxi_hgemm": (C.c_int, [_vp, _ch, _ch, _i64, _i64, _i64, _flt, _vp, _i64, _vp, _i64, _flt, _vp, _i64]),
    "oxi_gemm_lowp_update": (C.c_int, [_vp, C.c_int, _i64, _i64, _i64, _flt, _vp, _i64, _vp, _i64, _vp, _i64]),
    "oxi_sgetrf_lowp": (C.c_int, [_vp, C.c_int, _i64, _i64, _vp, _i64, _vp, _i64, C.POINTER(_i64)]),
    "oxi_sgetrf_tournament": (C.c_int, [_vp, C.c_int, _i64, _i64, _vp, _i64, _vp, _i64, C.POINTER(_i64)]),
    "oxi_sgetrs": (C.c_int, [_vp, _ch, _i64, _i64, _vp, _i64, _vp, _vp, _i64]),
    "oxi_dgesv_mixed": (C.c_int, [_vp, C.c_int, C.c_int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, C.c_int32,
                                  _dbl, C.c_int32, C.c_int32, C.POINTER(IRInfo)]),
    "oxi_dposv_mixed": (C.c_int, [_vp, C.c_int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, C.c_int32, _dbl,
                                  C.POINTER(IRInfo)]),
    "oxi_dequilibrate": (C.c_int, [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, C.c_int]),
    "oxi_profile_enable": (C.c_int, [_vp, C.c_int]),
    "oxi_profile_read": (C.c_int, [_vp, C.POINTER(_dbl), C.POINTER(_dbl), C.POINTER(_i64)]),
}

_lib = None
_lock = threading.Lock()


class OxiError(RuntimeError):
    """A runtime failure of the HIP library (OXI_ERR_*)."""

    def __init__(self, code, msg):
        super().__init__(f"oxiblas_mi355x error {code}: {msg}")
        self.code = code


def load():
    """Load the shared library and declare every entry point (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `make -C oxiblas.jl_amd/csrc` "
                    "(or __graft_entry__.build()); there is no CPU fallback")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


class Handle:
    """One device and the caller's stream (torch's current stream)."""

    def __init__(self, device=0):
        L = load()
        h = _vp()
        rc = L.oxi_create(device, C.byref(h))
        if rc != 0:
            raise OxiError(rc, f"oxi_create(device={device}) failed")
        self.ptr = h
        self.device = device
        self._L = L

    def bind_stream(self, stream_ptr):
        self._L.oxi_set_stream(self.ptr, _vp(stream_ptr))

    def last_error(self):
        return self._L.oxi_last_error(self.ptr).decode(errors="replace")

    def check(self, rc, what):
        """Map return codes to the reference's exception classes: negative =
        invalid argument (ArgumentError, nothing written)."""
        if rc == 0:
            return
        if rc < 0:
            raise ValueError(f"{what}: invalid argument {-rc}")
        raise OxiError(rc, f"{what}: {self.last_error()}")

    def close(self):
        if self.ptr:
            self._L.oxi_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_handles = {}


def handle(device=None):
    """Per-device handle bound to torch's current stream on that device."""
    import torch
    if device is None:
        device = torch.cuda.current_device()
    h = _handles.get(device)
    if h is None:
        h = Handle(device)
        _handles[device] = h
    h.bind_stream(torch.cuda.current_stream(device).cuda_stream)
    return h
