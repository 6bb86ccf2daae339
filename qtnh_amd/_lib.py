"""ctypes binding of libqtnh_hip.so (the C ABI declared in include/qtnh_hip.h).

The library is the only compute path: if it is missing, or there is no GPU,
every compute call raises.  There is no CPU fallback.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libqtnh_hip.so")

QTNH_OK, QTNH_EINVAL, QTNH_ENOMEM, QTNH_EHIP, QTNH_EUNSUPPORTED = range(5)
X = 0xFFFF
PREC_FP64, PREC_SPLIT = 0, 1
NOALLOC_SCRATCH_BYTES = 64 << 20  # QTNH_NOALLOC_SCRATCH_BYTES
RUN_PER_GATE, RUN_FUSED = 0, 1
QUERY_AMPLITUDE, QUERY_EXPECT_Z = 0, 1
SERVE_SEQUENTIAL, SERVE_BATCHED, SERVE_SEQUENTIAL_FUSED, SERVE_NETWORK = 0, 1, 2, 3
SERVE_SPLIT, SERVE_ALL_NETWORK = 0x100, 0x200  # flags of SERVE_NETWORK


class Query(ctypes.Structure):
    _fields_ = [("n_qubits", ctypes.c_int), ("depth", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("kind", ctypes.c_int), ("basis", ctypes.c_uint64), ("arrival_s", ctypes.c_double)]


class QueryResult(ctypes.Structure):
    _fields_ = [("start_s", ctypes.c_double), ("done_s", ctypes.c_double), ("value", ctypes.c_double * 2),
                ("zsum", ctypes.c_double), ("batch", ctypes.c_int), ("cached", ctypes.c_int), ("route", ctypes.c_int)]
DIST_NORMAL, DIST_UNIFORM = 0, 1

_p = ctypes.c_void_p
_i = ctypes.c_int
_u64 = ctypes.c_uint64
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_dp = ctypes.POINTER(ctypes.c_double)

# name -> argtypes; every function returns int (qtnh_status) unless listed in _RESTYPE.
SIGNATURES = {
    "qtnh_ctx_create": [_i, _p, ctypes.POINTER(_p)],
    "qtnh_ctx_destroy": [_p],
    "qtnh_ctx_trim": [_p],
    "qtnh_status_string": [_i],
    "qtnh_ctx_last_error": [_p],
    "qtnh_device_count": [ctypes.POINTER(_i)],
    "qtnh_buf_alloc": [_p, _u64, ctypes.POINTER(_p)],
    "qtnh_buf_free": [_p, _p],
    "qtnh_h2d": [_p, _p, _p, _u64],
    "qtnh_d2h": [_p, _p, _p, _u64],
    "qtnh_d2d": [_p, _p, _p, _u64],
    "qtnh_sync": [_p],
    "qtnh_fill_c128": [_p, _p, _u64, _u64, _i],
    "qtnh_default_repls": [_i, _i, _i, _i, _i, _u16p, _u16p, _u16p],
    "qtnh_symm_repls": [_i, _i, _i, _i, _i, _u16p, _u16p, _u16p, ctypes.POINTER(_i)],
    "qtnh_contract_dims": [_i, _u64p, _i, _u64p, _i, _u16p, _u16p, _u16p, ctypes.POINTER(_i), _u64p],
    "qtnh_permute_c128": [_p, _p, _p, _i, _u64p, _u16p],
    "qtnh_swap_c128": [_p, _p, _i, _u64p, _i, _i],
    "qtnh_gate_apply_c128": [_p, _p, _i, _u64p, _i, _u16p, _dp, _i],
    "qtnh_diag_apply_c128": [_p, _p, _i, _u64p, _i, _u16p, _dp],
    "qtnh_contract_c128": [_p, _p, _i, _u64p, _p, _i, _u64p, _i, _u16p, _u16p, _u16p, _p, _i, _dp],
    "qtnh_contract_c128_noalloc": [_p, _p, _i, _u64p, _p, _i, _u64p, _i, _u16p, _u16p, _u16p, _p, _i, _dp],
    "qtnh_norm2_c128": [_p, _p, _u64, _dp],
    "qtnh_scale_c128": [_p, _p, _u64, ctypes.c_double, ctypes.c_double],
    "qtnh_tune_permute": [_i, _i, _i, _i, _i, _i],
    "qtnh_tune_permute_auto": [_i],
    "qtnh_tune_zgemm": [_i],
    "qtnh_tune_split": [_i],
    "qtnh_tune_split_tile": [_i],
    "qtnh_tune_circuit": [_i, _i],
    "qtnh_tune_stream": [_i],
    "qtnh_tune_circuit_pipeline": [_i],
    "qtnh_circuit_create": [_i, ctypes.POINTER(_p)],
    "qtnh_circuit_destroy": [_p],
    "qtnh_circuit_gate": [_p, _i, _u16p, _dp],
    "qtnh_circuit_diag": [_p, _i, _u16p, _dp],
    "qtnh_circuit_qft": [_p, _dp],
    "qtnh_circuit_random": [_p, _i, _u64],
    "qtnh_circuit_info": [_p, ctypes.POINTER(_i), _u64p, _u64p],
    "qtnh_circuit_op": [_p, _u64, ctypes.POINTER(_i), ctypes.POINTER(_i), _u16p, _dp],
    "qtnh_circuit_run": [_p, _p, _p, _i],
    "qtnh_circuit_run_batch": [_p, _i, ctypes.POINTER(_p), ctypes.POINTER(_p)],
    "qtnh_expect_z_c128": [_p, _p, _i, _i, _u16p, _dp],
    "qtnh_circuit_grid": [_p, _i, _i, _i, _u64],
    "qtnh_net_create": [ctypes.POINTER(_p)],
    "qtnh_net_destroy": [_p],
    "qtnh_net_add_tensor": [_p, _i, _u64p, ctypes.POINTER(ctypes.c_int64), _dp, ctypes.POINTER(_i)],
    "qtnh_net_add_tensor_device": [_p, _i, _u64p, ctypes.POINTER(ctypes.c_int64), _p, ctypes.POINTER(_i)],
    "qtnh_net_plan_steps": [_p, _i, _p],
    "qtnh_net_plan_best": [_p, _u64, _i, _i, _u64p],
    "qtnh_net_batch_stats": [_p, _u64p, _u64p, _u64p],
    "qtnh_net_run_device": [_p, _p, _p],
    "qtnh_net_from_circuit": [_p, _p, ctypes.POINTER(ctypes.c_uint8)],
    "qtnh_net_plan": [_p, _u64, _i],
    "qtnh_net_set_precision": [_p, _i],
    "qtnh_net_info": [_p, _u64p, _u64p, ctypes.POINTER(_i), _dp, _dp, _u64p],
    "qtnh_net_result_dims": [_p, ctypes.POINTER(_i), _u64p],
    "qtnh_net_step": [_p, _u64, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)],
    "qtnh_net_run": [_p, _p, _dp, _u64],
    "qtnh_net_run_batch": [_p, _i, ctypes.POINTER(_p), ctypes.POINTER(_dp), _u64, ctypes.POINTER(_i)],
    "qtnh_serve": [_p, _i, _p, _p, _u64, _i],
    "qtnh_net_batch_create": [_p, _i, ctypes.POINTER(_p), _i, ctypes.POINTER(_p)],
    "qtnh_net_batch_set_leaf": [_p, _p, _i, _i, _dp],
    "qtnh_net_batch_run": [_p, _p, ctypes.POINTER(_dp)],
    "qtnh_net_batch_destroy": [_p],
    "qtnh_autotune_circuit": [_p, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), _dp],
    "qtnh_circuit_tune": [_p, _p, _p, ctypes.POINTER(_i), ctypes.POINTER(_i), _dp],
}
_RESTYPE = {"qtnh_status_string": ctypes.c_char_p, "qtnh_ctx_last_error": ctypes.c_char_p,
            "qtnh_tune_permute": None, "qtnh_tune_permute_auto": None, "qtnh_tune_zgemm": None,
            "qtnh_tune_split": None, "qtnh_tune_split_tile": None, "qtnh_tune_circuit": None,
            "qtnh_tune_stream": None, "qtnh_tune_circuit_pipeline": None}

_LIB = None


class QtnhError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("qtnh status %d (%s): %s" % (status, _STATUS_NAMES.get(status, "?"), msg))
        self.status = status


_STATUS_NAMES = {0: "ok", 1: "invalid argument", 2: "out of device memory", 3: "HIP runtime error",
                 4: "unsupported"}


def lib():
    """Load libqtnh_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libqtnh_hip.so not found at %s; build it with `python -m qtnh_amd.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, ctypes.c_int)
        _LIB = L
    return _LIB


def check(status, ctx=None):
    if status != QTNH_OK:
        msg = lib().qtnh_ctx_last_error(ctx).decode() if ctx else ""
        raise QtnhError(status, msg)


def device_count():
    n = ctypes.c_int(0)
    check(lib().qtnh_device_count(ctypes.byref(n)))
    return n.value
