"""Loader and ctypes declarations of libsfp_hip.so (include/sfp.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``sfparticles_amd/csrc/build.sh``).  Nothing here falls back to a CPU path:
every compute entry point raises ``SFPError`` when the library or a HIP
device is missing.
"""
import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsfp_hip.so")

SFP_OK = 0
SFP_E_HIP = 1
SFP_E_ARG = 2
SFP_E_CAPACITY = 3
SFP_E_STATE = 4

# store attributes (enum sfp_attr)
X, Y, Z, UX, UY, UZ, ID, PRUNED, TAU, SX, SY, SZ, INV_GAMMA, CHI, EX, EY, EZ, BX, BY, BZ, EVENT, DELTA, SHOT = range(23)
NATTR = 23
STORE_TAU = 1
STORE_SPIN = 2

RR_NONE, RR_PHOTON, RR_LL, RR_CLL = 0, 1, 2, 3
QED_RS, QED_OD = 0, 1
RUN_TIMING = 1
RUN_GRAPH = 2
RUN_AUTOTUNE = 4


class SFPError(RuntimeError):
    """An error reported by libsfp_hip.so (or its absence)."""


class CapacityError(SFPError):
    pass


class Field(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("p", C.c_double * 12)]


class RSTablesC(C.Structure):
    _fields_ = [("int_Ai", C.c_void_p), ("Aip", C.c_void_p), ("z_N", C.c_int64),
                ("z_min", C.c_double), ("z_max", C.c_double), ("z_delta", C.c_double),
                ("Ai", C.c_void_p)]


class ODTablesC(C.Structure):
    _fields_ = [("photon_total", C.c_void_p), ("pair_total", C.c_void_p),
                ("photon_cdf", C.c_void_p), ("pair_cdf", C.c_void_p), ("chi_N", C.c_int64),
                ("log_chi_min", C.c_double), ("log_chi_max", C.c_double),
                ("log_chi_delta", C.c_double), ("delta_N", C.c_int64),
                ("log_delta_min", C.c_double), ("log_delta_max", C.c_double),
                ("log_delta_delta", C.c_double), ("chi_grid", C.c_void_p),
                ("delta_grid", C.c_void_p), ("chi_min", C.c_double), ("chi_max", C.c_double)]


class SpeciesDesc(C.Structure):
    _fields_ = [("store", C.c_void_p), ("q", C.c_double), ("m", C.c_double), ("ae", C.c_double),
                ("rr", C.c_int32), ("push", C.c_int32), ("spin", C.c_int32),
                ("photon_dst", C.c_int32), ("pair_e_dst", C.c_int32), ("pair_p_dst", C.c_int32),
                ("pad", C.c_int32 * 2)]


class SimParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("photon_threshold", C.c_double), ("seed", C.c_uint64),
                ("qed_model", C.c_int32), ("precision", C.c_int32),
                ("spin_emission", C.c_int32), ("pad", C.c_int32)]


P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int32
U32 = C.c_uint32
U64 = C.c_uint64
D = C.c_double

# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "sfp_last_error": [],
    "sfp_version": [],
    "sfp_device_info": [C.POINTER(I32), C.c_char_p, I32],
    "sfp_malloc": [C.POINTER(P), I64],
    "sfp_free": [P],
    "sfp_memcpy_htod": [P, P, I64],
    "sfp_memcpy_dtoh": [P, P, I64],
    "sfp_synchronize": [],
    "sfp_mem_info": [C.POINTER(I64), C.POINTER(I64)],
    "sfp_tables_create": [P, P, C.POINTER(P)],
    "sfp_tables_destroy": [P],
    "sfp_eval_field": [P, I32, P, P, P, D, I64, P, P, P, P, P, P, P, P],
    "sfp_push_position": [P] * 7 + [I64, P, D, P],
    "sfp_boris": [P] * 10 + [D, I64, P, D, P],
    "sfp_boris_tbmt": [P] * 13 + [D, D, I64, P, D, P],
    "sfp_LL_push": [P] * 5 + [I64, P, D, P],
    "sfp_CLL_push": [P] * 5 + [I64, P, D, P],
    "sfp_update_chi": [P] * 11 + [I64, P, P],
    "sfp_pick_hard_photon": [P, P, P, D, I64, P],
    "sfp_photon_recoil": [P] * 6 + [I64, P, P],
    "sfp_create_photon": [P] * 16 + [I64, I64, P, U64, P, P, P],
    "sfp_create_pair": [P] * 17 + [I64, I64, I32, P, U64, P, P, P, P, P, P],
    "sfp_find_event_index": [P, I64, P, C.POINTER(I64), P],
    "sfp_bool_sum": [P, I64, C.POINTER(I64), P],
    "sfp_find_event_index_noalloc": [P, I64, P, P, P],
    "sfp_sim_stats_n": [P, I32, P, I32],
    "sfp_sim_set_flag_clear_every": [P, I64],
    "sfp_sim_set_shared_launches": [P, I32],
    "sfp_sim_qed_diag": [P, I32, P, C.POINTER(D)],
    "sfp_photon_from_rejection_sampling": [P, P, D, I64, P, P, P, P, U64, U64, U32, U32, P, P],
    "sfp_pair_from_rejection_sampling": [P, P, D, I64, P, P, P, P, U64, U64, U32, U32, P, P],
    "sfp_update_tau_e": [P, P, P, D, I64, P, P, P, P, U64, U64, U32, U32, P, P],
    "sfp_update_tau_gamma": [P, P, P, D, I64, P, P, P, P, U64, U64, U32, U32, P, P],
    "sfp_update_inv_gamma": [P, P, P, P, I64, P, I32, P],
    "sfp_math_eval": [I32, P, P, P, I64, P],
    "sfp_field_bound_eval": [P, I32, P, P, P, D, P, P, P, D, I64, P, P],
    "sfp_mlp_create": [C.POINTER(P)],
    "sfp_mlp_create_rate": [C.POINTER(P)],
    "sfp_mlp_shape": [P, C.POINTER(I32), C.POINTER(I32)],
    "sfp_mlp_rate": [P, P, P, I64, I32, P, P],
    "sfp_sim_set_rate_net": [P, P],
    "sfp_sim_set_recorder": [P, I32, I64, I64, I64, U64],
    "sfp_sim_get_records": [P, P, P, C.POINTER(I64)],
    "sfp_mlp_destroy": [P],
    "sfp_mlp_set_params": [P, P, P],
    "sfp_mlp_get_params": [P, P, P],
    "sfp_mlp_set_fallback": [P, C.c_uint64],
    "sfp_mlp_make_data": [P, P, I64, C.c_uint64, P, P],
    "sfp_mlp_train": [P, P, P, I64, I32, I32, C.c_float, P],
    "sfp_mlp_grad": [P, P, P, I64, P, P],
    "sfp_mlp_forward": [P, P, P, I64],
    "sfp_mlp_sample": [P, P, P, I64, I32, P, P],
    "sfp_mlp_device": [P, C.POINTER(P), C.POINTER(P)],
    "sfp_sim_set_sampler": [P, P],
    "sfp_sim_set_shots": [P, P, P, I32],
    "sfp_store_create": [I32, I64, C.POINTER(P)],
    "sfp_store_destroy": [P],
    "sfp_store_reserve": [P, I64],
    "sfp_store_capacity": [P, C.POINTER(I64)],
    "sfp_store_ptr": [P, I32, C.POINTER(P)],
    "sfp_store_set_count": [P, I64, U64],
    "sfp_store_get_count": [P, C.POINTER(I64), C.POINTER(U64)],
    "sfp_store_upload": [P, I32, P, I64, I64],
    "sfp_store_download": [P, I32, P, I64, I64],
    "sfp_store_prune": [P, C.POINTER(I64)],
    "sfp_sim_create": [P, I32, P, I32, P, P, C.POINTER(P)],
    "sfp_sim_destroy": [P],
    "sfp_sim_run": [P, C.POINTER(D), U64, I64, I32],
    "sfp_sim_kernel_time": [P, I32, I32, C.POINTER(D), C.POINTER(I64)],
    "sfp_sim_reset_timing": [P],
    "sfp_sim_set_timing_every": [P, I64],
    "sfp_sim_set_timing_kernel": [P, I32, I32],
    "sfp_sim_get_counts": [P, I32, P, P],
    "sfp_store_histogram": [P, I32, I32, I32, D, D, P],
    "sfp_sim_set_grid": [P, I32, I32, I32],
    "sfp_sim_set_shape": [P, I32, I32, I32, I32],
    "sfp_sim_launch_shape": [P, I32, I32, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32),
                             C.POINTER(I32)],
    "sfp_sim_stats": [P, I32, C.POINTER(I64)],
    "sfp_sim_reset_stats": [P],
    "sfp_copy_f64": [P, P, I64, P],
}
_RESTYPES = {"sfp_last_error": C.c_char_p}

_lock = threading.Lock()
_lib = None
_device_checked = False


def load(path=LIB_PATH):
    """Load the library (no device access)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise SFPError(
                    f"{path} is not built; run __graft_entry__.build() or "
                    "sfparticles_amd/csrc/build.sh (the HIP extension is required: "
                    "there is no CPU fallback)")
            L = C.CDLL(path)
            for name, args in SIGNATURES.items():
                fn = getattr(L, name)
                fn.argtypes = args
                fn.restype = _RESTYPES.get(name, C.c_int)
            _lib = L
    return _lib


def lib():
    """The library, after checking that a HIP device is usable."""
    global _device_checked
    L = load()
    if not _device_checked:
        cus = I32(0)
        name = C.create_string_buffer(64)
        rc = L.sfp_device_info(C.byref(cus), name, 64)
        if rc != SFP_OK:
            raise SFPError("no usable HIP device: " + L.sfp_last_error().decode())
        _device_checked = True
    return L


def check(rc):
    if rc != SFP_OK:
        msg = load().sfp_last_error().decode()
        if rc == SFP_E_CAPACITY:
            raise CapacityError(msg)
        raise SFPError(msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))


def device_info():
    cus = I32(0)
    name = C.create_string_buffer(64)
    check(lib().sfp_device_info(C.byref(cus), name, 64))
    return int(cus.value), name.value.decode()


class DeviceArray:
    """A library-allocated device buffer mirroring a host numpy array."""

    def __init__(self, nbytes):
        p = P()
        call("sfp_malloc", C.byref(p), int(nbytes))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.sfp_free(self.ptr)
        except Exception:
            pass
        self.ptr = None

    def upload(self, arr):
        assert arr.flags["C_CONTIGUOUS"] and arr.nbytes <= self.nbytes
        call("sfp_memcpy_htod", self.ptr, arr.ctypes.data, arr.nbytes)

    def download(self, arr):
        assert arr.flags["C_CONTIGUOUS"] and arr.nbytes <= self.nbytes
        call("sfp_memcpy_dtoh", arr.ctypes.data, self.ptr, arr.nbytes)

    @classmethod
    def of(cls, arr):
        d = cls(max(arr.nbytes, 1))
        if arr.nbytes:
            d.upload(arr)
        return d
