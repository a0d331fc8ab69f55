"""ctypes bindings of the CPU oracle (oracle.c).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()`` and
by the CPU-baseline leg of ``bench.py`` -- never by the product package.

Field terms are ``(kind, params)`` pairs with the reference builders' user
arguments (see ``orc_field`` in oracle.c); kinds: 0 static, 1 simple,
2 gaussian, 3 focused.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_libs = {}


def build():
    """Compile liboracle*.so from oracle.c (gcc)."""
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib(libm=False):
    name = "liboracle_libm.so" if libm else "liboracle.so"
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        _declare(L)
        _libs[name] = L
    return _libs[name]


P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int32
U32 = C.c_uint32
U64 = C.c_uint64
D = C.c_double


class OrcField(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("p", C.c_double * 12)]


class OrcRSTables(C.Structure):
    _fields_ = [("int_Ai", P), ("Aip", P), ("z_N", I64), ("z_min", D), ("z_max", D),
                ("z_delta", D), ("Ai", P)]


class OrcODTables(C.Structure):
    _fields_ = [("photon_total", P), ("pair_total", P), ("photon_cdf", P), ("pair_cdf", P),
                ("chi_N", I64), ("log_chi_min", D), ("log_chi_max", D), ("log_chi_delta", D),
                ("delta_N", I64), ("log_delta_min", D), ("log_delta_max", D),
                ("log_delta_delta", D), ("chi_grid", P), ("delta_grid", P), ("chi_min", D),
                ("chi_max", D)]


def _declare(L):
    L.orc_eval_field.argtypes = [P, I32, P, P, P, D, I64, P, P, P, P, P, P, P]
    L.orc_field_at.argtypes = [P, I32, D, D, D, D, P]
    L.orc_push_position.argtypes = [P] * 7 + [I64, P, D]
    L.orc_boris.argtypes = [P] * 10 + [D, I64, P, D]
    L.orc_boris_tbmt.argtypes = [P] * 13 + [D, D, I64, P, D]
    L.orc_LL_push.argtypes = [P] * 5 + [I64, P, D]
    L.orc_CLL_push.argtypes = [P] * 5 + [I64, P, D]
    L.orc_update_chi.argtypes = [P] * 11 + [I64, P]
    L.orc_photon_recoil.argtypes = [P] * 6 + [I64, P]
    L.orc_create_photon.argtypes = [P] * 16 + [I64, I64, P, U64, P, P]
    L.orc_create_pair.argtypes = [P] * 17 + [I64, I64, I32, P, U64, P, P, P, P, P]
    L.orc_find_event_index.argtypes = [P, I64, I64, P]
    L.orc_find_event_index.restype = I64
    L.orc_bool_sum.argtypes = [P, I64]
    L.orc_bool_sum.restype = I64
    L.orc_pick_hard_photon.argtypes = [P, P, P, D, I64]
    L.orc_photon_prob_rate.argtypes = [D, D, P]
    L.orc_photon_prob_rate.restype = D
    L.orc_pair_prob_rate.argtypes = [D, D, P]
    L.orc_pair_prob_rate.restype = D
    qed = [P, P, D, I64, P, P, P, P, U64, U64, U32, U32, P]
    L.orc_photon_from_rejection_sampling.argtypes = qed
    L.orc_pair_from_rejection_sampling.argtypes = qed
    L.orc_photon_from_rejection_sampling_spin.argtypes = qed + [D] + [P] * 12
    L.orc_photon_prob_rate_spin.argtypes = [D, D, D, P]
    L.orc_photon_prob_rate_spin.restype = D
    L.orc_spin_sigma.argtypes = [D] * 14
    L.orc_spin_sigma.restype = D
    tau = [P, P, P, D, I64, P, P, P, P, U64, U64, U32, U32, P]
    L.orc_update_tau_e.argtypes = tau
    L.orc_update_tau_gamma.argtypes = tau
    L.orc_od_total_rate.argtypes = [P, P, D]
    L.orc_od_total_rate.restype = D
    L.orc_od_inverse_cdf.argtypes = [P, P, D, D]
    L.orc_od_inverse_cdf.restype = D
    L.orc_philox4x32_10.argtypes = [P, P, P]
    L.orc_draw4.argtypes = [U64, U64, U64, U32, U32, P]
    L.orc_math.argtypes = [I32, P, P, P, I64]
    L.orc_num_threads.restype = I32


def ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data


def field_array(terms):
    """Pack [(kind, params), ...] into a ctypes array of OrcField."""
    arr = (OrcField * max(1, len(terms)))()
    for i, (kind, params) in enumerate(terms):
        arr[i].kind = int(kind)
        for j, v in enumerate(params):
            arr[i].p[j] = float(v)
    return arr


def rs_struct(T):
    s = OrcRSTables(T.int_Ai.ctypes.data, T.Aip.ctypes.data, T.z_N, T.z_min, T.z_max, T.z_delta,
                    T.Ai.ctypes.data)
    s._keep = T
    return s


def od_struct(T):
    s = OrcODTables(T.photon_total.ctypes.data, T.pair_total.ctypes.data,
                    T.photon_cdf.ctypes.data, T.pair_cdf.ctypes.data, T.chi_N, T.log_chi_min,
                    T.log_chi_max, T.log_chi_delta, T.delta_N, T.log_delta_min,
                    T.log_delta_max, T.log_delta_delta, T.chi_grid.ctypes.data,
                    T.delta_grid.ctypes.data, T.chi_min, T.chi_max)
    s._keep = T
    return s


def field_at(terms, x, y, z, t, libm=False):
    out = np.zeros(6)
    fa = field_array(terms)
    lib(libm).orc_field_at(C.byref(fa), len(terms), x, y, z, t, ptr(out))
    return out


def eval_field(terms, x, y, z, t, libm=False):
    """Fields at arrays of points (the per-particle field pass, fields.py:26-46);
    returns (Ex, Ey, Ez, Bx, By, Bz) arrays."""
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    n = x.shape[0]
    F = [np.zeros(n) for _ in range(6)]
    pr = np.zeros(n, np.uint8)
    fa = field_array(terms)
    lib(libm).orc_eval_field(C.byref(fa), len(terms), ptr(x), ptr(y), ptr(z), float(t), n, ptr(pr),
                             *[ptr(a) for a in F])
    return F


def math(fn, x, y=None):
    """Own elementary functions: fn in exp, log, sin, cos, atan, pow, exp10, log10."""
    names = ["exp", "log", "sin", "cos", "atan", "pow", "exp10", "log10"]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_math(names.index(fn), ptr(x), ptr(y), ptr(out), x.size)
    return out


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(ptr(c), ptr(k), ptr(out))
    return out


def draw4(seed, pid, step, tag, slot):
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_draw4(seed, pid, step, tag, slot, ptr(out))
    return out


def u01(hi, lo):
    k = (int(hi) << 20) | (int(lo) >> 12)
    return (k + 0.5) * 2.0 ** -52


def num_threads():
    return int(lib().orc_num_threads())
