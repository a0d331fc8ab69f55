"""The reference's backend callables (sfparticles/cpu.py:5-232) on the GPU.

Same names, argument order and in-place semantics as the reference's CPU
backend, for numpy arguments: inputs are copied to the device, the HIP kernel
runs, outputs are copied back into the caller's arrays.  ``Particles`` does
not use these wrappers -- it keeps its arrays on the device and calls the
library directly.
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _dev(a, dtype):
    a = np.asarray(a)
    assert a.dtype == dtype, f"expected {dtype}, got {a.dtype}"
    return L.DeviceArray.of(np.ascontiguousarray(a))


def _back(d, a):
    if isinstance(a, np.ndarray) and a.size:
        tmp = np.empty_like(a)
        d.download(tmp)
        a[...] = tmp


def _f(a):
    return _dev(a, np.float64)


def _b(a):
    return _dev(np.asarray(a).view(np.uint8) if np.asarray(a).dtype == np.bool_ else a, np.uint8)


def _back_bool(d, a):
    if a.size:
        tmp = np.empty(a.shape, dtype=np.uint8)
        d.download(tmp)
        a[...] = tmp.astype(bool)


class _Dev:
    """Device copies of one call's arguments.  Each copy lives as long as this
    object, i.e. to the end of the wrapper, after the synchronous copies back:
    a temporary whose pointer alone were passed would be freed before the
    asynchronous kernel reads it."""

    def __init__(self):
        self.keep = []

    def f(self, a):
        d = _f(a)
        self.keep.append(d)
        return d.ptr

    def b(self, a):
        d = _b(a)
        self.keep.append(d)
        return d.ptr

    def i64(self, a):
        d = _dev(a, np.int64)
        self.keep.append(d)
        return d.ptr

    def u64(self, a):
        d = _dev(a, np.uint64)
        self.keep.append(d)
        return d.ptr

    def out(self, a, kind="f"):
        """A device copy of an in/out array, copied back by back()."""
        d = _f(a) if kind == "f" else _b(a)
        self.keep.append(d)
        self._outs = getattr(self, "_outs", []) + [(d, a, kind)]
        return d.ptr

    def back(self):
        for d, a, kind in getattr(self, "_outs", []):
            (_back if kind == "f" else _back_bool)(d, a)


def push_position(x, y, z, ux, uy, uz, inv_gamma, N, to_be_pruned, dt):
    """cpu.py:5-19"""
    D = _Dev()
    L.call("sfp_push_position", D.out(x), D.out(y), D.out(z), D.f(ux), D.f(uy), D.f(uz),
           D.f(inv_gamma), int(N), D.b(to_be_pruned), float(dt), None)
    D.back()


def boris(ux, uy, uz, inv_gamma, Ex, Ey, Ez, Bx, By, Bz, q, N, to_be_pruned, dt):
    """cpu.py:22-31"""
    D = _Dev()
    L.call("sfp_boris", *[D.out(a) for a in (ux, uy, uz, inv_gamma)],
           *[D.f(a) for a in (Ex, Ey, Ez, Bx, By, Bz)], float(q), int(N), D.b(to_be_pruned),
           float(dt), None)
    D.back()


def boris_tbmt(ux, uy, uz, inv_gamma, sx, sy, sz, Ex, Ey, Ez, Bx, By, Bz, q, ae, N, to_be_pruned,
               dt):
    """cpu.py:33-42"""
    D = _Dev()
    L.call("sfp_boris_tbmt", *[D.out(a) for a in (ux, uy, uz, inv_gamma, sx, sy, sz)],
           *[D.f(a) for a in (Ex, Ey, Ez, Bx, By, Bz)], float(q), float(ae), int(N),
           D.b(to_be_pruned), float(dt), None)
    D.back()


def _ll(name, ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned, dt):
    D = _Dev()
    L.call(name, *[D.out(a) for a in (ux, uy, uz, inv_gamma)], D.f(chi_e), int(N),
           D.b(to_be_pruned), float(dt), None)
    D.back()


def LL_push(ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned, dt):
    """cpu.py:44-51"""
    _ll("sfp_LL_push", ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned, dt)


def CLL_push(ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned, dt):
    """cpu.py:53-60"""
    _ll("sfp_CLL_push", ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned, dt)


def update_chi(Ex, Ey, Ez, Bx, By, Bz, ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned):
    """cpu.py:62-75"""
    D = _Dev()
    L.call("sfp_update_chi", *[D.f(a) for a in (Ex, Ey, Ez, Bx, By, Bz, ux, uy, uz, inv_gamma)],
           D.out(chi_e), int(N), D.b(to_be_pruned), None)
    D.back()


def photon_recoil(ux, uy, uz, inv_gamma, event, photon_delta, N, to_be_pruned):
    """cpu.py:78-88"""
    D = _Dev()
    L.call("sfp_photon_recoil", *[D.out(a) for a in (ux, uy, uz, inv_gamma)], D.b(event),
           D.f(photon_delta), int(N), D.b(to_be_pruned), None)
    D.back()


def pick_hard_photon(event, delta, inv_gamma, threshold, N):
    """cpu.py:227-232"""
    D = _Dev()
    L.call("sfp_pick_hard_photon", D.out(event, "b"), D.f(delta), D.f(inv_gamma),
           float(threshold), int(N), None)
    D.back()


def create_photon(x_src, y_src, z_src, ux_src, uy_src, uz_src, x_dst, y_dst, z_dst, ux_dst, uy_dst,
                  uz_dst, inv_gamma_dst, photon_to_be_pruned, event_index, photon_delta, N_buffered,
                  N_photon):
    """cpu.py:90-118"""
    D = _Dev()
    L.call("sfp_create_photon", *[D.f(a) for a in (x_src, y_src, z_src, ux_src, uy_src, uz_src)],
           *[D.out(a) for a in (x_dst, y_dst, z_dst, ux_dst, uy_dst, uz_dst, inv_gamma_dst)],
           D.out(photon_to_be_pruned, "b"), D.i64(event_index), D.f(photon_delta),
           int(N_buffered), int(N_photon), None, 0, None, None, None)
    D.back()


def create_pair(x_src, y_src, z_src, ux_src, uy_src, uz_src, photon_to_be_pruned, x_dst, y_dst,
                z_dst, ux_dst, uy_dst, uz_dst, inv_gamma_dst, pair_to_be_pruned, event_index,
                pair_delta, N_buffered, N_pair, inverse_delta):
    """cpu.py:120-191"""
    D = _Dev()
    L.call("sfp_create_pair", *[D.f(a) for a in (x_src, y_src, z_src, ux_src, uy_src, uz_src)],
           D.out(photon_to_be_pruned, "b"),
           *[D.out(a) for a in (x_dst, y_dst, z_dst, ux_dst, uy_dst, uz_dst, inv_gamma_dst)],
           D.out(pair_to_be_pruned, "b"), D.i64(event_index), D.f(pair_delta), int(N_buffered),
           int(N_pair), int(bool(inverse_delta)), None, 0, None, None, None, None, None, None)
    D.back()


def find_event_index(event, N_event, N):
    """cpu.py:193-216 (ascending indices of the events)"""
    D = _Dev()
    out = L.DeviceArray(max(8 * int(N), 8))
    cnt = C.c_int64(0)
    L.call("sfp_find_event_index", D.b(event), int(N), out.ptr, C.byref(cnt), None)
    res = np.zeros(int(N_event), dtype=np.int64)
    n = min(int(cnt.value), int(N_event))
    if n:
        tmp = np.empty(int(cnt.value), dtype=np.int64)
        out.download(tmp)
        res[:n] = tmp[:n]
    return res


def bool_sum(event, N):
    """cpu.py:218-225"""
    D = _Dev()
    cnt = C.c_int64(0)
    L.call("sfp_bool_sum", D.b(event), int(N), C.byref(cnt), None)
    return int(cnt.value)


# --------------------------------------------------------------------------
# The QED model seam (particles.py:8-12): the reference's callables with its
# argument order, for numpy arguments.  The reference draws from numba's
# unseeded numpy.random; these draw Philox4x32-10 words keyed by ``seed``
# with counter (id, step, tag, slot), so the extra keyword arguments name the
# stream (defaults: ids 0..N-1, the package seed, step 0, tag 0 and the
# per-operation slot).
def _qed_stream(D, N, ids, seed, step, tag, slot):
    from . import qed
    ids = np.arange(int(N), dtype=np.uint64) if ids is None else np.asarray(ids, np.uint64)
    return (D.u64(ids), qed.seed if seed is None else int(seed), int(step), int(tag),
            qed.PER_OP_SLOT if slot is None else int(slot), qed.device_tables())


def _rs(name, inv_gamma, chi, dt, N, to_be_pruned, event, delta, ids, seed, step, tag, slot):
    D = _Dev()
    stream = _qed_stream(D, N, ids, seed, step, tag, slot)
    L.call(name, D.f(inv_gamma), D.f(chi), float(dt), int(N), D.b(to_be_pruned),
           D.out(event, "b"), D.out(delta), *stream, None)
    D.back()


def photon_from_rejection_sampling(inv_gamma, chi_e, dt, N, to_be_pruned, event, photon_delta,
                                   ids=None, seed=None, step=0, tag=0, slot=None):
    """qed/rejection_sampling.py:68-85"""
    _rs("sfp_photon_from_rejection_sampling", inv_gamma, chi_e, dt, N, to_be_pruned, event,
        photon_delta, ids, seed, step, tag, slot)


def pair_from_rejection_sampling(inv_gamma, chi_gamma, dt, N, to_be_pruned, event, pair_delta,
                                 ids=None, seed=None, step=0, tag=0, slot=None):
    """qed/rejection_sampling.py:87-101"""
    _rs("sfp_pair_from_rejection_sampling", inv_gamma, chi_gamma, dt, N, to_be_pruned, event,
        pair_delta, ids, seed, step, tag, slot)


def _od(name, tau, inv_gamma, chi, dt, N, to_be_pruned, event, delta, ids, seed, step, tag, slot):
    D = _Dev()
    stream = _qed_stream(D, N, ids, seed, step, tag, slot)
    L.call(name, D.out(tau), D.f(inv_gamma), D.f(chi), float(dt), int(N), D.b(to_be_pruned),
           D.out(event, "b"), D.out(delta), *stream, None)
    D.back()


def update_tau_e(tau_e, inv_gamma, chi_e, dt, N, to_be_pruned, event, photon_delta, ids=None,
                 seed=None, step=0, tag=0, slot=None):
    """qed/optical_depth.py:15-62"""
    _od("sfp_update_tau_e", tau_e, inv_gamma, chi_e, dt, N, to_be_pruned, event, photon_delta, ids,
        seed, step, tag, slot)


def update_tau_gamma(tau_gamma, inv_gamma, chi_gamma, dt, N, to_be_pruned, event, pair_delta,
                     ids=None, seed=None, step=0, tag=0, slot=None):
    """qed/optical_depth.py:64-110"""
    _od("sfp_update_tau_gamma", tau_gamma, inv_gamma, chi_gamma, dt, N, to_be_pruned, event,
        pair_delta, ids, seed, step, tag, slot)
