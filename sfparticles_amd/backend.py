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


def push_position(x, y, z, ux, uy, uz, inv_gamma, N, to_be_pruned, dt):
    """cpu.py:5-19"""
    dx, dy, dz = _f(x), _f(y), _f(z)
    L.call("sfp_push_position", dx.ptr, dy.ptr, dz.ptr, _f(ux).ptr, _f(uy).ptr, _f(uz).ptr,
           _f(inv_gamma).ptr, int(N), _b(to_be_pruned).ptr, float(dt), None)
    for d, a in ((dx, x), (dy, y), (dz, z)):
        _back(d, a)


def boris(ux, uy, uz, inv_gamma, Ex, Ey, Ez, Bx, By, Bz, q, N, to_be_pruned, dt):
    """cpu.py:22-31"""
    d = [_f(a) for a in (ux, uy, uz, inv_gamma)]
    F = [_f(a) for a in (Ex, Ey, Ez, Bx, By, Bz)]
    L.call("sfp_boris", *[v.ptr for v in d], *[v.ptr for v in F], float(q), int(N),
           _b(to_be_pruned).ptr, float(dt), None)
    for v, a in zip(d, (ux, uy, uz, inv_gamma)):
        _back(v, a)


def boris_tbmt(ux, uy, uz, inv_gamma, sx, sy, sz, Ex, Ey, Ez, Bx, By, Bz, q, ae, N, to_be_pruned,
               dt):
    """cpu.py:33-42"""
    arrs = (ux, uy, uz, inv_gamma, sx, sy, sz)
    d = [_f(a) for a in arrs]
    F = [_f(a) for a in (Ex, Ey, Ez, Bx, By, Bz)]
    L.call("sfp_boris_tbmt", *[v.ptr for v in d], *[v.ptr for v in F], float(q), float(ae),
           int(N), _b(to_be_pruned).ptr, float(dt), None)
    for v, a in zip(d, arrs):
        _back(v, a)


def _ll(name, ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned, dt):
    arrs = (ux, uy, uz, inv_gamma)
    d = [_f(a) for a in arrs]
    L.call(name, *[v.ptr for v in d], _f(chi_e).ptr, int(N), _b(to_be_pruned).ptr, float(dt), None)
    for v, a in zip(d, arrs):
        _back(v, a)


def LL_push(ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned, dt):
    """cpu.py:44-51"""
    _ll("sfp_LL_push", ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned, dt)


def CLL_push(ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned, dt):
    """cpu.py:53-60"""
    _ll("sfp_CLL_push", ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned, dt)


def update_chi(Ex, Ey, Ez, Bx, By, Bz, ux, uy, uz, inv_gamma, chi_e, N, to_be_pruned):
    """cpu.py:62-75"""
    dchi = _f(chi_e)
    L.call("sfp_update_chi", *[_f(a).ptr for a in (Ex, Ey, Ez, Bx, By, Bz, ux, uy, uz, inv_gamma)],
           dchi.ptr, int(N), _b(to_be_pruned).ptr, None)
    _back(dchi, chi_e)


def photon_recoil(ux, uy, uz, inv_gamma, event, photon_delta, N, to_be_pruned):
    """cpu.py:78-88"""
    arrs = (ux, uy, uz, inv_gamma)
    d = [_f(a) for a in arrs]
    L.call("sfp_photon_recoil", *[v.ptr for v in d], _b(event).ptr, _f(photon_delta).ptr, int(N),
           _b(to_be_pruned).ptr, None)
    for v, a in zip(d, arrs):
        _back(v, a)


def pick_hard_photon(event, delta, inv_gamma, threshold, N):
    """cpu.py:227-232"""
    de = _b(event)
    L.call("sfp_pick_hard_photon", de.ptr, _f(delta).ptr, _f(inv_gamma).ptr, float(threshold),
           int(N), None)
    _back_bool(de, event)


def create_photon(x_src, y_src, z_src, ux_src, uy_src, uz_src, x_dst, y_dst, z_dst, ux_dst, uy_dst,
                  uz_dst, inv_gamma_dst, photon_to_be_pruned, event_index, photon_delta, N_buffered,
                  N_photon):
    """cpu.py:90-118"""
    src = [_f(a) for a in (x_src, y_src, z_src, ux_src, uy_src, uz_src)]
    dst_arrs = (x_dst, y_dst, z_dst, ux_dst, uy_dst, uz_dst, inv_gamma_dst)
    dst = [_f(a) for a in dst_arrs]
    dp = _b(photon_to_be_pruned)
    L.call("sfp_create_photon", *[v.ptr for v in src], *[v.ptr for v in dst], dp.ptr,
           _dev(event_index, np.int64).ptr, _f(photon_delta).ptr, int(N_buffered), int(N_photon),
           None, 0, None, None, None)
    for v, a in zip(dst, dst_arrs):
        _back(v, a)
    _back_bool(dp, photon_to_be_pruned)


def create_pair(x_src, y_src, z_src, ux_src, uy_src, uz_src, photon_to_be_pruned, x_dst, y_dst,
                z_dst, ux_dst, uy_dst, uz_dst, inv_gamma_dst, pair_to_be_pruned, event_index,
                pair_delta, N_buffered, N_pair, inverse_delta):
    """cpu.py:120-191"""
    src = [_f(a) for a in (x_src, y_src, z_src, ux_src, uy_src, uz_src)]
    sp = _b(photon_to_be_pruned)
    dst_arrs = (x_dst, y_dst, z_dst, ux_dst, uy_dst, uz_dst, inv_gamma_dst)
    dst = [_f(a) for a in dst_arrs]
    dp = _b(pair_to_be_pruned)
    L.call("sfp_create_pair", *[v.ptr for v in src], sp.ptr, *[v.ptr for v in dst], dp.ptr,
           _dev(event_index, np.int64).ptr, _f(pair_delta).ptr, int(N_buffered), int(N_pair),
           int(bool(inverse_delta)), None, 0, None, None, None, None, None, None)
    for v, a in zip(dst, dst_arrs):
        _back(v, a)
    _back_bool(dp, pair_to_be_pruned)
    _back_bool(sp, photon_to_be_pruned)


def find_event_index(event, N_event, N):
    """cpu.py:193-216 (ascending indices of the events)"""
    de = _b(event)
    out = L.DeviceArray(max(8 * int(N), 8))
    cnt = C.c_int64(0)
    L.call("sfp_find_event_index", de.ptr, int(N), out.ptr, C.byref(cnt), None)
    res = np.zeros(int(N_event), dtype=np.int64)
    n = min(int(cnt.value), int(N_event))
    if n:
        tmp = np.empty(int(cnt.value), dtype=np.int64)
        out.download(tmp)
        res[:n] = tmp[:n]
    return res


def bool_sum(event, N):
    """cpu.py:218-225"""
    cnt = C.c_int64(0)
    L.call("sfp_bool_sum", _b(event).ptr, int(N), C.byref(cnt), None)
    return int(cnt.value)


# --------------------------------------------------------------------------
# The QED model seam (particles.py:8-12): the reference's callables with its
# argument order, for numpy arguments.  The reference draws from numba's
# unseeded numpy.random; these draw Philox4x32-10 words keyed by ``seed``
# with counter (id, step, tag, slot), so the extra keyword arguments name the
# stream (defaults: ids 0..N-1, the package seed, step 0, tag 0 and the
# per-operation slot).
def _qed_stream(N, ids, seed, step, tag, slot):
    from . import qed
    ids = np.arange(int(N), dtype=np.uint64) if ids is None else np.asarray(ids, np.uint64)
    return (_dev(ids, np.uint64), qed.seed if seed is None else int(seed), int(step), int(tag),
            qed.PER_OP_SLOT if slot is None else int(slot), qed.device_tables())


def _rs(name, inv_gamma, chi, dt, N, to_be_pruned, event, delta, ids, seed, step, tag, slot):
    did, sd, st, tg, sl, tables = _qed_stream(N, ids, seed, step, tag, slot)
    de, dd = _b(event), _f(delta)
    L.call(name, _f(inv_gamma).ptr, _f(chi).ptr, float(dt), int(N), _b(to_be_pruned).ptr, de.ptr,
           dd.ptr, did.ptr, sd, st, tg, sl, tables, None)
    _back_bool(de, event)
    _back(dd, delta)


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
    did, sd, st, tg, sl, tables = _qed_stream(N, ids, seed, step, tag, slot)
    dtau, de, dd = _f(tau), _b(event), _f(delta)
    L.call(name, dtau.ptr, _f(inv_gamma).ptr, _f(chi).ptr, float(dt), int(N), _b(to_be_pruned).ptr,
           de.ptr, dd.ptr, did.ptr, sd, st, tg, sl, tables, None)
    _back(dtau, tau)
    _back_bool(de, event)
    _back(dd, delta)


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
