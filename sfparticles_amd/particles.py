"""Species containers (sfparticles/particles.py) backed by device stores.

A ``Particles`` object starts with host (numpy) arrays exactly as the
reference builds them (particles.py:44-146).  The first GPU operation moves
them into a library-owned struct-of-arrays store on the device
(``sfp_store_*``); from then on the arrays live in HBM and attribute access
returns a read-only host copy (assigning an attribute uploads it).  The
per-species step operations (``_eval_field``, ``_calculate_chi``, ...) run the
corresponding HIP kernels; ``Simulation.start`` runs the fused step kernels on
the same stores.

Additions over the reference: every particle has a 64-bit ``id`` (seed
particles 0..N-1, appended ones ``next_id + rank``) that keys its Philox
stream, and each species has a ``tag`` (its index in the simulation).
"""
import ctypes as C
from enum import Enum, auto
from typing import Tuple
from weakref import ref

import numpy as np

from . import _lib as L
from . import qed
from .constants import c, e, epsilon_0, hbar, m_e, pi  # noqa: F401
from .fields import Fields, terms_to_c
# the reference's backend callables (particles.py:18-25), re-exported
from .backend import (push_position, boris, boris_tbmt, LL_push, CLL_push, update_chi,  # noqa: F401
                      pick_hard_photon, photon_recoil, create_pair, create_photon,
                      find_event_index, bool_sum)


class RadiationReactionType(Enum):
    """particles.py:27-40"""
    NONE = auto()
    PHOTON = auto()
    LL = auto()
    CLL = auto()


_RR_CODE = {RadiationReactionType.NONE: L.RR_NONE, RadiationReactionType.PHOTON: L.RR_PHOTON,
            RadiationReactionType.LL: L.RR_LL, RadiationReactionType.CLL: L.RR_CLL}

# array attributes and their store slots / dtypes
_ATTR = {
    "x": (L.X, np.float64), "y": (L.Y, np.float64), "z": (L.Z, np.float64),
    "ux": (L.UX, np.float64), "uy": (L.UY, np.float64), "uz": (L.UZ, np.float64),
    "id": (L.ID, np.uint64), "_to_be_pruned": (L.PRUNED, np.bool_),
    "tau": (L.TAU, np.float64), "sx": (L.SX, np.float64), "sy": (L.SY, np.float64),
    "sz": (L.SZ, np.float64), "inv_gamma": (L.INV_GAMMA, np.float64), "chi": (L.CHI, np.float64),
    "Ex": (L.EX, np.float64), "Ey": (L.EY, np.float64), "Ez": (L.EZ, np.float64),
    "Bx": (L.BX, np.float64), "By": (L.BY, np.float64), "Bz": (L.BZ, np.float64),
    "event": (L.EVENT, np.bool_), "photon_delta": (L.DELTA, np.float64),
    "pair_delta": (L.DELTA, np.float64), "shot": (L.SHOT, np.uint32),
}


# per-step scratch of the reference's passes, recomputed before every use
_SCRATCH = frozenset(("inv_gamma", "chi", "Ex", "Ey", "Ez", "Bx", "By", "Bz"))


class Particles(object):
    """particles.py:43-516"""

    def __init__(self, name: str, q: int, m: float, N: int = 0, props: Tuple = None,
                 RR: RadiationReactionType = RadiationReactionType.PHOTON, push=True) -> None:
        for k, v in (("_h", {}), ("_store", None), ("_cache", {}), ("_version", 0),
                     ("_ig_stale", False)):
            object.__setattr__(self, k, v)
        N = int(N)
        # invalid arguments raise AssertionError, as the reference's tests expect
        assert m >= 0, f"mass must be >= 0, got {m}"
        assert m != 0 or q == 0, "a massless species must be neutral"
        assert isinstance(RR, RadiationReactionType), \
            f"RR must be a RadiationReactionType, got {RR!r}"
        self.name, self.q, self.m, self.RR, self.push = name, q * e, m * m_e, RR, push
        self.bw = self.radiating = self.spin = False
        self.ae, self.tag, self._op_calls = 0.0, 0, 0

        # state: six arrays of length N, zero unless given
        if props:
            assert len(props) == 6, f"props holds x, y, z, ux, uy, uz: 6 arrays, got {len(props)}"
            state = prepare_props(props, N)
        else:
            assert m != 0 or N == 0, "photons need their momenta: give props, not only N"
            state = [np.zeros(N) for _ in range(6)]
        self.attrs = []
        for attr, arr in zip(("x", "y", "z", "ux", "uy", "uz"), state):
            setattr(self, attr, arr)
            self.attrs.append(attr)
        usq = state[3] ** 2 + state[4] ** 2 + state[5] ** 2
        if m == 0:
            assert (usq > 0).all(), "a photon's momentum must not be 0"
            self.inv_gamma = 1.0 / np.sqrt(usq)
        else:
            self.inv_gamma = 1.0 / np.sqrt(1.0 + usq)
        # the reference's per-step scratch (field, chi), then bookkeeping
        for attr in ("chi", "Ex", "Ey", "Ez", "Bx", "By", "Bz"):
            setattr(self, attr, np.zeros(N))
        self._to_be_pruned = np.zeros(N, dtype=bool)
        self.id = np.arange(N, dtype=np.uint64)
        self.attrs += ["inv_gamma", "chi", "Ex", "Ey", "Ez", "Bx", "By", "Bz", "_to_be_pruned",
                       "id"]
        self.N_buffered = self.buffer_size = self.next_id = N

    # ------------------------------------------------------ array storage --
    def __getattr__(self, name):
        if name in _ATTR:
            h = self.__dict__.get("_h")
            if h is not None and name in h:
                return h[name]
            if self.__dict__.get("_store") is not None and name in self.__dict__.get("attrs", ()):
                return self._download(name)
        raise AttributeError(f"'{type(self).__name__}' object has no attribute '{name}'")

    def __setattr__(self, name, value):
        if name in _ATTR:
            dtype = _ATTR[name][1]
            arr = np.ascontiguousarray(np.asarray(value, dtype=dtype))
            if name == "shot" and name not in self.attrs:
                # ensemble membership is optional: an attribute once set
                self.attrs.append(name)
            if self._store is None:
                self._h[name] = arr
            else:
                self._upload(name, arr)
            return
        object.__setattr__(self, name, value)

    @property
    def on_device(self) -> bool:
        return self._store is not None

    def _touch(self):
        object.__setattr__(self, "_version", self._version + 1)
        self._cache.clear()

    def _ensure_device(self):
        """Move the host arrays into a device store (particles.py:148-155).

        The per-step scratch arrays of the reference (1/gamma, chi, E, B) are
        not uploaded: the fused step keeps them in registers, and the per-op
        entry points allocate them on first use (1/gamma is then recomputed
        from u), so a store costs 57 bytes per particle plus its optional
        attributes instead of 129."""
        if self._store is not None:
            return
        L.lib()
        flags = (L.STORE_TAU if "tau" in self.attrs else 0) | (L.STORE_SPIN if self.spin else 0)
        h = C.c_void_p()
        L.call("sfp_store_create", int(flags), int(max(self.buffer_size, 1)), C.byref(h))
        object.__setattr__(self, "_store", h.value)
        host = self._h
        object.__setattr__(self, "_h", None)
        for name in self.attrs:
            if name in host and name not in _SCRATCH:
                self._upload(name, host[name])
        object.__setattr__(self, "_ig_stale", True)
        L.call("sfp_store_set_count", self._store, int(self.N_buffered), int(self.next_id))
        self._touch()

    def _capacity(self):
        cap = C.c_int64(0)
        L.call("sfp_store_capacity", self._store, C.byref(cap))
        return int(cap.value)

    def _ptr(self, name):
        if name == "inv_gamma" and self._ig_stale:
            self._refresh_inv_gamma()
        p = C.c_void_p()
        L.call("sfp_store_ptr", self._store, _ATTR[name][0], C.byref(p))
        return p.value

    def _refresh_inv_gamma(self):
        object.__setattr__(self, "_ig_stale", False)
        L.call("sfp_update_inv_gamma", self._ptr("ux"), self._ptr("uy"), self._ptr("uz"),
               self._ptr("inv_gamma"), int(self.N_buffered), self._ptr("_to_be_pruned"),
               int(self.m == 0), None)

    def _upload(self, name, arr):
        n = arr.shape[0]
        if n > self._capacity():
            L.call("sfp_store_reserve", self._store, int(n))
        src = arr.view(np.uint8) if arr.dtype == np.bool_ else arr
        L.call("sfp_store_upload", self._store, _ATTR[name][0], src.ctypes.data, 0, int(n))
        if name in ("ux", "uy", "uz"):
            pass
        if name == "inv_gamma":
            object.__setattr__(self, "_ig_stale", False)
        self._touch()

    def _download(self, name):
        hit = self._cache.get(name)
        if hit is not None:
            return hit
        if name == "inv_gamma" and self._ig_stale:
            self._refresh_inv_gamma()
        n = int(self.buffer_size)
        dtype = _ATTR[name][1]
        raw = np.empty(n, dtype=np.uint8 if dtype == np.bool_ else dtype)
        L.call("sfp_store_download", self._store, _ATTR[name][0], raw.ctypes.data, 0, n)
        arr = raw.view(np.bool_) if dtype == np.bool_ else raw
        arr.flags.writeable = False
        self._cache[name] = arr
        return arr

    def _sync_count_from_device(self):
        n = C.c_int64(0)
        nid = C.c_uint64(0)
        L.call("sfp_store_get_count", self._store, C.byref(n), C.byref(nid))
        self._set_count(int(n.value), int(nid.value))

    def _set_count(self, n: int, next_id: int):
        """The host's copy of the device count (after a run or a device read)."""
        self.N_buffered = n
        self.next_id = next_id
        if self.buffer_size < self.N_buffered:
            self.buffer_size = self.N_buffered
        self._touch()

    def _push_count_to_device(self):
        L.call("sfp_store_set_count", self._store, int(self.N_buffered), int(self.next_id))

    def _to_gpu(self) -> None:
        self._ensure_device()

    def _to_host(self) -> None:
        """Arrays stay in HBM; attribute access copies them to the host."""
        if self._store is not None:
            L.call("sfp_synchronize")

    def __del__(self):
        try:
            if self.__dict__.get("_store") is not None and L._lib is not None:
                L._lib.sfp_store_destroy(self._store)
        except Exception:
            pass

    # ----------------------------------------------------------- queries --
    @property
    def Npart(self) -> int:
        """particles.py:166-174"""
        if self._store is None:
            return self.N_buffered - int(np.count_nonzero(self._h["_to_be_pruned"][:self.N_buffered]))
        cnt = C.c_int64(0)
        L.call("sfp_bool_sum", self._ptr("_to_be_pruned"), int(self.N_buffered), C.byref(cnt), None)
        return self.N_buffered - int(cnt.value)

    @property
    def gamma(self):
        """particles.py:176-185"""
        return 1.0 / self.inv_gamma

    # ------------------------------------------------------- QED wiring --
    def set_photon(self, photon: "Particles") -> None:
        """particles.py:187-212: this species emits into `photon`."""
        assert self.m > 0, "only massive species emit photons"
        assert self.RR not in (RadiationReactionType.LL, RadiationReactionType.CLL), \
            "LL and CLL radiation reaction is continuous: no photons are emitted"
        assert isinstance(photon, Particles), "photon must be a Particles species"
        assert photon.m == 0 and photon.q == 0, "the photon species must be massless and neutral"
        self.radiating = True
        self.photon_ref = ref(photon)
        self._add_qed_arrays("photon_delta")

    def set_pair(self, electron: "Particles", positron: "Particles") -> None:
        """particles.py:215-242: this photon species converts into pairs."""
        assert self.m == 0, "only photons create Breit-Wheeler pairs"
        assert isinstance(electron, Particles) and isinstance(positron, Particles), \
            "electron and positron must be Particles species"
        assert electron.m == m_e and electron.q == -e, "the first pair species must be electrons"
        assert positron.m == m_e and positron.q == e, "the second pair species must be positrons"
        self.bw = True
        self.bw_electron_ref = ref(electron)
        self.bw_positron_ref = ref(positron)
        self._add_qed_arrays("pair_delta")

    def _add_qed_arrays(self, delta_name):
        new = {"event": np.full(self.buffer_size, False), delta_name: np.zeros(self.buffer_size)}
        if qed._use_optical_depth and "tau" not in self.attrs:
            new["tau"] = np.zeros(self.buffer_size)
        if self._store is not None:
            if "tau" in new:
                raise RuntimeError("enable the optical-depth model before moving species to the GPU")
        for k, v in new.items():
            if k not in self.attrs:
                self.attrs.append(k)
            setattr(self, k, v)

    # ------------------------------------------------ per-species steps --
    def _n(self):
        return int(self.N_buffered)

    def _push_momentum(self, dt: float) -> None:
        """particles.py:244-281"""
        if self.m == 0:
            return
        self._ensure_device()
        f = [self._ptr(k) for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz")]
        u = [self._ptr(k) for k in ("ux", "uy", "uz", "inv_gamma")]
        L.call("sfp_boris", *u, *f, float(self.q), self._n(), self._ptr("_to_be_pruned"), float(dt),
               None)
        self._touch()
        self._radiation_reaction(dt)

    def _radiation_reaction(self, dt):
        if self.RR in (RadiationReactionType.LL, RadiationReactionType.CLL):
            self._calculate_chi()
            name = "sfp_LL_push" if self.RR == RadiationReactionType.LL else "sfp_CLL_push"
            L.call(name, *[self._ptr(k) for k in ("ux", "uy", "uz", "inv_gamma", "chi")], self._n(),
                   self._ptr("_to_be_pruned"), float(dt), None)
            self._touch()

    def _push_position(self, dt: float) -> None:
        """particles.py:284-296"""
        self._ensure_device()
        L.call("sfp_push_position", *[self._ptr(k) for k in ("x", "y", "z", "ux", "uy", "uz",
                                                              "inv_gamma")],
               self._n(), self._ptr("_to_be_pruned"), float(dt), None)
        self._touch()

    def _eval_field(self, fields: Fields, t: float) -> None:
        """particles.py:299-323"""
        terms = fields.require_terms()
        self._ensure_device()
        arr = terms_to_c(terms)
        L.call("sfp_eval_field", arr, len(terms), self._ptr("x"), self._ptr("y"), self._ptr("z"),
               float(t), self._n(), self._ptr("_to_be_pruned"),
               *[self._ptr(k) for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz")], None)
        self._touch()

    def _calculate_chi(self) -> None:
        """particles.py:326-339"""
        self._ensure_device()
        L.call("sfp_update_chi", *[self._ptr(k) for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "ux",
                                                           "uy", "uz", "inv_gamma", "chi")],
               self._n(), self._ptr("_to_be_pruned"), None)
        self._touch()

    def _qed_args(self, dt, delta_name):
        step = self._op_calls
        self._op_calls += 1
        return step

    def _photon_event(self, dt: float):
        """particles.py:342-361"""
        self._ensure_device()
        step = self._qed_args(dt, "photon_delta")
        common = (float(dt), self._n(), self._ptr("_to_be_pruned"), self._ptr("event"),
                  self._ptr("photon_delta"), self._ptr("id"), int(qed.seed), int(step),
                  int(self.tag), int(qed.PER_OP_SLOT), qed.device_tables(), None)
        if qed._use_optical_depth:
            L.call("sfp_update_tau_e", self._ptr("tau"), self._ptr("inv_gamma"), self._ptr("chi"),
                   *common)
        else:
            L.call("sfp_photon_from_rejection_sampling", self._ptr("inv_gamma"), self._ptr("chi"),
                   *common)
        self._touch()
        return self.event, self.photon_delta

    def _pair_event(self, dt: float):
        """particles.py:364-383"""
        self._ensure_device()
        step = self._qed_args(dt, "pair_delta")
        common = (float(dt), self._n(), self._ptr("_to_be_pruned"), self._ptr("event"),
                  self._ptr("pair_delta"), self._ptr("id"), int(qed.seed), int(step),
                  int(self.tag), int(qed.PER_OP_SLOT), qed.device_tables(), None)
        if qed._use_optical_depth:
            L.call("sfp_update_tau_gamma", self._ptr("tau"), self._ptr("inv_gamma"),
                   self._ptr("chi"), *common)
        else:
            L.call("sfp_pair_from_rejection_sampling", self._ptr("inv_gamma"), self._ptr("chi"),
                   *common)
        self._touch()
        return self.event, self.pair_delta

    def _pick_hard_photon(self, threshold: float = 2.0) -> None:
        """particles.py:386-396"""
        self._ensure_device()
        L.call("sfp_pick_hard_photon", self._ptr("event"), self._ptr("photon_delta"),
               self._ptr("inv_gamma"), float(threshold), self._n(), None)
        self._touch()

    def _event_index(self):
        n = self._n()
        cnt = C.c_int64(0)
        L.call("sfp_bool_sum", self._ptr("event"), n, C.byref(cnt), None)
        k = int(cnt.value)
        if k == 0:
            return 0, None
        idx = L.DeviceArray(8 * k)
        cnt2 = C.c_int64(0)
        # room for every flag is not needed: the scatter writes exactly k entries
        L.call("sfp_find_event_index", self._ptr("event"), n, idx.ptr, C.byref(cnt2), None)
        return k, idx

    def _create_photon(self, pho: "Particles") -> None:
        """particles.py:399-431"""
        self._ensure_device()
        N_photon, event_index = self._event_index()
        if N_photon == 0:
            return
        if hasattr(self, "photon_ref"):
            pho._ensure_device()
            pho._extend(N_photon)
            L.call("sfp_create_photon",
                   *[self._ptr(k) for k in ("x", "y", "z", "ux", "uy", "uz")],
                   *[pho._ptr(k) for k in ("x", "y", "z", "ux", "uy", "uz", "inv_gamma",
                                           "_to_be_pruned")],
                   event_index.ptr, self._ptr("photon_delta"), int(pho.N_buffered), int(N_photon),
                   pho._ptr("id"), int(pho.next_id), pho._ptr("chi"),
                   pho._ptr("tau") if "tau" in pho.attrs else None, None)
            pho.N_buffered += N_photon
            pho.next_id += N_photon
            pho._push_count_to_device()
            pho._touch()
        if self.RR == RadiationReactionType.PHOTON:
            L.call("sfp_photon_recoil", *[self._ptr(k) for k in ("ux", "uy", "uz", "inv_gamma",
                                                                  "event", "photon_delta")],
                   self._n(), self._ptr("_to_be_pruned"), None)
            self._touch()

    def _create_pair(self, ele: "Particles", pos: "Particles") -> None:
        """particles.py:433-473"""
        self._ensure_device()
        N_pair, event_index = self._event_index()
        if N_pair == 0:
            return
        if hasattr(self, "bw_electron_ref"):
            for dst, inverse in ((ele, 0), (pos, 1)):
                dst._ensure_device()
                dst._extend(N_pair)
                spin = ([dst._ptr(k) for k in ("sx", "sy", "sz")] if dst.spin else [None] * 3)
                L.call("sfp_create_pair",
                       *[self._ptr(k) for k in ("x", "y", "z", "ux", "uy", "uz", "_to_be_pruned")],
                       *[dst._ptr(k) for k in ("x", "y", "z", "ux", "uy", "uz", "inv_gamma",
                                               "_to_be_pruned")],
                       event_index.ptr, self._ptr("pair_delta"), int(dst.N_buffered), int(N_pair),
                       inverse, dst._ptr("id"), int(dst.next_id), dst._ptr("chi"),
                       dst._ptr("tau") if "tau" in dst.attrs else None, *spin, None)
            for dst in (ele, pos):
                dst.N_buffered += N_pair
                dst.next_id += N_pair
                dst._push_count_to_device()
                dst._touch()
            self._touch()

    # ------------------------------------------------------- buffers ----
    def _extend(self, N_new: int) -> None:
        """particles.py:476-500 (same growth formula; new slots pruned, events false)"""
        if (self.buffer_size - self.N_buffered) < N_new:
            buffer_size_new = self.buffer_size + int(self.N_buffered / 4) + N_new
            if self._store is None:
                grow = buffer_size_new - self.buffer_size
                for attr in self.attrs:
                    self._h[attr] = np.resize(self._h[attr], buffer_size_new)
                    if attr == "event":
                        self._h["event"][-grow:] = False
                self._h["_to_be_pruned"][-grow:] = True
            else:
                L.call("sfp_store_reserve", self._store, int(buffer_size_new))
                self._touch()
            self.buffer_size = buffer_size_new

    def _prune(self) -> None:
        """particles.py:503-516"""
        if self._store is None:
            selected = ~self._h["_to_be_pruned"]
            for attr in self.attrs:
                self._h[attr] = np.ascontiguousarray(self._h[attr][selected])
            N = int(selected.sum())
        else:
            kept = C.c_int64(0)
            L.call("sfp_store_prune", self._store, C.byref(kept))
            N = int(kept.value)
            self._touch()
        self.buffer_size = N
        self.N_buffered = N


class SpinParticles(Particles):
    """particles.py:518-604 (props with 9 entries are accepted)."""

    def __init__(self, name: str, q: int, m: float, N: int = 0, props: Tuple = None,
                 RR: RadiationReactionType = RadiationReactionType.PHOTON, ae: float = 1.14e-3,
                 push: bool = True) -> None:
        # props: the six of Particles and sx, sy, sz (the reference's
        # constructor cannot take them, particles.py:566-567); spins 0 if not given
        if props:
            assert len(props) == 9, f"props holds x..uz and sx, sy, sz: 9 arrays, got {len(props)}"
            super().__init__(name, q, m, N, props[:6], RR, push)
            spins = prepare_props(props[6:], int(N))
        else:
            super().__init__(name, q, m, N, None, RR, push)
            spins = [np.zeros(int(N)) for _ in range(3)]
        self.sx, self.sy, self.sz = spins
        self.ae = ae
        self.spin = True
        self.attrs += ["sx", "sy", "sz"]

    def _push_momentum(self, dt: float) -> None:
        """particles.py:575-604 (LL only; the reference's spin species ignore CLL)"""
        if self.m == 0:
            return
        self._ensure_device()
        u = [self._ptr(k) for k in ("ux", "uy", "uz", "inv_gamma", "sx", "sy", "sz")]
        f = [self._ptr(k) for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz")]
        L.call("sfp_boris_tbmt", *u, *f, float(self.q), float(self.ae), self._n(),
               self._ptr("_to_be_pruned"), float(dt), None)
        self._touch()
        if self.RR == RadiationReactionType.LL:
            self._radiation_reaction(dt)


def prepare_props(props: Tuple, N: int) -> list:
    """particles.py:606-635: each property as a float64 vector of length N
    (an array as given, a list or tuple converted, a scalar only for N = 1)."""
    out = []
    for prop in props:
        if isinstance(prop, np.ndarray):
            assert prop.ndim == 1, "each property must be a 1-D array"
            arr = prop
        elif isinstance(prop, (int, float)):
            arr = np.asarray([prop], dtype=np.float64)
        elif isinstance(prop, (list, tuple)):
            arr = np.asarray(prop, dtype=np.float64)
        else:
            continue
        assert arr.shape[0] == N, f"a property has {arr.shape[0]} values for N = {N}"
        out.append(arr)
    return out
