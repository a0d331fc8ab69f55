"""Reference-structured CPU simulation driver built on the oracle passes.

Restates sfparticles/particles.py:43-635 (species container, buffer growth,
QED wiring, per-species operations) and sfparticles/simulation.py:12-161 (the
ordered step loop) with numpy arrays and the C passes of oracle.c.  Test
infrastructure only (tests, smoke check, CPU baseline).

Additions over the reference, all needed to replay the GPU bit for bit:
  * every particle carries a 64-bit id (seed particles 0..N-1, appended
    particles next_id + rank in append order) that keys its Philox stream;
  * the species' index in the simulation is the stream's species tag;
  * the QED model is chosen per simulation (``qed="rs"|"od"``) instead of by
    an import-time environment variable.
"""
import ctypes as C

import numpy as np

from . import lib, ptr, field_array, rs_struct, od_struct

C_LIGHT = 299792458.0
M_E = 9.1093837015e-31
E_CHARGE = 1.602176634e-19

PER_OP_SLOT = 0x80  # draws made by per-operation calls outside a simulation


def _rr_name(RR):
    return RR if isinstance(RR, str) else RR.name


class OParticles:
    """particles.py:43-146 (+ SpinParticles, particles.py:518-604)."""

    def __init__(self, name, q, m, N=0, props=None, RR="PHOTON", push=True, spin=False,
                 ae=1.14e-3, spin_props=None):
        self.name = name
        self.q = q * E_CHARGE
        self.m = m * M_E
        self.RR = _rr_name(RR)
        self.push = push
        self.bw = False
        self.radiating = False
        self.spin = spin
        self.ae = ae
        N = int(N)
        if props is None:
            assert m != 0 or N == 0
            props = [np.zeros(N)] * 6
        x, y, z, ux, uy, uz = [np.array(p, dtype=np.float64) for p in props]
        self.x, self.y, self.z, self.ux, self.uy, self.uz = x, y, z, ux, uy, uz
        self.attrs = ["x", "y", "z", "ux", "uy", "uz"]
        if m > 0:
            self.inv_gamma = 1.0 / np.sqrt(1 + self.ux ** 2 + self.uy ** 2 + self.uz ** 2)
        else:
            self.inv_gamma = 1.0 / np.sqrt(self.ux ** 2 + self.uy ** 2 + self.uz ** 2)
        self.attrs += ["inv_gamma"]
        self.chi = np.zeros(N)
        self.attrs += ["chi"]
        self.Ex, self.Ey, self.Ez = np.zeros(N), np.zeros(N), np.zeros(N)
        self.Bx, self.By, self.Bz = np.zeros(N), np.zeros(N), np.zeros(N)
        self.attrs += ["Ex", "Ey", "Ez", "Bx", "By", "Bz"]
        self.N_buffered = N
        self.buffer_size = N
        self._to_be_pruned = np.full(N, False)
        self.attrs += ["_to_be_pruned"]
        self.id = np.arange(N, dtype=np.uint64)
        self.attrs += ["id"]
        self.next_id = N
        if spin:
            if spin_props is None:
                spin_props = [np.zeros(N)] * 3
            self.sx, self.sy, self.sz = [np.array(p, dtype=np.float64) for p in spin_props]
            self.attrs += ["sx", "sy", "sz"]
        self.tau = None
        self.tag = 0
        self._op_calls = 0

    @property
    def Npart(self):
        return self.N_buffered - int(lib().orc_bool_sum(ptr(self._to_be_pruned), self.N_buffered))

    # particles.py:187-242
    def set_photon(self, photon, od=False):
        assert self.m > 0 and self.RR not in ("LL", "CLL")
        self.radiating = True
        self.photon_ref = photon
        self.event = np.full(self.buffer_size, False)
        self.photon_delta = np.zeros(self.buffer_size)
        self.attrs += ["event", "photon_delta"]
        if od:
            self.tau = np.zeros(self.buffer_size)
            self.attrs += ["tau"]

    def set_pair(self, electron, positron, od=False):
        assert self.m == 0
        self.bw = True
        self.bw_electron_ref = electron
        self.bw_positron_ref = positron
        self.event = np.full(self.buffer_size, False)
        self.pair_delta = np.zeros(self.buffer_size)
        self.attrs += ["event", "pair_delta"]
        if od:
            self.tau = np.zeros(self.buffer_size)
            self.attrs += ["tau"]

    # particles.py:476-500
    def _extend(self, N_new):
        if (self.buffer_size - self.N_buffered) < N_new:
            buffer_size_new = self.buffer_size + int(self.N_buffered / 4) + N_new
            for attr in self.attrs:
                setattr(self, attr, np.resize(getattr(self, attr), buffer_size_new))
                if attr == "event":
                    self.event[-(buffer_size_new - self.buffer_size):] = False
            self._to_be_pruned[-(buffer_size_new - self.buffer_size):] = True
            self.buffer_size = buffer_size_new

    # particles.py:503-516
    def _prune(self):
        selected = ~self._to_be_pruned
        for attr in self.attrs:
            setattr(self, attr, np.ascontiguousarray(getattr(self, attr)[selected]))
        N = int(selected.sum())
        self.buffer_size = N
        self.N_buffered = N


class OSimulation:
    """simulation.py:12-161 on the oracle passes."""

    def __init__(self, *all_particles, dt, fields, photon_threshold=2.0, t0=0.0, qed="rs",
                 seed=0, rs_tables=None, od_tables=None, libm=False, spin_emission=False):
        self.all_particles = all_particles
        self.spin_emission = bool(spin_emission)
        self.dt = dt
        self.terms = list(fields)
        self.fa = field_array(self.terms)
        self.photon_threshold = photon_threshold
        self.t = t0
        self.step = 0
        self.qed = qed
        self.seed = seed
        self.L = lib(libm)
        self.rs = rs_struct(rs_tables) if rs_tables is not None else None
        self.od = od_struct(od_tables) if od_tables is not None else None
        self.particles_push, self.particles_bw, self.particles_rad = [], [], []
        for i, p in enumerate(all_particles):
            p.tag = i
            if p.bw:
                self.particles_bw.append(p)
            if p.radiating:
                self.particles_rad.append(p)
            if p.push:
                self.particles_push.append(p)

    # ----------------------------------------------------- species operations
    def eval_field(self, p, t):
        self.L.orc_eval_field(self.fa, len(self.terms), ptr(p.x), ptr(p.y), ptr(p.z), t,
                              p.N_buffered, ptr(p._to_be_pruned), ptr(p.Ex), ptr(p.Ey), ptr(p.Ez),
                              ptr(p.Bx), ptr(p.By), ptr(p.Bz))

    def calculate_chi(self, p):
        self.L.orc_update_chi(ptr(p.Ex), ptr(p.Ey), ptr(p.Ez), ptr(p.Bx), ptr(p.By), ptr(p.Bz),
                              ptr(p.ux), ptr(p.uy), ptr(p.uz), ptr(p.inv_gamma), ptr(p.chi),
                              p.N_buffered, ptr(p._to_be_pruned))

    def qed_event(self, p, dt, step, slot_base=0):
        """particles.py:342-383"""
        delta = p.photon_delta if p.radiating else p.pair_delta
        if self.qed == "od":
            fn = self.L.orc_update_tau_e if p.radiating else self.L.orc_update_tau_gamma
            fn(ptr(p.tau), ptr(p.inv_gamma), ptr(p.chi), dt, p.N_buffered, ptr(p._to_be_pruned),
               ptr(p.event), ptr(delta), ptr(p.id), self.seed, step, p.tag, slot_base, C.byref(self.od))
        else:
            args = (ptr(p.inv_gamma), ptr(p.chi), dt, p.N_buffered, ptr(p._to_be_pruned),
                    ptr(p.event), ptr(delta), ptr(p.id), self.seed, step, p.tag, slot_base,
                    C.byref(self.rs))
            if p.radiating and p.spin and self.spin_emission:
                self.L.orc_photon_from_rejection_sampling_spin(
                    *args, float(p.q), ptr(p.ux), ptr(p.uy), ptr(p.uz), ptr(p.Ex), ptr(p.Ey),
                    ptr(p.Ez), ptr(p.Bx), ptr(p.By), ptr(p.Bz), ptr(p.sx), ptr(p.sy), ptr(p.sz))
            else:
                fn = (self.L.orc_photon_from_rejection_sampling if p.radiating
                      else self.L.orc_pair_from_rejection_sampling)
                fn(*args)

    def pick_hard_photon(self, p, threshold):
        self.L.orc_pick_hard_photon(ptr(p.event), ptr(p.photon_delta), ptr(p.inv_gamma),
                                    threshold, p.N_buffered)

    def find_events(self, p):
        n = int(self.L.orc_bool_sum(ptr(p.event), p.N_buffered))
        idx = np.zeros(n, dtype=np.int64)
        if n:
            self.L.orc_find_event_index(ptr(p.event), n, p.N_buffered, ptr(idx))
        return n, idx

    def create_photon(self, p, pho):
        """particles.py:399-431"""
        n, idx = self.find_events(p)
        if n == 0:
            return
        pho._extend(n)
        self.L.orc_create_photon(ptr(p.x), ptr(p.y), ptr(p.z), ptr(p.ux), ptr(p.uy), ptr(p.uz),
                                 ptr(pho.x), ptr(pho.y), ptr(pho.z), ptr(pho.ux), ptr(pho.uy),
                                 ptr(pho.uz), ptr(pho.inv_gamma), ptr(pho._to_be_pruned),
                                 ptr(idx), ptr(p.photon_delta), pho.N_buffered, n, ptr(pho.id),
                                 pho.next_id, ptr(pho.chi), ptr(pho.tau))
        pho.N_buffered += n
        pho.next_id += n
        if p.RR == "PHOTON":
            self.L.orc_photon_recoil(ptr(p.ux), ptr(p.uy), ptr(p.uz), ptr(p.inv_gamma),
                                     ptr(p.event), ptr(p.photon_delta), p.N_buffered,
                                     ptr(p._to_be_pruned))

    def create_pair(self, p, ele, pos):
        """particles.py:433-473"""
        n, idx = self.find_events(p)
        if n == 0:
            return
        ele._extend(n)
        pos._extend(n)
        for dst, inv in ((ele, 0), (pos, 1)):
            spin = (ptr(dst.sx), ptr(dst.sy), ptr(dst.sz)) if dst.spin else (None, None, None)
            self.L.orc_create_pair(ptr(p.x), ptr(p.y), ptr(p.z), ptr(p.ux), ptr(p.uy), ptr(p.uz),
                                   ptr(p._to_be_pruned), ptr(dst.x), ptr(dst.y), ptr(dst.z),
                                   ptr(dst.ux), ptr(dst.uy), ptr(dst.uz), ptr(dst.inv_gamma),
                                   ptr(dst._to_be_pruned), ptr(idx), ptr(p.pair_delta),
                                   dst.N_buffered, n, inv, ptr(dst.id), dst.next_id,
                                   ptr(dst.chi), ptr(dst.tau), *spin)
        ele.N_buffered += n
        ele.next_id += n
        pos.N_buffered += n
        pos.next_id += n

    def push_position(self, p, dt):
        self.L.orc_push_position(ptr(p.x), ptr(p.y), ptr(p.z), ptr(p.ux), ptr(p.uy), ptr(p.uz),
                                 ptr(p.inv_gamma), p.N_buffered, ptr(p._to_be_pruned), dt)

    def push_momentum(self, p, dt):
        """particles.py:244-281 and SpinParticles._push_momentum (particles.py:575-604)"""
        if p.m == 0:
            return
        if p.spin:
            self.L.orc_boris_tbmt(ptr(p.ux), ptr(p.uy), ptr(p.uz), ptr(p.inv_gamma), ptr(p.sx),
                                  ptr(p.sy), ptr(p.sz), ptr(p.Ex), ptr(p.Ey), ptr(p.Ez),
                                  ptr(p.Bx), ptr(p.By), ptr(p.Bz), p.q, p.ae, p.N_buffered,
                                  ptr(p._to_be_pruned), dt)
        else:
            self.L.orc_boris(ptr(p.ux), ptr(p.uy), ptr(p.uz), ptr(p.inv_gamma), ptr(p.Ex),
                             ptr(p.Ey), ptr(p.Ez), ptr(p.Bx), ptr(p.By), ptr(p.Bz), p.q,
                             p.N_buffered, ptr(p._to_be_pruned), dt)
        rr = p.RR
        if rr == "LL" or (rr == "CLL" and not p.spin):
            self.calculate_chi(p)
            fn = self.L.orc_LL_push if rr == "LL" else self.L.orc_CLL_push
            fn(ptr(p.ux), ptr(p.uy), ptr(p.uz), ptr(p.inv_gamma), ptr(p.chi), p.N_buffered,
               ptr(p._to_be_pruned), dt)

    # ------------------------------------------------------------ main loop
    def start(self, nstep, call_back=None, prune=True):
        """simulation.py:74-161 (without the progress bar)."""
        for istep in range(self.step, self.step + nstep):
            for p in self.particles_push:
                self.eval_field(p, self.t)
            for p in self.particles_bw:
                self.calculate_chi(p)
                self.qed_event(p, self.dt, istep)
            for p in self.particles_rad:
                self.calculate_chi(p)
                self.qed_event(p, self.dt, istep)
                self.pick_hard_photon(p, self.photon_threshold)
            for p in self.particles_rad:
                self.create_photon(p, p.photon_ref)
            for p in self.particles_bw:
                self.create_pair(p, p.bw_electron_ref, p.bw_positron_ref)
            for p in self.particles_push:
                self.push_position(p, 0.5 * self.dt)
                self.eval_field(p, self.t + 0.5 * self.dt)
                self.push_momentum(p, self.dt)
                self.push_position(p, 0.5 * self.dt)
            if call_back:
                call_back(istep)
            self.t += self.dt
            self.step += 1
        if prune:
            for p in self.all_particles:
                p._prune()
