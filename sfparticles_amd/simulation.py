"""Simulation driver (sfparticles/simulation.py) on the fused HIP step.

``Simulation.start`` advances every species with one fused kernel per species
and step (``sfp_sim_run``).  The order of work inside a step -- QED at time t
for every species, appends in species order, then the leapfrog push of old
and new particles -- is the reference's (simulation.py:102-144); particles
created in a step are pushed in that step and are QED-active from the next.

Callbacks and progress updates are served between chunks of steps, so a run
without a callback advances ``print_every`` steps per library call.
"""
import ctypes as C
from time import perf_counter_ns
from typing import Callable

import numpy as np

from . import _lib as L
from . import qed
from .fields import Fields, terms_to_c
from .particles import Particles, RadiationReactionType, _RR_CODE

try:
    from tqdm.auto import tqdm
except Exception:  # pragma: no cover
    tqdm = None


class Simulation(object):
    """simulation.py:12-161"""

    def __init__(self, *all_particles: Particles, dt: float, fields: Fields, print_every: int = 100,
                 t0=0.0, photon_threshold=2.0, seed: int = None, qed_model: str = None,
                 precision: str = "fp64", sampler=None, graph: bool = False,
                 autotune: bool = False, shots=None, spin_emission: bool = False,
                 rate_net=None) -> None:
        """precision: "fp64" (the reference's arithmetic, bit-identical to the CPU
        restatement) or "mixed" (fields, pushers and QED sampling in FP32;
        positions, time and chi in FP64).  sampler (mixed only): a
        qed.mlp.PhotonSampler, or "mlp" for the trained default, that draws the
        energy of emitted photons in place of the inverse-CDF table.
        graph: replay batches of 64 steps as a captured HIP graph.  autotune:
        pick each kernel's launch grid online.  Neither changes the results.
        shots: an ensemble of independent shots in one simulation, a list of
        Fields, one per shot; every particle moves in the field of its
        ``shot`` index (species without one get shot 0 for their particles)
        and what it creates inherits the index.  ``fields`` is then unused.
        spin_emission: photon emission of SpinParticles (rejection sampling)
        uses the spin-resolved rate, which depends on the projection of the
        spin on the direction of beta x dbeta/dt; the reference's emission
        ignores spin.  rate_net (mixed, spin_emission): a qed.mlp.EmissionRateNet,
        or "mlp" for the trained default, that gives the spin-resolved emission
        rate in place of the table."""
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
        if sampler is not None and precision != "mixed":
            raise ValueError("the learned sampler runs in the mixed-precision step")
        if isinstance(sampler, str):
            if sampler != "mlp":
                raise ValueError(f"unknown sampler {sampler!r}")
            from .qed.mlp import default_sampler
            sampler = default_sampler()
        self.precision = precision
        self.sampler = sampler
        self.spin_emission = bool(spin_emission)
        if rate_net is not None:
            if precision != "mixed" or not self.spin_emission:
                raise ValueError("the learned rate runs in the mixed-precision step with "
                                 "spin_emission=True")
            if isinstance(rate_net, str):
                if rate_net != "mlp":
                    raise ValueError(f"unknown rate network {rate_net!r}")
                from .qed.mlp import default_rate_net
                rate_net = default_rate_net()
        self.rate_net = rate_net
        self.graph = bool(graph)
        self.autotune = bool(autotune)
        self.shots = None if shots is None else list(shots)
        if self.shots is not None:
            if not self.shots:
                raise ValueError("an ensemble needs at least one shot")
            if fields is None:
                self.fields = fields = self.shots[0]
            for p in all_particles:
                if "shot" not in p.attrs:
                    p.shot = np.zeros(p.N_buffered, np.uint32)
                sh = np.asarray(p.shot)[:p.N_buffered]
                if sh.size and int(sh.max()) >= len(self.shots):
                    raise ValueError(f"{p.name}: shot index {int(sh.max())} without a field "
                                     f"({len(self.shots)} shots)")
        self.all_particles = all_particles
        self.particles_push = []
        self.particles_bw = []
        self.particles_rad = []
        self.dt = dt
        self.fields = fields
        self.print_every = print_every or None  # 0 behaves like None (reference divides by it)
        self.photon_threshold = photon_threshold
        self.seed = qed.seed if seed is None else int(seed)
        self.qed_model = qed.model() if qed_model is None else qed_model
        self.t = t0
        self.step = 0
        self.progress_bar = None
        self._sim = None
        self._sim_stores = None
        self.last_run_ns = 0

        for i, p in enumerate(all_particles):
            p.tag = i
            if p.bw:
                assert p.bw_electron_ref() in all_particles, \
                    f"{p.bw_electron_ref().name} not included in simulation."
                assert p.bw_positron_ref() in all_particles, \
                    f"{p.bw_positron_ref().name} not included in simulation."
                self.particles_bw.append(p)
            if p.radiating:
                assert p.photon_ref() in all_particles, f"{p.photon_ref().name} not included in simulation."
                self.particles_rad.append(p)
            if p.push:
                self.particles_push.append(p)

    # ------------------------------------------------------------ native --
    def _build(self):
        """Create (or re-create after a store change) the native simulation."""
        stores = tuple(p._store for p in self.all_particles)
        if self._sim is not None and stores == self._sim_stores:
            return
        self._destroy()
        terms = self.fields.require_terms()
        idx = {id(p): i for i, p in enumerate(self.all_particles)}
        n = len(self.all_particles)
        descs = (L.SpeciesDesc * n)()
        for i, p in enumerate(self.all_particles):
            d = descs[i]
            d.store = p._store
            d.q = float(p.q)
            d.m = float(p.m / 9.1093837015e-31) if p.m else 0.0
            d.ae = float(p.ae)
            d.rr = _RR_CODE[p.RR]
            d.push = int(bool(p.push))
            d.spin = int(bool(p.spin))
            d.photon_dst = idx[id(p.photon_ref())] if p.radiating else -1
            d.pair_e_dst = idx[id(p.bw_electron_ref())] if p.bw else -1
            d.pair_p_dst = idx[id(p.bw_positron_ref())] if p.bw else -1
        params = L.SimParams(float(self.dt), float(self.photon_threshold), int(self.seed),
                             L.QED_OD if self.qed_model == "od" else L.QED_RS,
                             1 if self.precision == "mixed" else 0, int(self.spin_emission), 0)
        arr = terms_to_c(terms)
        h = C.c_void_p()
        L.call("sfp_sim_create", descs, n, arr, len(terms), qed.device_tables(), C.byref(params),
               C.byref(h))
        self._sim = h.value
        self._sim_stores = stores
        if self.sampler is not None:
            L.call("sfp_sim_set_sampler", self._sim, self.sampler.handle)
        if self.rate_net is not None:
            L.call("sfp_sim_set_rate_net", self._sim, self.rate_net.handle)
        if self.shots is not None:
            per = [f.require_terms() for f in self.shots]
            flat = [t for ts in per for t in ts]
            counts = (C.c_int32 * len(per))(*[len(ts) for ts in per])
            L.call("sfp_sim_set_shots", self._sim, terms_to_c(flat), counts, len(per))

    def _destroy(self):
        if self._sim is not None and L._lib is not None:
            L._lib.sfp_sim_destroy(self._sim)
        self._sim = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass

    def _prepare(self):
        for p in self.all_particles:
            p._ensure_device()
            p._push_count_to_device()
        self._build()

    def run_steps(self, nsteps: int, timing: bool = False, timing_every: int = 1,
                  timing_kernel=None) -> None:
        """Advance nsteps without callbacks, progress or pruning.  With timing,
        the step kernels of the steps whose index timing_every divides are
        timed with HIP events (kernel_time), and the kernel timing_kernel =
        (species, phase) names at every step."""
        self._prepare()
        if timing:
            L.call("sfp_sim_set_timing_every", self._sim, int(timing_every))
            s_, ph_ = timing_kernel if timing_kernel is not None else (-1, 0)
            L.call("sfp_sim_set_timing_kernel", self._sim, int(s_), int(ph_))
        t = C.c_double(self.t)
        flags = (L.RUN_TIMING if timing else 0) | (L.RUN_GRAPH if self.graph else 0) | \
            (L.RUN_AUTOTUNE if self.autotune else 0)
        L.call("sfp_sim_run", self._sim, C.byref(t), int(self.step), int(nsteps), flags)
        self.t = t.value
        self.step += int(nsteps)
        for i, p in enumerate(self.all_particles):
            if nsteps > 0:  # the counts the run read at its end
                n, nid = C.c_int64(0), C.c_uint64(0)
                L.call("sfp_sim_get_counts", self._sim, i, C.byref(n), C.byref(nid))
                p._set_count(int(n.value), int(nid.value))
            else:
                p._sync_count_from_device()
            object.__setattr__(p, "_ig_stale", True)

    def kernel_time(self, species: int, phase: int):
        """(ms, launches) of a species' sweep (phase 0), QED (1) or emission (2) kernel."""
        ms = C.c_double(0)
        n = C.c_int64(0)
        L.call("sfp_sim_kernel_time", self._sim, int(species), int(phase), C.byref(ms),
               C.byref(n))
        return ms.value, int(n.value)

    def reset_timing(self):
        L.call("sfp_sim_reset_timing", self._sim)

    def record(self, particles, steps: int, n: int = None, every: int = 1):
        """Record x, y, z, ux, uy, uz of the first n particles of a lepton
        species on the device after every `every`-th of the next `steps`
        steps (the store_trajectory callback of example/trajectory.py:62-70,
        without a host round trip per step); read them with records()."""
        self._prepare()
        sp = self.all_particles.index(particles)
        n = int(particles.N_buffered if n is None else n)
        rows = max(1, int(steps) // int(every))
        self._rec_n = n
        L.call("sfp_sim_set_recorder", self._sim, sp, n, int(every), rows, int(self.step))

    def records(self) -> dict:
        """Recorded rows: t [rows] and x, y, z, ux, uy, uz [rows, n]."""
        r = C.c_int64(0)
        if self._sim is not None:
            L.call("sfp_sim_get_records", self._sim, None, None, C.byref(r))
        rows = int(r.value)
        n = getattr(self, "_rec_n", 0)
        t = np.empty(rows)
        buf = np.empty((rows, n, 6))
        if rows:
            L.call("sfp_sim_get_records", self._sim, t.ctypes.data, buf.ctypes.data, C.byref(r))
        out = {"t": t}
        for j, k in enumerate(("x", "y", "z", "ux", "uy", "uz")):
            out[k] = buf[:, :, j].copy()
        return out

    def set_grid(self, species: int, phase: int, grid: int):
        """Pin the workgroup count of a species' sweep (0) or QED (1) kernel;
        0 returns it to the tuner."""
        self._prepare()
        L.call("sfp_sim_set_grid", self._sim, int(species), int(phase), int(grid))

    def set_flag_clear_every(self, steps: int):
        """Clear the look-back flags between batches every `steps` steps (the
        library does so every 2^28 steps, before the 29-bit launch epoch
        could wrap; a small interval exercises the clearing)."""
        self._prepare()
        L.call("sfp_sim_set_flag_clear_every", self._sim, int(steps))

    def set_shared_launches(self, on: bool):
        """Run two consecutive lepton species that share QED kernels and a
        photon store (electrons, positrons) in one QED launch (the default),
        or one launch per species; the results are the same."""
        self._prepare()
        L.call("sfp_sim_set_shared_launches", self._sim, int(bool(on)))

    def set_shape(self, species: int, phase: int, grid: int, rows: int = 0):
        """Pin the workgroup count and the most rows of 256 particles per tile
        of a kernel (rows 0: the kernel's most); grid 0 returns both to the
        tuner."""
        self._prepare()
        L.call("sfp_sim_set_shape", self._sim, int(species), int(phase), int(grid), int(rows))

    def launch_shape(self, species: int, phase: int):
        """(grid, workgroup size, most tile rows, number of candidate shapes)
        of a kernel."""
        self._prepare()
        g, t, r, n = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        L.call("sfp_sim_launch_shape", self._sim, int(species), int(phase), C.byref(g),
               C.byref(t), C.byref(r), C.byref(n))
        return int(g.value), int(t.value), int(r.value), int(n.value)

    def stats(self, species: int) -> dict:
        """Counters of one species since the last reset: particle-steps (live
        particles after each step), particles read by its step kernels,
        particles they appended, current buffered count, QED-listed
        particles, and the run's host synchronisations, capacity replays,
        photon sleep counts, graph and tuner batches, dense steps, deferred
        conversions, run calls that kept the photon visit lists."""
        v = (C.c_int64 * 18)()
        L.call("sfp_sim_stats_n", self._sim, int(species), v, 18)
        return dict(psteps=int(v[0]), processed=int(v[1]), spawned=int(v[2]), n=int(v[3]),
                    hot=int(v[4]), syncs=int(v[5]), replays=int(v[6]), capacity=int(v[7]),
                    slept=int(v[8]), sleeps=int(v[9]), graph_batches=int(v[10]),
                    captures=int(v[11]), tune_batches=int(v[12]), emit=int(v[13]),
                    clock_restarts=int(v[14]), dense_steps=int(v[15]),
                    deferrals=int(v[16]), list_keeps=int(v[17]))

    def qed_diag(self, species: int) -> dict:
        """Diagnostics of the exact QED evaluations of a species since the last
        reset: rejection-sampling probabilities above 1 and the largest one,
        photon-emission z beyond the table, optical-depth chi above the
        table, NaN chi (no counterpart in the reference)."""
        v = (C.c_int64 * 4)()
        pm = C.c_double(0)
        L.call("sfp_sim_qed_diag", self._sim, int(species), v, C.byref(pm))
        return dict(p_over1=int(v[0]), z_clamp=int(v[1]), chi_clamp=int(v[2]), chi_nan=int(v[3]),
                    p_max=float(pm.value))

    def reset_stats(self):
        L.call("sfp_sim_reset_stats", self._sim)

    # ------------------------------------------------------------ driver --
    def start(self, nstep: int, total: int = None, call_back: Callable[[int], None] = None,
              callback_every: int = 1):
        """simulation.py:74-161.  ``call_back(istep)`` runs after the steps
        istep with (istep + 1) a multiple of ``callback_every`` (1: after
        every step, as in the reference), seeing ``t`` and ``step`` of that
        step's start as the reference's callback does; the steps in between
        run as one call, without a host round trip (and as graph replays with
        ``graph=True``)."""
        callback_every = max(1, int(callback_every))
        if total is None:
            total = nstep
        if self.progress_bar is None and self.print_every and tqdm is not None:
            self.progress_bar = tqdm(total=total, unit="step", smoothing=1)

        end = self.step + nstep
        tic = perf_counter_ns()
        while self.step < end:
            if call_back is not None:
                chunk = min(end - self.step, callback_every - self.step % callback_every)
            elif self.print_every:
                chunk = min(end - self.step, self.print_every - self.step % self.print_every)
            else:
                chunk = end - self.step
            t_last = self.t  # the clock at the start of the chunk's last step
            for _ in range(chunk - 1):
                t_last += self.dt  # the same additions the library makes
            self.run_steps(chunk)
            if call_back is not None and self.step % callback_every == 0:
                # as in the reference (simulation.py:140-144) the callback sees
                # the state after step istep, with the clock and the step
                # counter not yet advanced past it
                t_end, istep = self.t, self.step - 1
                self.t, self.step = t_last, istep
                try:
                    call_back(istep)
                finally:
                    self.t, self.step = t_end, istep + 1
            if self.print_every and self.step % self.print_every == 0:
                elapsed = perf_counter_ns() - tic
                tic = perf_counter_ns()
                Ntotal = sum(p.Npart for p in self.all_particles)
                ns_per_particle = elapsed / max(Ntotal, 1) / self.print_every
                if self.progress_bar is not None:
                    particle_num_str = ",".join(f"{p.Npart} {p.name}" for p in self.all_particles)
                    self.progress_bar.set_postfix_str(
                        f"{particle_num_str}, {ns_per_particle:.0f}ns/particle")
                    self.progress_bar.update(self.print_every)

        for p in self.all_particles:
            p._to_host()
            p._prune()
