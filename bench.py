"""Benchmark: particle-steps/s of the sfparticles step on one MI355X.

Workload (the head-on collision the metric is quoted on): 1e7 electrons,
gamma ~ N(1957, 5%), momentum along -x, x ~ N(12 um, 3 um), y, z ~ N(0, 1 um),
against a linearly polarised paraxial Gaussian pulse (lambda 0.8 um, a0 100,
w0 2 um, tau_FWHM 30 fs) starting at x0 = -12 um; dt = T/100; 3000 steps;
rejection-sampling photon emission + Breit-Wheeler pairs (electrons,
positrons, photons), photon threshold 2 m c^2, Philox seed 0; FP64.

A step is one pass of the fused step kernels over every species.  Inputs are
uploaded to HBM before the timed region.  With --steps 3000 (or more) the
timed region is the whole run from step 0; with fewer, the run is
fast-forwarded untimed to --steps / 2 steps before the collision centre (step
1500, where the pulse centre meets the bunch centre) and the window of
--steps steps around it is timed, closing with the catch-up that brings
sleeping photons to the window's last step.  --warmup steps run first on a
separate, smaller copy of the workload (kernel load, allocator warm-up).

Prints one JSON line.  Also reported: the dominant kernel's achieved
algorithmic HBM bandwidth against the 8 TB/s peak (HIP events around the
electron QED kernel's launches on every step of the timed window, and around
every kernel's on every --timing-every-th step), the whole step's
algorithmic bandwidth, and the CPU oracle timed on this machine's cores over
the same window on a sample of the same state.  --summarize-profile turns
the rocprofv3 output of a profiled run into the markdown kept under
profiles/.

--workload cascade runs the cascade configuration instead: 1e4 electrons at
rest at the node of a rotating-E standing wave of two counter-propagating
circularly polarised pulses (a0 1000), optical-depth model, 6 ctau / (c dt)
steps, FP64 and mixed precision.

--workload ensemble runs the batched-shot workload instead (not part of the
default run): --shots independent head-on shots with a0 ~ U(20, 500),
electron counts log-uniform in [1e3, --count-max], gamma_s ~ U(500, 5000)
(per electron N(gamma_s, 5%)), spins uniform on the sphere, in one simulation
(every kernel launch serves all shots), mixed precision with spin-resolved
emission and the learned emission rate, 64-step batches replayed as HIP
graphs with the launch-grid tuner on; then --check-shots of the smallest
shots run alone on the FP64 table path, and their photon and pair yields are
compared with the ensemble's.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

C_LIGHT = 299792458.0
LAMBDA = 0.8e-6
DT = LAMBDA / C_LIGHT / 100
CTAU = C_LIGHT * 30e-15 / math.sqrt(2 * math.log(2))
HBM_PEAK_GBS = 8000.0
# the CPU baseline times the oracle over at most this many steps (the
# collision centre when the GPU window is longer): about 10-30 s of CPU work
CPU_WINDOW = 60
# the kernel timed with HIP events at every step of the timed window (the
# electron QED kernel, which dominates the step); the others are timed on the
# steps --timing-every selects, and scaled by their own timed launches
TIMED_KERNEL = (0, 1)

# Algorithmic bytes of the two step kernels of a species (the persistent
# state read once and written once; transients stay in registers):
#   leptons  sweep: reads x, y, z, u and the pruned flag of every buffered
#            particle (49 B); writes x, y, z, u (48 B) of every particle it
#            pushes and a 4-byte index for every particle it lists.
#            QED kernel: per listed particle reads the index, x, y, z, u and
#            the id (60 B) and writes x, y, z, u (48 B).
#   photons  list kernel: reads the 4-byte wake step of every buffered photon
#            and writes a 4-byte index for every photon due this step.
#            photon kernel: per listed photon reads the index, the pruned flag,
#            x, y, z and u (53 B), the id where the exact QED front runs
#            (8 B); writes x, y, z (24 B) of every photon it pushes, a 4-byte
#            wake step per multi-step push, and per conversion the pruned flag
#            and wake step (5 B).
#   leptons with the learned sampler: the QED kernel writes an 88-byte record
#            per emission candidate (the sampler's inputs and the candidate's
#            state) instead of its state; the emission kernel reads the record
#            and writes the candidate's state (x, y, z, u: 48 B).
#   Optical depth: + tau read and written (16 B) per exact QED evaluation.
#   Per appended particle: x, y, z, u, id and flag (57 B) in each destination.
def kernel_bytes(photon, phase, st, od=False):
    hot = st["hot"]
    tau = 16 * hot if od else 0
    if photon:
        if phase == 0:
            return 4 * (st["processed"] + st["slept"]) + 4 * st["processed"]
        return (53 * st["processed"] + 8 * hot + tau + 24 * (st["processed"] - st["spawned"])
                + 4 * st["sleeps"] + 5 * st["spawned"] + 57 * 2 * st["spawned"])
    if phase == 0:
        return 49 * st["processed"] + 48 * (st["processed"] - hot) + 4 * hot
    emit = st.get("emit", 0)
    if phase == 2:
        return (88 + 48) * emit + 57 * st["spawned"]
    if emit:
        return 60 * hot + 48 * (hot - emit) + 88 * emit + tau
    return 60 * hot + 48 * hot + tau + 57 * st["spawned"]


# The survey's algorithmic bytes per particle-step (the persistent state read
# once and written once, id and status byte; transients in registers): e+-
# 105 B in FP64, 81 B in mixed precision (FP64 x, FP32 u); photons 81 B; and
# 57 B per appended particle.  They price the kernel that does a species'
# whole step: the lepton QED kernel on dense steps (it visits every lepton by
# index, the sweep stands down) and the photon kernel over the awake photons.
def survey_bytes(photon, phase, st, precision, steps):
    if photon and phase == 1:
        return 81 * st["processed"] + 57 * 2 * st["spawned"]
    if not photon and phase == 1 and st.get("dense_steps", 0) == steps and not st.get("emit"):
        return (105 if precision == "fp64" else 81) * st["hot"] + 57 * st["spawned"]
    return None


def make_case(n, nsteps, seed=0, field="gaussian"):
    rng = np.random.Generator(np.random.PCG64(seed))
    gamma = rng.normal(1957.0, 0.05 * 1957.0, n)
    d = C_LIGHT * DT * nsteps / 2  # 12 um for 3000 steps: the pulse meets the bunch half-way
    x = rng.normal(d, 3e-6, n)
    y = rng.normal(0.0, 1e-6, n)
    z = rng.normal(0.0, 1e-6, n)
    ux = -np.sqrt(gamma ** 2 - 1)
    if field == "focused":  # the same pulse in the focused expansion, focus at x = 0
        term = (3, (100.0, 2e-6, CTAU, 1, 0.0, -d, LAMBDA, 0.0, 0.0))
    else:
        term = (2, (100.0, 2e-6, CTAU, 1, -d, LAMBDA, 0.0, 0.0))
    return (x, y, z, ux, np.zeros(n), np.zeros(n)), [term]


def build_gpu(n, nsteps, reserve, path="fp64", field="gaussian"):
    import sfparticles_amd as sfa
    from sfparticles_amd import qed
    from sfparticles_amd.fields import Fields
    qed.set_model("rs")
    props, field = make_case(n, nsteps, field=field)
    e = sfa.Particles("electron", -1, 1, N=n, props=props)
    p = sfa.Particles("positron", 1, 1)
    g = sfa.Particles("photon", 0, 0)
    e.set_photon(g)
    p.set_photon(g)
    g.set_pair(e, p)
    sim = sfa.Simulation(e, p, g, dt=DT, fields=Fields(terms=field), print_every=None,
                         photon_threshold=2.0, seed=0,
                         precision="fp64" if path == "fp64" else "mixed",
                         sampler="mlp" if path == "mixed+mlp" else None)
    for s in (e, p, g):
        s._ensure_device()
    # preallocated SoA capacity (the reference's _extend), so the timed steps
    # grow no buffer
    for s, cap in zip((e, p, g), reserve):
        if cap > s.buffer_size:
            s._extend(cap - s.N_buffered)
    sim._build()
    return sim, (e, p, g)


def _settle(g):
    """Bring sleeping photons to the last step run (the store's catch-up
    callback runs before any read): a one-element download."""
    import ctypes as C
    from sfparticles_amd import _lib as L
    if g.N_buffered:
        one = np.empty(1)
        L.call("sfp_store_download", g._store, L.X, one.ctypes.data_as(C.c_void_p), 0, 1)


def photon_energy_sample(g, size=200000, seed=1, chunk=1 << 24):
    """|u| of `size` live photons drawn uniformly without replacement from the
    whole photon store (every live photon when there are fewer), read in
    chunks so that the store is never copied whole."""
    import ctypes as C
    from sfparticles_amd import _lib as L
    n = g.N_buffered
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for o in range(0, n, chunk):
        m = int(min(chunk, n - o))
        u = []
        for attr in (L.UX, L.UY, L.UZ, L.PRUNED):
            a = np.empty(m, np.uint8 if attr == L.PRUNED else np.float64)
            L.call("sfp_store_download", g._store, attr, a.ctypes.data_as(C.c_void_p), int(o), m)
            u.append(a)
        out.append(np.sqrt(u[0] ** 2 + u[1] ** 2 + u[2] ** 2)[u[3] == 0])
    e = np.concatenate(out) if out else np.zeros(0)
    if e.size > size:
        e = e[np.sort(rng.choice(e.size, size, replace=False))]
    return e


def sample_state(species, every):
    """Every `every`-th live particle of each species (all attributes the
    oracle needs, with ids), for the CPU baseline on the same window."""
    import ctypes as C
    from sfparticles_amd import _lib as L
    out = []
    for sp in species:
        n = sp.N_buffered
        d = {}
        for name, attr, dt in (("x", L.X, np.float64), ("y", L.Y, np.float64),
                               ("z", L.Z, np.float64), ("ux", L.UX, np.float64),
                               ("uy", L.UY, np.float64), ("uz", L.UZ, np.float64),
                               ("id", L.ID, np.uint64), ("pruned", L.PRUNED, np.uint8)):
            a = np.empty(n, dt)
            if n:
                L.call("sfp_store_download", sp._store, attr, a.ctypes.data_as(C.c_void_p), 0, n)
            d[name] = a
        keep = np.flatnonzero(d["pruned"] == 0)[::every]
        out.append({k: np.ascontiguousarray(v[keep]) for k, v in d.items() if k != "pruned"}
                   | {"next_id": int(sp.next_id) if hasattr(sp, "next_id") else 0})
    return out


def gpu_run(n, steps, warmup, nsteps_config, precision="fp64", chunk=250, timing_every=1,
            field="gaussian", cpu_every=0, shared=True):
    from sfparticles_amd import _lib as L
    if warmup > 0:
        wsim, _ = build_gpu(min(n, 100000), nsteps_config, (0, 0, 0), precision, field)
        wsim.run_steps(warmup)
        del wsim
    reserve = (2 * n + 4096, n + 4096, 28 * n + 4096)
    sim, species = build_gpu(n, nsteps_config, reserve, precision, field)
    if not shared:
        sim.set_shared_launches(False)
    # a run shorter than the geometry times the window of `steps` steps
    # centred on the collision (step nsteps_config / 2: the pulse centre
    # meets the bunch centre), after an untimed fast-forward
    start = 0 if steps >= nsteps_config else nsteps_config // 2 - steps // 2
    t0 = time.perf_counter()
    done = 0
    while done < start:
        m = min(chunk, start - done)
        sim.run_steps(m)
        done += m
        print(f"[bench] {precision}: fast-forward {done}/{start} steps, "
              f"{time.perf_counter() - t0:.1f} s, {[s.N_buffered for s in species]}",
              file=sys.stderr, flush=True)
    _settle(species[2])  # sleeping photons at the window's first step (untimed)
    # the CPU baseline's sample: this window's first step when the oracle can
    # run the whole window in seconds, else cpu_sample_at_centre()
    base = sample_state(species, cpu_every) if cpu_every and steps <= CPU_WINDOW else None
    sim.reset_stats()
    sim.reset_timing()
    L.call("sfp_synchronize")
    t0 = time.perf_counter()
    done = 0
    while done < steps:  # chunks, so that progress is visible during long runs
        m = min(chunk, steps - done)
        sim.run_steps(m, timing=True, timing_every=timing_every, timing_kernel=TIMED_KERNEL)
        done += m
        print(f"[bench] {precision}: {done}/{steps} steps, {time.perf_counter() - t0:.1f} s, "
              f"{[s.N_buffered for s in species]}", file=sys.stderr, flush=True)
    _settle(species[2])  # photons asleep in the window brought to its last step (timed)
    L.call("sfp_synchronize")
    t1 = time.perf_counter()
    stats = [sim.stats(i) for i in range(3)]
    qd = [sim.qed_diag(i) for i in range(3)]
    ktime = [[sim.kernel_time(i, ph) for ph in (0, 1, 2)] for i in range(3)]
    print(f"[bench] {precision}: photon spectrum sample from {species[2].N_buffered} photons",
          file=sys.stderr, flush=True)
    spectrum = photon_energy_sample(species[2])
    counts = [s.N_buffered for s in species]
    del sim, species
    return t1 - t0, stats, ktime, counts, spectrum, start, base, qd


def cpu_sample_at_centre(n, nsteps_config, every, field="gaussian", chunk=250):
    """For timed windows longer than CPU_WINDOW steps: an untimed FP64 GPU run
    of the same workload to CPU_WINDOW / 2 steps before the collision centre,
    and the CPU baseline's sample of its state there.  Returns (sample, start)."""
    reserve = (2 * n + 4096, n + 4096, 28 * n + 4096)
    sim, species = build_gpu(n, nsteps_config, reserve, "fp64", field)
    start = nsteps_config // 2 - CPU_WINDOW // 2
    done = 0
    while done < start:
        m = min(chunk, start - done)
        sim.run_steps(m)
        done += m
    print(f"[bench] cpu baseline: FP64 state at step {start} "
          f"{[s.N_buffered for s in species]}", file=sys.stderr, flush=True)
    _settle(species[2])
    base = sample_state(species, every)
    del sim, species
    return base, start


def cpu_baseline(base, start, steps, every):
    """The CPU oracle in reference-structured mode (one OpenMP pass per
    reference function, serial counts/compaction) over the same window of
    steps as the GPU line, on a bounded sample of the same state: every
    `every`-th live particle of each species as the FP64 GPU run held them at
    the window's first step (ids kept, so the draws are those of the same
    particles)."""
    import oracle
    from oracle.sim import OParticles, OSimulation
    from sfparticles_amd.qed.tables import load_rs_tables, load_od_tables
    _, field = make_case(1, 3000)
    made = []
    for d, (name, q, m) in zip(base, (("electron", -1, 1), ("positron", 1, 1), ("photon", 0, 0))):
        k = d["x"].size
        sp = OParticles(name, q, m, N=k, props=[d[a].copy() for a in
                                                ("x", "y", "z", "ux", "uy", "uz")])
        sp.id[:] = d["id"]
        sp.next_id = d["next_id"]
        made.append(sp)
    e, p, g = made
    e.set_photon(g)
    p.set_photon(g)
    g.set_pair(e, p)
    sim = OSimulation(e, p, g, dt=DT, fields=field, photon_threshold=2.0, qed="rs", seed=0,
                      rs_tables=load_rs_tables(), od_tables=load_od_tables())
    for _ in range(start):  # the time as the GPU run accumulates it
        sim.t += DT
    sim.step = start
    psteps = 0

    def cb(i):
        nonlocal psteps
        psteps += e.Npart + p.Npart + g.Npart
        if (i + 1) % 10 == 0:
            print(f"[bench] cpu baseline: {i + 1}/{steps} steps, "
                  f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    n0 = [s_.Npart for s_ in made]
    t0 = time.perf_counter()
    sim.start(steps, call_back=cb, prune=False)
    t1 = time.perf_counter()
    return psteps / (t1 - t0), t1 - t0, oracle.num_threads(), n0


def _short(name):
    """Kernel name without 'void ' and the parameter list."""
    name = name[5:] if name.startswith("void ") else name
    depth = 0
    for i, ch in enumerate(name):
        depth += ch == "<"
        depth -= ch == ">"
        if ch == "(" and depth == 0:
            return name[:i]
    return name


def _csv_rows(d, suffix):
    import csv
    import glob
    files = glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)
    rows = []
    for f in files:
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


_LEPTON = ("k_sweep<false", "k_sweep_pair", "k_hot", "k_emit", "k_step_prologue")


def _step_opens(seq):
    """Indices of the dispatches that open a step: k_step_prologue, or, where
    the first species' sweep takes the step's count snapshot itself (the
    electrons of this benchmark), a lepton sweep that follows no lepton
    kernel (the positrons' sweep follows the electrons' sweep, or their QED
    kernels when each species has its own)."""
    out = []
    for i, r in enumerate(seq):
        name = r[0]
        if name.startswith("k_step_prologue"):
            out.append(i)
        elif (name.startswith(("k_sweep<false", "k_sweep_pair"))
              and not (i and seq[i - 1][0].startswith(_LEPTON))):
            out.append(i)
    return out


def _by_species(seq):
    """Names of the per-species kernels tagged with their species: a step
    launches the lepton kernels for electrons, then for positrons, then the
    photon kernels (_step_opens).  With shared launches one launch sweeps both
    lepton species (k_sweep_pair) and one launch of the QED (and emission)
    kernel serves both: tagged [e-,e+]."""
    steps, cur = [], []
    opens = set(_step_opens(seq))
    for i, r in enumerate(seq):
        if i in opens and cur:
            steps.append(cur)
            cur = []
        cur.append(r)
    if cur:
        steps.append(cur)
    out = []
    for st in steps:
        n = {}
        for r in st:
            n[r[0]] = n.get(r[0], 0) + 1
        sweeps = sum(v for k, v in n.items() if k.startswith("k_sweep<false")) + 2 * sum(
            v for k, v in n.items() if k.startswith("k_sweep_pair"))
        seen = {}
        for r in st:
            name = r[0]
            if (name.startswith("k_sweep_pair")
                    or name.startswith(("k_hot", "k_emit")) and n[name] == 1 and sweeps == 2):
                name = f"{name} [e-,e+]"
            elif name.startswith(("k_sweep", "k_hot", "k_emit", "k_defer")):
                k = seen.get(name, 0)
                seen[name] = k + 1
                name = f"{name} [{('e-', 'e+')[min(k, 1)]}]"
            elif name.startswith(("k_awake", "k_photon")):
                name = f"{name} [photon]"
            out.append((name,) + tuple(r[1:]))
    return out


def _windows(seq, steps):
    """Timed windows of a dispatch sequence [(name, ...)]: the dispatches after
    one settle (k_catchup and the k_catchup_slow after it) up to and
    including the next, when exactly `steps` steps open between them (the
    settle before the timed steps, the timed steps, the settle that ends
    them)."""
    cu = [i for i, r in enumerate(seq) if r[0] == "k_catchup"]

    def end(i):  # the settle's last dispatch
        return i + 1 if i + 1 < len(seq) and seq[i + 1][0] == "k_catchup_slow" else i
    out = []
    for a, b in zip(cu, cu[1:]):
        seg = seq[end(a) + 1:end(b) + 1]
        if len(_step_opens(seg)) == steps:
            out.append(seg)
    return out


def summarize_profile(d):
    """Markdown summary of a profiled run of this benchmark: rocprofv3 output
    under d/kt (kernel trace of the bench command, its JSON line in d/kt.json),
    d/fetch, d/write (FETCH_SIZE, WRITE_SIZE), d/cal_f, d/cal_w (the same
    counters over the copy kernel of --calibrate-copy) and d/sq (SQ
    counters); commands in d/commands.txt.  Per timed window (one per path,
    in the order fp64, mixed, mixed+mlp) and kernel: dispatches, mean
    duration, HBM bytes per dispatch (counters scaled by the copy kernel's
    ratio of counted to known bytes) and the fraction of wave cycles spent
    waiting."""
    import collections
    out = []
    with open(os.path.join(d, "kt.json")) as fh:  # the bench line: the last JSON line
        line = json.loads([ln for ln in fh.read().splitlines() if ln.startswith("{")][-1])
    steps = line["steps"]
    paths = ["fp64", "mixed", "mixed+mlp"]
    cfg = line["config"]
    head = {"fp64-tables": "fp64", "mixed-fp32/fp64-tables": "mixed"}.get(cfg["path"], "mixed+mlp")
    if os.path.exists(os.path.join(d, "commands.txt")):
        out += ["Commands (repository root on one MI355X, `TMPDIR=/tmp`):", "", "```",
                open(os.path.join(d, "commands.txt")).read().strip(), "```", ""]
    out += [f"Bench line of the kernel-trace run: value {line['value']:.4g} particle-steps/s "
            f"({head}), {line['ms_per_step']} ms/step, window {cfg.get('window')}.", "",
            "Roofline in that line: `" + json.dumps(line["roofline"]) + "`", ""]
    # kernel trace
    kt = _csv_rows(os.path.join(d, "kt"), "kernel_trace.csv")
    kt.sort(key=lambda r: int(r["Start_Timestamp"]))
    seq = _by_species([(_short(r["Kernel_Name"]), int(r["Start_Timestamp"]),
                        int(r["End_Timestamp"])) for r in kt])
    wins = _windows(seq, steps)
    trace = {}
    for path, seg in zip(paths, wins):
        t = collections.defaultdict(list)
        for name, a, b in seg:
            t[name].append((b - a) / 1e3)
        # from the first step's start (host work before it is not step time)
        first = seg[_step_opens(seg)[0]][1]
        trace[path] = (t, (seg[-1][2] - first) / 1e3)

    def counters(sub, names):
        rows = _csv_rows(os.path.join(d, sub), "counter_collection.csv")
        byd = collections.OrderedDict()
        for r in rows:
            e = byd.setdefault(int(r["Dispatch_Id"]), [_short(r["Kernel_Name"]),
                                                       collections.defaultdict(float)])
            e[1][r["Counter_Name"]] += float(r["Counter_Value"])
        disp = _by_species([(v[0], v[1]) for _, v in sorted(byd.items())])
        res = {}
        for path, seg in zip(paths, _windows(disp, steps)):
            agg = collections.defaultdict(lambda: collections.defaultdict(list))
            for name, c in seg:
                for k in names:
                    if k in c:
                        agg[name][k].append(c[k])
            res[path] = agg
        return res, disp

    have = {k: os.path.isdir(os.path.join(d, k))
            for k in ("fetch", "write", "sq", "cal_f", "cal_w", "mfma")}
    scale = {}
    for sub, cname in (("cal_f", "FETCH_SIZE"), ("cal_w", "WRITE_SIZE")):
        if have[sub]:
            _, disp = counters(sub, [cname])
            v = [c[cname] for n, c in disp if n.startswith("k_copy")]
            if v:
                # --calibrate-copy: 2^27 doubles read and written per dispatch
                scale[cname] = float(np.median(v)) * 1024.0 / (8.0 * (1 << 27))
    fetch = counters("fetch", ["FETCH_SIZE"])[0] if have["fetch"] else {}
    write = counters("write", ["WRITE_SIZE"])[0] if have["write"] else {}
    sqn = ["SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY",
           "SQ_ACTIVE_INST_VALU", "SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD"]
    sq = counters("sq", sqn)[0] if have["sq"] else {}
    # BF16 matrix operations: 512 FLOPs per SQ_INSTS_VALU_MFMA_MOPS_BF16 count
    mf = counters("mfma", ["SQ_INSTS_VALU_MFMA_MOPS_BF16"])[0] if have["mfma"] else {}
    if scale:
        out += ["Counter scale (counted / known bytes of the copy kernel, 1 GiB read and "
                "1 GiB written per dispatch): " + ", ".join(f"{k} {v:.3f}" for k, v in scale.items())
                + "; HBM bytes below are the counters divided by it.", ""]
    for path in paths:
        if path not in trace:
            continue
        t, wall = trace[path]
        out += [f"## {path}: timed window ({steps} steps and the closing catch-up), "
                f"{wall / 1e3:.3f} ms from the first step's start to the last dispatch's end",
                "",
                "| kernel | dispatches | mean us | total us | HBM read MB / dispatch | "
                "HBM written MB / dispatch | wait / wave cycles | VALU active / wave cycles | "
                "busy cycles / dispatch | VALU insts / dispatch | VMEM read insts / dispatch | "
                "BF16 MFMA TFLOP/s |",
                "|---|---|---|---|---|---|---|---|---|---|---|---|"]
        step_rd = step_wr = 0.0
        for name, v in sorted(t.items(), key=lambda kv: -sum(kv[1])):
            rd = wr = wt = va = bz = vi = vm = mt = ""
            f = fetch.get(path, {}).get(name, {}).get("FETCH_SIZE")
            if f and "FETCH_SIZE" in scale:
                rd = f"{np.mean(f) * 1024 / scale['FETCH_SIZE'] / 1e6:.1f}"
                step_rd += sum(f) * 1024 / scale["FETCH_SIZE"]
            w = write.get(path, {}).get(name, {}).get("WRITE_SIZE")
            if w and "WRITE_SIZE" in scale:
                wr = f"{np.mean(w) * 1024 / scale['WRITE_SIZE'] / 1e6:.1f}"
                step_wr += sum(w) * 1024 / scale["WRITE_SIZE"]
            mo = mf.get(path, {}).get(name, {}).get("SQ_INSTS_VALU_MFMA_MOPS_BF16")
            if mo and sum(mo) > 0:
                mt = f"{np.mean(mo) * 512 / (np.mean(v) * 1e-6) / 1e12:.1f}"
            q = sq.get(path, {}).get(name, {})
            if q.get("SQ_WAVE_CYCLES") and q.get("SQ_WAIT_ANY"):
                wt = f"{sum(q['SQ_WAIT_ANY']) / max(sum(q['SQ_WAVE_CYCLES']), 1):.2f}"
            if q.get("SQ_WAVE_CYCLES") and q.get("SQ_ACTIVE_INST_VALU"):
                va = f"{sum(q['SQ_ACTIVE_INST_VALU']) / max(sum(q['SQ_WAVE_CYCLES']), 1):.2f}"
            if q.get("SQ_BUSY_CYCLES"):
                bz = f"{np.mean(q['SQ_BUSY_CYCLES']):.3g}"
            if q.get("SQ_INSTS_VALU"):
                vi = f"{np.mean(q['SQ_INSTS_VALU']):.3g}"
            if q.get("SQ_INSTS_VMEM_RD"):
                vm = f"{np.mean(q['SQ_INSTS_VMEM_RD']):.3g}"
            out.append(f"| `{name}` | {len(v)} | {np.mean(v):.1f} | {sum(v):.0f} | {rd} | {wr} | "
                       f"{wt} | {va} | {bz} | {vi} | {vm} | {mt} |")
        out.append("")
        if step_rd or step_wr:
            # the same window of the same workload in every run: the line's
            # algorithmic bytes per step apply to the counter runs' windows
            alg = line["roofline"]["whole_step"]["alg_bytes_per_step"] if path == head else None
            tr = (step_rd + step_wr) / steps
            out += [f"HBM traffic of the window, all kernels (counters): {tr / 1e9:.3f} GB per step "
                    f"({step_rd / steps / 1e9:.3f} read, {step_wr / steps / 1e9:.3f} written)"
                    + (f"; algorithmic (the survey's bytes per particle-step, the line's "
                       f"whole_step): {alg / 1e9:.3f} GB per step, traffic / algorithmic "
                       f"{tr / alg:.2f}." if alg else "."), ""]
        if path == head:
            rf = line["roofline"]
            kn = {"k_hot": "k_hot", "k_photon": "k_photon", "k_sweep": "k_sweep",
                  "k_awake": "k_awake", "k_emit": "k_emit"}
            key = next((k for k in kn if rf["kernel"].startswith(k)), None)
            # the roofline kernel serves one species per launch; the trace mean
            # over its dispatches of the species the line names
            sp_tag = "[photon]" if "(photon)" in rf["kernel"] else (
                "[e+]" if "(positron)" in rf["kernel"] else (
                    "[e-,e+]" if "positrons" in rf["kernel"] else "[e-]"))
            cand = [n for n in t if key and n.startswith(key) and n.endswith(sp_tag)]
            if cand:
                name = max(cand, key=lambda n: sum(t[n]))
                durs = t[name]
                mean_ms = float(np.mean(durs)) / 1e3
                ach = rf["alg_bytes_per_launch"] / (mean_ms / 1e3) / 1e9
                who = "its launches for the species the line names"
                out += [f"Roofline kernel `{name}`, {who}: trace mean {mean_ms:.4f} ms per "
                        f"launch against the "
                        f"line's {rf['avg_launch_ms']} ms (HIP events); "
                        f"{rf['alg_bytes_per_launch'] / 1e6:.1f} MB algorithmic per launch -> "
                        f"{ach:.0f} GB/s, frac {ach / HBM_PEAK_GBS:.4f} (line: {rf['frac']}).",
                        ""]
                f = fetch.get(path, {}).get(name, {}).get("FETCH_SIZE")
                w = write.get(path, {}).get(name, {}).get("WRITE_SIZE")
                if f and w and "FETCH_SIZE" in scale and "WRITE_SIZE" in scale:
                    tb = (np.mean(f) / scale["FETCH_SIZE"] + np.mean(w) / scale["WRITE_SIZE"]) * 1024
                    out += [f"Its HBM traffic (FETCH_SIZE + WRITE_SIZE, scaled): {tb / 1e6:.1f} MB "
                            f"per launch, {tb / rf['alg_bytes_per_launch']:.2f} x the algorithmic "
                            f"bytes; {tb / (mean_ms / 1e3) / 1e9:.0f} GB/s while it runs.", ""]
    return "\n".join(out)


def calibrate_copy(n_doubles, reps):
    """Plain streaming copy (for scaling the HBM byte counters)."""
    from sfparticles_amd import _lib as L
    src = L.DeviceArray(8 * n_doubles)
    dst = L.DeviceArray(8 * n_doubles)
    L.call("sfp_copy_f64", src.ptr, dst.ptr, n_doubles, None)
    L.call("sfp_synchronize")
    t0 = time.perf_counter()
    for _ in range(reps):
        L.call("sfp_copy_f64", src.ptr, dst.ptr, n_doubles, None)
    L.call("sfp_synchronize")
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes_read": 8 * n_doubles, "copy_bytes_written": 8 * n_doubles,
                      "reps": reps, "copy_GBps": round(16 * n_doubles / dt / 1e9, 1)}), flush=True)


def hbm_used_gb():
    """Device memory in use (total - free), GB."""
    import ctypes as C
    from sfparticles_amd import _lib as L
    fr, tot = C.c_int64(), C.c_int64()
    L.call("sfp_mem_info", C.byref(fr), C.byref(tot))
    return (tot.value - fr.value) / 1e9


def cascade_run(n, steps, precision, seed=0, x0_ctau=3.0, reserve=(0, 0, 0)):
    """The cascade workload: n electrons at rest uniformly in a ball of radius
    0.1 um at the node of a rotating-E standing wave (two counter-propagating
    circularly polarised plane pulses, each the sum of two linear pulses with
    pol 0 / pi/2 and cep 0 / pi/2; lambda 0.8 um, a0 1000, w0 3 um, tau_FWHM
    30 fs, starting x0_ctau ctau from the node), optical-depth model, photon
    threshold 20 m c^2.  Draw order (numpy PCG64(seed)): radius, direction."""
    import sfparticles_amd as sfa
    from sfparticles_amd import qed, _lib as L
    from sfparticles_amd.fields import Fields
    qed.set_model("od")
    rng = np.random.Generator(np.random.PCG64(seed))
    r = 0.1e-6 * rng.random(n) ** (1 / 3)
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    z0 = np.zeros(n)
    props = (r * v[0], r * v[1], r * v[2], z0, z0.copy(), z0.copy())
    x0 = x0_ctau * CTAU
    field = [(1, (1000.0, 3e-6, CTAU, 2, d, -d * x0, LAMBDA, pol, cep))
             for d in (1, -1) for pol, cep in ((0.0, 0.0), (math.pi / 2, math.pi / 2))]
    e = sfa.Particles("electron", -1, 1, N=n, props=props)
    p = sfa.Particles("positron", 1, 1)
    g = sfa.Particles("photon", 0, 0)
    e.set_photon(g)
    p.set_photon(g)
    g.set_pair(e, p)
    sim = sfa.Simulation(e, p, g, dt=DT, fields=Fields(terms=field), print_every=None,
                         photon_threshold=20.0, seed=0, precision=precision)
    for s_ in (e, p, g):
        s_._ensure_device()
    # preallocated SoA capacity (the reference's _extend), so that the growth
    # of the cascade needs no capacity replay
    for s_, cap in zip((e, p, g), reserve):
        if cap > s_.buffer_size:
            s_._extend(int(cap) - s_.N_buffered)
    sim._prepare()
    L.call("sfp_synchronize")
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        m = min(500, steps - done)
        sim.run_steps(m)
        done += m
        print(f"[bench] cascade {precision} {done}/{steps} steps, {time.perf_counter() - t0:.1f} s, "
              f"{[s_.N_buffered for s_ in (e, p, g)]}", file=sys.stderr, flush=True)
    L.call("sfp_synchronize")
    elapsed = time.perf_counter() - t0
    st = [sim.stats(i) for i in range(3)]
    return dict(value=round(sum(x["psteps"] for x in st) / elapsed, 1),
                ms_per_step=round(elapsed / steps * 1e3, 4),
                final_counts={nm: s_.N_buffered for nm, s_ in
                              zip(("electron", "positron", "photon"), (e, p, g))},
                particle_steps={nm: x["psteps"] for nm, x in
                                zip(("electron", "positron", "photon"), st)},
                capacity_replays=st[0]["replays"], seconds=round(elapsed, 2))


def ensemble_run(nshots, count_max, steps, check_shots, seed=0, ppe=64):
    """The batched-shot workload; returns the JSON line's fields."""
    import sfparticles_amd as sfa
    from sfparticles_amd import qed, _lib as L
    from sfparticles_amd.ensemble import head_on_shots, shot_histogram
    qed.set_model("rs")
    fields, props, shot, table = head_on_shots(nshots, steps, seed=seed, count=(1e3, count_max))
    n = shot.size
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    ids = np.arange(n, dtype=np.uint64)

    def build(sl, flds, precision, rate_net, graph):
        sub = tuple(np.ascontiguousarray(a[sl]) for a in props) + tuple(
            np.ascontiguousarray(a[sl]) for a in v)
        m = sub[0].size
        e = sfa.SpinParticles("electron", -1, 1, N=m, props=sub)
        e.shot = np.ascontiguousarray(shot[sl]) if len(flds) > 1 else np.zeros(m, np.uint32)
        e.id = ids[sl].copy()
        e.next_id = n
        p = sfa.Particles("positron", 1, 1)
        g = sfa.Particles("photon", 0, 0)
        p.shot = np.zeros(0, np.uint32)
        g.shot = np.zeros(0, np.uint32)
        e.set_photon(g)
        p.set_photon(g)
        g.set_pair(e, p)
        sim = sfa.Simulation(e, p, g, dt=DT, fields=None, shots=flds, print_every=None,
                             photon_threshold=2.0, seed=0, precision=precision,
                             spin_emission=True, rate_net=rate_net, graph=graph,
                             autotune=graph)
        for s_, cap in zip((e, p, g), (2 * m + 4096, m + 4096, ppe * m + 4096)):
            s_._ensure_device()
            s_._extend(cap - s_.N_buffered)
        sim._prepare()
        return sim, (e, p, g)

    sim, sp = build(slice(0, n), fields, "mixed", "mlp", True)
    L.call("sfp_synchronize")
    hbm_peak = hbm_used_gb()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        m = min(500, steps - done)
        sim.run_steps(m)
        done += m
        used = hbm_used_gb()
        hbm_peak = max(hbm_peak, used)
        print(f"[bench] ensemble {done}/{steps} steps, {time.perf_counter() - t0:.1f} s, "
              f"{[s_.N_buffered for s_ in sp]}, HBM in use {used:.1f} GB", file=sys.stderr,
              flush=True)
    L.call("sfp_synchronize")
    elapsed = time.perf_counter() - t0
    st = [sim.stats(i) for i in range(3)]
    psteps = sum(x["psteps"] for x in st)
    shapes = {f"{nm}_{phn}": sim.launch_shape(i, ph)[0]
              for i, nm in enumerate(("electron", "positron", "photon"))
              for ph, phn in ((0, "sweep"), (1, "hot")) if sim.launch_shape(i, ph)[0]}
    ens = [shot_histogram(s_, nshots)[0] for s_ in sp]
    counts = [s_.N_buffered for s_ in sp]
    del sim, sp
    # the smallest shots alone on the FP64 table path, same seed electrons and ids
    order = np.argsort(table["n"])[:check_shots]
    off = np.concatenate([[0], np.cumsum(table["n"])])
    checks = []
    for k in order:
        sl = slice(int(off[k]), int(off[k + 1]))
        s1, sp1 = build(sl, [fields[k]], "fp64", None, False)
        s1.run_steps(steps)
        alone = [s_.N_buffered for s_ in sp1]
        ph_e, ph_a = int(ens[2][k]), int(alone[2])
        pa_e, pa_a = int(ens[1][k]), int(alone[1])
        checks.append({"shot": int(k), "a0": round(float(table["a0"][k]), 1),
                       "electrons": int(table["n"][k]),
                       "gamma": round(float(table["gamma"][k]), 1),
                       "photons": [ph_e, ph_a], "pairs": [pa_e, pa_a],
                       "photon_rel_err": round((ph_e - ph_a) / max(ph_a, 1), 5),
                       "pair_z": round((pa_e - pa_a) / math.sqrt(max(pa_e + pa_a, 1)), 3)})
        del s1, sp1
    return dict(value=round(psteps / elapsed, 1), ms_per_step=round(elapsed / steps * 1e3, 4),
                electrons=int(n), shots=int(nshots), final_counts=counts,
                graph_batches=st[0]["graph_batches"], captures=st[0]["captures"],
                tuning_batches=st[0]["tune_batches"], grids=shapes,
                capacity_replays=st[0]["replays"], deferred_conversions=st[2]["deferrals"],
                hbm_used_gb_max=round(hbm_peak, 1), checks_vs_fp64_alone=checks)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=float, default=1e7)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--separate-launches", action="store_true",
                    help="one QED launch per lepton species (default: electrons and "
                         "positrons share theirs; the results are the same)")
    ap.add_argument("--cpu-every", type=int, default=20,
                    help="CPU baseline: every k-th particle of the GPU state at the window start")
    ap.add_argument("--field", choices=["gaussian", "focused"], default="gaussian",
                    help="head-on workload: the paraxial Gaussian pulse (the benchmark) or the "
                         "same pulse in the focused expansion")
    ap.add_argument("--timing-every", type=int, default=5,
                    help="time every step kernel with HIP events on every k-th step (default 5, "
                         "coprime to the photon list's 16-step rebuild period); the electron QED "
                         "kernel, which dominates the step, is timed at every step.  Events "
                         "around every launch of every step cost 5%% of the step time")
    ap.add_argument("--chunk", type=int, default=250,
                    help="head-on: steps per Simulation.run_steps call (a progress line each)")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--summarize-profile", metavar="DIR",
                    help="print a markdown summary of the rocprofv3 output of a profiled run "
                         "of this benchmark under DIR (no GPU work)")
    ap.add_argument("--workload", choices=["head-on", "ensemble", "cascade"], default="head-on")
    ap.add_argument("--shots", type=int, default=256)
    ap.add_argument("--count-max", type=float, default=1e6)
    ap.add_argument("--check-shots", type=int, default=8)
    ap.add_argument("--x0-ctau", type=float, default=3.0,
                    help="cascade: distance of the pulses' start from the node, in ctau")
    ap.add_argument("--reserve", default="0,0,0",
                    help="cascade: electron, positron, photon capacity reserved up front")
    ap.add_argument("--photons-per-electron", type=int, default=64,
                    help="ensemble: photon store capacity reserved per seed electron")
    ap.add_argument("--path", choices=["fp64", "mixed", "mixed+mlp", "all"], default="all",
                    help="fp64: the reference's arithmetic; mixed: FP32 fields, push and QED "
                         "with FP64 positions, time and chi; mixed+mlp: mixed with the learned "
                         "photon-energy sampler; all: the three, reporting mixed+mlp with its "
                         "throughput ratio and yield/spectrum differences to fp64")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: the particle step is measured on one "
              "MI355X (no multi-GPU path is built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.summarize_profile:
        print(summarize_profile(args.summarize_profile))
        return
    if args.calibrate_copy:
        calibrate_copy(1 << 27, 5)
        return

    if args.workload == "cascade":
        steps = args.steps if args.steps != 3000 else int(6 * CTAU / (C_LIGHT * DT))
        paths = ["fp64", "mixed"] if args.path == "all" else [args.path.replace("+mlp", "")]
        res = {pth: cascade_run(10000, steps, pth, x0_ctau=args.x0_ctau,
                                reserve=tuple(float(v) for v in args.reserve.split(",")))
               for pth in paths}
        head = res[paths[-1]]
        line = {"metric": "particle-steps/s (cascade, QED on, optical depth)",
                "value": head["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
                "warmup": 0, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None,
                "dtype": "f64" if paths[-1] == "fp64" else "f32+f64", "data": "synthetic",
                "config": dict(workload="rotating-E standing wave, a0 1000, 1e4 e- at rest at "
                                        f"the node, pulses from {args.x0_ctau:g} ctau, optical "
                                        "depth, photon threshold 20",
                               **{pth: r for pth, r in res.items()})}
        print(json.dumps(line), flush=True)
        return

    if args.workload == "ensemble":
        r = ensemble_run(args.shots, args.count_max, args.steps, args.check_shots,
                         ppe=args.photons_per_electron)
        line = {"metric": "particle-steps/s (ensemble of shots, QED on, spin-resolved)",
                "value": r.pop("value"), "unit": "particle-steps/s", "n_gpus": 1,
                "steps": args.steps, "warmup": 0, "ms_per_step": r.pop("ms_per_step"),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32+f64+bf16", "data": "synthetic",
                "config": dict(workload=f"{args.shots} head-on shots, a0 ~ U(20, 500), "
                                        f"electrons log-uniform in [1e3, {args.count_max:g}], "
                                        f"gamma ~ U(500, 5000); mixed precision, "
                                        f"spin-resolved emission with the learned rate, "
                                        f"graph batches, launch-grid tuner", **r)}
        print(json.dumps(line), flush=True)
        return

    n = int(args.n)
    nsteps_config = 3000
    names = ["electron", "positron", "photon"]

    cpu_every = 0 if args.no_cpu else max(1, int(args.cpu_every))

    def measure(precision):
        elapsed, stats, ktime, counts, spectrum, start, base, qd = gpu_run(
            n, args.steps, args.warmup, nsteps_config, precision, args.chunk, args.timing_every,
            args.field, cpu_every if precision == "fp64" else 0,
            shared=not args.separate_launches)
        # HIP events on the simulation's stream around the launches of every
        # timing_every-th step, and of every step for TIMED_KERNEL: a kernel's
        # time over the window is its mean timed launch times the steps
        def est(i, p):
            ms_, n_ = ktime[i][p]
            return ms_ / n_ * args.steps if n_ else 0.0
        psteps = sum(s["psteps"] for s in stats)
        asleep = stats[2]["slept"]
        # shared QED launches: the positrons' QED kernels run inside the
        # electrons' launches (no launch of their own was timed)
        shared = stats[1]["hot"] > 0 and ktime[0][1][1] > 0 and ktime[1][1][1] == 0
        # dominant kernel: largest time over (species, phase)
        dom, ph = max(((i, p) for i in range(3) for p in (0, 1, 2)), key=lambda ip: est(*ip))
        ms = est(dom, ph)
        both = shared and dom == 0 and ph > 0  # the launch serves both lepton species

        def launch_bytes(i):
            sb_ = survey_bytes(i == 2, ph, stats[i], precision, args.steps)
            return sb_, (sb_ if sb_ is not None else kernel_bytes(i == 2, ph, stats[i]))
        sb, per_launch = launch_bytes(dom)
        if both:
            sb1, _ = launch_bytes(1)
            sb = sb + sb1 if sb is not None and sb1 is not None else None
            per_launch = (sb if sb is not None else
                          kernel_bytes(False, ph, stats[0]) + kernel_bytes(False, ph, stats[1]))
        per_launch /= args.steps
        avg_s = ms / 1e3 / args.steps
        achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
        fk = "gaussian" if args.field == "gaussian" else "general"
        kname = ({(0, False): f"k_sweep<lepton, RS, {fk} field>",
                  (1, False): f"k_hot<RS, {fk} field>", (0, True): "k_awake",
                  (1, True): f"k_photon<RS, {fk} field>",
                  (2, False): f"k_emit<RS, {fk} field>"}[(ph, dom == 2)]
                 + f"{'' if precision == 'fp64' else ', ' + precision} "
                 + ("(electrons and positrons, one launch)" if both else f"({names[dom]})"))
        total_ms = sum(est(i, p) for i in range(3) for p in (0, 1, 2))
        # whole step at the survey's algorithmic bytes per particle-step: e+-
        # 105 B (FP64) or 81 B (mixed: FP64 x, FP32 u), photons 81 B per
        # executed (awake) photon-step; sleeping photons move no bytes
        b_lep = 105 if precision == "fp64" else 81
        step_bytes = b_lep * (stats[0]["psteps"] + stats[1]["psteps"]) + 81 * (
            stats[2]["psteps"] - asleep)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": None,
                    "kernel": kname, "avg_launch_ms": round(ms / args.steps, 4),
                    "alg_bytes_per_launch": int(per_launch),
                    "alg_bytes": "survey per-particle-step bytes" if sb is not None
                    else "kernel's own state bytes (list, re-reads)",
                    "kernel_time_share": round(ms / max(1e-9, total_ms), 3),
                    "timed_launches": ktime[dom][ph][1],
                    "whole_step": {"alg_bytes_per_step": int(step_bytes / args.steps),
                                   "GBps": round(step_bytes / elapsed / 1e9, 1),
                                   "frac": round(step_bytes / elapsed / 1e9 / HBM_PEAK_GBS, 4)}}
        detail = {"value": round(psteps / elapsed, 1),
                  "executed_value": round((psteps - asleep) / elapsed, 1),
                  "ms_per_step": round(elapsed / args.steps * 1e3, 4),
                  "window": [start, start + args.steps],
                  "final_counts": dict(zip(names, counts)),
                  "particle_steps": {names[i]: stats[i]["psteps"] for i in range(3)},
                  "photon_steps_asleep": asleep,
                  "qed_listed": {names[i]: stats[i]["hot"] for i in range(3)},
                  "photon_multistep_pushes": stats[2]["sleeps"],
                  "electron_dense_steps": stats[0]["dense_steps"],
                  "qed_diagnostics": {names[i]: qd[i] for i in range(3)},
                  "kernel_ms": {f"{names[i]}_{('sweep', 'hot', 'emit')[p]}": round(est(i, p), 2)
                                for i in range(3) for p in (0, 1, 2) if p == 0 or ktime[i][p][1]},
                  "shared_lepton_launches": shared,
                  "host_syncs": stats[0]["syncs"], "capacity_replays": stats[0]["replays"]}
        return detail, roofline, counts, spectrum, base, start

    paths = ["fp64", "mixed", "mixed+mlp"] if args.path == "all" else [args.path]
    res = {p: measure(p) for p in paths}

    def errors(path):
        """Yield and spectrum differences of a path to the FP64 table path on
        the same workload, with the statistical resolution of each: 3 sigma
        of the difference of two independent Poisson counts,
        3 sqrt(n64 + n) / n64 (photon ids, and with them the draws of every
        later photon, part once the two paths' decisions differ, so the pair
        counts are all but independent samples; for photon counts, which
        share most draws, this overstates the spread).  The gate: every
        yield within max(1 %, 3 sigma) and the photon spectra's two-sample
        KS p > 0.01."""
        from scipy import stats as st
        _, _, c64, s64, _, _ = res["fp64"]
        _, _, c, sp, _, _ = res[path]
        ks = st.ks_2samp(s64, sp)
        e = {"photon_yield": round((c[2] - c64[2]) / max(c64[2], 1), 6),
             "photon_yield_3sigma": round(3 * np.sqrt(c64[2] + c[2]) / max(c64[2], 1), 6),
             "pair_yield": round((c[1] - c64[1]) / max(c64[1], 1), 6),
             "pair_yield_3sigma": round(3 * np.sqrt(c64[1] + c[1]) / max(c64[1], 1), 6),
             "photons_fp64": int(c64[2]), "pairs_fp64": int(c64[1]),
             "photon_spectrum_ks_distance": round(float(ks.statistic), 6),
             "photon_spectrum_ks_p": float(ks.pvalue),
             "photon_spectrum_sample": [int(len(s64)), int(len(sp))]}
        e["gate"] = bool(abs(e["photon_yield"]) <= max(0.01, e["photon_yield_3sigma"])
                         and abs(e["pair_yield"]) <= max(0.01, e["pair_yield_3sigma"])
                         and e["photon_spectrum_ks_p"] > 0.01)
        return e

    # the headline is the last path named (mixed+MLP by default) if it passes
    # its error gate against FP64, else the FP64 table path
    errs = {p: errors(p) for p in paths if p != "fp64" and "fp64" in res}
    head = paths[-1]
    if head in errs and not errs[head]["gate"]:
        head = "fp64"
    detail, roofline, counts, spectrum, _, start = res[head]
    window = (f"steps {start}-{start + args.steps} (the collision centre)" if start
              else f"steps 0-{args.steps}")
    config = {"workload": f"head-on collision, {n} e-, a0 100, {args.field} pulse, RS tables, "
                          f"{nsteps_config}-step geometry, timed window {window}",
              "path": {"fp64": "fp64-tables", "mixed": "mixed-fp32/fp64-tables",
                       "mixed+mlp": "mixed-fp32/fp64 + bf16-mfma photon-energy network"}[head],
              "electrons": n, "steps_timed": args.steps}
    config.update(detail)
    for p_ in paths:
        if p_ != head:
            config[p_] = res[p_][0]
    if "fp64" in res and head != "fp64":
        config["speedup_vs_fp64"] = round(detail["value"] / res["fp64"][0]["value"], 3)
    if errs:
        config["error_vs_fp64"] = errs
    if "mixed" in res and "mixed+mlp" in res:
        config["sampler_gain"] = round(res["mixed+mlp"][0]["value"] / res["mixed"][0]["value"], 3)

    cpu = None
    if not args.no_cpu:
        base, start64 = (res["fp64"][4], res["fp64"][5]) if "fp64" in res else (None, 0)
        if base is not None:
            csteps = args.steps
            cwin = f"the same window ({window})"
        else:
            base, start64 = cpu_sample_at_centre(n, nsteps_config, cpu_every, args.field,
                                                 args.chunk)
            csteps = CPU_WINDOW
            cwin = (f"the {CPU_WINDOW} steps at the collision centre (steps {start64}-"
                    f"{start64 + CPU_WINDOW}) of the timed run")
        rate, secs, threads, n0 = cpu_baseline(base, start64, csteps, cpu_every)
        cpu = {"value": round(rate, 1), "unit": "particle-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"CPU oracle, reference-structured passes (OpenMP), FP64, over {cwin} "
                         f"on every {cpu_every}th live particle of the FP64 GPU state at its "
                         f"first step ({n0[0]} e-, {n0[1]} e+, {n0[2]} photons; {secs:.1f} s)"}

    line = {
        "metric": "particle-steps/s (1e7 e-, QED on): FP64 tables vs mixed+MLP at <=1% yield err",
        "value": detail["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": detail["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32+f64", "mixed+mlp": "f32+f64+bf16"}[head],
        "data": "synthetic",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
