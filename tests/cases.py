"""Shared test workloads: the same seeded inputs run through the GPU
implementation (sfparticles_amd) and through the CPU oracle (oracle.sim).

Workloads (synthetic inputs, numpy PCG64 streams; draw order: gamma, x, y, z):
  head-on   C1/C2 geometry -- electrons with gamma ~ N(1957, 5%) moving along
            -x, x ~ N(d, 3 um), y, z ~ N(0, 1 um), against a linearly polarised
            paraxial Gaussian pulse (lambda 0.8 um, w0 2 um, tau_FWHM 30 fs in
            intensity -> ctau = c tau / sqrt(2 ln 2) = 7.6386 um, focus x = 0)
            starting at x0 = -d; dt = T/100.  d is chosen so that the pulse
            centre meets the electrons half-way through the run.
  cascade   C3 geometry -- two counter-propagating circularly polarised plane
            pulses (each a sum of two linear pulses, pol 0 / pi/2 with cep
            0 / pi/2) forming a rotating-E standing wave; electrons at rest in
            a ball of radius r around the node.
"""
import math

import numpy as np

C_LIGHT = 299792458.0
UM = 1e-6
LAMBDA = 0.8e-6
DT = LAMBDA / C_LIGHT / 100  # T/100
CTAU_30FS = C_LIGHT * 30e-15 / math.sqrt(2 * math.log(2))


def head_on_case(n, nsteps, seed=0, a0=100.0, qed="rs", species=("e", "p", "g"), rr="PHOTON",
                 spin=False, threshold=2.0, philox_seed=0, w0=2 * UM):
    rng = np.random.Generator(np.random.PCG64(seed))
    gamma = rng.normal(1957.0, 0.05 * 1957.0, n)
    d = C_LIGHT * DT * nsteps / 2
    x = rng.normal(d, 3 * UM, n)
    y = rng.normal(0.0, 1 * UM, n)
    z = rng.normal(0.0, 1 * UM, n)
    ux = -np.sqrt(gamma ** 2 - 1)
    uy = np.zeros(n)
    uz = np.zeros(n)
    s = None
    if spin:
        v = rng.normal(size=(3, n))
        v /= np.linalg.norm(v, axis=0)
        s = (v[0].copy(), v[1].copy(), v[2].copy())
    field = [(2, (a0, w0, CTAU_30FS, 1, -d, LAMBDA, 0.0, 0.0))]
    return dict(props=(x, y, z, ux, uy, uz), spin_props=s, field=field, dt=DT, nsteps=nsteps,
                qed=qed, species=species, rr=rr, spin=spin, threshold=threshold,
                seed=philox_seed)


def cascade_case(n, nsteps, seed=0, a0=1000.0, radius=0.1 * UM, threshold=20.0, qed="od",
                 philox_seed=0, w0=3 * UM, ctau=CTAU_30FS, x0_ctau=3.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    # uniform in a ball
    r = radius * rng.random(n) ** (1 / 3)
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    x, y, z = r * v[0], r * v[1], r * v[2]
    zeros = np.zeros(n)
    x0 = x0_ctau * ctau
    field = []
    for direction in (1, -1):
        for pol, cep in ((0.0, 0.0), (math.pi / 2, math.pi / 2)):
            field.append((1, (a0, w0, ctau, 2, direction, -direction * x0, LAMBDA, pol, cep)))
    return dict(props=(x, y, z, zeros.copy(), zeros.copy(), zeros.copy()), spin_props=None,
                field=field, dt=DT, nsteps=nsteps, qed=qed, species=("e", "p", "g"),
                rr="PHOTON", spin=False, threshold=threshold, seed=philox_seed)


def _species_list(case, make):
    """Build (electrons, positrons, photons) as configured; returns dict name->obj."""
    return make(case)


def run_gpu_case(case, nsteps=None, chunks=1, timing=False, precision="fp64", sampler=None,
                 grids=None, reserve=None, shared=True, **sim_kw):
    """Run a case through sfparticles_amd.Simulation; returns per-species arrays."""
    sim, sp = build_gpu_case(case, precision=precision, sampler=sampler, grids=grids,
                             reserve=reserve, shared=shared, **sim_kw)
    steps = case["nsteps"] if nsteps is None else nsteps
    per = steps // chunks
    done = 0
    for k in range(chunks):
        m = per if k < chunks - 1 else steps - done
        sim.start(m)
        done += m
    return _collect(sp), sim


def build_gpu_case(case, precision="fp64", sampler=None, grids=None, reserve=None, shared=True,
                   **sim_kw):
    """The species and Simulation of a case, not yet run: (sim, [species]).
    shared=False: one QED launch per lepton species."""
    import sfparticles_amd as sfa
    from sfparticles_amd import qed as Q
    from sfparticles_amd.fields import Fields

    Q.set_model(case["qed"])
    RR = getattr(sfa.RadiationReactionType, case["rr"])
    n = case["props"][0].shape[0]
    if case["spin"]:
        props = tuple(case["props"]) + tuple(case["spin_props"])
        e = sfa.SpinParticles("electron", -1, 1, N=n, props=props, RR=RR)
    else:
        e = sfa.Particles("electron", -1, 1, N=n, props=case["props"], RR=RR)
    sp = [e]
    p = g = None
    if "p" in case["species"]:
        p = sfa.Particles("positron", 1, 1, RR=RR if RR != sfa.RadiationReactionType.LL else RR)
        sp.append(p)
    if "g" in case["species"]:
        g = sfa.Particles("photon", 0, 0)
        sp.append(g)
        if RR in (sfa.RadiationReactionType.PHOTON, sfa.RadiationReactionType.NONE):
            e.set_photon(g)
            if p is not None:
                p.set_photon(g)
        if p is not None:
            g.set_pair(e, p)
    sim = sfa.Simulation(*sp, dt=case["dt"], fields=Fields(terms=case["field"]), print_every=None,
                         photon_threshold=case["threshold"], seed=case["seed"], precision=precision,
                         sampler=sampler, **sim_kw)
    for (si, ph), shape in (grids or {}).items():
        if isinstance(shape, tuple):
            sim.set_shape(si, ph, *shape)
        else:
            sim.set_grid(si, ph, shape)
    if not shared:
        sim.set_shared_launches(False)
    # capacities reserved up front (stores otherwise start at the seed count
    # and grow through capacity halts)
    for s, cap in zip(sp, reserve or ()):
        s._ensure_device()
        if cap > s.buffer_size:
            s._extend(cap - s.N_buffered)
    return sim, sp


def _collect(sp):
    out = {}
    for s in sp:
        d = {"n": int(s.N_buffered)}
        for a in ("x", "y", "z", "ux", "uy", "uz", "id"):
            d[a] = np.array(getattr(s, a)[: s.N_buffered])
        if "tau" in s.attrs:
            d["tau"] = np.array(s.tau[: s.N_buffered])
        if getattr(s, "spin", False):
            for a in ("sx", "sy", "sz"):
                d[a] = np.array(getattr(s, a)[: s.N_buffered])
        out[s.name] = d
    return out


def run_oracle_case(case, nsteps=None, chunks=1, libm=False, spin_emission=False, subset=None):
    """Run a case through the oracle.  subset: indices of the seed electrons
    to run (with their ids; particles created get ids from the case's full
    seed count on), e.g. to replay part of a large GPU run."""
    sim, sp = build_oracle_case(case, libm=libm, spin_emission=spin_emission, subset=subset)
    steps = case["nsteps"] if nsteps is None else nsteps
    per = steps // chunks
    done = 0
    for k in range(chunks):
        m = per if k < chunks - 1 else steps - done
        sim.start(m)
        done += m
    out = {}
    for s in sp:
        d = {"n": int(s.N_buffered)}
        for a in ("x", "y", "z", "ux", "uy", "uz", "id"):
            d[a] = np.array(getattr(s, a)[: s.N_buffered])
        if s.tau is not None:
            d["tau"] = np.array(s.tau[: s.N_buffered])
        if s.spin:
            for a in ("sx", "sy", "sz"):
                d[a] = np.array(getattr(s, a)[: s.N_buffered])
        out[s.name] = d
    return out


def build_oracle_case(case, libm=False, spin_emission=False, subset=None):
    """The oracle's species and simulation of a case, not yet run."""
    import oracle
    from oracle.sim import OParticles, OSimulation
    from sfparticles_amd.qed.tables import load_rs_tables, load_od_tables

    od = case["qed"] == "od"
    n_all = case["props"][0].shape[0]
    pick = (lambda a: a.copy()) if subset is None else (lambda a: np.ascontiguousarray(a[subset]))
    n = n_all if subset is None else len(subset)
    rr = case["rr"]
    e = OParticles("electron", -1, 1, N=n, props=[pick(a) for a in case["props"]], RR=rr,
                   spin=case["spin"], spin_props=[pick(a) for a in case["spin_props"]]
                   if case["spin"] else None)
    if subset is not None:
        e.id[:] = np.asarray(subset, dtype=np.uint64)
        e.next_id = n_all
    sp = [e]
    p = g = None
    if "p" in case["species"]:
        p = OParticles("positron", 1, 1, RR=rr)
        sp.append(p)
    if "g" in case["species"]:
        g = OParticles("photon", 0, 0)
        sp.append(g)
        if rr in ("PHOTON", "NONE"):
            e.set_photon(g, od=od)
            if p is not None:
                p.set_photon(g, od=od)
        if p is not None:
            g.set_pair(e, p, od=od)
    sim = OSimulation(*sp, dt=case["dt"], fields=case["field"], photon_threshold=case["threshold"],
                      qed=case["qed"], seed=case["seed"], rs_tables=load_rs_tables(),
                      od_tables=load_od_tables(), libm=libm, spin_emission=spin_emission)
    return sim, sp


def compare_final(gpu, ref, rtol=0.0):
    """Counts and ids must agree exactly; floating-point arrays bit for bit
    (rtol == 0) or within rtol relative to the array's scale."""
    assert set(gpu) == set(ref), (set(gpu), set(ref))
    for name in ref:
        g, r = gpu[name], ref[name]
        assert g["n"] == r["n"], f"{name}: count {g['n']} != {r['n']}"
        assert np.array_equal(g["id"], r["id"]), f"{name}: ids differ"
        for a in r:
            if a in ("n", "id"):
                continue
            if rtol == 0.0:
                bad = ~((g[a] == r[a]) | (np.isnan(g[a]) & np.isnan(r[a])))
                assert not bad.any(), (
                    f"{name}.{a}: {bad.sum()} of {bad.size} differ, first at "
                    f"{np.argmax(bad)}: {g[a][bad][:3]} vs {r[a][bad][:3]}")
            else:
                scale = np.maximum(np.abs(r[a]), np.max(np.abs(r[a])) * 1e-6 + 1e-300)
                err = np.max(np.abs(g[a] - r[a]) / scale) if r[a].size else 0.0
                assert err <= rtol, f"{name}.{a}: max relative error {err:.3e} > {rtol:.1e}"
