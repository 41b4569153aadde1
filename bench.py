"""Benchmark: FP64 zone-updates/s of HARM's conservative timestep on one MI355X.

Workload (one "step" = one full step_ch(): 4 RK substages + dt update):
Fishbone-Moncrief torus around a Kerr hole a=0.9375 in modified Kerr-Schild
coordinates, 256x256 zones (r from inside the horizon to 40 M, hslope 0.3),
rin=6, rmax=12, beta=100 poloidal loop, 4% seeded u perturbation, PPM
reconstruction, HLL fluxes, Toth flux-CT, 2D Newton con-to-prim, floors,
RK4, Courant 0.9, adaptive dt.  Synthetic initial state, all inputs resident
in GPU memory before the timed region.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  `value` is the graph-replayed whole-step throughput;
`roofline` comes from HIP events recorded around every kernel launch on the
launch stream over an eagerly launched copy of the timed region;
`cpu_baseline` is the oracle (FP64 C restatement, OpenMP) on the host cores,
on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 zone-updates/s at 256² torus; batched & adjoint-step cost vs plain step"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit for each hot kernel (read + write of the data
# the kernel must touch at least once; the unit is the kernel's work item).
#   k_flux_single, k_flux  per face (both directions in one launch):
#              p of one zone (64) + packed face geometry (16 doubles, 128)
#              + F (64) + EMF quotient q (8)                         = 264 B
#   k_update   per zone: pi, pb (128) + CENT geometry (gcov upper
#              triangle, gcon, gdet, two floors, corner gdet: 30 doubles,
#              240) + connection (Gamma^k_{lj}, l <= j: 40 doubles, 320)
#              + F1, F2 (128) + q1, q2 (16) + ulast, unew read (128)
#              + pf, ulast, unew written (192) + pflag (4)          = 1156 B
BYTES_PER_UNIT = {"flux": 264, "update": 1156}


def units_per_launch(stage, n1, n2):
    return {"flux": (n1 + 1) * (n2 + 2) + (n1 + 2) * (n2 + 1), "update": n1 * n2}[stage]


def kernel_of(stage, flux_variant):
    if stage == "update":
        return "k_update"
    import harm_ramt_amd as H
    return {H.FLUX_PAIR: "k_flux", H.FLUX_SHARED: "k_flux_shared"}.get(flux_variant, "k_flux_single")


F64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (and matrix) peak


def profile_of(kernel, variants):
    """The committed counter profile of `kernel` (profiles/pmc_roofline.json:
    rocprofv3 --pmc runs of this workload with these kernel variants; bytes
    from FETCH_SIZE and WRITE_SIZE scaled by a copy kernel of known size,
    FP64 flops from SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes), or None if none was
    taken for these variants."""
    path = os.path.join(ROOT, "profiles", "pmc_roofline.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        if d.get("variants") != variants:
            return None
        return d.get("kernels", {}).get(kernel)
    except (OSError, ValueError):
        return None


def setup(n1, n2):
    import harm_ramt_amd as H
    g = H.grid_mks(n1, n2, a=0.9375, Rout=40.0, hslope=0.3)
    geom = H.Geometry(g)
    p0 = H.init_fishmon(geom, gam=4.0 / 3.0, rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=7)
    prm = H.make_params(n1, n2, g.dx1, g.dx2, lim=H.PARA, fluxmethod=H.HLLFLUX, timeorder=4, cour=0.9)
    return g, geom, p0, prm


def grid_scan(sizes=(512, 1024), steps=40):
    """The same torus at larger grids (not part of `value`): zone-updates/s
    once the grid holds more wavefronts than the chip has SIMDs (256^2 is
    1,024 update wavefronts for 1,024 SIMDs)"""
    import harm_ramt_amd as H
    out = {}
    for n in sizes:
        g, geom, p0, prm = setup(n, n)
        h = H.Harm(prm, geom)
        h.set_array(H.P, p0)
        h.bound_prim(H.P)
        h.set_time(0.0, 1e-5, 2000.0)
        h.step_n(8)                        # autotuner, graphs
        h.get_time()
        t0 = time.perf_counter()
        h.step_n(steps)
        h.get_time()
        el = time.perf_counter() - t0
        fv, bv = h.get_tuning()
        # the flux and update kernels' HBM rates at this size, from HIP events
        # around each launch of an eagerly launched continuation of the run
        h.prof_enable(True)
        h.prof_reset()
        h.step_n(min(steps, 20))
        h.get_time()
        kernels = {}
        for kid, name in H.STAGE_NAMES.items():
            ms, cnt = h.prof_get(kid)
            if cnt and name in BYTES_PER_UNIT:
                avg = 1e3 * ms / cnt
                gbs = BYTES_PER_UNIT[name] * units_per_launch(name, n, n) / (avg * 1e-6) / 1e9
                kernels[kernel_of(name, fv)] = {"avg_launch_us": round(avg, 3), "achieved_GBs": round(gbs, 1),
                                                "frac": round(gbs / HBM_PEAK_GBS, 4)}
        h.prof_enable(False)
        out[f"{n}x{n}"] = {"zone_updates_per_s": steps * n * n / el, "ms_per_step": 1e3 * el / steps,
                           "steps": steps, "variants": [fv, bv], "kernels": kernels}
        h.close()
        del geom, p0
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _oracle_rate(geom, p0, prm, n1, n2, threads, seconds):
    """zone-updates/s of the oracle on `threads` OpenMP threads, stepping the
    torus from t=0 for at least `seconds` (the first step, a first touch of
    the arrays, untimed)"""
    import harm_ramt_amd as H
    from oracle.oracle import Oracle
    o = Oracle(H.params_dict(prm), nthreads=threads)
    o.set_geometry(*geom.tables())
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, 1e-5, 2000.0)
    o.step()
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step()
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds and steps >= 3:
            break
    return steps * n1 * n2 / el, steps, el


def cpu_baseline(geom, p0, prm, n1, n2, seconds):
    """The oracle on the host: every core the OS reports (os.cpu_count()) and
    the process's own CPU set are both tried for a few seconds, and the faster
    thread count is timed for `seconds`; both rates are reported."""
    ncpu = os.cpu_count() or 1
    try:
        nset = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        nset = ncpu
    trial = {}
    for th in sorted({ncpu, nset}):
        trial[th] = _oracle_rate(geom, p0, prm, n1, n2, th, 3.0)[0]
    best = max(trial, key=trial.get)
    rate, steps, el = _oracle_rate(geom, p0, prm, n1, n2, best, seconds)
    tried = ", ".join(f"{k} threads {v:.3g}" for k, v in sorted(trial.items()))
    return {"value": rate, "unit": "zone-updates/s", "cores": best, "kind": "port", "nproc": ncpu,
            "affinity_cpus": nset, "cpu_model": cpu_model(),
            "sample": f"{steps} steps of the same {n1}x{n2} torus from t=0 (steps 2..{steps + 1}), "
                      f"{el:.1f} s on {best} OpenMP threads (the faster of the 3 s trials: {tried} zone-updates/s)"}


def batch_draws(nruns, seed=2024):
    """The batched sweep's runs, drawn from one seed: grids n1 in [64, 512]
    and n2 in [64, 256] (multiples of 16), spin in [0, 0.99], r_max in
    [10, 16], beta in [30, 300], and the seed of each run's +-2 % u
    perturbation."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [dict(n1=int(rng.integers(4, 33)) * 16, n2=int(rng.integers(4, 17)) * 16, a=float(rng.uniform(0, 0.99)),
                 rmax=float(rng.uniform(10, 16)), beta=float(rng.uniform(30, 300)), seed=int(rng.integers(1, 2**31)))
            for _ in range(nruns)]


def build_torus(d):
    """(geometry, initial state, parameters) of one drawn torus run: MKS
    Rout 40 M, hslope 0.3, rin 6, PPM, HLL, RK4, Courant 0.9"""
    import harm_ramt_amd as H
    g = H.grid_mks(d["n1"], d["n2"], a=d["a"], Rout=40.0, hslope=0.3)
    geom = H.Geometry(g)
    p0 = H.init_fishmon(geom, rin=6.0, rmax=d["rmax"], beta=d["beta"], randfact=4e-2, seed=d["seed"])
    prm = H.make_params(d["n1"], d["n2"], g.dx1, g.dx2, lim=H.PARA, fluxmethod=H.HLLFLUX, timeorder=4, cour=0.9)
    return geom, p0, prm


def batch_sweep(nruns, steps, warmup, seed=2024, tuning=None):
    """Batched sweep of independent FM tori (grids between 64x64 and 512x256,
    spin in [0, 0.99], r_max in [10, 16], beta in [30, 300], +-2 % u
    perturbations, all drawn from one seed): zone-updates/s of the batch
    advanced by one graph per batch step, against the same contexts stepped
    one after another (each with its own graph).  Both legs time `steps`
    steps; the batched leg is timed before and after the separate leg and
    averaged, so the slow drift of the states does not favour either."""
    from concurrent.futures import ThreadPoolExecutor
    import harm_ramt_amd as H
    draws = batch_draws(nruns, seed)
    build = build_torus

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        built = list(ex.map(build, draws))
    runs = []
    for geom, p0, prm in built:
        h = H.Harm(prm, geom)
        h.set_array(H.P, p0)
        h.bound_prim(H.P)
        h.set_time(0.0, 1e-5, 2000.0)
        runs.append(h)
    del built
    zones = sum(d["n1"] * d["n2"] for d in draws)
    b = H.Batch(runs)
    if tuning:                             # forced variants (profiling runs): no autotuner
        b.set_tuning(tuning[0] if tuning[0] in (H.FLUX_FACE_64x4, H.FLUX_FACE_64x2) else H.FLUX_FACE_64x2,
                     tuning[1] if tuning[1] in (H.BOUND_KERNELS, H.BOUND_BLOCK) else H.BOUND_KERNELS)
        for h in runs:
            h.set_tuning(*tuning)
    b.step_n(warmup)                       # includes the batch autotuner
    for h in runs:
        h.step_n(1)                        # each context's own autotuner and graph
    t0 = time.perf_counter()
    b.step_n(steps)
    tb1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    for h in runs:
        h.step_n(steps)
    ts = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.step_n(steps)
    tb2 = time.perf_counter() - t0
    failed = sum(h.failures()[1] for h in runs)
    fv, bv = b.get_tuning()
    vb = steps * zones / (0.5 * (tb1 + tb2))
    vs = steps * zones / ts
    b.close()
    for h in runs:
        h.close()
    return {"runs": nruns, "zones_per_batch_step": zones, "steps": steps, "seed": seed,
            "grids": [f"{d['n1']}x{d['n2']}" for d in draws],
            "batched_zone_updates_per_s": vb, "separate_zone_updates_per_s": vs, "speedup": vb / vs,
            "batched_ms_per_step": 1e3 * 0.5 * (tb1 + tb2) / steps, "separate_ms_per_step": 1e3 * ts / steps,
            "variants": {"flux": fv, "boundary": bv}, "failed_runs": failed}


def ffde_measure(steps, warmup, nspins=16, n=512, a_single=0.9):
    """Force-free split monopole (EOMFFDE, MKS n x n, Rout 100 M, hslope 0.1,
    MC, LAXF, midpoint RK2, flux-CT; the reference's force-free settings,
    initbase.c:151-157, init.bz.c:44-55): zone-updates/s of one run at spin
    a_single, and of nspins runs with spins evenly spaced in [0.1, 0.99]
    advanced as one batch against the same contexts stepped one after
    another (each with its own graph).  Synthetic initial state: B^r from the
    split monopole, zero drift velocity."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import harm_ramt_amd as H

    def build(a):
        g = H.grid_mks(n, n, a=a, Rout=100.0, hslope=0.1)
        geom = H.Geometry(g)
        p0 = H.init_monopole(geom)
        prm = H.make_params(n, n, g.dx1, g.dx2, lim=H.MC, fluxmethod=H.LAXFFLUX, timeorder=2, eomtype=H.EOMFFDE,
                            poledeath=0)
        h = H.Harm(prm, geom)
        h.set_array(H.P, p0)
        h.bound_prim(H.P)
        h.set_time(0.0, 1e-3, 1e30)
        return h

    h = build(a_single)
    h.step_n(warmup)
    h.get_time()
    t0 = time.perf_counter()
    h.step_n(steps)
    h.get_time()
    t_single = (time.perf_counter() - t0) / steps
    single = {"spin": a_single, "zone_updates_per_s": n * n / t_single, "ms_per_step": 1e3 * t_single,
              "tuning": list(h.get_tuning())}
    h.close()
    spins = [float(x) for x in np.linspace(0.1, 0.99, nspins)]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        runs = list(ex.map(build, spins))
    b = H.Batch(runs)
    b.step_n(warmup)
    for r in runs:
        r.step_n(1)
    zones = nspins * n * n
    t0 = time.perf_counter()
    b.step_n(steps)
    tb1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    for r in runs:
        r.step_n(steps)
    ts = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.step_n(steps)
    tb2 = time.perf_counter() - t0
    failed = sum(r.failures()[1] for r in runs)
    vb, vs = steps * zones / (0.5 * (tb1 + tb2)), steps * zones / ts
    out = {"workload": f"force-free split monopole, MKS {n}x{n}, Rout 100, MC+LAXF+RK2, flux-CT",
           "single": single,
           "batch": {"runs": nspins, "spins": [round(x, 6) for x in spins], "steps": steps,
                     "batched_zone_updates_per_s": vb, "separate_zone_updates_per_s": vs, "speedup": vb / vs,
                     "batched_ms_per_step": 1e3 * 0.5 * (tb1 + tb2) / steps, "separate_ms_per_step": 1e3 * ts / steps,
                     "variants": list(b.get_tuning()), "failed_runs": failed}}
    b.close()
    for r in runs:
        r.close()
    return out


def spin_derivative_measure(n=512, nsteps=100, a0=0.9, r_flux=20.0, growth_steps=400):
    """d L / d a for the force-free split monopole (config of ffde_measure at
    spin a0): L = the outgoing energy flux through the zone row nearest
    r_flux after nsteps steps.  The tangent-linear run (harm_tl: the step on
    value-plus-tangent numbers, dt sequence of the values) starts from the
    spin tangents of the geometry tables, the grid spacing and the initial
    state (fourth-order differences of the host generators); checked against
    the central differences of two plain runs at a0 -+ h (h = 1e-5, 1e-7) with
    their own grids and initial states and the base run's dt sequence
    imposed.  The run is then continued to growth_steps steps recording the
    largest state tangent: the linearised force-free step amplifies
    perturbations at the equatorial current sheet by a roughly constant
    factor per step, which bounds the horizon over which a derivative of this
    problem is meaningful.  Also the cost of a tangent-linear step against a
    plain (graph) step."""
    import numpy as np
    import harm_ramt_amd as H

    def grid(a):
        return H.grid_mks(n, n, a=a, Rout=100.0, hslope=0.1)

    def ctx(a):
        g = grid(a)
        geom = H.Geometry(g)
        p0 = H.init_monopole(geom)
        prm = H.make_params(n, n, g.dx1, g.dx2, lim=H.MC, fluxmethod=H.LAXFFLUX, timeorder=2, eomtype=H.EOMFFDE,
                            poledeath=0)
        h = H.Harm(prm, geom)
        h.set_array(H.P, p0)
        h.bound_prim(H.P)
        h.set_time(0.0, 1e-3, 1e30)
        h.set_tuning(H.FLUX_FACE_64x2, H.BOUND_FUSED)
        return g, p0, h

    g, p0, base = ctx(a0)
    row = int(np.argmin([abs(H.zone_coords(g, i, n // 2)[2] - r_flux) for i in range(n)]))
    t0 = time.perf_counter()
    dgcov, dgcon, dgdet, dconn, dp0, ddx1, ddx2 = H.spin_tangents(grid, H.init_monopole, a0)
    t_tables = time.perf_counter() - t0
    tl = H.Tangent(base)
    tl.set_geometry_tangent(dgcov, dgcon, dgdet, dconn)
    tl.set_spacing_tangent(ddx1, ddx2)
    tl.set_state_tangent(dp0)
    del dgcov, dgcon, dgdet, dconn
    dts = []
    t0 = time.perf_counter()
    for _ in range(nsteps):
        dts.append(base.get_time()[1])
        tl.step_n(1)
    t_tl_run = time.perf_counter() - t0
    L, dL = tl.energy_flux(row)

    def fd_run(a):
        _, _, h = ctx(a)
        tt = 0.0
        for dt in dts:
            h.set_time(tt, dt, 1e30)
            h.step_ch()
            tt = h.get_time()[0]
        tz = H.Tangent(h)
        v = tz.energy_flux(row)[0]
        tz.close()
        h.close()
        return v

    fd = {f"{hh:g}": (fd_run(a0 + hh) - fd_run(a0 - hh)) / (2 * hh) for hh in (1e-5, 1e-7)}
    # growth of the tangent beyond the horizon: largest |dp/da| over the
    # interior every 50 steps, and where it sits at the end
    growth = []
    for s in range(nsteps, growth_steps, 50):
        tl.step_n(50)
        d = np.abs(tl.state_tangent()[3:-3, 3:-3])
        growth.append([s + 50, float(np.nanmax(d)) if np.isfinite(d).any() else None])
    d = np.nan_to_num(np.abs(tl.state_tangent()[3:-3, 3:-3]), nan=np.inf).max(axis=2)
    big = np.argwhere(d >= np.sort(d, axis=None)[-min(100, d.size)])
    # steady-state cost: 50 tangent-linear steps against 50 plain graph steps
    # continuing the same run
    k = 50
    base.step_n(2)
    base.get_time()
    t0 = time.perf_counter()
    base.step_n(k)
    base.get_time()
    t_plain = (time.perf_counter() - t0) / k
    t0 = time.perf_counter()
    tl.step_n(k)
    t_tl = (time.perf_counter() - t0) / k
    tl.close()
    base.close()
    return {"workload": f"force-free split monopole a={a0}, MKS {n}x{n}, MC+LAXF+RK2, {nsteps} steps",
            "functional": f"outgoing energy flux through zone row {row} (r ~ {r_flux} M) after the last step",
            "L": L, "dL_da_tangent_linear": dL, "dL_da_central_difference": fd,
            "rel_diff": {k: abs(dL - v) / abs(v) for k, v in fd.items()},
            "tangent_max_by_step": growth,
            "largest_tangents_zone_range": {"i": [int(big[:, 0].min()), int(big[:, 0].max())],
                                            "j": [int(big[:, 1].min()), int(big[:, 1].max())]},
            "tangent_step_ms": 1e3 * t_tl, "plain_step_ms": 1e3 * t_plain, "tangent_cost_vs_plain_step": t_tl / t_plain,
            "tangent_run_s": t_tl_run, "tangent_tables_s": t_tables}


def adjoint_measure(n=256, nsteps=256, K=None, seed=3):
    """Discrete adjoint on the Orszag-Tang problem (Minkowski, periodic, MC,
    HLL, RK4, flux-CT) at n x n over nsteps steps with checkpoints every K:
    J = magnetic energy at the end, gradients with respect to the field
    amplitude B0, the velocity amplitude and a seeded Gaussian density
    perturbation field (sigma 1e-3, also part of the state).  Reports the
    cost of forward + reverse sweep against the plain step (steady state: step
    graphs captured and kernel variants chosen before the timed sweep), and the adjoint
    against central finite differences (h = 1e-6, same time steps).  K
    (checkpoint interval) defaults to nsteps when every substage state of the
    run fits in 32 GB of HBM: the forward run then keeps them all and the
    reverse sweep recomputes nothing."""
    import numpy as np
    import harm_ramt_amd as H
    g = H.grid_cartesian(n, n)
    geom = H.Geometry(g)
    def params(lim):
        return H.make_params(n, n, g.dx1, g.dx2, lim=lim, fluxmethod=H.HLLFLUX, timeorder=4, gam=5.0 / 3.0,
                             cour=0.8, bc=(H.PERIODIC,) * 4, fixupflux=0, poledeath=0)
    prm = params(H.MC)
    rng = np.random.default_rng(seed)
    drho = np.zeros((n + 6, n + 6))
    drho[3:-3, 3:-3] = rng.normal(0, 1e-3, (n, n))

    def state(B0, vamp, gam=5.0 / 3.0):
        i = np.arange(-3, n + 3)[:, None]
        j = np.arange(-3, n + 3)[None, :]
        x = g.startx1 + (i + 0.5) * g.dx1 + 0 * j
        y = g.startx2 + (j + 0.5) * g.dx2 + 0 * i
        v1, v2 = -vamp * np.sin(2 * np.pi * y), vamp * np.sin(2 * np.pi * x)
        gamma = 1 / np.sqrt(1 - v1 * v1 - v2 * v2)
        p = np.zeros((n + 6, n + 6, 8))
        p[..., 0] = 25 / (36 * np.pi) + drho
        p[..., 1] = 5 / (12 * np.pi) / (gam - 1)
        p[..., 2], p[..., 3] = gamma * v1, gamma * v2
        p[..., 5], p[..., 6] = -B0 * np.sin(2 * np.pi * y), B0 * np.sin(4 * np.pi * x)
        return p

    B0, vamp = 1.0 / np.sqrt(4 * np.pi), 0.5
    p0 = state(B0, vamp)
    if K is None:
        state_bytes = (n + 6) * (n + 6) * 8 * 8
        K = nsteps if (nsteps * 4 + 1) * state_bytes <= 32 * 2**30 else 16

    def ctx(p, pr=None):
        h = H.Harm(pr or prm, geom)
        h.set_array(H.P, p)
        h.bound_prim(H.P)
        h.set_time(0.0, 1e-4, 1e30)
        return h

    # plain step (graph replay) on the same problem
    h = ctx(p0)
    h.step_n(3)
    h.get_time()
    tuning = h.get_tuning()
    t0 = time.perf_counter()
    h.step_n(nsteps)
    h.get_time()
    t_plain = (time.perf_counter() - t0) / nsteps
    h.close()
    # adjoint: forward with checkpoints, then the reverse sweep, in steady
    # state: one untimed sweep first captures the step graphs (the kernel
    # variants are the plain run's), then the state is reset and timed
    h = ctx(p0)
    h.set_tuning(*tuning)
    a = H.Adjoint(h, nsteps, K)
    a.forward()
    a.seed_magnetic_energy()
    a.backward()
    h.set_array(H.P, p0)
    h.bound_prim(H.P)
    h.set_time(0.0, 1e-4, 1e30)
    h.get_time()
    t0 = time.perf_counter()
    a.forward()
    t_fwd = time.perf_counter() - t0
    J = a.seed_magnetic_energy()
    t0 = time.perf_counter()
    lam = a.backward()
    t_bwd = time.perf_counter() - t0
    dts = a.dts()
    nonsmooth = a.nonsmooth()
    a.close()
    h.close()

    def J_of(p, pr, dts, keep=False):
        hh = ctx(p, pr)
        hh.set_time(0.0, dts[0], 1e30)
        aa = H.Adjoint(hh, nsteps, nsteps)
        aa.forward(dts)
        val = aa.seed_magnetic_energy()
        if keep:  # every state kept (checkpoint_every = nsteps) for branch_switches
            return val, (hh, aa)
        aa.close()
        hh.close()
        return val

    def J_pair(name, e, pr, dts):
        """J(+e), J(-e) and the limiter branches that differ between the two
        forward runs (every substage's reconstruction input)"""
        Jp, rp = J_of(perturbed(name, e), pr, dts, keep=True)
        Jm, rm = J_of(perturbed(name, -e), pr, dts, keep=True)
        sw, cmp_ = rp[1].branch_switches(rm[1])
        for hh, aa in (rp, rm):
            aa.close()
            hh.close()
        return Jp, Jm, sw, cmp_

    hh = 1e-6
    w = np.zeros_like(p0)
    w[3:-3, 3:-3, 0] = rng.normal(0, 1, (n, n))

    def perturbed(name, e):
        if name == "B0":
            return state(B0 + e, vamp)
        if name == "v_amplitude":
            return state(B0, vamp + e)
        return p0 + e * w

    def checks_of(lam, pr, dts, sweep=False):
        out = {}
        for name in ("B0", "v_amplitude", "rho_perturbation_direction"):
            dpdq = w if name == "rho_perturbation_direction" else (perturbed(name, hh) - perturbed(name, -hh)) / (2 * hh)
            ad = float((lam * dpdq).sum())
            Jp, Jm, sw, ncmp = J_pair(name, hh, pr, dts)
            fd = (Jp - Jm) / (2 * hh)
            scale = float(np.abs(lam * dpdq).sum())
            out[name] = {"adjoint": ad, "finite_difference": fd, "abs_diff_over_term_scale": abs(ad - fd) / scale,
                         "rel_diff": abs(ad - fd) / abs(fd), "limiter_branch_switches": sw,
                         "slopes_compared": ncmp}
            if sweep:
                # a kink of J (a limiter branch switch between the +h and -h
                # runs) adds an error that shrinks with h, unlike a wrong
                # derivative: the differences and the switches per step size
                J0 = J_of(p0 if name == "rho_perturbation_direction" else perturbed(name, 0.0), pr, dts)
                out[name]["one_sided_plus"] = (Jp - J0) / hh
                out[name]["one_sided_minus"] = (J0 - Jm) / hh
                rels, sws = {}, {}
                for e in (1e-7, 1e-8, 1e-9):
                    jp, jm, swe, _ = J_pair(name, e, pr, dts)
                    f = (jp - jm) / (2 * e)
                    rels[f"{e:g}"] = abs(ad - f) / abs(f)
                    sws[f"{e:g}"] = swe
                out[name]["rel_diff_vs_fd_step"] = rels
                out[name]["limiter_branch_switches_vs_fd_step"] = sws
        return out

    checks = checks_of(lam, prm, dts, sweep=True)
    # MC's slope switches are kinks of J: finite differences straddle them.
    # The same checks with a smooth limiter (NLIMCENT), where central
    # differences converge to the derivative
    prm_s = params(H.NLIMCENT)
    hs = ctx(p0, prm_s)
    as_ = H.Adjoint(hs, nsteps, K)
    as_.forward()
    as_.seed_magnetic_energy()
    lam_s = as_.backward()
    dts_s = as_.dts()
    ns_s = as_.nonsmooth()
    as_.close()
    hs.close()
    checks_smooth = checks_of(lam_s, prm_s, dts_s)
    checks_smooth["limiter"] = "NLIMCENT"
    checks_smooth["nonsmooth_points"] = ns_s[0]
    return {"workload": f"Orszag-Tang {n}x{n}, Minkowski, periodic, MC+HLL+RK4, flux-CT, {nsteps} steps, "
                        f"checkpoints every {K}, J = magnetic energy at the end",
            "J": J, "plain_ms_per_step": 1e3 * t_plain,
            "adjoint_forward_ms_per_step": 1e3 * t_fwd / nsteps, "adjoint_reverse_ms_per_step": 1e3 * t_bwd / nsteps,
            "adjoint_cost_vs_plain_step": (t_fwd + t_bwd) / nsteps / t_plain,
            "reverse_sweep_cost_vs_plain_step": t_bwd / nsteps / t_plain,
            "nonsmooth_points": nonsmooth[0], "failed_inversions_in_recompute": nonsmooth[1],
            "gradient_checks": checks, "gradient_checks_smooth_limiter": checks_smooth}


def copy_calibration(mib):
    """A plain copy of known size: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
    this run gives the counter reading for mib MiB read and mib MiB written."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.arange(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(10):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 10}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n1", type=int, default=256)
    ap.add_argument("--n2", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-runs", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-grid-scan", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-acct", action="store_true")
    ap.add_argument("--no-ffde", action="store_true")
    ap.add_argument("--ffde-steps", type=int, default=20)
    ap.add_argument("--no-spin", action="store_true")
    ap.add_argument("--spin-steps", type=int, default=100)
    ap.add_argument("--eager", action="store_true",
                    help="launch the timed steps eagerly instead of replaying a graph (for profilers)")
    ap.add_argument("--tuning", default=None,
                    help="force the kernel variants as FLUX,BOUND (harm_set_tuning); default: autotuned")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB ten times (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the path is a single-GPU grid; "
              "multi-GPU halo exchange is design-only). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import harm_ramt_amd as H
    n1, n2 = args.n1, args.n2
    g, geom, p0, prm = setup(n1, n2)
    h = H.Harm(prm, geom)
    h.set_array(H.P, p0)
    h.bound_prim(H.P)
    h.set_time(0.0, 1e-5, 2000.0)   # initbase.c:164 dt=1e-5, tf well beyond the run
    if args.tuning:
        h.set_tuning(*[int(v) for v in args.tuning.split(",")])

    if args.eager:
        h.prof_enable(True)         # eager launches, per-stage events
    h.step_n(args.warmup)           # the first call also times the kernel variants and keeps the fastest
    h.get_time()                    # synchronises
    fv, bv = h.get_tuning()
    ns0, wm0 = h.newton_spread()    # synchronises
    t0 = time.perf_counter()
    h.step_n(args.steps)            # HIP-graph replay, no host sync inside
    t_now, dt_now = h.get_time()    # synchronises
    el = time.perf_counter() - t0
    value = args.steps * n1 * n2 / el
    ns1, wm1 = h.newton_spread()
    # the Newton iterations of the timed steps: needed per zone inversion, and
    # issued (each wavefront runs its slowest zone's count) over needed
    ninv = args.steps * 4 * n1 * n2
    newton = {"iterations_per_inversion": (ns1 - ns0) / ninv,
              "wavefront_max_per_inversion": 64 * (wm1 - wm0) / ninv,
              "issued_over_needed": 64 * (wm1 - wm0) / max(1, ns1 - ns0)}
    nfail, failed = h.failures()
    if failed:
        print(f"bench.py: the run set the failure flag (t={t_now})", file=sys.stderr)
        sys.exit(1)

    # per-stage HIP-event timing over an eagerly launched copy of the timed region
    h.prof_enable(True)
    h.prof_reset()
    prof_steps = min(args.steps, 100)
    t1 = time.perf_counter()
    h.step_n(prof_steps)
    h.get_time()
    el_prof = time.perf_counter() - t1
    stages = {}
    for kid, name in H.STAGE_NAMES.items():
        ms, cnt = h.prof_get(kid)
        if cnt:
            stages[name] = {"total_ms": ms, "launches": cnt, "avg_us": 1e3 * ms / cnt}
    h.prof_enable(False)
    dom = max((k for k in stages if k in BYTES_PER_UNIT), key=lambda k: stages[k]["total_ms"])
    kname = kernel_of(dom, fv)
    bytes_launch = BYTES_PER_UNIT[dom] * units_per_launch(dom, n1, n2)
    achieved = bytes_launch / (stages[dom]["avg_us"] * 1e-6) / 1e9
    variants = {"flux": fv, "boundary": bv}
    prof = profile_of(kname, variants)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": prof["traffic_bytes_per_launch"] if prof else None,
                "kernel": kname, "avg_launch_us": round(stages[dom]["avg_us"], 3),
                "algorithmic_bytes_per_launch": bytes_launch}
    if prof:  # the FP64 VALU rate beside the HBM one: profiled flops per launch / live launch time
        tf = prof["f64_flops_per_launch"] / (stages[dom]["avg_us"] * 1e-6) / 1e12
        roofline["valu_f64"] = {"achieved": round(tf, 3), "peak": F64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac": round(tf / F64_PEAK_TFLOPS, 4),
                                "flops_per_launch": prof["f64_flops_per_launch"]}

    # the same run continued with the reference's conservation accounting on
    # (diag_flux, diag_source, diag_fixup every step; not part of `value`)
    accounting = None
    if not args.no_acct:
        ep, df = H.ener_regions(g, 0.4)
        h.acct_setup(ep, df)
        h.step_n(2)
        h.get_time()
        na = min(args.steps, 50)
        t2 = time.perf_counter()
        h.step_n(na)
        h.get_time()
        ms_acct = 1e3 * (time.perf_counter() - t2) / na
        # the run continued with the tree sums (harm_acct_set_order)
        h.acct_setup(ep, df)
        h.acct_set_order(H.ACCT_TREE)
        h.step_n(2)
        h.get_time()
        t2 = time.perf_counter()
        h.step_n(na)
        h.get_time()
        ms_tree = 1e3 * (time.perf_counter() - t2) / na
        plain = 1e3 * el / args.steps
        accounting = {"steps": na, "ms_per_step": ms_acct, "vs_plain_step": ms_acct / plain, "regions": 3,
                      "tree_sums": {"ms_per_step": ms_tree, "vs_plain_step": ms_tree / plain}}
        h.acct_off()

    h.close()
    forced = [int(v) for v in args.tuning.split(",")] if args.tuning else None
    batched = None if args.no_batch else batch_sweep(args.batch_runs, args.batch_steps, 3, tuning=forced)
    scan = None if args.no_grid_scan else grid_scan()
    adjoint = None if args.no_adjoint else adjoint_measure()
    ffde = None if args.no_ffde else ffde_measure(args.ffde_steps, 3)
    if ffde is not None and not args.no_spin:
        ffde["spin_derivative"] = spin_derivative_measure(nsteps=args.spin_steps)
    # the spin derivative's growth diagnostic continues the tangent run
    cpu = None if args.no_cpu else cpu_baseline(geom, p0, prm, n1, n2, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Fishbone-Moncrief torus, Kerr a=0.9375, MKS {n1}x{n2}, PPM+HLL+RK4, "
                               f"flux-CT, 2D Newton con-to-prim, floors, adaptive dt",
                   "n1": n1, "n2": n2, "timeorder": 4, "lim": "PARA", "flux": "HLL", "t_end": t_now,
                   "dt_end": dt_now, "failed_inversions_repaired": nfail,
                   "kernel_variants": {"flux": {H.FLUX_FACE_64x4: "face_64x4", H.FLUX_FACE_64x2: "face_64x2",
                                                H.FLUX_PAIR: "lane_pair", H.FLUX_SHARED: "zone_shared"}[fv],
                                       "boundary": {H.BOUND_KERNELS: "wide_launches",
                                                    H.BOUND_BLOCK: "one_workgroup",
                                                    H.BOUND_KERNELS_SEQCHECK: "wide_launches_seqcheck",
                                                    H.BOUND_FIXONE: "wide_fills_one_workgroup_fixup",
                                                    H.BOUND_FUSED: "one_launch_fills_one_workgroup_fixup",
                                                    H.BOUND_INUPDATE: "first_fill_in_update_one_workgroup_fixup"}[bv]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "grid_scan": scan,
        "adjoint": adjoint,
        "ffde": ffde,
        "accounting": accounting,
        "stages": stages,
        "newton": newton,
        "eager_timed_ms_per_step": 1e3 * el_prof / prof_steps,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
