"""Full-size check and CPU calibration against the compiled reference (run in
the build container, where oracle/_ref/ref_ks256 is built from the
reference's sources by oracle/ref/Makefile).

The benchmark's workload -- the 256x256 Fishbone-Moncrief torus of bench.py
(PPM, HLL, RK4, dt0 = 1e-5) -- is stepped by the reference's own step_ch()
and by the oracle.  Printed as one JSON line:
  - whether every step's state, t and dt agree bit for bit,
  - the reference's zone-updates/s on one core (time of n2 steps minus n1
    steps, so that set-up and the harness's per-step output cancel),
  - the oracle's on one thread and on all threads of this machine.

    python oracle/ref/calibrate.py [--n1 4] [--n2 12]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import harm_ramt_amd as H  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_ks256")
N = 256


def setup():
    g = H.grid_mks(N, N, a=0.9375, Rout=40.0, hslope=0.3)
    geom = H.Geometry(g)
    p0 = H.init_fishmon(geom, gam=4.0 / 3.0, rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=7)
    prm = H.make_params(N, N, g.dx1, g.dx2, lim=H.PARA, fluxmethod=H.HLLFLUX, timeorder=4, cour=0.9)
    return g, geom, p0, prm


def run_ref(g, p0, nsteps, work):
    p0.tofile(os.path.join(work, "p0.bin"))
    cfg = dict(prim=os.path.join(work, "p0.bin"), dt=1e-5, tf=2000.0, nsteps=nsteps, out=os.path.join(work, "o.bin"),
               lim=H.PARA, fluxmethod=H.HLLFLUX, timeorder=4, a=g.a, Rout=g.Rout, hslope=g.hslope, R0=g.R0,
               geom="ref", gam=4.0 / 3.0, cour=0.9)
    path = os.path.join(work, "run.cfg")
    with open(path, "w") as f:
        for k, v in cfg.items():
            f.write(f"{k} {v!r}\n" if isinstance(v, float) else f"{k} {v}\n")
    t0 = time.perf_counter()
    subprocess.run([REF, path], cwd=work, check=True, capture_output=True)
    el = time.perf_counter() - t0
    d = np.fromfile(os.path.join(work, "o.bin"), dtype=np.float64)[4:]
    nz = (N + 6) * (N + 6) * 8
    recs = [d[s * (4 + nz):s * (4 + nz) + 4] for s in range(nsteps)]
    states = [d[s * (4 + nz) + 4:(s + 1) * (4 + nz)].reshape(N + 6, N + 6, 8) for s in range(nsteps)]
    return el, recs, states


def run_oracle(geom, p0, prm, nsteps, nthreads, keep=False):
    o = Oracle(H.params_dict(prm), nthreads=nthreads)
    o.set_geometry(*geom.tables())
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, 1e-5, 2000.0)
    out = []
    t0 = time.perf_counter()
    for _ in range(nsteps):
        o.step()
        if keep:
            out.append((o.get_time(), o.get_array(0).copy()))
    return time.perf_counter() - t0, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n1", type=int, default=4)
    ap.add_argument("--n2", type=int, default=12)
    args = ap.parse_args()
    g, geom, p0, prm = setup()
    with tempfile.TemporaryDirectory() as work:
        e1, _, _ = run_ref(g, p0, args.n1, work)
        e2, recs, states = run_ref(g, p0, args.n2, work)
    ref_rate = (args.n2 - args.n1) * N * N / (e2 - e1)
    _, mine = run_oracle(geom, p0, prm, args.n2, 1, keep=True)
    same = all(np.array_equal(s, m[1]) and r[0] == m[0][0] and r[1] == m[0][1]
               for r, s, m in zip(recs, states, mine))
    t1a, _ = run_oracle(geom, p0, prm, args.n1, 1)
    t1b, _ = run_oracle(geom, p0, prm, args.n2, 1)
    ncpu = os.cpu_count()
    tna, _ = run_oracle(geom, p0, prm, args.n1, ncpu)
    tnb, _ = run_oracle(geom, p0, prm, args.n2, ncpu)
    print(json.dumps({
        "workload": "FM torus 256x256, PPM+HLL+RK4, dt0=1e-5 (bench.py)", "steps": args.n2,
        "bitwise_equal_every_step": bool(same),
        "reference_zone_updates_per_s_1core": ref_rate,
        "oracle_zone_updates_per_s_1thread": (args.n2 - args.n1) * N * N / (t1b - t1a),
        f"oracle_zone_updates_per_s_{ncpu}threads": (args.n2 - args.n1) * N * N / (tnb - tna),
    }))


if __name__ == "__main__":
    main()
