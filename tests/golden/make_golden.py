"""Generate golden vectors from the COMPILED REFERENCE (run in the build
container, where /root/reference exists; the GPU box only reads the .npz).

For each case the script builds the input state and geometry with
harm_ramt_amd's host generators, runs the reference's own step_ch() through
oracle/_ref/ref_* (see oracle/ref/: one build per problem family and
compile-time option), and stores
  - sha256 of the input state and of the geometry tables (pins the inputs),
  - per step: t, dt, Newton iteration count (nstroke), failure code
    (failed + 2 x number of zones with a nonzero FLAGUTOPRIMFAIL pflag; the
    force-free inversion never writes pflag, so there it counts whatever the
    reference's pflag array holds),
  - the full primitive state p[N1+6][N2+6][8] after the listed steps.
The reference's set_grid() tables for the torus grid are stored as a sha256
plus a sample, pinning the product's geometry generator.

    python tests/golden/make_golden.py [CASE ...]   (default: every case)
"""
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import harm_ramt_amd as H  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")
OUT_DIR = os.path.dirname(os.path.abspath(__file__))

# name: (kind, n1, n2, lim, fluxmethod, timeorder, nsteps, dt0, saved steps)
CASES = {
    "torus_para_hll_rk4": ("ks", 32, 64, H.PARA, H.HLLFLUX, 4, 60, 1e-3, (1, 20, 60)),
    "torus_mc_laxf_rk2": ("ks", 32, 64, H.MC, H.LAXFFLUX, 2, 10, 1e-3, (10,)),
    "torus_vanl_hll_rk3": ("ks", 32, 64, H.VANL, H.HLLFLUX, 3, 5, 1e-3, (5,)),
    "torus_minm_laxf_rk1": ("ks", 32, 64, H.MINM, H.LAXFFLUX, 1, 5, 1e-3, (5,)),
    "torus_csslope_hll_rk4": ("ks", 32, 64, H.CSSLOPE, H.HLLFLUX, 4, 3, 1e-3, (3,)),
    "ot_mc_hll_rk4": ("mink", 32, 32, H.MC, H.HLLFLUX, 4, 40, 1e-4, (1, 40)),
    "ot_para_laxf_rk2": ("mink", 32, 32, H.PARA, H.LAXFFLUX, 2, 10, 1e-4, (10,)),
    "ot_donor_hll_rk1": ("mink", 32, 32, H.DONOR, H.HLLFLUX, 1, 5, 1e-4, (5,)),
    "ot_nlimcent_hll_rk3": ("mink", 32, 32, H.NLIMCENT, H.HLLFLUX, 3, 5, 1e-4, (5,)),
    "ffde_monopole_mc_laxf_rk2": ("ffde", 32, 64, H.MC, H.LAXFFLUX, 2, 200, 1e-3, (1, 100, 200)),
    "ffde_monopole_para_hll_rk4": ("ffde", 32, 64, H.PARA, H.HLLFLUX, 4, 40, 1e-3, (40,)),
    # torus threaded by a weak split-monopole field, CHECKSOLUTION 1: b^2 flags
    # in the magnetised atmosphere, zones failed by fixup_checksolution
    "torus_field_checksolution_rk4": ("ks_mono", 32, 64, H.PARA, H.HLLFLUX, 4, 40, 1e-3, (1, 10, 40)),
    # DOENO 1 (U converted centre -> average -> centre around the update) and
    # DOENO 2 (flux divergence converted average -> centre)
    "torus_eno_u_para_hll_rk4": ("ks", 32, 64, H.PARA, H.HLLFLUX, 4, 30, 1e-3, (1, 30)),
    "torus_eno_du_mc_laxf_rk2": ("ks", 32, 64, H.MC, H.LAXFFLUX, 2, 30, 1e-3, (30,)),
    "ot_eno_u_mc_hll_rk4": ("mink", 32, 32, H.MC, H.HLLFLUX, 4, 30, 1e-4, (1, 30)),
    "ot_eno_du_para_hll_rk3": ("mink", 32, 32, H.PARA, H.HLLFLUX, 3, 30, 1e-4, (30,)),
}
# bounds.fishmon.c's outflow extrapolation EXTRAP of the cases that set it
# (the reference build ksx1: a configured copy of bounds.fishmon.c)
EXTRAP = {"torus_extrap1_para_hll_rk4": 1, "torus_extrap2_mc_hll_rk2": 2}
CASES["torus_extrap1_para_hll_rk4"] = ("ks", 32, 64, H.PARA, H.HLLFLUX, 4, 30, 1e-3, (1, 30))
CASES["torus_extrap2_mc_hll_rk2"] = ("ks", 32, 64, H.MC, H.HLLFLUX, 2, 30, 1e-3, (1, 30))
# DOENO of the cases that set it (the reference is built once per value)
ENO = {"torus_eno_u_para_hll_rk4": 1, "torus_eno_du_mc_laxf_rk2": 2, "ot_eno_u_mc_hll_rk4": 1,
       "ot_eno_du_para_hll_rk3": 2}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def problem(kind, n1, n2):
    """Input state, geometry and the parameter dict shared by reference, oracle, GPU."""
    if kind in ("ks", "ks_mono"):
        g = H.grid_mks(n1, n2)
        geom = H.Geometry(g)
        p0 = H.init_fishmon(geom)
        if kind == "ks_mono":
            p0[..., 5] += H.init_monopole(geom, 1e-2)[..., 5]
        gam, cour, bc, fix, pole = 4.0 / 3.0, 0.9, (H.OUTFLOW, H.OUTFLOW, H.POLARAXIS, H.POLARAXIS), 1, 1
    elif kind == "ffde":
        g = H.grid_mks(n1, n2, a=0.9, Rout=100.0, hslope=0.1)
        geom = H.Geometry(g)
        p0 = H.init_monopole(geom)
        gam, cour, bc, fix, pole = 4.0 / 3.0, 0.9, (H.OUTFLOW, H.OUTFLOW, H.POLARAXIS, H.POLARAXIS), 1, 0
    else:
        g = H.grid_cartesian(n1, n2)
        geom = H.Geometry(g)
        p0 = H.init_orszag_tang(geom)
        gam, cour, bc, fix, pole = 5.0 / 3.0, 0.8, (H.PERIODIC,) * 4, 0, 0
    return g, geom, p0, dict(gam=gam, cour=cour, bc=bc, fixupflux=fix, poledeath=pole)


def bc_of(kind):
    """BCtype[X1UP..X2DN] of a problem family (periodic: the stand-in value,
    which the reference's BCtype tests treat as neither OUTFLOW nor POLARAXIS)."""
    if kind == "mink":
        return (H.PERIODIC,) * 4
    return (H.OUTFLOW, H.OUTFLOW, H.POLARAXIS, H.POLARAXIS)


def run_reference(kind, g, geom, p0, lim, flux, to, nsteps, dt0, work, eno=0, rdump=False, diag=False,
                  gammamax=None, dump=False, binary=None):
    p0.tofile(os.path.join(work, "p0.bin"))
    cfg = dict(prim=os.path.join(work, "p0.bin"), dt=dt0, nsteps=nsteps, out=os.path.join(work, "o.bin"),
               lim=lim, fluxmethod=flux, timeorder=to)
    if gammamax is not None:
        cfg["gammamax"] = gammamax
    for n, b in enumerate(bc_of(kind)):
        cfg[f"bc{n}"] = b
    if rdump:
        cfg["rdump"] = 1
        os.makedirs(os.path.join(work, "dumps"), exist_ok=True)
    if diag:
        cfg.update(diag=1, diag_out=os.path.join(work, "diag.bin"), h_over_r_jet=0.4)
    if dump:
        cfg["dump"] = 1
        os.makedirs(os.path.join(work, "dumps"), exist_ok=True)
    if kind in ("ks", "ks_mono", "ffde"):
        cfg.update(a=g.a, Rout=g.Rout, hslope=g.hslope, R0=g.R0, geom="ref", gam=4.0 / 3.0, cour=0.9,
                   tables_out=os.path.join(work, "tables.bin"))
    else:
        names = ["gcov", "gcon", "gdet", "conn"]
        for nm, a in zip(names, geom.tables()[:4]):
            a.tofile(os.path.join(work, f"mk_{nm}.bin"))
        cfg.update(geom=os.path.join(work, "mk_gcov.bin"), geom_gcon=os.path.join(work, "mk_gcon.bin"),
                   geom_gdet=os.path.join(work, "mk_gdet.bin"), geom_conn=os.path.join(work, "mk_conn.bin"),
                   Rin=1.0, R0=0.0, Rout=float(np.e), startx1=0.0, startx2=0.0, dx1=g.dx1, dx2=g.dx2, hslope=1.0,
                   gam=5.0 / 3.0, cour=0.8)
    path = os.path.join(work, "run.cfg")
    with open(path, "w") as f:
        for k, v in cfg.items():
            f.write(f"{k} {v!r}\n" if isinstance(v, float) else f"{k} {v}\n")
    binary = binary or {"ks_mono": "ref_kscs"}.get(kind, "ref_" + kind) + (f"eno{eno}" if eno else "")
    subprocess.run([os.path.join(REF_DIR, binary), path], cwd=work, check=True, capture_output=True)
    d = np.fromfile(os.path.join(work, "o.bin"), dtype=np.float64)[4:]
    nz = (g.n1 + 6) * (g.n2 + 6) * 8
    recs, states = [], []
    for s in range(nsteps):
        r = d[s * (4 + nz):(s + 1) * (4 + nz)]
        recs.append(r[:4])
        states.append(r[4:].reshape(g.n1 + 6, g.n2 + 6, 8))
    return np.array(recs), states


# restart files written by the reference's restart_write() after the steps
RDUMP_CASES = {"rdump_ot_mc_hll_rk4": ("mink", 32, 32, H.MC, H.HLLFLUX, 4, 3, 1e-4)}


def make_rdump(name, work):
    kind, n1, n2, lim, flux, to, nsteps, dt0 = RDUMP_CASES[name]
    g, geom, p0, extra = problem(kind, n1, n2)
    recs, states = run_reference(kind, g, geom, p0, lim, flux, to, nsteps, dt0, work, rdump=True)
    text = open(os.path.join(work, "dumps", "rdump-0"), "rb").read()
    binary = open(os.path.join(work, "dumps", "rdump-1.bin"), "rb").read()
    np.savez_compressed(os.path.join(OUT_DIR, f"{name}.npz"), kind=kind, n1=n1, n2=n2, lim=lim, fluxmethod=flux,
                        timeorder=to, nsteps=nsteps, dt0=dt0, input_sha=sha(p0), t=recs[:, 0], dt=recs[:, 1],
                        p=states[-1], rdump_text=np.frombuffer(text, dtype=np.uint8),
                        rdump_binary=np.frombuffer(binary, dtype=np.uint8))
    print(name, "text", len(text), "bytes, binary", len(binary), "bytes")


# conservation accounting of the reference's steps (diag_flux, diag_source,
# diag_fixup accumulators of every energy region, written after the steps)
# and its dump-time reductions divbmaxavg() and constotal() of the final state
DIAG_CASES = {
    "diag_torus_para_hll_rk4": ("ks", 32, 64, H.PARA, H.HLLFLUX, 4, 20, 1e-3),
    "diag_ot_mc_hll_rk2": ("mink", 32, 32, H.MC, H.HLLFLUX, 2, 10, 1e-4),
}
# GAMMAMAX lowered (initbase.c:218) so that limit_gamma acts on face states
# and zones: its accounting (COUNTLIMITGAMMAACT) is exercised
DIAG_GAMMAMAX = {"diag_ot_gammamax_mc_hll_rk2": ("mink", 32, 32, H.MC, H.HLLFLUX, 2, 10, 1e-4, 1.1)}
# the other equation and conversion modes with accounting: force-free
# (EOMFFDE), CHECKSOLUTION 1 (b^2 flags and checksolution repairs), DOENO 1
# and 2 (the source and flux terms of the converted update)
DIAG_CASES.update({
    "diag_ffde_monopole_mc_laxf_rk2": ("ffde", 32, 64, H.MC, H.LAXFFLUX, 2, 20, 1e-3),
    "diag_torus_field_checksolution_rk4": ("ks_mono", 32, 64, H.PARA, H.HLLFLUX, 4, 20, 1e-3),
    "diag_torus_eno_u_para_hll_rk4": ("ks", 32, 64, H.PARA, H.HLLFLUX, 4, 12, 1e-3),
    "diag_ot_eno_du_mc_hll_rk2": ("mink", 32, 32, H.MC, H.HLLFLUX, 2, 10, 1e-4),
})
DIAG_ENO = {"diag_torus_eno_u_para_hll_rk4": 1, "diag_ot_eno_du_mc_hll_rk2": 2}
# force-free: fixup1zone() books COUNTFLOORACT from locals it never sets under
# EOMFFDE (fixup.c:283-313); the ffdezi build zero-initialises them
DIAG_BINARY = {"diag_ffde_monopole_mc_laxf_rk2": "ref_ffdezi"}
NREG, NFL, NSRC, NFLOOR = 3, 7, 7, 9
DIAG_LAYOUT = [("enerpos", (NREG, 4)), ("doflux", (NREG, 4)), ("pdot", (NREG, 4, 8)),
               ("pdotterms", (NREG, 4, NFL, 8)), ("pcum", (NREG, 4, 8)), ("fladd", (NREG, 8)),
               ("fladdterms", (NREG, NFLOOR, 8)), ("sourceadd", (NREG, 8)), ("sourceaddterms", (NREG, NSRC, 8)),
               ("pdottermsjet2", (4, NFL, 8)), ("frdot", ("n1", 8)), ("lumvsr", ("n1",)), ("divbmax", ()),
               ("divbavg", ()), ("constotal", (NREG, 8)), ("dVF", ())]


def read_diag(path, n1):
    d = np.fromfile(path, dtype=np.float64)
    out, o = {}, 0
    for name, shape in DIAG_LAYOUT:
        shape = tuple(n1 if s == "n1" else s for s in shape)
        n = int(np.prod(shape)) if shape else 1
        out[name] = d[o:o + n].reshape(shape) if shape else d[o]
        o += n
    assert o == d.size, (o, d.size)
    out["enerpos"] = out["enerpos"].astype(np.int32)
    out["doflux"] = out["doflux"].astype(np.int32)
    return out


def make_diag(name, work):
    if name in DIAG_GAMMAMAX:
        kind, n1, n2, lim, flux, to, nsteps, dt0, gmax = DIAG_GAMMAMAX[name]
    else:
        (kind, n1, n2, lim, flux, to, nsteps, dt0), gmax = DIAG_CASES[name], None
    g, geom, p0, extra = problem(kind, n1, n2)
    eno = DIAG_ENO.get(name, 0)
    recs, states = run_reference(kind, g, geom, p0, lim, flux, to, nsteps, dt0, work, eno=eno, diag=True,
                                 gammamax=gmax, binary=DIAG_BINARY.get(name))
    dg = read_diag(os.path.join(work, "diag.bin"), n1)
    more = {} if gmax is None else {"gammamax": gmax}
    if eno:
        more["eno"] = eno
    if name in DIAG_BINARY:
        more["zero_init_locals"] = 1
    np.savez_compressed(os.path.join(OUT_DIR, f"{name}.npz"), kind=kind, n1=n1, n2=n2, lim=lim, fluxmethod=flux,
                        timeorder=to, nsteps=nsteps, dt0=dt0, input_sha=sha(p0), t=recs[:, 0], dt=recs[:, 1],
                        fail=recs[:, 3], p=states[-1], h_over_r_jet=0.4, **more, **dg)
    print(name, "fladd", dg["fladd"][0], "divb", dg["divbmax"], dg["divbavg"])


# the reference's 56-column dump (dump.c) of the final state, text and
# binary, with the current and Faraday columns its steps computed
DUMP_CASES = {"dump_torus_para_hll_rk4": ("ks", 32, 64, H.PARA, H.HLLFLUX, 4, 5, 1e-3)}


def make_dump(name, work):
    kind, n1, n2, lim, flux, to, nsteps, dt0 = DUMP_CASES[name]
    g, geom, p0, extra = problem(kind, n1, n2)
    recs, states = run_reference(kind, g, geom, p0, lim, flux, to, nsteps, dt0, work, dump=True)
    d = os.path.join(work, "dumps")
    text = open(os.path.join(d, "dump0000"), "rb").read()
    binary = open(os.path.join(d, "dump0001.bin"), "rb").read()
    np.savez_compressed(os.path.join(OUT_DIR, f"{name}.npz"), kind=kind, n1=n1, n2=n2, lim=lim, fluxmethod=flux,
                        timeorder=to, nsteps=nsteps, dt0=dt0, input_sha=sha(p0), t=recs[:, 0], dt=recs[:, 1],
                        p=states[-1], dump_text=np.frombuffer(text, dtype=np.uint8),
                        dump_binary=np.frombuffer(binary, dtype=np.uint8))
    print(name, "text", len(text), "bytes, binary", len(binary), "bytes")


# the reference's default problem as shipped (init.c PROBLEMTYPE 7, bounds.c,
# CYLMINKMETRIC, EOMFFDE) from its own init(): the initial state, panalytic
# and omegafanalytic it sets up, its set_grid() tables (pinned by sha256
# against harm_grid_cylmink), and its steps.  t0 > 15 (TSTATIC) takes
# bound_disk's outflow branch for B1, B3.
DISK_CASES = {"disk_default_mc_laxf_rk2": (32, 64, 40, 1e-3, 0.0, (1, 10, 40)),
              "disk_late_mc_laxf_rk2": (32, 64, 20, 1e-2, 20.0, (1, 20))}
# the same problem with bounds.c's polar axis on both x2 sides (u^2, b^2
# antisymmetric across the pole, POLEDEATH 1) from the first step on
DISK_BC = {"disk_polar_mc_laxf_rk2": {"bc2": H.POLARAXIS, "bc3": H.POLARAXIS}}
DISK_CASES["disk_polar_mc_laxf_rk2"] = (32, 64, 30, 1e-3, 0.0, (1, 30))


def make_disk(name, work):
    n1, n2, nsteps, dt0, t0, saved = DISK_CASES[name]
    ref = "/root/reference/nu.75_m.5.txt"
    link = os.path.join(work, "nu.75_m.5.txt")
    if not os.path.exists(link):
        os.symlink(ref, link)
    cfg = dict(tables_out=os.path.join(work, "dtables.bin"), init_out=os.path.join(work, "dinit.bin"),
               params_out=os.path.join(work, "dparams.txt"), dt=dt0, t0=t0, nsteps=nsteps,
               out=os.path.join(work, "do.bin"), **DISK_BC.get(name, {}))
    path = os.path.join(work, "disk.cfg")
    with open(path, "w") as f:
        for k, v in cfg.items():
            f.write(f"{k} {v!r}\n" if isinstance(v, float) else f"{k} {v}\n")
    subprocess.run([os.path.join(REF_DIR, "ref_disk"), path], cwd=work, check=True, capture_output=True)
    m1, m2 = n1 + 6, n2 + 6
    init = np.fromfile(os.path.join(work, "dinit.bin")).reshape(3, m1, m2, 8)
    prm = dict(l.split() for l in open(os.path.join(work, "dparams.txt")))
    tb = np.fromfile(os.path.join(work, "dtables.bin"))
    nz = m1 * m2
    tables = tb[:nz * (64 + 64 + 4 + 64)]
    g = H.grid_cylmink(n1, n2, float(prm["Rin"]), float(prm["Rout"]), float(prm["Zin"]), float(prm["Zout"]))
    mine = np.concatenate([a.ravel() for a in H.Geometry(g).tables()[:4]])
    assert np.array_equal(mine.view(np.int64), tables.view(np.int64)), "harm_grid_cylmink differs from set_grid()"
    d = np.fromfile(os.path.join(work, "do.bin"))[4:]
    recs = d.reshape(nsteps, 4 + nz * 8)
    out = dict(kind="disk", n1=n1, n2=n2, lim=int(prm["lim"]), fluxmethod=int(prm["fluxmethod"]),
               timeorder=int(prm["timeorder"]), fluxb=int(prm["FLUXB"]), cour=float(prm["cour"]),
               gam=float(prm["gam"]), bc=np.array([int(prm[f"bc{k}"]) for k in range(4)]),
               Rin=float(prm["Rin"]), Rout=float(prm["Rout"]), Zin=float(prm["Zin"]), Zout=float(prm["Zout"]),
               gammamax=float(prm["GAMMAMAX"]), nsteps=nsteps, dt0=dt0, t0=t0,
               p0=init[0], panalytic=init[1], omegaf=init[2][..., 0],
               tables_sha=hashlib.sha256(tables.tobytes()).hexdigest(),
               t=recs[:, 0], dt=recs[:, 1], nstroke=recs[:, 2], failcode=recs[:, 3], saved_steps=np.array(saved))
    for s in saved:
        out[f"p_{s}"] = recs[s - 1, 4:].reshape(m1, m2, 8)
    np.savez_compressed(os.path.join(OUT_DIR, f"{name}.npz"), **out)
    print(name, "t", recs[-1, 0], "dt", recs[-1, 1])


def main():
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle", "ref"), "-s"])
    with tempfile.TemporaryDirectory() as work:
        rd = [a for a in sys.argv[1:] if a in RDUMP_CASES] or ([] if sys.argv[1:] else list(RDUMP_CASES))
        for name in rd:
            make_rdump(name, work)
        du = [a for a in sys.argv[1:] if a in DUMP_CASES] or ([] if sys.argv[1:] else list(DUMP_CASES))
        for name in du:
            make_dump(name, work)
        dk = [a for a in sys.argv[1:] if a in DISK_CASES] or ([] if sys.argv[1:] else list(DISK_CASES))
        for name in dk:
            make_disk(name, work)
        alld = {**DIAG_CASES, **DIAG_GAMMAMAX}
        dc = [a for a in sys.argv[1:] if a in alld] or ([] if sys.argv[1:] else list(alld))
        for name in dc:
            make_diag(name, work)
        names = [a for a in sys.argv[1:] if a not in RDUMP_CASES and a not in alld and a not in DUMP_CASES
                 and a not in DISK_CASES] \
            if sys.argv[1:] else list(CASES)
        for name in names:
            kind, n1, n2, lim, flux, to, nsteps, dt0, saved = CASES[name]
            g, geom, p0, extra = problem(kind, n1, n2)
            eno = ENO.get(name, 0)
            extrap = EXTRAP.get(name, 0)
            recs, states = run_reference(kind, g, geom, p0, lim, flux, to, nsteps, dt0, work, eno,
                                         binary=f"ref_ksx{extrap}" if extrap else None)
            out = dict(kind=kind, n1=n1, n2=n2, lim=lim, fluxmethod=flux, timeorder=to, nsteps=nsteps, dt0=dt0,
                       eno=eno, extrap=extrap,
                       input_sha=sha(p0), geom_sha=sha(np.concatenate([a.ravel() for a in geom.tables()])),
                       t=recs[:, 0], dt=recs[:, 1], nstroke=recs[:, 2], failcode=recs[:, 3],
                       saved_steps=np.array(saved))
            for s in saved:
                out[f"p_{s}"] = states[s - 1]
            np.savez_compressed(os.path.join(OUT_DIR, f"{name}.npz"), **out)
            print(name, "nstroke", int(recs[:, 2].sum()), "failures", int((recs[:, 3] // 2).sum()))
        if sys.argv[1:]:
            return
        # reference set_grid() tables for the 32x64 torus grid
        kind, n1, n2 = "ks", 32, 64
        g, geom, p0, extra = problem(kind, n1, n2)
        tb = np.fromfile(os.path.join(work, "tables.bin"), dtype=np.float64)
        nz = (n1 + 6) * (n2 + 6)
        sizes = [nz * 64, nz * 64, nz * 4, nz * 64]
        parts, o = [], 0
        for sz in sizes:
            parts.append(tb[o:o + sz])
            o += sz
        rin, sx1, sx2, dx1, dx2 = tb[o:o + 5]
        allt = np.concatenate(parts)
        np.savez_compressed(os.path.join(OUT_DIR, "ks_geometry_32x64.npz"), sha=hashlib.sha256(allt.tobytes()).hexdigest(),
                            sample_idx=np.arange(0, allt.size, 997), sample=allt[::997],
                            Rin=rin, startx=np.array([sx1, sx2]), dx=np.array([dx1, dx2]),
                            sizes=np.array(sizes))
        print("geometry tables sha", hashlib.sha256(allt.tobytes()).hexdigest()[:16])


if __name__ == "__main__":
    main()
