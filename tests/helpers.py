"""Problem set-up shared by the oracle tests and the GPU parity tests."""
import glob
import hashlib
import os

import numpy as np

import harm_ramt_amd as H

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def problem(kind, n1, n2, lim=H.PARA, fluxmethod=H.HLLFLUX, timeorder=4, seed=7, eno=0):
    """(grid, geometry, initial state, harm_params) for a torus ('ks'), a
    torus threaded by a weak split-monopole field with CHECKSOLUTION on
    ('ks_mono'), an Orszag-Tang ('mink') or a force-free split-monopole
    ('ffde') problem."""
    if kind in ("ks", "ks_mono"):
        g = H.grid_mks(n1, n2)
        geom = H.Geometry(g)
        p0 = H.init_fishmon(geom, seed=seed)
        if kind == "ks_mono":   # torus threaded by a weak split-monopole field, CHECKSOLUTION on
            p0[..., 5] += H.init_monopole(geom, 1e-2)[..., 5]
        prm = H.make_params(n1, n2, g.dx1, g.dx2, lim=lim, fluxmethod=fluxmethod, timeorder=timeorder,
                            checksolution=1 if kind == "ks_mono" else 0, eno=eno)
    elif kind == "ffde":
        g = H.grid_mks(n1, n2, a=0.9, Rout=100.0, hslope=0.1)
        geom = H.Geometry(g)
        p0 = H.init_monopole(geom)
        prm = H.make_params(n1, n2, g.dx1, g.dx2, lim=lim, fluxmethod=fluxmethod, timeorder=timeorder,
                            eomtype=H.EOMFFDE, poledeath=0, eno=eno)
    else:
        g = H.grid_cartesian(n1, n2)
        geom = H.Geometry(g)
        p0 = H.init_orszag_tang(geom)
        prm = H.make_params(n1, n2, g.dx1, g.dx2, lim=lim, fluxmethod=fluxmethod, timeorder=timeorder,
                            gam=5.0 / 3.0, cour=0.8, bc=(H.PERIODIC,) * 4, fixupflux=0, poledeath=0, eno=eno)
    return g, geom, p0, prm


def golden_cases():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not os.path.basename(p).startswith(("ks_geometry", "rdump_", "diag_", "dump_", "disk_")))


def disk_cases():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "disk_*.npz")))


def disk_problem(gd):
    """The reference's default problem from a disk golden: the cylindrical-
    Minkowski grid, the state, panalytic and omegafanalytic its init() set
    up (omegaf expanded to [N1M][N2M][NPR], component 0), and the parameters
    it ran with (bounds.c: FIXED sides from panalytic, bound_disk, whose
    first call the reference's init() already made)."""
    n1, n2 = int(gd["n1"]), int(gd["n2"])
    g = H.grid_cylmink(n1, n2, float(gd["Rin"]), float(gd["Rout"]), float(gd["Zin"]), float(gd["Zout"]))
    geom = H.Geometry(g)
    om = np.zeros_like(gd["panalytic"])
    om[..., 0] = gd["omegaf"]
    prm = H.make_params(n1, n2, g.dx1, g.dx2, lim=int(gd["lim"]), fluxmethod=int(gd["fluxmethod"]),
                        timeorder=int(gd["timeorder"]), fluxb=int(gd["fluxb"]), cour=float(gd["cour"]),
                        gam=float(gd["gam"]), gammamax=float(gd["gammamax"]), eomtype=H.EOMFFDE,
                        bc=tuple(int(b) for b in gd["bc"]),
                        poledeath=1 if H.POLARAXIS in tuple(int(b) for b in gd["bc"][2:]) else 0,
                        boundsfile=H.BOUNDS_DISK, extrap=1,
                        disk_first=0)
    return g, geom, np.array(gd["p0"]), np.array(gd["panalytic"]), om, prm


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def golden_problem(gd):
    """problem() with the settings a golden vector file records."""
    eno = int(gd["eno"]) if "eno" in gd.files else 0
    g, geom, p0, prm = problem(str(gd["kind"]), int(gd["n1"]), int(gd["n2"]), lim=int(gd["lim"]),
                               fluxmethod=int(gd["fluxmethod"]), timeorder=int(gd["timeorder"]), eno=eno)
    if "extrap" in gd.files:
        prm.extrap = int(gd["extrap"])
    if "gammamax" in gd.files:  # GAMMAMAX and GAMMAFAIL = 100 GAMMAMAX (initbase.c:218-219)
        prm.gammamax = float(gd["gammamax"])
        prm.gammafail = 100.0 * prm.gammamax
    return g, geom, p0, prm


def same(a, b):
    """IEEE equality elementwise (+0 == -0), NaN never equal."""
    return np.array_equal(a, b)
