"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg -- never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
NG, NPR = 3, 8
P, PK0, PK1, F1, F2, ULAST, UNEW, EMF = 0, 1, 2, 3, 4, 5, 6, 7


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_int), ("n2", ctypes.c_int),
        ("eomtype", ctypes.c_int), ("lim", ctypes.c_int), ("fluxmethod", ctypes.c_int), ("fluxb", ctypes.c_int),
        ("timeorder", ctypes.c_int), ("bc", ctypes.c_int * 4), ("fixupflux", ctypes.c_int),
        ("poledeath", ctypes.c_int), ("checksolution", ctypes.c_int), ("eno", ctypes.c_int),
        ("gam", ctypes.c_double), ("cour", ctypes.c_double), ("safe", ctypes.c_double), ("defcon", ctypes.c_double),
        ("gammamax", ctypes.c_double), ("gammafail", ctypes.c_double), ("rhominlimit", ctypes.c_double),
        ("uuminlimit", ctypes.c_double), ("dx1", ctypes.c_double), ("dx2", ctypes.c_double),
        ("boundsfile", ctypes.c_int), ("extrap", ctypes.c_int), ("disk_first", ctypes.c_int),
        ("rescaletype", ctypes.c_int), ("nthreads", ctypes.c_int),
    ]


_lib = None


NREG, NUMFLUXTERMS, NUMSOURCES, NUMFAILFLOORFLAGS = 3, 7, 7, 9


class OracleAcct(ctypes.Structure):
    """or_acct: the regions and accumulators of defs.h:218-239"""
    _fields_ = [("enerpos", (ctypes.c_int * 4) * NREG), ("doflux", (ctypes.c_int * 4) * NREG),
                ("dVF", ctypes.c_double),
                ("pdot", ((ctypes.c_double * 8) * 4) * NREG),
                ("pdotterms", (((ctypes.c_double * 8) * NUMFLUXTERMS) * 4) * NREG),
                ("pcum", ((ctypes.c_double * 8) * 4) * NREG),
                ("fladd", (ctypes.c_double * 8) * NREG),
                ("fladdterms", ((ctypes.c_double * 8) * NUMFAILFLOORFLAGS) * NREG),
                ("sourceadd", (ctypes.c_double * 8) * NREG),
                ("sourceaddterms", ((ctypes.c_double * 8) * NUMSOURCES) * NREG),
                ("pdottermsjet2", ((ctypes.c_double * 8) * NUMFLUXTERMS) * 4)]


def build():
    subprocess.check_call(["make", "-C", HERE, "-s"])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    sig = {
        "or_create": ([ctypes.POINTER(OracleParams)], vp),
        "or_destroy": ([vp], None),
        "or_set_geometry": ([vp, dp, dp, dp, dp, dp], None),
        "or_array": ([vp, ctypes.c_int], dp),
        "or_pflag": ([vp], ip),
        "or_set_time": ([vp, ctypes.c_double, ctypes.c_double, ctypes.c_double], None),
        "or_get_time": ([vp, dp, dp], None),
        "or_nstroke": ([vp], ctypes.c_long),
        "or_failed": ([vp], ctypes.c_int),
        "or_fluxcalc": ([vp, ctypes.c_int, ctypes.c_int, dp], ctypes.c_int),
        "or_fix_flux": ([vp], None),
        "or_flux_ct": ([vp], ctypes.c_int),
        "or_advance": ([vp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, dp],
                       ctypes.c_int),
        "or_bound_prim": ([vp, ctypes.c_int], ctypes.c_int),
        "or_post_fixup": ([vp, ctypes.c_int], ctypes.c_int),
        "or_step": ([vp], ctypes.c_int),
        "or_utoprim_2d_final": ([vp, ctypes.c_int, ctypes.c_int, dp, dp, ip], ctypes.c_int),
        "or_utoprimgen": ([vp, ctypes.c_int, ctypes.c_int, dp, dp, ip], ctypes.c_int),
        "or_primtoU": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp], ctypes.c_int),
        "or_primtoflux": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp], ctypes.c_int),
        "or_vchar": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, dp], ctypes.c_int),
        "or_source": ([vp, ctypes.c_int, ctypes.c_int, dp, dp], ctypes.c_int),
        "or_slope_lim": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                          ctypes.c_double, ctypes.c_double, dp, dp, dp], None),
        "or_eno_line": ([ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ip, ip], ctypes.c_double),
        "or_eno_weight": ([ctypes.c_int] * 4, ctypes.c_double),
        "or_acct_setup": ([vp, ip, ip, ctypes.c_double], None),
        "or_acct_off": ([vp], None),
        "or_acct_ffde_pr0_zero": ([vp, ctypes.c_int], None),
        "or_acct_get": ([vp, ctypes.POINTER(OracleAcct), dp, dp], None),
        "or_diag_flux": ([vp, ctypes.c_double], None),
        "or_divbmaxavg": ([vp, ctypes.c_int, dp, dp, ip, ip], None),
        "or_constotal": ([vp, ctypes.c_int, ctypes.c_int, dp], None),
        "or_current_enable": ([vp, ctypes.c_int], None),
        "or_current": ([vp], None),
        "or_current_get": ([vp, dp, dp], None),
        "or_dump_content": ([vp, dp], None),
        "or_set_analytic": ([vp, dp, dp], None),
        "or_set_radius": ([vp, dp], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _d(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """CPU restatement of the step on one grid (reference layout arrays)."""

    def __init__(self, params: dict, nthreads=0):
        self.lib = load()
        prm = OracleParams()
        for f, _ in OracleParams._fields_:
            if f == "nthreads":
                continue
            v = params[f]
            if f == "bc":
                for k in range(4):
                    prm.bc[k] = v[k]
            else:
                setattr(prm, f, v)
        prm.nthreads = nthreads
        self.prm = prm
        self.h = self.lib.or_create(ctypes.byref(prm))
        self.shape = (prm.n1 + 2 * NG, prm.n2 + 2 * NG, NPR)
        self.nz = self.shape[0] * self.shape[1]

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_destroy(self.h)
            self.h = None

    def set_analytic(self, panalytic, omegaf=None):
        pa = np.ascontiguousarray(panalytic, dtype=np.float64)
        om = None if omegaf is None else np.ascontiguousarray(omegaf, dtype=np.float64)
        self.lib.or_set_analytic(self.h, _d(pa), _d(om) if om is not None else None)

    def set_radius(self, r):
        """EXTRAP 2: r of every zone centre [N1M][N2M]"""
        self.lib.or_set_radius(self.h, _d(np.ascontiguousarray(r, dtype=np.float64)))

    def set_geometry(self, gcov, gcon, gdet, conn, floors):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (gcov, gcon, gdet, conn, floors)]
        self.lib.or_set_geometry(self.h, *[_d(a) for a in arrs])

    def view(self, sel):
        ptr = self.lib.or_array(self.h, sel)
        n = self.nz if sel == EMF else self.nz * NPR
        a = np.ctypeslib.as_array(ptr, shape=(n,))
        return a.reshape(self.shape[:2]) if sel == EMF else a.reshape(self.shape)

    def set_array(self, sel, a):
        self.view(sel)[...] = a

    def get_array(self, sel):
        return self.view(sel).copy()

    def pflag(self):
        ptr = self.lib.or_pflag(self.h)
        return np.ctypeslib.as_array(ptr, shape=(self.nz * 5,)).reshape(self.shape[:2] + (5,)).copy()

    def set_time(self, t, dt, tf):
        self.lib.or_set_time(self.h, t, dt, tf)

    def get_time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        self.lib.or_get_time(self.h, ctypes.byref(t), ctypes.byref(dt))
        return t.value, dt.value

    def fluxcalc(self, sel, dir):
        ndt = ctypes.c_double()
        self.lib.or_fluxcalc(self.h, sel, dir, ctypes.byref(ndt))
        return ndt.value

    def fix_flux(self):
        self.lib.or_fix_flux(self.h)

    def flux_ct(self):
        self.lib.or_flux_ct(self.h)

    def advance(self, pi, pb, CUf, pf, Cunew, timeorder, numtimeorders):
        cu = np.ascontiguousarray(CUf, dtype=np.float64)
        cn = np.ascontiguousarray(Cunew, dtype=np.float64)
        ndt = ctypes.c_double()
        self.lib.or_advance(self.h, pi, pb, _d(cu), pf, _d(cn), timeorder, numtimeorders, ctypes.byref(ndt))
        return ndt.value

    def bound_prim(self, sel=P):
        return self.lib.or_bound_prim(self.h, sel)

    def step(self):
        return self.lib.or_step(self.h)

    def nstroke(self):
        return self.lib.or_nstroke(self.h)

    def failed(self):
        return self.lib.or_failed(self.h)

    # single-zone entry points
    def utoprimgen(self, i, j, U, pr):
        U = np.ascontiguousarray(U, dtype=np.float64)
        out = np.ascontiguousarray(pr, dtype=np.float64).copy()
        fl = ctypes.c_int()
        ns = self.lib.or_utoprimgen(self.h, i, j, _d(U), _d(out), ctypes.byref(fl))
        return out, fl.value, ns

    def utoprim_2d_final(self, i, j, Ug, pr):
        Ug = np.ascontiguousarray(Ug, dtype=np.float64)
        out = np.ascontiguousarray(pr, dtype=np.float64).copy()
        fl = ctypes.c_int()
        ns = self.lib.or_utoprim_2d_final(self.h, i, j, _d(Ug), _d(out), ctypes.byref(fl))
        return out, fl.value, ns

    def primtoU(self, i, j, loc, pr):
        pr = np.ascontiguousarray(pr, dtype=np.float64)
        U = np.zeros(NPR)
        self.lib.or_primtoU(self.h, i, j, loc, _d(pr), _d(U))
        return U

    def primtoflux(self, i, j, loc, pr, dir):
        pr = np.ascontiguousarray(pr, dtype=np.float64)
        F = np.zeros(NPR)
        self.lib.or_primtoflux(self.h, i, j, loc, _d(pr), dir, _d(F))
        return F

    def vchar(self, i, j, loc, pr, dir):
        pr = np.ascontiguousarray(pr, dtype=np.float64)
        vmax, vmin = ctypes.c_double(), ctypes.c_double()
        fail = self.lib.or_vchar(self.h, i, j, loc, _d(pr), dir, ctypes.byref(vmax), ctypes.byref(vmin))
        return vmax.value, vmin.value, fail

    def source(self, i, j, pr):
        pr = np.ascontiguousarray(pr, dtype=np.float64)
        dU = np.zeros(NPR)
        self.lib.or_source(self.h, i, j, _d(pr), _d(dU))
        return dU

    # conservation accounting (diag.c, fixup.c, dump_ener.c)
    def acct_setup(self, enerpos, doflux, dx3=2 * np.pi):
        ep = np.ascontiguousarray(enerpos, dtype=np.int32)
        df = np.ascontiguousarray(doflux, dtype=np.int32)
        assert ep.shape == (NREG, 4) and df.shape == (NREG, 4)
        ip = ctypes.POINTER(ctypes.c_int)
        self.lib.or_acct_setup(self.h, ep.ctypes.data_as(ip), df.ctypes.data_as(ip), dx3)

    def acct_off(self):
        self.lib.or_acct_off(self.h)

    def acct_ffde_pr0_zero(self, on=True):
        """force-free fixup1zone events measured from pr0 = 0, as the
        reference books them with zero-initialised locals"""
        self.lib.or_acct_ffde_pr0_zero(self.h, int(on))

    def acct(self):
        """the accumulators as a dict of arrays (plus frdot[n1][8], lumvsr[n1])"""
        a = OracleAcct()
        frdot = np.zeros((self.prm.n1, NPR))
        lum = np.zeros(self.prm.n1)
        self.lib.or_acct_get(self.h, ctypes.byref(a), _d(frdot), _d(lum))
        out = {name: np.ctypeslib.as_array(getattr(a, name)).copy() for name, _ in OracleAcct._fields_
               if name != "dVF"}
        out["dVF"] = a.dVF
        out["frdot"], out["lumvsr"] = frdot, lum
        return out

    def diag_flux(self, Dt):
        self.lib.or_diag_flux(self.h, Dt)

    def divbmaxavg(self, sel=P):
        mx, av = ctypes.c_double(), ctypes.c_double()
        im, jm = ctypes.c_int(), ctypes.c_int()
        self.lib.or_divbmaxavg(self.h, sel, ctypes.byref(mx), ctypes.byref(av), ctypes.byref(im), ctypes.byref(jm))
        return mx.value, av.value, im.value, jm.value

    def current_enable(self, on=True):
        self.lib.or_current_enable(self.h, int(on))

    def current_step(self):
        """the step's current computation on the current state and dt"""
        self.lib.or_current(self.h)

    def current(self):
        """(jcon[N1M][N2M][4], fcon[N1M][N2M][6])"""
        j = np.zeros(self.shape[:2] + (4,))
        f = np.zeros(self.shape[:2] + (6,))
        self.lib.or_current_get(self.h, _d(j), _d(f))
        return j, f

    def dump_content(self):
        """dump_content's 50 physics columns per interior zone [n1][n2][50]"""
        out = np.zeros((self.prm.n1, self.prm.n2, 50))
        self.lib.or_dump_content(self.h, _d(out))
        return out

    def constotal(self, region, sel=P):
        v = np.zeros(NPR)
        self.lib.or_constotal(self.h, sel, region, _d(v))
        return v


def slope_lim(lim, k, yll, yl, yc, yr, yrr):
    lib = load()
    dq, l, r = ctypes.c_double(np.nan), ctypes.c_double(np.nan), ctypes.c_double(np.nan)
    lib.or_slope_lim(lim, k, yll, yl, yc, yr, yrr, ctypes.byref(dq), ctypes.byref(l), ctypes.byref(r))
    return dq.value, l.value, r.value


ENO_A2C, ENO_C2A = 0, 1


def eno_line(kind, line, pos):
    """cvtij of one value: kind ENO_A2C (averages to centre) or ENO_C2A, along
    the line of interior values; returns (value, order, shift)."""
    lib = load()
    a = np.ascontiguousarray(line, dtype=np.float64)
    order, shift = ctypes.c_int(), ctypes.c_int()
    v = lib.or_eno_line(kind, _d(a), len(a), pos, ctypes.byref(order), ctypes.byref(shift))
    return v, order.value, shift.value


def eno_weight(kind, order, s, d):
    return load().or_eno_weight(kind, order, s, d)
