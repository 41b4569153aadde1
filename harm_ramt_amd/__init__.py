"""HARM conservative timestep on MI355X (gfx950): Python host mirror of the C-ABI.

The product is the shared library ``_lib/libharm_amd.so`` (HIP kernels + the
C-ABI declared in ``include/harm_amd.h``).  This module binds it with ctypes and
mirrors the reference's call surface -- ``fluxcalc``, ``flux_ct``, ``advance``,
``bound_prim``, ``step_ch``, ``Utoprimgen`` (step_ch.c, advance.c,
bounds.fishmon.c) -- on numpy arrays in the reference layout
``p[N1+6][N2+6][8]``.  There is no CPU fallback: if the library is missing,
importing the bindings raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libharm_amd.so")

NPR = 8
NG = 3
NUMPFLAGS = 5
# mnemonics (same values as the reference)
DONOR, VANL, MINM, MC, PARA, CSSLOPE = 0, 1, 2, 3, 4, 90
NLIMCENT, NLIMUP, NLIMDOWN = 101, 102, 103
LAXFFLUX, HLLFLUX = 0, 1
FLUXCTHLL, FLUXCTTOTH = 0, 1
EOMGRMHD, EOMFFDE = 0, 1
OUTFLOW, FIXED, POLARAXIS, PERIODIC = 0, 3, 4, 100
BOUNDS_FISHMON, BOUNDS_DISK = 0, 1
X1UP, X1DN, X2UP, X2DN = 0, 1, 2, 3
P, PK0, PK1, F1, F2, ULAST, UNEW = 0, 1, 2, 3, 4, 5, 6
METRIC_MINK, METRIC_KS = 0, 2
FLUX_FACE_64x4, FLUX_FACE_64x2, FLUX_PAIR, FLUX_SHARED = 0, 1, 2, 3
BOUND_KERNELS, BOUND_BLOCK, BOUND_KERNELS_SEQCHECK, BOUND_FIXONE, BOUND_FUSED, BOUND_INUPDATE = 0, 1, 2, 3, 4, 5
ACCT_ORDERED, ACCT_TREE = 0, 1
K_FLUX, K_EMF, K_UPDATE, K_BOUNDARY = 0, 2, 3, 4
STAGE_NAMES = {K_FLUX: "flux", K_EMF: "flux_ct", K_UPDATE: "update", K_BOUNDARY: "boundary"}


class HarmParams(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_int), ("n2", ctypes.c_int),
        ("eomtype", ctypes.c_int), ("lim", ctypes.c_int), ("fluxmethod", ctypes.c_int),
        ("fluxb", ctypes.c_int), ("timeorder", ctypes.c_int),
        ("bc", ctypes.c_int * 4),
        ("fixupflux", ctypes.c_int), ("poledeath", ctypes.c_int), ("checksolution", ctypes.c_int),
        ("eno", ctypes.c_int),
        ("gam", ctypes.c_double), ("cour", ctypes.c_double), ("safe", ctypes.c_double),
        ("defcon", ctypes.c_double), ("gammamax", ctypes.c_double), ("gammafail", ctypes.c_double),
        ("rhominlimit", ctypes.c_double), ("uuminlimit", ctypes.c_double),
        ("dx1", ctypes.c_double), ("dx2", ctypes.c_double),
        ("boundsfile", ctypes.c_int), ("extrap", ctypes.c_int), ("disk_first", ctypes.c_int),
        ("rescaletype", ctypes.c_int),
    ]


NUMENERREGIONS, NUMDIRS, NUMTSCALES, NUMFAILFLOORFLAGS = 3, 4, 4, 9


_d, _l, _i = ctypes.c_double, ctypes.c_long, ctypes.c_int


class RestartHeader(ctypes.Structure):
    """harm_restart_header: the fields of the reference's restart header
    (restart.c:554-682), named as its globals."""
    _fields_ = [
        ("totalsize", _i * 2), ("t", _d), ("tf", _d), ("nstep", _l), ("a", _d), ("gam", _d), ("cour", _d),
        ("DTd", _d), ("DTavg", _d), ("DTener", _d), ("DTi", _d), ("DTr", _l), ("DTdebug", _d),
        ("dump_cnt", _l), ("image_cnt", _l), ("rdump_cnt", _l), ("avg_cnt", _l), ("debug_cnt", _l),
        ("dt", _d), ("lim", _i), ("TIMEORDER", _i), ("fluxmethod", _i), ("FLUXB", _i), ("UTOPRIMVERSION", _i),
        ("failed", _i), ("R0", _d), ("Rin", _d), ("Rout", _d), ("hslope", _d), ("defcoord", _i),
        ("BCtype", _i * 4), ("realnstep", _l), ("debugfail", _i), ("whichrestart", _i), ("cooling", _i),
        ("restartsteps", _l * 2),
    ] + [(n, _i) for n in ("GAMMIEDUMP", "GAMMIEIMAGE", "GAMMIEENER", "DODIAGS", "DOENERDIAG", "DOGDUMPDIAG",
                           "DORDUMPDIAG", "DODUMPDIAG", "DOAVGDIAG", "DOIMAGEDIAG", "DOAREAMAPDIAG", "POSDEFMETRIC",
                           "periodicx1", "periodicx2", "binaryoutput", "sortedoutput")] + [
        (n, _d) for n in ("defcon", "SAFE", "RHOMIN", "UUMIN", "RHOMINLIMIT", "UUMINLIMIT", "BSQORHOLIMIT",
                          "BSQOULIMIT", "GAMMAMAX", "GAMMADAMP", "GAMMAFAIL")] + [
        ("pcumreg_tot", _d * 8 * NUMDIRS * NUMENERREGIONS), ("fladdreg_tot", _d * 8 * NUMENERREGIONS),
        ("sourceaddreg_tot", _d * 8 * NUMENERREGIONS), ("Ureg_init_tot", _d * 8 * NUMENERREGIONS),
        ("failfloorcount_tot", _d * NUMFAILFLOORFLAGS * NUMTSCALES), ("prMAX", _d * 8), ("rescaletype", _i),
        ("lumvsr_tot", ctypes.POINTER(_d)),
    ]


def rdump_write(path, header, p, binary=False):
    """Write a restart file of the interior of the host array p[N1+6][N2+6][8]
    (restart_write, restart.c:182).  header: RestartHeader, its lumvsr_tot
    pointing at N1 doubles or NULL."""
    p = np.ascontiguousarray(p, dtype=np.float64)
    n1, n2 = p.shape[0] - 2 * NG, p.shape[1] - 2 * NG
    rc = load_library().harm_rdump_write_host(os.fsencode(path), int(binary), ctypes.byref(header), n1, n2, _dptr(p))
    if rc:
        raise RuntimeError(f"harm_rdump_write_host {path}: error {rc}")


def rdump_read(path, n1, n2, binary=False, p=None):
    """Read a restart file (restart_read, restart.c:234) into the interior of
    p (zeros when not given); returns (RestartHeader, p, lumvsr)."""
    h = RestartHeader()
    lum = np.zeros(n1)
    h.lumvsr_tot = lum.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    p = np.zeros((n1 + 2 * NG, n2 + 2 * NG, NPR)) if p is None else np.ascontiguousarray(p, dtype=np.float64)
    rc = load_library().harm_rdump_read_host(os.fsencode(path), int(binary), ctypes.byref(h), n1, n2, _dptr(p))
    if rc:
        raise RuntimeError(f"harm_rdump_read_host {path}: error {rc}")
    h.lumvsr_tot = ctypes.POINTER(ctypes.c_double)()
    return h, p, lum


class HarmGrid(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_int), ("n2", ctypes.c_int), ("metric", ctypes.c_int), ("defcoord", ctypes.c_int),
        ("a", ctypes.c_double), ("R0", ctypes.c_double), ("Rin", ctypes.c_double), ("Rout", ctypes.c_double),
        ("hslope", ctypes.c_double), ("startx1", ctypes.c_double), ("startx2", ctypes.c_double),
        ("dx1", ctypes.c_double), ("dx2", ctypes.c_double), ("rhomin", ctypes.c_double),
        ("uumin", ctypes.c_double), ("Zin", ctypes.c_double), ("Zout", ctypes.c_double),
        ("npow", ctypes.c_double),
    ]


NUMENERREGIONS, NUMFLUXTERMS, NUMSOURCES, NUMFAILFLOORFLAGS = 3, 7, 7, 9


class EnerRegion(ctypes.Structure):
    """harm_ener_region: one energy region's accumulators (defs.h:218-239)"""
    _fields_ = [("pdot", (ctypes.c_double * 8) * 4),
                ("pdotterms", ((ctypes.c_double * 8) * NUMFLUXTERMS) * 4),
                ("pcum", (ctypes.c_double * 8) * 4),
                ("fladd", ctypes.c_double * 8),
                ("fladdterms", (ctypes.c_double * 8) * NUMFAILFLOORFLAGS),
                ("sourceadd", ctypes.c_double * 8),
                ("sourceaddterms", (ctypes.c_double * 8) * NUMSOURCES)]


DUMP_COLUMNS = 56


class DumpHeader(ctypes.Structure):
    """harm_dump_header: dump_header()'s fields (dump.c:48-80)"""
    _fields_ = [("t", ctypes.c_double), ("totalsize", ctypes.c_int * 2), ("startx", ctypes.c_double * 2),
                ("dx", ctypes.c_double * 2), ("realnstep", ctypes.c_long), ("gam", ctypes.c_double),
                ("a", ctypes.c_double), ("R0", ctypes.c_double), ("Rin", ctypes.c_double),
                ("Rout", ctypes.c_double), ("hslope", ctypes.c_double), ("dt", ctypes.c_double),
                ("defcoord", ctypes.c_int)]


def dump_write(path, header, cols, binary=False):
    """the reference's dump file from cols[n1][n2][56] (harm_dump_write_host)"""
    lib = load_library()
    cols = np.ascontiguousarray(cols, dtype=np.float64)
    n1, n2 = cols.shape[:2]
    assert cols.shape[2] == DUMP_COLUMNS
    rc = lib.harm_dump_write_host(os.fsencode(path), int(binary), ctypes.byref(header), n1, n2, _dptr(cols))
    if rc != 0:
        raise HarmError(f"harm_dump_write_host {path}: error {rc}")


def zone_coords(grid, i, j):
    """(X1, X2, r, th) at the centre of zone (i,j)"""
    v = [ctypes.c_double() for _ in range(4)]
    load_library().harm_zone_coords(ctypes.byref(grid), i, j, *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def ener_regions(grid, h_over_r_jet=0.4):
    """setflux()+setjetflux() (initbase.c:736-980) on a grid: (enerpos,
    doflux), int32 arrays [3][4] indexed [region][X1UP,X1DN,X2UP,X2DN]"""
    lib = load_library()
    ep = np.zeros((NUMENERREGIONS, 4), dtype=np.int32)
    df = np.zeros((NUMENERREGIONS, 4), dtype=np.int32)
    rc = lib.harm_ener_regions(ctypes.byref(grid), h_over_r_jet, _iptr(ep), _iptr(df))
    if rc != 0:
        raise HarmError(f"harm_ener_regions: error {rc}")
    return ep, df


# every symbol include/harm_amd.h declares (checked by the CPU tests)
EXPORTS = [
    "harm_create", "harm_destroy", "harm_last_error", "harm_set_stream", "harm_set_geometry",
    "harm_set_array", "harm_get_array", "harm_array_ptr", "harm_get_pflag", "harm_set_time",
    "harm_get_time", "harm_fluxcalc", "harm_fix_flux", "harm_flux_ct", "harm_advance",
    "harm_bound_prim", "harm_pre_fixup", "harm_post_fixup", "harm_step", "harm_step_n", "harm_utoprimgen", "harm_utoprimgen_noalloc", "harm_get_failures",
    "harm_tl_create", "harm_tl_destroy", "harm_tl_set_geometry_tangent", "harm_tl_set_spacing_tangent",
    "harm_tl_set_state_tangent",
    "harm_tl_get_state_tangent", "harm_tl_step_n", "harm_tl_energy_flux",
    "harm_get_nstroke", "harm_get_newton_spread", "harm_prof_enable", "harm_prof_get", "harm_prof_reset",
    "harm_set_tuning", "harm_get_tuning",
    "harm_batch_create", "harm_batch_destroy", "harm_batch_step_n", "harm_batch_set_tuning",
    "harm_batch_get_tuning",
    "harm_adj_create", "harm_adj_destroy", "harm_adj_forward", "harm_adj_get_dts",
    "harm_adj_seed_magnetic_energy", "harm_adj_seed", "harm_adj_backward", "harm_adj_get_lambda",
    "harm_adj_get_nonsmooth", "harm_adj_branch_switches",
    "harm_grid_mks", "harm_grid_cartesian", "harm_grid_cylmink", "harm_geometry", "harm_init_fishmon",
    "harm_set_analytic", "harm_set_radius",
    "harm_init_orszag_tang", "harm_init_monopole",
    "harm_rdump_write_host", "harm_rdump_read_host", "harm_rdump_write", "harm_rdump_read",
    "harm_ener_regions", "harm_acct_setup", "harm_acct_off", "harm_acct_set_order", "harm_acct_read",
    "harm_diag_flux",
    "harm_divbmaxavg", "harm_constotal",
    "harm_current_enable", "harm_get_current", "harm_dump_write_host", "harm_dump_write", "harm_zone_coords",
]

_lib = None


def load_library(path=LIB_PATH):
    """Load libharm_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run python -m harm_ramt_amd.build)")
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    sig = {
        "harm_create": [ctypes.POINTER(HarmParams), ctypes.POINTER(vp)],
        "harm_destroy": [vp],
        "harm_set_stream": [vp, vp],
        "harm_set_geometry": [vp, dp, dp, dp, dp, dp],
        "harm_set_array": [vp, ctypes.c_int, dp],
        "harm_get_array": [vp, ctypes.c_int, dp],
        "harm_array_ptr": [vp, ctypes.c_int, ctypes.POINTER(vp)],
        "harm_get_pflag": [vp, ip],
        "harm_set_time": [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double],
        "harm_get_time": [vp, dp, dp],
        "harm_fluxcalc": [vp, ctypes.c_int, ctypes.c_int, dp],
        "harm_fix_flux": [vp],
        "harm_flux_ct": [vp],
        "harm_advance": [vp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, dp],
        "harm_bound_prim": [vp, ctypes.c_int],
        "harm_pre_fixup": [vp, ctypes.c_int],
        "harm_post_fixup": [vp, ctypes.c_int],
        "harm_step": [vp],
        "harm_step_n": [vp, ctypes.c_int],
        "harm_utoprimgen": [vp, ctypes.c_int, ip, dp, dp, ip],
        "harm_utoprimgen_noalloc": [vp, ctypes.c_int, ip, dp, dp, ip],
        "harm_tl_create": [vp, ctypes.POINTER(vp)],
        "harm_tl_destroy": [vp],
        "harm_tl_set_geometry_tangent": [vp, dp, dp, dp, dp],
        "harm_tl_set_spacing_tangent": [vp, ctypes.c_double, ctypes.c_double],
        "harm_tl_set_state_tangent": [vp, dp],
        "harm_tl_get_state_tangent": [vp, dp],
        "harm_tl_step_n": [vp, ctypes.c_int],
        "harm_tl_energy_flux": [vp, ctypes.c_int, dp, dp],
        "harm_get_failures": [vp, ip, ip],
        "harm_get_nstroke": [vp, ctypes.POINTER(ctypes.c_longlong)],
        "harm_get_newton_spread": [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
        "harm_prof_enable": [vp, ctypes.c_int],
        "harm_prof_get": [vp, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_longlong)],
        "harm_prof_reset": [vp],
        "harm_set_tuning": [vp, ctypes.c_int, ctypes.c_int],
        "harm_get_tuning": [vp, ip, ip],
        "harm_batch_create": [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)],
        "harm_batch_destroy": [vp],
        "harm_batch_step_n": [vp, ctypes.c_int],
        "harm_batch_set_tuning": [vp, ctypes.c_int, ctypes.c_int],
        "harm_batch_get_tuning": [vp, ip, ip],
        "harm_adj_create": [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)],
        "harm_adj_destroy": [vp],
        "harm_adj_forward": [vp, dp],
        "harm_adj_get_dts": [vp, dp],
        "harm_adj_seed_magnetic_energy": [vp, dp],
        "harm_adj_seed": [vp, dp],
        "harm_adj_backward": [vp],
        "harm_adj_get_lambda": [vp, dp],
        "harm_adj_get_nonsmooth": [vp, ip, ip],
        "harm_adj_branch_switches": [vp, vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
        "harm_grid_mks": [ctypes.POINTER(HarmGrid), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                          ctypes.c_double, ctypes.c_double, ctypes.c_double],
        "harm_grid_cartesian": [ctypes.POINTER(HarmGrid), ctypes.c_int, ctypes.c_int],
        "harm_grid_cylmink": [ctypes.POINTER(HarmGrid), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, ctypes.c_double],
        "harm_set_analytic": [vp, dp, dp],
        "harm_set_radius": [vp, dp],
        "harm_geometry": [ctypes.POINTER(HarmGrid), dp, dp, dp, dp, dp],
        "harm_init_fishmon": [ctypes.POINTER(HarmGrid), dp, dp, dp, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              ctypes.c_ulong, dp],
        "harm_init_orszag_tang": [ctypes.POINTER(HarmGrid), ctypes.c_double, dp],
        "harm_init_monopole": [ctypes.POINTER(HarmGrid), dp, ctypes.c_double, dp],
        "harm_rdump_write_host": [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(RestartHeader), ctypes.c_int,
                                  ctypes.c_int, dp],
        "harm_rdump_read_host": [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(RestartHeader), ctypes.c_int,
                                 ctypes.c_int, dp],
        "harm_rdump_write": [vp, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(RestartHeader)],
        "harm_rdump_read": [vp, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(RestartHeader)],
        "harm_ener_regions": [ctypes.POINTER(HarmGrid), ctypes.c_double, ip, ip],
        "harm_acct_setup": [vp, ip, ip, ctypes.c_double],
        "harm_acct_off": [vp],
        "harm_acct_set_order": [vp, ctypes.c_int],
        "harm_acct_read": [vp, ctypes.POINTER(EnerRegion), dp, dp, dp],
        "harm_diag_flux": [vp, ctypes.c_double],
        "harm_divbmaxavg": [vp, ctypes.c_int, dp, dp, ip, ip],
        "harm_constotal": [vp, ctypes.c_int, ctypes.c_int, dp],
        "harm_current_enable": [vp, ctypes.c_int],
        "harm_get_current": [vp, dp, dp],
        "harm_dump_write_host": [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(DumpHeader), ctypes.c_int,
                                 ctypes.c_int, dp],
        "harm_dump_write": [vp, ctypes.POINTER(HarmGrid), ctypes.c_char_p, ctypes.c_int, ctypes.c_long],
        "harm_zone_coords": [ctypes.POINTER(HarmGrid), ctypes.c_int, ctypes.c_int, dp, dp, dp, dp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    lib.harm_last_error.argtypes = []
    lib.harm_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class HarmError(RuntimeError):
    pass


def _chk(rc, what):
    if rc != 0:
        raise HarmError(f"{what}: {load_library().harm_last_error().decode()}")


# ---------------------------------------------------------------------------
# host-side problem generation (no GPU)
# ---------------------------------------------------------------------------
class Geometry:
    """Geometry tables of set_grid() in the reference layout."""

    def __init__(self, grid: HarmGrid):
        n1m, n2m = grid.n1 + 2 * NG, grid.n2 + 2 * NG
        self.grid = grid
        self.gcov = np.zeros((n1m, n2m, 4, 4, 4))
        self.gcon = np.zeros((n1m, n2m, 4, 4, 4))
        self.gdet = np.zeros((n1m, n2m, 4))
        self.conn = np.zeros((n1m, n2m, 4, 4, 4))
        self.floors = np.zeros((n1m, n2m, 2))
        lib = load_library()
        _chk(lib.harm_geometry(ctypes.byref(grid), _dptr(self.gcov), _dptr(self.gcon), _dptr(self.gdet),
                               _dptr(self.conn), _dptr(self.floors)), "harm_geometry")

    def tables(self):
        return self.gcov, self.gcon, self.gdet, self.conn, self.floors

    @property
    def radius(self):
        """r of every zone centre [N1M][N2M] (harm_zone_coords), for EXTRAP 2"""
        if getattr(self, "_radius", None) is None:
            g = self.grid
            r = np.zeros((g.n1 + 2 * NG, g.n2 + 2 * NG))
            for i in range(-NG, g.n1 + NG):
                for j in range(-NG, g.n2 + NG):
                    r[i + NG, j + NG] = zone_coords(g, i, j)[2]
            self._radius = r
        return self._radius


def grid_mks(n1, n2, a=0.9375, Rout=40.0, hslope=0.3, R0=0.0):
    g = HarmGrid()
    _chk(load_library().harm_grid_mks(ctypes.byref(g), n1, n2, a, Rout, hslope, R0), "harm_grid_mks")
    return g


def grid_cartesian(n1, n2):
    g = HarmGrid()
    _chk(load_library().harm_grid_cartesian(ctypes.byref(g), n1, n2), "harm_grid_cartesian")
    return g


def grid_cylmink(n1, n2, Rin=1.0, Rout=1e3, Zin=-1e3, Zout=1e3):
    """Cylindrical Minkowski (R, z) on the bi-exponential grid defcoord 666,
    the reference's default problem's grid (init.c:54-60)."""
    g = HarmGrid()
    _chk(load_library().harm_grid_cylmink(ctypes.byref(g), n1, n2, Rin, Rout, Zin, Zout), "harm_grid_cylmink")
    return g


def init_fishmon(geom: Geometry, gam=4.0 / 3.0, rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0,
                 randfact=4e-2, seed=7):
    g = geom.grid
    p = np.zeros((g.n1 + 2 * NG, g.n2 + 2 * NG, NPR))
    _chk(load_library().harm_init_fishmon(ctypes.byref(g), _dptr(geom.gcov), _dptr(geom.gcon), _dptr(geom.gdet),
                                          gam, rin, rmax, kappa, beta, randfact, seed, _dptr(p)),
         "harm_init_fishmon")
    return p


def init_orszag_tang(geom: Geometry, gam=5.0 / 3.0):
    g = geom.grid
    p = np.zeros((g.n1 + 2 * NG, g.n2 + 2 * NG, NPR))
    _chk(load_library().harm_init_orszag_tang(ctypes.byref(g), gam, _dptr(p)), "harm_init_orszag_tang")
    return p


def init_monopole(geom: Geometry, B0=1.0):
    """Force-free split monopole (init.bz.c:146-155) on a Kerr grid."""
    g = geom.grid
    p = np.zeros((g.n1 + 2 * NG, g.n2 + 2 * NG, NPR))
    _chk(load_library().harm_init_monopole(ctypes.byref(g), _dptr(geom.gdet), B0, _dptr(p)), "harm_init_monopole")
    return p


def make_params(n1, n2, dx1, dx2, *, lim=PARA, fluxmethod=HLLFLUX, timeorder=4, bc=None, fixupflux=1,
                poledeath=1, checksolution=0, eno=0, gam=4.0 / 3.0, cour=0.9, safe=1.3, defcon=1.0, gammamax=25.0,
                gammafail=None, rhominlimit=1e-20, uuminlimit=1e-20, fluxb=FLUXCTTOTH, eomtype=EOMGRMHD,
                boundsfile=0, extrap=0, disk_first=0, rescaletype=1):
    """Parameter block with the reference's GRMHD defaults (initbase.c:119-243)."""
    prm = HarmParams()
    prm.n1, prm.n2 = n1, n2
    prm.eomtype, prm.lim, prm.fluxmethod, prm.fluxb, prm.timeorder = eomtype, lim, fluxmethod, fluxb, timeorder
    b = bc if bc is not None else (OUTFLOW, OUTFLOW, POLARAXIS, POLARAXIS)
    for k in range(4):
        prm.bc[k] = b[k]
    prm.fixupflux, prm.poledeath, prm.checksolution, prm.eno = fixupflux, poledeath, checksolution, eno
    prm.gam, prm.cour, prm.safe, prm.defcon = gam, cour, safe, defcon
    prm.gammamax = gammamax
    prm.gammafail = 100.0 * gammamax if gammafail is None else gammafail
    prm.rhominlimit, prm.uuminlimit = rhominlimit, uuminlimit
    prm.dx1, prm.dx2 = dx1, dx2
    prm.boundsfile, prm.extrap, prm.disk_first, prm.rescaletype = boundsfile, extrap, disk_first, rescaletype
    return prm


def params_dict(prm: HarmParams):
    return {f: (list(getattr(prm, f)) if f == "bc" else getattr(prm, f)) for f, _ in HarmParams._fields_}


# ---------------------------------------------------------------------------
# the GPU context: the reference's call surface
# ---------------------------------------------------------------------------
class Harm:
    """One HARM grid resident on the GPU (a harm_ctx)."""

    def __init__(self, prm: HarmParams, geom: Geometry = None):
        self.lib = load_library()
        self.prm = prm
        self.shape = (prm.n1 + 2 * NG, prm.n2 + 2 * NG, NPR)
        h = ctypes.c_void_p()
        _chk(self.lib.harm_create(ctypes.byref(prm), ctypes.byref(h)), "harm_create")
        self.h = h
        if geom is not None:
            self.set_geometry(*geom.tables())
            if prm.extrap == 2:
                self.set_radius(geom.radius)

    def close(self):
        if self.h:
            self.lib.harm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_radius(self, r):
        """EXTRAP 2: r of every zone centre [N1M][N2M] (Geometry.radius)"""
        r = np.ascontiguousarray(r, dtype=np.float64)
        _chk(self.lib.harm_set_radius(self.h, _dptr(r)), "harm_set_radius")

    def set_geometry(self, gcov, gcon, gdet, conn, floors):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (gcov, gcon, gdet, conn, floors)]
        _chk(self.lib.harm_set_geometry(self.h, *[_dptr(a) for a in arrs]), "harm_set_geometry")

    def set_array(self, sel, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape
        _chk(self.lib.harm_set_array(self.h, sel, _dptr(a)), "harm_set_array")

    def set_analytic(self, panalytic, omegaf=None):
        """harm_set_analytic: the analytic solution FIXED boundaries and
        bound_disk() read (panalytic, omegafanalytic; [N1M][N2M][NPR])."""
        pa = np.ascontiguousarray(panalytic, dtype=np.float64)
        assert pa.shape == self.shape
        om = None if omegaf is None else np.ascontiguousarray(omegaf, dtype=np.float64)
        if om is not None:
            assert om.shape == self.shape
        _chk(self.lib.harm_set_analytic(self.h, _dptr(pa), _dptr(om) if om is not None else None),
             "harm_set_analytic")

    def get_array(self, sel):
        a = np.zeros(self.shape)
        _chk(self.lib.harm_get_array(self.h, sel, _dptr(a)), "harm_get_array")
        return a

    def array_ptr(self, sel):
        v = ctypes.c_void_p()
        _chk(self.lib.harm_array_ptr(self.h, sel, ctypes.byref(v)), "harm_array_ptr")
        return v.value

    def get_pflag(self):
        a = np.zeros(self.shape[:2] + (NUMPFLAGS,), dtype=np.int32)
        _chk(self.lib.harm_get_pflag(self.h, _iptr(a)), "harm_get_pflag")
        return a

    def set_time(self, t, dt, tf):
        _chk(self.lib.harm_set_time(self.h, t, dt, tf), "harm_set_time")

    def get_time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        _chk(self.lib.harm_get_time(self.h, ctypes.byref(t), ctypes.byref(dt)), "harm_get_time")
        return t.value, dt.value

    # step_ch.c:741
    def fluxcalc(self, pr_sel, dir):
        ndt = ctypes.c_double()
        _chk(self.lib.harm_fluxcalc(self.h, pr_sel, dir, ctypes.byref(ndt)), "harm_fluxcalc")
        return ndt.value

    # fixup.c:1357
    def fix_flux(self):
        _chk(self.lib.harm_fix_flux(self.h), "harm_fix_flux")

    # step_ch.c:1159
    def flux_ct(self):
        _chk(self.lib.harm_flux_ct(self.h), "harm_flux_ct")

    # advance.c:15
    def advance(self, pi, pb, CUf, pf, Cunew, timeorder, numtimeorders):
        cu = np.ascontiguousarray(CUf, dtype=np.float64)
        cn = np.ascontiguousarray(Cunew, dtype=np.float64)
        ndt = ctypes.c_double()
        _chk(self.lib.harm_advance(self.h, pi, pb, _dptr(cu), pf, _dptr(cn), timeorder, numtimeorders,
                                   ctypes.byref(ndt)), "harm_advance")
        return ndt.value

    # bounds.fishmon.c:13
    def bound_prim(self, sel=P):
        _chk(self.lib.harm_bound_prim(self.h, sel), "harm_bound_prim")

    # fixup.c:25
    def pre_fixup(self, sel=P):
        _chk(self.lib.harm_pre_fixup(self.h, sel), "harm_pre_fixup")

    # fixup.c:37
    def post_fixup(self, sel=P):
        _chk(self.lib.harm_post_fixup(self.h, sel), "harm_post_fixup")

    # step_ch.c:16
    def step_ch(self):
        _chk(self.lib.harm_step(self.h), "harm_step")

    def step_n(self, n):
        _chk(self.lib.harm_step_n(self.h, n), "harm_step_n")

    # restart.c:182 restart_write / :234 restart_read on HARM_P
    def rdump_write(self, path, header, binary=False):
        _chk(self.lib.harm_rdump_write(self.h, os.fsencode(path), int(binary), ctypes.byref(header)),
             "harm_rdump_write")

    def rdump_read(self, path, binary=False):
        """Reads the interior of HARM_P (call bound_prim afterwards, as
        restart_init does); returns (RestartHeader, lumvsr)."""
        h = RestartHeader()
        lum = np.zeros(self.prm.n1)
        h.lumvsr_tot = lum.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        _chk(self.lib.harm_rdump_read(self.h, os.fsencode(path), int(binary), ctypes.byref(h)), "harm_rdump_read")
        h.lumvsr_tot = ctypes.POINTER(ctypes.c_double)()
        return h, lum

    # conservation accounting (diag.c diag_flux/diag_source, dump_ener.c)
    def acct_setup(self, enerpos, doflux, dx3=2 * np.pi):
        ep = np.ascontiguousarray(enerpos, dtype=np.int32)
        df = np.ascontiguousarray(doflux, dtype=np.int32)
        if ep.shape != (NUMENERREGIONS, 4) or df.shape != (NUMENERREGIONS, 4):
            raise ValueError("enerpos and doflux are [3][4]")
        _chk(self.lib.harm_acct_setup(self.h, _iptr(ep), _iptr(df), dx3), "harm_acct_setup")

    def acct_off(self):
        _chk(self.lib.harm_acct_off(self.h), "harm_acct_off")

    def acct_set_order(self, order):
        """ACCT_ORDERED (the reference's serial sums, bitwise) or ACCT_TREE
        (source and fixup totals by reduction trees, for throughput)"""
        _chk(self.lib.harm_acct_set_order(self.h, order), "harm_acct_set_order")

    def acct(self):
        """the accumulators as a dict of arrays with a leading region axis,
        plus pdottermsjet2[4][7][8], frdot[n1][8] and lumvsr[n1]"""
        regs = (EnerRegion * NUMENERREGIONS)()
        jet2 = np.zeros((4, NUMFLUXTERMS, 8))
        frdot = np.zeros((self.prm.n1, 8))
        lum = np.zeros(self.prm.n1)
        _chk(self.lib.harm_acct_read(self.h, regs, _dptr(jet2), _dptr(frdot), _dptr(lum)), "harm_acct_read")
        out = {name: np.array([np.ctypeslib.as_array(getattr(regs[r], name)) for r in range(NUMENERREGIONS)])
               for name, _ in EnerRegion._fields_}
        out["pdottermsjet2"], out["frdot"], out["lumvsr"] = jet2, frdot, lum
        return out

    def diag_flux(self, Dt):
        _chk(self.lib.harm_diag_flux(self.h, Dt), "harm_diag_flux")

    def divbmaxavg(self, sel=P):
        """(divbmax, divbavg, imax, jmax) of dump_ener.c:449 divbmaxavg"""
        mx, av = ctypes.c_double(), ctypes.c_double()
        im, jm = ctypes.c_int(), ctypes.c_int()
        _chk(self.lib.harm_divbmaxavg(self.h, sel, ctypes.byref(mx), ctypes.byref(av), ctypes.byref(im),
                                      ctypes.byref(jm)), "harm_divbmaxavg")
        return mx.value, av.value, im.value, jm.value

    # the current J^mu, F^{mu nu} (phys.c:1124-1460) and the dump (dump.c)
    def current_enable(self, on=True):
        _chk(self.lib.harm_current_enable(self.h, int(on)), "harm_current_enable")

    def current(self):
        """(jcon[N1M][N2M][4], fcon[N1M][N2M][6])"""
        n1m, n2m = self.prm.n1 + 6, self.prm.n2 + 6
        j = np.zeros((n1m, n2m, 4))
        f = np.zeros((n1m, n2m, 6))
        _chk(self.lib.harm_get_current(self.h, _dptr(j), _dptr(f)), "harm_get_current")
        return j, f

    def dump_write(self, path, grid, binary=False, realnstep=0):
        _chk(self.lib.harm_dump_write(self.h, ctypes.byref(grid), os.fsencode(path), int(binary), realnstep),
             "harm_dump_write")

    def constotal(self, region, sel=P):
        v = np.zeros(8)
        _chk(self.lib.harm_constotal(self.h, sel, region, _dptr(v)), "harm_constotal")
        return v

    # step_ch.c:468 for a batch of zones
    def utoprimgen(self, zones, U, pr):
        zones = np.ascontiguousarray(zones, dtype=np.int32)
        U = np.ascontiguousarray(U, dtype=np.float64)
        out = np.ascontiguousarray(pr, dtype=np.float64).copy()
        fl = np.zeros(len(zones), dtype=np.int32)
        _chk(self.lib.harm_utoprimgen(self.h, len(zones), _iptr(zones), _dptr(U), _dptr(out), _iptr(fl)),
             "harm_utoprimgen")
        return out, fl

    def failures(self):
        n, f = ctypes.c_int(), ctypes.c_int()
        _chk(self.lib.harm_get_failures(self.h, ctypes.byref(n), ctypes.byref(f)), "harm_get_failures")
        return n.value, f.value

    def newton_spread(self):
        """(Newton iterations, sum over wavefronts of each one's largest count)"""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        _chk(self.lib.harm_get_newton_spread(self.h, ctypes.byref(a), ctypes.byref(b)), "harm_get_newton_spread")
        return a.value, b.value

    def nstroke(self):
        v = ctypes.c_longlong()
        _chk(self.lib.harm_get_nstroke(self.h, ctypes.byref(v)), "harm_get_nstroke")
        return v.value

    def set_tuning(self, flux_variant, bound_variant):
        _chk(self.lib.harm_set_tuning(self.h, flux_variant, bound_variant), "harm_set_tuning")

    def get_tuning(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _chk(self.lib.harm_get_tuning(self.h, ctypes.byref(a), ctypes.byref(b)), "harm_get_tuning")
        return a.value, b.value

    def prof_enable(self, on=True):
        _chk(self.lib.harm_prof_enable(self.h, 1 if on else 0), "harm_prof_enable")

    def prof_reset(self):
        _chk(self.lib.harm_prof_reset(self.h), "harm_prof_reset")

    def prof_get(self, kernel):
        ms, cnt = ctypes.c_double(), ctypes.c_longlong()
        _chk(self.lib.harm_prof_get(self.h, kernel, ctypes.byref(ms), ctypes.byref(cnt)), "harm_prof_get")
        return ms.value, cnt.value


class Batch:
    """Many Harm contexts advanced together (harm_batch): one launch per stage
    over every run's tiles, one graph replay per batch step.  Keeps references
    to the contexts so they outlive the batch."""

    def __init__(self, runs):
        self.lib = load_library()
        self.runs = list(runs)
        arr = (ctypes.c_void_p * len(self.runs))(*[r.h.value for r in self.runs])
        h = ctypes.c_void_p()
        _chk(self.lib.harm_batch_create(arr, len(self.runs), ctypes.byref(h)), "harm_batch_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.harm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step_n(self, nsteps):
        _chk(self.lib.harm_batch_step_n(self.h, nsteps), "harm_batch_step_n")

    def set_tuning(self, flux_variant, bound_variant):
        _chk(self.lib.harm_batch_set_tuning(self.h, flux_variant, bound_variant), "harm_batch_set_tuning")

    def get_tuning(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _chk(self.lib.harm_batch_get_tuning(self.h, ctypes.byref(a), ctypes.byref(b)), "harm_batch_get_tuning")
        return a.value, b.value


class Tangent:
    """Tangent-linear run of a Harm context (harm_tl): the context's state
    and its derivative with respect to one parameter advance together."""

    def __init__(self, harm: "Harm"):
        self.lib = load_library()
        self.harm = harm
        h = ctypes.c_void_p()
        _chk(self.lib.harm_tl_create(harm.h, ctypes.byref(h)), "harm_tl_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.harm_tl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_geometry_tangent(self, dgcov, dgcon, dgdet, dconn):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (dgcov, dgcon, dgdet, dconn)]
        _chk(self.lib.harm_tl_set_geometry_tangent(self.h, *[_dptr(x) for x in a]), "harm_tl_set_geometry_tangent")

    def set_spacing_tangent(self, ddx1, ddx2=0.0):
        _chk(self.lib.harm_tl_set_spacing_tangent(self.h, ddx1, ddx2), "harm_tl_set_spacing_tangent")

    def set_state_tangent(self, dp):
        d = np.ascontiguousarray(dp, dtype=np.float64)
        assert d.shape == self.harm.shape
        _chk(self.lib.harm_tl_set_state_tangent(self.h, _dptr(d)), "harm_tl_set_state_tangent")

    def state_tangent(self):
        d = np.zeros(self.harm.shape)
        _chk(self.lib.harm_tl_get_state_tangent(self.h, _dptr(d)), "harm_tl_get_state_tangent")
        return d

    def step_n(self, nsteps):
        _chk(self.lib.harm_tl_step_n(self.h, nsteps), "harm_tl_step_n")

    def energy_flux(self, i):
        L, dL = ctypes.c_double(), ctypes.c_double()
        _chk(self.lib.harm_tl_energy_flux(self.h, i, ctypes.byref(L), ctypes.byref(dL)), "harm_tl_energy_flux")
        return L.value, dL.value


def spin_tangents(grid_of_spin, init_of_geometry, a, h=1e-4):
    """d/da of the geometry tables and of the initial state at spin a, by the
    fourth-order central difference of the host generators,
    (8 (T(a+h) - T(a-h)) - (T(a+2h) - T(a-2h))) / (12 h): the generators are
    smooth in a (the horizon zone, which fixes Rin's formula, does not depend
    on a), so the error is O(h^4) plus rounding of order 1e-16 / h.
    grid_of_spin(a) -> HarmGrid, init_of_geometry(Geometry) -> state.
    Returns (dgcov, dgcon, dgdet, dconn, dp0, ddx1, ddx2): the grid spacing
    dx1 moves with a through Rin."""
    def tabs(x):
        g = Geometry(grid_of_spin(x))
        return g.gcov, g.gcon, g.gdet, g.conn, init_of_geometry(g), g.grid.dx1, g.grid.dx2
    tp, tm, t2p, t2m = tabs(a + h), tabs(a - h), tabs(a + 2 * h), tabs(a - 2 * h)
    return tuple((8.0 * (p - m) - (p2 - m2)) / (12.0 * h) for p, m, p2, m2 in zip(tp, tm, t2p, t2m))


class Adjoint:
    """Discrete adjoint of n steps of a Harm context (harm_adj)."""

    def __init__(self, harm: "Harm", nsteps, checkpoint_every=16):
        self.lib = load_library()
        self.harm = harm
        self.nsteps = nsteps
        h = ctypes.c_void_p()
        _chk(self.lib.harm_adj_create(harm.h, nsteps, checkpoint_every, ctypes.byref(h)), "harm_adj_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.harm_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, dts=None):
        if dts is None:
            _chk(self.lib.harm_adj_forward(self.h, None), "harm_adj_forward")
        else:
            d = np.ascontiguousarray(dts, dtype=np.float64)
            assert d.shape == (self.nsteps,)
            _chk(self.lib.harm_adj_forward(self.h, _dptr(d)), "harm_adj_forward")

    def dts(self):
        d = np.zeros(self.nsteps)
        _chk(self.lib.harm_adj_get_dts(self.h, _dptr(d)), "harm_adj_get_dts")
        return d

    def seed_magnetic_energy(self):
        J = ctypes.c_double()
        _chk(self.lib.harm_adj_seed_magnetic_energy(self.h, ctypes.byref(J)), "harm_adj_seed_magnetic_energy")
        return J.value

    def seed(self, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        assert lam.shape == self.harm.shape
        _chk(self.lib.harm_adj_seed(self.h, _dptr(lam)), "harm_adj_seed")

    def backward(self):
        _chk(self.lib.harm_adj_backward(self.h), "harm_adj_backward")
        lam = np.zeros(self.harm.shape)
        _chk(self.lib.harm_adj_get_lambda(self.h, _dptr(lam)), "harm_adj_get_lambda")
        return lam

    def nonsmooth(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _chk(self.lib.harm_adj_get_nonsmooth(self.h, ctypes.byref(a), ctypes.byref(b)), "harm_adj_get_nonsmooth")
        return a.value, b.value

    def branch_switches(self, other: "Adjoint"):
        """(limiter branches that differ between this forward run and
        other's, slopes compared); both with checkpoint_every = nsteps"""
        n, c = ctypes.c_longlong(), ctypes.c_longlong()
        _chk(self.lib.harm_adj_branch_switches(self.h, other.h, ctypes.byref(n), ctypes.byref(c)),
             "harm_adj_branch_switches")
        return n.value, c.value
