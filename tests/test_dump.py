"""The 56-column dump (dump.c:23-191) and the current the reference's step
computes for it (WHICHCURRENTCALC 1, phys.c:1124-1460), against the files
the compiled reference wrote itself (tests/golden/dump_torus_para_hll_rk4.npz:
its dump() after 5 torus steps, text and binary, so that the current's time
derivative has run on its three-step form).

- The oracle's 50 physics columns (state, div B, u, b, wave speeds, gdet,
  J, F and their lowered forms) are the reference's to the bit; the six
  coordinate columns from the host grid are the reference's.
- Writing the reference's header and columns reproduces both files byte for
  byte.
- Closed forms: a uniform state has zero current; a fluid at rest in flat
  space has F^{rh}, F^{rp}, F^{hp} = B3, -B2, B1 and no F^{ti}.
- The device path (-m gpu): the context's dump after the same steps is the
  reference's file byte for byte, and its current matches the oracle's on a
  larger grid.
"""
import struct

import numpy as np
import pytest

import harm_ramt_amd as H
from oracle.oracle import Oracle
from tests.helpers import golden_problem, load_golden, problem, same, sha

NAME = "dump_torus_para_hll_rk4"
HDR_BYTES = 116  # dump_header() in binary: 12 doubles, 3 ints, 1 long


@pytest.fixture(scope="module")
def gd():
    return load_golden(NAME)


def _ref_columns(gd):
    n1, n2 = int(gd["n1"]), int(gd["n2"])
    b = gd["dump_binary"].tobytes()
    assert len(b) == HDR_BYTES + n1 * n2 * H.DUMP_COLUMNS * 8
    return np.frombuffer(b[HDR_BYTES:], dtype=np.float64).reshape(n2, n1, H.DUMP_COLUMNS).transpose(1, 0, 2)


def _ref_header(gd):
    b = gd["dump_binary"].tobytes()[:HDR_BYTES]
    v = struct.unpack("<dii4dq7di", b)
    h = H.DumpHeader()
    h.t, h.totalsize[0], h.totalsize[1] = v[0], v[1], v[2]
    h.startx[0], h.startx[1], h.dx[0], h.dx[1] = v[3:7]
    h.realnstep = v[7]
    h.gam, h.a, h.R0, h.Rin, h.Rout, h.hslope, h.dt = v[8:15]
    h.defcoord = v[15]
    return h


def _oracle(gd):
    g, geom, p0, prm = golden_problem(gd)
    assert sha(p0) == str(gd["input_sha"])
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, float(gd["dt0"]), 1e30)
    o.current_enable(True)
    for _ in range(int(gd["nsteps"])):
        o.step()
    return g, o


def test_oracle_dump_columns_match_reference(gd):
    g, o = _oracle(gd)
    assert same(o.get_array(0), gd["p"])
    ref = _ref_columns(gd)
    mine = o.dump_content()
    for c in range(50):
        assert same(mine[..., c], ref[..., 6 + c]), f"column {6 + c}"
    assert np.count_nonzero(ref[..., 41:45])  # the current is not trivially zero


def test_coordinate_columns_and_header(gd):
    n1, n2 = int(gd["n1"]), int(gd["n2"])
    g, geom, p0, prm = golden_problem(gd)
    ref = _ref_columns(gd)
    for i in range(n1):
        for j in range(n2):
            X1, X2, r, th = H.zone_coords(g, i, j)
            assert list(ref[i, j, :6]) == [i, j, X1, X2, r, th], (i, j)
    h = _ref_header(gd)
    assert h.totalsize[0] == n1 and h.totalsize[1] == n2 and h.realnstep == int(gd["nsteps"])
    assert h.t == gd["t"][-1] and h.dt == gd["dt"][-1] and h.gam == 4.0 / 3.0
    assert (h.startx[0], h.startx[1], h.dx[0], h.dx[1]) == (g.startx1, g.startx2, g.dx1, g.dx2)
    assert (h.a, h.R0, h.Rin, h.Rout, h.hslope, h.defcoord) == (g.a, g.R0, g.Rin, g.Rout, g.hslope, g.defcoord)


@pytest.mark.parametrize("binary", [False, True])
def test_writer_reproduces_reference_files(gd, tmp_path, binary):
    out = str(tmp_path / ("d.bin" if binary else "d"))
    H.dump_write(out, _ref_header(gd), _ref_columns(gd), binary=binary)
    ref = (gd["dump_binary"] if binary else gd["dump_text"]).tobytes()
    assert open(out, "rb").read() == ref


def test_oracle_current_uniform_and_at_rest():
    n1, n2 = 8, 8
    g, geom, p0, prm = problem("mink", n1, n2)
    p = np.zeros_like(p0)
    p[..., 0], p[..., 1] = 1.0, 0.5
    B = (0.3, -0.7, 0.2)
    p[..., 5], p[..., 6], p[..., 7] = B
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_array(0, p)
    o.set_time(0.0, 1e-3, 1e30)
    o.current_enable(True)
    for _ in range(3):  # three calls: the time-derivative form runs on the last
        o.current_step()
    j, f = o.current()
    assert np.all(j == 0)
    inner = f[3:-3, 3:-3]
    assert np.all(inner[..., :3] == 0)
    assert np.all(inner[..., 3] == B[2]) and np.all(inner[..., 4] == -B[1]) and np.all(inner[..., 5] == B[0])


# ---------------------------------------------------------------- GPU ----

@pytest.mark.gpu
@pytest.mark.parametrize("binary", [False, True])
def test_gpu_dump_matches_reference_file(gd, tmp_path, binary):
    print(f"[case] {NAME} binary={binary}", flush=True)
    g, geom, p0, prm = golden_problem(gd)
    c = H.Harm(prm, geom)
    c.set_array(H.P, p0)
    c.bound_prim(H.P)
    c.set_time(0.0, float(gd["dt0"]), 1e30)
    c.current_enable(True)
    for _ in range(int(gd["nsteps"])):
        c.step_ch()
    assert same(c.get_array(H.P), gd["p"])
    out = str(tmp_path / "gpu")
    c.dump_write(out, g, binary=binary, realnstep=int(gd["nsteps"]))
    assert open(out, "rb").read() == (gd["dump_binary"] if binary else gd["dump_text"]).tobytes()


@pytest.mark.gpu
def test_gpu_current_matches_oracle_larger_grid():
    n1, n2, nsteps = 64, 128, 6
    print(f"[case] current torus {n1}x{n2}", flush=True)
    g, geom, p0, prm = problem("ks", n1, n2)
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, 1e-3, 1e30)
    o.current_enable(True)
    c = H.Harm(prm, geom)
    c.set_array(H.P, p0)
    c.bound_prim(H.P)
    c.set_time(0.0, 1e-3, 1e30)
    c.current_enable(True)
    for s in range(nsteps):
        o.step()
        if s < 2:  # the first call (kernel variants timed) and the second, on the two-step form
            c.step_ch()
            jo, fo = o.current()
            jc, fc = c.current()
            assert same(jc, jo) and same(fc, fo), f"step {s + 1}"
    c.step_n(nsteps - 2)  # graph replays carry the call count and dt history on the device
    jo, fo = o.current()
    jc, fc = c.current()
    assert same(jc, jo) and same(fc, fo)
    assert same(c.get_array(H.P), o.get_array(0))
