"""The Fortran host and the restart / history I/O on the GPU.

* The Fortran driver (wcns_amd/fortran/wcns_drive: flang, bind(C) module and
  shims named as the reference's subroutines) runs 10 steps of the reference's
  CPU-runnable case (64x32x16, Euler, periodic) from files the case generator
  writes (param.dat, PLOT3D grid, 1976 boundary file, restart file).  Its
  residual history equals the ctypes path's bit for bit, in its stage-by-stage
  mode and in its graph mode, and so does the restart file it leaves.
* A run that saves its state after N steps and is resumed from that file in a
  new context continues bit for bit like the uninterrupted run (Euler, and SA
  with its .turb file), when the saved temperature is kept (the reference's
  restart recomputes it from the primitives, which the default load does).
* The force history recorded on the device equals wcns_force step by step.
"""
import os
import subprocess

import numpy as np
import pytest

from wcns_amd import build, cases

pytestmark = pytest.mark.gpu


def _full_history(path):
    rows = [ln.split() for ln in open(path) if ln.strip()]
    return np.array([float(r[1]) for r in rows]), np.array([float(r[2]) for r in rows])


def _records(path):
    raw, out, o = open(path, "rb").read(), [], 0
    while o < len(raw):
        n = int(np.frombuffer(raw[o:o + 4], "<i4")[0])
        out.append(raw[o + 4:o + 4 + n])
        o += 8 + n
    return out


def _write_case(case, d, **kw):
    cases.write_plot3d(case, os.path.join(d, "grid.x"))
    cases.write_inp(case, os.path.join(d, "grid.inp"))
    cases.write_param(case, os.path.join(d, "param.dat"), **kw)
    cases.write_flow(case, os.path.join(d, "flow.dat"))


@pytest.mark.parametrize("mode", ["stages", "graph"])
def test_fortran_driver_matches_ctypes_path(tmp_path, mode):
    from wcns_amd.solver import Solver
    case = cases.cartesian_periodic_case(ni=64, nj=32, nk=16, seed=1)
    _write_case(case, str(tmp_path), nomax=10, nwerror=5 if mode == "graph" else 1, ndisk=1000, nstart=1,
                tecname="flow.plt", nplot=2)
    print(f"fortran driver, {mode}", flush=True)
    p = subprocess.run([build.FDRIVE, "param.dat", mode], cwd=tmp_path, capture_output=True, text=True,
                       timeout=240)
    print(p.stdout[-2000:], flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    rms_f, max_f = _full_history(tmp_path / "res.dat.full")
    s = Solver(case)
    hist = s.step(10)
    hs = s.residual_history(10)
    assert np.array_equal(rms_f, hist)
    assert np.array_equal(max_f, hs[1])
    # the restart file the driver wrote at its last step holds the ctypes
    # path's state (r, u, v, w, p; an Euler run carries T from its input file)
    s.save_flow(str(tmp_path / "flow_py.dat"), nout=10)
    rf, rp = _records(tmp_path / "flow.dat"), _records(tmp_path / "flow_py.dat")
    assert rf[0] == rp[0] and len(rf) == len(rp) == 2
    qf, qp = (np.frombuffer(r[1], "<f8").reshape(-1, 6) for r in (rf, rp))
    assert np.array_equal(qf[:, :5], qp[:, :5])
    assert len(open(tmp_path / "res.dat").read().splitlines()) == (10 if mode == "stages" else 2)
    # nplot = 2: the volume Tecplot file of the last step, named as the
    # reference names it, equal to the ctypes path's but for T (as above)
    s.tecplot(str(tmp_path / "py_vol.plt"), 2)
    hf, names, _, df = parse_plt(open(tmp_path / "flow_vol.plt", "rb").read())
    hp, _, _, dp = parse_plt(open(tmp_path / "py_vol.plt", "rb").read())
    same = hf == hp and all(np.array_equal(np.delete(a[3], names.index("T"), 0), np.delete(b[3], names.index("T"), 0))
                            for a, b in zip(df, dp))
    assert same
    s.close()


@pytest.mark.parametrize("turb", [-1, 1], ids=["euler", "sa"])
def test_restart_resumes_bit_for_bit(tmp_path, turb):
    from wcns_amd.solver import Solver
    if turb == 1:
        case = cases.multiblock_case(nvis=1, wall=True, periodic_i=False, ilo_bc=5, ihi_bc=6, jhi_bc=3)
        case.params.nlamtur = 1
    else:
        case = cases.wavy_periodic_case(ni=24, nj=16, nk=12)
    print(f"restart, model {turb}", flush=True)
    a = Solver(case)
    a.step(6, history=False)
    path = str(tmp_path / "flow.dat")
    a.save_flow(path, nout=6, wholetime=0.5)
    ha = a.step(5)
    b = Solver(case)
    nout, wt = b.load_flow(path, t_from_file=True)
    assert (nout, wt) == (6, 0.5)
    hb = b.step(5)
    assert np.array_equal(ha, hb)
    names = ["r", "u", "v", "w", "p"] + (["qke", "vist"] if turb == 1 else [])
    for nb in range(len(case.blocks)):
        for n in names:
            x, y = a.field(nb, n)[3:-3, 3:-3, 3:-3], b.field(nb, n)[3:-3, 3:-3, 3:-3]
            assert np.array_equal(x, y), (nb, n)
    if turb == 1:
        assert os.path.exists(path + ".turb")
    a.close()
    b.close()


def test_generated_restart_file_loads_the_initial_state(tmp_path):
    """cases.write_flow's file, loaded with wcns_load_flow, gives the case's
    initial interior state bit for bit."""
    from wcns_amd.solver import Solver
    case = cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=4)
    cases.write_flow(case, str(tmp_path / "f.dat"), nout=3)
    a = Solver(case)
    b = Solver(cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=9))
    assert b.load_flow(str(tmp_path / "f.dat"))[0] == 3
    for n in ("r", "u", "v", "w", "p"):
        assert np.array_equal(a.field(0, n)[3:-3, 3:-3, 3:-3], b.field(0, n)[3:-3, 3:-3, 3:-3])
    a.close()
    b.close()


def test_force_history_on_device_matches_wcns_force():
    from wcns_amd.solver import Solver
    case = cases.bump_channel_case(ni=32, nj=24, nk=12, dy_wall=1e-3, growth=1.15, reynolds=1e4)
    s = Solver(case)
    ref = []
    for _ in range(4):
        s.step(1, history=False)
        s.force_record()
        ref.append(s.force())
    hist, cnt = s.force_history(4)
    assert cnt == 4
    assert np.array_equal(hist, np.array(ref))
    s.close()


def parse_plt(buf):
    """Header bytes (through the end-of-header marker), variable names, zones
    (name, imax, jmax, kmax) and per zone (data formats, shared variables,
    shared connectivity, values [variable, node]) of a Tecplot binary file as
    mod_tecios writes it (parallel.f90:1-173)."""
    import struct
    b = bytes(buf)
    assert b[:8] == b"#!TDV102"
    o = [8]

    def take(fmt):
        v = struct.unpack_from(fmt, b, o[0])[0]
        o[0] += 4
        return v

    def text():
        out = []
        c = take("<i")
        while c != 0:
            out.append(chr(c))
            c = take("<i")
        return "".join(out)

    assert take("<i") == 1
    assert text() == "WCNS_SOLVER"
    nv = take("<i")
    names = [text() for _ in range(nv)]
    zones = []
    while True:
        m = take("<f")
        if m == 357.0:
            break
        assert m == 299.0
        zn = text()
        assert [take("<i") for _ in range(5)] == [-1, 0, 0, 0, 0]
        dims = [take("<i") for _ in range(3)]
        assert take("<i") == 0
        zones.append((zn, *dims))
    head = b[:o[0]]
    data = []
    for zn, im, jm, km in zones:
        assert take("<f") == 299.0
        forms = [take("<i") for _ in range(nv)]
        share, conn = take("<i"), take("<i")
        n = im * jm * km
        a = np.frombuffer(b, dtype="<f8", count=nv * n, offset=o[0]).reshape(nv, n)
        o[0] += 8 * nv * n
        data.append((forms, share, conn, a))
    assert o[0] == len(b)
    return head, names, zones, data


def test_tecplot_files_match_reference(tmp_path):
    """The three Tecplot writers of post_processing (tecout_bin_surface, _wall,
    _volume, PostProcess.f90:97-772) against the files the reference's own
    routines wrote for the same seeded two-block SA state after its
    boundary_all and initial_q_dq_c_t (tests/golden/tecplot_golden.npz):
    identical headers, zones and data formats, and every value within 1e-12
    of the reference's (identical in practice).  The reference's wall heat
    flux qw reads a local cp that it never sets (PostProcess.f90:363, 475);
    in the run that wrote the fixture it held 1/((gama-1) moo^2), the value
    the library uses (sub_tgh.f90:868), and the columns agree."""
    import sys
    from wcns_amd.solver import Solver
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_golden_tecplot import tecplot_case, tecplot_state
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tecplot_golden.npz"))
    case = tecplot_case()
    s = Solver(case)
    for nb in range(len(case.blocks)):
        for name, a in zip(("r", "u", "v", "w", "p", "qke", "vist"), tecplot_state(case, nb)):
            s.set_field(nb, name, a)
    s.boundary_all()
    s.initial_q_dq_c_t()
    for kind, name in enumerate(("surface", "wall", "volume")):
        path = str(tmp_path / f"{name}.plt")
        s.tecplot(path, kind)
        mine, ref = parse_plt(open(path, "rb").read()), parse_plt(g[name])
        assert mine[0] == ref[0], name
        worst = 0.0
        for (fm, sm, cm, am), (fr, sr, cr, ar) in zip(mine[3], ref[3]):
            assert fm == fr and sm == sr and cm == cr
            for v, vname in enumerate(ref[1]):
                scale = max(np.max(np.abs(ar[v])), 1e-300)
                worst = max(worst, float(np.max(np.abs(am[v] - ar[v]))) / scale)
        print(name, "zones", len(ref[2]), "variables", ref[1], "max rel diff", worst, flush=True)
        assert worst <= 1e-12, (name, worst)
    s.close()
