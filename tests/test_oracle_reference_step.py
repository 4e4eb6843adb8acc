"""The oracle against the reference's own step routines, compiled from source
in their serial configuration (oracle/_ref/libref_step.so, built by
`make -C oracle ref` where the reference sources are mounted).

Each stage of the pseudo-time step -- boundary_all, initial_q_dq_c_t,
timestep_tgh, r_h_s_vis (with the SA / SST step), r_h_s_invis, l_h_s_tgh
(LU-SGS, SGS, update_limited) and residual_curve -- is run on both from the
same state, and every field the stage writes must agree bit for bit over the
reference's bounds of that field.  The serial reference has no interface
averaging (communicate_dq_npp / communicate_pv_npp exist only in its MPI
build, parallel.f90:3020-3326), so the oracle's averaging is left out of these
sequences; it is pinned by tests/test_oracle_invariants.py.

The wall force (force + force_dif, PostProcess.f90:1415-1680) is compared
after the steps of the viscous cases.

Forced branches: a repaired node and the too-many-repairs stop of
update_limited (Sub2.f90:661-792), the WCNS positivity fallback
(WCNSE5.f90:2224-2245) and the inviscid wall (BC.f90:1412-1486)."""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref_step as R
from wcns_amd import cases

pytestmark = pytest.mark.skipif(not R.available(), reason="reference step library not built (make -C oracle ref)")

G3 = ("r", "u", "v", "w", "p", "t", "vist", "qke", "qkw", "q", "dq")
G1 = ("c", "sra", "srb", "src", "srva", "srvb", "srvc", "visl", "vol", "kcx", "kcy", "kcz", "etx", "ety", "etz",
      "ctx", "cty", "ctz")
N1 = ("dtdt", "dist")

STAGE_FIELDS = {
    "boundary_all": ["r", "u", "v", "w", "p", "vist", "qke", "qkw"],
    "initial_q_dq_c_t": ["c", "t", "q", "dq", "visl"],
    "timestep": ["sra", "srb", "src", "srva", "srvb", "srvc", "dtdt"],
    "rhs_vis": ["dq", "qke", "qkw", "vist"],
    "rhs_invis": ["dq"],
    "lhs": ["dq", "q", "r", "u", "v", "w", "p"],
}


def _window(name, a):
    """The part of a [k+2, j+2, i+2] array inside the reference's bounds."""
    if name in N1:
        return a[3:-3, 3:-3, 3:-3]
    if name in G1:
        return a[1:-2, 1:-2, 1:-2]
    return a


def _fields_for(case, names):
    out = []
    for n in names:
        if n in ("qke", "vist") and case.params.nlamtur <= 0:
            continue
        if n == "qkw" and case.params.nlamtur != 2:
            continue
        if n in ("t", "visl", "srva", "srvb", "srvc") and case.params.nvis != 1:
            continue
        out.append(n)
    return out


def _check(case, o, r, names, tag):
    for nb in range(len(case.blocks)):
        for n in _fields_for(case, names):
            a = _window(n, r.field(nb, n))
            b = _window(n, o.field(nb, n))
            if not np.array_equal(a, b, equal_nan=True):
                d = np.abs(a - b)
                raise AssertionError(f"{tag}: block {nb} field {n} differs, max |diff| {np.nanmax(d):.3e}")


def _pair(case):
    o = O.OracleCase(case)
    dist = [o.field(nb, "dist") for nb in range(len(case.blocks))] if case.params.nlamtur > 0 else None
    r = R.RefStep(case, dist=dist)
    return o, r


def _step(case, o, r, tag, check=True):
    stages = ["boundary_all", "initial_q_dq_c_t", "timestep"] + (["rhs_vis"] if case.params.nvis == 1 else []) + \
        ["rhs_invis", "lhs"]
    for st in stages:
        r.stage(st)
        o.stage(st)
        if check:
            _check(case, o, r, STAGE_FIELDS[st], f"{tag} {st}")
    ro, rr = o.residual(), r.residual()
    assert ro[0] == rr[0] and ro[1] == rr[1] and list(ro[2]) == list(rr[2]), (tag, ro, rr)
    return ro[0]


def test_setup_metrics_images_and_turbulence_inflow():
    """GRID_DERIVATIVE_gcl's metrics, read_bc_connect's image maps and
    inflow_turbulence / init_SA / init_SST equal the oracle's set-up."""
    case = cases.multiblock_case(nx=(12, 14), ny=(11, 13), nz=(10, 9), nvis=1, wall=True, amp_grid=0.05)
    for model in (1, 2):
        case.params.nlamtur = model
        o, r = _pair(case)
        _check(case, o, r, list(G1[8:]) + ["qke", "qkw", "vist"], f"setup model {model}")
        muoo, koo, omgoo, vtoo = r.turb_inflow()
        assert muoo == 0.1 and vtoo == 1.0e-6
        for nb, blk in enumerate(case.blocks):
            for nr, reg in enumerate(blk.regions):
                if reg.bctype >= 0:
                    continue
                box, io, jo, ko = o.region_images(nb, nr)
                ir, jr, kr = r.images(nb, nr)
                assert np.array_equal(io, ir) and np.array_equal(jo, jr) and np.array_equal(ko, kr), (nb, nr)
        o.close()


def test_cartesian_64x32x16_100_steps_bit_exact():
    """The reference's CPU-runnable configuration (64x32x16 nodes, i and k
    periodic, far field in j, Euler, WCNS-E5 scheme 41, Steger-Warming, M 0.8,
    CFL 5, 8 Fourier modes): 100 whole steps, every residual-monitor entry and
    the final state bit for bit; the first step stage by stage."""
    case = cases.cartesian_periodic_case(ni=64, nj=32, nk=16, seed=1)
    o, r = _pair(case)
    hist = [_step(case, o, r, "step 1", check=True)]
    for s in range(99):
        hist.append(_step(case, o, r, f"step {s + 2}", check=False))
    _check(case, o, r, ["r", "u", "v", "w", "p", "q", "dq", "dtdt"], "after 100 steps")
    assert hist[-1] < hist[0]
    o.close()


@pytest.mark.parametrize("nflux,nscheme", [(3, 41), (1, 4)])
def test_viscous_bump_steps(nflux, nscheme):
    """Laminar bump channel (no-slip adiabatic wall BC 2, far field BC 4,
    periodic span): 5 steps stage by stage."""
    case = cases.bump_channel_case(ni=24, nj=40, nk=8, dy_wall=1e-3, growth=1.15, nflux=nflux, nscheme=nscheme)
    o, r = _pair(case)
    for s in range(5):
        _step(case, o, r, f"step {s + 1}")
    # force + force_dif (PostProcess.f90:1415-1680) on the state the BCs leave
    r.stage("boundary_all")
    o.stage("boundary_all")
    assert np.array_equal(o.force(), r.force())
    o.close()


@pytest.mark.parametrize("model", [1, 2], ids=["SA", "SST"])
def test_turbulent_multiblock_mixed_bc(model):
    """2x2 blocks with a no-slip wall, symmetry (BC 3) on top, fixed inflow
    (BC 5) and outflow (BC 6), 1:1 interfaces and a periodic span; SA or SST:
    4 steps stage by stage."""
    case = cases.multiblock_case(nvis=1, wall=True, periodic_i=False, ilo_bc=5, ihi_bc=6, jhi_bc=3, reynolds=1e5)
    case.params.nlamtur = model
    case.params.nwholefield = 0
    o, r = _pair(case)
    for s in range(4):
        _step(case, o, r, f"step {s + 1}")
    r.stage("boundary_all")
    o.stage("boundary_all")
    assert np.array_equal(o.force(), r.force())
    o.close()


def test_inviscid_wall_and_supersonic():
    """Euler with a slip wall (BC 2, nvis = 0: inviscid_wall) at M 2, Roe."""
    case = cases.multiblock_case(nx=(14, 12), ny=(12,), nz=(10,), nvis=0, wall=True, moo=2.0, nflux=3)
    o, r = _pair(case)
    for s in range(4):
        _step(case, o, r, f"step {s + 1}")
    o.close()


def _poke(case, o, r, name, nodes, value):
    for nb, i, j, k in nodes:
        for side in (o, r):
            a = side.field(nb, name)
            a[k + 2, j + 2, i + 2] = value
            side.set_field(nb, name, a)


def test_forced_repair_and_wcns_fallback():
    """A few interior nodes with density above rmax_limit (1e3) are repaired by
    update_limited's neighbour average, and near-vacuum pressure pockets send
    the WCNS interpolation to its positivity fallback; both agree bit for bit."""
    case = cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=3)
    o, r = _pair(case)
    _poke(case, o, r, "r", [(0, 6, 6, 6), (0, 15, 9, 4)], 2.0e3)
    _poke(case, o, r, "p", [(0, 10, 4, 8), (0, 11, 4, 8)], 1.0e-7)
    n0 = o.counter(1)
    _step(case, o, r, "forced step")
    assert o.counter(1) - n0 >= 2, "no node was repaired"
    assert o.counter(0) > 0, "the WCNS fallback was not reached"
    _step(case, o, r, "next step")
    o.close()


def test_forced_abort_stops_both():
    """More than n/64 abnormal nodes: the reference's update_limited stops the
    run (Sub2.f90:785-790) and the oracle returns its error code.  The
    reference runs in a child process, since its STOP ends the process."""
    script = textwrap.dedent("""
        import numpy as np
        from oracle import ref_step as R
        from wcns_amd import cases
        case = cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=3)
        r = R.RefStep(case)
        a = r.field(0, "r")
        a[4:9, 4:11, 3:12] = 2.0e3
        r.set_field(0, "r", a)
        for st in ("boundary_all", "initial_q_dq_c_t", "timestep", "rhs_invis", "lhs"):
            r.stage(st)
        print("LHS RETURNED")
    """)
    env = dict(os.environ, PYTHONPATH=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    p = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, env=env, timeout=120)
    assert "LHS RETURNED" not in p.stdout, p.stdout
    case = cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=3)
    o = O.OracleCase(case)
    a = o.field(0, "r")
    a[4:9, 4:11, 3:12] = 2.0e3
    o.set_field(0, "r", a)
    for st in ("boundary_all", "initial_q_dq_c_t", "timestep", "rhs_invis"):
        o.stage(st)
    assert O.lib().o_lhs(o.h) == -3
    assert o.counter(1) > (24 * 16 * 12) // 64
    o.close()


def _char_cases(nscheme, nflux):
    """A periodic Cartesian box (interfaces to itself in i and k), two blocks
    with a slip wall at M 2, and four viscous blocks with a no-slip wall,
    symmetry top and inflow / outflow."""
    a = cases.cartesian_periodic_case(ni=20, nj=14, nk=12, seed=3, nscheme=nscheme, nflux=nflux)
    b = cases.multiblock_case(nx=(14, 12), ny=(12,), nz=(10,), nvis=0, wall=True, moo=2.0, nflux=nflux,
                              nscheme=nscheme)
    c = cases.multiblock_case(nvis=1, wall=True, periodic_i=False, ilo_bc=5, ihi_bc=6, jhi_bc=3, reynolds=1e5,
                              nscheme=nscheme, nflux=nflux)
    return {"periodic": a, "wall-M2": b, "viscous-4-blocks": c}


@pytest.mark.parametrize("nscheme,nflux", [(45, 1), (45, 2), (45, 3), (47, 1), (47, 2), (47, 3)])
def test_characteristic_schemes_against_reference(nscheme, nflux):
    """WCNS_E_5_45 (characteristic jumps through the Roe-averaged face
    eigenvectors, WCNSE5.f90:3442-4155) and WCNS_E_5_47 (characteristic
    variables of each node's own state, WCNSE5.f90:4158-4854), whose pencils
    read four ghost metrics at each end from the interface partner
    (Convect.f90:180-297): two steps stage by stage on periodic, multi-block,
    wall and viscous cases."""
    for tag, case in _char_cases(nscheme, nflux).items():
        o, r = _pair(case)
        for s in range(2):
            _step(case, o, r, f"{nscheme}/{nflux} {tag} step {s + 1}")
        o.close()


def _step_poisoned(case, o, r, tag, value):
    for st in ["boundary_all", "initial_q_dq_c_t", "timestep"] + (["rhs_vis"] if case.params.nvis == 1 else []) + \
            ["rhs_invis", "lhs"]:
        if st == "rhs_invis":
            r.poison(0, value)
        r.stage(st)
        o.stage(st)
        _check(case, o, r, STAGE_FIELDS[st], f"{tag} {st}")


@pytest.mark.parametrize("nscheme", [42, 47])
def test_roemx_flux_against_reference(nscheme):
    """flux_roemx (flux.f90:246-393) with the minimum pressure ratio around
    each face (Convect.f90:78-177).  At the faces 0 the reference reads
    pressure-ratio entries its loops never write; with the heap block those
    entries come from filled with 1e300 beforehand (ref_step_poison), its
    steps equal the oracle's bit for bit, the ratios that exist deciding the
    minimum as in the oracle.  Filled with 0, the results differ: the reads
    are reached."""
    case = cases.cartesian_periodic_case(ni=20, nj=14, nk=12, seed=3, nscheme=nscheme, nflux=4)
    o, r = _pair(case)
    for s in range(2):
        _step_poisoned(case, o, r, f"roemx {nscheme} step {s + 1}", 1.0e300)
    o.close()
    o, r = _pair(case)
    for st in ["boundary_all", "initial_q_dq_c_t", "timestep"]:
        r.stage(st)
        o.stage(st)
    r.poison(0, 0.0)
    r.stage("rhs_invis")
    o.stage("rhs_invis")
    assert not np.array_equal(r.field(0, "dq"), o.field(0, "dq"))
    o.close()
    # equal-sized blocks reuse the block the previous one freed
    case = cases.multiblock_case(nx=(12, 12), ny=(12, 12), nz=(10,), nvis=1, wall=True, periodic_i=False, ilo_bc=5,
                                 ihi_bc=6, jhi_bc=3, reynolds=1e5, nscheme=nscheme, nflux=4)
    o, r = _pair(case)
    for s in range(2):
        _step_poisoned(case, o, r, f"roemx {nscheme} viscous step {s + 1}", 1.0e300)
    o.close()
