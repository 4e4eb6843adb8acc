"""GPU library (through its C-ABI) against the CPU oracle on identical seeded
inputs: stage by stage, then whole steps.

Tolerances: FP64 path, relative 1e-12 against the field's scale (the stated
bar); stages that involve no transcendental function are expected to be
bit-identical and are checked with 0 ulp where noted.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from wcns_amd import cases

pytestmark = pytest.mark.gpu

REL = 1e-12


def _solver(case):
    from wcns_amd.solver import Solver
    return Solver(case)


def _rel_err(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def _interior(a, blk, ghost=0):
    g = 3 - ghost
    return a[g:a.shape[0] - g, g:a.shape[1] - g, g:a.shape[2] - g]


def _cmp_prim(s, o, blk, tol, ghost=3):
    for name in ("r", "u", "v", "w", "p"):
        a = _interior(s.field(0, name), blk, ghost)
        b = _interior(o.field(0, name), blk, ghost)
        e = _rel_err(a, b)
        assert e <= tol, f"{name}: rel err {e:.3e}"


CASES = [
    dict(ni=24, nj=16, nk=12, nscheme=41, nflux=1),
    dict(ni=24, nj=16, nk=12, nscheme=4, nflux=1),
    dict(ni=20, nj=14, nk=13, nscheme=41, nflux=3),
    dict(ni=16, nj=20, nk=12, nscheme=4, nflux=3),
    # van Leer (nflux 2) and WCNS_E_5_42 (nscheme 42), Convect.f90:20-37
    dict(ni=24, nj=16, nk=12, nscheme=41, nflux=2),
    dict(ni=20, nj=14, nk=13, nscheme=42, nflux=1),
    dict(ni=24, nj=16, nk=12, nscheme=42, nflux=2),
    dict(ni=16, nj=20, nk=12, nscheme=42, nflux=3),
    dict(ni=16, nj=20, nk=12, nscheme=4, nflux=2),
]


@pytest.mark.parametrize("kw", CASES, ids=lambda k: f"{k['ni']}x{k['nj']}x{k['nk']}-s{k['nscheme']}-f{k['nflux']}")
def test_stage_by_stage(kw):
    case = cases.cartesian_periodic_case(seed=3, amp=0.02, **kw)
    blk = case.blocks[0]
    print(f"case {kw}", flush=True)
    s = _solver(case)
    o = O.OracleCase(case)
    # metrics (GCL, setup)
    for m in ("kcx", "kcy", "kcz", "etx", "ety", "etz", "ctx", "cty", "ctz", "vol"):
        e = _rel_err(_interior(s.field(0, m), blk), _interior(o.field(0, m), blk))
        assert e == 0.0, f"metric {m} differs: {e:.3e}"
    for step in range(3):
        s.boundary_all(); o.stage("boundary_all")
        _cmp_prim(s, o, blk, REL)
        s.initial_q_dq_c_t(); o.stage("initial_q_dq_c_t")
        s.timestep_tgh(); o.stage("timestep")
        for m in ("sra", "srb", "src", "dtdt"):
            e = _rel_err(_interior(s.field(0, m), blk), _interior(o.field(0, m), blk))
            assert e <= REL, f"{m}: {e:.3e}"
        s.r_h_s_invis(); o.stage("rhs_invis")
        rg = _interior(s.field(0, "rhs"), blk)
        ro = _interior(o.field(0, "dq"), blk)
        assert _rel_err(rg, ro) <= REL, f"rhs step {step}: {_rel_err(rg, ro):.3e}"
        s.communicate_dq_npp(); o.stage("communicate_dq")
        rg = _interior(s.field(0, "rhs"), blk)
        ro = _interior(o.field(0, "dq"), blk)
        assert _rel_err(rg, ro) <= REL, f"rhs after averaging step {step}: {_rel_err(rg, ro):.3e}"
        s.l_h_s_tgh(); o.stage("lhs")
        dg = _interior(s.field(0, "dq"), blk)
        do = _interior(o.field(0, "dq"), blk)
        assert _rel_err(dg, do) <= REL, f"dq step {step}: {_rel_err(dg, do):.3e}"
        s.communicate_pv_npp(); o.stage("communicate_pv")
        _cmp_prim(s, o, blk, REL, ghost=0)
        rs = s.residual_curve()
        ro_ = o.residual()
        assert abs(rs[0] - ro_[0]) <= 1e-12 * ro_[0]
        assert abs(rs[1] - ro_[1]) <= 1e-12 * ro_[1]
        assert list(rs[2]) == list(ro_[2])


def test_cartesian_64x32x16_100_steps():
    """The reference's CPU-runnable configuration: 64x32x16 nodes, i and k
    periodic, far field in j, Euler, WCNS-E5 (scheme 41) + Steger-Warming,
    M 0.8, CFL 5, 8 Fourier modes of 1 % on rho and p, 100 LU-SGS steps."""
    case = cases.cartesian_periodic_case(ni=64, nj=32, nk=16, seed=1)
    blk = case.blocks[0]
    s = _solver(case)
    o = O.OracleCase(case)
    hs = s.step(100)
    ho = o.step(100)
    rel = np.abs(hs - ho) / ho
    print("residual history rel err max", rel.max(), "drop", ho[0] / ho[-1], flush=True)
    assert rel.max() <= 1e-10
    _cmp_prim(s, o, blk, 1e-11, ghost=0)
    assert s.counter(0) == o.counter(0)
    assert s.counter(1) == o.counter(1)


VIS_CASES = [
    dict(ni=24, nj=16, nk=12, nflux=3, nscheme=41),
    dict(ni=20, nj=18, nk=13, nflux=1, nscheme=4),
    dict(ni=24, nj=16, nk=12, nflux=2, nscheme=42),
]


@pytest.mark.parametrize("kw", VIS_CASES, ids=lambda k: f"{k['ni']}x{k['nj']}x{k['nk']}-s{k['nscheme']}-f{k['nflux']}")
def test_viscous_stage_by_stage(kw):
    """Laminar bump channel: adiabatic no-slip wall (BC 2), far field (BC 4),
    periodic span -- viscous residual, viscous spectral radii, wall ghosts."""
    case = cases.bump_channel_case(dy_wall=1e-3, growth=1.2, reynolds=1e3, **kw)
    blk = case.blocks[0]
    print(f"viscous case {kw}", flush=True)
    s = _solver(case)
    o = O.OracleCase(case)
    for step in range(3):
        s.boundary_all(); o.stage("boundary_all")
        _cmp_prim(s, o, blk, REL)
        s.initial_q_dq_c_t(); o.stage("initial_q_dq_c_t")
        for m in ("t", "visl"):
            e = _rel_err(_interior(s.field(0, m), blk, 1 if m == "t" else 0),
                         _interior(o.field(0, m), blk, 1 if m == "t" else 0))
            assert e <= REL, f"{m}: {e:.3e}"
        s.timestep_tgh(); o.stage("timestep")
        for m in ("srva", "srvb", "srvc", "dtdt"):
            e = _rel_err(_interior(s.field(0, m), blk), _interior(o.field(0, m), blk))
            assert e <= REL, f"{m}: {e:.3e}"
        s.r_h_s_vis(); o.stage("rhs_vis")
        rg = _interior(s.field(0, "rhs"), blk)
        ro = _interior(o.field(0, "dq"), blk)
        assert _rel_err(rg, ro) <= REL, f"viscous rhs step {step}: {_rel_err(rg, ro):.3e}"
        s.communicate_dq_npp(); o.stage("communicate_dq")
        s.r_h_s_invis(); o.stage("rhs_invis")
        s.communicate_dq_npp(); o.stage("communicate_dq")
        rg = _interior(s.field(0, "rhs"), blk)
        ro = _interior(o.field(0, "dq"), blk)
        assert _rel_err(rg, ro) <= REL, f"rhs step {step}: {_rel_err(rg, ro):.3e}"
        s.l_h_s_tgh(); o.stage("lhs")
        dg = _interior(s.field(0, "dq"), blk)
        do = _interior(o.field(0, "dq"), blk)
        assert _rel_err(dg, do) <= REL, f"dq step {step}: {_rel_err(dg, do):.3e}"
        s.communicate_pv_npp(); o.stage("communicate_pv")
        _cmp_prim(s, o, blk, REL, ghost=0)


def test_viscous_bump_30_steps():
    case = cases.bump_channel_case(ni=32, nj=24, nk=12, dy_wall=1e-3, growth=1.15, reynolds=1e4)
    blk = case.blocks[0]
    s = _solver(case)
    o = O.OracleCase(case)
    hs = s.step(30)
    ho = o.step(30)
    rel = np.abs(hs - ho) / ho
    print("viscous residual history rel err max", rel.max(), flush=True)
    assert rel.max() <= 1e-10
    _cmp_prim(s, o, blk, 1e-11, ghost=0)


def test_tiled_sweeps_match_plane_sweeps(monkeypatch):
    """The pipelined tile sweep and one-launch-per-hyperplane sweep give identical bits."""
    case = cases.wavy_periodic_case(ni=40, nj=37, nk=13, nflux=3)
    monkeypatch.setenv("WCNS_SWEEP", "plane")
    a = _solver(case)
    monkeypatch.setenv("WCNS_SWEEP", "tiled")
    b = _solver(case)
    ha = a.step(4)
    hb = b.step(4)
    assert np.array_equal(ha, hb)
    for name in ("r", "u", "v", "w", "p"):
        assert np.array_equal(a.field(0, name), b.field(0, name))


@pytest.mark.parametrize("lhs", [0, 1], ids=["fp64", "fp32"])
def test_paired_sgs_helpers_match_unpaired(monkeypatch, lhs):
    """The SGS sweeps with three helper wavefronts forming two operand parts
    each give the bits of six helpers forming one, on an uneven multi-block SA
    case, in both implicit-operator precisions."""
    from wcns_amd.solver import Solver
    case = _uneven_sa()
    monkeypatch.setenv("WCNS_SGS_PAIR", "0")
    a = Solver(case, lhs_precision=lhs)
    monkeypatch.setenv("WCNS_SGS_PAIR", "1")
    b = Solver(case, lhs_precision=lhs)
    assert not a.sweep_layout()[2] and b.sweep_layout()[2]
    ha = a.step(4)
    hb = b.step(4)
    assert np.array_equal(ha, hb)
    for nb in range(len(case.blocks)):
        for name in ("r", "u", "v", "w", "p"):
            assert np.array_equal(a.field(nb, name), b.field(nb, name)), f"block {nb} {name}"


@pytest.mark.parametrize("model", [1, 2], ids=["sa", "sst"])
def test_staged_turbulence_sweeps_match_field_pool_sweeps(monkeypatch, model):
    """The scalar turbulence sweeps on operands staged plane-major give the
    bits of the same sweeps reading the field pools, over four steps of an
    uneven multi-block case with SA and with SST (two variables)."""
    case = _uneven_sa()
    case.params.nlamtur = model
    monkeypatch.setenv("WCNS_SA_STAGED", "0")
    a = _solver(case)
    monkeypatch.setenv("WCNS_SA_STAGED", "1")
    b = _solver(case)
    monkeypatch.delenv("WCNS_SA_STAGED")
    ha = a.step(4)
    hb = b.step(4)
    assert np.array_equal(ha, hb)
    for nb in range(len(case.blocks)):
        for name in ("r", "u", "v", "w", "p", "qke", "vist") + (("qkw",) if model == 2 else ()):
            assert np.array_equal(a.field(nb, name), b.field(nb, name)), f"block {nb} {name}"


def test_inviscid_register_budgets_match(monkeypatch):
    """The inviscid walk under the two-wave register budget gives the bits of
    the one-wave budget (each forced with WCNS_INV_WAVES), over four steps of
    an uneven multi-block SA case with SW and with Roe fluxes."""
    for nflux in (1, 3):
        case = _uneven_sa()
        case.params.nflux = nflux
        monkeypatch.setenv("WCNS_INV_WAVES", "1")
        a = _solver(case)
        monkeypatch.setenv("WCNS_INV_WAVES", "2")
        b = _solver(case)
        monkeypatch.delenv("WCNS_INV_WAVES")
        ha = a.step(4)
        hb = b.step(4)
        assert np.array_equal(ha, hb)
        for nb in range(len(case.blocks)):
            for name in ("r", "u", "v", "w", "p"):
                assert np.array_equal(a.field(nb, name), b.field(nb, name)), f"flux {nflux} block {nb} {name}"


def test_inviscid_stage_sees_primitives_written_after_viscous_stage():
    """The inviscid stage reuses the viscous stage's i<->j copies of the
    primitives only while nothing has written them: a pressure changed with
    set_field between the two stages must reach the inviscid residual.  Run A
    (viscous, set_field, inviscid) against run B (set_field, inviscid), both
    from a zeroed residual: identical bits."""
    case = cases.bump_channel_case(ni=32, nj=24, nk=12, dy_wall=1e-3, growth=1.15, reynolds=1e4)
    a = _solver(case)
    b = _solver(case)
    for s in (a, b):
        s.boundary_all()
        s.initial_q_dq_c_t()
        s.timestep_tgh()
    p = a.field(0, "p")
    p2 = p * (1.0 + 0.01 * np.sin(np.arange(p.size).reshape(p.shape) * 0.37))
    a.r_h_s_vis()                       # copies the primitives for the i direction
    a.set_field(0, "p", p2)
    a.initial_q_dq_c_t()                # R = 0 again
    a.r_h_s_invis()
    b.set_field(0, "p", p2)
    b.initial_q_dq_c_t()
    b.r_h_s_invis()
    ra, rb = a.field(0, "rhs")[3:-3, 3:-3, 3:-3], b.field(0, "rhs")[3:-3, 3:-3, 3:-3]
    assert np.array_equal(ra, rb)
    assert np.any(ra != 0.0)


def test_mixed_precision_lhs_converges_like_fp64():
    """FP32 implicit operator (FP64 residual, weights and update): same residual
    drop within a few steps of the FP64 operator on the same case."""
    from wcns_amd.solver import Solver
    case = cases.wavy_periodic_case(ni=32, nj=24, nk=12)
    a = Solver(case)
    b = Solver(case, lhs_precision=1)
    ha = a.step(150)
    hb = b.step(150)
    print("fp64 drop", ha[0] / ha[-1], "fp32-lhs drop", hb[0] / hb[-1], flush=True)
    assert np.all(np.isfinite(hb))
    assert hb[-1] / hb[0] < 3.0 * ha[-1] / ha[0]
    pa = a.field(0, "p")[3:-3, 3:-3, 3:-3]
    pb = b.field(0, "p")[3:-3, 3:-3, 3:-3]
    print("pressure difference", np.max(np.abs(pa - pb)) / np.max(np.abs(pa)), flush=True)


def test_wall_force_matches_oracle():
    """Force and moment coefficients on the bump wall after 5 viscous steps
    (pressure plus Stokes stress, four-triangle quadrature) within 1e-10."""
    case = cases.bump_channel_case(ni=32, nj=24, nk=12, dy_wall=1e-3, growth=1.15, reynolds=1e4)
    case.params.xref, case.params.yref = 0.25, 0.1
    case.params.sref, case.params.lfref, case.params.attack = 0.5, 1.0, 2.0
    s = _solver(case)
    o = O.OracleCase(case)
    s.step(5)
    o.step(5)
    fs = s.force()
    fo = o.force()
    print("force gpu", fs, "oracle", fo, flush=True)
    scale = np.max(np.abs(fo[:8]))
    assert np.max(np.abs(fs[:8] - fo[:8])) <= 1e-10 * scale
    assert abs(fs[8] - fo[8]) <= 1e-10 * abs(fo[8])


def _steps_to_drop(s, orders, chunk=100, limit=4000):
    """Steps the solver until its residual monitor has dropped `orders` decades
    from its first value (checked every `chunk` steps); returns the steps."""
    h0, n = None, 0
    while n < limit:
        h = s.step(chunk)
        n += chunk
        h0 = h[0] if h0 is None else h0
        if np.log10(h0 / h[-1]) >= orders:
            return n, np.log10(h0 / h[-1])
    raise AssertionError(f"no {orders}-order drop in {limit} steps")


def test_converged_wall_force_matches_oracle():
    """Converged state: the steady Euler bump channel (48x24x12, slip top,
    CFL 20) run to an 8-order residual drop.  The FP64 path's wall-force
    coefficients match the oracle's after the same steps within 1e-10; the
    solver with the FP32 implicit operator reaches the same 8-order drop and
    its coefficients match within 1e-8."""
    from wcns_amd.solver import Solver
    case = cases.bump_channel_case(ni=48, nj=24, nk=12, dy_wall=0.02, growth=1.1, top_bc=3, cfl=20.0)
    case.params.nvis = 0
    s = Solver(case)
    n64, d64 = _steps_to_drop(s, 8.0)
    o = O.OracleCase(case)
    o.set_threads(min(16, os.cpu_count() or 1))
    ho = o.step(n64)
    fs, fo = s.force(), o.force()
    scale = np.max(np.abs(fo[:8]))
    err64 = np.max(np.abs(fs[:8] - fo[:8])) / scale
    b = Solver(case, lhs_precision=1)
    n32, d32 = _steps_to_drop(b, 8.0)
    fb = b.force()
    err32 = np.max(np.abs(fb[:8] - fo[:8])) / scale
    print(f"fp64: {d64:.2f} orders in {n64} steps, oracle {np.log10(ho[0] / ho[-1]):.2f}; force rel err {err64:.2e}"
          f" | fp32 operator: {d32:.2f} orders in {n32} steps, force rel err {err32:.2e}", flush=True)
    assert err64 <= 1e-10 and abs(fs[8] - fo[8]) <= 1e-10 * abs(fo[8])
    assert err32 <= 1e-8 and abs(fb[8] - fo[8]) <= 1e-8 * abs(fo[8])
    o.close()
    b.close()
    s.close()


def test_pipelined_sweeps_match_plane_sweeps_full_size(monkeypatch):
    """At the benchmark size (256x128x64 bump channel, 8x16 tiles of 16x4
    columns) the pipelined sweeps and the per-hyperplane sweeps agree bit for
    bit over two whole steps."""
    case = cases.bump_channel_case(ni=256, nj=128, nk=64)
    monkeypatch.setenv("WCNS_SWEEP", "plane")
    a = _solver(case)
    monkeypatch.setenv("WCNS_SWEEP", "tiled")
    b = _solver(case)
    ha = a.step(2)
    hb = b.step(2)
    assert np.array_equal(ha, hb)
    for name in ("r", "u", "v", "w", "p"):
        assert np.array_equal(a.field(0, name), b.field(0, name))


def _cmp_all_blocks(s, o, case, names, tol, ghost=0):
    for nb, blk in enumerate(case.blocks):
        for name in names:
            a = _interior(s.field(nb, name), blk, ghost)
            b = _interior(o.field(nb, name), blk, ghost)
            e = _rel_err(a, b)
            assert e <= tol, f"block {nb} {name}: rel err {e:.3e}"


MB_CASES = [
    dict(nx=(14, 18), ny=(13, 15), nz=(12,), nvis=0, nflux=3),
    dict(nx=(14, 16), ny=(13, 14), nz=(12, 13), nvis=1, wall=True, nflux=1, nscheme=4),
]


@pytest.mark.parametrize("kw", MB_CASES, ids=["2x2x1-euler-roe", "2x2x2-wall-sw"])
def test_multiblock_stage_by_stage(kw):
    """Uneven blocks with 1:1 interfaces in every direction (plus periodic
    faces and, in the second case, a no-slip wall): every stage of three
    steps against the oracle, all blocks."""
    case = cases.multiblock_case(**kw)
    print(f"multi-block case {kw}", flush=True)
    _stages_all_blocks(case, 3)


def _stages_all_blocks(case, nsteps):
    s = _solver(case)
    o = O.OracleCase(case)
    prim = ("r", "u", "v", "w", "p")
    kw = {"nvis": case.params.nvis}
    for step in range(nsteps):
        s.boundary_all(); o.stage("boundary_all")
        _cmp_all_blocks(s, o, case, prim, REL, ghost=3)
        s.initial_q_dq_c_t(); o.stage("initial_q_dq_c_t")
        s.timestep_tgh(); o.stage("timestep")
        _cmp_all_blocks(s, o, case, ("sra", "srb", "src", "dtdt"), REL)
        if kw["nvis"]:
            s.r_h_s_vis(); o.stage("rhs_vis")
            s.communicate_dq_npp(); o.stage("communicate_dq")
        s.r_h_s_invis(); o.stage("rhs_invis")
        s.communicate_dq_npp(); o.stage("communicate_dq")
        for nb, blk in enumerate(case.blocks):
            e = _rel_err(_interior(s.field(nb, "rhs"), blk), _interior(o.field(nb, "dq"), blk))
            assert e <= REL, f"block {nb} rhs step {step}: {e:.3e}"
        s.l_h_s_tgh(); o.stage("lhs")
        _cmp_all_blocks(s, o, case, ("dq",), REL)
        s.communicate_pv_npp(); o.stage("communicate_pv")
        _cmp_all_blocks(s, o, case, prim, REL)
        rs = s.residual_curve()
        ro = o.residual()
        assert abs(rs[0] - ro[0]) <= 1e-12 * ro[0]
        assert list(rs[2]) == list(ro[2])


VARIANT_CASES = [
    ("periodic", 45, 1), ("periodic", 47, 3), ("periodic", 42, 4), ("periodic", 47, 4),
    ("blocks", 45, 3), ("blocks", 47, 2), ("blocks", 47, 4),
    ("wall", 45, 1), ("wall", 47, 3), ("wall", 42, 4),
    ("slip-M2", 45, 3), ("slip-M2", 47, 1),
]


def _variant_case(kind, nscheme, nflux):
    """The flow is oblique to every grid direction (attack 12 deg, sideslip
    8 deg).  With a velocity component that is exactly zero, the faces normal
    to it carry a normal Mach number of round-off size after the
    characteristic round trip of scheme 47, and RoeMx's mia**(1 - rpmin)
    (flux.f90:345) turns last-bit differences of its inputs (the GPU's pow in
    the far-field ghosts, BC.f90:460-607) into relative differences of the
    order of 1e-4 at such faces: the formula is ill-conditioned there, on the
    reference as here."""
    if kind == "periodic":
        c = cases.cartesian_periodic_case(ni=20, nj=14, nk=13, seed=3, amp=0.02, nscheme=nscheme, nflux=nflux)
    elif kind == "blocks":
        c = cases.multiblock_case(nx=(14, 18), ny=(13, 15), nz=(12,), nvis=0, nflux=nflux, nscheme=nscheme)
    elif kind == "wall":
        c = cases.multiblock_case(nx=(14, 16), ny=(13, 14), nz=(12, 13), nvis=1, wall=True, nflux=nflux,
                                  nscheme=nscheme)
    else:
        c = cases.multiblock_case(nx=(14, 12), ny=(12,), nz=(12,), nvis=0, wall=True, moo=2.0, nflux=nflux,
                                  nscheme=nscheme)
    c.params.attack, c.params.sideslip = 12.0, 8.0
    return c


@pytest.mark.parametrize("kind,nscheme,nflux", VARIANT_CASES, ids=lambda v: str(v))
def test_characteristic_schemes_and_roemx(kind, nscheme, nflux):
    """WCNS_E_5_45 / _47 (characteristic interpolation, ghost metrics from the
    interface partners, Convect.f90:180-297) and the RoeMx flux (nflux 4,
    pressure-ratio sensor, flux.f90:246-393): every stage of two steps
    against the oracle on the periodic box, uneven blocks, a no-slip wall and
    a slip wall at M 2, then five whole steps."""
    case = _variant_case(kind, nscheme, nflux)
    print(f"variant {kind} scheme {nscheme} flux {nflux}", flush=True)
    _stages_all_blocks(case, 2)
    s = _solver(case)
    o = O.OracleCase(case)
    hs, ho = s.step(5), o.step(5)
    assert np.max(np.abs(hs - ho) / ho) <= 1e-10
    _cmp_all_blocks(s, o, case, ("r", "u", "v", "w", "p"), 1e-11)
    assert s.counter(0) == o.counter(0)
    s.close()
    o.close()


def test_variant_parameters_rejected():
    """RoeMx only from schemes 42 and 47 (invcode, Convect.f90:34, 47); the
    tangent and adjoint cover schemes 4, 41 and 42 with fluxes 1-3."""
    from wcns_amd.solver import Solver, WcnsError
    with pytest.raises(WcnsError, match="nflux 4"):
        Solver(cases.cartesian_periodic_case(ni=20, nj=14, nk=13, nscheme=41, nflux=4))
    s = Solver(cases.cartesian_periodic_case(ni=20, nj=14, nk=13, nscheme=47, nflux=1))
    # rejected before any array is read (null addresses)
    with pytest.raises(WcnsError, match="tangent and adjoint"):
        s.adj_residual(0, None, None, None)
    with pytest.raises(WcnsError, match="tangent and adjoint"):
        s.adj_step_jvp(None, None)
    s.close()


def test_multiblock_40_steps():
    case = cases.multiblock_case(nx=(14, 17, 12), ny=(13, 15), nz=(12, 14), nvis=1, wall=True)
    s = _solver(case)
    o = O.OracleCase(case)
    hs = s.step(40)
    ho = o.step(40)
    rel = np.abs(hs - ho) / ho
    print("multi-block residual history rel err max", rel.max(), flush=True)
    assert rel.max() <= 1e-10
    _cmp_all_blocks(s, o, case, ("r", "u", "v", "w", "p"), 1e-11)


def _uneven_sa():
    case = cases.uneven_blocks_case(nblk=(3, 2, 2), lo=13, hi=22, seed=3, nvis=1, wall=True)
    case.params.nlamtur = 1
    return case


@pytest.mark.parametrize("make", ["uneven-sa", "mixed-bc-sa"])
def test_surface_levels_match_region_sequence(monkeypatch, make):
    """Boundary conditions and interface averaging launched as dependency
    levels of all regions give the bits of one launch per region in the
    reference's order (WCNS_SURFACE=sequence), over five whole steps."""
    case = _uneven_sa() if make == "uneven-sa" else _sa_mixed_bc()
    print(f"surface levels vs sequence: {make}, {len(case.blocks)} blocks", flush=True)
    monkeypatch.setenv("WCNS_SURFACE", "sequence")
    a = _solver(case)
    monkeypatch.delenv("WCNS_SURFACE")
    b = _solver(case)
    ha = a.step(5)
    hb = b.step(5)
    assert np.array_equal(ha, hb)
    for nb in range(len(case.blocks)):
        for name in ("r", "u", "v", "w", "p", "qke", "vist"):
            assert np.array_equal(a.field(nb, name), b.field(nb, name)), f"block {nb} {name}"


# ---------------------------------------------------------------- Spalart-Allmaras
def _sa_bump(**kw):
    case = cases.bump_channel_case(dy_wall=1e-3, growth=1.2, reynolds=1e5, **kw)
    case.params.nlamtur = 1
    return case


def _sa_mixed_bc():
    """Two x two blocks, no-slip wall at j = 1, symmetry on top, fixed
    freestream inflow (BC 5) and outflow (BC 6) on the i faces, k periodic."""
    case = cases.multiblock_case(nvis=1, wall=True, periodic_i=False, ilo_bc=5, ihi_bc=6, jhi_bc=3)
    case.params.nlamtur = 1
    return case


SA_CASES = [dict(ni=24, nj=16, nk=12, nflux=3, nscheme=41), dict(ni=20, nj=18, nk=13, nflux=1, nscheme=4)]


@pytest.mark.parametrize("kw", SA_CASES, ids=lambda k: f"{k['ni']}x{k['nj']}x{k['nk']}-s{k['nscheme']}-f{k['nflux']}")
def test_sa_stage_by_stage(kw):
    """SA bump channel: wall distance and initial field at set-up; each step the
    SA ghosts (wall, far field, periodic), the radii with mu_t, the SA step
    inside r_h_s_vis (nu~ and mu_t after it) and the viscous residual with mu_t."""
    case = _sa_bump(**kw)
    blk = case.blocks[0]
    print(f"SA case {kw}", flush=True)
    s = _solver(case)
    o = O.OracleCase(case)
    e = _rel_err(_interior(s.field(0, "dist"), blk), _interior(o.field(0, "dist"), blk))
    assert e <= REL, f"wall distance: {e:.3e}"
    for name in ("qke", "vist"):
        e = _rel_err(_interior(s.field(0, name), blk, 1), _interior(o.field(0, name), blk, 1))
        assert e == 0.0, f"initial {name}: {e:.3e}"
    for step in range(3):
        s.boundary_all(); o.stage("boundary_all")
        _cmp_prim(s, o, blk, REL)
        for name in ("qke", "vist"):
            e = _rel_err(_interior(s.field(0, name), blk, 1), _interior(o.field(0, name), blk, 1))
            assert e <= REL, f"{name} after BCs, step {step}: {e:.3e}"
        s.initial_q_dq_c_t(); o.stage("initial_q_dq_c_t")
        s.timestep_tgh(); o.stage("timestep")
        for m in ("srva", "srvb", "srvc", "dtdt"):
            e = _rel_err(_interior(s.field(0, m), blk), _interior(o.field(0, m), blk))
            assert e <= REL, f"{m}: {e:.3e}"
        s.r_h_s_vis(); o.stage("rhs_vis")
        for name in ("qke", "vist"):
            e = _rel_err(_interior(s.field(0, name), blk, 1), _interior(o.field(0, name), blk, 1))
            assert e <= REL, f"{name} after the SA step, step {step}: {e:.3e}"
        rg = _interior(s.field(0, "rhs"), blk)
        ro = _interior(o.field(0, "dq"), blk)
        assert _rel_err(rg, ro) <= REL, f"viscous rhs step {step}: {_rel_err(rg, ro):.3e}"
        s.communicate_dq_npp(); o.stage("communicate_dq")
        s.r_h_s_invis(); o.stage("rhs_invis")
        s.communicate_dq_npp(); o.stage("communicate_dq")
        s.l_h_s_tgh(); o.stage("lhs")
        dg = _interior(s.field(0, "dq"), blk)
        do = _interior(o.field(0, "dq"), blk)
        assert _rel_err(dg, do) <= REL, f"dq step {step}: {_rel_err(dg, do):.3e}"
        s.communicate_pv_npp(); o.stage("communicate_pv")
        _cmp_prim(s, o, blk, REL, ghost=0)


def test_sa_bump_30_steps():
    case = _sa_bump(ni=32, nj=24, nk=12)
    blk = case.blocks[0]
    s = _solver(case)
    o = O.OracleCase(case)
    hs = s.step(30)
    ho = o.step(30)
    rel = np.abs(hs - ho) / ho
    print("SA residual history rel err max", rel.max(), flush=True)
    assert rel.max() <= 1e-10
    _cmp_prim(s, o, blk, 1e-11, ghost=0)
    for name in ("qke", "vist"):
        e = _rel_err(_interior(s.field(0, name), blk), _interior(o.field(0, name), blk))
        assert e <= 1e-11, f"{name}: {e:.3e}"


def test_sa_multiblock_mixed_bcs_20_steps():
    """Four blocks: SA across 1:1 interfaces, symmetry, fixed inflow and outflow
    ghosts, and the batched scalar sweeps over all blocks at once."""
    case = _sa_mixed_bc()
    print(f"SA multiblock case: {len(case.blocks)} blocks", flush=True)
    s = _solver(case)
    o = O.OracleCase(case)
    for nb, blk in enumerate(case.blocks):
        e = _rel_err(_interior(s.field(nb, "dist"), blk), _interior(o.field(nb, "dist"), blk))
        assert e <= REL, f"wall distance block {nb}: {e:.3e}"
    hs = s.step(20)
    ho = o.step(20)
    rel = np.abs(hs - ho) / ho
    print("SA multiblock residual history rel err max", rel.max(), flush=True)
    assert rel.max() <= 1e-10
    _cmp_all_blocks(s, o, case, ("r", "u", "v", "w", "p", "qke", "vist"), 1e-11)


# ---------------------------------------------------------------- graph replay
GRAPH_CASES = {
    "bump-laminar": lambda: cases.bump_channel_case(ni=24, nj=16, nk=12, dy_wall=1e-3, growth=1.2, reynolds=1e4),
    "sa-multiblock": _sa_mixed_bc,
    "cart-euler": lambda: cases.cartesian_periodic_case(ni=24, nj=16, nk=12, seed=3, amp=0.02),
    "sst-multiblock": lambda: _sst_mixed_bc(),
}


@pytest.mark.parametrize("name", list(GRAPH_CASES))
def test_graph_replay_matches_plain_launches(name):
    """wcns_step_graph replays a captured two-step graph: with plain steps and
    host reads between the replays, every field equals the plain launches'."""
    case = GRAPH_CASES[name]()
    print(f"graph case {name}", flush=True)
    a = _solver(case)
    b = _solver(case)
    a.step(9, history=False)
    b.step_graph(4)
    b.field(0, "r")                  # host read between replays
    b.step(1, history=False)
    b.step_graph(3)                  # odd: one graph pair + one plain step
    b.step_graph(1)
    names = ["r", "u", "v", "w", "p"] + {1: ["qke", "vist"], 2: ["qke", "qkw", "vist"]}.get(case.params.nlamtur, [])
    for nb in range(len(case.blocks)):
        for nm in names:
            np.testing.assert_array_equal(a.field(nb, nm), b.field(nb, nm), err_msg=f"block {nb} {nm}")
    ra, rb = a.residual_curve(), b.residual_curve()
    assert ra[0] == rb[0] and ra[1] == rb[1]
    # every step's monitor went to the device history, graph replays included
    ha, hb = a.residual_history(9), b.residual_history(9)
    assert ha[3] == 9 and hb[3] == 9
    for x, y in zip(ha[:3], hb[:3]):
        np.testing.assert_array_equal(x, y)
    assert ha[0][-1] == ra[0] and ha[1][-1] == ra[1]


@pytest.mark.parametrize("name", ["bump-laminar", "sa-multiblock"])
def test_autotuned_tiles_give_identical_steps(name):
    """The autotuner only changes the sweep tile shape: the steps after it are
    bit-identical to those of an untuned solver, and it leaves the state alone."""
    case = GRAPH_CASES[name]()
    print(f"autotune case {name}", flush=True)
    a = _solver(case)
    b = _solver(case)
    a.step(2, history=False)
    r_before = a.field(0, "r")
    tj, tk, ms = a.autotune(reps=1)
    print("tile", tj, tk, "ms", ms, flush=True)
    assert 1 <= tj * tk <= 64 and ms > 0
    np.testing.assert_array_equal(a.field(0, "r"), r_before)
    a.step(3, history=False)
    b.step(5, history=False)
    names = ["r", "u", "v", "w", "p"] + {1: ["qke", "vist"], 2: ["qke", "qkw", "vist"]}.get(case.params.nlamtur, [])
    for nb in range(len(case.blocks)):
        for nm in names:
            np.testing.assert_array_equal(a.field(nb, nm), b.field(nb, nm), err_msg=f"block {nb} {nm}")


# ---------------------------------------------------------------- SST k-omega
def _sst_bump(**kw):
    case = cases.bump_channel_case(dy_wall=1e-3, growth=1.2, reynolds=1e5, **kw)
    case.params.nlamtur = 2
    return case


def _sst_mixed_bc():
    case = cases.multiblock_case(nvis=1, wall=True, periodic_i=False, ilo_bc=5, ihi_bc=6, jhi_bc=3)
    case.params.nlamtur = 2
    return case


@pytest.mark.parametrize("kw", SA_CASES, ids=lambda k: f"{k['ni']}x{k['nj']}x{k['nk']}-s{k['nscheme']}-f{k['nflux']}")
def test_sst_stage_by_stage(kw):
    """SST bump channel: initial k, omega, mu_t; each step the k/omega ghosts
    (wall omega value, far field, periodic), the SST step inside r_h_s_vis and
    the viscous residual with its mu_t."""
    case = _sst_bump(**kw)
    blk = case.blocks[0]
    print(f"SST case {kw}", flush=True)
    s = _solver(case)
    o = O.OracleCase(case)
    for name in ("qke", "qkw", "vist"):
        e = _rel_err(_interior(s.field(0, name), blk, 1), _interior(o.field(0, name), blk, 1))
        assert e == 0.0, f"initial {name}: {e:.3e}"
    for step in range(3):
        s.boundary_all(); o.stage("boundary_all")
        _cmp_prim(s, o, blk, REL)
        for name in ("qke", "qkw", "vist"):
            e = _rel_err(_interior(s.field(0, name), blk, 1), _interior(o.field(0, name), blk, 1))
            assert e <= REL, f"{name} after BCs, step {step}: {e:.3e}"
        s.initial_q_dq_c_t(); o.stage("initial_q_dq_c_t")
        s.timestep_tgh(); o.stage("timestep")
        s.r_h_s_vis(); o.stage("rhs_vis")
        for name in ("qke", "qkw", "vist"):
            e = _rel_err(_interior(s.field(0, name), blk, 1), _interior(o.field(0, name), blk, 1))
            assert e <= REL, f"{name} after the SST step, step {step}: {e:.3e}"
        rg = _interior(s.field(0, "rhs"), blk)
        ro = _interior(o.field(0, "dq"), blk)
        assert _rel_err(rg, ro) <= REL, f"viscous rhs step {step}: {_rel_err(rg, ro):.3e}"
        s.communicate_dq_npp(); o.stage("communicate_dq")
        s.r_h_s_invis(); o.stage("rhs_invis")
        s.communicate_dq_npp(); o.stage("communicate_dq")
        s.l_h_s_tgh(); o.stage("lhs")
        s.communicate_pv_npp(); o.stage("communicate_pv")
        _cmp_prim(s, o, blk, REL, ghost=0)


def test_sst_multiblock_mixed_bcs_20_steps():
    case = _sst_mixed_bc()
    print(f"SST multiblock case: {len(case.blocks)} blocks", flush=True)
    s = _solver(case)
    o = O.OracleCase(case)
    hs = s.step(20)
    ho = o.step(20)
    rel = np.abs(hs - ho) / ho
    print("SST multiblock residual history rel err max", rel.max(), flush=True)
    assert rel.max() <= 1e-10
    _cmp_all_blocks(s, o, case, ("r", "u", "v", "w", "p", "qke", "qkw", "vist"), 1e-11)


@pytest.mark.parametrize("name", ["bump-laminar", "sa-multiblock"])
def test_transposed_i_direction_matches_direct(name, monkeypatch):
    """The i-direction residuals run as j pencils of transposed copies: the steps
    equal those of the direct i-pencil kernels (WCNS_IDIR=direct) bit for bit."""
    case = GRAPH_CASES[name]()
    a = _solver(case)
    monkeypatch.setenv("WCNS_IDIR", "direct")
    b = _solver(case)
    a.step(5, history=False)
    b.step(5, history=False)
    names = ["r", "u", "v", "w", "p"] + {1: ["qke", "vist"], 2: ["qke", "qkw", "vist"]}.get(case.params.nlamtur, [])
    for nb in range(len(case.blocks)):
        for nm in names:
            np.testing.assert_array_equal(a.field(nb, nm), b.field(nb, nm), err_msg=f"block {nb} {nm}")
