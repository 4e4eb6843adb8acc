"""The CPU oracle against golden vectors written by the reference's own Fortran
(WCNSE5.f90 / flux.f90 compiled from source by oracle/Makefile, vectors made by
tests/golden/make_golden.py), and, where the reference library is present in
this container, against it directly on fresh seeded inputs.

The bar is bit-exactness: both sides round every product and sum once
(-ffp-contract=off, no FMA in the x86-64 baseline ISA) in the same order.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "pencil_golden.npz"))


@pytest.mark.parametrize("n", [3, 4, 5, 6, 9, 17])
def test_value_half_node_golden(n):
    out = O.value_half_node(GOLD[f"vhn_in_{n}"])
    np.testing.assert_array_equal(out, GOLD[f"vhn_out_{n}"])


@pytest.mark.parametrize("n", [3, 4, 5, 6, 9, 17])
def test_flux_dxyz_golden(n):
    out = O.flux_dxyz(GOLD[f"fdx_in_{n}"])
    np.testing.assert_array_equal(out, GOLD[f"fdx_out_{n}"])


def test_flux_sw_golden():
    for row in GOLD["flux_sw"]:
        pl, pr, nv, f = row[:5], row[5:10], row[10:14], row[14:19]
        np.testing.assert_array_equal(O.flux_sw(pl, pr, nv), f)


@pytest.mark.parametrize("key", [str(k) for k in GOLD["wcns_cases"]])
def test_wcns_line_golden(key):
    scheme = 41 if key.startswith("wcns41") else 4
    df, ne = O.wcns_line(scheme, 1, GOLD[key + "_tr"], GOLD[key + "_q"])
    np.testing.assert_array_equal(df, GOLD[key + "_df"])
    if key.endswith("fallback"):
        assert ne > 0


REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref.so")


@pytest.mark.skipif(not os.path.exists(REF), reason="reference library not built here")
def test_wcns_line_against_reference_fresh_inputs():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    lib = mg.load()
    rng = np.random.default_rng(777)
    for scheme, fn in ((4, "wcns_e_5_"), (41, "wcns_e_5_41_")):
        for n in (5, 7, 12, 20, 64):
            for bl in (0, 1):
                for br in (0, 1):
                    q = mg.smooth_line(rng, n, bl, br, amp=0.2, rough=bool(rng.integers(0, 2)))
                    tr = mg.metrics_line(rng, n)
                    ref = mg.wcns_line(lib, fn, tr, q)
                    mine, _ = O.wcns_line(scheme, 1, tr, q)
                    np.testing.assert_array_equal(mine, ref)


# ---- serial build of the reference: flux_Roe, WCNS with Roe, the SA block step
SER = np.load(os.path.join(HERE, "golden", "serial_golden.npz"))
REF_SA = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_sa.so")


def _serial_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mgs", os.path.join(HERE, "golden", "make_golden_serial.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_flux_roe_golden():
    for row in SER["flux_roe"]:
        pl, pr, nv, f = row[:5], row[5:10], row[10:14], row[14:19]
        np.testing.assert_array_equal(O.flux_roe(pl, pr, nv), f)


@pytest.mark.parametrize("key", [str(k) for k in SER["roe_cases"]])
def test_wcns_line_roe_golden(key):
    scheme = 41 if key.startswith("roe41") else 4
    df, _ = O.wcns_line(scheme, 3, SER[key + "_tr"], SER[key + "_q"])
    np.testing.assert_array_equal(df, SER[key + "_df"])


def _oracle_sa_step(mgs, dims, fields):
    oc = O.OracleCase(mgs.sa_case(*dims))
    for n in mgs.SA_FIELDS:
        oc.set_field(0, n, fields[n])
    oc.turbulence(0)
    q, v = oc.field(0, "qke"), oc.field(0, "vist")
    oc.close()
    return q, v


@pytest.mark.parametrize("tag", [str(k) for k in SER["sa_cases"]])
def test_sa_block_step_golden(tag):
    """Sources, diffusion, NND convection, scalar LU-SGS, update, cal_vist_SA and
    max_control_turbulence of one block, bit for bit against the reference."""
    mgs = _serial_module()
    dims = tuple(int(d) for d in SER[f"{tag}_dims"])
    fields = {n: SER[f"{tag}_in_{n}"] for n in mgs.SA_FIELDS}
    q, v = _oracle_sa_step(mgs, dims, fields)
    assert (q != fields["qke"]).sum() > 0
    np.testing.assert_array_equal(q, SER[f"{tag}_qke"])
    np.testing.assert_array_equal(v, SER[f"{tag}_vist"])


@pytest.mark.skipif(not os.path.exists(REF_SA), reason="serial reference library not built here")
@pytest.mark.parametrize("seed,kick,dims,steps", [(11, 0.5, (16, 12, 9), 3), (12, 3.0, (13, 11, 10), 1),
                                                  (13, 6.0, (12, 10, 8), 2)])
def test_sa_block_step_against_reference_fresh_inputs(seed, kick, dims, steps):
    mgs = _serial_module()
    lib = mgs.load()
    case = mgs.sa_case(*dims)
    f = mgs.sa_inputs(case, steps, seed, kick)
    qr, vr = mgs.sa_step(lib, dims, mgs.sa_scal(case.params), f)
    q, v = _oracle_sa_step(mgs, dims, f)
    np.testing.assert_array_equal(q, qr)
    np.testing.assert_array_equal(v, vr)


@pytest.mark.skipif(not os.path.exists(REF_SA), reason="serial reference library not built here")
def test_flux_roe_against_reference_fresh_inputs():
    mgs = _serial_module()
    lib = mgs.load()
    rng = np.random.default_rng(4242)
    for _ in range(500):
        pl = np.array([0.2 + 2 * rng.random(), *(3 * rng.standard_normal(3)), 0.1 + 3 * rng.random()])
        pr = np.array([0.2 + 2 * rng.random(), *(3 * rng.standard_normal(3)), 0.1 + 3 * rng.random()])
        nv = np.array([*rng.standard_normal(3), 0.0])
        np.testing.assert_array_equal(O.flux_roe(pl, pr, nv), mgs.flux_roe(lib, pl, pr, nv))


@pytest.mark.skipif(not os.path.exists(REF_SA), reason="serial reference library not built here")
def test_flux_roe_c2_nonpositive_stops_both():
    """A Roe average with c^2 <= 0 (a negative face pressure, as an
    extrapolated WCNS state can give): the reference's flux_Roe stops the run
    (flux.f90:178-181), so it runs in a child process; the oracle raises."""
    import subprocess
    import sys
    import textwrap
    pl = [1.0, 0.3, 0.0, 0.0, -2.0]
    pr = [1.0, 0.2, 0.1, 0.0, 0.7]
    nv = [1.0, 0.0, 0.0, 0.0]
    with pytest.raises(ArithmeticError):
        O.flux_roe(np.array(pl), np.array(pr), nv)
    script = textwrap.dedent(f"""
        import importlib.util, numpy as np
        spec = importlib.util.spec_from_file_location("mgs", {os.path.join(HERE, "golden", "make_golden_serial.py")!r})
        m = importlib.util.module_from_spec(spec); spec.loader.exec_module(m)
        lib = m.load()
        m.flux_roe(lib, np.array({pl}), np.array({pr}), np.array({nv}))
        print("FLUX RETURNED")
    """)
    env = dict(os.environ, PYTHONPATH=os.path.dirname(HERE))
    p = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, env=env, timeout=120)
    assert "math error in subroutine flux_Roe" in p.stdout
    assert "FLUX RETURNED" not in p.stdout


@pytest.mark.skipif(not os.path.exists(REF_SA), reason="serial reference library not built here")
def test_flux_vanleer_against_reference_fresh_inputs():
    """van Leer's splitting (flux.f90:395-501) on states that cover the
    subsonic branch and both supersonic branches of each half."""
    mgs = _serial_module()
    lib = mgs.load()
    rng = np.random.default_rng(4343)
    for _ in range(600):
        pl = np.array([0.2 + 2 * rng.random(), *(3 * rng.standard_normal(3)), 0.1 + 3 * rng.random()])
        pr = np.array([0.2 + 2 * rng.random(), *(3 * rng.standard_normal(3)), 0.1 + 3 * rng.random()])
        nv = np.array([*rng.standard_normal(3), 0.0])
        ref = np.zeros(5)
        lib.flux_vanleer_(mgs.dp(pl.copy()), mgs.dp(pr.copy()), mgs.ip(5), mgs.dv(nv[0]), mgs.dv(nv[1]),
                          mgs.dv(nv[2]), mgs.dv(nv[3]), mgs.dp(ref), mgs.dv(mgs.EFIX), mgs.dv(mgs.GAMA))
        np.testing.assert_array_equal(O.flux_vanleer(pl, pr, nv), ref)


@pytest.mark.skipif(not os.path.exists(REF_SA), reason="serial reference library not built here")
@pytest.mark.parametrize("scheme,fn", [(4, "wcns_e_5_"), (41, "wcns_e_5_41_"), (42, "wcns_e_5_42_")])
@pytest.mark.parametrize("flux,ffn", [(1, "flux_sw_"), (2, "flux_vanleer_"), (3, "flux_roe_")])
def test_wcns_schemes_and_fluxes_against_reference(scheme, fn, flux, ffn):
    """Whole WCNS lines (WCNS_E_5, _41, _42 of WCNSE5.f90) with each flux the
    reference dispatches to them (Convect.f90:20-37), on smooth and rough
    lines with and without physical boundaries at either end."""
    import importlib.util
    mgs = _serial_module()
    lib = mgs.load()
    spec = importlib.util.spec_from_file_location("mg", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    rng = np.random.default_rng(1000 * scheme + flux)
    for n in (5, 7, 12, 20, 33):
        for bl in (0, 1):
            for br in (0, 1):
                q = mg.smooth_line(rng, n, bl, br, amp=0.2, rough=bool(rng.integers(0, 2)))
                tr = mg.metrics_line(rng, n)
                ref = mgs.wcns_line_flux(lib, fn, ffn, tr, q)
                mine, _ = O.wcns_line(scheme, flux, tr, q)
                np.testing.assert_array_equal(mine, ref)


def _oracle_sst_step(mgs, dims, fields):
    oc = O.OracleCase(mgs.sst_case(*dims))
    for n in mgs.SA_FIELDS + ["qkw"]:
        oc.set_field(0, n, fields[n])
    oc.turbulence(0)
    out = oc.field(0, "qke"), oc.field(0, "qkw"), oc.field(0, "vist")
    oc.close()
    return out


@pytest.mark.parametrize("tag", [str(k) for k in SER["sst_cases"]])
def test_sst_block_step_golden(tag):
    """The SST k-omega step of one block (sources with the F1 blending, diffusion,
    diagonal, NND convection and scalar LU-SGS per variable, update, mu_t of
    cal_vist_SST, max_control), bit for bit against the reference."""
    mgs = _serial_module()
    dims = tuple(int(d) for d in SER[f"{tag}_dims"])
    fields = {n: SER[f"{tag}_in_{n}"] for n in mgs.SA_FIELDS + ["qkw"]}
    k, w, v = _oracle_sst_step(mgs, dims, fields)
    assert (k != fields["qke"]).sum() > 0 and (w != fields["qkw"]).sum() > 0
    np.testing.assert_array_equal(k, SER[f"{tag}_qke"])
    np.testing.assert_array_equal(w, SER[f"{tag}_qkw"])
    np.testing.assert_array_equal(v, SER[f"{tag}_vist"])


@pytest.mark.skipif(not os.path.exists(REF_SA), reason="serial reference library not built here")
@pytest.mark.parametrize("seed,kick,dims,steps", [(21, 0.5, (16, 12, 9), 3), (22, 2.5, (13, 11, 10), 1)])
def test_sst_block_step_against_reference_fresh_inputs(seed, kick, dims, steps):
    mgs = _serial_module()
    lib = mgs.load()
    case = mgs.sst_case(*dims)
    f = mgs.sst_inputs(case, steps, seed, kick)
    kr, wr, vr = mgs.sst_step(lib, dims, mgs.sst_scal(case.params), f)
    k, w, v = _oracle_sst_step(mgs, dims, f)
    np.testing.assert_array_equal(k, kr)
    np.testing.assert_array_equal(w, wr)
    np.testing.assert_array_equal(v, vr)


@pytest.mark.skipif(not os.path.exists(REF_SA), reason="serial reference library not built here")
@pytest.mark.parametrize("nlamtur", [1, 2])
def test_wall_turbulence_values_against_reference(nlamtur):
    """The oracle's wall values of nu~ (SA) or k and omega (SST) after
    boundary_all equal the reference's BC_wall_turbulence + BC_face_dif_turbulence
    applied to the same pre-BC state, at the face points no other region touches."""
    import ctypes as C
    mgs = _serial_module()
    lib = mgs.load()
    case = mgs.sst_case(14, 10, 8) if nlamtur == 2 else mgs.sa_case(14, 10, 8)
    blk = case.blocks[0]
    oc = O.OracleCase(case)
    oc.step(2)
    rng = np.random.default_rng(9)
    names = ["qke"] + (["qkw"] if nlamtur == 2 else [])
    for nm in names:
        f = oc.field(0, nm)
        oc.set_field(0, nm, f * (1.0 + 0.2 * rng.random(f.shape)))
    pre = {nm: np.ascontiguousarray(oc.field(0, nm)) for nm in ["r", "t", "qke", "qkw", "dist"]}
    oc.stage("boundary_all")
    # wall j = 1: ghost j = 0, first interior j = 2; points i = 2..ni-1, k = 2..nk-1
    idx = []
    for k in range(2, blk.nk):
        for i in range(2, blk.ni):
            idx += [i, 0, k, i, 2, k, i, 1, k]
    idx = np.array(idx, dtype=np.int32)
    qo, wo = np.zeros_like(pre["r"]), np.zeros_like(pre["r"])
    ptrs = (C.c_void_p * 5)(*[pre[n].ctypes.data for n in ["r", "t", "qke", "qkw", "dist"]])
    scal = (C.c_double * 2)(case.params.reynolds, 110.4 / case.params.tref)
    lib.ref_wall_turb.restype = C.c_int
    rc = lib.ref_wall_turb((C.c_int * 3)(blk.ni, blk.nj, blk.nk), C.c_int(nlamtur), scal, ptrs,
                           C.c_int(len(idx) // 9), idx.ctypes.data_as(C.c_void_p),
                           qo.ctypes.data_as(C.c_void_p), wo.ctypes.data_as(C.c_void_p))
    assert rc == 0
    # G3 index = node + 2: nodes k = 2..nk-1, j = 0..1 (ghost, face), i = 2..ni-1
    sl = (slice(2 + 2, 2 + blk.nk), slice(2 + 0, 2 + 2), slice(2 + 2, 2 + blk.ni))
    np.testing.assert_array_equal(oc.field(0, "qke")[sl], qo[sl])
    if nlamtur == 2:
        np.testing.assert_array_equal(oc.field(0, "qkw")[sl], wo[sl])
        assert np.any(wo[sl][:, 0, :] != 0.0)


# ------------------------------------------------ whole steps of the reference
STEP_GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "step_golden.npz")


@pytest.mark.parametrize("tag", ["cart", "bump", "sa", "repair", "s45", "s47wall", "s45vis", "roemx42", "roemx47"])
def test_whole_steps_against_reference_fixtures(tag):
    """The oracle's steps against the reference's own serial step routines
    (tests/golden/make_golden_step.py): every residual-monitor entry (rms, max
    and its location), the final interior state (SHA-256 of its bytes) and the
    wall force, bit for bit; the same stage sequence, without the interface
    averaging that only the reference's MPI build has."""
    import hashlib
    from tests.golden.make_golden_step import golden_cases, poke_repair, state_hash, STAGES_INV, STAGES_VIS
    g = np.load(STEP_GOLDEN)
    case, nsteps, vis = golden_cases()[tag]
    o = O.OracleCase(case)
    o.set_threads(4)
    if tag == "repair":
        poke_repair(o)
    hist = []
    for _ in range(nsteps):
        for st in (STAGES_VIS if vis else STAGES_INV):
            o.stage(st)
        rms, rmax, loc = o.residual()
        hist.append([rms, rmax] + [float(x) for x in loc])
    assert np.array_equal(np.array(hist), g[f"{tag}_hist"])
    assert np.array_equal(state_hash(case, o), g[f"{tag}_hash"])
    if vis:
        o.stage("boundary_all")
        assert np.array_equal(o.force(), g[f"{tag}_force"])
    o.close()
