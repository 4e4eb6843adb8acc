"""Checks that pin the oracle where the reference's own compiled routines
cannot be run here (flux_Roe, GCL metrics, interface topology, the implicit
stage): closed forms, invariants and special cases chosen so that a dropped
term, a wrong sign or a transposed index fails one of them."""
import numpy as np
import pytest

from oracle import oracle as O
from wcns_amd import cases

GAMA = 1.4


def phys_flux(q, n):
    r, u, v, w, p = q
    e = p / (GAMA - 1) + 0.5 * r * (u * u + v * v + w * w)
    cit = n[0] * u + n[1] * v + n[2] * w
    return np.array([r * cit, r * cit * u + n[0] * p, r * cit * v + n[1] * p, r * cit * w + n[2] * p,
                     (e + p) * cit])


def rand_state(rng):
    return np.array([0.5 + rng.random(), rng.standard_normal(), rng.standard_normal(), rng.standard_normal(),
                     0.5 + rng.random()])


def test_roe_consistency_and_upwinding():
    """flux_Roe(q,q) = F(q) exactly; for a supersonic left state moving along
    the normal the Roe flux (efix -> 0) is the left physical flux."""
    rng = np.random.default_rng(1)
    for _ in range(50):
        q = rand_state(rng)
        n = np.append(rng.standard_normal(3), 0.0)
        np.testing.assert_allclose(O.flux_roe(q, q, n), phys_flux(q, n[:3]), rtol=1e-13, atol=1e-13)
    ql = np.array([1.0, 3.0, 0.1, -0.2, 1.0])           # Mach ~ 2.5 along +x
    qr = ql * np.array([1.001, 1.0, 1.0, 1.0, 1.002])
    f = O.flux_roe(ql, qr, [1.0, 0.0, 0.0, 0.0], efix=0.0)
    np.testing.assert_allclose(f, phys_flux(ql, [1, 0, 0]), rtol=1e-12)


def test_roe_raises_on_negative_c2():
    with pytest.raises(ArithmeticError):
        O.flux_roe(np.array([1.0, 0, 0, 0, -5.0]), np.array([1.0, 0, 0, 0, -5.0]), [1, 0, 0, 0])


def test_steger_warming_split_sums_to_flux():
    """With efix = 0 the Steger-Warming halves add up to F(q)."""
    rng = np.random.default_rng(2)
    for _ in range(50):
        q = rand_state(rng)
        n = np.append(rng.standard_normal(3), 0.0)
        np.testing.assert_allclose(O.flux_sw(q, q, n, efix=0.0), phys_flux(q, n[:3]), rtol=1e-12, atol=1e-12)


def test_mxdq_split_jacobian_sums_to_jacobian():
    """A+ dq - A- dq with rad = |lambda|max reconstructs A dq (finite difference of F)."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        prim = rand_state(rng)
        g = np.append(rng.standard_normal(3), 0.0)
        dq = rng.standard_normal(5) * 1e-6
        fp = O.mxdq_std(prim, g, dq, 7.0, 1)
        fm = O.mxdq_std(prim, g, dq, 7.0, -1)
        r, u, v, w, p = prim
        q = np.array([r, r * u, r * v, r * w, p / (GAMA - 1) + 0.5 * r * (u * u + v * v + w * w)])

        def q2p(qq):
            rr = qq[0]
            uu, vv, ww = qq[1:4] / rr
            pp = (GAMA - 1) * (qq[4] - 0.5 * rr * (uu * uu + vv * vv + ww * ww))
            return np.array([rr, uu, vv, ww, pp])
        adq = phys_flux(q2p(q + dq), g[:3]) - phys_flux(q2p(q - dq), g[:3])
        np.testing.assert_allclose(fp + fm, 0.5 * adq, rtol=1e-5, atol=1e-12)


def test_gcl_metrics_cartesian_closed_form():
    c = cases.cartesian_periodic_case(ni=16, nj=12, nk=10, stretch=1.05)
    o = O.OracleCase(c)
    b = c.blocks[0]
    dx = b.x[1, 0, 0] - b.x[0, 0, 0]
    dz = b.z[0, 0, 1] - b.z[0, 0, 0]
    kcx = o.field(0, "kcx")[3:-3, 3:-3, 3:-3]
    vol = o.field(0, "vol")[3:-3, 3:-3, 3:-3]
    for zero in ("kcy", "kcz", "etx", "etz", "ctx", "cty"):
        assert np.max(np.abs(o.field(0, zero)[3:-3, 3:-3, 3:-3])) < 1e-14
    # kcx = y_eta * z_zeta, ctz = x_xi * y_eta, vol = x_xi * y_eta * z_zeta (with the scheme's y_eta)
    ety = o.field(0, "ety")[3:-3, 3:-3, 3:-3]
    np.testing.assert_allclose(ety, dx * dz, rtol=1e-13)
    np.testing.assert_allclose(vol, kcx * dx, rtol=1e-13)


def test_freestream_preserved_on_curvilinear_grid():
    """GCL metrics + WCNS-E5 + FLUX_DXYZ: a uniform flow gives R = 0 to round-off
    on a curved grid, through the periodic interfaces, for both schemes and fluxes."""
    for scheme, flux in ((41, 1), (4, 1), (41, 3)):
        c = cases.wavy_periodic_case(ni=20, nj=14, nk=13, nscheme=scheme, nflux=flux)
        c.init = cases.uniform_init
        o = O.OracleCase(c)
        o.stage("boundary_all")
        o.stage("initial_q_dq_c_t")
        o.stage("timestep")
        o.stage("rhs_invis")
        r = o.field(0, "dq")[3:-3, 3:-3, 3:-3]
        # the far-field faces (BC 4) regenerate the freestream through pow(); keep the interior rows
        inner = r[:, 4:-4, :]
        assert np.max(np.abs(inner)) < 1e-12, (scheme, flux, np.max(np.abs(inner)))


def test_interface_images_periodic():
    c = cases.cartesian_periodic_case(ni=12, nj=10, nk=14)
    o = O.OracleCase(c)
    box, img, jmg, kmg = o.region_images(0, 0)          # i-min face -> i-max face
    assert (img == 12).all()
    jj, kk = np.meshgrid(np.arange(1, 11), np.arange(1, 15), indexing="ij")
    np.testing.assert_array_equal(jmg, jj.T.ravel())
    np.testing.assert_array_equal(kmg, kk.T.ravel())
    box, img, jmg, kmg = o.region_images(0, 5)          # k-max face -> k-min face
    assert (kmg == 1).all()


def test_multipoint_groups_periodic_box():
    """i and k periodic: the nodes (1|ni, j, 1|nk) of each j form one group of four."""
    ni, nj, nk = 12, 10, 14
    c = cases.cartesian_periodic_case(ni=ni, nj=nj, nk=nk)
    o = O.OracleCase(c)
    npts, pts = o.multipoint()
    assert len(npts) == nj and (npts == 4).all()
    for g in range(nj):
        s = {tuple(p) for p in pts[4 * g:4 * g + 4]}
        j = pts[4 * g][2]
        assert s == {(1, 1, j, 1), (1, ni, j, 1), (1, 1, j, nk), (1, ni, j, nk)}


def test_interface_average_is_ordered_not_symmetric():
    """Same-process pairs are averaged in place and in order (parallel.f90:3083-3098):
    after a, then b: a' = (a+b)/2, b' = (a'+b)/2 -- not the symmetric mean."""
    c = cases.cartesian_periodic_case(ni=12, nj=10, nk=14)
    o = O.OracleCase(c)
    dq = np.zeros_like(o.field(0, "dq"))
    # rhs on the i faces only, away from the k faces (no multi-points there)
    dq[3 + 5, 3 + 4, 3 + 0, :] = 1.0     # node (1, 5, 6)
    dq[3 + 5, 3 + 4, 3 + 11, :] = 3.0    # node (12, 5, 6)
    o.set_field(0, "dq", dq)
    o.stage("communicate_dq")
    out = o.field(0, "dq")
    a = out[3 + 5, 3 + 4, 3 + 0, 0]
    b = out[3 + 5, 3 + 4, 3 + 11, 0]
    assert abs(a - 2.0) < 1e-13 and abs(b - 2.5) < 1e-13


def test_residual_monitor_definition():
    c = cases.cartesian_periodic_case(ni=12, nj=10, nk=14)
    o = O.OracleCase(c)
    o.stage("boundary_all"); o.stage("initial_q_dq_c_t"); o.stage("timestep")
    dq = np.zeros_like(o.field(0, "dq"))
    dq[3 + 4, 3 + 3, 3 + 2, 2] = 5.0        # node (3, 4, 5), m = 3
    o.set_field(0, "dq", dq)
    rms, rmax, loc = o.residual()
    dt = o.field(0, "dtdt")[3 + 4, 3 + 3, 3 + 2]
    vol = o.field(0, "vol")[3 + 4, 3 + 3, 3 + 2]
    val = 5.0 * (1.0 / (dt * vol))
    assert rmax == val
    assert list(loc) == [3, 4, 5, 3, 1]
    np.testing.assert_allclose(rms, np.sqrt(val * val / (10 * 8 * 12 * 5)), rtol=1e-15)


def test_oracle_runs_cartesian_case_and_converges():
    c = cases.cartesian_periodic_case(ni=24, nj=16, nk=12)
    o = O.OracleCase(c)
    h = o.step(30)
    assert np.all(np.isfinite(h)) and h[-1] < h[0]


def test_planted_sign_error_breaks_freestream():
    """A wrong sign in a metric (planted on the oracle's metric array) destroys
    freestream preservation -- the invariant above is sensitive to such a bug."""
    c = cases.wavy_periodic_case(ni=20, nj=14, nk=13)
    c.init = cases.uniform_init
    o = O.OracleCase(c)
    kcy = o.field(0, "kcy")
    o.set_field(0, "kcy", -kcy)
    o.stage("boundary_all"); o.stage("initial_q_dq_c_t"); o.stage("timestep"); o.stage("rhs_invis")
    r = o.field(0, "dq")[3:-3, 7:-7, 3:-3]
    assert np.max(np.abs(r)) > 1e-6


def _uniform_grid_case(ni=14, nj=13, nk=12, stretch=False):
    xi = np.arange(ni) * 0.1
    yj = np.arange(nj) * 0.07
    zk = np.arange(nk) * 0.05
    X, Y, Z = np.meshgrid(xi, yj, zk, indexing="ij")
    regs = [cases.Region((1, 1, 1, nj, 1, nk), 6), cases.Region((ni, ni, 1, nj, 1, nk), 6),
            cases.Region((1, ni, 1, 1, 1, nk), 6), cases.Region((1, ni, nj, nj, 1, nk), 6),
            cases.Region((1, ni, 1, nj, 1, 1), 6), cases.Region((1, ni, 1, nj, nk, nk), 6)]
    blk = cases.Block(ni, nj, nk, X, Y, Z, regs)
    prm = cases.Params(moo=0.5, reynolds=100.0, nvis=1, nflux=3)
    return cases.Case(prm, [blk])


def test_viscous_linear_velocity_gives_zero_momentum_residual():
    """u = A x with constant rho, p (so constant T and mu): the stress is uniform,
    so every interior momentum component of the viscous residual vanishes, while
    the energy component carries the dissipation tau:grad(u) / Re exactly."""
    A = np.array([[0.3, 0.7, -0.2], [0.1, -0.5, 0.4], [-0.6, 0.2, 0.2]])
    c = _uniform_grid_case()
    c.init = cases.linear_field_init(A)
    o = O.OracleCase(c)
    o.stage("initial_q_dq_c_t")          # T and mu (no BC call: ghosts hold the linear field)
    o.stage("rhs_vis")
    r = -o.field(0, "dq")[3:-3, 3:-3, 3:-3]
    inner = r[3:-3, 3:-3, 3:-3]
    assert np.max(np.abs(inner[..., 1:4])) < 1e-11
    assert np.max(np.abs(inner[..., 0])) == 0.0
    # energy: d/dx_j (u_i tau_ij) = tau_ij A_ij (constant), times vol, over Re
    mu = o.field(0, "visl")[6, 6, 6]
    S = A + A.T
    tau = mu * (S - 2.0 / 3.0 * np.trace(A) * np.eye(3))
    vol = 0.1 * 0.07 * 0.05
    expect = np.sum(tau * A) * vol / 100.0
    np.testing.assert_allclose(inner[..., 4], expect, rtol=1e-9)


def test_viscous_uniform_flow_no_residual():
    c = cases.bump_channel_case(ni=20, nj=14, nk=12, dy_wall=1e-3, growth=1.2)
    o = O.OracleCase(c)
    o.stage("initial_q_dq_c_t")
    o.stage("rhs_vis")
    r = o.field(0, "dq")[3:-3, 3:-3, 3:-3]
    assert np.max(np.abs(r[3:-3, 3:-3, :])) < 1e-12


def test_viscous_bump_runs_and_converges():
    c = cases.bump_channel_case(ni=24, nj=16, nk=12, dy_wall=1e-3, growth=1.2, reynolds=1e3)
    o = O.OracleCase(c)
    h = o.step(20)
    assert np.all(np.isfinite(h))


# ------------------------------------------------------------- wall force
def _flat_wall(**prm):
    """Flat no-slip wall y = 0, x in [-1.5, 2.5], z in [0, 0.5] (area 2), uniform spacing."""
    c = cases.bump_channel_case(ni=12, nj=10, nk=8, bump=0.0, dy_wall=0.1, growth=1.0)
    for k, v in prm.items():
        setattr(c.params, k, v)
    return c


def _set_state(o, c, u, v, w, p, mu):
    b = c.blocks[0]
    shp = (b.nk + 6, b.nj + 6, b.ni + 6)
    x = np.zeros(shp); y = np.zeros(shp)
    x[3:-3, 3:-3, 3:-3] = b.x.transpose(2, 1, 0)
    y[3:-3, 3:-3, 3:-3] = b.y.transpose(2, 1, 0)
    o.set_field(0, "r", np.ones(shp))
    for name, f in (("u", u), ("v", v), ("w", w), ("p", p)):
        o.set_field(0, name, f(x, y) * np.ones(shp))
    o.set_field(0, "visl", np.full(shp, mu))


def test_force_constant_pressure_closed_form():
    """Uniform overpressure c on the flat wall: force -2c*A on the wall normal,
    moments from the plate centroid, centre of pressure at the plate middle."""
    c = _flat_wall(sref=1.0, lfref=1.0, nwholefield=1)
    o = O.OracleCase(c)
    poo = o.freestream()[4]
    cc = 0.03
    _set_state(o, c, lambda x, y: 0.0, lambda x, y: 0.0, lambda x, y: 0.0, lambda x, y: poo + cc, 1.0)
    f = o.force()
    # int x dA = 1.0, int z dA = 0.5 over the plate
    assert abs(f[0]) < 1e-14 and abs(f[2]) < 1e-14
    assert f[1] == pytest.approx(-4.0 * cc, rel=1e-12)
    assert f[5] == pytest.approx(-2.0 * cc * 1.0, rel=1e-12)      # cmz = int x fy
    assert f[3] == pytest.approx(2.0 * cc * 0.5, rel=1e-12)       # cmx = -int z fy
    assert f[7] == pytest.approx(f[1], rel=1e-12)                  # cl at zero incidence
    assert f[8] == pytest.approx(0.5, rel=1e-10)                   # xcp = plate middle


def test_force_scaling_by_reference_values():
    c1 = _flat_wall(sref=1.0, lfref=1.0, nwholefield=1)
    c2 = _flat_wall(sref=2.0, lfref=0.5, nwholefield=0, attack=30.0)
    out = []
    for c in (c1, c2):
        o = O.OracleCase(c)
        poo = o.freestream()[4]
        _set_state(o, c, lambda x, y: 0.0, lambda x, y: 0.0, lambda x, y: 0.0, lambda x, y: poo + 0.01 * (1 + x), 1.0)
        out.append(o.force())
    a, b = out
    np.testing.assert_allclose(b[:3], a[:3] * 2 / 2.0, rtol=1e-12)           # (2-0)/sref
    np.testing.assert_allclose(b[3:6], a[3:6] * 2 / (2.0 * 0.5), rtol=1e-12)  # (2-0)/(sref lfref)
    al = 30.0 * 3.14159265 / 180.0
    assert b[7] == pytest.approx(b[1] * np.cos(al) - b[0] * np.sin(al), rel=1e-12)
    assert b[6] == pytest.approx(b[1] * np.sin(al) + b[0] * np.cos(al), rel=1e-12)


def test_force_viscous_stress_closed_form():
    """u = a y, v = b x + d y, w = e y over a uniform grid: every one-sided and
    central difference is exact, so the wall traction is the Stokes stress:
    fx = 2 re mu (a+b), fy = 4 re mu c23 d, fz = 2 re mu e per unit area."""
    c = _flat_wall(sref=1.0, lfref=1.0, nwholefield=1, yref=0.3)
    o = O.OracleCase(c)
    poo = o.freestream()[4]
    a, bb, d, e, mu = 0.3, -0.2, 0.05, 0.7, 1.7
    _set_state(o, c, lambda x, y: a * y, lambda x, y: bb * x + d * y, lambda x, y: e * y, lambda x, y: poo, mu)
    re = 1.0 / c.params.reynolds
    c23 = 0.66666666667
    f = o.force()
    A = 2.0
    assert f[0] == pytest.approx(2 * re * mu * (a + bb) * A, rel=1e-10)
    assert f[1] == pytest.approx(4 * re * mu * c23 * d * A, rel=1e-10)
    assert f[2] == pytest.approx(2 * re * mu * e * A, rel=1e-10)
    # the wall is y = 0: cmz = int x fy dA - (0 - yref) int fx dA
    assert f[5] == pytest.approx(4 * re * mu * c23 * d * 1.0 + 0.3 * f[0], rel=1e-9)


# ------------------------------------------------------------ multi-block
def test_freestream_preserved_across_block_interfaces():
    """2 x 2 blocks of uneven sizes on a curved grid, 1:1 interfaces in i and j,
    periodic i and k: a uniform flow gives R = 0 to round-off in every block,
    in the rows next to the interfaces too (far-field rows excluded)."""
    for scheme, flux in ((41, 3), (4, 1)):
        c = cases.multiblock_case(amp=0.0, nscheme=scheme, nflux=flux)
        c.init = cases.uniform_init
        o = O.OracleCase(c)
        for st in ("boundary_all", "initial_q_dq_c_t", "timestep", "rhs_invis"):
            o.stage(st)
        for nb in range(len(c.blocks)):
            r = o.field(nb, "dq")[3:-3, 3:-3, 3:-3]
            row = (nb // 2) % 2                      # 0: far field at j = 1, 1: far field at j = nj
            sel = r[:, 4:, :] if row == 0 else r[:, :-4, :]
            assert np.max(np.abs(sel)) < 1e-13, (scheme, flux, nb, np.max(np.abs(sel)))


def test_block_interface_images_match_coordinates():
    """analyze_bc_connect's affine image of every interface point lands on the
    partner node with the same coordinates (a period lx apart across the
    periodic i faces)."""
    c = cases.multiblock_case(nx=(14, 17, 12), ny=(13, 15), nz=(12, 14))
    o = O.OracleCase(c)
    lx = 2.0
    for nb, blk in enumerate(c.blocks):
        for nr, reg in enumerate(blk.regions):
            if reg.bctype != -1:
                continue
            box, img, jmg, kmg = o.region_images(nb, nr)
            tb = c.blocks[reg.nbt - 1]
            pts = [(i, j, k) for k in range(box[2], box[5] + 1) for j in range(box[1], box[4] + 1)
                   for i in range(box[0], box[3] + 1)]
            assert len(pts) == len(img)
            for (i, j, k), it, jt, kt in zip(pts, img, jmg, kmg):
                src = np.array([blk.x[i - 1, j - 1, k - 1], blk.y[i - 1, j - 1, k - 1], blk.z[i - 1, j - 1, k - 1]])
                dst = np.array([tb.x[it - 1, jt - 1, kt - 1], tb.y[it - 1, jt - 1, kt - 1], tb.z[it - 1, jt - 1, kt - 1]])
                d = dst - src
                if nr < 2 and abs(abs(d[0]) - lx) < 1e-9:
                    d[0] = 0.0
                if nr >= 4:                         # periodic k: z apart by lz
                    d[2] = 0.0 if abs(abs(d[2]) - 0.6) < 1e-9 else d[2]
                assert np.max(np.abs(d)) < 1e-12, (nb, nr, (i, j, k), (it, jt, kt), d)


# ---------------------------------------------------------------- turbulence models
def _turb_case(nlamtur, **kw):
    kw.setdefault("dy_wall", 0.01)
    kw.setdefault("growth", 1.2)
    case = cases.bump_channel_case(ni=kw.pop("ni", 14), nj=kw.pop("nj", 10), nk=kw.pop("nk", 8),
                                   reynolds=kw.pop("reynolds", 1e5), **kw)
    case.params.nlamtur = nlamtur
    return case


def test_wall_distance_is_brute_force_minimum():
    """calc_wall_dists_parallel (parallel.f90:4313-4468): sqrt of the least
    squared distance to any node of a no-slip wall region, plus small -- on a
    two-block case with the wall split between the blocks."""
    case = cases.multiblock_case(nvis=1, wall=True, periodic_i=False, ilo_bc=5, ihi_bc=6, jhi_bc=3)
    case.params.nlamtur = 1
    oc = O.OracleCase(case)
    wall = []
    for blk in case.blocks:
        for reg in blk.regions:
            if reg.bctype == 2:
                s = reg.s_raw
                sl = (slice(abs(s[0]) - 1, abs(s[1])), slice(abs(s[2]) - 1, abs(s[3])), slice(abs(s[4]) - 1, abs(s[5])))
                wall.append(np.stack([blk.x[sl].ravel(), blk.y[sl].ravel(), blk.z[sl].ravel()], 1))
    wall = np.concatenate(wall)
    for nb, blk in enumerate(case.blocks):
        pts = np.stack([blk.x.ravel(), blk.y.ravel(), blk.z.ravel()], 1)
        d2 = ((pts[:, None, :] - wall[None, :, :]) ** 2).sum(-1).min(1)
        want = (np.sqrt(d2) + 1e-38).reshape(blk.x.shape)
        got = oc.field(nb, "dist")[3:-3, 3:-3, 3:-3].transpose(2, 1, 0)
        np.testing.assert_allclose(got, want, rtol=2e-16, atol=0)
    # a plant: dropping the z term changes the distance of off-plane nodes
    assert not np.allclose(np.sqrt(((pts[:, None, :2] - wall[None, :, :2]) ** 2).sum(-1).min(1)).reshape(blk.x.shape),
                           want, rtol=1e-6)


def test_sa_freestream_values_and_initial_field():
    """inflow_turbulence/init_SA: nu~ = 0.1 and mu_t = nu~ fv1(nu~ r/mu) with the
    reference's constants (cv1 = 7.1) on nodes -1..n+1."""
    case = _turb_case(1)
    oc = O.OracleCase(case)
    q = oc.field(0, "qke")[1:-2, 1:-2, 1:-2]          # nodes -1..n+1
    v = oc.field(0, "vist")[1:-2, 1:-2, 1:-2]
    assert np.all(q == 0.1)
    chi = 0.1
    np.testing.assert_allclose(v, chi * chi ** 3 / (chi ** 3 + 7.1 ** 3), rtol=1e-15)


def test_sst_freestream_values():
    """inflow_turbulence for SST: omega = 10 lref/lfref, k = omega mu_t/Re with
    mu_t = 1e-6, Re rescaled by lref/lfref (PreProcess.f90:994)."""
    case = _turb_case(2)
    case.params.lfref = 2.0
    case.params.lref = 0.5
    oc = O.OracleCase(case)
    re = 1e5 / 2.0 * 0.5
    np.testing.assert_allclose(oc.field(0, "qkw")[5, 5, 5], 10.0 * 0.25, rtol=1e-15)
    np.testing.assert_allclose(oc.field(0, "qke")[5, 5, 5], 2.5 * 1e-6 / re, rtol=1e-15)
    np.testing.assert_allclose(oc.field(0, "vist")[5, 5, 5], 1e-6, rtol=1e-15)


def test_sst_wall_ghosts_closed_form():
    """BC_wall_turbulence for SST: k odd in the first ghost, omega there =
    1600 mu(T_face)/(rho_face Re d_1^2) - omega_1, face values the mean of ghost
    and first interior node (the face node's rho before the wall BC sets it)."""
    case = _turb_case(2)
    oc = O.OracleCase(case)
    oc.step(2)
    blk = case.blocks[0]
    # perturb k and omega so that the ghost values are not trivially uniform
    rng = np.random.default_rng(3)
    for nm in ("qke", "qkw"):
        f = oc.field(0, nm)
        oc.set_field(0, nm, f * (1.0 + 0.1 * rng.random(f.shape)))
    r0, t0 = oc.field(0, "r").copy(), oc.field(0, "t").copy()
    k0, w0, d0 = oc.field(0, "qke").copy(), oc.field(0, "qkw").copy(), oc.field(0, "dist")
    oc.stage("boundary_all")
    k1, w1 = oc.field(0, "qke"), oc.field(0, "qkw")
    # the wall is j = 1 (G3 index 3), interior node j = 2 (index 4), ghost j = 0 (index 2)
    sl = (slice(3 + 1, 3 + blk.nk - 1), slice(None), slice(3 + 1, 3 + blk.ni - 1))   # away from the far-field edges
    face, inner, ghost = (lambda a, jj: a[sl[0], jj, sl[2]]), 4, 2
    visc = 110.4 / case.params.tref
    tm = face(t0, 3)
    mu = tm * np.sqrt(tm) * (1.0 + visc) / (tm + visc)
    want_w = 1600.0 * mu / face(r0, 3) / 1e5 / (face(d0, inner) * face(d0, inner)) - face(w0, inner)
    np.testing.assert_array_equal(face(w1, ghost), want_w)
    np.testing.assert_array_equal(face(k1, ghost), -face(k0, inner))
    np.testing.assert_array_equal(face(w1, 3), 0.5 * (face(w0, inner) + want_w))
    np.testing.assert_array_equal(face(k1, 3), 0.5 * (face(k0, inner) - face(k0, inner)))


@pytest.mark.parametrize("nlamtur", [1, 2])
def test_turbulence_models_run_and_stay_bounded(nlamtur):
    case = _turb_case(nlamtur, ni=24, nj=16, nk=10, dy_wall=1e-3)
    oc = O.OracleCase(case)
    h = oc.step(10)
    assert np.all(np.isfinite(h)) and h[-1] < h[2]
    for nm in ("qke", "vist") + (("qkw",) if nlamtur == 2 else ()):
        f = oc.field(0, nm)[3:-3, 3:-3, 3:-3]
        assert np.all(np.isfinite(f)) and f.min() >= 0.0
