"""Tangent and adjoint of the residual and of the wall-pressure force, through
the C-ABI (device arrays) and the torch.autograd wrappers.

* primal: the residual that wcns_adj_residual forms from (P, X) equals the
  oracle's r_h_s_vis + r_h_s_invis on the same ghosted primitives (relative
  1e-12), and the solver's own stages bit for bit;
* dot-product test: <W, J (dP, dX)> = <J^T W, (dP, dX)> between the forward-mode
  tangent (wcns_adj_jvp) and the adjoint (wcns_adj_vjp), relative 1e-12;
* central finite differences (h = 1e-7) of the residual itself along seeded
  random directions of P and X against the tangent.  Node-by-node random
  directions are rough, and the WCNS weights (1/(1e-6 + IS)^2) are strongly
  nonlinear at that scale, so the plain difference quotient carries a
  truncation error that falls as h^2 (about 3e-6 of max|dR| at h = 1e-7, 3e-8
  at 1e-8, on the 24x16x12 bump): it is held to 1e-4.  Its Richardson
  extrapolation from h and h/2 cancels the h^2 term and leaves the rounding
  of the quotient, eps * max|R| / h (R sums large cancelling fluxes): held
  to 5e-6 (measured 1e-7 on the small cases, 1.2e-6 on 128x96x64).  The
  tangent and adjoint are held to each other at 1e-12 by the dot products;
* the same for J = omega . (pressure force and moment coefficients).
"""
import numpy as np
import pytest

from oracle import oracle as O
from wcns_amd import cases

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _solver(case):
    from wcns_amd.solver import Solver
    return Solver(case)


def _state(case, steps=4, nb=0):
    """A solver advanced a few steps, its ghosts refreshed from the interior."""
    from wcns_amd import adjoint as A
    s = _solver(case)
    s.step(steps, history=False)
    s.boundary_all()
    s.sync()
    P, X = A.state_tensors(s, nb)
    return s, P, X


def _directions(P, X, seed, xscale):
    """Random (dP, dX): dP per variable on the variable's scale, dX of size xscale."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    scale = P.abs().amax(dim=(1, 2, 3), keepdim=True).cpu()
    dP = (torch.randn(P.shape, generator=g, dtype=torch.float64) * scale).to(P.device)
    dX = (torch.randn(X.shape, generator=g, dtype=torch.float64) * xscale).to(X.device)
    return dP, dX


def _laminar(**kw):
    base = dict(ni=24, nj=16, nk=12, dy_wall=1e-3, growth=1.2, reynolds=1e4)
    base.update(kw)
    return cases.bump_channel_case(**base)


def _sa(**kw):
    c = _laminar(reynolds=1e5, **kw)
    c.params.nlamtur = 1
    return c


def _sst(**kw):
    c = _laminar(reynolds=1e5, **kw)
    c.params.nlamtur = 2
    return c


CASES = {
    "laminar-roe-s41": lambda: _laminar(),
    "laminar-roe-s4": lambda: _laminar(nscheme=4),
    "laminar-sw-s41": lambda: _laminar(nflux=1, ni=20, nj=18, nk=12),
    "laminar-vl-s42": lambda: _laminar(nflux=2, nscheme=42, ni=20, nj=18, nk=12),
    "sa-frozen-roe-s41": lambda: _sa(),
    "sst-frozen-roe-s41": lambda: _sst(),
}


@pytest.mark.parametrize("name", ["laminar-roe-s41", "laminar-sw-s41", "laminar-vl-s42"])
def test_adjoint_primal_is_the_oracle_residual(name):
    from wcns_amd import adjoint as A
    case = CASES[name]()
    print(f"case {name}", flush=True)
    s, P, X = _state(case)
    R = A.residual(s, 0, P, X).cpu().numpy()
    # the solver's own stages on the same state, bit for bit
    s.initial_q_dq_c_t()
    s.r_h_s_vis()
    s.r_h_s_invis()
    rs = s.field(0, "rhs")[3:-3, 3:-3, 3:-3]
    assert np.array_equal(R, rs.transpose(3, 0, 1, 2)), "adjoint primal differs from the solver's stages"
    # the oracle on the same ghosted primitives
    o = O.OracleCase(case)
    Pn = P.cpu().numpy()
    for m, f in enumerate(("r", "u", "v", "w", "p")):
        o.set_field(0, f, Pn[m])
    o.stage("initial_q_dq_c_t")
    o.stage("rhs_vis")
    o.stage("rhs_invis")
    ro = o.field(0, "dq")[3:-3, 3:-3, 3:-3].transpose(3, 0, 1, 2)
    err = np.abs(R - ro).max() / np.abs(ro).max()
    print(f"residual vs oracle rel err {err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("name", list(CASES))
def test_dot_product(name):
    from wcns_amd import adjoint as A
    case = CASES[name]()
    print(f"case {name}", flush=True)
    s, P, X = _state(case)
    for seed in (1, 2):
        dP, dX = _directions(P, X, seed, 1e-4)
        g = torch.Generator(device="cpu").manual_seed(100 + seed)
        W = torch.randn((5,) + tuple(X.shape[1:]), generator=g, dtype=torch.float64).to(P.device)
        _, dR = A.jvp(s, 0, P, X, dP, dX)
        gP, gX = A.vjp(s, 0, P, X, W)
        lhs = float((W * dR).sum())
        rp, rx = float((gP * dP).sum()), float((gX * dX).sum())
        rhs = rp + rx
        rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
        print(f"seed {seed}: <W,Jd> {lhs:.15e}  <J'W,d> {rhs:.15e} (P {rp:.3e}, X {rx:.3e})  rel {rel:.2e}",
              flush=True)
        assert rel <= 1e-12
        # the X part alone, so that a wrong metric adjoint cannot hide behind the P part
        _, dRx = A.jvp(s, 0, P, X, None, dX)
        lx = float((W * dRx).sum())
        assert abs(lx - rx) <= 1e-12 * max(abs(lx), abs(rx))


def test_multiblock_interface_block():
    """Block 1 of a 2x2 viscous wall case: its interface ghosts are inputs like
    any other; dot products at 1e-12 and Richardson differences."""
    from wcns_amd import adjoint as A
    case = cases.multiblock_case(nvis=1, wall=True, reynolds=1e4)
    s, P, X = _state(case, nb=1)
    for seed in (1, 2):
        dP, dX = _directions(P, X, seed, 1e-4)
        W = torch.randn((5,) + tuple(X.shape[1:]), dtype=torch.float64,
                        generator=torch.Generator().manual_seed(50 + seed)).to(P.device)
        _, dR = A.jvp(s, 1, P, X, dP, dX)
        gP, gX = A.vjp(s, 1, P, X, W)
        lhs = float((W * dR).sum())
        rhs = float((gP * dP).sum()) + float((gX * dX).sum())
        rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
        print(f"block 1 seed {seed}: dot-product rel {rel:.2e}", flush=True)
        assert rel <= 1e-12
    dP, dX = _directions(P, X, 9, 1e-2)
    _, dR = A.jvp(s, 1, P, X, dP, dX)

    def cd(h):
        return (A.residual(s, 1, P + h * dP, X + h * dX) - A.residual(s, 1, P - h * dP, X - h * dX)) / (2 * h)
    f1, f2 = cd(1e-7), cd(5e-8)
    rich = float(((4 * f2 - f1) / 3 - dR).abs().max() / dR.abs().max())
    print(f"block 1 Richardson rel err {rich:.2e}", flush=True)
    assert rich <= 5e-6


def _fd_check(s, A, P, X, dP, dX, h=1e-7):
    """(plain central difference at h, its Richardson extrapolation from h and
    h/2), each as max|FD - dR| / max|dR|."""
    _, dR = A.jvp(s, 0, P, X, dP, dX)

    def cd(hh):
        Rp = A.residual(s, 0, P + hh * dP, X + hh * dX)
        Rm = A.residual(s, 0, P - hh * dP, X - hh * dX)
        return (Rp - Rm) / (2 * hh)
    f1, f2 = cd(h), cd(h / 2)
    rich = (4.0 * f2 - f1) / 3.0
    scale = dR.abs().max()
    return float((f1 - dR).abs().max() / scale), float((rich - dR).abs().max() / scale)


@pytest.mark.parametrize("name", list(CASES))
def test_finite_differences(name):
    from wcns_amd import adjoint as A
    case = CASES[name]()
    print(f"case {name}", flush=True)
    s, P, X = _state(case)
    worst = np.zeros(2)
    for seed in range(8):
        dP, dX = _directions(P, X, 10 + seed, 1e-2)
        worst = np.maximum(worst, _fd_check(s, A, P, X, dP, dX))
    print(f"max rel err vs tangent: central FD {worst[0]:.3e}, Richardson {worst[1]:.3e}", flush=True)
    assert worst[0] <= 1e-4 and worst[1] <= 5e-6


def test_sa_wall_block_128x96x64_200_directions():
    """The adjoint workload: a 128x96x64 wall-bounded curvilinear block, SA
    (frozen mu_t), Mach 0.7; FD (h = 1e-7) on 200 seeded directions and two
    dot-product tests."""
    from wcns_amd import adjoint as A
    case = cases.bump_channel_case(ni=128, nj=96, nk=64, dy_wall=1e-4, growth=1.08, reynolds=3e6, moo=0.7)
    case.params.nlamtur = 1
    s, P, X = _state(case, steps=3)
    worst = np.zeros(2)
    for seed in range(200):
        dP, dX = _directions(P, X, 1000 + seed, 1e-2)
        worst = np.maximum(worst, _fd_check(s, A, P, X, dP, dX))
        if seed % 50 == 49:
            print(f"{seed + 1} directions: central FD {worst[0]:.3e}, Richardson {worst[1]:.3e}", flush=True)
    assert worst[0] <= 1e-4 and worst[1] <= 5e-6
    for seed in (1, 2):
        dP, dX = _directions(P, X, seed, 1e-5)
        W = torch.randn((5,) + tuple(X.shape[1:]), dtype=torch.float64,
                        generator=torch.Generator().manual_seed(seed)).to(P.device)
        _, dR = A.jvp(s, 0, P, X, dP, dX)
        gP, gX = A.vjp(s, 0, P, X, W)
        lhs = float((W * dR).sum())
        rhs = float((gP * dP).sum()) + float((gX * dX).sum())
        rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
        print(f"dot-product rel {rel:.2e}", flush=True)
        assert rel <= 1e-12


OMEGA = np.array([0.3, -1.1, 0.2, 0.05, -0.4, 0.7])


def test_force_primal_is_the_solvers_force():
    """Inviscid walls: wcns_force is the pressure part alone."""
    from wcns_amd import adjoint as A
    case = _laminar()
    case.params.nvis = 0
    s, P, X = _state(case)
    J = A.force(s, 0, P, X, OMEGA)
    c = s.force()[:6]
    assert abs(J - float(OMEGA @ c)) <= 1e-13 * float(np.abs(OMEGA) @ np.abs(c))


@pytest.mark.parametrize("name", ["laminar-roe-s41", "sa-frozen-roe-s41"])
def test_force_gradient(name):
    from wcns_amd import adjoint as A
    case = CASES[name]()
    s, P, X = _state(case)
    J, gP, gX = A.force(s, 0, P, X, OMEGA, grad=True)
    h = 1e-7
    for seed in range(4):
        dP, dX = _directions(P, X, 50 + seed, 1e-2)
        lin = float((gP * dP).sum()) + float((gX * dX).sum())
        fd = (A.force(s, 0, P + h * dP, X + h * dX, OMEGA) - A.force(s, 0, P - h * dP, X - h * dX, OMEGA)) / (2 * h)
        rel = abs(fd - lin) / max(abs(lin), 1e-300)
        print(f"seed {seed}: dJ {lin:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= 1e-6


def test_autograd_wrappers():
    from wcns_amd import adjoint as A
    case = CASES["laminar-roe-s41"]()
    s, P, X = _state(case)
    W = torch.randn((5,) + tuple(X.shape[1:]), dtype=torch.float64,
                    generator=torch.Generator().manual_seed(7)).to(P.device)
    Pg, Xg = P.clone().requires_grad_(True), X.clone().requires_grad_(True)
    R = A.Residual.apply(s, 0, Pg, Xg)
    (R * W).sum().backward()
    gP, gX = A.vjp(s, 0, P, X, W)
    assert torch.equal(Pg.grad, gP) and torch.equal(Xg.grad, gX)
    Pg.grad = None
    Xg.grad = None
    J = A.WallForce.apply(s, 0, Pg, Xg, OMEGA)
    (2.0 * J).backward()
    _, fP, fX = A.force(s, 0, P, X, OMEGA, grad=True)
    assert torch.equal(Pg.grad, 2.0 * fP) and torch.equal(Xg.grad, 2.0 * fX)
