"""Tangent and adjoint of a whole pseudo-time step, and the multi-step adjoint
with checkpoints (wcns_adj_step_jvp / wcns_adj_step_vjp / wcns_adj_multistep),
through the C-ABI, on inviscid cases and on viscous ones (laminar, and SA /
SST with the turbulence model held: frozen mu_t):

* the primal of the differentiated step equals the solver's own step
  (wcns_step) bit for bit, and does not read the ghost values of its input;
* dot-product test <Wn, G' dP> = <G'^T Wn, dP> at 1e-12 (the tangent runs the
  solver's arithmetic in dual numbers; the adjoint is the hand-transposed
  sweeps, averagings, update and boundary conditions, and the residual's own
  adjoint), on periodic / far-field, multi-block inviscid-wall (Mach 2),
  inflow / outflow / symmetry / wall cases with multi-point interface groups,
  and their viscous counterparts (no-slip wall reading the lagged T, viscous
  radii in the diagonal and the LU sweeps);
* the tangent against central differences of the solver's own step (inviscid
  and laminar: with a turbulence model the solver's step moves mu_t too),
  one step and a chain of three (Richardson extrapolation of h = 1e-7 and h/2: the
  differences' error falls as h^2, 2.2e-4 at h = 1e-5 to 2.1e-8 at 1e-7, the
  WCNS weights being strongly nonlinear at the scale of these directions);
* the 5-step reverse pass of the wall force (Euler and laminar bump
  channels): checkpoints every 2 steps give the store-every-step gradient bit
  for bit, J equals wcns_force's pressure coefficient, and <dJ/dP0, d>
  matches central differences of J; a too-small budget and the FP32 operator
  are rejected; a 64-block SST case runs 10 steps back.
"""
import numpy as np
import pytest

from wcns_amd import cases

pytestmark = pytest.mark.gpu


def _interior_mask(solver, total):
    import torch
    _, offs = solver.pool_layout()
    mask = torch.zeros(total, dtype=torch.bool)
    for nb, b in enumerate(solver.case.blocks):
        m = torch.zeros((b.nk + 6, b.nj + 6, b.ni + 6), dtype=torch.bool)
        m[3:-3, 3:-3, 3:-3] = True
        n = m.numel()
        mask[offs[nb]:offs[nb] + n] = m.ravel()
    return mask


def _direction(solver, P, seed, smooth=False):
    """A direction on the interior nodes of the differentiated rows (nvar),
    scaled per variable like the state."""
    import torch
    total = P.shape[1]
    nv = solver.state_layout()[1]
    mask = _interior_mask(solver, total).to(P.device)
    g = torch.Generator(device="cpu").manual_seed(seed)
    if smooth:
        _, offs = solver.pool_layout()
        d = torch.zeros((nv, total), dtype=torch.float64)
        for nb, b in enumerate(solver.case.blocks):
            K, J, I = torch.meshgrid(torch.arange(b.nk + 6), torch.arange(b.nj + 6), torch.arange(b.ni + 6),
                                     indexing="ij")
            n = (b.ni + 6) * (b.nj + 6) * (b.nk + 6)
            for m in range(nv):
                ph = torch.rand(3, generator=g) * 6.283
                f = (torch.sin(6.283 * I / (b.ni + 5) + ph[0]) * torch.cos(3.1416 * J / (b.nj + 5) + ph[1])
                     * torch.cos(6.283 * K / (b.nk + 5) + ph[2]))
                d[m, offs[nb]:offs[nb] + n] = f.ravel().double()
    else:
        d = torch.rand((nv, total), generator=g, dtype=torch.float64) - 0.5
    d = d.to(P.device)
    scale = torch.tensor([1.0, 1.0, 1.0, 1.0, 2.0, 1.0][:nv], dtype=torch.float64, device=P.device)[:, None]
    return (d * scale * mask).contiguous(), mask


def _set_pool(solver, P):
    from wcns_amd.adjoint import _state_names, pool_block
    Pc = P.cpu()
    for nb in range(len(solver.case.blocks)):
        blk = pool_block(solver, Pc, nb).numpy()
        for m, f in enumerate(_state_names(solver)):
            solver.set_field(nb, f, blk[m])


def _euler_cases():
    a = cases.wavy_periodic_case(ni=24, nj=16, nk=12)
    b = cases.multiblock_case(nx=(14, 12), ny=(12,), nz=(12,), nvis=0, wall=True, moo=2.0, nflux=3)
    c = cases.multiblock_case(nx=(12, 13), ny=(12, 13), nz=(12,), nvis=0, wall=True, moo=0.6, nflux=3,
                              periodic_i=False, ilo_bc=5, ihi_bc=6, jhi_bc=3)
    return {"periodic-farfield-sw": a, "wall-mach2-2blocks": b, "inflow-outflow-symmetry-4blocks": c}


def _viscous_cases():
    # a slip top: over a far-field face that the flow runs along, boundary_4
    # switches between its inflow and outflow values where the normal velocity
    # changes sign, and central differences across that switch measure the
    # jump, not the derivative
    d = cases.bump_channel_case(ni=24, nj=24, nk=12, dy_wall=1e-3, growth=1.15, reynolds=1e4, top_bc=3)
    out = {"laminar-bump": d}
    for model, tag in ((-1, "laminar"), (1, "sa-frozen"), (2, "sst-frozen")):
        c = cases.multiblock_case(nx=(12, 13), ny=(12, 13), nz=(12,), nvis=1, wall=True, moo=0.6, nflux=3,
                                  periodic_i=False, ilo_bc=5, ihi_bc=6, jhi_bc=3, reynolds=1e4)
        c.params.nlamtur = model
        out[f"{tag}-4blocks-wall-symmetry"] = c
    return out


CASES = {**_euler_cases(), **_viscous_cases()}
FD_CASES = ["periodic-farfield-sw", "inflow-outflow-symmetry-4blocks", "laminar-bump",
            "laminar-4blocks-wall-symmetry"]


def _advanced(case, n=3):
    """A solver whose state is n steps into the run (so that every term is live)."""
    from wcns_amd.solver import Solver
    s = Solver(case)
    s.step(n, history=False)
    return s


@pytest.mark.parametrize("name", list(CASES))
def test_step_primal_equals_solver_step(name):
    import torch
    from wcns_amd.adjoint import pool_state, step_jvp
    case = CASES[name]
    print(f"step primal, {name}", flush=True)
    s = _advanced(case)
    P = pool_state(s)
    ref = _advanced(case)
    ref.step(1, history=False)
    Pref = pool_state(ref)
    nv = s.state_layout()[1]
    zero = torch.zeros((nv, P.shape[1]), dtype=torch.float64, device=P.device)
    Pn, _ = step_jvp(s, P, zero)
    mask = _interior_mask(s, P.shape[1]).to(P.device)
    assert torch.equal(Pn[:, mask], Pref[:, mask])
    # the input's ghost values of r..p (and T) are not part of the state
    Pg = P.clone()
    Pg[:nv, ~mask] = 7.0
    Pn2, _ = step_jvp(s, Pg, zero)
    assert torch.equal(Pn2[:, mask], Pn[:, mask])
    s.close()
    ref.close()


@pytest.mark.parametrize("name", list(CASES))
def test_step_dot_product(name):
    import torch
    from wcns_amd.adjoint import pool_state, step_jvp, step_vjp
    case = CASES[name]
    print(f"step dot product, {name}", flush=True)
    s = _advanced(case)
    P = pool_state(s)
    worst = 0.0
    for seed in range(3):
        dP, mask = _direction(s, P, seed)
        Wn, _ = _direction(s, P, 100 + seed)
        _, dPn = step_jvp(s, P, dP)
        gP = step_vjp(s, P, Wn)
        a = float((Wn * dPn).sum())
        b = float((gP * dP).sum())
        rel = abs(a - b) / max(abs(a), abs(b))
        worst = max(worst, rel)
        assert torch.count_nonzero(gP[:, ~mask]) == 0
    print(f"dot product rel diff {worst:.2e}", flush=True)
    assert worst <= 1e-12
    s.close()


def _fd_step(case, P, d, h, nsteps):
    """Central difference of nsteps solver steps along d, Richardson h, h/2."""
    from wcns_amd.adjoint import pool_state
    from wcns_amd.solver import Solver
    out = []
    for hh in (h, h / 2):
        res = []
        for sgn in (1, -1):
            s = Solver(case)
            Pp = P.clone()
            Pp[:d.shape[0]] += sgn * hh * d
            _set_pool(s, Pp)
            s.step(nsteps, history=False)
            res.append(pool_state(s))
            s.close()
        out.append((res[0] - res[1]) / (2 * hh))
    return (4 * out[1] - out[0]) / 3


@pytest.mark.parametrize("name", FD_CASES)
def test_step_tangent_matches_differences(name):
    from wcns_amd.adjoint import pool_state, step_jvp
    case = CASES[name]
    print(f"step tangent vs differences, {name}", flush=True)
    s = _advanced(case)
    P = pool_state(s)
    d, mask = _direction(s, P, 7, smooth=True)
    if name == "laminar-bump":
        # the bump channel's flow has w = 0 (flat in z): there the radius's
        # |V . grad zeta| has no derivative, and central differences across
        # the kink measure its mean; a direction that leaves w alone stays on it
        d[3] = 0.0
    # one step and a chain of three
    Pc, dc = P, d
    for n in range(3):
        Pc, dc = step_jvp(s, Pc, dc)
        if n in (0, 2):
            fd = _fd_step(case, P, d, 1e-7, n + 1)
            fd = fd[:dc.shape[0]]
            err = float((dc - fd)[:, mask].abs().max() / dc[:, mask].abs().max())
            print(f"{n + 1} step(s): tangent vs differences {err:.2e}", flush=True)
            assert err <= 1e-6
    s.close()


def _bump(viscous):
    case = cases.bump_channel_case(ni=32, nj=24, nk=12, dy_wall=1e-3, growth=1.15, top_bc=3, reynolds=1e4)
    if not viscous:
        case.params.nvis = 0
    return case


@pytest.mark.parametrize("viscous", [False, True], ids=["euler", "laminar"])
def test_multistep_checkpoints_force_and_differences(viscous):
    import torch
    from wcns_amd.adjoint import multistep_gradient, pool_state
    from wcns_amd.solver import Solver, WcnsError
    case = _bump(viscous)
    print(f"multistep adjoint, bump channel, viscous {viscous}", flush=True)
    s = _advanced(case, 2)
    P0 = pool_state(s)
    om = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0]
    J1, g1, used1 = multistep_gradient(s, P0, 5, 1, om)
    J2, g2, used2 = multistep_gradient(s, P0, 5, 2, om)
    print(f"J {J1:.12e}, store-all {used1:.3f} GB, every 2 {used2:.3f} GB", flush=True)
    assert J1 == J2 and torch.equal(g1, g2)
    assert used2 < used1
    # J is the pressure part of wcns_force's cfy after 5 steps
    ref = Solver(case)
    _set_pool(ref, P0)
    ref.step(5, history=False)
    if not viscous:
        cfy = ref.force()[1]
        assert abs(J1 - cfy) <= 1e-13 * max(1.0, abs(cfy))
    ref.close()
    # <dJ/dP0, d> against central differences of J (Richardson)
    d, mask = _direction(s, P0, 11, smooth=True)
    vals = []
    for h in (1e-7, 5e-8):
        Pp, Pm = P0.clone(), P0.clone()
        Pp[:d.shape[0]] += h * d
        Pm[:d.shape[0]] -= h * d
        jp, _, _ = multistep_gradient(s, Pp.contiguous(), 5, 5, om)
        jm, _, _ = multistep_gradient(s, Pm.contiguous(), 5, 5, om)
        vals.append((jp - jm) / (2 * h))
    fd = (4 * vals[1] - vals[0]) / 3
    ad = float((g1 * d).sum())
    print(f"dJ: adjoint {ad:.10e}, differences {fd:.10e}", flush=True)
    assert abs(ad - fd) <= 1e-6 * abs(ad)
    with pytest.raises(WcnsError) as ei:
        multistep_gradient(s, P0, 5, 2, om, budget_gb=1e-6)
    assert ei.value.code == -1
    s.close()


def test_multistep_sst_64_blocks_runs():
    """The 64-block SST configuration at reduced block sides (12-20 nodes,
    Re 1e7, M 0.85, frozen mu_t): 10 steps back with checkpoints every 5; the
    gradient is finite, zero on ghosts, and the budget check reports the
    memory."""
    import torch
    from wcns_amd.adjoint import multistep_gradient, pool_state
    from wcns_amd.solver import Solver
    case = cases.uneven_blocks_case(nblk=(4, 4, 4), lo=12, hi=20, seed=11, nvis=1, wall=True, nflux=3,
                                    reynolds=1e7, moo=0.85, periodic_i=False, ilo_bc=5, ihi_bc=6)
    case.params.nlamtur = 2
    print(f"multistep adjoint, 64 blocks SST, {case.nodes} nodes", flush=True)
    s = Solver(case)
    s.step(3, history=False)
    P0 = pool_state(s)
    J, g, used = multistep_gradient(s, P0, 10, 5, [0.0, 1.0, 0.0, 0.0, 0.0, 0.0])
    mask = _interior_mask(s, P0.shape[1]).to(P0.device)
    print(f"J {J:.6e}, |g| {float(g.norm()):.6e}, {used:.3f} GB", flush=True)
    assert bool(torch.isfinite(g).all()) and float(g.norm()) > 0.0
    assert torch.count_nonzero(g[:, ~mask]) == 0
    s.close()


def test_step_adjoint_rejects_fp32_operator():
    import torch
    from wcns_amd.adjoint import pool_state, step_jvp
    from wcns_amd.solver import Solver, WcnsError
    case = _bump(True)
    s = Solver(case, lhs_precision=1)
    P = pool_state(s)
    with pytest.raises(WcnsError) as ei:
        step_jvp(s, P, torch.zeros((6, P.shape[1]), dtype=torch.float64, device=P.device))
    assert ei.value.code == -9
    s.close()
