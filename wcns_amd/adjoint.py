"""Differentiable residual and wall force of one block, as torch.autograd
Functions over the gfx950 tangent and adjoint kernels (libwcns_amd.so:
wcns_adj_residual / wcns_adj_jvp / wcns_adj_vjp / wcns_adj_force).

Layouts (float64 tensors on the solver's GPU):
  P  (5, nk+6, nj+6, ni+6)  primitives r, u, v, w, p of nodes -2..n+3 (ghosts
                             included, as the boundary conditions leave them)
  X  (3, nk, nj, ni)        node coordinates x, y, z
  R  (5, nk, nj, ni)        residual at the interior nodes (RHS.f90:2-72, the
                             mean-flow terms; mu_t held at its stored value)

The residual's derivative w.r.t. the grid goes through the conservative
metrics (WCNSE5.f90:1478-1728) formed on the device from X.  Every call sets
the block's state on the device to (P, X).
"""
import torch

from .solver import Solver


def _blk(solver, nb):
    b = solver.case.blocks[nb]
    return b.ni, b.nj, b.nk


def state_tensors(solver, nb, device="cuda"):
    """(P, X) of block nb as the solver holds them now."""
    ni, nj, nk = _blk(solver, nb)
    P = torch.stack([torch.from_numpy(solver.field(nb, f)) for f in ("r", "u", "v", "w", "p")]).to(device)
    b = solver.case.blocks[nb]
    X = torch.stack([torch.from_numpy(a.transpose(2, 1, 0).copy()) for a in (b.x, b.y, b.z)]).to(device)
    return P.contiguous(), X.contiguous()


def _check_tensor(name, t, shape, solver):
    """t must be a contiguous float64 tensor of this shape on the solver's GPU:
    the library reads and writes through its raw address."""
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float64 GPU tensor")
    dev = getattr(solver, "device", None)
    if dev is not None and t.device.index != dev:
        raise ValueError(f"{name} is on cuda:{t.device.index}, the solver on cuda:{dev}")


def _check(P, X, solver, nb):
    ni, nj, nk = _blk(solver, nb)
    _check_tensor("P", P, (5, nk + 6, nj + 6, ni + 6), solver)
    _check_tensor("X", X, (3, nk, nj, ni), solver)


def residual(solver: Solver, nb, P, X):
    _check(P, X, solver, nb)
    ni, nj, nk = _blk(solver, nb)
    R = torch.empty((5, nk, nj, ni), dtype=torch.float64, device=P.device)
    torch.cuda.synchronize(P.device)
    solver.adj_residual(nb, P.data_ptr(), X.data_ptr(), R.data_ptr())
    return R


def jvp(solver: Solver, nb, P, X, dP=None, dX=None):
    """(R, dR): the residual and its derivative along (dP, dX)."""
    _check(P, X, solver, nb)
    ni, nj, nk = _blk(solver, nb)
    R = torch.empty((5, nk, nj, ni), dtype=torch.float64, device=P.device)
    dR = torch.empty_like(R)
    dP = None if dP is None else dP.contiguous()
    dX = None if dX is None else dX.contiguous()
    if dP is not None:
        _check_tensor("dP", dP, P.shape, solver)
    if dX is not None:
        _check_tensor("dX", dX, X.shape, solver)
    torch.cuda.synchronize(P.device)
    solver.adj_jvp(nb, P.data_ptr(), X.data_ptr(), 0 if dP is None else dP.data_ptr(),
                   0 if dX is None else dX.data_ptr(), R.data_ptr(), dR.data_ptr())
    return R, dR


def vjp(solver: Solver, nb, P, X, W):
    """(gP, gX) = ((dR/dP)^T W, (dR/dX)^T W)."""
    _check(P, X, solver, nb)
    ni, nj, nk = _blk(solver, nb)
    W = W.contiguous()
    _check_tensor("W", W, (5, nk, nj, ni), solver)
    gP = torch.empty_like(P)
    gX = torch.empty_like(X)
    torch.cuda.synchronize(P.device)
    solver.adj_vjp(nb, P.data_ptr(), X.data_ptr(), W.data_ptr(), gP.data_ptr(), gX.data_ptr())
    return gP, gX


def force(solver: Solver, nb, P, X, omega, grad=False):
    """J = omega . (pressure part of cfx, cfy, cfz, cmx, cmy, cmz) over the
    block's no-slip walls; with grad, also (dJ/dP, dJ/dX)."""
    _check(P, X, solver, nb)
    torch.cuda.synchronize(P.device)
    if not grad:
        return solver.adj_force(nb, P.data_ptr(), X.data_ptr(), omega)
    gP = torch.empty_like(P)
    gX = torch.empty_like(X)
    J = solver.adj_force(nb, P.data_ptr(), X.data_ptr(), omega, gP.data_ptr(), gX.data_ptr())
    return J, gP, gX


class Residual(torch.autograd.Function):
    """R = Residual.apply(solver, nb, P, X); backward gives dL/dP and dL/dX."""

    @staticmethod
    def forward(ctx, solver, nb, P, X):
        ctx.solver, ctx.nb = solver, nb
        ctx.save_for_backward(P, X)
        return residual(solver, nb, P.detach().contiguous(), X.detach().contiguous())

    @staticmethod
    def backward(ctx, gR):
        P, X = ctx.saved_tensors
        gP, gX = vjp(ctx.solver, ctx.nb, P.detach().contiguous(), X.detach().contiguous(), gR)
        return None, None, gP, gX


class WallForce(torch.autograd.Function):
    """J = WallForce.apply(solver, nb, P, X, omega) (a 0-d tensor)."""

    @staticmethod
    def forward(ctx, solver, nb, P, X, omega):
        J, gP, gX = force(solver, nb, P.detach().contiguous(), X.detach().contiguous(), omega, grad=True)
        ctx.save_for_backward(gP, gX)
        return torch.tensor(J, dtype=torch.float64, device=P.device)

    @staticmethod
    def backward(ctx, gJ):
        gP, gX = ctx.saved_tensors
        return None, None, gJ * gP, gJ * gX, None


# ------------------------------------------------ a whole step
# The state is the field pool's arrays of every block: (nstate, total)
# float64 tensors, block nb's (nk+6, nj+6, ni+6) ghosted nodes (i fastest)
# from offset[nb] on.  r, u, v, w, p, then for a viscous case T, then with a
# turbulence model nu~ or k, mu_t (and SST omega); the first nvar
# (Solver.state_layout) are differentiated, at the interior nodes: directions
# and weights are (nvar, total), ghost values of inputs are not read and
# those of outputs are zero.

STATE_FIELDS = ("r", "u", "v", "w", "p", "t", "qke", "vist", "qkw")


def _state_names(solver):
    prm = solver.case.params
    names = ["r", "u", "v", "w", "p"]
    if prm.nvis == 1:
        names.append("t")
    if prm.nlamtur in (1, 2):
        names += ["qke", "vist"]
    if prm.nlamtur == 2:
        names.append("qkw")
    return names


def pool_state(solver, device="cuda"):
    """The solver's current state as a pool-layout (nstate, total) tensor."""
    total, offs = solver.pool_layout()
    names = _state_names(solver)
    P = torch.zeros((len(names), total), dtype=torch.float64)
    for nb, blk in enumerate(solver.case.blocks):
        n = (blk.ni + 6) * (blk.nj + 6) * (blk.nk + 6)
        for m, f in enumerate(names):
            P[m, offs[nb]:offs[nb] + n] = torch.from_numpy(solver.field(nb, f).ravel())
    return P.to(device).contiguous()


def pool_block(solver, A, nb):
    """Block nb's part of a pool-layout tensor as (rows, nk+6, nj+6, ni+6)."""
    total, offs = solver.pool_layout()
    blk = solver.case.blocks[nb]
    n = (blk.ni + 6) * (blk.nj + 6) * (blk.nk + 6)
    return A[:, offs[nb]:offs[nb] + n].reshape(A.shape[0], blk.nk + 6, blk.nj + 6, blk.ni + 6)


def _check_pool(name, t, solver, rows):
    total, _ = solver.pool_layout()
    _check_tensor(name, t, (rows, total), solver)


def step_jvp(solver: Solver, P, dP):
    """(G(P), G'(P) dP) for one pseudo-time step G (wcns_adj_step_jvp)."""
    ns, nv = solver.state_layout()
    _check_pool("P", P, solver, ns)
    _check_pool("dP", dP, solver, nv)
    Pn, dPn = torch.empty_like(P), torch.empty_like(dP)
    torch.cuda.synchronize(P.device)
    solver.adj_step_jvp(P.data_ptr(), dP.data_ptr(), Pn.data_ptr(), dPn.data_ptr())
    return Pn, dPn


def step_vjp(solver: Solver, P, Wn):
    """G'(P)^T Wn (wcns_adj_step_vjp)."""
    ns, nv = solver.state_layout()
    _check_pool("P", P, solver, ns)
    _check_pool("Wn", Wn, solver, nv)
    gP = torch.empty_like(Wn)
    torch.cuda.synchronize(P.device)
    solver.adj_step_vjp(P.data_ptr(), Wn.data_ptr(), gP.data_ptr())
    return gP


def multistep_gradient(solver: Solver, P0, nsteps, every, omega, budget_gb=0.0):
    """J = omega . (pressure force coefficients) after nsteps steps from P0 and
    dJ/dP0, with checkpoints every `every` steps (wcns_adj_multistep); returns
    (J, gP0, GB of checkpoints and work arrays)."""
    ns, nv = solver.state_layout()
    _check_pool("P0", P0, solver, ns)
    g = torch.empty((nv, P0.shape[1]), dtype=torch.float64, device=P0.device)
    torch.cuda.synchronize(P0.device)
    J, used = solver.adj_multistep(P0.data_ptr(), nsteps, every, omega, budget_gb, g.data_ptr())
    return J, g, used


class _Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, P, solver):
        ns, nv = solver.state_layout()
        Pn, _ = step_jvp(solver, P, torch.zeros((nv, P.shape[1]), dtype=P.dtype, device=P.device))
        ctx.save_for_backward(P)
        ctx.solver = solver
        return Pn

    @staticmethod
    def backward(ctx, Wn):
        (P,) = ctx.saved_tensors
        ns, nv = ctx.solver.state_layout()
        g = torch.zeros_like(P)
        g[:nv] = step_vjp(ctx.solver, P, Wn[:nv].contiguous())
        return g, None


def step(solver: Solver, P):
    """One pseudo-time step as a differentiable function of the state (the
    turbulence model's rows carried along with zero derivative)."""
    return _Step.apply(P, solver)
