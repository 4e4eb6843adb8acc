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
