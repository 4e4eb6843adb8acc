"""Synthetic multi-block cases in the reference's own terms: namelist parameters
(param.dat), PLOT3D node coordinates per block, and boundary regions exactly as
written in the .inp "1976" file (reference src/parallel.f90:852-929).

Arrays are indexed [i-1, j-1, k-1] for node coordinates and [i+2, j+2, k+2] for
ghosted ("G3", bounds -2..n+3) fields, matching the Fortran bounds of
src/PreProcess.f90:720-786.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class Params:
    """Namelist values that the hot path reads (reference src/param.dat groups
    inflow, step, flowtype, turbulence, technic)."""
    moo: float = 0.8
    reynolds: float = 1.0e5
    attack: float = 0.0
    sideslip: float = 0.0
    tref: float = 288.15
    twall: float = -1.0
    cfl: float = 5.0
    timedt_rate: float = 1.1
    efix: float = 0.1
    csrv: float = 1.0
    gama: float = 1.4
    prl: float = 0.72
    prt: float = 0.90
    nvis: int = 0
    nlamtur: int = -1
    nscheme: int = 41
    nflux: int = 1
    ntmst: int = 0


@dataclass
class Region:
    """One boundary region line of the .inp file: s_raw = (s_st1, s_ed1, s_st2,
    s_ed2, s_st3, s_ed3) with the file's signs; for bctype -1 also the target
    line t_raw, the target block nbt (1-based) and its region index ibcwin."""
    s_raw: tuple
    bctype: int
    t_raw: Optional[tuple] = None
    nbt: int = 0
    ibcwin: int = 0


@dataclass
class Block:
    ni: int
    nj: int
    nk: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    regions: List[Region] = field(default_factory=list)
    name: str = "blk"


@dataclass
class Case:
    params: Params
    blocks: List[Block]
    init: object = None          # callable(case, nb) -> (r, u, v, w, p) G3 arrays
    name: str = "case"
    seed: int = 0

    @property
    def nodes(self):
        return sum(b.ni * b.nj * b.nk for b in self.blocks)

    def freestream(self):
        """Non-dimensional freestream of src/PreProcess.f90:968-1113 (height < 0)."""
        pai = 3.14159265          # src/Globalvar.f90:236
        a = self.params.attack * pai / 180.0
        s = self.params.sideslip * pai / 180.0
        uoo = np.cos(a) * np.cos(s)
        voo = np.sin(a) * np.cos(s)
        woo = np.sin(s)
        poo = 1.0 / (self.params.gama * self.params.moo * self.params.moo)
        return 1.0, uoo, voo, woo, poo

    def initial_prim(self, nb):
        if self.init is not None:
            return self.init(self, nb)
        b = self.blocks[nb]
        roo, uoo, voo, woo, poo = self.freestream()
        shp = (b.ni + 6, b.nj + 6, b.nk + 6)
        return (np.full(shp, roo), np.full(shp, uoo), np.full(shp, voo), np.full(shp, woo),
                np.full(shp, poo))


def _periodic_regions(ni, nj, nk, i_periodic, j_bc, k_periodic, j_bc_hi=None, i_bc=None, k_bc=None):
    """Six face regions of a single block; periodic faces are self-connected
    (node 1 coincides with node n, src/Sub2.f90:206-322)."""
    regs = []
    if i_periodic:
        regs.append(Region((1, 1, 1, nj, 1, nk), -1, (ni, ni, 1, nj, 1, nk), 1, 2))
        regs.append(Region((ni, ni, 1, nj, 1, nk), -1, (1, 1, 1, nj, 1, nk), 1, 1))
    else:
        lo, hi = i_bc
        regs.append(Region((1, 1, 1, nj, 1, nk), lo))
        regs.append(Region((ni, ni, 1, nj, 1, nk), hi))
    regs.append(Region((1, ni, 1, 1, 1, nk), j_bc))
    regs.append(Region((1, ni, nj, nj, 1, nk), j_bc if j_bc_hi is None else j_bc_hi))
    if k_periodic:
        base = len(regs)
        regs.append(Region((1, ni, 1, nj, 1, 1), -1, (1, ni, 1, nj, nk, nk), 1, base + 2))
        regs.append(Region((1, ni, 1, nj, nk, nk), -1, (1, ni, 1, nj, 1, 1), 1, base + 1))
    else:
        lo, hi = k_bc
        regs.append(Region((1, ni, 1, nj, 1, 1), lo))
        regs.append(Region((1, ni, 1, nj, nk, nk), hi))
    return regs


def fourier_perturbed_init(modes, amp):
    """rho, p = freestream * (1 + amp * sum_m a_m sin(2 pi k_m.x / L + phi_m))."""
    def init(case, nb):
        b = case.blocks[nb]
        roo, uoo, voo, woo, poo = case.freestream()
        shp = (b.ni + 6, b.nj + 6, b.nk + 6)
        r = np.full(shp, roo)
        p = np.full(shp, poo)
        u = np.full(shp, uoo)
        v = np.full(shp, voo)
        w = np.full(shp, woo)
        X, Y, Z = b.x, b.y, b.z
        Lx = X.max() - X.min()
        Ly = Y.max() - Y.min()
        Lz = Z.max() - Z.min()
        pr = np.zeros_like(X)
        pp = np.zeros_like(X)
        for (kx, ky, kz, a, phr, php) in modes:
            arg = 2 * np.pi * (kx * (X - X.min()) / Lx + ky * (Y - Y.min()) / Ly + kz * (Z - Z.min()) / Lz)
            pr += a * np.sin(arg + phr)
            pp += a * np.sin(arg + php)
        r[3:-3, 3:-3, 3:-3] = roo * (1.0 + amp * pr)
        p[3:-3, 3:-3, 3:-3] = poo * (1.0 + amp * pp)
        return r, u, v, w, p
    return init


def cartesian_periodic_case(ni=64, nj=32, nk=16, seed=1, nscheme=41, nflux=1, cfl=5.0, moo=0.8,
                            stretch=1.05, lx=2.0, ly=1.0, lz=0.5, nmodes=8, amp=0.01):
    """Single Cartesian block, i and k periodic (self-connected), far field (BC 4)
    on both j faces; y = (ly/2) tanh(stretch (2 eta - 1)) / tanh(stretch).
    Inviscid Euler, freestream Mach `moo`, rho and p carry `nmodes` seeded
    Fourier modes of total amplitude `amp` (integer wavenumbers in i and k)."""
    rng = np.random.default_rng(seed)
    xi = np.arange(ni) / (ni - 1) * lx
    eta = np.arange(nj) / (nj - 1)
    yj = 0.5 * ly * np.tanh(stretch * (2 * eta - 1)) / np.tanh(stretch)
    zk = np.arange(nk) / (nk - 1) * lz
    X, Y, Z = np.meshgrid(xi, yj, zk, indexing="ij")
    blk = Block(ni, nj, nk, X.copy(), Y.copy(), Z.copy(),
                _periodic_regions(ni, nj, nk, True, 4, True), name="cart")
    a = rng.random(nmodes)
    a = a / a.sum()
    modes = [(int(rng.integers(1, 4)), float(rng.random() * 1.5), int(rng.integers(0, 3)), float(a[m]),
              float(rng.random() * 2 * np.pi), float(rng.random() * 2 * np.pi)) for m in range(nmodes)]
    prm = Params(moo=moo, cfl=cfl, nscheme=nscheme, nflux=nflux, nvis=0)
    return Case(prm, [blk], init=fourier_perturbed_init(modes, amp),
                name=f"cart{ni}x{nj}x{nk}", seed=seed)
