"""Writes tests/golden/tecplot_golden.npz: the Tecplot files that the
reference's own tecout_bin_surface, tecout_bin_wall and tecout_bin_volume
(PostProcess.f90:97-772, with mod_tecios, parallel.f90:1-173) write for a
seeded state of a two-block SA case with a no-slip wall, a symmetry top,
inflow / outflow and a block interface.  The routines run in
oracle/_ref/libref_step.so (the reference's serial step routines compiled
from source by `make -C oracle ref`), after the reference's boundary_all and
initial_q_dq_c_t on the state of `tecplot_state`.

Stored: the bytes of each file (uint8 arrays surface, wall, volume).

Run from the repository root: python tests/golden/make_golden_tecplot.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from wcns_amd import cases  # noqa: E402


def tecplot_case():
    case = cases.multiblock_case(nx=(12, 13), ny=(12,), nz=(12,), nvis=1, wall=True, periodic_i=False,
                                 ilo_bc=5, ihi_bc=6, jhi_bc=3, reynolds=1.0e5)
    case.params.nlamtur = 1
    return case


def tecplot_state(case, nb):
    """Seeded smooth perturbation of the freestream, every node of block nb
    including the ghosts, in the [k+2, j+2, i+2] frame: r, u, v, w, p, nu~, mu_t."""
    b = case.blocks[nb]
    k, j, i = np.meshgrid(np.arange(b.nk + 6), np.arange(b.nj + 6), np.arange(b.ni + 6), indexing="ij")
    rng = np.random.default_rng(17 + nb)
    ph = rng.random(8) * 2 * np.pi
    s = lambda a, q: np.sin(0.37 * i * (q + 1) + 0.29 * j + 0.23 * k * (2 - q) + ph[q])
    r0, u0, v0, w0, p0 = (np.asarray(a, dtype=np.float64).transpose(2, 1, 0) for a in case.initial_prim(nb))
    r = r0 * (1.0 + 0.04 * s(None, 0))
    u = u0 * (1.0 + 0.10 * s(None, 1))
    v = v0 + 0.05 * s(None, 2)
    w = w0 + 0.03 * s(None, 3)
    p = p0 * (1.0 + 0.06 * s(None, 4))
    nut = 1.0e-3 * (1.5 + s(None, 5))
    mut = 2.0e-2 * (1.5 + s(None, 6))
    return [np.ascontiguousarray(a) for a in (r, u, v, w, p, nut, mut)]


def main():
    from oracle import ref_step as R
    if not R.available():
        raise SystemExit("oracle/_ref/libref_step.so is not built (make -C oracle ref)")
    case = tecplot_case()
    ref = R.RefStep(case)
    for nb in range(len(case.blocks)):
        for name, a in zip(("r", "u", "v", "w", "p", "qke", "vist"), tecplot_state(case, nb)):
            ref.set_field(nb, name, a)
    ref.stage("boundary_all")
    ref.stage("initial_q_dq_c_t")
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for kind, name in enumerate(("surface", "wall", "volume")):
            path = ref.tecout(kind, os.path.join(td, "flow.plt"))
            out[name] = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "tecplot_golden.npz"), **out)
    print({k: v.size for k, v in out.items()})


if __name__ == "__main__":
    main()
