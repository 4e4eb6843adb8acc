"""ctypes driver of _ref/libref_step.so -- TEST INFRASTRUCTURE ONLY.

The library is the reference's own serial step routines (boundary_all,
initial_q_dq_c_t, timestep_tgh, r_h_s_vis, r_h_s_invis, l_h_s_tgh,
residual_curve and the metrics of GRID_DERIVATIVE_gcl), compiled from
/root/reference/src by oracle/Makefile (target `ref`) together with
ref_step_bridge.f90.  It exists only where the reference sources are mounted
and is never shipped; tests skip when it is absent.

The Fortran module state is global to the process, so one RefStep case is live
at a time (a new one replaces the previous).
"""
import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "_ref", "libref_step.so")

STAGES = {"boundary_all": 0, "initial_q_dq_c_t": 1, "timestep": 2, "rhs_vis": 3, "rhs_invis": 4, "lhs": 5}
FIELDS = {"r": 0, "u": 1, "v": 2, "w": 3, "p": 4, "t": 5, "c": 6, "dtdt": 7, "sra": 8, "srb": 9, "src": 10,
          "srva": 11, "srvb": 12, "srvc": 13, "visl": 14, "vist": 15, "vol": 16, "kcx": 17, "kcy": 18,
          "kcz": 19, "etx": 20, "ety": 21, "etz": 22, "ctx": 23, "cty": 24, "ctz": 25, "qke": 26,
          "dist": 27, "qkw": 28, "dq": 100, "q": 101}
PAI = 3.14159265          # Globalvar.f90:236

_lib = None


def available():
    return os.path.exists(LIBPATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIBPATH, mode=os.RTLD_LAZY | os.RTLD_GLOBAL if hasattr(os, "RTLD_LAZY") else 0)
        dptr, iptr = C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.ref_step_init.argtypes = [dptr, iptr]
        L.ref_step_block.argtypes = [C.c_int, iptr, dptr, dptr, dptr]
        L.ref_step_region.argtypes = [C.c_int, C.c_int, C.c_int, iptr, C.c_int, iptr, C.c_int, C.c_char_p]
        L.ref_step_bcindex.argtypes = []
        L.ref_step_metrics.argtypes = []
        L.ref_step_turb_init.argtypes = []
        L.ref_step_initial_t.argtypes = []
        L.ref_step_turb_inflow.argtypes = [dptr]
        L.ref_step_images.argtypes = [C.c_int, C.c_int, iptr, iptr, iptr]
        L.ref_step_field.argtypes = [C.c_int, C.c_int, C.c_int, dptr]
        L.ref_step_poison.argtypes = [C.c_long, C.c_double]
        L.ref_step_stage.argtypes = [C.c_int]
        L.ref_step_residual.argtypes = [dptr, iptr]
        L.ref_step_force.argtypes = [dptr, C.c_int, C.c_char_p, dptr]
        L.ref_step_tecout.argtypes = [C.c_int, C.c_char_p]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class RefStep:
    """A wcns_amd.cases.Case loaded into the reference's module variables.

    The metrics come from the reference's GRID_DERIVATIVE_gcl; the initial
    primitives (ghosts included) are the case's; with a turbulence model the
    uniform initial field is the reference's init_SA / init_SST and the wall
    distance is taken from `dist` (per block, G3 frame) when given."""

    def __init__(self, case, dist=None):
        L = lib()
        prm = case.params
        self.case = case
        att = prm.attack * PAI / 180.0
        ssl = prm.sideslip * PAI / 180.0
        poo = 1.0 / (prm.gama * prm.moo * prm.moo)
        scal = np.array([prm.moo, prm.reynolds / prm.lfref * prm.lref, att, ssl, prm.tref, prm.twall, prm.cfl,
                         prm.timedt_rate, prm.efix, prm.csrv, prm.gama, prm.prl, prm.prt,
                         1.0, np.cos(att) * np.cos(ssl), np.sin(att) * np.cos(ssl), np.sin(ssl), poo,
                         np.sqrt(prm.gama * poo / 1.0), 1.0, 110.4 / prm.tref, prm.lfref, prm.lref,
                         prm.kmaxlim, prm.timedt_turb, prm.xk, prm.xb], dtype=np.float64)
        iscal = np.array([len(case.blocks), prm.nvis, prm.nlamtur, prm.nscheme, prm.nflux, prm.ntmst],
                         dtype=np.int32)
        assert L.ref_step_init(_d(scal), _i(iscal)) == 0
        for nb, blk in enumerate(case.blocks):
            dims = np.array([blk.ni, blk.nj, blk.nk], dtype=np.int32)
            x, y, z = (np.ascontiguousarray(a.transpose(2, 1, 0)).ravel() for a in (blk.x, blk.y, blk.z))
            assert L.ref_step_block(nb + 1, _i(dims), _d(x), _d(y), _d(z)) == 0
        with tempfile.TemporaryDirectory() as td:
            fn = os.path.join(td, "bcline.txt").encode()
            for nb, blk in enumerate(case.blocks):
                for nr, reg in enumerate(blk.regions):
                    s = np.array(reg.s_raw, dtype=np.int32)
                    t = np.array(reg.t_raw if reg.t_raw is not None else [0] * 6, dtype=np.int32)
                    assert L.ref_step_region(nb + 1, nr + 1, len(blk.regions), _i(s), reg.bctype, _i(t),
                                             reg.nbt, fn) == 0
        assert L.ref_step_bcindex() == 0
        assert L.ref_step_metrics() == 0
        for nb in range(len(case.blocks)):
            for name, a in zip(("r", "u", "v", "w", "p"), case.initial_prim(nb)):
                self.set_field(nb, name, np.ascontiguousarray(np.asarray(a, dtype=np.float64).transpose(2, 1, 0)))
        if prm.nvis == 1:
            assert L.ref_step_initial_t() == 0
        if prm.nlamtur > 0:
            assert L.ref_step_turb_init() == 0
            if dist is not None:
                for nb in range(len(case.blocks)):
                    self.set_field(nb, "dist", dist[nb])

    def _shape(self, nb):
        b = self.case.blocks[nb]
        return (b.nk + 6, b.nj + 6, b.ni + 6)

    def field(self, nb, name):
        """[k+2, j+2, i+2] array (the oracle's frame); nodes outside the
        reference's bounds of that field are 0."""
        shp = self._shape(nb)
        if name in ("dq", "q"):
            out = np.zeros(shp + (5,))
        else:
            out = np.zeros(shp)
        assert lib().ref_step_field(nb + 1, FIELDS[name], 0, _d(out)) == 0
        return out

    def set_field(self, nb, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert lib().ref_step_field(nb + 1, FIELDS[name], 1, _d(a)) == 0

    def stage(self, name):
        assert lib().ref_step_stage(STAGES[name]) == 0

    def poison(self, nb, value):
        """Leave `value` in the heap block that the next r_h_s_invis with
        flux 4 allocates for block nb's pressure ratios (rpres), so that the
        entries the reference reads without writing hold it."""
        blk = self.case.blocks[nb]
        n = 3 * (blk.ni + 3) * (blk.nj + 3) * (blk.nk + 3)
        assert lib().ref_step_poison(n, value) == 0

    def residual(self):
        out = np.zeros(2)
        loc = np.zeros(5, dtype=np.int32)
        assert lib().ref_step_residual(_d(out), _i(loc)) == 0
        return out[0], out[1], loc

    def force(self):
        """cfx, cfy, cfz, cmx, cmy, cmz, cd, cl, xcp of the reference's force."""
        prm = self.case.params
        fref = np.array([prm.sref, prm.xref, prm.yref, prm.zref], dtype=np.float64)
        out = np.zeros(9)
        with tempfile.TemporaryDirectory() as td:
            fn = os.path.join(td, "force.dat").encode()
            assert lib().ref_step_force(_d(fref), prm.nwholefield, fn, _d(out)) == 0
        return out

    def tecout(self, kind, tecname):
        """The reference's Tecplot writer `kind` (0 surface, 1 wall, 2 volume)
        on the loaded state, with tecname set to `tecname` (a path); returns
        the path of the file it wrote."""
        assert lib().ref_step_tecout(kind, tecname.encode()) == 0
        stem = tecname[:tecname.rfind(".")] if "." in os.path.basename(tecname) else tecname
        return tecname if kind == 0 else stem + ("_wall.plt" if kind == 1 else "_vol.plt")

    def turb_inflow(self):
        out = np.zeros(4)
        lib().ref_step_turb_inflow(_d(out))
        return out

    def images(self, nb, nr):
        reg = self.case.blocks[nb].regions[nr]
        s = np.abs(np.array(reg.s_raw)).reshape(3, 2)
        n = int(np.prod(s.max(1) - s.min(1) + 1))
        img, jmg, kmg = (np.zeros(n, np.int32) for _ in range(3))
        assert lib().ref_step_images(nb + 1, nr + 1, _i(img), _i(jmg), _i(kmg)) == 0
        return img, jmg, kmg
