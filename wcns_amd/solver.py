"""Host-side mirror of the reference's hot-path interface over the C-ABI of
libwcns_amd.so (include/wcns_amd.h).

The reference drives one pseudo-time step from src/Main.f90:85-109 by calling
argument-less subroutines over module globals.  `Solver` keeps the same names
and order -- boundary_all, initial_q_dq_c_t, timestep_tgh, r_h_s_vis,
communicate_dq_npp, r_h_s_invis, l_h_s_tgh, communicate_pv_npp,
residual_curve -- each one a call into the GPU library on device-resident
state.  Errors the reference turns into `stop` raise WcnsError here.

There is no CPU fallback: if the gfx950 library is missing, construction fails.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libwcns_amd.so")

E_NAMES = {-1: "WCNS_E_ARG", -2: "WCNS_E_STATE", -3: "WCNS_E_HIP", -4: "WCNS_E_NOMEM", -5: "WCNS_E_BC",
           -6: "WCNS_E_JACOBIAN", -7: "WCNS_E_ROE", -8: "WCNS_E_REPAIR", -9: "WCNS_E_UNSUPPORTED"}

FIELDS = {"r": 0, "u": 1, "v": 2, "w": 3, "p": 4, "t": 5, "dtdt": 7, "sra": 8, "srb": 9, "src": 10,
          "srva": 11, "srvb": 12, "srvc": 13, "visl": 14, "vol": 16, "kcx": 17, "kcy": 18, "kcz": 19,
          "etx": 20, "ety": 21, "etz": 22, "ctx": 23, "cty": 24, "ctz": 25, "dq": 100, "rhs": 102,
          "vist": 15, "qke": 26, "dist": 27, "qkw": 28}

EXPORTS = ["wcns_create", "wcns_destroy", "wcns_last_error", "wcns_add_block", "wcns_add_region", "wcns_setup",
           "wcns_set_lhs_precision", "wcns_set_prim", "wcns_get_field", "wcns_set_field", "wcns_boundary_all",
           "wcns_initial_q_dq_c_t", "wcns_timestep", "wcns_r_h_s_vis", "wcns_r_h_s_invis", "wcns_communicate_dq",
           "wcns_l_h_s", "wcns_communicate_pv", "wcns_residual", "wcns_step", "wcns_get_counter", "wcns_sync",
           "wcns_region_images", "wcns_multipoint", "wcns_timing", "wcns_timing_get", "wcns_force",
           "wcns_step_graph", "wcns_autotune", "wcns_adj_residual", "wcns_adj_jvp", "wcns_adj_vjp",
           "wcns_adj_force", "wcns_sweep_layout", "wcns_residual_history",
           "wcns_residual_noalloc", "wcns_force_record", "wcns_force_history", "wcns_save_flow", "wcns_load_flow",
           "wcns_pool_layout", "wcns_adj_step_jvp", "wcns_adj_step_vjp", "wcns_adj_multistep", "wcns_tecplot"]

TIMING_SLOTS = ["bc", "init", "timestep", "vis_deriv", "rhs_inv_i", "rhs_inv_j", "rhs_inv_k", "comm_dq",
                "sweep_lu_l", "sweep_lu_u", "sweep_gs_l", "sweep_gs_u", "update", "comm_pv", "residual",
                "sweep_stage", "rhs_vis_i", "rhs_vis_j", "rhs_vis_k", "sa_rhs", "sa_sweep_l", "sa_sweep_u",
                "sa_update", "transpose", "vis_rows"]


class WcnsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{E_NAMES.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("moo", "reynolds", "attack", "sideslip", "tref", "twall",
                                          "cfl", "timedt_rate", "efix", "csrv", "gama", "prl", "prt")] + \
               [(n, C.c_int) for n in ("nvis", "nlamtur", "nscheme", "nflux", "ntmst")] + \
               [(n, C.c_double) for n in ("sref", "lfref", "xref", "yref", "zref")] + [("nwholefield", C.c_int)] + \
               [(n, C.c_double) for n in ("timedt_turb", "kmaxlim", "xk", "xb", "lref")]


_lib = None


def load_library(path=LIBPATH):
    """Load libwcns_amd.so (built by wcns_amd/build.py); raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"gfx950 library not built: {path} (run python -m wcns_amd.build)")
    L = C.CDLL(path)
    dptr, iptr, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
    L.wcns_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(Params)]
    L.wcns_destroy.argtypes = [vp]
    L.wcns_last_error.restype = C.c_char_p
    L.wcns_last_error.argtypes = [vp]
    L.wcns_add_block.argtypes = [vp, C.c_int, C.c_int, C.c_int, dptr, dptr, dptr, iptr]
    L.wcns_add_region.argtypes = [vp, C.c_int, iptr, C.c_int, iptr, C.c_int, C.c_int]
    L.wcns_setup.argtypes = [vp]
    L.wcns_set_lhs_precision.argtypes = [vp, C.c_int]
    L.wcns_set_prim.argtypes = [vp, C.c_int, dptr, dptr, dptr, dptr, dptr]
    L.wcns_get_field.argtypes = [vp, C.c_int, C.c_int, dptr]
    L.wcns_set_field.argtypes = [vp, C.c_int, C.c_int, dptr]
    for n in ("wcns_boundary_all", "wcns_initial_q_dq_c_t", "wcns_timestep", "wcns_r_h_s_vis", "wcns_r_h_s_invis",
              "wcns_communicate_dq", "wcns_l_h_s", "wcns_communicate_pv", "wcns_sync"):
        getattr(L, n).argtypes = [vp]
    L.wcns_residual.argtypes = [vp, dptr, dptr, iptr]
    L.wcns_residual_noalloc.argtypes = [vp, dptr, dptr, iptr]
    L.wcns_step.argtypes = [vp, C.c_int, dptr]
    L.wcns_residual_history.argtypes = [vp, C.c_int, dptr, dptr, iptr, C.POINTER(C.c_longlong)]
    L.wcns_get_counter.argtypes = [vp, C.c_int, C.POINTER(C.c_longlong)]
    L.wcns_region_images.argtypes = [vp, C.c_int, C.c_int, iptr, iptr, iptr, iptr]
    L.wcns_multipoint.argtypes = [vp, iptr, iptr, iptr, C.c_int]
    L.wcns_timing.argtypes = [vp, C.c_int]
    L.wcns_timing_get.argtypes = [vp, C.c_int, dptr, C.POINTER(C.c_longlong)]
    L.wcns_force.argtypes = [vp, dptr]
    L.wcns_force_record.argtypes = [vp]
    L.wcns_force_history.argtypes = [vp, C.c_int, dptr, C.POINTER(C.c_longlong)]
    L.wcns_save_flow.argtypes = [vp, C.c_char_p, C.c_int, C.c_double]
    L.wcns_load_flow.argtypes = [vp, C.c_char_p, C.c_int, iptr, dptr]
    L.wcns_tecplot.argtypes = [vp, C.c_char_p, C.c_int]
    L.wcns_step_graph.argtypes = [vp, C.c_int]
    L.wcns_autotune.argtypes = [vp, C.c_int, iptr, iptr, dptr]
    L.wcns_sweep_layout.argtypes = [vp, iptr, iptr, iptr]
    # adjoint entries take device pointers (vp)
    L.wcns_adj_residual.argtypes = [vp, C.c_int, vp, vp, vp]
    L.wcns_adj_jvp.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp]
    L.wcns_adj_vjp.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
    L.wcns_adj_force.argtypes = [vp, C.c_int, vp, vp, dptr, dptr, vp, vp]
    L.wcns_pool_layout.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long), iptr, iptr]
    L.wcns_adj_step_jvp.argtypes = [vp, vp, vp, vp, vp]
    L.wcns_adj_step_vjp.argtypes = [vp, vp, vp, vp]
    L.wcns_adj_multistep.argtypes = [vp, vp, C.c_int, C.c_int, dptr, C.c_double, dptr, vp, dptr]
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def g3(a):
    """[i+2, j+2, k+2]-indexed array -> flat G3 order (i fastest)."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).transpose(2, 1, 0)).ravel()


class Solver:
    """One GPU context holding a multi-block case (wcns_amd.cases.Case)."""

    def __init__(self, case, device=0, lhs_precision=0):
        self.L = load_library()
        self.case = case
        self.device = device
        prm = Params()
        for k, _ in Params._fields_:
            setattr(prm, k, getattr(case.params, k))
        self.h = C.c_void_p()
        self._check(self.L.wcns_create(C.byref(self.h), device, C.byref(prm)), "wcns_create")
        for blk in case.blocks:
            x, y, z = (np.ascontiguousarray(a.transpose(2, 1, 0)).ravel() for a in (blk.x, blk.y, blk.z))
            nb = C.c_int()
            self._check(self.L.wcns_add_block(self.h, blk.ni, blk.nj, blk.nk, _d(x), _d(y), _d(z), C.byref(nb)))
        for nb, blk in enumerate(case.blocks):
            for reg in blk.regions:
                s = np.array(reg.s_raw, dtype=np.int32)
                t = np.array(reg.t_raw if reg.t_raw is not None else [0] * 6, dtype=np.int32)
                self._check(self.L.wcns_add_region(self.h, nb, _i(s), reg.bctype, _i(t), reg.nbt, reg.ibcwin))
        self._check(self.L.wcns_setup(self.h), "wcns_setup")
        if lhs_precision:
            self._check(self.L.wcns_set_lhs_precision(self.h, lhs_precision))
        for nb in range(len(case.blocks)):
            prim = [g3(a) for a in case.initial_prim(nb)]
            self._check(self.L.wcns_set_prim(self.h, nb, *[_d(a) for a in prim]))

    def _check(self, rc, what=""):
        if rc < 0:
            msg = self.L.wcns_last_error(self.h) if self.h else b""
            raise WcnsError(rc, f"{what} {msg.decode(errors='replace') if msg else ''}")
        return rc

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.wcns_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- the reference's step routines (src/Main.f90:85-109)
    def boundary_all(self):
        self._check(self.L.wcns_boundary_all(self.h), "boundary_all")

    def initial_q_dq_c_t(self):
        self._check(self.L.wcns_initial_q_dq_c_t(self.h), "initial_q_dq_c_t")

    def timestep_tgh(self):
        self._check(self.L.wcns_timestep(self.h), "timestep_tgh")

    def r_h_s_vis(self):
        self._check(self.L.wcns_r_h_s_vis(self.h), "r_h_s_vis")

    def r_h_s_invis(self):
        self._check(self.L.wcns_r_h_s_invis(self.h), "r_h_s_invis")

    def communicate_dq_npp(self):
        self._check(self.L.wcns_communicate_dq(self.h), "communicate_dq_npp")

    def l_h_s_tgh(self):
        self._check(self.L.wcns_l_h_s(self.h), "l_h_s_tgh")

    def communicate_pv_npp(self):
        self._check(self.L.wcns_communicate_pv(self.h), "communicate_pv_npp")

    def residual_curve(self):
        rms, rmax = C.c_double(), C.c_double()
        loc = np.zeros(5, dtype=np.int32)
        self._check(self.L.wcns_residual(self.h, C.byref(rms), C.byref(rmax), _i(loc)), "residual_curve")
        return rms.value, rmax.value, loc

    def step(self, n=1, history=True):
        """n whole pseudo-time steps; returns the RMS history when requested."""
        if history:
            hist = np.zeros(max(n, 1))
            self._check(self.L.wcns_step(self.h, n, _d(hist)), "step")
            return hist[:n]
        self._check(self.L.wcns_step(self.h, n, None), "step")
        return None

    def residual_history(self, n):
        """The monitor of the last n steps run by step / step_graph, oldest first:
        (rms (n,), max (n,), loc (n, 5) = i, j, k, m, nb), and the number of
        steps monitored since set-up (wcns_residual_history)."""
        rms, rmax = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        loc = np.zeros((max(n, 1), 5), dtype=np.int32)
        cnt = C.c_longlong()
        self._check(self.L.wcns_residual_history(self.h, n, _d(rms), _d(rmax), _i(loc), C.byref(cnt)),
                    "residual_history")
        m = min(n, cnt.value, 8192)
        return rms[:m], rmax[:m], loc[:m], cnt.value

    def autotune(self, reps=3):
        """Pick the sweep tile shape by timing (wcns_autotune); returns (TJ, TK, ms per sweep set)."""
        tj, tk, ms = C.c_int(), C.c_int(), C.c_double()
        self._check(self.L.wcns_autotune(self.h, reps, C.byref(tj), C.byref(tk), C.byref(ms)), "autotune")
        return tj.value, tk.value, ms.value

    def sweep_layout(self):
        """(TJ, TK, SGS helpers paired) of the pipelined sweeps in use (wcns_sweep_layout)."""
        tj, tk, pr = C.c_int(), C.c_int(), C.c_int()
        self._check(self.L.wcns_sweep_layout(self.h, C.byref(tj), C.byref(tk), C.byref(pr)), "sweep layout")
        return tj.value, tk.value, bool(pr.value)

    def step_graph(self, n=1):
        """n whole steps replayed from a captured HIP graph (wcns_step_graph)."""
        self._check(self.L.wcns_step_graph(self.h, n), "step_graph")

    def sync(self):
        self._check(self.L.wcns_sync(self.h))

    # ---- inspection
    def field(self, nb, name):
        blk = self.case.blocks[nb]
        n = (blk.ni + 6) * (blk.nj + 6) * (blk.nk + 6)
        ncomp = 5 if name in ("dq", "rhs") else 1
        out = np.zeros(ncomp * n)
        self._check(self.L.wcns_get_field(self.h, nb, FIELDS[name], _d(out)), "get_field")
        shape = (blk.nk + 6, blk.nj + 6, blk.ni + 6)
        if ncomp == 5:
            return out.reshape((5,) + shape).transpose(1, 2, 3, 0)
        return out.reshape(shape)

    def set_field(self, nb, name, arr):
        a = np.asarray(arr, dtype=np.float64)
        if name in ("dq", "rhs"):
            a = np.ascontiguousarray(a.transpose(3, 0, 1, 2))
        a = np.ascontiguousarray(a).ravel()
        self._check(self.L.wcns_set_field(self.h, nb, FIELDS[name], _d(a)), "set_field")

    def timing(self, on=True):
        """Start (clear) or stop recording HIP events around every kernel slot."""
        self._check(self.L.wcns_timing(self.h, 1 if on else 0))

    def timings(self):
        """{slot: (total_ms, launches)} since the last timing(True)."""
        out = {}
        for i, name in enumerate(TIMING_SLOTS):
            ms, n = C.c_double(), C.c_longlong()
            self._check(self.L.wcns_timing_get(self.h, i, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def force(self):
        """cfx, cfy, cfz, cmx, cmy, cmz, cd, cl, xcp over the no-slip walls
        (src/PostProcess.f90:1415-1503 force)."""
        out = np.zeros(9)
        self._check(self.L.wcns_force(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def force_record(self):
        """Append the force coefficients to the on-device history (no host wait)."""
        self._check(self.L.wcns_force_record(self.h), "force_record")

    def force_history(self, n):
        """The last n force records, oldest first: (array (m, 9), records so far)."""
        out = np.zeros((max(n, 1), 9))
        cnt = C.c_longlong()
        self._check(self.L.wcns_force_history(self.h, n, _d(out), C.byref(cnt)), "force_history")
        return out[:min(n, cnt.value, 8192)], cnt.value

    def save_flow(self, path, nout=0, wholetime=0.0):
        """Write the restart file (and path.turb with a turbulence model) in the reference's format."""
        self._check(self.L.wcns_save_flow(self.h, os.fsencode(path), nout, wholetime), "save_flow")

    def tecplot(self, path, kind):
        """Write a Tecplot binary file as the reference's post_processing does:
        kind 0 surfaces (tecout_bin_surface), 1 walls (tecout_bin_wall),
        2 volume (tecout_bin_volume)."""
        self._check(self.L.wcns_tecplot(self.h, os.fsencode(path), kind), "tecplot")

    def load_flow(self, path, t_from_file=False):
        """Read a restart file written by save_flow or by the reference; returns
        (nout, wholetime).  T follows initial_t as in the reference's restart, or
        with t_from_file stays as saved (an exact resume)."""
        no, wt = C.c_int(), C.c_double()
        self._check(self.L.wcns_load_flow(self.h, os.fsencode(path), int(bool(t_from_file)), C.byref(no),
                                          C.byref(wt)), "load_flow")
        return no.value, wt.value

    def counter(self, which):
        v = C.c_longlong()
        self._check(self.L.wcns_get_counter(self.h, which, C.byref(v)))
        return v.value

    def region_images(self, nb, nr):
        blk = self.case.blocks[nb]
        n = (blk.ni + 1) * (blk.nj + 1) * (blk.nk + 1)
        img, jmg, kmg = (np.zeros(n, np.int32) for _ in range(3))
        box = np.zeros(6, np.int32)
        self._check(self.L.wcns_region_images(self.h, nb, nr, _i(img), _i(jmg), _i(kmg), _i(box)))
        cnt = int(np.prod(box[3:] - box[:3] + 1))
        return box, img[:cnt], jmg[:cnt], kmg[:cnt]

    # ---- tangent and adjoint of the residual; arguments are device addresses
    # (int, e.g. torch.Tensor.data_ptr(); 0 for an absent optional array)
    def adj_residual(self, nb, P, X, R):
        self._check(self.L.wcns_adj_residual(self.h, nb, P, X, R), "adj_residual")

    def adj_jvp(self, nb, P, X, dP, dX, R, dR):
        self._check(self.L.wcns_adj_jvp(self.h, nb, P, X, dP or None, dX or None, R or None, dR), "adj_jvp")

    def adj_vjp(self, nb, P, X, W, gP, gX):
        self._check(self.L.wcns_adj_vjp(self.h, nb, P, X, W, gP or None, gX or None), "adj_vjp")

    def adj_force(self, nb, P, X, omega, gP=0, gX=0):
        om = np.ascontiguousarray(omega, dtype=np.float64)
        J = np.zeros(1)
        self._check(self.L.wcns_adj_force(self.h, nb, P, X, _d(om), _d(J), gP or None, gX or None), "adj_force")
        return float(J[0])

    # ---- tangent and adjoint of a whole step over the field pool: arrays of
    # `total` elements, block nb's ghosted nodes from offsets[nb]; nstate of
    # them carry the state, the first nvar are differentiated
    def pool_layout(self):
        total = C.c_long()
        offs = np.zeros(len(self.case.blocks), dtype=np.int64)
        ns, nv = C.c_int(), C.c_int()
        self._check(self.L.wcns_pool_layout(self.h, C.byref(total), offs.ctypes.data_as(C.POINTER(C.c_long)),
                                            C.byref(ns), C.byref(nv)), "pool_layout")
        return total.value, offs

    def state_layout(self):
        """(nstate, nvar) of the step's state (wcns_pool_layout)."""
        total = C.c_long()
        ns, nv = C.c_int(), C.c_int()
        self._check(self.L.wcns_pool_layout(self.h, C.byref(total), None, C.byref(ns), C.byref(nv)), "pool_layout")
        return ns.value, nv.value

    def adj_step_jvp(self, P, dP, Pn=0, dPn=0):
        self._check(self.L.wcns_adj_step_jvp(self.h, P or None, dP, Pn or None, dPn or None), "adj_step_jvp")

    def adj_step_vjp(self, P, Wn, gP):
        self._check(self.L.wcns_adj_step_vjp(self.h, P or None, Wn, gP), "adj_step_vjp")

    def adj_multistep(self, P0, nsteps, every, omega, budget_gb, gP0):
        om = np.ascontiguousarray(omega, dtype=np.float64)
        J, used = np.zeros(1), np.zeros(1)
        self._check(self.L.wcns_adj_multistep(self.h, P0, nsteps, every, _d(om), budget_gb, _d(J), gP0, _d(used)),
                    "adj_multistep")
        return float(J[0]), float(used[0])

    def multipoint(self):
        ng = C.c_int()
        npts = np.zeros(1 << 16, np.int32)
        cap = 4 << 18
        pts = np.zeros(cap, np.int32)
        tot = self._check(self.L.wcns_multipoint(self.h, C.byref(ng), _i(npts), _i(pts), cap))
        return npts[:ng.value].copy(), pts[:4 * tot].reshape(-1, 4).copy()
