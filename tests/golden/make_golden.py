"""Write golden vectors for the pencil-level routines from the reference's own
Fortran, compiled from source into oracle/_ref/libref.so (see oracle/Makefile).

Routines exercised (reference src/ file:line):
  VALUE_HALF_NODE  WCNSE5.f90:799-852
  FLUX_DXYZ        WCNSE5.f90:209-254
  flux_SW          flux.f90:2-77
  WCNS_E_5         WCNSE5.f90:2-206     (with flux_SW)
  WCNS_E_5_41      WCNSE5.f90:2087-2759 (with flux_SW)
flux_Roe is not called: it needs flux() from Convect.f90, which cannot be
compiled here without an MPI library, so Roe is pinned by closed forms instead.

Usage:  python tests/golden/make_golden.py   (needs oracle/_ref/libref.so)
Output: tests/golden/pencil_golden.npz
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(REPO, "oracle", "_ref", "libref.so")

GAMA = 1.4
EFIX = 0.1


def load():
    # lazy binding: routines of other reference files that are never called stay unresolved
    lib = C.CDLL(LIB, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    # module variables of global_const (Globalvar.f90:211-282); Main.f90:20-21 sets 4
    C.c_int.in_dll(lib, "_QMglobal_constEnijk2nd").value = 4
    C.c_int.in_dll(lib, "_QMglobal_constEnijk2d").value = 4
    C.c_double.in_dll(lib, "_QMglobal_constEgama").value = GAMA
    return lib


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ip(v):
    return C.byref(C.c_int(v))


def dv(v):
    return C.byref(C.c_double(v))


def value_half_node(lib, q):
    n, ni = q.shape[1], q.shape[0]
    qf = np.asfortranarray(q.T.copy())                 # q(n, nmax)
    out = np.zeros((ni + 1, n)).T.copy(order="F")      # q_half(n, 0:nmax)
    lib.value_half_node_(ip(n), ip(ni), ip(ni), dp(qf), dp(out))
    return out.T.copy()


def flux_dxyz(lib, f):
    nl, ni = f.shape[1], f.shape[0] - 1
    ff = np.asfortranarray(f.T.copy())
    df = np.zeros((nl, ni), order="F")
    lib.flux_dxyz_(ip(nl), ip(ni), ip(ni), dp(ff), dp(df))
    return df.T.copy()


def flux_sw(lib, pl, pr, nvec):
    f = np.zeros(5)
    lib.flux_sw_(dp(pl.copy()), dp(pr.copy()), ip(5), dv(nvec[0]), dv(nvec[1]), dv(nvec[2]),
                 dv(nvec[3]), dp(f), dv(EFIX), dv(GAMA), dv(GAMA))
    return f


def wcns_line(lib, fn, trxyz, qline):
    """trxyz (ni,5); qline (ni+6, 5) prim over -2..ni+3; returns df (ni,5)."""
    ni = trxyz.shape[0]
    nmax = ni
    tr = np.asfortranarray(trxyz.T.copy())
    q = np.zeros((6, nmax + 6, 2), order="F")
    q[:5, :, 0] = qline.T
    q[5, :, :] = GAMA
    df = np.zeros((5, nmax), order="F")
    flux_sw_ptr = C.cast(lib.flux_sw_, C.c_void_p)
    fnptr = getattr(lib, fn)
    fnptr.argtypes = None
    fnptr(dv(0.0), dv(0.0), ip(ni), ip(nmax), ip(5), ip(1), ip(2), flux_sw_ptr, C.c_void_p(0),
          dv(EFIX), dp(tr), dp(q), dp(df))
    return df.T.copy()


def smooth_line(rng, n, bc_left, bc_right, amp=0.05, rough=False):
    x = np.arange(-2, n + 4, dtype=float) / n
    rho = 1.0 + amp * np.sin(2 * np.pi * (x * rng.integers(1, 4) + rng.random()))
    p = 1.0 / (GAMA * 0.8 ** 2) * (1.0 + amp * np.cos(2 * np.pi * (x * rng.integers(1, 4) + rng.random())))
    u = 0.8 + 0.1 * np.sin(2 * np.pi * (x + rng.random()))
    v = 0.05 * np.cos(2 * np.pi * (2 * x + rng.random()))
    w = 0.02 * np.sin(2 * np.pi * (3 * x + rng.random()))
    if rough:
        jump = n // 2
        rho[jump + 2:] *= 1.8
        p[jump + 2:] *= 2.5
    q = np.stack([rho, u, v, w, p], axis=1)
    if bc_left:
        q[0, 0] = -abs(q[0, 0])
    if bc_right:
        q[-1, 0] = -abs(q[-1, 0])
    return q


def metrics_line(rng, n):
    t = np.zeros((n, 5))
    t[:, 0] = 1.0 + 0.1 * rng.random(n)
    t[:, 1] = 0.2 * rng.standard_normal(n)
    t[:, 2] = 0.1 * rng.standard_normal(n)
    t[:, 3] = 0.0
    t[:, 4] = 1.0 + 0.05 * rng.random(n)
    return t


def main():
    lib = load()
    rng = np.random.default_rng(20261203)
    out = {}
    # VALUE_HALF_NODE / FLUX_DXYZ on random data, incl. the n<=4 branch
    for n in (3, 4, 5, 6, 9, 17):
        q = rng.standard_normal((n, 5))
        out[f"vhn_in_{n}"] = q
        out[f"vhn_out_{n}"] = value_half_node(lib, q)
        f = rng.standard_normal((n + 1, 5))
        out[f"fdx_in_{n}"] = f
        out[f"fdx_out_{n}"] = flux_dxyz(lib, f)
    # flux_SW on random states and normals
    fl = []
    for _ in range(64):
        pl = np.array([1 + 0.3 * rng.random(), rng.standard_normal(), rng.standard_normal(),
                       rng.standard_normal(), 0.5 + rng.random()])
        pr = np.array([1 + 0.3 * rng.random(), rng.standard_normal(), rng.standard_normal(),
                       rng.standard_normal(), 0.5 + rng.random()])
        nv = np.array([rng.standard_normal(), rng.standard_normal(), rng.standard_normal(), 0.0])
        fl.append(np.concatenate([pl, pr, nv, flux_sw(lib, pl, pr, nv)]))
    out["flux_sw"] = np.array(fl)
    # full WCNS lines (scheme 4 and 41), physical-boundary flags on either side,
    # smooth and with a jump, plus a line that triggers the positivity fallback
    cases = []
    for fn, tag in (("wcns_e_5_", 4), ("wcns_e_5_41_", 41)):
        for n in (6, 9, 16, 33):
            for bl in (False, True):
                for br in (False, True):
                    for rough in (False, True):
                        q = smooth_line(rng, n, bl, br, rough=rough)
                        tr = metrics_line(rng, n)
                        df = wcns_line(lib, fn, tr, q)
                        key = f"wcns{tag}_n{n}_{int(bl)}{int(br)}{int(rough)}"
                        out[key + "_q"] = q
                        out[key + "_tr"] = tr
                        out[key + "_df"] = df
                        cases.append(key)
        n = 12
        q = smooth_line(rng, n, True, True)
        q[4:7, 4] = 1e-9          # pressure near zero -> reconstruction fallback
        q[5, 0] = 2e-9
        tr = metrics_line(rng, n)
        key = f"wcns{tag}_fallback"
        out[key + "_q"] = q
        out[key + "_tr"] = tr
        out[key + "_df"] = wcns_line(lib, fn, tr, q)
        cases.append(key)
    out["wcns_cases"] = np.array(cases)
    path = os.path.join(HERE, "pencil_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(cases)} WCNS lines, {len(out)} arrays")


if __name__ == "__main__":
    sys.exit(main())
