"""Build recipe of the gfx950 library wcns_amd/libwcns_amd.so (hipcc, in-tree)."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwcns_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = os.path.join(HERE, "fortran")
FDRIVE = os.path.join(FDIR, "wcns_drive")
# the bind(C) module, the shims of the reference's subroutine names, the driver
FSOURCES = ["wcns_amd_c.f90", "wcns_shims.f90", "wcns_drive.f90"]

# -ffp-contract=off: every a*b+c rounds twice, as in the reference's x86-64 build.
# packed FP32 ops are disabled for all device code.
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-variable"]

SOURCES = ["wcns_api.hip", "wcns_rhs_inv.hip", "wcns_rhs_vis.hip", "wcns_step.hip", "wcns_metrics.hip", "wcns_force.hip",
           "wcns_sweep.hip", "wcns_sa.hip", "wcns_transpose.hip", "wcns_adjoint.hip",
           "wcns_adstep.hip"]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
    return out


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_fortran(verbose)
    return LIB


def build_fortran(verbose=True):
    """The Fortran host side (wcns_amd/fortran): the ISO_C_BINDING module, the
    shims and the driver program wcns_drive, linked against libwcns_amd.so."""
    mod = os.path.join(OBJ, "fortran")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for src in FSOURCES:
        out = os.path.join(mod, os.path.splitext(src)[0] + ".o")
        cmd = [FLANG, "-O2", "-fPIC", "-module-dir", mod, "-c", os.path.join(FDIR, src), "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(f"flang failed for {src}:\n{r.stderr[-4000:]}")
        objs.append(out)
    tmp = FDRIVE + ".tmp"
    cmd = [FLANG, "-o", tmp] + objs + ["-L" + HERE, "-lwcns_amd", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"Fortran link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, FDRIVE)
    if verbose:
        print(f"built {FDRIVE}")
    return FDRIVE


if __name__ == "__main__":
    build()
    sys.exit(0)
