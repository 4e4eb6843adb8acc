"""The Fortran host side of the drop-in boundary (wcns_amd/fortran): the
ISO_C_BINDING module, the shims with the reference's subroutine names and the
driver program, compiled with flang and linked against libwcns_amd.so by
wcns_amd/build.py.  On the CPU: the driver reads the reference's input files
written by wcns_amd.cases and stops cleanly where there is no GPU; the GPU
runs are in tests/test_gpu_host_io.py."""
import os
import subprocess

import numpy as np
import pytest

from wcns_amd import build, cases

DRIVE = build.FDRIVE


def _write_case(case, d, **kw):
    cases.write_plot3d(case, os.path.join(d, "grid.x"))
    cases.write_inp(case, os.path.join(d, "grid.inp"))
    cases.write_param(case, os.path.join(d, "param.dat"), **kw)
    cases.write_flow(case, os.path.join(d, "flow.dat"))


@pytest.mark.skipif(not os.path.exists(DRIVE), reason="Fortran driver not built")
def test_driver_reads_inputs_and_fails_cleanly_without_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    _write_case(cases.cartesian_periodic_case(ni=24, nj=16, nk=12), str(tmp_path), nomax=3)
    p = subprocess.run([DRIVE, "param.dat"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert p.returncode == 2
    assert "wcns_create failed" in p.stdout


def test_writers_follow_the_reference_layouts(tmp_path):
    """Record structure of the PLOT3D grid and the restart file (4-byte
    markers around each record) and the 1976 boundary file's line layout."""
    case = cases.multiblock_case(nx=(12, 14), ny=(12,), nz=(12,))
    _write_case(case, str(tmp_path))
    raw = open(tmp_path / "grid.x", "rb").read()
    recs, o = [], 0
    while o < len(raw):
        n = int(np.frombuffer(raw[o:o + 4], "<i4")[0])
        assert int(np.frombuffer(raw[o + 4 + n:o + 8 + n], "<i4")[0]) == n
        recs.append(raw[o + 4:o + 4 + n])
        o += 8 + n
    assert np.frombuffer(recs[0], "<i4")[0] == 2
    assert list(np.frombuffer(recs[1], "<i4")) == [12, 12, 12]
    xyz = np.frombuffer(recs[3], "<f8").reshape(3, 12, 12, 12)
    assert np.array_equal(xyz[0], case.blocks[0].x.transpose(2, 1, 0))
    flow = open(tmp_path / "flow.dat", "rb").read()
    n0 = int(np.frombuffer(flow[:4], "<i4")[0])
    assert n0 == 4 + 7 * 8
    body = np.frombuffer(flow[8 + n0 + 4:8 + n0 + 4 + 12 ** 3 * 48], "<f8").reshape(12, 12, 12, 6)
    r0 = case.initial_prim(0)[0][3:-3, 3:-3, 3:-3].transpose(2, 1, 0)
    assert np.array_equal(body[..., 0], r0)
    lines = open(tmp_path / "grid.inp").read().split("\n")
    assert lines[0].split() == ["1976"] and lines[2].split() == ["2"]
    assert lines[4].split() == ["12", "12", "12"]
