"""The ISO_C_BINDING module compiles with flang and a Fortran caller links
against libecalj_gpu.so; the host-only QP update is called from Fortran and
checked against the oracle (no GPU needed)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import ecalj_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_fortran_module_binds_and_calls(tmp_path, orc):
    mod = os.path.join(ROOT, "ecalj_amd", "fortran", "ecalj_gpu_mod.f90")
    prog = os.path.join(ROOT, "tests", "fortran", "qp_call.f90")
    libdir = os.path.dirname(ecalj_amd.LIB_PATH)
    exe = str(tmp_path / "qp_call")
    subprocess.run([FLANG, "-c", mod, "-module-dir", str(tmp_path), "-o", str(tmp_path / "m.o")], check=True)
    subprocess.run([FLANG, prog, str(tmp_path / "m.o"), "-I", str(tmp_path), "-module-dir", str(tmp_path),
                    f"-L{libdir}", "-lecalj_gpu", f"-Wl,-rpath,{libdir}", "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == 0
    eqp = np.array([float(v) for v in out[1:]])
    n = 3
    sec = np.diag([complex(-0.1 * i, -0.01 * i) for i in range(1, n + 1)])
    ref = orc.qp_energy([-1.5, 0.2, 3.1], [-0.5, -0.3, -0.1], [-0.01, -0.02, 0.0], np.asfortranarray(sec),
                        [-12.0, -10.0, -9.0])
    np.testing.assert_allclose(eqp, ref, rtol=1e-15)
