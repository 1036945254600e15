"""A Fortran caller (tests/fortran/zminv_call.f90, built by build()) drives
the library through the ISO_C_BINDING module on the GPU: zminv on a matrix
that needs row exchanges and the real-axis W build, checked against numpy
and the FP64 oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "zminv_call")


def crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran caller not built (build() needs flang)")
def test_fortran_caller_zminv_and_real_axis_w(orc, tmp_path):
    print("fortran caller: zminv n=9, wc_build_real ngb=24 nw=3", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(77)
    n, ngb, nw = 9, 24, 3
    a = crand(rng, n, n)
    a[0, 0] = 0.0
    chi = crand(rng, ngb, ngb, nw) * 0.01
    vc = rng.uniform(0.5, 20.0, ngb)
    zc, _ = np.linalg.qr(crand(rng, ngb, ngb))
    zc = np.asfortranarray(zc)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([n, ngb, nw], np.int32).tofile(f)
        for x in (a, chi):
            x.ravel(order="F").astype(np.complex128).tofile(f)
        vc.astype(np.float64).tofile(f)
        zc.ravel(order="F").astype(np.complex128).tofile(f)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    out = np.fromfile(fout, np.complex128)
    inv = out[:n * n].reshape((n, n), order="F")
    wc = out[n * n:].reshape((ngb, ngb, nw), order="F")
    np.testing.assert_allclose(inv, np.linalg.inv(a), rtol=0, atol=1e-10 * np.abs(inv).max())
    ref = orc.wc_build(chi, np.sqrt(vc), zc, nspin=1, is_x0_m_basis=True, is_wc_m_basis=True)
    assert np.abs(wc - ref).max() / np.abs(ref).max() < 1e-10
