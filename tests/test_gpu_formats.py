"""W records between device memory and the fixed-length __WVI/__WVR files
(ecalj_wv_write_dev / ecalj_wv_read_dev), checked against the byte layout a
Fortran direct-access write produces (tests/fortran/formats_call.f90)."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "formats_call")


def crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran caller not built (build() needs flang)")
def test_wv_records_device_roundtrip(tmp_path):
    import ecalj_amd

    print("wv records ld=13 nrec=3 from rec 2", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(5)
    ld = 13
    w = crand(rng, ld, ld, 5)
    # Fortran writes records 1..5
    s = tmp_path / "in.bin"
    with open(s, "wb") as f:
        np.array([ld, 5], np.int32).tofile(f)
        w.ravel(order="F").tofile(f)
    subprocess.run([EXE, "wv_w", str(tmp_path / "__WVR.1"), str(s)], check=True, timeout=60)
    h = ecalj_amd.Handle(0, "f64")
    try:
        d = h.array((ld, ld, 3), np.complex128)
        h.wv_read_dev(tmp_path / "__WVR.1", 2, d, ld)          # records 2..4
        np.testing.assert_array_equal(d.download(), w[:, :, 1:4])
        # overwrite records 3..5 from the device; record 1-2 stay
        d2 = h.to_dev(w[:, :, :3] * 2.0)
        h.wv_write_dev(tmp_path / "__WVR.1", 3, d2, ld)
    finally:
        h.close()
    raw = np.fromfile(tmp_path / "__WVR.1", np.complex128).reshape((ld, ld, 5), order="F")
    np.testing.assert_array_equal(raw[:, :, :2], w[:, :, :2])
    np.testing.assert_array_equal(raw[:, :, 2:], 2.0 * w[:, :, :3])
