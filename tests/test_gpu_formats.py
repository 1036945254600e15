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


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran caller not built (build() needs flang)")
def test_wv4_records_with_fortran(tmp_path):
    """complex(4) W records (the reference's mixed-precision build): the
    library reads a file that a Fortran direct-access write of complex(4)
    records (recl 8*ld^2) made, widened exactly; it writes records that the
    Fortran read of those records returns as the complex(4) rounding of the
    device's complex128 W; records it does not touch stay."""
    import ecalj_amd

    print("wv4 records ld=11 nrec=4", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(6)
    ld, nrec = 11, 4
    w4 = crand(rng, ld, ld, nrec).astype(np.complex64)
    s = tmp_path / "in.bin"
    with open(s, "wb") as f:
        np.array([ld, nrec], np.int32).tofile(f)
        w4.ravel(order="F").tofile(f)
    fname = tmp_path / "__WVI.3"
    subprocess.run([EXE, "w4_w", str(fname), str(s)], check=True, timeout=60)
    assert os.path.getsize(fname) == 8 * ld * ld * nrec
    h = ecalj_amd.Handle(0, "f64")
    w8 = crand(rng, ld, ld, 2)
    try:
        d = h.array((ld, ld, 3), np.complex128)
        h.wv4_read_dev(fname, 2, d, ld)                      # records 2..4
        np.testing.assert_array_equal(d.download(), w4[:, :, 1:4].astype(np.complex128))
        h.wv4_write_dev(fname, 2, h.to_dev(w8), ld)           # records 2..3 from complex128
    finally:
        h.close()
    out = tmp_path / "out.bin"
    with open(out, "wb") as f:
        np.array([ld, nrec], np.int32).tofile(f)
    subprocess.run([EXE, "w4_r", str(fname), str(out)], check=True, timeout=60)
    back = np.fromfile(out, np.complex64).reshape((ld, ld, nrec), order="F")
    np.testing.assert_array_equal(back[:, :, 0], w4[:, :, 0])
    np.testing.assert_array_equal(back[:, :, 1:3], w8.astype(np.complex64))
    np.testing.assert_array_equal(back[:, :, 3], w4[:, :, 3])
