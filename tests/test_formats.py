"""The files shared with the unmodified Fortran steps: the library's readers
and writers (ecalj_vcoud_*, ecalj_sec2_*, ecalj_wv_d_*) against a program
built by flang with the record statements of hvccfp0, hsfp0 and WVIllwI
(tests/fortran/formats_call.f90), both directions.  No GPU."""
import os
import subprocess

import numpy as np
import pytest

import ecalj_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "formats_call")
needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran caller not built (build() needs flang)")


def crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def _run(mode, f, s):
    subprocess.run([EXE, mode, str(f), str(s)], check=True, timeout=60)


@needs_exe
def test_vcoud_both_directions(tmp_path):
    rng = np.random.default_rng(1)
    ngb = 37
    q = np.array([0.1, -0.2, 0.35])
    v = rng.uniform(0.1, 50.0, ngb)
    z = crand(rng, ngb, ngb)
    # Fortran writes, the library reads
    s = tmp_path / "in.bin"
    with open(s, "wb") as f:
        np.array([ngb], np.int32).tofile(f)
        q.tofile(f)
        v.tofile(f)
        z.ravel(order="F").tofile(f)
    _run("vc_w", tmp_path / "__Vcoud.1", s)
    q2, v2, z2 = ecalj_amd.vcoud_read(tmp_path / "__Vcoud.1")
    np.testing.assert_array_equal(q2, q)
    np.testing.assert_array_equal(v2, v)
    np.testing.assert_array_equal(z2, z)
    # the library writes, Fortran reads
    ecalj_amd.vcoud_write(tmp_path / "__Vcoud.2", q, v, z)
    _run("vc_r", tmp_path / "__Vcoud.2", tmp_path / "out.bin")
    raw = np.fromfile(tmp_path / "out.bin", np.uint8)
    assert int(raw[:4].view(np.int32)[0]) == ngb
    np.testing.assert_array_equal(raw[4:28].view(np.float64), q)
    np.testing.assert_array_equal(raw[28:28 + 8 * ngb].view(np.float64), v)
    np.testing.assert_array_equal(raw[28 + 8 * ngb:].view(np.complex128), z.ravel(order="F"))
    # byte-identical files
    assert (tmp_path / "__Vcoud.1").read_bytes() != b"" and \
        (tmp_path / "__Vcoud.2").read_bytes() == open(tmp_path / "__Vcoud.1", "rb").read()


@needs_exe
def test_sec2_both_directions(tmp_path):
    rng = np.random.default_rng(2)
    nq, ntq = 3, 5
    h8 = np.array([1, nq, ntq, 64, 8, 4, 4, 4], np.int32)
    isp = np.ones(nq, np.int32)
    q = np.asfortranarray(rng.standard_normal((3, nq)))
    z = crand(rng, ntq, ntq, nq)
    s = tmp_path / "in.bin"
    with open(s, "wb") as f:
        h8.tofile(f)
        isp.tofile(f)
        q.ravel(order="F").tofile(f)
        z.ravel(order="F").tofile(f)
    _run("s2_w", tmp_path / "SEC2U_f", s)
    h, i, q2, z2 = ecalj_amd.sec2_read(tmp_path / "SEC2U_f")
    np.testing.assert_array_equal(h, h8)
    np.testing.assert_array_equal(i, isp)
    np.testing.assert_array_equal(q2, q)
    np.testing.assert_array_equal(z2, z)
    ecalj_amd.sec2_write(tmp_path / "SEC2U_l", h8, isp, q, z)
    assert (tmp_path / "SEC2U_l").read_bytes() == (tmp_path / "SEC2U_f").read_bytes()


@needs_exe
def test_wv_record_layout(tmp_path):
    """Fortran direct-access records of 16*ld^2 bytes: record r at byte (r-1)*16*ld^2."""
    rng = np.random.default_rng(3)
    ld, nrec = 11, 4
    w = crand(rng, ld, ld, nrec)
    s = tmp_path / "in.bin"
    with open(s, "wb") as f:
        np.array([ld, nrec], np.int32).tofile(f)
        w.ravel(order="F").tofile(f)
    _run("wv_w", tmp_path / "__WVI.1", s)
    raw = np.fromfile(tmp_path / "__WVI.1", np.complex128)
    assert raw.size == ld * ld * nrec
    np.testing.assert_array_equal(raw, w.ravel(order="F"))


def test_wv_d_roundtrip(tmp_path):
    ecalj_amd.wv_d_write(tmp_path / "__WV.d", 8, 16 * 700 ** 2, 700, 244, 16, 10, 0)
    np.testing.assert_array_equal(ecalj_amd.wv_d_read(tmp_path / "__WV.d"), [8, 16 * 700 ** 2, 700, 244, 16, 10, 0])
    line = (tmp_path / "__WV.d").read_text()
    assert line.startswith(" ") and len(line.rstrip("\n")) == 1 + 7 * 14   # (1x,10i14)


def test_vcoud_read_reports_size_first(tmp_path):
    rng = np.random.default_rng(4)
    ecalj_amd.vcoud_write(tmp_path / "v", [0, 0, 0.5], rng.uniform(1, 2, 6), crand(rng, 6, 6))
    import ctypes as C
    n = C.c_int(0)
    rc = ecalj_amd.lib().ecalj_vcoud_read(str(tmp_path / "v").encode(), C.byref(n), None, None, None, 0)
    assert rc == 1 and n.value == 6
