"""The front end's files the eigenvector readers start from: @MNLA_CPHI,
__QGpsi and __EValue, written by a flang-built program with the record
statements of sugw and mkqg (tests/fortran/formats_call.f90), read by the
library (ecalj_mnla_read, ecalj_qgpsi_read, ecalj_evalue_read).  No GPU."""
import os
import subprocess

import numpy as np
import pytest

import ecalj_amd
from ecalj_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "formats_call")
needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran caller not built (build() needs flang)")


def _lib_or_skip():
    try:
        ecalj_amd.lib()
    except (RuntimeError, OSError):
        pytest.skip("libecalj_gpu.so not built")


def _run(mode, f, s):
    subprocess.run([EXE, mode, str(f), str(s)], check=True, timeout=60)


def write_mnla(path, tmp, case):
    """@MNLA_CPHI of a RotationCase through the Fortran writer."""
    rows = []
    for io in range(len(case.l_tbl)):
        l = int(case.l_tbl[io])
        for m in range(-l, l + 1):
            rows.append([m, case.k_tbl[io], l, case.ibas_tbl[io], case.offset_tbl[io] + m + l + 1, io + 1])
    s = tmp / "mn.bin"
    with open(s, "wb") as f:
        np.array([len(rows)], np.int32).tofile(f)
        np.array(rows, np.int32).tofile(f)
    _run("mn_w", path, s)


def write_qgpsi(path, tmp, case, cut=2.3):
    s = tmp / "qg.bin"
    nq = case.qtt.shape[1]
    with open(s, "wb") as f:
        np.array([nq, case.ngpmx], np.int32).tofile(f)
        np.array([cut]).tofile(f)
        np.array([64, case.nqi, case.imx, case.nqi], np.int32).tofile(f)
        for iq in range(nq):
            case.qtt[:, iq].astype(np.float64).tofile(f)
            np.array([case.ngp[iq], 1 if iq < case.nqi else 0], np.int32).tofile(f)
            n = max(int(case.ngp[iq]), 1)
            case.ngvecp[:, :n, iq].ravel(order="F").astype(np.int32).tofile(f)
            case.ngvecprev[:, :, :, iq].ravel(order="F").astype(np.int32).tofile(f)
    _run("qg_w", path, s)


@needs_exe
def test_mnla_read(tmp_path):
    _lib_or_skip()
    c = synth.rotation_case()
    write_mnla(tmp_path / "@MNLA_CPHI", tmp_path, c)
    ib, l, k, off = ecalj_amd.mnla_read(tmp_path / "@MNLA_CPHI")
    np.testing.assert_array_equal(ib, c.ibas_tbl)
    np.testing.assert_array_equal(l, c.l_tbl)
    np.testing.assert_array_equal(k, c.k_tbl)
    np.testing.assert_array_equal(off, c.offset_tbl)


@needs_exe
def test_qgpsi_read(tmp_path):
    _lib_or_skip()
    c = synth.rotation_case()
    write_qgpsi(tmp_path / "__QGpsi", tmp_path, c)
    r = ecalj_amd.qgpsi_read(tmp_path / "__QGpsi")
    assert (r["nqnum"], r["ngpmx"], r["nqbz"], r["nqi"], r["imx"], r["nqibz"]) == \
        (c.qtt.shape[1], c.ngpmx, 64, c.nqi, c.imx, c.nqi)
    assert r["qpgcut"] == 2.3
    np.testing.assert_array_equal(r["qtt"], c.qtt)
    np.testing.assert_array_equal(r["ngp"], c.ngp)
    np.testing.assert_array_equal(r["irr"], [1] * c.nqi + [0] * (c.qtt.shape[1] - c.nqi))
    np.testing.assert_array_equal(r["ngvecp"], c.ngvecp)
    np.testing.assert_array_equal(r["ngvecprev"], c.ngvecprev)


@needs_exe
@pytest.mark.parametrize("nspx", [1, 2])
def test_evalue_read(tmp_path, nspx):
    _lib_or_skip()
    rng = np.random.default_rng(nspx)
    nband, nqi = 23, 4
    q = rng.standard_normal((3, nqi))
    ev = rng.standard_normal((nband, nqi, nspx))
    s = tmp_path / "ev.bin"
    with open(s, "wb") as f:
        np.array([nband, nqi, nspx, 1, nspx], np.int32).tofile(f)
        q.ravel(order="F").tofile(f)
        ev.ravel(order="F").tofile(f)
    _run("ev_w", tmp_path / "__EValue", s)
    r = ecalj_amd.evalue_read(tmp_path / "__EValue")
    assert (r["nband"], r["nqi"], r["nsp"], r["nspc"], r["nspx"]) == (nband, nqi, nspx, 1, nspx)
    np.testing.assert_array_equal(r["qtti"], q)
    np.testing.assert_array_equal(r["evud"], ev)


def test_readers_reject_missing_and_short_files(tmp_path):
    _lib_or_skip()
    with pytest.raises(ecalj_amd.EcaljError):
        ecalj_amd.mnla_read(tmp_path / "absent")
    (tmp_path / "short").write_bytes(b"\x20\x00\x00\x00" + b"\x00" * 8)
    with pytest.raises(ecalj_amd.EcaljError):
        ecalj_amd.qgpsi_read(tmp_path / "short")
    with pytest.raises(ecalj_amd.EcaljError):
        ecalj_amd.evalue_read(tmp_path / "short")
