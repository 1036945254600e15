"""Eigenfunctions on the full mesh on the GPU (ecalj_readcphi_dev,
ecalj_readgeig_dev) against the oracle's rotmto / rotipw, which are pinned by
the real-space rotation of a Bloch function (tests/test_rotwave_oracle.py);
and the file path: @MNLA_CPHI, __QGpsi and fixed-length __CPHI / __GEIG
records written by the flang-built writer, read by the library, rotated on
the device.  The data movement is exact up to one complex multiply per
element, so the bar is 1e-14 relative."""
import os
import subprocess
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "formats_call")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


@pytest.fixture(scope="module")
def case():
    return synth.rotation_case()


@pytest.fixture(scope="module")
def h(case):
    import ecalj_amd

    hd = ecalj_amd.Handle(0, "f64")
    hd.set_hamindex(case)
    hd.set_mnla(case.ibas_tbl, case.l_tbl, case.k_tbl, case.offset_tbl)
    yield hd
    hd.close()


def _oracle_cphi(orc, case, irr, iq, nspc, nband):
    nd = case.ndima
    out = np.zeros((nd * nspc, nband), complex)
    x = irr[:nd * nspc, :, case.iqimap[iq - 1] - 1]
    for s in range(nspc):
        out[s * nd:(s + 1) * nd] = orc.rotmto(case, iq, x[s * nd:(s + 1) * nd])
    return out


def _oracle_geig(orc, case, irr, iq, nspc, nband):
    n, mx = case.ngp[case.iqmap[iq - 1] - 1], case.ngpmx
    out = np.zeros((mx * nspc, nband), complex)
    x = irr[:mx * nspc, :, case.iqimap[iq - 1] - 1]
    for s in range(nspc):
        out[s * mx:s * mx + n] = orc.rotipw(case, iq, np.asfortranarray(x[s * mx:s * mx + n]))
    return out


@pytest.mark.parametrize("nspc,pad", [(1, 0), (2, 3)])
def test_readcphi_dev(h, orc, case, nspc, pad):
    say("readcphi_dev", nspc, pad)
    rng = np.random.default_rng(nspc)
    nband = 11
    ldc = case.ndima * nspc + pad
    irr = crand(rng, ldc, nband, case.nqi)
    iq = np.arange(1, case.qtt.shape[1] + 1)
    din = h.to_dev(irr)
    dout = h.to_dev(np.full((ldc, nband, len(iq)), 5.0 + 0j, order="F"))
    h.readcphi_dev(iq, nspc, nband, din, ldc, dout)
    h.sync()
    got = dout.download()
    for t, q in enumerate(iq):
        want = _oracle_cphi(orc, case, irr, q, nspc, nband)
        assert relerr(got[:case.ndima * nspc, :, t], want) < 1e-14, q
        assert np.all(got[case.ndima * nspc:, :, t] == 5.0)   # padding rows untouched


@pytest.mark.parametrize("nspc,pad", [(1, 0), (2, 2)])
def test_readgeig_dev(h, orc, case, nspc, pad):
    say("readgeig_dev", nspc, pad)
    rng = np.random.default_rng(10 + nspc)
    nband = 9
    ldg = case.ngpmx * nspc + pad
    irr = crand(rng, ldg, nband, case.nqi)
    iq = np.arange(1, case.qtt.shape[1] + 1)[::-1].copy()
    din = h.to_dev(irr)
    dout = h.array((ldg, nband, len(iq)))
    h.readgeig_dev(iq, nspc, nband, din, ldg, dout)
    h.sync()
    got = dout.download()
    for t, q in enumerate(iq):
        want = _oracle_geig(orc, case, irr, q, nspc, nband)
        assert relerr(got[:case.ngpmx * nspc, :, t], want) < 1e-14, q
        assert np.all(got[case.ngpmx * nspc:, :, t] == 0)


def test_readgeig_rejects_unmapped_g(orc, case):
    import dataclasses

    import ecalj_amd

    say("readgeig unmapped G")
    bad = dataclasses.replace(case, ngvecprev=np.full_like(case.ngvecprev, 9999))
    hd = ecalj_amd.Handle(0, "f64")
    try:
        hd.set_hamindex(bad)
        d = hd.array((case.ngpmx, 2, case.nqi))
        o = hd.array((case.ngpmx, 2, 1))
        with pytest.raises(ecalj_amd.EcaljError):
            hd.readgeig_dev([case.nqi + 1], 1, 2, d, case.ngpmx, o)
        hd.sync()
    finally:
        hd.close()


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran writer not built (build() needs flang)")
def test_files_to_rotated_eigenvectors(orc, case, tmp_path):
    """@MNLA_CPHI + __QGpsi + __CPHI/__GEIG records -> readers -> device rotation."""
    import ecalj_amd
    from test_eigen_formats import write_mnla, write_qgpsi

    say("files -> readcphi/readgeig")
    rng = np.random.default_rng(7)
    nband, nsp, isp = 6, 2, 2
    write_mnla(tmp_path / "@MNLA_CPHI", tmp_path, case)
    write_qgpsi(tmp_path / "__QGpsi", tmp_path, case)
    cphi = crand(rng, case.ndima, nband, case.nqi, nsp)
    geig = crand(rng, case.ngpmx, nband, case.nqi, nsp)
    for name, arr, ld in (("__CPHI", cphi, case.ndima), ("__GEIG", geig, case.ngpmx)):
        recl = 16 * ld * (nband + 2)            # records padded as with nbandmx > nband
        s = tmp_path / (name + ".bin")
        with open(s, "wb") as f:
            np.array([recl, case.nqi * nsp, ld * nband], np.int32).tofile(f)
            for iqi in range(case.nqi):
                for sp in range(nsp):           # record ikpisp = isp + nsp*(iqi-1)
                    arr[:, :, iqi, sp].ravel(order="F").tofile(f)
        subprocess.run([EXE, "cp_w", str(tmp_path / name), str(s)], check=True, timeout=60)
    ib, l, k, off = ecalj_amd.mnla_read(tmp_path / "@MNLA_CPHI")
    qg = ecalj_amd.qgpsi_read(tmp_path / "__QGpsi")
    import dataclasses

    c2 = dataclasses.replace(case, qtt=qg["qtt"], ngp=qg["ngp"], ngvecp=qg["ngvecp"], ngvecprev=qg["ngvecprev"],
                             ngpmx=qg["ngpmx"], imx=qg["imx"])
    hd = ecalj_amd.Handle(0, "f64")
    try:
        hd.set_hamindex(c2)
        hd.set_mnla(ib, l, k, off)
        dc = hd.array((case.ndima, nband, case.nqi))
        dg = hd.array((case.ngpmx, nband, case.nqi))
        hd.rec_read_dev(tmp_path / "__CPHI", 16 * case.ndima * (nband + 2), isp, nsp, case.nqi,
                        16 * case.ndima * nband, dc)
        hd.rec_read_dev(tmp_path / "__GEIG", 16 * case.ngpmx * (nband + 2), isp, nsp, case.nqi,
                        16 * case.ngpmx * nband, dg)
        iq = np.arange(1, case.qtt.shape[1] + 1)
        oc = hd.array((case.ndima, nband, len(iq)))
        og = hd.array((case.ngpmx, nband, len(iq)))
        hd.readcphi_dev(iq, 1, nband, dc, case.ndima, oc)
        hd.readgeig_dev(iq, 1, nband, dg, case.ngpmx, og)
        hd.sync()
        gc, gg = oc.download(), og.download()
    finally:
        hd.close()
    for t, q in enumerate(iq):
        assert relerr(gc[:, :, t], _oracle_cphi(orc, case, cphi[:, :, :, isp - 1], q, 1, nband)) < 1e-14
        assert relerr(gg[:, :, t], _oracle_geig(orc, case, geig[:, :, :, isp - 1], q, 1, nband)) < 1e-14
