"""Interstitial plane-wave rows of zmel on the GPU (ecalj_zmel_ipw / _dev)
against the oracle (orc_zmel_ipw, pinned by a numpy triple sum), both
contraction orders, every precision; and the whole zmel of a real material
layout: MT rows + IPW rows, then the M->E transform and conjugation
(ecalj_zmel_assemble_dev) against the oracle composition.
Bars: FP64 1e-12 of the largest element, complex64 storage 2e-5."""
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-12, "f32": 2e-5, "bf16x3": 2e-5}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
@pytest.mark.parametrize("seed,ngc,ngp1,ngp2,nm,ntp0", [(1, 20, 15, 18, 7, 4), (2, 27, 19, 14, 3, 6),
                                                       (3, 64, 41, 45, 33, 9)])
def test_zmel_ipw_vs_oracle(orc, prec, seed, ngc, ngp1, ngp2, nm, ntp0):
    import ecalj_amd

    say("zmel_ipw", prec, seed, ngc, ngp1, ngp2, nm, ntp0)
    c = synth.ipw_case(seed, ngc, ngp1, ngp2, nm, ntp0)
    ref = orc.zmel_ipw(c)
    h = ecalj_amd.Handle(0, prec)
    try:
        h.set_ppovl(c)
        got = h.zmel_ipw(c)
        # into rows 5.. of a larger array: the other rows stay as they were
        big = np.full((ngc + 9, nm, ntp0), 3.0 - 2.0j, np.complex128, order="F")
        h.zmel_ipw(c, zmel=big, row0=5)
    finally:
        h.close()
    assert relerr(got, ref) < TOL[prec]
    assert relerr(big[5:5 + ngc], ref) < TOL[prec]
    assert np.all(big[:5] == 3.0 - 2.0j) and np.all(big[5 + ngc:] == 3.0 - 2.0j)


def test_zmel_ipw_rejects_missing_combined_vector(orc):
    import ecalj_amd

    say("zmel_ipw missing combined vector")
    c = synth.ipw_case(5, 12, 10, 9, 6, 3)
    bad = synth.IPWCase(**{**c.__dict__, "igcgp2i": np.zeros_like(c.igcgp2i)})
    h = ecalj_amd.Handle(0, "f64")
    try:
        h.set_ppovl(bad)
        with pytest.raises(ecalj_amd.EcaljError) as e:
            h.zmel_ipw(bad)
        assert e.value.status == 4
    finally:
        h.close()


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_zmel_mt_plus_ipw_m2e(orc, prec):
    """zmelt = [MT rows; IPW rows] (S/m_zmel.f90:323-538), zmel = zcousq^H zmelt
    (:559-560), conj (:590-594) -- the layout of a real material's zmel."""
    import ecalj_amd

    say("zmel MT + IPW + M->E", prec)
    rng = np.random.default_rng(11)
    nv, md = [4, 6], [5, 7]
    nbloch = sum(md)
    ndima = sum(nv)
    c = synth.ipw_case(6, 21, 13, 16, 9, 4)
    ngc, nm, ntp0 = c.ngveccR.shape[1], c.dgeigqk.shape[1], c.geigq.shape[1]
    nbb = nbloch + ngc
    ppbs = [rng.standard_normal((v, v, m)) for v, m in zip(nv, md)]
    ppb = np.concatenate([p.ravel(order="F") for p in ppbs])
    phase = np.exp(1j * rng.uniform(0, 6, 2))
    cm = np.asfortranarray(rng.standard_normal((ndima, nm)) + 1j * rng.standard_normal((ndima, nm)))
    cq = np.asfortranarray(rng.standard_normal((ndima, ntp0)) + 1j * rng.standard_normal((ndima, ntp0)))
    m2e, _ = np.linalg.qr(rng.standard_normal((nbb, nbb)) + 1j * rng.standard_normal((nbb, nbb)))
    m2e = np.asfortranarray(m2e)
    mt = orc.zmel(nv, md, phase, ppb, cm, cq)
    ref = (m2e.conj().T @ np.concatenate([mt, orc.zmel_ipw(c)], axis=0).reshape(nbb, -1, order="F")).conj()
    h = ecalj_amd.Handle(0, prec)
    try:
        h.set_ppb(nv, md, ppb)
        h.set_coulomb(np.ones(nbb), m2e)
        h.set_ppovl(c)
        st = np.complex128 if prec == "f64" else np.complex64
        d_cm, d_cq = h.to_dev(cm, np.complex128), h.to_dev(cq, np.complex128)
        d_mt = h.array((nbloch, nm, ntp0), st)
        d_ipw = h.array((ngc, nm, ntp0), st)
        d_out = h.array((nbb, nm, ntp0), st)
        h.zmel_dev(d_cm, d_cq, d_mt, nm, ntp0, ndima, phase=phase, is_m_basis=True)
        d_geq, d_dgk = h.to_dev(c.geigq, np.complex128), h.to_dev(c.dgeigqk, np.complex128)
        h.zmel_ipw_dev(c, d_geq, d_dgk, d_ipw, ngc)
        h.zmel_assemble_dev(d_mt, nbloch, d_ipw, ngc, nm * ntp0, d_out, is_m_basis=False, zmelconjg=True)
        got = d_out.download().astype(np.complex128).reshape(nbb, -1, order="F")
    finally:
        h.close()
    assert relerr(got, ref) < TOL[prec]
