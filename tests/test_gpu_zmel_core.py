"""zmel with core middle states and the rotated-atom map on the GPU
(ecalj_set_ppb_full + ecalj_zmel_init, the CoreValence rows of
S/m_zmel.f90:374-390 and iatomp of :295-305, 334) against the oracle
(orc_zmel_init); then Sigma_c over a middle-state range that starts in the
core states, with the core branch of the imaginary-axis weights
(S/m_sxcf_gemm.f90:437-443), against the oracle chain.

Bars: FP64 1e-12 relative; f32 2e-6, bf16x3 1e-5 on zmel; Sigma_c 1e-10
(FP64) and 1e-4 (mixed)."""
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu

ZTOL = {"f64": 1e-12, "f32": 2e-6, "bf16x3": 1e-5}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _crand(rng, *s):
    return np.asfortranarray(rng.standard_normal(s) + 1j * rng.standard_normal(s))


def _case(seed, natom=4, nc=(3, 1, 0, 2), nv=(14, 12, 16, 12), md=(30, 26, 34, 26)):
    rng = np.random.default_rng(seed)
    nlnmx = max(c + v for c, v in zip(nc, nv))
    mdimx = max(md)
    ppb = np.zeros((nlnmx, nlnmx, mdimx, natom), order="F")
    for a in range(natom):
        n = nc[a] + nv[a]
        ppb[:n, :n, :md[a], a] = rng.standard_normal((n, n, md[a])) / np.sqrt(sum(md))
    return rng, list(nc), list(nv), list(md), ppb


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
@pytest.mark.parametrize("ns1,ns2,nq", [(1, 40, 9), (2, 5, 30), (7, 60, 5), (1, 4, 12)])
def test_zmel_core_rows(orc, prec, ns1, ns2, nq):
    import ecalj_amd

    rng, nc, nv, md, ppb = _case(ns1 * 100 + ns2)
    nctot = sum(nc)
    ndima = sum(nv)
    nmv = max(0, ns2 - max(ns1, nctot + 1) + 1)
    cm, cq = _crand(rng, ndima, nmv), _crand(rng, ndima, nq)
    ph = np.exp(1j * rng.uniform(0, 6, len(nc)))
    h = ecalj_amd.Handle(0, prec)
    try:
        h.set_ppb_full(nc, nv, md, ppb)
        for conj in (False, True):
            got = h.zmel_init(ns1, ns2, nctot, cm, cq, phase=ph, zmelconjg=conj)
            ref = orc.zmel_init(nc, nv, md, ppb, ns1, ns2, nctot, cm, cq, phase=ph, conjg=conj)
            say(f"zmel core {prec} ns={ns1}..{ns2} nq={nq} conj={conj}: rel {rel(got, ref):.2e}")
            assert rel(got, ref) < ZTOL[prec]
    finally:
        h.close()


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_zmel_iatomp_and_m2e(orc, prec):
    """A permutation of equivalent atoms, and the product-basis-to-Coulomb-eigenbasis transform."""
    import ecalj_amd

    rng, nc, nv, md, ppb = _case(5, natom=3, nc=(2, 2, 2), nv=(10, 10, 10), md=(20, 20, 20))
    nctot, ndima, ngb = 6, 30, 60
    cm, cq = _crand(rng, ndima, 11), _crand(rng, ndima, 7)
    ph = np.exp(1j * rng.uniform(0, 6, 3))
    perm = np.array([3, 1, 2], np.int32)
    q, _ = np.linalg.qr(_crand(rng, ngb, ngb))
    vc = rng.uniform(0.5, 20.0, ngb)
    h = ecalj_amd.Handle(0, prec)
    try:
        h.set_ppb_full(nc, nv, md, ppb)
        h.set_coulomb(vc, q)
        got = h.zmel_init(1, 17, nctot, cm, cq, phase=ph, iatomp=perm)
        ref = orc.zmel_init(nc, nv, md, ppb, 1, 17, nctot, cm, cq, phase=ph, iatomp=perm)
        say(f"zmel iatomp {prec}: rel {rel(got, ref):.2e}")
        assert rel(got, ref) < ZTOL[prec]
        got = h.zmel_init(1, 17, nctot, cm, cq, phase=ph, iatomp=perm, is_m_basis=False, zmelconjg=True)
        ref = orc.zmel_init(nc, nv, md, ppb, 1, 17, nctot, cm, cq, phase=ph, iatomp=perm, m2e=np.asfortranarray(q),
                            conjg=True)
        say(f"zmel iatomp m2e {prec}: rel {rel(got, ref):.2e}")
        assert rel(got, ref) < ZTOL[prec]
        with pytest.raises(ecalj_amd.EcaljError):     # nctot must be ppb_full's core count
            h.zmel_init(1, 17, nctot + 1, cm, cq)
    finally:
        h.close()


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_sigc_with_core_middle_states(orc, prec):
    """Sigma_c over middle states 1..nctot+nbv (core first): zmel with core
    rows, the core branch of the weights (it <= nctot), the contraction."""
    import ecalj_amd

    rng, nc, nv, md, ppb = _case(9)
    nctot, ndima, ngb = sum(nc), sum(nv), sum(md)
    nbv, ntq, niw = 24, 5, 10
    ns1, ns2 = 1, nctot + nbv
    cm, cq = _crand(rng, ndima, nbv), _crand(rng, ndima, ntq)
    cm /= np.linalg.norm(cm, axis=0)
    cq /= np.linalg.norm(cq, axis=0)
    ek = np.concatenate([np.sort(rng.uniform(-40.0, -20.0, nctot)), np.sort(rng.uniform(-1.0, 3.0, nbv))])
    om = np.sort(rng.uniform(-0.8, 0.4, ntq))
    x, wx = np.polynomial.legendre.leggauss(niw)
    x, wx = 0.5 * (x + 1), 0.5 * wx
    expa = np.exp(-((1 - x) / x) ** 2)
    a = _crand(rng, ngb, ngb, niw + 1) * 0.02
    wc = np.asfortranarray(-(np.einsum("ijn,kjn->ikn", a, a.conj())))
    ref_z = orc.zmel_init(nc, nv, md, ppb, ns1, ns2, nctot, cm, cq)
    ref_g = orc.wgtim(om, ek, 0.003, 0.25, x, wx, expa, ns1=ns1, nctot=nctot)
    ref = orc.sigc_imag(ref_z, wc, ref_g)
    h = ecalj_amd.Handle(0, prec)
    try:
        h.set_ppb_full(nc, nv, md, ppb)
        z = h.zmel_init(ns1, ns2, nctot, cm, cq)
        g = h.sigc_weights(om, ek, 0.003, 0.25, x, wx, expa, ns1=ns1, nctot=nctot)
        assert np.abs(g - ref_g).max() < 1e-13 * np.abs(ref_g).max()
        got = h.sigc(z, wc, g)
        say(f"sigc with {nctot} core states {prec}: rel {rel(got, ref):.2e}")
        assert rel(got, ref) < (1e-10 if prec == "f64" else 1e-4)
    finally:
        h.close()
