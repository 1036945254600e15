"""The Gamma point of the W build on the GPU against the FP64 oracle:
the q = 0 body (ecalj_wc_build_gamma), the head and wing at the offset
points (ecalj_eps_head), modifyWV0 (ecalj_wv0_patch), and the chain
chi(q0i) -> llw -> getw0 -> W(q=0) records with W(1,1) = w0.

Bars (FP64): norm-wise relative 1e-10 where epsilon is well conditioned,
max(1e-10, 1e-15 cond(eps)) on the real axis (two backward-stable inverses
agree to about cond * machine epsilon).  Every case prints its name and shape
before touching the GPU."""
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


@pytest.fixture(scope="module")
def h():
    import ecalj_amd

    hd = ecalj_amd.Handle(0, "f64")
    yield hd
    hd.close()


def _coulomb(rng, n, head=None):
    vc = rng.uniform(0.5, 40.0, n)
    if head is not None:
        vc[0] = head
    Q, _ = np.linalg.qr(crand(rng, n, n))
    return vc, np.asfortranarray(Q)


def _chi_imag(rng, n, nw, scale=0.01):
    """Hermitian negative semidefinite chi(i nu) (the imaginary-axis form)."""
    c = np.zeros((n, n, nw), complex, order="F")
    for i in range(nw):
        X = crand(rng, n, max(n // 3, 2)) * scale / (1 + i)
        c[:, :, i] = -(X @ X.conj().T)
    return c


def _chi_real(orc, rng, n, nw):
    fr = orc.frhis(0.5)[1]
    nwhis = len(fr) - 1
    r = np.zeros((n, n, nwhis + 1), complex, order="F")
    for b in range(1, nwhis + 1):
        X = crand(rng, n, 3) * 0.02
        r[:, :, b] = X @ X.conj().T
    freqr = np.concatenate([[0.0], 0.5 * (fr[:nw - 1] + fr[1:nw])])
    return orc.dpsion_real(r, fr, orc.rmatt(freqr, fr))


def _cond(chi, vc, Q, nspin, xm, vcou1=None):
    s = np.sqrt(vc)
    fac = 2.0 if nspin == 1 else 1.0
    k = 1.0
    for i in range(chi.shape[2]):
        x = fac * chi[:, :, i]
        if xm:
            x = Q.conj().T @ x @ Q
        e = np.eye(len(vc)) - s[:, None] * x * s[None, :]
        if vcou1 is not None:
            e[0, 0] = 1 - vcou1 * x[0, 0]
        k = max(k, np.linalg.cond(e))
    return k


@pytest.mark.parametrize("n,nw,xm,nspin,real", [(40, 3, True, 1, False), (131, 2, False, 2, False),
                                                (200, 4, True, 1, False), (64, 3, True, 1, True),
                                                (150, 2, False, 2, True)])
def test_wc_build_gamma(h, orc, n, nw, xm, nspin, real):
    say("wc_build_gamma", n, nw, xm, nspin, real)
    rng = np.random.default_rng(n + nw)
    vc, Q = _coulomb(rng, n, head=1e6)
    chi = _chi_real(orc, rng, n, nw) if real else _chi_imag(rng, n, nw)
    h.set_coulomb(vc, Q)
    g = h.wc_build_gamma(chi, nspin=nspin, is_x0_m_basis=xm, real_axis=real)
    o = orc.wc_build_gamma(chi, np.sqrt(vc), Q, nspin=nspin, is_x0_m_basis=xm)
    assert np.all(g[0, :, :] == 0) and np.all(g[:, 0, :] == 0)
    bar = max(1e-10, 1e-15 * _cond(chi, vc, Q, nspin, xm)) if real else 1e-10
    assert relerr(g, o) < bar


def test_wc_build_gamma_dev_ldw(h, orc):
    say("wc_build_gamma_dev ldw", 77, 3)
    n, nw, ldw = 77, 3, 90
    rng = np.random.default_rng(3)
    vc, Q = _coulomb(rng, n)
    chi = _chi_imag(rng, n, nw)
    h.set_coulomb(vc, Q)
    dc = h.to_dev(chi)
    dw = h.to_dev(np.full((ldw, n, nw), 7.0 + 0j, order="F"))
    h.wc_build_gamma_dev(dc, n, nw, dw, ldw=ldw, nspin=1, is_x0_m_basis=True)
    h.sync()
    w = dw.download()
    assert np.all(w[n:, :, :] == 7.0)              # padding rows untouched
    o = orc.wc_build_gamma(chi, np.sqrt(vc), Q, nspin=1, is_x0_m_basis=True)
    assert relerr(w[:n], o) < 1e-10


@pytest.mark.parametrize("n,nw,xm,nspin,real,wing_iw", [(40, 3, True, 1, False, -1), (129, 2, False, 2, False, 1),
                                                        (64, 3, True, 1, True, 0), (201, 2, True, 1, True, 0)])
def test_eps_head(h, orc, n, nw, xm, nspin, real, wing_iw):
    say("eps_head", n, nw, xm, nspin, real, wing_iw)
    rng = np.random.default_rng(n * 3 + nw)
    vc, Q = _coulomb(rng, n)
    vcou1 = vc[0] * 1.02
    chi = _chi_real(orc, rng, n, nw) if real else _chi_imag(rng, n, nw)
    h.set_coulomb(vc, Q)
    llw, wing = h.eps_head(chi, vcou1, nspin=nspin, is_x0_m_basis=xm, real_axis=real, wing_iw=wing_iw)
    lo, wo = orc.eps_head(chi, np.sqrt(vc), Q, vcou1, nspin=nspin, is_x0_m_basis=xm, wing_iw=wing_iw)
    bar = max(1e-10, 1e-15 * _cond(chi, vc, Q, nspin, xm, vcou1))
    assert relerr(llw, lo) < bar
    if wing_iw >= 0:
        assert relerr(wing, wo) < bar
    else:
        assert np.all(wing == 0)


def test_eps_head_rejects_bad_vcou1(h, orc):
    import ecalj_amd

    say("eps_head bad vcou1")
    n = 8
    rng = np.random.default_rng(0)
    vc, Q = _coulomb(rng, n)
    h.set_coulomb(vc, Q)
    chi = _chi_imag(rng, n, 1)
    for v in (0.0, -1.0, float("inf")):
        with pytest.raises(ecalj_amd.EcaljError):
            h.eps_head(chi, v)


@pytest.mark.parametrize("wm", [True, False])
def test_wv0_patch(h, orc, wm):
    say("wv0_patch", wm)
    n, nw = 96, 5
    rng = np.random.default_rng(12)
    vc, Q = _coulomb(rng, n)
    h.set_coulomb(vc, Q)
    w = crand(rng, n, n, nw)
    w0 = crand(rng, nw)
    g = h.wv0_patch(w, w0, is_wc_m_basis=wm)
    o = orc.wv0_patch(w, w0, Q, is_wc_m_basis=wm)
    assert relerr(g, o) < 1e-13


@pytest.mark.parametrize("real", [False, True])
def test_gamma_chain(h, orc, real):
    """Six offset points -> llw -> getw0 (host) -> W(q=0) -> modifyWV0."""
    import ecalj_amd

    n, nw = 90, 3
    say("gamma chain", n, nw, real)
    rng = np.random.default_rng(44 + real)
    gc = synth.gamma_cell(lmxax=4)
    llw_g = np.zeros((nw, 6), complex)
    llw_o = np.zeros((nw, 6), complex)
    kap = 1.0
    for i in range(6):
        vc, Q = _coulomb(rng, n)
        qn = np.linalg.norm(gc.q0i[:, i])
        vcou1 = 4 * np.pi / (qn * qn)
        vc[0] = vcou1 * 0.999
        chi = _chi_real(orc, rng, n, nw) if real else _chi_imag(rng, n, nw)
        chi[0, :, :] /= np.sqrt(vcou1)
        chi[:, 0, :] /= np.sqrt(vcou1)          # chi(1,1) ~ q^2, the wings ~ q
        h.set_coulomb(vc, Q)
        llw_g[:, i], _ = h.eps_head(chi, vcou1, nspin=1, is_x0_m_basis=True, real_axis=real)
        llw_o[:, i], _ = orc.eps_head(chi, np.sqrt(vc), Q, vcou1, nspin=1, is_x0_m_basis=True)
        kap = max(kap, _cond(chi, vc, Q, 1, True, vcou1))
    assert relerr(llw_g, llw_o) < max(1e-10, 1e-15 * kap)
    ii = 0 if real else 1
    w0_g, lm_g = ecalj_amd.getw0(ii, llw_g, gc.dmlx, gc.epinvq0i, gc.wklm, gc.wbz, gc.lmxax, gc.epinv)
    w0_o, lm_o = orc.getw0(ii, llw_o, gc.dmlx, gc.epinvq0i, gc.wklm, gc.wbz, gc.lmxax, gc.epinv)
    assert relerr(w0_g, w0_o) < max(1e-10, 1e-14 * kap)
    vc, Q = _coulomb(rng, n, head=0.0)
    chi = _chi_real(orc, rng, n, nw) if real else _chi_imag(rng, n, nw)
    h.set_coulomb(vc, Q)
    wg = h.wc_build_gamma(chi, nspin=1, is_x0_m_basis=True, real_axis=real)
    wg = h.wv0_patch(wg, w0_g, is_wc_m_basis=True)
    wo = orc.wv0_patch(orc.wc_build_gamma(chi, np.sqrt(vc), Q, nspin=1, is_x0_m_basis=True), w0_o, Q)
    assert relerr(wg, wo) < max(1e-10, 1e-14 * kap)


@pytest.mark.parametrize("prec,tol", [("f32", 2e-5), ("bf16x3", 1e-4)])
def test_gamma_mixed_forms(orc, prec, tol):
    """The mixed forms: M <-> E transforms on the form's core around the FP64
    inverse (the reference's __MP split); bar = the form's per-result bar
    times cond(eps)."""
    import ecalj_amd

    n, nw = 140, 3
    say("gamma mixed", prec, n, nw)
    rng = np.random.default_rng(77)
    vc, Q = _coulomb(rng, n)
    vcou1 = vc[0] * 1.02
    chi = _chi_imag(rng, n, nw)
    hd = ecalj_amd.Handle(0, prec)
    try:
        hd.set_coulomb(vc, Q)
        llw, _ = hd.eps_head(chi, vcou1, nspin=1, is_x0_m_basis=True)
        w = hd.wc_build_gamma(chi, nspin=1, is_x0_m_basis=True)
        w0 = crand(rng, nw)
        wp = hd.wv0_patch(w, w0, is_wc_m_basis=True)
    finally:
        hd.close()
    lo, _ = orc.eps_head(chi, np.sqrt(vc), Q, vcou1, nspin=1, is_x0_m_basis=True)
    wo = orc.wc_build_gamma(chi, np.sqrt(vc), Q, nspin=1, is_x0_m_basis=True)
    kap = _cond(chi, vc, Q, 1, True, vcou1)
    assert relerr(llw, lo) < tol * kap
    assert relerr(w, wo) < tol * kap
    assert relerr(wp, orc.wv0_patch(w, w0, Q)) < tol
