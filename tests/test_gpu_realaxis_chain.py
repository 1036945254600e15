"""Real-axis chain on the GPU against the FP64 oracle: dpsion_chiq's real
branch (ecalj_dpsion_real), the WVRllwR W build with pivoted inversion
(ecalj_wc_build_real), the zminv seam (ecalj_zminv), and the chain
rcxq -> chi(omega) -> W_c(omega) -> Sigma_c pole term.

Bars: FP64 1e-10 (norm-wise relative); complex64 storage of rcxq 1e-6.
Every case prints its name and shape before touching the GPU."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def _freqr(fr, nw):
    return np.concatenate([[0.0], 0.5 * (fr[:nw] + fr[1:nw + 1])])


@pytest.fixture(scope="module")
def h():
    import ecalj_amd

    hd = ecalj_amd.Handle(0, "f64")
    yield hd
    hd.close()


# ------------------------------------------------------------------ dpsion, real axis
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
@pytest.mark.parametrize("npr,npr_col,nw,ecut", [(37, 29, 45, 1e10), (64, 64, 7, 0.4), (5, 3, 100, 1e10)])
def test_dpsion_real(orc, prec, tol, npr, npr_col, nw, ecut):
    import ecalj_amd

    say("dpsion_real", prec, npr, npr_col, nw, ecut)
    rng = np.random.default_rng(npr + nw)
    fr = orc.frhis(1.2)[1]
    nwhis = len(fr) - 1
    freqr = _freqr(fr, nw)
    r = crand(rng, npr, npr_col, nwhis + 1)
    r[:, :, 0] *= 0.3  # slot 0 is read (the i*rcxq(:,:,0) term) but never scaled
    if prec == "f32":
        r = np.asfortranarray(r.astype(np.complex64).astype(np.complex128))
    o_r = r.copy(order="F")
    zo = orc.dpsion_real(o_r, fr, orc.rmatt(freqr, fr), ecut=ecut)
    hd = ecalj_amd.Handle(0, prec)
    try:
        zg = hd.dpsion_real(r, fr, freqr, ecut=ecut)
        # the already-scaled array with rcxq_scaled: same result
        zs = hd.dpsion_real(o_r, fr, freqr, ecut=ecut, rcxq_scaled=True)
    finally:
        hd.close()
    assert relerr(zg, zo) < tol
    assert relerr(zs, zo) < tol


def test_dpsion_imag_many_frequencies(h, orc):
    """niw > 32 runs in output chunks; the chunk that stores the scaled rcxq runs last."""
    say("dpsion_imag niw=41")
    rng = np.random.default_rng(41)
    fr = orc.frhis(0.9)[1]
    nwhis = len(fr) - 1
    nu = np.linspace(0.0, 3.0, 41)
    r = crand(rng, 19, 23, nwhis + 1)
    g_r = r.copy(order="F")
    zg = h.dpsion_imag(g_r, fr, nu, scale_rcxq=True)
    o_r = r.copy(order="F")
    zo = orc.dpsion_imag(o_r, fr, orc.imatt(nu, fr))
    assert relerr(zg, zo) < 1e-12
    assert relerr(g_r[:, :, 1:], o_r[:, :, 1:]) < 1e-14


def test_dpsion_real_rejects_bad_mesh(h, orc):
    import ecalj_amd

    say("dpsion_real errors")
    fr = orc.frhis(0.3)[1]
    nwhis = len(fr) - 1
    r = np.zeros((2, 2, nwhis + 1), complex, order="F")
    with pytest.raises(ecalj_amd.EcaljError) as e:  # nw >= nwhis (S/dpsion5.f90:203)
        h.dpsion_real(r, fr, _freqr(np.concatenate([fr, fr[-1:] * 1.1]), nwhis))
    assert e.value.status == 1
    bad = _freqr(fr, 5)
    bad[0] = 1e-3
    with pytest.raises(ecalj_amd.EcaljError) as e:  # freq_r(0) /= 0 (S/dpsion5.f90:204)
        h.dpsion_real(r, fr, bad)
    assert e.value.status == 4


# ------------------------------------------------------------------ zminv (pivoted)
def _needs_pivots(rng, n):
    a = crand(rng, n, n)
    a[0, 0] = 0.0  # in-order elimination divides by zero here
    a[2, 2] = 1e-13
    if n > 40:
        a[33, :] *= 1e-9  # small row inside the second pivot block
    return a


@pytest.mark.parametrize("n", [1, 5, 32, 33, 64, 100, 257])
def test_zminv(h, n):
    say("zminv", n)
    rng = np.random.default_rng(n)
    a = _needs_pivots(rng, n) if n > 2 else crand(rng, n, n)
    g = h.zminv(a)
    ref = np.linalg.inv(a)
    assert relerr(g, ref) < 1e-9
    assert np.abs(g @ a - np.eye(n)).max() < 1e-8


def test_zminv_batched_dev(h):
    say("zminv_dev batched n=70 x 5, lda=75")
    rng = np.random.default_rng(70)
    n, lda, nb = 70, 75, 5
    a = np.zeros((lda, n, nb), complex, order="F")
    for b in range(nb):
        a[:n, :, b] = _needs_pivots(rng, n)
        a[n:, :, b] = 7.0 + 7.0j  # padding rows must stay untouched
    d = h.to_dev(a)
    h.zminv_dev(d, n, nbatch=nb, lda=lda)
    g = d.download().reshape((lda, n, nb), order="F")
    for b in range(nb):
        assert relerr(g[:n, :, b], np.linalg.inv(a[:n, :, b])) < 1e-9
    assert np.all(g[n:] == 7.0 + 7.0j)


def test_zminv_singular_reports_error(h):
    import ecalj_amd

    say("zminv singular")
    a = crand(np.random.default_rng(0), 12, 12)
    a[:, 5] = 0.0
    with pytest.raises(ecalj_amd.EcaljError) as e:
        h.zminv(a)
    assert e.value.status == 2
    h.sync()


# ------------------------------------------------------------------ W, real axis
def _chi_real(orc, rng, n, nw):
    """chi(omega) from a positive semidefinite binned rcxq through the oracle's
    real-axis branch: a response with the reference's structure."""
    fr = orc.frhis(0.5)[1]
    nwhis = len(fr) - 1
    r = np.zeros((n, n, nwhis + 1), complex, order="F")
    for b in range(1, nwhis + 1):
        X = crand(rng, n, 3) * 0.02
        r[:, :, b] = X @ X.conj().T
    freqr = _freqr(fr, nw - 1)
    return orc.dpsion_real(r, fr, orc.rmatt(freqr, fr))


def _cond_eps(chi, vc, Q, nspin, xm):
    s = np.sqrt(vc)
    fac = 2.0 if nspin == 1 else 1.0
    k = 1.0
    for i in range(chi.shape[2]):
        x = fac * chi[:, :, i]
        if xm:
            x = Q.conj().T @ x @ Q
        k = max(k, np.linalg.cond(np.eye(len(vc)) - s[:, None] * x * s[None, :]))
    return k


@pytest.mark.parametrize("n,nw,xm,wm,nspin", [(40, 3, True, True, 1), (97, 2, False, False, 2),
                                             (130, 2, True, False, 1), (300, 2, False, True, 1)])
def test_wc_build_real(h, orc, n, nw, xm, wm, nspin):
    say("wc_build_real", n, nw, xm, wm, nspin)
    rng = np.random.default_rng(n + 1)
    vc = rng.uniform(0.5, 40.0, n)
    Q, _ = np.linalg.qr(crand(rng, n, n))
    Q = np.asfortranarray(Q)
    chi = _chi_real(orc, rng, n, nw)
    assert np.abs(chi[:, :, -1] - chi[:, :, -1].conj().T).max() > 1e-3  # not Hermitian
    h.set_coulomb(vc, Q)
    g = h.wc_build_real(chi, nspin=nspin, is_x0_m_basis=xm, is_wc_m_basis=wm)
    o = orc.wc_build(chi, np.sqrt(vc), Q, nspin=nspin, is_x0_m_basis=xm, is_wc_m_basis=wm)
    # two backward-stable inversions agree to about cond(eps) * eps_machine
    assert relerr(g, o) < max(1e-10, 1e-15 * _cond_eps(chi, vc, Q, nspin, xm))


def test_wc_build_real_zero_leading_pivot(h, orc):
    """eps(1,1) = 0: in-order elimination fails, the real-axis build pivots."""
    import ecalj_amd

    say("wc_build_real zero pivot")
    n = 70
    rng = np.random.default_rng(9)
    vc = rng.uniform(0.5, 4.0, n)
    chi = crand(rng, n, n)[:, :, None] * 0.01
    chi = np.asfortranarray(chi)
    chi[0, 0, 0] = 1.0 / vc[0]  # nspin=2: eps(1,1) = 1 - vc(1) chi(1,1) = 0
    h.set_coulomb(vc, np.eye(n, dtype=complex))
    g = h.wc_build_real(chi, nspin=2, is_x0_m_basis=False, is_wc_m_basis=False)
    o = orc.wc_build(chi, np.sqrt(vc), np.eye(n, dtype=complex), nspin=2, is_x0_m_basis=False,
                     is_wc_m_basis=False)
    assert relerr(g, o) < 1e-9
    # the imaginary-axis build eliminates in order: eps(1,1) is zero up to
    # rounding, so it either reports a zero pivot or loses the result
    try:
        g2 = h.wc_build(chi, nspin=2, is_x0_m_basis=False, is_wc_m_basis=False)
        assert relerr(g2, o) > 1e-6
    except ecalj_amd.EcaljError as e:
        assert e.status == 2
    h.sync()


# ------------------------------------------------------------------ chain
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_chain_rcxq_to_sigma_real(orc, prec):
    """rcxq -> chi(omega) -> W_c(omega) -> real-axis Sigma_c pole term, each
    step on the device, against the same chain in the oracle.  For complex64
    storage the oracle gets the same rounded rcxq; the device then keeps chi_E
    and W_E in complex64 around the FP64 inverse.  Near a zero of eps(omega)
    (the plasmon) cond(eps) is large; the bars scale with it."""
    import ecalj_amd

    n, nm, ntq = 48, 13, 4
    say("chain real axis", prec, n, nm, ntq)
    rng = np.random.default_rng(123)
    fr = orc.frhis(1.5)[1]
    nwhis = len(fr) - 1
    nw = int(np.searchsorted(fr, 1.0))
    freqr = _freqr(fr, nw)
    r = np.zeros((n, n, nwhis + 1), complex, order="F")
    for b in range(1, nwhis + 1, 5):
        X = crand(rng, n, 2) * 0.02
        r[:, :, b] = X @ X.conj().T
    if prec == "f32":
        r = np.asfortranarray(r.astype(np.complex64).astype(np.complex128))
    vc = rng.uniform(0.5, 30.0, n)
    Q, _ = np.linalg.qr(crand(rng, n, n))
    Q = np.asfortranarray(Q)
    z = crand(rng, n, nm, ntq)
    ekc = np.sort(rng.uniform(-0.8, 0.6, nm))
    omega = rng.uniform(-0.5, 0.4, ntq)
    ef, esmr, wk = 0.01, 0.01, 0.125
    # oracle chain
    chi_o = orc.dpsion_real(r.copy(order="F"), fr, orc.rmatt(freqr, fr))
    w_o = orc.wc_build(chi_o, np.sqrt(vc), Q, nspin=1, is_x0_m_basis=True, is_wc_m_basis=True)
    s_o, ne = orc.sigc_real(z, w_o, freqr, omega, ekc, ef, esmr, wk)
    assert ne > 0
    tol = max(1e-10, 1e-15 * _cond_eps(chi_o, vc, Q, 1, True))
    hd = ecalj_amd.Handle(0, prec)
    try:
        hd.set_coulomb(vc, Q)
        d_r = hd.to_dev(r, np.complex128 if prec == "f64" else np.complex64)
        d_chi = hd.array((n, n, nw + 1), np.complex128)
        d_w = hd.array((n, n, nw + 1), np.complex128)
        hd.dpsion_real_dev(d_r, n, n, nwhis, fr, freqr, d_chi)
        hd.wc_build_real_dev(d_chi, d_w, n, nw + 1, nspin=1)
        chi_g = d_chi.download()
        w_g = d_w.download()
    finally:
        hd.close()
    h64 = ecalj_amd.Handle(0, "f64")  # the pole term itself on the FP64 path
    try:
        s_g = h64.sigc_real(z, w_g, freqr, omega, ekc, ef, esmr, wk)
    finally:
        h64.close()
    # step by step: each device step against the oracle step on the device's input
    e_chi = relerr(chi_g, chi_o)
    assert e_chi < 1e-10
    w_o2 = orc.wc_build(chi_g, np.sqrt(vc), Q, nspin=1, is_x0_m_basis=True, is_wc_m_basis=True)
    kappa = tol / 1e-15
    if prec == "f64":
        assert relerr(w_g, w_o2) < tol
    else:
        # the complex64 form stores chi_E and W_E in complex64 around the FP64
        # inverse (the reference's __MP split): per frequency the error is the
        # FP32 rounding (2^-24, with a 16x margin for the K-long FP32 sums)
        # amplified by that frequency's cond(eps)
        s = np.sqrt(vc)
        for i in range(nw + 1):
            x = Q.conj().T @ (2 * chi_g[:, :, i]) @ Q
            k_i = np.linalg.cond(np.eye(n) - s[:, None] * x * s[None, :])
            assert relerr(w_g[:, :, i], w_o2[:, :, i]) < max(1e-6, 1e-6 * k_i), i
    s_o2, _ = orc.sigc_real(z, w_g, freqr, omega, ekc, ef, esmr, wk)
    assert relerr(s_g, s_o2) < 1e-10
    # end to end: the chi difference (summation order over ~300 bins) is
    # amplified by cond(eps) near its zero
    e_w = 10 * kappa * e_chi if prec == "f64" else 1e-6 * kappa
    assert relerr(w_g, w_o) < max(1e-10, e_w)
    assert relerr(s_g, s_o) < max(1e-10, e_w)
