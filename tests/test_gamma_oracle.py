"""Pins of the oracle's Gamma-point pieces (oracle/gamma.c) and the
product's host getw0 against it.  No GPU.

* sylm/sylmnc: the reference's own S/sylm.f90 built into oracle/_ref;
* Gaunt table: the reference's scg (S/scg.f90 in oracle/_ref) read through
  rotcg's table layout; sympy's real_gaunt; closed forms;
* getw0: the isotropic closed form wklm(1) (4 pi)^1.5/wbz (1/eps - 1)
  (the reference's own check, S/m_w0w0i.f90:278-279); the Galerkin equations
  of 1 = llw(q) K(q) integrated on a dense sphere rule with llw(q) evaluated
  from the tensors themselves; llmat reproducing llw at the q0i directions
  (the check of S/m_w0w0i.f90:195-207);
* head / wing: diagonal chi (no local fields: llw = eps(1,1)); the Schur
  complement eps11 - eps1r eps_rr^-1 eps_r1 and wing = -eps1r eps_rr^-1;
* q = 0 body: numpy inverse of the mu, nu >= 2 block; modifyWV0 by numpy.
"""
import ctypes as C
import math

import numpy as np
import pytest

from ecalj_amd import synth


def crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


# ------------------------------------------------------------------ harmonics and Gaunt
def test_sylm_matches_reference_build(orc):
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(1)
    lmx = 6
    n = (lmx + 1) ** 2
    cy = np.zeros(n)
    ref.sylmnc_(cy.ctypes.data_as(C.POINTER(C.c_double)), C.byref(C.c_int(lmx)))
    for t in range(40):
        r = rng.standard_normal(3)
        if t == 0:
            r = np.array([0.0, 0.0, 1.3])   # the z-axis branch
        if t == 1:
            r = np.array([0.7, -0.4, 0.0])  # the equatorial branch
        yl = np.zeros(n)
        r2 = C.c_double(0.0)
        ref.sylm_(r.ctypes.data_as(C.POINTER(C.c_double)), yl.ctypes.data_as(C.POINTER(C.c_double)),
                  C.byref(C.c_int(lmx)), C.byref(r2))
        want = yl * cy / np.linalg.norm(r) ** np.array([int(math.isqrt(i)) for i in range(n)])
        np.testing.assert_allclose(orc.sylm(r, lmx), want, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(synth.real_ylm(lmx, r), want, rtol=1e-12, atol=1e-12)


def _ref_cg_table(ref, lmax):
    """cgn(ilm, ilma, ilmb) of rotcg (S/rotcg.f90:37-53) from the reference's scg."""
    lnj, lnx = 62200, 7400
    cg = np.zeros(lnj)
    jcg = np.zeros(lnj, np.int32)
    indxcg = np.zeros(lnx, np.int32)
    ref._QMm_scgPscg(C.byref(C.c_int(lmax)), cg.ctypes.data_as(C.POINTER(C.c_double)),
                     indxcg.ctypes.data_as(C.POINTER(C.c_int)), jcg.ctypes.data_as(C.POINTER(C.c_int)))
    n1, n2 = (lmax + 1) ** 2, (2 * lmax + 1) ** 2
    t = np.zeros((n2, n1, n1))
    for a in range(1, n1 + 1):
        for b in range(1, n1 + 1):
            ii = max(a, b)
            indx = (ii * (ii - 1)) // 2 + min(a, b)
            for icg in range(indxcg[indx - 1], indxcg[indx]):
                t[jcg[icg - 1] - 1, a - 1, b - 1] = cg[icg - 1]
    return t


def test_gaunt_matches_reference_scg(orc):
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    lmax = 4
    t = _ref_cg_table(ref, lmax)
    g = orc.gaunt(lmax, 2 * lmax)       # L1, L2 <= lmax; L3 <= 2 lmax
    n1, n2 = (lmax + 1) ** 2, (2 * lmax + 1) ** 2
    want = np.transpose(t, (1, 2, 0))   # [L1, L2, L3]
    assert np.abs(want).max() > 0.1
    # scg carries sqrt(pi) to 10 digits (1.772453851), hence 1e-9, not 1e-13
    np.testing.assert_allclose(g[:n1, :n1, :n2], want, rtol=0, atol=1e-9)


def test_gaunt_sympy_and_closed_forms(orc):
    sw = pytest.importorskip("sympy.physics.wigner")
    g = orc.gaunt(3, 2)
    assert abs(g[0, 0, 0] - 1 / math.sqrt(4 * math.pi)) < 1e-14
    n = 16
    np.testing.assert_allclose(g[:n, :n, 0], np.eye(n) / math.sqrt(4 * math.pi), atol=1e-14)
    assert abs(g[6, 6, 6] - 2 / 7 * math.sqrt(5 / (4 * math.pi))) < 1e-14
    # magnitudes against sympy (its real harmonics may differ in sign by m)
    for (l1, m1, l2, m2, l3, m3) in [(1, 1, 1, 1, 2, 2), (1, -1, 1, 1, 2, -2), (2, 1, 1, 0, 1, 1), (3, -2, 1, 1, 2, -1),
                                     (2, -1, 2, 1, 2, -2), (3, 3, 3, -3, 2, -2), (3, 0, 1, 0, 2, 0)]:
        want = abs(float(sw.real_gaunt(l1, l2, l3, m1, m2, m3)))
        got = abs(g[l1 * l1 + l1 + m1, l2 * l2 + l2 + m2, l3 * l3 + l3 + m3])
        assert abs(got - want) < 1e-13, (l1, m1, l2, m2, l3, m3, got, want)


# ------------------------------------------------------------------ getw0
def _llw_of_q(gc, a, pts):
    """llw(e) = sum_j a_j e.epinv_j.e (the fitted quadratic form)."""
    f = np.einsum("pa,abj,pb->pj", pts, gc.epinv, pts)
    return f @ a


def test_getw0_isotropic_closed_form(orc):
    gc = synth.gamma_cell(lmxax=6)
    eps = np.array([3.0 + 0.2j, 11.5 - 0.7j, 1.0 + 0.0j])
    llw = np.repeat(eps[:, None], 6, axis=1)
    w0, _ = orc.getw0(1, llw, gc.dmlx, gc.epinvq0i, gc.wklm, gc.wbz, gc.lmxax, gc.epinv)
    want = gc.wklm[0] * (4 * math.pi) ** 1.5 / gc.wbz * (1.0 / eps - 1.0)
    np.testing.assert_allclose(w0, want, rtol=1e-12, atol=1e-12 * abs(want).max())
    assert abs(w0[2]) < 1e-10 * abs(want).max()


@pytest.mark.parametrize("lmxax", [2, 4, 6])
def test_getw0_galerkin_equations(orc, lmxax):
    """K_L = 4 pi K/wbz extracted with unit wklm; int Y_L1 (llw(q) K(q) - 1) = 0
    for every even L1 <= lmxax, on a dense sphere rule."""
    rng = np.random.default_rng(lmxax)
    gc = synth.gamma_cell(lmxax=lmxax)
    nlk = (lmxax + 1) ** 2
    llw = (6.0 + 2.0 * rng.standard_normal((2, 6))) + 0.5j * rng.standard_normal((2, 6))
    even = [lm for lm in range(nlk) if math.isqrt(lm) % 2 == 0]
    K = np.zeros((2, nlk), complex)
    for lm in even:
        wk = np.zeros(nlk)
        wk[lm] = 1.0
        w0, _ = orc.getw0(1, llw, gc.dmlx, gc.epinvq0i, wk, gc.wbz, lmxax, gc.epinv)
        if lm == 0:
            w0 = w0 + (4 * math.pi) ** 1.5 / gc.wbz
        K[:, lm] = w0 * gc.wbz / (4 * math.pi)
    pts, wts = synth.sphere_rule(20)
    Y = synth.real_ylm(lmxax, pts)
    for iw in range(2):
        a = np.linalg.solve(gc.epinvq0i, llw[iw])
        Lq = _llw_of_q(gc, a, pts)
        Kq = Y @ K[iw]
        res = (wts * (Lq * Kq - 1.0)) @ Y[:, even]
        assert np.abs(res).max() < 1e-11, res
        # llw(q) reproduces llw at the six directions
        e = (gc.q0i / np.linalg.norm(gc.q0i, axis=0)).T
        np.testing.assert_allclose(_llw_of_q(gc, a, e), llw[iw], rtol=1e-12)


def test_getw0_llmat(orc):
    gc = synth.gamma_cell()
    rng = np.random.default_rng(3)
    llw = 4.0 + rng.standard_normal((3, 6)) + 0.1j * rng.standard_normal((3, 6))
    _, llmat = orc.getw0(-1, llw, gc.dmlx, gc.epinvq0i, gc.wklm, gc.wbz, gc.lmxax, gc.epinv)
    e = (gc.q0i / np.linalg.norm(gc.q0i, axis=0)).T
    got = np.einsum("ia,ab,ib->i", e, llmat, e)    # e_i.llmat.e_i at iw = 0 (index 1 of -1..1)
    np.testing.assert_allclose(got, llw[1], rtol=1e-12)


def test_product_getw0_matches_oracle(orc):
    import ecalj_amd

    try:
        ecalj_amd.lib()
    except (RuntimeError, OSError):
        pytest.skip("libecalj_gpu.so not built")
    rng = np.random.default_rng(11)
    for lmxax in (0, 2, 4, 8):
        gc = synth.gamma_cell(seed=lmxax + 2, lmxax=lmxax)
        llw = 5.0 + rng.standard_normal((7, 6)) + 0.3j * rng.standard_normal((7, 6))
        w_o, m_o = orc.getw0(-3, llw, gc.dmlx, gc.epinvq0i, gc.wklm, gc.wbz, lmxax, gc.epinv)
        w_g, m_g = ecalj_amd.getw0(-3, llw, gc.dmlx, gc.epinvq0i, gc.wklm, gc.wbz, lmxax, gc.epinv)
        assert relerr(w_g, w_o) < 1e-12, lmxax
        assert relerr(m_g, m_o) < 1e-13


# ------------------------------------------------------------------ head, q = 0 body, modifyWV0
def _coulomb(rng, n):
    vc = rng.uniform(0.5, 40.0, n)
    Q, _ = np.linalg.qr(crand(rng, n, n))
    return vc, np.asfortranarray(Q)


def _eps(chi, vc, Q, vcou1, nspin, xm):
    s = np.sqrt(vc)
    fac = 2.0 if nspin == 1 else 1.0
    x = fac * chi
    if xm:
        x = Q.conj().T @ x @ Q
    e = np.eye(len(vc)) - s[:, None] * x * s[None, :]
    if vcou1 is not None:
        e[0, 0] = 1.0 - vcou1 * x[0, 0]
    return e


def test_eps_head_diagonal_chi(orc):
    n = 12
    rng = np.random.default_rng(4)
    vc = rng.uniform(0.5, 4.0, n)
    chi = np.zeros((n, n, 3), complex, order="F")
    for i in range(3):
        chi[:, :, i] = np.diag(-rng.uniform(0.01, 0.2, n) + 0.01j * rng.standard_normal(n))
    vcou1 = 350.0
    llw, _ = orc.eps_head(chi, np.sqrt(vc), np.eye(n), vcou1, nspin=2, is_x0_m_basis=False)
    np.testing.assert_allclose(llw, 1.0 - vcou1 * chi[0, 0, :], rtol=1e-14)


@pytest.mark.parametrize("nspin,xm", [(1, True), (2, False)])
def test_eps_head_schur_complement(orc, nspin, xm):
    n, nw = 30, 3
    rng = np.random.default_rng(5 + nspin)
    vc, Q = _coulomb(rng, n)
    chi = crand(rng, n, n, nw) * 0.01
    vcou1 = vc[0] * 1.3
    llw, wing = orc.eps_head(chi, np.sqrt(vc), Q, vcou1, nspin=nspin, is_x0_m_basis=xm, wing_iw=1)
    for iw in range(nw):
        e = _eps(chi[:, :, iw], vc, Q, vcou1, nspin, xm)
        s = np.linalg.solve(e[1:, 1:].T, e[0, 1:])          # eps1r eps_rr^-1
        assert abs(llw[iw] - (e[0, 0] - s @ e[1:, 0])) < 1e-12 * abs(llw[iw])
        if iw == 1:
            np.testing.assert_allclose(wing, -s, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("nspin,xm", [(1, True), (2, False)])
def test_wc_build_gamma_block(orc, nspin, xm):
    n, nw = 25, 2
    rng = np.random.default_rng(8)
    vc, Q = _coulomb(rng, n)
    chi = crand(rng, n, n, nw) * 0.01
    w = orc.wc_build_gamma(chi, np.sqrt(vc), Q, nspin=nspin, is_x0_m_basis=xm)
    s = np.sqrt(vc)
    for iw in range(nw):
        e = _eps(chi[:, :, iw], vc, Q, None, nspin, xm)
        inv = np.linalg.inv(e[1:, 1:])
        want = s[1:, None] * inv * s[None, 1:] - np.diag(vc[1:])
        np.testing.assert_allclose(w[1:, 1:, iw], want, rtol=0, atol=1e-11 * np.abs(want).max())
        assert np.all(w[0, :, iw] == 0) and np.all(w[:, 0, iw] == 0)


def test_wv0_patch(orc):
    n, nw = 9, 3
    rng = np.random.default_rng(2)
    _, Q = _coulomb(rng, n)
    w = crand(rng, n, n, nw)
    w0 = crand(rng, nw)
    got = orc.wv0_patch(w, w0, Q, is_wc_m_basis=True)
    for iw in range(nw):
        x = w[:, :, iw].copy()
        x[0, 0] = w0[iw]
        np.testing.assert_allclose(got[:, :, iw], Q @ x @ Q.conj().T, atol=1e-13)
    got2 = orc.wv0_patch(w, w0, Q, is_wc_m_basis=False)
    assert np.all(got2[0, 0, :] == w0) and np.all(got2[1:, :, :] == w[1:, :, :])
