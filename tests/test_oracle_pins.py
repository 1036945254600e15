"""Pins of the FP64 CPU oracle against mathematics and the reference's own
compiled routines.  Each test is chosen so that a plausible slip in the
restatement (a dropped term, a sign, a transposed operand, a wrong index)
fails it.  No GPU needed."""
import math

import numpy as np
import pytest
from scipy import integrate

from ecalj_amd import synth


def crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


# ---------------------------------------------------------------- gauss/freq
@pytest.mark.parametrize("n", [6, 10, 12, 16, 20, 24, 32, 40, 48])
def test_gauss_matches_legendre(orc, n):
    x, w = orc.gauss(n, 0.0, 1.0)
    t, wt = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(x, 0.5 * (t + 1.0), rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, 0.5 * wt, rtol=0, atol=2e-15)
    # exact for polynomials of degree < 2n on [0,1]
    for j in (0, 1, 2 * n - 1):
        assert abs((w * x ** j).sum() - 1.0 / (j + 1)) < 1e-13


def test_gauss_rejects_unsupported(orc):
    with pytest.raises(ValueError):
        orc.gauss(7)


def test_freq01(orc):
    fx, fw, wx, ex = orc.freq01(12, 1.3)
    np.testing.assert_allclose(fw, (1 - fx) / fx, rtol=1e-15)
    np.testing.assert_allclose(ex, np.exp(-1.3 ** 2 * fw ** 2), rtol=1e-15)


def test_frhis_closed_form(orc):
    n, f = orc.frhis(2.0, 1.03, 1e-5)
    aa, bb = 0.03, 1e-5 / 0.03
    ref = bb * (np.exp(aa * np.arange(n + 1)) - 1.0)
    np.testing.assert_allclose(f, ref, rtol=1e-14, atol=0)
    # nwhis = first i with edge > omg2max, plus 2 margin bins
    first = int(np.argmax(bb * (np.exp(aa * np.arange(1000)) - 1) > 2.0 + 1e-6)) + 1
    assert n == first + 2


# ---------------------------------------------------------------- hilbertmat
def _edges(nwhis, fr):
    L = np.concatenate([-fr[nwhis:0:-1], [0.0], fr[0:nwhis]])
    R = np.concatenate([-fr[nwhis - 1::-1], [0.0], fr[1:nwhis + 1]])
    return L, 0.5 * (L + R), R


def test_hilbertmat_matches_reference_build(orc):
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes as C

    fr = orc.frhis(0.7)[1]
    nwhis = len(fr) - 1
    L, Cc, R = _edges(nwhis, fr)
    for zz in (0.0 + 0.0j, 0.05j, 0.8j, 3.5j, 0.31 + 0.0j, 0.2 + 0.1j):
        ours = orc.hilbertmat(zz, L, Cc, R)
        out = np.zeros(2 * nwhis + 1, np.complex128)
        z = np.array([zz], np.complex128)
        nn = C.c_int(nwhis)
        dp = C.POINTER(C.c_double)
        ref.hilbertmat_(z.ctypes.data_as(dp), C.byref(nn), L.ctypes.data_as(dp), Cc.ctypes.data_as(dp),
                        R.ctypes.data_as(dp), out.ctypes.data_as(dp))
        # The reference formula forms zz*log((zz-b)/(zz-a)) + (b-a) and then
        # divides by the bin width: a cancellation of relative size
        # |zz|/width.  flang's and glibc's clog differ by an ulp, so the two
        # builds agree to ~eps*|zz|/w_min of the scale (2e-8 at |zz| = 3.5 Ha
        # with 1e-5 Ha bins), for the reference as much as for this oracle.
        w_min = fr[1] - fr[0]
        tol = 1e-13 * (1.0 + abs(zz) / w_min)
        np.testing.assert_allclose(ours, out, rtol=0, atol=tol * np.abs(out).max())


def test_hilbertmat_is_hilbert_transform_of_linear_interp(orc):
    """sum_i rmat(i) f(i) = int f(w)/(w-z) dw for f piecewise linear through
    (C_i, f_i), zero at w=0 and at the outer edges."""
    fr = orc.frhis(0.3)[1]
    nwhis = len(fr) - 1
    L, Cc, R = _edges(nwhis, fr)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(2 * nwhis + 1)
    f[nwhis] = 0.0
    xs = np.concatenate([[L[0]], Cc[:nwhis], [0.0], Cc[nwhis + 1:], [R[-1]]])
    ys = np.concatenate([[0.0], f[:nwhis], [0.0], f[nwhis + 1:], [0.0]])
    for z in (0.07j, 0.4j, 1.9j):
        rm = orc.hilbertmat(z, L, Cc, R)
        lhs = (rm * f).sum()
        g = lambda w: np.interp(w, xs, ys)
        re = integrate.quad(lambda w: (g(w) * (w - z.real)) / ((w - z.real) ** 2 + z.imag ** 2), xs[0], xs[-1],
                            points=list(xs), limit=2000, epsabs=1e-12)[0]
        im = integrate.quad(lambda w: (g(w) * z.imag) / ((w - z.real) ** 2 + z.imag ** 2), xs[0], xs[-1],
                            points=list(xs), limit=2000, epsabs=1e-12)[0]
        assert abs(lhs - (re + 1j * im)) < 1e-8 * max(1.0, abs(lhs))


def test_imatt_definition(orc):
    fr = orc.frhis(0.5)[1]
    nwhis = len(fr) - 1
    L, Cc, R = _edges(nwhis, fr)
    nu = np.array([0.0, 0.1, 1.7])
    im = orc.imatt(nu, fr)
    for i, v in enumerate(nu):
        rr = orc.hilbertmat(1j * v, L, Cc, R)
        np.testing.assert_allclose(im[i], (rr[nwhis + 1:] - rr[nwhis - 1::-1]).real / math.pi, rtol=1e-14, atol=1e-15)


# ---------------------------------------------------------------- wfacx
def test_wfac_matches_reference_build(orc):
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes as C

    rng = np.random.default_rng(5)
    for _ in range(200):
        e1, e2, ek = rng.uniform(-1, 1, 3)
        esmr = rng.choice([0.0, 0.003, 0.05])
        args = [C.c_double(v) for v in (e1, e2, ek, esmr)]
        r1 = ref._QMm_wfacPwfacx2(*[C.byref(a) for a in args])
        assert orc.wfacx2(e1, e2, ek, esmr) == pytest.approx(r1, abs=1e-15)
        if esmr > 0 and r1 > 1e-8:
            r2 = ref._QMm_wfacPweavx2(*[C.byref(a) for a in args])
            assert orc.weavx2(e1, e2, ek, esmr) == pytest.approx(r2, rel=1e-12, abs=1e-14)


# ---------------------------------------------------------------- zmel
def test_zmel_einsum_multiatom(orc):
    rng = np.random.default_rng(7)
    nv, md = [3, 5], [4, 6]
    ndima, ngb, nm, nq = 9, 10, 3, 4
    ppbs = [rng.standard_normal((v, v, m)) for v, m in zip(nv, md)]
    ppb = np.concatenate([p.ravel(order="F") for p in ppbs])
    phase = np.exp(1j * rng.uniform(0, 6, 2))
    cm, cq = crand(rng, ndima, nm), crand(rng, ndima, nq)
    ref = np.zeros((ngb, nm, nq), complex)
    ias = ims = 0
    for a in range(2):
        v, m = nv[a], md[a]
        # zmel(I,n,m) = phase * sum_{i1,i2} conj(cphim(i2,n)) ppb(i1,i2,I) cphiq(i1,m)
        ref[ims:ims + m] = phase[a] * np.einsum("bn,cbI,cm->Inm", cm[ias:ias + v].conj(), ppbs[a], cq[ias:ias + v])
        ias += v
        ims += m
    out = orc.zmel(nv, md, phase, ppb, cm, cq)
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(orc.zmel(nv, md, phase, ppb, cm, cq, conjg=True), ref.conj(), rtol=1e-13, atol=1e-13)
    # M->E with a unitary m2e preserves the column norms
    m2e, _ = np.linalg.qr(crand(rng, ngb, ngb))
    e = orc.zmel(nv, md, phase, ppb, cm, cq, m2e=np.asfortranarray(m2e))
    np.testing.assert_allclose(e, np.einsum("IJ,Jnm->Inm", m2e.conj().T, ref), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(np.linalg.norm(e, axis=0), np.linalg.norm(ref, axis=0), rtol=1e-12)


# ---------------------------------------------------------------- x0gemm
def _x0_lists(rng, npairs, nwhis, ns_range, nq):
    itc = rng.integers(ns_range[0] - 1, ns_range[1] + 2, npairs).astype(np.int32)  # includes out-of-range
    itpc = rng.integers(1, nq + 1, npairs).astype(np.int32)
    iwini = rng.integers(1, nwhis - 2, npairs).astype(np.int32)
    nhw = rng.integers(1, 4, npairs).astype(np.int32)
    iwend = (iwini + nhw - 1).astype(np.int32)
    icouini = np.concatenate([[1], 1 + np.cumsum(nhw)[:-1]]).astype(np.int32)
    whwc = rng.uniform(0.1, 1.0, int(nhw.sum()))
    jpmc = np.ones(npairs, np.int32)
    return itc, itpc, iwini, iwend, icouini, whwc, jpmc


def test_x0gemm_bruteforce(orc):
    rng = np.random.default_rng(9)
    npr, nwhis, ns1, ns2, nq, npairs = 11, 8, 2, 5, 3, 40
    z = crand(rng, npr, ns2 - ns1 + 1, nq)
    itc, itpc, iwini, iwend, icouini, whwc, jpmc = _x0_lists(rng, npairs, nwhis, (ns1, ns2), nq)
    for ipr_col, npr_col in ((1, npr), (4, 5)):
        rc = np.zeros((npr, npr_col, nwhis + 1), complex, order="F")
        lo, hi = 3, npairs - 2
        orc.x0gemm(rc, npr, ipr_col, npr_col, nwhis, 1, ns1, ns2, z, lo, hi, iwini, iwend, itc, itpc, jpmc,
                   icouini, whwc)
        ref = np.zeros_like(rc)
        for ic in range(lo - 1, hi):
            it, itp = itc[ic], itpc[ic]
            if it < ns1 or it > ns2:
                continue
            col = z[:, it - ns1, itp - 1]
            for iw in range(iwini[ic], iwend[ic] + 1):
                w = whwc[iw - iwini[ic] + icouini[ic] - 1]
                ref[:, :, iw] += np.outer(col.conj(), col[ipr_col - 1:ipr_col - 1 + npr_col]) * w
        np.testing.assert_allclose(rc, ref, rtol=1e-13, atol=1e-13)
        assert np.all(rc[:, :, 0] == 0)


def test_x0gemm_npm2_negative_bins(orc):
    rng = np.random.default_rng(10)
    npr, nwhis, nq = 5, 6, 2
    z = crand(rng, npr, 2, nq)
    itc = np.array([1, 2, 1], np.int32)
    itpc = np.array([1, 2, 2], np.int32)
    jpmc = np.array([1, 2, 2], np.int32)
    iwini = np.array([2, 3, 5], np.int32)
    iwend = np.array([2, 4, 5], np.int32)
    icouini = np.array([1, 2, 4], np.int32)
    whwc = np.array([0.5, 0.25, 0.75, 2.0])
    rc = np.zeros((npr, npr, 2 * nwhis + 1), complex, order="F")
    orc.x0gemm(rc, npr, 1, npr, nwhis, 2, 1, 2, z, 1, 3, iwini, iwend, itc, itpc, jpmc, icouini, whwc)
    ref = np.zeros_like(rc)
    o = lambda it, itp: np.outer(z[:, it - 1, itp - 1].conj(), z[:, it - 1, itp - 1])
    ref[:, :, nwhis + 2] += 0.5 * o(1, 1)
    ref[:, :, nwhis - 3] += 0.25 * o(2, 2)
    ref[:, :, nwhis - 4] += 0.75 * o(2, 2)
    ref[:, :, nwhis - 5] += 2.0 * o(1, 2)
    np.testing.assert_allclose(rc, ref, rtol=1e-14, atol=1e-14)


# ---------------------------------------------------------------- dpsion / chi0
def test_dpsion_and_composed_form(orc):
    rng = np.random.default_rng(11)
    npr, ns, nq = 7, 3, 4
    fr = orc.frhis(1.0)[1]
    nwhis = len(fr) - 1
    nu = np.array([0.0, 0.2, 0.9, 4.0])
    im = orc.imatt(nu, fr)
    z = crand(rng, npr, ns, nq)
    itc, itpc, iwini, iwend, icouini, whwc, jpmc = _x0_lists(rng, 30, nwhis, (1, ns), nq)
    itc = np.clip(itc, 1, ns).astype(np.int32)
    rc = np.zeros((npr, npr, nwhis + 1), complex, order="F")
    orc.x0gemm(rc, npr, 1, npr, nwhis, 1, 1, ns, z, 1, 30, iwini, iwend, itc, itpc, jpmc, icouini, whwc)
    rc0 = rc.copy()
    chi = orc.dpsion_imag(rc, fr, im)
    # numpy restatement of the scaling and the bin GEMM
    hl, hr = fr[:-1], fr[1:]
    scale = -1.0 / (hr - hl)
    ref = np.einsum("ijb,vb->ijv", rc0[:, :, 1:] * scale[None, None, :], im)
    np.testing.assert_allclose(chi, ref, rtol=1e-12, atol=1e-14)
    # composed direct-frequency weights give the same chi
    cols, wts = [], []
    for ic in range(30):
        w = np.zeros(len(nu))
        for iw in range(iwini[ic], iwend[ic] + 1):
            w += whwc[iw - iwini[ic] + icouini[ic] - 1] * scale[iw - 1] * im[:, iw - 1]
        cols.append(z[:, itc[ic] - 1, itpc[ic] - 1])
        wts.append(w)
    chi2 = orc.chi0_freq(np.array(cols).T, np.array(wts).T)
    np.testing.assert_allclose(chi2, chi, rtol=1e-11, atol=1e-14)


def test_chi_delta_limit_is_lindhard(orc):
    """A pair of weight pi*w in a narrow window at D gives chi(i nu) ->
    -w * 2D/(D^2+nu^2) |z|^2 (the reference's sign convention) as the
    histogram mesh is refined; the error is first order in bin width / D."""
    D, w = 0.83, 0.125
    nu = np.array([0.0, 0.3, 1.0, 3.0])
    lind = -w * 2 * D / (D ** 2 + nu ** 2)
    errs = []
    for ratio in (1.03, 1.003):
        fr = orc.frhis(2.0, ratio, 1e-5)[1]
        nwhis = len(fr) - 1
        im = orc.imatt(nu, fr)
        lo, hi = D * 0.9999, D * 1.0001
        b0 = np.searchsorted(fr, lo, "right")
        b1 = np.searchsorted(fr, hi, "right")
        whw = [math.pi * w * (min(hi, fr[b]) - max(lo, fr[b - 1])) / (hi - lo) for b in range(b0, b1 + 1)]
        z = np.ones((1, 1, 1), complex, order="F")
        rc = np.zeros((1, 1, nwhis + 1), complex, order="F")
        orc.x0gemm(rc, 1, 1, 1, nwhis, 1, 1, 1, z, 1, 1, [b0], [b1], [1], [1], [1], [1], whw)
        chi = orc.dpsion_imag(rc, fr, im)[0, 0]
        errs.append(np.abs(chi.real / lind - 1).max())
        assert np.abs(chi.imag).max() < 1e-12
    assert errs[0] < 1e-2
    assert errs[1] < 1.5e-3 and errs[1] < errs[0] / 3


# ---------------------------------------------------------------- W build
def test_zinv(orc):
    rng = np.random.default_rng(12)
    a = crand(rng, 37, 37)
    np.testing.assert_allclose(orc.zinv(a), np.linalg.inv(a), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("xm,wm", [(True, True), (False, False), (True, False)])
def test_wc_build_textbook(orc, xm, wm):
    rng = np.random.default_rng(13)
    n, nw = 23, 3
    vc = rng.uniform(0.5, 30.0, n)
    Q, _ = np.linalg.qr(crand(rng, n, n))
    chi = np.zeros((n, n, nw), complex, order="F")
    for i in range(nw):
        X = crand(rng, n, 6) * 0.05
        chi[:, :, i] = -X @ X.conj().T
    wc = orc.wc_build(chi, np.sqrt(vc), np.asfortranarray(Q), nspin=1, is_x0_m_basis=xm, is_wc_m_basis=wm)
    for i in range(nw):
        c = 2.0 * chi[:, :, i]
        cE = Q.conj().T @ c @ Q if xm else c
        v = np.diag(vc)
        # W_c = (v^-1 - chi)^-1 - v  (Dyson form, different algebra from the reference's)
        We = np.linalg.inv(np.linalg.inv(v) - cE) - v
        W = Q @ We @ Q.conj().T if wm else We
        np.testing.assert_allclose(wc[:, :, i], W, rtol=1e-9, atol=1e-9 * np.abs(W).max())
        # Hermitian, negative semidefinite when chi is
        assert np.abs(W - W.conj().T).max() < 1e-9 * np.abs(W).max()
        assert np.linalg.eigvalsh(0.5 * (We + We.conj().T)).max() < 1e-9 * np.abs(W).max()


# ---------------------------------------------------------------- Sigma_c weights
def _wint_exact(we, esmr, ua, f):
    sig2 = 2 * (0.5 * esmr) ** 2
    g = lambda w: -we / math.pi * (1 - math.exp(-(w * w + we * we) / sig2)) / (w * w + we * we) * f(w)
    pts = [0.0, abs(we), 10 * math.sqrt(sig2)]
    a = integrate.quad(g, 0, 50, points=pts, limit=2000, epsabs=1e-13, epsrel=1e-12)[0]
    b = integrate.quad(g, 50, np.inf, limit=500, epsabs=1e-13)[0]
    return a + b


@pytest.mark.parametrize("we_ry", [-1.3, -0.21, -0.004, 0.0007, 0.05, 0.9, 2.4])
def test_wgtim_exact_on_singular_part(orc, we_ry):
    """For f(w') = exp(-ua^2 w'^2) the weights reproduce the integral
    -(1/pi) int_0^inf we (1-exp(-(w'^2+we^2)/sig2))/(w'^2+we^2) f dw' exactly
    (the subtracted singular part is integrated analytically, Eq.(57) of
    PRB 76, 165106)."""
    niw, ua, esmr = 16, 1.0, 0.003
    x, fw, wt, ex = orc.freq01(niw, ua)
    omega = np.array([we_ry])  # ekc = 0 -> we = 0.5*omega (Hartree)
    g = orc.wgtim(omega, [0.0], esmr, 1.0, x, wt, ex, ua=ua)[:, 0, 0]
    we = 0.5 * we_ry
    f = lambda w: math.exp(-ua * ua * w * w)
    lhs = g[0] * 1.0 + (g[1:] * np.exp(-ua * ua * fw * fw)).sum()
    assert lhs == pytest.approx(_wint_exact(we, esmr, ua, f), rel=1e-9, abs=1e-12)


def test_wgtim_smooth_integrand_converges(orc):
    ua, esmr, we = 1.0, 0.003, 0.37
    f = lambda w: 1.0 / (1.0 + (w / 0.8) ** 2)
    exact = _wint_exact(we, esmr, ua, f)
    errs = []
    for niw in (12, 24, 48):
        x, fw, wt, ex = orc.freq01(niw, ua)
        g = orc.wgtim(np.array([2 * we]), [0.0], esmr, 1.0, x, wt, ex, ua=ua)[:, 0, 0]
        errs.append(abs(g[0] * f(0.0) + (g[1:] * np.array([f(v) for v in fw])).sum() - exact))
    assert errs[2] < 1e-6 and errs[2] <= errs[0]


def test_wgtim_wkkr_and_layout(orc):
    x, fw, wt, ex = orc.freq01(12, 1.0)
    om = np.array([0.1, -0.3])
    ek = np.array([-0.5, 0.2, 1.1])
    g1 = orc.wgtim(om, ek, 0.003, 1.0, x, wt, ex)
    g2 = orc.wgtim(om, ek, 0.003, 0.25, x, wt, ex)
    np.testing.assert_allclose(g2, 0.25 * g1, rtol=1e-15)
    single = orc.wgtim(om[1:], ek[2:], 0.003, 1.0, x, wt, ex)
    np.testing.assert_allclose(g1[:, 2, 1], single[:, 0, 0], rtol=1e-15)


# ---------------------------------------------------------------- Sigma_c
def test_sigc_imag_gemm_chain(orc):
    rng = np.random.default_rng(14)
    ngb, nm, ntq, niw = 9, 4, 3, 2
    z = crand(rng, ngb, nm, ntq)
    wc = crand(rng, ngb, ngb, niw + 1)
    g = rng.standard_normal((niw + 1, nm, ntq))
    zs0 = crand(rng, ntq, ntq) * 0.1
    out = orc.sigc_imag(z, wc, np.asfortranarray(g), zs0)
    ref = zs0.copy()
    cz = np.zeros((ngb, nm, ntq), complex)
    for iw in range(niw + 1):
        # keepwv unpack: wc(i,j)=upper, then wc(j,i)=conj(upper) -- for i=j the
        # second store wins, so the diagonal is conjugated (S/m_sxcf_gemm.f90:471-472)
        U1 = np.triu(wc[:, :, iw], 1)
        Wh = U1 + U1.conj().T + np.diag(wc[:, :, iw].diagonal().conj())
        cz += np.einsum("JI,Jnm->Inm", Wh.conj(), z * g[iw][None])
    ref += np.einsum("Inm,Inp->mp", cz.conj(), z)
    d = np.diag(ref).copy()
    d = d.real + 1j * np.minimum(d.imag, 0.0)
    np.fill_diagonal(ref, d)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-13)


def test_qp_energy_and_herm(orc):
    rng = np.random.default_rng(15)
    n = 5
    sec = crand(rng, n, n)
    elda, sex, sexc, vxc = (rng.standard_normal(n) for _ in range(4))
    h = 2 * 13.605693122994
    np.testing.assert_allclose(orc.qp_energy(elda, sex, sexc, sec, vxc),
                               elda + (sex + sexc) * h + sec.diagonal().real * h - vxc, rtol=1e-14)
    a, b, c = crand(rng, n, n), crand(rng, n, n), crand(rng, n, n)
    se = orc.sigma_herm(a, b, sec, c)
    np.testing.assert_allclose(se, a + b + 0.5 * (sec + sec.conj().T) - c, rtol=1e-14)


# ---------------------------------------------------------------- synth inputs
def test_synth_index_lists_consistent():
    qp = synth.make_qpoint(synth.CONFIGS["T2"])
    fr = synth.frhis(qp)
    pw = synth.pair_weights(qp, fr)
    ix = synth.x0kf_index(pw)
    cfg = qp.cfg
    assert len(ix.itc) == cfg.nk * cfg.no * cfg.nu
    # total weight per pair is pi * w_k
    for ic in range(0, len(ix.itc), 7):
        s = ix.whwc[ix.icouini[ic] - 1: ix.icouini[ic] - 1 + ix.iwend[ic] - ix.iwini[ic] + 1].sum()
        assert s == pytest.approx(math.pi * qp.wk[ix.kc[ic] - 1], rel=1e-12)
    assert ix.iwini.min() >= 1 and ix.iwend.max() <= len(fr) - 1
    assert np.all(ix.itpc >= 1) and np.all(ix.itpc <= cfg.nu)
