"""Parity of the HIP library (through its C ABI) with the FP64 CPU oracle.

Bar for the FP64 path: max|gpu - oracle| <= 1e-10 * max|oracle| per result
(norm-wise relative), index handling bit-exact.  Every case prints its name
and shape before touching the GPU."""
import math
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-10


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


# ------------------------------------------------------------------ zmel
@pytest.mark.parametrize("nv,md,nm,nq,nbatch", [
    ([12], [16], 2, 5, 1),       # m-first
    ([12], [16], 7, 3, 1),       # q-first
    ([5, 9, 3], [7, 11, 4], 4, 6, 3),
    ([70], [130], 9, 40, 2),
])
def test_zmel(h, orc, nv, md, nm, nq, nbatch):
    say("zmel", nv, md, nm, nq, nbatch)
    rng = np.random.default_rng(sum(nv) + nm)
    ndima = sum(nv) + 2
    ppbs = [rng.standard_normal((v, v, m)) for v, m in zip(nv, md)]
    ppb = np.concatenate([p.ravel(order="F") for p in ppbs])
    phase = np.exp(1j * rng.uniform(0, 6, len(nv)))
    h.set_ppb(nv, md, ppb)
    ngb = sum(md)
    m2e, _ = np.linalg.qr(crand(rng, ngb, ngb))
    h.set_coulomb(rng.uniform(0.1, 5.0, ngb), np.asfortranarray(m2e))
    cm = crand(rng, ndima, nm, nbatch)
    cq = crand(rng, ndima, nq, nbatch)
    for is_m, conj in ((True, False), (True, True), (False, True)):
        g = h.zmel(cm if nbatch > 1 else cm[:, :, 0], cq if nbatch > 1 else cq[:, :, 0], phase=phase,
                   is_m_basis=is_m, zmelconjg=conj)
        for b in range(nbatch):
            ref = orc.zmel(nv, md, phase, ppb, cm[:, :, b], cq[:, :, b], m2e=None if is_m else m2e, conjg=conj)
            gb = g[..., b] if nbatch > 1 else g
            assert relerr(gb, ref) < TOL


# ------------------------------------------------------------------ x0gemm
def _lists(rng, npairs, nwhis, ns1, ns2, nq, npm):
    itc = rng.integers(ns1 - 1, ns2 + 2, npairs).astype(np.int32)
    itpc = rng.integers(1, nq + 1, npairs).astype(np.int32)
    iwini = rng.integers(1, nwhis - 2, npairs).astype(np.int32)
    nhw = rng.integers(1, 4, npairs).astype(np.int32)
    iwend = (iwini + nhw - 1).astype(np.int32)
    icouini = np.concatenate([[1], 1 + np.cumsum(nhw)[:-1]]).astype(np.int32)
    whwc = rng.uniform(0.1, 1.0, int(nhw.sum()))
    jpmc = rng.integers(1, npm + 1, npairs).astype(np.int32)
    return iwini, iwend, itc, itpc, jpmc, icouini, whwc


@pytest.mark.parametrize("npr,ipr_col,npr_col,npm,npairs", [
    (70, 1, 70, 1, 300),     # square: upper tiles + mirror
    (70, 9, 40, 1, 200),     # column block of a rank
    (33, 1, 33, 2, 150),     # negative-frequency bins
    (130, 1, 130, 1, 3),     # nearly empty bins
])
def test_x0gemm(h, orc, npr, ipr_col, npr_col, npm, npairs):
    say("x0gemm", npr, ipr_col, npr_col, npm, npairs)
    rng = np.random.default_rng(npr + npairs)
    nwhis, ns1, ns2, nq = 12, 2, 6, 5
    z = crand(rng, npr, ns2 - ns1 + 1, nq)
    iwini, iwend, itc, itpc, jpmc, icouini, whwc = _lists(rng, npairs, nwhis, ns1, ns2, nq, npm)
    nsl = nwhis + 1 if npm == 1 else 2 * nwhis + 1
    r0 = crand(rng, npr, npr_col, nsl) * 0.01
    lo, hi = 2, npairs
    rg = h.x0gemm(r0.copy(order="F"), npr, ipr_col, npr_col, nwhis, npm, ns1, ns2, z, lo, hi, iwini, iwend, itc,
                  itpc, jpmc, icouini, whwc)
    ro = orc.x0gemm(r0.copy(order="F"), npr, ipr_col, npr_col, nwhis, npm, ns1, ns2, z, lo, hi, iwini, iwend, itc,
                    itpc, jpmc, icouini, whwc)
    assert relerr(rg, ro) < TOL


def test_x0gemm_rejects_bad_index(h):
    import ecalj_amd

    say("x0gemm bad index")
    z = np.zeros((8, 2, 3), np.complex128, order="F")
    r = np.zeros((8, 8, 5), np.complex128, order="F")
    with pytest.raises(ecalj_amd.EcaljError):  # iwend beyond nwhis is refused before any launch
        h.x0gemm(r, 8, 1, 8, 4, 1, 1, 2, z, 1, 1, [3], [5], [1], [1], [1], [1], [1.0, 1.0, 1.0])


# ------------------------------------------------------------------ dpsion / chi0
def test_dpsion_imag(h, orc):
    say("dpsion_imag")
    rng = np.random.default_rng(3)
    fr = orc.frhis(1.2)[1]
    nwhis = len(fr) - 1
    nu = np.array([0.0, 0.05, 0.4, 2.0, 9.0])
    r = crand(rng, 37, 29, nwhis + 1)
    g_r = r.copy(order="F")
    zg = h.dpsion_imag(g_r, fr, nu, scale_rcxq=True)
    o_r = r.copy(order="F")
    zo = orc.dpsion_imag(o_r, fr, orc.imatt(nu, fr))
    assert relerr(zg, zo) < TOL
    assert relerr(g_r[:, :, 1:], o_r[:, :, 1:]) < 1e-14


def test_chi0_freq(h, orc):
    say("chi0_freq")
    rng = np.random.default_rng(4)
    z = crand(rng, 83, 211)
    w = rng.standard_normal((7, 211))
    assert relerr(h.chi0_freq(z, w), orc.chi0_freq(z, w)) < TOL


# ------------------------------------------------------------------ W
# large n, including a ragged last pivot block (520 = 16 x 32 + 8)
@pytest.mark.parametrize("n,nw,xm,wm,nspin", [(72, 3, True, True, 1), (100, 2, False, False, 2),
                                             (33, 4, True, False, 1), (150, 2, True, True, 1),
                                             (520, 2, True, True, 1), (640, 1, False, True, 1)])
def test_wc_build(h, orc, n, nw, xm, wm, nspin):
    say("wc_build", n, nw, xm, wm, nspin)
    rng = np.random.default_rng(n)
    vc = rng.uniform(0.5, 40.0, n)
    Q, _ = np.linalg.qr(crand(rng, n, n))
    Q = np.asfortranarray(Q)
    chi = np.zeros((n, n, nw), complex, order="F")
    for i in range(nw):
        X = crand(rng, n, n // 3) * 0.03
        chi[:, :, i] = -X @ X.conj().T
    h.set_coulomb(vc, Q)
    g = h.wc_build(chi, nspin=nspin, is_x0_m_basis=xm, is_wc_m_basis=wm)
    o = orc.wc_build(chi, np.sqrt(vc), Q, nspin=nspin, is_x0_m_basis=xm, is_wc_m_basis=wm)
    assert relerr(g, o) < TOL


def test_wc_build_singular_reports_error(h):
    """eps = 1 - v^1/2 chi v^1/2 = 0 when chi = 1/v: refused with ECALJ_ERR_SINGULAR."""
    import ecalj_amd

    say("wc_build singular")
    n = 40
    vc = np.linspace(1.0, 2.0, n)
    h.set_coulomb(vc, np.eye(n, dtype=complex))
    chi = np.zeros((n, n, 1), complex, order="F")
    chi[:, :, 0] = np.diag(1.0 / vc)
    with pytest.raises(ecalj_amd.EcaljError) as e:
        h.wc_build(chi, nspin=2, is_x0_m_basis=False, is_wc_m_basis=False)
    assert e.value.status == 2
    h.sync()  # the flag was cleared: the handle keeps working


# ------------------------------------------------------------------ Sigma_c
def test_sigc_weights(h, orc):
    say("sigc_weights")
    rng = np.random.default_rng(5)
    x, fw, wt, ex = orc.freq01(16, 1.0)
    om = rng.uniform(-1.0, 4.0, 6)
    ek = np.concatenate([rng.uniform(-1.0, 4.0, 20), om[:2] + 1e-5])
    for nctot in (0, 3):
        g = h.sigc_weights(om, ek, 0.003, 0.37, x, wt, ex, ua=1.0, nctot=nctot)
        o = orc.wgtim(om, ek, 0.003, 0.37, x, wt, ex, ua=1.0, nctot=nctot)
        assert relerr(g, o) < TOL


@pytest.mark.parametrize("ngb,nm,ntq,niw", [(40, 7, 3, 6), (130, 21, 8, 12), (64, 64, 2, 10)])
def test_sigc(h, orc, ngb, nm, ntq, niw):
    say("sigc", ngb, nm, ntq, niw)
    rng = np.random.default_rng(ngb + nm)
    z = crand(rng, ngb, nm, ntq)
    wc = crand(rng, ngb, ngb, niw + 1)  # non-Hermitian on purpose: checks the upper-triangle unpack
    g = np.asfortranarray(rng.standard_normal((niw + 1, nm, ntq)))
    zs0 = crand(rng, ntq, ntq)
    assert relerr(h.sigc(z, wc, g, zs0), orc.sigc_imag(z, wc, g, zs0)) < TOL


# ------------------------------------------------------------------ whole q-point
@pytest.mark.parametrize("name", ["T1", "T2", "T3"])
def test_qpoint_pipeline(orc, name):
    from ecalj_amd.qsgw import QPointGPU
    from oracle import pipeline

    say("qpoint", name)
    qp = synth.make_qpoint(synth.CONFIGS[name])
    ref = pipeline.run(qp)
    g = QPointGPU(qp, "f64")
    g.chi0()
    g.wbuild()
    g.make_wk()
    g.sigc()
    chi, wc, zsec, eqp = g.results()
    assert relerr(chi, ref.chi) < TOL
    assert relerr(wc, ref.wc) < TOL
    assert relerr(zsec, ref.zsec) < TOL
    assert np.abs(eqp - ref.eqp).max() < 1e-9  # eV


@pytest.mark.parametrize("name", ["T1", "T2", "T3"])
def test_qpoint_pipeline_real_axis(orc, name):
    """The q-point with the real-axis chain: chi(omega), W_c(omega) (WVRllwR)
    and the pole term of Sigma_c in the same contraction as the imaginary axis."""
    from ecalj_amd.qsgw import QPointGPU
    from oracle import pipeline

    say("qpoint real axis", name)
    qp = synth.make_qpoint(synth.CONFIGS[name])
    ref = pipeline.run(qp, real_axis=True)
    cfg = qp.cfg
    g = QPointGPU(qp, "f64", real_axis=True)
    g.chi0()
    g.wbuild()
    g.make_wk()
    g.sigc()
    chi, wc, zsec, eqp = g.results()
    chi_r, w_r = g.chi_r.download(), g.w_r.download()
    g.h.close()
    e_chi = relerr(chi_r, ref.chi_r)
    assert e_chi < TOL
    s, fac = qp.vcousq, 2.0 if cfg.nspin == 1 else 1.0
    kappa = max(np.linalg.cond(np.eye(cfg.ni) - s[:, None] * (fac * qp.zcousq.conj().T @ ref.chi_r[:, :, i]
                                                                @ qp.zcousq) * s[None, :])
                for i in range(ref.chi_r.shape[2]))
    bar = max(TOL, 10 * kappa * e_chi)
    assert relerr(w_r, ref.w_r) < bar
    assert relerr(zsec, ref.zsec) < bar
    assert np.abs(eqp - ref.eqp).max() < 1e-6  # eV


def test_qpoint_real_axis_mixed_within_1meV():
    from ecalj_amd.qsgw import QPointGPU

    say("qpoint real axis mixed T3")
    qp = synth.make_qpoint(synth.CONFIGS["T3"])
    e = {}
    for prec in ("f64", "bf16x3", "f32"):
        g = QPointGPU(qp, prec, real_axis=True)
        g.chi0()
        g.wbuild()
        g.make_wk()
        g.sigc()
        e[prec] = g.results()[3]
        g.h.close()
    assert np.abs(e["bf16x3"] - e["f64"]).max() < 1e-3
    assert np.abs(e["f32"] - e["f64"]).max() < 1e-3
