"""Mixed-precision forms (complex64 storage; FP32 MFMA or bf16x3 split MFMA,
FP32 accumulation) against the FP64 oracle.  Bars, stated per result as
max|gpu - oracle| / max|oracle|:
  f32    : 2e-5   (FP32 rounding of data and of K-long sums)
  bf16x3 : 1e-4   (hi/lo bf16 split keeps ~16 mantissa bits, 2^-16 ~ 1.5e-5 per product)
and |eQP(gpu) - eQP(oracle)| < 1 meV for whole q-points, the accuracy bar of
the mixed form."""
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu
TOL = {"f32": 2e-5, "bf16x3": 1e-4}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.fixture(scope="module", params=["f32", "bf16x3"])
def hp(request):
    import ecalj_amd

    hd = ecalj_amd.Handle(0, request.param)
    yield hd, request.param
    hd.close()


def test_zmel_mixed(hp, orc):
    h, prec = hp
    say("zmel", prec)
    rng = np.random.default_rng(21)
    nv, md = [5, 9, 3], [7, 11, 4]
    ppb = np.concatenate([rng.standard_normal((v, v, m)).ravel(order="F") for v, m in zip(nv, md)])
    phase = np.exp(1j * rng.uniform(0, 6, 3))
    h.set_ppb(nv, md, ppb)
    m2e, _ = np.linalg.qr(crand(rng, 22, 22))
    h.set_coulomb(rng.uniform(0.1, 5.0, 22), np.asfortranarray(m2e))
    # both contraction orders: occupied side first (nm <= nq) and end side first (nm > nq)
    for nmc, nqc in ((4, 37), (37, 4)):
        cm, cq = crand(rng, 19, nmc, 3), crand(rng, 19, nqc, 3)
        for is_m, conj in ((True, True), (False, False), (True, False)):
            g = h.zmel(cm, cq, phase=phase, is_m_basis=is_m, zmelconjg=conj)
            for b in range(3):
                ref = orc.zmel(nv, md, phase, ppb, cm[:, :, b], cq[:, :, b], m2e=None if is_m else m2e, conjg=conj)
                assert relerr(g[..., b], ref) < TOL[prec], (nmc, nqc, is_m, conj)


def test_x0gemm_mixed(hp, orc):
    h, prec = hp
    say("x0gemm", prec)
    rng = np.random.default_rng(22)
    npr, nwhis, ns1, ns2, nq, npairs = 97, 10, 1, 4, 6, 400
    z = crand(rng, npr, ns2 - ns1 + 1, nq)
    itc = rng.integers(1, ns2 + 1, npairs).astype(np.int32)
    itpc = rng.integers(1, nq + 1, npairs).astype(np.int32)
    iwini = rng.integers(1, nwhis - 2, npairs).astype(np.int32)
    nhw = rng.integers(1, 3, npairs).astype(np.int32)
    iwend = (iwini + nhw - 1).astype(np.int32)
    icouini = np.concatenate([[1], 1 + np.cumsum(nhw)[:-1]]).astype(np.int32)
    whwc = rng.uniform(0.1, 1.0, int(nhw.sum()))
    jpmc = np.ones(npairs, np.int32)
    r0 = np.zeros((npr, npr, nwhis + 1), complex, order="F")
    args = (npr, 1, npr, nwhis, 1, ns1, ns2, z, 1, npairs, iwini, iwend, itc, itpc, jpmc, icouini, whwc)
    g = h.x0gemm(r0.copy(order="F"), *args)
    o = orc.x0gemm(r0.copy(order="F"), *args)
    assert relerr(g, o) < TOL[prec]


@pytest.mark.parametrize("npm,ipr_col,npr_col", [(2, 1, 97), (1, 5, 40), (2, 9, 88)])
def test_x0gemm_mixed_blocks(hp, orc, npm, ipr_col, npr_col):
    """Column blocks of rcxq (ipr_col > 1: no Hermitian mirror) and both
    signs of the frequency (npm = 2: slices -nwhis..nwhis)."""
    h, prec = hp
    say("x0gemm blocks", prec, npm, ipr_col, npr_col)
    rng = np.random.default_rng(100 + 10 * npm + ipr_col)
    npr, nwhis, ns1, ns2, nq, npairs = 97, 9, 2, 5, 5, 300
    z = crand(rng, npr, ns2 - ns1 + 1, nq)
    itc = rng.integers(ns1, ns2 + 1, npairs).astype(np.int32)
    itpc = rng.integers(1, nq + 1, npairs).astype(np.int32)
    iwini = rng.integers(1, nwhis - 2, npairs).astype(np.int32)
    nhw = rng.integers(1, 3, npairs).astype(np.int32)
    iwend = (iwini + nhw - 1).astype(np.int32)
    icouini = np.concatenate([[1], 1 + np.cumsum(nhw)[:-1]]).astype(np.int32)
    whwc = rng.uniform(0.1, 1.0, int(nhw.sum()))
    jpmc = rng.integers(1, npm + 1, npairs).astype(np.int32)
    nsl = nwhis + 1 if npm == 1 else 2 * nwhis + 1
    r0 = np.asfortranarray(crand(rng, npr, npr_col, nsl) * 0.1)
    args = (npr, ipr_col, npr_col, nwhis, npm, ns1, ns2, z, 1, npairs, iwini, iwend, itc, itpc, jpmc, icouini, whwc)
    g = h.x0gemm(r0.copy(order="F"), *args)
    o = orc.x0gemm(r0.copy(order="F"), *args)
    assert relerr(g, o) < TOL[prec]


def test_x0gemm_bf16x3_groups(orc, monkeypatch):
    """The bf16x3 x0gemm gathers each bin's entries into scratch planes in
    groups of bounded size; with tiny groups (64 entries) bins are split over
    groups and accumulate into rcxq: same result as one group, to rounding."""
    import ecalj_amd

    rng = np.random.default_rng(23)
    npr, nwhis, ns1, ns2, nq, npairs = 130, 7, 1, 4, 6, 500
    z = crand(rng, npr, ns2 - ns1 + 1, nq)
    itc = rng.integers(1, ns2 + 1, npairs).astype(np.int32)
    itpc = rng.integers(1, nq + 1, npairs).astype(np.int32)
    iwini = rng.integers(1, nwhis - 2, npairs).astype(np.int32)
    nhw = rng.integers(1, 3, npairs).astype(np.int32)
    iwend = (iwini + nhw - 1).astype(np.int32)
    icouini = np.concatenate([[1], 1 + np.cumsum(nhw)[:-1]]).astype(np.int32)
    whwc = rng.uniform(0.1, 1.0, int(nhw.sum()))
    jpmc = np.ones(npairs, np.int32)
    r0 = np.zeros((npr, npr, nwhis + 1), complex, order="F")
    args = (npr, 1, npr, nwhis, 1, ns1, ns2, z, 1, npairs, iwini, iwend, itc, itpc, jpmc, icouini, whwc)
    o = orc.x0gemm(r0.copy(order="F"), *args)
    h = ecalj_amd.Handle(0, "bf16x3")
    try:
        one = h.x0gemm(r0.copy(order="F"), *args)
        monkeypatch.setenv("ECALJ_X0_GROUP_ENTRIES", "64")
        say("x0gemm bf16x3 in groups of 64 entries")
        many = h.x0gemm(r0.copy(order="F"), *args)
    finally:
        h.close()
    assert relerr(one, o) < TOL["bf16x3"]
    assert relerr(many, o) < TOL["bf16x3"]
    assert relerr(many, one) < 1e-5


# (ngb, nm, ntq, niw): ragged sizes, tile-aligned sizes, and more than 32
# frequencies (the bf16x3 kernel then accumulates over two launches)
@pytest.mark.parametrize("ngb,nm,ntq,niw", [(130, 21, 8, 12), (64, 64, 2, 10), (161, 13, 5, 40), (1, 3, 1, 2)])
def test_sigc_mixed(hp, orc, ngb, nm, ntq, niw):
    h, prec = hp
    say("sigc", prec, ngb, nm, ntq, niw)
    rng = np.random.default_rng(ngb)
    z = crand(rng, ngb, nm, ntq)
    wc = crand(rng, ngb, ngb, niw + 1)
    g = np.asfortranarray(rng.standard_normal((niw + 1, nm, ntq)))
    assert relerr(h.sigc(z, wc, g), orc.sigc_imag(z, wc, g)) < TOL[prec]


@pytest.mark.parametrize("mode", [5, 6, 7, 8, 9, 10, 11, 12])
@pytest.mark.parametrize("ngb,nm,ntq,niw", [(130, 21, 8, 12), (64, 64, 2, 10), (161, 13, 5, 40), (1, 3, 1, 2),
                                            (300, 40, 4, 0)])
def test_sigc_bf3_frequency_inner_forms(orc, mode, ngb, nm, ntq, niw):
    """The frequency-inner variants, forced (5, 6: three-product form; 7, 8:
    four-product form; 9, 10: three-product form, 8 waves side by side; 11,
    12: three-product form on the 32x32x16 MFMA, 4 waves): ragged and
    tile-aligned shapes, more than 24 frequencies (two launches that
    accumulate), one frequency, one row."""
    import ecalj_amd

    say("sigc bf16x3 3m variant", mode, ngb, nm, ntq, niw)
    rng = np.random.default_rng(ngb + niw)
    z = crand(rng, ngb, nm, ntq)
    wc = crand(rng, ngb, ngb, niw + 1)
    g = np.asfortranarray(rng.standard_normal((niw + 1, nm, ntq)))
    h = ecalj_amd.Handle(0, "bf16x3")
    try:
        h.set_tuning(mode)
        assert relerr(h.sigc(z, wc, g), orc.sigc_imag(z, wc, g)) < TOL["bf16x3"]
    finally:
        h.close()


def test_sigc_bf3_tuner_variants(orc):
    """Every kernel variant of the tuner gives the oracle's Sigma_c; the variants
    of one form (four products frequency-outer; three products in either wave
    layout; three products on the 32x32x16 MFMA; four products
    frequency-inner) are bit-identical to each other
    (same sums in the same order);
    auto mode records its choice and returns the chosen variant's result."""
    import ecalj_amd

    rng = np.random.default_rng(7)
    ngb, nm, ntq, niw = 203, 19, 9, 11
    z = crand(rng, ngb, nm, ntq)
    wc = crand(rng, ngb, ngb, niw + 1)
    g = np.asfortranarray(rng.standard_normal((niw + 1, nm, ntq)))
    ref = orc.sigc_imag(z, wc, g)
    h = ecalj_amd.Handle(0, "bf16x3")
    try:
        outs = []
        for mode in (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12):
            say("sigc bf16x3 tuner variant", mode)
            h.set_tuning(mode)
            outs.append(h.sigc(z, wc, g))
            assert relerr(outs[-1], ref) < TOL["bf16x3"]
        for o in outs[1:4]:
            assert np.array_equal(outs[0], o)
        assert np.array_equal(outs[4], outs[5])
        assert np.array_equal(outs[6], outs[7])
        assert np.array_equal(outs[8], outs[9])
        assert np.array_equal(outs[4], outs[8])  # three products: 2 x 4 and 1 x 8 wave layouts
        assert np.array_equal(outs[10], outs[11])  # three products on the 32x32x16 MFMA
        h.set_tuning(0)
        say("sigc bf16x3 tuner auto")
        a = h.sigc(z, wc, g)
        b = h.sigc(z, wc, g)  # cached choice
        rep = h.tuning_report()
        assert rep.count("sigc_bf3") == 1 and "*" in rep
        chosen = [ln for ln in rep.splitlines() if "sigc_bf3" in ln][0]
        star = chosen[:chosen.index("*")].split()[-3]  # the name before "<ms> ms *"
        expect = (outs[10] if star.startswith("3m32") else outs[8] if "/w8" in star else
                  outs[4] if star.startswith("3m") else outs[6] if star.startswith("vin4") else outs[0])
        assert np.array_equal(a, expect) and np.array_equal(b, expect)
    finally:
        h.close()


@pytest.mark.parametrize("name", ["T3", "C1"])
@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
def test_qpoint_mixed_qp_energy_within_1meV(orc, name, prec):
    from ecalj_amd.qsgw import QPointGPU
    from oracle import pipeline

    say("qpoint mixed", name, prec)
    qp = synth.make_qpoint(synth.CONFIGS[name])
    ref = pipeline.run(qp)
    g = QPointGPU(qp, prec)
    g.chi0()
    g.wbuild()
    g.make_wk()
    g.sigc()
    chi, wc, zsec, eqp = g.results()
    err_mev = np.abs(eqp - ref.eqp).max() * 1e3
    say(f"  max |dEqp| = {err_mev:.4f} meV, chi relerr {relerr(chi, ref.chi):.2e}, W relerr {relerr(wc, ref.wc):.2e}")
    assert err_mev < 1.0
    assert relerr(chi, ref.chi) < TOL[prec]


@pytest.mark.parametrize("real", [False, True])
def test_wc_build_mixed(hp, orc, real):
    """The reference's __MP split of the W build: M <-> E transforms on the
    precision's GEMM core (complex64 chi_E, W_E), epsilon and its inverse in
    FP64.  Bar: the storage rounding of chi_E and W_E times cond(eps)."""
    h, prec = hp
    n, nw = 150, 3
    say("wc_build mixed", prec, n, nw, real)
    rng = np.random.default_rng(31)
    vc = rng.uniform(0.5, 40.0, n)
    Q, _ = np.linalg.qr(crand(rng, n, n))
    Q = np.asfortranarray(Q)
    chi = np.zeros((n, n, nw), complex, order="F")
    for i in range(nw):
        X = crand(rng, n, 40) * 0.01 / (1 + i)
        chi[:, :, i] = -(X @ X.conj().T)
        if real:
            chi[:, :, i] += 0.3j * (X[:, :20] @ X[:, :20].conj().T)
    h.set_coulomb(vc, Q)
    for wm in (True, False):
        g = (h.wc_build_real if real else h.wc_build)(chi, nspin=1, is_x0_m_basis=True, is_wc_m_basis=wm)
        o = orc.wc_build(chi, np.sqrt(vc), Q, nspin=1, is_x0_m_basis=True, is_wc_m_basis=wm)
        kap = 1.0
        s = np.sqrt(vc)
        for i in range(nw):
            e = np.eye(n) - s[:, None] * (Q.conj().T @ (2 * chi[:, :, i]) @ Q) * s[None, :]
            kap = max(kap, np.linalg.cond(e))
        err = relerr(g, o)
        say(f"  wm={wm} relerr {err:.2e} cond {kap:.1f}")
        assert err < TOL[prec] * kap
